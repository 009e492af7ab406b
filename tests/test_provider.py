"""Host-side provider (sonicsolver_amd/host/sf_provider.cpp): recipe names,
time recipes and the compiled explicit plan, through the C ABI in
include/sonic_amd_host.h.  No GPU is touched."""
import pytest

from sonicsolver_amd import (SonicError, describe_plan, recipe_from_pair, resolve_recipe,
                             resolve_time_recipe)


@pytest.mark.parametrize("name,expected", [
    ("weno5Rusanov", (0, 0, 3)), ("teno5Rusanov", (1, 0, 3)),
    ("weno5HLLC", (0, 1, 3)), ("teno5HLLC", (1, 1, 3)),
    # the reference's other registered ids (toString(TermRecipeId)), with
    # the halo of each scheme (2 WENO3, 3 WENO5 / TENO5, 4 WENO7)
    ("weno3Steger", (2, 2, 2)), ("weno3Rusanov", (2, 0, 2)), ("weno3LaxFriedrichs", (2, 3, 2)),
    ("weno3Roe", (2, 4, 2)), ("weno5Steger", (0, 2, 3)), ("weno5LaxFriedrichs", (0, 3, 3)),
    ("weno5Roe", (0, 4, 3)), ("teno5Steger", (1, 2, 3)), ("teno5LaxFriedrichs", (1, 3, 3)),
    ("teno5Roe", (1, 4, 3)), ("weno7Steger", (3, 2, 4)), ("weno7Rusanov", (3, 0, 4)),
    ("weno7LaxFriedrichs", (3, 3, 4)), ("weno7Roe", (3, 4, 4)),
    ("weno3LaxWendroff", (2, 5, 2)), ("weno5LaxWendroff", (0, 5, 3)),
    ("teno5LaxWendroff", (1, 5, 3)), ("weno7LaxWendroff", (3, 5, 4)),
    ("weno3HLLC", (2, 1, 2)), ("weno7HLLC", (3, 1, 4)),
    # the reference's token normalisation: quotes, '_', '-', spaces, case
    ("WENO5_Rusanov", (0, 0, 3)), ('"teno5-hllc"', (1, 1, 3)), (" Weno5 HLLC ", (0, 1, 3)),
])
def test_served_recipes_resolve(name, expected):
    assert resolve_recipe(name) == expected


@pytest.mark.parametrize("name", ["primitiveUpwind1", "PRIMITIVE_UPWIND1"])
def test_reference_recipes_not_served_are_rejected_explicitly(name):
    with pytest.raises(SonicError) as e:
        resolve_recipe(name)
    assert e.value.code == -1 and "not served" in str(e.value)


def test_unknown_recipe_rejected():
    with pytest.raises(SonicError) as e:
        resolve_recipe("weno9Magic")
    assert "Unsupported convection term recipe" in str(e.value)


def test_legacy_pair_form():
    assert recipe_from_pair("TENO5", "Rusanov") == (1, 0, 3)
    assert recipe_from_pair("", "llf") == (0, 0, 3)          # empty scheme selects WENO5
    assert recipe_from_pair("weno5", "HLLC") == (0, 1, 3)
    assert recipe_from_pair("weno5", "") == (0, 2, 3)        # empty flux selects Steger-Warming
    assert recipe_from_pair("weno7", "sw") == (3, 2, 4)
    assert recipe_from_pair("WENO3", "lxf") == (2, 3, 2)
    assert recipe_from_pair("teno5", "Roe") == (1, 4, 3)
    assert recipe_from_pair("weno5", "LaxWendroff") == (0, 5, 3)
    assert recipe_from_pair("weno3", "lw") == (2, 5, 2)
    for scheme, flux in (("weno5", "upwindAdvection"), ("upwind1", "rusanov"), ("weno9", "roe"),
                         ("weno5", "godunov")):
        with pytest.raises(SonicError):
            recipe_from_pair(scheme, flux)


def test_time_recipes():
    assert resolve_time_recipe("forwardEuler") == (0, 1)
    assert resolve_time_recipe("SSPRK3") == (1, 3)
    assert resolve_time_recipe("classical_RK4") == (2, 4)
    assert resolve_time_recipe("") == (2, 4)                  # empty selects classicalRK4
    # obsolete aliases are rejected (test/test_explicitStages.cpp:23-30,89-90)
    for bad in ("Euler", "RK4"):
        with pytest.raises(SonicError):
            resolve_time_recipe(bad)


def test_plan_shape():
    text = describe_plan("weno5HLLC", "SSPRK3")
    lines = [l.strip() for l in text.strip().splitlines()]
    assert lines == [
        "Sequence Explicit.step",
        "Update flow.step.prepare [flow.conservative]",
        "Update flow.dt.compute [flow.conservative]",
        "Update flow.step.begin [flow.conservative]",
        "StageLoop Explicit.stages x3",
        "Update explicit.stage.execute [flow.conservative]",
        "Commit flow.step.commit [flow.conservative]",
        "Commit time.commit [flow.conservative]",
    ]
    assert "StageLoop Explicit.stages x4" in describe_plan("teno5Rusanov", "classicalRK4")
    assert "StageLoop Explicit.stages x1" in describe_plan("teno5Rusanov", "forwardEuler")
