/// @file SF_deviceExplicitOps.cpp
/// @brief The explicit ops of the "flow.conservative" provider on the
/// MI355X library; see SF_deviceExplicitOps.h.  Compiled inside the
/// sonicSolver tree (its include directories plus <repo>/include) and
/// linked with -L<repo>/sonicsolver_amd -lsonic_amd.
#include "SF_deviceExplicitOps.h"

#include "sonic_amd.h"

#include "core/config/types/SF_termRecipe.h"
#include "core/mesh/SF_dimension.h"
#include "core/system/SF_operationIds.h"

#include <cmath>
#include <cstdlib>
#include <set>
#include <stdexcept>
#include <string>

namespace SF::SolverAlgorithm {
namespace {

void check(int rc, const char* what) {
    if (rc == 0) return;
    char msg[512];
    sf_last_error(msg, sizeof msg);
    throw std::runtime_error(std::string(what) + ": " + msg);
}

[[noreturn]] void notServed(const std::string& why) {
    throw std::runtime_error("device explicit path: " + why + "; bind the CPU ops instead");
}

int schemeOf(FDM::ConvectionScheme scheme) {
    switch (scheme) {
    case FDM::ConvectionScheme::WENO3: return SF_SCHEME_WENO3;
    case FDM::ConvectionScheme::WENO5: return SF_SCHEME_WENO5;
    case FDM::ConvectionScheme::TENO5: return SF_SCHEME_TENO5;
    case FDM::ConvectionScheme::WENO7: return SF_SCHEME_WENO7;
    case FDM::ConvectionScheme::Upwind1: break;
    }
    notServed("the primitive upwind recipe is not a characteristic WENO recipe");
}

// By the flux's canonical token (FDM::toString), so that a tree with the
// HLLC recipes registered (hllc_recipes.patch) maps FluxSplitter::HLLC
// without a change here; every other value is rejected, never defaulted.
int fluxOf(FDM::FluxSplitter flux) {
    const std::string t = FDM::toString(flux);
    if (t == "StegerWarming") return SF_FLUX_STEGER_WARMING;
    if (t == "Rusanov") return SF_FLUX_RUSANOV;
    if (t == "LaxFriedrichs") return SF_FLUX_LAX_FRIEDRICHS;
    if (t == "Roe") return SF_FLUX_ROE;
    if (t == "LaxWendroff") return SF_FLUX_LAX_WENDROFF;
    if (t == "HLLC") return SF_FLUX_HLLC;
    notServed("face flux '" + t + "' has no device implementation");
}

int timeOf(FDM::TimeRecipeId id) {
    switch (id) {
    case FDM::TimeRecipeId::ForwardEuler: return SF_TIME_EULER;
    case FDM::TimeRecipeId::SSPRK3: return SF_TIME_SSPRK3;
    case FDM::TimeRecipeId::ClassicalRK4: return SF_TIME_RK4;
    }
    notServed("unknown time recipe");
}

int uniqueNodes(const Field& f, int axis) {
    return axis == 0 ? f.NX() : axis == 1 ? f.NY() : f.NZ();
}

// Whether every node of a boundary set lies on the plane of the first
// (side 0) or last (side 1) unique node of `axis`.
bool onEnd(const Field& f, const std::vector<int>& nodes, int axis, int side) {
    if (nodes.empty()) return false;
    const int plane = f.NG() + (side ? uniqueNodes(f, axis) - 1 : 0);
    for (int idx : nodes) {
        const int i = idx % f.MX(), j = (idx / f.MX()) % f.MY(), k = idx / (f.MX() * f.MY());
        if ((axis == 0 ? i : axis == 1 ? j : k) != plane) return false;
    }
    return true;
}

// How Boundary::Applicator::apply (SF_applicator.h:51-83) treats one end:
// 1 = zeroGradient for density, velocity and pressure (the node is a
// solver-boundary point and is refilled, SF_zeroGradient.cpp:88-130);
// 0 = no setting names a set on it (the applicator neither marks nor
// fills it, SF_applicator.h:162-187).  Anything else is not served.
int endKind(const Field& f, const FDM::BoundaryConfig& bc, int axis, int side) {
    std::set<std::string> names;
    for (const auto& [name, nodes] : f.getAllSets())
        if (onEnd(f, nodes, axis, side)) names.insert(name);
    int zeroGradient[3] = {0, 0, 0};
    auto count = [&](const auto& settings, int slot) {
        for (const auto& s : settings) {
            if (!names.count(s.name)) continue;
            if (s.type != ZERO_GRADIENT)
                notServed("boundary set '" + s.name + "' carries a condition other than zeroGradient");
            ++zeroGradient[slot];
        }
    };
    count(bc.density, 0);
    count(bc.velocity, 1);
    count(bc.pressure, 2);
    for (const auto& s : bc.thermal)
        if (names.count(s.name) && s.type != ThermalBCType::Empty)
            notServed("boundary set '" + s.name + "' carries a thermal condition");
    if (!zeroGradient[0] && !zeroGradient[1] && !zeroGradient[2]) return 0;
    if (zeroGradient[0] && zeroGradient[1] && zeroGradient[2]) return 1;
    notServed("an end is zeroGradient for some of density, velocity and pressure only");
}

// The metric of `axis` at a node, and whether the node's metrics are those
// of a Cartesian grid (no cross terms).
double diagonalMetric(const Field& f, int axis, int i, int j, int k, bool& cartesian) {
    const double m[3][3] = {{f.XiX(i, j, k), f.XiY(i, j, k), f.XiZ(i, j, k)},
                            {f.EtX(i, j, k), f.EtY(i, j, k), f.EtZ(i, j, k)},
                            {f.ZeX(i, j, k), f.ZeY(i, j, k), f.ZeZ(i, j, k)}};
    for (int b = 0; b < 3; ++b)
        if (b != axis && m[axis][b] != 0.0) cartesian = false;
    return m[axis][axis];
}

}  // namespace

bool deviceExplicitRequested(const System::CompiledNumericalSystem& numerics) {
    if (std::getenv("SF_DEVICE_EXPLICIT")) return true;
    return numerics.recipes.convection &&
           std::string(FDM::toString(numerics.recipes.convection->flux())) == "HLLC";
}

int deviceExplicitOrdinal() {
    const char* v = std::getenv("SF_DEVICE_EXPLICIT");
    return v && *v ? std::atoi(v) : 0;
}

DeviceExplicitPath::~DeviceExplicitPath() {
    if (grid_) sf_grid_destroy(grid_);
}

void DeviceExplicitPath::upload(const Field& f) {
    const int ng = f.NG();
    const size_t N = (size_t)n_[0] * n_[1] * n_[2];
    std::vector<double> soa(5 * N);
    for (int k = 0; k < n_[2]; ++k)
        for (int j = 0; j < n_[1]; ++j)
            for (int i = 0; i < n_[0]; ++i)
                for (int v = 0; v < 5; ++v)
                    soa[v * N + ((size_t)k * n_[1] + j) * n_[0] + i] = f(i + ng, j + ng, k + ng, v);
    check(sf_state_upload(grid_, soa.data()), "sf_state_upload");
    uploaded_ = true;
}

void DeviceExplicitPath::downloadTo(Field& f) const {
    const int ng = f.NG();
    const size_t N = (size_t)n_[0] * n_[1] * n_[2];
    std::vector<double> soa(5 * N);
    check(sf_state_download(grid_, soa.data()), "sf_state_download");
    const int planes[3] = {f.NX(), f.NY(), f.NZ()};
    for (int k = 0; k < planes[2]; ++k)
        for (int j = 0; j < planes[1]; ++j)
            for (int i = 0; i < planes[0]; ++i) {
                const int c[3] = {active_[0] ? i : 0, active_[1] ? j : 0, active_[2] ? k : 0};
                const size_t u = ((size_t)c[2] * n_[1] + c[1]) * n_[0] + c[0];
                for (int v = 0; v < 5; ++v) f(i + ng, j + ng, k + ng, v) = soa[v * N + u];
            }
}

std::shared_ptr<DeviceExplicitPath> bindDeviceExplicitOps(
        Run::OpRegistry& operations, const System::CompiledNumericalSystem& numerics,
        const FDM::BoundaryConfig& boundaries, State::StateBundle& state,
        const double& maximumTimeStep, int device) {
    if (state.patches.size() != 1 || !state.patches.front())
        notServed("the device path advances exactly one patch");
    const Field& f = *state.patches.front();
    if (f.NVar() != 5) notServed("the state is not the five single-fluid conserved variables");
    if (!numerics.recipes.convection) notServed("no convection recipe is selected");
    const FDM::TermRecipe& recipe = *numerics.recipes.convection;
    if (recipe.role() != FDM::TermRole::Convection ||
        recipe.reconstruction() != FDM::ReconstructionVariable::Characteristic)
        notServed("the convection recipe is not a characteristic reconstruction");
    if (numerics.recipes.diffusion) notServed("diffusion terms run on the CPU path");
    if (boundaries.ilwEnabled) notServed("ILW boundary reconstruction runs on the CPU path");
    if (!state.stateModel || !state.stateModel->perfectGasGamma())
        notServed("the thermodynamic closure is not a perfect gas");

    sf_grid_desc d{};
    d.scheme = schemeOf(recipe.convection());
    d.flux = fluxOf(recipe.flux());
    d.time = timeOf(numerics.time.recipe.id());
    d.gamma = *state.stateModel->perfectGasGamma();
    d.device = device;
    auto path = std::shared_ptr<DeviceExplicitPath>(new DeviceExplicitPath());
    const int ng = f.NG(), halo = recipe.haloWidth();
    if (halo > ng) notServed("the patch has fewer ghost layers than the recipe's halo");
    bool cartesian = true;
    for (int a = 0; a < 3; ++a) {
        path->active_[a] = Math::isDirectionActiveIndex(a);
        path->n_[a] = path->active_[a] ? uniqueNodes(f, a) : 1;
        d.n[a] = path->n_[a];
        d.active[a] = path->active_[a] ? 1 : 0;
        d.d[a] = 1.0;
        if (!path->active_[a]) continue;
        // uniform Cartesian: the same diagonal metric at every unique node
        const double m0 = diagonalMetric(f, a, ng, ng, ng, cartesian);
        for (int k = 0; k < uniqueNodes(f, 2); ++k)
            for (int j = 0; j < uniqueNodes(f, 1); ++j)
                for (int i = 0; i < uniqueNodes(f, 0); ++i)
                    if (diagonalMetric(f, a, i + ng, j + ng, k + ng, cartesian) != m0)
                        notServed("the mesh spacing is not uniform");
        if (!cartesian || !(m0 > 0.0) || !std::isfinite(m0))
            notServed("the mesh is not Cartesian");
        d.d[a] = 1.0 / m0;
        const int lo = endKind(f, boundaries, a, 0), hi = endKind(f, boundaries, a, 1);
        d.bc[a] = lo && hi ? SF_BC_TRANSMISSIVE : lo ? SF_BC_NONE_HI : hi ? SF_BC_NONE_LO
                                                                          : SF_BC_NONE;
        // an end without a condition keeps its ghost layers as the case set
        // them; the device holds them as copies of the end node
        for (int side = 0; side < 2; ++side) {
            if ((side ? hi : lo) != 0) continue;
            const int node = ng + (side ? path->n_[a] - 1 : 0), dir = side ? 1 : -1;
            for (int k = 0; k < uniqueNodes(f, 2); ++k)
                for (int j = 0; j < uniqueNodes(f, 1); ++j)
                    for (int i = 0; i < uniqueNodes(f, 0); ++i) {
                        int c[3] = {i + ng, j + ng, k + ng};
                        if (c[a] != node) continue;
                        for (int l = 1; l <= halo; ++l) {
                            int g[3] = {c[0], c[1], c[2]};
                            g[a] += dir * l;
                            for (int v = 0; v < 5; ++v)
                                if (f(g[0], g[1], g[2], v) != f(c[0], c[1], c[2], v))
                                    notServed("the ghosts of an end without a condition differ "
                                              "from its node");
                        }
                    }
        }
    }
    check(sf_grid_create(&d, &path->grid_), "sf_grid_create");

    const double cfl = numerics.dt.cfl, maxDeltaT = numerics.dt.maxDeltaT;
    // FlowOperations::bindExplicitOps, SF_flowOperations.cpp:466-522
    operations.bind(System::OpIds::FlowStepPrepare, "flow.conservative",
                    [path, &state, &maximumTimeStep] {
        if (!std::isfinite(maximumTimeStep) || maximumTimeStep <= 0.0)
            throw std::runtime_error(
                "SolverAlgorithm::FlowOperations requires a finite positive time-step limit.");
        if (!path->uploaded_) path->upload(state.singlePatch());
        check(sf_step_prepare(path->grid_), "flow.step.prepare");
    });
    operations.bind(System::OpIds::FlowDtCompute, "flow.conservative",
                    [path, &state, cfl, maxDeltaT, &maximumTimeStep] {
        double dt = 0.0;
        check(sf_compute_dt(path->grid_, cfl, maxDeltaT, maximumTimeStep, &dt),
              "flow.dt.compute");
        state.dt = dt;
    });
    operations.bind(System::OpIds::FlowStepBegin, "flow.conservative",
                    [path] { check(sf_step_begin(path->grid_), "flow.step.begin"); });
    operations.bind(System::OpIds::ExplicitStageExecute, "flow.conservative",
                    [path](const Run::ExecutionContext& context) {
        check(sf_stage(path->grid_, context.stageIndex), "explicit.stage.execute");
    });
    // FlowOperations::bindOperations, SF_flowOperations.cpp:647-652: the
    // boundary state at t + dt; the reference's fail-fast checks
    // (validateEquationState after every stage) become one check of the
    // device's fault record per step, which leaves the state of the start
    // of a faulting step on the device and in the patch.
    operations.bind(System::OpIds::FlowStepCommit, "flow.conservative", [path, &state] {
        check(sf_step_commit(path->grid_), "flow.step.commit");
        const int rc = sf_fault_check(path->grid_, nullptr, nullptr, nullptr);
        path->downloadTo(state.singlePatch());
        check(rc, "flow.step.commit");
    });
    return path;
}

}  // namespace SF::SolverAlgorithm
