// sf_provider.cpp -- see sf_provider.h.
#include "sf_provider.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <sstream>

namespace sfamd {
// Shared with the grid ABI (sf_capi.hip): the calling thread's last error.
void set_last_error(const std::string& message);
}  // namespace sfamd

namespace sfamd::host {

// normalizeToken, app/application/model/parsing/SF_parserCommon.h:17-27.
std::string normalizeToken(std::string value) {
    value.erase(std::remove(value.begin(), value.end(), '"'), value.end());
    value.erase(std::remove_if(value.begin(), value.end(),
                               [](unsigned char c) {
                                   return c == '_' || c == '-' || std::isspace(c);
                               }),
                value.end());
    std::transform(value.begin(), value.end(), value.begin(),
                   [](unsigned char c) { return (char)std::tolower(c); });
    return value;
}

namespace {

const char* schemeName(ConvectionScheme s) {
    switch (s) {
    case ConvectionScheme::TENO5: return "teno5";
    case ConvectionScheme::WENO3: return "weno3";
    case ConvectionScheme::WENO7: return "weno7";
    default: return "weno5";
    }
}

// the flux part of toString(TermRecipeId) (SF_termRecipe.h:242-258)
const char* fluxName(FluxSplitter f) {
    switch (f) {
    case FluxSplitter::HLLC: return "HLLC";
    case FluxSplitter::StegerWarming: return "Steger";
    case FluxSplitter::LaxFriedrichs: return "LaxFriedrichs";
    case FluxSplitter::Roe: return "Roe";
    case FluxSplitter::LaxWendroff: return "LaxWendroff";
    default: return "Rusanov";
    }
}

std::string servedList() {
    std::string out;
    for (const auto& r : servedConvectionRecipes()) out += (out.empty() ? "" : ", ") + r.id;
    return out;
}

// The name the reference registers (toString(TermRecipeId),
// SF_termRecipe.h:242-258) that this provider does not serve: the primitive
// upwind recipe.
bool registeredElsewhere(const std::string& token) { return token == "primitiveupwind1"; }

}  // namespace

// The reference's registration order (TermRecipeId, SF_termRecipe.h:52-62:
// schemes WENO3, WENO5, TENO5, WENO7; fluxes Steger, Rusanov,
// LaxFriedrichs, Roe, LaxWendroff), then the HLLC recipes, which have no
// reference counterpart.
const std::vector<TermRecipe>& servedConvectionRecipes() {
    static const std::vector<TermRecipe> recipes = [] {
        std::vector<TermRecipe> r;
        const ConvectionScheme schemes[] = {ConvectionScheme::WENO3, ConvectionScheme::WENO5,
                                            ConvectionScheme::TENO5, ConvectionScheme::WENO7};
        for (ConvectionScheme sc : schemes)
            for (FluxSplitter f : {FluxSplitter::StegerWarming, FluxSplitter::Rusanov,
                                   FluxSplitter::LaxFriedrichs, FluxSplitter::Roe,
                                   FluxSplitter::LaxWendroff})
                r.push_back(builtInConvectionRecipe(sc, f));
        for (ConvectionScheme sc : schemes)
            r.push_back(builtInConvectionRecipe(sc, FluxSplitter::HLLC));
        return r;
    }();
    return recipes;
}

// builtInConvectionRecipe, SF_termRecipe.h:204-221: halo 2 for WENO3, 4 for
// WENO7, 3 otherwise.
TermRecipe builtInConvectionRecipe(ConvectionScheme scheme, FluxSplitter flux) {
    const int halo = scheme == ConvectionScheme::WENO3 ? 2 : scheme == ConvectionScheme::WENO7 ? 4 : 3;
    return TermRecipe{std::string(schemeName(scheme)) + fluxName(flux), scheme, flux, halo};
}

// resolveConvectionTermRecipe, SF_numericsParser.h:140-162.
TermRecipe resolveConvectionTermRecipe(const std::string& name) {
    const std::string token = normalizeToken(name);
    for (const auto& r : servedConvectionRecipes())
        if (token == normalizeToken(r.id)) return r;
    if (registeredElsewhere(token))
        throw std::invalid_argument("Convection term recipe '" + name +
                                    "' is registered in sonicSolver but not served by the "
                                    "MI355X provider. Served recipes: " + servedList() + ".");
    throw std::invalid_argument("Unsupported convection term recipe '" + name +
                                "'. Select a registered recipe such as " + servedList() + ".");
}

// parseConvectionScheme, SF_numericsParser.h:11-25 (empty selects WENO5).
ConvectionScheme parseConvectionScheme(const std::string& value) {
    const std::string t = normalizeToken(value);
    if (t.empty() || t == "weno5") return ConvectionScheme::WENO5;
    if (t == "teno5") return ConvectionScheme::TENO5;
    if (t == "weno3") return ConvectionScheme::WENO3;
    if (t == "weno7") return ConvectionScheme::WENO7;
    if (t == "upwind1")
        throw std::invalid_argument("Convection scheme '" + value +
                                    "' is not served by the MI355X provider (WENO3, WENO5, "
                                    "TENO5, WENO7).");
    throw std::invalid_argument("Unsupported convection scheme '" + value +
                                "'. Supported schemes: WENO3, WENO5, TENO5, WENO7.");
}

// parseFluxSplitter, SF_numericsParser.h:77-94 (empty selects
// Steger-Warming).
FluxSplitter parseFluxSplitter(const std::string& value) {
    const std::string t = normalizeToken(value);
    if (t.empty() || t == "stegerwarming" || t == "sw") return FluxSplitter::StegerWarming;
    if (t == "rusanov" || t == "localaxfriedrichs" || t == "llf") return FluxSplitter::Rusanov;
    if (t == "laxfriedrichs" || t == "lf" || t == "lxf" || t == "laxf")
        return FluxSplitter::LaxFriedrichs;
    if (t == "roe") return FluxSplitter::Roe;
    if (t == "hllc") return FluxSplitter::HLLC;
    if (t == "laxwendroff" || t == "lw") return FluxSplitter::LaxWendroff;
    if (t == "upwindadvection")
        throw std::invalid_argument("Flux method '" + value +
                                    "' is not served by the MI355X provider (StegerWarming, "
                                    "Rusanov, LaxFriedrichs, Roe, LaxWendroff, HLLC).");
    throw std::invalid_argument("Unsupported inviscid flux method '" + value +
                                "'. Supported methods: StegerWarming, Rusanov, LaxFriedrichs, "
                                "Roe, LaxWendroff, HLLC.");
}

// resolveTimeRecipe, SF_numericsParser.h:115-125 (empty selects classicalRK4).
TimeRecipeId resolveTimeRecipe(const std::string& value) {
    const std::string t = normalizeToken(value);
    if (t.empty() || t == "classicalrk4") return TimeRecipeId::ClassicalRK4;
    if (t == "forwardeuler") return TimeRecipeId::ForwardEuler;
    if (t == "ssprk3") return TimeRecipeId::SSPRK3;
    throw std::invalid_argument("Unsupported time recipe '" + value +
                                "'. Supported recipes: forwardEuler, SSPRK3, classicalRK4.");
}

// ForwardEulerMethod / SSPRK3Method / ClassicalRK4Method::compile,
// SF_builtinProviders.cpp:85-127: (abscissa, base weight, increment weight).
CompiledTimeRecipe compileTimeRecipe(TimeRecipeId id) {
    CompiledTimeRecipe r;
    r.id = id;
    switch (id) {
    case TimeRecipeId::ForwardEuler:
        r.name = "forwardEuler";
        r.order = 1;
        r.stages = {{0.0, 0.0, 1.0}};
        break;
    case TimeRecipeId::SSPRK3:
        r.name = "SSPRK3";
        r.order = 3;
        r.stages = {{0.0, 0.0, 1.0}, {1.0, 0.75, 0.25}, {0.5, 1.0 / 3.0, 2.0 / 3.0}};
        break;
    case TimeRecipeId::ClassicalRK4:
        r.name = "classicalRK4";
        r.order = 4;
        r.stages = {{0.0, 0.0, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.0, 1.0}, {1.0, 0.0, 0.0}};
        r.finalWeights = {1.0, 2.0, 2.0, 1.0};
        r.finalDivisor = 6.0;
        break;
    }
    return r;
}

// explicitFragment (SF_builtinProviders.cpp:63-81) for the single
// ConservativeResidual participant, plus its commit lifecycle leaves
// (ConservativeResidualMethod::lowerLifecycle, :166-175).
PlanNode explicitFragment(const TermRecipe& convection, const CompiledTimeRecipe& time) {
    const std::string provider = "flow.conservative";
    auto leaf = [&](PlanNodeKind kind, const char* id, const char* name, const char* op) {
        PlanNode n;
        n.kind = kind;
        n.id = id;
        n.name = name;
        n.operation = op;
        n.provider = provider;
        return n;
    };
    PlanNode root;
    root.kind = PlanNodeKind::Sequence;
    root.id = "Explicit.step";
    root.name = "explicit time step (" + convection.id + ")";
    root.children.push_back(leaf(PlanNodeKind::Update, "Explicit.prepare",
                                 "prepare physical step", OpIds::FlowStepPrepare));
    root.children.push_back(leaf(PlanNodeKind::Update, "Explicit.dt", "compute time step",
                                 OpIds::FlowDtCompute));
    root.children.push_back(leaf(PlanNodeKind::Update, "Explicit.begin", "begin explicit step",
                                 OpIds::FlowStepBegin));
    PlanNode stages;
    stages.kind = PlanNodeKind::StageLoop;
    stages.id = "Explicit.stages";
    stages.name = time.name + " fused explicit stages";
    stages.repetitions = time.stageCount();
    stages.children.push_back(leaf(PlanNodeKind::Update, "Explicit.stage.execute",
                                   "execute explicit stage", OpIds::ExplicitStageExecute));
    root.children.push_back(std::move(stages));
    root.children.push_back(leaf(PlanNodeKind::Commit, "Explicit.commit",
                                 "commit physical state", OpIds::FlowStepCommit));
    root.children.push_back(leaf(PlanNodeKind::Commit, "Explicit.time.commit",
                                 "commit physical clock", OpIds::TimeCommit));
    return root;
}

std::string describePlan(const PlanNode& root) {
    std::ostringstream out;
    std::function<void(const PlanNode&, int)> walk = [&](const PlanNode& n, int depth) {
        out << std::string(2 * depth, ' ');
        switch (n.kind) {
        case PlanNodeKind::Sequence: out << "Sequence " << n.id; break;
        case PlanNodeKind::StageLoop: out << "StageLoop " << n.id << " x" << n.repetitions; break;
        case PlanNodeKind::Update: out << "Update " << n.operation << " [" << n.provider << "]"; break;
        case PlanNodeKind::Commit: out << "Commit " << n.operation << " [" << n.provider << "]"; break;
        }
        out << "\n";
        for (const auto& c : n.children) walk(c, depth + 1);
    };
    walk(root, 0);
    return out.str();
}

void OpRegistry::bind(const std::string& opId, const std::string& provider, Op op) {
    ops_[opId] = Entry{provider, std::move(op)};
}

void OpRegistry::invoke(const std::string& opId, const ExecutionContext& context) const {
    auto it = ops_.find(opId);
    if (it == ops_.end())
        throw std::runtime_error("Plan leaf '" + opId + "' has no bound operation.");
    it->second.op(context);
}

void OpRegistry::validateBindings(const PlanNode& plan) const {
    if ((plan.kind == PlanNodeKind::Update || plan.kind == PlanNodeKind::Commit) &&
        !bound(plan.operation))
        throw std::runtime_error("Plan leaf '" + plan.operation + "' (" + plan.id +
                                 ") has no bound operation.");
    for (const auto& c : plan.children) validateBindings(c);
}

// PlanExecutor::execute (SF_planExecutor.cpp:81-104): a StageLoop hands the
// stage index and count to its leaves.
void PlanExecutor::execute(const PlanNode& node, const OpRegistry& ops, ExecutionContext ctx) {
    switch (node.kind) {
    case PlanNodeKind::Sequence:
        for (const auto& c : node.children) execute(c, ops, ctx);
        return;
    case PlanNodeKind::StageLoop:
        for (int s = 0; s < node.repetitions; ++s) {
            ExecutionContext inner{s, node.repetitions};
            for (const auto& c : node.children) execute(c, ops, inner);
        }
        return;
    case PlanNodeKind::Update:
    case PlanNodeKind::Commit:
        ops.invoke(node.operation, ctx);
        return;
    }
}

namespace {

void check(int rc) {
    if (rc != 0) {
        char msg[512];
        sf_last_error(msg, sizeof msg);
        throw std::runtime_error(msg);
    }
}

}  // namespace

FlowConservativeProvider::FlowConservativeProvider(const sf_grid_desc& geometry,
                                                   const TermRecipe& convection,
                                                   const CompiledTimeRecipe& time, double cfl,
                                                   double maxDeltaT)
    : convection_(convection), time_(time), cfl_(cfl), maxDeltaT_(maxDeltaT) {
    if (!(cfl > 0.0) || !std::isfinite(cfl))
        throw std::invalid_argument("CFL number must be positive and finite.");
    if (!(maxDeltaT > 0.0))
        throw std::invalid_argument("maxDeltaT must be positive.");
    sf_grid_desc d = geometry;
    d.scheme = (int)convection.convection;
    d.flux = (int)convection.flux;
    d.time = (int)time.id;
    check(sf_grid_create(&d, &grid_));
    plan_ = explicitFragment(convection_, time_);
}

FlowConservativeProvider::~FlowConservativeProvider() {
    if (grid_) sf_grid_destroy(grid_);
}

// FlowOperations::bindExplicitOps (SF_flowOperations.cpp:466-522) and the
// commit leaves (:648-652, SF_singleFluidStepper.cpp:40-46).
void FlowConservativeProvider::bindExplicitOps(OpRegistry& ops) {
    const std::string p = "flow.conservative";
    ops.bind(OpIds::FlowStepPrepare, p, [this](const ExecutionContext&) {
        check(sf_step_prepare(grid_));
    });
    ops.bind(OpIds::FlowDtCompute, p, [this](const ExecutionContext&) {
        double dt = 0.0;
        check(sf_compute_dt(grid_, cfl_, maxDeltaT_, state_.limit, &dt));
        state_.dt = dt;
    });
    ops.bind(OpIds::FlowStepBegin, p, [this](const ExecutionContext&) {
        check(sf_step_begin(grid_));
    });
    ops.bind(OpIds::ExplicitStageExecute, p, [this](const ExecutionContext& c) {
        if (c.stageIndex < 0 || c.stageCount != time_.stageCount())
            throw std::runtime_error("explicit.stage.execute outside its stage loop.");
        check(sf_stage(grid_, c.stageIndex));
    });
    // The reference throws at the first non-physical stage
    // (SF_equationState.h:25-57); here the step's fault record is checked
    // once the step is committed, before time.commit, so a faulting step is
    // never counted and the grid keeps the state at its start.
    ops.bind(OpIds::FlowStepCommit, p, [this](const ExecutionContext&) {
        check(sf_step_commit(grid_));
        check(sf_fault_check(grid_, nullptr, nullptr, nullptr));
    });
    ops.bind(OpIds::TimeCommit, p, [this](const ExecutionContext&) {
        state_.time += state_.dt;
        ++state_.step;
        dts_.push_back(state_.dt);
    });
}

// Time::Driver::run, SF_time.cpp:42-93.
int runDriver(const RunControl& c, FlowConservativeProvider& provider, const OpRegistry& ops,
              const std::function<bool(int, double, double)>& onOutput) {
    if (!std::isfinite(c.startTime) || !std::isfinite(c.endTime) || c.endTime < c.startTime ||
        c.endStep < 0)
        throw std::runtime_error("Time::Driver received invalid runTime bounds.");
    if (c.writeByStep ? c.writeIntervalSteps <= 0
                      : (!std::isfinite(c.writeIntervalTime) || c.writeIntervalTime <= 0.0))
        throw std::runtime_error("Time::Driver requires a positive output interval.");
    ops.validateBindings(provider.plan());
    double time = c.startTime;
    int step = 0;
    int nextOutputStep = c.writeIntervalSteps;
    double nextOutputTime = c.startTime + c.writeIntervalTime;
    DriverState& st = provider.state();
    st.time = time;
    const double eps = std::numeric_limits<double>::epsilon();
    while (true) {
        const double remaining = c.endTime - time;
        const double roundoff =
            64.0 * eps * std::max({std::abs(time), std::abs(c.endTime), 1.0});
        const bool finished = remaining <= roundoff || (c.endStep > 0 && step >= c.endStep);
        if (finished) break;
        ++step;
        double limit = remaining;
        if (!c.writeByStep) limit = std::min(limit, nextOutputTime - time);
        if (!std::isfinite(limit) || limit <= 0.0)
            throw std::runtime_error("Time::Driver produced a non-positive time-step limit.");
        st.limit = limit;
        PlanExecutor::execute(provider.plan(), ops);
        const double tol = 1.0e-11 * std::max({std::abs(time), std::abs(st.time), 1.0});
        if (!std::isfinite(st.dt) || st.dt <= 0.0 || st.dt > limit + tol ||
            !std::isfinite(st.time) || std::abs(st.time - (time + st.dt)) > tol)
            throw std::runtime_error("Time::Driver stepper returned inconsistent dt/time.");
        time = st.time;
        bool stop = false;
        if (c.writeByStep && step == nextOutputStep) {
            if (onOutput) stop = !onOutput(step, time, st.dt);
            nextOutputStep += c.writeIntervalSteps;
        } else if (!c.writeByStep && time >= nextOutputTime - tol) {
            if (onOutput) stop = !onOutput(step, time, st.dt);
            nextOutputTime += c.writeIntervalTime;
        }
        if (stop) break;
    }
    return step;
}

}  // namespace sfamd::host

// ---------------------------------------------------------------------------
// C ABI (include/sonic_amd_host.h)
// ---------------------------------------------------------------------------
using namespace sfamd::host;

struct sf_solver {
    std::unique_ptr<FlowConservativeProvider> provider;
    OpRegistry ops;
};

namespace {

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::invalid_argument& e) {
        sfamd::set_last_error(e.what());
        return SF_E_ARG;
    } catch (const std::exception& e) {
        sfamd::set_last_error(e.what());
        return SF_E_NUMERIC;
    }
}

}  // namespace

extern "C" {

int sf_recipe_resolve(const char* name, int* scheme, int* flux, int* halo) {
    return guarded([&] {
        if (!name) throw std::invalid_argument("sf_recipe_resolve: null recipe name");
        const TermRecipe r = resolveConvectionTermRecipe(name);
        if (scheme) *scheme = (int)r.convection;
        if (flux) *flux = (int)r.flux;
        if (halo) *halo = r.haloWidth;
        return 0;
    });
}

int sf_recipe_from_pair(const char* scheme_token, const char* flux_token, int* scheme,
                        int* flux, int* halo) {
    return guarded([&] {
        const TermRecipe r = builtInConvectionRecipe(
            parseConvectionScheme(scheme_token ? scheme_token : ""),
            parseFluxSplitter(flux_token ? flux_token : ""));
        if (scheme) *scheme = (int)r.convection;
        if (flux) *flux = (int)r.flux;
        if (halo) *halo = r.haloWidth;
        return 0;
    });
}

int sf_time_recipe_resolve(const char* token, int* time, int* stages) {
    return guarded([&] {
        const CompiledTimeRecipe r = compileTimeRecipe(resolveTimeRecipe(token ? token : ""));
        if (time) *time = (int)r.id;
        if (stages) *stages = r.stageCount();
        return 0;
    });
}

int sf_plan_describe(const char* conv, const char* time, char* buf, size_t len) {
    return guarded([&] {
        if (!conv || !buf || len == 0) throw std::invalid_argument("sf_plan_describe: null argument");
        const std::string text = describePlan(explicitFragment(
            resolveConvectionTermRecipe(conv), compileTimeRecipe(resolveTimeRecipe(time ? time : ""))));
        if (text.size() + 1 > len) throw std::invalid_argument("sf_plan_describe: buffer too small");
        std::memcpy(buf, text.c_str(), text.size() + 1);
        return 0;
    });
}

int sf_solver_create(const sf_grid_desc* geometry, const char* conv, const char* time,
                     double cfl, double max_dt, sf_solver** out) {
    return guarded([&] {
        if (!geometry || !conv || !out) throw std::invalid_argument("sf_solver_create: null argument");
        *out = nullptr;
        auto s = std::make_unique<sf_solver>();
        s->provider = std::make_unique<FlowConservativeProvider>(
            *geometry, resolveConvectionTermRecipe(conv),
            compileTimeRecipe(resolveTimeRecipe(time ? time : "")), cfl, max_dt);
        s->provider->bindExplicitOps(s->ops);
        s->ops.validateBindings(s->provider->plan());
        *out = s.release();
        return 0;
    });
}

int sf_solver_destroy(sf_solver* s) {
    delete s;
    return 0;
}

int sf_solver_grid(sf_solver* s, sf_grid** grid) {
    if (!s || !grid) {
        sfamd::set_last_error("sf_solver_grid: null argument");
        return SF_E_ARG;
    }
    *grid = s->provider->grid();
    return 0;
}

int sf_solver_run(sf_solver* s, const sf_run_control* c, sf_output_fn on_output, void* user,
                  int* steps_done, double* final_time) {
    return guarded([&] {
        if (!s || !c) throw std::invalid_argument("sf_solver_run: null argument");
        RunControl rc;
        rc.startTime = c->start_time;
        rc.endTime = c->end_time;
        rc.endStep = c->end_step;
        rc.writeByStep = c->write_by_step != 0;
        rc.writeIntervalSteps = c->write_interval_steps;
        rc.writeIntervalTime = c->write_interval_time;
        const int steps = runDriver(rc, *s->provider, s->ops, [&](int step, double t, double dt) {
            return on_output ? on_output(user, step, t, dt) == 0 : true;
        });
        if (steps_done) *steps_done = steps;
        if (final_time) *final_time = s->provider->state().time;
        uint32_t flags = 0;
        if (sf_error_flags(s->provider->grid(), &flags, 0) != 0 || flags != 0)
            throw std::runtime_error("device raised numerical fault flags during the run");
        return 0;
    });
}

int sf_solver_dt_history(sf_solver* s, double* out, int capacity, int* count) {
    if (!s) {
        sfamd::set_last_error("sf_solver_dt_history: null solver");
        return SF_E_ARG;
    }
    const auto& h = s->provider->dtHistory();
    if (count) *count = (int)h.size();
    if (out)
        for (int i = 0; i < capacity && i < (int)h.size(); ++i) out[i] = h[(size_t)i];
    return 0;
}

}  // extern "C"
