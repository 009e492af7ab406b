// sf_provider.h -- host-side runtime provider of the explicit Euler path:
// the WHICH / HOW contracts of sonicSolver for this path, restated in C++
// over the device grid ABI (include/sonic_amd.h).
//
//   TermRecipe / resolveConvectionTermRecipe   core/config/types/SF_termRecipe.h:52-221,
//                                              SF_numericsParser.h:11-162
//   TimeRecipe / CompiledTimeRecipe / tableaux core/config/types/SF_timeRecipe.h,
//                                              solver/system/SF_compiledTimeRecipe.h,
//                                              solver/system/SF_builtinProviders.cpp:85-127
//   explicit plan fragment                     SF_builtinProviders.cpp:29-82,162-175
//   OpRegistry / PlanExecutor                  solver/run/SF_planExecutor.h:33-72,
//                                              SF_planExecutor.cpp:81-104,190-208
//   flow.conservative op binding               solver/algorithm/SF_flowOperations.cpp:466-522,648
//   Time::Driver                               solver/algorithm/time/SF_time.cpp:42-93
#pragma once

#include "sonic_amd.h"
#include "sonic_amd_host.h"

#include <array>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sfamd::host {

std::string normalizeToken(std::string value);

// ---- WHICH ------------------------------------------------------------------
enum class ConvectionScheme { WENO5 = SF_SCHEME_WENO5, TENO5 = SF_SCHEME_TENO5,
                              WENO3 = SF_SCHEME_WENO3, WENO7 = SF_SCHEME_WENO7 };
enum class FluxSplitter { Rusanov = SF_FLUX_RUSANOV, HLLC = SF_FLUX_HLLC,
                          StegerWarming = SF_FLUX_STEGER_WARMING,
                          LaxFriedrichs = SF_FLUX_LAX_FRIEDRICHS, Roe = SF_FLUX_ROE,
                          LaxWendroff = SF_FLUX_LAX_WENDROFF };

struct TermRecipe {
    std::string id;             // canonical name, e.g. "weno5Rusanov"
    ConvectionScheme convection;
    FluxSplitter flux;
    int haloWidth;              // 2 WENO3, 3 WENO5 / TENO5, 4 WENO7
};

// Recipes served by this provider, in registration order.
const std::vector<TermRecipe>& servedConvectionRecipes();
TermRecipe resolveConvectionTermRecipe(const std::string& name);
ConvectionScheme parseConvectionScheme(const std::string& token);
FluxSplitter parseFluxSplitter(const std::string& token);
TermRecipe builtInConvectionRecipe(ConvectionScheme scheme, FluxSplitter flux);

enum class TimeRecipeId { ForwardEuler = SF_TIME_EULER, SSPRK3 = SF_TIME_SSPRK3,
                          ClassicalRK4 = SF_TIME_RK4 };

struct TimeStage {
    double abscissa = 0.0;
    double baseWeight = 0.0;
    double incrementWeight = 0.0;
};

struct CompiledTimeRecipe {
    TimeRecipeId id = TimeRecipeId::ForwardEuler;
    std::string name;
    int order = 1;
    std::vector<TimeStage> stages;
    std::array<double, 4> finalWeights{};
    double finalDivisor = 1.0;
    int stageCount() const { return (int)stages.size(); }
};

TimeRecipeId resolveTimeRecipe(const std::string& token);
CompiledTimeRecipe compileTimeRecipe(TimeRecipeId id);

// ---- HOW --------------------------------------------------------------------
namespace OpIds {
inline constexpr const char* FlowStepPrepare = "flow.step.prepare";
inline constexpr const char* FlowDtCompute = "flow.dt.compute";
inline constexpr const char* FlowStepBegin = "flow.step.begin";
inline constexpr const char* ExplicitStageExecute = "explicit.stage.execute";
inline constexpr const char* FlowStepCommit = "flow.step.commit";
inline constexpr const char* TimeCommit = "time.commit";
}  // namespace OpIds

enum class PlanNodeKind { Sequence, StageLoop, Update, Commit };

struct PlanNode {
    PlanNodeKind kind = PlanNodeKind::Sequence;
    std::string id, name, operation, provider;
    int repetitions = 1;
    std::vector<PlanNode> children;
};

// The explicit step for the ConservativeResidual method ("flow.conservative").
PlanNode explicitFragment(const TermRecipe& convection, const CompiledTimeRecipe& time);
std::string describePlan(const PlanNode& root);

struct ExecutionContext {
    int stageIndex = -1;
    int stageCount = 0;
};

class OpRegistry {
public:
    using Op = std::function<void(const ExecutionContext&)>;
    void bind(const std::string& opId, const std::string& provider, Op op);
    bool bound(const std::string& opId) const { return ops_.count(opId) != 0; }
    void invoke(const std::string& opId, const ExecutionContext& context) const;
    // Reject a plan with an unbound leaf (PlanExecutor::validateBindings).
    void validateBindings(const PlanNode& plan) const;

private:
    struct Entry {
        std::string provider;
        Op op;
    };
    std::map<std::string, Entry> ops_;
};

class PlanExecutor {
public:
    static void execute(const PlanNode& node, const OpRegistry& ops,
                        ExecutionContext context = {});
};

// ---- runtime provider -----------------------------------------------------------
struct DriverState {
    double time = 0.0;
    double dt = 0.0;
    int step = 0;
    double limit = 0.0;         // time-step limit from the driver
};

class FlowConservativeProvider {
public:
    FlowConservativeProvider(const sf_grid_desc& geometry, const TermRecipe& convection,
                             const CompiledTimeRecipe& time, double cfl, double maxDeltaT);
    ~FlowConservativeProvider();
    FlowConservativeProvider(const FlowConservativeProvider&) = delete;
    FlowConservativeProvider& operator=(const FlowConservativeProvider&) = delete;

    void bindExplicitOps(OpRegistry& ops);
    sf_grid* grid() const { return grid_; }
    DriverState& state() { return state_; }
    const std::vector<double>& dtHistory() const { return dts_; }
    const PlanNode& plan() const { return plan_; }

private:
    sf_grid* grid_ = nullptr;
    TermRecipe convection_;
    CompiledTimeRecipe time_;
    double cfl_, maxDeltaT_;
    DriverState state_;
    PlanNode plan_;
    std::vector<double> dts_;
};

struct RunControl {
    double startTime = 0.0;
    double endTime = 0.0;
    int endStep = 0;
    bool writeByStep = false;
    int writeIntervalSteps = 0;
    double writeIntervalTime = 0.0;
};

// Time::Driver::run: returns the number of steps taken.
int runDriver(const RunControl& control, FlowConservativeProvider& provider,
              const OpRegistry& ops,
              const std::function<bool(int step, double time, double dt)>& onOutput);

}  // namespace sfamd::host
