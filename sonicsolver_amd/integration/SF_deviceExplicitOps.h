#pragma once
/// @file SF_deviceExplicitOps.h
/// @brief Reference-side binding of the explicit density-based step to the
/// MI355X library (libsonic_amd.so, C ABI in include/sonic_amd.h).
///
/// A sonicSolver maintainer adds this file (and its .cpp) to
/// src/solver/algorithm/ and calls bindDeviceExplicitOps from
/// FlowOperations::bindOperations in place of bindExplicitOps when the
/// compiled numerics are served by the device (the patch
/// sonicsolver_amd/integration/flowOperations.patch).  It binds the op ids
/// flow.step.prepare, flow.dt.compute, flow.step.begin,
/// explicit.stage.execute and flow.step.commit for the "flow.conservative"
/// provider, as FlowOperations::bindExplicitOps and bindOperations do
/// (SF_flowOperations.cpp:466-522,647-652); time.commit stays with the
/// stepper (SF_singleFluidStepper.cpp:40-46).
#include "SF_field.h"
#include "core/config/SF_configTypes.h"
#include "core/state/SF_stateBundle.h"
#include "solver/run/SF_planExecutor.h"
#include "solver/system/SF_numericalSystem.h"

#include <memory>
#include <vector>

struct sf_grid;

namespace SF::SolverAlgorithm {

/// Device-resident state of one patch and the op bodies that advance it.
class DeviceExplicitPath {
public:
    ~DeviceExplicitPath();
    DeviceExplicitPath(const DeviceExplicitPath&) = delete;
    DeviceExplicitPath& operator=(const DeviceExplicitPath&) = delete;

    /// Copy the device state back into the patch (its unique nodes; the
    /// planes of an inactive direction all take the one device plane).
    void downloadTo(Field& field) const;

private:
    friend std::shared_ptr<DeviceExplicitPath> bindDeviceExplicitOps(
        Run::OpRegistry&, const System::CompiledNumericalSystem&,
        const FDM::BoundaryConfig&, State::StateBundle&, const double&, int);
    DeviceExplicitPath() = default;
    void upload(const Field& field);

    sf_grid* grid_ = nullptr;
    int n_[3] = {1, 1, 1};
    bool active_[3] = {false, false, false};
    bool uploaded_ = false;
};

/// Whether the explicit ops go to the device: when the environment names a
/// device (SF_DEVICE_EXPLICIT = its ordinal), or when the selected
/// convection recipe's flux is HLLC, which only the device serves.
bool deviceExplicitRequested(const System::CompiledNumericalSystem& numerics);
/// The device ordinal SF_DEVICE_EXPLICIT names (0 when unset).
int deviceExplicitOrdinal();

/// Binds the explicit ops of the single patch of `state` to the device.
/// Throws std::runtime_error (binding nothing) when the compiled numerics
/// are not served by the device: a recipe other than characteristic
/// WENO3/WENO5/TENO5/WENO7 with a Steger-Warming, Rusanov, Lax-Friedrichs,
/// Roe, Lax-Wendroff or HLLC flux; a mesh that is not uniform Cartesian;
/// a boundary end other than zeroGradient for density, velocity and
/// pressure or one whose sets carry no condition; ILW; more than one
/// patch.  FlowOperations lets the exception stop the run: a device run is
/// asked for explicitly, and the CPU path has no HLLC flux.
std::shared_ptr<DeviceExplicitPath> bindDeviceExplicitOps(
    Run::OpRegistry& operations, const System::CompiledNumericalSystem& numerics,
    const FDM::BoundaryConfig& boundaries, State::StateBundle& state,
    const double& maximumTimeStep, int device = 0);

}  // namespace SF::SolverAlgorithm
