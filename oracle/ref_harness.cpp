// ref_harness.cpp -- drives sonicSolver's own hot-path code on
// programmatically built Fields.  TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile (target `ref`) from the reference sources where
// they lie; the output goes to oracle/_ref/libsfref.so.  It is used to pin
// the C restatement (sfcpu.c) and as the timed CPU baseline.  Nothing in
// the shipped package links or loads it.
//
// What is the reference's code here: Field (SF_field.cpp), the boundary
// applicator with linear zeroGradient (SF_applicator.h, SF_boundary.cpp,
// SF_thermodynamicClosure.cpp, SF_zeroGradient.cpp), divDispatch /
// computeAllFluxes / characteristicWENOFaceFlux (SF_div.h, SF_WENO.h),
// weno5_core / teno5_core / buildCharacteristicMatrix (SF_WENO5.cpp,
// SF_TENO5.cpp, SF_riemann.cpp), RusanovEOS::faceFlux / deltaT
// (SF_rusanovEOS.cpp), validateEquationState (SF_equationState.h) and
// Time::Explicit::begin / executeStage (SF_explicit.cpp) and Time::Driver
// (SF_time.cpp).
//
// What this file supplies itself:
//  - the explicit tableaux, restated from SF_builtinProviders.cpp:85-127,
//    because that translation unit pulls the whole case compiler (and
//    through it HYPRE) into the link;
//  - uniform Cartesian metrics and named boundary sets (the values that
//    Mesh::computeMetrics and the mesh generator produce on a uniform
//    grid, SF_mesh.cpp:320-338, SF_meshGen.cpp:514-541);
//  - periodic ghost fill and the HLLC flux, which the reference does not
//    have (the HLLC flux is the oracle's, plugged into the reference's
//    characteristicWENOFaceFlux through its FaceFlux parameter).
#include "core/field/SF_field.h"
#include "core/flux/SF_flux.h"
#include "core/residual/SF_residual.h"
#include "solver/discretization/convection/SF_div.h"
#include "solver/discretization/convection/SF_WENO.h"
#include "methods/numerics/Flux/SF_rusanovEOS.h"
#include "methods/numerics/state/SF_equationState.h"
#include "methods/numerics/structured/SF_residualAssembly.h"
#include "models/physics/fluidStateModel/SF_factory.h"
#include "solver/algorithm/time/SF_explicit.h"
#include "solver/algorithm/time/SF_time.h"
#include "solver/boundary/SF_applicator.h"

#include "sfcpu.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using Model = SF::Physics::FluidStateModel::Model;

struct Case {
    sfo_desc d;
    double dt = 0.0;   // the step's dt (state.dt), read by the Lax-Wendroff flux
    std::shared_ptr<const Model> eq;
    SF::Field field;
    SF::Boundary::Applicator bc;
    std::vector<SF::SolverAlgorithm::PatchWorkspace> ws;
    // A z slab of a decomposed grid: its z ghosts come from the neighbouring
    // slabs (the exchange below) instead of the periodic fill.
    std::function<void()> z_exchange;
};

void configure_directions(const sfo_desc& d) {
    SF::Math::resetActiveDirections();
    for (int a = 0; a < 3; ++a)
        if (!d.active[a]) SF::Math::deactivateDirection(a);
}

// zeroGradient on the lower (side 0) or upper (side 1) end of an axis.
bool zg_end(const sfo_desc& d, int a, int side) {
    if (!d.active[a]) return false;
    if (d.bc[a] == SFO_TRANSMISSIVE) return true;
    if (d.bc[a] == SFO_NONE_LO) return side == 1;
    if (d.bc[a] == SFO_NONE_HI) return side == 0;
    return false;
}

// The boundary sets a case file would declare zeroGradient for rho, U and
// p (an end without a condition is simply not named).  The y and z sets
// name both ends at once, so there an end without a condition needs the
// other end to lack one too.
SF::FDM::BoundaryConfig boundary_config(const sfo_desc& d) {
    SF::FDM::BoundaryConfig cfg;
    for (int a = 1; a < 3; ++a)
        if (d.active[a] && (d.bc[a] == SFO_NONE_LO || d.bc[a] == SFO_NONE_HI))
            throw std::runtime_error("one end of the y or z set cannot go without a condition");
    std::vector<std::string> sets;
    if (zg_end(d, 0, 0)) sets.push_back("Left");
    if (zg_end(d, 0, 1)) sets.push_back("Right");
    if (zg_end(d, 1, 0) || zg_end(d, 1, 1)) sets.push_back("sideWalls");
    if (zg_end(d, 2, 0) || zg_end(d, 2, 1)) sets.push_back("frontBack");
    for (const auto& s : sets) {
        cfg.density.push_back({s, SF::ZERO_GRADIENT, 0.0});
        cfg.velocity.push_back({s, SF::ZERO_GRADIENT, SF::Vector3()});
        cfg.pressure.push_back({s, SF::ZERO_GRADIENT, 0.0});
    }
    return cfg;
}

std::unique_ptr<Case> make_case(const sfo_desc& d, const double* soa) {
    auto c = std::make_unique<Case>();
    c->d = d;
    c->eq = SF::Physics::FluidStateModel::makeSingleFluidPerfectGas(
        d.gamma, 287.05, 0.0, 0.72);
    configure_directions(d);
    SF::Field& f = c->field;
    f.setStateModel(c->eq);
    const int ng = sfo_halo(d.scheme);
    f.setup(d.n[0], d.n[1], d.n[2], ng, 5);
    const double jac = 1.0 / (d.d[0] * d.d[1] * d.d[2]);
    for (int k = 0; k < f.MZ(); ++k)
        for (int j = 0; j < f.MY(); ++j)
            for (int i = 0; i < f.MX(); ++i) {
                f.X(i, j, k) = (i - ng) * d.d[0];
                f.Y(i, j, k) = (j - ng) * d.d[1];
                f.Z(i, j, k) = (k - ng) * d.d[2];
                f.Jac(i, j, k) = jac;
                f.XiX(i, j, k) = 1.0 / d.d[0];
                f.EtY(i, j, k) = 1.0 / d.d[1];
                f.ZeZ(i, j, k) = 1.0 / d.d[2];
            }
    std::map<std::string, std::vector<int>> sets;
    const int nx = d.n[0], ny = d.n[1], nz = d.n[2];
    for (int k = ng; k < ng + nz; ++k)
        for (int j = ng; j < ng + ny; ++j) {
            sets["Left"].push_back(f.getIdx(ng, j, k));
            sets["Right"].push_back(f.getIdx(ng + nx - 1, j, k));
        }
    for (int k = ng; k < ng + nz; ++k)
        for (int i = ng; i < ng + nx; ++i) {
            sets["sideWalls"].push_back(f.getIdx(i, ng, k));
            sets["sideWalls"].push_back(f.getIdx(i, ng + ny - 1, k));
        }
    for (int j = ng; j < ng + ny; ++j)
        for (int i = ng; i < ng + nx; ++i) {
            sets["frontBack"].push_back(f.getIdx(i, j, ng));
            sets["frontBack"].push_back(f.getIdx(i, j, ng + nz - 1));
        }
    for (auto& kv : sets) {
        std::sort(kv.second.begin(), kv.second.end());
        kv.second.erase(std::unique(kv.second.begin(), kv.second.end()),
                        kv.second.end());
    }
    f.setBoundarySets(sets);
    c->bc = SF::Boundary::Applicator(boundary_config(d));
    const size_t N = (size_t)nx * ny * nz;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                for (int v = 0; v < 5; ++v)
                    f(i + ng, j + ng, k + ng, v) =
                        soa[v * N + ((size_t)k * ny + j) * nx + i];
    // Ghost layers of an end without a condition: nothing refills them, so
    // they keep the values they start with, here copies of the end node.
    // This reproduces the reference's frozen WENO7 Sod baseline
    // (test/Sod/sodCase_weno7_t0p2, whose Left set carries no rho/U/p
    // condition).
    for (int a = 0; a < 3; ++a)
        for (int side = 0; side < 2; ++side) {
            if (!d.active[a] || d.bc[a] == SFO_TRANSMISSIVE || d.bc[a] == SFO_PERIODIC) continue;
            if (zg_end(d, a, side)) continue;
            int p[3];
            for (p[2] = ng; p[2] < ng + nz; ++p[2])
                for (p[1] = ng; p[1] < ng + ny; ++p[1])
                    for (p[0] = ng; p[0] < ng + nx; ++p[0]) {
                        const int end = side == 0 ? ng : ng + d.n[a] - 1;
                        if (p[a] != end) continue;
                        for (int l = 1; l <= ng; ++l) {
                            int gp[3] = {p[0], p[1], p[2]};
                            gp[a] += side == 0 ? -l : l;
                            for (int v = 0; v < 5; ++v)
                                f(gp[0], gp[1], gp[2], v) = f(p[0], p[1], p[2], v);
                        }
                    }
        }
    f.invalidateThermodynamicCache();
    c->ws.resize(1);
    c->ws[0].ensureFor(f);
    return c;
}

void periodic_fill(Case& c) {
    SF::Field& f = c.field;
    const int ng = f.NG();
    for (int a = 0; a < 3; ++a) {
        if (!c.d.active[a] || c.d.bc[a] != SFO_PERIODIC) continue;
        if (a == 2 && c.z_exchange) continue;
        const int n = c.d.n[a];
        int lo[3] = {ng, ng, ng}, hi[3] = {ng + c.d.n[0], ng + c.d.n[1], ng + c.d.n[2]};
        int p[3];
        const int e1 = (a + 1) % 3, e2 = (a + 2) % 3;
        for (p[e2] = lo[e2]; p[e2] < hi[e2]; ++p[e2])
            for (p[e1] = lo[e1]; p[e1] < hi[e1]; ++p[e1])
                for (int l = 0; l < ng; ++l) {
                    int gl[3] = {p[0], p[1], p[2]}, sl[3] = {p[0], p[1], p[2]};
                    int gh[3] = {p[0], p[1], p[2]}, sh[3] = {p[0], p[1], p[2]};
                    gl[a] = l;          sl[a] = ng + ((l - ng) % n + n) % n;
                    gh[a] = ng + n + l; sh[a] = ng + l % n;
                    for (int v = 0; v < 5; ++v) {
                        f(gl[0], gl[1], gl[2], v) = f(sl[0], sl[1], sl[2], v);
                        f(gh[0], gh[1], gh[2], v) = f(sh[0], sh[1], sh[2], v);
                    }
                }
    }
    f.invalidateThermodynamicCache();
}

void prepare_boundary(Case& c) {
    c.bc.apply(c.field);
    periodic_fill(c);
    if (c.z_exchange) c.z_exchange();
    c.field.invalidateThermodynamicCache();
}

void hllc_face_flux(const double qL[5], const double qR[5], const double n[3],
                    double out[5], double gamma) {
    if (sfo_hllc(qL, qR, n, gamma, out) != 0) {
        std::cerr << "[SF FATAL] HLLC received an invalid state." << std::endl;
        std::exit(1);
    }
}

SF::FDM::ConvectionScheme scheme_of(int scheme) {
    switch (scheme) {
    case SFO_TENO5: return SF::FDM::ConvectionScheme::TENO5;
    case SFO_WENO3: return SF::FDM::ConvectionScheme::WENO3;
    case SFO_WENO7: return SF::FDM::ConvectionScheme::WENO7;
    }
    return SF::FDM::ConvectionScheme::WENO5;
}

SF::FDM::FluxSplitter splitter_of(int flux) {
    switch (flux) {
    case SFO_STEGER: return SF::FDM::FluxSplitter::StegerWarming;
    case SFO_LAXF: return SF::FDM::FluxSplitter::LaxFriedrichs;
    case SFO_ROE: return SF::FDM::FluxSplitter::Roe;
    case SFO_LAXW: return SF::FDM::FluxSplitter::LaxWendroff;
    }
    return SF::FDM::FluxSplitter::Rusanov;
}

// The HLLC face flux inside the reference's characteristic WENO loop.
template <int NS, class Core>
void hllc_fluxes(Case& c, Core core) {
    auto& w = c.ws[0];
    const double gamma = c.d.gamma;
    SF::Flux::computeAllFluxes<NS>(
        c.field, w.convectiveFlux,
        [&](const auto& s, const auto& m, auto* fl, int, int, int, SF::Math::Dir) {
            SF::Flux::characteristicWENOFaceFlux<NS>(
                s, m, fl, core,
                [&](const double qL[5], const double qR[5], const double n[3],
                    double out[5]) { hllc_face_flux(qL, qR, n, out, gamma); },
                gamma);
        },
        SF::Flux::LowOrderFlux::Method::FirstOrderRoe,
        SF::FDM::IBMBoundaryScheme::LowOrder, 0, gamma, c.eq.get());
}

// ConservativeRHS per-patch assembly for one stage (SF_conservativeRHS.cpp
// :137-188 -> Compressible::System::convection -> divDispatch).
void assemble(Case& c) {
    prepare_boundary(c);
    auto& w = c.ws[0];
    w.clear();
    const double gamma = c.d.gamma;
    if (c.d.flux != SFO_HLLC) {
        SF::divDispatch(c.field, w.convectiveFlux, w.residual, scheme_of(c.d.scheme),
                        splitter_of(c.d.flux), c.dt,
                        SF::FDM::IBMBoundaryScheme::LowOrder, 0, gamma, *c.eq);
        return;
    }
    switch (c.d.scheme) {
    case SFO_TENO5: hllc_fluxes<6>(c, SF::TENO5::teno5_core); break;
    case SFO_WENO3: hllc_fluxes<4>(c, SF::WENO3::weno3_core); break;
    case SFO_WENO7: hllc_fluxes<8>(c, SF::WENO7::weno7_core); break;
    default: hllc_fluxes<6>(c, SF::WENO5::weno5_core);
    }
    SF::Math::assembleConvectiveFluxResidual(c.field, w.convectiveFlux, w.residual);
}

SF::System::CompiledTimeRecipe time_recipe(int time) {
    using SF::System::CompiledTimeRecipe;
    using SF::System::ExplicitStageBackend;
    using SF::FDM::TimeRecipeId;
    if (time == SFO_EULER)
        return CompiledTimeRecipe::fromMethod(
            SF::FDM::builtInTimeRecipe(TimeRecipeId::ForwardEuler),
            {{{0.0, 0.0, 1.0}}}, 1, ExplicitStageBackend::ForwardEuler);
    if (time == SFO_SSPRK3)
        return CompiledTimeRecipe::fromMethod(
            SF::FDM::builtInTimeRecipe(TimeRecipeId::SSPRK3),
            {{{0.0, 0.0, 1.0}, {1.0, 0.75, 0.25}, {0.5, 1.0 / 3.0, 2.0 / 3.0}}}, 3,
            ExplicitStageBackend::SSPRK3);
    return CompiledTimeRecipe::fromMethod(
        SF::FDM::builtInTimeRecipe(TimeRecipeId::ClassicalRK4),
        {{{0.0, 0.0, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.0, 1.0}, {1.0, 0.0, 0.0}}}, 4,
        ExplicitStageBackend::ClassicalRK4, {{1.0, 2.0, 2.0, 1.0}}, 6.0);
}

void export_state(const Case& c, double* soa) {
    const SF::Field& f = c.field;
    const int ng = f.NG();
    const int nx = c.d.n[0], ny = c.d.n[1], nz = c.d.n[2];
    const size_t N = (size_t)nx * ny * nz;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                for (int v = 0; v < 5; ++v)
                    soa[v * N + ((size_t)k * ny + j) * nx + i] =
                        f(i + ng, j + ng, k + ng, v);
}

// The reference reports orchestration errors by throwing; a C caller gets -1.
template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "[sfref] " << e.what() << std::endl;
        return -1;
    }
}

}  // namespace

extern "C" {

double sfref_weno5(const double v[5]) { return SF::WENO5::weno5_core(v); }
double sfref_teno5(const double v[5]) { return SF::TENO5::teno5_core(v); }
double sfref_weno3(const double v[4]) { return SF::WENO3::weno3_core(v); }
double sfref_weno7(const double v[7]) { return SF::WENO7::weno7_core(v); }


// StegerWarmingFlux / LaxFriedrichsFlux / RoeFlux ::flux by id.
int sfref_split_flux(int flux, const double qL[5], const double qR[5], const double n[3],
                     double gamma, double out[5]) {
    return guarded([&] {
        if (flux == SFO_STEGER) SF::Flux::StegerWarmingFlux::flux(qL, qR, n, out, gamma);
        else if (flux == SFO_LAXF) SF::Flux::LaxFriedrichsFlux::flux(qL, qR, n, out, gamma);
        else if (flux == SFO_ROE) SF::Flux::RoeFlux::flux(qL, qR, n, out, gamma);
        else throw std::runtime_error("not a splitting flux id");
    });
}

int sfref_lax_wendroff(const double qL[5], const double qR[5], const double n[3], double gamma,
                       double dt, double spacing, double out[5]) {
    return guarded([&] {
        SF::Flux::LaxWendroffFlux::flux(qL, qR, n, dt, spacing, out, gamma);
    });
}

int sfref_char_matrix(const double q[5], const double n[3], double gamma,
                      double L[25], double R[25]) {
    return guarded([&] { SF::Riemann::buildCharacteristicMatrix(q, n, gamma, L, R); });
}

int sfref_rusanov(const double qL[5], const double qR[5], const double n[3],
                  double gamma, double out[5]) {
    return guarded([&] {
        auto eq = SF::Physics::FluidStateModel::makeSingleFluidPerfectGas(
            gamma, 287.05, 0.0, 0.72);
        SF::Numerics::RusanovEOS::faceFlux(*eq, qL, qR, {n[0], n[1], n[2]}, out, 5);
    });
}

}  // extern "C"

namespace {
template <int NS, class Core>
void char_face(const double* st, const double cof[3], Core core, int flux, double gamma,
               double dt, double spacing, double out[5]) {
    double stencil[NS][5];
    for (int s = 0; s < NS; ++s)
        for (int v = 0; v < 5; ++v) stencil[s][v] = st[5 * s + v];
    const double metrics[4] = {cof[0], cof[1], cof[2], 1.0};
    auto eq = SF::Physics::FluidStateModel::makeSingleFluidPerfectGas(gamma, 287.05, 0.0, 0.72);
    auto face = [&](const double qL[5], const double qR[5], const double n[3], double o[5]) {
        switch (flux) {
        case SFO_HLLC: hllc_face_flux(qL, qR, n, o, gamma); return;
        case SFO_STEGER: SF::Flux::StegerWarmingFlux::flux(qL, qR, n, o, gamma); return;
        case SFO_LAXF: SF::Flux::LaxFriedrichsFlux::flux(qL, qR, n, o, gamma); return;
        case SFO_ROE: SF::Flux::RoeFlux::flux(qL, qR, n, o, gamma); return;
        case SFO_LAXW: SF::Flux::LaxWendroffFlux::flux(qL, qR, n, dt, spacing, o, gamma); return;
        }
        SF::Numerics::RusanovEOS::faceFlux(*eq, qL, qR, {n[0], n[1], n[2]}, o, 5);
    };
    SF::Flux::characteristicWENOFaceFlux<NS>(stencil, metrics, out, core, face, gamma);
}

}  // namespace

extern "C" {

// characteristicWENOFaceFlux<NS> for any scheme / flux id; the stencil
// holds sfo_stencil_size(scheme) nodes.
int sfref_char_face_flux_dt(const double* st, const double cof[3], int scheme, int flux,
                            double gamma, double dt, double spacing, double out[5]) {
    return guarded([&] {
        switch (scheme) {
        case SFO_TENO5: char_face<6>(st, cof, SF::TENO5::teno5_core, flux, gamma, dt, spacing, out); break;
        case SFO_WENO3: char_face<4>(st, cof, SF::WENO3::weno3_core, flux, gamma, dt, spacing, out); break;
        case SFO_WENO7: char_face<8>(st, cof, SF::WENO7::weno7_core, flux, gamma, dt, spacing, out); break;
        default: char_face<6>(st, cof, SF::WENO5::weno5_core, flux, gamma, dt, spacing, out);
        }
    });
}

int sfref_char_face_flux_n(const double* st, const double cof[3], int scheme, int flux,
                           double gamma, double out[5]) {
    return sfref_char_face_flux_dt(st, cof, scheme, flux, gamma, 0.0, 0.0, out);
}

int sfref_char_face_flux(const double st[30], const double cof[3], int scheme,
                         int flux, double gamma, double out[5]) {
    return sfref_char_face_flux_n(st, cof, scheme, flux, gamma, out);
}

// Residual of one state after boundary preparation, on the unique nodes.
int sfref_residual(const sfo_desc* d, const double* soa_in, double* soa_out) {
  return guarded([&] {
    auto c = make_case(*d, soa_in);
    assemble(*c);
    const SF::Field& f = c->field;
    const int ng = f.NG();
    const int nx = d->n[0], ny = d->n[1], nz = d->n[2];
    const size_t N = (size_t)nx * ny * nz;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const bool solved = !f.isSolverBoundaryPoint(i + ng, j + ng, k + ng);
                for (int v = 0; v < 5; ++v)
                    soa_out[v * N + ((size_t)k * ny + j) * nx + i] = solved
                        ? SF::Math::spatialResidual(f, c->ws[0].residual,
                                                    i + ng, j + ng, k + ng, v)
                        : 0.0;
            }
  });
}

namespace {
// One physical step on an assembled case: the body shared by sfref_run and
// the time driver's advance callback.
double step_case(Case& c, SF::State::StateBundle& state,
                 const SF::System::CompiledTimeRecipe& recipe, double cfl, double max_dt,
                 double limit) {
    prepare_boundary(c);
    double dt = SF::Numerics::RusanovEOS::deltaT(c.field, *c.eq, cfl);
    dt = std::min({dt, max_dt, limit});
    state.dt = dt;
    c.dt = dt;
    SF::Time::Explicit::Workspace workspace;
    SF::Time::Explicit::begin(workspace, state.patches, state, recipe, nullptr);
    for (int st = 0; st < recipe.stageCount(); ++st) {
        SF::Time::Explicit::executeStage(
            workspace, st, state.patches, c.ws, state, nullptr,
            [&](const std::vector<SF::Field*>&,
                std::vector<SF::SolverAlgorithm::PatchWorkspace>&, double) { assemble(c); },
            [&](const std::vector<SF::Field*>&) {},
            [&](const std::vector<SF::Field*>& patches, const char* label) {
                for (SF::Field* p : patches)
                    SF::Numerics::validateEquationState(*p, *c.eq, label);
            });
    }
    prepare_boundary(c);
    state.time += dt;
    ++state.step;
    return dt;
}
}  // namespace

extern "C" {

// Time::Driver::run (SF_time.cpp:42-93) with the explicit step as its
// advance callback; returns the number of steps.
int sfref_run_driver(const sfo_desc* d, double* soa_io, double start, double end,
                     int end_step, int write_by_step, int write_steps, double write_time,
                     double cfl, double max_dt, double* dt_seq, int cap, int* nsteps) {
  return guarded([&] {
    auto c = make_case(*d, soa_io);
    SF::State::StateBundle state;
    state.patches = {&c->field};
    state.stateModel = c->eq;
    state.time = start;
    const auto recipe = time_recipe(d->time);
    SF::Time::RunControl rc;
    rc.startTime = start;
    rc.endTime = end;
    rc.endStep = end_step;
    rc.writeByStep = write_by_step != 0;
    rc.writeIntervalSteps = write_steps;
    rc.writeIntervalTime = write_time;
    SF::Time::DriverCallbacks cb;
    int n = 0;
    cb.advance = [&](double limit, int) {
        const double dt = step_case(*c, state, recipe, cfl, max_dt, limit);
        if (dt_seq && n < cap) dt_seq[n] = dt;
        ++n;
        SF::Time::AdvanceResult r;
        r.dt = dt;
        r.time = state.time;
        return r;
    };
    SF::Time::Driver(rc).run(cb);
    *nsteps = n;
    export_state(*c, soa_io);
  });
}

}  // extern "C"

// flow.step.prepare -> flow.dt.compute -> flow.step.begin ->
// explicit.stage.execute x stages -> flow.step.commit, repeated nsteps
// times (SF_flowOperations.cpp:466-522,648-652).
int sfref_run(const sfo_desc* d, double* soa_io, int nsteps, double cfl,
              double max_dt, double* dt_seq) {
  return guarded([&] {
    auto c = make_case(*d, soa_io);
    SF::State::StateBundle state;
    state.patches = {&c->field};
    state.stateModel = c->eq;
    state.time = 0.0;
    const auto recipe = time_recipe(d->time);
    for (int s = 0; s < nsteps; ++s) {
        prepare_boundary(*c);
        double dt = SF::Numerics::RusanovEOS::deltaT(c->field, *c->eq, cfl);
        dt = std::min(dt, max_dt);
        state.dt = dt;
        c->dt = dt;
        if (dt_seq) dt_seq[s] = dt;
        SF::Time::Explicit::Workspace workspace;
        SF::Time::Explicit::begin(workspace, state.patches, state, recipe, nullptr);
        for (int st = 0; st < recipe.stageCount(); ++st) {
            SF::Time::Explicit::executeStage(
                workspace, st, state.patches, c->ws, state, nullptr,
                [&](const std::vector<SF::Field*>&,
                    std::vector<SF::SolverAlgorithm::PatchWorkspace>&, double) {
                    assemble(*c);
                },
                [&](const std::vector<SF::Field*>&) {},
                [&](const std::vector<SF::Field*>& patches, const char* label) {
                    for (SF::Field* p : patches)
                        SF::Numerics::validateEquationState(*p, *c->eq, label);
                });
        }
        prepare_boundary(*c);
        state.time += dt;
        ++state.step;
    }
    export_state(*c, soa_io);
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The same step on every host core: the grid cut into z slabs, one thread
// per slab, each slab the reference's own Field and code path.  Before every
// ghost fill the slabs swap their ng outermost interior planes (the halo
// exchange of the reference's parallel.split decomposition, done here in
// shared memory) and the step's dt is the minimum over the slabs (its
// Allreduce MIN).  The arithmetic per node is the serial run's, so the
// result is the serial run's bit for bit.
namespace {

// A barrier that releases every waiter with an error once any slab failed.
class SlabBarrier {
public:
    explicit SlabBarrier(int n) : n_(n) {}
    void wait() {
        std::unique_lock<std::mutex> lk(m_);
        if (failed_) throw std::runtime_error("another slab failed");
        const unsigned long gen = gen_;
        if (++count_ == n_) {
            count_ = 0;
            ++gen_;
            cv_.notify_all();
            return;
        }
        cv_.wait(lk, [&] { return gen_ != gen || failed_; });
        if (failed_) throw std::runtime_error("another slab failed");
    }
    void fail() {
        std::lock_guard<std::mutex> lk(m_);
        failed_ = true;
        cv_.notify_all();
    }

private:
    std::mutex m_;
    std::condition_variable cv_;
    int n_, count_ = 0;
    unsigned long gen_ = 0;
    bool failed_ = false;
};

}  // namespace

extern "C" {

// nsteps physical steps of a grid periodic in z, cut into `nslabs` z slabs
// of at least ng nodes each, one thread per slab; `seconds` (if not null)
// receives the wall time of the steps alone, without building the slabs.
// Returns 0, or -1 if the description does not allow the cut or a slab
// failed.
int sfref_run_slabs(const sfo_desc* d, double* soa_io, int nsteps, double cfl,
                    double max_dt, int nslabs, double* dt_seq, double* seconds) {
  return guarded([&] {
    const int ng = sfo_halo(d->scheme);
    const int nz = d->n[2];
    if (!d->active[2] || d->bc[2] != SFO_PERIODIC || nslabs < 1 || nz < nslabs * ng)
        throw std::runtime_error("slab run needs an active periodic z axis with at least "
                                 "ng nodes per slab");
    const int nx = d->n[0], ny = d->n[1];
    const size_t plane = (size_t)nx * ny, N = plane * nz;
    std::vector<int> z0(nslabs + 1);
    for (int s = 0; s <= nslabs; ++s) z0[s] = (int)((long long)nz * s / nslabs);
    configure_directions(*d);
    std::vector<std::unique_ptr<Case>> cs(nslabs);
    for (int s = 0; s < nslabs; ++s) {
        sfo_desc sd = *d;
        sd.n[2] = z0[s + 1] - z0[s];
        std::vector<double> part((size_t)5 * plane * sd.n[2]);
        for (int v = 0; v < 5; ++v)
            std::copy(soa_io + v * N + plane * z0[s], soa_io + v * N + plane * z0[s + 1],
                      part.begin() + (size_t)v * plane * sd.n[2]);
        cs[s] = make_case(sd, part.data());
    }
    SlabBarrier bar(nslabs);
    std::vector<double> dts(nslabs);
    for (int s = 0; s < nslabs; ++s) {
        Case& c = *cs[s];
        const Case& lo = *cs[(s + nslabs - 1) % nslabs];
        const Case& hi = *cs[(s + 1) % nslabs];
        c.z_exchange = [&c, &lo, &hi, &bar, ng, nx, ny] {
            bar.wait();   // every slab has finished writing its interior
            SF::Field& f = c.field;
            const int nzs = c.d.n[2], nzl = lo.d.n[2];
            for (int l = 0; l < ng; ++l)
                for (int j = ng; j < ng + ny; ++j)
                    for (int i = ng; i < ng + nx; ++i)
                        for (int v = 0; v < 5; ++v) {
                            f(i, j, l, v) = lo.field(i, j, nzl + l, v);
                            f(i, j, ng + nzs + l, v) = hi.field(i, j, ng + l, v);
                        }
            bar.wait();   // every slab has read its neighbours
        };
    }
    const auto recipe = time_recipe(d->time);
    auto body = [&](int s) {
        Case& c = *cs[s];
        SF::State::StateBundle state;
        state.patches = {&c.field};
        state.stateModel = c.eq;
        state.time = 0.0;
        for (int step = 0; step < nsteps; ++step) {
            prepare_boundary(c);
            dts[s] = SF::Numerics::RusanovEOS::deltaT(c.field, *c.eq, cfl);
            bar.wait();
            double dt = *std::min_element(dts.begin(), dts.end());
            dt = std::min(dt, max_dt);
            bar.wait();   // dts is rewritten only after every slab read it
            state.dt = dt;
            c.dt = dt;
            if (s == 0 && dt_seq) dt_seq[step] = dt;
            SF::Time::Explicit::Workspace workspace;
            SF::Time::Explicit::begin(workspace, state.patches, state, recipe, nullptr);
            for (int st = 0; st < recipe.stageCount(); ++st) {
                SF::Time::Explicit::executeStage(
                    workspace, st, state.patches, c.ws, state, nullptr,
                    [&](const std::vector<SF::Field*>&,
                        std::vector<SF::SolverAlgorithm::PatchWorkspace>&, double) {
                        assemble(c);
                    },
                    [&](const std::vector<SF::Field*>&) {},
                    [&](const std::vector<SF::Field*>& patches, const char* label) {
                        for (SF::Field* p : patches)
                            SF::Numerics::validateEquationState(*p, *c.eq, label);
                    });
            }
            prepare_boundary(c);
            state.time += dt;
            ++state.step;
        }
    };
    std::vector<std::thread> th;
    std::vector<std::string> err(nslabs);
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < nslabs; ++s)
        th.emplace_back([&, s] {
            try {
                body(s);
            } catch (const std::exception& e) {
                err[s] = e.what();
                bar.fail();
            }
        });
    for (auto& t : th) t.join();
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int s = 0; s < nslabs; ++s)
        if (!err[s].empty()) throw std::runtime_error("slab " + std::to_string(s) + ": " + err[s]);
    for (int s = 0; s < nslabs; ++s) {
        const Case& c = *cs[s];
        const int nzs = c.d.n[2];
        for (int k = 0; k < nzs; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    for (int v = 0; v < 5; ++v)
                        soa_io[v * N + ((size_t)(z0[s] + k) * ny + j) * nx + i] =
                            c.field(i + ng, j + ng, k + ng, v);
    }
  });
}

}  // extern "C"
