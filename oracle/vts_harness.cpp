// vts_harness.cpp -- the reference's own structured-grid output
// (ResultWriter::save -> writeToDisk + recordAndWritePVD,
// infrastructure/io/SF_vtkOutput.cpp:59-67,113-125,
// SF_vtkStructuredWriter.cpp:22-208, SF_vtkIndexWriter.cpp:283-321) over a
// reference Field built programmatically, for byte comparison with the
// library's writer.  TEST INFRASTRUCTURE ONLY.
//
// What is the reference's code here: ResultWriter (the four
// infrastructure/io/SF_vtk*.cpp units), the block points
// (MeshGen::generateBlockPoints, SF_meshGen.cpp:32-74), Field and its
// thermodynamic state (SF_field.cpp, the single-fluid perfect-gas model).
// What is this file's: building the Field from a block, its node sets and
// a state, and the list of output arrays the single-fluid application
// hands the writer -- the fluid fields of configureOutput's provider
// (app/application/model/SF_output.cpp:8-26, renamed by the case's field
// names rho / U / p) and the extra scalars of
// app/application/execution/SF_singleFluid.cpp:390-410 -- in the name order
// the provider's registry returns them.
#include "core/field/SF_field.h"
#include "infrastructure/io/SF_resultWriter.h"
#include "infrastructure/mesh/createMesh/SF_meshGen.h"
#include "models/physics/fluidStateModel/SF_factory.h"

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::vector<std::string> split(const char* text) {
    std::vector<std::string> out;
    std::stringstream ss(text ? text : "");
    std::string item;
    while (std::getline(ss, item, ','))
        if (!item.empty()) out.push_back(item);
    return out;
}

}  // namespace

extern "C" {

// nodes: block node layers (inactive axes included); vertex[3 v + a]: the
// scaled hex vertices; active: the axes the state varies along; states:
// count states of N x 5 (SoA over the unique nodes of the active axes);
// sets: nsets names with masks[s * nodes] over the block nodes (i fastest);
// rho_sets / u_sets / p_sets: comma-separated set names of each field's
// boundary settings, in order.  Writes <dir>/<job>_<step>.vts (by_step) or
// <dir>/<job>_t<time>.vts and <dir>/<job>.pvd.  Returns 0, or 1 on an
// exception (message in err).
int sfref_write_vts(const int nodes[3], const double* vertex, const int active[3], double gamma,
                    double gas_constant, int count, const int* steps, const double* times,
                    const double* states, int nsets, const char* const* set_names,
                    const uint8_t* masks, const char* rho_sets, const char* u_sets,
                    const char* p_sets, const char* dir, const char* job, int by_step,
                    char* err, size_t err_len) {
    try {
        const int nx = nodes[0], ny = nodes[1], nz = nodes[2];
        const int ng = 3;
        auto model = SF::Physics::FluidStateModel::makeSingleFluidPerfectGas(gamma, gas_constant,
                                                                              0.0, 0.72);
        SF::Field f;
        f.setStateModel(model);
        f.setup(nx, ny, nz, ng, 5);
        double v[8][3];
        for (int k = 0; k < 8; ++k)
            for (int a = 0; a < 3; ++a) v[k][a] = vertex[3 * k + a];
        const int nc[3] = {nx - 1, ny - 1, nz - 1};
        const double grade[3] = {1.0, 1.0, 1.0};
        std::vector<double> X, Y, Z;
        SF::MeshGen::generateBlockPoints(v, nc, grade, X, Y, Z);
        size_t p = 0;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i, ++p) {
                    f.X(i + ng, j + ng, k + ng) = X[p];
                    f.Y(i + ng, j + ng, k + ng) = Y[p];
                    f.Z(i + ng, j + ng, k + ng) = Z[p];
                }
        std::map<std::string, std::vector<int>> sets;
        const size_t total = (size_t)nx * ny * nz;
        for (int s = 0; s < nsets; ++s) {
            std::vector<int>& list = sets[set_names[s]];
            size_t n = 0;
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i, ++n)
                        if (masks[(size_t)s * total + n]) list.push_back(f.getIdx(i + ng, j + ng, k + ng));
        }
        f.setBoundarySets(sets);

        SF::ResultWriterConfig cfg;
        cfg.caseDir = dir;
        cfg.outputDir = ".";
        cfg.jobName = job;
        for (const auto& s : split(rho_sets)) cfg.densityBoundary.push_back({s, SF::ZERO_GRADIENT, 0.0});
        for (const auto& s : split(u_sets))
            cfg.velocityBoundary.push_back({s, SF::ZERO_GRADIENT, SF::Vector3()});
        for (const auto& s : split(p_sets)) cfg.pressureBoundary.push_back({s, SF::ZERO_GRADIENT, 0.0});
        SF::ResultWriter writer(cfg);

        using Array = SF::ResultWriter::ScalarField;
        std::vector<Array> arrays;
        auto scalar = [&](const char* name, std::function<double(int, int, int)> fn) {
            Array a;
            a.name = name;
            a.valueAt = std::move(fn);
            arrays.push_back(std::move(a));
        };
        scalar("MomentumX", [&f](int i, int j, int k) { return f(i, j, k, SF::RU); });
        scalar("MomentumY", [&f](int i, int j, int k) { return f(i, j, k, SF::RV); });
        scalar("MomentumZ", [&f](int i, int j, int k) { return f(i, j, k, SF::RW); });
        scalar("SoundSpeed", [&f](int i, int j, int k) { return f.thermodynamicState(i, j, k).soundSpeed; });
        scalar("Temperature", [&f](int i, int j, int k) { return f.thermodynamicState(i, j, k).temperature; });
        scalar("TotalEnergyDensity", [&f](int i, int j, int k) { return f(i, j, k, SF::E); });
        {
            Array u;
            u.name = "U";
            u.components = 3;
            u.componentAt = [&f](int i, int j, int k, int c) {
                return f.thermodynamicState(i, j, k).velocity[c];
            };
            arrays.push_back(std::move(u));
        }
        scalar("p", [&f](int i, int j, int k) { return f.thermodynamicState(i, j, k).pressure; });
        scalar("rho", [&f](int i, int j, int k) { return f.thermodynamicState(i, j, k).density; });

        const int rn[3] = {active[0] ? nx : 1, active[1] ? ny : 1, active[2] ? nz : 1};
        const size_t N = (size_t)rn[0] * rn[1] * rn[2];
        for (int w = 0; w < count; ++w) {
            const double* soa = states + (size_t)w * 5 * N;
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        const size_t u = ((size_t)(active[2] ? k : 0) * rn[1] + (active[1] ? j : 0)) * rn[0] +
                                         (active[0] ? i : 0);
                        for (int var = 0; var < 5; ++var) f(i + ng, j + ng, k + ng, var) = soa[var * N + u];
                    }
            f.invalidateThermodynamicCache();
            // timeStep output names the file by step and indexes it at the
            // physical time (saveStep); runTime output names it by time
            if (by_step) writer.saveStep(f, steps[w], times[w], arrays);
            else writer.save(f, times[w], arrays);
        }
        return 0;
    } catch (const std::exception& e) {
        if (err && err_len) std::snprintf(err, err_len, "%s", e.what());
        return 1;
    }
}

}  // extern "C"
