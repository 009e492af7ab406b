// sf_vts.cpp -- see sf_vts.h.
#include "sf_vts.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace sfamd::host {

namespace {

std::string g17(double v) {
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

// formatStepName / formatTimeName (SF_vtkOutput.cpp:23-41)
std::string step_name(const std::string& job, int step) {
    std::ostringstream o;
    o << job << "_" << std::setw(6) << std::setfill('0') << step;
    return o.str();
}

std::string time_name(const std::string& job, double t) {
    std::ostringstream o;
    o << job << "_t" << std::fixed << std::setprecision(6) << t;
    std::string s = o.str();
    if (s.find('.') != std::string::npos) {
        while (!s.empty() && s.back() == '0') s.pop_back();
        if (!s.empty() && s.back() == '.') s.pop_back();
    }
    return s;
}

// Trilinear block points of the structured generator for uniform grading
// (MeshGen::generateBlockPoints, infrastructure/mesh/createMesh/
// SF_meshGen.cpp:32-75): xi = i/(n-1), weights summed over the 8 vertices in
// (kk, jj, ii) order.
void block_points(const VtsLayout& L, std::vector<double>& x, std::vector<double>& y,
                  std::vector<double>& z) {
    const int* np = L.nodes;
    const size_t total = (size_t)np[0] * np[1] * np[2];
    x.resize(total);
    y.resize(total);
    z.resize(total);
    static const int map[2][2][2] = {{{0, 1}, {3, 2}}, {{4, 5}, {7, 6}}};
    auto param = [](int i, int n) { return n <= 1 ? 0.0 : (double)i / (n - 1); };
    auto N = [](int i, double t) { return i ? t : (1.0 - t); };
    size_t idx = 0;
    for (int k = 0; k < np[2]; ++k) {
        const double zeta = param(k, np[2]);
        for (int j = 0; j < np[1]; ++j) {
            const double eta = param(j, np[1]);
            for (int i = 0; i < np[0]; ++i) {
                const double xi = param(i, np[0]);
                double px = 0, py = 0, pz = 0;
                for (int kk = 0; kk < 2; ++kk)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int ii = 0; ii < 2; ++ii) {
                            const double w = N(ii, xi) * N(jj, eta) * N(kk, zeta);
                            const int vi = map[kk][jj][ii];
                            px += w * L.vertex[vi][0];
                            py += w * L.vertex[vi][1];
                            pz += w * L.vertex[vi][2];
                        }
                x[idx] = px;
                y[idx] = py;
                z[idx] = pz;
                ++idx;
            }
        }
    }
}

// SingleFluidStateModel::close + PerfectGasEOS::fromDensityEnergy
// (SF_singleFluidStateModel.cpp:49-71, SF_perfectGasEOS.h:26-32).
struct Closed {
    double rho, u[3], p, T, c;
};

Closed close(const double q[5], double gamma, double R) {
    Closed s;
    s.rho = q[0];
    for (int d = 0; d < 3; ++d) s.u[d] = q[1 + d] / s.rho;
    const double kinetic = 0.5 * s.rho * (s.u[0] * s.u[0] + s.u[1] * s.u[1] + s.u[2] * s.u[2]);
    const double eint = q[4] - kinetic;
    const double e = eint / s.rho;
    const double thermal = e - 0.0;
    s.p = (gamma - 1.0) * s.rho * thermal;
    s.T = thermal * (gamma - 1.0) / R;
    s.c = std::sqrt(gamma * s.p / s.rho);
    return s;
}

}  // namespace

VtsSeries::VtsSeries(const std::string& directory, const std::string& job, const VtsLayout& layout,
                     bool by_step)
    : dir_(directory), job_(job), lay_(layout), by_step_(by_step) {
    std::filesystem::create_directories(dir_);
    block_points(lay_, x_, y_, z_);
}

void VtsSeries::write(const double* soa, int step, double time) {
    const std::string base = by_step_ ? step_name(job_, step) : time_name(job_, time);
    const std::string file = base + ".vts";
    const std::string path = dir_ + "/" + file;
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot open " + path);
    const int nx = lay_.nodes[0], ny = lay_.nodes[1], nz = lay_.nodes[2];
    // unique-node index of the state for a block node (empty axes collapse)
    const int rn[3] = {lay_.active[0] ? nx : 1, lay_.active[1] ? ny : 1, lay_.active[2] ? nz : 1};
    const size_t N = (size_t)rn[0] * rn[1] * rn[2];
    auto state_index = [&](int i, int j, int k) {
        const int a = lay_.active[0] ? i : 0, b = lay_.active[1] ? j : 0, c = lay_.active[2] ? k : 0;
        return ((size_t)c * rn[1] + b) * rn[0] + a;
    };

    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    out << "  <StructuredGrid WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1
        << "\">\n";
    out << "    <Piece Extent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1 << "\">\n";
    out << "      <Points>\n";
    out << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    size_t idx = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i, ++idx)
                out << g17(x_[idx]) << " " << g17(y_[idx]) << " " << g17(z_[idx]) << " ";
            out << "\n";
        }
    out << "        </DataArray>\n";
    out << "      </Points>\n";
    out << "      <PointData>\n";

    // BC_ID arrays: the sets the boundary settings reference, numbered in
    // the order density, velocity, pressure settings name them.
    std::vector<std::string> names;
    for (const char* f : {"rho", "U", "p"}) {
        auto it = lay_.field_sets.find(f);
        if (it == lay_.field_sets.end()) continue;
        for (const std::string& s : it->second)
            if (std::find(names.begin(), names.end(), s) == names.end()) names.push_back(s);
    }
    auto id_of = [&](const std::string& s) {
        return (int)(std::find(names.begin(), names.end(), s) - names.begin()) + 1;
    };
    const size_t total = (size_t)nx * ny * nz;
    std::map<std::string, std::vector<int>> masks;
    for (const char* f : {"rho", "U", "p"}) {
        std::vector<int> m(total, 0);
        auto it = lay_.field_sets.find(f);
        if (it != lay_.field_sets.end())
            for (const std::string& s : it->second) {
                const int id = id_of(s);
                size_t n = 0;
                for (int k = 0; k < nz; ++k)
                    for (int j = 0; j < ny; ++j)
                        for (int i = 0; i < nx; ++i, ++n)
                            if (lay_.in_set(s, i, j, k)) m[n] = id;
            }
        masks[f] = std::move(m);
    }
    std::vector<int> all = masks["rho"];
    for (size_t n = 0; n < total; ++n) {
        if (masks["U"][n]) all[n] = masks["U"][n];
        if (masks["p"][n]) all[n] = masks["p"][n];
    }
    auto write_ids = [&](const char* name, const std::vector<int>& m) {
        out << "        <DataArray type=\"Int32\" Name=\"" << name << "\" format=\"ascii\">\n";
        size_t n = 0;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                for (int i = 0; i < nx; ++i, ++n) out << m[n] << " ";
                out << "\n";
            }
        out << "        </DataArray>\n";
    };
    write_ids("BC_ID", all);
    write_ids("BC_ID_rho", masks["rho"]);
    write_ids("BC_ID_U", masks["U"]);
    write_ids("BC_ID_p", masks["p"]);
    out << "        <!-- BC_ID legend: 0=interior/internal-shared/unconfigured";
    for (size_t s = 0; s < names.size(); ++s) out << ", " << (s + 1) << "=" << names[s];
    out << "; variable BC_ID arrays use the same legend and follow per-variable boundaryField "
           "order on overlaps -->\n";

    // Point arrays, sorted by name as the output registry orders them.
    struct Array {
        const char* name;
        int components;
        std::function<double(const double*, const Closed&, int)> value;
    };
    const std::vector<Array> arrays = {
        {"MomentumX", 1, [](const double* q, const Closed&, int) { return q[1]; }},
        {"MomentumY", 1, [](const double* q, const Closed&, int) { return q[2]; }},
        {"MomentumZ", 1, [](const double* q, const Closed&, int) { return q[3]; }},
        {"SoundSpeed", 1, [](const double*, const Closed& s, int) { return s.c; }},
        {"Temperature", 1, [](const double*, const Closed& s, int) { return s.T; }},
        {"TotalEnergyDensity", 1, [](const double* q, const Closed&, int) { return q[4]; }},
        {"U", 3, [](const double*, const Closed& s, int c) { return s.u[c]; }},
        {"p", 1, [](const double*, const Closed& s, int) { return s.p; }},
        {"rho", 1, [](const double*, const Closed& s, int) { return s.rho; }},
    };
    for (const Array& a : arrays) {
        out << "        <DataArray type=\"Float64\" Name=\"" << a.name << "\" NumberOfComponents=\""
            << a.components << "\" format=\"ascii\">\n";
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const size_t u = state_index(i, j, k);
                    const double q[5] = {soa[u], soa[N + u], soa[2 * N + u], soa[3 * N + u],
                                         soa[4 * N + u]};
                    const Closed s = close(q, lay_.gamma, lay_.gas_constant);
                    for (int c = 0; c < a.components; ++c) out << g17(a.value(q, s, c)) << ' ';
                    out << '\n';
                }
        out << "        </DataArray>\n";
    }
    out << "      </PointData>\n";
    out << "    </Piece>\n";
    out << "  </StructuredGrid>\n";
    out << "</VTKFile>\n";
    if (!out) throw std::runtime_error("write failed: " + path);
    out.close();
    files_.push_back(file);
    if (pvd_.empty() || pvd_.back().second != file || std::fabs(pvd_.back().first - time) >= 1.0e-14)
        pvd_.emplace_back(time, file);
    flush_pvd();
}

void VtsSeries::flush_pvd() const {
    if (pvd_.empty()) return;
    const std::string path = dir_ + "/" + job_ + ".pvd";
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot open " + path);
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    out << "  <Collection>\n";
    for (const auto& e : pvd_) {
        char t[64];
        std::snprintf(t, sizeof t, "%.12e", e.first);
        out << "    <DataSet timestep=\"" << t << "\" group=\"\" part=\"0\" file=\"" << e.second
            << "\"/>\n";
    }
    out << "  </Collection>\n";
    out << "</VTKFile>\n";
}

void VtsSeries::finish() { flush_pvd(); }

}  // namespace sfamd::host
