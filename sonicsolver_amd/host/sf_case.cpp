// sf_case.cpp -- SonicFile case directory -> the explicit density-based
// path (include/sonic_amd_case.h).
//
// Reading follows SF::CaseIO::read (infrastructure/io/case/SF_case.cpp:
// 251-290): case.yaml, then the registries mesh/mesh.yaml,
// fields/fields.yaml, solvers/solvers.yaml and models/models.yaml; every
// registry entry {type, file} names a document in the registry's own folder
// whose SonicFile header must carry the folder's object and the entry's
// name as type.  Decoding follows CaseAdapter (SF_nativeDecode.cpp:906-1072
// runtime / numerics, 1323-1556 fields) and FieldIO::decode
// (infrastructure/io/field/SF_field.cpp:35-57); named node sets are those of
// the structured generator (infrastructure/mesh/createMesh/SF_meshGen.cpp:
// 514-541).
#include "sonic_amd_case.h"
#include "sf_provider.h"
#include "sf_vts.h"
#include "sf_yaml.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <functional>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace sfamd {
void set_last_error(const std::string& message);
}

namespace fs = std::filesystem;
using sfamd::host::yaml::Node;

namespace {

[[noreturn]] void reject(const std::string& where, const std::string& what) {
    throw std::runtime_error(where + ": " + what);
}

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

struct Document {
    std::string rel;   // path relative to the case directory
    Node body;
};

// Read a SonicFile document and check its header.
Document read_sonic(const fs::path& dir, const std::string& rel, const std::string& object,
                    const std::string& type) {
    const fs::path full = dir / rel;
    if (!fs::exists(full)) reject(rel, "missing file");
    Node n = sfamd::host::yaml::parseFile(full.string());
    if (!n.isMap()) reject(rel, "a SonicFile document is a mapping");
    const Node* h = n.find("SonicFile");
    if (!h || !h->isMap()) reject(rel, "missing SonicFile header");
    const Node* o = h->find("object");
    const Node* t = h->find("type");
    if (!o || !t || o->string() != object || t->string() != type)
        reject(rel, "expected object " + object + " type " + type);
    Document d{rel, Node{}};
    d.body.kind = Node::Map;
    for (auto& m : n.members)
        if (m.first != "SonicFile") d.body.members.push_back(std::move(m));
    return d;
}

// Registry entry -> referenced document (must stay in the registry folder).
Document follow(const fs::path& dir, const Document& reg, const std::string& folder,
                const std::string& object, const std::string& entry) {
    const Node* e = reg.body.find(entry);
    if (!e) reject(reg.rel, "no '" + entry + "' entry");
    if (!e->isMap()) reject(reg.rel, "entry '" + entry + "' must be {type, file}");
    const Node* file = e->find("file");
    if (!file) reject(reg.rel, "entry '" + entry + "' has no file");
    const fs::path rel = fs::path(file->string()).lexically_normal();
    if (rel.is_absolute() || rel.parent_path() != fs::path(folder))
        reject(reg.rel, "entry '" + entry + "': registered object " + object + " must stay in " +
                            folder + "/");
    return read_sonic(dir, rel.string(), object, entry);
}

double number_at(const Node& map, const std::string& key, const std::string& where, double dflt,
                 bool required = false) {
    const Node* n = map.find(key);
    if (!n || n->isNull()) {
        if (required) reject(where, "missing " + key);
        return dflt;
    }
    try {
        return n->number();
    } catch (const std::exception& e) {
        reject(where, key + ": " + e.what());
    }
}

std::string string_at(const Node& map, const std::string& key) {
    const Node* n = map.find(key);
    return (n && n->isScalar()) ? n->text : std::string();
}

struct Vec3 {
    double v[3] = {0, 0, 0};
};

// Initial-value setting of one field: FieldIO::decode splits the
// internalField entry into the "all" value (default / value / initial) and
// named sets (every other key: a value, or {type, value}).
struct Setting {
    std::string set;
    Vec3 value;
};

struct FieldInit {
    int components = 1;
    std::vector<Setting> settings;  // "all" first, then the named sets
};

// The structured block and its named node sets.
struct Mesh {
    int nodes[3] = {1, 1, 1};   // all node layers, empty axes included
    double vertex[8][3] = {};   // scaled block vertices, hex order
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    double d[3] = {1, 1, 1};

    // Node-set membership (SF_meshGen.cpp:514-541).
    bool in_set(const std::string& name, int i, int j, int k) const {
        const int nx = nodes[0], ny = nodes[1], nz = nodes[2];
        if (name == "Left" || name == "inlet") return i == 0;
        if (name == "Right" || name == "outlet") return i == nx - 1;
        if (name == "sideWalls") return ny > 1 && (j == 0 || j == ny - 1);
        if (name == "frontBack") return nz > 1 && (k == 0 || k == nz - 1);
        if (name == "walls")
            return (ny > 1 && (j == 0 || j == ny - 1)) || (nz > 1 && (k == 0 || k == nz - 1));
        if (name == "leftHalf") return i < nx / 2;
        if (name == "rightHalf") return i >= nx / 2;
        if (name == "all") return true;
        return false;
    }
    static bool known_set(const std::string& name) {
        static const std::set<std::string> names = {"Left",   "Right",    "inlet",    "outlet",
                                                    "sideWalls", "frontBack", "walls",
                                                    "leftHalf", "rightHalf", "all"};
        return names.count(name) != 0;
    }
};

// Which axis ends a boundary set touches.
std::vector<std::pair<int, int>> set_ends(const std::string& name, const Mesh& m) {
    std::vector<std::pair<int, int>> e;
    if (name == "Left" || name == "inlet") e.push_back({0, 0});
    if (name == "Right" || name == "outlet") e.push_back({0, 1});
    if ((name == "sideWalls" || name == "walls") && m.nodes[1] > 1) {
        e.push_back({1, 0});
        e.push_back({1, 1});
    }
    if ((name == "frontBack" || name == "walls") && m.nodes[2] > 1) {
        e.push_back({2, 0});
        e.push_back({2, 1});
    }
    return e;
}

}  // namespace

struct sf_case {
    std::string dir, name, job, output_dir;
    Mesh mesh;
    sf_grid_desc desc{};
    std::string convection, time;
    int ghosts = -1;                             // nGhost of the mesh (-1: not given)
    double cfl = 1.0;
    double max_dt = std::numeric_limits<double>::max();
    double gamma = 1.4, gas_constant = 0.0;
    sf_run_control control{};
    bool write_initial = true;
    std::map<std::string, FieldInit> fields;     // rho, U, p
    std::vector<std::string> boundary_sets;      // in file order, for BC_ID arrays
    std::map<std::string, std::vector<std::string>> field_boundary_sets;
    std::vector<double> state;                   // lowered unique-node SoA
    std::vector<std::string> notes;
};

namespace {

void read_mesh(sf_case& c, const fs::path& dir) {
    const Document reg = read_sonic(dir, "mesh/mesh.yaml", "mesh", "registry");
    if (!reg.body.find("generation"))
        reject(reg.rel, "this reader builds the mesh from a generation entry (structuredMesh)");
    const Node* ge = reg.body.find("generation");
    if (ge->isMap() && ge->find("type") && ge->find("type")->text != "structuredMesh")
        reject(reg.rel, "generation type must be structuredMesh");
    const Document gen = follow(dir, reg, "mesh", "mesh", "generation");
    const Node& g = gen.body;
    const double scale = number_at(g, "scale", gen.rel, 1.0);
    const Node* verts = g.find("vertices");
    const Node* blocks = g.find("blocks");
    if (!verts || !verts->isSeq() || verts->items.empty()) reject(gen.rel, "missing vertices");
    if (!blocks || !blocks->isSeq() || blocks->items.size() != 1)
        reject(gen.rel, "exactly one block is supported");
    const Node& b = blocks->items[0];
    const Node* bv = b.find("vertices");
    const Node* cells = b.find("cells");
    if (!bv || !bv->isSeq() || bv->items.size() != 8) reject(gen.rel, "a block has 8 vertices");
    if (!cells || !cells->isSeq() || cells->items.size() != 3) reject(gen.rel, "a block has 3 cell counts");
    if (const Node* gr = b.find("grading")) {
        for (const Node& x : gr->items)
            if (x.number() != 1.0) reject(gen.rel, "only uniform grading (1.0) is supported");
    }
    if (const Node* edges = g.find("edges"))
        if (edges->isSeq() && !edges->items.empty()) reject(gen.rel, "curved edges are not supported");
    auto vertex = [&](int k) {
        const long long id = bv->items[(size_t)k].integer();
        if (id < 0 || id >= (long long)verts->items.size()) reject(gen.rel, "vertex index out of range");
        const Node& p = verts->items[(size_t)id];
        if (!p.isSeq() || p.items.size() != 3) reject(gen.rel, "a vertex has 3 coordinates");
        Vec3 r;
        for (int a = 0; a < 3; ++a) r.v[a] = p.items[(size_t)a].number() * scale;
        return r;
    };
    // hex vertex order 0..7: (x0,y0,z0) (x1,y0,z0) (x1,y1,z0) (x0,y1,z0), then z1
    const Vec3 v0 = vertex(0), v6 = vertex(6);
    static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int k = 0; k < 8; ++k) {
        const Vec3 p = vertex(k);
        for (int a = 0; a < 3; ++a) c.mesh.vertex[k][a] = p.v[a];
        for (int a = 0; a < 3; ++a) {
            const double want = corner[k][a] ? v6.v[a] : v0.v[a];
            if (std::fabs(p.v[a] - want) > 1e-12 * std::max(1.0, std::fabs(want)))
                reject(gen.rel, "the block must be an axis-aligned box");
        }
    }
    for (int a = 0; a < 3; ++a) {
        const long long nc = cells->items[(size_t)a].integer();
        if (nc < 1) reject(gen.rel, "cell counts must be positive");
        c.mesh.nodes[a] = (int)nc + 1;
        c.mesh.lo[a] = v0.v[a];
        c.mesh.hi[a] = v6.v[a];
        if (!(c.mesh.hi[a] > c.mesh.lo[a])) reject(gen.rel, "the box must have positive extent");
        c.mesh.d[a] = (c.mesh.hi[a] - c.mesh.lo[a]) / (double)nc;
    }
    if (const Node* ng = g.find("nGhost")) c.ghosts = (int)ng->integer();
}

void read_fields(sf_case& c, const fs::path& dir) {
    const Document reg = read_sonic(dir, "fields/fields.yaml", "fields", "registry");
    std::map<std::string, int> components;
    for (const auto& m : reg.body.members) {
        if (m.first == "internalField" || m.first == "boundaries") continue;
        const std::string type = m.second.isMap() ? string_at(m.second, "type") : "";
        components[m.first] = type == "vector" ? 3 : 1;
    }
    for (const char* need : {"rho", "U", "p"})
        if (!components.count(need))
            reject(reg.rel, std::string("field ") + need + " is not registered; the density-based "
                            "single-fluid state needs rho, U and p");
    if (components["U"] != 3) reject(reg.rel, "U must be a vector field");
    const Document init = follow(dir, reg, "fields", "fields", "internalField");
    const Document bnd = follow(dir, reg, "fields", "fields", "boundaries");
    for (const char* fname : {"rho", "U", "p"}) {
        const std::string where = init.rel + " " + fname;
        const Node* f = init.body.find(fname);
        if (!f) reject(init.rel, std::string("no initial value for ") + fname);
        FieldInit fi;
        fi.components = components[fname];
        auto value_of = [&](const Node& n, const std::string& ctx) {
            Vec3 v;
            if (fi.components == 1) {
                v.v[0] = n.number();
            } else {
                if (!n.isSeq() || n.items.size() != 3) reject(ctx, "a vector value has 3 components");
                for (int a = 0; a < 3; ++a) v.v[a] = n.items[(size_t)a].number();
            }
            return v;
        };
        if (!f->isMap()) {
            fi.settings.push_back({"all", value_of(*f, where)});
        } else {
            const Node* all = nullptr;
            for (const char* k : {"default", "value", "initial", "uniform"})
                if (!all) all = f->find(k);
            if (!all) reject(where, "missing initial value (default / value / initial)");
            fi.settings.push_back({"all", value_of(*all, where)});
            std::vector<Setting> named;
            for (const auto& m : f->members) {
                if (m.first == "default" || m.first == "value" || m.first == "initial" ||
                    m.first == "uniform" || m.first == "type")
                    continue;
                if (!Mesh::known_set(m.first))
                    reject(where, "set '" + m.first + "' is not one of the structured generator's "
                                  "node sets");
                const Node* val = &m.second;
                if (m.second.isMap()) {
                    const std::string t = lower(string_at(m.second, "type"));
                    if (!t.empty() && t != "fixedvalue")
                        reject(where, "set '" + m.first + "': only fixedValue initial sets are supported");
                    val = m.second.find("value");
                    if (!val) reject(where, "set '" + m.first + "' has no value");
                }
                named.push_back({m.first, value_of(*val, where + " " + m.first)});
            }
            // the decoded object iterates its keys in sorted order
            std::sort(named.begin(), named.end(),
                      [](const Setting& a, const Setting& b) { return a.set < b.set; });
            fi.settings.insert(fi.settings.end(), named.begin(), named.end());
        }
        c.fields[fname] = fi;
    }
    // Boundaries: each end of an active axis carries zeroGradient for rho,
    // U and p, or no condition for any of the three (the reference then
    // neither marks it a solver boundary nor refills its ghosts,
    // SF_applicator.h:51-83,162-187); a set marked empty deactivates its
    // axis.
    int state[3][2];   // 0 unset, 1 zeroGradient, 2 empty
    std::memset(state, 0, sizeof state);
    bool first_field = true;
    for (const char* fname : {"rho", "U", "p"}) {
        const std::string where = bnd.rel + " " + fname;
        const Node* f = bnd.body.find(fname);
        if (!f || !f->isMap()) reject(bnd.rel, std::string("no boundary conditions for ") + fname);
        int fstate[3][2];
        std::memset(fstate, 0, sizeof fstate);
        std::vector<std::pair<std::string, std::string>> conds;
        const Node* dflt = f->find("default");
        for (const auto& m : f->members) {
            if (m.first == "default") continue;
            std::string cond = m.second.isMap() ? string_at(m.second, "type") : m.second.text;
            conds.push_back({m.first, cond});
        }
        if (dflt) {
            const std::string cond = dflt->isMap() ? string_at(*dflt, "type") : dflt->text;
            for (const char* s : {"Left", "Right", "sideWalls", "frontBack"}) {
                bool have = false;
                for (auto& cd : conds) have = have || cd.first == s;
                if (!have) conds.push_back({s, cond});
            }
        }
        for (const auto& [set, cond] : conds) {
            if (!Mesh::known_set(set))
                reject(where, "boundary set '" + set + "' is not one of the structured generator's sets");
            const std::string lc = lower(cond);
            int code = 0;
            if (lc == "zerogradient") code = 1;
            else if (lc == "empty") code = 2;
            else
                reject(where, "boundary '" + set + "': condition '" + cond + "' is not served by this "
                              "path (zeroGradient and empty are)");
            for (auto [axis, end] : set_ends(set, c.mesh)) {
                if (fstate[axis][end] && fstate[axis][end] != code)
                    reject(where, "conflicting conditions on one axis end");
                fstate[axis][end] = code;
            }
            auto& lst = c.field_boundary_sets[fname];
            lst.push_back(set);
            if (std::find(c.boundary_sets.begin(), c.boundary_sets.end(), set) == c.boundary_sets.end())
                c.boundary_sets.push_back(set);
        }
        for (int a = 0; a < 3; ++a)
            for (int e = 0; e < 2; ++e) {
                if (c.mesh.nodes[a] <= 1) continue;
                if (!first_field && state[a][e] != fstate[a][e])
                    reject(bnd.rel, std::string("rho, U and p disagree on the ") +
                                        (e ? "upper" : "lower") + " end of axis " + "xyz"[a] +
                                        " (a condition for some of them and none for the "
                                        "others is not served)");
                state[a][e] = fstate[a][e];
            }
        first_field = false;
    }
    for (int a = 0; a < 3; ++a) {
        if (c.mesh.nodes[a] <= 1) {
            c.desc.active[a] = 0;
            continue;
        }
        if ((state[a][0] == 2) != (state[a][1] == 2))
            reject(bnd.rel, std::string("axis ") + "xyz"[a] + " is empty on one end only");
        c.desc.active[a] = state[a][0] != 2;
        if (!c.desc.active[a]) continue;
        const bool lo = state[a][0] == 1, hi = state[a][1] == 1;
        c.desc.bc[a] = lo && hi ? SF_BC_TRANSMISSIVE : hi ? SF_BC_NONE_LO
                     : lo ? SF_BC_NONE_HI : SF_BC_NONE;
        if (!(lo && hi))
            c.notes.push_back(std::string("axis ") + "xyz"[a] + ": the " +
                              (lo ? "upper" : hi ? "lower" : "lower and upper") +
                              " end carries no rho/U/p condition (node advanced, ghosts hold "
                              "the initial end state)");
    }
    if (!c.desc.active[0]) reject(bnd.rel, "the x axis must be active");
    if (c.desc.active[2] && !c.desc.active[1]) reject(bnd.rel, "z active with y empty is not supported");
}

void read_solvers(sf_case& c, const fs::path& dir) {
    const Document reg = read_sonic(dir, "solvers/solvers.yaml", "solver", "registry");
    // runtime (SF_nativeDecode.cpp:906-960, 962-982)
    const Document rt = follow(dir, reg, "solvers", "solver", "runtime");
    const Node& r = rt.body;
    c.control.start_time = number_at(r, "startTime", rt.rel, 0.0);
    c.control.end_time = number_at(r, "endTime", rt.rel, 0.0);
    c.control.end_step = (int)std::max(0.0, number_at(r, "endStep", rt.rel, 0.0));
    c.cfl = number_at(r, "CFL", rt.rel, 1.0);
    c.max_dt = number_at(r, "maxDeltaT", rt.rel, std::numeric_limits<double>::max());
    if (const Node* wi = r.find("writeInitial")) c.write_initial = wi->boolean();
    const std::string wc = lower(string_at(r, "writeControl"));
    if (wc != "timestep" && wc != "runtime")
        reject(rt.rel, "unsupported writeControl '" + string_at(r, "writeControl") +
                           "'. Supported values are timeStep and runTime.");
    if (const Node* wv = r.find("writeInterval")) {
        if (wc == "timestep") {
            c.control.write_by_step = 1;
            c.control.write_interval_steps = (int)std::max(1.0, wv->number());
        } else {
            c.control.write_by_step = 0;
            c.control.write_interval_time = wv->number();
        }
    }
    const Node* out = r.find("output");
    if (!out || !out->isMap()) reject(rt.rel, "must declare output with jobName and outputDir.");
    c.job = string_at(*out, "jobName");
    c.output_dir = string_at(*out, "outputDir");
    if (c.job.empty()) reject(rt.rel, "output must set jobName.");
    if (c.output_dir.empty()) reject(rt.rel, "output must set outputDir.");
    if (const Node* par = r.find("parallel"))
        if (const Node* en = par->find("enabled"))
            if (en->boolean()) c.notes.push_back("parallel.enabled ignored: one device owns the grid");
    // numerics (SF_nativeDecode.cpp:984-1072)
    const Document nu = follow(dir, reg, "solvers", "solver", "numerics");
    const Node& n = nu.body;
    using namespace sfamd::host;
    TermRecipe recipe;
    bool have_recipe = false;
    if (const Node* terms = n.find("terms")) {
        const std::string conv = string_at(*terms, "convection");
        if (!normalizeToken(conv).empty()) {
            recipe = resolveConvectionTermRecipe(conv);
            have_recipe = true;
        }
        if (!normalizeToken(string_at(*terms, "diffusion")).empty())
            reject(nu.rel, "diffusion terms are not part of this (inviscid) path");
    }
    ConvectionScheme scheme = ConvectionScheme::WENO5;
    FluxSplitter flux = FluxSplitter::Rusanov;
    bool have_scheme = false, have_flux = false;
    if (const Node* conv = n.find("convection")) {
        const std::string d = string_at(*conv, "default");
        if (!d.empty()) {
            scheme = parseConvectionScheme(d);
            have_scheme = true;
        }
        const std::string fl = string_at(*conv, "flux");
        if (!fl.empty()) {
            flux = parseFluxSplitter(fl);
            have_flux = true;
        }
        const std::string rec = lower(string_at(*conv, "reconstruction"));
        if (!rec.empty() && rec != "characteristic")
            reject(nu.rel, "convection.reconstruction '" + rec + "': this path reconstructs in "
                           "characteristic variables");
        const std::string form = normalizeToken(string_at(*conv, "formulation"));
        if (!form.empty() && form != "conservativefluxdifference")
            reject(nu.rel, "convection.formulation must be conservativeFluxDifference");
    }
    if (have_recipe) {
        if (have_scheme && scheme != recipe.convection)
            reject(nu.rel, "terms.convection and convection.default disagree");
    } else {
        if (!have_scheme) reject(nu.rel, "no convection recipe (terms.convection or convection.default)");
        // parseFluxSplitter("") is Steger-Warming (SF_numericsParser.h:77-80)
        recipe = builtInConvectionRecipe(scheme, have_flux ? flux : parseFluxSplitter(""));
    }
    c.convection = recipe.id;
    std::string time_token;
    if (const Node* t = n.find("time")) time_token = string_at(*t, "default");
    const TimeRecipeId tid = resolveTimeRecipe(time_token);
    c.time = compileTimeRecipe(tid).name;
    c.desc.scheme = (int)recipe.convection;
    c.desc.flux = (int)recipe.flux;
    c.desc.time = (int)tid;
    // algorithm: the density-based explicit path
    if (reg.body.find("algorithm")) {
        const Document al = follow(dir, reg, "solvers", "solver", "algorithm");
        const std::string type = string_at(al.body, "type");
        if (!type.empty() && normalizeToken(type) != "densitybase")
            reject(al.rel, "algorithm type '" + type + "': this path is densityBase");
    }
    if (reg.body.find("state")) {
        const Document st = follow(dir, reg, "solvers", "solver", "state");
        if (const Node* use = st.body.find("use")) {
            std::set<std::string> u;
            for (const Node& x : use->items) u.insert(x.text);
            if (u != std::set<std::string>{"rho", "rhoU", "rhoE"})
                reject(st.rel, "the state must be use: [rho, rhoU, rhoE]");
        }
    }
}

void read_models(sf_case& c, const fs::path& dir) {
    Document th;
    const fs::path reg_path = dir / "models/models.yaml";
    if (fs::exists(reg_path)) {
        const Document reg = read_sonic(dir, "models/models.yaml", "models", "registry");
        if (reg.body.find("thermoDynamics"))
            th = follow(dir, reg, "models", "models", "thermoDynamics");
        else
            th = read_sonic(dir, "models/thermoDynamics.yaml", "models", "thermoDynamics");
    } else {
        th = read_sonic(dir, "models/thermoDynamics.yaml", "models", "thermoDynamics");
    }
    // parseThermophysical (app/application/model/SF_nativeBinding.cpp:64-104)
    const Node* sel = th.body.find("thermoDynamics");
    const Node* props = th.body.find("properties");
    if (!sel || !props) reject(th.rel, "needs thermoDynamics and properties");
    const std::string eos = string_at(*sel, "equationOfState");
    if (eos != "perfectGas") reject(th.rel, "equationOfState '" + eos + "': this path is perfectGas");
    const Node* e = props->find("equationOfState");
    if (!e) reject(th.rel, "missing properties.equationOfState");
    c.gamma = number_at(*e, "gamma", th.rel + " properties.equationOfState", 0.0, true);
    c.gas_constant = number_at(*e, "R", th.rel + " properties.equationOfState", 0.0, true);
    if (!(c.gamma > 1.0) || !std::isfinite(c.gamma)) reject(th.rel, "gamma must exceed 1");
    if (const Node* tr = props->find("transport"))
        if (tr->find("mu") && tr->find("mu")->number() != 0.0)
            c.notes.push_back("transport.mu ignored: the explicit path is inviscid");
}

// Initial conserved state on the lowered grid: the settings of each field
// in order ("all", then sets by name), then rho U -> momentum and p ->
// total energy as totalEnergyFromPressure forms it
// (SF_singleFluidStateModel.cpp:73-109).
void build_state(sf_case& c) {
    const Mesh& m = c.mesh;
    int n[3];
    for (int a = 0; a < 3; ++a) n[a] = c.desc.active[a] ? m.nodes[a] : 1;
    const size_t N = (size_t)n[0] * n[1] * n[2];
    c.state.assign(5 * N, 0.0);
    auto value_at = [&](const FieldInit& f, int i, int j, int k, int comp) {
        double v = 0.0;
        for (const Setting& s : f.settings)
            if (m.in_set(s.set, i, j, k)) v = s.value.v[comp];
        return v;
    };
    const FieldInit& rho = c.fields.at("rho");
    const FieldInit& U = c.fields.at("U");
    const FieldInit& p = c.fields.at("p");
    // an empty axis carries no variation: every layer must agree
    for (int a = 0; a < 3; ++a) {
        if (c.desc.active[a] || m.nodes[a] <= 1) continue;
        for (const FieldInit* f : {&rho, &U, &p})
            for (const Setting& s : f->settings) {
                if (s.set == "frontBack" && a == 2) reject("fields/internalField.yaml",
                                                          "an initial set on an empty axis' planes");
                if (s.set == "sideWalls" && a == 1) reject("fields/internalField.yaml",
                                                          "an initial set on an empty axis' planes");
                if (s.set == "walls") reject("fields/internalField.yaml",
                                             "an initial set on an empty axis' planes");
            }
    }
    const double g = c.gamma;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const size_t u = ((size_t)k * n[1] + j) * n[0] + i;
                const double r = value_at(rho, i, j, k, 0);
                const double vel[3] = {value_at(U, i, j, k, 0), value_at(U, i, j, k, 1),
                                       value_at(U, i, j, k, 2)};
                const double pr = value_at(p, i, j, k, 0);
                if (!(r > 0.0) || !(pr > 0.0))
                    reject("fields/internalField.yaml", "initial density and pressure must be positive "
                                                        "on every node");
                const double m1 = r * vel[0], m2 = r * vel[1], m3 = r * vel[2];
                const double e = 0.0 + pr / ((g - 1.0) * r);
                const double kinetic = 0.5 * (1.0 / r) * (m1 * m1 + m2 * m2 + m3 * m3);
                c.state[u] = r;
                c.state[N + u] = m1;
                c.state[2 * N + u] = m2;
                c.state[3 * N + u] = m3;
                c.state[4 * N + u] = r * e + kinetic;
            }
}

int fail(int code, const std::string& msg) {
    sfamd::set_last_error(msg);
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        return fail(SF_E_ARG, e.what());
    } catch (...) {
        return fail(SF_E_ARG, "unknown error");
    }
}

void copy_str(const std::string& s, char* buf, size_t len) {
    if (!buf || len == 0) return;
    const size_t n = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
}

}  // namespace

namespace {

// The output layout of a case: its block, active axes, gas and the sets
// its rho / U / p boundary settings name.
sfamd::host::VtsLayout vts_layout(const sf_case* c) {
    sfamd::host::VtsLayout lay;
    for (int a = 0; a < 3; ++a) {
        lay.nodes[a] = c->mesh.nodes[a];
        lay.active[a] = c->desc.active[a];
    }
    std::memcpy(lay.vertex, c->mesh.vertex, sizeof lay.vertex);
    lay.gamma = c->gamma;
    lay.gas_constant = c->gas_constant;
    // the decoded boundary object iterates its sets in sorted order
    for (auto [f, sets] : c->field_boundary_sets) {
        std::sort(sets.begin(), sets.end());
        sets.erase(std::unique(sets.begin(), sets.end()), sets.end());
        lay.field_sets[f] = sets;
    }
    lay.in_set = [mesh = c->mesh](const std::string& s, int i, int j, int k) {
        return mesh.in_set(s, i, j, k);
    };
    return lay;
}

}  // namespace

extern "C" {

int sf_case_open(const char* case_dir, sf_case** out) {
    if (!case_dir || !out) return fail(SF_E_ARG, "sf_case_open: null argument");
    *out = nullptr;
    return guarded([&] {
        auto c = std::make_unique<sf_case>();
        const fs::path dir = fs::path(case_dir);
        if (!fs::is_directory(dir)) reject(case_dir, "not a directory");
        c->dir = dir.string();
        // case.yaml (SF_case.cpp:278-283)
        const Document meta = read_sonic(dir, "case.yaml", "case", "registry");
        const Node* fv = meta.body.find("formatVersion");
        if (!fv || fv->integer() != 1) reject("case.yaml", "unsupported or missing formatVersion");
        c->name = string_at(meta.body, "name");
        if (c->name.empty()) reject("case.yaml", "missing name");
        read_mesh(*c, dir);
        read_fields(*c, dir);
        read_solvers(*c, dir);
        read_models(*c, dir);
        for (int a = 0; a < 3; ++a) {
            c->desc.n[a] = c->desc.active[a] ? c->mesh.nodes[a] : 1;
            if (!c->desc.active[a]) c->desc.bc[a] = SF_BC_TRANSMISSIVE;
            c->desc.d[a] = c->mesh.d[a];
        }
        // checkGhostDepth (SF_stencil.h:23-34): the mesh's ghost layers must
        // cover the recipe's halo
        const int halo = c->desc.scheme == SF_SCHEME_WENO3 ? 2
                       : c->desc.scheme == SF_SCHEME_WENO7 ? 4 : 3;
        if (c->ghosts >= 0 && c->ghosts < halo)
            reject("mesh/generation.yaml", "nGhost " + std::to_string(c->ghosts) +
                                               " is below the halo " + std::to_string(halo) +
                                               " that " + c->convection + " needs");
        c->desc.gamma = c->gamma;
        c->desc.device = 0;
        if (c->control.end_step == 0 && !(c->control.end_time > c->control.start_time))
            reject("solvers/runtime.yaml", "needs endTime after startTime or an endStep");
        build_state(*c);
        *out = c.release();
        return 0;
    });
}

int sf_case_close(sf_case* c) {
    delete c;
    return 0;
}

int sf_case_grid(const sf_case* c, sf_grid_desc* desc) {
    if (!c || !desc) return fail(SF_E_ARG, "sf_case_grid: null argument");
    *desc = c->desc;
    return 0;
}

int sf_case_state(const sf_case* c, double* soa) {
    if (!c || !soa) return fail(SF_E_ARG, "sf_case_state: null argument");
    std::memcpy(soa, c->state.data(), sizeof(double) * c->state.size());
    return 0;
}

int sf_case_numerics(const sf_case* c, char* convection, size_t convection_len, char* time,
                     size_t time_len, double* cfl, double* max_dt) {
    if (!c) return fail(SF_E_ARG, "sf_case_numerics: null case");
    copy_str(c->convection, convection, convection_len);
    copy_str(c->time, time, time_len);
    if (cfl) *cfl = c->cfl;
    if (max_dt) *max_dt = c->max_dt;
    return 0;
}

int sf_case_run_control(const sf_case* c, sf_run_control* control) {
    if (!c || !control) return fail(SF_E_ARG, "sf_case_run_control: null argument");
    *control = c->control;
    return 0;
}

int sf_case_describe(const sf_case* c, char* buf, size_t len) {
    if (!c) return fail(SF_E_ARG, "sf_case_describe: null case");
    std::ostringstream o;
    o.precision(17);
    o << "case " << c->name << "\n";
    o << "grid " << c->desc.n[0] << " x " << c->desc.n[1] << " x " << c->desc.n[2] << " active "
      << c->desc.active[0] << c->desc.active[1] << c->desc.active[2] << " spacing " << c->desc.d[0]
      << " " << c->desc.d[1] << " " << c->desc.d[2] << "\n";
    o << "mesh nodes " << c->mesh.nodes[0] << " " << c->mesh.nodes[1] << " " << c->mesh.nodes[2]
      << " box [" << c->mesh.lo[0] << "," << c->mesh.hi[0] << "]x[" << c->mesh.lo[1] << ","
      << c->mesh.hi[1] << "]x[" << c->mesh.lo[2] << "," << c->mesh.hi[2] << "]\n";
    o << "convection " << c->convection << " time " << c->time << " CFL " << c->cfl
      << " maxDeltaT " << c->max_dt << "\n";
    o << "gamma " << c->gamma << " R " << c->gas_constant << "\n";
    o << "run start " << c->control.start_time << " end " << c->control.end_time << " endStep "
      << c->control.end_step << (c->control.write_by_step ? " write every " : " write interval ")
      << (c->control.write_by_step ? (double)c->control.write_interval_steps
                                   : c->control.write_interval_time)
      << " writeInitial " << c->write_initial << "\n";
    o << "output " << c->output_dir << "/" << c->job << "\n";
    for (const auto& [name, f] : c->fields) {
        o << "field " << name << ":";
        for (const Setting& s : f.settings) {
            o << " " << s.set << "=";
            if (f.components == 1) o << s.value.v[0];
            else o << "(" << s.value.v[0] << "," << s.value.v[1] << "," << s.value.v[2] << ")";
        }
        o << "\n";
    }
    for (const std::string& n : c->notes) o << "note " << n << "\n";
    copy_str(o.str(), buf, len);
    return 0;
}

int sf_case_mesh(const sf_case* c, int nodes[3], double vertex[24], double* gas_constant) {
    if (!c) return fail(SF_E_ARG, "sf_case_mesh: null case");
    for (int a = 0; a < 3; ++a)
        if (nodes) nodes[a] = c->mesh.nodes[a];
    if (vertex)
        for (int v = 0; v < 8; ++v)
            for (int a = 0; a < 3; ++a) vertex[3 * v + a] = c->mesh.vertex[v][a];
    if (gas_constant) *gas_constant = c->gas_constant;
    return 0;
}

int sf_case_boundary_sets(const sf_case* c, const char* field, char* buf, size_t len) {
    if (!c || !field || !buf || len == 0) return fail(SF_E_ARG, "sf_case_boundary_sets: invalid argument");
    const sfamd::host::VtsLayout lay = vts_layout(c);
    std::string out;
    auto it = lay.field_sets.find(field);
    if (it != lay.field_sets.end())
        for (const std::string& s : it->second) out += (out.empty() ? "" : ",") + s;
    if (out.size() + 1 > len) return fail(SF_E_ARG, "sf_case_boundary_sets: buffer too small");
    std::memcpy(buf, out.c_str(), out.size() + 1);
    return 0;
}

int sf_case_set_mask(const sf_case* c, const char* set, uint8_t* mask) {
    if (!c || !set || !mask) return fail(SF_E_ARG, "sf_case_set_mask: invalid argument");
    const Mesh& m = c->mesh;
    size_t n = 0;
    for (int k = 0; k < m.nodes[2]; ++k)
        for (int j = 0; j < m.nodes[1]; ++j)
            for (int i = 0; i < m.nodes[0]; ++i, ++n) mask[n] = m.in_set(set, i, j, k) ? 1 : 0;
    return 0;
}

int sf_case_write_vts(const sf_case* c, const char* output_dir, int count, const int* steps,
                      const double* times, const double* states) {
    if (!c || !output_dir || count < 1 || !steps || !times || !states)
        return fail(SF_E_ARG, "sf_case_write_vts: invalid argument");
    const size_t per = c->state.size();
    try {
        sfamd::host::VtsSeries series(output_dir, c->job, vts_layout(c),
                                      c->control.write_by_step != 0);
        for (int w = 0; w < count; ++w) series.write(states + per * (size_t)w, steps[w], times[w]);
        series.finish();
    } catch (const std::exception& e) {
        return fail(SF_E_ARG, e.what());
    }
    return 0;
}

int sf_case_run(sf_case* c, const char* output_dir, double* final_soa, int* steps,
                double* end_time) {
    if (!c) return fail(SF_E_ARG, "sf_case_run: null case");
    sf_solver* solver = nullptr;
    if (int rc = sf_solver_create(&c->desc, c->convection.c_str(), c->time.c_str(), c->cfl,
                                  c->max_dt, &solver))
        return rc;
    sf_grid* grid = nullptr;
    int rc = sf_solver_grid(solver, &grid);
    if (!rc) rc = sf_state_upload(grid, c->state.data());
    struct Out {
        sf_case* c;
        sf_grid* grid;
        std::unique_ptr<sfamd::host::VtsSeries> series;
        std::vector<double> host;
        int status = 0;
    } out{c, grid, nullptr, std::vector<double>(c->state.size()), 0};
    if (!rc && output_dir) {
        try {
            out.series = std::make_unique<sfamd::host::VtsSeries>(output_dir, c->job, vts_layout(c),
                                                                  c->control.write_by_step != 0);
            if (c->write_initial) out.series->write(c->state.data(), 0, c->control.start_time);
        } catch (const std::exception& e) {
            rc = fail(SF_E_ARG, e.what());
        }
    }
    int done = 0;
    double t = 0.0;
    if (!rc) {
        auto cb = [](void* user, int step, double time, double) -> int {
            auto* o = static_cast<Out*>(user);
            if (!o->series) return 0;
            if (sf_state_download(o->grid, o->host.data()) != 0) {
                o->status = SF_E_HIP;
                return 1;
            }
            try {
                o->series->write(o->host.data(), step, time);
            } catch (const std::exception& e) {
                sfamd::set_last_error(e.what());
                o->status = SF_E_ARG;
                return 1;
            }
            return 0;
        };
        rc = sf_solver_run(solver, &c->control, cb, &out, &done, &t);
        if (!rc && out.status) rc = out.status;
        if (!rc && out.series) {
            try {
                out.series->finish();
            } catch (const std::exception& e) {
                rc = fail(SF_E_ARG, e.what());
            }
        }
        if (!rc && final_soa) rc = sf_state_download(grid, final_soa);
    }
    sf_solver_destroy(solver);
    if (steps) *steps = done;
    if (end_time) *end_time = t;
    return rc;
}

}  // extern "C"
