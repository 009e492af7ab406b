// sf_vts.h -- structured-grid VTK output of a single-fluid state in the
// reference's layout: ResultWriter::writeToDisk
// (infrastructure/io/SF_vtkStructuredWriter.cpp:22-208) with the arrays the
// single-fluid application registers (app/application/model/SF_output.cpp:
// 8-26, app/application/execution/SF_singleFluid.cpp:390-410), file names
// of SF_vtkOutput.cpp:23-41 and the PVD index of
// SF_vtkIndexWriter.cpp:283-321.
#pragma once

#include <functional>
#include <map>
#include <string>
#include <vector>

namespace sfamd::host {

struct VtsLayout {
    int nodes[3] = {1, 1, 1};      // node layers of the block, empty axes included
    double vertex[8][3] = {};      // scaled block vertices (hex order)
    int active[3] = {1, 0, 0};     // axes the state varies along
    double gamma = 1.4;
    double gas_constant = 287.05;
    // boundary settings of rho, U and p: set names in the order the
    // reference iterates them
    std::map<std::string, std::vector<std::string>> field_sets;
    std::function<bool(const std::string&, int, int, int)> in_set;
};

class VtsSeries {
public:
    // by_step: name files <job>_<step, 6 digits> (writeControl timeStep),
    // otherwise <job>_t<time> (runTime).
    VtsSeries(const std::string& directory, const std::string& job, const VtsLayout& layout,
              bool by_step);
    // soa: the state on the unique nodes of the active axes.
    void write(const double* soa, int step, double time);
    void finish();
    const std::vector<std::string>& files() const { return files_; }

private:
    std::string dir_, job_;
    VtsLayout lay_;
    bool by_step_;
    std::vector<double> x_, y_, z_;
    std::vector<std::pair<double, std::string>> pvd_;
    std::vector<std::string> files_;
    void flush_pvd() const;
};

}  // namespace sfamd::host
