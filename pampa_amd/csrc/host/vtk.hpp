// Legacy ASCII VTK output (vtk.cxx:15-170): one file output_<n/dn>.vtk per
// written step, the mesh first, cell fields appended.
#pragma once

#include "common.hpp"

namespace pampa {
namespace vtk {

extern bool on;  // -vtk or `vtk 1`
extern int dn;   // -vtk_interval or `vtk 1 <dn>`

int write_mesh(const std::string& prefix, int n, const std::vector<double>& points,
               const std::vector<int>& cpt_ptr, const std::vector<int>& cpts,
               const std::vector<int>& mat, const std::vector<int>& nodal);

// Field with layout [cell][group][direction] (num_groups/num_directions >= 1).
int write_field(const std::string& prefix, int n, const std::string& name, const double* data,
                int num_cells, int num_groups = 1, int num_directions = 1);

}  // namespace vtk
}  // namespace pampa
