#include "vtk.hpp"

namespace pampa {
namespace vtk {

bool on = false;
int dn = 1;

static std::string file_name(const std::string& prefix, int n) {
   return n < 0 ? prefix + ".vtk" : prefix + "_" + std::to_string(n / dn) + ".vtk";
}

int write_mesh(const std::string& prefix, int n, const std::vector<double>& points,
               const std::vector<int>& cpt_ptr, const std::vector<int>& cpts,
               const std::vector<int>& mat, const std::vector<int>& nodal) {
   if (!on || (n % dn != 0)) return 0;
   std::string fname = file_name(prefix, n);
   std::ofstream f(fname);
   PAMPA_CHECK(!f.is_open(), "unable to open " + fname);
   f << std::scientific << std::setprecision(6);
   f << "# vtk DataFile Version 3.0\nFVM mesh\nASCII\nDATASET UNSTRUCTURED_GRID\n\n";
   size_t np = points.size() / 3;
   f << "POINTS " << np << " double\n";
   for (size_t i = 0; i < np; i++)
      f << points[3 * i] << " " << points[3 * i + 1] << " " << points[3 * i + 2] << "\n";
   f << "\n";
   size_t nc = cpt_ptr.size() - 1;
   f << "CELLS " << nc << " " << nc + cpts.size() << "\n";
   for (size_t i = 0; i < nc; i++) {
      f << cpt_ptr[i + 1] - cpt_ptr[i];
      for (int q = cpt_ptr[i]; q < cpt_ptr[i + 1]; q++) f << " " << cpts[q];
      f << "\n";
   }
   f << "\nCELL_TYPES " << nc << "\n";
   for (size_t i = 0; i < nc; i++) {
      int k = cpt_ptr[i + 1] - cpt_ptr[i];
      int t = k == 2 ? 3 : k == 3 ? 5 : k == 4 ? 9 : k == 6 ? 13 : k == 8 ? 12 : k == 12 ? 16 : -1;
      PAMPA_CHECK(t < 0, "wrong cell type");
      f << t << "\n";
   }
   f << "\nCELL_DATA " << nc << "\n\n";
   f << "SCALARS materials double 1\nLOOKUP_TABLE default\n";
   for (size_t i = 0; i < nc; i++) f << mat[i] + 1 << "\n";
   f << "\n";
   if (!nodal.empty()) {
      f << "SCALARS nodal_indices double 1\nLOOKUP_TABLE default\n";
      for (size_t i = 0; i < nc; i++) f << nodal[i] << "\n";
      f << "\n";
   }
   return 0;
}

int write_field(const std::string& prefix, int n, const std::string& name, const double* data,
                int num_cells, int num_groups, int num_directions) {
   if (!on || (n % dn != 0)) return 0;
   std::string fname = prefix + "_" + std::to_string(n / dn) + ".vtk";
   std::ofstream f(fname, std::ios_base::app);
   PAMPA_CHECK(!f.is_open(), "unable to open " + fname);
   f << std::scientific << std::setprecision(6);
   for (int g = 0; g < num_groups; g++)
      for (int m = 0; m < num_directions; m++) {
         f << "SCALARS " << name;
         if (num_groups > 1) f << "_" << (g + 1);
         if (num_directions > 1) f << "_" << (m + 1);
         f << " double 1\nLOOKUP_TABLE default\n";
         long stride = (long)num_groups * num_directions;
         for (long i = 0, iv = (long)g * num_directions + m; i < num_cells; i++, iv += stride)
            f << data[iv] << "\n";
         f << "\n";
      }
   return 0;
}

}  // namespace vtk
}  // namespace pampa
