// Materials with constant nuclear data (Material.*, ConstantNuclearData.*,
// PrecursorData.*): the per-group tables the neutronic solvers read, with the
// reference's derived quantities (ConstantNuclearData.cxx:129-177).
#pragma once

#include "common.hpp"

namespace pampa {

struct Material {
   std::string name;
   bool has_nuclear_data = false;
   int G = -1;
   std::vector<double> sigma_total, nu_sigma_fission, kappa_sigma_fission, sigma_transport;
   std::vector<double> sigma_scattering;  // [from][to]
   std::vector<double> diffusion_coefficient, chi_prompt, chi_delayed, chi_effective, velocity;
   int num_precursor_groups = 0;
   std::vector<double> lambda, beta;
   double beta_total = 0.0;
   bool fuel = false, bc = false, split = false;
   // temperature-dependent data (nuclear-data-set, FeedbackNuclearData.cxx:4-53):
   // one constant data set per tabulated temperature; the material's own
   // fields hold the first set
   std::vector<double> temperatures;
   std::vector<Material> sets;

   int read(std::istream& in);
   int read(const std::string& filename, const std::string& base_dir);
   // derived data after reading (ConstantNuclearData::check(beta))
   int derive();
   // data needed by a solver with G groups (ConstantNuclearData::check(G, diffusion, transient))
   int check(int num_groups, bool diffusion, bool transient) const;
};

}  // namespace pampa
