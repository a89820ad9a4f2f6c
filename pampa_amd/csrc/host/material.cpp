#include "material.hpp"

#include "input.hpp"

namespace pampa {

static int read_nuclear_data(Material& m, std::istream& in) {
   while (true) {
      auto line = input::next_line(in);
      if (line.empty() || line[0] == "}") break;
      const std::string& k = line[0];
      auto vec = [&](std::vector<double>& v, const char* what) -> int {
         PAMPA_CHECK(line.size() != 1, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(m.G < 1, "missing number of energy groups");
         PAMPA_CHECK(input::read(v, m.G, 0.0, DBL_MAX, in), std::string("wrong ") + what);
         return 0;
      };
      if (k == "energy-groups") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(input::read(m.G, 1, INT_MAX, line[1]), "wrong number of energy groups");
      } else if (k == "sigma-total") {
         PAMPA_CHECK(vec(m.sigma_total, "total cross sections"), "wrong nuclear data");
      } else if (k == "nu-sigma-fission") {
         PAMPA_CHECK(vec(m.nu_sigma_fission, "nu-fission cross sections"), "wrong nuclear data");
      } else if (k == "kappa-sigma-fission") {
         PAMPA_CHECK(vec(m.kappa_sigma_fission, "kappa-fission cross sections"), "wrong nuclear data");
      } else if (k == "sigma-transport") {
         PAMPA_CHECK(vec(m.sigma_transport, "transport cross sections"), "wrong nuclear data");
      } else if (k == "sigma-scattering") {
         PAMPA_CHECK(line.size() != 1, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(input::read_rows(m.sigma_scattering, m.G, m.G, 0.0, DBL_MAX, in),
                     "wrong scattering cross sections");
      } else if (k == "diffusion-coefficient") {
         PAMPA_CHECK(vec(m.diffusion_coefficient, "diffusion coefficients"), "wrong nuclear data");
      } else if (k == "fission-spectrum" || k == "fission-spectrum-prompt") {
         PAMPA_CHECK(vec(m.chi_prompt, "prompt fission spectrum"), "wrong nuclear data");
      } else if (k == "fission-spectrum-delayed") {
         PAMPA_CHECK(vec(m.chi_delayed, "delayed fission spectrum"), "wrong nuclear data");
      } else if (k == "neutron-velocity") {
         PAMPA_CHECK(vec(m.velocity, "neutron velocity"), "wrong nuclear data");
      } else {
         PAMPA_CHECK(true, "unrecognized keyword '" + k + "'");
      }
   }
   m.has_nuclear_data = true;
   return 0;
}

static int read_precursor_data(Material& m, std::istream& in) {
   while (true) {
      auto line = input::next_line(in);
      if (line.empty() || line[0] == "}") break;
      const std::string& k = line[0];
      if (k == "precursor-groups") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(input::read(m.num_precursor_groups, 1, INT_MAX, line[1]),
                     "wrong number of delayed-neutron precursor groups");
      } else if (k == "lambda") {
         PAMPA_CHECK(input::read(m.lambda, m.num_precursor_groups, 0.0, DBL_MAX, in),
                     "wrong precursor decay constants");
      } else if (k == "beta") {
         PAMPA_CHECK(input::read(m.beta, m.num_precursor_groups, 0.0, DBL_MAX, in),
                     "wrong precursor fractions");
         for (double b : m.beta) m.beta_total += b;
      } else {
         PAMPA_CHECK(true, "unrecognized keyword '" + k + "'");
      }
   }
   return 0;
}

int Material::read(std::istream& in) {
   while (true) {
      auto line = input::next_line(in);
      if (line.empty() || line[0] == "}") break;
      const std::string& k = line[0];
      if (k == "nuclear-data") {
         PAMPA_CHECK(line.size() != 2 || line[1] != "{", "missing opening '{' for constant nuclear data");
         PAMPA_CHECK(read_nuclear_data(*this, in), "unable to read the constant nuclear data");
      } else if (k == "nuclear-data-set") {
         PAMPA_CHECK(line.size() != 2 || line[1] != "{", "missing opening '{' for nuclear data");
         while (true) {
            auto l2 = input::next_line(in);
            if (l2.empty() || l2[0] == "}") break;
            if (l2[0] == "temperature") {
               PAMPA_CHECK(l2.size() != 2, "wrong number of arguments for keyword 'temperature'");
               int n = 0;
               PAMPA_CHECK(input::read(n, 0, INT_MAX, l2[1]), "wrong number of temperatures");
               PAMPA_CHECK(input::read(temperatures, n, 0.0, DBL_MAX, in), "wrong temperature data");
            } else if (l2[0] == "nuclear-data") {
               PAMPA_CHECK(l2.size() != 2 || l2[1] != "{", "missing opening '{' for nuclear data");
               Material set;
               PAMPA_CHECK(read_nuclear_data(set, in), "unable to read the nuclear data");
               sets.push_back(set);
            } else {
               PAMPA_CHECK(true, "unrecognized keyword '" + l2[0] + "'");
            }
         }
         PAMPA_CHECK(sets.empty() || sets.size() != temperatures.size(),
                     "one nuclear-data block per temperature is needed");
         for (size_t k = 1; k < temperatures.size(); k++)
            PAMPA_CHECK(!(temperatures[k] > temperatures[k - 1]), "temperatures must increase");
         has_nuclear_data = true;
      } else if (k == "precursor-data") {
         PAMPA_CHECK(line.size() != 2 || line[1] != "{", "missing opening '{' for precursor data");
         PAMPA_CHECK(read_precursor_data(*this, in), "unable to read the precursor data");
      } else if (k == "thermal-properties") {
         // thermal properties belong to the heat-conduction solver; accepted and ignored
         PAMPA_CHECK(line.size() < 2, "wrong number of arguments for keyword '" + k + "'");
      } else if (k == "fuel") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(input::read(fuel, line[1]), "wrong switch for fuel materials");
      } else if (k == "bc") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(input::read(bc, line[1]), "wrong switch for boundary-condition materials");
      } else if (k == "split") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(input::read(split, line[1]), "wrong switch for split materials");
      } else {
         PAMPA_CHECK(true, "unrecognized keyword '" + k + "'");
      }
   }
   if (!sets.empty()) {
      // every set derived with this material's delayed fraction; the material's
      // own fields take the first set
      for (auto& st : sets) {
         st.beta_total = beta_total;
         PAMPA_CHECK(st.derive(), "wrong nuclear data");
      }
      const std::vector<double> temps = temperatures;
      std::vector<Material> all = sets;
      const std::string nm = name;
      const bool f = fuel, b = bc, sp = split;
      const double bt = beta_total;
      const int np = num_precursor_groups;
      const std::vector<double> lam = lambda, bet = beta;
      *this = all[0];
      name = nm;
      fuel = f;
      bc = b;
      split = sp;
      beta_total = bt;
      num_precursor_groups = np;
      lambda = lam;
      beta = bet;
      temperatures = temps;
      sets = all;
   } else if (has_nuclear_data) {
      PAMPA_CHECK(derive(), "wrong nuclear data");
   }
   return 0;
}

int Material::read(const std::string& filename, const std::string& base_dir) {
   std::ifstream in;
   PAMPA_CHECK(!input::open(in, filename, base_dir), "unable to open " + filename);
   PAMPA_CHECK(read(in), "unable to read the material from " + filename);
   return 0;
}

int Material::derive() {
   if (!nu_sigma_fission.empty() || !chi_prompt.empty()) {
      PAMPA_CHECK(nu_sigma_fission.empty(), "missing nu-fission cross sections");
      PAMPA_CHECK(chi_prompt.empty(), "missing fission spectrum");
   }
   if (nu_sigma_fission.empty()) nu_sigma_fission.assign(G, 0.0);
   if (kappa_sigma_fission.empty()) {
      // default fission energy over neutron yield (ConstantNuclearData.hxx:27)
      const double f = 3.2e-11 / 2.4355;
      kappa_sigma_fission = nu_sigma_fission;
      for (auto& v : kappa_sigma_fission) v *= f;
   }
   if (chi_prompt.empty()) chi_prompt.assign(G, 0.0);
   if (chi_delayed.empty()) chi_delayed = chi_prompt;
   if (beta_total > 0.0) {
      chi_effective.assign(G, 0.0);
      for (int g = 0; g < G; g++)
         chi_effective[g] = (1.0 - beta_total) * chi_prompt[g] + beta_total * chi_delayed[g];
   } else {
      chi_effective = chi_prompt;
   }
   if (!sigma_transport.empty()) {
      PAMPA_CHECK(!diffusion_coefficient.empty(),
                  "transport cross sections can only be defined if diffusion coefficients are not");
      diffusion_coefficient.assign(G, 0.0);
      for (int g = 0; g < G; g++) diffusion_coefficient[g] = 1.0 / (3.0 * sigma_transport[g]);
   }
   return 0;
}

int Material::check(int num_groups, bool diffusion, bool transient) const {
   PAMPA_CHECK(!has_nuclear_data, "missing nuclear data");
   for (const auto& st : sets) PAMPA_CHECK(st.check(num_groups, diffusion, transient), "wrong nuclear data");
   PAMPA_CHECK(G != num_groups, "wrong number of energy groups");
   PAMPA_CHECK(sigma_total.empty(), "missing total cross sections");
   PAMPA_CHECK(sigma_scattering.empty(), "missing scattering cross sections");
   PAMPA_CHECK(chi_effective.empty(), "missing effective fission spectrum");
   if (diffusion) PAMPA_CHECK(diffusion_coefficient.empty(), "missing diffusion coefficients");
   if (transient) {
      PAMPA_CHECK(chi_prompt.empty(), "missing prompt fission spectrum");
      PAMPA_CHECK(chi_delayed.empty(), "missing delayed fission spectrum");
      PAMPA_CHECK(velocity.empty(), "missing neutron velocities");
   }
   return 0;
}

}  // namespace pampa
