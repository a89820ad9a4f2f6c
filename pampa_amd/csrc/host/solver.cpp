#include "solver.hpp"

#include <cstdint>
#include <cstring>
#include <fstream>

#include <algorithm>
#include <cmath>

#include "input.hpp"
#include "vtk.hpp"

namespace pampa {

SolverOptions options;

// ---- Solver ---------------------------------------------------------------------------

int Solver::getField(double* v, const std::string& fname) const {
   for (const auto& f : fields)
      if (f.name == fname) {
         PAMPA_CHECK(!f.vec || !*f.vec, "field '" + fname + "' is not allocated");
         PK_CHECK(pk_vec_get_values(*f.vec, v));
         return 0;
      }
   PAMPA_CHECK(true, "unable to find field '" + fname + "'");
}

int Solver::setField(const double* v, const std::string& fname) {
   for (const auto& f : fields)
      if (f.name == fname) {
         PAMPA_CHECK(!f.vec || !*f.vec, "field '" + fname + "' is not allocated");
         PK_CHECK(pk_vec_set_values(*f.vec, v));
         return 0;
      }
   PAMPA_CHECK(true, "unable to find field '" + fname + "'");
}

int Solver::fieldDevice(const std::string& fname, void** dev, long* n) const {
   for (const auto& f : fields)
      if (f.name == fname) {
         PAMPA_CHECK(!f.vec || !*f.vec, "field '" + fname + "' is not allocated");
         PK_CHECK(pk_vec_device_ptr(*f.vec, dev));
         *n = pk_vec_size(*f.vec);
         return 0;
      }
   PAMPA_CHECK(true, "unable to find field '" + fname + "'");
}

int Solver::getOperator(const std::string&, long*, long*, int*, int*, double*) const {
   PAMPA_CHECK(true, "no operator available for solver '" + name + "'");
}

int Solver::getStats(int*, long*, double*, double*) const {
   PAMPA_CHECK(true, "no statistics available for solver '" + name + "'");
}

// ---- NeutronicSolver ----------------------------------------------------------------

int NeutronicSolver::readKeyword(const std::vector<std::string>& line, std::istream& in) {
   const std::string& k = line[0];
   if (k == "energy-groups") {
      PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
      PAMPA_CHECK(input::read(G, 1, INT_MAX, line[1]), "wrong number of energy groups");
   } else if (k == "bc") {
      PAMPA_CHECK(line.size() < 3, "wrong number of arguments for keyword '" + k + "'");
      if (bcs.empty()) bcs.resize(1 + mesh->boundaries.size());
      int b = mesh->findBoundary(line[1]);
      PAMPA_CHECK(b < 0, "wrong boundary name");
      unsigned i = 2;
      PAMPA_CHECK(input::read(bcs[b + 1], line, i, in), "wrong boundary condition");
   } else if (k == "power") {
      PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
      PAMPA_CHECK(input::read(power, 0.0, DBL_MAX, line[1]), "wrong power level");
   } else if (k == "convergence") {
      // convergence norms drive the coupling loop of multiphysics runs, which is
      // not part of this build: the line is checked and accepted
      PAMPA_CHECK(line.size() != 5, "wrong number of arguments for keyword '" + k + "'");
      PAMPA_CHECK(line[1] != "power" && line[1] != "production-rate", "wrong field");
   } else {
      PAMPA_CHECK(true, "unrecognized keyword '" + k + "'");
   }
   return 0;
}

int NeutronicSolver::read(std::istream& in) {
   while (true) {
      auto line = input::next_line(in);
      if (line.empty() || line[0] == "}") break;
      PAMPA_CHECK(readKeyword(line, in), "wrong solver definition");
   }
   return 0;
}

int NeutronicSolver::boundaryTable(double t) {
   const std::vector<BoundaryCondition>& src = bcs.empty() ? mesh->bcs : bcs;
   size_t nb = 1 + mesh->boundaries.size();
   bc_type.assign(nb, PK_BC_NONE);
   bc_value.assign(nb, 0.0);
   for (size_t b = 0; b < nb && b < src.size(); b++) {
      bc_type[b] = (int)src[b].type;
      if (!src[b].f.empty()) bc_value[b] = src[b].f[0](t);
   }
   return 0;
}

int NeutronicSolver::materialTables(bool diffusion, XSTables& t) const {
   const int nm = (int)materials.size();
   PAMPA_CHECK(nm == 0, "no materials defined");
   t = XSTables();
   t.nm = nm;
   t.G = G;
   auto add = [&](const Material& m) {
      t.st.insert(t.st.end(), m.sigma_total.begin(), m.sigma_total.end());
      t.ss.insert(t.ss.end(), m.sigma_scattering.begin(), m.sigma_scattering.end());
      t.nf.insert(t.nf.end(), m.nu_sigma_fission.begin(), m.nu_sigma_fission.end());
      t.kf.insert(t.kf.end(), m.kappa_sigma_fission.begin(), m.kappa_sigma_fission.end());
      t.ch.insert(t.ch.end(), m.chi_effective.begin(), m.chi_effective.end());
      if (diffusion) t.D.insert(t.D.end(), m.diffusion_coefficient.begin(), m.diffusion_coefficient.end());
      if (transient) {
         t.vel.insert(t.vel.end(), m.velocity.begin(), m.velocity.end());
         t.chip.insert(t.chip.end(), m.chi_prompt.begin(), m.chi_prompt.end());
         t.chid.insert(t.chid.end(), m.chi_delayed.begin(), m.chi_delayed.end());
      }
   };
   for (const auto& m : materials) {
      PAMPA_CHECK(m.check(G, diffusion, transient), "wrong nuclear data in material '" + m.name + "'");
      if (m.sets.empty()) {
         t.nsets.push_back(1);
         t.temps.push_back(0.0);
         add(m);
      } else {
         t.nsets.push_back((int)m.sets.size());
         t.temps.insert(t.temps.end(), m.temperatures.begin(), m.temperatures.end());
         for (const auto& st : m.sets) add(st);
      }
   }
   return 0;
}

int NeutronicSolver::uploadMaterials(bool diffusion) {
   XSTables t;
   PAMPA_CHECK(materialTables(diffusion, t), "wrong material data");
   for (int i = 0; i < mesh->num_cells; i++)
      PAMPA_CHECK(mesh->mat[i] < 0 || mesh->mat[i] >= t.nm, "wrong material index in the mesh");
   PK_CHECK(pk_xs_create_sets(t.nm, G, t.nsets.data(), t.temps.data(), t.st.data(), t.ss.data(), t.nf.data(),
                              t.kf.data(), t.ch.data(), diffusion ? t.D.data() : nullptr, &dxs));
   feedback = t.feedback();
   return 0;
}

int NeutronicSolver::checkMaterials(bool transient_) {
   transient = transient_;
   PAMPA_CHECK(G < 1, "missing number of energy groups");
   const bool diffusion = name == "diffusion";
   for (const auto& m : materials)
      PAMPA_CHECK(m.check(G, diffusion, transient), "wrong nuclear data in material '" + m.name + "'");
   return 0;
}

int NeutronicSolver::initialize(bool transient_) {
   PAMPA_CHECK(checkMaterials(transient_), "wrong material data");
   PAMPA_CHECK(build(), "unable to build the solver");
   return 0;
}

// q_i, P_i, S_i = beta P_i and the total power; with normalize the flux is
// first scaled to the requested power (NeutronicSolver.cxx:46-116).
int NeutronicSolver::powerAndProduction(bool normalize) {
   double tot = 0.0;
   PK_CHECK(pk_normalize_flux(dmesh, dxs, phi, normalize ? power : -1.0, keff, q, P, &tot));
   power_total = tot;
   // delayed source S = beta P (per material, NeutronicSolver.cxx:102), on the
   // device with each cell's beta (fixed once the materials are read)
   if (!beta) {
      const long C = mesh->num_cells;
      std::vector<double> hb(C);
      for (long i = 0; i < C; i++) hb[i] = materials[mesh->mat[i]].beta_total;
      PK_CHECK(pk_vec_create(C, &beta));
      PK_CHECK(pk_vec_set_values(beta, hb.data()));
   }
   PK_CHECK(pk_vec_pointwise_mult(S, beta, P));
   return 0;
}

int NeutronicSolver::output(const std::string& path, int n, bool write_mesh) {
   if (n == 0) output::print("Effective multiplication factor", keff, false, 6);
   output::print("Power", power_total, true, 3);
   if (write_mesh) PAMPA_CHECK(mesh->writeVTK(path + "/output", n), "unable to write the mesh");
   PAMPA_CHECK(writeVTK(path, n), "unable to write the solution in .vtk format");
   PAMPA_CHECK(writePETSc(n), "unable to write the solution in PETSc format");
   return 0;
}

// A vector in PETSc's binary format, as VecView on a binary viewer writes it:
// big-endian int32 class id 1211214 and length, then the values as big-endian
// IEEE doubles; the viewer's companion <file>.info is written empty.  File
// <prefix>_<n>.ptc in the working directory (petsc.cxx:491-511).
static int write_petsc_vec(const std::string& prefix, int n, pk_vec v) {
   if (!options.dump) return 0;
   const long N = pk_vec_size(v);
   std::vector<double> h(N);
   PK_CHECK(pk_vec_get_values(v, h.data()));
   const std::string fname = prefix + "_" + std::to_string(n) + ".ptc";
   std::ofstream f(fname, std::ios::binary);
   PAMPA_CHECK(!f.is_open(), "unable to open " + fname);
   auto be32 = [&](uint32_t x) {
      unsigned char b[4] = {(unsigned char)(x >> 24), (unsigned char)(x >> 16), (unsigned char)(x >> 8),
                            (unsigned char)x};
      f.write((const char*)b, 4);
   };
   be32(1211214u);
   be32((uint32_t)N);
   std::vector<unsigned char> buf(8 * (size_t)N);
   for (long i = 0; i < N; i++) {
      uint64_t x;
      std::memcpy(&x, &h[i], 8);
      for (int k = 0; k < 8; k++) buf[8 * i + k] = (unsigned char)(x >> (56 - 8 * k));
   }
   f.write((const char*)buf.data(), (std::streamsize)buf.size());
   PAMPA_CHECK(!f.good(), "unable to write " + fname);
   std::ofstream info(fname + ".info");
   PAMPA_CHECK(!info.is_open(), "unable to open " + fname + ".info");
   return 0;
}

int DiffusionSolver::writePETSc(int n) const { return write_petsc_vec("scalar_flux", n, phi); }

int SNSolver::writePETSc(int n) const { return write_petsc_vec("angular_flux", n, psi); }

int NeutronicSolver::finalize() {
   pk_vec* vs[] = {&phi, &phi0, &T, &S, &q, &P, &beta};
   for (auto* v : vs) PK_CHECK(pk_vec_destroy(v));
   PK_CHECK(pk_xs_destroy(&dxs));
   PK_CHECK(pk_mesh_destroy(&dmesh));
   return 0;
}

int NeutronicSolver::getStats(int* outer, long* inner, double* residual, double* seconds) const {
   if (outer) *outer = stat_outer;
   if (inner) *inner = stat_inner;
   if (residual) *residual = stat_residual;
   if (seconds) *seconds = stat_seconds;
   return 0;
}

// ---- DiffusionSolver ------------------------------------------------------------------

int DiffusionSolver::assembleBase(pk_mat* Rb, pk_mat* Fb) {
   PAMPA_CHECK(boundaryTable(0.0), "wrong boundary conditions");
   PK_CHECK(pk_diffusion_assemble(dmesh, dxs, (int)bc_type.size(), bc_type.data(), bc_value.data(), Rb, Fb));
   return 0;
}

int DiffusionSolver::build() {
   PAMPA_CHECK(uploadMaterials(true), "wrong material data");
   PAMPA_CHECK(mesh->toDevice(&dmesh), "unable to upload the mesh");
   long C = mesh->num_cells;
   PK_CHECK(pk_vec_create(C * G, &phi));
   PK_CHECK(pk_vec_create(C, &T));
   PK_CHECK(pk_vec_create(C, &S));
   PK_CHECK(pk_vec_create(C, &q));
   PK_CHECK(pk_vec_create(C, &P));
   PK_CHECK(pk_vec_create(C * G, &b));
   // temperature-dependent cross sections are evaluated at this field
   PK_CHECK(pk_xs_set_temperature(dxs, T));
   fields = {{"temperature", &T, true, false},   {"delayed-source", &S, true, false},
             {"scalar-flux", &phi, false, false}, {"power", &q, false, true},
             {"production-rate", &P, false, true}};
   return 0;
}

int DiffusionSolver::solve(int n, double dt, double t) {
   output::print("Run " + name + " solver...", true);
   output::indent(true);
   PAMPA_CHECK(boundaryTable(t), "wrong boundary conditions");
   const int nb = (int)bc_type.size();
   if (n > n0 + 1) {
      if (!phi0) PK_CHECK(pk_vec_create((long)mesh->num_cells * G, &phi0));
      PK_CHECK(pk_vec_copy(phi, phi0));
      n0++;
   }
   if (n == 0) {
      // DiffusionSolver::buildMatrices + EPS solve (DiffusionSolver.cxx:88-350):
      // the values of the previous steady assembly recomputed in place when the
      // boundary table is the same (cross sections and temperatures may differ),
      // else a new assembly
      if (steady_ops && R && F && bc_type == steady_bct && bc_value == steady_bcv) {
         PK_CHECK(pk_diffusion_refill(dmesh, dxs, R, F));
      } else {
         PK_CHECK(pk_mat_destroy(&R));
         PK_CHECK(pk_mat_destroy(&F));
         PK_CHECK(pk_diffusion_assemble(dmesh, dxs, nb, bc_type.data(), bc_value.data(), &R, &F));
         steady_ops = true;
         steady_bct = bc_type;
         steady_bcv = bc_value;
      }
      // the reference recreates the EPS for every steady solve
      // (NeutronicSolver.cxx:14-19); here the solver keeps its device memory and
      // captured graphs across solves of the same problem and takes the new R, F
      if (eps && pk_eig_set_operators(eps, R, F)) PK_CHECK(pk_eig_destroy(&eps));
      if (!eps) PK_CHECK(pk_eig_create(R, F, &eps));
      PK_CHECK(pk_eig_set_tolerances(eps, options.eps_tol, options.eps_max_it));
      PK_CHECK(pk_eig_set_precision(eps, options.precision));
      PK_CHECK(pk_eig_set_inner(eps, options.gmres_restart, options.inner_reduction, 2000));
      PK_CHECK(pk_eig_set_shift(eps, options.wielandt_shift));
      PK_CHECK(pk_eig_set_acceleration(eps, options.outer_acceleration));
      PK_CHECK(pk_eig_set_preconditioner(eps, options.preconditioner < 0 ? 1 : options.preconditioner));
      PK_CHECK(pk_eig_set_initial(eps, phi));
      int rc = pk_eig_solve(eps);
      if (rc) {
         report_error(__FILE__, __LINE__, __func__, pk_last_error());
         pk_eig_destroy(&eps);
         PAMPA_CHECK(true, "unable to solve the eigensystem");
      }
      double lambda = 0.0;
      PK_CHECK(pk_eig_get_eigenpair(eps, &lambda, phi));
      PK_CHECK(pk_eig_get_stats(eps, &stat_outer, &stat_inner, &stat_residual, &stat_seconds));
      keff = 1.0 / lambda;
      PAMPA_CHECK(powerAndProduction(true), "unable to normalize the scalar flux");
   } else {
      // transient rows and KSP solve (DiffusionSolver.cxx:94-148,168-175,344)
      XSTables t;
      PAMPA_CHECK(materialTables(true, t), "wrong material data");
      std::vector<double>& vel = t.vel;
      std::vector<double>& chip = t.chip;
      std::vector<double>& chid = t.chid;
      std::vector<double> beta;
      for (const auto& m : materials) beta.push_back(m.beta_total);
      PK_CHECK(pk_mat_destroy(&R));
      PK_CHECK(pk_mat_destroy(&F));
      steady_ops = false;
      PK_CHECK(pk_diffusion_assemble_transient(dmesh, dxs, nb, bc_type.data(), bc_value.data(),
                                               vel.data(), dt, chip.data(), chid.data(), beta.data(),
                                               keff, S, phi0, &R, b));
      pk_lin ksp = nullptr;
      PK_CHECK(pk_lin_create(R, &ksp));
      PK_CHECK(pk_lin_set_tolerances(ksp, options.ksp_rtol, options.ksp_atol, options.ksp_max_it));
      int rc = pk_lin_solve(ksp, b, phi);
      if (rc) {
         report_error(__FILE__, __LINE__, __func__, pk_last_error());
         pk_lin_destroy(&ksp);
         PAMPA_CHECK(true, "unable to solve the linear system");
      }
      PK_CHECK(pk_lin_get_stats(ksp, &stat_inner, &stat_residual));
      PK_CHECK(pk_lin_destroy(&ksp));
      PAMPA_CHECK(powerAndProduction(false), "unable to calculate the thermal power");
   }
   output::outdent(true);
   output::print("Done.", true);
   return 0;
}

int DiffusionSolver::writeVTK(const std::string& path, int n) {
   if (!vtk::on || n % vtk::dn != 0) return 0;
   long C = mesh->num_cells;
   std::vector<double> hphi(C * G), hq(C);
   PK_CHECK(pk_vec_get_values(phi, hphi.data()));
   PK_CHECK(pk_vec_get_values(q, hq.data()));
   PAMPA_CHECK(vtk::write_field(path + "/output", n, "flux", hphi.data(), (int)C, G), "unable to write the scalar flux");
   PAMPA_CHECK(vtk::write_field(path + "/output", n, "power", hq.data(), (int)C), "unable to write the thermal power");
   return 0;
}

int DiffusionSolver::finalize() {
   PK_CHECK(pk_eig_destroy(&eps));
   PK_CHECK(pk_mat_destroy(&R));
   PK_CHECK(pk_mat_destroy(&F));
   steady_ops = false;
   PK_CHECK(pk_vec_destroy(&b));
   return NeutronicSolver::finalize();
}

int DiffusionSolver::timeKernel(int reps, int fp32, double* seconds, double* bytes) {
   PAMPA_CHECK(!R, "the operator is not assembled (solve first)");
   long n = 0, nnz = 0;
   PK_CHECK(pk_mat_get_size(R, &n, &nnz));
   pk_vec x = nullptr, y = nullptr;
   PK_CHECK(pk_vec_create(n, &x));
   PK_CHECK(pk_vec_create(n, &y));
   PK_CHECK(pk_vec_set(x, 1.0));
   int rc = pk_mat_time_mult(R, x, y, reps, fp32, seconds);
   pk_vec_destroy(&x);
   pk_vec_destroy(&y);
   PK_CHECK(rc);
   const double s = fp32 ? 4.0 : 8.0;
   *bytes = (double)nnz * (s + 4.0) + (double)(n + 1) * 4.0 + 2.0 * n * s;
   return 0;
}

int DiffusionSolver::getOperator(const std::string& which, long* n, long* nnz, int* ptr, int* col,
                                 double* val) const {
   if (which == "b") {
      // right-hand side of the last time step: n values in val, no pattern
      PAMPA_CHECK(!b, "the time-step right-hand side is not assembled");
      *n = pk_vec_size(b);
      *nnz = 0;
      if (val) PK_CHECK(pk_vec_get_values(b, val));
      return 0;
   }
   pk_mat M = which == "R" ? R : which == "F" ? F : nullptr;
   PAMPA_CHECK(!M, "operator '" + which + "' is not assembled");
   PK_CHECK(pk_mat_get_size(M, n, nnz));
   if (ptr || col || val) PK_CHECK(pk_mat_get_csr(M, ptr, col, val));
   return 0;
}

// ---- quadrature -----------------------------------------------------------------------

int build_quadrature(int order, std::vector<double>& dirs, std::vector<double>& wts,
                     std::vector<int>& refl) {
   std::vector<std::array<double, 3>> mu;
   std::vector<double> w;
   // first-octant level-symmetric tables with the reference's 7-digit
   // constants (AngularQuadratureSet.cxx:17-155)
   switch (order) {
      case 2: {
         double a = 1.0 / std::sqrt(3.0);
         mu = {{a, a, a}};
         w = {1.0 / 8.0};
         break;
      }
      case 4: {
         double a = 0.3500212, b = 0.8688903;
         mu = {{a, a, b}, {a, b, a}, {b, a, a}};
         w.assign(3, (1.0 / 3.0) / 8.0);
         break;
      }
      case 6: {
         double a = 0.2666355, b = 0.6815076, c = 0.9261808;
         double w0 = 0.1761263, w1 = 0.1572071;
         mu = {{a, a, c}, {a, c, a}, {c, a, a}, {a, b, b}, {b, a, b}, {b, b, a}};
         w = {w0 / 8.0, w0 / 8.0, w0 / 8.0, w1 / 8.0, w1 / 8.0, w1 / 8.0};
         break;
      }
      case 8: {
         double a = 0.2182179, b = 0.5773503, c = 0.7867958, d = 0.9511897;
         double w0 = 0.1209877, w1 = 0.0907407, w2 = 0.0925926;
         mu = {{a, a, d}, {a, d, a}, {d, a, a}, {a, b, c}, {a, c, b},
               {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}, {b, b, b}};
         w = {w0 / 8, w0 / 8, w0 / 8, w1 / 8, w1 / 8, w1 / 8, w1 / 8, w1 / 8, w1 / 8, w2 / 8};
         break;
      }
      default:
         PAMPA_CHECK(true, "SN order not implemented");
   }
   const int M = order * (order + 2), q = M / 8;
   dirs.assign(3L * M, 0.0);
   wts.assign(M, 0.0);
   for (int o = 0; o < 8; o++)
      for (int m = 0; m < q; m++) {
         int l = o * q + m;
         dirs[3 * l + 0] = (o & 1) ? -mu[m][0] : mu[m][0];
         dirs[3 * l + 1] = (o & 2) ? -mu[m][1] : mu[m][1];
         dirs[3 * l + 2] = (o & 4) ? -mu[m][2] : mu[m][2];
         wts[l] = w[m];
      }
   refl.assign(3L * M, -1);
   for (int ax = 0; ax < 3; ax++)
      for (int m = 0; m < M; m++) {
         double d[3] = {dirs[3 * m], dirs[3 * m + 1], dirs[3 * m + 2]};
         double a = -2.0 * d[ax];
         d[ax] = a * 1.0 + d[ax];
         for (int m2 = 0; m2 < M; m2++) {
            double dot = dirs[3 * m2] * d[0] + dirs[3 * m2 + 1] * d[1] + dirs[3 * m2 + 2] * d[2];
            if (dot > 1.0 - 1.0e-6) {
               PAMPA_CHECK(refl[3 * m + ax] != -1, "multiple reflected directions found");
               refl[3 * m + ax] = m2;
            }
         }
         PAMPA_CHECK(refl[3 * m + ax] == -1, "reflected direction not found");
      }
   return 0;
}

// ---- SNSolver -----------------------------------------------------------------------

int SNSolver::readKeyword(const std::vector<std::string>& line, std::istream& in) {
   const std::string& k = line[0];
   if (k == "order") {
      PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
      PAMPA_CHECK(input::read(order, 1, INT_MAX, line[1]), "wrong SN order");
   } else if (k == "mixed-face-interpolation") {
      PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
      PAMPA_CHECK(input::read(delta, 0.0, 1.0, line[1]),
                  "wrong weight between upwind and linear interpolation");
   } else if (k == "least-squares-boundary-interpolation") {
      PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
      PAMPA_CHECK(input::read(ls, line[1]), "wrong switch for least-squares boundary interpolation");
   } else {
      return NeutronicSolver::readKeyword(line, in);
   }
   return 0;
}

int SNSolver::build() {
   PAMPA_CHECK(build_quadrature(order, dirs, wts, refl), "unable to build the angular quadrature set");
   M = (int)wts.size();
   PAMPA_CHECK(uploadMaterials(false), "wrong material data");
   PAMPA_CHECK(mesh->toDevice(&dmesh), "unable to upload the mesh");
   long C = mesh->num_cells;
   PK_CHECK(pk_vec_create(C * G, &phi));
   PK_CHECK(pk_vec_create(C * G * (long)M, &psi));
   PK_CHECK(pk_vec_create(C, &T));
   PK_CHECK(pk_vec_create(C, &S));
   PK_CHECK(pk_vec_create(C, &q));
   PK_CHECK(pk_vec_create(C, &P));
   PK_CHECK(pk_xs_set_temperature(dxs, T));
   fields = {{"temperature", &T, true, false},    {"delayed-source", &S, true, false},
             {"scalar-flux", &phi, false, false},  {"angular-flux", &psi, false, false},
             {"power", &q, false, true},           {"production-rate", &P, false, true}};
   // matrix-free sweeps reproduce the assembled operator exactly when the face
   // interpolation is pure upwind, the LS correction is off and every boundary
   // is vacuum (no coupling to downwind cells or to other directions)
   PAMPA_CHECK(boundaryTable(0.0), "wrong boundary conditions");
   bool all_vacuum = true;
   for (size_t f = 0; f < mesh->fnb.size(); f++)
      if (mesh->fnb[f] < 0 && bc_type[-mesh->fnb[f]] != PK_BC_VACUUM) all_vacuum = false;
   PAMPA_CHECK(feedback, "temperature-dependent nuclear data is supported by the diffusion solver only");
   sweepable = (delta == 1.0) && !ls && all_vacuum;
   if (sweepable) {
      PK_CHECK(pk_sn_create(dmesh, dxs, M, dirs.data(), wts.data(), delta, (int)bc_type.size(),
                            bc_type.data(), &sn));
      PK_CHECK(pk_sn_set_precision(sn, options.precision));
      PK_CHECK(pk_sn_set_tolerances(sn, options.sn_tol_k, options.sn_tol_phi, options.sn_max_outer));
      PK_CHECK(pk_sn_set_contraction(sn, options.contraction));
   }
   return 0;
}

// phi = 4 pi sum_m w_m psi (SNSolver.cxx:272-299)
int SNSolver::scalarFluxFromAngular() {
   long C = mesh->num_cells;
   std::vector<double> hpsi(C * G * (long)M), hphi(C * G, 0.0);
   PK_CHECK(pk_vec_get_values(psi, hpsi.data()));
   for (long l = 0, p = 0; l < C * G; l++) {
      double s = 0.0;
      for (int m = 0; m < M; m++) s += wts[m] * hpsi[p++];
      hphi[l] = s * 4.0 * M_PI;
   }
   PK_CHECK(pk_vec_set_values(phi, hphi.data()));
   return 0;
}

int SNSolver::solve(int n, double dt, double t) {
   (void)dt;
   (void)t;
   output::print("Run " + name + " solver...", true);
   output::indent(true);
   PAMPA_CHECK(n > 0, "SN transients are not supported");
   if (sweepable) {
      // the moment sum_m w_m psi comes with the solve: phi = 4 pi phiw on the
      // device (no host pass over the angular flux)
      double k = 1.0;
      int rc = pk_sn_solve(sn, &k, psi, phi);
      if (!rc) rc = pk_vec_scale(phi, 4.0 * M_PI);
      if (rc) {
         report_error(__FILE__, __LINE__, __func__, pk_last_error());
         PAMPA_CHECK(true, "unable to solve the eigensystem");
      }
      keff = k;
      long sweeps = 0;
      PK_CHECK(pk_sn_get_stats(sn, &stat_outer, &sweeps, &stat_seconds));
      stat_inner = sweeps;
   } else {
      PAMPA_CHECK(assembleReferenceForm(), "unable to build the coefficient matrices");
      pk_eig eps = nullptr;
      PK_CHECK(pk_eig_create(R, F, &eps));
      PK_CHECK(pk_eig_set_tolerances(eps, options.eps_tol, options.eps_max_it));
      PK_CHECK(pk_eig_set_precision(eps, options.precision));
      PK_CHECK(pk_eig_set_inner(eps, options.gmres_restart, options.inner_reduction, 20000));
      PK_CHECK(pk_eig_set_shift(eps, options.wielandt_shift));
      PK_CHECK(pk_eig_set_acceleration(eps, options.outer_acceleration));
      // with downwind face interpolation (delta < 1) the workgroup-block sweep
      // of the preconditioner amplifies instead of smoothing (|I - A S| ~ 40 on
      // such cores, where B^-1 alone converges): cell-block Jacobi there
      int prec = options.preconditioner < 0 ? (delta < 1.0 ? 2 : 1) : options.preconditioner;
      PK_CHECK(pk_eig_set_preconditioner(eps, prec));
      PK_CHECK(pk_eig_set_initial(eps, psi));
      int rc = pk_eig_solve(eps);
      if (rc) {
         report_error(__FILE__, __LINE__, __func__, pk_last_error());
         pk_eig_destroy(&eps);
         PAMPA_CHECK(true, "unable to solve the eigensystem");
      }
      double lambda = 0.0;
      PK_CHECK(pk_eig_get_eigenpair(eps, &lambda, psi));
      PK_CHECK(pk_eig_get_stats(eps, &stat_outer, &stat_inner, &stat_residual, &stat_seconds));
      PK_CHECK(pk_eig_destroy(&eps));
      keff = 1.0 / lambda;
   }
   // phi = 4 pi sum w psi; phi scaled to the power; psi scaled with sum w psi kapf V
   if (!sweepable) PAMPA_CHECK(scalarFluxFromAngular(), "unable to calculate the scalar flux");
   // the power of the flux before its normalisation (sum phi kapf V, on the
   // device: power < 0 leaves phi as it is)
   double p0 = 0.0;
   PK_CHECK(pk_normalize_flux(dmesh, dxs, phi, -1.0, keff, q, P, &p0));
   PAMPA_CHECK(!(p0 != 0.0), "zero power in the scalar-flux solution");
   PAMPA_CHECK(powerAndProduction(true), "unable to normalize the scalar flux");
   // angular normalisation: p0 of psi is p0 of phi / (4 pi)
   PK_CHECK(pk_vec_scale(psi, power / (p0 / (4.0 * M_PI))));
   {
      // Entries that are exactly zero in the exact eigenvector (e.g. the
      // directions leaving a vacuum boundary into a source-free, non-scattering
      // group) come out of an iterative solve as residual-level noise of either
      // sign; values within 10 x tol of the largest one are that noise floor
      // and are set to zero before the reference's negativity check
      // (SNSolver.cxx:329).
      long negatives = 0;
      PK_CHECK(pk_vec_clip_noise(psi, 10.0 * options.eps_tol, &negatives));
      PAMPA_CHECK(negatives > 0, "negative values in the angular-flux solution");
   }
   output::outdent(true);
   output::print("Done.", true);
   return 0;
}

int SNSolver::writeVTK(const std::string& path, int n) {
   if (!vtk::on || n % vtk::dn != 0) return 0;
   long C = mesh->num_cells;
   std::vector<double> hphi(C * G), hpsi(C * G * (long)M), hq(C);
   PK_CHECK(pk_vec_get_values(phi, hphi.data()));
   PK_CHECK(pk_vec_get_values(psi, hpsi.data()));
   PK_CHECK(pk_vec_get_values(q, hq.data()));
   PAMPA_CHECK(vtk::write_field(path + "/output", n, "flux", hphi.data(), (int)C, G), "unable to write the scalar flux");
   PAMPA_CHECK(vtk::write_field(path + "/output", n, "flux", hpsi.data(), (int)C, G, M), "unable to write the angular flux");
   PAMPA_CHECK(vtk::write_field(path + "/output", n, "power", hq.data(), (int)C), "unable to write the thermal power");
   return 0;
}

int SNSolver::finalize() {
   PK_CHECK(pk_sn_destroy(&sn));
   PK_CHECK(pk_mat_destroy(&R));
   PK_CHECK(pk_mat_destroy(&F));
   PK_CHECK(pk_vec_destroy(&psi));
   return NeutronicSolver::finalize();
}

int SNSolver::timeKernel(int reps, int fp32, double* seconds, double* bytes) {
   PAMPA_CHECK(!sn, "timing needs the swept (matrix-free) S_N operator");
   PK_CHECK(pk_sn_set_precision(sn, fp32 ? PK_PREC_MIXED : PK_PREC_FP64));
   int rc = pk_sn_time_sweep(sn, reps, seconds, bytes);
   PK_CHECK(pk_sn_set_precision(sn, options.precision));
   PK_CHECK(rc);
   return 0;
}

int SNSolver::timeSource(int reps, int fp32, int mode, double* seconds, double* bytes) {
   PAMPA_CHECK(!sn, "timing needs the swept (matrix-free) S_N operator");
   PK_CHECK(pk_sn_set_precision(sn, fp32 ? PK_PREC_MIXED : PK_PREC_FP64));
   PK_CHECK(pk_sn_set_contraction(sn, mode));
   int rc = pk_sn_time_source(sn, reps, seconds, bytes);
   PK_CHECK(pk_sn_set_contraction(sn, options.contraction));
   PK_CHECK(pk_sn_set_precision(sn, options.precision));
   PK_CHECK(rc);
   return 0;
}

int SNSolver::getOperator(const std::string& which, long* n, long* nnz, int* ptr, int* col,
                          double* val) const {
   pk_mat Mx = which == "R" ? R : which == "F" ? F : nullptr;
   PAMPA_CHECK(!Mx, "operator '" + which + "' is not assembled");
   PK_CHECK(pk_mat_get_size(Mx, n, nnz));
   if (ptr || col || val) PK_CHECK(pk_mat_get_csr(Mx, ptr, col, val));
   return 0;
}

}  // namespace pampa
