// Solver API of the drop-in (Solver.hxx:69-96, PhysicsSolver.hxx:27-40,
// NeutronicSolver.hxx:49-52): read / initialize / solve / output / finalize,
// plus named input/output fields.  The numerics go through the pk_* C-ABI
// (device-resident operators, vectors and solvers).
#pragma once

#include "material.hpp"
#include "mesh.hpp"

namespace pampa {

// Options that the reference takes from the PETSc options database
// (petsc.cxx:80-88,99,165,200) and that the pk_* layer honours.
struct SolverOptions {
   double eps_tol = 1e-8;     // eigen residual |R x - F x/k| / |F x/k|
   int eps_max_it = 20000;
   int precision = PK_PREC_FP64;
   int gmres_restart = 30;  // PETSc's default GMRES restart
   double inner_reduction = 0.1;
   double wielandt_shift = 0.1;
   int outer_acceleration = 3;  // Anderson depth of the outer iterates (0: plain updates)
   // 0 cell-block Jacobi + workgroup-block sweep (bj), 1 with one Richardson
   // correction (bj2), 2 cell-block Jacobi alone (cell); -1: the solver's choice
   int preconditioner = -1;
   // S_N scattering + fission source: 0 matrix cores at 16 groups, else the
   // vector kernel (auto); 1 vector kernel; 2 matrix cores
   int contraction = 0;
   double ksp_rtol = 1e-5, ksp_atol = 1e-50;  // PETSc KSP defaults
   int ksp_max_it = 10000;
   double sn_tol_k = 1e-9, sn_tol_phi = 1e-8;
   int sn_max_outer = 100000;
   bool verbose_solver = false;
   bool dump = false;  // -petsc_dump / `petsc dump 1`: solution vectors in PETSc binary format
};
extern SolverOptions options;

// Per-material tables in the layout of pk_xs_create_sets: material m owns
// nsets[m] sets at temperatures temps (one set for constant data); tables
// [sets][G] and sigs [sets][G][G]; the transient tables likewise.
struct XSTables {
   int nm = 0, G = 0;
   std::vector<int> nsets;
   std::vector<double> temps;
   std::vector<double> st, ss, nf, kf, ch, D;
   std::vector<double> vel, chip, chid;
   bool feedback() const { return (int)temps.size() > nm; }
};

struct Field {
   std::string name;
   pk_vec* vec = nullptr;
   bool input = false, output = false;
};

class Solver {
  public:
   Solver(const std::string& name) : name(name) {}
   virtual ~Solver() {}
   const std::string name;
   virtual int read(std::istream& in) = 0;
   // input checks that need no device (PhysicsSolver::initialize -> checkMaterials)
   virtual int checkMaterials(bool transient) = 0;
   virtual int initialize(bool transient) = 0;
   virtual int solve(int n, double dt, double t) = 0;
   virtual int output(const std::string& path, int n, bool write_mesh = true) = 0;
   virtual int finalize() = 0;
   int getField(double* v, const std::string& name) const;
   int setField(const double* v, const std::string& name);
   // device address and length of a field (for coupled codes on the same GPU)
   int fieldDevice(const std::string& name, void** dev, long* n) const;
   // solution summary used by the C-ABI extensions and tests
   virtual double getKeff() const { return -1.0; }
   virtual double getPower() const { return 0.0; }
   virtual int getOperator(const std::string& which, long* n, long* nnz, int* ptr, int* col,
                           double* val) const;
   virtual int getStats(int* outer, long* inner, double* residual, double* seconds) const;
   virtual void getSizes(int* cells, int* groups, int* directions) const = 0;
   virtual int timeKernel(int reps, int fp32, double* seconds, double* bytes) = 0;
   // the S_N source contraction alone (mode as SolverOptions::contraction)
   virtual int timeSource(int reps, int fp32, int mode, double* seconds, double* bytes) {
      PAMPA_CHECK(true, "the solver has no source contraction kernel");
      return 0;
   }
   // forget any state a steady solve starts from (other than the fields)
   virtual int coldStart() { return 0; }

  protected:
   std::vector<Field> fields;
};

class NeutronicSolver : public Solver {
  public:
   NeutronicSolver(const std::string& name, Mesh* mesh, std::vector<Material>& mats)
       : Solver(name), mesh(mesh), materials(mats) {}
   int read(std::istream& in) override;
   int checkMaterials(bool transient) override;
   int initialize(bool transient) override;
   int output(const std::string& path, int n, bool write_mesh) override;
   int finalize() override;
   double getKeff() const override { return keff; }
   double getPower() const override { return power_total; }
   int getStats(int* outer, long* inner, double* residual, double* seconds) const override;
   void getSizes(int* cells, int* groups, int* directions) const override {
      if (cells) *cells = mesh->num_cells;
      if (groups) *groups = G;
      if (directions) *directions = numDirections();
   }
   virtual int numDirections() const { return 1; }
   int materialTables(bool diffusion, XSTables& t) const;
   pk_mesh deviceMesh() const { return dmesh; }
   long numCells() const { return mesh->num_cells; }
   int groups() const { return G; }

  protected:
   virtual int readKeyword(const std::vector<std::string>& line, std::istream& in);
   virtual int build() = 0;
   virtual int writeVTK(const std::string& path, int n) = 0;
   // the solution vector in PETSc's binary format (petsc.cxx:491-511), with -petsc_dump
   virtual int writePETSc(int n) const = 0;
   int uploadMaterials(bool diffusion);
   bool feedback = false;  // some material has temperature-dependent data
   int powerAndProduction(bool normalize);  // NeutronicSolver.cxx:46-116

   Mesh* mesh;
   std::vector<Material>& materials;
   int G = -1;
   std::vector<BoundaryCondition> bcs;  // solver-level overrides (indexed by boundary + 1)
   double power = 1.0, power_total = 0.0, keff = -1.0;
   bool transient = false;
   int n0 = -1;
   pk_mesh dmesh = nullptr;
   pk_xs dxs = nullptr;
   pk_vec phi = nullptr, phi0 = nullptr, T = nullptr, S = nullptr, q = nullptr, P = nullptr;
   pk_vec beta = nullptr;  // delayed neutron fraction of each cell's material (built on first use)
   std::vector<int> bc_type;
   std::vector<double> bc_value;
   int stat_outer = 0;
   long stat_inner = 0;
   double stat_residual = 0.0, stat_seconds = 0.0;
   int boundaryTable(double t);
};

class DiffusionSolver : public NeutronicSolver {
  public:
   DiffusionSolver(Mesh* mesh, std::vector<Material>& mats) : NeutronicSolver("diffusion", mesh, mats) {}
   int solve(int n, double dt, double t) override;
   int finalize() override;
   int getOperator(const std::string& which, long* n, long* nnz, int* ptr, int* col,
                   double* val) const override;
   int timeKernel(int reps, int fp32, double* seconds, double* bytes) override;
   // R and F of the input as given (steady state, t = 0), owned by the caller
   int assembleBase(pk_mat* Rb, pk_mat* Fb);

  protected:
   int build() override;
   int writeVTK(const std::string& path, int n) override;
   int writePETSc(int n) const override;

  private:
   pk_mat R = nullptr, F = nullptr;
   pk_vec b = nullptr;
   pk_eig eps = nullptr;  // kept across steady solves of this problem
   // R and F hold a steady assembly with this boundary table: the next steady
   // solve recomputes their values in place (same pattern and layout)
   bool steady_ops = false;
   std::vector<int> steady_bct;
   std::vector<double> steady_bcv;
};

class SNSolver : public NeutronicSolver {
  public:
   SNSolver(Mesh* mesh, std::vector<Material>& mats) : NeutronicSolver("sn", mesh, mats) {}
   int solve(int n, double dt, double t) override;
   int finalize() override;
   int getOperator(const std::string& which, long* n, long* nnz, int* ptr, int* col,
                   double* val) const override;
   int numDirections() const override { return M; }
   int timeKernel(int reps, int fp32, double* seconds, double* bytes) override;
   int timeSource(int reps, int fp32, int mode, double* seconds, double* bytes) override;
   // the sweep handle (null when the input is not sweepable)
   pk_sn sweepHandle() const { return sweepable ? sn : nullptr; }
   int coldStart() override { return sn ? pk_sn_reset(sn) : 0; }

  protected:
   int readKeyword(const std::vector<std::string>& line, std::istream& in) override;
   int build() override;
   int writeVTK(const std::string& path, int n) override;
   int writePETSc(int n) const override;

  private:
   int order = -1;
   double delta = 0.1;  // mixed-face-interpolation (SNSolver.hxx:16)
   bool ls = false;     // least-squares-boundary-interpolation
   int M = 0;
   std::vector<double> dirs, wts;
   std::vector<int> refl;
   bool sweepable = false;
   pk_sn sn = nullptr;
   pk_mat R = nullptr, F = nullptr;  // reference-form operator (non-sweepable inputs)
   pk_vec psi = nullptr;
   int assembleReferenceForm();
   int scalarFluxFromAngular();
};

// Level-symmetric S_N sets (AngularQuadratureSet.cxx:4-209): directions [M][3],
// weights [M] (sum 1), reflected direction per axis [M][3].
int build_quadrature(int order, std::vector<double>& dirs, std::vector<double>& wts,
                     std::vector<int>& refl);

}  // namespace pampa
