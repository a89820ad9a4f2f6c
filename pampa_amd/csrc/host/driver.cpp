#include "driver.hpp"

#include <chrono>
#include <cstdlib>

#include "input.hpp"
#include "vtk.hpp"

namespace pampa {

namespace output {
bool verbose = false, silent = false;
int padding = 0;

void print(const std::string& message, bool info) {
   if ((!info || verbose) && !silent) std::cout << std::string(3 * padding, ' ') << message << std::endl;
}

void print(const std::string& name, double x, bool scientific, int precision, bool info) {
   if ((!info || verbose) && !silent) {
      std::ostringstream s;
      if (scientific)
         s << std::scientific;
      else
         s << std::fixed;
      s << std::setprecision(precision) << x;
      std::cout << std::string(3 * padding, ' ') << name << ": " << s.str() << "." << std::endl;
   }
}

void indent(bool info) {
   if ((!info || verbose) && !silent) padding++;
}

void outdent(bool info) {
   if ((!info || verbose) && !silent) padding--;
}
}  // namespace output

// Options of the PETSc options database that map onto the pk_* solvers
// (`-name value` on the command line or `petsc name value` in the input).
int Driver::setOption(const std::string& name, const std::string& value) {
   std::string n = name;
   if (!n.empty() && n[0] == '-') n = n.substr(1);
   if (n == "eps_tol") {
      PAMPA_CHECK(input::read(options.eps_tol, 0.0, 1.0, value), "wrong eps_tol");
   } else if (n == "eps_max_it") {
      PAMPA_CHECK(input::read(options.eps_max_it, 1, INT_MAX, value), "wrong eps_max_it");
   } else if (n == "ksp_rtol") {
      PAMPA_CHECK(input::read(options.ksp_rtol, 0.0, 1.0, value), "wrong ksp_rtol");
   } else if (n == "ksp_atol") {
      PAMPA_CHECK(input::read(options.ksp_atol, 0.0, DBL_MAX, value), "wrong ksp_atol");
   } else if (n == "ksp_max_it") {
      PAMPA_CHECK(input::read(options.ksp_max_it, 1, INT_MAX, value), "wrong ksp_max_it");
   } else if (n == "pk_preconditioner") {
      PAMPA_CHECK(value != "bj" && value != "bj2" && value != "cell",
                  "wrong pk_preconditioner (bj, bj2 or cell)");
      options.preconditioner = value == "bj2" ? 1 : (value == "cell" ? 2 : 0);
   } else if (n == "pk_contraction") {
      PAMPA_CHECK(value != "auto" && value != "vector" && value != "mfma",
                  "wrong pk_contraction (auto, vector or mfma)");
      options.contraction = value == "vector" ? 1 : (value == "mfma" ? 2 : 0);
   } else if (n == "ksp_gmres_restart") {
      PAMPA_CHECK(input::read(options.gmres_restart, 1, 64, value), "wrong ksp_gmres_restart");
   } else if (n == "pk_precision") {
      PAMPA_CHECK(value != "fp64" && value != "mixed", "wrong pk_precision (fp64 or mixed)");
      options.precision = value == "mixed" ? PK_PREC_MIXED : PK_PREC_FP64;
   } else if (n == "pk_shift") {
      PAMPA_CHECK(input::read(options.wielandt_shift, 0.0, DBL_MAX, value), "wrong pk_shift");
   } else if (n == "pk_outer_acceleration") {
      PAMPA_CHECK(input::read(options.outer_acceleration, 0, 3, value), "wrong pk_outer_acceleration");
   } else if (n == "pk_inner_reduction") {
      PAMPA_CHECK(input::read(options.inner_reduction, 1e-12, 0.99, value), "wrong pk_inner_reduction");
   } else if (n == "pk_sn_tol") {
      PAMPA_CHECK(input::read(options.sn_tol_phi, 0.0, 1.0, value), "wrong pk_sn_tol");
      options.sn_tol_k = options.sn_tol_phi * 0.1;
   } else {
      // other PETSc/SLEPc options have no counterpart in the pk_* solvers
      output::print("Warning: option '" + name + "' is not used.", true);
   }
   return 0;
}

int Driver::parse(const std::string& filename, std::vector<double>& dt) {
   std::ifstream in(filename);
   PAMPA_CHECK(!in.is_open(), "unable to open " + filename);
   const std::string base = input::dir_of(filename);
   bool mesh_ready = false;
   while (true) {
      auto line = input::next_line(in);
      if (line.empty()) break;
      const std::string& k = line[0];
      if (k == "mesh") {
         PAMPA_CHECK(line.size() != 3, "wrong number of arguments for keyword '" + k + "'");
         if (line[1] == "cartesian")
            mesh.reset(new CartesianMesh());
         else if (line[1] == "unstructured")
            mesh.reset(new UnstructuredExtrudedMesh());
         else if (line[1] == "partitioned")
            mesh.reset(new PartitionedMesh());
         else
            PAMPA_CHECK(true, "wrong mesh type");
         PAMPA_CHECK(mesh->read(line[2], base), "unable to read the mesh from " + line[2]);
         PAMPA_CHECK(mesh->build(), "unable to build the mesh");
      } else if (k == "mesh-nodal") {
         PAMPA_CHECK(true, "nodal meshes belong to the heat-conduction solver, which is not part of this build");
      } else if (k == "material") {
         PAMPA_CHECK(line.size() != 3, "wrong number of arguments for keyword '" + k + "'");
         Material m;
         m.name = line[1];
         if (line[2] == "{")
            PAMPA_CHECK(m.read(in), "unable to read the material from " + filename);
         else
            PAMPA_CHECK(m.read(line[2], base), "unable to read the material from " + line[2]);
         // a boundary-condition material names a boundary of the mesh (Parser.cxx:93)
         if (m.bc) {
            PAMPA_CHECK(!mesh, "missing mesh definition before a boundary-condition material");
            mesh->addBoundary(m.name);
         }
         materials.push_back(m);
      } else if (k == "solver") {
         PAMPA_CHECK(line.size() < 3, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(!mesh, "missing mesh definition before the solver");
         if (!mesh_ready) {
            // split materials, then boundary-condition materials out of the mesh
            // (Parser.cxx:107-130)
            bool split = false, bc = false;
            for (const auto& m : materials) {
               split |= m.split;
               bc |= m.bc;
            }
            if (split) PAMPA_CHECK(split_materials(*mesh, materials), "unable to split materials in the mesh");
            if (bc)
               PAMPA_CHECK(remove_bc_materials(*mesh, materials),
                           "unable to remove boundary-condition materials from the mesh");
         }
         mesh_ready = true;
         std::unique_ptr<Solver> s;
         if (line[1] == "diffusion")
            s.reset(new DiffusionSolver(mesh.get(), materials));
         else if (line[1] == "sn")
            s.reset(new SNSolver(mesh.get(), materials));
         else
            PAMPA_CHECK(true, "solver '" + line[1] + "' is not part of this build");
         if (line[2] == "{") {
            PAMPA_CHECK(s->read(in), "unable to read the solver from " + filename);
         } else {
            std::ifstream sf;
            PAMPA_CHECK(!input::open(sf, line[2], base), "unable to open " + line[2]);
            PAMPA_CHECK(s->read(sf), "unable to read the solver from " + line[2]);
         }
         solvers.push_back(std::move(s));
      } else if (k == "dt") {
         PAMPA_CHECK(line.size() != 2, "wrong number of arguments for keyword '" + k + "'");
         int nt = 0;
         PAMPA_CHECK(input::read(nt, -INT_MAX, INT_MAX, line[1]), "wrong number of time steps");
         if (nt > 0) {
            PAMPA_CHECK(input::read(dt, nt, 0.0, DBL_MAX, in), "wrong dt data");
         } else {
            PAMPA_CHECK(input::read(dt, 1, 0.0, DBL_MAX, in), "wrong dt data");
            dt.assign(-nt, dt[0]);
         }
      } else if (k == "vtk") {
         PAMPA_CHECK(line.size() < 2 || line.size() > 3, "wrong number of arguments for keyword '" + k + "'");
         PAMPA_CHECK(input::read(vtk::on, line[1]), "wrong switch for .vtk output");
         if (line.size() == 3) PAMPA_CHECK(input::read(vtk::dn, 1, INT_MAX, line[2]), "wrong .vtk output interval");
      } else if (k == "petsc") {
         PAMPA_CHECK(line.size() != 3, "wrong number of arguments for keyword '" + k + "'");
         if (line[1] == "verbose") {
            PAMPA_CHECK(input::read(options.verbose_solver, line[2]), "wrong switch for verbose PETSc output");
         } else if (line[1] == "dump") {
            PAMPA_CHECK(input::read(options.dump, line[2]), "wrong switch to write the solution in PETSc format");
         } else {
            PAMPA_CHECK(setOption(line[1], line[2]), "wrong PETSc/SLEPc option");
         }
      } else if (k == "include") {
         PAMPA_CHECK(line.size() < 2, "wrong number of arguments for keyword '" + k + "'");
         const std::string& inc = line[1];
         std::ifstream here(inc);
         std::string path = here.is_open() || inc[0] == '/' ? inc : base + "/" + inc;
         PAMPA_CHECK(parse(path, dt), "unable to parse " + inc);
      } else {
         PAMPA_CHECK(true, "unrecognized keyword '" + k + "'");
      }
   }
   (void)mesh_ready;
   return 0;
}

int Driver::initialize(int argc, char* argv[], std::vector<double>& dt) {
   PAMPA_CHECK(argc < 2, "missing input file");
   std::string filename(argv[1]);
   // process-global switches start from their defaults for every calculation
   options = SolverOptions();
   output::verbose = output::silent = false;
   output::padding = 0;
   vtk::on = false;
   vtk::dn = 1;
   // own flags and PETSc-style options (output.cxx:21-24, vtk.cxx:18-21, petsc.cxx:29-32)
   for (int a = 2; a < argc; a++) {
      std::string s(argv[a]);
      if (s == "-verbose")
         output::verbose = true;
      else if (s == "-silent")
         output::silent = true;
      else if (s == "-vtk")
         vtk::on = true;
      else if (s == "-vtk_interval" && a + 1 < argc)
         PAMPA_CHECK(input::read(vtk::dn, 1, INT_MAX, argv[++a]), "wrong .vtk output interval");
      else if (s == "-petsc_verbose")
         options.verbose_solver = true;
      else if (s == "-petsc_dump")
         options.dump = true;
      else if (s.size() > 1 && s[0] == '-' && a + 1 < argc && argv[a + 1][0] != '-')
         PAMPA_CHECK(setOption(s, argv[++a]), "wrong command-line option");
      else if (s.size() > 1 && s[0] == '-')
         PAMPA_CHECK(setOption(s, "1"), "wrong command-line option");
   }
   output::print("\nInitialize...");
   output::indent();
   output::print("Parse the input file...", true);
   PAMPA_CHECK(parse(filename, dt), "unable to parse " + filename);
   output::print("Done.", true);
   output::print("Initialize the solver...", true);
   PAMPA_CHECK(solvers.empty(), "no solvers defined");
   PAMPA_CHECK(solvers.size() > 1, "coupled multiphysics runs are not part of this build");
   solver = solvers[0].get();
   PAMPA_CHECK(solver->checkMaterials(!dt.empty()), "wrong material data");
   // the device is touched only once the input is known to be valid
   PK_CHECK(pk_initialize(0));
   PAMPA_CHECK(solver->initialize(!dt.empty()), "unable to initialize the solver");
   output::print("Done.", true);
   output::outdent();
   output::print("Done.");
   initialized = true;
   return 0;
}

int Driver::solve(int n, double dt, double t) {
   PAMPA_CHECK(!initialized, "the calculation is not initialized");
   output::print("\n--------------------------------");
   if (n == 0)
      output::print("\nSolve steady state...\n");
   else
      output::print("\nSolve time step " + std::to_string(n) + "...\n");
   auto t1 = std::chrono::steady_clock::now();
   PAMPA_CHECK(solver->solve(n, dt, t), "unable to get the solution");
   auto t2 = std::chrono::steady_clock::now();
   output::print("", !options.verbose_solver);
   output::print("Solution time", std::chrono::duration<double>(t2 - t1).count(), true, 3, true);
   if (options.verbose_solver) {
      int outer = 0;
      long inner = 0;
      double res = 0.0, sec = 0.0;
      if (solver->getStats(&outer, &inner, &res, &sec) == 0) {
         output::print("Solver information:");
         output::indent();
         output::print("Number of outer iterations: " + std::to_string(outer) + ".");
         output::print("Number of inner iterations: " + std::to_string(inner) + ".");
         output::print("Residual norm", res, true, 3);
         output::outdent();
      }
   }
   if (n > 0) {
      output::print("Time step size", dt, true, 3);
      output::print("Physical time", t, true, 3);
   }
   PAMPA_CHECK(solver->output(".", n), "unable to output the solution");
   output::print("\nDone.");
   return 0;
}

int Driver::finalize() {
   output::print("\n--------------------------------");
   output::print("\nFinalize...");
   output::indent();
   output::print("Finalize the solver...", true);
   if (solver) PAMPA_CHECK(solver->finalize(), "unable to finalize the solver");
   output::print("Done.", true);
   solvers.clear();
   solver = nullptr;
   materials.clear();
   mesh.reset();
   initialized = false;
   output::outdent();
   output::print("Done.\n");
   return 0;
}

int Driver::getField(double* v, const std::string& name) const {
   PAMPA_CHECK(!solver, "the calculation is not initialized");
   return solver->getField(v, name);
}

int Driver::setField(const double* v, const std::string& name) {
   PAMPA_CHECK(!solver, "the calculation is not initialized");
   return solver->setField(v, name);
}

}  // namespace pampa
