// Calculation driver of the drop-in (Driver.hxx:13-58, Parser.cxx:4-260):
// parse the input, initialise the main solver, solve steps, output, finalise.
#pragma once

#include <memory>

#include "solver.hpp"

namespace pampa {

class Driver {
  public:
   int initialize(int argc, char* argv[], std::vector<double>& dt);
   int solve(int n, double dt, double t);
   int finalize();
   int getField(double* v, const std::string& name) const;
   int setField(const double* v, const std::string& name);
   Solver* mainSolver() const { return solver; }

  private:
   int parse(const std::string& filename, std::vector<double>& dt);
   int setOption(const std::string& name, const std::string& value);
   std::unique_ptr<Mesh> mesh;
   std::vector<Material> materials;
   std::vector<std::unique_ptr<Solver>> solvers;
   Solver* solver = nullptr;
   bool initialized = false;
};

}  // namespace pampa
