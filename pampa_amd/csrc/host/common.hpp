// Host-side utilities of the pampa_amd drop-in: error convention, terminal
// output protocol and small containers.
#pragma once

#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/pk.h"

namespace pampa {

// Errors print "file:line:function(): error: message." and propagate an int
// status (the reference's PAMPA_CHECK convention, utils.hxx:41-47).
void report_error(const char* file, int line, const char* func, const std::string& msg);

#define PAMPA_CHECK(cond, msg)                                             \
   do {                                                                    \
      if (cond) {                                                          \
         ::pampa::report_error(__FILE__, __LINE__, __func__, (msg));       \
         return 1;                                                         \
      }                                                                    \
   } while (0)

#define PK_CHECK(call)                                                                   \
   do {                                                                                  \
      if ((call) != 0) {                                                                 \
         ::pampa::report_error(__FILE__, __LINE__, __func__, std::string(pk_last_error())); \
         return 1;                                                                       \
      }                                                                                  \
   } while (0)

// Terminal output (output.cxx:31-93): 3 spaces per indentation level, info
// messages only with -verbose, nothing with -silent.
namespace output {
extern bool verbose, silent;
extern int padding;
void print(const std::string& message, bool info = false);
void print(const std::string& name, double x, bool scientific, int precision, bool info = false);
void indent(bool info = false);
void outdent(bool info = false);
}  // namespace output

// Boundary conditions (utils.hxx:103-119), constant parameter (time tables
// are evaluated at t by the Function rules of Function.hxx:34-52).
enum BCType { BC_NONE = 0, BC_VACUUM = 1, BC_REFLECTIVE = 2, BC_ROBIN = 3, BC_DIRICHLET = 4,
              BC_ADIABATIC = 5, BC_CONVECTION = 6 };

struct Function {
   std::vector<double> t, y;
   double operator()(double x) const;
};

struct BoundaryCondition {
   BCType type = BC_NONE;
   std::vector<Function> f;
};

}  // namespace pampa
