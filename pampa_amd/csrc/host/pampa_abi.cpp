// The drop-in C-ABI (pampa.hxx:6-27): one process-global driver.
#include <cstdio>
#include <cstring>

#include "../../../include/pampa.h"
#include "driver.hpp"

static pampa::Driver g_driver;

static int initialize_impl(int argc, char* argv[], std::vector<double>& dt) {
   // the arguments are copied first, as the reference does for Fortran callers
   std::vector<std::string> args(argv, argv + argc);
   std::vector<char*> ptrs;
   for (auto& a : args) ptrs.push_back(&a[0]);
   return g_driver.initialize(argc, ptrs.data(), dt);
}

extern "C" {

void pampa_initialize(int argc, char* argv[], double** dt, int* ndt, int* error) {
   std::vector<double> dt0;
   *error = initialize_impl(argc, argv, dt0);
   if (*error > 0) {
      printf("Error in pampa.initialize().\n");
      return;
   }
   *ndt = (int)dt0.size();
   *dt = new double[dt0.size() ? dt0.size() : 1];
   for (size_t i = 0; i < dt0.size(); i++) (*dt)[i] = dt0[i];
}

void pampa_initialize_steady_state(int argc, char* argv[], int* error) {
   std::vector<double> dt0;
   *error = initialize_impl(argc, argv, dt0);
   if (*error > 0) printf("Error in pampa.initialize().\n");
}

void pampa_solve(int n, double dt, double t, int* error) {
   *error = g_driver.solve(n, dt, t);
   if (*error > 0) printf("Error in pampa.solve().\n");
}

void pampa_solve_steady_state(int* error) {
   pampa_solve(0, 0.0, 0.0, error);
   if (*error > 0) printf("Error in pampa_solve().\n");
}

void pampa_finalize(double** dt, int* error) {
   *error = g_driver.finalize();
   if (*error > 0) {
      printf("Error in pampa.finalize().\n");
      return;
   }
   delete[] * dt;
   *dt = nullptr;
}

void pampa_finalize_steady_state(int* error) {
   *error = g_driver.finalize();
   if (*error > 0) printf("Error in pampa.finalize().\n");
}

void pampa_get_field(double* v, const char name[], int* error) {
   *error = g_driver.getField(v, std::string(name));
   if (*error > 0) printf("Error in pampa.getField().\n");
}

void pampa_set_field(const double* v, const char name[], int* error) {
   *error = g_driver.setField(v, std::string(name));
   if (*error > 0) printf("Error in pampa.setField().\n");
}

void pampa_ext_get_result(double* keff, double* power, int* error) {
   pampa::Solver* s = g_driver.mainSolver();
   *error = s ? 0 : 1;
   if (!s) return;
   if (keff) *keff = s->getKeff();
   if (power) *power = s->getPower();
}

void pampa_ext_get_operator(const char name[], long* n, long* nnz, int* row_ptr, int* col_idx,
                            double* val, int* error) {
   pampa::Solver* s = g_driver.mainSolver();
   *error = s ? s->getOperator(std::string(name), n, nnz, row_ptr, col_idx, val) : 1;
}

void pampa_ext_get_stats(int* outer, long* inner, double* residual, double* seconds, int* error) {
   pampa::Solver* s = g_driver.mainSolver();
   *error = s ? s->getStats(outer, inner, residual, seconds) : 1;
}

void pampa_ext_time_kernel(int reps, int fp32, double* seconds, double* bytes, int* error) {
   pampa::Solver* s = g_driver.mainSolver();
   *error = s ? s->timeKernel(reps, fp32, seconds, bytes) : 1;
}

void pampa_ext_time_source(int reps, int fp32, int mode, double* seconds, double* bytes, int* error) {
   pampa::Solver* s = g_driver.mainSolver();
   *error = s ? s->timeSource(reps, fp32, mode, seconds, bytes) : 1;
}

void pampa_ext_cold_start(int* error) {
   pampa::Solver* s = g_driver.mainSolver();
   *error = s ? s->coldStart() : 1;
}

void pampa_ext_field_device(const char name[], void** dev, long* n, int* error) {
   pampa::Solver* s = g_driver.mainSolver();
   *error = s ? s->fieldDevice(std::string(name), dev, n) : 1;
}

void pampa_ext_get_sizes(int* num_cells, int* num_groups, int* num_directions, int* error) {
   pampa::Solver* s = g_driver.mainSolver();
   *error = s ? 0 : 1;
   if (s) s->getSizes(num_cells, num_groups, num_directions);
}

}  // extern "C"
