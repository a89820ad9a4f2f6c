/*
 * pampa.h -- drop-in C-ABI of pampa_amd.
 *
 * The first eight entry points keep the reference's library interface verbatim
 * (pampa.hxx:6-27, implemented at pampa.cxx:11-113): one process-global
 * calculation driven by an input file, *error > 0 on failure with a one-line
 * message on stdout, and the time-step array allocated by pampa_initialize and
 * released by pampa_finalize.  Field names are those of the reference's solver
 * registries (DiffusionSolver.cxx:384-403, SNSolver.cxx:721-747):
 * "temperature", "delayed-source" (inputs), "scalar-flux", "angular-flux",
 * "power", "production-rate"; get/set copy exactly the field's length.
 *
 * The pampa_ext_* entry points have no counterpart in the reference: they read
 * results that the reference only prints (keff, total power), the assembled
 * operators (for parity checks) and solver statistics.
 */
#ifndef PAMPA_AMD_PAMPA_H
#define PAMPA_AMD_PAMPA_H

#ifdef __cplusplus
extern "C" {
#endif

/* pampa.hxx:9 / pampa.cxx:11-35 */
void pampa_initialize(int argc, char* argv[], double** dt, int* ndt, int* error);
/* pampa.hxx:12 / pampa.cxx:38-56 */
void pampa_initialize_steady_state(int argc, char* argv[], int* error);
/* pampa.hxx:15 / pampa.cxx:59-65 */
void pampa_solve(int n, double dt, double t, int* error);
/* pampa.hxx:18 / pampa.cxx:68-74 */
void pampa_solve_steady_state(int* error);
/* pampa.hxx:21 / pampa.cxx:77-86 */
void pampa_finalize(double** dt, int* error);
/* pampa.hxx:24 / pampa.cxx:89-95 */
void pampa_finalize_steady_state(int* error);
/* pampa.hxx:27 / pampa.cxx:98-104 */
void pampa_get_field(double* v, const char name[], int* error);
/* pampa.hxx:30 / pampa.cxx:107-113 */
void pampa_set_field(const double* v, const char name[], int* error);

/* Extensions (no reference counterpart). */
/* keff and total power of the last solve (the values printed by printLog,
 * NeutronicSolver.cxx:119-127). */
void pampa_ext_get_result(double* keff, double* power, int* error);
/* Assembled operator "R" or "F" of the main solver as CSR: call with null
 * arrays to get n and nnz, then with arrays of n+1 / nnz / nnz entries. */
void pampa_ext_get_operator(const char name[], long* n, long* nnz, int* row_ptr, int* col_idx,
                            double* val, int* error);
/* Outer iterations, inner iterations (Krylov steps or sweeps), final residual
 * and device seconds of the last solve. */
void pampa_ext_get_stats(int* outer, long* inner, double* residual, double* seconds, int* error);
/* Number of cells and energy groups of the main solver's mesh (field sizes). */
void pampa_ext_get_sizes(int* num_cells, int* num_groups, int* num_directions, int* error);
/* Device time of the main solver's hot kernel, averaged over `reps` back-to-back
 * launches (HIP events on the library stream), and its algorithmic bytes per
 * launch: the SpMV of R (diffusion) or one full transport sweep (S_N, summed
 * over its level steps).  fp32 != 0 times the FP32 variant. */
void pampa_ext_time_kernel(int reps, int fp32, double* seconds, double* bytes, int* error);

#ifdef __cplusplus
}
#endif

#endif
