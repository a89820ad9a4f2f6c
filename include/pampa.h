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

/* pampa.hxx:6 / pampa.cxx:11-35 */
void pampa_initialize(int argc, char* argv[], double** dt, int* ndt, int* error);
/* pampa.hxx:9 / pampa.cxx:38-56 */
void pampa_initialize_steady_state(int argc, char* argv[], int* error);
/* pampa.hxx:12 / pampa.cxx:59-65 */
void pampa_solve(int n, double dt, double t, int* error);
/* pampa.hxx:15 / pampa.cxx:68-74 */
void pampa_solve_steady_state(int* error);
/* pampa.hxx:18 / pampa.cxx:77-86 */
void pampa_finalize(double** dt, int* error);
/* pampa.hxx:21 / pampa.cxx:89-95 */
void pampa_finalize_steady_state(int* error);
/* pampa.hxx:24 / pampa.cxx:98-104 */
void pampa_get_field(double* v, const char name[], int* error);
/* pampa.hxx:27 / pampa.cxx:107-113 */
void pampa_set_field(const double* v, const char name[], int* error);

/* Extensions (no reference counterpart). */
/* keff and total power of the last solve (the values printed by printLog,
 * NeutronicSolver.cxx:119-127). */
void pampa_ext_get_result(double* keff, double* power, int* error);
/* Assembled operator "R" or "F" of the main solver as CSR: call with null
 * arrays to get n and nnz, then with arrays of n+1 / nnz / nnz entries.
 * "b" is the right-hand side of the last time step (n > 0): nnz = 0 and the
 * n values in val (row_ptr and col_idx are not used).  After a time step "R"
 * is that step's operator (DiffusionSolver.cxx:134-175). */
void pampa_ext_get_operator(const char name[], long* n, long* nnz, int* row_ptr, int* col_idx,
                            double* val, int* error);
/* Device address (HIP device memory of the library's GPU) and length of a
 * field of the main solver, for a coupled code that exchanges fields on the
 * same GPU without host copies (the fields of pampa_get_field/set_field,
 * Solver.cxx:18-43).  Writes through it take effect at the next solve; the
 * address holds until pampa_finalize. */
void pampa_ext_field_device(const char name[], void** dev, long* n, int* error);
/* Outer iterations, inner iterations (Krylov steps or sweeps), final residual
 * and device seconds of the last solve. */
void pampa_ext_get_stats(int* outer, long* inner, double* residual, double* seconds, int* error);
/* The next steady solve starts cold: the S_N solver forgets its previous
 * solution (a diffusion solve starts from the scalar-flux field, which the
 * caller sets). */
void pampa_ext_cold_start(int* error);
/* Number of cells and energy groups of the main solver's mesh (field sizes). */
void pampa_ext_get_sizes(int* num_cells, int* num_groups, int* num_directions, int* error);
/* Device time of the main solver's hot kernel, averaged over `reps` back-to-back
 * launches (HIP events on the library stream), and its algorithmic bytes per
 * launch: the SpMV of R (diffusion) or one full transport sweep (S_N, summed
 * over its level steps).  fp32 != 0 times the FP32 variant. */
void pampa_ext_time_kernel(int reps, int fp32, double* seconds, double* bytes, int* error);
/* The same for the S_N solver's scattering + fission source contraction
 * (pk_sn_time_source) with contraction mode 1 (vector kernel) or 2 (matrix
 * cores, 16 groups); the solver keeps its own mode (-pk_contraction). */
void pampa_ext_time_source(int reps, int fp32, int mode, double* seconds, double* bytes, int* error);

/* ---- serving (no counterpart in the reference; for callers that solve many
 * perturbed copies of a few base problems as requests arrive).  A server owns
 * one device eigen pool sized to budget_bytes of device memory (<= 0: 80% of
 * the free device memory) with max_slots problems at once, and a set of
 * resident base cases (diffusion inputs and sweepable S_N inputs).  Diffusion
 * requests are admitted in arrival order as slots and memory free up and are
 * batched into every kernel launch with the others in flight (continuous
 * batching); S_N requests go in arrival order to the case's workers (host
 * threads, each with its own sweep handle and stream, running beside the
 * pool).  A result does not depend on what else was in flight.  Independent
 * of the process-global calculation of pampa_initialize. */
typedef struct pampa_server_s* pampa_server;
void pampa_serve_create(double budget_bytes, int max_slots, int precision, pampa_server* S, int* error);
/* eigen tolerance (relative residual), inner GMRES reduction, Wielandt shift */
void pampa_serve_set_options(pampa_server S, double tol, double reduction, double shift, int* error);
/* parse and upload an input; rows = cells x groups */
void pampa_serve_add_case(pampa_server S, const char* input, int* case_id, long* rows, int* groups,
                          int* error);
/* the cross-section tables a request (seed, rel) of the case solves with:
 * sigt, D, nusf [nm][G], sigs [nm][G][G] (NULL: skip) */
void pampa_serve_case_tables(pampa_server S, int case_id, unsigned long long seed, double rel,
                             double* sigt, double* sigs, double* D, double* nusf, int* error);
/* device bytes one request of the case holds in the pool, and the pool's size */
void pampa_serve_footprint(pampa_server S, int case_id, double* bytes, double* pool_bytes, int* error);
/* n requests (open loop): request r arrives arrival[r] seconds after the call
 * starts (non-decreasing) and asks for case case_id[r] with every cross section
 * scaled by its own U[1-rel, 1+rel] factor from seed[r] (0: none); warm != 0
 * starts it from the case's latest converged flux (S_N: the worker's previous
 * solution).  Per request: keff, latency (completion - arrival, s),
 * outer/inner iterations (S_N: source iterations, sweeps), converged, and,
 * where flux[r] != NULL, the eigenvector (cells x groups; diffusion: unit
 * 2-norm; S_N: the moment sum_m w_m psi). */
void pampa_serve_run(pampa_server S, int n, const double* arrival, const int* case_id,
                     const unsigned long long* seed, double rel, int warm, double* keff, double* latency,
                     int* outer, long* inner, int* converged, double* const* flux, int* error);
void pampa_serve_forget_warm(pampa_server S, int* error);
/* workers (sweep handles) per S_N case added after this call (default 1) */
void pampa_serve_set_sn_workers(pampa_server S, int workers, int* error);
/* seconds > 0: pampa_serve_run prints a progress line on stderr that often */
void pampa_serve_set_progress(pampa_server S, double seconds, int* error);
/* Every case once in the pool, and the SpMV of all their inner operators in
 * one launch timed (average seconds over reps, algorithmic CSR bytes, rows). */
void pampa_serve_time_spmv(pampa_server S, int reps, double* seconds, double* bytes, double* rows, int* error);
void pampa_serve_stats(pampa_server S, double* seconds, long* ticks, int* max_live, double* pool_bytes,
                       int* error);
void pampa_serve_destroy(pampa_server* S, int* error);

#ifdef __cplusplus
}
#endif

#endif
