/*
 * pk.h -- C-ABI numerics layer of pampa_amd (MI355X / gfx950).
 *
 * This is the boundary that replaces the reference's PETSc/SLEPc usage on the
 * neutronics hot path: the petsc:: wrapper (petsc.hxx:20-89) and the raw
 * Mat/Vec/EPS calls made by the solvers (DiffusionSolver.cxx:96-98,293-316,
 * 329-334; SNSolver.cxx:352-354,512-627,640-645).  Objects are opaque handles
 * over device memory; every call returns 0 on success and a positive code on
 * failure, with a one-line message available from pk_last_error() (the
 * reference's int-return convention, utils.hxx:38-47).  Nothing in this header
 * uses HIP or C++ types.
 *
 * Not re-entrant across threads for the same handles; one HIP stream per
 * process-wide context (the reference is single-threaded per rank, pampa.cxx:8).
 */
#ifndef PAMPA_AMD_PK_H
#define PAMPA_AMD_PK_H

#ifdef __cplusplus
extern "C" {
#endif

#define PK_INSERT_VALUES 0 /* MatSetValues INSERT_VALUES */
#define PK_ADD_VALUES 1    /* MatSetValues ADD_VALUES */

#define PK_PREC_FP64 0  /* every solve in FP64 (the plain drop-in) */
#define PK_PREC_MIXED 1 /* FP32 inner solves / sweeps, FP64 outer quantities */

/* Boundary-condition kinds, numbered as the reference's BC::Type (utils.hxx:103-105). */
#define PK_BC_NONE 0
#define PK_BC_VACUUM 1
#define PK_BC_REFLECTIVE 2
#define PK_BC_ROBIN 3

typedef struct pk_vec_s* pk_vec;
typedef struct pk_mat_s* pk_mat;
typedef struct pk_mesh_s* pk_mesh;
typedef struct pk_xs_s* pk_xs;
typedef struct pk_eig_s* pk_eig;
typedef struct pk_lin_s* pk_lin;
typedef struct pk_batch_s* pk_batch;
typedef struct pk_sn_s* pk_sn;

/* ---- runtime ------------------------------------------------------------
 * pk_initialize replaces petsc::initialize (petsc.cxx:18-36) for the numerics;
 * pk_finalize replaces petsc::finalize. */
int pk_initialize(int device);
int pk_finalize(void);
int pk_synchronize(void);
const char* pk_last_error(void);
int pk_device_memory(double* free_bytes, double* total_bytes); /* hipMemGetInfo of the device */
/* STREAM-style copy of n floats (16-byte accesses): average seconds per launch
 * over reps launches; moves 8 n bytes.  Context for the HBM figures and the
 * calibration kernel of the hardware traffic counters. */
int pk_stream_copy(long n, int reps, double* avg_seconds);

/* ---- vectors (petsc::create(Vec&), destroy, set, copy; petsc.cxx:113-140,
 *      236-345) ------------------------------------------------------------ */
int pk_vec_create(long n, pk_vec* v);
int pk_vec_destroy(pk_vec* v);
long pk_vec_size(pk_vec v);
int pk_vec_set(pk_vec v, double x);                     /* petsc::set / VecSet */
int pk_vec_set_values(pk_vec v, const double* host);    /* petsc::copy(const double*, Vec*) */
int pk_vec_get_values(pk_vec v, double* host);          /* petsc::copy(const Vec*, double*) */
int pk_vec_copy(pk_vec src, pk_vec dst);                /* VecCopy */
int pk_vec_scale(pk_vec v, double a);                   /* VecScale */
int pk_vec_axpy(pk_vec y, double a, pk_vec x);          /* VecAXPY */
int pk_vec_dot(pk_vec x, pk_vec y, double* r);          /* VecDot (fixed-order FP64 reduction) */
int pk_vec_norm2(pk_vec x, double* r);                  /* VecNorm(NORM_2) */
int pk_vec_pointwise_mult(pk_vec w, pk_vec x, pk_vec y); /* VecPointwiseMult: w_i = x_i y_i (the delayed
                                                            source S = beta P, NeutronicSolver.cxx:102) */
int pk_vec_device_ptr(pk_vec v, void** dev);            /* device address (interop) */
/* Entries in [-rel max|v|, 0) set to 0 (round-off of zeros of an iterative
 * solution); *negatives = entries below -rel max|v| (the check of
 * SNSolver.cxx:329 after that clean-up). */
int pk_vec_clip_noise(pk_vec v, double rel, long* negatives);

/* ---- matrices (petsc::create(Mat&) petsc.cxx:91-111, MatSetValues,
 *      MatAssemblyBegin/End).  CSR with PETSc-AIJ semantics: sorted columns,
 *      explicit zeros kept, INSERT (last value wins) or ADD. ---------------- */
int pk_mat_create(long n, int prealloc, pk_mat* M);
int pk_mat_set_values(pk_mat M, int nrows, const int* rows, int ncols, const int* cols,
                      const double* vals, int mode);
int pk_mat_assemble(pk_mat M);
int pk_mat_destroy(pk_mat* M);
int pk_mat_get_size(pk_mat M, long* n, long* nnz);
int pk_mat_get_csr(pk_mat M, int* row_ptr, int* col_idx, double* val); /* host copies */
int pk_mat_mult(pk_mat M, pk_vec x, pk_vec y);                          /* MatMult */
int pk_mat_set_block_size(pk_mat M, int bs); /* cell block (energy groups) for preconditioning */
/* Time one MatMult kernel: average device time per launch over `reps` launches
 * (HIP events on the library stream). */
int pk_mat_time_mult(pk_mat M, pk_vec x, pk_vec y, int reps, int fp32, double* avg_seconds);

/* ---- mesh, cross sections and device assembly ---------------------------
 * The single-domain mesh (Cells/Faces, Mesh.hxx:12-52) in ragged face layout:
 * cell i owns faces face_ptr[i] .. face_ptr[i+1]-1.  fnb >= 0 is a cell,
 * < 0 is -(boundary_index + 1).  cen/fcen/fnorm are xyz triples. */
int pk_mesh_create(int num_cells, const int* face_ptr, const double* vol, const double* cen,
                   const int* mat, const double* farea, const double* fcen, const double* fnorm,
                   const int* fnb, pk_mesh* mesh);
int pk_mesh_destroy(pk_mesh* mesh);
/* Per-material tables (ConstantNuclearData accessors, Material.hxx:129-169):
 * sigt/nusf/kapf/chi/D are [num_mat][G]; sigs is [num_mat][G from][G to].
 * D may be NULL (SN). */
int pk_xs_create(int num_mat, int G, const double* sigt, const double* sigs, const double* nusf,
                 const double* kapf, const double* chi, const double* D, pk_xs* xs);
/* Temperature-dependent tables (FeedbackNuclearData.hxx:8-150): material m
 * owns num_sets[m] consecutive sets tabulated at increasing temperatures
 * temps[] (one set: constant data); every table is [sets][G] (sigs
 * [sets][G][G]).  Coefficients are evaluated at each cell's temperature
 * (pk_xs_set_temperature) by linear interpolation between the two sets around
 * it, the end set outside the tabulated range.  pk_xs_create is the one-set
 * case.  The transient tables of pk_diffusion_assemble_transient (velocity,
 * prompt and delayed spectra) are then per set as well. */
int pk_xs_create_sets(int num_mat, int G, const int* num_sets, const double* temps, const double* sigt,
                      const double* sigs, const double* nusf, const double* kapf, const double* chi,
                      const double* D, pk_xs* xs);
/* The cell temperatures the tables are evaluated at, a vector of one value per
 * cell read at each assembly and normalisation (the solver's "temperature"
 * field, which a coupled code sets through pampa_set_field); NULL: 0 K. */
int pk_xs_set_temperature(pk_xs xs, pk_vec T);
int pk_xs_destroy(pk_xs* xs);
/* new values for the same table sizes (a perturbed copy of a material set; [sets][G]) */
int pk_xs_set(pk_xs xs, const double* sigt, const double* sigs, const double* nusf, const double* kapf,
              const double* chi, const double* D);
/* R and F of DiffusionSolver::buildMatrices (DiffusionSolver.cxx:88-320, n = 0)
 * assembled on the device; bc_type/bc_value are indexed by bc id =
 * boundary_index + 1 (size num_bc).  Replaces the MatSetValues loop. */
int pk_diffusion_assemble(pk_mesh mesh, pk_xs xs, int num_bc, const int* bc_type,
                          const double* bc_value, pk_mat* R, pk_mat* F);
/* The values of R and F (from pk_diffusion_assemble on this mesh) recomputed for
 * new cross sections: same mesh and boundary table, so the same pattern and
 * device layout; nothing is allocated. */
int pk_diffusion_refill(pk_mesh mesh, pk_xs xs, pk_mat R, pk_mat F);
/* Transient rows (DiffusionSolver.cxx:134-148,168-175): R gains
 * V/(v dt) on the diagonal and the prompt fission -(1-beta) chi_p nuSf V / keff;
 * b = chi_d S + V/(v dt) phi0.  velocity/chip/chid are [num_mat][G], beta is
 * [num_mat]; S [num_cells]; phi0 and b are pk_vec of size num_cells*G.  With
 * temperature-dependent tables, velocity/chip/chid are [sets][G]. */
int pk_diffusion_assemble_transient(pk_mesh mesh, pk_xs xs, int num_bc, const int* bc_type,
                                    const double* bc_value, const double* velocity, double dt,
                                    const double* chip, const double* chid, const double* beta,
                                    double keff, pk_vec S, pk_vec phi0, pk_mat* R, pk_vec b);

/* R and F of SNSolver::buildMatrices (SNSolver.cxx:344-631, n = 0) in the
 * reference form, assembled on the device: row l = (i G + g) M + m holds the
 * cell's dense G M block (total reaction, isotropic scattering; F: chi nusf w V)
 * and one entry per physical neighbour (face-interpolation weights of
 * SNSolver.cxx:159-208 with this delta; least-squares boundary correction when
 * ls_coef != NULL; reflected directions on reflective boundaries), values
 * accumulated in the reference's MatSetValues order.  dirs [M][3], wts [M],
 * refl [M][3] (reflected direction per axis, AngularQuadratureSet.cxx:185-205);
 * ls_coef [num_faces][3] per face slot (SNSolver.cxx:212-268, boundary cells),
 * bc_type by bc id.  Replaces the MatSetValues loop of the S_N solver. */
int pk_sn_assemble(pk_mesh mesh, pk_xs xs, int M, const double* dirs, const double* wts, const int* refl,
                   double delta, const double* ls_coef, int num_bc, const int* bc_type, pk_mat* R, pk_mat* F);

/* Power normalisation and power/production rates on the device
 * (NeutronicSolver.cxx:46-116): phi scaled so that sum phi kapf V = power
 * (power < 0: phi left as is, rates only); q[i] = sum_g phi kapf V,
 * P[i] = sum_g phi nusf V / keff.  Returns 1 with an error if a flux is
 * negative. */
int pk_normalize_flux(pk_mesh mesh, pk_xs xs, pk_vec phi, double power, double keff, pk_vec q,
                      pk_vec P, double* total_power);

/* pk_normalize_flux without allocation or library scratch: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same kernels and results as pk_normalize_flux; the reductions' partials,
 * scalars and counters live in a fixed 16.9 KB device array per GPU (resolved
 * by pk_initialize, which must have run), and the call allocates nothing.
 * Cells x groups must stay below 2^31.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (from several host threads)
 * race; the counters are cleared at the start of each call. */
int pk_normalize_flux_static(pk_mesh mesh, pk_xs xs, pk_vec phi, double power, double keff, pk_vec q,
                             pk_vec P, double* total_power);

/* ---- eigen solver (petsc::create(EPS&) petsc.cxx:175-204, EPSSetInitialSpace,
 *      petsc::solve(EPS) petsc.cxx:428-488, EPSGetEigenpair) -------------
 * Solves R x = lambda F x for the eigenvalue of largest 1/lambda (keff) by
 * Wielandt-shifted power iteration with restarted GMRES and cell-block Jacobi
 * inner solves.  tol is the relative residual |R x - F x / k| / |F x / k|. */
int pk_eig_create(pk_mat R, pk_mat F, pk_eig* E);
/* new R and F of the same sizes (the next steady solve of the same problem):
 * the solver keeps its device memory and captured graphs */
int pk_eig_set_operators(pk_eig E, pk_mat R, pk_mat F);
int pk_eig_set_tolerances(pk_eig E, double tol, int max_outer);
int pk_eig_set_precision(pk_eig E, int precision);
int pk_eig_set_inner(pk_eig E, int restart, double reduction, int max_inner);
/* first Wielandt shift distance (default 0.1; 0: no shift): k_s = k + delta
 * once k settles, then per problem halved once down to 0.05 when an inner
 * solve took <= 24 GMRES steps, and down to 0.004 when the outer residual
 * stalls (three updates in a row reducing it by less than 0.8) */
int pk_eig_set_shift(pk_eig E, double delta);
/* Anderson mixing of the outer iterates (Walker and Ni 2011): the update
 * phi -> g = phi + e (e the inner solve's correction) is mixed over the last
 * `depth` differences once the shift is applied, 0 <= depth <= 3 (default 3;
 * 0: the plain update x = phi + e); the eigenpair is that of the plain update */
int pk_eig_set_acceleration(pk_eig E, int depth);
/* inner preconditioner, built on S = cell-block Jacobi B^-1 followed by one
 * block-Jacobi sweep over each 256-row workgroup's diagonal block of A:
 * 0 = S; 1 (default) = S (2 I - A S), one Richardson correction (one more SpMV
 * per GMRES step, about half the steps); 2 = B^-1 alone (no sweep: for
 * operators on which the sweep amplifies, such as S_N with downwind
 * face interpolation, mixed-face-interpolation < 1) */
int pk_eig_set_preconditioner(pk_eig E, int poly);
int pk_eig_set_initial(pk_eig E, pk_vec x);
int pk_eig_solve(pk_eig E);
int pk_eig_get_eigenpair(pk_eig E, double* lambda, pk_vec x); /* lambda = 1/keff */
int pk_eig_get_stats(pk_eig E, int* outer, long* inner, double* residual, double* seconds);
int pk_eig_destroy(pk_eig* E);

/* ---- batched eigen solver: `num` independent problems R_b x = lambda F_b x
 *      (each one what pk_eig solves; petsc::solve(EPS) petsc.cxx:428-488 called
 *      for many inputs) advanced together, one kernel launch per step for all
 *      of them, every convergence decision taken per problem on the device.
 *      Members may differ in size and block size; F_b must hold one cell block
 *      per row (the diffusion fission operator).  Options as for pk_eig. */
int pk_batch_create(int num, const pk_mat* R, const pk_mat* F, pk_batch* B);
int pk_batch_set_member(pk_batch B, int index, pk_mat R, pk_mat F); /* must fit the pool */
int pk_batch_set_tolerances(pk_batch B, double tol, int max_outer);
int pk_batch_set_precision(pk_batch B, int precision);
int pk_batch_set_inner(pk_batch B, int restart, double reduction, int max_inner);
int pk_batch_set_shift(pk_batch B, double delta);
int pk_batch_set_acceleration(pk_batch B, int depth); /* as pk_eig_set_acceleration; pools too */
int pk_batch_set_preconditioner(pk_batch B, int poly); /* as pk_eig_set_preconditioner */
int pk_batch_set_initial(pk_batch B, int index, pk_vec x); /* all-zero x: no initial space */
int pk_batch_solve(pk_batch B);
int pk_batch_get_eigenpair(pk_batch B, int index, double* lambda, pk_vec x);
int pk_batch_get_stats(pk_batch B, int index, int* outer, long* inner, double* residual,
                       double* seconds);
int pk_batch_destroy(pk_batch* B);

/* ---- eigen pool: the engine behind pk_batch, for continuous batching.  A pool
 *      of fixed capacity (rows, stored entries of R and F, block-inverse values,
 *      slots) holds problems that are admitted and released while others
 *      iterate; each step advances every live problem by one tick (an outer
 *      update where due, then one GMRES restart cycle).  Results of a problem do
 *      not depend on what else is in the pool. */
#define PK_POOL_FULL 2 /* pk_pool_admit: no room now, nothing changed */
/* device bytes a pool of this capacity allocates */
double pk_pool_bytes(long max_rows, long max_entries, long max_fentries, long max_binv, int max_slots,
                     int precision, int restart);
int pk_pool_create(long max_rows, long max_entries, long max_fentries, long max_binv, int max_slots,
                   int precision, int restart, pk_batch* P);
int pk_pool_set_options(pk_batch P, double tol, int max_outer, double reduction, int max_inner,
                        double delta);
/* copies R, F (and x0 when not NULL) into the pool; the caller may destroy them after */
int pk_pool_admit(pk_batch P, pk_mat R, pk_mat F, pk_vec x0, int* slot);
/* *fits = 1 when pk_pool_admit(R, F) would find room now */
int pk_pool_fits(pk_batch P, pk_mat R, pk_mat F, int* fits);
/* Autotuner of the pool's operator layout (no counterpart in the reference,
 * whose AIJ storage has no tunable layout; petsc.cxx:91-111 only
 * preallocates rows).  The first problem of each size class (rows, cell
 * block, nonzeros, precision) admitted to any pool times the sliced-ELL SpMV
 * of its R for slice heights 32, 64, 128 and workgroups of 128, 256, 512,
 * 1024 rows, and keeps the fastest pair for every problem of the class (the
 * matrix is laid out again with that height; the plain SpMV takes that
 * workgroup).  Every choice gives the same bits.  PK_TUNE=0 disables it,
 * PK_SLICE=32|64|128 forces a height.  pk_tune_get reports class i in the
 * order tuned: its key, the choice, and the 12 timings in microseconds
 * ([height 32/64/128][workgroup 128/256/512/1024], 0 = height skipped because
 * its padding exceeds the 64-row layout's). */
int pk_tune_count(int* n);
int pk_tune_get(int i, long* rows, int* block, long* nnz, int* precision, int* slice_height, int* spmv_block,
                double* us);
int pk_tune_reset(void);
/* device bytes one problem (R, F) holds in a pool of this precision */
int pk_pool_footprint(pk_mat R, pk_mat F, int precision, int restart, double* bytes);
int pk_pool_step(pk_batch P, int ticks, int* live, int* done);
int pk_pool_collect(pk_batch P, int max, int* slots, int* count); /* finished, admission order */
int pk_pool_result(pk_batch P, int slot, double* keff, pk_vec x, int* outer, long* inner,
                   double* residual, int* converged);
int pk_pool_release(pk_batch P, int slot);
int pk_pool_capacity(pk_batch P, long* free_rows, long* free_entries, int* free_slots, long* ticks);
/* Time the SpMV of every live problem's inner operator in one launch (HIP
 * events on the library stream): average seconds per launch over `reps`, and
 * the algorithmic CSR bytes of one launch (unpadded nonzeros). */
int pk_pool_time_spmv(pk_batch P, int reps, double* avg_seconds, double* bytes);

/* ---- linear solver (petsc::create(KSP&) petsc.cxx:147-172, petsc::solve(KSP)
 *      petsc.cxx:382-425): GMRES + cell-block Jacobi, nonzero initial guess. */
int pk_lin_create(pk_mat A, pk_lin* L);
int pk_lin_set_tolerances(pk_lin L, double rtol, double atol, int max_it);
int pk_lin_solve(pk_lin L, pk_vec b, pk_vec x);
int pk_lin_get_stats(pk_lin L, long* iterations, double* residual);
int pk_lin_destroy(pk_lin* L);

/* ---- discrete ordinates (matrix-free form of SNSolver::buildMatrices in the
 *      upwind regime: mixed-face-interpolation 1, LS off, vacuum boundaries).
 * dirs[M][3], wts[M] from the level-symmetric set (AngularQuadratureSet.cxx:4-209).
 * The operator R psi = lambda F psi of SNSolver.cxx:344-631 is applied as
 * upwind sweeps over level schedules; psi layout is the reference's
 * [cell][group][direction]. */
int pk_sn_create(pk_mesh mesh, pk_xs xs, int M, const double* dirs, const double* wts, double delta,
                 int num_bc, const int* bc_type, pk_sn* S);
int pk_sn_get_schedule(pk_sn S, int* num_levels, int* level_of_cell); /* [M], [M][C] */
int pk_sn_set_precision(pk_sn S, int precision);
int pk_sn_set_tolerances(pk_sn S, double tol_k, double tol_phi, int max_outer);
/* One sweep psi = L^-1 q for all (g, m); q[cell][group] includes the volume.
 * phiw[cell][group] = sum_m w_m psi.  Host arrays, for parity checks. */
int pk_sn_sweep_host(pk_sn S, const double* q, double* psi, double* phiw);
/* Source iteration inside power iteration (NeutronicSolver::solve,
 * NeutronicSolver.cxx:4-43, in the sweepable regime).  Every handle runs on a
 * stream of its own with its own scratch: solves of different handles may run
 * at the same time from different host threads, and beside the library
 * stream.  The call returns when the handle's work is complete.  A handle
 * starts from its previous solution (warm) unless reset. */
int pk_sn_solve(pk_sn S, double* keff, pk_vec psi, pk_vec phiw);
/* New cross sections for the handle's tables (a perturbed request on the same
 * mesh; no counterpart in the reference, which rebuilds the solver): total,
 * scattering [from][to], nu-fission, kappa-fission and chi per material and
 * group, uploaded on the handle's stream; the sweep records are rebuilt on the
 * next sweep, the schedules and value slots are kept. */
int pk_sn_set_xs(pk_sn S, const double* sigt, const double* sigs, const double* nusf, const double* kapf,
                 const double* chi);
/* Allocate the buffers and build the sweep records of the current precision
 * now (they are otherwise made by the first solve), so that later solves and
 * cross-section changes allocate nothing. */
int pk_sn_prepare(pk_sn S);
/* The scalar-flux moment phiw[cell][group] = sum_m w_m psi of the last solve
 * (host copy on the handle's stream; production-normalised as iterated). */
int pk_sn_get_moment(pk_sn S, double* phiw);
/* Anderson acceleration of the source iteration (Walker and Ni 2011) over the
 * last `depth` iterates, 0 <= depth <= 8 (0: plain source iteration; default
 * 8).  The fixed point, so the eigenpair at convergence, is the same. */
int pk_sn_set_acceleration(pk_sn S, int depth);
/* Forget the previous solution: the next solve starts cold. */
int pk_sn_reset(pk_sn S);
int pk_sn_get_stats(pk_sn S, int* outer, long* sweeps, double* seconds);
int pk_sn_time_sweep(pk_sn S, int reps, double* avg_seconds, double* bytes_per_sweep);
/* The isotropic scattering + fission source of each source iteration,
 * q_i = V_i (sum_g2 Sigma_s(g2->g) phiw_i,g2 + chi_g sum_g2 nuSigma_f,g2 phiw_i,g2 / k)
 * (SNSolver.cxx:417-439 moved to the right-hand side), is a per-material
 * G x G contraction.  mode 0 (default): on the matrix cores (FP64 MFMA) when
 * G = 16, else the vector kernel; 1: always the vector kernel; 2: always the
 * matrix cores (G = 16 only).  The two agree to rounding. */
int pk_sn_set_contraction(pk_sn S, int mode);
/* That source in FP64 for a given moment phiw[cell][group] and k (host
 * arrays), through the handle's contraction: a parity entry point. */
int pk_sn_source_host(pk_sn S, const double* phiw, double keff, double* q);
/* Average device time of the source kernel alone in the handle's precision
 * (reps calls after one warm-up) and its algorithmic bytes per call.  It
 * overwrites the handle's moment, so the next solve starts cold. */
int pk_sn_time_source(pk_sn S, int reps, double* avg_seconds, double* bytes_per_call);
int pk_sn_destroy(pk_sn* S);

/* ---- kernel profiler (replaces the timing printed by petsc::solve with
 *      -petsc_verbose, petsc.cxx:382-425,428-488): with profiling on, the hot
 *      kernels ("gmres_spmv_dot", "sn_sweep_step", "spmv") are bracketed by HIP
 *      events on the library stream; per kernel name this returns the launch
 *      count, summed device seconds and summed algorithmic bytes. */
int pk_profile_enable(int on);
int pk_profile_reset(void);
int pk_profile_get(const char* name, long* launches, double* seconds, double* bytes);

#ifdef __cplusplus
}
#endif

#endif
