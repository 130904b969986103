/*
 * w90gpu.h — C ABI of the MI355X wannierisation library (libw90gpu.so).
 *
 * Scope: the spread minimisation of wannier90 without disentanglement
 * (wann_main, non-gamma, no preconditioner / site symmetry / SLWF / guiding
 * centres).  Every entry point takes plain pointers and sizes; complex
 * arrays are interleaved (re, im) doubles in Fortran (column-major) order,
 * exactly the memory of wannier90's complex(kind=dp) arrays, so a Fortran
 * caller passes its arrays through ISO_C_BINDING without copies and a C or
 * Python caller passes numpy arrays in order='F'.
 *
 * Status convention (src/error.F90:27-36): 0 ok, >0 error, <0 warning.
 *
 * Threading: a context may be used by one host thread at a time.  Sessions of
 * one context share its device and run on their own HIP streams.
 */
#ifndef W90GPU_H
#define W90GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: same values as w90_error codes (src/error.F90:27-36) */
#define W90GPU_OK 0
#define W90GPU_ERR_FATAL 1
#define W90GPU_ERR_ALLOC 2
#define W90GPU_ERR_DEALLOC 3
#define W90GPU_ERR_INPUT 5
#define W90GPU_ERR_FILE 6
#define W90GPU_WARN_UNCONVERGED (-1)
#define W90GPU_WARN (-2)

#define W90GPU_NUM_NNMAX 12      /* types.F90:116 num_nnmax */
#define W90GPU_MAX_WANN 256      /* largest num_wann the kernels accept */

/* arithmetic of the rotation / gradient pass */
#define W90GPU_PREC_FP64 0       /* plain drop-in: everything in FP64 */
#define W90GPU_PREC_MIXED 1      /* FP32 MFMA bulk + FP64 phases, retraction and polish */

typedef struct w90gpu_ctx w90gpu_ctx;
typedef struct w90gpu_session w90gpu_session;

/* Control parameters of the minimiser on this path; field meaning and
 * defaults follow wann_control_type (src/wannier90_types.F90:170-192) and
 * the .win keyword defaults (src/wannier90_readwrite.F90:628-804). */
typedef struct {
  int num_iter;          /* num_iter            (100)   */
  int num_cg_steps;      /* num_cg_steps        (5)     */
  int conv_window;       /* conv_window         (-1: never stop early) */
  int lfixstep;          /* fixed_step given -> 1 */
  double conv_tol;       /* conv_tol            (1e-10) */
  double trial_step;     /* trial_step          (2.0)   */
  double fixed_step;     /* fixed_step          (-999)  */
  double conv_noise_amp; /* must be <= 0 on this path  */
  int precond;           /* must be 0 */
  int guiding_centres;   /* must be 0 */
  int selective_loc;     /* must be 0 */
  int precision;         /* W90GPU_PREC_FP64 or W90GPU_PREC_MIXED */
} w90gpu_control;

/* One wannierisation problem: sizes and the k-mesh tables of kmesh_info_type
 * (src/types.F90:161-176).  Arrays are read during the call only. */
typedef struct {
  int num_wann;
  int num_kpts;
  int nntot;
  const int* nnlist;     /* (num_kpts, nntot), 1-based k indices */
  const double* bk;      /* (3, nntot, num_kpts) Cartesian b-vectors, 1/Ang */
  const double* wb;      /* (nntot) weights, Ang^2 */
  double wbtot;          /* sum of wb */
} w90gpu_problem;

/* Spread decomposition and loop statistics (wann_omega_type and the
 * localisation_vars_type of src/wannierise.F90:35-43). */
typedef struct {
  double omega_total;     /* Omega = Omega_I + Omega_D + Omega_OD        */
  double omega_invariant; /* Omega_I                                     */
  double omega_d;         /* Omega_D                                     */
  double omega_od;        /* Omega_OD                                    */
  double omega_tilde;     /* Omega_D + Omega_OD                          */
  double gcnorm1;         /* last squared gradient norm (RMS^2 on CONV)  */
  double delta_omega;     /* last iteration's change of Omega            */
  double solve_seconds;   /* wall time of the loop on the device         */
  int iterations;         /* iterations performed                        */
  int converged;          /* 1 if the conv_window test was satisfied     */
  int unitarity_ok;       /* 1 if the final U passed the 1e-5 check      */
  int status;             /* status of the solve                         */
  int bulk_iterations;    /* iterations of the FP32 bulk phase (mixed)   */
  int held_second_steps; /* parabolic steps launched by the host after a block
                            without second-step kernels (engine diagnostics) */
} w90gpu_result;

/* ---------------------------------------------------------------- setup */

/* Fill defaults of wann_control_type. */
void w90gpu_control_defaults(w90gpu_control* ctl);

/* Create a context on HIP device `device`.  Replaces the library state
 * object of w90_create (src/c_interface.F90:10) for the GPU path. */
int w90gpu_create(int device, w90gpu_ctx** ctx);
int w90gpu_destroy(w90gpu_ctx* ctx);

/* Message of the last error on this context (empty if none). */
const char* w90gpu_last_error(const w90gpu_ctx* ctx);

/* Library version string. */
const char* w90gpu_version(void);

/* -------------------------------------------------------------- k-mesh */

/* Shells, B1 weights and neighbour lists; replaces kmesh_get
 * (src/kmesh.F90:61-740).  real_lattice(3,3): row i is lattice vector a_i
 * (Fortran real_lattice(i,:)), Ang.  kpt_latt(3,num_kpts) fractional.
 * num_shells <= 0 selects shells automatically (kmesh_shell_automatic);
 * otherwise shell_list(num_shells) is used (kmesh_shell_fixed).
 * Outputs must hold W90GPU_NUM_NNMAX neighbours; they are written packed
 * with the returned nntot: nnlist(num_kpts,nntot) 1-based,
 * nncell(3,num_kpts,nntot), bk(3,nntot,num_kpts), wb(nntot), wbtot. */
int w90gpu_kmesh_get(const double* real_lattice, int num_kpts, const double* kpt_latt,
                     int search_shells, double tol, int num_shells, const int* shell_list,
                     int skip_b1_tests, int* nntot, int* nnlist, int* nncell, double* bk,
                     double* wb, double* wbtot, char* errmsg, int errlen);

/* -------------------------------------------------------- file formats */

/* Reads path (a .mmn file) into m_matrix(num_bands,num_bands,nntot,num_kpts)
 * complex; replaces the .mmn part of overlap_read (src/overlap.F90:234-308).
 * Each block "k k' G" goes to the b with nnlist(k,b) = k' and
 * nncell(:,k,b) = G (nnlist(num_kpts,nntot) 1-based, nncell(3,num_kpts,nntot),
 * the layout of w90gpu_kmesh_get); the later of two blocks with one
 * destination wins, destinations no block names are zero.  The file is
 * parsed by nthreads host threads (<= 0: OMP_NUM_THREADS, else all cores);
 * values are converted with correct rounding, as strtod does.
 * W90GPU_ERR_FILE with the reference's message in errmsg when the file
 * cannot be opened or read, its sizes differ, or a block has no (or more
 * than one) matching neighbour. */
int w90gpu_read_mmn(const char* path, int num_bands, int num_kpts, int nntot, const int* nnlist,
                    const int* nncell, double* m_matrix, int nthreads, char* errmsg, int errlen);

/* Reads path (a .amn file) into a_matrix(num_bands,num_proj,num_kpts)
 * complex; replaces the .amn part of overlap_read (src/overlap.F90:323-366)
 * for num_proj = num_wann projections, all of them used.  Entries no record
 * names are zero; the later of two records with one destination wins.
 * Threads, conversion and errors as w90gpu_read_mmn (a record whose
 * indices are out of range is a read error). */
int w90gpu_read_amn(const char* path, int num_bands, int num_kpts, int num_proj, double* a_matrix,
                    int nthreads, char* errmsg, int errlen);

/* ------------------------------------------------ one-shot drop-in solve */

/* The spread minimisation loop of wann_main (src/wannierise.F90:48-1072)
 * for one problem.  m_matrix (num_wann,num_wann,nntot,num_kpts) complex and
 * u_matrix (num_wann,num_wann,num_kpts) complex are updated in place, as
 * w90_wannierise does with the caller's arrays
 * (src/library_interface.F90:464-530).  centres(3,num_wann) Ang and
 * spreads(num_wann) Ang^2 receive wannier_data%centres/spreads.
 * Returns W90GPU_WARN_UNCONVERGED never (the reference does not flag a
 * loop that ran out of iterations); >0 on error with a message in
 * w90gpu_last_error. */
int w90gpu_wann_main(w90gpu_ctx* ctx, const w90gpu_problem* prob, const w90gpu_control* ctl,
                     double* m_matrix, double* u_matrix, double* centres, double* spreads,
                     w90gpu_result* result);

/* ------------------------------------------- per-phase entry points (FP64) */

/* wann_omega (src/wannierise.F90:1979-2284), non-SLWF branch, csheet=1,
 * sheet=0.  If first_pass, Omega_I is computed from M; otherwise
 * omega_invariant is used.  Outputs: rave(3,num_wann), r2ave(num_wann),
 * omega[4] = {Omega_I, Omega_D, Omega_OD, Omega_total}. */
int w90gpu_omega(w90gpu_ctx* ctx, const w90gpu_problem* prob, const double* m_matrix,
                 int first_pass, double omega_invariant, double* rave, double* r2ave,
                 double* omega);

/* No counterpart in the reference.  w90gpu_omega for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same result, computed with every
 * device buffer in a static array of W90GPU_NOALLOC_SCRATCH_BYTES per device
 * (taken when the library's code is loaded on a device, held until the
 * process ends), on the null stream; it calls no hipMalloc, hipFree or other
 * allocating routine.  A problem whose buffers exceed the array is rejected
 * with W90GPU_ERR_INPUT.  Calls on one device share the array: calls that
 * overlap on one GPU race. */
#define W90GPU_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int w90gpu_omega_noalloc(w90gpu_ctx* ctx, const w90gpu_problem* prob, const double* m_matrix,
                         int first_pass, double omega_invariant, double* rave, double* r2ave,
                         double* omega);

/* wann_domega (src/wannierise.F90:2287-2569), non-SLWF gradient:
 * cdodq(num_wann,num_wann,num_kpts) complex. */
int w90gpu_domega(w90gpu_ctx* ctx, const w90gpu_problem* prob, const double* m_matrix,
                  double* cdodq);

/* internal_new_u_and_m (src/wannierise.F90:1599-1751): with cdq the
 * anti-Hermitian step (num_wann,num_wann,num_kpts), U(k) <- U(k) exp(cdq_k)
 * and M(k,b) <- exp(cdq_k)^+ M(k,b) exp(cdq_{k+b}), in place. */
int w90gpu_new_u_and_m(w90gpu_ctx* ctx, const w90gpu_problem* prob, const double* cdq,
                       double* u_matrix, double* m_matrix);

/* overlap_project (src/overlap.F90:758-952) for num_bands = num_wann: the
 * Lowdin-orthonormalised U(k) = Z V^+ of the SVD A(k) = Z S V^+ (the unitary
 * polar factor of A), then M(k,b) <- U(k)^+ M(k,b) U(k+b).
 * a_matrix (num_wann,num_wann,num_kpts) complex in, u_matrix out, m_matrix
 * (num_wann,num_wann,nntot,num_kpts) rotated in place.  Fails like the
 * reference when the resulting U is not unitary to 1e-5 (singular A). */
int w90gpu_project(w90gpu_ctx* ctx, const w90gpu_problem* prob, const double* a_matrix,
                   double* u_matrix, double* m_matrix);

/* ------------------------------------------------ device-resident sessions */

/* A session keeps one problem's M, U and k-mesh tables resident in HBM and
 * reruns the minimisation on them without host traffic.  Used by the
 * drop-in solve, by benchmarks (inputs resident in HBM) and by serving. */
int w90gpu_session_create(w90gpu_ctx* ctx, const w90gpu_problem* prob, const w90gpu_control* ctl,
                          w90gpu_session** sess);
int w90gpu_session_destroy(w90gpu_session* sess);
/* Upload the initial M and U (host, Fortran layout). */
int w90gpu_session_load(w90gpu_session* sess, const double* m_matrix, const double* u_matrix);
/* Run the minimisation from the loaded initial state (which is kept). */
int w90gpu_session_solve(w90gpu_session* sess, w90gpu_result* result);
/* Enqueue the solve without waiting; w90gpu_session_wait completes it. */
int w90gpu_session_launch(w90gpu_session* sess);
int w90gpu_session_wait(w90gpu_session* sess, w90gpu_result* result);
/* Download the final state of the last solve. */
int w90gpu_session_fetch(w90gpu_session* sess, double* m_matrix, double* u_matrix,
                         double* centres, double* spreads);
/* Per-iteration record of the last solve: row i (i = 0 .. iterations) holds
 * {Omega, Omega_D, Omega_OD, gcnorm1} after iteration i (row 0: initial
 * state), i.e. the numbers of the CONV/SPRD lines of wann_main
 * (src/wannierise.F90:703-756).  Writes min(rows, iterations+1) rows;
 * returns the number written or a negative status. */
int w90gpu_session_history(w90gpu_session* sess, double* out, int rows);
/* Device bytes held by a session of this problem shape. */
size_t w90gpu_session_bytes(int num_wann, int num_kpts, int nntot);
/* Kernel timing: the M-rotation kernel of this problem (the fused rotation
 * and omega partial sums of one trial step) launched `reps` times back to
 * back on the final state of the last solve, replayed from a HIP graph as
 * in the solve and bracketed by HIP events on the session's stream;
 * *avg_ms = elapsed / reps.  The solver state is not changed (each launch
 * rewrites the same scratch buffer). */
int w90gpu_session_rotation_time(w90gpu_session* sess, int reps, double* avg_ms);
/* The same for one phase of an iteration: W90GPU_PHASE_ROTATION (as
 * above), W90GPU_PHASE_GRADIENT (wann_domega: G and its per-k norms),
 * W90GPU_PHASE_RETRACTION (exp of the scaled search direction and the U
 * update of the parabolic step) or W90GPU_PHASE_REDUCTION (the final
 * reduction of wann_omega over k, without the control logic that follows
 * it).  Each launch rewrites scratch buffers only (the other M or U buffer,
 * G, exp(dW), the other centres buffer), so the returned M, U and centres
 * of the solve are unchanged. */
#define W90GPU_PHASE_ROTATION 0
#define W90GPU_PHASE_GRADIENT 1
#define W90GPU_PHASE_RETRACTION 2
#define W90GPU_PHASE_REDUCTION 3
int w90gpu_session_phase_time(w90gpu_session* sess, int phase, int reps, double* avg_ms);

/* ------------------------------------------------ batches and serving
 *
 * No counterpart in the reference (which runs one problem per process):
 * many independent problems of any num_wann / k-mesh solved together.  A
 * batch owns one device pool of pool_bytes; a problem is admitted when its
 * w90gpu_problem_bytes fit in the free part of the pool.  Every phase of an
 * iteration is one launch over all running problems; each problem keeps its
 * own loop state, convergence test and precision, and its results are the
 * same as those of a session running it alone (the kernels sum in an order
 * that does not depend on the other problems).  Problems may be added while
 * others run (continuous batching): w90gpu_batch_step starts them at the next
 * block of iterations.  Not thread-safe: one host thread per batch. */
typedef struct w90gpu_batch w90gpu_batch;

/* Device bytes one problem holds (inputs, state, workspaces, and the
 * per-iteration record of num_iter iterations). */
size_t w90gpu_problem_bytes(int num_wann, int num_kpts, int nntot, int num_iter);
int w90gpu_batch_create(w90gpu_ctx* ctx, size_t pool_bytes, w90gpu_batch** batch);
int w90gpu_batch_destroy(w90gpu_batch* batch);
/* Admit a problem with its initial M and U (host, Fortran layout; copied
 * before the call returns).  *id >= 0 identifies it.  Returns
 * W90GPU_ERR_ALLOC, with the batch unchanged, when the pool has no room. */
int w90gpu_batch_add(w90gpu_batch* batch, const w90gpu_problem* prob, const w90gpu_control* ctl,
                     const double* m_matrix, const double* u_matrix, int* id);
/* Start the added problems and run max_blocks blocks of iterations over all
 * running problems (0: only start them).  ids of problems finished since
 * the last call go to finished[0 .. *nfinished) (at most cap; the rest are
 * reported by the next call). */
int w90gpu_batch_step(w90gpu_batch* batch, int max_blocks, int* finished, int cap, int* nfinished);
/* Run until every added problem has finished. */
int w90gpu_batch_run(w90gpu_batch* batch);
/* Queue every finished problem again from its initial M and U. */
int w90gpu_batch_restart(w90gpu_batch* batch);
/* Result and final state of a finished problem (M and U stay on the device
 * until removed). */
int w90gpu_batch_result(w90gpu_batch* batch, int id, w90gpu_result* result);
int w90gpu_batch_fetch(w90gpu_batch* batch, int id, double* m_matrix, double* u_matrix, double* centres,
                       double* spreads);
/* Release a problem that is not running; its bytes return to the pool. */
int w90gpu_batch_remove(w90gpu_batch* batch, int id);
int w90gpu_batch_info(w90gpu_batch* batch, size_t* pool_bytes, size_t* used_bytes, int* running, int* queued);

/* ------------------------------------------------------ disentanglement
 *
 * dis_main (src/disentangle.F90:31-284) for k-point meshes, followed by
 * setup_m_loc (:286-370): the optimally connected num_wann-dimensional
 * subspace of the num_bands bands inside the outer energy window, with the
 * bands of the inner (frozen) window kept, found by the iteration of
 * dis_extract (:1893-2864).  Not covered: gamma_only, site symmetry and
 * dis_spheres (W90GPU_ERR_INPUT).  Limits: num_bands <= W90GPU_DIS_MAX_BANDS
 * (each k's subspace problems run in one workgroup: in its LDS up to 64
 * bands, above that in an area of global memory per workgroup). */
#define W90GPU_DIS_MAX_BANDS 256

/* dis_control_type and dis_manifold_type (src/wannier90_types.F90:127-136,
 * src/types.F90:179-195): field meaning and defaults as there. */
typedef struct {
  double win_min, win_max;   /* outer window, eV (default: every band)   */
  double froz_min, froz_max; /* inner window, eV                          */
  int frozen_states;         /* 1 if dis_froz_max is given (default 0)   */
  int num_iter;              /* dis_num_iter    (200)                     */
  double mix_ratio;          /* dis_mix_ratio   (0.5), in (0, 1]          */
  double conv_tol;           /* dis_conv_tol    (1e-10)                   */
  int conv_window;           /* dis_conv_window (3)                       */
} w90gpu_dis_control;

typedef struct {
  double omega_invariant; /* final Omega_I of the optimal subspace (Ang^2) */
  double seconds;         /* wall time of the call on the device            */
  double sweeps_mean;     /* Jacobi sweeps per extraction eigenproblem      */
  int iterations;         /* dis_extract iterations performed               */
  int converged;          /* 1 if the dis_conv_window test was satisfied    */
  int status;             /* status of the call                             */
  int sweeps_max;         /* most Jacobi sweeps of one extraction eigenproblem */
} w90gpu_dis_result;

void w90gpu_dis_control_defaults(w90gpu_dis_control* ctl);

/* Disentangle one problem.  prob gives num_wann, num_kpts and the k-mesh
 * tables (bk is not used).  Inputs, Fortran layout: eigval (num_bands,
 * num_kpts) in eV, ascending per k; a_matrix (num_bands, num_wann,
 * num_kpts), the projections that the reference passes in u_matrix_opt;
 * m_matrix_orig (num_bands, num_bands, nntot, num_kpts).  Outputs:
 * u_matrix_opt (num_bands, num_wann, num_kpts), rows relative to each
 * k's window, zero below it; u_matrix (num_wann, num_wann, num_kpts), the
 * initial U of wannierise (internal_find_u); m_matrix (num_wann, num_wann,
 * nntot, num_kpts), the overlaps in the gauge of u_matrix_opt and
 * u_matrix (the input of w90gpu_wann_main); eigval_opt (num_bands,
 * num_kpts), the eigenvalues of the Hamiltonian in the optimal subspace
 * in rows 1..num_wann and the window's eigenvalues above; nfirstwin and
 * ndimwin (num_kpts), the 1-based first band and size of each window;
 * history (3, rows) per iteration Omega_I(i-1) (the estimate from the new
 * eigenvalues), Omega_I(i) and the fractional change, as on the DIS lines
 * of the .wout (may be NULL; rows beyond history_rows are not written).
 * The error messages of dis_windows, dis_project and dis_proj_froz are
 * kept. */
int w90gpu_disentangle(w90gpu_ctx* ctx, const w90gpu_problem* prob, int num_bands, const w90gpu_dis_control* ctl,
                       const double* eigval, const double* a_matrix, const double* m_matrix_orig,
                       double* u_matrix_opt, double* u_matrix, double* m_matrix, double* eigval_opt,
                       int* nfirstwin, int* ndimwin, double* history, int history_rows, w90gpu_dis_result* result);

/* ------------------------------------------------ rotation tile shapes
 *
 * No counterpart in the reference.  The M rotation has several kernels, by
 * tile shape, per num_wann class: "w16" (up to 16 bands), "fused32" or
 * "fused64" (17-32), "fused64", "direct64" or "fused48" (33-64, fused48 up
 * to 48), "gemm64" or "gemm32" (65-256).  A problem's kernel is fixed by (num_wann, nntot) when it is
 * admitted: a choice made earlier in the process (w90gpu_tune_set, or a
 * measurement of the online tuner), else the table measured on MI355X that
 * is compiled into the library, else the class default.  With the
 * environment variable W90GPU_TUNE=1 when a context or batch is created, a
 * key without a choice is measured at the start of its first solve: every
 * candidate is timed on the problem's own data, and one replaces the
 * default only if it is at least 3% faster.  W90GPU_TUNE_FILE=<path> reads
 * choices from, and appends measurements to, a text file.  Results depend
 * on the kernel only through rounding. */
/* Name of the kernel that a problem of this shape admitted now would use
 * (NUL-terminated, at most len bytes). */
int w90gpu_tune_get(int num_wann, int nntot, char* name, size_t len);
/* Fix the kernel of (num_wann, nntot) for problems admitted from now on;
 * name NULL or "" drops a fixed choice.  W90GPU_ERR_INPUT if the kernel
 * does not serve num_wann's class. */
int w90gpu_tune_set(int num_wann, int nntot, const char* name);

#ifdef __cplusplus
}
#endif

#endif /* W90GPU_H */
