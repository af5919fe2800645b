/*
 * quac_hip.h -- the thin C-ABI between the QuaC-compatible C host layer
 * (include/quac.h) and the hand-written gfx950 kernels of libquac.so.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a
 * negative QH_E* code; qh_last_error() returns a message for the last failure.
 * Device objects (qh_mat, qh_vec, qh_batch) live in HBM on the current device
 * and are driven on one HIP stream per process (qh_stream()).
 *
 * Each entry point names the reference interface whose work it replaces.
 */
#ifndef QUAC_HIP_H
#define QUAC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { QH_OK = 0, QH_EINVAL = -1, QH_ENOMEM = -2, QH_EHIP = -3, QH_ERANGE = -4, QH_ENODEV = -5 };

/* ------------------------------------------------------------------ runtime */
/* replaces the device side of QuaC_initialize/QuaC_finalize (src/quac.c:25-58, 89-141) */
int qh_init(int device);
int qh_finalize(void);
const char *qh_last_error(void);
int qh_device_count(void);
int qh_synchronize(void);
/* free and total device memory of the current device (after the library stream drains) */
int qh_mem_info(size_t *free_bytes, size_t *total_bytes);
/* allocate / free raw device memory (plumbing for the host layer and tests) */
int qh_dmalloc(void **p, size_t bytes);
int qh_dfree(void *p);
int qh_h2d(void *dst, const void *src, size_t bytes);
int qh_d2h(void *dst, const void *src, size_t bytes);

/* ----------------------------------------------------- superoperator assembly */
/*
 * A monomial operator for device assembly.  Every QuaC operator (and a VEC
 * pair |p1><p2|) maps a basis level to at most one level, so it is described
 * by a per-level table: next[l] (-1: no entry), prev[l] (inverse map), value.
 * stride = n_after = total_levels / (levels * n_before) (kron.c:315).
 */
typedef struct { int64_t stride; int32_t levels; int32_t tab; } qh_op;
typedef struct { int32_t next, prev; double re, im; } qh_level;
/* term kinds = the dispatch of _add_ops_to_mat (kron.c:953-993) */
enum { QH_HAM_DM = 0, QH_LIN_DM = 1, QH_HAM_ONLY = 2, QH_LIN_MCWF = 3 };
typedef struct { int32_t kind, nops, op0, zero_a; double a_re, a_im; } qh_term;
enum { QH_ASM_SCALE_MINUS_I = 1, QH_ASM_CONJ_A = 2, QH_ASM_FP32 = 4 };

typedef struct qh_mat qh_mat;
/*
 * Build the CSR of construct_matrix (qsystem.c:593-676; kron.c:638-1036) on the
 * device: rows in PETSc order with sorted columns, explicit zeros kept, every
 * diagonal stored, duplicates summed in insertion order (term, row, part).
 * dim = rows (N^2 for DM, N for WF/MCWF).  flags: QH_ASM_*.
 */
int qh_assemble(int64_t tot, int64_t dim, int nterms, const qh_term *terms, int nops, const qh_op *ops,
                int nlev, const qh_level *levels, int flags, qh_mat **out);
int qh_mat_destroy(qh_mat *m);
/* a device matrix from a host CSR (square; rowptr[0] = 0, 0 < nnz < 2^31;
 * strictly increasing columns in a row; complex128 values interleaved), checked
 * on the host before the upload -- PETSc's MatCreateSeqAIJWithArrays */
int qh_mat_from_csr(int64_t rows, const int64_t *rowptr, const int64_t *col, const double *val, qh_mat **out);
int qh_mat_info(const qh_mat *m, int64_t *rows, int64_t *nnz);
/* copy the CSR back (int64 like PetscInt; values complex128 interleaved) */
int qh_mat_export(const qh_mat *m, int64_t *rowptr, int64_t *col, double *val);
/* make / refresh the complex64 copy of the values used by the FP32 path */
int qh_mat_make_fp32(qh_mat *m);
/* block-diagonal concatenation of n matrices (batched parameter sweeps) */
int qh_mat_blockdiag(int n, qh_mat *const *mats, qh_mat **out);
/* select SpMV variant: lanes per row G in {1,2,4,8,16,32,64} of the CSR kernel,
 * 128..136 = slot-aligned sliced ELL (one wavefront per 64 rows; 2/4/8 slots per
 * pass, 131/132 with vector-loaded uniform values, 133/134 with 2/4 slots per pass
 * and the next pass's loads in flight, 135/136 = 129/131 with the streamed data
 * under the non-temporal cache policy); FP32 only: 140/141/142 = row-pair sliced
 * layout (one wavefront per 128 rows, two rows per lane, 16-byte loads; 2/4/2
 * slots per pass, 142 reading the slots that stay inside each 2048-row window
 * from an LDS copy; 143/144 = 140/142 with the streamed data under the
 * non-temporal cache policy); 0 = autotune */
int qh_mat_set_lanes(qh_mat *m, int lanes64, int lanes32);
int qh_mat_get_lanes(const qh_mat *m, int *lanes64, int *lanes32);
/* sliced layout statistics (built on demand): slices, slots, stored complex
 * values (of complex slots), stored real values (of real-or-imaginary slots),
 * and the metadata bytes */
int qh_mat_layout_info(qh_mat *m, int64_t *slices, int64_t *slots, int64_t *stored_vals, int64_t *stored_reals,
                       int64_t *meta_bytes);
/* FP32 row-pair sliced layout statistics (built on demand; zeros when it does
 * not apply): 128-row slices, slots, stored complex pairs, stored real pairs,
 * metadata bytes (slot records, per-slice offsets, separable-block tables),
 * stored single reals (slots whose two rows share each value), and the slots of
 * a separable diagonal (no stored values: g_i + h_j formed in the kernel) */
int qh_mat_pair_layout_info(qh_mat *m, int64_t *slices, int64_t *slots, int64_t *stored_pairs,
                            int64_t *stored_real_pairs, int64_t *meta_bytes, int64_t *stored_real_singles,
                            int64_t *sep_slots);

/* ------------------------------------------------------------------ vectors */
typedef struct qh_vec qh_vec;
enum { QH_FP64 = 64, QH_FP32 = 32 };
/* complex vector of n entries; FP32 vectors carry a compensation word */
int qh_vec_create(int64_t n, int prec, qh_vec **out);
int qh_vec_destroy(qh_vec *v);
int qh_vec_size(const qh_vec *v, int64_t *n, int *prec);
int qh_vec_set(qh_vec *v, const double *host_c128);        /* whole vector, complex128 */
int qh_vec_get(const qh_vec *v, double *host_c128);
int qh_vec_zero(qh_vec *v);
int qh_vec_add_entry(qh_vec *v, int64_t loc, double re, double im);  /* VecSetValue(ADD_VALUES) */
int qh_vec_get_entry(const qh_vec *v, int64_t loc, double *re, double *im);
int qh_vec_copy(const qh_vec *src, qh_vec *dst);           /* any precision to any */
int qh_vec_scale(qh_vec *v, double re, double im);
/* raw device pointer (double2* for FP64, float2* hi for FP32) */
void *qh_vec_data(qh_vec *v);

/* y = A x (MatMult, used by qsystem.c:937,954,1310) */
int qh_spmv(const qh_mat *A, const qh_vec *x, qh_vec *y);

/* --------------------------------------------------------------- stepping */
/*
 * Fixed-step classic RK4 on dx/dt = A x (time_step_sys with -ts_rk_type 4
 * -ts_adapt_type none, qsystem.c:895-1098).  Steps run as fused SpMV+stage
 * kernels captured in a HIP graph; x (FP64 or FP32) is advanced in place.
 * renorm_every > 0 (FP32 only): every that many steps the trace of the density
 * matrix (dimension tot) is computed in FP64 and rho is rescaled to unit trace.
 */
int qh_rk4(qh_mat *A, qh_vec *x, double dt, int64_t nsteps, int64_t tot, int renorm_every);
/* same steps launched eagerly with a HIP event pair around every stage kernel;
 * the mean stage-kernel duration is then returned by qh_last_stage_ms */
int qh_rk4_timed(qh_mat *A, qh_vec *x, double dt, int64_t nsteps, int64_t tot, int renorm_every);
/* the set-up of qh_rk4 with the same arguments, without stepping: SpMV variant
 * chosen (autotuned on a first call), stage buffers allocated, step graph
 * captured and instantiated (the TSSetUp of qsystem.c:937-1023) */
int qh_rk4_prepare(qh_mat *A, qh_vec *x, double dt, int64_t tot, int renorm_every);
/*
 * Time-dependent RHS (qsystem.c:1288-1330): y = A x + sum_k s_k A_k x, with the
 * entries of all the A_k merged into one CSR (built once and cached on A; a
 * dictionary of the distinct (term, value) pairs in place of per-entry values
 * where there are few).  A's part runs on its autotuned sliced layout followed
 * by one pass over the merged A_k, or, where A has none, fused with it in one
 * CSR kernel.  s_k = s_re_im[2k] + i s_re_im[2k+1] given by the host,
 * |s_k| <= 1e-10 skipped (qsystem.c:34,1320).  ntd <= 150.  FP64 or FP32
 * vectors (x, y one precision, distinct).
 */
int qh_spmv_td(const qh_mat *A, int ntd, qh_mat *const *Ak, const double *s_re_im,
               const qh_vec *x, qh_vec *y);
/* One classic RK4 step of size h of dx/dt = A x + sum_k s_k(t) A_k x, in place
 * (FP64, or the FP32 hi+lo state): per stage A's sliced stage kernel and the A_k
 * pass (or one fused CSR kernel); s4_re_im holds the coefficients of the stage
 * times t, t+h/2, t+h/2, t+h (4 x ntd complex).  The RHS evaluations of
 * time_step_sys with time-dependent terms (qsystem.c:946-950, 1288-1330); the
 * stage buffers are kept with the cached merged structure. */
int qh_rk4_td_step(qh_mat *A, int ntd, qh_mat *const *Ak, const double *s4_re_im, qh_vec *x, double h);
/* nsteps consecutive RK4 steps of size h (the run of fixed steps of
 * time_step_sys between monitor calls and events): s_re_im holds, for step j,
 * the coefficients of its four stage times t_j, t_j + h/2, t_j + h/2, t_j + h
 * (nsteps x 4 x ntd complex).  A generator of at most 800 rows (no variant
 * forced, dictionary form) runs the whole run in one workgroup with the state in
 * LDS; otherwise qh_rk4_td_step per step. */
int qh_rk4_td_steps(qh_mat *A, int ntd, qh_mat *const *Ak, const double *s_re_im, int64_t nsteps, qh_vec *x,
                    double h);
/* Switches of the two functions above (no counterpart in the reference; for
 * tuning and for tests that compare the forms): dict_max = the largest
 * (term, value) dictionary (0: every entry stores its term and value; default
 * 4096, at most 8192; applies from the next merged build), sliced = 1 runs A's
 * part on its sliced layout where it has one (default), 0 the fused CSR kernel;
 * lanes = how the A_k entries join a sliced stage: 0 = summed inside A's stage
 * kernel where its variant allows (128, 134, 136, 140, 142, 143), else as 1; 1 = a
 * pass of their own, one thread per row over the sliced (64-row, entry-major)
 * form of the dictionary entries; 4, 8, 16 = a pass with that many lanes per row
 * over their CSR (0 and 1 fall back to 4 or 8 lanes, by the rows' average
 * length, where the sliced form does not exist). */
int qh_td_config(int dict_max, int sliced, int lanes);
/* generic stage helpers for host-driven integrators (RK3BS, td RK4) */
int qh_axpy_to(qh_vec *y, const qh_vec *x, double c_re, double c_im, const qh_vec *k); /* y = x + c k */
int qh_lincomb(qh_vec *y, int n, const double *c_re_im, qh_vec *const *v);             /* y = sum c_i v_i */
/* weighted RMS error of RK3BS: sqrt(mean(|e|^2/(atol+rtol max(|x|,|xn|))^2)) */
int qh_rk3bs_err(const qh_vec *x, const qh_vec *xn, const qh_vec *k1, const qh_vec *k2,
                 const qh_vec *k3, const qh_vec *k4, double h, double atol, double rtol, double *err);
/* One Bogacki-Shampine 3(2) trial step of dx/dt = A x (TSRK3BS tableau, FSAL:
 * k1 = A x is taken as given when have_k1): writes k2..k4, the candidate xn and
 * the weighted RMS error estimate (atol + rtol max(|x|, |xn|)) into *err; y is
 * workspace.  The step of the default time_step_sys integrator
 * (src/qsystem.c:1003-1005, PETSc TSRK3BS); on the sliced layout its stages are
 * three fused launches. */
int qh_rk3bs_step(qh_mat *A, const qh_vec *x, qh_vec *xn, qh_vec *k1, qh_vec *k2, qh_vec *k3, qh_vec *k4,
                  qh_vec *y, double h, double atol, double rtol, int have_k1, double *err);
/* The same trial step of dx/dt = A x + sum_k s_k(t) A_k x (the default
 * integrator of time_step_sys with time-dependent terms, qsystem.c:1003-1005,
 * 1288-1330): s4_re_im holds the coefficients of the stage times t, t + h/2,
 * t + 3h/4, t + h (4 x ntd complex; the first unused when have_k1).  Where A has
 * a one-row sliced FP64 variant (128, 134, 136) the merged entries are summed in
 * the four stage launches of qh_rk3bs_step; otherwise four qh_spmv_td
 * evaluations and the combinations. */
int qh_rk3bs_td_step(qh_mat *A, int ntd, qh_mat *const *Ak, const double *s4_re_im, const qh_vec *x, qh_vec *xn,
                     qh_vec *k1, qh_vec *k2, qh_vec *k3, qh_vec *k4, qh_vec *y, double h, double atol, double rtol,
                     int have_k1, double *err);
/* The whole Bogacki-Shampine integration of dx/dt = A x from t0 to tmax for a
 * generator of at most 640 rows (FP64 x in place) in one workgroup, the step
 * size controller included (the time_step_sys loop without monitor, events or
 * time-dependent terms, qsystem.c:995-1005,1078): first trial step h0; adaptive
 * != 0: fac = 0.9 err^(-1/3) in [0.1, 10], next step in [hmin, hmax], accepted
 * when err <= 1 or h <= hmin; adaptive == 0: fixed steps of h0, the last cut at
 * tmax.  At most steps_max accepted steps; returns the final time, the accepted
 * steps and the next step size. */
int qh_rk3bs_small(qh_mat *A, qh_vec *x, double t0, double tmax, double h0, int64_t steps_max, double atol,
                   double rtol, double hmin, double hmax, int adaptive, double *t_out, int64_t *steps_out,
                   double *h_out);

/* ------------------------------------------------------------- observables */
/* Tr(ops rho) (qvec_utilities.c:1197-1281), FP64 accumulation on the device */
int qh_expect_dm(const qh_vec *x, int64_t tot, int nops, const qh_op *ops, int nlev,
                 const qh_level *levels, double *re, double *im);
int qh_expect_wf(const qh_vec *x, int64_t tot, int nops, const qh_op *ops, int nlev,
                 const qh_level *levels, double *re, double *im);
/*
 * qh_expect_dm without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and the
 * bit-identical result of qh_expect_dm; it calls no hipMalloc/hipFree and keeps
 * its op chain, level table and partial sums in a fixed 105.25 KiB device array
 * per GPU (held until the process ends).  Limits: nops <= 64, nlev <= 4096
 * (QH_ERANGE beyond).  Calls on one GPU share that array: calls that overlap
 * (several host threads at once) race.
 */
int qh_expect_dm_noalloc(const qh_vec *x, int64_t tot, int nops, const qh_op *ops, int nlev,
                         const qh_level *levels, double *re, double *im);
/* trace of a DM (get_trace_qvec, qvec_utilities.c:378-392) */
int qh_trace(const qh_vec *x, int64_t tot, double *re, double *im);
/* <n_s> for every subsystem s at once: Re rho_ii weighted by level (populations) */
int qh_populations(const qh_vec *x, int nsub, const int64_t *levels, const int64_t *strides, int dm,
                   double *pops);

/* ------------------------------------------------------------------- gates */
/*
 * rho <- U rho U^dagger in place (U on the row index, U* on the column index),
 * or psi <- U psi for a WF.  nq = 1 or 2; u = 2x2 or 4x4 complex128 row-major;
 * strides = n_after of the subsystems the gate acts on (first = most
 * significant bit of the gate basis).  Replaces _apply_gate_sys
 * (quantum_circuits.c:435-481) and _apply_gate2 (:696-772).
 */
int qh_gate(qh_vec *x, int64_t tot, int dm, int nq, const int64_t *strides, const double *u);

/* ------------------------------------------------------- batched sweeps (FP32) */
/*
 * A batch of independent instances of uneven size advanced together: one
 * block-diagonal CSR, one fused FP32 SpMV+stage kernel per RK stage over all
 * instances, hipGraph-captured stepping, per-instance FP64 trace renormalisation
 * and populations.  Replaces running time_step_sys once per instance.
 */
typedef struct qh_batch qh_batch;
int qh_batch_create(int n, qh_mat *const *mats, const int64_t *tot, int prec, qh_batch **out);
int qh_batch_destroy(qh_batch *b);
int qh_batch_set_state(qh_batch *b, int i, const double *host_c128);
int qh_batch_get_state(qh_batch *b, int i, double *host_c128);
int qh_batch_rk4(qh_batch *b, double dt, int64_t nsteps, int renorm_every);
/* populations of each instance: nsub[i] subsystems, levels/strides packed */
int qh_batch_populations(qh_batch *b, const int *nsub, const int64_t *levels, const int64_t *strides,
                         double *pops);
int qh_batch_info(const qh_batch *b, int64_t *rows, int64_t *nnz);
/* the batch's block-diagonal matrix, concatenated state, and instance offsets */
qh_mat *qh_batch_matrix(qh_batch *b);
qh_vec *qh_batch_state_vec(qh_batch *b);
int64_t qh_batch_offset(qh_batch *b, int i);
/* device-to-device copy of instance i's state from / into a vector of any
 * precision with the instance's row count (no host round trip) */
int qh_batch_set_state_vec(qh_batch *b, int i, const qh_vec *src);
int qh_batch_get_state_vec(qh_batch *b, int i, qh_vec *dst);
/* serial number of a device matrix: distinct for every matrix the process
 * creates (an address can be reused after a destroy, a serial number is not) */
uint64_t qh_mat_uid(const qh_mat *m);

/* ------------------------------------------------------------ steady state */
/*
 * GMRES(restart) on A' x = e0 with A' = A + e0 (trace row), solver.c:33-220;
 * FP64, CGS2 Gram-Schmidt, optional Jacobi right preconditioner; x holds the
 * initial guess.  Stops at ||b - A'x|| <= tol or maxit iterations.
 */
int qh_gmres(const qh_mat *A, int64_t tot, qh_vec *x, int restart, double tol, int64_t maxit,
             int jacobi, int64_t *iters, double *res);
/*
 * Hamiltonian-eigenbasis preconditioner of the steady-state solve (replaces the
 * role of PETSc's PCASM/ILU(0) at solver.c:182-183; no operator-level
 * counterpart in the reference).  heff is the N x N effective Hamiltonian
 * H - (i/2) sum_k a_k C_k^dag C_k, assembled by qh_assemble with QH_HAM_ONLY
 * (Hamiltonian) and QH_LIN_MCWF (Lindblad) terms and no flags.  M^-1 divides
 * the Liouvillian's no-jump part in the eigenbasis of the Hermitian part of heff.
 */
typedef struct qh_pc qh_pc;
int qh_pc_create_hbasis(const qh_mat *heff, qh_pc **out);
/* out = M^-1 in, both FP64 vectors of N*N entries (vec layout of an N x N matrix) */
int qh_pc_apply(const qh_pc *pc, const qh_vec *in, qh_vec *out);
int qh_pc_destroy(qh_pc *pc);
/* qh_gmres with a chosen right preconditioner: pc_kind 0 none, 1 Jacobi,
 * 2 the Hamiltonian-eigenbasis operator pc */
#define QH_PC_NONE 0
#define QH_PC_JACOBI 1
#define QH_PC_HBASIS 2
int qh_gmres_pc(const qh_mat *A, int64_t tot, qh_vec *x, int restart, double tol, int64_t maxit,
                int pc_kind, const qh_pc *pc, int64_t *iters, double *res);
/* qh_gmres_pc with the search space of every cycle augmented by u, a predicted
 * solution (augmented GMRES): the residual is minimised over x + M^-1 K_k + span(u),
 * so a cycle needs at most the iterations of the plain one.  The steady-state
 * solve seeded by the learned warm start (no counterpart in the reference, whose
 * solve starts from zero, solver.c:143-198). */
int qh_gmres_aug(const qh_mat *A, int64_t tot, qh_vec *x, const qh_vec *u, int restart, double tol, int64_t maxit,
                 int pc_kind, const qh_pc *pc, int64_t *iters, double *res);
/* eigenvalues w (ascending) and, if V != NULL, eigenvectors (column-major) of a
 * real symmetric n x n host matrix, on the device (rocSOLVER); used by the
 * density-matrix fidelity of get_fidelity_qvec (qvec_utilities.c:2339-2430) */
int qh_sym_eig(int64_t n, const double *A, double *w, double *V);

/* ---------------------------------------------------------- MCWF ensembles */
/*
 * Monte Carlo wavefunction trajectories (time_step_sys under use_mcwf_solver,
 * qsystem.c:1024-1079; norm test and jumps qsystem.c:1138-1235).  An ensemble
 * is one FP64 vector of M trajectories x N amplitudes, trajectory i at entries
 * [i*N, (i+1)*N).  qh_vec_view aliases a range of an FP64 vector (destroying the
 * view does not free the storage).
 */
int qh_vec_view(qh_vec *base, int64_t offset, int64_t n, qh_vec **out);
int qh_ens_broadcast(const qh_vec *init, qh_vec *ens, int64_t M);
/* nsteps classic RK4 steps of size dt of every trajectory under the N x N generator A */
int qh_ens_rk4(const qh_mat *A, qh_vec *ens, int64_t M, double dt, int64_t nsteps);
/* squared norm of each trajectory into norms2[M] (host) */
int qh_ens_norms(const qh_vec *ens, int64_t M, double *norms2);
/* mean over the trajectories of <psi|O|psi>/<psi|psi> (qvec_utilities.c:1353-1420) */
int qh_ens_expect(const qh_vec *ens, int64_t M, int64_t tot, int nops, const qh_op *ops, int nlev,
                  const qh_level *levels, double *re, double *im);
/* fits = 1 when the qh_mcwf_* calls below accept A and nj jump operators */
int qh_mcwf_fits(const qh_mat *A, int nj, int *fits);
/* Trajectories of at most 1024 amplitudes (rows of A with at most 16 entries):
 * every trajectory i with jstep[i] < J takes classic RK4 steps (size dt; step
 * J-1 has size hlast; the first step of a call is shortened by tsub[i], the time
 * already covered inside step jstep[i]) and stops at the first norm crossing:
 * when its squared norm falls below thr[i] during a step (the jump test,
 * qsystem.c:1138-1153), the crossing time inside the step is located (the norm
 * of the RK4 step as a polynomial of the step size, the event location of
 * qsystem.c:1017-1019) and the state is the RK4 step re-taken to that time.  On
 * return jstep[i] / tsub[i] give where trajectory i stands (tsub[i] = 0 unless it
 * stopped inside a step) and jump[i] = 1 if it stopped for a jump (host arrays of M). */
int qh_mcwf_advance(const qh_mat *A, qh_vec *ens, int64_t M, const double *thr, int32_t *jstep, double *tsub,
                    int32_t J, double dt, double hlast, int32_t *jump);
/* The norm crossing of one trajectory inside an RK4 step of size h from psi0
 * (any dimension): *tau = the first time in (0, h] at which the squared norm of
 * the RK4 step falls to thr (h when it stays above), out = that step re-taken to
 * *tau.  psi0 and out are distinct FP64 vectors (views of an ensemble allowed). */
int qh_mcwf_cross(const qh_mat *A, const qh_vec *psi0, qh_vec *out, double h, double thr, double *tau);
/* p[q*nj + k] = ||C_k psi_{idx[q]}||^2 for q < n, k < nj (C: nj <= 16 N x N
 * operators, N <= 1024; the jump probabilities of qsystem.c:1159-1200) */
int qh_mcwf_jump_norms(int nj, qh_mat *const *C, const qh_vec *ens, int64_t N, const int64_t *idx, int32_t n,
                       double *p);
/* psi_{idx[q]} <- (sre[q] + i sim[q]) C_{k[q]} psi_{idx[q]} for q < n, the idx[q]
 * distinct (the jump, qsystem.c:1201-1235) */
int qh_mcwf_jump_apply(int nj, qh_mat *const *C, qh_vec *ens, int64_t N, const int64_t *idx, int32_t n,
                       const int32_t *k, const double *sre, const double *sim);

/* --------------------------------------------------- warm-start network */
/*
 * The learned warm start of the steady-state solve: a fully connected network
 * x -> [Linear -> tanh] x (nlayers-1) -> Linear -> y in FP32 with hand-written
 * forward, backward and Adam kernels.  No counterpart in the reference (QuaC has
 * no learned component).  dims: nlayers+1 widths; parameters are one FP32 array
 * W_0 (dims[1] x dims[0], row-major), b_0, W_1, b_1, ...  x, t, y are device
 * arrays of batch x width, row-major.  Loss: mean over the batch of
 * sum_j 0.5 w_j (y_j - t_j)^2 with per-output weights w (default 1).
 */
typedef struct qh_mlp qh_mlp;
int qh_mlp_create(int nlayers, const int *dims, int max_batch, qh_mlp **out);
int qh_mlp_destroy(qh_mlp *net);
int qh_mlp_num_params(const qh_mlp *net, int64_t *n);
int qh_mlp_set_params(qh_mlp *net, const float *host);     /* also resets the Adam state */
int qh_mlp_get_params(const qh_mlp *net, float *host);
int qh_mlp_set_output_weights(qh_mlp *net, const float *host);
int qh_mlp_forward(qh_mlp *net, const float *x, int batch, float *y);
int qh_mlp_loss_grad(qh_mlp *net, const float *x, const float *t, int batch, float *loss, float *grad_host);
int qh_mlp_train_step(qh_mlp *net, const float *x, const float *t, int batch, float lr, float beta1, float beta2,
                      float eps, float *loss);

/* ------------------------------------------------------------- timing hooks */
/* average duration (ms) of the last qh_rk4 stage kernels, measured with events */
int qh_last_stage_ms(double *spmv_ms, int64_t *launches);
/* set the number of RK steps captured per graph launch (default 8) */
int qh_set_graph_steps(int steps);
/* mean duration (ms) of reps plain SpMV launches y = A x on the chosen precision */
int qh_bench_spmv(qh_mat *A, int prec, int reps, double *avg_ms);
/* achievable HBM bandwidth: reps one-pass streaming copies (bytes read + bytes
 * written, 16 bytes per thread) and reps read-only passes over a fresh buffer of
 * `bytes` (a multiple of 16, at least 4096); GB/s of each */
int qh_bench_stream(int64_t bytes, int reps, double *copy_gbs, double *read_gbs);
/* HIP event pair on the library stream around a caller's region; qh_timer_end
 * waits for the region to finish and returns its device time in ms */
int qh_timer_begin(void);
int qh_timer_end(double *ms);

#ifdef __cplusplus
}
#endif
#endif
