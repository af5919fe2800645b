/*
 * qo.h -- CPU oracle for the QuaC Lindblad/Schrodinger hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * is compared against; it is never linked into, loaded by or called from the
 * product library (quac_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * It is a plain-C, FP64, single-process restatement of the reference's
 * algorithms (every function cites the reference file:line it follows):
 *   - operator-chain row rules           src/kron.c:161-207, 294-538, 540-635
 *   - superoperator assembly with PETSc  src/kron.c:638-1036, src/qsystem.c:593-676
 *     MatSetValue(ADD_VALUES) semantics  (sorted rows, duplicates summed in
 *                                        insertion order, explicit zeros kept)
 *   - CSR MatMult, classic RK4, RK3BS    src/qsystem.c:895-1098 (TS wrapper)
 *   - expectation values / trace         src/qvec_utilities.c:322-392, 1197-1281
 *   - gate application (U* x U superop   src/quantum_circuits.c:435-481, 851-1067,
 *     path and in-place path)            1771-2236, 2484-2739, 2903-2980
 *   - steady-state GMRES                 src/solver.c:33-220
 *
 * Pinning (what checks each function; see tests/):
 *   qo_construct       48 pristine PETSc-binary fixtures (bit-exact structure and
 *                      values) + dense brute-force Lindbladian for VEC/OTHER ops
 *   qo_spmv/qo_rk4     degree-4 Taylor polynomial of exp(hA) on tiny dense A,
 *                      T1-decay closed forms (1/e), trace conservation
 *   qo_rk3bs_trial     closed forms of one step on a linear ODE: the candidate is
 *                      (1 + z + z^2/2 + z^3/6) x and the embedded error
 *                      -(z^3 + z^4)/48 x (z = hA), on tiny dense A
 *   qo_rk3bs           T1-decay closed form (reference tolerance 1e-4)
 *   qo_rk4_td          analytic Rabi drive with a decaying envelope
 *                      (time_dep_tests.c:46-64,142-213), WF and DM forms; with
 *                      f = 1 equal to qo_rk4 on the summed matrix
 *   qo_expect_dm       dense Tr(O rho) brute force; rand_dm.dat known answers
 *   qo_expect_wf       dense <psi|O|psi> brute force; the |1,1> known answers of
 *                      qvec_utilities_tests.c:232-278
 *   qo_gate_*          textbook U rho U^dagger on dense matrices; the reference's
 *                      hard-coded rows of every gate in quantum_gates_test_sys.c
 *                      (sigma x/y/z on one and two qubits, H, CNOT, CXZ and CZ in
 *                      both orders); RX/RY/RZ/U1-3/EYE against exp(-i theta P/2)
 *                      and the OpenQASM U3 definition
 *   qo_gmres           residual ||L'x - e0|| and the analytic driven-damped
 *                      two-level steady state
 */
#ifndef QUAC_ORACLE_QO_H
#define QUAC_ORACLE_QO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* op_type values (src/operators_p.h:6-16) */
enum { QO_RAISE = -1, QO_NUMBER = 0, QO_LOWER = 1, QO_VEC = 2, QO_SIGMA_X = 3,
       QO_SIGMA_Y = 4, QO_SIGMA_Z = 5, QO_IDENTITY = 6, QO_OTHER = 7 };
/* mat_term_type values (src/qsystem_p.h:16-21) */
enum { QO_TD_LIN = -2, QO_TD_HAM = -1, QO_HAM = 1, QO_LIN = 2 };
/* tensor_control_enum (src/kron_p.h:8-12) */
enum { QO_TENSOR_IG = -1, QO_TENSOR_GG = 0, QO_TENSOR_GI = 1 };

/* error codes */
enum { QO_OK = 0, QO_ERR_ARG = -1, QO_ERR_PAULI = -2, QO_ERR_VEC = -3, QO_ERR_MEM = -4 };

/* One operator of a term: 4 int64 {type, levels, n_before, position}. */
typedef struct { int64_t type, levels, n_before, position; } qo_op;

typedef struct {
  int64_t n, nnz;
  int64_t *rowptr;   /* n+1 */
  int64_t *col;      /* nnz */
  double  *val;      /* 2*nnz interleaved re,im */
} qo_csr;

/* chain of ops applied to index i (kron.c:161-207). Returns j (-1: no entry). */
int qo_chain(int64_t tot, int64_t i, int nops, const qo_op *ops, int tc,
             double *val_re, double *val_im, int64_t *j);

/*
 * Assemble the system matrix exactly as construct_matrix (qsystem.c:593-676).
 * terms: kind[t], nops[t], off[t] (index of first op in ops[]), a[2t],a[2t+1].
 * Time-dependent terms (kind<0) enter with a=0 after every time-independent term
 * (qsystem.c:645-653).  dm: density-matrix equations; mcwf: MCWF solver.
 * conj_a: 1 = current kron.c:723 form (conj(a) on the G^T x I part);
 *         0 = the form the pristine fixtures were written with (plain a).
 * td_index >= 0: build instead the cached time-dependent matrix of term
 *         td_index with a=1 (qsystem.c:647-648, no -i scale).
 */
int qo_construct(int64_t tot, int dm, int mcwf, int conj_a, int nterms,
                 const int64_t *kind, const int64_t *nops, const int64_t *off,
                 const double *a, const int64_t *ops, int td_index, qo_csr **out);
int64_t qo_csr_rows(const qo_csr *m);
int64_t qo_csr_nnz(const qo_csr *m);
void qo_csr_export(const qo_csr *m, int64_t *rowptr, int64_t *col, double *val);
void qo_csr_free(qo_csr *m);

/* y = A x  (row-ordered accumulation, as MatMult_SeqAIJ) */
void qo_spmv(const qo_csr *m, const double *x, double *y);
/* classic RK4, fixed step: K1=Ax, K2=A(x+h/2K1), K3=A(x+h/2K2), K4=A(x+hK3) */
int qo_rk4(const qo_csr *m, double *x, double dt, int64_t nsteps, int nthreads);
/* time-dependent RHS (qsystem.c:1300-1330) is handled by qo_rk4_td */
typedef void (*qo_tdfunc)(double t, int k, double *re, double *im);
int qo_rk4_td(const qo_csr *m, int ntd, qo_csr *const *td, const double *td_a, int wf_scale,
              qo_tdfunc f, double *x, double t0, double dt, int64_t nsteps);
/* one Bogacki-Shampine 3(2) trial step: k1 (given when have_k1), k2, k3, k4,
 * candidate xn and the weighted RMS norm of the embedded error */
int qo_rk3bs_trial(const qo_csr *m, const double *x, double h, int have_k1, double *k1, double *k2,
                   double *k3, double *k4, double *xn, double atol, double rtol, double *err);
/* Bogacki-Shampine 3(2) adaptive, PETSc-TSAdaptBasic-style controller */
int qo_rk3bs(const qo_csr *m, double *x, double t0, double tmax, double dt0,
             int64_t steps_max, double atol, double rtol, double hmin, double hmax,
             int64_t *steps_taken, double *t_end);

/* Tr(ops * rho) for a DM vec(rho) (qvec_utilities.c:1197-1281) */
int qo_expect_dm(int64_t tot, const double *x, int nops, const qo_op *ops, double *re, double *im);
/* <psi| ops |psi> for a WF */
int qo_expect_wf(int64_t tot, const double *x, int nops, const qo_op *ops, double *re, double *im);

/* gate types (src/quantum_gates.h:11-33) */
enum { QO_CNOT = -1, QO_CXZ = -2, QO_CZ = -3, QO_HADAMARD = 1, QO_SIGMAX = 2, QO_SIGMAY = 3,
       QO_SIGMAZ = 4, QO_EYE = 5, QO_RX = 6, QO_RY = 7, QO_RZ = 8, QO_U1 = 9, QO_U2 = 10, QO_U3 = 11 };
/*
 * Superoperator gate path (_apply_gate_sys, quantum_circuits.c:435-481):
 * rho <- G rho with G = U* x U built row by row; qubit q -> subsystem q.
 * levels[nsub], nbefore[nsub] describe the subsystems; dm=1 -> DM superop.
 */
int qo_gate_superop(int nsub, const int64_t *levels, const int64_t *nbefore, int dm,
                    int gate, const int64_t *qubits, double theta, double phi, double lambda,
                    double *x);
/*
 * In-place path (_apply_gate2, quantum_circuits.c:696-772): qubit q -> stride
 * prod(dims[0..q-1]); for DM the column copy uses q + nsub and the conjugate
 * gate.  rz_conj_bug=1 reproduces RZ_gate_func_wf ignoring conj_gate.
 */
int qo_gate_inplace(int nsub, const int64_t *dims, int dm, int gate, const int64_t *qubits,
                    double theta, double phi, double lambda, int rz_conj_bug, double *x);

/*
 * Steady state (solver.c:33-220 semantics): A' = A with row 0 += 1 at columns
 * i*(N+1) (trace row), b = e0; restarted GMRES(restart), CGS2 orthogonalisation,
 * FP64, optional Jacobi (diagonal) right preconditioner; x holds the initial
 * guess on entry.  Returns iterations in *iters, final true residual norm in *res.
 */
int qo_gmres(const qo_csr *m, int64_t tot, double *x, int restart, double rtol_abs,
             int64_t maxit, int jacobi, int64_t *iters, double *res);

#ifdef __cplusplus
}
#endif
#endif
