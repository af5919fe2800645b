/*
 * quac.h -- QuaC-compatible C API of libquac.so (drop-in for the qsystem path).
 *
 * Same function names, argument order and meaning as the reference headers
 * (src/qsystem.h, src/qvec_utilities.h, src/quantum_circuits.h,
 * src/quantum_gates.h, src/quac.h).  PETSc is not used: PetscScalar/PetscInt/
 * PetscReal keep their complex128/int64/double meaning, and Vec/Mat/TS are
 * handles to device objects (HBM-resident vectors and CSR matrices, and the
 * stepping context passed to monitors).  All entry points return void and, like
 * the reference, print "ERROR! ..." and exit() on misuse, unless
 * quac_set_error_mode(QUAC_ERRORS_RETURN) was called, in which case the call
 * returns early and quac_last_error() describes the failure.
 */
#ifndef QUAC_H_
#define QUAC_H_

#include <complex.h>
#include <stdint.h>

#include "quac_hip.h"

#ifdef __cplusplus
#error "quac.h is a C header (its 'operator' type name follows the reference C API)"
#endif

/* ------------------------------------------------------------------ types */
typedef int64_t PetscInt;           /* 64-bit indices, quick_install_instructions:9-16 */
typedef double PetscReal;
typedef double _Complex PetscScalar; /* complex build, quac.c:32-34 */
typedef int PetscErrorCode;
typedef enum { PETSC_FALSE = 0, PETSC_TRUE = 1 } PetscBool;
typedef struct qh_vec *Vec;         /* device vector (FP64 or FP32 hi+lo) */
typedef struct qh_mat *Mat;         /* device CSR */
typedef struct quac_ts *TS;         /* stepping context handed to monitors */

#ifndef PETSC_i
#define PETSC_i _Complex_I
#endif
#define PetscRealPart(a) creal(a)
#define PetscImaginaryPart(a) cimag(a)
#define PetscConjComplex(a) conj(a)
#define PetscAbsComplex(a) cabs(a)
#define PETSC_COMM_WORLD 0
#define PetscFunctionReturn(x) return (x)
int PetscPrintf(int comm, const char *fmt, ...);

/* op_type (src/operators_p.h:6-16) -- values are ABI */
typedef enum { RAISE = -1, NUMBER = 0, LOWER = 1, VEC = 2, SIGMA_X = 3, SIGMA_Y = 4, SIGMA_Z = 5,
               IDENTITY = 6, OTHER = 7 } op_type;
/* mat_term_type (src/qsystem_p.h:16-21) */
typedef enum { TD_LINDBLAD = -2, TD_HAM = -1, HAM = 1, LINDBLAD = 2 } mat_term_type;
/* qvec_type (src/qsystem_p.h:23-27) */
typedef enum { DENSITY_MATRIX = -1, WAVEFUNCTION = 1, WF_ENSEMBLE = 2 } qvec_type;
/* gate_type (src/quantum_gates.h:11-33) */
typedef enum { NULL_GATE = -1000, SQiSWP = -8, CZ_ARP = -7, CUSTOM2QGATE = -6, CZX = -5, CmZ = -4, CZ = -3,
               CXZ = -2, CNOT = -1, HADAMARD = 1, SIGMAX = 2, SIGMAY = 3, SIGMAZ = 4, EYE = 5, RX = 6, RY = 7,
               RZ = 8, U1 = 9, U2 = 10, U3 = 11, CUSTOM1QGATE = 12 } gate_type;

/* struct operator (src/operators.h:8-31) -- same fields */
typedef struct operator {
  double initial_pop;
  PetscInt initial_exc;
  PetscInt n_before;
  PetscInt my_levels;
  op_type my_op_type;
  PetscInt pos_in_sys_hspace;
  struct operator *dag, *n, *sig_x, *sig_y, *sig_z, *eye, *other;
  PetscInt position;                 /* VEC only */
  struct operator **vec_op_list;     /* vec[0] only */
  PetscScalar evals[2];
  PetscScalar evecs[2][2];
} *operator;
typedef operator *vec_op;

/* struct mat_term (src/qsystem_p.h:29-37) */
typedef struct {
  PetscInt num_ops;
  mat_term_type my_term_type;
  PetscScalar a;
  void *ctx;
  Mat mat_A;
  PetscScalar (*time_dep_func)(PetscReal, void *);
  operator *ops;
  /* FFI form of time_dep_func returning (re, im) through pointers (new) */
  void (*time_dep_func_ri)(PetscReal, void *, PetscReal *, PetscReal *);
} mat_term;

/* struct qvec (src/qsystem_p.h:53-59) */
typedef struct qvec {
  qvec_type my_type;
  PetscBool ens_spawned;
  PetscInt n, Istart, Iend, total_levels, *hspace_dims, ndims_hspace, n_ops, n_ensemble, ens_i;
  Vec data;
  Vec *ens_datas;
} *qvec;

struct quantum_gate_struct;
typedef void (*custom_gate_func_type)(PetscScalar *, PetscInt, PetscInt, void *);
/* struct quantum_gate_struct (src/quantum_gates.h:38-48); row-generator pointers omitted */
struct quantum_gate_struct {
  PetscReal time, run_time;
  gate_type my_gate_type;
  PetscInt *qubit_numbers, num_qubits;
  custom_gate_func_type custom_func;
  PetscReal theta, phi, lambda;
  void *gate_ctx;
};
struct gate_layer_struct {
  PetscReal time;
  PetscInt num_gates;
  struct quantum_gate_struct *gate_list;
};
/* struct circuit (src/qsystem_p.h:43-48) */
typedef struct circuit {
  PetscInt num_gates, gate_list_size, current_gate, num_layers, current_layer;
  PetscReal start_time;
  struct quantum_gate_struct *gate_list;
  struct gate_layer_struct *layer_list;
} circuit;
typedef struct { PetscScalar gate_data[4][4]; } custom_gate_data;

/* struct qsystem (src/qsystem_p.h:62-113): fields apps read keep their names */
typedef struct qsystem {
  PetscInt num_time_indep, num_time_dep;
  PetscInt alloc_time_indep, alloc_time_dep;
  PetscScalar min_time_dep;
  mat_term *time_indep, *time_dep;
  PetscInt hspace_frozen;
  PetscBool dm_equations, mcwf_solver, time_step_called;
  PetscInt mat_allocated;
  Mat mat_A;
  PetscInt num_subsystems, alloc_subsystems;
  operator *subsystem_list;
  PetscInt total_levels, dim;
  qvec solution_qvec;
  PetscInt np, nid, Istart, Iend, my_num;
  PetscInt num_tot_trajs, num_local_trajs, seed, ens_i;
  PetscReal current_time;
  PetscErrorCode (*ts_monitor)(TS, PetscInt, PetscReal, Vec, void *);
  void *ts_ctx;
  PetscInt num_circuits, circuit_list_size, current_circuit;
  circuit *circuit_list;
  /* MCWF state (after the fields above; apps do not read these): one threshold
   * per trajectory (qsystem.c:1085-1091), the jump operators C_k as N x N
   * matrices with their coefficients, the uniform generator's state */
  PetscReal *rand_number;
  Mat *mcwf_jump;
  PetscScalar *mcwf_jump_a;
  PetscInt mcwf_njump;
  unsigned long long mcwf_rng[4];
} *qsystem;

/* ------------------------------------------------------------ src/quac.h */
void QuaC_initialize(int argc, char **args);   /* quac.c:25-58  */
void QuaC_finalize(void);                      /* quac.c:89-141 */
void QuaC_clear(void);                         /* quac.c:65-82  */

/* --------------------------------------------------------- src/qsystem.h */
void initialize_system(qsystem *);                                    /* qsystem.c:16-65   */
void destroy_system(qsystem *);                                       /* qsystem.c:67-91   */
void create_op_sys(qsystem, PetscInt, operator *);                    /* qsystem.c:153-230 */
void destroy_op_sys(operator *);                                      /* qsystem.c:338-349 */
void create_vec_op_sys(qsystem, PetscInt, vec_op *);                  /* qsystem.c:265-306 */
void destroy_vec_op_sys(vec_op *);                                    /* qsystem.c:308-316 */
void use_mcwf_solver(qsystem, PetscInt, PetscInt);                    /* qsystem.c:358-403 */
void add_ham_term(qsystem, PetscScalar, PetscInt, ...);               /* qsystem.c:407-434 */
void add_lin_term(qsystem, PetscScalar, PetscInt, ...);               /* qsystem.c:436-462 */
void add_lin_term_list(qsystem, PetscScalar, PetscInt, operator *);   /* qsystem.c:464-489 */
void add_ham_term_time_dep(qsystem, PetscScalar, void *, PetscScalar (*)(PetscReal, void *), PetscInt, ...); /* :491-521 */
void add_lin_term_time_dep(qsystem, PetscScalar, void *, PetscScalar (*)(PetscReal, void *), PetscInt, ...); /* :523-552 */
void clear_mat_terms_sys(qsystem);                                    /* qsystem.c:557-591 */
void construct_matrix(qsystem);                                       /* qsystem.c:593-676 */
void construct_op_matrix_wf_list(qsystem, PetscScalar, Mat *, PetscInt, operator *); /* qsystem.c:836-863 */
void time_step_sys(qsystem, qvec, PetscReal, PetscReal, PetscReal, PetscInt);        /* qsystem.c:895-1098 */
void set_ts_monitor_sys(qsystem, PetscErrorCode (*)(TS, PetscInt, PetscReal, Vec, void *), void *); /* :875-879 */
void set_init_excited_op(operator, PetscInt);                         /* qsystem.c:1340-1349 */
void set_qvec_from_init_excited_op(qsystem, qvec);                    /* qvec_utilities.c */
/* list forms of the variadic term adders (new; add_lin_term_list exists upstream) */
void add_ham_term_list(qsystem, PetscScalar, PetscInt, operator *);
typedef void (*quac_tdfunc_ri)(PetscReal t, void *ctx, PetscReal *re, PetscReal *im);
void add_ham_term_time_dep_list(qsystem, PetscScalar, void *, quac_tdfunc_ri, PetscInt, operator *);
void add_lin_term_time_dep_list(qsystem, PetscScalar, void *, quac_tdfunc_ri, PetscInt, operator *);

/* ------------------------------------------------- src/qvec_utilities.h */
void create_qvec_sys(qsystem, qvec *);                  /* qvec_utilities.c:633-648 */
void create_dm_sys(qsystem, qvec *);                    /* :654-705 */
void create_wf_sys(qsystem, qvec *);                    /* :710-760 */
void create_wf_ensemble_sys(qsystem, qvec *);           /* :765-817 */
void create_arb_qvec(qvec *, PetscInt, qvec_type);
void destroy_qvec(qvec *);                              /* :1076-1090 */
void add_to_qvec(qvec, PetscScalar, ...);               /* :919-965 */
void add_to_qvec_loc(qvec, PetscScalar, PetscInt);      /* :879-891 */
void add_to_qvec_fock_op(PetscScalar, qvec, PetscInt, ...);                 /* :817-842 */
void add_to_qvec_fock_op_list(PetscScalar, qvec, PetscInt, operator *, PetscInt *);
void get_qvec_loc_fock_op_list(qvec, PetscInt *, PetscInt, operator *, PetscInt *); /* :996-1033 */
void assemble_qvec(qvec);                               /* :1047-1064 */
void get_expectation_value_qvec(qvec, PetscScalar *, PetscInt, ...);        /* :1133-1151 */
void get_expectation_value_qvec_list(qvec, PetscScalar *, PetscInt, operator *); /* :1176-1195 */
void get_trace_qvec(qvec, PetscScalar *);               /* :378-392 */
void get_dm_element_qvec(qvec, PetscInt, PetscInt, PetscScalar *);  /* :322-337 */
void get_wf_element_qvec(qvec, PetscInt, PetscScalar *);
void get_bitstring_probs(qvec, PetscInt *, PetscReal **, PetscReal **); /* :1682-1717 */
void read_qvec_dm_binary(qvec *, const char[]);         /* :132-166 */
void read_qvec_wf_binary(qvec *, const char[]);         /* :97-126 */
void copy_qvec(qvec, qvec);                              /* :31-50 (WF -> DM as |psi><psi|) */
void print_qvec(qvec);
void create_arb_qvec_dims(qvec *, PetscInt, PetscInt *, qvec_type);  /* :523-570 */
void copy_qvec_wf_to_dm(qvec, qvec);                     /* :57-90 */
void check_qvec_equal(qvec, qvec, PetscBool *);          /* :26-30 */
void ptrace_over_list_qvec(qvec, PetscInt, PetscInt *, qvec *);      /* :1591-1656 */
void get_fidelity_qvec(qvec, qvec, PetscReal *, PetscReal *);        /* :2159-2245 */
void get_superfidelity_qvec(qvec, qvec, PetscReal *);    /* :2102-2141 */
void get_hilbert_schmidt_dist_qvec(qvec, qvec, PetscReal *);         /* :1657-1680 */
void print_dm_qvec(qvec);                                /* :291-305 */
void print_wf_qvec(qvec);                                /* :397-412 */
void get_qvec_loc_fock_op(qvec, PetscInt *, PetscInt, ...);          /* :971-993 */
void get_dm_element_qvec_local(qvec, PetscInt, PetscInt, PetscScalar *);        /* :351-359 */
void get_wf_element_qvec_local(qvec, PetscInt, PetscScalar *);                  /* :416-431 */
void get_wf_element_ens_i_qvec_local(qvec, PetscInt, PetscInt, PetscScalar *);  /* :434-449 */
void combine_circuit_to_mat_sys(qsystem, Mat *, circuit);         /* quantum_circuits.c:257-314 */
void apply_op_to_qvec_mat(qvec, PetscScalar **, operator);         /* qvec_utilities.c:1553-1589 */
void qvec_mat_mult(Mat, qvec);                                      /* qvec_utilities.c:169-195 */
void add_to_wf_ens_loc(qvec, PetscInt, PetscScalar, PetscInt);      /* qvec_utilities.c:899-911 */
void print_wf_ens_i_qvec(qvec);                                     /* qvec_utilities.c:456-466 */
void get_qvec_local_idxs(qvec, PetscInt, PetscInt *);                           /* :1427-1438 */
void check_qvec_consistent(qvec, qvec);                                         /* :10-24 */
void change_qvec_dims(qvec, PetscInt, PetscInt *);                              /* :582-626 */
/* cross-entropy benchmark scores of bitstring distributions (:1720-1991) */
void get_linear_xeb_fidelity(qvec, qvec, PetscReal *, PetscReal *);
void get_linear_xeb_unbiased_fidelity(qvec, qvec, PetscReal *, PetscReal *);
void get_log_xeb_fidelity(qvec, qvec, PetscReal *, PetscReal *);
void get_hog_score(qvec, qvec, PetscReal *, PetscReal *);
void get_hog_score_fidelity(qvec, qvec, PetscReal *, PetscReal *);
void get_linear_xeb_fidelity_probs(PetscReal *, PetscReal *, PetscInt, PetscReal *, PetscReal *, PetscInt, PetscInt,
                                   PetscReal *, PetscReal *);
void get_linear_xeb_unbiased_fidelity_probs(PetscReal *, PetscReal *, PetscInt, PetscReal *, PetscReal *, PetscInt,
                                            PetscInt, PetscReal *, PetscReal *);
void get_log_xeb_fidelity_probs(PetscReal *, PetscReal *, PetscInt, PetscReal *, PetscReal *, PetscInt, PetscInt,
                                PetscReal *, PetscReal *);
void get_hog_score_probs(PetscReal *, PetscReal *, PetscInt, PetscReal *, PetscReal *, PetscInt, PetscInt,
                         PetscReal *, PetscReal *);
void get_hog_score_fidelity_probs(PetscReal *, PetscReal *, PetscInt, PetscReal *, PetscReal *, PetscInt, PetscInt,
                                  PetscReal *, PetscReal *);
void print_qvec_file(qvec, const char[]);               /* :203-211 */
void print_dm_qvec_file(qvec, const char[]);            /* :213-241 */
void print_wf_qvec_file(qvec, const char[]);            /* :246-268 */
/* PETSc binary writers (MatView / VecView on a binary viewer, the fixture
 * format of tests/t_helpers.c:7-13); return 0 on success */
int quac_write_mat_petsc_binary(Mat, const char[]);
int quac_write_qvec_petsc_binary(qvec, const char[]);

/* ------------------------------------------ src/quantum_circuits.h/gates.h */
void create_circuit(circuit *, PetscInt);                       /* quantum_gates.c:610-628 */
void destroy_circuit(circuit *);                                /* quantum_gates.c:630-649 */
void add_gate_to_circuit_sys(circuit *, PetscReal, gate_type, ...); /* quantum_circuits.c:172-255 */
void schedule_circuit_layers(qsystem, circuit *);               /* quantum_circuits.c:316-404 */
void apply_circuit_to_sys(qsystem, circuit *, PetscReal);       /* quantum_circuits.c:407-431 */
void apply_circuit_to_qvec(qsystem, circuit, qvec);             /* quantum_circuits.c:143-159 */
void apply_circuit_to_qvec2(circuit, qvec);                      /* quantum_circuits.c:123-140 */
void apply_gate2(qvec, gate_type, ...);                         /* quantum_circuits.c:776-847 */
void apply_single_qb_measurement_error_probs(PetscReal *, PetscInt, PetscReal, PetscReal, PetscInt); /* :620-644 */
void apply_projective_measurement_qubit(qvec, PetscScalar *, PetscInt, ...);                       /* :484-502 */
void apply_projective_measurement_qubit_list(qvec, PetscScalar *, PetscInt, operator *);           /* :504-582 */
void apply_single_qb_measurement_error_state(qvec, PetscReal, PetscReal, PetscInt);                 /* :646-688 */

/* ------------------------------------------------------------ src/solver.h */
/* steady state of a qsystem (solver.c:33-220 semantics; the reference has it
 * only for the legacy global API).  x receives rho_ss. */
void steady_state_sys(qsystem, qvec);
/* explicit solver settings: GMRES(restart) to ||b - A'x|| <= tol or maxit
 * iterations with the right preconditioner precond (QUAC_SS_PC_*); x holds the
 * initial guess on entry */
enum { QUAC_SS_PC_NONE = 0, QUAC_SS_PC_JACOBI = 1, QUAC_SS_PC_HBASIS = 2 };
/* the N x N effective Hamiltonian H - (i/2) sum a C^dag C of the time-independent
 * terms as a device CSR (destroy with qh_mat_destroy); returns 0 on success */
int quac_assemble_heff(qsystem, Mat *out);
void steady_state_sys_opts(qsystem, qvec, PetscInt restart, PetscReal tol, PetscInt maxit, PetscInt precond,
                           PetscInt *iters, PetscReal *residual);
/* steady_state_sys_opts with the GMRES search space augmented by u, a predicted
 * steady state (the learned warm start; no counterpart in the reference): each
 * restart cycle minimises the residual over its Krylov space plus span(u). */
void steady_state_sys_aug(qsystem, qvec x, qvec u, PetscInt restart, PetscReal tol, PetscInt maxit, PetscInt precond,
                          PetscInt *iters, PetscReal *residual);

/* ------------------------------------------------------ new in this library */
enum { QUAC_ERRORS_EXIT = 0, QUAC_ERRORS_RETURN = 1 };
void quac_set_error_mode(int mode);
const char *quac_last_error(void);   /* "" when the last call succeeded */
void quac_clear_error(void);
/* options normally given as PETSc command-line flags: -ts_rk_type {3bs,4},
 * -ts_adapt_type {basic,none}, -quac_precision {fp64,fp32}, -quac_batch_precision {fp32,fp64}, -quac_renorm N,
 * -quac_kron_conj_a {0,1}, -ksp_rtol X, -ksp_gmres_restart N, -quac_jacobi {0,1},
 * -quac_small_path {1,0} (0: generators of at most 640 rows take the host-driven
 * RK3BS loop instead of the one-workgroup integration) */
void quac_set_option(const char *name, const char *value);
const char *quac_get_option(const char *name);
/* last time_step_sys or time_step_sys_batch: steps taken and final time */
void quac_ts_stats(PetscInt *steps, PetscReal *t_final);
/* Batched parameter sweeps (no counterpart in the reference, which steps one
 * qsystem per time_step_sys call, qsystem.c:895-1098): n constructed
 * density-matrix systems without time-dependent terms, circuits or MCWF,
 * advanced together by fixed-step classic RK4 from init_time to time_max in
 * steps of dt (the last one shortened to land on time_max; at most steps_max),
 * states[i] in place.  One block-diagonal device matrix and one fused stage
 * kernel per RK stage over every instance, replayed as a HIP graph; the batch is
 * kept for the next call with the same matrices.  -quac_batch_precision
 * {fp32 (default), fp64}; FP32 rescales every instance to unit trace every
 * -quac_renorm steps (0: never).  Monitors are not called. */
void time_step_sys_batch(PetscInt n, qsystem *systems, qvec *states, PetscReal init_time, PetscReal time_max,
                         PetscReal dt, PetscInt steps_max);
/* frees the batch kept by time_step_sys_batch (QuaC_finalize does too) */
void time_step_sys_batch_release(void);
/* TS accessors usable from monitors */
void TSGetTime(TS, PetscReal *);
void TSGetStepNumber(TS, PetscInt *);
/* non-variadic forms of the variadic gate calls (for FFI callers) */
void add_gate_to_circuit_sys_list(circuit *, PetscReal time, gate_type, const PetscInt *qubits, PetscReal theta,
                                  PetscReal phi, PetscReal lambda, void *gate_ctx);
void apply_gate2_list(qvec, gate_type, const PetscInt *qubits, PetscReal theta, PetscReal phi, PetscReal lambda,
                      void *gate_ctx);
void apply_gate_sys_list(qsystem, qvec, gate_type, const PetscInt *qubits, PetscReal theta, PetscReal phi,
                         PetscReal lambda, void *gate_ctx);
/* accessors for FFI callers that do not mirror the structs */
Mat quac_system_matrix(qsystem);
Mat quac_system_td_matrix(qsystem, PetscInt k);
PetscInt quac_system_total_levels(qsystem);
PetscInt quac_system_dim(qsystem);
PetscBool quac_system_dm(qsystem);
Vec quac_qvec_data(qvec);
void quac_qvec_info(qvec, PetscInt *n, PetscInt *total_levels, qvec_type *type);

#endif
