/*
 * quac_legacy.h -- the original global-operator API of QuaC (one implicit
 * system per process), kept for programs written before qsystem:
 * src/operators.h:43-69, src/dm_utilities.h:10-39, src/solver.h:8-15.
 *
 * Every call works on one process-wide qsystem that the first create_op /
 * create_vec creates; the matrix is assembled on the device the first time a
 * state is created or evolved, and again when terms were added since.  States
 * are the Vec handles made by create_full_dm / create_dm.  Differences from the
 * reference (documented with the implementation): add_lin_mat (a Lindblad term
 * from an arbitrary matrix) and the stiff solver are refused with an error, as
 * the reference's time_step refuses the stiff solver; get_dm_element stores the
 * element (the reference reads it but never writes *val).
 */
#ifndef QUAC_LEGACY_H
#define QUAC_LEGACY_H

#include "quac.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MAX_SUB 100                           /* operators.h:63 */
extern int nid;                               /* rank (0: one process per GPU) */
extern int np;                                /* number of processes (1) */
extern PetscInt total_levels;                 /* operators.c:26 */
extern int num_subsystems;                    /* operators.c:27 */
extern operator subsystem_list[MAX_SUB];      /* operators.h:65 */

/* operators.c */
void create_op(int, operator *);                                       /* :59-200   */
void create_vec(int, vec_op *);                                        /* :210-242  */
void no_lindblad_terms(void);                                          /* :278-280  */
void add_to_ham(PetscScalar, operator);                                /* :290-337  */
void add_to_ham_stiff(PetscScalar, operator);                          /* :348-392  */
void add_to_ham_time_dep(double (*)(double), int, ...);                /* :253-276  */
void add_to_ham_mult2(PetscScalar, operator, operator);                /* :405-488  */
void add_to_ham_stiff_mult2(PetscScalar, operator, operator);          /* :499-585  */
void add_to_ham_mult3(PetscScalar, operator, operator, operator);      /* :601-663  */
void add_lin(PetscScalar, operator);                                   /* :679-719  */
void add_lin_mult2(PetscScalar, operator, operator);                   /* :743-821  */
void add_lin_mat(PetscScalar, Mat);                                    /* :837-901  */
void print_dense_ham(void);                                            /* :39-48    */
void set_initial_pop(operator, double);                                /* :1386-1398 */
void combine_ops_to_mat(Mat *, PetscInt, int, ...);                    /* :923-994  */

/* dm_utilities.c */
void create_dm(Vec *, PetscInt);                                       /* :616-623  */
void create_full_dm(Vec *);                                            /* :635-655  */
void destroy_dm(Vec);                                                  /* :1029-1032 */
void set_dm_from_initial_pop(Vec);                                     /* :665-804  */
void get_dm_element(Vec, PetscInt, PetscInt, PetscScalar *);           /* :965-981  */
void add_value_to_dm(Vec, PetscInt, PetscInt, PetscScalar);            /* :1049-1061 */
void assemble_dm(Vec);                                                 /* :1073-1076 */
int get_num_populations(void);                                         /* :1127-1143 */
void get_populations(Vec, double **);                                  /* :1157-1325 */
void get_expectation_value(Vec, PetscScalar *, int, ...);              /* :1352-1469 */
void trace_dm(PetscScalar *, Vec);                                     /* :1766-1786 */
void partial_trace_over(Vec, Vec, int, ...);                           /* :150-260  */
void partial_trace_keep(Vec, Vec, int, ...);                           /* :468-603  */
void get_fidelity(Vec, Vec, double *);                                 /* :1618-1722 */
void get_bipartite_concurrence(Vec, double *);                         /* :1487-1602 */
void print_dm(Vec, int);                                               /* :13-25    */
void print_psi(Vec, int);                                              /* :120-132  */
void print_dm_sparse(Vec, int);                                        /* :32-43    */
void print_dm_sparse_to_file(Vec, int, char[]);                        /* :51-66    */
void print_mat_sparse_to_file(Mat, char[]);                            /* :71-92    */
void print_mat_sparse(Mat);                                            /* :98-113   */
void get_dm_element_local(Vec, PetscInt, PetscInt, PetscScalar *);     /* :999-1016 */
void measure_dm(Vec, operator);                                        /* :270-308  */
void partial_trace_over_one(Vec, Vec, PetscInt, PetscInt, PetscInt, PetscInt); /* :1078-1117 */

/* solver.c */
void steady_state(Vec);                                                /* :33-221   */
void time_step(Vec, PetscReal, PetscReal, PetscReal, PetscInt);        /* :236-514  */
void set_ts_monitor(PetscErrorCode (*)(TS, PetscInt, PetscReal, Vec, void *));            /* :526-529 */
void set_ts_monitor_ctx(PetscErrorCode (*)(TS, PetscInt, PetscReal, Vec, void *), void *); /* :540-543 */

/* quantum_gates.h:68-72: circuits of the legacy API (create_circuit and
 * destroy_circuit are in quac.h); gate arguments as add_gate_to_circuit_sys */
void add_gate_to_circuit(circuit *, PetscReal, gate_type, ...);       /* quantum_gates.c:651-714 */
void add_circuit_to_circuit(circuit *, circuit, PetscReal);           /* quantum_gates.c:716-749 */
void start_circuit_at_time(circuit *, PetscReal);                     /* quantum_gates.c:755-760 */

/* quac.c:119-141: free an operator (with its dag and n) / a VEC operator array */
void destroy_op(operator *);
void destroy_vec(vec_op *);

/* the process-wide system behind the calls above (NULL before the first create_op) */
qsystem quac_legacy_system(void);

#ifdef __cplusplus
}
#endif
#endif
