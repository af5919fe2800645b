"""quac_amd -- Python mirror of the QuaC-compatible C API of libquac.so.

Every call goes through the C-ABI of ``quac_amd/libquac.so`` (include/quac.h and
include/quac_hip.h); nothing is computed in Python.  The library must have been
built (``python -m quac_amd.build``); importing raises if it is missing.

The class and method names follow the reference API (src/qsystem.h,
src/qvec_utilities.h, src/quantum_circuits.h):

    qs = System()                      # initialize_system
    q, c = qs.create_op(2), qs.create_op(10)
    qs.add_ham_term(1.0, q.n)          # add_ham_term(qsys, 1.0, 1, q->n)
    qs.add_lin_term(1e-2, q)
    qs.construct_matrix()
    rho = qs.create_qvec(); rho.add(1.0, 10, 10)
    qs.time_step(rho, 0.0, 1.0, 0.01, 1000)
    rho.expectation(q.n)
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libquac.so")

RAISE, NUMBER, LOWER, VEC, SIGMA_X, SIGMA_Y, SIGMA_Z, IDENTITY, OTHER = -1, 0, 1, 2, 3, 4, 5, 6, 7
DENSITY_MATRIX, WAVEFUNCTION = -1, 1
GATES = {"CNOT": -1, "CXZ": -2, "CZ": -3, "CmZ": -4, "CZX": -5, "CUSTOM2Q": -6, "CZ_ARP": -7, "SQiSWP": -8,
         "H": 1, "HADAMARD": 1, "SIGMAX": 2, "X": 2, "SIGMAY": 3, "Y": 3, "SIGMAZ": 4, "Z": 4, "EYE": 5, "RX": 6,
         "RY": 7, "RZ": 8, "U1": 9, "U2": 10, "U3": 11, "CUSTOM1Q": 12}
FP64, FP32 = 64, 32


class QuacError(RuntimeError):
    pass


class Cx(ctypes.Structure):
    """PetscScalar (double _Complex) by value: two doubles in two SSE registers."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]

    @classmethod
    def of(cls, z):
        z = complex(z)
        return cls(z.real, z.imag)

    def value(self):
        return complex(self.re, self.im)


class OperatorStruct(ctypes.Structure):
    pass


OpPtr = ctypes.POINTER(OperatorStruct)
OperatorStruct._fields_ = [
    ("initial_pop", ctypes.c_double), ("initial_exc", ctypes.c_int64), ("n_before", ctypes.c_int64),
    ("my_levels", ctypes.c_int64), ("my_op_type", ctypes.c_int), ("pos_in_sys_hspace", ctypes.c_int64),
    ("dag", OpPtr), ("n", OpPtr), ("sig_x", OpPtr), ("sig_y", OpPtr), ("sig_z", OpPtr), ("eye", OpPtr),
    ("other", OpPtr), ("position", ctypes.c_int64), ("vec_op_list", ctypes.POINTER(OpPtr)),
    ("evals", Cx * 2), ("evecs", Cx * 4)]


class QhOp(ctypes.Structure):
    """qh_op of include/quac_hip.h: one operator of a monomial chain"""
    _fields_ = [("stride", ctypes.c_int64), ("levels", ctypes.c_int32), ("tab", ctypes.c_int32)]


class QhLevel(ctypes.Structure):
    """qh_level of include/quac_hip.h: next/prev level and value of one table entry"""
    _fields_ = [("next", ctypes.c_int32), ("prev", ctypes.c_int32), ("re", ctypes.c_double), ("im", ctypes.c_double)]


class GateStruct(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("run_time", ctypes.c_double), ("my_gate_type", ctypes.c_int),
                ("qubit_numbers", ctypes.POINTER(ctypes.c_int64)), ("num_qubits", ctypes.c_int64),
                ("custom_func", ctypes.c_void_p), ("theta", ctypes.c_double), ("phi", ctypes.c_double),
                ("lambda_", ctypes.c_double), ("gate_ctx", ctypes.c_void_p)]


class LayerStruct(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("num_gates", ctypes.c_int64), ("gate_list", ctypes.POINTER(GateStruct))]


class CircuitStruct(ctypes.Structure):
    _fields_ = [("num_gates", ctypes.c_int64), ("gate_list_size", ctypes.c_int64), ("current_gate", ctypes.c_int64),
                ("num_layers", ctypes.c_int64), ("current_layer", ctypes.c_int64), ("start_time", ctypes.c_double),
                ("gate_list", ctypes.POINTER(GateStruct)), ("layer_list", ctypes.POINTER(LayerStruct))]


MONITOR = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p,
                           ctypes.c_void_p)
TDFUNC_RI = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double))

_lib = None


def _setup(L):
    P, vp, i64, dbl, i32 = ctypes.POINTER, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
    flt = ctypes.c_float

    def sig(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    # host API (include/quac.h)
    sig("QuaC_initialize", None, i32, P(ctypes.c_char_p))
    sig("QuaC_finalize", None)
    sig("quac_set_error_mode", None, i32)
    sig("quac_last_error", ctypes.c_char_p)
    sig("quac_clear_error", None)
    sig("quac_set_option", None, ctypes.c_char_p, ctypes.c_char_p)
    sig("quac_get_option", ctypes.c_char_p, ctypes.c_char_p)
    sig("initialize_system", None, P(vp))
    sig("destroy_system", None, P(vp))
    sig("create_op_sys", None, vp, i64, P(OpPtr))
    sig("destroy_op_sys", None, P(OpPtr))
    sig("create_vec_op_sys", None, vp, i64, P(P(OpPtr)))
    sig("destroy_vec_op_sys", None, P(P(OpPtr)))
    sig("use_mcwf_solver", None, vp, i64, i64)
    sig("add_ham_term_list", None, vp, Cx, i64, P(OpPtr))
    sig("add_lin_term_list", None, vp, Cx, i64, P(OpPtr))
    sig("clear_mat_terms_sys", None, vp)
    sig("construct_matrix", None, vp)
    sig("construct_op_matrix_wf_list", None, vp, Cx, P(vp), i64, P(OpPtr))
    sig("time_step_sys", None, vp, vp, dbl, dbl, dbl, i64)
    sig("time_step_sys_batch", None, i64, P(vp), P(vp), dbl, dbl, dbl, i64)
    sig("time_step_sys_batch_release", None)
    sig("qh_batch_set_state_vec", i32, vp, i32, vp)
    sig("qh_batch_get_state_vec", i32, vp, i32, vp)
    sig("qh_mat_uid", ctypes.c_uint64, vp)
    sig("set_ts_monitor_sys", None, vp, MONITOR, vp)
    sig("quac_ts_stats", None, P(i64), P(dbl))
    sig("create_qvec_sys", None, vp, P(vp))
    sig("create_dm_sys", None, vp, P(vp))
    sig("create_wf_sys", None, vp, P(vp))
    sig("create_arb_qvec", None, P(vp), i64, i32)
    sig("destroy_qvec", None, P(vp))
    sig("add_to_qvec_loc", None, vp, Cx, i64)
    sig("assemble_qvec", None, vp)
    sig("get_expectation_value_qvec_list", None, vp, P(Cx), i64, P(OpPtr))
    sig("get_trace_qvec", None, vp, P(Cx))
    sig("get_dm_element_qvec", None, vp, i64, i64, P(Cx))
    sig("get_bitstring_probs", None, vp, P(i64), P(P(dbl)), P(P(dbl)))
    sig("read_qvec_dm_binary", None, P(vp), ctypes.c_char_p)
    sig("read_qvec_wf_binary", None, P(vp), ctypes.c_char_p)
    sig("quac_write_mat_petsc_binary", i32, vp, ctypes.c_char_p)
    sig("quac_write_qvec_petsc_binary", i32, vp, ctypes.c_char_p)
    sig("print_qvec_file", None, vp, ctypes.c_char_p)
    sig("set_init_excited_op", None, OpPtr, i64)
    sig("set_qvec_from_init_excited_op", None, vp, vp)
    sig("create_circuit", None, P(CircuitStruct), i64)
    sig("destroy_circuit", None, P(CircuitStruct))
    sig("add_gate_to_circuit_sys_list", None, P(CircuitStruct), dbl, i32, P(i64), dbl, dbl, dbl, vp)
    sig("schedule_circuit_layers", None, vp, P(CircuitStruct))
    sig("apply_circuit_to_sys", None, vp, P(CircuitStruct), dbl)
    sig("apply_circuit_to_qvec", None, vp, CircuitStruct, vp)
    sig("apply_gate2_list", None, vp, i32, P(i64), dbl, dbl, dbl, vp)
    sig("apply_gate_sys_list", None, vp, vp, i32, P(i64), dbl, dbl, dbl, vp)
    sig("steady_state_sys_opts", None, vp, vp, i64, dbl, i64, i64, P(i64), P(dbl))
    sig("steady_state_sys_aug", None, vp, vp, vp, i64, dbl, i64, i64, P(i64), P(dbl))
    sig("quac_system_matrix", vp, vp)
    sig("quac_system_td_matrix", vp, vp, i64)
    sig("quac_system_total_levels", i64, vp)
    sig("quac_system_dim", i64, vp)
    sig("quac_qvec_data", vp, vp)
    sig("quac_qvec_info", None, vp, P(i64), P(i64), P(i32))
    sig("add_ham_term_time_dep_list", None, vp, Cx, vp, TDFUNC_RI, i64, P(OpPtr))
    sig("add_lin_term_time_dep_list", None, vp, Cx, vp, TDFUNC_RI, i64, P(OpPtr))
    # device C-ABI (include/quac_hip.h)
    sig("qh_init", i32, i32)
    sig("qh_last_error", ctypes.c_char_p)
    sig("qh_device_count", i32)
    sig("qh_synchronize", i32)
    sig("qh_mat_info", i32, vp, P(i64), P(i64))
    sig("qh_mat_export", i32, vp, P(i64), P(i64), P(dbl))
    sig("qh_mat_destroy", i32, vp)
    sig("qh_mat_make_fp32", i32, vp)
    sig("qh_mat_set_lanes", i32, vp, i32, i32)
    sig("qh_mat_get_lanes", i32, vp, P(i32), P(i32))
    sig("qh_mat_layout_info", i32, vp, P(i64), P(i64), P(i64), P(i64), P(i64))
    sig("qh_mat_pair_layout_info", i32, vp, P(i64), P(i64), P(i64), P(i64), P(i64), P(i64), P(i64))
    sig("qh_vec_create", i32, i64, i32, P(vp))
    sig("qh_vec_destroy", i32, vp)
    sig("qh_vec_set", i32, vp, P(dbl))
    sig("qh_vec_get", i32, vp, P(dbl))
    sig("qh_vec_copy", i32, vp, vp)
    sig("qh_vec_size", i32, vp, P(i64), P(i32))
    sig("qh_spmv", i32, vp, vp, vp)
    sig("qh_rk3bs_step", i32, vp, vp, vp, vp, vp, vp, vp, vp, dbl, dbl, dbl, i32, P(dbl))
    sig("qh_rk4", i32, vp, vp, dbl, i64, i64, i32)
    sig("qh_rk4_prepare", i32, vp, vp, dbl, i64, i32)
    sig("create_arb_qvec_dims", None, P(vp), i64, P(i64), i32)
    sig("destroy_qvec", None, P(vp))
    sig("copy_qvec", None, vp, vp)
    sig("copy_qvec_wf_to_dm", None, vp, vp)
    sig("check_qvec_equal", None, vp, vp, P(i32))
    sig("ptrace_over_list_qvec", None, vp, i64, P(i64), P(vp))
    sig("get_fidelity_qvec", None, vp, vp, P(dbl), P(dbl))
    sig("get_superfidelity_qvec", None, vp, vp, P(dbl))
    sig("get_hilbert_schmidt_dist_qvec", None, vp, vp, P(dbl))
    sig("qvec_mat_mult", None, vp, vp)
    sig("apply_op_to_qvec_mat", None, vp, P(P(Cx)), vp)
    sig("combine_circuit_to_mat_sys", None, vp, P(vp), CircuitStruct)
    sig("add_to_wf_ens_loc", None, vp, i64, Cx, i64)
    sig("get_wf_element_ens_i_qvec_local", None, vp, i64, i64, P(Cx))
    sig("qh_ens_rk4", i32, vp, vp, i64, dbl, i64)
    sig("qh_ens_norms", i32, vp, i64, P(dbl))
    sig("qh_mcwf_fits", i32, vp, i32, P(i32))
    sig("qh_mcwf_advance", i32, vp, vp, i64, P(dbl), P(i32), P(dbl), i32, dbl, dbl, P(i32))
    sig("qh_mcwf_jump_norms", i32, i32, P(vp), vp, i64, P(i64), i32, P(dbl))
    sig("qh_mcwf_jump_apply", i32, i32, P(vp), vp, i64, P(i64), i32, P(i32), P(dbl), P(dbl))
    sig("qh_rk4_timed", i32, vp, vp, dbl, i64, i64, i32)
    sig("qh_last_stage_ms", i32, P(dbl), P(i64))
    sig("qh_set_graph_steps", i32, i32)
    sig("qh_bench_spmv", i32, vp, i32, i32, P(dbl))
    sig("qh_bench_stream", i32, i64, i32, P(dbl), P(dbl))
    sig("qh_mem_info", i32, P(ctypes.c_size_t), P(ctypes.c_size_t))
    sig("qh_td_config", i32, i32, i32, i32)
    sig("qh_pc_create_hbasis", i32, vp, P(vp))
    sig("qh_pc_apply", i32, vp, vp, vp)
    sig("qh_pc_destroy", i32, vp)
    sig("qh_gmres_pc", i32, vp, i64, vp, i32, dbl, i64, i32, vp, P(i64), P(dbl))
    sig("quac_assemble_heff", i32, vp, P(vp))
    sig("qh_mlp_create", i32, i32, P(i32), i32, P(vp))
    sig("qh_mlp_destroy", i32, vp)
    sig("qh_mlp_num_params", i32, vp, P(i64))
    sig("qh_mlp_set_params", i32, vp, P(flt))
    sig("qh_mlp_get_params", i32, vp, P(flt))
    sig("qh_mlp_set_output_weights", i32, vp, P(flt))
    sig("qh_mlp_forward", i32, vp, vp, i32, vp)
    sig("qh_mlp_loss_grad", i32, vp, vp, vp, i32, P(flt), P(flt))
    sig("qh_mlp_train_step", i32, vp, vp, vp, i32, flt, flt, flt, flt, P(flt))
    sig("qh_timer_begin", i32)
    sig("qh_timer_end", i32, P(dbl))
    sig("qh_trace", i32, vp, i64, P(dbl), P(dbl))
    sig("qh_expect_dm", i32, vp, i64, i32, P(QhOp), i32, P(QhLevel), P(dbl), P(dbl))
    sig("qh_expect_wf", i32, vp, i64, i32, P(QhOp), i32, P(QhLevel), P(dbl), P(dbl))
    sig("qh_expect_dm_noalloc", i32, vp, i64, i32, P(QhOp), i32, P(QhLevel), P(dbl), P(dbl))
    sig("qh_populations", i32, vp, i32, P(i64), P(i64), i32, P(dbl))
    sig("qh_gate", i32, vp, i64, i32, i32, P(i64), P(dbl))
    sig("qh_gmres", i32, vp, i64, vp, i32, dbl, i64, i32, P(i64), P(dbl))
    sig("qh_batch_create", i32, i32, P(vp), P(i64), i32, P(vp))
    sig("qh_batch_destroy", i32, vp)
    sig("qh_batch_set_state", i32, vp, i32, P(dbl))
    sig("qh_batch_get_state", i32, vp, i32, P(dbl))
    sig("qh_batch_rk4", i32, vp, dbl, i64, i32)
    sig("qh_batch_populations", i32, vp, P(i32), P(i64), P(i64), P(dbl))
    sig("qh_batch_info", i32, vp, P(i64), P(i64))
    sig("qh_batch_matrix", vp, vp)
    sig("qh_batch_state_vec", vp, vp)


def lib():
    """Load libquac.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise QuacError(f"{LIBPATH} is missing: build it with `python -m quac_amd.build`")
        L = ctypes.CDLL(LIBPATH)
        _setup(L)
        L.quac_set_error_mode(1)  # QUAC_ERRORS_RETURN: failures raise QuacError here
        _lib = L
    return _lib


def _check():
    L = lib()
    msg = L.quac_last_error()
    if msg:
        L.quac_clear_error()
        raise QuacError(msg.decode())


def _qh(rc):
    if rc != 0:
        raise QuacError(lib().qh_last_error().decode())


def device_count():
    return lib().qh_device_count()


def init(args=()):
    """QuaC_initialize(argc, argv) with PETSc-style options, e.g. ("-ts_rk_type", "4")."""
    argv = [b"quac"] + [str(a).encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    L = lib()
    L.quac_clear_error()
    L.QuaC_initialize(len(argv), arr)
    _check()


def set_option(name, value):
    lib().quac_set_option(name.encode(), str(value).encode())


def get_option(name):
    v = lib().quac_get_option(name.encode())
    return None if v is None else v.decode()


def synchronize():
    _qh(lib().qh_synchronize())


class Op:
    """A reference-counted view of a ``struct operator`` owned by the library."""

    def __init__(self, ptr):
        self.ptr = ptr

    def _link(self, name):
        p = getattr(self.ptr.contents, name)
        return Op(p) if p else None

    dag = property(lambda s: s._link("dag"))
    n = property(lambda s: s._link("n"))
    eye = property(lambda s: s._link("eye"))
    sig_x = property(lambda s: s._link("sig_x"))
    sig_y = property(lambda s: s._link("sig_y"))
    sig_z = property(lambda s: s._link("sig_z"))
    other = property(lambda s: s._link("other"))
    levels = property(lambda s: s.ptr.contents.my_levels)
    n_before = property(lambda s: s.ptr.contents.n_before)
    op_type = property(lambda s: s.ptr.contents.my_op_type)
    position = property(lambda s: s.ptr.contents.position)


def _oparray(ops):
    arr = (OpPtr * max(1, len(ops)))()
    for i, o in enumerate(ops):
        arr[i] = o.ptr
    return arr


class Matrix:
    """Borrowed handle to a device CSR (qh_mat)."""

    def __init__(self, handle, owned=False):
        self.h = handle
        self.owned = owned

    def info(self):
        r, z = ctypes.c_int64(), ctypes.c_int64()
        _qh(lib().qh_mat_info(self.h, ctypes.byref(r), ctypes.byref(z)))
        return r.value, z.value

    def write_petsc_binary(self, path):
        rc = lib().quac_write_mat_petsc_binary(self.h, str(path).encode())
        _check()
        if rc:
            raise QuacError(lib().quac_last_error().decode())

    def export(self):
        """(rowptr int64, col int64, val complex128) copied from the device."""
        rows, nnz = self.info()
        rp = np.empty(rows + 1, np.int64)
        col = np.empty(max(nnz, 1), np.int64)
        val = np.empty(2 * max(nnz, 1), np.float64)
        _qh(lib().qh_mat_export(self.h, rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                col.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                val.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return rp, col[:nnz], val.view(np.complex128)[:nnz]

    def set_lanes(self, l64, l32):
        _qh(lib().qh_mat_set_lanes(self.h, l64, l32))

    def lanes(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _qh(lib().qh_mat_get_lanes(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def layout_info(self):
        """sliced layout: dict(slices, slots, stored_vals (complex), stored_reals, meta_bytes)"""
        v = [ctypes.c_int64() for _ in range(5)]
        _qh(lib().qh_mat_layout_info(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("slices", "slots", "stored_vals", "stored_reals", "meta_bytes"), [x.value for x in v]))

    def pair_layout_info(self):
        """FP32 row-pair sliced layout: dict(slices, slots, stored_pairs, stored_real_pairs, meta_bytes,
        stored_real_singles, sep_slots)"""
        v = [ctypes.c_int64() for _ in range(7)]
        _qh(lib().qh_mat_pair_layout_info(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("slices", "slots", "stored_pairs", "stored_real_pairs", "meta_bytes", "stored_real_singles",
                         "sep_slots"), [x.value for x in v]))

    def __del__(self):
        if getattr(self, "owned", False) and self.h and _lib is not None:
            _lib.qh_mat_destroy(self.h)
            self.h = None


class DeviceVec:
    """A qh_vec (complex128 or complex64 hi+lo) in HBM."""

    def __init__(self, n=None, prec=FP64, handle=None, owned=True):
        if handle is None:
            h = ctypes.c_void_p()
            _qh(lib().qh_vec_create(n, prec, ctypes.byref(h)))
            handle = h
        self.h = handle if isinstance(handle, ctypes.c_void_p) else ctypes.c_void_p(handle)
        self.owned = owned
        nn, pp = ctypes.c_int64(), ctypes.c_int()
        _qh(lib().qh_vec_size(self.h, ctypes.byref(nn), ctypes.byref(pp)))
        self.n, self.prec = nn.value, pp.value

    def set(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        assert x.size == self.n
        _qh(lib().qh_vec_set(self.h, x.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def get(self):
        out = np.empty(self.n, np.complex128)
        _qh(lib().qh_vec_get(self.h, out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def __del__(self):
        if getattr(self, "owned", False) and self.h and _lib is not None:
            _lib.qh_vec_destroy(self.h)
            self.h = None


class QVec:
    """qvec: a density matrix or wavefunction (src/qsystem_p.h:53-59)."""

    def __init__(self, handle, system=None):
        self.h = handle
        self.system = system
        n, tl, ty = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        lib().quac_qvec_info(self.h, ctypes.byref(n), ctypes.byref(tl), ctypes.byref(ty))
        self.n, self.total_levels, self.type = n.value, tl.value, ty.value

    @property
    def is_dm(self):
        return self.type == DENSITY_MATRIX

    def write_petsc_binary(self, path):
        rc = lib().quac_write_qvec_petsc_binary(self.h, str(path).encode())
        _check()
        if rc:
            raise QuacError(lib().quac_last_error().decode())

    def print_file(self, path):
        lib().print_qvec_file(self.h, str(path).encode())
        _check()

    def data(self):
        return DeviceVec(handle=lib().quac_qvec_data(self.h), owned=False)

    def add_loc(self, alpha, loc):
        lib().add_to_qvec_loc(self.h, Cx.of(alpha), loc)
        _check()

    def add(self, alpha, i, j=None):
        """add_to_qvec: DM location N*i + j (qvec_utilities.c:919-946); WF location i."""
        loc = self.total_levels * i + j if self.is_dm else i
        self.add_loc(alpha, loc)

    def assemble(self):
        lib().assemble_qvec(self.h)

    def get(self):
        return self.data().get()

    def set(self, x):
        self.data().set(x)

    def expectation(self, *ops):
        out = Cx()
        lib().get_expectation_value_qvec_list(self.h, ctypes.byref(out), len(ops), _oparray(ops))
        _check()
        return out.value()

    def trace(self):
        out = Cx()
        lib().get_trace_qvec(self.h, ctypes.byref(out))
        _check()
        return out.value()

    def element(self, row, col):
        out = Cx()
        lib().get_dm_element_qvec(self.h, row, col, ctypes.byref(out))
        _check()
        return out.value()

    def bitstring_probs(self):
        n = ctypes.c_int64()
        p = ctypes.POINTER(ctypes.c_double)()
        v = ctypes.POINTER(ctypes.c_double)()
        lib().get_bitstring_probs(self.h, ctypes.byref(n), ctypes.byref(p), ctypes.byref(v))
        _check()
        out = np.array([p[i] for i in range(n.value)])
        libc = ctypes.CDLL(None)
        libc.free(p)
        libc.free(v)
        return out

    def populations(self, levels, strides):
        out = np.zeros(len(levels))
        lv = np.array(levels, np.int64)
        st = np.array(strides, np.int64)
        _qh(lib().qh_populations(ctypes.c_void_p(lib().quac_qvec_data(self.h)), len(levels),
                                 lv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(self.is_dm),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def apply_gate2(self, gate, qubits, theta=0.0, phi=0.0, lam=0.0, ctx=None):
        q = (ctypes.c_int64 * len(qubits))(*qubits)
        lib().apply_gate2_list(self.h, GATES[gate] if isinstance(gate, str) else gate, q, theta, phi, lam, ctx)
        _check()

    def destroy(self):
        if self.h:
            h = ctypes.c_void_p(self.h) if not isinstance(self.h, ctypes.c_void_p) else self.h
            lib().destroy_qvec(ctypes.byref(h))
            self.h = None


def read_qvec_dm_binary(path):
    h = ctypes.c_void_p()
    lib().read_qvec_dm_binary(ctypes.byref(h), path.encode())
    _check()
    return QVec(h)


def read_qvec_wf_binary(path):
    h = ctypes.c_void_p()
    lib().read_qvec_wf_binary(ctypes.byref(h), path.encode())
    _check()
    return QVec(h)


class Circuit:
    """circuit (src/qsystem_p.h:43-48) with add_gate / schedule / apply."""

    def __init__(self, num_gates):
        self.c = CircuitStruct()
        lib().create_circuit(ctypes.byref(self.c), num_gates)

    def add_gate(self, time, gate, qubits, theta=0.0, phi=0.0, lam=0.0, ctx=None, run_time=None):
        q = (ctypes.c_int64 * len(qubits))(*qubits)
        lib().add_gate_to_circuit_sys_list(ctypes.byref(self.c), time, GATES[gate] if isinstance(gate, str) else gate,
                                           q, theta, phi, lam, ctx)
        _check()
        if run_time is not None:
            self.c.gate_list[self.c.num_gates - 1].run_time = run_time

    @property
    def num_layers(self):
        return self.c.num_layers

    def layer_times(self):
        return [self.c.layer_list[i].time for i in range(self.c.num_layers)]

    def destroy(self):
        if self.c.gate_list:
            lib().destroy_circuit(ctypes.byref(self.c))


class System:
    """qsystem (src/qsystem_p.h:62-113)."""

    def __init__(self):
        h = ctypes.c_void_p()
        L = lib()
        L.quac_clear_error()
        L.initialize_system(ctypes.byref(h))
        _check()
        self.h = h
        self.ops = []
        self._keep = []

    def create_op(self, levels):
        p = OpPtr()
        lib().create_op_sys(self.h, levels, ctypes.byref(p))
        _check()
        op = Op(p)
        self.ops.append(op)
        return op

    def create_vec_op(self, levels):
        arr = ctypes.POINTER(OpPtr)()
        lib().create_vec_op_sys(self.h, levels, ctypes.byref(arr))
        _check()
        self._keep.append(arr)
        return [Op(arr[i]) for i in range(levels)]

    def add_ham_term(self, a, *ops):
        lib().add_ham_term_list(self.h, Cx.of(a), len(ops), _oparray(ops))
        _check()

    def add_lin_term(self, a, *ops):
        lib().add_lin_term_list(self.h, Cx.of(a), len(ops), _oparray(ops))
        _check()

    def _td_cb(self, func):
        def _f(t, ctx, re, im):
            z = complex(func(t))
            re[0] = z.real
            im[0] = z.imag
        cb = TDFUNC_RI(_f)
        self._keep.append(cb)
        return cb

    def add_ham_term_time_dep(self, a, func, *ops):
        """add_ham_term_time_dep with func(t) -> complex (kept alive by the system)."""
        lib().add_ham_term_time_dep_list(self.h, Cx.of(a), None, self._td_cb(func), len(ops), _oparray(ops))
        _check()

    def add_lin_term_time_dep(self, a, func, *ops):
        lib().add_lin_term_time_dep_list(self.h, Cx.of(a), None, self._td_cb(func), len(ops), _oparray(ops))
        _check()

    def use_mcwf_solver(self, ntraj, seed):
        lib().use_mcwf_solver(self.h, ntraj, seed)
        _check()

    def construct_matrix(self):
        lib().construct_matrix(self.h)
        _check()

    def matrix(self):
        h = lib().quac_system_matrix(self.h)
        if not h:
            raise QuacError("matrix not constructed")
        return Matrix(ctypes.c_void_p(h))

    def td_matrix(self, k):
        return Matrix(ctypes.c_void_p(lib().quac_system_td_matrix(self.h, k)))

    def op_matrix_wf(self, a, *ops):
        h = ctypes.c_void_p()
        lib().construct_op_matrix_wf_list(self.h, Cx.of(a), ctypes.byref(h), len(ops), _oparray(ops))
        _check()
        return Matrix(h, owned=True)

    @property
    def total_levels(self):
        return lib().quac_system_total_levels(self.h)

    @property
    def dim(self):
        return lib().quac_system_dim(self.h)

    def create_qvec(self):
        h = ctypes.c_void_p()
        lib().create_qvec_sys(self.h, ctypes.byref(h))
        _check()
        return QVec(h, self)

    def create_dm(self):
        h = ctypes.c_void_p()
        lib().create_dm_sys(self.h, ctypes.byref(h))
        _check()
        return QVec(h, self)

    def create_wf(self):
        h = ctypes.c_void_p()
        lib().create_wf_sys(self.h, ctypes.byref(h))
        _check()
        return QVec(h, self)

    def set_monitor(self, fn):
        """fn(step, t, DeviceVec) -> None, called each step (TSMonitor semantics)."""
        if fn is None:
            cb = ctypes.cast(None, MONITOR)
        else:
            def _cb(ts, step, t, vec, ctx):
                try:
                    fn(step, t, DeviceVec(handle=vec, owned=False))
                    return 0
                except Exception:  # noqa: BLE001 - the monitor's failure is reported as an error code
                    return 1
            cb = MONITOR(_cb)
        self._monitor = cb
        lib().set_ts_monitor_sys(self.h, cb, None)

    def time_step(self, x, t0, tmax, dt, steps_max):
        lib().time_step_sys(self.h, x.h, t0, tmax, dt, steps_max)
        _check()
        s, t = ctypes.c_int64(), ctypes.c_double()
        lib().quac_ts_stats(ctypes.byref(s), ctypes.byref(t))
        return s.value, t.value

    def steady_state(self, x, restart=100, tol=1e-11, maxit=10000, precond=None, jacobi=None, aug=None):
        """GMRES steady state into x (x is the initial guess); precond: "hbasis"
        (default), "jacobi" or "none"; jacobi=True/False is the older spelling;
        aug: a QVec holding a predicted steady state that augments every cycle's
        search space (steady_state_sys_aug)"""
        if precond is None:
            precond = "hbasis" if jacobi is None else ("jacobi" if jacobi else "none")
        pc = {"none": 0, "jacobi": 1, "hbasis": 2}[precond]
        it, res = ctypes.c_int64(), ctypes.c_double()
        if aug is None:
            lib().steady_state_sys_opts(self.h, x.h, restart, tol, maxit, pc, ctypes.byref(it), ctypes.byref(res))
        else:
            lib().steady_state_sys_aug(self.h, x.h, aug.h, restart, tol, maxit, pc, ctypes.byref(it),
                                       ctypes.byref(res))
        _check()
        return it.value, res.value

    def apply_circuit(self, circ, start_time):
        lib().apply_circuit_to_sys(self.h, ctypes.byref(circ.c), start_time)
        _check()

    def schedule(self, circ):
        lib().schedule_circuit_layers(self.h, ctypes.byref(circ.c))
        _check()

    def apply_circuit_to_qvec(self, circ, x):
        lib().apply_circuit_to_qvec(self.h, circ.c, x.h)
        _check()

    def apply_gate(self, x, gate, qubits, theta=0.0, phi=0.0, lam=0.0, ctx=None):
        """one gate with the superoperator convention (qubit q = subsystem q)"""
        q = (ctypes.c_int64 * len(qubits))(*qubits)
        lib().apply_gate_sys_list(self.h, x.h, GATES[gate] if isinstance(gate, str) else gate, q, theta, phi, lam,
                                  ctx)
        _check()

    def clear_terms(self):
        lib().clear_mat_terms_sys(self.h)
        _check()

    def destroy(self):
        if self.h:
            lib().destroy_system(ctypes.byref(self.h))
            self.h = None


class Batch:
    """Many independent density-matrix instances advanced together (qh_batch)."""

    def __init__(self, matrices, tots, prec=FP32):
        n = len(matrices)
        arr = (ctypes.c_void_p * n)(*[m.h.value if isinstance(m.h, ctypes.c_void_p) else m.h for m in matrices])
        t = (ctypes.c_int64 * n)(*tots)
        h = ctypes.c_void_p()
        _qh(lib().qh_batch_create(n, arr, t, prec, ctypes.byref(h)))
        self.h = h
        self.n = n
        self.tots = list(tots)

    def set_state(self, i, x):
        x = np.ascontiguousarray(x, np.complex128)
        _qh(lib().qh_batch_set_state(self.h, i, x.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def get_state(self, i):
        out = np.empty(self.tots[i] ** 2, np.complex128)
        _qh(lib().qh_batch_get_state(self.h, i, out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def rk4(self, dt, nsteps, renorm_every=0):
        _qh(lib().qh_batch_rk4(self.h, dt, nsteps, renorm_every))

    def populations(self, subs):
        """subs[i] = list of (levels, stride) of instance i -> list of arrays"""
        nsub = (ctypes.c_int * self.n)(*[len(s) for s in subs])
        lv = np.array([L for s in subs for (L, _) in s], np.int64)
        st = np.array([S for s in subs for (_, S) in s], np.int64)
        out = np.zeros(len(lv))
        _qh(lib().qh_batch_populations(self.h, nsub, lv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        res, k = [], 0
        for s in subs:
            res.append(out[k:k + len(s)])
            k += len(s)
        return res

    def info(self):
        r, z = ctypes.c_int64(), ctypes.c_int64()
        _qh(lib().qh_batch_info(self.h, ctypes.byref(r), ctypes.byref(z)))
        return r.value, z.value

    def matrix(self):
        return Matrix(ctypes.c_void_p(lib().qh_batch_matrix(self.h)))

    def state(self):
        return DeviceVec(handle=lib().qh_batch_state_vec(self.h), owned=False)

    def destroy(self):
        if self.h:
            _qh(lib().qh_batch_destroy(self.h))
            self.h = None


def time_step_batch(systems, states, t0, tmax, dt, steps_max):
    """time_step_sys_batch: the systems advanced together (fixed-step RK4), each
    state in place; returns (steps, final time) of quac_ts_stats"""
    n = len(systems)
    hs = (ctypes.c_void_p * n)(*[s.h for s in systems])
    xs = (ctypes.c_void_p * n)(*[x.h for x in states])
    lib().time_step_sys_batch(n, hs, xs, t0, tmax, dt, steps_max)
    _check()
    st, t = ctypes.c_int64(), ctypes.c_double()
    lib().quac_ts_stats(ctypes.byref(st), ctypes.byref(t))
    return st.value, t.value


def rk4_device(mat, vec, dt, nsteps, tot=0, renorm_every=0, timed=False):
    """qh_rk4 / qh_rk4_timed on raw device objects; returns avg stage-kernel ms if timed."""
    f = lib().qh_rk4_timed if timed else lib().qh_rk4
    _qh(f(mat.h, vec.h, dt, nsteps, tot, renorm_every))
    if timed:
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _qh(lib().qh_last_stage_ms(ctypes.byref(ms), ctypes.byref(n)))
        return ms.value
    return None


def rk4_prepare(mat, vec, dt, tot=0, renorm_every=0):
    """qh_rk4_prepare: autotune, stage buffers and step-graph capture of a later
    rk4_device call with the same arguments"""
    _qh(lib().qh_rk4_prepare(mat.h, vec.h, dt, tot, renorm_every))


def set_graph_steps(n):
    _qh(lib().qh_set_graph_steps(n))


def bench_spmv(mat, prec, reps=20):
    """mean ms of a plain SpMV launch on the matrix (events on the library stream)"""
    ms = ctypes.c_double()
    _qh(lib().qh_bench_spmv(mat.h, prec, reps, ctypes.byref(ms)))
    return ms.value


def mem_info():
    """(free, total) device bytes"""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _qh(lib().qh_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def td_config(dict_max=4096, sliced=1, lanes=0):
    """switches of the time-dependent RHS (qh_td_config): the (term, value)
    dictionary limit (0 = none), A's part on its sliced layout, lanes per row of
    the A_k pass (0 = auto)"""
    _qh(lib().qh_td_config(dict_max, sliced, lanes))


def bench_stream(nbytes=1 << 30, reps=10):
    """(copy GB/s, read GB/s) of one-pass streaming kernels over nbytes buffers"""
    c, r = ctypes.c_double(), ctypes.c_double()
    _qh(lib().qh_bench_stream(nbytes, reps, ctypes.byref(c), ctypes.byref(r)))
    return c.value, r.value


class StreamTimer:
    """HIP event pair on the library stream around a region (device ms in .ms)"""

    def __enter__(self):
        _qh(lib().qh_timer_begin())
        return self

    def __exit__(self, *exc):
        if exc[0] is None:
            ms = ctypes.c_double()
            _qh(lib().qh_timer_end(ctypes.byref(ms)))
            self.ms = ms.value
        return False


class HBasisPC:
    """Hamiltonian-eigenbasis preconditioner of a system's steady-state solve"""

    def __init__(self, system):
        m = ctypes.c_void_p()
        rc = lib().quac_assemble_heff(system.h, ctypes.byref(m))
        _check()
        if rc:
            raise QuacError(lib().quac_last_error().decode())
        try:
            pc = ctypes.c_void_p()
            _qh(lib().qh_pc_create_hbasis(m, ctypes.byref(pc)))
            self.h = pc
        finally:
            lib().qh_mat_destroy(m)

    def apply(self, vin, vout):
        _qh(lib().qh_pc_apply(self.h, vin.h, vout.h))

    def destroy(self):
        if self.h:
            lib().qh_pc_destroy(self.h)
            self.h = None


class DeviceF32:
    """raw FP32 device array (plumbing for the warm-start network)"""

    def __init__(self, n):
        self.n = int(n)
        p = ctypes.c_void_p()
        _qh(lib().qh_dmalloc(ctypes.byref(p), max(4 * self.n, 4)))
        self.h = p

    def set(self, a):
        a = np.ascontiguousarray(a, np.float32).reshape(-1)
        assert a.size == self.n
        _qh(lib().qh_h2d(self.h, a.ctypes.data_as(ctypes.c_void_p), 4 * self.n))

    def get(self):
        out = np.empty(self.n, np.float32)
        _qh(lib().qh_d2h(out.ctypes.data_as(ctypes.c_void_p), self.h, 4 * self.n))
        return out

    def __del__(self):
        try:
            if self.h:
                lib().qh_dfree(self.h)
                self.h = None
        except Exception:
            pass


class MLP:
    """the warm-start network on the device: x -> [Linear -> tanh]* -> Linear"""

    def __init__(self, dims, max_batch):
        self.dims = [int(d) for d in dims]
        arr = (ctypes.c_int * len(self.dims))(*self.dims)
        h = ctypes.c_void_p()
        _qh(lib().qh_mlp_create(len(self.dims) - 1, arr, int(max_batch), ctypes.byref(h)))
        self.h = h
        n = ctypes.c_int64()
        _qh(lib().qh_mlp_num_params(h, ctypes.byref(n)))
        self.nparam = n.value
        self.max_batch = int(max_batch)

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparam
        _qh(lib().qh_mlp_set_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def get_params(self):
        p = np.empty(self.nparam, np.float32)
        _qh(lib().qh_mlp_get_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return p

    def set_output_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.dims[-1]
        _qh(lib().qh_mlp_set_output_weights(self.h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def forward(self, x_dev, batch, y_dev):
        _qh(lib().qh_mlp_forward(self.h, x_dev.h, int(batch), y_dev.h))

    def loss_grad(self, x_dev, t_dev, batch):
        loss = ctypes.c_float()
        g = np.empty(self.nparam, np.float32)
        _qh(lib().qh_mlp_loss_grad(self.h, x_dev.h, t_dev.h, int(batch), ctypes.byref(loss),
                                   g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return loss.value, g

    def train_step(self, x_dev, t_dev, batch, lr, beta1=0.9, beta2=0.999, eps=1e-8, want_loss=False):
        loss = ctypes.c_float()
        _qh(lib().qh_mlp_train_step(self.h, x_dev.h, t_dev.h, int(batch), lr, beta1, beta2, eps,
                                    ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def destroy(self):
        if self.h:
            lib().qh_mlp_destroy(self.h)
            self.h = None
