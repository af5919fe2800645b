"""Edge cases of the device path: zero steps, empty and minimal systems,
single-instance batches, mismatched sizes and bad gate strides (errors are
reported, nothing runs), the smallest MCWF ensemble, and the bandwidth
calibration's argument checks and ordinary vector kernels on a size that is
not a multiple of the block."""
import ctypes

import numpy as np
import pytest

import oracle
import quac_amd as qa
import recipes
from qbuild import build_system, populations_np

pytestmark = pytest.mark.gpu


def _fixed_rk4():
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")


def test_zero_steps_leave_the_state_unchanged():
    _fixed_rk4()
    r = recipes.spin_chain_recipe(5, seed=1)
    N, loc = recipes.spin_chain_init(5)
    qs, _, _ = build_system(r)
    x = qs.create_qvec()
    x.add_loc(1.0, loc)
    before = x.get().copy()
    steps, t = qs.time_step(x, 0.0, 0.0, 0.01, 100)
    assert steps == 0 and t == 0.0
    assert np.array_equal(x.get(), before)
    v = qa.DeviceVec(N * N)
    v.set(before)
    qa.rk4_device(qs.matrix(), v, 0.01, 0)
    assert np.array_equal(v.get(), before)
    qs.destroy()


def test_system_without_terms_keeps_rho():
    """no Hamiltonian or Lindblad term: the superoperator is the stored zero
    diagonal (PETSc keeps every diagonal entry), and stepping leaves rho as it is"""
    _fixed_rk4()
    r = {"subs": [("op", 3), ("op", 2)], "terms": [("lin", 0.0, [(0, "lower")])], "mcwf": False}
    qs, _, _ = build_system(r)
    N = 6
    rp, col, val = qs.matrix().export()
    assert np.count_nonzero(val) == 0
    x = qs.create_qvec()
    rng = np.random.default_rng(1)
    a = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    rho = a @ a.conj().T
    rho /= np.trace(rho)
    x.set(rho.reshape(-1, order="F"))
    qs.time_step(x, 0.0, 0.5, 0.01, 1000)
    assert np.abs(x.get() - rho.reshape(-1, order="F")).max() < 1e-15
    qs.destroy()


def test_single_instance_batch_matches_the_plain_path():
    _fixed_rk4()
    p = recipes.tc_instances(1, 5)[0]
    r = recipes.tc_recipe(*p)
    qs, _, _ = build_system(r)
    N, loc = recipes.tc_init(p[0], p[1])
    x0 = np.zeros(N * N, np.complex128)
    x0[loc] = 1
    b = qa.Batch([qs.matrix()], [N], prec=qa.FP64)
    b.set_state(0, x0)
    b.rk4(0.01, 60)
    v = qa.DeviceVec(N * N)
    v.set(x0)
    qa.rk4_device(qs.matrix(), v, 0.01, 60)
    assert np.abs(b.get_state(0) - v.get()).max() < 1e-13
    b.destroy()
    qs.destroy()


def test_size_mismatch_and_bad_gate_strides_are_errors():
    r = recipes.spin_chain_recipe(4, seed=2)
    qs, _, _ = build_system(r)
    A = qs.matrix()
    n = A.info()[0]
    wrong = qa.DeviceVec(n + 2)
    with pytest.raises(qa.QuacError, match="size"):
        qa.rk4_device(A, wrong, 0.01, 3)
    with pytest.raises(qa.QuacError):
        qa.rk4_device(A, qa.DeviceVec(n), 0.01, -1)
    L = qa.lib()
    v = qa.DeviceVec(16 * 16)
    u = np.eye(2, dtype=np.complex128).view(np.float64).ravel()
    strides = (ctypes.c_int64 * 1)(3)      # not a 2-level subsystem of a 16-level space
    rc = L.qh_gate(v.h, 16, 1, 1, strides, u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc != 0 and b"stride" in L.qh_last_error()
    qs.destroy()


def test_smallest_mcwf_ensemble():
    """use_mcwf_solver refuses fewer than two trajectories, as the reference
    does (qsystem.c:372-373, whose message says "more than 2"); the smallest
    ensemble, two trajectories in one wavefront, of a decaying qubit with no
    drive has jumped to the ground state by t = 20 (the chance one has not is
    2 e^-10), so its excited population is zero"""
    _fixed_rk4()
    qs = qa.System()
    q = qs.create_op(2)
    qs.add_ham_term(1.0, q.n)
    qs.add_lin_term(0.5, q)
    with pytest.raises(qa.QuacError, match="more than 2"):
        qs.use_mcwf_solver(1, 7)
    qs.use_mcwf_solver(2, 7)
    qs.construct_matrix()
    x = qs.create_qvec()
    x.add(1.0, 1)
    steps, t = qs.time_step(x, 0.0, 20.0, 0.01, 10 ** 6)
    assert steps == 2000 and abs(t - 20.0) < 1e-9
    e = x.expectation(q.n)
    assert abs(e.real) < 1e-9 and abs(e.imag) < 1e-12
    qs.destroy()


def test_mcwf_with_time_dependent_terms_is_refused():
    """the MCWF trajectories step with the time-independent effective
    Hamiltonian on the device; time_step_sys refuses an ensemble whose system has
    time-dependent terms with an error, not a silent result without them"""
    _fixed_rk4()
    qs = qa.System()
    q = qs.create_op(2)
    qs.add_ham_term(1.0, q.n)
    qs.add_lin_term(0.5, q)
    qs.add_ham_term_time_dep(0.3, lambda t: 1.0, q)
    qs.use_mcwf_solver(4, 7)
    qs.construct_matrix()
    x = qs.create_qvec()
    x.add(1.0, 1)
    with pytest.raises(qa.QuacError, match="time-dependent"):
        qs.time_step(x, 0.0, 0.1, 0.01, 100)
    qs.destroy()


def test_stream_calibration_and_ragged_vector_kernels():
    c, r = qa.bench_stream(1 << 24, 2)
    assert 100.0 < c < 20000.0 and 100.0 < r < 20000.0
    with pytest.raises(qa.QuacError):
        qa.bench_stream(1000, 2)          # below 4096 bytes and not a multiple of 16
    # FP64 -> FP32 hi + lo split and back over one element per thread, with a
    # last partial block
    n = 256 * 37 + 5
    rng = np.random.default_rng(4)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    v64 = qa.DeviceVec(n, qa.FP64)
    v64.set(x)
    v32 = qa.DeviceVec(n, qa.FP32)
    qa._qh(qa.lib().qh_vec_copy(v64.h, v32.h))
    back = qa.DeviceVec(n, qa.FP64)
    qa._qh(qa.lib().qh_vec_copy(v32.h, back.h))
    y = back.get()
    assert np.max(np.abs(y - x)) <= 1e-14 * np.max(np.abs(x))
