"""Time stepping on the device against the CPU oracle and closed forms.

Tolerances (written here, not hidden): the FP64 drop-in path must match the
oracle's FP64 RK4 to 1e-10 in every population (and 1e-10 max-abs on the
whole density matrix for the small cases); the FP32 bulk path must keep every
population within 1e-6 of the FP64 path."""
import ctypes
import math

import numpy as np
import pytest

import oracle
import quac_amd as qa
import recipes
from qbuild import build_system, populations_np, subsystem_layout

pytestmark = pytest.mark.gpu


def _fixed_rk4():
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")


def _adaptive():
    qa.set_option("ts_rk_type", "3bs")
    qa.set_option("ts_adapt_type", "basic")


def _oracle_run(recipe, x0, dt, steps):
    tot, terms = recipes.oracle_terms(recipe)
    m = oracle.construct(tot, terms, dm=True)
    return m.rk4(x0, dt, steps, nthreads=4)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_jc_rk4_vs_oracle(prec):
    """Jaynes-Cummings 2x10 (400-row superoperator), dt = 0.01, 2000 steps."""
    _fixed_rk4()
    qa.set_option("quac_precision", prec)
    try:
        r = recipes.jc_recipe()
        N, loc = recipes.jc_init()
        qs, subs, op = build_system(r)
        x = qs.create_qvec()
        x.add_loc(1.0, loc)
        steps, tend = qs.time_step(x, 0.0, 20.0, 0.01, 100000)
        assert steps == 2000 and abs(tend - 20.0) < 1e-9
        got = x.get()
        x0 = np.zeros(N * N, np.complex128)
        x0[loc] = 1
        ref = _oracle_run(r, x0, 0.01, 2000)
        pg, pr = populations_np(got, r), populations_np(ref, r)
        if prec == "fp64":
            assert np.abs(got - ref).max() < 1e-10
            assert np.abs(pg - pr).max() < 1e-10
        else:
            assert np.abs(pg - pr).max() < 1e-6
        # device observables agree with numpy on the same state
        e = x.expectation(subs[1][1].n)
        assert abs(e.real - pr[1]) < (1e-10 if prec == "fp64" else 1e-6)
        qs.destroy()
    finally:
        qa.set_option("quac_precision", "fp64")


def test_spin_chain6_fp64_and_fp32_vs_oracle():
    _fixed_rk4()
    r = recipes.spin_chain_recipe(6, seed=1)
    N, loc = recipes.spin_chain_init(6)
    x0 = np.zeros(N * N, np.complex128)
    x0[loc] = 1
    ref = _oracle_run(r, x0, 0.005, 400)
    pr = populations_np(ref, r)
    for prec, tol in (("fp64", 1e-10), ("fp32", 1e-6)):
        qa.set_option("quac_precision", prec)
        qs, subs, op = build_system(r)
        x = qs.create_qvec()
        x.add_loc(1.0, loc)
        qs.time_step(x, 0.0, 2.0, 0.005, 100000)
        pg = x.populations([L for L, _ in subsystem_layout(r)], [s for _, s in subsystem_layout(r)])
        assert np.abs(pg - pr).max() < tol, (prec, np.abs(pg - pr).max())
        assert abs(x.trace() - 1.0) < (1e-12 if prec == "fp64" else 1e-6)
        qs.destroy()
    qa.set_option("quac_precision", "fp64")


def test_spin_chain10_full_size_vs_oracle():
    """The 10-qubit dissipative chain at full size (rho 1024 x 1024, 1 048 576
    superoperator rows), 20 RK4 steps of 0.005 through time_step_sys in FP64 and
    in FP32, against the oracle's FP64 RK4 on its own CSR: FP64 1e-10 on every
    population (and 1e-12 on the whole state), FP32 1e-6."""
    _fixed_rk4()
    r = recipes.spin_chain_recipe(10, seed=1)
    N, loc = recipes.spin_chain_init(10)
    x0 = np.zeros(N * N, np.complex128)
    x0[loc] = 1
    tot, terms = recipes.oracle_terms(r)
    ref = oracle.construct(tot, terms, dm=True).rk4(x0, 0.005, 20, nthreads=8)
    pr = populations_np(ref, r)
    lay = subsystem_layout(r)
    try:
        for prec, tol in (("fp64", 1e-10), ("fp32", 1e-6)):
            qa.set_option("quac_precision", prec)
            qs, subs, op = build_system(r)
            x = qs.create_qvec()
            x.add_loc(1.0, loc)
            steps, tend = qs.time_step(x, 0.0, 0.1, 0.005, 100000)
            assert steps == 20 and abs(tend - 0.1) < 1e-12
            pg = x.populations([L for L, _ in lay], [s for _, s in lay])
            assert np.abs(pg - pr).max() < tol, (prec, np.abs(pg - pr).max())
            if prec == "fp64":
                assert np.abs(x.get() - ref).max() < 1e-12
            assert abs(x.trace() - 1.0) < (1e-12 if prec == "fp64" else 1e-6)
            qs.destroy()
    finally:
        qa.set_option("quac_precision", "fp64")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_rk4_after_a_matrix_is_destroyed_and_another_created(prec):
    """Two chains of the same size with different couplings, the second system
    built right after the first is destroyed (its matrix object may take the
    freed address): the step graph and stage buffers kept between calls must
    follow the new matrix, so each run matches the oracle for its own system."""
    _fixed_rk4()
    pr = qa.FP64 if prec == "fp64" else qa.FP32
    tol = 1e-10 if prec == "fp64" else 1e-6
    N, loc = recipes.spin_chain_init(5)
    x0 = np.zeros(N * N, np.complex128)
    x0[loc] = 1
    for seed in (1, 2, 3):
        r = recipes.spin_chain_recipe(5, seed=seed)
        qs, _, _ = build_system(r)
        A = qs.matrix()
        v = qa.DeviceVec(N * N, pr)
        v.set(x0)
        qa.rk4_device(A, v, 0.01, 20, tot=N, renorm_every=10 if pr == qa.FP32 else 0)
        ref = _oracle_run(r, x0, 0.01, 20)
        err = np.abs(populations_np(v.get(), r) - populations_np(ref, r)).max()
        assert err < tol, (seed, prec, err)
        del v
        qs.destroy()


def test_rk4_device_direct_matches_oracle_and_lanes():
    """qh_rk4 on the assembled matrix for every lanes-per-row variant."""
    r = recipes.spin_chain_recipe(5, seed=3)
    N, loc = recipes.spin_chain_init(5)
    qs, _, _ = build_system(r)
    A = qs.matrix()
    x0 = np.zeros(N * N, np.complex128)
    x0[loc] = 1
    ref = _oracle_run(r, x0, 0.01, 37)
    # 128..136 = sliced layout; 140..144 = FP32 row-pair sliced layout (FP64 keeps 8 lanes per row there)
    for g in (1, 2, 4, 8, 16, 32, 64, 128, 129, 130, 131, 132, 133, 134, 135, 136, 140, 141, 142, 143, 144):
        A.set_lanes(g if g < 140 else 8, g)
        v = qa.DeviceVec(N * N, qa.FP64)
        v.set(x0)
        qa.rk4_device(A, v, 0.01, 37)
        assert np.abs(v.get() - ref).max() < 1e-12, g
        w = qa.DeviceVec(N * N, qa.FP32)
        w.set(x0)
        qa.rk4_device(A, w, 0.01, 37)
        assert np.abs(w.get() - ref).max() < 1e-6, g
    A.set_lanes(0, 0)
    qs.destroy()


def _odd_recipe():
    """3-level x 5-level system: 225 rows, so the last lane of the last
    128-row slice owns one row only"""
    return {"subs": [("op", 3), ("op", 5)],
            "terms": [("ham", 1.0, [(0, "n")]), ("ham", 0.7, [(1, "n")]),
                      ("ham", 0.3, [(0, "raise"), (1, "lower")]), ("ham", 0.3, [(0, "lower"), (1, "raise")]),
                      ("lin", 0.1, [(1, "lower")]), ("lin", 0.05, [(0, "lower")])], "mcwf": False}


@pytest.mark.parametrize("name", ["chain10", "tc_batch", "odd", "tc_single"])
def test_row_pair_layout_fp32_spmv(name):
    """FP32 SpMV on the row-pair sliced layout (140 .. 144) against the FP64 CSR
    product: every row within 1e-6 of sum_k |a_k| |x_k| (complex64 values and
    inputs, FP32 pass sums)"""
    rng = np.random.default_rng(7)
    keep = []
    if name == "chain10":
        qs, _, _ = build_system(recipes.spin_chain_recipe(10, seed=1))
        A, keep = qs.matrix(), [qs]
    elif name == "odd":
        qs, _, _ = build_system(_odd_recipe())
        A, keep = qs.matrix(), [qs]
    elif name == "tc_single":
        # one Tavis-Cummings system (rho 160 x 160): its diagonal separates, so the
        # layout forms it from g_i + h_j (one block: the whole matrix)
        qs, _, _ = build_system(recipes.tc_recipe(3, 20, 0.1, 0.05, 5e-3, 5e-3))
        A, keep = qs.matrix(), [qs]
        assert A.pair_layout_info()["sep_slots"] > 0
    else:
        mats, tots = [], []
        for p in recipes.tc_instances(40, 3):
            q2, _, _ = build_system(recipes.tc_recipe(*p))
            keep.append(q2)
            mats.append(q2.matrix())
            tots.append(recipes.tc_init(p[0], p[1])[0])
        b = qa.Batch(mats, tots, prec=qa.FP32)
        A = b.matrix()
        keep.append(b)
    n = A.info()[0]
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64).astype(np.complex128)
    xv = qa.DeviceVec(n, qa.FP32)
    xv.set(x)
    rp, col, val = A.export()
    rows = np.repeat(np.arange(n), np.diff(rp))
    ref = np.zeros(n, np.complex128)
    np.add.at(ref, rows, val * x[col])
    mag = np.zeros(n)
    np.add.at(mag, rows, np.abs(val) * np.abs(x[col]))
    for g in (140, 141, 142, 143, 144):
        A.set_lanes(8, g)
        yv = qa.DeviceVec(n, qa.FP32)
        qa._qh(qa.lib().qh_spmv(A.h, xv.h, yv.h))
        err = np.abs(yv.get() - ref)
        assert np.all(err <= 1e-6 * mag + 1e-30), (g, float((err / np.maximum(mag, 1e-300)).max()))
    A.set_lanes(0, 0)
    for k in keep:
        k.destroy()


def _sep_slices(rp, col, val, blocks):
    """per 128-row slice: True when a block of `blocks` (row offsets, then the row
    count) holds the whole slice and its diagonal separates -- an N x N density
    matrix block, N even, starting at an even row, whose diagonal entry at row i,
    column j equals D(i, 0) + D(0, j) - D(0, 0) to 1e-12 relative"""
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    d = np.full(n, np.nan + 0j)
    on = col == rows
    d[rows[on]] = val[on]
    ok_blk = []
    for b in range(len(blocks) - 1):
        off, m = blocks[b], blocks[b + 1] - blocks[b]
        N = int(round(np.sqrt(m)))
        ok = N >= 2 and N * N == m and N % 2 == 0 and off % 2 == 0
        if ok:
            D = d[off:off + m].reshape(N, N, order="F")        # D[i, j]: rho row i, column j
            g, h = D[:, 0], D[0, :] - D[0, 0]
            err = np.abs((D - (g[:, None] + h[None, :])).real) + np.abs((D - (g[:, None] + h[None, :])).imag)
            sc = (np.abs(D.real) + np.abs(D.imag) + (np.abs(g.real) + np.abs(g.imag))[:, None]
                  + (np.abs(h.real) + np.abs(h.imag))[None, :])
            ok = bool(np.all(np.isfinite(D)) and np.all(err <= 1e-12 * sc))
        ok_blk.append(ok)
    out = []
    for s0 in range(0, n, 128):
        s1 = min(s0 + 128, n)
        b = int(np.searchsorted(blocks, s0, side="right")) - 1
        out.append(ok_blk[b] and s1 <= blocks[b + 1])
    return out


def _pair_slot_classes(rp, col, val, blocks=None):
    """The row-pair layout restated in Python: per lane (rows 2l, 2l+1) each entry
    of row 2l at column c pairs with an unused entry of row 2l+1 at c+1 (load
    x[c], x[c+1], halves (0, 1)), else c-1 (load from c-1, halves (1, 0)), else c
    (halves (0, 0), or (1, 1) from c-1 at the last column); an unmatched entry
    loads alone with the in-order halves (row 2l from c, row 2l+1 from c-1) unless
    that leaves the vector.  Slots of a 128-row slice are the distinct
    (|b - 2l|, h0, h1, rank) over its lanes; a slot is uniform when every active
    lane holds bit-identical (value row 2l, value row 2l+1) (zero for an absent
    row), scalar when all those are real or all imaginary, else complex; a
    scalar slot stores one real per lane when every lane's two rows hold the same
    FP32 value (an absent row holds zero), else a real pair per lane.  In a
    slice inside a block with a separable diagonal (_sep_slices) the slot
    (0, 0, 1, 0) stores nothing when it is not uniform and every active lane holds
    the diagonal entries of both its rows.
    Returns (slots, complex pairs stored, real pairs stored, single reals stored,
    separable-diagonal slots)."""
    n = len(rp) - 1
    slots = ncpx = nreal = nsingle = nsep = 0
    z = np.complex128(0)
    sep_ok = _sep_slices(rp, col, val, blocks if blocks is not None else [0, n])
    for s0 in range(0, n, 128):
        per_slot = {}
        diag_lane = {}
        for r0 in range(s0, min(n, s0 + 128), 2):
            r1 = r0 + 1
            e0 = list(range(rp[r0], rp[r0 + 1]))
            e1 = list(range(rp[r1], rp[r1 + 1])) if r1 < n else []
            used = set()
            loads = []
            for a in e0:
                c = int(col[a])
                f, how = None, None
                for how_, want in ((0, c + 1), (1, c - 1), (2, c)):
                    for b in e1:
                        if b not in used and int(col[b]) == want:
                            f, how = b, how_
                            break
                    if f is not None:
                        break
                if f is not None:
                    used.add(f)
                last = c + 1 >= n
                if how == 0:
                    loads.append((c, 0, 1, a, f))
                elif how == 1:
                    loads.append((c - 1, 1, 0, a, f))
                elif how == 2:
                    loads.append((c - 1, 1, 1, a, f) if last else (c, 0, 0, a, f))
                else:
                    loads.append((c - 1, 1, 0, a, None) if last else (c, 0, 1, a, None))
            for b in e1:
                if b in used:
                    continue
                c = int(col[b])
                loads.append((0, 1, 0, None, b) if c == 0 else (c - 1, 0, 1, None, b))
            seen = {}
            for (bb, h0, h1, a, b) in sorted(loads, key=lambda t: (abs(t[0] - r0), t[1], t[2], t[0] - r0 > 0)):
                k = (abs(bb - r0), h0, h1)
                rank = seen.get(k, 0)
                seen[k] = rank + 1
                v = (val[a] if a is not None else z, val[b] if b is not None else z)
                per_slot.setdefault(k + (rank,), []).append(v)
                dg = a is not None and b is not None and int(col[a]) == r0 and int(col[b]) == r1
                diag_lane.setdefault(k + (rank,), []).append(dg)
        for key, vs in per_slot.items():
            slots += 1
            bits = {(v0.real.tobytes(), v0.imag.tobytes(), v1.real.tobytes(), v1.imag.tobytes()) for v0, v1 in vs}
            if len(bits) == 1:
                continue
            if sep_ok[s0 // 128] and key == (0, 0, 1, 0) and all(diag_lane[key]):
                nsep += 1
                continue
            isre = all(v0.imag == 0 and v1.imag == 0 for v0, v1 in vs)
            if isre or all(v0.real == 0 and v1.real == 0 for v0, v1 in vs):
                part = (lambda z: np.float32(z.real)) if isre else (lambda z: np.float32(z.imag))
                if all(part(v0) == part(v1) for v0, v1 in vs):
                    nsingle += len(vs)
                else:
                    nreal += len(vs)
            else:
                ncpx += len(vs)
    return slots, ncpx, nreal, nsingle, nsep


@pytest.mark.parametrize("name", ["odd", "tc3"])
def test_row_pair_layout_counts(name):
    """slot and stored-value counts of the device-built row-pair layout against the
    Python restatement of its pairing and classification"""
    keep = []
    if name == "odd":
        qs, _, _ = build_system(_odd_recipe())
        A, keep = qs.matrix(), [qs]
    else:
        mats, tots = [], []
        for p in recipes.tc_instances(3, 3):
            q2, _, _ = build_system(recipes.tc_recipe(*p))
            keep.append(q2)
            mats.append(q2.matrix())
            tots.append(recipes.tc_init(p[0], p[1])[0])
        b = qa.Batch(mats, tots, prec=qa.FP32)
        A = b.matrix()
        keep.append(b)
    rp, col, val = A.export()
    lay = A.pair_layout_info()
    got = (lay["slots"], lay["stored_pairs"], lay["stored_real_pairs"], lay["stored_real_singles"],
           lay["sep_slots"])
    blocks = None
    if name == "tc3":
        blocks = np.concatenate([[0], np.cumsum([t * t for t in tots])])
    want = _pair_slot_classes(rp, col, val, blocks)
    assert got == want, (got, want)
    if name == "tc3":
        assert want[3] > 0 and want[4] > 0   # equal-scalar and separable-diagonal slots
    for k in keep:
        k.destroy()


def test_row_pair_layout_rk4_odd_rows():
    """FP32 RK4 on the row-pair layout for a matrix with an odd number of rows
    (the last lane's second row absent): populations within 1e-6 of the oracle"""
    r = _odd_recipe()
    qs, _, _ = build_system(r)
    A = qs.matrix()
    N = 15
    x0 = np.zeros(N * N, np.complex128)
    x0[7 * N + 7] = 1.0
    ref = _oracle_run(r, x0, 0.01, 200)
    for g in (140, 141, 142, 143, 144):
        A.set_lanes(8, g)
        w = qa.DeviceVec(N * N, qa.FP32)
        w.set(x0)
        qa.rk4_device(A, w, 0.01, 200)
        assert np.abs(w.get() - ref).max() < 1e-6, g
    A.set_lanes(0, 0)
    qs.destroy()


def _slot_classes(rp, col, val):
    """Slots of the sliced layout restated in numpy: an entry of row r at column
    r + d sits in slot (r // 64, |d|, rank), rank 1 only for the d > 0 entry of a
    row that also holds r - |d|.  A slot is uniform when all its values are
    bit-identical, scalar when they are all real or all imaginary, else complex.
    Returns (slots, complex values stored, real values stored)."""
    n = len(rp) - 1
    row = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    col = col.astype(np.int64)
    d = col - row
    ad = np.abs(d)
    keys = np.sort(row * n + col)
    partner = row * n + (row - ad)
    pos = np.minimum(np.searchsorted(keys, partner), len(keys) - 1)
    rank = ((d > 0) & (keys[pos] == partner)).astype(np.int64)
    sid = (row // 64) * (2 * n + 2) + 2 * ad + rank
    order = np.argsort(sid, kind="stable")
    sid, v = sid[order], val[order]
    starts = np.flatnonzero(np.r_[True, sid[1:] != sid[:-1]])
    size = np.diff(np.r_[starts, len(sid)])
    re_bits, im_bits = v.real.view(np.int64), v.imag.view(np.int64)
    uni = ((np.maximum.reduceat(re_bits, starts) == np.minimum.reduceat(re_bits, starts)) &
           (np.maximum.reduceat(im_bits, starts) == np.minimum.reduceat(im_bits, starts)))
    real = np.maximum.reduceat(np.abs(v.imag), starts) == 0
    imag = np.maximum.reduceat(np.abs(v.real), starts) == 0
    sca = ~uni & (real | imag)
    cpx = ~uni & ~sca
    return len(starts), int(size[cpx].sum()), int(size[sca].sum())


@pytest.mark.parametrize("name", ["chain10", "tc_batch"])
def test_sliced_ell_spmv_matches_csr(name):
    """the sliced layout (slot union, uniform / real-or-imaginary / complex slot
    groups, compacted values, sign masks) against the CSR product, every variant"""
    rng = np.random.default_rng(5)
    if name == "chain10":
        qs, _, _ = build_system(recipes.spin_chain_recipe(10, seed=1))
        A, keep = qs.matrix(), [qs]
    else:
        keep, mats, tots = [], [], []
        for p in recipes.tc_instances(40, 3):
            q2, _, _ = build_system(recipes.tc_recipe(*p))
            keep.append(q2)
            mats.append(q2.matrix())
            tots.append(recipes.tc_init(p[0], p[1])[0])
        b = qa.Batch(mats, tots, prec=qa.FP64)
        A = b.matrix()
        keep.append(b)
    n = A.info()[0]
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    xv = qa.DeviceVec(n)
    xv.set(x)
    outs = []
    for g in (8, 128, 129, 130, 131, 132, 133, 134, 135, 136):
        A.set_lanes(g, g)
        yv = qa.DeviceVec(n)
        qa._qh(qa.lib().qh_spmv(A.h, xv.h, yv.h))
        outs.append(yv.get())
    rp, col, val = A.export()
    ref = np.zeros(n, np.complex128)
    np.add.at(ref, np.repeat(np.arange(n), np.diff(rp)), val * x[col])
    for g, o in zip((8, 128, 129, 130, 131, 132, 133, 134, 135, 136), outs):
        assert np.abs(o - ref).max() <= 1e-12 * np.abs(ref).max(), g
    lay = A.layout_info()
    slots, ncpx, nreal = _slot_classes(rp, col, val)
    assert (lay["slots"], lay["stored_vals"], lay["stored_reals"]) == (slots, ncpx, nreal)
    if name == "tc_batch":
        assert nreal > ncpx > 0     # the sweep's superoperators store mostly real-or-imaginary slots
    for k in keep:
        k.destroy()


@pytest.mark.parametrize("name", ["chain6", "tc_batch"])
def test_rk3bs_step_fused_matches_separate_launches(name):
    """qh_rk3bs_step on the sliced layout (three fused stage launches, error sum
    in the last epilogue) against the same step on the CSR kernel (separate
    SpMV / axpy / lincomb / error launches): k2..k4 and the candidate to 1e-12
    relative, the error estimate to 1e-10 relative; with and without a given k1"""
    rng = np.random.default_rng(11)
    keep = []
    if name == "chain6":
        qs, _, _ = build_system(recipes.spin_chain_recipe(6, seed=2))
        A, keep = qs.matrix(), [qs]
    else:
        mats, tots = [], []
        for p in recipes.tc_instances(12, 4):
            q2, _, _ = build_system(recipes.tc_recipe(*p))
            keep.append(q2)
            mats.append(q2.matrix())
            tots.append(recipes.tc_init(p[0], p[1])[0])
        b = qa.Batch(mats, tots, prec=qa.FP64)
        A = b.matrix()
        keep.append(b)
    n = A.info()[0]
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    h, atol, rtol = 0.05, 5e-7, 5e-7
    results = {}
    for g in (8, 134):
        A.set_lanes(g, 0)
        xv = qa.DeviceVec(n)
        xv.set(x)
        v = {k: qa.DeviceVec(n) for k in ("xn", "k1", "k2", "k3", "k4", "y")}
        out = []
        for have_k1 in (0, 1):
            err = ctypes.c_double()
            qa._qh(qa.lib().qh_rk3bs_step(A.h, xv.h, v["xn"].h, v["k1"].h, v["k2"].h, v["k3"].h, v["k4"].h,
                                          v["y"].h, h, atol, rtol, have_k1, ctypes.byref(err)))
            out.append((err.value, {k: v[k].get() for k in ("xn", "k1", "k2", "k3", "k4")}))
        results[g] = out
    for (e8, v8), (e134, v134) in zip(results[8], results[134]):
        for k in v8:
            assert np.abs(v8[k] - v134[k]).max() <= 1e-12 * np.abs(v8[k]).max(), (name, k)
        assert abs(e8 - e134) <= 1e-10 * e8, (e8, e134)
    A.set_lanes(0, 0)
    for k in keep:
        k.destroy()


@pytest.mark.parametrize("lanes", [8, 134])
def test_rk3bs_step_matches_oracle_trial(lanes):
    """qh_rk3bs_step (fused sliced-layout launches, and the CSR kernels) against
    the oracle's qo_rk3bs_trial, which is pinned by the closed forms of one
    Bogacki-Shampine step: k1..k4 and the candidate to 1e-12 relative, the
    weighted error norm to 1e-10 relative, with and without a given k1"""
    rng = np.random.default_rng(21)
    r = recipes.spin_chain_recipe(5, seed=3)
    qs, _, _ = build_system(r)
    A = qs.matrix()
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    n = A.info()[0]
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    h, atol, rtol = 0.08, 1e-6, 1e-5
    A.set_lanes(lanes, 0)
    xv = qa.DeviceVec(n)
    xv.set(x)
    v = {k: qa.DeviceVec(n) for k in ("xn", "k1", "k2", "k3", "k4", "y")}
    for have_k1 in (0, 1):
        err = ctypes.c_double()
        qa._qh(qa.lib().qh_rk3bs_step(A.h, xv.h, v["xn"].h, v["k1"].h, v["k2"].h, v["k3"].h, v["k4"].h,
                                      v["y"].h, h, atol, rtol, have_k1, ctypes.byref(err)))
        xn, e_or, ks = m.rk3bs_trial(x, h, atol=atol, rtol=rtol)
        got = {"xn": v["xn"].get(), "k1": v["k1"].get(), "k2": v["k2"].get(), "k3": v["k3"].get(),
               "k4": v["k4"].get()}
        want = {"xn": xn, "k1": ks[0], "k2": ks[1], "k3": ks[2], "k4": ks[3]}
        for k in got:
            assert np.abs(got[k] - want[k]).max() <= 1e-12 * np.abs(want[k]).max(), (lanes, have_k1, k)
        assert abs(err.value - e_or) <= 1e-10 * e_or, (err.value, e_or)
    A.set_lanes(0, 0)
    qs.destroy()


def test_spmv_matches_oracle():
    r = recipes.kerr_recipe(recipes.kerr_params(4, 1)[0], levels=4)
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    rng = np.random.default_rng(0)
    x = rng.normal(size=tot * tot) + 1j * rng.normal(size=tot * tot)
    qs, _, _ = build_system(r)
    xv, yv = qa.DeviceVec(tot * tot), qa.DeviceVec(tot * tot)
    xv.set(x)
    qa._qh(qa.lib().qh_spmv(qs.matrix().h, xv.h, yv.h))
    assert np.abs(yv.get() - m.spmv(x)).max() < 1e-12
    qs.destroy()


@pytest.mark.parametrize("vec", (False, True))
def test_t1_decay_adaptive_default(vec):
    """time_step_tests.c:24-101: T1 decay with the default adaptive RK3BS, <n>(1)=1/e +- 1e-4."""
    _adaptive()
    qs = qa.System()
    if vec:
        v = qs.create_vec_op(2)
        qs.add_lin_term(1.0, v[0], v[1])
    else:
        o = qs.create_op(2)
        qs.add_lin_term(1.0, o)
    x = qs.create_qvec()
    x.add(1.0, 1, 1)
    qs.construct_matrix()
    steps, tend = qs.time_step(x, 0.0, 1.0, 0.05, 200)
    assert abs(tend - 1.0) < 1e-9
    e = x.expectation(v[1], v[1]) if vec else x.expectation(o.n)
    assert abs(e.real - math.exp(-1)) < 1e-4
    qs.destroy()


def test_cascade_and_circuit_decay():
    """time_step_tests.c:337-396: SIGMAX on qubit 0 at t=0.1 hits the last-created
    subsystem (in-place convention), qubit 1 at 0.2; <n> = e^-1 and e^-0.9 at t=1.1."""
    _adaptive()
    qs = qa.System()
    o2 = qs.create_op(2)
    o3 = qs.create_op(2)
    qs.add_ham_term(1.0, o2.n)
    qs.add_lin_term(1.0, o2)
    qs.add_lin_term(1.0, o3)
    x = qs.create_qvec()
    x.add(1.0, 0, 0)
    qs.construct_matrix()
    c = qa.Circuit(5)
    c.add_gate(0.0, "SIGMAX", [0])
    c.add_gate(0.1, "SIGMAX", [1])
    qs.apply_circuit(c, 0.1)
    steps, tend = qs.time_step(x, 0.0, 1.1, 0.05, 200)
    assert abs(x.expectation(o2.n).real - math.exp(-0.9)) < 5e-4
    assert abs(x.expectation(o3.n).real - math.exp(-1.0)) < 5e-4
    qs.destroy()


def test_time_dependent_rabi_drive():
    """time_dep_tests.c:46-64,142-213: H(t) = Omega(t)/2 sigma_x with
    Omega(t) = Omega0 e^{-alpha t}: <sigma_z>(t) = cos(Omega0 (1 - e^{-alpha t})/alpha)."""
    _fixed_rk4()
    om, al = 2.0, 0.5
    qs = qa.System()
    o = qs.create_op(2)
    qs.add_ham_term_time_dep(1.0, lambda t: om / 2 * math.exp(-al * t), o.sig_x)
    qs.add_lin_term(0.0, o)
    x = qs.create_qvec()
    x.add(1.0, 0, 0)
    qs.construct_matrix()
    qs.time_step(x, 0.0, 2.0, 0.001, 100000)
    want = math.cos(om * (1 - math.exp(-al * 2.0)) / al)
    assert abs(x.expectation(o.sig_z).real - want) < 1e-6
    qs.destroy()


def test_monitor_called_every_step():
    _fixed_rk4()
    r = recipes.jc_recipe(levels=4)
    qs, subs, op = build_system(r)
    N = 8
    x = qs.create_qvec()
    x.add_loc(1.0, 4 * N + 4)
    seen = []
    qs.set_monitor(lambda step, t, v: seen.append((step, round(t, 9))))
    qs.time_step(x, 0.0, 0.1, 0.01, 1000)
    qs.set_monitor(None)
    assert [s for s, _ in seen] == list(range(11))
    assert abs(seen[-1][1] - 0.1) < 1e-9
    qs.destroy()


def test_steps_max_stops_early():
    """'Did not reach final time' path: steps_max smaller than tmax/dt."""
    _fixed_rk4()
    qs, _, _ = build_system(recipes.jc_recipe(levels=3))
    x = qs.create_qvec()
    x.add_loc(1.0, 0)
    steps, tend = qs.time_step(x, 0.0, 1.0, 0.01, 7)
    assert steps == 7 and abs(tend - 0.07) < 1e-12
    qs.destroy()


def test_small_system_single_workgroup_matches_regular_kernels():
    """generators of <= 800 rows run all steps in one LDS-resident workgroup;
    same result as the multi-launch kernels and the oracle"""
    r = recipes.jc_recipe()
    N, loc = recipes.jc_init()
    qs, _, _ = build_system(r)
    A = qs.matrix()
    x0 = np.zeros(N * N, np.complex128)
    x0[loc] = 1
    ref = _oracle_run(r, x0, 0.01, 300)
    v = qa.DeviceVec(N * N, qa.FP64)
    v.set(x0)
    qa.rk4_device(A, v, 0.01, 300)
    assert np.abs(v.get() - ref).max() < 1e-12
    A.set_lanes(8, 8)
    w = qa.DeviceVec(N * N, qa.FP64)
    w.set(x0)
    qa.rk4_device(A, w, 0.01, 300)
    assert np.abs(w.get() - ref).max() < 1e-12
    A.set_lanes(0, 0)
    f = qa.DeviceVec(N * N, qa.FP32)
    f.set(x0)
    qa.rk4_device(A, f, 0.01, 300, tot=N, renorm_every=50)
    assert np.abs(populations_np(f.get(), r) - populations_np(ref, r)).max() < 1e-6
    qs.destroy()


@pytest.mark.parametrize("adapt", ["basic", "none"])
def test_small_system_rk3bs_in_one_workgroup_matches_host_loop(adapt):
    """the default integrator on a generator of at most 640 rows runs in one
    workgroup with the step-size controller on the device; -quac_small_path 0
    keeps the host-driven loop (one trial step per launch sequence and an error
    read-back), which must take the same steps to the same state"""
    qa.set_option("ts_rk_type", "3bs")
    qa.set_option("ts_adapt_type", adapt)
    r = recipes.jc_recipe()
    N, loc = recipes.jc_init()
    qs, _, _ = build_system(r)
    xs = qs.create_qvec()
    xs.add_loc(1.0, loc)
    s1, t1 = qs.time_step(xs, 0.0, 20.0, 0.05, 100000)
    xh = qs.create_qvec()
    xh.add_loc(1.0, loc)
    qa.set_option("quac_small_path", "0")
    try:
        s2, t2 = qs.time_step(xh, 0.0, 20.0, 0.05, 100000)
    finally:
        qa.set_option("quac_small_path", "1")
    assert abs(t1 - 20.0) < 1e-9 and abs(t2 - 20.0) < 1e-9
    assert s1 == s2, (s1, s2)
    assert np.abs(xs.get() - xh.get()).max() < 1e-10
    # against the oracle's fixed-step RK4 at a fine step (the adaptive error
    # control holds the local error near atol = rtol = 5e-7)
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    x0 = np.zeros(N * N, np.complex128)
    x0[loc] = 1.0
    ref = m.rk4(x0, 0.005, 4000)
    assert np.abs(populations_np(xs.get(), r) - populations_np(ref, r)).max() < 1e-4
    _adaptive()
    qs.destroy()


def test_small_system_rk3bs_with_monitor_and_events_matches_host_loop():
    """the one-workgroup integration stops at each gate time and, with a
    monitor, after every accepted step: the cascade-and-circuit case of
    time_step_tests.c:337-396 gives the host loop's state, and the monitor sees
    increasing accepted step numbers ending at t = 1.1"""
    _adaptive()

    def run(small):
        qa.set_option("quac_small_path", "1" if small else "0")
        qs = qa.System()
        o2 = qs.create_op(2)
        o3 = qs.create_op(2)
        qs.add_ham_term(1.0, o2.n)
        qs.add_lin_term(1.0, o2)
        qs.add_lin_term(1.0, o3)
        x = qs.create_qvec()
        x.add(1.0, 0, 0)
        qs.construct_matrix()
        c = qa.Circuit(5)
        c.add_gate(0.0, "SIGMAX", [0])
        c.add_gate(0.1, "SIGMAX", [1])
        qs.apply_circuit(c, 0.1)
        seen = []
        qs.set_monitor(lambda step, t, v: seen.append((step, t)))
        qs.time_step(x, 0.0, 1.1, 0.05, 200)
        qs.set_monitor(None)
        out = (x.get(), x.expectation(o2.n).real, x.expectation(o3.n).real, seen)
        qs.destroy()
        return out

    try:
        a, a2, a3, seen = run(True)
        b, b2, b3, _ = run(False)
    finally:
        qa.set_option("quac_small_path", "1")
    assert np.abs(a - b).max() < 1e-9
    assert abs(a2 - math.exp(-0.9)) < 5e-4 and abs(a3 - math.exp(-1.0)) < 5e-4
    steps = [s for s, _ in seen]
    assert steps[0] == 0 and all(q > p for p, q in zip(steps, steps[1:]))
    assert abs(seen[-1][1] - 1.1) < 1e-9


@pytest.mark.parametrize("small", ["1", "0"])
def test_adaptive_integrator_reports_a_non_finite_state(small):
    """a NaN in the state makes the error estimate NaN, which no step size can
    bring below 1: both the one-workgroup integration and the host loop stop at
    the first trial step with an error instead of retrying"""
    _adaptive()
    qa.set_option("quac_small_path", small)
    try:
        qs, _, _ = build_system(recipes.jc_recipe(levels=3))
        x = qs.create_qvec()
        x.add_loc(1.0, 0)
        x.add_loc(float("nan"), 7)
        with pytest.raises(qa.QuacError, match="non-finite"):
            qs.time_step(x, 0.0, 1.0, 0.05, 1000)
        qs.destroy()
    finally:
        qa.set_option("quac_small_path", "1")
