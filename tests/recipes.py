"""System recipes shared by the oracle and the product in tests and bench.

A recipe is a neutral description of a QuaC qsystem:
  subs:  list of ("op", levels) | ("vec", levels) in creation order
         (create_op_sys / create_vec_op_sys, src/qsystem.c:153-230, 265-306)
  terms: list of (kind, a, [(sub, opname), ...]) with kind "ham"|"lin"|"td_ham"|"td_lin"
         and opname in lower|raise|n|eye|sx|sy|sz|other|("vec", position)
  mcwf:  use_mcwf_solver semantics for assembly (kron.c:664-700)

The fixture recipes below restate the call sequences of the reference's
matrix-construction tests (tests/matrix_construction_test_{ham,lin,ham_only,
ham_lin}_op.c) so that the pristine matrices can be rebuilt and compared.
"""
import math

OPTYPE = {"raise": -1, "n": 0, "lower": 1, "vec": 2, "sx": 3, "sy": 4, "sz": 5, "eye": 6, "other": 7}
KIND = {"ham": 1, "lin": 2, "td_ham": -1, "td_lin": -2}


def layout(recipe):
    """(tot, [(levels, n_before)]) for the subsystems (qsystem.c:93-151, n_before = tot at creation)."""
    tot = 1
    out = []
    for (_, L) in recipe["subs"]:
        out.append((L, tot))
        tot *= L
    return tot, out


def is_dm(recipe):
    return any(k in ("lin", "td_lin") for (k, _, _) in recipe["terms"]) or recipe.get("dm", False)


def oracle_terms(recipe):
    """Terms as (kind, a, [(type, levels, n_before, position)]) for oracle.construct."""
    tot, lay = layout(recipe)
    terms = []
    for (k, a, ops) in recipe["terms"]:
        olist = []
        for (s, name) in ops:
            L, nb = lay[s]
            if isinstance(name, tuple):
                olist.append((OPTYPE["vec"], L, nb, name[1]))
            else:
                olist.append((OPTYPE[name], L, nb, -1))
        terms.append((KIND[k], complex(a), olist))
    return tot, terms


# ---------------------------------------------------------------------------
# pristine-matrix recipes
# ---------------------------------------------------------------------------
_B1 = [(0, 2.0, [["lower"], ["raise"], ["n"]]), (1, 1.0, [["lower"], ["raise"], ["n"]]),
       (2, 0.5, [["lower"], ["raise"], ["n"]])]


def _basic_terms(nop):
    """basic (b*) recipes: ops op2,op3,op4 = subsystems 0,1,2 (levels 2,3,4)."""
    w2, w3, w4, w5 = 2.0, 1.0, 0.5, 0.25
    if nop == 1:
        t = []
        for s, w in ((0, w2), (1, w3), (2, w4)):
            t += [(w, [(s, "lower")]), (w, [(s, "raise")]), (w, [(s, "n")])]
        return t
    if nop == 2:
        return [(w2, [(0, "lower"), (0, "raise")]), (w2, [(0, "raise"), (0, "lower")]),
                (w2, [(0, "n"), (0, "raise")]),
                (w3, [(1, "lower"), (1, "lower")]), (w3, [(1, "raise"), (1, "raise")]),
                (w3, [(1, "n"), (1, "lower")]),
                (w4, [(2, "lower"), (2, "n")]), (w4, [(2, "raise"), (2, "n")]), (w4, [(2, "n"), (2, "n")]),
                (w5, [(2, "lower"), (1, "n")]), (w5, [(1, "raise"), (0, "n")]), (w5, [(0, "n"), (2, "n")])]
    return [(w2, [(0, "lower"), (0, "raise"), (0, "n")]), (w2, [(0, "raise"), (0, "lower"), (0, "n")]),
            (w2, [(0, "n"), (0, "raise"), (0, "lower")]),
            (w3, [(1, "lower"), (1, "lower"), (1, "lower")]), (w3, [(1, "raise"), (1, "raise"), (1, "raise")]),
            (w3, [(1, "n"), (1, "lower"), (1, "n")]),
            (w4, [(2, "lower"), (2, "n"), (2, "raise")]), (w4, [(2, "raise"), (2, "n"), (2, "lower")]),
            (w4, [(2, "n"), (2, "n"), (2, "n")]),
            (w5, [(2, "lower"), (1, "n"), (0, "raise")]), (w5, [(1, "raise"), (0, "n"), (2, "lower")]),
            (w5, [(0, "n"), (2, "n"), (1, "n")])]


def _pauli_terms(nop):
    w = 2.0
    if nop == 1:
        return [(w, [(0, "sx")]), (w, [(0, "sy")]), (w, [(0, "sz")]), (w, [(0, "eye")])]
    if nop == 2:
        return [(w, [(0, "sx"), (0, "sy")]), (w, [(0, "sy"), (0, "sz")]), (w, [(0, "sz"), (0, "sx")]),
                (w, [(0, "eye"), (0, "eye")])]
    return [(w, [(0, "sx"), (0, "sx"), (0, "sz")]), (w, [(0, "sy"), (0, "sy"), (0, "sx")]),
            (w, [(0, "sz"), (0, "sz"), (0, "sy")]), (w, [(0, "eye"), (0, "eye"), (0, "eye")])]


def fixture_recipe(family, nop, flavor):
    """family in ham|lin|ham_only|ham_lin|mcwf; nop in 1,2,3; flavor in br|bc|pr|pc."""
    basic = flavor[0] == "b"
    cplx = 1j if flavor[1] == "c" else 1.0
    subs = [("op", 2), ("op", 3), ("op", 4)] if basic else [("op", 2)]
    base = _basic_terms(nop) if basic else _pauli_terms(nop)
    terms = []
    if family in ("ham", "ham_only", "ham_lin", "mcwf"):
        terms += [("ham", w * cplx, ops) for (w, ops) in base]
    if family in ("lin", "ham_lin", "mcwf"):
        terms += [("lin", w * cplx, ops) for (w, ops) in base]
    if family == "ham":
        # add_lin_term(qsys,0.0,1,opX->n) forces the DM space (..._ham_op.c:46,212,384,80)
        terms.append(("lin", 0.0, [(2 if (basic and nop == 1) else (1 if basic else 0), "n")]))
    return {"subs": subs, "terms": terms, "mcwf": family == "mcwf"}


FIXTURES = [(fam, nop, fl) for fam in ("ham", "lin", "ham_only", "mcwf") for nop in (1, 2, 3)
            for fl in ("br", "bc", "pr", "pc")]


# ---------------------------------------------------------------------------
# workload recipes (synthetic inputs, seeded)
# ---------------------------------------------------------------------------
class SplitMix:
    """SplitMix64 seeding xoshiro256**; uniform(a,b) = a + (b-a)*(x>>11)*2^-53."""

    def __init__(self, seed):
        self.s = []
        z = seed & 0xFFFFFFFFFFFFFFFF
        for _ in range(4):
            z = (z + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
            t = z
            t = ((t ^ (t >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
            t = ((t ^ (t >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
            self.s.append(t ^ (t >> 31))

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & 0xFFFFFFFFFFFFFFFF

    def next(self):
        s = self.s
        r = (self._rotl((s[1] * 5) & 0xFFFFFFFFFFFFFFFF, 7) * 9) & 0xFFFFFFFFFFFFFFFF
        t = (s[1] << 17) & 0xFFFFFFFFFFFFFFFF
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return r

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * 2.0 ** -53)

    def randint(self, lo, hi):
        """integer uniform in [lo, hi]"""
        return lo + int(self.uniform() * (hi - lo + 1))


def jc_recipe(levels=10, wq=1.0, wc=1.0, g=0.1, kappa=1e-2, gamma=1e-2):
    """Jaynes-Cummings (examples/simple_jc_test.c:34-55 call shape): emitter (sub 0), cavity (sub 1)."""
    return {"subs": [("op", 2), ("op", levels)],
            "terms": [("ham", wq, [(0, "n")]), ("ham", wc, [(1, "n")]),
                      ("ham", g, [(0, "raise"), (1, "lower")]), ("ham", g, [(1, "raise"), (0, "lower")]),
                      ("lin", gamma, [(0, "lower")]), ("lin", kappa, [(1, "lower")])],
            "mcwf": False}


def jc_init(levels=10):
    """emitter excited, cavity vacuum: Hilbert index 1*levels + 0, DM loc = idx*N + idx."""
    N = 2 * levels
    idx = 1 * levels + 0
    return N, idx * N + idx


def spin_chain_recipe(nq=10, seed=1, gamma1=1e-3, gamma2=5e-3):
    """H = sum h_i X_i + sum J_i Z_i Z_{i+1}; L: gamma1 L[sigma-_i] + gamma2 L[sigma_z_i]."""
    rng = SplitMix(seed)
    h = [rng.uniform(0.5, 1.5) for _ in range(nq)]
    J = [rng.uniform(0.5, 1.5) for _ in range(nq - 1)]
    terms = [("ham", h[i], [(i, "sx")]) for i in range(nq)]
    terms += [("ham", J[i], [(i, "sz"), (i + 1, "sz")]) for i in range(nq - 1)]
    terms += [("lin", gamma1, [(i, "lower")]) for i in range(nq)]
    terms += [("lin", gamma2, [(i, "sz")]) for i in range(nq)]
    return {"subs": [("op", 2)] * nq, "terms": terms, "mcwf": False}


def spin_chain_init(nq=10):
    """|0101...>: qubit i (sub i, most significant first) in state i%2."""
    N = 2 ** nq
    idx = 0
    for i in range(nq):
        idx = idx * 2 + (i % 2)
    return N, idx * N + idx


def tc_instances(count=512, seed=2):
    """Tavis-Cummings sweep parameters: (nq, nc, g, delta, kappa, gamma) per instance."""
    rng = SplitMix(seed)
    out = []
    for _ in range(count):
        nq = rng.randint(1, 4)
        nc = rng.randint(5, 30)
        g = rng.uniform(0.05, 0.2)
        delta = rng.uniform(-0.1, 0.1)
        kappa = rng.uniform(1e-3, 1e-2)
        gamma = rng.uniform(1e-3, 1e-2)
        out.append((nq, nc, g, delta, kappa, gamma))
    return out


def tc_recipe(nq, nc, g, delta, kappa, gamma):
    """H = delta a^dag a + sum_i g (sigma_i^dag a + sigma_i a^dag) (rotating frame);
    L: kappa L[a] + gamma L[sigma_i].  Subsystems: qubits 0..nq-1, cavity last."""
    subs = [("op", 2)] * nq + [("op", nc)]
    c = nq
    terms = [("ham", delta, [(c, "n")])]
    for i in range(nq):
        terms.append(("ham", g, [(i, "raise"), (c, "lower")]))
        terms.append(("ham", g, [(c, "raise"), (i, "lower")]))
    terms.append(("lin", kappa, [(c, "lower")]))
    for i in range(nq):
        terms.append(("lin", gamma, [(i, "lower")]))
    return {"subs": subs, "terms": terms, "mcwf": False}


def tc_init(nq, nc):
    """all qubits excited, cavity vacuum."""
    N = (2 ** nq) * nc
    idx = 0
    for _ in range(nq):
        idx = idx * 2 + 1
    idx = idx * nc
    return N, idx * N + idx


def noisy_circuit_recipe(nq=12, t1=1e-3, tphi=2e-3):
    """per-qubit T1 decay and dephasing (no Hamiltonian)."""
    terms = [("lin", t1, [(i, "lower")]) for i in range(nq)]
    terms += [("lin", tphi, [(i, "sz")]) for i in range(nq)]
    return {"subs": [("op", 2)] * nq, "terms": terms, "mcwf": False}


def random_circuit(nq=12, ngates=300, seed=3):
    """seeded random circuit: H, CNOT(q,q+1), RZ(theta), RX(theta); theta in [0, 2pi)."""
    rng = SplitMix(seed)
    gates = []
    for _ in range(ngates):
        k = rng.randint(0, 3)
        if k == 0:
            gates.append(("H", (rng.randint(0, nq - 1),), 0.0))
        elif k == 1:
            q = rng.randint(0, nq - 2)
            gates.append(("CNOT", (q, q + 1), 0.0))
        elif k == 2:
            gates.append(("RZ", (rng.randint(0, nq - 1),), rng.uniform(0.0, 2 * math.pi)))
        else:
            gates.append(("RX", (rng.randint(0, nq - 1),), rng.uniform(0.0, 2 * math.pi)))
    return gates


def kerr_params(seed, count):
    """3 coupled driven Kerr cavities: 10 features per set
    (delta x3, K x3, F x3, J) and cavity loss kappa (x3, shared input layer via mean)."""
    rng = SplitMix(seed)
    out = []
    for _ in range(count):
        d = [rng.uniform(-1, 1) for _ in range(3)]
        K = [rng.uniform(0.05, 0.5) for _ in range(3)]
        F = [rng.uniform(0, 1) for _ in range(3)]
        J = rng.uniform(0.1, 0.5)
        kap = [rng.uniform(0.05, 0.2) for _ in range(3)]
        out.append({"delta": d, "K": K, "F": F, "J": J, "kappa": kap})
    return out


def kerr_sweep_params(seed, count):
    """parameter sets of the learned steady-state warm start: detunings in
    U[-1, 1], Kerr nonlinearities in U[0.05, 0.5] and drives in U[0, 1] per
    cavity, hopping in U[0.1, 0.5], one cavity loss in U[0.05, 0.2] shared by
    the three cavities"""
    rng = SplitMix(seed)
    out = []
    for _ in range(count):
        d = [rng.uniform(-1, 1) for _ in range(3)]
        K = [rng.uniform(0.05, 0.5) for _ in range(3)]
        F = [rng.uniform(0, 1) for _ in range(3)]
        J = rng.uniform(0.1, 0.5)
        kap = rng.uniform(0.05, 0.2)
        out.append({"delta": d, "K": K, "F": F, "J": J, "kappa": [kap] * 3})
    return out


KERR_CENTER = {"delta": [0.3, -0.2, 0.1], "K": [0.2, 0.3, 0.25], "F": [0.5, 0.6, 0.4], "J": 0.3,
               "kappa": [0.1] * 3}


def kerr_local_params(seed, count, rel=0.05, center=None):
    """parameter sets of a local sweep: every parameter of the centre set (inside
    the ranges of kerr_sweep_params) scaled by an independent factor in
    U[1 - rel, 1 + rel], the shared cavity loss included"""
    c = center or KERR_CENTER
    rng = SplitMix(seed)
    out = []
    for _ in range(count):
        f = lambda v: v * (1.0 + rng.uniform(-rel, rel))
        d = [f(v) for v in c["delta"]]
        K = [f(v) for v in c["K"]]
        F = [f(v) for v in c["F"]]
        J = f(c["J"])
        kap = f(c["kappa"][0])
        out.append({"delta": d, "K": K, "F": F, "J": J, "kappa": [kap] * 3})
    return out


def kerr_features(p):
    """the 10 network inputs: every Hamiltonian parameter in units of the loss
    (the steady state of L = -i[H, .] + kappa sum D[a_k] depends on H / kappa only)"""
    k = p["kappa"][0]
    return [x / k for x in p["delta"]] + [x / k for x in p["K"]] + [x / k for x in p["F"]] + [p["J"] / k]


def kerr_recipe(p, levels=8):
    """H = sum d_k n_k + K_k a^dag a^dag a a + F_k (a_k + a_k^dag) + J sum (a_k^dag a_{k+1} + h.c.);
    L: kappa_k L[a_k]."""
    terms = []
    for k in range(3):
        terms.append(("ham", p["delta"][k], [(k, "n")]))
        terms.append(("ham", p["K"][k], [(k, "raise"), (k, "raise"), (k, "lower"), (k, "lower")]))
        terms.append(("ham", p["F"][k], [(k, "lower")]))
        terms.append(("ham", p["F"][k], [(k, "raise")]))
    for k in range(2):
        terms.append(("ham", p["J"], [(k, "raise"), (k + 1, "lower")]))
        terms.append(("ham", p["J"], [(k + 1, "raise"), (k, "lower")]))
    for k in range(3):
        terms.append(("lin", p["kappa"][k], [(k, "lower")]))
    return {"subs": [("op", levels)] * 3, "terms": terms, "mcwf": False}
