"""Build a product System (through libquac.so's C-ABI) from a neutral recipe."""
import numpy as np

import quac_amd as qa
import recipes


def build_system(recipe, construct=True):
    qs = qa.System()
    subs = []
    for (kind, L) in recipe["subs"]:
        subs.append(("op", qs.create_op(L)) if kind == "op" else ("vec", qs.create_vec_op(L)))

    def op(ref):
        s, name = ref
        kind, o = subs[s]
        if isinstance(name, tuple):
            return o[name[1]]
        return {"lower": lambda: o, "raise": lambda: o.dag, "n": lambda: o.n, "eye": lambda: o.eye,
                "sx": lambda: o.sig_x, "sy": lambda: o.sig_y, "sz": lambda: o.sig_z,
                "other": lambda: o.other}[name]()

    if recipe.get("mcwf"):
        # use_mcwf_solver needs a Lindblad term first (qsystem.c:366-370): add terms, then switch
        pass
    for (k, a, ops) in recipe["terms"]:
        plist = [op(r) for r in ops]
        if k == "ham":
            qs.add_ham_term(a, *plist)
        elif k == "lin":
            qs.add_lin_term(a, *plist)
    if recipe.get("mcwf"):
        qs.use_mcwf_solver(16, 1)
    if construct:
        qs.construct_matrix()
    return qs, subs, op


def dm_state(qs, N, loc):
    x = qs.create_qvec()
    x.add_loc(1.0, loc)
    return x


def subsystem_layout(recipe):
    """(levels, stride) per subsystem for populations"""
    tot, lay = recipes.layout(recipe)
    return [(L, tot // (L * nb)) for (L, nb) in lay]


def populations_np(x, recipe):
    """numpy populations <n_s> from a DM vector (reference for the device kernel)"""
    tot, lay = recipes.layout(recipe)
    diag = np.real(x[np.arange(tot) * (tot + 1)])
    out = []
    for (L, nb) in lay:
        st = tot // (L * nb)
        lev = (np.arange(tot) // st) % L
        out.append(float(np.sum(lev * diag)))
    return np.array(out)
