"""Dense textbook references for small systems (numpy), used to pin the oracle.

vec(rho) is column-major (loc = N*col + row, qvec_utilities.c:326), so
vec(A X B) = kron(B.T, A) vec(X).  Subsystem 0 is the most significant factor.
"""
import numpy as np

import recipes


def single(name, L):
    M = np.zeros((L, L), np.complex128)
    if isinstance(name, tuple):                         # VEC pair element |p><.|
        raise ValueError
    if name == "lower":
        for i in range(L - 1):
            M[i, i + 1] = np.sqrt(i + 1.0)
    elif name == "raise":
        for i in range(L - 1):
            M[i + 1, i] = np.sqrt(i + 1.0)
    elif name == "n":
        M = np.diag(np.arange(L).astype(np.complex128))
    elif name == "eye":
        M = np.eye(L, dtype=np.complex128)
    elif name == "sx":
        M[0, 1] = M[1, 0] = 1
    elif name == "sy":
        M[0, 1] = -1j
        M[1, 0] = 1j
    elif name == "sz":
        M[0, 0] = 1
        M[1, 1] = -1
    elif name == "other":
        for i in range(2, L):
            M[i, i] = 1
    return M


def embed(M, s, lay):
    nb = lay[s][1]
    tot = int(np.prod([L for (L, _) in lay]))
    na = tot // (lay[s][0] * nb)
    return np.kron(np.eye(nb), np.kron(M, np.eye(na)))


def chain_matrix(recipe, ops):
    """Dense product ops[0] @ ops[1] @ ... ; consecutive VEC entries form |p1><p2|."""
    tot, lay = recipes.layout(recipe)
    G = np.eye(tot, dtype=np.complex128)
    k = 0
    while k < len(ops):
        s, name = ops[k]
        if isinstance(name, tuple):
            s2, name2 = ops[k + 1]
            L = lay[s][0]
            M = np.zeros((L, L), np.complex128)
            M[name[1], name2[1]] = 1.0
            k += 2
        else:
            M = single(name, lay[s][0])
            k += 1
        G = G @ embed(M, s, lay)
    return G


def superop(recipe, conj_a=False):
    """Dense system matrix for a recipe (time-independent terms)."""
    tot, _ = recipes.layout(recipe)
    I = np.eye(tot)
    dm = recipes.is_dm(recipe)
    mcwf = recipe.get("mcwf", False)
    if dm and not mcwf:
        A = np.zeros((tot * tot, tot * tot), np.complex128)
        for (k, a, ops) in recipe["terms"]:
            if k.startswith("td"):
                continue
            G = chain_matrix(recipe, ops)
            if k == "ham":
                a2 = np.conj(a) if conj_a else a
                A += -1j * a * np.kron(I, G) + 1j * a2 * np.kron(G.T, I)
            else:
                CdC = G.conj().T @ G
                A += a * (np.kron(G.conj(), G) - 0.5 * np.kron(I, CdC) - 0.5 * np.kron(CdC.T, I))
        return A
    H = np.zeros((tot, tot), np.complex128)
    for (k, a, ops) in recipe["terms"]:
        G = chain_matrix(recipe, ops)
        if k == "ham":
            H += a * G
        else:
            H += -0.5j * a * (G.conj().T @ G)
    return -1j * H


def vec(rho):
    return np.asarray(rho).flatten(order="F")


def unvec(v, N):
    return np.asarray(v).reshape(N, N, order="F")
