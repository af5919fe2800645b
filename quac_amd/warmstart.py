"""Learned warm start for the steady-state solve (steady_state_sys).

A small network maps the parameters of a family of Lindblad systems to the
coefficients of their steady state in a low-rank basis learned from solved
examples; the prediction seeds GMRES, which then finishes every solve to the
requested residual (and starts again from zero if a seeded solve does not
converge).  Pieces:

  * training data: steady states that the device solver itself computes
    (System.steady_state) for sampled parameter sets;
  * basis: the leading principal directions of the centred solutions
    (rank r), from the eigen-decomposition of their Gram matrix;
  * network: features -> [Linear -> tanh] x 3 -> 2r outputs (real and
    imaginary parts of the coefficients, standardised), trained on the device
    with the hand-written forward / backward / Adam kernels of qh_mlp.hip
    (quac_amd.MLP); the loss weights each output by its variance, so that it
    measures the squared distance of the reconstructed density matrices;
  * prediction: coefficients -> mean + basis combination, made Hermitian and
    rescaled to unit trace.
"""
import time

import numpy as np

from . import MLP, DeviceF32


def _glorot(dims, rng):
    parts = []
    for din, dout in zip(dims[:-1], dims[1:]):
        lim = np.sqrt(6.0 / (din + dout))
        parts.append(rng.uniform(-lim, lim, size=din * dout))
        parts.append(np.zeros(dout))
    return np.concatenate(parts).astype(np.float32)


def hermitian_unit_trace(x, N):
    """(rho + rho^dag)/2 rescaled to trace 1, in the column-major vec layout"""
    rho = np.asarray(x).reshape(N, N, order="F")
    rho = 0.5 * (rho + rho.conj().T)
    tr = np.trace(rho).real
    if tr != 0.0:
        rho = rho / tr
    return rho.reshape(-1, order="F")


class WarmStart:
    def __init__(self, N, rank=128, hidden=(256, 256, 256), seed=0):
        self.N = int(N)
        self.rank = int(rank)
        self.hidden = tuple(int(h) for h in hidden)
        self.seed = seed
        self.net = None

    # ------------------------------------------------------------- fitting
    def fit(self, feats, sols, epochs=2000, batch=250, lr=2e-3, lr_min=1e-5, reshuffle=25, log=None):
        """feats: (n, f) parameters; sols: (n, N*N) complex steady states"""
        t0 = time.perf_counter()
        feats = np.asarray(feats, np.float64)
        n = feats.shape[0]
        batch = max(1, min(batch, n))          # a training set smaller than one batch is one batch
        self.fmean = feats.mean(0)
        self.fstd = feats.std(0) + 1e-12
        X = np.asarray(sols, np.complex64)
        self.mean = X.mean(0)
        Xc = X - self.mean
        G = Xc @ Xc.conj().T                                  # n x n Gram matrix
        lam, V = np.linalg.eigh(G.astype(np.complex128))
        order = np.argsort(lam)[::-1]
        lam, V = lam[order], V[:, order]
        r = min(self.rank, n, int(np.sum(lam > lam[0] * 1e-12)))
        self.rank = r
        sig = np.sqrt(np.maximum(lam[:r], 1e-300))
        self.basis = ((V[:, :r].conj().T @ Xc) / sig[:, None].astype(np.complex64)).astype(np.complex64)  # r x d
        C = Xc @ self.basis.conj().T                          # n x r coefficients
        T = np.hstack([C.real, C.imag]).astype(np.float64)
        self.tmean = T.mean(0)
        self.tstd = T.std(0) + 1e-12
        Tn = ((T - self.tmean) / self.tstd).astype(np.float32)
        Fn = ((feats - self.fmean) / self.fstd).astype(np.float32)
        w = (self.tstd ** 2) / np.sum(self.tstd ** 2) * T.shape[1]
        t_pca = time.perf_counter() - t0

        dims = [feats.shape[1], *self.hidden, 2 * r]
        rng = np.random.default_rng(self.seed)
        self.net = MLP(dims, batch)
        self.net.set_params(_glorot(dims, rng))
        self.net.set_output_weights(w.astype(np.float32))
        nb = max(1, n // batch)
        xd = DeviceF32(nb * batch * dims[0])
        td = DeviceF32(nb * batch * dims[-1])
        xs = [None] * nb
        ts = [None] * nb
        t1 = time.perf_counter()
        losses = []
        step = 0
        total = epochs * nb
        for ep in range(epochs):
            if ep % reshuffle == 0:
                perm = rng.permutation(n)[: nb * batch]
                xd.set(Fn[perm])
                td.set(Tn[perm])
                for b in range(nb):
                    xs[b] = _View(xd, b * batch * dims[0])
                    ts[b] = _View(td, b * batch * dims[-1])
            for b in range(nb):
                lr_t = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * step / total))
                want = (b == nb - 1) and (ep % max(1, epochs // 20) == 0 or ep == epochs - 1)
                loss = self.net.train_step(xs[b], ts[b], batch, lr_t, want_loss=want)
                if want:
                    losses.append((ep, loss))
                    if log:
                        log(f"epoch {ep} loss {loss:.4e}")
                step += 1
        t_train = time.perf_counter() - t1
        # reconstruction error of the basis and of the network on the training set
        proj = (C @ self.basis) + self.mean
        pred = self.predict_raw(feats)
        nrm = np.linalg.norm(X, axis=1)
        return {"rank": r, "pca_s": t_pca, "train_s": t_train, "steps": total,
                "losses": losses,
                "train_basis_err": float(np.mean(np.linalg.norm(proj - X, axis=1) / nrm)),
                "train_net_err": float(np.mean(np.linalg.norm(pred - X, axis=1) / nrm))}

    # ---------------------------------------------------------- prediction
    def predict_raw(self, feats):
        """network reconstruction (before the Hermitian / trace projection), (m, d) complex64"""
        feats = np.atleast_2d(np.asarray(feats, np.float64))
        m = feats.shape[0]
        Fn = ((feats - self.fmean) / self.fstd).astype(np.float32)
        out = np.empty((m, 2 * self.rank), np.float32)
        B = self.net.max_batch
        xd = DeviceF32(B * feats.shape[1])
        yd = DeviceF32(B * 2 * self.rank)
        for i in range(0, m, B):
            k = min(B, m - i)
            buf = np.zeros((B, feats.shape[1]), np.float32)
            buf[:k] = Fn[i:i + k]
            xd.set(buf)
            self.net.forward(xd, k, yd)
            out[i:i + k] = yd.get().reshape(B, -1)[:k]
        T = out.astype(np.float64) * self.tstd + self.tmean
        C = (T[:, :self.rank] + 1j * T[:, self.rank:]).astype(np.complex64)
        return C @ self.basis + self.mean

    def predict(self, feats):
        """warm starts (m, N*N) complex128: Hermitian, unit trace"""
        raw = self.predict_raw(feats)
        return np.array([hermitian_unit_trace(x.astype(np.complex128), self.N) for x in raw])

    def project(self, sols):
        """the best approximation of given solutions in the learned basis (same projection)"""
        X = np.atleast_2d(np.asarray(sols, np.complex64))
        C = (X - self.mean) @ self.basis.conj().T
        raw = C @ self.basis + self.mean
        return np.array([hermitian_unit_trace(x.astype(np.complex128), self.N) for x in raw])

    def destroy(self):
        if self.net is not None:
            self.net.destroy()
            self.net = None


class _View:
    """a device FP32 array starting at an element offset of another (no ownership)"""

    def __init__(self, base, offset):
        import ctypes
        self.h = ctypes.c_void_p(base.h.value + 4 * offset)
