// qh_gmres.hip -- steady state by restarted GMRES in FP64 on the device.
//
// Semantics of steady_state (solver.c:33-220): the trace row is added to row 0
// of the Liouvillian (row 0 += 1 at the diagonal columns i*(N+1)) so that the
// solution of A' x = e0 has unit trace; GMRES with restart 100.  The Krylov
// basis lives in HBM (column-major, one contiguous vector per basis column);
// classical Gram-Schmidt with one re-orthogonalisation (CGS2) is done with two
// streaming kernels per pass (all dots of a pass in one launch, all updates in
// one launch), the Hessenberg column stays on the device until the single
// device-to-host copy per iteration; Givens rotations and the triangular solve
// run on the host (they are O(restart^2)).  Right preconditioners: none, the
// inverse diagonal (Jacobi), or the Hamiltonian-eigenbasis operator of
// qh_precond.hip.
#include <cmath>
#include <complex>
#include <vector>

#include "qh_common.h"

namespace {

typedef std::complex<double> cd;

__global__ void k_diag_inv(const int32_t *rp, const int32_t *col, const double2 *val, int64_t n, double2 *dinv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double dr = 0.0, di = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; k++)
      if (col[k] == i) {
        dr = val[k].x;
        di = val[k].y;
      }
    if (i == 0) dr += 1.0;  // stabilisation entry at (0,0)
    const double dd = dr * dr + di * di;
    dinv[i] = dd == 0.0 ? make_double2(1.0, 0.0) : make_double2(dr / dd, -di / dd);
  }
}

// z = dinv .* v
__global__ void k_cmul_vec(const double2 *a, const double2 *v, double2 *z, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 p = a[i], q = v[i];
    z[i] = make_double2(p.x * q.x - p.y * q.y, p.x * q.y + p.y * q.x);
  }
}
// z = v ./ dinv  (x -> u for the initial guess)
__global__ void k_cdiv_vec(const double2 *a, const double2 *v, double2 *z, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 p = a[i], q = v[i];
    double dd = p.x * p.x + p.y * p.y;
    z[i] = make_double2((q.x * p.x + q.y * p.y) / dd, (q.y * p.x - q.x * p.y) / dd);
  }
}

// y[0] += sum_i z[i*(N+1)]   (single block)
__global__ void k_add_trace_row0(const double2 *z, int64_t tot, double2 *y) {
  __shared__ double sr[256], si[256];
  double ar = 0.0, ai = 0.0;
  for (int64_t i = threadIdx.x; i < tot; i += 256) {
    double2 v = z[i * (tot + 1)];
    ar += v.x;
    ai += v.y;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sr[threadIdx.x] += sr[threadIdx.x + o];
      si[threadIdx.x] += si[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) y[0] = make_double2(y[0].x + sr[0], y[0].y + si[0]);
}

// r = b - r with b = e0
__global__ void k_resid(double2 *r, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 v = r[i];
    r[i] = make_double2((i == 0 ? 1.0 : 0.0) - v.x, -v.y);
  }
}

const int DOT_BLOCKS = 256;

// partial[j*DOT_BLOCKS + bx] = sum over the block's rows of conj(V_j) w
__global__ void __launch_bounds__(256) k_multidot(const double2 *V, int64_t ld, const double2 *w, int64_t n,
                                                  double2 *partial) {
  const int j = blockIdx.y;
  const double2 *v = V + ld * j;
  double ar = 0.0, ai = 0.0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double2 a = v[i], b = w[i];
    ar += a.x * b.x + a.y * b.y;
    ai += a.x * b.y - a.y * b.x;
  }
  __shared__ double sr[256], si[256];
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sr[threadIdx.x] += sr[threadIdx.x + o];
      si[threadIdx.x] += si[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(int64_t)j * gridDim.x + blockIdx.x] = make_double2(sr[0], si[0]);
}

// h[j] = sum_bx partial[j][bx]  (one block per j)
__global__ void k_reduce_rows(const double2 *partial, int nb, double2 *h) {
  const int j = blockIdx.x;
  __shared__ double sr[256], si[256];
  double ar = 0.0, ai = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    double2 v = partial[(int64_t)j * nb + b];
    ar += v.x;
    ai += v.y;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sr[threadIdx.x] += sr[threadIdx.x + o];
      si[threadIdx.x] += si[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) h[j] = make_double2(sr[0], si[0]);
}

// w -= sum_j h_j V_j
__global__ void __launch_bounds__(256) k_multiaxpy(const double2 *V, int64_t ld, int m, const double2 *h, double2 *w,
                                                   int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = w[i];
    for (int j = 0; j < m; j++) {
      const double2 hj = h[j], v = V[ld * j + i];
      acc.x -= hj.x * v.x - hj.y * v.y;
      acc.y -= hj.x * v.y + hj.y * v.x;
    }
    w[i] = acc;
  }
}

// v = w / ||w||  with the norm read from device memory (|.|^2 in h.x)
__global__ void k_normalize(const double2 *w, const double2 *nrm2, double2 *v, int64_t n) {
  const double s = nrm2[0].x > 0.0 ? 1.0 / sqrt(nrm2[0].x) : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = make_double2(w[i].x * s, w[i].y * s);
}

// x += sum_j y_j V_j
__global__ void k_update(double2 *x, const double2 *V, int64_t ld, int m, const double2 *y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = x[i];
    for (int j = 0; j < m; j++) {
      const double2 c = y[j], v = V[ld * j + i];
      acc.x += c.x * v.x - c.y * v.y;
      acc.y += c.x * v.y + c.y * v.x;
    }
    x[i] = acc;
  }
}

// y += x
__global__ void k_axpy1(const double2 *x, double2 *y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = make_double2(y[i].x + x[i].x, y[i].y + x[i].y);
}

unsigned gr(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

int qh_pc_apply_dev(const qh_pc *pc, const double2 *v, double2 *z);
int64_t qh_pc_dim(const qh_pc *pc);

// Restarted right-preconditioned GMRES on A' M^-1 (M: none, Jacobi or the
// Hamiltonian-eigenbasis operator).  Each cycle starts from the true residual
// r = e0 - A'x of the caller's x, builds the Krylov basis of A' M^-1 from
// r/||r||, and ends with x += M^-1 (V y).
//
// With an augmentation vector u (a predicted solution) every cycle minimises the
// residual over x + M^-1 span(V_k) + span(u) instead (augmented GMRES, Saad 1997):
// with c = A' u split as V_{k+1} g + c_perp (g built by orthogonalising c against
// each new basis vector, L = ||c_perp||), and the Givens rotations of the
// Hessenberg matrix applied to g as well, the residual of the best (y, alpha) is
// |rho| L / sqrt(|gamma|^2 + L^2) <= |rho| (rho: the plain GMRES residual, gamma:
// the rotated g's last entry), at alpha = conj(gamma) rho / (|gamma|^2 + L^2).  A
// cycle therefore never needs more iterations than the plain one from the same x.
static int gmres_impl(const qh_mat *A, int64_t tot, qh_vec *x, const qh_vec *aug, int restart, double tol,
                      int64_t maxit, int pck, const qh_pc *pc, int64_t *iters, double *res) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && x && x->prec == QH_FP64 && x->n == A->rows && tot * tot == A->rows, QH_EINVAL,
             "qh_gmres: FP64 density-matrix vector of the matrix size required");
  QH_REQUIRE(restart >= 1 && restart <= 1000, QH_EINVAL, "qh_gmres: restart in [1,1000]");
  QH_REQUIRE(pck >= 0 && pck <= 2, QH_EINVAL, "qh_gmres: preconditioner kind must be 0, 1 or 2");
  QH_REQUIRE(pck != 2 || (pc && qh_pc_dim(pc) == tot), QH_EINVAL,
             "qh_gmres: the Hamiltonian-eigenbasis preconditioner must match the Hilbert dimension");
  QH_REQUIRE(!aug || (aug->prec == QH_FP64 && aug->n == A->rows), QH_EINVAL,
             "qh_gmres: the augmentation vector must be an FP64 vector of the matrix size");
  const int64_t n = A->rows;
  const int R = restart;
  hipStream_t s = qh_stream();
  double2 *V = nullptr, *w = nullptr, *z = nullptr, *t = nullptr, *dinv = nullptr, *part = nullptr, *h = nullptr,
          *yv = nullptr, *cu = nullptr, *crem = nullptr;
  double2 *xd = (double2 *)x->d;
  rc = qh_alloc_t(&V, (size_t)n * (R + 1));
  if (!rc) rc = qh_alloc_t(&w, n);
  if (!rc) rc = qh_alloc_t(&z, n);
  if (!rc) rc = qh_alloc_t(&t, n);
  if (!rc) rc = qh_alloc_t(&part, (size_t)DOT_BLOCKS * (R + 2));
  if (!rc) rc = qh_alloc_t(&h, 2 * (R + 2) + 2);  // CGS2 dots, norms; g and L^2 of the augmentation
  if (!rc && aug) rc = qh_alloc_t(&cu, n);
  if (!rc && aug) rc = qh_alloc_t(&crem, n);
  if (!rc) rc = qh_alloc_t(&yv, R + 1);
  if (!rc && pck == 1) rc = qh_alloc_t(&dinv, n);
  if (rc) goto done;
  {
    if (pck == 1)
      hipLaunchKernelGGL(k_diag_inv, dim3(gr(n)), dim3(256), 0, s, A->rowptr, A->col, A->val, n, dinv);
    // out = M^-1 v
    auto precond = [&](const double2 *v, double2 *out) -> int {
      if (pck == 1) {
        hipLaunchKernelGGL(k_cmul_vec, dim3(gr(n)), dim3(256), 0, s, dinv, v, out, n);
        QH_HIP(hipGetLastError());
        return QH_OK;
      }
      if (pck == 2) return qh_pc_apply_dev(pc, v, out);
      QH_HIP(hipMemcpyAsync(out, v, 16 * n, hipMemcpyDeviceToDevice, s));
      return QH_OK;
    };
    // out = A' v  (A plus the trace functional on row 0)
    auto apply_a = [&](const double2 *v, double2 *out) -> int {
      qh_vec iv, ov;
      iv.n = ov.n = n;
      iv.prec = ov.prec = QH_FP64;
      iv.d = const_cast<double2 *>(v);
      ov.d = out;
      int r2 = qh_spmv(A, &iv, &ov);
      iv.d = ov.d = nullptr;
      if (r2) return r2;
      hipLaunchKernelGGL(k_add_trace_row0, dim3(1), dim3(256), 0, s, v, tot, out);
      QH_HIP(hipGetLastError());
      return QH_OK;
    };
    auto dots = [&](int m, const double2 *vec, double2 *hout) {
      hipLaunchKernelGGL(k_multidot, dim3(DOT_BLOCKS, m), dim3(256), 0, s, V, n, vec, n, part);
      hipLaunchKernelGGL(k_reduce_rows, dim3(m), dim3(256), 0, s, part, DOT_BLOCKS, hout);
    };
    auto norm2 = [&](const double2 *vec, double2 *hout) {
      // ||vec||^2 via the multidot kernel with V := vec
      hipLaunchKernelGGL(k_multidot, dim3(DOT_BLOCKS, 1), dim3(256), 0, s, vec, n, vec, n, part);
      hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(256), 0, s, part, DOT_BLOCKS, hout);
    };

    // the augmentation's g_j = v_j^H c_rem and ||c_rem||^2 after removing v_j, into
    // the two slots after the CGS2 ones
    double2 *const hg = h + 2 * (R + 2);
    auto aug_step = [&](const double2 *v) {
      hipLaunchKernelGGL(k_multidot, dim3(DOT_BLOCKS, 1), dim3(256), 0, s, v, n, crem, n, part);
      hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(256), 0, s, part, DOT_BLOCKS, hg);
      hipLaunchKernelGGL(k_multiaxpy, dim3(gr(n)), dim3(256), 0, s, v, n, 1, hg, crem, n);
      norm2(crem, hg + 1);
    };
    if (aug) {
      rc = apply_a((const double2 *)aug->d, cu);
      if (rc) goto done;
    }

    int64_t it = 0;
    double rn = 0.0;
    std::vector<cd> H((size_t)(R + 1) * R), cs(R), sn(R), g(R + 1), y(R), Gp(R + 1);
    std::vector<double2> hh(2 * (R + 2) + 2);
    double L2 = 0.0;
    for (;;) {
      rc = apply_a(xd, w);
      if (rc) break;
      hipLaunchKernelGGL(k_resid, dim3(gr(n)), dim3(256), 0, s, w, n);
      norm2(w, h);
      double2 hn2;
      QH_HIP(hipMemcpyAsync(&hn2, h, 16, hipMemcpyDeviceToHost, s));
      QH_HIP(hipStreamSynchronize(s));
      rn = sqrt(hn2.x);
      if (rn <= tol || it >= maxit) break;
      hipLaunchKernelGGL(k_normalize, dim3(gr(n)), dim3(256), 0, s, w, h, V, n);
      std::fill(H.begin(), H.end(), cd(0.0, 0.0));
      std::fill(g.begin(), g.end(), cd(0.0, 0.0));
      std::fill(Gp.begin(), Gp.end(), cd(0.0, 0.0));
      g[0] = rn;
      if (aug) {
        QH_HIP(hipMemcpyAsync(crem, cu, 16 * n, hipMemcpyDeviceToDevice, s));
        aug_step(V);
        QH_HIP(hipMemcpyAsync(hh.data() + 2 * (R + 2), hg, sizeof(double2) * 2, hipMemcpyDeviceToHost, s));
        QH_HIP(hipStreamSynchronize(s));
        Gp[0] = cd(hh[2 * (R + 2)].x, hh[2 * (R + 2)].y);
        L2 = hh[2 * (R + 2) + 1].x;
      }
      // the residual norm of the cycle's best combination so far
      auto cycle_res = [&](int kk) {
        const double rho = std::abs(g[kk]);
        if (!aug) return rho;
        const double den = std::norm(Gp[kk]) + L2;
        return den > 0.0 ? rho * sqrt(L2 / den) : rho;
      };
      int k = 0;
      for (; k < R && it < maxit; k++, it++) {
        rc = precond(V + n * k, z);
        if (!rc) rc = apply_a(z, w);
        if (rc) break;
        // CGS2: two passes of (all dots, all updates); h[0..k] and h[R+2 .. R+2+k]
        dots(k + 1, w, h);
        hipLaunchKernelGGL(k_multiaxpy, dim3(gr(n)), dim3(256), 0, s, V, n, k + 1, h, w, n);
        dots(k + 1, w, h + (R + 2));
        hipLaunchKernelGGL(k_multiaxpy, dim3(gr(n)), dim3(256), 0, s, V, n, k + 1, h + (R + 2), w, n);
        norm2(w, h + (k + 1));
        hipLaunchKernelGGL(k_normalize, dim3(gr(n)), dim3(256), 0, s, w, h + (k + 1), V + n * (k + 1), n);
        if (aug) aug_step(V + n * (k + 1));
        QH_HIP(hipMemcpyAsync(hh.data(), h, sizeof(double2) * (2 * (R + 2) + 2), hipMemcpyDeviceToHost, s));
        QH_HIP(hipStreamSynchronize(s));
        for (int jj = 0; jj <= k; jj++)
          H[(size_t)jj * R + k] = cd(hh[jj].x + hh[R + 2 + jj].x, hh[jj].y + hh[R + 2 + jj].y);
        H[(size_t)(k + 1) * R + k] = cd(sqrt(hh[k + 1].x), 0.0);
        for (int jj = 0; jj < k; jj++) {
          cd a = H[(size_t)jj * R + k], b = H[(size_t)(jj + 1) * R + k];
          H[(size_t)jj * R + k] = std::conj(cs[jj]) * a + std::conj(sn[jj]) * b;
          H[(size_t)(jj + 1) * R + k] = -sn[jj] * a + cs[jj] * b;
        }
        cd a = H[(size_t)k * R + k], b = H[(size_t)(k + 1) * R + k];
        double den = sqrt(std::norm(a) + std::norm(b));
        if (den == 0.0) {
          cs[k] = 1.0;
          sn[k] = 0.0;
        } else {
          cs[k] = a / den;
          sn[k] = b / den;
        }
        cd gk = g[k];
        g[k] = std::conj(cs[k]) * gk;
        g[k + 1] = -sn[k] * gk;
        H[(size_t)k * R + k] = std::conj(cs[k]) * a + std::conj(sn[k]) * b;
        H[(size_t)(k + 1) * R + k] = 0.0;
        if (aug) {
          // the new entry of g, then rotation k on its rows k, k+1
          const cd gnew(hh[2 * (R + 2)].x, hh[2 * (R + 2)].y), gp = Gp[k];
          Gp[k] = std::conj(cs[k]) * gp + std::conj(sn[k]) * gnew;
          Gp[k + 1] = -sn[k] * gp + cs[k] * gnew;
          L2 = hh[2 * (R + 2) + 1].x;
        }
        if (cycle_res(k + 1) <= tol) {
          k++;
          it++;
          break;
        }
      }
      if (rc) break;
      // the augmentation's coefficient, and the right-hand side it leaves for y
      cd alpha(0.0, 0.0);
      if (aug) {
        const double den = std::norm(Gp[k]) + L2;
        if (den > 0.0) alpha = std::conj(Gp[k]) * g[k] / den;
      }
      for (int i = k - 1; i >= 0; i--) {
        cd sacc = g[i] - Gp[i] * alpha;
        for (int jj = i + 1; jj < k; jj++) sacc -= H[(size_t)i * R + jj] * y[jj];
        y[i] = sacc / H[(size_t)i * R + i];
      }
      if (aug && alpha != cd(0.0, 0.0)) {
        // x += alpha u
        const double2 ah = make_double2(alpha.real(), alpha.imag());
        QH_HIP(hipMemcpyAsync(yv, &ah, sizeof(double2), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_update, dim3(gr(n)), dim3(256), 0, s, xd, (const double2 *)aug->d, n, 1, yv, n);
        QH_HIP(hipStreamSynchronize(s));
      }
      if (k > 0) {
        // x += M^-1 (V y)
        std::vector<double2> yh(k);
        for (int i = 0; i < k; i++) yh[i] = make_double2(y[i].real(), y[i].imag());
        QH_HIP(hipMemcpyAsync(yv, yh.data(), sizeof(double2) * k, hipMemcpyHostToDevice, s));
        QH_HIP(hipMemsetAsync(t, 0, 16 * n, s));
        hipLaunchKernelGGL(k_update, dim3(gr(n)), dim3(256), 0, s, t, V, n, k, yv, n);
        rc = precond(t, z);
        if (rc) break;
        hipLaunchKernelGGL(k_axpy1, dim3(gr(n)), dim3(256), 0, s, z, xd, n);
        QH_HIP(hipStreamSynchronize(s));
      }
    }
    if (!rc) {
      QH_HIP(hipStreamSynchronize(s));
      if (iters) *iters = it;
      if (res) *res = rn;
    }
  }
done:
  qh_free_ptr(V);
  qh_free_ptr(w);
  qh_free_ptr(z);
  qh_free_ptr(t);
  qh_free_ptr(dinv);
  qh_free_ptr(part);
  qh_free_ptr(h);
  qh_free_ptr(yv);
  qh_free_ptr(cu);
  qh_free_ptr(crem);
  return rc;
}

extern "C" int qh_gmres(const qh_mat *A, int64_t tot, qh_vec *x, int restart, double tol, int64_t maxit, int jacobi,
                        int64_t *iters, double *res) {
  return gmres_impl(A, tot, x, nullptr, restart, tol, maxit, jacobi ? 1 : 0, nullptr, iters, res);
}

extern "C" int qh_gmres_pc(const qh_mat *A, int64_t tot, qh_vec *x, int restart, double tol, int64_t maxit,
                           int pc_kind, const qh_pc *pc, int64_t *iters, double *res) {
  return gmres_impl(A, tot, x, nullptr, restart, tol, maxit, pc_kind, pc, iters, res);
}

extern "C" int qh_gmres_aug(const qh_mat *A, int64_t tot, qh_vec *x, const qh_vec *u, int restart, double tol,
                            int64_t maxit, int pc_kind, const qh_pc *pc, int64_t *iters, double *res) {
  return gmres_impl(A, tot, x, u, restart, tol, maxit, pc_kind, pc, iters, res);
}
