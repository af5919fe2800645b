// qh_precond.hip -- the Hamiltonian-eigenbasis preconditioner of the
// steady-state GMRES solve.
//
// The Liouvillian of a density-matrix system splits as L = L0 + J, where
//   L0 rho = -i Heff rho + i rho Heff^dag,   Heff = H - (i/2) sum_k a_k C_k^dag C_k
// is the no-jump part and J rho = sum_k a_k C_k rho C_k^dag the jumps.  In the
// column-major vec layout L0 is the Kronecker sum I (x) (-i Heff) + conj(Heff) (x) I,
// so with Heff ~= W (E + i diag(g)) W^dag, W the eigenvectors of the Hermitian part
// of Heff and g_i = <w_i| (Heff - Heff^dag)/(2i) |w_i>, the operator
//   M rho = W [ (W^dag rho W) .* Dm ] W^dag,   Dm_ij = (g_i + g_j) - i (E_i - E_j)
// approximates L0 and is inverted exactly by dividing by Dm instead.  GMRES is
// then run on L' M^-1 (right preconditioning): the jump part and the dropped
// off-diagonal decay couplings are what is left for the Krylov iteration.
//
// Set-up: Heff is assembled on the device as an N x N CSR by the caller,
// densified, split into Hermitian / anti-Hermitian parts, and the Hermitian part
// diagonalised with rocSOLVER.  When that part is real (a Hamiltonian with real
// matrix elements, the usual case) the real symmetric divide-and-conquer solver
// runs first (dsyevd; 11.5 ms at N = 512 on an MI355X), else the complex Jacobi /
// divide-and-conquer solver (zheevdj; 24 ms); the others follow as fallbacks
// (zheevd returned non-finite vectors for N >= 216 with this rocSOLVER, and
// zheevj takes 50 ms at N = 512).  Every result is checked (finite, W^dag W = I
// to 1e-8) before it is used.  Each application costs four N^3 complex GEMMs
// (rocBLAS) and one elementwise pass, or, with a real eigenbasis, four real
// GEMMs of 2N x N by N x N, two transposes and the elementwise pass.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "qh_common.h"

struct qh_pc {
  int64_t n = 0;             // Hilbert dimension N (vectors have N*N entries)
  double2 *W = nullptr;      // eigenvectors, N x N column-major
  double2 *D = nullptr;      // denominators Dm, N x N column-major
  double2 *T = nullptr;      // workspace N x N
  // real eigenbasis (a Hamiltonian with real matrix elements): W as real N x N,
  // Dm transposed, a second workspace; the apply then runs real GEMMs
  double *Wr = nullptr;
  double2 *Dt = nullptr;
  double2 *T2 = nullptr;
  rocblas_handle blas = nullptr;
};

namespace {

#define QH_RB(call)                                                                   \
  do {                                                                                \
    rocblas_status st_ = (call);                                                      \
    if (st_ != rocblas_status_success) {                                              \
      qh_set_error("%s:%d %s: rocBLAS status %d", __FILE__, __LINE__, #call, (int)st_); \
      return QH_EHIP;                                                                 \
    }                                                                                 \
  } while (0)

unsigned grid1(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  return (unsigned)(b < 1 ? 1 : b);
}

__global__ void k_fill0(double2 *a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = make_double2(0.0, 0.0);
}

// dense column-major copy of an N x N CSR matrix
__global__ void k_densify(const int32_t *rp, const int32_t *col, const double2 *val, int64_t n, double2 *A) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int32_t k = rp[r]; k < rp[r + 1]; k++) A[r + n * (int64_t)col[k]] = val[k];
}

// Hermitian part (A + A^dag)/2 into Hh, anti-Hermitian part (A - A^dag)/(2i) into Ha
__global__ void k_split_herm(const double2 *A, int64_t n, double2 *Hh, double2 *Ha) {
  const int64_t total = n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n, j = t / n;
    const double2 a = A[i + n * j], b = A[j + n * i];  // b^* = (A^dag)_ij
    Hh[t] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    // (a - conj(b)) / (2i) = ((a.y + b.y) - i (a.x - b.x)) / 2
    Ha[t] = make_double2(0.5 * (a.y + b.y), -0.5 * (a.x - b.x));
  }
}

// g_i = Re sum_r conj(W[r,i]) T[r,i]   (T = Ha W); one block per column
__global__ void __launch_bounds__(256) k_coldot(const double2 *W, const double2 *T, int64_t n, double *g) {
  __shared__ double red[256];
  const int64_t i = blockIdx.x;
  double acc = 0.0;
  for (int64_t r = threadIdx.x; r < n; r += 256) {
    const double2 w = W[r + n * i], t = T[r + n * i];
    acc += w.x * t.x + w.y * t.y;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) g[i] = red[0];
}

// Dm_ij = (g_i + g_j) - i (E_i - E_j), with |Dm| floored at floor_ (kept on the
// stable side: a non-positive real part)
__global__ void k_denoms(const double *E, const double *g, int64_t n, double floor_, double2 *D) {
  const int64_t total = n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n, j = t / n;
    double re = g[i] + g[j], im = -(E[i] - E[j]);
    if (re * re + im * im < floor_ * floor_) re = -floor_;
    D[t] = make_double2(re, im);
  }
}

__global__ void k_cdiv_inplace(double2 *a, const double2 *d, int64_t total) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = a[t], y = d[t];
    const double den = y.x * y.x + y.y * y.y;
    a[t] = make_double2((x.x * y.x + x.y * y.y) / den, (x.y * y.x - x.x * y.y) / den);
  }
}

// flag[0] = 1 if any eigenvalue or eigenvector entry is not finite
__global__ void k_nonfinite(const double2 *W, const double *E, int64_t n, int *flag) {
  const int64_t total = n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const double2 w = W[t];
    if (!isfinite(w.x) || !isfinite(w.y) || (t < n && !isfinite(E[t]))) flag[0] = 1;
  }
}

// flag[0] = 1 if W^dag W (in T) departs from the identity by more than tol
__global__ void k_orth_check(const double2 *T, int64_t n, double tol, int *flag) {
  const int64_t total = n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = T[t];
    const double re = v.x - ((t % n) == (t / n) ? 1.0 : 0.0);
    if (!(fabs(re) <= tol && fabs(v.y) <= tol)) flag[0] = 1;
  }
}

// flag[0] = 1 if any entry has a nonzero imaginary part
__global__ void k_imag_flag(const double2 *A, int64_t total, int *flag) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
    if (A[t].y != 0.0) flag[0] = 1;
}

__global__ void k_real_part(const double2 *A, int64_t total, double *R) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
    R[t] = A[t].x;
}

__global__ void k_to_complex(const double *R, int64_t total, double2 *W) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
    W[t] = make_double2(R[t], 0.0);
}

// out = in^T for N x N complex matrices (32 x 32 tiles through LDS, both sides
// coalesced); block (32, 8)
__global__ void __launch_bounds__(256) k_ctranspose(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                    int64_t n) {
  __shared__ double2 tile[32][33];
  const int64_t bx = (int64_t)blockIdx.x * 32, by = (int64_t)blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int64_t r = bx + threadIdx.x, c = by + k;
    if (r < n && c < n) tile[k][threadIdx.x] = in[r + n * c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int64_t r = by + threadIdx.x, c = bx + k;
    if (r < n && c < n) out[r + n * c] = tile[threadIdx.x][k];
  }
}

const rocblas_double_complex ONE(1.0, 0.0), ZERO(0.0, 0.0);

// eigensolvers in the order they are tried
enum { EIG_DSYEVD, EIG_DSYEVDJ, EIG_ZHEEVDJ, EIG_ZHEEVD, EIG_ZHEEVJ, EIG_ZHEEV };

}  // namespace

extern "C" int qh_pc_destroy(qh_pc *pc) {
  if (!pc) return QH_OK;
  qh_free_ptr(pc->W);
  qh_free_ptr(pc->D);
  qh_free_ptr(pc->T);
  qh_free_ptr(pc->Wr);
  qh_free_ptr(pc->Dt);
  qh_free_ptr(pc->T2);
  if (pc->blas) rocblas_destroy_handle(pc->blas);
  delete pc;
  return QH_OK;
}

extern "C" int qh_pc_create_hbasis(const qh_mat *heff, qh_pc **out) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(heff && out && heff->rows > 0 && heff->rows <= 16384, QH_EINVAL,
             "qh_pc_create_hbasis: N x N effective Hamiltonian with N in [1, 16384] required");
  *out = nullptr;
  const int64_t n = heff->rows, nn = n * n;
  hipStream_t s = qh_stream();
  qh_pc *pc = new qh_pc();
  pc->n = n;
  double2 *A = nullptr, *Ha = nullptr;
  double *E = nullptr, *work = nullptr, *g = nullptr;
  rocblas_int *info = nullptr;
  rc = qh_alloc_t(&pc->W, nn);
  if (!rc) rc = qh_alloc_t(&pc->D, nn);
  if (!rc) rc = qh_alloc_t(&pc->T, nn);
  if (!rc) rc = qh_alloc_t(&A, nn);
  if (!rc) rc = qh_alloc_t(&Ha, nn);
  if (!rc) rc = qh_alloc_t(&E, n);
  if (!rc) rc = qh_alloc_t(&work, n);
  if (!rc) rc = qh_alloc_t(&g, n);
  if (!rc) rc = qh_alloc_t(&info, 2);
  auto cleanup = [&]() {
    qh_free_ptr(A);
    qh_free_ptr(Ha);
    qh_free_ptr(E);
    qh_free_ptr(work);
    qh_free_ptr(g);
    qh_free_ptr(info);
  };
  if (rc) {
    cleanup();
    qh_pc_destroy(pc);
    return rc;
  }
  auto fail = [&](int code) {
    cleanup();
    qh_pc_destroy(pc);
    return code;
  };
  if (rocblas_create_handle(&pc->blas) != rocblas_status_success) {
    qh_set_error("rocblas_create_handle failed");
    return fail(QH_EHIP);
  }
  if (rocblas_set_stream(pc->blas, s) != rocblas_status_success) {
    qh_set_error("rocblas_set_stream failed");
    return fail(QH_EHIP);
  }
  hipLaunchKernelGGL(k_fill0, dim3(grid1(nn)), dim3(256), 0, s, A, nn);
  hipLaunchKernelGGL(k_densify, dim3(grid1(n)), dim3(256), 0, s, heff->rowptr, heff->col, heff->val, n, A);
  hipLaunchKernelGGL(k_split_herm, dim3(grid1(nn)), dim3(256), 0, s, A, n, pc->W, Ha);
  if (hipGetLastError() != hipSuccess) {
    qh_set_error("qh_pc_create_hbasis: set-up kernel launch failed");
    return fail(QH_EHIP);
  }
  // Hermitian eigensolver: each candidate of the list runs on a fresh copy of
  // the Hermitian part (kept in A); the first whose result is finite and
  // orthonormal is used
  {
    int *flag = nullptr;
    double *R = nullptr;
    if (qh_alloc_t(&flag, 1)) return fail(QH_ENOMEM);
    QH_HIP(hipMemcpyAsync(A, pc->W, sizeof(double2) * nn, hipMemcpyDeviceToDevice, s));
    int himag = 0;
    QH_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_imag_flag, dim3(grid1(nn)), dim3(256), 0, s, A, nn, flag);
    QH_HIP(hipMemcpyAsync(&himag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    QH_HIP(hipStreamSynchronize(s));
    std::vector<int> methods;
    if (!himag) methods = {EIG_DSYEVD, EIG_DSYEVDJ};
    for (int m : {EIG_ZHEEVDJ, EIG_ZHEEVD, EIG_ZHEEVJ, EIG_ZHEEV}) methods.push_back(m);
    if (!himag && qh_alloc_t(&R, nn)) {
      qh_free_ptr(flag);
      return fail(QH_ENOMEM);
    }
    int ok = 0;
    for (size_t mi = 0; mi < methods.size() && !ok; mi++) {
      const int method = methods[mi];
      const bool real = method == EIG_DSYEVD || method == EIG_DSYEVDJ;
      rocblas_status st;
      rocblas_double_complex *Wd = (rocblas_double_complex *)pc->W;
      if (real) {
        hipLaunchKernelGGL(k_real_part, dim3(grid1(nn)), dim3(256), 0, s, A, nn, R);
        if (method == EIG_DSYEVD)
          st = rocsolver_dsyevd(pc->blas, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, R,
                                (rocblas_int)n, E, work, info);
        else
          st = rocsolver_dsyevdj(pc->blas, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, R,
                                 (rocblas_int)n, E, info);
        hipLaunchKernelGGL(k_to_complex, dim3(grid1(nn)), dim3(256), 0, s, R, nn, pc->W);
      } else {
        QH_HIP(hipMemcpyAsync(pc->W, A, sizeof(double2) * nn, hipMemcpyDeviceToDevice, s));
        if (method == EIG_ZHEEVDJ)
          st = rocsolver_zheevdj(pc->blas, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, Wd,
                                 (rocblas_int)n, E, info);
        else if (method == EIG_ZHEEVD)
          st = rocsolver_zheevd(pc->blas, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, Wd,
                                (rocblas_int)n, E, work, info);
        else if (method == EIG_ZHEEVJ)
          st = rocsolver_zheevj(pc->blas, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_lower,
                                (rocblas_int)n, Wd, (rocblas_int)n, 0.0, work, 100, info + 1, E, info);
        else
          st = rocsolver_zheev(pc->blas, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, Wd,
                               (rocblas_int)n, E, work, info);
      }
      if (st != rocblas_status_success) continue;
      int hinfo0 = 0, hflag = 0;
      QH_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
      hipLaunchKernelGGL(k_nonfinite, dim3(grid1(nn)), dim3(256), 0, s, pc->W, E, n, flag);
      if (rocblas_zgemm(pc->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, (rocblas_int)n,
                        (rocblas_int)n, (rocblas_int)n, &ONE, Wd, (rocblas_int)n, Wd, (rocblas_int)n, &ZERO,
                        (rocblas_double_complex *)pc->T, (rocblas_int)n) != rocblas_status_success)
        continue;
      hipLaunchKernelGGL(k_orth_check, dim3(grid1(nn)), dim3(256), 0, s, pc->T, n, 1e-8, flag);
      QH_HIP(hipMemcpyAsync(&hinfo0, info, sizeof(int), hipMemcpyDeviceToHost, s));
      QH_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
      QH_HIP(hipStreamSynchronize(s));
      ok = (hinfo0 == 0 && hflag == 0);
      if (ok && real) {  // keep the real eigenvectors for the real-GEMM apply
        pc->Wr = R;
        R = nullptr;
      }
    }
    qh_free_ptr(flag);
    qh_free_ptr(R);
    if (!ok) {
      qh_set_error("qh_pc_create_hbasis: no Hermitian eigensolver produced a finite eigenbasis");
      return fail(QH_EHIP);
    }
  }
  // g_i = <w_i| Ha |w_i>
  if (rocblas_zgemm(pc->blas, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                    (rocblas_int)n, &ONE, (const rocblas_double_complex *)Ha, (rocblas_int)n,
                    (const rocblas_double_complex *)pc->W, (rocblas_int)n, &ZERO, (rocblas_double_complex *)A,
                    (rocblas_int)n) != rocblas_status_success) {
    qh_set_error("rocblas_zgemm failed");
    return fail(QH_EHIP);
  }
  hipLaunchKernelGGL(k_coldot, dim3((unsigned)n), dim3(256), 0, s, pc->W, A, n, g);
  int hinfo = 0;
  std::vector<double> hE(n), hg(n);
  if (hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(hE.data(), E, sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(hg.data(), g, sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    qh_set_error("qh_pc_create_hbasis: device copy failed");
    return fail(QH_EHIP);
  }
  if (hinfo != 0) {
    qh_set_error("qh_pc_create_hbasis: the eigensolver did not converge (info %d)", hinfo);
    return fail(QH_EHIP);
  }
  // floor for |Dm|: a small fraction of the operator's scale
  double scale = 0.0;
  for (int64_t i = 0; i < n; i++) scale = std::max(scale, std::max(std::fabs(hE[i]), std::fabs(hg[i])));
  const double floor_ = (scale > 0.0 ? scale : 1.0) * 1e-10;
  hipLaunchKernelGGL(k_denoms, dim3(grid1(nn)), dim3(256), 0, s, E, g, n, floor_, pc->D);
  if (pc->Wr) {
    if (qh_alloc_t(&pc->Dt, nn) || qh_alloc_t(&pc->T2, nn)) return fail(QH_ENOMEM);
    const dim3 tg((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
    hipLaunchKernelGGL(k_ctranspose, tg, dim3(32, 8), 0, s, pc->D, pc->Dt, n);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    qh_set_error("qh_pc_create_hbasis: denominator kernel failed");
    return fail(QH_EHIP);
  }
  cleanup();
  *out = pc;
  return QH_OK;
}

// Eigenvalues w (ascending) and, if V, orthonormal eigenvectors of the real
// symmetric n x n host matrix A (either storage order; V column-major, column
// k the eigenvector of w[k]) on the device: rocSOLVER dsyevd, then dsyevdj if
// the result is not finite.
extern "C" int qh_sym_eig(int64_t n, const double *A, double *w, double *V) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && w && n >= 1 && n <= 32768, QH_EINVAL, "qh_sym_eig: 1 <= n <= 32768 required");
  hipStream_t s = qh_stream();
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success || rocblas_set_stream(h, s) != rocblas_status_success) {
    if (h) rocblas_destroy_handle(h);
    qh_set_error("qh_sym_eig: rocBLAS handle");
    return QH_EHIP;
  }
  double *dA = nullptr, *dw = nullptr, *de = nullptr;
  rocblas_int *info = nullptr;
  rc = qh_alloc_t(&dA, (size_t)(n * n));
  if (!rc) rc = qh_alloc_t(&dw, (size_t)n);
  if (!rc) rc = qh_alloc_t(&de, (size_t)n);
  if (!rc) rc = qh_alloc_t(&info, 1);
  std::vector<double> hv;
  int ok = 0;
  for (int method = 0; method < 2 && !rc && !ok; method++) {
    if (hipMemcpyAsync(dA, A, sizeof(double) * n * n, hipMemcpyHostToDevice, s) != hipSuccess) {
      rc = QH_EHIP;
      break;
    }
    const rocblas_evect ev = V ? rocblas_evect_original : rocblas_evect_none;
    rocblas_status st = method == 0
                            ? rocsolver_dsyevd(h, ev, rocblas_fill_lower, (rocblas_int)n, dA, (rocblas_int)n, dw, de, info)
                            : rocsolver_dsyevdj(h, ev, rocblas_fill_lower, (rocblas_int)n, dA, (rocblas_int)n, dw, info);
    if (st != rocblas_status_success) continue;
    int hinfo = 1;
    if (hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(w, dw, sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        (V && hipMemcpyAsync(V, dA, sizeof(double) * n * n, hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = QH_EHIP;
      break;
    }
    ok = hinfo == 0;
    for (int64_t i = 0; ok && i < n; i++) ok = std::isfinite(w[i]);
    for (int64_t i = 0; ok && V && i < n * n; i++) ok = std::isfinite(V[i]);
  }
  qh_free_ptr(dA);
  qh_free_ptr(dw);
  qh_free_ptr(de);
  qh_free_ptr(info);
  rocblas_destroy_handle(h);
  if (rc) {
    qh_set_error("qh_sym_eig: device copy failed");
    return rc;
  }
  QH_REQUIRE(ok, QH_EHIP, "qh_sym_eig: no eigensolver produced a finite result");
  return QH_OK;
}

// z = M^-1 v for vec layouts of N x N matrices (column-major, loc = N*col + row)
int qh_pc_apply_dev(const qh_pc *pc, const double2 *v, double2 *z) {
  const rocblas_int n = (rocblas_int)pc->n;
  hipStream_t s = qh_stream();
  if (pc->Wr) {
    // W real: a complex N x N matrix column-major is a real 2N x N matrix
    // (rows re, im interleaved), and X W = its real product with W.  With
    // Y = W^T V W and Z = Y ./ Dm:  U = V W;  A = (U^T W) ./ Dm^T = Z^T;
    // B = A W^T = (W Z)^T;  z = B^T W^T = W Z W^T.
    const double one = 1.0, zero = 0.0;
    const rocblas_int m2 = 2 * n;
    const dim3 tg((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
    double *Z = (double *)z, *T = (double *)pc->T, *T2 = (double *)pc->T2;
    QH_RB(rocblas_dgemm(pc->blas, rocblas_operation_none, rocblas_operation_none, m2, n, n, &one,
                        (const double *)v, m2, pc->Wr, n, &zero, Z, m2));
    hipLaunchKernelGGL(k_ctranspose, tg, dim3(32, 8), 0, s, z, pc->T, pc->n);
    QH_RB(rocblas_dgemm(pc->blas, rocblas_operation_none, rocblas_operation_none, m2, n, n, &one, T, m2, pc->Wr, n,
                        &zero, Z, m2));
    hipLaunchKernelGGL(k_cdiv_inplace, dim3(grid1(pc->n * pc->n)), dim3(256), 0, s, z, pc->Dt, pc->n * pc->n);
    QH_RB(rocblas_dgemm(pc->blas, rocblas_operation_none, rocblas_operation_transpose, m2, n, n, &one, Z, m2,
                        pc->Wr, n, &zero, T, m2));
    hipLaunchKernelGGL(k_ctranspose, tg, dim3(32, 8), 0, s, pc->T, pc->T2, pc->n);
    QH_RB(rocblas_dgemm(pc->blas, rocblas_operation_none, rocblas_operation_transpose, m2, n, n, &one, T2, m2,
                        pc->Wr, n, &zero, Z, m2));
    QH_HIP(hipGetLastError());
    return QH_OK;
  }
  const rocblas_double_complex *W = (const rocblas_double_complex *)pc->W;
  rocblas_double_complex *T = (rocblas_double_complex *)pc->T;
  rocblas_double_complex *Z = (rocblas_double_complex *)z;
  // T = W^dag V ; Z = T W ; Z ./= Dm ; T = W Z ; Z = T W^dag
  QH_RB(rocblas_zgemm(pc->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, n, n, n, &ONE, W, n,
                      (const rocblas_double_complex *)v, n, &ZERO, T, n));
  QH_RB(rocblas_zgemm(pc->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &ONE, T, n, W, n, &ZERO, Z,
                      n));
  hipLaunchKernelGGL(k_cdiv_inplace, dim3(grid1(pc->n * pc->n)), dim3(256), 0, s, z, pc->D, pc->n * pc->n);
  QH_RB(rocblas_zgemm(pc->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &ONE, W, n, Z, n, &ZERO, T,
                      n));
  QH_RB(rocblas_zgemm(pc->blas, rocblas_operation_none, rocblas_operation_conjugate_transpose, n, n, n, &ONE, T, n, W,
                      n, &ZERO, Z, n));
  QH_HIP(hipGetLastError());
  return QH_OK;
}

int64_t qh_pc_dim(const qh_pc *pc) { return pc ? pc->n : 0; }

extern "C" int qh_pc_apply(const qh_pc *pc, const qh_vec *in, qh_vec *out) {
  QH_REQUIRE(pc && in && out && in != out && in->prec == QH_FP64 && out->prec == QH_FP64 &&
                 in->n == pc->n * pc->n && out->n == in->n,
             QH_EINVAL, "qh_pc_apply: two distinct FP64 vectors of N*N entries required");
  return qh_pc_apply_dev(pc, (const double2 *)in->d, (double2 *)out->d);
}
