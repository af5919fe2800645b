// qh_mcwf.hip -- Monte Carlo wavefunction trajectories advanced as one ensemble.
//
// The ensemble is one FP64 vector of M trajectories x N amplitudes
// (trajectory i at entries [i*N, (i+1)*N)).  All trajectories share the N x N
// MCWF generator A = -i Heff (construct_matrix under use_mcwf_solver), so an RK4
// step is four launches of one stage kernel over all M*N entries: each thread
// owns one (trajectory, row), reads the row of A (shared by every trajectory,
// hence cache resident) and gathers within its own trajectory.  Per-trajectory
// squared norms (the jump test of _sys_MCWF_EventFunction, qsystem.c:1138-1153)
// come from one block per trajectory.  The jump itself (choice of the operator
// and renormalisation, qsystem.c:1159-1235) is decided on the host.
#include <vector>

#include "qh_common.h"

namespace {

enum { ENS_PLAIN = 0, ENS_FINAL = 1 };

// MODE plain:  kout = A (x + ck kin)          (per trajectory)
// MODE final:  xout = x + h/6 (k1 + 2 k2 + 2 k3 + A (x + ck kin))
template <int MODE>
__global__ void __launch_bounds__(256) k_ens_stage(int32_t N, int64_t total, const int32_t *__restrict__ rp,
                                                   const int32_t *__restrict__ col, const double2 *__restrict__ val,
                                                   const double2 *__restrict__ x, const double2 *__restrict__ kin,
                                                   double ck, double2 *__restrict__ kout,
                                                   const double2 *__restrict__ k1, const double2 *__restrict__ k2,
                                                   const double2 *__restrict__ k3, double2 *__restrict__ xout,
                                                   double h) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t base = (t / N) * N;
    const int32_t r = (int32_t)(t - base);
    double are = 0.0, aim = 0.0;
    for (int32_t k = rp[r]; k < rp[r + 1]; k++) {
      const int64_t c = base + col[k];
      double2 v = x[c];
      if (kin) {
        const double2 q = kin[c];
        v.x += ck * q.x;
        v.y += ck * q.y;
      }
      const double2 a = val[k];
      are = fma(a.x, v.x, fma(-a.y, v.y, are));
      aim = fma(a.x, v.y, fma(a.y, v.x, aim));
    }
    if (MODE == ENS_PLAIN) {
      kout[t] = make_double2(are, aim);
    } else {
      const double s6 = h * (1.0 / 6.0), s3 = h * (1.0 / 3.0);
      const double2 a = k1[t], b = k2[t], c = k3[t], xo = x[t];
      xout[t] = make_double2(xo.x + s6 * a.x + s3 * b.x + s3 * c.x + s6 * are,
                             xo.y + s6 * a.y + s3 * b.y + s3 * c.y + s6 * aim);
    }
  }
}

// norms2[i] = sum_r |psi_i[r]|^2, one block per trajectory
__global__ void __launch_bounds__(256) k_ens_norms(const double2 *__restrict__ e, int64_t N, double *norms2) {
  __shared__ double red[256];
  const double2 *p = e + (int64_t)blockIdx.x * N;
  double acc = 0.0;
  for (int64_t r = threadIdx.x; r < N; r += 256) acc += p[r].x * p[r].x + p[r].y * p[r].y;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) norms2[blockIdx.x] = red[0];
}

// every trajectory <- the same initial state
__global__ void k_ens_broadcast(const double2 *__restrict__ src, int64_t N, int64_t total, double2 *__restrict__ e) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
    e[t] = src[t % N];
}

unsigned eg(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 16384) b = 16384;
  return (unsigned)(b < 1 ? 1 : b);
}

// stage workspace, kept between calls (k1, k2, k3, alternate state)
struct EnsWork {
  int64_t n = 0;
  double2 *buf = nullptr;
  double *norms = nullptr;
  int64_t nn = 0;
} g_ew;

int ens_work(int64_t n) {
  if (g_ew.n >= n) return QH_OK;
  qh_free_ptr(g_ew.buf);
  g_ew.buf = nullptr;
  g_ew.n = 0;
  int rc = qh_alloc_t(&g_ew.buf, (size_t)4 * n);
  if (rc) return rc;
  g_ew.n = n;
  return QH_OK;
}

}  // namespace

extern "C" int qh_ens_broadcast(const qh_vec *init, qh_vec *ens, int64_t M) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(init && ens && init->prec == QH_FP64 && ens->prec == QH_FP64 && M >= 1 && ens->n == M * init->n,
             QH_EINVAL, "qh_ens_broadcast: FP64 ensemble of M copies of the initial state required");
  hipLaunchKernelGGL(k_ens_broadcast, dim3(eg(ens->n)), dim3(256), 0, qh_stream(), (const double2 *)init->d,
                     init->n, ens->n, (double2 *)ens->d);
  QH_HIP(hipGetLastError());
  return QH_OK;
}

// nsteps classic RK4 steps of size dt on every trajectory (A: N x N, ens: M*N)
extern "C" int qh_ens_rk4(const qh_mat *A, qh_vec *ens, int64_t M, double dt, int64_t nsteps) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && ens && ens->prec == QH_FP64 && M >= 1 && ens->n == M * A->rows && A->rows <= (1 << 30),
             QH_EINVAL, "qh_ens_rk4: FP64 ensemble of M trajectories of the matrix dimension required");
  QH_REQUIRE(nsteps >= 0, QH_EINVAL, "qh_ens_rk4: negative step count");
  const int64_t n = ens->n;
  rc = ens_work(n);
  if (rc) return rc;
  hipStream_t s = qh_stream();
  double2 *k1 = g_ew.buf, *k2 = k1 + n, *k3 = k2 + n, *alt = k3 + n;
  double2 *x = (double2 *)ens->d;
  double2 *cur = x, *nxt = alt;
  const int32_t N = (int32_t)A->rows;
  const unsigned g = eg(n);
  for (int64_t st = 0; st < nsteps; st++) {
    hipLaunchKernelGGL(k_ens_stage<ENS_PLAIN>, dim3(g), dim3(256), 0, s, N, n, A->rowptr, A->col, A->val, cur,
                       (const double2 *)nullptr, 0.0, k1, nullptr, nullptr, nullptr, nullptr, 0.0);
    hipLaunchKernelGGL(k_ens_stage<ENS_PLAIN>, dim3(g), dim3(256), 0, s, N, n, A->rowptr, A->col, A->val, cur, k1,
                       0.5 * dt, k2, nullptr, nullptr, nullptr, nullptr, 0.0);
    hipLaunchKernelGGL(k_ens_stage<ENS_PLAIN>, dim3(g), dim3(256), 0, s, N, n, A->rowptr, A->col, A->val, cur, k2,
                       0.5 * dt, k3, nullptr, nullptr, nullptr, nullptr, 0.0);
    hipLaunchKernelGGL(k_ens_stage<ENS_FINAL>, dim3(g), dim3(256), 0, s, N, n, A->rowptr, A->col, A->val, cur, k3,
                       dt, nullptr, k1, k2, k3, nxt, dt);
    double2 *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  QH_HIP(hipGetLastError());
  if (cur != x) QH_HIP(hipMemcpyAsync(x, cur, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
  return QH_OK;
}

// squared norm of each of the M trajectories (host array of M)
extern "C" int qh_ens_norms(const qh_vec *ens, int64_t M, double *norms2) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(ens && norms2 && ens->prec == QH_FP64 && M >= 1 && ens->n % M == 0 && M <= (1 << 30), QH_EINVAL,
             "qh_ens_norms: FP64 ensemble of M equal trajectories required");
  hipStream_t s = qh_stream();
  if (g_ew.nn < M) {
    qh_free_ptr(g_ew.norms);
    g_ew.norms = nullptr;
    g_ew.nn = 0;
    rc = qh_alloc_t(&g_ew.norms, (size_t)M);
    if (rc) return rc;
    g_ew.nn = M;
  }
  hipLaunchKernelGGL(k_ens_norms, dim3((unsigned)M), dim3(256), 0, s, (const double2 *)ens->d, ens->n / M,
                     g_ew.norms);
  QH_HIP(hipGetLastError());
  QH_HIP(hipMemcpyAsync(norms2, g_ew.norms, sizeof(double) * M, hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  return QH_OK;
}
