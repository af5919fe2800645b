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
#include <cstring>
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

// ----------------------------------------------- trajectories of up to 1024 amplitudes
// A wavefront integrates 64 / S trajectories side by side, one segment of S
// lanes per trajectory (S a power of two >= N, or S = 0: segments of exactly N
// lanes, floor(64 / N) of them, where that fits more trajectories into a
// wavefront than the power of two does; the lanes past the last segment idle):
// lane r of a segment holds
// amplitude r and row r of the generator (NZ entries in registers), and
// the products exchange the amplitudes through the segment's part of the
// wavefront's LDS row.  A trajectory runs classic RK4 steps (size dt, the last
// of the run hlast) from its step index jst[i] up to J and stops right after a
// step at which its squared norm fell below thr[i] (the jump test of
// _sys_MCWF_EventFunction, qsystem.c:1138-1153): jump[i] = 1.  The wavefront
// steps until all its trajectories have stopped; a stopped segment keeps its
// amplitudes.
const int ENS_W = 64;
const int ENS_MAXNZ = 16;

// row product over all NZ register entries (entries past the row's end have
// coefficient zero and point at an amplitude of the segment), so that the LDS
// reads of a stage are issued together
template <int NZ>
__device__ inline double2 ens_row_product(const int32_t *c, const double2 *a, const double2 *s) {
  double2 q[NZ];
#pragma unroll
  for (int e = 0; e < NZ; e++) q[e] = s[c[e]];
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int e = 0; e < NZ; e++) {
    re = fma(a[e].x, q[e].x, fma(-a[e].y, q[e].y, re));
    im = fma(a[e].x, q[e].y, fma(a[e].y, q[e].x, im));
  }
  return make_double2(re, im);
}

// v of the lane that DPP control CTRL selects within the lane's row of 16
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// sum of v over the aligned segment of S lanes, the bit-identical value in every
// lane of the segment (each exchange adds two partial sums that the partner
// lane adds in the other order): DPP quad_perm [1,0,3,2], [2,3,0,1], the
// half-row mirror and the row mirror inside a row of 16, then lane shuffles
template <int S>
__device__ inline double seg_sum(double v) {
  if (S >= 2) v += dpp_f64<0xB1>(v);
  if (S >= 4) v += dpp_f64<0x4E>(v);
  if (S >= 8) v += dpp_f64<0x141>(v);
  if (S >= 16) v += dpp_f64<0x140>(v);
  if (S >= 32) v += __shfl_xor(v, 16, 64);
  if (S >= 64) v += __shfl_xor(v, 32, 64);
  return v;
}

// v of the lane DPP control CTRL selects, 0 where the source lane lies outside
// the row or the lane's row is not in ROWS
template <int CTRL, int ROWS>
__device__ inline double dpp_f64_or0(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xF, false);
  return __hiloint2double(hi, lo);
}

// inclusive prefix sum of v over the 64 lanes: Hillis-Steele inside each row of
// 16 (row_shr 1, 2, 4, 8), then the row totals carried forward (row_bcast 15
// into rows 1 and 3, row_bcast 31 into rows 2 and 3)
__device__ inline double wave_prefix_f64(double v) {
  v += dpp_f64_or0<0x111, 0xF>(v);
  v += dpp_f64_or0<0x112, 0xF>(v);
  v += dpp_f64_or0<0x114, 0xF>(v);
  v += dpp_f64_or0<0x118, 0xF>(v);
  v += dpp_f64_or0<0x142, 0xA>(v);
  v += dpp_f64_or0<0x143, 0xC>(v);
  return v;
}

// Location of a norm crossing inside an RK4 step.  On the linear system
// d psi/dt = A psi a classic RK4 step of size tau from psi0 is exactly
// P(tau) = v0 + tau v1 + tau^2 v2 + tau^3 v3 + tau^4 v4 with v_k = A^k psi0 / k!,
// so ||P(tau)||^2 = sum_m c_m tau^m with c_m = sum_{j+k=m} Re<v_j, v_k>, a
// degree-8 polynomial.  The crossing ||P(tau)||^2 = thr (the event
// norm - r = 0 of _sys_MCWF_EventFunction, qsystem.c:1138-1153, located by
// PETSc's event handler to 1e-9 at qsystem.c:1017-1019) is found on (0, h] by
// the Illinois variant of regula falsi, and the state there is P(tau): the RK4
// step re-taken to the crossing.
// lane partial of the nine c_m from the lane's parts of v0..v4
__device__ inline void cross_partials(const double2 v[5], double c[9]) {
#pragma unroll
  for (int m = 0; m < 9; m++) c[m] = 0.0;
#pragma unroll
  for (int j = 0; j < 5; j++)
#pragma unroll
    for (int k = j; k < 5; k++) c[j + k] += (j == k ? 1.0 : 2.0) * (v[j].x * v[k].x + v[j].y * v[k].y);
}
__device__ inline double poly8(const double c[9], double t) {
  double f = c[8];
#pragma unroll
  for (int m = 7; m >= 0; m--) f = f * t + c[m];
  return f;
}
// the first tau in (0, h] with ||P(tau)||^2 <= thr (h when the polynomial stays
// above thr by rounding although the stage arithmetic fell below it)
__device__ inline double cross_tau(const double c[9], double thr, double h) {
  double a = 0.0, fa = poly8(c, 0.0) - thr, b = h, fb = poly8(c, h) - thr;
  if (fb > 0.0 || fa <= 0.0) return h;
  int side = 0;
  for (int it = 0; it < 200 && b - a > 1e-15 * h; it++) {
    const double t = b - fb * (b - a) / (fb - fa);
    const double ft = poly8(c, t) - thr;
    if (ft > 0.0) {
      a = t;
      fa = ft;
      if (side == 1) fb *= 0.5;
      side = 1;
    } else {
      b = t;
      fb = ft;
      if (side == -1) fa *= 0.5;
      side = -1;
      if (ft == 0.0) break;
    }
  }
  return b;
}
__device__ inline double2 horner4(const double2 v[5], double t) {
  double2 r = v[4];
#pragma unroll
  for (int k = 3; k >= 0; k--) r = make_double2(r.x * t + v[k].x, r.y * t + v[k].y);
  return r;
}

template <int S, int NZ>
__global__ void __launch_bounds__(256) k_mcwf_run(int32_t N, int64_t M, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ col, const double2 *__restrict__ val,
                                                  double2 *__restrict__ ens, const double *__restrict__ thr,
                                                  int32_t *__restrict__ jst, double *__restrict__ tsub, int32_t J,
                                                  double dt, double hlast, int32_t *__restrict__ jump) {
  const int SL = S ? S : N;             // lanes per segment
  const int P = S ? ENS_W / S : ENS_W / N;  // trajectories per wavefront
  __shared__ double2 sh[4][ENS_W];
  __shared__ double red[4][ENS_W];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int seg = lane / SL, r = lane - seg * SL;
  const int64_t i = ((int64_t)blockIdx.x * 4 + w) * P + seg;
  if (((int64_t)blockIdx.x * 4 + w) * P >= M) return;  // whole wavefront
  const bool have = seg < P && i < M;  // seg == P: a lane past the last segment
  // sum over the lane's segment, the bit-identical value in all its lanes:
  // DPP for power-of-two segments, else every lane adds the segment's values in
  // lane order from the wavefront's LDS row (an idle lane sums segment 0); used
  // for the nine polynomial coefficients of a crossing, whose terms have either
  // sign
  // the per-step squared norm on segments of N lanes: the difference of two
  // entries of the wavefront's prefix sum, read by every lane of the segment
  // from the same two lanes (the values are non-negative, so the difference
  // keeps full relative accuracy at the scale of the segment's own sum)
  auto nsum = [&](double v) -> double {
    if constexpr (S != 0) return seg_sum<S>(v);
    const double pf = wave_prefix_f64(v);
    const int b = (seg < P ? seg : 0) * N;
    const double pe = __shfl(pf, b + N - 1, 64);
    const double ps = __shfl(pf, b > 0 ? b - 1 : 0, 64);
    return b > 0 ? pe - ps : pe;
  };
  auto ssum = [&](double v) -> double {
    if constexpr (S != 0) return seg_sum<S>(v);
    red[w][lane] = v;
    __builtin_amdgcn_wave_barrier();
    const double *q = red[w] + (seg < P ? seg : 0) * N;
    double acc = 0.0;
    for (int k = 0; k < N; k++) acc += q[k];
    __builtin_amdgcn_wave_barrier();
    return acc;
  };
  int32_t j = have ? jst[i] : J;
  double sub = have ? tsub[i] : 0.0;  // time already covered inside step j
  const bool on = have && r < N;
  int32_t c[NZ];
  double2 a[NZ];
  const int32_t beg = on ? rp[r] : 0;
  const int nz = on ? rp[r + 1] - beg : 0;
#pragma unroll
  for (int e = 0; e < NZ; e++) {
    c[e] = r;
    a[e] = make_double2(0.0, 0.0);
    if (e < nz) {
      c[e] = col[beg + e];
      a[e] = val[beg + e];
    }
  }
  double2 *sg = sh[w] + seg * SL;  // idle lanes: their own lane's entry
  double2 psi = on ? ens[i * N + r] : make_double2(0.0, 0.0);
  const double t = have ? thr[i] : 0.0;
  // A v for this lane's row; the LDS row is written and read by this wavefront
  // only, in program order
  auto apply = [&](double2 v) {
    sg[r] = v;
    __builtin_amdgcn_wave_barrier();
    const double2 q = ens_row_product<NZ>(c, a, sg);
    __builtin_amdgcn_wave_barrier();
    return q;
  };
  bool act = j < J;
  int32_t jumped = 0;
  while (__ballot(act)) {
    const double h = ((j == J - 1) ? hlast : dt) - sub;
    const double2 k1 = apply(psi);
    const double2 k2 = apply(make_double2(psi.x + 0.5 * h * k1.x, psi.y + 0.5 * h * k1.y));
    const double2 k3 = apply(make_double2(psi.x + 0.5 * h * k2.x, psi.y + 0.5 * h * k2.y));
    const double2 k4 = apply(make_double2(psi.x + h * k3.x, psi.y + h * k3.y));
    const double s6 = h * (1.0 / 6.0), s3 = h * (1.0 / 3.0);
    const double2 nx = make_double2(psi.x + s6 * k1.x + s3 * k2.x + s3 * k3.x + s6 * k4.x,
                                    psi.y + s6 * k1.y + s3 * k2.y + s3 * k3.y + s6 * k4.y);
    const double n2 = nsum(on ? nx.x * nx.x + nx.y * nx.y : 0.0);
    const bool crossed = act && n2 < t;
    double2 pc = nx;
    double tau = h;
    if (__ballot(crossed)) {  // wave-uniform: every segment takes the same products
      double2 v[5];
      v[0] = psi;
      v[1] = k1;
      const double2 a2 = apply(k1);
      v[2] = make_double2(0.5 * a2.x, 0.5 * a2.y);
      const double2 a3 = apply(v[2]);
      v[3] = make_double2(a3.x * (1.0 / 3.0), a3.y * (1.0 / 3.0));
      const double2 a4 = apply(v[3]);
      v[4] = make_double2(0.25 * a4.x, 0.25 * a4.y);
      if (!on) v[0] = v[1] = v[2] = v[3] = v[4] = make_double2(0.0, 0.0);
      double cm[9];
      cross_partials(v, cm);
#pragma unroll
      for (int m = 0; m < 9; m++) cm[m] = ssum(cm[m]);
      tau = crossed ? cross_tau(cm, t, h) : h;
      pc = horner4(v, tau);
    }
    if (act) {
      if (crossed) {
        jumped = 1;
        act = false;
        if (tau < h * (1.0 - 1e-14)) {  // the jump falls inside the step: resume there
          psi = pc;
          sub += tau;
        } else {
          psi = nx;
          j++;
          sub = 0.0;
        }
      } else {
        psi = nx;
        j++;
        sub = 0.0;
        if (j >= J) act = false;
      }
    }
  }
  if (on) ens[i * N + r] = psi;
  if (have && r == 0) {
    jst[i] = j;
    tsub[i] = sub;
    jump[i] = jumped;
  }
}

// the jump operators of an ensemble (CSR, N x N each)
const int ENS_MAXJ = 16;
struct JumpOps {
  const int32_t *rp[ENS_MAXJ];
  const int32_t *col[ENS_MAXJ];
  const double2 *val[ENS_MAXJ];
};

// p[q * nj + k] = || C_k psi_{idx[q]} ||^2, one wavefront per (q, k), rows
// strided over its lanes
__global__ void __launch_bounds__(256) k_mcwf_jump_norms(int32_t N, const double2 *__restrict__ ens,
                                                        const int64_t *__restrict__ idx, int32_t n, int32_t nj,
                                                        JumpOps ops, double *__restrict__ p) {
  const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (g >= (int64_t)n * nj) return;  // whole wavefront
  const int64_t q = g / nj;
  const int k = (int)(g % nj);
  const double2 *psi = ens + idx[q] * N;
  double v = 0.0;
  for (int32_t r = lane; r < N; r += 64) {
    double re = 0.0, im = 0.0;
    for (int32_t e = ops.rp[k][r]; e < ops.rp[k][r + 1]; e++) {
      const double2 a = ops.val[k][e], x = psi[ops.col[k][e]];
      re = fma(a.x, x.x, fma(-a.y, x.y, re));
      im = fma(a.x, x.y, fma(a.y, x.x, im));
    }
    v += re * re + im * im;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) p[g] = v;
}

// psi_{idx[q]} <- sc[q] C_{kk[q]} psi_{idx[q]}, one workgroup per q: psi is
// staged in LDS, then every row is rewritten
const int ENS_WMAX = 1024;  // amplitudes per trajectory
__global__ void __launch_bounds__(256) k_mcwf_jump_apply(int32_t N, double2 *__restrict__ ens,
                                                        const int64_t *__restrict__ idx, int32_t n,
                                                        const int32_t *__restrict__ kk,
                                                        const double2 *__restrict__ sc, JumpOps ops) {
  __shared__ double2 s[ENS_WMAX];
  const int64_t q = blockIdx.x;
  if (q >= n) return;  // whole workgroup
  double2 *psi = ens + idx[q] * N;
  for (int32_t r = threadIdx.x; r < N; r += 256) s[r] = psi[r];
  __syncthreads();
  const int k = kk[q];
  const double2 f = sc[q];
  for (int32_t r = threadIdx.x; r < N; r += 256) {
    double re = 0.0, im = 0.0;
    for (int32_t e = ops.rp[k][r]; e < ops.rp[k][r + 1]; e++) {
      const double2 a = ops.val[k][e], x = s[ops.col[k][e]];
      re = fma(a.x, x.x, fma(-a.y, x.y, re));
      im = fma(a.x, x.y, fma(a.y, x.x, im));
    }
    psi[r] = make_double2(f.x * re - f.y * im, f.x * im + f.y * re);
  }
}

// Trajectories of 65..1024 amplitudes: one workgroup per trajectory, thread r
// holding amplitude r and row r of the generator; the products exchange the
// amplitudes through LDS between barriers, and the squared norm is a block sum
// that every thread forms from the per-wavefront partials in the same order, so
// the loop test is the same in every thread.  Steps, stop rule and outputs as
// k_mcwf_run.
template <int NZ, int LB>
__global__ void __launch_bounds__(LB) k_mcwf_run_wg(int32_t N, const int32_t *__restrict__ rp,
                                                     const int32_t *__restrict__ col,
                                                     const double2 *__restrict__ val, double2 *__restrict__ ens,
                                                     const double *__restrict__ thr, int32_t *__restrict__ jst,
                                                     double *__restrict__ tsub, int32_t J, double dt, double hlast,
                                                     int32_t *__restrict__ jump) {
  __shared__ double2 sh[ENS_WMAX];
  __shared__ double red[ENS_WMAX / 64][9];
  const int64_t i = blockIdx.x;
  const int r = threadIdx.x, w = r >> 6, nw = blockDim.x >> 6;
  int32_t j = jst[i];
  double sub = tsub[i];
  if (j >= J) {  // the same in every thread of the workgroup
    if (r == 0) jump[i] = 0;
    return;
  }
  const bool on = r < N;
  int32_t c[NZ];
  double2 a[NZ];
  const int32_t beg = on ? rp[r] : 0;
  const int nz = on ? rp[r + 1] - beg : 0;
#pragma unroll
  for (int e = 0; e < NZ; e++) {
    c[e] = 0;
    a[e] = make_double2(0.0, 0.0);
    if (e < nz) {
      c[e] = col[beg + e];
      a[e] = val[beg + e];
    }
  }
  double2 psi = on ? ens[i * N + r] : make_double2(0.0, 0.0);
  const double t = thr[i];
  auto apply = [&](double2 v) {
    __syncthreads();  // the previous product's reads are done
    if (on) sh[r] = v;
    __syncthreads();
    return ens_row_product<NZ>(c, a, sh);
  };
  // the block sum of nv values per thread, the same order (hence the same
  // value) in every thread
  auto block_sum = [&](double *v, int nv) {
#pragma unroll
    for (int m = 0; m < 9; m++)
      if (m < nv) {
        const double part = seg_sum<64>(v[m]);
        if ((r & 63) == 0) red[w][m] = part;
      }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 9; m++)
      if (m < nv) {
        double acc = 0.0;
        for (int k = 0; k < nw; k++) acc += red[k][m];
        v[m] = acc;
      }
    __syncthreads();  // red is rewritten by the next sum
  };
  int32_t jumped = 0;
  while (true) {
    const double h = ((j == J - 1) ? hlast : dt) - sub;
    const double2 k1 = apply(psi);
    const double2 k2 = apply(make_double2(psi.x + 0.5 * h * k1.x, psi.y + 0.5 * h * k1.y));
    const double2 k3 = apply(make_double2(psi.x + 0.5 * h * k2.x, psi.y + 0.5 * h * k2.y));
    const double2 k4 = apply(make_double2(psi.x + h * k3.x, psi.y + h * k3.y));
    const double s6 = h * (1.0 / 6.0), s3 = h * (1.0 / 3.0);
    const double2 nx = make_double2(psi.x + s6 * k1.x + s3 * k2.x + s3 * k3.x + s6 * k4.x,
                                    psi.y + s6 * k1.y + s3 * k2.y + s3 * k3.y + s6 * k4.y);
    double n2[9];
    n2[0] = on ? nx.x * nx.x + nx.y * nx.y : 0.0;
    block_sum(n2, 1);
    if (n2[0] < t) {
      // the crossing inside the step (cross_tau), the same in every thread
      double2 v[5];
      v[0] = psi;
      v[1] = k1;
      const double2 a2 = apply(k1);
      v[2] = make_double2(0.5 * a2.x, 0.5 * a2.y);
      const double2 a3 = apply(v[2]);
      v[3] = make_double2(a3.x * (1.0 / 3.0), a3.y * (1.0 / 3.0));
      const double2 a4 = apply(v[3]);
      v[4] = make_double2(0.25 * a4.x, 0.25 * a4.y);
      if (!on) v[0] = v[1] = v[2] = v[3] = v[4] = make_double2(0.0, 0.0);
      double cm[9];
      cross_partials(v, cm);
      block_sum(cm, 9);
      const double tau = cross_tau(cm, t, h);
      jumped = 1;
      if (tau < h * (1.0 - 1e-14)) {
        psi = horner4(v, tau);
        sub += tau;
      } else {
        psi = nx;
        j++;
        sub = 0.0;
      }
      break;
    }
    psi = nx;
    j++;
    sub = 0.0;
    if (j >= J) break;
  }
  if (on) ens[i * N + r] = psi;
  if (r == 0) {
    jst[i] = j;
    tsub[i] = sub;
    jump[i] = jumped;
  }
}

// The crossing of one trajectory of any size (the step-by-step loop): one
// workgroup forms v1..v4 in the global scratch v (4 N entries), the nine c_m,
// tau = cross_tau on thread 0, and out = P(tau).
__global__ void __launch_bounds__(1024) k_mcwf_cross(int32_t N, const int32_t *__restrict__ rp,
                                                    const int32_t *__restrict__ col,
                                                    const double2 *__restrict__ val, const double2 *psi0,
                                                    double2 *out, double2 *v, double h, double thr,
                                                    double *tau_out) {
  __shared__ double red[16][9];
  __shared__ double tau_s;
  const int tid = threadIdx.x, w = tid >> 6, nw = blockDim.x >> 6;
  const double inv[4] = {1.0, 0.5, 1.0 / 3.0, 0.25};
  for (int k = 0; k < 4; k++) {
    const double2 *src = k == 0 ? psi0 : v + (int64_t)(k - 1) * N;
    double2 *dst = v + (int64_t)k * N;
    for (int32_t r = tid; r < N; r += blockDim.x) {
      double re = 0.0, im = 0.0;
      for (int32_t e = rp[r]; e < rp[r + 1]; e++) {
        const double2 a = val[e], x = src[col[e]];
        re = fma(a.x, x.x, fma(-a.y, x.y, re));
        im = fma(a.x, x.y, fma(a.y, x.x, im));
      }
      dst[r] = make_double2(inv[k] * re, inv[k] * im);
    }
    __syncthreads();
  }
  double cm[9];
#pragma unroll
  for (int m = 0; m < 9; m++) cm[m] = 0.0;
  for (int32_t r = tid; r < N; r += blockDim.x) {
    double2 q[5];
    q[0] = psi0[r];
    for (int k = 0; k < 4; k++) q[k + 1] = v[(int64_t)k * N + r];
    double p[9];
    cross_partials(q, p);
#pragma unroll
    for (int m = 0; m < 9; m++) cm[m] += p[m];
  }
#pragma unroll
  for (int m = 0; m < 9; m++) {
    const double part = seg_sum<64>(cm[m]);
    if ((tid & 63) == 0) red[w][m] = part;
  }
  __syncthreads();
  if (tid == 0) {
    double c[9];
    for (int m = 0; m < 9; m++) {
      c[m] = 0.0;
      for (int k = 0; k < nw; k++) c[m] += red[k][m];
    }
    tau_s = cross_tau(c, thr, h);
    tau_out[0] = tau_s;
  }
  __syncthreads();
  const double tau = tau_s;
  for (int32_t r = tid; r < N; r += blockDim.x) {
    double2 q[5];
    q[0] = psi0[r];
    for (int k = 0; k < 4; k++) q[k + 1] = v[(int64_t)k * N + r];
    out[r] = horner4(q, tau);
  }
}

int jump_ops(int nj, qh_mat *const *C, int64_t N, JumpOps *ops) {
  QH_REQUIRE(nj >= 1 && nj <= ENS_MAXJ && C, QH_EINVAL, "MCWF jumps: 1..16 jump operators required");
  for (int k = 0; k < nj; k++) {
    QH_REQUIRE(C[k] && C[k]->rows == N, QH_EINVAL, "MCWF jumps: operator %d is not N x N", k);
    ops->rp[k] = C[k]->rowptr;
    ops->col[k] = C[k]->col;
    ops->val[k] = C[k]->val;
  }
  return QH_OK;
}

}  // namespace

extern "C" int qh_mcwf_advance(const qh_mat *A, qh_vec *ens, int64_t M, const double *thr, int32_t *jstep,
                               double *tsub, int32_t J, double dt, double hlast, int32_t *jump) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && ens && thr && jstep && tsub && jump && ens->prec == QH_FP64 && M >= 1 && M <= (1 << 30) &&
                 ens->n == M * A->rows && A->rows >= 1 && A->rows <= ENS_WMAX && J >= 0,
             QH_EINVAL, "qh_mcwf_advance: FP64 ensemble of M trajectories of at most 1024 amplitudes required");
  const int32_t N = (int32_t)A->rows;
  rc = qh_mat_max_row(const_cast<qh_mat *>(A));  // cached on the matrix
  if (rc) return rc;
  QH_REQUIRE(A->max_row <= ENS_MAXNZ, QH_EINVAL, "qh_mcwf_advance: a row with more than %d entries", ENS_MAXNZ);
  hipStream_t s = qh_stream();
  // device copies: thresholds and covered step times, step indices, jump flags
  char *d = nullptr, *h = nullptr;
  const size_t bt = sizeof(double) * M, bj = sizeof(int32_t) * M;
  rc = qh_scratch(2 * bt + 2 * bj + 256, (void **)&d, 2 * bt + 2 * bj, (void **)&h);
  if (rc) return rc;
  double *dthr = (double *)d, *dsub = (double *)(d + bt);
  int32_t *djs = (int32_t *)(d + 2 * bt), *djump = (int32_t *)(d + 2 * bt + bj);
  memcpy(h, thr, bt);
  memcpy(h + bt, tsub, bt);
  memcpy(h + 2 * bt, jstep, bj);
  QH_HIP(hipMemcpyAsync(d, h, 2 * bt + bj, hipMemcpyHostToDevice, s));
  int S = N <= 2 ? 2 : N <= 4 ? 4 : N <= 8 ? 8 : N <= 16 ? 16 : N <= 32 ? 32 : 64;
  if (ENS_W / N > ENS_W / S) S = 0;  // segments of exactly N lanes fit more trajectories
  const int per = S ? ENS_W / S : ENS_W / N;
  const int64_t waves = (M + per - 1) / per;
  const dim3 grid((unsigned)((waves + 3) / 4));
  // generator entries per lane: the row length rounded up to 2, 3, 4 or 16
  // (each padded entry costs four FP64 FMAs per stage)
  const int nzc = A->max_row <= 2 ? 2 : A->max_row <= 3 ? 3 : A->max_row <= 4 ? 4 : ENS_MAXNZ;
  if (N > ENS_W) {  // one workgroup of whole wavefronts per trajectory
    const dim3 wg((unsigned)((N + 63) / 64 * 64)), tg((unsigned)M);
    // a 512-thread bound leaves 256 VGPRs per lane (the 16-entry rows spill
    // under the 1024-thread bound's 128)
#define QH_MCWF_WG(NZ, LB)                                                                                     \
  hipLaunchKernelGGL((k_mcwf_run_wg<NZ, LB>), tg, wg, 0, s, N, A->rowptr, A->col, A->val, (double2 *)ens->d,   \
                     dthr, djs, dsub, J, dt, hlast, djump)
    const int nzw = A->max_row <= 4 ? 4 : A->max_row <= 8 ? 8 : ENS_MAXNZ;
    if (N <= 512) {
      if (nzw == 4) QH_MCWF_WG(4, 512);
      else if (nzw == 8) QH_MCWF_WG(8, 512);
      else QH_MCWF_WG(ENS_MAXNZ, 512);
    } else {
      if (nzw == 4) QH_MCWF_WG(4, 1024);
      else if (nzw == 8) QH_MCWF_WG(8, 1024);
      else QH_MCWF_WG(ENS_MAXNZ, 1024);
    }
#undef QH_MCWF_WG
  } else {
#define QH_MCWF_RUN_NZ(SS, NZ)                                                                                   \
  hipLaunchKernelGGL((k_mcwf_run<SS, NZ>), grid, dim3(256), 0, s, N, M, A->rowptr, A->col, A->val,               \
                     (double2 *)ens->d, dthr, djs, dsub, J, dt, hlast, djump)
#define QH_MCWF_RUN(SS)                                                                                          \
  if (S == SS) {                                                                                                  \
    if (nzc == 2)                                                                                                 \
      QH_MCWF_RUN_NZ(SS, 2);                                                                                      \
    else if (nzc == 3)                                                                                            \
      QH_MCWF_RUN_NZ(SS, 3);                                                                                      \
    else if (nzc == 4)                                                                                            \
      QH_MCWF_RUN_NZ(SS, 4);                                                                                      \
    else                                                                                                          \
      QH_MCWF_RUN_NZ(SS, ENS_MAXNZ);                                                                              \
  }
  QH_MCWF_RUN(0)
  QH_MCWF_RUN(2)
  QH_MCWF_RUN(4)
  QH_MCWF_RUN(8)
  QH_MCWF_RUN(16)
  QH_MCWF_RUN(32)
  QH_MCWF_RUN(64)
#undef QH_MCWF_RUN
#undef QH_MCWF_RUN_NZ
  }
  QH_HIP(hipGetLastError());
  QH_HIP(hipMemcpyAsync(h + bt, dsub, bt + 2 * bj, hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  memcpy(tsub, h + bt, bt);
  memcpy(jstep, h + 2 * bt, bj);
  memcpy(jump, h + 2 * bt + bj, bj);
  return QH_OK;
}

extern "C" int qh_mcwf_fits(const qh_mat *A, int nj, int *fits) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && fits, QH_EINVAL, "qh_mcwf_fits: null argument");
  *fits = 0;
  if (A->rows < 1 || A->rows > ENS_WMAX || nj < 1 || nj > ENS_MAXJ) return QH_OK;
  rc = qh_mat_max_row(const_cast<qh_mat *>(A));
  if (rc) return rc;
  *fits = A->max_row <= ENS_MAXNZ;
  return QH_OK;
}

extern "C" int qh_mcwf_jump_norms(int nj, qh_mat *const *C, const qh_vec *ens, int64_t N, const int64_t *idx,
                                  int32_t n, double *p) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(ens && idx && p && ens->prec == QH_FP64 && N >= 1 && N <= ENS_WMAX && ens->n % N == 0 && n >= 0,
             QH_EINVAL, "qh_mcwf_jump_norms: FP64 ensemble of at most 1024 amplitudes per trajectory required");
  JumpOps ops;
  rc = jump_ops(nj, C, N, &ops);
  if (rc || n == 0) return rc;
  for (int32_t q = 0; q < n; q++)
    QH_REQUIRE(idx[q] >= 0 && idx[q] < ens->n / N, QH_EINVAL, "qh_mcwf_jump_norms: trajectory out of range");
  hipStream_t s = qh_stream();
  char *d = nullptr, *h = nullptr;
  const size_t bi = sizeof(int64_t) * n, bp = sizeof(double) * n * nj;
  rc = qh_scratch(bi + bp + 256, (void **)&d, bi + bp, (void **)&h);
  if (rc) return rc;
  memcpy(h, idx, bi);
  QH_HIP(hipMemcpyAsync(d, h, bi, hipMemcpyHostToDevice, s));
  const int64_t waves = (int64_t)n * nj;
  double *dp = (double *)(d + ((bi + 255) & ~(size_t)255));
  hipLaunchKernelGGL(k_mcwf_jump_norms, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, (int32_t)N,
                     (const double2 *)ens->d, (const int64_t *)d, n, (int32_t)nj, ops, dp);
  QH_HIP(hipGetLastError());
  QH_HIP(hipMemcpyAsync(h, dp, bp, hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  memcpy(p, h, bp);
  return QH_OK;
}

extern "C" int qh_mcwf_jump_apply(int nj, qh_mat *const *C, qh_vec *ens, int64_t N, const int64_t *idx, int32_t n,
                                  const int32_t *k, const double *sre, const double *sim) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(ens && idx && k && sre && sim && ens->prec == QH_FP64 && N >= 1 && N <= ENS_WMAX && ens->n % N == 0 &&
                 n >= 0,
             QH_EINVAL, "qh_mcwf_jump_apply: FP64 ensemble of at most 1024 amplitudes per trajectory required");
  JumpOps ops;
  rc = jump_ops(nj, C, N, &ops);
  if (rc || n == 0) return rc;
  for (int32_t q = 0; q < n; q++) {
    QH_REQUIRE(idx[q] >= 0 && idx[q] < ens->n / N, QH_EINVAL, "qh_mcwf_jump_apply: trajectory out of range");
    QH_REQUIRE(k[q] >= 0 && k[q] < nj, QH_EINVAL, "qh_mcwf_jump_apply: operator index out of range");
  }
  hipStream_t s = qh_stream();
  char *d = nullptr, *h = nullptr;
  const size_t bi = (sizeof(int64_t) * n + 255) & ~(size_t)255, bk = (sizeof(int32_t) * n + 255) & ~(size_t)255,
               bs = sizeof(double2) * n;
  rc = qh_scratch(bi + bk + bs, (void **)&d, bi + bk + bs, (void **)&h);
  if (rc) return rc;
  memcpy(h, idx, sizeof(int64_t) * n);
  memcpy(h + bi, k, sizeof(int32_t) * n);
  double2 *hs = (double2 *)(h + bi + bk);
  for (int32_t q = 0; q < n; q++) hs[q] = make_double2(sre[q], sim[q]);
  QH_HIP(hipMemcpyAsync(d, h, bi + bk + bs, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mcwf_jump_apply, dim3((unsigned)n), dim3(256), 0, s, (int32_t)N,
                     (double2 *)ens->d, (const int64_t *)d, n, (const int32_t *)(d + bi),
                     (const double2 *)(d + bi + bk), ops);
  QH_HIP(hipGetLastError());
  QH_HIP(hipStreamSynchronize(s));
  return QH_OK;
}

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

extern "C" int qh_mcwf_cross(const qh_mat *A, const qh_vec *psi0, qh_vec *out, double h, double thr,
                             double *tau) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && psi0 && out && tau && psi0->prec == QH_FP64 && out->prec == QH_FP64 && psi0->n == A->rows &&
                 out->n == A->rows && psi0->d != out->d && h > 0.0 && A->rows < ((int64_t)1 << 31),
             QH_EINVAL, "qh_mcwf_cross: FP64 state and output of the generator's dimension, h > 0");
  const int64_t N = A->rows;
  char *d = nullptr, *hp = nullptr;
  rc = qh_scratch(sizeof(double2) * 4 * (size_t)N + 256, (void **)&d, sizeof(double), (void **)&hp);
  if (rc) return rc;
  double2 *v = (double2 *)(d + 256);
  double *dtau = (double *)d;
  hipStream_t s = qh_stream();
  hipLaunchKernelGGL(k_mcwf_cross, dim3(1), dim3(1024), 0, s, (int32_t)N, A->rowptr, A->col, A->val,
                     (const double2 *)psi0->d, (double2 *)out->d, v, h, thr, dtau);
  QH_HIP(hipGetLastError());
  QH_HIP(hipMemcpyAsync(hp, dtau, sizeof(double), hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  *tau = *(double *)hp;
  return QH_OK;
}
