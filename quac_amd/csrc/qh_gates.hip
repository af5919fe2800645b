// qh_gates.hip -- in-place gate application rho <- U rho U^dagger (or psi <- U psi).
//
// The reference builds a fresh dim x dim AIJ gate superoperator U* x U for every
// gate and multiplies (quantum_circuits.c:435-481), or runs QuEST-style pair
// updates twice over vec(rho) (:696-772).  Here one pass does both sides: each
// thread owns the 2x2 (1-qubit) or 4x4 (2-qubit) block of rho addressed by the
// gate's row bits and column bits, loads it once, applies U on the left and U^dagger on
// the right in registers, and stores it back: 2 * rows * sizeof(entry) bytes per
// gate, the minimum for an in-place update.
#include <algorithm>

#include "qh_common.h"

namespace {

struct U4 {
  double re[16], im[16];
};

// insert a zero digit of a 2-level subsystem at stride s
__device__ inline int64_t ins0(int64_t t, int64_t s) { return (t / s) * (2 * s) + t % s; }
// the same for a power-of-two stride s = 2^k: shifts and a mask instead of the
// emulated 64-bit division
__device__ inline int64_t ins0p(int64_t t, int k) { return ((t >> k) << (k + 1)) | (t & ((1LL << k) - 1)); }

// entry i of the state: complex128, or the FP32 hi + lo pair (F32)
template <bool F32>
__device__ inline double2 ldv(const void *d, const float2 *lo, int64_t i) {
  if (!F32) return ((const double2 *)d)[i];
  float2 h = ((const float2 *)d)[i], l = lo[i];
  return make_double2((double)h.x + (double)l.x, (double)h.y + (double)l.y);
}
template <bool F32>
__device__ inline void stv(void *d, float2 *lo, int64_t i, double2 v) {
  if (!F32) {
    ((double2 *)d)[i] = v;
    return;
  }
  float hr = (float)v.x, hi = (float)v.y;
  ((float2 *)d)[i] = make_float2(hr, hi);
  lo[i] = make_float2((float)(v.x - (double)hr), (float)(v.y - (double)hi));
}

// K = 2 (1 qubit) or 4 (2 qubits); DM: X' = U X U^dagger on a KxK block; WF: psi' = U psi.
// P2: every inserted stride is a power of two, passed as its exponent in lg.
// F32: the state is the FP32 hi + lo pair (x the high words, lo the low words).
template <int K, bool DM, bool P2, bool F32>
__global__ void __launch_bounds__(256) k_gate(void *x, float2 *lo, int64_t nthreads, int nins, int64_t s0,
                                              int64_t s1, int64_t s2, int64_t s3, int4 lg, int64_t bs0, int64_t bs1,
                                              int64_t cs0, int64_t cs1, U4 u) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nthreads) return;
  // insert the gate's zero digits, lowest stride first
  int64_t base = t;
  if (P2) {
    const int ks[4] = {lg.x, lg.y, lg.z, lg.w};
#pragma unroll
    for (int q = 0; q < (DM ? 2 * (K / 2) : K / 2); q++) base = ins0p(base, ks[q]);
  } else {
    const int64_t ss[4] = {s0, s1, s2, s3};
    for (int q = 0; q < nins; q++) base = ins0(base, ss[q]);
  }
  // offsets of the K basis states of the gate (first qubit = most significant)
  int64_t ro[K], co[K];
#pragma unroll
  for (int a = 0; a < K; a++) {
    if (K == 2) {
      ro[a] = a * bs0;
      co[a] = a * cs0;
    } else {
      ro[a] = ((a >> 1) & 1) * bs0 + (a & 1) * bs1;
      co[a] = ((a >> 1) & 1) * cs0 + (a & 1) * cs1;
    }
  }
  if (!DM) {
    double2 p[K];
#pragma unroll
    for (int a = 0; a < K; a++) p[a] = ldv<F32>(x, lo, base + ro[a]);
#pragma unroll
    for (int a = 0; a < K; a++) {
      double r = 0.0, i = 0.0;
#pragma unroll
      for (int c = 0; c < K; c++) {
        r += u.re[a * K + c] * p[c].x - u.im[a * K + c] * p[c].y;
        i += u.re[a * K + c] * p[c].y + u.im[a * K + c] * p[c].x;
      }
      stv<F32>(x, lo, base + ro[a], make_double2(r, i));
    }
    return;
  }
  // X[c][d] = rho[row c, col d] at base + ro[c] + co[d]
  double2 X[K][K];
#pragma unroll
  for (int c = 0; c < K; c++)
#pragma unroll
    for (int d = 0; d < K; d++) X[c][d] = ldv<F32>(x, lo, base + ro[c] + co[d]);
  // row a of Y = U X, then row a of Z = Y U^dagger (Z[a][b] = sum_d Y[a][d]
  // conj(U[b][d])) stored at once: one row of Y lives at a time, so the block
  // and one row are the registers held
#pragma unroll
  for (int a = 0; a < K; a++) {
    double2 Y[K];
#pragma unroll
    for (int d = 0; d < K; d++) {
      double r = 0.0, i = 0.0;
#pragma unroll
      for (int c = 0; c < K; c++) {
        r += u.re[a * K + c] * X[c][d].x - u.im[a * K + c] * X[c][d].y;
        i += u.re[a * K + c] * X[c][d].y + u.im[a * K + c] * X[c][d].x;
      }
      Y[d] = make_double2(r, i);
    }
#pragma unroll
    for (int b = 0; b < K; b++) {
      double r = 0.0, i = 0.0;
#pragma unroll
      for (int d = 0; d < K; d++) {
        const double ur = u.re[b * K + d], ui = -u.im[b * K + d];
        r += Y[d].x * ur - Y[d].y * ui;
        i += Y[d].x * ui + Y[d].y * ur;
      }
      stv<F32>(x, lo, base + ro[a] + co[b], make_double2(r, i));
    }
  }
}

// Two-qubit gate on a density matrix: a workgroup takes 64 consecutive 4x4
// blocks, and wavefront q of its four handles column q of all of them.  Lane t
// of wavefront q loads column q of block t and forms column q of Y = U X; the
// workgroup transposes Y through LDS; then the lane forms and stores row q of
// Z = Y U^dagger for block t.  Every load and store instruction of a wavefront
// covers 64 consecutive blocks (one contiguous run where the gate's qubits are
// not the lowest ones), and each lane holds 4 entries, not 16.
template <bool P2, bool F32>
__global__ void __launch_bounds__(256) k_gate2_dm(void *x, float2 *lo, int64_t nblk, int64_t s0, int64_t s1,
                                                  int64_t s2, int64_t s3, int4 lg, int64_t bs0, int64_t bs1,
                                                  int64_t cs0, int64_t cs1, U4 u) {
  __shared__ double2 sh[4][256];  // sh[a][q * 64 + t] = Y[a][q] of block t
  const int tid = threadIdx.x, q = tid >> 6, tl = tid & 63;
  const int64_t t = (int64_t)blockIdx.x * 64 + tl;
  const bool on = t < nblk;
  int64_t base = on ? t : 0;
  if (P2) {
    base = ins0p(base, lg.x);
    base = ins0p(base, lg.y);
    base = ins0p(base, lg.z);
    base = ins0p(base, lg.w);
  } else {
    base = ins0(base, s0);
    base = ins0(base, s1);
    base = ins0(base, s2);
    base = ins0(base, s3);
  }
  int64_t ro[4];
#pragma unroll
  for (int a = 0; a < 4; a++) ro[a] = ((a >> 1) & 1) * bs0 + (a & 1) * bs1;
  const int64_t coq = ((q >> 1) & 1) * cs0 + (q & 1) * cs1;
  // a lane past the last block loads block 0 and stores nothing:
  // unconditional loads, all in flight together
  double2 X[4];
#pragma unroll
  for (int c = 0; c < 4; c++) X[c] = ldv<F32>(x, lo, base + ro[c] + coq);
  // column q of Y = U X
#pragma unroll
  for (int a = 0; a < 4; a++) {
    double r = 0.0, i = 0.0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      r += u.re[a * 4 + c] * X[c].x - u.im[a * 4 + c] * X[c].y;
      i += u.re[a * 4 + c] * X[c].y + u.im[a * 4 + c] * X[c].x;
    }
    sh[a][q * 64 + tl] = make_double2(r, i);
  }
  __syncthreads();
  // row q of Y: Y[q][d] is entry q of column d, written by wavefront d
  double2 Y[4];
#pragma unroll
  for (int d = 0; d < 4; d++) Y[d] = sh[q][d * 64 + tl];
  if (!on) return;
  const int64_t rq = base + ((q >> 1) & 1) * bs0 + (q & 1) * bs1;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    double r = 0.0, i = 0.0;
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const double ur = u.re[b * 4 + d], ui = -u.im[b * 4 + d];
      r += Y[d].x * ur - Y[d].y * ui;
      i += Y[d].x * ui + Y[d].y * ur;
    }
    const int64_t cb = ((b >> 1) & 1) * cs0 + (b & 1) * cs1;
    stv<F32>(x, lo, rq + cb, make_double2(r, i));
  }
}

}  // namespace

extern "C" int qh_gate(qh_vec *x, int64_t tot, int dm, int nq, const int64_t *strides, const double *u) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(x && strides && u && (nq == 1 || nq == 2), QH_EINVAL, "qh_gate: bad arguments");
  QH_REQUIRE(dm ? tot * tot == x->n : tot == x->n, QH_EINVAL, "qh_gate: vector size mismatch");
  for (int q = 0; q < nq; q++)
    QH_REQUIRE(strides[q] > 0 && tot % (2 * strides[q]) == 0, QH_EINVAL,
               "qh_gate: stride %lld is not a 2-level subsystem of dimension %lld", (long long)strides[q],
               (long long)tot);
  QH_REQUIRE(nq == 1 || strides[0] != strides[1], QH_EINVAL, "qh_gate: both qubits are the same subsystem");
  U4 m;
  const int K = nq == 1 ? 2 : 4;
  for (int i = 0; i < K * K; i++) {
    m.re[i] = u[2 * i];
    m.im[i] = u[2 * i + 1];
  }
  int64_t ins[4];
  int nins = 0;
  for (int q = 0; q < nq; q++) ins[nins++] = strides[q];
  if (dm)
    for (int q = 0; q < nq; q++) ins[nins++] = strides[q] * tot;
  std::sort(ins, ins + nins);
  for (int q = nins; q < 4; q++) ins[q] = 1;
  const int64_t total = dm ? tot * tot : tot;
  const int64_t nthreads = total >> nins;
  const int64_t bs0 = strides[0], bs1 = nq == 2 ? strides[1] : 0;
  const int64_t cs0 = dm ? strides[0] * tot : 0, cs1 = (dm && nq == 2) ? strides[1] * tot : 0;
  const dim3 grid(qh_blocks(nthreads, 256));
  hipStream_t s = qh_stream();
  // exponents when every inserted stride is a power of two
  bool p2 = true;
  int lgv[4] = {0, 0, 0, 0};
  for (int q = 0; q < nins; q++) {
    if (ins[q] & (ins[q] - 1)) p2 = false;
    int k = 0;
    while (k < 62 && ((int64_t)1 << k) < ins[q]) k++;
    lgv[q] = k;
  }
  const int4 lg = make_int4(lgv[0], lgv[1], lgv[2], lgv[3]);
#define QH_GATE(KK, DMM, PP)                                                                                      \
  do {                                                                                                            \
    if (x->lo)                                                                                                    \
      hipLaunchKernelGGL((k_gate<KK, DMM, PP, true>), grid, dim3(256), 0, s, x->d, x->lo, nthreads, nins, ins[0], \
                         ins[1], ins[2], ins[3], lg, bs0, bs1, cs0, cs1, m);                                      \
    else                                                                                                          \
      hipLaunchKernelGGL((k_gate<KK, DMM, PP, false>), grid, dim3(256), 0, s, x->d, x->lo, nthreads, nins,        \
                         ins[0], ins[1], ins[2], ins[3], lg, bs0, bs1, cs0, cs1, m);                              \
  } while (0)
  if (nq == 1 && dm) {
    if (p2) QH_GATE(2, true, true); else QH_GATE(2, true, false);
  } else if (nq == 1) {
    if (p2) QH_GATE(2, false, true); else QH_GATE(2, false, false);
  } else if (dm) {
    // 64 blocks per workgroup of four wavefronts (nthreads blocks, always 4
    // inserted strides)
    const dim3 g4(qh_blocks(nthreads, 64));
    if (p2) {
      if (x->lo)
        hipLaunchKernelGGL((k_gate2_dm<true, true>), g4, dim3(256), 0, s, x->d, x->lo, nthreads, ins[0], ins[1],
                           ins[2], ins[3], lg, bs0, bs1, cs0, cs1, m);
      else
        hipLaunchKernelGGL((k_gate2_dm<true, false>), g4, dim3(256), 0, s, x->d, x->lo, nthreads, ins[0], ins[1],
                           ins[2], ins[3], lg, bs0, bs1, cs0, cs1, m);
    } else {
      if (x->lo)
        hipLaunchKernelGGL((k_gate2_dm<false, true>), g4, dim3(256), 0, s, x->d, x->lo, nthreads, ins[0], ins[1],
                           ins[2], ins[3], lg, bs0, bs1, cs0, cs1, m);
      else
        hipLaunchKernelGGL((k_gate2_dm<false, false>), g4, dim3(256), 0, s, x->d, x->lo, nthreads, ins[0], ins[1],
                           ins[2], ins[3], lg, bs0, bs1, cs0, cs1, m);
    }
  } else {
    if (p2) QH_GATE(4, false, true); else QH_GATE(4, false, false);
  }
#undef QH_GATE
  QH_HIP(hipGetLastError());
  return QH_OK;
}
