// qh_spmv.hip -- the hot path: complex CSR SpMV fused with the RK4 stage updates.
//
// Each RK4 step of dx/dt = A x is four launches of one kernel template:
//   S1: K1 = A x
//   S2: K2 = A (x + h/2 K1)        (stage input formed on the fly in the gather)
//   S3: K3 = A (x + h/2 K2)
//   S4: x' = x + h/6 (K1 + 2 K2 + 2 K3 + A (x + h K3))   (update fused in the epilogue)
// so the stage vectors Y_s are never materialised: per step the matrix is
// streamed 4 times and 14 vector passes are made (4 gathers of x, 3 of K, 3 K
// writes, 3 K reads + x read + x' write in the epilogue).
//
// Layout and mapping (gfx950, wave64): a group of G lanes owns one row (64/G rows
// per wavefront, G autotuned per matrix); lanes stride through the row so every
// load instruction of a wave reads contiguous col/val bytes; the matrix streams
// use non-temporal loads so they do not evict x from L2/MALL; the group reduces
// with cross-lane shuffles and lane 0 runs the fused epilogue.  Block indices
// are remapped so each XCD sweeps one contiguous band of rows.
//
// FP64 path: complex128 values, FP64 accumulation (the drop-in path).
// FP32 path: complex64 values and vectors; each lane accumulates with TwoSum
// compensation, the state carries a complex64 compensation word (x = hi + lo)
// so that the RK update does not lose the low bits of x.
#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <vector>

#include "qh_common.h"
#include "qh_ks.h"

namespace {

template <typename R>
struct C2;
template <>
struct C2<double> {
  typedef double2 T;
  typedef double __attribute__((ext_vector_type(2))) V;
};
template <>
struct C2<float> {
  typedef float2 T;
  typedef float __attribute__((ext_vector_type(2))) V;
};

template <typename R>
__device__ inline typename C2<R>::T ld_stream(const typename C2<R>::T *p) {
  typename C2<R>::V v = __builtin_nontemporal_load(reinterpret_cast<const typename C2<R>::V *>(p));
  typename C2<R>::T t;
  t.x = v.x;
  t.y = v.y;
  return t;
}

// error-free sum: s + e == a + b exactly
__device__ inline void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

template <typename R>
struct Acc;

template <>
struct Acc<double> {
  double re = 0.0, im = 0.0;
  __device__ inline void add(double pr, double pi) {
    re += pr;
    im += pi;
  }
  template <int G>
  __device__ inline void reduce() {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
      re += __shfl_xor(re, off, G);
      im += __shfl_xor(im, off, G);
    }
  }
  __device__ inline double2 value() const { return make_double2(re, im); }
};

template <>
struct Acc<float> {
  float re = 0.f, im = 0.f, cre = 0.f, cim = 0.f;
  __device__ inline void add(float pr, float pi) {
    float e;
    two_sum(re, pr, re, e);
    cre += e;
    two_sum(im, pi, im, e);
    cim += e;
  }
  template <int G>
  __device__ inline void reduce() {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
      float ore = __shfl_xor(re, off, G), oim = __shfl_xor(im, off, G);
      float ocr = __shfl_xor(cre, off, G), oci = __shfl_xor(cim, off, G);
      float e;
      two_sum(re, ore, re, e);
      cre += e + ocr;
      two_sum(im, oim, im, e);
      cim += e + oci;
    }
  }
  __device__ inline float2 value() const { return make_float2(re + cre, im + cim); }
};

template <typename R, int G, int MODE>
__global__ void __launch_bounds__(256) k_rk_stage(
    int32_t nrows, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
    const typename C2<R>::T *__restrict__ val, const typename C2<R>::T *__restrict__ x,
    const typename C2<R>::T *__restrict__ kin, R ck, typename C2<R>::T *__restrict__ kout,
    const typename C2<R>::T *__restrict__ k1, const typename C2<R>::T *__restrict__ k2,
    const typename C2<R>::T *__restrict__ k3, typename C2<R>::T *__restrict__ xout, R h,
    float2 *__restrict__ xlo) {
  typedef typename C2<R>::T T;
  const unsigned b = qh_xcd_block(blockIdx.x, gridDim.x);
  const int64_t tid = (int64_t)b * 256 + threadIdx.x;
  const int32_t row = (int32_t)(tid / G);
  const int lane = (int)(tid % G);
  if (row >= nrows) return;  // G divides 256: a group never straddles the bound
  const int32_t beg = rp[row], end = rp[row + 1];
  Acc<R> acc;
  int32_t k = beg + lane;
  // two independent loads in flight per lane
  for (; k + G < end; k += 2 * G) {
    const int32_t c0 = __builtin_nontemporal_load(col + k);
    const int32_t c1 = __builtin_nontemporal_load(col + k + G);
    const T a0 = ld_stream<R>(val + k);
    const T a1 = ld_stream<R>(val + k + G);
    T x0 = x[c0], x1 = x[c1];
    if (MODE != MODE_SPMV) {
      const T q0 = kin[c0], q1 = kin[c1];
      x0.x += ck * q0.x;
      x0.y += ck * q0.y;
      x1.x += ck * q1.x;
      x1.y += ck * q1.y;
    }
    acc.add(a0.x * x0.x - a0.y * x0.y, a0.x * x0.y + a0.y * x0.x);
    acc.add(a1.x * x1.x - a1.y * x1.y, a1.x * x1.y + a1.y * x1.x);
  }
  if (k < end) {
    const int32_t c0 = __builtin_nontemporal_load(col + k);
    const T a0 = ld_stream<R>(val + k);
    T x0 = x[c0];
    if (MODE != MODE_SPMV) {
      const T q0 = kin[c0];
      x0.x += ck * q0.x;
      x0.y += ck * q0.y;
    }
    acc.add(a0.x * x0.x - a0.y * x0.y, a0.x * x0.y + a0.y * x0.x);
  }
  acc.template reduce<G>();
  if (lane != 0) return;
  const T y = acc.value();
  if (MODE != MODE_FINAL) {
    kout[row] = y;
    return;
  }
  // x' = x + h/6 (K1 + 2K2 + 2K3 + K4)
  const T a = k1[row], bq = k2[row], cq = k3[row];
  const R s6 = h * (R)(1.0 / 6.0), s3 = h * (R)(1.0 / 3.0);
  const R dre = s6 * a.x + s3 * bq.x + s3 * cq.x + s6 * y.x;
  const R dim = s6 * a.y + s3 * bq.y + s3 * cq.y + s6 * y.y;
  const T xo = x[row];
  if (sizeof(R) == 8) {
    T r;
    r.x = xo.x + dre;
    r.y = xo.y + dim;
    xout[row] = r;
  } else {
    // compensated update of (hi, lo): hi' + lo' = hi + lo + d to ~2^-48
    float2 lo = xlo[row];
    float sr, er, si, ei;
    two_sum((float)xo.x, (float)dre, sr, er);
    two_sum((float)xo.y, (float)dim, si, ei);
    float lr = lo.x + er, li = lo.y + ei;
    float hr = sr + lr, hi = si + li;
    lr = lr - (hr - sr);
    li = li - (hi - si);
    T r;
    r.x = hr;
    r.y = hi;
    xout[row] = r;
    xlo[row] = make_float2(lr, li);
  }
}

// ------------------------------------------------- slot-aligned sliced ELL
// One wavefront owns a slice of 64 consecutive rows, one lane per row.  Slot k
// of the slice holds, for every lane that has it, the entry at column row +- ad_k;
// the per-slot metadata is wave-uniform (scalar loads), the values of the active
// lanes are compacted (no padding, no column-index stream), and the x gathers of
// a slot are contiguous across the wavefront.
//
// The slot loop is branch-free: the metadata's 64-bit lane masks select per lane
// directly (inverse ballot), inactive lanes load through an out-of-range buffer
// offset and read exact zeros, so the U gathers of a pass are all in flight
// together.  Products and sums are formed in FP64 for both storage precisions
// (FP64 FMA issues at the FP32 rate on CDNA4).

// AUX: the cache policy bits of the buffer load (0, or 2 = non-temporal)
template <typename R, int AUX = 0>
__device__ inline typename C2<R>::T ks_load(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
  if constexpr (sizeof(R) == 4)
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, AUX));
  else
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, AUX));
}

template <typename R, int AUX = 0>
__device__ inline R ks_load_real(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
  if constexpr (sizeof(R) == 4)
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, AUX));
  else
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, AUX));
}

// One pass over U slots of a slice starting at mp, split into the issue of its
// loads (ks_issue*) and the products of the loaded values (ks_fma*), so that a
// group can keep the loads of the next pass in flight while it multiplies the
// current one.  cnt valid slots when TAIL (the metadata array carries
// QH_KS_PAD empty slots past its end, so the reads of a tail pass never leave
// it; slots past cnt get an empty lane mask).
// KIND: uniform slots take their value from the metadata, scalar slots load one
// real per lane (vr over the scalar values) and multiply it by (vre, vim),
// complex slots load one complex value per lane (vr over the complex values).
// VV: the uniform value of a slot is read with a vector load (broadcast to all
// lanes) instead of a scalar load, which keeps the metadata held in SGPRs at six
// dwords per slot and lets a pass carry 8 slots; mr covers the metadata array
// and k is the slot index of mp[0].
template <typename R, int KIND, int U, bool VV>
struct KsBuf {
  typename C2<R>::T xv[U], vv[U];
  R sv[U];
  double ur[U], ui[U];  // uniform value, or the (1, 0) / (0, 1) multiplier of a scalar slot
};

template <typename R, int KIND, bool TAIL, int U, bool VV, int AUX = 0>
__device__ __forceinline__ void ks_issue(const qh_kmeta *__restrict__ mp, int32_t cnt, uint32_t r8,
                                         __amdgpu_buffer_rsrc_t xr, __amdgpu_buffer_rsrc_t vr,
                                         __amdgpu_buffer_rsrc_t mr, int32_t k, KsBuf<R, KIND, U, VV> &b) {
  typedef typename C2<R>::T T;
  qh_kmeta m[U];
  uint64_t mask[U];
#pragma unroll
  for (int j = 0; j < U; j++) {
    m[j] = mp[j];
    mask[j] = (!TAIL || j < cnt) ? m[j].mask : 0ull;
  }
  if (KIND == KS_UNI && VV) {
#pragma unroll
    for (int j = 0; j < U; j++) {
      const double2 u = __builtin_bit_cast(
          double2, __builtin_amdgcn_raw_buffer_load_b128(
                       mr, (uint32_t)(k + j) * (uint32_t)sizeof(qh_kmeta) + (uint32_t)offsetof(qh_kmeta, vre), 0, 0));
      b.ur[j] = u.x;
      b.ui[j] = u.y;
    }
  } else if (KIND != KS_CPX) {
#pragma unroll
    for (int j = 0; j < U; j++) {
      b.ur[j] = m[j].vre;
      b.ui[j] = m[j].vim;
    }
  }
#pragma unroll
  for (int j = 0; j < U; j++) {
    const bool on = __builtin_amdgcn_inverse_ballot_w64(mask[j]);
    const bool neg = __builtin_amdgcn_inverse_ballot_w64(m[j].sign);
    const uint32_t ad8 = m[j].ad * (uint32_t)sizeof(T);
    uint32_t off = neg ? r8 - ad8 : r8 + ad8;
    off = on ? off : QH_KS_OOB;
    b.xv[j] = ks_load<R>(xr, off);
    if (KIND != KS_UNI) {
      // rank of the lane among the slot's active lanes
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask[j] >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)mask[j], 0u));
      if (KIND == KS_SCA)
        b.sv[j] = ks_load_real<R, AUX>(vr, (m[j].base + rank) * (uint32_t)sizeof(R));
      else
        b.vv[j] = ks_load<R, AUX>(vr, (m[j].base + rank) * (uint32_t)sizeof(T));
    }
  }
}

template <typename R, int KIND, int U, bool VV>
__device__ __forceinline__ void ks_fma(const KsBuf<R, KIND, U, VV> &b, double &are, double &aim) {
#pragma unroll
  for (int j = 0; j < U; j++) {
    const double xre = (double)b.xv[j].x, xim = (double)b.xv[j].y;
    double ar, ai;
    if (KIND == KS_CPX) {
      ar = (double)b.vv[j].x;
      ai = (double)b.vv[j].y;
    } else if (KIND == KS_SCA) {
      ar = (double)b.sv[j] * b.ur[j];
      ai = (double)b.sv[j] * b.ui[j];
    } else {
      ar = b.ur[j];
      ai = b.ui[j];
    }
    are = fma(ar, xre, fma(-ai, xim, are));
    aim = fma(ar, xim, fma(ai, xre, aim));
  }
}

// FP32 pass over U slots of the 32-byte metadata: one scalar load per slot, the
// offsets pre-scaled, products summed in FP32 within the pass and the pass sum
// added to the FP64 row sum.
template <int KIND, int U>
struct KsBuf32 {
  float2 xv[U], vv[U];
  float sv[U];
  float ur[U], ui[U];
};

template <int KIND, bool TAIL, int U, int AUX = 0>
__device__ __forceinline__ void ks_issue32(const qh_kmeta32 *__restrict__ mp, int32_t cnt, uint32_t r8,
                                           __amdgpu_buffer_rsrc_t xr, __amdgpu_buffer_rsrc_t vr,
                                           KsBuf32<KIND, U> &b) {
  qh_kmeta32 m[U];
  uint64_t mask[U];
#pragma unroll
  for (int j = 0; j < U; j++) {
    m[j] = mp[j];
    mask[j] = (!TAIL || j < cnt) ? m[j].mask : 0ull;
    b.ur[j] = m[j].vre;
    b.ui[j] = m[j].vim;
  }
#pragma unroll
  for (int j = 0; j < U; j++) {
    const bool on = __builtin_amdgcn_inverse_ballot_w64(mask[j]);
    const bool neg = __builtin_amdgcn_inverse_ballot_w64(m[j].sign);
    uint32_t off;
    if (KIND != KS_UNI)
      off = neg ? r8 - m[j].ad8 : r8 + m[j].ad8;
    else
      off = r8 + (neg ? m[j].base : m[j].ad8);  // base holds -ad8 for uniform slots
    off = on ? off : QH_KS_OOB;
    b.xv[j] = ks_load<float>(xr, off);
    if (KIND != KS_UNI) {
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask[j] >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)mask[j], 0u));
      if (KIND == KS_SCA)
        b.sv[j] = ks_load_real<float, AUX>(vr, (m[j].base + rank) * 4u);
      else
        b.vv[j] = ks_load<float, AUX>(vr, (m[j].base + rank) * 8u);
    }
  }
}

template <int KIND, int U>
__device__ __forceinline__ void ks_fma32(const KsBuf32<KIND, U> &b, double &are, double &aim) {
  float pr = 0.f, pi = 0.f;
#pragma unroll
  for (int j = 0; j < U; j++) {
    float ar, ai;
    if (KIND == KS_CPX) {
      ar = b.vv[j].x;
      ai = b.vv[j].y;
    } else if (KIND == KS_SCA) {
      ar = b.sv[j] * b.ur[j];
      ai = b.sv[j] * b.ui[j];
    } else {
      ar = b.ur[j];
      ai = b.ui[j];
    }
    pr = fmaf(ar, b.xv[j].x, fmaf(-ai, b.xv[j].y, pr));
    pi = fmaf(ar, b.xv[j].y, fmaf(ai, b.xv[j].x, pi));
  }
  are += (double)pr;
  aim += (double)pi;
}

// Issue / multiply policies over the two metadata layouts, so that one group
// driver serves both; W is the number of slots of a pass.
template <int KIND, int AUX = 0>
struct KsOps32 {
  typedef KsAcc Acc;
  template <int W>
  using Buf = KsBuf32<KIND, W>;
  const qh_kmeta32 *mp;
  uint32_t r8;
  __amdgpu_buffer_rsrc_t xr, vr;
  template <int W, bool TAIL = false>
  __device__ __forceinline__ void issue(Buf<W> &b, int32_t cnt = W) {
    ks_issue32<KIND, TAIL, W, AUX>(mp, cnt, r8, xr, vr, b);
    mp += W;
  }
  template <int W>
  __device__ __forceinline__ void fma(const Buf<W> &b, Acc &a) {
    ks_fma32<KIND, W>(b, a.re, a.im);
  }
  __device__ __forceinline__ void flush(Acc &) {}
};

template <typename R, int KIND, bool VV, int AUX = 0>
struct KsOps {
  typedef KsAcc Acc;
  template <int W>
  using Buf = KsBuf<R, KIND, W, VV>;
  const qh_kmeta *mp;
  uint32_t r8;
  __amdgpu_buffer_rsrc_t xr, vr, mr;
  int32_t k;
  template <int W, bool TAIL = false>
  __device__ __forceinline__ void issue(Buf<W> &b, int32_t cnt = W) {
    ks_issue<R, KIND, TAIL, W, VV, AUX>(mp, cnt, r8, xr, vr, mr, k, b);
    mp += W;
    k += W;
  }
  template <int W>
  __device__ __forceinline__ void fma(const Buf<W> &b, Acc &a) {
    ks_fma<R, KIND, W, VV>(b, a.re, a.im);
  }
  __device__ __forceinline__ void flush(Acc &) {}
};

template <int KIND, int U, bool PIPE, int AUX = 0>
__device__ __forceinline__ void ks_group32(const qh_kmeta32 *__restrict__ meta, int32_t k, int32_t kend,
                                           uint32_t r8, __amdgpu_buffer_rsrc_t xr, __amdgpu_buffer_rsrc_t vr,
                                           KsAcc &acc) {
  KsOps32<KIND, AUX> ops{meta + k, r8, xr, vr};
  ks_run<U, PIPE>(ops, kend - k, acc);
}

template <typename R, int KIND, int U, bool VV, bool PIPE, int AUX = 0>
__device__ __forceinline__ void ks_group(const qh_kmeta *__restrict__ meta, int32_t k, int32_t kend, uint32_t r8,
                                         __amdgpu_buffer_rsrc_t xr, __amdgpu_buffer_rsrc_t vr,
                                         __amdgpu_buffer_rsrc_t mr, KsAcc &acc) {
  KsOps<R, KIND, VV, AUX> ops{meta + k, r8, xr, vr, mr, k};
  ks_run<U, PIPE>(ops, kend - k, acc);
}

// Stage kernel on the sliced layout.  The stage input Y_s = x + c_s K_{s-1} is
// materialised by the previous stage's epilogue, so every stage gathers one
// vector:
//   MODE_SPMV : kout = A xin
//   MODE_STAGE: kout = A xin;  ynext = x + cn * kout          (stages 1-3)
//   MODE_FINAL: xout = x + h/6 (K1 + 2 K2 + 2 K3 + A xin)    (stage 4)
// U slots are processed per pass (all metadata loads, then all gathers, then
// the FMAs); U trades memory-level parallelism against the SGPRs the metadata
// occupies, and is chosen by the autotuner.  The FP32 state is x = hi + lo
// (complex64 pair); its updates are formed in FP64 and split again.
// entry r of a complex vector (scalar type R), read or written with the
// non-temporal hint when NT
template <bool NT, typename R>
__device__ inline typename C2<R>::T ld_nt(const typename C2<R>::T *p, int64_t r) {
  if constexpr (NT) {
    typedef typename C2<R>::V V;
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(p + r));
    typename C2<R>::T t;
    t.x = v[0];
    t.y = v[1];
    return t;
  } else {
    return p[r];
  }
}
template <bool NT, typename R>
__device__ inline void st_nt(typename C2<R>::T *p, int64_t r, typename C2<R>::T t) {
  if constexpr (NT) {
    typedef typename C2<R>::V V;
    V v;
    v[0] = t.x;
    v[1] = t.y;
    __builtin_nontemporal_store(v, reinterpret_cast<V *>(p + r));
  } else {
    p[r] = t;
  }
}

// NT: the matrix values, the state words, K1..K3 and the K and low-word stores
// of the RK4 stages under the non-temporal cache policy (the gathered input and
// the next stage input keep the default one)
// TDP: the row also sums its merged A_k entries of a time-dependent RHS (the
// sliced form over the same 64-row slices), before the epilogue
template <typename R, int MODE, int U, bool VV, bool PIPE, bool NT = false, bool TDP = false>
__global__ void __launch_bounds__(256) k_ks_stage(
    int32_t nrows, int32_t nslices, const int32_t *__restrict__ moff, const int32_t *__restrict__ umid,
    const int32_t *__restrict__ cmid, const qh_kmeta *__restrict__ meta, const qh_kmeta32 *__restrict__ meta32,
    int32_t nmeta, int32_t nval, const typename C2<R>::T *__restrict__ val, int32_t nsval,
    const R *__restrict__ sval,
    const typename C2<R>::T *__restrict__ xin, typename C2<R>::T *__restrict__ kout,
    typename C2<R>::T *__restrict__ ynext, R cn, const typename C2<R>::T *__restrict__ x,
    const typename C2<R>::T *__restrict__ k1, const typename C2<R>::T *__restrict__ k2,
    const typename C2<R>::T *__restrict__ k3, typename C2<R>::T *__restrict__ xout, R h,
    float2 *__restrict__ xlo, double atol, double rtol, double *__restrict__ eacc, qh_td_ell td) {
  typedef typename C2<R>::T T;
  const unsigned b = qh_xcd_block(blockIdx.x, gridDim.x);
  const int s = (int)(b * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6));
  if (s >= nslices) return;  // whole wavefront
  const int lane = threadIdx.x & 63;
  const int32_t r = s * 64 + lane;
  const int32_t k0 = moff[s], km = umid[s], kc = cmid[s], kend = moff[s + 1];
  const __amdgpu_buffer_rsrc_t xr = ks_rsrc(xin, (int64_t)nrows * (int64_t)sizeof(T));
  const __amdgpu_buffer_rsrc_t vr = ks_rsrc(val, (int64_t)nval * (int64_t)sizeof(T));
  const __amdgpu_buffer_rsrc_t sr = ks_rsrc(sval, (int64_t)nsval * (int64_t)sizeof(R));
  const __amdgpu_buffer_rsrc_t mr = ks_rsrc(meta, (int64_t)nmeta * (int64_t)sizeof(qh_kmeta));
  const uint32_t r8 = (uint32_t)r * (uint32_t)sizeof(T);
  KsAcc acc{0.0, 0.0};
  constexpr int AUX = NT ? 2 : 0;
  if (sizeof(R) == 4 && !VV) {
    ks_group32<KS_UNI, U, PIPE>(meta32, k0, km, r8, xr, vr, acc);
    ks_group32<KS_SCA, U, PIPE, AUX>(meta32, km, kc, r8, xr, sr, acc);
    ks_group32<KS_CPX, U, PIPE, AUX>(meta32, kc, kend, r8, xr, vr, acc);
  } else {
    ks_group<R, KS_UNI, U, VV, PIPE>(meta, k0, km, r8, xr, vr, mr, acc);
    ks_group<R, KS_SCA, U, VV, PIPE, AUX>(meta, km, kc, r8, xr, sr, mr, acc);
    ks_group<R, KS_CPX, U, VV, PIPE, AUX>(meta, kc, kend, r8, xr, vr, mr, acc);
  }
  double are = acc.re, aim = acc.im;
  if constexpr (TDP) {
    // (rows past nrows in the last slice read padding: column 0, coefficient 0)
    const int32_t w = td.w[s];
    const int64_t base = td.off[s] + lane;
#pragma unroll 4
    for (int32_t i = 0; i < w; i++) {
      const int64_t e = base + 64 * (int64_t)i;
      const T v = xin[td.col[e]];
      const double2 c = td.cf[td.idx[e]];
      are += c.x * (double)v.x - c.y * (double)v.y;
      aim += c.x * (double)v.y + c.y * (double)v.x;
    }
  }
  if (MODE == MODE_BS3E) {
    // Bogacki-Shampine error stage: k4 = A xn and the weighted squared error of
    // h (b1 - b1^) k1 + ... + h (b4 - b4^) k4 with the weights atol + rtol max(|x|, |xn|)
    // (as qh_rk3bs_err), summed over the wavefront into eacc[slice]
    double v = 0.0;
    if (r < nrows) {
      const double c1 = 2.0 / 9.0 - 7.0 / 24.0, c2 = 1.0 / 3.0 - 1.0 / 4.0, c3 = 4.0 / 9.0 - 1.0 / 3.0,
                   c4 = -1.0 / 8.0;
      const T a = k1[r], bq = k2[r], cq = k3[r], xo = x[r], xn = xin[r];
      const double hh = (double)h;
      const double er = hh * (c1 * a.x + c2 * bq.x + c3 * cq.x + c4 * are);
      const double ei = hh * (c1 * a.y + c2 * bq.y + c3 * cq.y + c4 * aim);
      const double ax = sqrt((double)xo.x * xo.x + (double)xo.y * xo.y);
      const double an = sqrt((double)xn.x * xn.x + (double)xn.y * xn.y);
      const double sc = atol + rtol * (ax > an ? ax : an);
      v = (er * er + ei * ei) / (sc * sc);
      T y;
      y.x = (R)are;
      y.y = (R)aim;
      kout[r] = y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) eacc[s] = v;
    return;
  }
  if (r >= nrows) return;
  T y;
  y.x = (R)are;
  y.y = (R)aim;
  if (MODE == MODE_SPMV) {
    kout[r] = y;
    return;
  }
  if (MODE == MODE_BS3) {
    // Bogacki-Shampine stage 3: k3 = A Y3 and x_n = x + h (2/9 k1 + 1/3 k2 + 4/9 k3)
    kout[r] = y;
    const T xo = x[r], a = k1[r], bq = k2[r];
    const double hh = (double)h;
    T o;
    o.x = (R)((double)xo.x + hh * (2.0 / 9.0) * a.x + hh / 3.0 * bq.x + hh * (4.0 / 9.0) * are);
    o.y = (R)((double)xo.y + hh * (2.0 / 9.0) * a.y + hh / 3.0 * bq.y + hh * (4.0 / 9.0) * aim);
    xout[r] = o;
    return;
  }
  const T xo = ld_nt<NT, R>(x, r);
  if (MODE == MODE_STAGE) {
    st_nt<NT, R>(kout, r, y);
    T yn;
    if (sizeof(R) == 8) {
      yn.x = xo.x + cn * y.x;
      yn.y = xo.y + cn * y.y;
    } else {
      // the complex64 stage input does not need the low word of the state (it
      // would change it by at most one rounding); stage 4 uses both words
      yn.x = (R)((double)xo.x + (double)cn * are);
      yn.y = (R)((double)xo.y + (double)cn * aim);
    }
    ynext[r] = yn;
    return;
  }
  const T a = ld_nt<NT, R>(k1, r), bq = ld_nt<NT, R>(k2, r), cq = ld_nt<NT, R>(k3, r);
  if (sizeof(R) == 8) {
    const R s6 = h * (R)(1.0 / 6.0), s3 = h * (R)(1.0 / 3.0);
    const R dre = s6 * a.x + s3 * bq.x + s3 * cq.x + s6 * y.x;
    const R dim = s6 * a.y + s3 * bq.y + s3 * cq.y + s6 * y.y;
    T o;
    o.x = xo.x + dre;
    o.y = xo.y + dim;
    xout[r] = o;
  } else {
    // x' = hi + lo + d in FP64, split again into (hi', lo')
    const double s6 = (double)h * (1.0 / 6.0), s3 = (double)h * (1.0 / 3.0);
    const double dre = s6 * (double)a.x + s3 * (double)bq.x + s3 * (double)cq.x + s6 * are;
    const double dim = s6 * (double)a.y + s3 * (double)bq.y + s3 * (double)cq.y + s6 * aim;
    const float2 lo = ld_nt<NT, float>(xlo, r);
    const double nr = ((double)xo.x + (double)lo.x) + dre, ni = ((double)xo.y + (double)lo.y) + dim;
    T o;
    o.x = (R)nr;
    o.y = (R)ni;
    xout[r] = o;
    st_nt<NT, float>(xlo, r, make_float2((float)(nr - (double)o.x), (float)(ni - (double)o.y)));
  }
}

// Build of the sliced layout from the CSR: one 64-lane block per slice.
// Each lane sorts its row's entries by key = (|col-row| << 1 | rank) where rank
// separates the two entries row-ad and row+ad of one row; the slice's slots are
// the union of the lanes' keys in increasing order (a wave-wide min merge).
const int KS_MAXROW = 64;

// Per-slice counts (FILL false): cnt[0*ns + s] slots, [1] complex values,
// [2] uniform slots, [3] scalar slots, [4] scalar values.  FILL writes the
// metadata (uniform, scalar, complex groups at moff, moff + uniform,
// moff + uniform + scalar) and the values (complex at voff, scalar at svoff).
template <bool FILL>
__global__ void __launch_bounds__(64) k_ks_build(int32_t nrows, const int32_t *__restrict__ rp,
                                                 const int32_t *__restrict__ col, const double2 *__restrict__ val,
                                                 int32_t *cnt, int32_t ns, const int32_t *moff,
                                                 const int32_t *voff, const int32_t *svoff, qh_kmeta *meta,
                                                 double2 *kval, double *sval, int *overflow) {
  __shared__ uint64_t keys[64][KS_MAXROW];
  const int s = blockIdx.x, lane = threadIdx.x;
  const int32_t r = s * 64 + lane;
  int len = 0, beg = 0;
  if (r < nrows) {
    beg = rp[r];
    len = rp[r + 1] - beg;
    if (len > KS_MAXROW) {
      overflow[0] = 1;
      len = 0;
    }
  }
  uint64_t *kk = keys[lane];
  // key0 = ad << 1 | (d > 0), packed with the entry index in the low 6 bits
  for (int e = 0; e < len; e++) {
    const int32_t d = col[beg + e] - r;
    const uint64_t ad = (uint64_t)(d < 0 ? -(int64_t)d : (int64_t)d);
    const uint64_t key0 = (ad << 1) | (d > 0 ? 1u : 0u);
    uint64_t v = (key0 << 6) | (uint64_t)e;
    int p = e;
    while (p > 0 && kk[p - 1] > v) {
      kk[p] = kk[p - 1];
      p--;
    }
    kk[p] = v;
  }
  // rank: the second entry with the same |d| gets rank 1
  uint64_t prev_ad = ~0ull;
  for (int e = 0; e < len; e++) {
    const uint64_t ad = kk[e] >> 7;
    const uint64_t rank = (ad == prev_ad) ? 1u : 0u;
    kk[e] = (((ad << 1) | rank) << 6) | (kk[e] & 63u);
    prev_ad = ad;
  }
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int p = 0, slot = 0, uslot = 0, sslot = 0, cslot = 0;
  int32_t base = FILL ? voff[s] : 0;    // complex values of this slice start here
  int32_t sbase = FILL ? svoff[s] : 0;  // scalar values of this slice start here
  const int32_t mo = FILL ? moff[s] : 0;
  const int32_t ms = FILL ? mo + cnt[2 * ns + s] : 0;  // scalar slots follow the uniform ones
  const int32_t mc = FILL ? ms + cnt[3 * ns + s] : 0;  // complex slots follow the scalar ones
  for (;;) {
    const uint64_t cur = p < len ? (kk[p] >> 6) : ~0ull;
    const uint64_t m = wave_min_u64(cur);
    if (m == ~0ull) break;
    const bool act = (cur == m);
    const uint64_t mask = __ballot(act);
    const int e = act ? (int)(kk[p] & 63u) : 0;
    const double2 v = act ? val[beg + e] : make_double2(0.0, 0.0);
    // uniform slot: every active lane holds the bit-identical value of the first one
    const int first = __ffsll((unsigned long long)mask) - 1;
    const double fre = __shfl(v.x, first, 64), fim = __shfl(v.y, first, 64);
    const bool same = !act || (__double_as_longlong(v.x) == __double_as_longlong(fre) &&
                               __double_as_longlong(v.y) == __double_as_longlong(fim));
    const bool uni = __ballot(same) == ~0ull;
    // scalar slot: all values real (or all imaginary)
    const bool isre = !uni && __ballot(act && v.y != 0.0) == 0ull;
    const bool isim = !uni && !isre && __ballot(act && v.x != 0.0) == 0ull;
    const bool sca = isre || isim;
    if (FILL) {
      const bool neg = act && (col[beg + e] < r);
      const uint64_t sign = __ballot(neg);
      const uint32_t rank = (uint32_t)__popcll(mask & lt);
      if (lane == 0) {
        qh_kmeta md;
        md.mask = mask;
        md.sign = sign;
        md.ad = (uint32_t)(m >> 1);
        md.base = uni ? QH_KS_UNIFORM : (uint32_t)(sca ? sbase : base);
        md.vre = uni ? fre : sca ? (isre ? 1.0 : 0.0) : 0.0;
        md.vim = uni ? fim : sca ? (isre ? 0.0 : 1.0) : 0.0;
        meta[uni ? mo + uslot : sca ? ms + sslot : mc + cslot] = md;
      }
      if (act && sca) sval[sbase + rank] = isre ? v.x : v.y;
      if (act && !uni && !sca) kval[base + rank] = v;
    }
    if (sca) sbase += __popcll(mask);
    else if (!uni) base += __popcll(mask);
    if (act) p++;
    slot++;
    if (uni)
      uslot++;
    else if (sca)
      sslot++;
    else
      cslot++;
  }
  if (!FILL && lane == 0) {
    cnt[s] = slot;
    cnt[ns + s] = base;
    cnt[2 * ns + s] = uslot;
    cnt[3 * ns + s] = sslot;
    cnt[4 * ns + s] = sbase;
  }
}

__global__ void k_narrow_ks(const double2 *a, float2 *b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = make_float2((float)a[i].x, (float)a[i].y);
}
__global__ void k_narrow_ks_real(const double *a, float *b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

int ks_build(qh_mat *A) {
  if (A->ks_state) return QH_OK;
  hipStream_t s = qh_stream();
  const int32_t ns = (int32_t)((A->rows + 63) / 64);
  // the stage kernels address vectors and the value streams through buffer
  // resources: byte offsets must stay below 2^31 (the value streams are checked
  // once their sizes are known)
  if (A->rows * (int64_t)sizeof(double2) >= ((int64_t)1 << 31)) {
    A->ks_state = -1;
    return QH_OK;
  }
  int32_t *cnt = nullptr, *voff = nullptr, *svoff = nullptr;
  int *ovf = nullptr;
  int rc = qh_alloc_t(&cnt, (size_t)5 * ns + 1);
  if (!rc) rc = qh_alloc_t(&voff, (size_t)ns + 1);
  if (!rc) rc = qh_alloc_t(&svoff, (size_t)ns + 1);
  if (!rc) rc = qh_alloc_t(&ovf, 1);
  if (!rc) rc = qh_alloc_t(&A->ks_moff, (size_t)ns + 1);
  if (!rc) rc = qh_alloc_t(&A->ks_umid, (size_t)ns);
  if (!rc) rc = qh_alloc_t(&A->ks_cmid, (size_t)ns);
  auto cleanup = [&]() {
    qh_free_ptr(cnt);
    qh_free_ptr(voff);
    qh_free_ptr(svoff);
    qh_free_ptr(ovf);
  };
  auto drop = [&]() {
    qh_free_ptr(A->ks_moff);
    qh_free_ptr(A->ks_umid);
    qh_free_ptr(A->ks_cmid);
    A->ks_moff = A->ks_umid = A->ks_cmid = nullptr;
  };
  if (rc) {
    cleanup();
    drop();
    return rc;
  }
  QH_HIP(hipMemsetAsync(ovf, 0, sizeof(int), s));
  hipLaunchKernelGGL((k_ks_build<false>), dim3(ns), dim3(64), 0, s, (int32_t)A->rows, A->rowptr, A->col, A->val,
                     cnt, ns, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ovf);
  QH_HIP(hipGetLastError());
  int hovf = 0;
  std::vector<int32_t> hc((size_t)5 * ns);
  QH_HIP(hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost, s));
  QH_HIP(hipMemcpyAsync(hc.data(), cnt, sizeof(int32_t) * 5 * ns, hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  const int32_t *hw = hc.data(), *hn = hw + ns, *hu = hn + ns, *hs = hu + ns, *hsv = hs + ns;
  int64_t nstored = 0, nsstored = 0;
  for (int32_t i = 0; i < ns; i++) {
    nstored += hn[i];
    nsstored += hsv[i];
  }
  if (nstored * (int64_t)sizeof(double2) >= ((int64_t)1 << 31) ||
      nsstored * (int64_t)sizeof(double) >= ((int64_t)1 << 31))
    hovf = 1;
  if (hovf) {  // a row longer than the slice builder handles, or too many stored values: keep the CSR kernels
    A->ks_state = -1;
    cleanup();
    drop();
    return QH_OK;
  }
  std::vector<int32_t> ho(ns + 1), hv(ns + 1), hsvo(ns + 1), hm(ns), hcm(ns);
  int64_t tot = 0, totv = 0, tots = 0;
  for (int32_t i = 0; i < ns; i++) {
    ho[i] = (int32_t)tot;
    hv[i] = (int32_t)totv;
    hsvo[i] = (int32_t)tots;
    hm[i] = (int32_t)(tot + hu[i]);
    hcm[i] = (int32_t)(tot + hu[i] + hs[i]);
    tot += hw[i];
    totv += hn[i];
    tots += hsv[i];
  }
  ho[ns] = (int32_t)tot;
  hv[ns] = (int32_t)totv;
  hsvo[ns] = (int32_t)tots;
  QH_HIP(hipMemcpyAsync(A->ks_moff, ho.data(), sizeof(int32_t) * (ns + 1), hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(A->ks_umid, hm.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(A->ks_cmid, hcm.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(voff, hv.data(), sizeof(int32_t) * (ns + 1), hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(svoff, hsvo.data(), sizeof(int32_t) * (ns + 1), hipMemcpyHostToDevice, s));
  rc = qh_alloc_t(&A->ks_meta, (size_t)tot + QH_KS_PAD);
  if (!rc) QH_HIP(hipMemsetAsync(A->ks_meta + tot, 0, sizeof(qh_kmeta) * QH_KS_PAD, s));
  if (!rc) rc = qh_alloc_t(&A->ks_val, (size_t)(totv > 0 ? totv : 1));
  if (!rc) rc = qh_alloc_t(&A->ks_sval, (size_t)(tots > 0 ? tots : 1));
  if (rc) {
    cleanup();
    return rc;
  }
  hipLaunchKernelGGL((k_ks_build<true>), dim3(ns), dim3(64), 0, s, (int32_t)A->rows, A->rowptr, A->col, A->val,
                     cnt, ns, A->ks_moff, voff, svoff, A->ks_meta, A->ks_val, A->ks_sval, ovf);
  QH_HIP(hipGetLastError());
  QH_HIP(hipStreamSynchronize(s));
  cleanup();
  A->nslices = ns;
  A->ks_slots = tot;
  A->ks_nval = totv;
  A->ks_nsval = tots;
  A->ks_state = 1;
  return QH_OK;
}

__global__ void k_make_meta32(const qh_kmeta *m, int64_t n, qh_kmeta32 *o) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const qh_kmeta a = m[i];
    qh_kmeta32 b;
    b.mask = a.mask;
    b.sign = a.sign;
    b.ad8 = a.ad * 8u;
    b.base = a.base == QH_KS_UNIFORM ? (uint32_t)(-(int32_t)(a.ad * 8u)) : a.base;
    b.vre = (float)a.vre;
    b.vim = (float)a.vim;
    o[i] = b;
  }
}

int ks_make_fp32(qh_mat *A) {
  if (A->ks_state != 1 || (A->ks_val32 && A->ks_sval32 && A->ks_meta32)) return QH_OK;
  int rc = QH_OK;
  if (!A->ks_val32) {
    rc = qh_alloc_t(&A->ks_val32, (size_t)(A->ks_nval > 0 ? A->ks_nval : 1));
    if (rc) return rc;
    if (A->ks_nval)
      hipLaunchKernelGGL(k_narrow_ks, dim3(qh_blocks(A->ks_nval, 256)), dim3(256), 0, qh_stream(), A->ks_val,
                         A->ks_val32, A->ks_nval);
  }
  if (!A->ks_sval32) {
    rc = qh_alloc_t(&A->ks_sval32, (size_t)(A->ks_nsval > 0 ? A->ks_nsval : 1));
    if (rc) return rc;
    if (A->ks_nsval)
      hipLaunchKernelGGL(k_narrow_ks_real, dim3(qh_blocks(A->ks_nsval, 256)), dim3(256), 0, qh_stream(),
                         A->ks_sval, A->ks_sval32, A->ks_nsval);
  }
  if (!A->ks_meta32) {
    const int64_t nm = A->ks_slots + QH_KS_PAD;
    rc = qh_alloc_t(&A->ks_meta32, (size_t)nm);
    if (rc) return rc;
    hipLaunchKernelGGL(k_make_meta32, dim3(qh_blocks(nm, 256)), dim3(256), 0, qh_stream(), A->ks_meta, nm,
                       A->ks_meta32);
  }
  QH_HIP(hipGetLastError());
  return QH_OK;
}

// launch the sliced-layout kernel with the slot unroll encoded in the variant code G
template <typename R, int MODE>
int launch_ks(int G, hipStream_t s, const qh_mat *A, const void *xin, void *kout, void *ynext, R cn, const void *x,
              const void *k1, const void *k2, const void *k3, void *xout, R h, float2 *xlo, double atol = 0.0,
              double rtol = 0.0, double *eacc = nullptr, const qh_td_ell *td = nullptr) {
  typedef typename C2<R>::T T;
  const T *kv = sizeof(R) == 8 ? (const T *)A->ks_val : (const T *)A->ks_val32;
  const R *sv = sizeof(R) == 8 ? (const R *)A->ks_sval : (const R *)A->ks_sval32;
  const dim3 grid(qh_blocks(A->nslices, 4)), blk(256);
  const int32_t n = (int32_t)A->rows, ns = A->nslices;
  const qh_td_ell none{nullptr, nullptr, nullptr, nullptr, nullptr};
#define QH_LKT(UU, VV, PP, NN, TT, TD)                                                                             \
  hipLaunchKernelGGL((k_ks_stage<R, MODE, UU, VV, PP, NN, TT>), grid, blk, 0, s, n, ns, A->ks_moff, A->ks_umid,     \
                     A->ks_cmid,                                                                              \
                     A->ks_meta, A->ks_meta32, (int32_t)(A->ks_slots + QH_KS_PAD), (int32_t)A->ks_nval, kv,    \
                     (int32_t)A->ks_nsval, sv, (const T *)xin, (T *)kout, (T *)ynext, cn, (const T *)x,       \
                     (const T *)k1, (const T *)k2, (const T *)k3, (T *)xout, h, xlo, atol, rtol, eacc, TD)
#define QH_LK(UU, VV, PP, NN) QH_LKT(UU, VV, PP, NN, false, none)
  if constexpr (MODE == MODE_SPMV || MODE == MODE_STAGE || MODE == MODE_FINAL ||
                ((MODE == MODE_BS3 || MODE == MODE_BS3E) && sizeof(R) == 8)) {
    if (td) {  // the time-dependent entries summed in the same kernel (variants 128, 134, 136)
      switch (G) {
        case 128: QH_LKT(2, false, false, false, true, *td); return QH_OK;
        case 134: QH_LKT(4, false, true, false, true, *td); return QH_OK;
        case 136: QH_LKT(4, true, false, true, true, *td); return QH_OK;
        default:
          qh_set_error("variant %d has no time-dependent form", G);
          return QH_EINVAL;
      }
    }
  } else if (td) {
    qh_set_error("no time-dependent form of this stage");
    return QH_EINVAL;
  }
  switch (G) {
    case 128: QH_LK(2, false, false, false); break;
    case 129: QH_LK(4, false, false, false); break;
    case 130: QH_LK(8, false, false, false); break;
    case 131: QH_LK(4, true, false, false); break;
    case 132: QH_LK(8, true, false, false); break;
    case 133: QH_LK(2, false, true, false); break;
    case 134: QH_LK(4, false, true, false); break;
    case 135: QH_LK(4, false, false, true); break;
    case 136: QH_LK(4, true, false, true); break;
    default:
      qh_set_error("unsupported sliced-layout variant %d", G);
      return QH_EINVAL;
  }
#undef QH_LK
#undef QH_LKT
  return QH_OK;
}

// a stage on either sliced layout (the row-pair one is FP32 only)
template <typename R, int MODE>
int launch_sliced(int G, hipStream_t s, const qh_mat *A, const void *xin, void *kout, void *ynext, R cn,
                  const void *x, const void *k1, const void *k2, const void *k3, void *xout, R h, float2 *xlo) {
  if (QH_IS_KS2(G)) {
    if (sizeof(R) != 4) {
      qh_set_error("the row-pair sliced layout (variant %d) is FP32 only", G);
      return QH_EINVAL;
    }
    return ks2_launch(G, MODE, s, A, xin, kout, ynext, (float)cn, x, k1, k2, k3, xout, (float)h, xlo);
  }
  return launch_ks<R, MODE>(G, s, A, xin, kout, ynext, cn, x, k1, k2, k3, xout, h, xlo);
}

template <typename R, int MODE>
int launch_stage_g(int G, hipStream_t s, const qh_mat *A, const typename C2<R>::T *val, const void *x,
                   const void *kin, R ck, void *kout, const void *k1, const void *k2, const void *k3, void *xout,
                   R h, float2 *xlo) {
  typedef typename C2<R>::T T;
  const int32_t n = (int32_t)A->rows;
  if (QH_IS_KSELL(G) || QH_IS_KS2(G)) {  // plain y = A x on a sliced layout (stages go through launch_sliced)
    if (MODE != MODE_SPMV) {
      qh_set_error("sliced-layout stages go through launch_sliced");
      return QH_EINVAL;
    }
    return launch_sliced<R, MODE_SPMV>(G, s, A, x, kout, nullptr, (R)0, nullptr, nullptr, nullptr, nullptr, nullptr,
                                       (R)0, nullptr);
  }
  const unsigned grid = qh_blocks((int64_t)n * G, 256);
#define QH_L(GG)                                                                                                 \
  hipLaunchKernelGGL((k_rk_stage<R, GG, MODE>), dim3(grid), dim3(256), 0, s, n, A->rowptr, A->col, val,          \
                     (const T *)x, (const T *)kin, ck, (T *)kout, (const T *)k1, (const T *)k2, (const T *)k3,   \
                     (T *)xout, h, xlo)
  switch (G) {
    case 1: QH_L(1); break;
    case 2: QH_L(2); break;
    case 4: QH_L(4); break;
    case 8: QH_L(8); break;
    case 16: QH_L(16); break;
    case 32: QH_L(32); break;
    case 64: QH_L(64); break;
    default:
      qh_set_error("unsupported lanes-per-row %d", G);
      return QH_EINVAL;
  }
#undef QH_L
  return QH_OK;
}

// ------------------------------------------------------------ autotuner
// Time each variant (CSR lanes per row, sliced layouts) on the real matrix with
// one plain SpMV launch plus one RK stage launch (the stage epilogue's vector
// traffic and registers are part of what the stepping runs) and keep the fastest.
template <typename R>
int autotune(qh_mat *A, int *out_g) {
  typedef typename C2<R>::T T;
  const typename C2<R>::T *val = sizeof(R) == 8 ? (const T *)A->val : (const T *)A->val32;
  // the buffers of one RK4 step: state, K1..K3, the two stage inputs, the new state
  void *x = nullptr, *kk[3] = {nullptr, nullptr, nullptr}, *ya = nullptr, *yb = nullptr, *xb = nullptr;
  float2 *lo = nullptr;
  const size_t vb = (size_t)A->rows * sizeof(T);
  int rc = qh_alloc(&x, vb);
  for (int i = 0; i < 3 && !rc; i++) rc = qh_alloc(&kk[i], vb);
  if (!rc) rc = qh_alloc(&ya, vb);
  if (!rc) rc = qh_alloc(&yb, vb);
  if (!rc) rc = qh_alloc(&xb, vb);
  if (!rc && sizeof(R) == 4) rc = qh_alloc_t(&lo, (size_t)A->rows);
  auto release = [&]() {
    qh_free_ptr(x);
    for (void *p : kk) qh_free_ptr(p);
    qh_free_ptr(ya);
    qh_free_ptr(yb);
    qh_free_ptr(xb);
    qh_free_ptr(lo);
  };
  if (rc) {
    release();
    return rc;
  }
  hipStream_t s = qh_stream();
  QH_HIP(hipMemsetAsync(x, 0, vb, s));
  if (lo) QH_HIP(hipMemsetAsync(lo, 0, (size_t)A->rows * sizeof(float2), s));
  // one whole RK4 step with the launch sequence of rk4_step (three stage
  // launches and the final one; the final stage moves the most data), so the
  // choice reflects the mix a step runs
  const R h = (R)0.01;
  auto run = [&](int g) {
    if (QH_IS_KSELL(g) || QH_IS_KS2(g)) {
      const void *ins[4] = {x, ya, yb, ya};
      void *outs[3] = {ya, yb, ya};
      const R cns[3] = {h / 2, h / 2, h};
      for (int st = 0; st < 3; st++)
        launch_sliced<R, MODE_STAGE>(g, s, A, ins[st], kk[st], outs[st], cns[st], x, nullptr, nullptr, nullptr,
                                     nullptr, h, lo);
      launch_sliced<R, MODE_FINAL>(g, s, A, ins[3], nullptr, nullptr, (R)0, x, kk[0], kk[1], kk[2], xb, h, lo);
    } else {
      launch_stage_g<R, MODE_SPMV>(g, s, A, val, x, nullptr, (R)0, kk[0], nullptr, nullptr, nullptr, nullptr, h,
                                   lo);
      launch_stage_g<R, MODE_STAGE>(g, s, A, val, x, kk[0], h / 2, kk[1], nullptr, nullptr, nullptr, nullptr, h,
                                    lo);
      launch_stage_g<R, MODE_STAGE>(g, s, A, val, x, kk[1], h / 2, kk[2], nullptr, nullptr, nullptr, nullptr, h,
                                    lo);
      launch_stage_g<R, MODE_FINAL>(g, s, A, val, x, kk[2], h, nullptr, kk[0], kk[1], kk[2], xb, h, lo);
    }
  };
  hipEvent_t e0, e1;
  QH_HIP(hipEventCreate(&e0));
  QH_HIP(hipEventCreate(&e1));
  const double avg = (double)A->nnz / (double)(A->rows > 0 ? A->rows : 1);
  std::vector<int> cands;
  {
    int brc = ks_build(A);
    if (!brc && A->ks_state == 1 && sizeof(R) == 4) brc = ks_make_fp32(A);
    if (brc) return brc;
    if (A->ks_state == 1)
      for (int u = 128; u <= 136; u++) cands.push_back(u);
    if (sizeof(R) == 4) {
      brc = ks2_build(A);
      if (brc) return brc;
      if (A->ks2_state == 1)
        for (int u = 140; u <= 144; u++) cands.push_back(u);
    }
  }
  // the CSR kernels only when no sliced layout exists or the matrix is small:
  // on a large one they take 5-30x the sliced step and are never chosen
  const bool sliced = A->ks_state == 1 || A->ks2_state == 1;
  if (!sliced || A->rows < (1 << 16))
    for (int g : {2, 4, 8, 16, 32, 64})
      if (g <= 4 * avg + 4) cands.push_back(g);
  if (cands.empty()) cands.push_back(2);
  float best = 1e30f;
  int bestg = cands[0];
  for (int g : cands) {
    for (int w = 0; w < 2; w++) run(g);
    QH_HIP(hipEventRecord(e0, s));
    const int reps = 5;
    for (int w = 0; w < reps; w++) run(g);
    QH_HIP(hipEventRecord(e1, s));
    QH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    QH_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (getenv("QH_AUTOTUNE_VERBOSE"))
      fprintf(stderr, "qh autotune: %s rows %lld nnz %lld variant %d: %.2f us per RK4 step\n",
              sizeof(R) == 8 ? "fp64" : "fp32", (long long)A->rows, (long long)A->nnz, g, 1e3 * ms / reps);
    // a later candidate has to beat the current choice by 2 % (1 % once a rep
    // takes over 100 us, where the launch-to-launch noise is well below that),
    // so that variants within timing noise of each other do not flip from run to run
    const float margin = best / reps > 0.1f ? 0.99f : 0.98f;
    if (ms < margin * best) {
      best = ms;
      bestg = g;
    }
  }
  QH_HIP(hipEventDestroy(e0));
  QH_HIP(hipEventDestroy(e1));
  release();
  *out_g = bestg;
  return QH_OK;
}

// The autotuner's choice is remembered per (precision, rows, nnz) for the rest
// of the process, so the matrices of a parameter sweep (one structure, other
// values) are timed once; the variant decides only the speed (every variant
// computes the same product).
struct TuneMemo {
  int prec;
  int64_t rows, nnz;
  int g;
};
static TuneMemo g_tune[64];
static int g_ntune = 0;

template <typename R>
int autotune_memo(qh_mat *A, int *out_g) {
  const int prec = sizeof(R) == 8 ? QH_FP64 : QH_FP32;
  for (int i = 0; i < g_ntune && i < 64; i++)
    if (g_tune[i].prec == prec && g_tune[i].rows == A->rows && g_tune[i].nnz == A->nnz) {
      *out_g = g_tune[i].g;
      return QH_OK;
    }
  int rc = autotune<R>(A, out_g);
  if (rc) return rc;
  g_tune[g_ntune % 64] = TuneMemo{prec, A->rows, A->nnz, *out_g};
  g_ntune++;
  return QH_OK;
}

int lanes_for(qh_mat *A, int prec, int *g) {
  int rc = QH_OK;
  if (prec == QH_FP64) {
    if (!A->lanes64) rc = autotune_memo<double>(A, &A->lanes64);
    if (!rc && QH_IS_KSELL(A->lanes64)) rc = ks_build(A);
    if (!rc && QH_IS_KSELL(A->lanes64) && A->ks_state != 1) A->lanes64 = 8;
    *g = A->lanes64;
  } else {
    if (!A->val32) rc = qh_mat_make_fp32(A);
    if (!rc && !A->lanes32) rc = autotune_memo<float>(A, &A->lanes32);
    if (!rc && QH_IS_KSELL(A->lanes32)) rc = ks_build(A);
    if (!rc && QH_IS_KSELL(A->lanes32)) rc = ks_make_fp32(A);
    if (!rc && QH_IS_KSELL(A->lanes32) && A->ks_state != 1) A->lanes32 = 8;
    if (!rc && QH_IS_KS2(A->lanes32)) rc = ks2_build(A);
    if (!rc && QH_IS_KS2(A->lanes32) && A->ks2_state != 1) A->lanes32 = 8;
    *g = A->lanes32;
  }
  return rc;
}

// ------------------------------------------------------------ trace renorm
// FP64 trace of a complex64 (hi+lo) density matrix: per-block partial sums,
// then one block sums them and writes 1/Tr; the scale kernel reads it.
__global__ void __launch_bounds__(256) k_trace_partial_f32(const float2 *hi, const float2 *lo, int64_t tot,
                                                           int64_t base, double *partial) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < tot; i += (int64_t)gridDim.x * 256) {
    int64_t loc = base + i * (tot + 1);
    acc += (double)hi[loc].x + (double)lo[loc].x;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_trace_final(const double *partial, int n, double *inv) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) inv[0] = (sh[0] != 0.0) ? 1.0 / sh[0] : 1.0;
}

__global__ void k_scale_by_dev(float2 *hi, float2 *lo, int64_t n, const double *inv) {
  const double s = inv[0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double xr = ((double)hi[i].x + (double)lo[i].x) * s, xi = ((double)hi[i].y + (double)lo[i].y) * s;
    float hr = (float)xr, hm = (float)xi;
    hi[i] = make_float2(hr, hm);
    lo[i] = make_float2((float)(xr - (double)hr), (float)(xi - (double)hm));
  }
}

}  // namespace

// ------------------------------------------------------------ stepper state
struct Stepper {
  const qh_mat *A = nullptr;
  uint64_t Auid = 0;  // A's serial number: a new matrix may reuse a destroyed one's address
  int prec = 0;
  int64_t n = 0;
  void *k[3] = {nullptr, nullptr, nullptr};
  void *y[2] = {nullptr, nullptr};  // materialised stage inputs (sliced-layout path)
  void *xalt = nullptr;  // double buffer for the state (hi part for FP32)
  double *partial = nullptr, *inv = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  // parameters the graph was captured with
  void *gx = nullptr;
  float2 *glo = nullptr;
  double gdt = 0;
  int gsteps = 0, grenorm = 0, gG = 0;
  int64_t gtot = 0;
};

static Stepper g_st;
static int g_graph_steps = 8;
static double g_last_stage_ms = 0.0;
static int64_t g_last_launches = 0;

static void stepper_release() {
  if (g_st.exec) (void)hipGraphExecDestroy(g_st.exec);
  if (g_st.graph) (void)hipGraphDestroy(g_st.graph);
  for (auto &p : g_st.k) qh_free_ptr(p);
  for (auto &p : g_st.y) qh_free_ptr(p);
  qh_free_ptr(g_st.xalt);
  qh_free_ptr(g_st.partial);
  qh_free_ptr(g_st.inv);
  g_st = Stepper();
}

static int stepper_prepare(const qh_mat *A, int prec, int64_t n) {
  if (g_st.A == A && g_st.Auid == A->uid && g_st.prec == prec && g_st.n == n && g_st.k[0]) return QH_OK;
  (void)hipStreamSynchronize(qh_stream());
  stepper_release();
  size_t es = prec == QH_FP64 ? 16 : 8;
  int rc = QH_OK;
  for (auto &p : g_st.k)
    if (!rc) rc = qh_alloc(&p, (size_t)n * es);
  for (auto &p : g_st.y)
    if (!rc) rc = qh_alloc(&p, (size_t)n * es);
  if (!rc) rc = qh_alloc(&g_st.xalt, (size_t)n * es);
  if (!rc) rc = qh_alloc_t(&g_st.partial, 1024);
  if (!rc) rc = qh_alloc_t(&g_st.inv, 1);
  if (rc) {
    stepper_release();
    return rc;
  }
  g_st.A = A;
  g_st.Auid = A->uid;
  g_st.prec = prec;
  g_st.n = n;
  return QH_OK;
}

// one RK4 step from xa into xb (xb != xa), launched on stream s
template <typename R>
static int rk4_step(hipStream_t s, const qh_mat *A, int G, const void *xa, void *xb, float2 *xlo, R h,
                    hipEvent_t *ev) {
  typedef typename C2<R>::T T;
  const T *val = sizeof(R) == 8 ? (const T *)A->val : (const T *)A->val32;
  void **k = g_st.k;
  int rc;
  if (QH_IS_KSELL(G) || QH_IS_KS2(G)) {
    // stage inputs materialised in the epilogues: x -> Ya -> Yb -> Ya
    void *Ya = g_st.y[0], *Yb = g_st.y[1];
    const void *ins[4] = {xa, Ya, Yb, Ya};
    void *outs[3] = {Ya, Yb, Ya};
    const R cns[3] = {h / 2, h / 2, h};
    for (int st = 0; st < 3; st++) {
      if (ev) QH_HIP(hipEventRecord(ev[st], s));
      rc = launch_sliced<R, MODE_STAGE>(G, s, A, ins[st], k[st], outs[st], cns[st], xa, nullptr, nullptr, nullptr,
                                        nullptr, h, xlo);
      if (rc) return rc;
    }
    if (ev) QH_HIP(hipEventRecord(ev[3], s));
    rc = launch_sliced<R, MODE_FINAL>(G, s, A, ins[3], nullptr, nullptr, (R)0, xa, k[0], k[1], k[2], xb, h, xlo);
    if (ev) QH_HIP(hipEventRecord(ev[4], s));
    return rc;
  }
  if (ev) QH_HIP(hipEventRecord(ev[0], s));
  rc = launch_stage_g<R, MODE_SPMV>(G, s, A, val, xa, nullptr, (R)0, k[0], nullptr, nullptr, nullptr, nullptr, h,
                                    xlo);
  if (ev) QH_HIP(hipEventRecord(ev[1], s));
  if (!rc)
    rc = launch_stage_g<R, MODE_STAGE>(G, s, A, val, xa, k[0], h / 2, k[1], nullptr, nullptr, nullptr, nullptr, h,
                                       xlo);
  if (ev) QH_HIP(hipEventRecord(ev[2], s));
  if (!rc)
    rc = launch_stage_g<R, MODE_STAGE>(G, s, A, val, xa, k[1], h / 2, k[2], nullptr, nullptr, nullptr, nullptr, h,
                                       xlo);
  if (ev) QH_HIP(hipEventRecord(ev[3], s));
  if (!rc)
    rc = launch_stage_g<R, MODE_FINAL>(G, s, A, val, xa, k[2], h, nullptr, k[0], k[1], k[2], xb, h, xlo);
  if (ev) QH_HIP(hipEventRecord(ev[4], s));
  return rc;
}

static int renorm_launch(hipStream_t s, qh_vec *x, void *hi, int64_t tot, int64_t base) {
  const int nb = 512;
  hipLaunchKernelGGL(k_trace_partial_f32, dim3(nb), dim3(256), 0, s, (const float2 *)hi, x->lo, tot, base,
                     g_st.partial);
  hipLaunchKernelGGL(k_trace_final, dim3(1), dim3(256), 0, s, g_st.partial, nb, g_st.inv);
  hipLaunchKernelGGL(k_scale_by_dev, dim3(std::min<int64_t>(qh_blocks(x->n, 256), 8192)), dim3(256), 0, s,
                     (float2 *)hi, x->lo, x->n, g_st.inv);
  QH_HIP(hipGetLastError());
  return QH_OK;
}

extern "C" int qh_set_graph_steps(int steps) {
  QH_REQUIRE(steps >= 1 && steps <= 1024, QH_EINVAL, "graph steps must be in [1,1024]");
  g_graph_steps = steps;
  return QH_OK;
}

extern "C" int qh_last_stage_ms(double *spmv_ms, int64_t *launches) {
  if (spmv_ms) *spmv_ms = g_last_stage_ms;
  if (launches) *launches = g_last_launches;
  return QH_OK;
}

// Advance x in place by nsteps RK4 steps.  Steps are grouped in pairs (x -> alt
// -> x) so a captured graph always starts and ends on the caller's buffer.
// ------------------------------------------------ small systems in one workgroup
// For a generator of at most QH_SMALL_ROWS rows the whole RK4 integration runs
// in one workgroup with the state and the three stage vectors in LDS (5 x 16 B
// per row): every stage is one pass over the CSR rows followed by a barrier, so
// a step costs four barriers instead of four kernel launches.  The state is kept
// in FP64 in LDS for both storage precisions; the matrix values are FP64.
#define QH_SMALL_ROWS 800

// MAXNZ > 0: every thread owns at most one row and keeps its (at most MAXNZ)
// entries in registers; MAXNZ == 0: rows are read from the CSR in memory.
template <bool F32, int MAXNZ>
__global__ void __launch_bounds__(1024) k_rk4_small(int32_t n, const int32_t *__restrict__ rp,
                                                    const int32_t *__restrict__ col, const double2 *__restrict__ val,
                                                    void *xg, float2 *xlo, double h, int64_t nsteps, int64_t tot,
                                                    int renorm_every) {
  extern __shared__ double2 sm[];
  double2 *xa = sm, *xb = sm + n, *k1 = sm + 2 * n, *k2 = sm + 3 * n, *k3 = sm + 4 * n;
  __shared__ double2 scale;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int r = tid; r < n; r += nt) {
    if (F32) {
      const float2 hi = ((const float2 *)xg)[r], lo = xlo[r];
      xa[r] = make_double2((double)hi.x + (double)lo.x, (double)hi.y + (double)lo.y);
    } else {
      xa[r] = ((const double2 *)xg)[r];
    }
  }
  __syncthreads();
  // register copy of this thread's row (MAXNZ > 0: n <= blockDim.x, one row per thread)
  int32_t rc_[MAXNZ > 0 ? MAXNZ : 1];
  double2 rv_[MAXNZ > 0 ? MAXNZ : 1];
  int rlen = 0;
  if (MAXNZ > 0 && tid < n) {
    const int32_t b = rp[tid];
    rlen = rp[tid + 1] - b;
#pragma unroll
    for (int k = 0; k < (MAXNZ > 0 ? MAXNZ : 1); k++) {
      rc_[k] = k < rlen ? col[b + k] : 0;
      rv_[k] = k < rlen ? val[b + k] : make_double2(0.0, 0.0);
    }
  }
  auto rowdot = [&](int r, const double2 *xs, const double2 *kin, double c) {
    double are = 0.0, aim = 0.0;
    if (MAXNZ > 0) {
#pragma unroll
      for (int k = 0; k < (MAXNZ > 0 ? MAXNZ : 1); k++) {
        if (k < rlen) {
          double2 v = xs[rc_[k]];
          if (kin) {
            v.x += c * kin[rc_[k]].x;
            v.y += c * kin[rc_[k]].y;
          }
          const double2 a = rv_[k];
          are = fma(a.x, v.x, fma(-a.y, v.y, are));
          aim = fma(a.x, v.y, fma(a.y, v.x, aim));
        }
      }
      return make_double2(are, aim);
    }
    for (int32_t k = rp[r]; k < rp[r + 1]; k++) {
      const int32_t cc = col[k];
      double2 v = xs[cc];
      if (kin) {
        v.x += c * kin[cc].x;
        v.y += c * kin[cc].y;
      }
      const double2 a = val[k];
      are = fma(a.x, v.x, fma(-a.y, v.y, are));
      aim = fma(a.x, v.y, fma(a.y, v.x, aim));
    }
    return make_double2(are, aim);
  };
  const double s6 = h * (1.0 / 6.0), s3 = h * (1.0 / 3.0);
  for (int64_t st = 0; st < nsteps; st++) {
    for (int r = tid; r < n; r += nt) k1[r] = rowdot(r, xa, nullptr, 0.0);
    __syncthreads();
    for (int r = tid; r < n; r += nt) k2[r] = rowdot(r, xa, k1, 0.5 * h);
    __syncthreads();
    for (int r = tid; r < n; r += nt) k3[r] = rowdot(r, xa, k2, 0.5 * h);
    __syncthreads();
    for (int r = tid; r < n; r += nt) {
      const double2 k4 = rowdot(r, xa, k3, h), a = k1[r], b = k2[r], c = k3[r], xo = xa[r];
      xb[r] = make_double2(xo.x + s6 * a.x + s3 * b.x + s3 * c.x + s6 * k4.x,
                           xo.y + s6 * a.y + s3 * b.y + s3 * c.y + s6 * k4.y);
    }
    __syncthreads();
    double2 *t = xa;
    xa = xb;
    xb = t;
    if (renorm_every > 0 && (st + 1) % renorm_every == 0) {
      if (tid == 0) {
        double tr = 0.0, ti = 0.0;
        for (int64_t i = 0; i < tot; i++) {
          tr += xa[i * (tot + 1)].x;
          ti += xa[i * (tot + 1)].y;
        }
        const double d = tr * tr + ti * ti;
        scale = d > 0.0 ? make_double2(tr / d, -ti / d) : make_double2(1.0, 0.0);
      }
      __syncthreads();
      const double2 sc = scale;
      for (int r = tid; r < n; r += nt) {
        const double2 v = xa[r];
        xa[r] = make_double2(sc.x * v.x - sc.y * v.y, sc.x * v.y + sc.y * v.x);
      }
      __syncthreads();
    }
  }
  for (int r = tid; r < n; r += nt) {
    const double2 v = xa[r];
    if (F32) {
      const float hr = (float)v.x, hm = (float)v.y;
      ((float2 *)xg)[r] = make_float2(hr, hm);
      xlo[r] = make_float2((float)(v.x - (double)hr), (float)(v.y - (double)hm));
    } else {
      ((double2 *)xg)[r] = v;
    }
  }
}

// The adaptive Bogacki-Shampine integration of time_step_sys (TSRK3BS with
// TSAdapt basic, qsystem.c:995-1005) in one workgroup for a generator of at most
// QH_SMALL_BS3_ROWS rows: x, the candidate and K1..K4 in LDS (6 x 16 B per row),
// the step-size controller on the device (thread 0, broadcast through LDS), so a
// trial step costs five barriers and no host round trip.  The same stage
// formulas, error weights and controller as the host-driven loop: K2 = A(x +
// h/2 K1), K3 = A(x + 3h/4 K2), xn = x + h(2/9 K1 + 1/3 K2 + 4/9 K3), K4 = A xn,
// err = RMS of h sum (b_i - b^_i) K_i over atol + rtol max(|x|, |xn|);
// fac = 0.9 err^(-1/3) in [0.1, 10], h' = h fac in [hmin, hmax]; accepted when
// err <= 1 or h <= hmin (1 + 1e-12), with K1 of the next step = K4 (FSAL).
// adaptive == 0: fixed steps of h0 (the last cut at tmax), K1 recomputed.
// out: final time, accepted steps, the next step size.
#define QH_SMALL_BS3_ROWS 640

template <int MAXNZ>
__global__ void __launch_bounds__(QH_SMALL_BS3_ROWS) k_rk3bs_small(int32_t n, const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ col,
                                                      const double2 *__restrict__ val, double2 *xg, double t0,
                                                      double tmax, double h0, int64_t steps_max, double atol,
                                                      double rtol, double hmin, double hmax, int adaptive, double eps,
                                                      double *out) {
  extern __shared__ double2 sm[];
  double2 *xa = sm, *xn = sm + n, *k1 = sm + 2 * n, *k2 = sm + 3 * n, *k3 = sm + 4 * n, *k4 = sm + 5 * n;
  __shared__ double red[QH_SMALL_BS3_ROWS / 64];
  __shared__ double s_hnew;
  __shared__ int s_acc, s_bad;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int r = tid; r < n; r += nt) xa[r] = xg[r];
  int32_t rc_[MAXNZ > 0 ? MAXNZ : 1];
  double2 rv_[MAXNZ > 0 ? MAXNZ : 1];
  int rlen = 0;
  if (MAXNZ > 0 && tid < n) {
    const int32_t b = rp[tid];
    rlen = rp[tid + 1] - b;
#pragma unroll
    for (int k = 0; k < (MAXNZ > 0 ? MAXNZ : 1); k++) {
      rc_[k] = k < rlen ? col[b + k] : 0;
      rv_[k] = k < rlen ? val[b + k] : make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  auto rowdot = [&](int r, const double2 *xs, const double2 *kin, double c) {
    double are = 0.0, aim = 0.0;
    if (MAXNZ > 0) {
#pragma unroll
      for (int k = 0; k < (MAXNZ > 0 ? MAXNZ : 1); k++) {
        if (k < rlen) {
          double2 v = xs[rc_[k]];
          if (kin) {
            v.x += c * kin[rc_[k]].x;
            v.y += c * kin[rc_[k]].y;
          }
          const double2 a = rv_[k];
          are = fma(a.x, v.x, fma(-a.y, v.y, are));
          aim = fma(a.x, v.y, fma(a.y, v.x, aim));
        }
      }
      return make_double2(are, aim);
    }
    for (int32_t k = rp[r]; k < rp[r + 1]; k++) {
      const int32_t cc = col[k];
      double2 v = xs[cc];
      if (kin) {
        v.x += c * kin[cc].x;
        v.y += c * kin[cc].y;
      }
      const double2 a = val[k];
      are = fma(a.x, v.x, fma(-a.y, v.y, are));
      aim = fma(a.x, v.y, fma(a.y, v.x, aim));
    }
    return make_double2(are, aim);
  };
  const double c1 = 2.0 / 9.0 - 7.0 / 24.0, c2 = 1.0 / 3.0 - 1.0 / 4.0, c3 = 4.0 / 9.0 - 1.0 / 3.0, c4 = -1.0 / 8.0;
  double t = t0, h = h0;
  int64_t step = 0;
  bool have_k1 = false;
  if (tid == 0) s_bad = 0;
  __syncthreads();
  // a bound on the trial steps: a rejection shrinks the step by at least 0.9
  // and a step at hmin is always accepted, so fewer than 64 trials per accepted
  // step are possible while the state is finite; the loop ends in every case
  const int64_t trial_cap = 64 * (steps_max < ((int64_t)1 << 40) ? steps_max : ((int64_t)1 << 40)) + 64;
  int64_t trials = 0;
  while (t < tmax - eps && step < steps_max && trials++ < trial_cap) {
    const double hh = (t + h > tmax) ? tmax - t : h;
    if (!have_k1) {
      for (int r = tid; r < n; r += nt) k1[r] = rowdot(r, xa, nullptr, 0.0);
      __syncthreads();
    }
    for (int r = tid; r < n; r += nt) k2[r] = rowdot(r, xa, k1, 0.5 * hh);
    __syncthreads();
    for (int r = tid; r < n; r += nt) k3[r] = rowdot(r, xa, k2, 0.75 * hh);
    __syncthreads();
    for (int r = tid; r < n; r += nt) {
      const double2 xo = xa[r], a = k1[r], bq = k2[r], cq = k3[r];
      xn[r] = make_double2(xo.x + hh * (2.0 / 9.0) * a.x + hh / 3.0 * bq.x + hh * (4.0 / 9.0) * cq.x,
                           xo.y + hh * (2.0 / 9.0) * a.y + hh / 3.0 * bq.y + hh * (4.0 / 9.0) * cq.y);
    }
    __syncthreads();
    double v = 0.0;
    for (int r = tid; r < n; r += nt) {
      const double2 d = rowdot(r, xn, nullptr, 0.0), a = k1[r], bq = k2[r], cq = k3[r], xo = xa[r], xw = xn[r];
      k4[r] = d;
      const double er = hh * (c1 * a.x + c2 * bq.x + c3 * cq.x + c4 * d.x);
      const double ei = hh * (c1 * a.y + c2 * bq.y + c3 * cq.y + c4 * d.y);
      const double ax = sqrt(xo.x * xo.x + xo.y * xo.y), an = sqrt(xw.x * xw.x + xw.y * xw.y);
      const double sc = atol + rtol * (ax > an ? ax : an);
      v += (er * er + ei * ei) / (sc * sc);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      double sum = 0.0;
      for (int w = 0; w < (nt + 63) / 64; w++) sum += red[w];
      const double err = sqrt(sum / (double)n);
      // a non-finite error estimate (a NaN or infinite state) ends the
      // integration: the controller could not leave the step otherwise
      s_bad = !(err <= 1.7976931348623157e308);
      double fac = 0.9 * pow(err > 1e-300 ? err : 1e-300, -1.0 / 3.0);
      fac = fac < 0.1 ? 0.1 : (fac > 10.0 ? 10.0 : fac);
      double hn = hh * fac;
      hn = hn < hmin ? hmin : (hn > hmax ? hmax : hn);
      s_hnew = hn;
      s_acc = (!adaptive || err <= 1.0 || hh <= hmin * (1 + 1e-12)) ? 1 : 0;
    }
    __syncthreads();
    if (s_bad) break;
    const int acc = s_acc;
    if (acc) {
      double2 *tp = xa;
      xa = xn;
      xn = tp;
      t += hh;
      step++;
    }
    if (adaptive) {
      if (acc) {
        double2 *tk = k1;
        k1 = k4;
        k4 = tk;
      }
      have_k1 = true;
      // a step cut to land on tmax (the end, or an event the caller stops at)
      // leaves the step size as it was, as the host loop does after an event
      if (!(acc && hh < h && t >= tmax - eps)) h = s_hnew;
    }
    // the next trial writes K2 (and K1 at fixed steps) only after every thread
    // has read s_acc / s_hnew: the barrier after its first stage orders that
    __syncthreads();
  }
  for (int r = tid; r < n; r += nt) xg[r] = xa[r];
  if (tid == 0) {
    out[0] = t;
    out[1] = (double)step;
    out[2] = h;
    out[3] = s_bad ? 1.0 : 0.0;
  }
}

extern "C" int qh_rk3bs_small(qh_mat *A, qh_vec *x, double t0, double tmax, double h0, int64_t steps_max, double atol,
                              double rtol, double hmin, double hmax, int adaptive, double *t_out, int64_t *steps_out,
                              double *h_out) {
  QH_REQUIRE(A && x && t_out && steps_out && h_out && h0 > 0.0, QH_EINVAL, "qh_rk3bs_small: bad arguments");
  QH_REQUIRE(x->prec == QH_FP64 && x->n == A->rows, QH_EINVAL, "qh_rk3bs_small: an FP64 vector of the matrix size");
  QH_REQUIRE(A->rows >= 1 && A->rows <= QH_SMALL_BS3_ROWS, QH_EINVAL,
             "qh_rk3bs_small: %lld rows (the one-workgroup path takes at most %d)", (long long)A->rows,
             QH_SMALL_BS3_ROWS);
  int rc = qh_mat_max_row(A);
  if (rc) return rc;
  double *dout = nullptr, *hout = nullptr;
  rc = qh_scratch(4 * sizeof(double), (void **)&dout, 4 * sizeof(double), (void **)&hout);
  if (rc) return rc;
  const int32_t n = (int32_t)A->rows;
  int bt = ((n + 63) / 64) * 64;
  const size_t lds = (size_t)6 * n * sizeof(double2);
  const double eps = 1e-12 * (fabs(tmax) + fabs(h0) + 1.0);
  hipStream_t s = qh_stream();
  // rows of at most 8 entries in registers (one row per thread); longer rows read
  // from the CSR, which the caches hold for a generator this small (register
  // copies of 16 or 32 entries would spill)
  const int mz = A->max_row <= 8 ? 8 : 0;
#define QH_LB(Z)                                                                                                \
  hipLaunchKernelGGL((k_rk3bs_small<Z>), dim3(1), dim3(bt), lds, s, n, A->rowptr, A->col, A->val,             \
                     (double2 *)x->d, t0, tmax, h0, steps_max, atol, rtol, hmin, hmax, adaptive, eps, dout)
  if (mz == 8)
    QH_LB(8);
  else
    QH_LB(0);
#undef QH_LB
  QH_HIP(hipGetLastError());
  QH_HIP(hipMemcpyAsync(hout, dout, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  *t_out = hout[0];
  *steps_out = (int64_t)hout[1];
  *h_out = hout[2];
  QH_REQUIRE(hout[3] == 0.0, QH_EINVAL, "qh_rk3bs_small: non-finite error estimate at t = %g (the state is not finite)",
             hout[0]);
  return QH_OK;
}

// longest row of a CSR: per-block maxima over a grid-stride range, combined
// with an atomic max into out[0] (zeroed by the caller)
__global__ void k_max_row(const int32_t *__restrict__ rp, int64_t n, int32_t *out) {
  __shared__ int32_t red[256];
  int32_t m = 0;
  for (int64_t r = blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256)
    m = max(m, rp[r + 1] - rp[r]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, red[0]);
}

int qh_mat_max_row(qh_mat *A) {
  if (A->max_row > 0) return QH_OK;
  int32_t *d = nullptr;
  int rc = qh_alloc_t(&d, 1);
  if (rc) return rc;
  hipStream_t s = qh_stream();
  QH_HIP(hipMemsetAsync(d, 0, sizeof(int32_t), s));
  const unsigned nb = (unsigned)std::min<int64_t>(qh_blocks(A->rows, 256), 1024);
  hipLaunchKernelGGL(k_max_row, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, A->rowptr, A->rows, d);
  QH_HIP(hipGetLastError());
  int32_t h = 0;
  QH_HIP(hipMemcpyAsync(&h, d, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  qh_free_ptr(d);
  A->max_row = h > 0 ? h : 1;
  return QH_OK;
}

static int rk4_small(qh_mat *A, qh_vec *x, double dt, int64_t nsteps, int64_t tot, int renorm_every) {
  int rc0 = qh_mat_max_row(A);
  if (rc0) return rc0;
  const int32_t n = (int32_t)A->rows;
  int bt = ((n + 63) / 64) * 64;
  if (bt > 1024) bt = 1024;
  const size_t lds = (size_t)5 * n * sizeof(double2);
  hipStream_t s = qh_stream();
  const int mz = A->max_row <= 8 ? 8 : A->max_row <= 16 ? 16 : A->max_row <= 32 ? 32 : 0;
#define QH_LS(F, Z)                                                                                             \
  hipLaunchKernelGGL((k_rk4_small<F, Z>), dim3(1), dim3(bt), lds, s, n, A->rowptr, A->col, A->val, x->d, x->lo, \
                     dt, nsteps, tot, renorm_every)
  const bool f32 = x->prec == QH_FP32;
  switch (mz) {
    case 8: if (f32) QH_LS(true, 8); else QH_LS(false, 8); break;
    case 16: if (f32) QH_LS(true, 16); else QH_LS(false, 16); break;
    case 32: if (f32) QH_LS(true, 32); else QH_LS(false, 32); break;
    default: if (f32) QH_LS(true, 0); else QH_LS(false, 0); break;
  }
#undef QH_LS
  QH_HIP(hipGetLastError());
  return QH_OK;
}

// steps per captured graph: pairs of steps (the state ping-pongs back into the
// caller's buffer), exactly renorm_every (rounded up to a pair) when the trace
// is renormalised
static int graph_pair_steps(int renorm_every) {
  int pair_steps = g_graph_steps & ~1;
  if (pair_steps < 2) pair_steps = 2;
  if (renorm_every > 0) pair_steps = renorm_every + (renorm_every & 1);
  return pair_steps;
}

// capture and instantiate the graph of pair_steps steps of x under A, unless the
// one held was captured with the same parameters
static int graph_prepare(const qh_mat *A, qh_vec *x, double dt, int64_t tot, int renorm_every, int G,
                         int pair_steps) {
  hipStream_t s = qh_stream();
  const bool recapture = !g_st.exec || g_st.gx != x->d || g_st.glo != x->lo || g_st.gdt != dt ||
                         g_st.gsteps != pair_steps || g_st.grenorm != renorm_every || g_st.gG != G ||
                         g_st.gtot != tot;
  if (!recapture) return QH_OK;
  if (g_st.exec) (void)hipGraphExecDestroy(g_st.exec);
  if (g_st.graph) (void)hipGraphDestroy(g_st.graph);
  g_st.exec = nullptr;
  g_st.graph = nullptr;
  QH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int crc = QH_OK;
  for (int st = 0; st < pair_steps && !crc; st++) {
    void *xa = (st & 1) ? g_st.xalt : x->d, *xb = (st & 1) ? x->d : g_st.xalt;
    crc = x->prec == QH_FP64 ? rk4_step<double>(s, A, G, xa, xb, x->lo, dt, nullptr)
                             : rk4_step<float>(s, A, G, xa, xb, x->lo, (float)dt, nullptr);
  }
  if (!crc && renorm_every > 0) crc = renorm_launch(s, x, x->d, tot, 0);
  hipGraph_t g = nullptr;
  hipError_t ce = hipStreamEndCapture(s, &g);
  if (crc) {
    if (g) (void)hipGraphDestroy(g);
    return crc;
  }
  QH_HIP(ce);
  g_st.graph = g;
  QH_HIP(hipGraphInstantiate(&g_st.exec, g, nullptr, nullptr, 0));
  g_st.gx = x->d;
  g_st.glo = x->lo;
  g_st.gdt = dt;
  g_st.gsteps = pair_steps;
  g_st.grenorm = renorm_every;
  g_st.gG = G;
  g_st.gtot = tot;
  return QH_OK;
}

static int rk4_impl(qh_mat *A, qh_vec *x, double dt, int64_t nsteps, int64_t tot, int renorm_every, int timed) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && x && x->n == A->rows, QH_EINVAL, "qh_rk4: vector size %lld != matrix rows %lld",
             x ? (long long)x->n : -1LL, A ? (long long)A->rows : -1LL);
  QH_REQUIRE(nsteps >= 0, QH_EINVAL, "qh_rk4: negative step count");
  QH_REQUIRE(renorm_every == 0 || (x->prec == QH_FP32 && tot > 0 && tot * tot == x->n), QH_EINVAL,
             "qh_rk4: trace renormalisation needs an FP32 density matrix of dimension tot");
  // small generators (and no variant forced by qh_mat_set_lanes): one workgroup, LDS-resident
  if (!timed && A->rows <= QH_SMALL_ROWS && !A->lanes_forced) {
    if (nsteps == 0) return QH_OK;
    return rk4_small(A, x, dt, nsteps, tot, renorm_every);
  }
  int G = 0;
  rc = lanes_for(A, x->prec, &G);
  if (rc) return rc;
  rc = stepper_prepare(A, x->prec, x->n);
  if (rc) return rc;
  hipStream_t s = qh_stream();
  if (nsteps == 0) return QH_OK;

  if (timed) {
    // eager launches with events around every stage kernel: per-launch durations
    std::vector<hipEvent_t> ev(5 * nsteps);
    for (auto &e : ev) QH_HIP(hipEventCreate(&e));
    void *bufs[2] = {x->d, g_st.xalt};
    for (int64_t st = 0; st < nsteps; st++) {
      void *xa = bufs[st & 1], *xb = bufs[(st + 1) & 1];
      rc = x->prec == QH_FP64 ? rk4_step<double>(s, A, G, xa, xb, x->lo, dt, &ev[5 * st])
                              : rk4_step<float>(s, A, G, xa, xb, x->lo, (float)dt, &ev[5 * st]);
      if (rc) return rc;
      if (renorm_every > 0 && (st + 1) % renorm_every == 0) renorm_launch(s, x, xb, tot, 0);
    }
    if (nsteps & 1) QH_HIP(hipMemcpyAsync(x->d, g_st.xalt, x->n * (x->prec == QH_FP64 ? 16 : 8),
                                          hipMemcpyDeviceToDevice, s));
    QH_HIP(hipStreamSynchronize(s));
    double tot_ms = 0.0;
    for (int64_t st = 0; st < nsteps; st++)
      for (int q = 0; q < 4; q++) {
        float ms = 0.f;
        QH_HIP(hipEventElapsedTime(&ms, ev[5 * st + q], ev[5 * st + q + 1]));
        tot_ms += ms;
      }
    g_last_stage_ms = tot_ms / (4.0 * (double)nsteps);
    g_last_launches = 4 * nsteps;
    for (auto &e : ev) QH_HIP(hipEventDestroy(e));
    return QH_OK;
  }

  // graph path: capture 2*P steps (P pairs) once, replay
  const int pair_steps = graph_pair_steps(renorm_every);
  const int64_t full = nsteps / pair_steps;
  if (full > 0) {
    rc = graph_prepare(A, x, dt, tot, renorm_every, G, pair_steps);
    if (rc) return rc;
    for (int64_t r = 0; r < full; r++) QH_HIP(hipGraphLaunch(g_st.exec, s));
  }
  // remainder: eager launches
  const int64_t rem = nsteps - full * pair_steps;
  for (int64_t st = 0; st < rem; st++) {
    void *xa = (st & 1) ? g_st.xalt : x->d, *xb = (st & 1) ? x->d : g_st.xalt;
    rc = x->prec == QH_FP64 ? rk4_step<double>(s, A, G, xa, xb, x->lo, dt, nullptr)
                            : rk4_step<float>(s, A, G, xa, xb, x->lo, (float)dt, nullptr);
    if (rc) return rc;
  }
  if (rem & 1)
    QH_HIP(hipMemcpyAsync(x->d, g_st.xalt, x->n * (x->prec == QH_FP64 ? 16 : 8), hipMemcpyDeviceToDevice, s));
  QH_HIP(hipGetLastError());
  return QH_OK;
}

extern "C" int qh_rk4(qh_mat *A, qh_vec *x, double dt, int64_t nsteps, int64_t tot, int renorm_every) {
  return rk4_impl(A, x, dt, nsteps, tot, renorm_every, 0);
}

extern "C" int qh_rk4_timed(qh_mat *A, qh_vec *x, double dt, int64_t nsteps, int64_t tot, int renorm_every) {
  return rk4_impl(A, x, dt, nsteps, tot, renorm_every, 1);
}

// the set-up of a qh_rk4 call without its steps: the SpMV variant chosen (the
// autotuner runs here on a first call), the stage buffers allocated and the
// step graph captured and instantiated, so that a later qh_rk4 call with the
// same arguments only replays it
extern "C" int qh_rk4_prepare(qh_mat *A, qh_vec *x, double dt, int64_t tot, int renorm_every) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && x && x->n == A->rows, QH_EINVAL, "qh_rk4_prepare: vector size != matrix rows");
  QH_REQUIRE(renorm_every == 0 || (x->prec == QH_FP32 && tot > 0 && tot * tot == x->n), QH_EINVAL,
             "qh_rk4_prepare: trace renormalisation needs an FP32 density matrix of dimension tot");
  if (A->rows <= QH_SMALL_ROWS && !A->lanes_forced) return QH_OK;  // one-workgroup path: nothing to capture
  int G = 0;
  rc = lanes_for(A, x->prec, &G);
  if (!rc) rc = stepper_prepare(A, x->prec, x->n);
  if (!rc) rc = graph_prepare(A, x, dt, tot, renorm_every, G, graph_pair_steps(renorm_every));
  if (!rc) QH_HIP(hipStreamSynchronize(qh_stream()));
  return rc;
}

extern "C" int qh_mat_set_lanes(qh_mat *m, int l64, int l32) {
  QH_REQUIRE(m, QH_EINVAL, "null matrix");
  auto ok = [](int g) {
    return g == 0 || g == 1 || g == 2 || g == 4 || g == 8 || g == 16 || g == 32 || g == 64 || QH_IS_KSELL(g);
  };
  QH_REQUIRE(ok(l64) && (ok(l32) || QH_IS_KS2(l32)), QH_EINVAL,
             "lanes per row must be a power of two <= 64, 128..136 (sliced layout), or for FP32 140..144 (row-pair "
             "sliced layout)");
  m->lanes64 = l64;
  m->lanes32 = l32;
  m->lanes_forced = (l64 != 0 || l32 != 0);
  if (g_st.A == m) g_st.gG = -1;  // force a re-capture
  return QH_OK;
}

extern "C" int qh_mat_layout_info(qh_mat *m, int64_t *slices, int64_t *slots, int64_t *stored, int64_t *sreals,
                                  int64_t *mbytes) {
  QH_REQUIRE(m, QH_EINVAL, "null matrix");
  int rc = ks_build(m);
  if (rc) return rc;
  const bool ok = m->ks_state == 1;
  if (slices) *slices = ok ? m->nslices : 0;
  if (slots) *slots = ok ? m->ks_slots : 0;
  if (stored) *stored = ok ? m->ks_nval : 0;
  if (sreals) *sreals = ok ? m->ks_nsval : 0;
  if (mbytes) *mbytes = ok ? (int64_t)(m->ks_slots * sizeof(qh_kmeta) + 12 * m->nslices + 4) : 0;
  return QH_OK;
}

extern "C" int qh_mat_pair_layout_info(qh_mat *m, int64_t *slices, int64_t *slots, int64_t *stored_pairs,
                                       int64_t *stored_real_pairs, int64_t *mbytes, int64_t *stored_real_singles,
                                       int64_t *sep_slots) {
  QH_REQUIRE(m, QH_EINVAL, "null matrix");
  int rc = ks2_build(m);
  if (rc) return rc;
  const bool ok = m->ks2_state == 1;
  if (slices) *slices = ok ? m->ks2_nslices : 0;
  if (slots) *slots = ok ? m->ks2_slots : 0;
  if (stored_pairs) *stored_pairs = ok ? m->ks2_nval : 0;
  if (stored_real_pairs) *stored_real_pairs = ok ? m->ks2_nsval : 0;
  if (stored_real_singles) *stored_real_singles = ok ? m->ks2_nsval1 : 0;
  if (sep_slots) *sep_slots = ok ? m->ks2_nsep : 0;
  if (mbytes)
    *mbytes = ok ? (int64_t)(m->ks2_slots * sizeof(qh_kmeta2) + 28 * m->ks2_nslices + 4 +
                             m->ks2_nsepblk * sizeof(qh_sepblk) + m->ks2_ngh * sizeof(double2))
                 : 0;
  return QH_OK;
}

extern "C" int qh_mat_get_lanes(const qh_mat *m, int *l64, int *l32) {
  QH_REQUIRE(m, QH_EINVAL, "null matrix");
  if (l64) *l64 = m->lanes64;
  if (l32) *l32 = m->lanes32;
  return QH_OK;
}

// mean duration of `reps` plain SpMV launches (y = A x) measured with events
extern "C" int qh_bench_spmv(qh_mat *A, int prec, int reps, double *avg_ms) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && avg_ms && reps > 0 && (prec == QH_FP64 || prec == QH_FP32), QH_EINVAL, "qh_bench_spmv: bad args");
  int G = 0;
  rc = lanes_for(A, prec, &G);
  if (rc) return rc;
  const size_t es = prec == QH_FP64 ? 16 : 8;
  void *x = nullptr, *y = nullptr;
  rc = qh_alloc(&x, (size_t)A->rows * es);
  if (!rc) rc = qh_alloc(&y, (size_t)A->rows * es);
  if (rc) {
    qh_free_ptr(x);
    return rc;
  }
  hipStream_t s = qh_stream();
  QH_HIP(hipMemsetAsync(x, 0, (size_t)A->rows * es, s));
  std::vector<hipEvent_t> ev(reps + 1);
  for (auto &e : ev) QH_HIP(hipEventCreate(&e));
  for (int r = 0; r < reps && !rc; r++) {
    QH_HIP(hipEventRecord(ev[r], s));
    rc = prec == QH_FP64 ? launch_stage_g<double, MODE_SPMV>(G, s, A, A->val, x, nullptr, 0.0, y, nullptr, nullptr,
                                                             nullptr, nullptr, 0.0, nullptr)
                         : launch_stage_g<float, MODE_SPMV>(G, s, A, A->val32, x, nullptr, 0.f, y, nullptr, nullptr,
                                                            nullptr, nullptr, 0.f, nullptr);
  }
  QH_HIP(hipEventRecord(ev[reps], s));
  QH_HIP(hipEventSynchronize(ev[reps]));
  double tot = 0.0;
  for (int r = 0; r < reps; r++) {
    float ms = 0.f;
    QH_HIP(hipEventElapsedTime(&ms, ev[r], ev[r + 1]));
    tot += ms;
  }
  *avg_ms = tot / reps;
  for (auto &e : ev) QH_HIP(hipEventDestroy(e));
  qh_free_ptr(x);
  qh_free_ptr(y);
  return rc;
}

extern "C" int qh_spmv(const qh_mat *A_, const qh_vec *x, qh_vec *y) {
  qh_mat *A = const_cast<qh_mat *>(A_);
  QH_REQUIRE(A && x && y && x->n == A->rows && y->n == A->rows && x->prec == y->prec && x->d != y->d, QH_EINVAL,
             "qh_spmv: bad arguments");
  int G = 0;
  int rc = lanes_for(A, x->prec, &G);
  if (rc) return rc;
  hipStream_t s = qh_stream();
  if (x->prec == QH_FP64)
    rc = launch_stage_g<double, MODE_SPMV>(G, s, A, A->val, x->d, nullptr, 0.0, y->d, nullptr, nullptr, nullptr,
                                           nullptr, 0.0, nullptr);
  else {
    rc = launch_stage_g<float, MODE_SPMV>(G, s, A, A->val32, x->d, nullptr, 0.f, y->d, nullptr, nullptr, nullptr,
                                          nullptr, 0.f, nullptr);
    if (!rc) QH_HIP(hipMemsetAsync(y->lo, 0, (size_t)y->n * 8, s));
  }
  QH_HIP(hipGetLastError());
  return rc;
}

// ------------------------------------------------ generic vector helpers (FP64)
namespace {
__global__ void k_axpy_to(double2 *y, const double2 *x, double cr, double ci, const double2 *k, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 a = x[i], b = k[i];
    y[i] = make_double2(a.x + cr * b.x - ci * b.y, a.y + cr * b.y + ci * b.x);
  }
}
struct LinArgs {
  const double2 *v[8];
  double cr[8], ci[8];
  int n;
};
__global__ void k_lincomb(double2 *y, LinArgs a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int j = 0; j < a.n; j++) {
      double2 b = a.v[j][i];
      re += a.cr[j] * b.x - a.ci[j] * b.y;
      im += a.cr[j] * b.y + a.ci[j] * b.x;
    }
    y[i] = make_double2(re, im);
  }
}
__global__ void __launch_bounds__(256) k_rk3bs_err(const double2 *x, const double2 *xn, const double2 *k1,
                                                   const double2 *k2, const double2 *k3, const double2 *k4,
                                                   double h, double atol, double rtol, int64_t n, double *partial) {
  __shared__ double sh[256];
  double acc = 0.0;
  const double c1 = 2.0 / 9.0 - 7.0 / 24.0, c2 = 1.0 / 3.0 - 1.0 / 4.0, c3 = 4.0 / 9.0 - 1.0 / 3.0, c4 = -1.0 / 8.0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double er = h * (c1 * k1[i].x + c2 * k2[i].x + c3 * k3[i].x + c4 * k4[i].x);
    double ei = h * (c1 * k1[i].y + c2 * k2[i].y + c3 * k3[i].y + c4 * k4[i].y);
    double ax = sqrt(x[i].x * x[i].x + x[i].y * x[i].y), an = sqrt(xn[i].x * xn[i].x + xn[i].y * xn[i].y);
    double sc = atol + rtol * (ax > an ? ax : an);
    acc += (er * er + ei * ei) / (sc * sc);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
unsigned gridn(int64_t n) { return (unsigned)std::min<int64_t>(qh_blocks(n, 256), 8192); }
}  // namespace

namespace {
// sum of n partial sums into out[0] (one block)
__global__ void __launch_bounds__(1024) k_sum_partials(const double *p, int32_t n, double *out) {
  __shared__ double sh[1024];
  double a = 0.0;
  for (int32_t i = threadIdx.x; i < n; i += 1024) a += p[i];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}
}  // namespace

// One Bogacki-Shampine 3(2) trial step (PETSc TSRK3BS tableau, FSAL): k1 = A x
// unless have_k1, k2 = A(x + h/2 k1), k3 = A(x + 3h/4 k2), xn = x + h(2/9 k1 +
// 1/3 k2 + 4/9 k3), k4 = A xn, and the weighted RMS error estimate.  On the
// sliced layout the stages are three fused launches (the stage input Y3, the
// candidate xn and the error sums are formed in the epilogues; y holds Y2, k4
// holds Y3 until the last stage); on the CSR kernels the same sequence runs as
// separate launches.
// The Bogacki-Shampine trial step on a one-row sliced variant G: four stage
// launches (k1 = f(x) unless have_k1; y = x + h/2 k1; k2 = f(y), Y = x + 3h/4 k2;
// k3 = f(Y), xn = x + h (2/9 k1 + 1/3 k2 + 4/9 k3); k4 = f(xn) and the weighted
// error norm), f = A plus, with td (four stage tables, may be null), the merged
// time-dependent entries summed in the same kernels
static int rk3bs_ks(int G, qh_mat *A, const qh_vec *x, qh_vec *xn, qh_vec *k1, qh_vec *k2, qh_vec *k3, qh_vec *k4,
                    qh_vec *y, double h, double atol, double rtol, int have_k1, double *err, const qh_td_ell *td) {
  hipStream_t s = qh_stream();
  // per-slice error partial sums, then their total at dacc[nslices]
  double *dacc = nullptr, *hacc = nullptr;
  int rc = qh_scratch(sizeof(double) * ((size_t)A->nslices + 1), (void **)&dacc, sizeof(double), (void **)&hacc);
  if (rc) return rc;
  if (!have_k1)
    rc = launch_ks<double, MODE_SPMV>(G, s, A, x->d, k1->d, nullptr, 0.0, nullptr, nullptr, nullptr, nullptr,
                                      nullptr, 0.0, nullptr, 0.0, 0.0, nullptr, td ? &td[0] : nullptr);
  if (!rc) rc = qh_axpy_to(y, x, 0.5 * h, 0.0, k1);
  if (!rc)
    rc = launch_ks<double, MODE_STAGE>(G, s, A, y->d, k2->d, k4->d, 0.75 * h, x->d, nullptr, nullptr, nullptr,
                                       nullptr, h, nullptr, 0.0, 0.0, nullptr, td ? &td[1] : nullptr);
  if (!rc)
    rc = launch_ks<double, MODE_BS3>(G, s, A, k4->d, k3->d, nullptr, 0.0, x->d, k1->d, k2->d, nullptr, xn->d, h,
                                     nullptr, 0.0, 0.0, nullptr, td ? &td[2] : nullptr);
  if (rc) return rc;
  rc = launch_ks<double, MODE_BS3E>(G, s, A, xn->d, k4->d, nullptr, 0.0, x->d, k1->d, k2->d, k3->d, xn->d, h,
                                    nullptr, atol, rtol, dacc, td ? &td[3] : nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, dacc, (int32_t)A->nslices, dacc + A->nslices);
  QH_HIP(hipGetLastError());
  QH_HIP(hipMemcpyAsync(hacc, dacc + A->nslices, sizeof(double), hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  *err = sqrt(hacc[0] / (double)A->rows);
  return QH_OK;
}

static int rk3bs_check(qh_mat *A, const qh_vec *x, qh_vec *xn, qh_vec *k1, qh_vec *k2, qh_vec *k3, qh_vec *k4,
                       qh_vec *y, double *err, const char *who) {
  QH_REQUIRE(A && x && xn && k1 && k2 && k3 && k4 && y && err, QH_EINVAL, "%s: null argument", who);
  const qh_vec *vs[7] = {x, xn, k1, k2, k3, k4, y};
  for (const qh_vec *v : vs)
    QH_REQUIRE(v->prec == QH_FP64 && v->n == A->rows, QH_EINVAL, "%s: FP64 vectors of the matrix dimension required",
               who);
  return QH_OK;
}

extern "C" int qh_rk3bs_step(qh_mat *A, const qh_vec *x, qh_vec *xn, qh_vec *k1, qh_vec *k2, qh_vec *k3, qh_vec *k4,
                             qh_vec *y, double h, double atol, double rtol, int have_k1, double *err) {
  int rc = rk3bs_check(A, x, xn, k1, k2, k3, k4, y, err, "qh_rk3bs_step");
  if (rc) return rc;
  int G = 0;
  rc = lanes_for(A, QH_FP64, &G);
  if (rc) return rc;
  if (!QH_IS_KSELL(G)) {
    if (!have_k1) rc = qh_spmv(A, x, k1);
    if (!rc) rc = qh_axpy_to(y, x, 0.5 * h, 0.0, k1);
    if (!rc) rc = qh_spmv(A, y, k2);
    if (!rc) rc = qh_axpy_to(y, x, 0.75 * h, 0.0, k2);
    if (!rc) rc = qh_spmv(A, y, k3);
    if (!rc) {
      const double c[8] = {1.0, 0.0, h * 2.0 / 9.0, 0.0, h / 3.0, 0.0, h * 4.0 / 9.0, 0.0};
      qh_vec *v[4] = {const_cast<qh_vec *>(x), k1, k2, k3};
      rc = qh_lincomb(xn, 4, c, v);
    }
    if (!rc) rc = qh_spmv(A, xn, k4);
    if (!rc) rc = qh_rk3bs_err(x, xn, k1, k2, k3, k4, h, atol, rtol, err);
    return rc;
  }
  return rk3bs_ks(G, A, x, xn, k1, k2, k3, k4, y, h, atol, rtol, have_k1, err, nullptr);
}

int qh_rk3bs_td_sliced(qh_mat *A, const qh_vec *x, qh_vec *xn, qh_vec *k1, qh_vec *k2, qh_vec *k3, qh_vec *k4,
                       qh_vec *y, double h, double atol, double rtol, int have_k1, double *err,
                       const qh_td_ell *td4, int *done) {
  *done = 0;
  int rc = rk3bs_check(A, x, xn, k1, k2, k3, k4, y, err, "qh_rk3bs_td_step");
  if (rc) return rc;
  if (A->rows <= QH_SMALL_ROWS && !A->lanes_forced) return QH_OK;
  int G = 0;
  rc = lanes_for(A, QH_FP64, &G);
  if (rc) return rc;
  if (!(G == 128 || G == 134 || G == 136) || A->ks_state != 1) return QH_OK;
  *done = 1;
  return rk3bs_ks(G, A, x, xn, k1, k2, k3, k4, y, h, atol, rtol, have_k1, err, td4);
}

extern "C" int qh_axpy_to(qh_vec *y, const qh_vec *x, double cr, double ci, const qh_vec *k) {
  QH_REQUIRE(y && x && k && y->n == x->n && k->n == x->n && x->prec == QH_FP64 && y->prec == QH_FP64 &&
                 k->prec == QH_FP64,
             QH_EINVAL, "qh_axpy_to: FP64 vectors of equal size required");
  hipLaunchKernelGGL(k_axpy_to, dim3(gridn(x->n)), dim3(256), 0, qh_stream(), (double2 *)y->d,
                     (const double2 *)x->d, cr, ci, (const double2 *)k->d, x->n);
  QH_HIP(hipGetLastError());
  return QH_OK;
}

extern "C" int qh_lincomb(qh_vec *y, int n, const double *c, qh_vec *const *v) {
  QH_REQUIRE(y && n >= 1 && n <= 8 && y->prec == QH_FP64, QH_EINVAL, "qh_lincomb: 1..8 FP64 vectors");
  LinArgs a;
  a.n = n;
  for (int j = 0; j < n; j++) {
    QH_REQUIRE(v[j] && v[j]->n == y->n && v[j]->prec == QH_FP64, QH_EINVAL, "qh_lincomb: size mismatch");
    a.v[j] = (const double2 *)v[j]->d;
    a.cr[j] = c[2 * j];
    a.ci[j] = c[2 * j + 1];
  }
  hipLaunchKernelGGL(k_lincomb, dim3(gridn(y->n)), dim3(256), 0, qh_stream(), (double2 *)y->d, a, y->n);
  QH_HIP(hipGetLastError());
  return QH_OK;
}

extern "C" int qh_rk3bs_err(const qh_vec *x, const qh_vec *xn, const qh_vec *k1, const qh_vec *k2,
                            const qh_vec *k3, const qh_vec *k4, double h, double atol, double rtol, double *err) {
  QH_REQUIRE(x && xn && k1 && k2 && k3 && k4 && err && x->prec == QH_FP64, QH_EINVAL, "qh_rk3bs_err: bad args");
  const int nb = 512;
  double *partial = nullptr, *hp = nullptr;
  int rc = qh_scratch(sizeof(double) * nb, (void **)&partial, sizeof(double) * nb, (void **)&hp);
  if (rc) return rc;
  hipLaunchKernelGGL(k_rk3bs_err, dim3(nb), dim3(256), 0, qh_stream(), (const double2 *)x->d,
                     (const double2 *)xn->d, (const double2 *)k1->d, (const double2 *)k2->d,
                     (const double2 *)k3->d, (const double2 *)k4->d, h, atol, rtol, x->n, partial);
  QH_HIP(hipMemcpyAsync(hp, partial, sizeof(double) * nb, hipMemcpyDeviceToHost, qh_stream()));
  QH_HIP(hipStreamSynchronize(qh_stream()));
  double sum = 0.0;
  for (int i = 0; i < nb; i++) sum += hp[i];
  *err = sqrt(sum / (double)x->n);
  return QH_OK;
}

// ------------------------------------------------------------------ batches
struct qh_batch {
  qh_mat *A = nullptr;            // block-diagonal matrix of all instances
  qh_vec *x = nullptr;            // concatenated states
  int n = 0;
  std::vector<int64_t> off, rows, tot;
  double *partial = nullptr;      // per-instance trace scratch
  double *inv = nullptr;
  int64_t *doff = nullptr, *dtot = nullptr;
};

namespace {
// per-instance FP64 trace (one block per instance) and rescale
__global__ void __launch_bounds__(256) k_batch_trace(const float2 *hi, const float2 *lo, const int64_t *off,
                                                     const int64_t *tot, double *inv) {
  __shared__ double sh[256];
  const int b = blockIdx.x;
  const int64_t N = tot[b], base = off[b];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < N; i += 256) {
    int64_t loc = base + i * (N + 1);
    acc += (double)hi[loc].x + (double)lo[loc].x;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) inv[b] = sh[0] != 0.0 ? 1.0 / sh[0] : 1.0;
}
__global__ void k_batch_scale(float2 *hi, float2 *lo, const int64_t *off, const double *inv, int n, int64_t total) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    // binary search of the instance owning entry i
    int lo_i = 0, hi_i = n - 1;
    while (lo_i < hi_i) {
      int mid = (lo_i + hi_i + 1) / 2;
      if (off[mid] <= i) lo_i = mid;
      else hi_i = mid - 1;
    }
    const double s = inv[lo_i];
    double xr = ((double)hi[i].x + (double)lo[i].x) * s, xi = ((double)hi[i].y + (double)lo[i].y) * s;
    float hr = (float)xr, hm = (float)xi;
    hi[i] = make_float2(hr, hm);
    lo[i] = make_float2((float)(xr - (double)hr), (float)(xi - (double)hm));
  }
}
}  // namespace

extern "C" int qh_batch_create(int n, qh_mat *const *mats, const int64_t *tot, int prec, qh_batch **out) {
  QH_REQUIRE(n > 0 && mats && tot && out, QH_EINVAL, "qh_batch_create: bad arguments");
  qh_batch *b = new qh_batch();
  b->n = n;
  int64_t o = 0;
  for (int i = 0; i < n; i++) {
    b->off.push_back(o);
    b->rows.push_back(mats[i]->rows);
    b->tot.push_back(tot[i]);
    if (tot[i] * tot[i] != mats[i]->rows) {
      delete b;
      qh_set_error("qh_batch_create: instance %d is not a density-matrix superoperator of dimension %lld", i,
                   (long long)tot[i]);
      return QH_EINVAL;
    }
    o += mats[i]->rows;
  }
  int rc = qh_mat_blockdiag(n, mats, &b->A);
  if (!rc && prec == QH_FP32) rc = qh_mat_make_fp32(b->A);
  if (!rc) rc = qh_vec_create(o, prec, &b->x);
  if (!rc) rc = qh_alloc_t(&b->inv, n);
  if (!rc) rc = qh_alloc_t(&b->doff, n);
  if (!rc) rc = qh_alloc_t(&b->dtot, n);
  if (!rc) rc = qh_h2d(b->doff, b->off.data(), 8 * n);
  if (!rc) rc = qh_h2d(b->dtot, b->tot.data(), 8 * n);
  if (rc) {
    qh_batch_destroy(b);
    return rc;
  }
  *out = b;
  return QH_OK;
}

extern "C" int qh_batch_destroy(qh_batch *b) {
  if (!b) return QH_OK;
  if (g_st.A == b->A) {
    (void)hipStreamSynchronize(qh_stream());
    stepper_release();
  }
  qh_mat_destroy(b->A);
  qh_vec_destroy(b->x);
  qh_free_ptr(b->inv);
  qh_free_ptr(b->doff);
  qh_free_ptr(b->dtot);
  delete b;
  return QH_OK;
}

extern "C" int qh_batch_info(const qh_batch *b, int64_t *rows, int64_t *nnz) { return qh_mat_info(b->A, rows, nnz); }

extern "C" int qh_batch_set_state(qh_batch *b, int i, const double *host) {
  QH_REQUIRE(b && i >= 0 && i < b->n, QH_EINVAL, "qh_batch_set_state: bad instance");
  // set through a temporary FP64 vector and copy into the slice
  qh_vec *t = nullptr;
  int rc = qh_vec_create(b->rows[i], QH_FP64, &t);
  if (rc) return rc;
  rc = qh_vec_set(t, host);
  if (!rc) {
    qh_vec slice;
    slice.n = b->rows[i];
    slice.prec = b->x->prec;
    slice.d = (char *)b->x->d + b->off[i] * (b->x->prec == QH_FP64 ? 16 : 8);
    slice.lo = b->x->lo ? b->x->lo + b->off[i] : nullptr;
    rc = qh_vec_copy(t, &slice);
    slice.d = nullptr;
    slice.lo = nullptr;
  }
  QH_HIP(hipStreamSynchronize(qh_stream()));
  qh_vec_destroy(t);
  return rc;
}

extern "C" int qh_batch_get_state(qh_batch *b, int i, double *host) {
  QH_REQUIRE(b && i >= 0 && i < b->n, QH_EINVAL, "qh_batch_get_state: bad instance");
  qh_vec *t = nullptr;
  int rc = qh_vec_create(b->rows[i], QH_FP64, &t);
  if (rc) return rc;
  qh_vec slice;
  slice.n = b->rows[i];
  slice.prec = b->x->prec;
  slice.d = (char *)b->x->d + b->off[i] * (b->x->prec == QH_FP64 ? 16 : 8);
  slice.lo = b->x->lo ? b->x->lo + b->off[i] : nullptr;
  rc = qh_vec_copy(&slice, t);
  slice.d = nullptr;
  slice.lo = nullptr;
  if (!rc) rc = qh_vec_get(t, host);
  qh_vec_destroy(t);
  return rc;
}

// the batch state slice of instance i as a qh_vec view (owns nothing)
static qh_vec batch_slice(qh_batch *b, int i) {
  qh_vec slice;
  slice.n = b->rows[i];
  slice.prec = b->x->prec;
  slice.d = (char *)b->x->d + b->off[i] * (b->x->prec == QH_FP64 ? 16 : 8);
  slice.lo = b->x->lo ? b->x->lo + b->off[i] : nullptr;
  return slice;
}

extern "C" int qh_batch_set_state_vec(qh_batch *b, int i, const qh_vec *src) {
  QH_REQUIRE(b && src && i >= 0 && i < b->n, QH_EINVAL, "qh_batch_set_state_vec: bad instance");
  QH_REQUIRE(src->n == b->rows[i], QH_EINVAL, "qh_batch_set_state_vec: state size %lld, instance %d has %lld rows",
             (long long)src->n, i, (long long)b->rows[i]);
  qh_vec slice = batch_slice(b, i);
  const int rc = qh_vec_copy(src, &slice);
  slice.d = nullptr;
  slice.lo = nullptr;
  return rc;
}

extern "C" int qh_batch_get_state_vec(qh_batch *b, int i, qh_vec *dst) {
  QH_REQUIRE(b && dst && i >= 0 && i < b->n, QH_EINVAL, "qh_batch_get_state_vec: bad instance");
  QH_REQUIRE(dst->n == b->rows[i], QH_EINVAL, "qh_batch_get_state_vec: state size %lld, instance %d has %lld rows",
             (long long)dst->n, i, (long long)b->rows[i]);
  qh_vec slice = batch_slice(b, i);
  const int rc = qh_vec_copy(&slice, dst);
  slice.d = nullptr;
  slice.lo = nullptr;
  return rc;
}

extern "C" uint64_t qh_mat_uid(const qh_mat *m) { return m ? m->uid : 0; }

extern "C" int qh_batch_rk4(qh_batch *b, double dt, int64_t nsteps, int renorm_every) {
  QH_REQUIRE(b, QH_EINVAL, "null batch");
  QH_REQUIRE(renorm_every == 0 || b->x->prec == QH_FP32, QH_EINVAL, "batch renormalisation needs FP32");
  if (renorm_every == 0) return rk4_impl(b->A, b->x, dt, nsteps, 0, 0, 0);
  int64_t done = 0;
  while (done < nsteps) {
    int64_t chunk = std::min<int64_t>(renorm_every, nsteps - done);
    int rc = rk4_impl(b->A, b->x, dt, chunk, 0, 0, 0);
    if (rc) return rc;
    if (chunk == renorm_every) {
      hipLaunchKernelGGL(k_batch_trace, dim3(b->n), dim3(256), 0, qh_stream(), (const float2 *)b->x->d, b->x->lo,
                         b->doff, b->dtot, b->inv);
      hipLaunchKernelGGL(k_batch_scale, dim3(gridn(b->x->n)), dim3(256), 0, qh_stream(), (float2 *)b->x->d,
                         b->x->lo, b->doff, b->inv, b->n, b->x->n);
      QH_HIP(hipGetLastError());
    }
    done += chunk;
  }
  return QH_OK;
}

extern "C" qh_vec *qh_batch_state_vec(qh_batch *b) { return b ? b->x : nullptr; }
extern "C" qh_mat *qh_batch_matrix(qh_batch *b) { return b ? b->A : nullptr; }
extern "C" int64_t qh_batch_offset(qh_batch *b, int i) { return (b && i >= 0 && i < b->n) ? b->off[i] : -1; }

// ------------------------------------------------------------ time-dependent stages
int qh_td_sliced_prepare(qh_mat *A, int prec, int *G) {
  *G = 0;
  if (A->rows <= QH_SMALL_ROWS && !A->lanes_forced) return QH_OK;
  int g = 0;
  const int rc = lanes_for(A, prec, &g);
  if (rc) return rc;
  if (QH_IS_KSELL(g) || QH_IS_KS2(g)) *G = g;
  return QH_OK;
}

int qh_td_sliced_stage(int G, const qh_mat *A, int prec, int mode, const void *xin, void *kout, void *ynext,
                       double cn, const void *x, const void *k1, const void *k2, const void *k3, void *xout,
                       double h, float2 *xlo, const qh_td_ell *td, int *fused) {
  hipStream_t s = qh_stream();
  *fused = 0;
  int rc;
  if (td && prec == QH_FP32 && (G == 140 || G == 142 || G == 143) && A->ks2_state == 1) {
    *fused = 1;
    rc = ks2_launch(G, mode, s, A, xin, kout, ynext, (float)cn, x, k1, k2, k3, xout, (float)h, xlo, td);
  } else if (td && (G == 128 || G == 134 || G == 136) && A->ks_state == 1) {
    *fused = 1;
#define QH_TF(R, M) \
  launch_ks<R, M>(G, s, A, xin, kout, ynext, (R)cn, x, k1, k2, k3, xout, (R)h, xlo, 0.0, 0.0, nullptr, td)
    if (prec == QH_FP64)
      rc = mode == MODE_SPMV ? QH_TF(double, MODE_SPMV) : mode == MODE_STAGE ? QH_TF(double, MODE_STAGE)
                                                                            : QH_TF(double, MODE_FINAL);
    else
      rc = mode == MODE_SPMV ? QH_TF(float, MODE_SPMV) : mode == MODE_STAGE ? QH_TF(float, MODE_STAGE)
                                                                           : QH_TF(float, MODE_FINAL);
#undef QH_TF
  } else {
#define QH_TS(R, M) \
  launch_sliced<R, M>(G, s, A, xin, kout, ynext, (R)cn, x, k1, k2, k3, xout, (R)h, xlo)
    if (prec == QH_FP64)
      rc = mode == MODE_SPMV ? QH_TS(double, MODE_SPMV) : mode == MODE_STAGE ? QH_TS(double, MODE_STAGE)
                                                                            : QH_TS(double, MODE_FINAL);
    else
      rc = mode == MODE_SPMV ? QH_TS(float, MODE_SPMV) : mode == MODE_STAGE ? QH_TS(float, MODE_STAGE)
                                                                           : QH_TS(float, MODE_FINAL);
#undef QH_TS
  }
  if (rc) return rc;
  QH_HIP(hipGetLastError());
  return QH_OK;
}
