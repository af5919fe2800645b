// qh_ks2.hip -- the FP32 stage kernel on the row-pair sliced layout.
//
// A wavefront owns a slice of 128 consecutive rows, lane l the two rows 2l and
// 2l+1.  Every x gather is one 16-byte load per lane of a complex64 pair
// x[b], x[b+1] that serves both rows of the lane (row 2l takes half h0, row
// 2l+1 half h1), and every value and epilogue access is 8 or 16 bytes per lane.
// The texture data path of a CU returns one 4-lane quad per clock, so 16-byte
// lane accesses move twice the bytes per clock of the 8-byte accesses of the
// one-row-per-lane layout (qh_spmv.hip), which that layout is bound by.
//
// Pairing: for each entry of row 2l at column c the builder looks for an entry
// of row 2l+1 at c+1 (straight, b = c), c-1 (crossed, b = c-1) or c (same);
// unmatched entries load alone (one row of the lane inactive, zero value), with
// the in-order halves where the vector allows it.  A slot collects the lanes
// whose loads share (|b - 2l|, h0, h1, rank);
// the rest of the machinery (sign masks, uniform / scalar / complex slot groups,
// compacted values, exact tails, two passes in flight) is that of the
// one-row-per-lane layout, with one more group: scalar slots whose two rows
// hold the same value in every lane store one real per lane (4 bytes, not 8).  Products are summed in FP32 within a pass and the
// pass sums in FP64; the state is hi + lo as in the other FP32 kernels.
#include <algorithm>
#include <cmath>
#include <vector>

#include "qh_common.h"
#include "qh_ks.h"

namespace {

const int KS2_MAXROW = 32;  // longest row the pair builder takes (the CSR kernels serve longer rows)
// rows of the LDS window of the windowed kernel: 16 slices, one workgroup of 16
// wavefronts (for a density matrix of dimension 1024, two columns of rho: every
// left-multiplication entry of their rows, and the right-multiplication entries
// that flip the lowest column bit)
const int KS2_WIN = 2048;

// ------------------------------------------------------------------ build
// Per-slice counts (FILL false): cnt[0*ns + s] slots, [1] complex pairs, [2]
// uniform slots, [3] scalar slots with a value per row, [4] their value pairs,
// [5] uniform slots whose rows take the halves in order (row 2l half 0, row
// 2l+1 half 1: no select in the kernel), [6] those of them whose loads all stay
// inside the slice's KS2_WIN-row window, [7] scalar slots whose two rows hold
// the same value in every lane, [8] their values (one per lane), [9] slots of a
// separable diagonal (stored nothing: the kernel forms g_i + h_j).  FILL writes
// the metadata (uniform in-order window slots at moff, the other uniform
// in-order slots, the other uniform slots, then the scalar, the equal-scalar,
// the complex and the separable-diagonal groups) and the values (complex pairs
// at voff, scalar pairs at svoff, equal scalars at s1off).  slice_blk[s]: the
// separable block that holds the whole slice, or -1.
template <bool FILL>
__global__ void __launch_bounds__(64) k_ks2_build(int32_t n, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ col, const double2 *__restrict__ val,
                                                  int32_t *cnt, int32_t ns, const int32_t *moff, const int32_t *voff,
                                                  const int32_t *svoff, const int32_t *s1off, qh_kmeta2 *meta,
                                                  float4 *kval, float2 *sval, float *sval1,
                                                  const int32_t *__restrict__ slice_blk, int *overflow) {
  // per lane up to 2 * KS2_MAXROW pair loads: key << 12 | e0 << 6 | e1 (63 = no entry)
  __shared__ uint64_t keys[64][2 * KS2_MAXROW];
  const int s = blockIdx.x, lane = threadIdx.x;
  const int64_t r0 = (int64_t)s * 128 + 2 * lane, r1 = r0 + 1;
  int len0 = 0, len1 = 0, beg0 = 0, beg1 = 0;
  if (r0 < n) {
    beg0 = rp[r0];
    len0 = rp[r0 + 1] - beg0;
  }
  if (r1 < n) {
    beg1 = rp[r1];
    len1 = rp[r1 + 1] - beg1;
  }
  if (len0 > KS2_MAXROW || len1 > KS2_MAXROW) {
    overflow[0] = 1;
    len0 = len1 = 0;
  }
  uint64_t *kk = keys[lane];
  int np = 0;
  uint32_t used1 = 0;
  auto push = [&](int64_t b, int h0, int h1, int e0, int e1) {
    const int64_t o = b - r0;
    const uint64_t ad = (uint64_t)(o < 0 ? -o : o);
    // which rows are active is not part of the key: a row without the entry gets
    // a zero value, so single-row loads share the slots of the pairs beside them
    const uint64_t pat = (uint64_t)((h0 << 2) | (h1 << 3));
    // negative offsets first so that the rank below separates b = r0 - ad from r0 + ad
    const uint64_t key = (((ad << 4) | pat) << 1) | (o > 0 ? 1u : 0u);
    const uint64_t v = (key << 12) | ((uint64_t)e0 << 6) | (uint64_t)e1;
    int p = np;
    while (p > 0 && kk[p - 1] > v) {
      kk[p] = kk[p - 1];
      p--;
    }
    kk[p] = v;
    np++;
  };
  for (int e0 = 0; e0 < len0; e0++) {
    const int64_t c = col[beg0 + e0];
    int f = -1, how = 0;
    for (int pass = 0; pass < 3 && f < 0; pass++) {
      const int64_t want = pass == 0 ? c + 1 : pass == 1 ? c - 1 : c;
      for (int e1 = 0; e1 < len1; e1++)
        if (!((used1 >> e1) & 1u) && col[beg1 + e1] == want) {
          f = e1;
          how = pass;
          break;
        }
    }
    if (f >= 0) used1 |= 1u << f;
    if (f >= 0 && how == 0) {
      push(c, 0, 1, e0, f);  // straight: x[c], x[c+1]
    } else if (f >= 0 && how == 1) {
      push(c - 1, 1, 0, e0, f);  // crossed: row 2l+1 reads x[c-1]
    } else if (f >= 0) {
      // same column for both rows: keep the 16-byte load inside the vector
      const bool last = c + 1 >= n;
      push(last ? c - 1 : c, last ? 1 : 0, last ? 1 : 0, e0, f);
    } else {
      // row 2l alone: in-order halves (row 2l on half 0) unless at the end of the vector
      const bool last = c + 1 >= n;
      push(last ? c - 1 : c, last ? 1 : 0, last ? 0 : 1, e0, 63);
    }
  }
  for (int e1 = 0; e1 < len1; e1++) {
    if ((used1 >> e1) & 1u) continue;
    // row 2l+1 alone: in-order halves (row 2l+1 on half 1) unless at column 0
    const int64_t c = col[beg1 + e1];
    const bool first = c == 0;
    push(first ? 0 : c - 1, first ? 1 : 0, first ? 0 : 1, 63, e1);
  }
  // rank: the second load of a lane with the same (ad, pattern) gets rank 1
  uint64_t prev = ~0ull;
  for (int e = 0; e < np; e++) {
    const uint64_t adp = kk[e] >> 13;
    const uint64_t rank = (adp == prev) ? 1u : 0u;
    kk[e] = (((adp << 1) | rank) << 12) | (kk[e] & 4095u);
    prev = adp;
  }
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int p = 0, slot = 0, uslot = 0, sslot = 0, s1slot = 0, cslot = 0, xslot = 0, ustr = 0, uwin = 0;
  const int32_t sblk = slice_blk ? slice_blk[s] : -1;
  const int64_t w0 = ((int64_t)s * 128 / KS2_WIN) * KS2_WIN;  // first row of the slice's window
  int32_t base = FILL ? voff[s] : 0;
  int32_t sbase = FILL ? svoff[s] : 0;
  int32_t s1base = FILL ? s1off[s] : 0;
  const int32_t mo = FILL ? moff[s] : 0;
  const int32_t ms = FILL ? mo + cnt[2 * ns + s] : 0;
  const int32_t ms1 = FILL ? ms + cnt[3 * ns + s] : 0;
  const int32_t mc = FILL ? ms1 + cnt[7 * ns + s] : 0;
  const int32_t mx = FILL ? mo + cnt[s] - cnt[9 * ns + s] : 0;  // separable-diagonal slots
  const int32_t mu2 = FILL ? mo + cnt[5 * ns + s] : 0;  // uniform slots that need the half selects
  const int32_t mw = FILL ? mo + cnt[6 * ns + s] : 0;   // in-order uniform slots outside the window
  for (;;) {
    const uint64_t cur = p < np ? (kk[p] >> 12) : ~0ull;
    const uint64_t m = wave_min_u64(cur);
    if (m == ~0ull) break;
    const bool act = (cur == m);
    const uint64_t mask = __ballot(act);
    const int e0 = act ? (int)((kk[p] >> 6) & 63u) : 63, e1 = act ? (int)(kk[p] & 63u) : 63;
    const double2 z = make_double2(0.0, 0.0);
    const double2 v0 = (act && e0 != 63) ? val[beg0 + e0] : z;
    const double2 v1 = (act && e1 != 63) ? val[beg1 + e1] : z;
    const int first = __ffsll((unsigned long long)mask) - 1;
    const double f0r = __shfl(v0.x, first, 64), f0i = __shfl(v0.y, first, 64);
    const double f1r = __shfl(v1.x, first, 64), f1i = __shfl(v1.y, first, 64);
    auto bits = [](double a) { return __double_as_longlong(a); };
    const bool same = !act || (bits(v0.x) == bits(f0r) && bits(v0.y) == bits(f0i) && bits(v1.x) == bits(f1r) &&
                               bits(v1.y) == bits(f1i));
    const bool uni = __ballot(same) == ~0ull;
    const bool isre = !uni && __ballot(act && (v0.y != 0.0 || v1.y != 0.0)) == 0ull;
    const bool isim = !uni && !isre && __ballot(act && (v0.x != 0.0 || v1.x != 0.0)) == 0ull;
    // separable diagonal: the slot (|d| = 0, in-order halves, rank 0) holds the
    // diagonal entries of both rows of every active lane, in a separable block
    const bool dgl = act && e0 != 63 && e1 != 63 && col[beg0 + e0] == r0 && col[beg1 + e1] == r1;
    const bool sep = sblk >= 0 && !uni && (m >> 5) == 0 && ((m >> 1) & 15u) == 8u && (m & 1u) == 0 &&
                     __ballot(act && !dgl) == 0ull;
    const bool sca = (isre || isim) && !sep;
    // equal scalars: both rows of every active lane hold the same FP32 value
    auto part = [&](double2 v) { return (float)(isre ? v.x : v.y); };
    const bool eq1 = sca && __ballot(act && part(v0) != part(v1)) == 0ull;
    // in-order halves: an active row 2l takes half 0, an active row 2l+1 half 1
    const uint32_t pat_ = (uint32_t)((m >> 1) & 15u);
    const bool inorder = !(pat_ & 4u) && (pat_ & 8u);  // h0 = 0, h1 = 1
    // the pair starts at b = (column of row 2l) - h0, or (column of row 2l+1) - h1
    const int64_t bcol = !act ? r0
                         : e0 != 63 ? (int64_t)col[beg0 + e0] - ((pat_ >> 2) & 1u)
                                    : (int64_t)col[beg1 + e1] - ((pat_ >> 3) & 1u);
    const bool inwin = __ballot(act && (bcol < w0 || bcol + 1 >= w0 + KS2_WIN)) == 0ull;
    const bool win = uni && inorder && inwin;
    if (FILL) {
      // slot key m = (ad << 4 | pattern) << 1 | rank
      const uint32_t pat = (uint32_t)((m >> 1) & 15u);
      const uint32_t h0 = (pat >> 2) & 1u, h1 = (pat >> 3) & 1u;
      const uint64_t ad = m >> 5;
      const uint64_t sign = __ballot(act && bcol < r0);
      const uint32_t rank = (uint32_t)__popcll(mask & lt);
      if (lane == 0) {
        qh_kmeta2 md;
        md.mask = mask;
        md.sign = sign;
        md.ad8 = (uint32_t)(ad * 8u) | h0 | (h1 << 1);
        md.base = uni ? 0u : sep ? (uint32_t)sblk : (uint32_t)(eq1 ? s1base : sca ? sbase : base);
        if (uni) {
          md.v0re = (float)f0r;
          md.v0im = (float)f0i;
          md.v1re = (float)f1r;
          md.v1im = (float)f1i;
        } else {
          md.v0re = md.v1re = sca ? (isre ? 1.f : 0.f) : 0.f;
          md.v0im = md.v1im = sca ? (isre ? 0.f : 1.f) : 0.f;
        }
        meta[uni ? (win ? mo + uwin : inorder ? mw + (ustr - uwin) : mu2 + (uslot - ustr))
                 : sep ? mx + xslot : eq1 ? ms1 + s1slot : sca ? ms + sslot : mc + cslot] = md;
      }
      if (act && eq1)
        sval1[s1base + rank] = part(v0);
      else if (act && sca)
        sval[sbase + rank] = make_float2(part(v0), part(v1));
      if (act && !uni && !sca && !sep)
        kval[base + rank] = make_float4((float)v0.x, (float)v0.y, (float)v1.x, (float)v1.y);
    }
    if (sep) {
      // nothing stored
    } else if (eq1)
      s1base += __popcll(mask);
    else if (sca)
      sbase += __popcll(mask);
    else if (!uni)
      base += __popcll(mask);
    if (act) p++;
    slot++;
    if (uni) {
      uslot++;
      if (inorder) ustr++;
      if (win) uwin++;
    } else if (sep) {
      xslot++;
    } else if (eq1) {
      s1slot++;
    } else if (sca) {
      sslot++;
    } else {
      cslot++;
    }
  }
  if (!FILL && lane == 0) {
    cnt[s] = slot;
    cnt[ns + s] = base;
    cnt[2 * ns + s] = uslot;
    cnt[3 * ns + s] = sslot;
    cnt[4 * ns + s] = sbase;
    cnt[5 * ns + s] = ustr;
    cnt[6 * ns + s] = uwin;
    cnt[7 * ns + s] = s1slot;
    cnt[8 * ns + s] = s1base;
    cnt[9 * ns + s] = xslot;
  }
}

// ------------------------------------------------------------------ stage
// row sums of rows 2l and 2l+1: FP32 partial sums of the last passes, flushed
// into the FP64 sums every second pass
struct KsAcc2 {
  double r0, i0, r1, i1;
  float p0r, p0i, p1r, p1i;
};

template <int KIND, int W>
struct Ks2Buf {
  float4 xv[W], vv[W];
  float2 sv[W];
  float s1[W];
  float u0r[W], u0i[W], u1r[W], u1i[W];
  uint32_t h[W];
};

// INORDER: every slot of the group has row 2l on half 0 and row 2l+1 on half 1
// (known at compile time, no selects)
template <int KIND, bool INORDER = false, int AUX = 0>
struct Ks2Ops {
  typedef KsAcc2 Acc;
  template <int W>
  using Buf = Ks2Buf<KIND, W>;
  const qh_kmeta2 *mp;
  uint32_t r8;  // byte offset of x[2l] within the stage input
  __amdgpu_buffer_rsrc_t xr, vr;
  template <int W>
  __device__ __forceinline__ void issue(Buf<W> &b) {
    qh_kmeta2 m[W];
#pragma unroll
    for (int j = 0; j < W; j++) {
      m[j] = mp[j];
      b.h[j] = m[j].ad8 & 3u;
      b.u0r[j] = m[j].v0re;
      b.u0i[j] = m[j].v0im;
      b.u1r[j] = m[j].v1re;
      b.u1i[j] = m[j].v1im;
    }
#pragma unroll
    for (int j = 0; j < W; j++) {
      const bool on = __builtin_amdgcn_inverse_ballot_w64(m[j].mask);
      const bool neg = __builtin_amdgcn_inverse_ballot_w64(m[j].sign);
      const uint32_t ad8 = m[j].ad8 & ~7u;
      // r8 + ad8 + (neg ? -2 ad8 : 0): one select against zero and a three-way add
      const uint32_t back = neg ? (0u - 2u * ad8) : 0u;
      uint32_t off = r8 + ad8 + back;
      off = on ? off : QH_KS_OOB;
      b.xv[j] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(xr, off, 0, 0));
      if (KIND != KS_UNI) {
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m[j].mask >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m[j].mask, 0u));
        if (KIND == KS_SC1)
          b.s1[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(vr, (m[j].base + rank) * 4u, 0, AUX));
        else if (KIND == KS_SCA)
          b.sv[j] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(vr, (m[j].base + rank) * 8u, 0, AUX));
        else
          b.vv[j] =
              __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(vr, (m[j].base + rank) * 16u, 0, AUX));
      }
    }
    mp += W;
  }
  template <int W>
  __device__ __forceinline__ void fma(const Buf<W> &b, Acc &acc) {
    float p0r = acc.p0r, p0i = acc.p0i, p1r = acc.p1r, p1i = acc.p1i;
#pragma unroll
    for (int j = 0; j < W; j++) {
      const float4 xv = b.xv[j];
      float x0r, x0i, x1r, x1i;
      if (INORDER) {
        x0r = xv.x;
        x0i = xv.y;
        x1r = xv.z;
        x1i = xv.w;
      } else {
        const bool h0 = b.h[j] & 1u, h1 = b.h[j] & 2u;
        x0r = h0 ? xv.z : xv.x;
        x0i = h0 ? xv.w : xv.y;
        x1r = h1 ? xv.z : xv.x;
        x1i = h1 ? xv.w : xv.y;
      }
      float a0r, a0i, a1r, a1i;
      if (KIND == KS_CPX) {
        a0r = b.vv[j].x;
        a0i = b.vv[j].y;
        a1r = b.vv[j].z;
        a1i = b.vv[j].w;
      } else if (KIND == KS_SC1) {
        a0r = a1r = b.s1[j] * b.u0r[j];
        a0i = a1i = b.s1[j] * b.u0i[j];
      } else if (KIND == KS_SCA) {
        a0r = b.sv[j].x * b.u0r[j];
        a0i = b.sv[j].x * b.u0i[j];
        a1r = b.sv[j].y * b.u0r[j];
        a1i = b.sv[j].y * b.u0i[j];
      } else {
        a0r = b.u0r[j];
        a0i = b.u0i[j];
        a1r = b.u1r[j];
        a1i = b.u1i[j];
      }
      p0r = fmaf(a0r, x0r, fmaf(-a0i, x0i, p0r));
      p0i = fmaf(a0r, x0i, fmaf(a0i, x0r, p0i));
      p1r = fmaf(a1r, x1r, fmaf(-a1i, x1i, p1r));
      p1i = fmaf(a1r, x1i, fmaf(a1i, x1r, p1i));
    }
    acc.p0r = p0r;
    acc.p0i = p0i;
    acc.p1r = p1r;
    acc.p1i = p1i;
  }
  __device__ __forceinline__ void flush(Acc &acc) {
    acc.r0 += (double)acc.p0r;
    acc.i0 += (double)acc.p0i;
    acc.r1 += (double)acc.p1r;
    acc.i1 += (double)acc.p1i;
    acc.p0r = acc.p0i = acc.p1r = acc.p1i = 0.f;
  }
};

// In-order uniform slots whose loads all stay inside the workgroup's window of
// KS2_WIN rows: the pair is read from the LDS copy of the window (xs, KS2_WIN + 2
// complex64, the last two zero for inactive lanes) instead of global memory.
struct Ks2OpsWin {
  typedef KsAcc2 Acc;
  template <int W>
  using Buf = Ks2Buf<KS_UNI, W>;
  const qh_kmeta2 *mp;
  uint32_t w8;  // byte offset of row 2l within the window
  const float2 *xs;
  template <int W>
  __device__ __forceinline__ void issue(Buf<W> &b) {
    qh_kmeta2 m[W];
#pragma unroll
    for (int j = 0; j < W; j++) {
      m[j] = mp[j];
      b.h[j] = 0u;
      b.u0r[j] = m[j].v0re;
      b.u0i[j] = m[j].v0im;
      b.u1r[j] = m[j].v1re;
      b.u1i[j] = m[j].v1im;
    }
#pragma unroll
    for (int j = 0; j < W; j++) {
      const bool on = __builtin_amdgcn_inverse_ballot_w64(m[j].mask);
      const bool neg = __builtin_amdgcn_inverse_ballot_w64(m[j].sign);
      const uint32_t ad8 = m[j].ad8 & ~7u;
      const uint32_t back = neg ? (0u - 2u * ad8) : 0u;
      uint32_t off = w8 + ad8 + back;
      off = on ? off : (uint32_t)(KS2_WIN * 8);
      const float2 *q = (const float2 *)((const char *)xs + off);
      const float2 lo = q[0], hi = q[1];
      b.xv[j] = make_float4(lo.x, lo.y, hi.x, hi.y);
    }
    mp += W;
  }
  template <int W>
  __device__ __forceinline__ void fma(const Buf<W> &b, Acc &acc) {
    Ks2Ops<KS_UNI, true>{}.fma(b, acc);
  }
  __device__ __forceinline__ void flush(Acc &acc) { Ks2Ops<KS_UNI, true>{}.flush(acc); }
};

// Separable-diagonal slots (at most one per slice): the lane's own pair x[2l],
// x[2l+1] times the diagonal entries g_i + h_j and g_(i+1) + h_j, formed in FP64
// from the block's g and h (t = 2l - off, j = t / N, i = t - j N; i is even, so
// i + 1 stays in the column) and rounded to FP32 as the stored diagonal would be.
struct Ks2OpsSep {
  typedef KsAcc2 Acc;
  template <int W>
  struct Buf {
    float4 xv[W];
    double2 g0[W], g1[W], h[W];
  };
  const qh_kmeta2 *mp;
  uint32_t r8;  // byte offset of x[2l] within the stage input
  int64_t r;    // row 2l
  __amdgpu_buffer_rsrc_t xr, ghr;
  const qh_sepblk *bp;
  template <int W>
  __device__ __forceinline__ void issue(Buf<W> &b) {
#pragma unroll
    for (int j = 0; j < W; j++) {
      const qh_kmeta2 m = mp[j];
      const bool on = __builtin_amdgcn_inverse_ballot_w64(m.mask);
      b.xv[j] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(xr, on ? r8 : QH_KS_OOB, 0, 0));
      const qh_sepblk q = bp[m.base];
      const uint32_t t = (uint32_t)(r - q.off);
      const uint32_t jj = (uint32_t)(((uint64_t)__umulhi(t, q.m) + t) >> q.sh);
      const uint32_t ii = t - jj * q.n;
      const uint32_t go = on ? (q.gbase + ii) * 16u : QH_KS_OOB, ho = on ? (q.hbase + jj) * 16u : QH_KS_OOB;
      b.g0[j] = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(ghr, go, 0, 0));
      b.g1[j] = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(ghr, on ? go + 16u : go, 0, 0));
      b.h[j] = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(ghr, ho, 0, 0));
    }
    mp += W;
  }
  template <int W>
  __device__ __forceinline__ void fma(const Buf<W> &b, Acc &acc) {
    float p0r = acc.p0r, p0i = acc.p0i, p1r = acc.p1r, p1i = acc.p1i;
#pragma unroll
    for (int j = 0; j < W; j++) {
      const float a0r = (float)(b.g0[j].x + b.h[j].x), a0i = (float)(b.g0[j].y + b.h[j].y);
      const float a1r = (float)(b.g1[j].x + b.h[j].x), a1i = (float)(b.g1[j].y + b.h[j].y);
      const float4 xv = b.xv[j];
      p0r = fmaf(a0r, xv.x, fmaf(-a0i, xv.y, p0r));
      p0i = fmaf(a0r, xv.y, fmaf(a0i, xv.x, p0i));
      p1r = fmaf(a1r, xv.z, fmaf(-a1i, xv.w, p1r));
      p1i = fmaf(a1r, xv.w, fmaf(a1i, xv.z, p1i));
    }
    acc.p0r = p0r;
    acc.p0i = p0i;
    acc.p1r = p1r;
    acc.p1i = p1i;
  }
  __device__ __forceinline__ void flush(Acc &acc) { Ks2Ops<KS_UNI, true>{}.flush(acc); }
};

// the two complex64 of rows (r, r+1) at byte offset o of a vector; the second is
// zero (and not loaded / stored) when r is the last row.  AUX: the cache policy
// bits of the buffer instruction (0, or 2 = non-temporal: streamed data that is
// not read again soon, so it does not displace the gathered stage input from
// the caches)
template <int AUX = 0>
__device__ inline float4 ld_pair(__amdgpu_buffer_rsrc_t rs, uint32_t o, bool two) {
  if (two) return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, o, 0, AUX));
  const float2 a = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, o, 0, AUX));
  return make_float4(a.x, a.y, 0.f, 0.f);
}
template <int AUX = 0>
__device__ inline void st_pair(__amdgpu_buffer_rsrc_t rs, uint32_t o, bool two, float4 v) {
  if (two)
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v), rs, o,
                                           0, AUX);
  else
    __builtin_amdgcn_raw_buffer_store_b64(
        __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, make_float2(v.x, v.y)), rs, o, 0, AUX);
}

//   MODE_SPMV : kout = A xin
//   MODE_STAGE: kout = A xin;  ynext = x + cn * kout          (stages 1-3)
//   MODE_FINAL: xout = x + h/6 (K1 + 2 K2 + 2 K3 + A xin)    (stage 4)
// with the FP32 state x = hi + lo updated in FP64 and split again.
// WIN: workgroups of KS2_WIN / 128 wavefronts; each first copies its window of
// the stage input into LDS and reads the window slots from there.
// NT: the matrix values, the state words, K1..K3 and the K and low-word stores
// are streamed with the non-temporal cache policy (the gathered stage input and
// the next stage input Y keep the default policy)
// SEP: the matrix has separable-diagonal slots (the code for them is left out
// of the kernels of matrices without)
// TDP: each row also sums its merged A_k entries of a time-dependent RHS (the
// sliced form over 64-row slices: a lane's two rows are adjacent entries of one
// slice), before the epilogue
template <int MODE, int U, bool WIN, bool NT = false, bool SEP = false, bool TDP = false>
__global__ void __launch_bounds__(WIN ? KS2_WIN / 2 : 256, WIN ? 8 : 1) k_ks2_stage(
    int32_t n, int32_t nslices, const int32_t *__restrict__ moff, const int32_t *__restrict__ wmid,
    const int32_t *__restrict__ smid, const int32_t *__restrict__ umid, const int32_t *__restrict__ emid,
    const int32_t *__restrict__ cmid, const int32_t *__restrict__ xmid, const qh_sepblk *__restrict__ sepb,
    int32_t ngh, const double2 *__restrict__ gh, const qh_kmeta2 *__restrict__ meta, int32_t nval,
                                                   const float4 *__restrict__ val, int32_t nsval,
                                                   const float2 *__restrict__ sval, int32_t nsval1,
                                                   const float *__restrict__ sval1, const float2 *xin, float2 *kout,
                                                   float2 *ynext, float cn, const float2 *x, const float2 *k1,
                                                   const float2 *k2, const float2 *k3, float2 *xout, float h,
                                                   float2 *xlo, qh_td_ell td) {
  const unsigned b = qh_xcd_block(blockIdx.x, gridDim.x);
  const int s = (int)(b * (WIN ? KS2_WIN / 128 : 4) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6));
  const int64_t vb = (int64_t)n * 8;
  const __amdgpu_buffer_rsrc_t xr = ks_rsrc(xin, vb);
  __shared__ float2 xs[WIN ? KS2_WIN + 2 : 2];
  // WIN: the window rows [b KS2_WIN, (b+1) KS2_WIN), two per thread (zeros past
  // n), are loaded first but written to LDS only after the slots read from
  // global memory have been processed, so the copy's latency hides behind them;
  // every wavefront of the workgroup reaches the barrier
  float4 wv = make_float4(0.f, 0.f, 0.f, 0.f);
  if (WIN) {
    const int64_t w = (int64_t)b * KS2_WIN + 2 * threadIdx.x;
    if (w + 1 < n)
      wv = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(xr, (uint32_t)w * 8u, 0, 0));
    else if (w < n)
      wv = ld_pair(xr, (uint32_t)w * 8u, false);
  } else if (s >= nslices) {
    return;  // whole wavefront
  }
  const bool live = s < nslices;
  const int lane = threadIdx.x & 63;
  const int32_t r = s * 128 + 2 * lane;
  const int32_t k0 = live ? moff[s] : 0, kw = live ? wmid[s] : 0, ku = live ? smid[s] : 0,
                km = live ? umid[s] : 0, ke = live ? emid[s] : 0, kc = live ? cmid[s] : 0,
                kend = live ? moff[s + 1] : 0;
  const int32_t kx = (SEP && live) ? xmid[s] : kend;
  const __amdgpu_buffer_rsrc_t vr = ks_rsrc(val, (int64_t)nval * 16);
  const __amdgpu_buffer_rsrc_t sr = ks_rsrc(sval, (int64_t)nsval * 8);
  const __amdgpu_buffer_rsrc_t s1r = ks_rsrc(sval1, (int64_t)nsval1 * 4);
  const uint32_t r8 = (uint32_t)r * 8u;
  KsAcc2 acc{0.0, 0.0, 0.0, 0.0, 0.f, 0.f, 0.f, 0.f};
  // (a wavefront past the last slice has empty slot ranges: k0 = ... = kend = 0)
  ks_run<U, true>(Ks2Ops<KS_UNI, true>{meta + kw, r8, xr, vr}, ku - kw, acc);
  if (!WIN) ks_run<U, true>(Ks2Ops<KS_UNI, true>{meta + k0, r8, xr, vr}, kw - k0, acc);
  constexpr int AUX = NT ? 2 : 0;
  ks_run<U, true>(Ks2Ops<KS_UNI>{meta + ku, r8, xr, vr}, km - ku, acc);
  ks_run<U, true>(Ks2Ops<KS_SCA, false, AUX>{meta + km, r8, xr, sr}, ke - km, acc);
  ks_run<U, true>(Ks2Ops<KS_SC1, false, AUX>{meta + ke, r8, xr, s1r}, kc - ke, acc);
  ks_run<U, true>(Ks2Ops<KS_CPX, false, AUX>{meta + kc, r8, xr, vr}, kx - kc, acc);
  if (SEP && kend > kx) {  // the one separable-diagonal slot of the slice, after the pipelined groups
    Ks2OpsSep op{meta + kx, r8, (int64_t)r, xr, ks_rsrc(gh, (int64_t)ngh * 16), sepb};
    Ks2OpsSep::Buf<1> sb;
    op.issue<1>(sb);
    op.fma<1>(sb, acc);
    op.flush(acc);
  }
  if (WIN) {
    xs[2 * threadIdx.x] = make_float2(wv.x, wv.y);
    xs[2 * threadIdx.x + 1] = make_float2(wv.z, wv.w);
    if (threadIdx.x == 0) xs[KS2_WIN] = xs[KS2_WIN + 1] = make_float2(0.f, 0.f);
    __syncthreads();
    ks_run<U, true>(Ks2OpsWin{meta + k0, (uint32_t)(r - (int32_t)b * KS2_WIN) * 8u, xs}, kw - k0, acc);
  }
  if (!live) return;
  if (r >= n) return;
  const bool two = r + 1 < n;
  double a0r = acc.r0, a0i = acc.i0, a1r = acc.r1, a1i = acc.i1;
  if constexpr (TDP) {
    // r is even and below n, so its 64-row slice exists and the pair's columns
    // and indices are adjacent (8- and 4-byte aligned); a row r + 1 >= n reads
    // padding (column 0, coefficient 0)
    const int32_t sl = r >> 6;
    const int32_t w = td.w[sl];
    const int64_t base = td.off[sl] + (r & 63);
#pragma unroll 2
    for (int32_t i = 0; i < w; i++) {
      const int64_t e = base + 64 * (int64_t)i;
      const int2 c2 = *reinterpret_cast<const int2 *>(td.col + e);
      const uint32_t i2 = *reinterpret_cast<const uint32_t *>(td.idx + e);
      const float2 v0 = xin[c2.x], v1 = xin[c2.y];
      const double2 f0 = td.cf[i2 & 0xffffu], f1 = td.cf[i2 >> 16];
      a0r += f0.x * (double)v0.x - f0.y * (double)v0.y;
      a0i += f0.x * (double)v0.y + f0.y * (double)v0.x;
      a1r += f1.x * (double)v1.x - f1.y * (double)v1.y;
      a1i += f1.x * (double)v1.y + f1.y * (double)v1.x;
    }
  }
  const float4 y = make_float4((float)a0r, (float)a0i, (float)a1r, (float)a1i);
  if (MODE == MODE_SPMV) {
    st_pair(ks_rsrc(kout, vb), r8, two, y);
    return;
  }
  const float4 xo = ld_pair<AUX>(ks_rsrc(x, vb), r8, two);
  if (MODE == MODE_STAGE) {
    // the next stage input is stored in complex64: the low word of the state
    // (below half an ulp of the high word) would change it by at most one
    // rounding, so only the high word is read here (the update in stage 4
    // uses both)
    st_pair<AUX>(ks_rsrc(kout, vb), r8, two, y);
    const double c = (double)cn;
    const float4 yn = make_float4((float)((double)xo.x + c * a0r), (float)((double)xo.y + c * a0i),
                                  (float)((double)xo.z + c * a1r), (float)((double)xo.w + c * a1i));
    st_pair(ks_rsrc(ynext, vb), r8, two, yn);
    return;
  }
  const __amdgpu_buffer_rsrc_t lr = ks_rsrc(xlo, vb);
  const float4 lo = ld_pair<AUX>(lr, r8, two);
  const float4 a = ld_pair<AUX>(ks_rsrc(k1, vb), r8, two), bq = ld_pair<AUX>(ks_rsrc(k2, vb), r8, two),
               cq = ld_pair<AUX>(ks_rsrc(k3, vb), r8, two);
  const double s6 = (double)h * (1.0 / 6.0), s3 = (double)h * (1.0 / 3.0);
  const double n0r = ((double)xo.x + (double)lo.x) + (s6 * (double)a.x + s3 * (double)bq.x + s3 * (double)cq.x + s6 * a0r);
  const double n0i = ((double)xo.y + (double)lo.y) + (s6 * (double)a.y + s3 * (double)bq.y + s3 * (double)cq.y + s6 * a0i);
  const double n1r = ((double)xo.z + (double)lo.z) + (s6 * (double)a.z + s3 * (double)bq.z + s3 * (double)cq.z + s6 * a1r);
  const double n1i = ((double)xo.w + (double)lo.w) + (s6 * (double)a.w + s3 * (double)bq.w + s3 * (double)cq.w + s6 * a1i);
  const float4 o = make_float4((float)n0r, (float)n0i, (float)n1r, (float)n1i);
  st_pair(ks_rsrc(xout, vb), r8, two, o);
  st_pair<AUX>(lr, r8, two,
          make_float4((float)(n0r - (double)o.x), (float)(n0i - (double)o.y), (float)(n1r - (double)o.z),
                      (float)(n1i - (double)o.w)));
}

}  // namespace

// the diagonal entry of every row (NaN where the row has none)
__global__ void k_ks2_diag(int32_t n, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                           const double2 *__restrict__ val, double2 *__restrict__ d) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double2 v = make_double2(__builtin_nan(""), __builtin_nan(""));
  for (int32_t k = rp[r]; k < rp[r + 1]; k++)
    if (col[k] == r) v = val[k];
  d[r] = v;
}

// Separability of the diagonal of each candidate block (N x N density matrix
// at block offset off): row r = off + N j + i is bad when its diagonal entry
// D(i, j) is not finite or differs from D(i, 0) + D(0, j) - D(0, 0) by more than
// 1e-12 of the magnitudes involved.  The block of r by binary search over boff.
__global__ void k_ks2_sep_check(int32_t n, const double2 *__restrict__ d, int32_t nb,
                                const int64_t *__restrict__ boff, const int32_t *__restrict__ bn,
                                int32_t *__restrict__ bad) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (boff[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  const int32_t N = bn[lo];
  if (N == 0) return;
  const int64_t off = boff[lo];
  const uint32_t t = (uint32_t)(r - off), j = t / (uint32_t)N, i = t - j * (uint32_t)N;
  const double2 dij = d[r], di0 = d[off + i], d0j = d[off + (int64_t)N * j], d00 = d[off];
  const double hr = d0j.x - d00.x, hi_ = d0j.y - d00.y;
  const double er = dij.x - (di0.x + hr), ei = dij.y - (di0.y + hi_);
  const double sc = fabs(dij.x) + fabs(dij.y) + fabs(di0.x) + fabs(di0.y) + fabs(hr) + fabs(hi_);
  const bool ok = isfinite(dij.x) && isfinite(dij.y) && fabs(er) + fabs(ei) <= 1e-12 * sc;
  if (!ok) bad[lo] = 1;
}

// g_i = D(i, 0) and h_j = D(0, j) - D(0, 0) of one separable block into gh
__global__ void k_ks2_sep_gather(const double2 *__restrict__ d, int64_t off, int32_t N, double2 *__restrict__ g,
                                 double2 *__restrict__ h) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  g[k] = d[off + k];
  const double2 a = d[off + (int64_t)N * k], z = d[off];
  h[k] = make_double2(a.x - z.x, a.y - z.y);
}

// The blocks of A (the instances of a batch, else A itself) whose diagonal
// separates: an N x N density-matrix block (N even, starting at an even row)
// whose diagonal entry at row i, column j is g_i + h_j (g_i = D(i, 0),
// h_j = D(0, j) - D(0, 0), every entry checked to 1e-12 relative on the device).
// Fills the block parameters, their g and h (device, *gh_out of *ngh_out
// entries), and for every slice the block that holds it whole (-1: none).
static int ks2_sep_prepare(qh_mat *A, int32_t ns, std::vector<qh_sepblk> &blks, double2 **gh_out, int64_t *ngh_out,
                           std::vector<int32_t> &slice_blk) {
  const int64_t n = A->rows;
  std::vector<int64_t> bo = A->blk_off;
  if (bo.size() < 2 || bo.back() != n) bo = {0, n};
  const int32_t nb = (int32_t)bo.size() - 1;
  std::vector<int32_t> bn((size_t)nb, 0);
  int ncand = 0;
  for (int32_t b = 0; b < nb; b++) {
    const int64_t off = bo[b], rows = bo[b + 1] - bo[b];
    const int64_t N = (int64_t)std::llround(std::sqrt((double)rows));
    if (N < 2 || N * N != rows || (N & 1) || (off & 1) || N > 46340) continue;
    bn[b] = (int32_t)N;
    ncand++;
  }
  *gh_out = nullptr;
  *ngh_out = 0;
  std::vector<int32_t> blk_of((size_t)nb, -1);
  double2 *dd = nullptr, *gh = nullptr;
  int64_t *dboff = nullptr;
  int32_t *dbn = nullptr, *dbad = nullptr;
  hipStream_t s = qh_stream();
  int rc = QH_OK;
  if (ncand > 0) {
    rc = qh_alloc_t(&dd, (size_t)n);
    if (!rc) rc = qh_alloc_t(&dboff, bo.size());
    if (!rc) rc = qh_alloc_t(&dbn, (size_t)nb);
    if (!rc) rc = qh_alloc_t(&dbad, (size_t)nb);
    std::vector<int32_t> hbad((size_t)nb, 0);
    hipError_t e = hipSuccess;
    if (!rc) {
      hipLaunchKernelGGL(k_ks2_diag, dim3(qh_blocks(n, 256)), dim3(256), 0, s, (int32_t)n, A->rowptr, A->col, A->val,
                         dd);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(dboff, bo.data(), sizeof(int64_t) * bo.size(), hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipMemcpyAsync(dbn, bn.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipMemsetAsync(dbad, 0, sizeof(int32_t) * nb, s);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ks2_sep_check, dim3(qh_blocks(n, 256)), dim3(256), 0, s, (int32_t)n, dd, nb, dboff, dbn,
                           dbad);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipMemcpyAsync(hbad.data(), dbad, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        qh_set_error("separable-diagonal check: %s", hipGetErrorString(e));
        rc = QH_EHIP;
      }
    }
    int64_t total = 0;
    for (int32_t b = 0; !rc && b < nb; b++) {
      if (bn[b] == 0 || hbad[b]) continue;
      const int64_t N = bn[b];
      if (total + 2 * N >= ((int64_t)1 << 27)) continue;
      qh_sepblk q;
      q.off = bo[b];
      q.n = (uint32_t)N;
      uint32_t sh = 0;
      while (sh < 32 && ((uint64_t)1 << sh) < (uint64_t)N) sh++;
      q.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << sh) - (uint64_t)N)) / (uint64_t)N + 1);
      q.sh = sh;
      q.gbase = (uint32_t)total;
      q.hbase = (uint32_t)(total + N);
      q.pad = 0;
      blk_of[b] = (int32_t)blks.size();
      blks.push_back(q);
      total += 2 * N;
    }
    if (!rc && total > 0) {
      rc = qh_alloc_t(&gh, (size_t)total);
      for (size_t k = 0; !rc && k < blks.size(); k++) {
        const qh_sepblk &q = blks[k];
        hipLaunchKernelGGL(k_ks2_sep_gather, dim3(qh_blocks(q.n, 256)), dim3(256), 0, s, dd, q.off, (int32_t)q.n,
                           gh + q.gbase, gh + q.hbase);
      }
      if (!rc) {
        const hipError_t e2 = hipGetLastError();
        if (e2 == hipSuccess) {
          *gh_out = gh;
          *ngh_out = total;
          gh = nullptr;
        } else {
          qh_set_error("separable-diagonal gather: %s", hipGetErrorString(e2));
          rc = QH_EHIP;
        }
      }
    }
    if (!rc) {
      const hipError_t e3 = hipStreamSynchronize(s);  // dd is freed below
      if (e3 != hipSuccess) {
        qh_set_error("separable-diagonal gather: %s", hipGetErrorString(e3));
        rc = QH_EHIP;
      }
    }
  }
  qh_free_ptr(dd);
  qh_free_ptr(gh);
  qh_free_ptr(dboff);
  qh_free_ptr(dbn);
  qh_free_ptr(dbad);
  if (rc) {
    blks.clear();
    qh_free_ptr(*gh_out);
    *gh_out = nullptr;
    *ngh_out = 0;
    return rc;
  }
  slice_blk.assign((size_t)ns, -1);
  size_t b = 0;
  for (int32_t sl = 0; sl < ns; sl++) {
    const int64_t lo = (int64_t)sl * 128, hi = std::min<int64_t>(lo + 128, n);
    while (b + 1 < bo.size() && bo[b + 1] <= lo) b++;
    if (b + 1 < bo.size() && hi <= bo[b + 1]) slice_blk[sl] = blk_of[b];
  }
  return QH_OK;
}

// Build the row-pair layout of A (FP32 values) once; ks2_state -1 when it does
// not apply (fewer than 2 rows, a row longer than KS2_MAXROW, or byte offsets
// beyond the 2^31 reach of the buffer loads).
int ks2_build(qh_mat *A) {
  if (A->ks2_state) return QH_OK;
  hipStream_t s = qh_stream();
  if (A->rows < 2 || A->rows * (int64_t)sizeof(float2) >= ((int64_t)1 << 31)) {
    A->ks2_state = -1;
    return QH_OK;
  }
  const int32_t ns = (int32_t)((A->rows + 127) / 128);
  const int NC = 10;  // per-slice counts of k_ks2_build
  int32_t *cnt = nullptr, *voff = nullptr, *svoff = nullptr, *s1off = nullptr, *sblk = nullptr;
  int *ovf = nullptr;
  std::vector<qh_sepblk> blks;
  double2 *ghd = nullptr;  // device g/h of the separable blocks (owned by A once built)
  int64_t ngh = 0;
  std::vector<int32_t> hsb;
  int rc = ks2_sep_prepare(A, ns, blks, &ghd, &ngh, hsb);
  if (!rc) rc = qh_alloc_t(&cnt, (size_t)NC * ns);
  if (!rc) rc = qh_alloc_t(&voff, (size_t)ns + 1);
  if (!rc) rc = qh_alloc_t(&svoff, (size_t)ns + 1);
  if (!rc) rc = qh_alloc_t(&s1off, (size_t)ns + 1);
  if (!rc) rc = qh_alloc_t(&sblk, (size_t)ns);
  if (!rc) rc = qh_alloc_t(&ovf, 1);
  auto cleanup = [&]() {
    qh_free_ptr(cnt);
    qh_free_ptr(voff);
    qh_free_ptr(svoff);
    qh_free_ptr(s1off);
    qh_free_ptr(sblk);
    qh_free_ptr(ovf);
    qh_free_ptr(ghd);  // null once A owns it
    ghd = nullptr;
  };
  if (rc) {
    cleanup();
    return rc;
  }
  QH_HIP(hipMemcpyAsync(sblk, hsb.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  QH_HIP(hipMemsetAsync(ovf, 0, sizeof(int), s));
  hipLaunchKernelGGL((k_ks2_build<false>), dim3(ns), dim3(64), 0, s, (int32_t)A->rows, A->rowptr, A->col, A->val,
                     cnt, ns, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, sblk, ovf);
  QH_HIP(hipGetLastError());
  int hovf = 0;
  std::vector<int32_t> hc((size_t)NC * ns);
  QH_HIP(hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost, s));
  QH_HIP(hipMemcpyAsync(hc.data(), cnt, sizeof(int32_t) * NC * ns, hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  const int32_t *hw = hc.data(), *hn = hw + ns, *hu = hn + ns, *hs = hu + ns, *hsv = hs + ns, *hst = hsv + ns,
                *hwn = hst + ns, *hs1 = hwn + ns, *hs1v = hs1 + ns, *hxs = hs1v + ns;
  std::vector<int32_t> ho(ns + 1), hv(ns + 1), hso(ns + 1), hs1o(ns + 1), hm(ns), hem(ns), hcm(ns), hsm(ns), hwm(ns),
      hxm(ns);
  int64_t tot = 0, totv = 0, tots = 0, tots1 = 0, totx = 0;
  for (int32_t i = 0; i < ns; i++) {
    ho[i] = (int32_t)tot;
    hv[i] = (int32_t)totv;
    hso[i] = (int32_t)tots;
    hs1o[i] = (int32_t)tots1;
    hwm[i] = (int32_t)(tot + hwn[i]);
    hsm[i] = (int32_t)(tot + hst[i]);
    hm[i] = (int32_t)(tot + hu[i]);
    hem[i] = (int32_t)(tot + hu[i] + hs[i]);
    hcm[i] = (int32_t)(tot + hu[i] + hs[i] + hs1[i]);
    hxm[i] = (int32_t)(tot + hw[i] - hxs[i]);
    totx += hxs[i];
    tot += hw[i];
    totv += hn[i];
    tots += hsv[i];
    tots1 += hs1v[i];
  }
  if (hovf || totv * 16 >= ((int64_t)1 << 31) || tots * 8 >= ((int64_t)1 << 31) ||
      tots1 * 4 >= ((int64_t)1 << 31) || tot >= ((int64_t)1 << 31)) {
    A->ks2_state = -1;
    cleanup();
    return QH_OK;
  }
  ho[ns] = (int32_t)tot;
  hv[ns] = (int32_t)totv;
  hso[ns] = (int32_t)tots;
  hs1o[ns] = (int32_t)tots1;
  rc = qh_alloc_t(&A->ks2_moff, (size_t)ns + 1);
  if (!rc) rc = qh_alloc_t(&A->ks2_wmid, (size_t)ns);
  if (!rc) rc = qh_alloc_t(&A->ks2_smid, (size_t)ns);
  if (!rc) rc = qh_alloc_t(&A->ks2_umid, (size_t)ns);
  if (!rc) rc = qh_alloc_t(&A->ks2_emid, (size_t)ns);
  if (!rc) rc = qh_alloc_t(&A->ks2_cmid, (size_t)ns);
  if (!rc) rc = qh_alloc_t(&A->ks2_xmid, (size_t)ns);
  if (!rc) rc = qh_alloc_t(&A->ks2_sep, blks.empty() ? 1 : blks.size());
  if (!rc) {
    if (ghd) {
      A->ks2_gh = ghd;
      ghd = nullptr;
    } else {
      rc = qh_alloc_t(&A->ks2_gh, 1);
    }
  }
  if (!rc) rc = qh_alloc_t(&A->ks2_meta, (size_t)(tot > 0 ? tot : 1));
  if (!rc) rc = qh_alloc_t(&A->ks2_val, (size_t)(totv > 0 ? totv : 1));
  if (!rc) rc = qh_alloc_t(&A->ks2_sval, (size_t)(tots > 0 ? tots : 1));
  if (!rc) rc = qh_alloc_t(&A->ks2_sval1, (size_t)(tots1 > 0 ? tots1 : 1));
  if (rc) {
    cleanup();
    return rc;
  }
  QH_HIP(hipMemcpyAsync(A->ks2_moff, ho.data(), sizeof(int32_t) * (ns + 1), hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(A->ks2_wmid, hwm.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(A->ks2_smid, hsm.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(A->ks2_umid, hm.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(A->ks2_emid, hem.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(A->ks2_cmid, hcm.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(A->ks2_xmid, hxm.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  if (!blks.empty())
    QH_HIP(hipMemcpyAsync(A->ks2_sep, blks.data(), sizeof(qh_sepblk) * blks.size(), hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(voff, hv.data(), sizeof(int32_t) * (ns + 1), hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(svoff, hso.data(), sizeof(int32_t) * (ns + 1), hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(s1off, hs1o.data(), sizeof(int32_t) * (ns + 1), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL((k_ks2_build<true>), dim3(ns), dim3(64), 0, s, (int32_t)A->rows, A->rowptr, A->col, A->val,
                     cnt, ns, A->ks2_moff, voff, svoff, s1off, A->ks2_meta, A->ks2_val, A->ks2_sval, A->ks2_sval1,
                     sblk, ovf);
  QH_HIP(hipGetLastError());
  QH_HIP(hipStreamSynchronize(s));
  cleanup();
  A->ks2_nslices = ns;
  A->ks2_slots = tot;
  A->ks2_nval = totv;
  A->ks2_nsval = tots;
  A->ks2_nsval1 = tots1;
  A->ks2_nsep = totx;
  A->ks2_ngh = ngh;
  A->ks2_nsepblk = (int64_t)blks.size();
  A->ks2_state = 1;
  return QH_OK;
}

// One stage launch on the row-pair layout (G = 140 / 141: 2 / 4 slots per pass;
// 142: 2 slots per pass, window slots read from LDS; 143 / 144: 140 / 142 with
// the non-temporal cache policy on the streamed data);
// mode as in the one-row-per-lane kernels.
int ks2_launch(int G, int mode, hipStream_t s, const qh_mat *A, const void *xin, void *kout, void *ynext, float cn,
               const void *x, const void *k1, const void *k2, const void *k3, void *xout, float h, float2 *xlo,
               const qh_td_ell *td) {
  if (A->ks2_state != 1) {
    qh_set_error("row-pair sliced layout not built for this matrix");
    return QH_EINVAL;
  }
  const int32_t n = (int32_t)A->rows, ns = A->ks2_nslices;
  const bool win = G == 142 || G == 144;
  const dim3 grid(qh_blocks(ns, win ? KS2_WIN / 128 : 4)), blk(win ? KS2_WIN / 2 : 256);
  const qh_td_ell none{nullptr, nullptr, nullptr, nullptr, nullptr};
#define QH_LK2T(MM, UU, WW, NN, SS, TT, TD)                                                                      \
  hipLaunchKernelGGL((k_ks2_stage<MM, UU, WW, NN, SS, TT>), grid, blk, 0, s, n, ns, A->ks2_moff, A->ks2_wmid,       \
                     A->ks2_smid,                                                                                \
                     A->ks2_umid, A->ks2_emid, A->ks2_cmid, A->ks2_xmid, A->ks2_sep,                               \
                     (int32_t)A->ks2_ngh,                                                                        \
                     A->ks2_gh, A->ks2_meta, (int32_t)A->ks2_nval, A->ks2_val,                                    \
                     (int32_t)A->ks2_nsval, A->ks2_sval, (int32_t)A->ks2_nsval1, A->ks2_sval1,                  \
                     (const float2 *)xin, (float2 *)kout, (float2 *)ynext, cn, (const float2 *)x,                \
                     (const float2 *)k1, (const float2 *)k2, (const float2 *)k3, (float2 *)xout, h, xlo, TD)
#define QH_LK2S(MM, UU, WW, NN, SS) QH_LK2T(MM, UU, WW, NN, SS, false, none)
#define QH_LK2(MM, UU, WW, NN)            \
  do {                                    \
    if (A->ks2_nsep)                      \
      QH_LK2S(MM, UU, WW, NN, true);      \
    else                                  \
      QH_LK2S(MM, UU, WW, NN, false);     \
  } while (0)
  if (!QH_IS_KS2(G) || mode < MODE_SPMV || mode > MODE_FINAL) {
    qh_set_error("unsupported row-pair variant %d / mode %d", G, mode);
    return QH_EINVAL;
  }
  if (td) {  // the time-dependent entries summed in the same kernel (variants 140, 142, 143)
#define QH_LK2D(MM, UU, WW, NN)                                                             \
  do {                                                                                      \
    if (A->ks2_nsep)                                                                        \
      QH_LK2T(MM, UU, WW, NN, true, true, *td);                                            \
    else                                                                                    \
      QH_LK2T(MM, UU, WW, NN, false, true, *td);                                           \
  } while (0)
#define QH_LK2DG(MM)                                                                        \
  switch (G) {                                                                              \
    case 140: QH_LK2D(MM, 2, false, false); break;                                          \
    case 142: QH_LK2D(MM, 2, true, false); break;                                           \
    case 143: QH_LK2D(MM, 2, false, true); break;                                           \
    default:                                                                                \
      qh_set_error("row-pair variant %d has no time-dependent form", G);                    \
      return QH_EINVAL;                                                                     \
  }
    if (mode == MODE_SPMV) {
      QH_LK2DG(MODE_SPMV)
    } else if (mode == MODE_STAGE) {
      QH_LK2DG(MODE_STAGE)
    } else {
      QH_LK2DG(MODE_FINAL)
    }
#undef QH_LK2DG
#undef QH_LK2D
    return QH_OK;
  }
#define QH_LK2G(MM)                      \
  switch (G) {                           \
    case 140: QH_LK2(MM, 2, false, false); break; \
    case 141: QH_LK2(MM, 4, false, false); break; \
    case 142: QH_LK2(MM, 2, true, false); break;  \
    case 143: QH_LK2(MM, 2, false, true); break;  \
    default: QH_LK2(MM, 2, true, true); break;    \
  }
  if (mode == MODE_SPMV) {
    QH_LK2G(MODE_SPMV)
  } else if (mode == MODE_STAGE) {
    QH_LK2G(MODE_STAGE)
  } else {
    QH_LK2G(MODE_FINAL)
  }
#undef QH_LK2G
#undef QH_LK2S
#undef QH_LK2T
#undef QH_LK2
  QH_HIP(hipGetLastError());
  return QH_OK;
}
