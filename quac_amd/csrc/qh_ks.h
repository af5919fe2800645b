// qh_ks.h -- device pieces shared by the sliced-layout stage kernels
// (qh_spmv.hip: one row per lane; qh_ks2.hip: two rows per lane).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qh_common.h"

namespace {

// stage epilogues: SpMV only; RK4 stages 1-3; RK4 stage 4; Bogacki-Shampine
// stage 3 (k3 and the 3rd-order candidate); Bogacki-Shampine error stage (k4 and
// the weighted error norm)
enum { MODE_SPMV = 0, MODE_STAGE = 1, MODE_FINAL = 2, MODE_BS3 = 3, MODE_BS3E = 4 };

// Raw buffer resource over n elements (gfx9 descriptor word 3: 32-bit data
// format).  A buffer load whose byte offset is at or beyond n * sizeof(T)
// returns zero instead of touching memory: inactive lanes load with offset
// QH_KS_OOB and so contribute exact zeros without any per-lane select.
#define QH_KS_OOB 0x80000000u
__device__ inline __amdgpu_buffer_rsrc_t ks_rsrc(const void *p, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}

// slot groups of a slice (KS_SC1: real values, one per lane for both of its rows)
enum { KS_UNI = 0, KS_SCA = 1, KS_CPX = 2, KS_SC1 = 3 };

// row sum of a one-row-per-lane pass (FP64)
struct KsAcc {
  double re, im;
};

// The last rem (< U) slots of a group: for U <= 4 a pass of exactly rem slots
// (no load is issued for a slot that is not there), else a masked pass of U.
// before() runs between the issue and the products (the pipelined driver
// multiplies its pending pass there).
template <int U, typename Ops, typename F>
__device__ __forceinline__ void ks_tail(Ops &ops, int32_t rem, F &&before, typename Ops::Acc &acc) {
  if constexpr (U <= 4) {
    if (rem == 1) {
      typename Ops::template Buf<1> t;
      ops.template issue<1>(t);
      __builtin_amdgcn_sched_barrier(0);
      before();
      ops.fma(t, acc);
    }
    if constexpr (U > 2) {
      if (rem == 2) {
        typename Ops::template Buf<2> t;
        ops.template issue<2>(t);
        __builtin_amdgcn_sched_barrier(0);
        before();
        ops.fma(t, acc);
      }
    }
    if constexpr (U > 3) {
      if (rem == 3) {
        typename Ops::template Buf<3> t;
        ops.template issue<3>(t);
        __builtin_amdgcn_sched_barrier(0);
        before();
        ops.fma(t, acc);
      }
    }
  } else {
    typename Ops::template Buf<U> t;
    ops.template issue<U, true>(t, rem);
    __builtin_amdgcn_sched_barrier(0);
    before();
    ops.fma(t, acc);
  }
}

// all n slots of one group of a slice, U per pass.  PIPE: the loads of pass p+1
// are issued before the products of pass p (two passes in flight; the passes
// alternate between two buffers so no loaded register is copied while its load
// is outstanding, and the scheduling barriers keep the compiler from moving
// the products of a pass above the issue of the next one).  The products are
// summed in pass order either way; ops.flush(acc) marks where the ops may move
// partial sums into their final accumulators (after every second pass and at
// the end).
template <int U, bool PIPE, typename Ops>
__device__ __forceinline__ void ks_run(Ops ops, int32_t n, typename Ops::Acc &acc) {
  static_assert(U - 1 <= QH_KS_PAD, "tail pass reads past the padding");
  typename Ops::template Buf<U> a, b;
  const int32_t np = n / U, rem = n % U;
  if (!PIPE || np == 0) {
    for (int32_t p = 0; p < np; p++) {
      ops.template issue<U>(a);
      ops.fma(a, acc);
      ops.flush(acc);
    }
    if (rem) ks_tail<U>(ops, rem, [] {}, acc);
    ops.flush(acc);
    return;
  }
  ops.template issue<U>(a);
  int32_t p = 1;
  for (; p + 1 < np; p += 2) {
    ops.template issue<U>(b);
    __builtin_amdgcn_sched_barrier(0);
    ops.fma(a, acc);
    ops.template issue<U>(a);
    __builtin_amdgcn_sched_barrier(0);
    ops.fma(b, acc);
    ops.flush(acc);
  }
  if (p < np) {  // a in flight, one full pass left
    ops.template issue<U>(b);
    __builtin_amdgcn_sched_barrier(0);
    ops.fma(a, acc);
    if (rem)
      ks_tail<U>(ops, rem, [&] { ops.fma(b, acc); }, acc);
    else
      ops.fma(b, acc);
  } else {  // a in flight, no full pass left
    if (rem)
      ks_tail<U>(ops, rem, [&] { ops.fma(a, acc); }, acc);
    else
      ops.fma(a, acc);
  }
  ops.flush(acc);
}

__device__ inline uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}


}  // namespace
