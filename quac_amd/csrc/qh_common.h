// qh_common.h -- internal definitions shared by the gfx950 kernels of libquac.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>
#include <cstdio>

#include "../../include/quac_hip.h"

// ----------------------------------------------------------------- errors
void qh_set_error(const char *fmt, ...);

#define QH_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      qh_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return QH_EHIP;                                                                \
    }                                                                                \
  } while (0)

#define QH_REQUIRE(cond, code, ...)   \
  do {                                \
    if (!(cond)) {                    \
      qh_set_error(__VA_ARGS__);      \
      return (code);                  \
    }                                 \
  } while (0)

hipStream_t qh_stream();
int qh_ensure_init();
int qh_num_cus();

// ----------------------------------------------------------------- objects
// Slot metadata of the slot-aligned sliced-ELL SpMV layout (one slice = 64
// consecutive rows = one wavefront).  Every entry of the slice whose column is
// row +- ad (and that is the rank-th such entry of its row) sits in one slot;
// mask marks the lanes (rows) that have it, sign the lanes whose column is
// row - ad, and base is the index of the slot's first value in the compacted
// value array (values of active lanes are stored consecutively in lane order).
// When every active lane of the slot holds the bit-identical value, the value
// is kept in the metadata (vre, vim) and base is QH_KS_UNIFORM: no per-lane
// value is stored or loaded for that slot.  When the values differ but are all
// real, or all imaginary, the slot is a scalar slot: one real number per lane
// is stored in the scalar value array (index base + lane rank) and the value of
// the lane is that number times (vre, vim) = (1, 0) or (0, 1).  The remaining
// slots store complex values.  A slice lists its uniform slots first, then its
// scalar slots, then its complex slots, each group in increasing (ad, rank)
// order.
struct qh_kmeta {
  uint64_t mask, sign;
  uint32_t ad, base;
  double vre, vim;
};
#define QH_KS_UNIFORM 0xFFFFFFFFu
// The FP32 path's copy of the slot metadata, 32 B per slot (one scalar load):
// the byte offset ad8 = 8 |d| of a complex64 gather, and in `base` the
// value index of a scalar or complex slot or, for a uniform slot, the negated
// offset -ad8; (vre, vim) in FP32.
struct qh_kmeta32 {
  uint64_t mask, sign;
  uint32_t ad8, base;
  float vre, vim;
};
// Slot metadata of the FP32 row-pair sliced layout (one slice = 128 consecutive
// rows = one wavefront, lane l owns rows 2l and 2l+1 of the slice).  A slot is one
// 16-byte load per lane of the complex64 pair x[b], x[b+1] with b = 2l +- ad
// (sign as above); row 2l takes half h0 of the pair and row 2l+1 half h1, and the
// slot's lanes share (ad, h0, h1); a row of a lane without the entry has value
// zero.  ad8 holds 8 ad | h0 | h1 << 1.  Uniform slots (both rows' values the same
// across the active lanes) keep them in (v0re, v0im), (v1re, v1im); scalar slots
// store two reals per active lane (value = real * (v0re, v0im)); complex slots
// store two complex64 per active lane.
struct qh_kmeta2 {
  uint64_t mask, sign;
  uint32_t ad8, base;
  float v0re, v0im, v1re, v1im;
};
// empty slots after the last one, so that a pass of up to 8 slots may read
// its metadata without a bounds check
#define QH_KS_PAD 16

// A block of the row-pair layout whose diagonal separates: the diagonal entry of
// row off + N j + i (density-matrix row i, column j of an N x N block, N even) is
// g_i + h_j, with g at gh[gbase ..] and h at gh[hbase ..] (complex128); m, sh
// divide a row index within the block by N (q = (mulhi(t, m) + t) >> sh).
struct qh_sepblk {
  int64_t off;
  uint32_t n, m, sh, gbase, hbase, pad;
};

struct qh_td_cache;  // merged time-dependent matrices and stage buffers (qh_td.hip)
void qh_td_cache_free(qh_td_cache *c);

// a process-wide serial number for every qh_mat created (caches keyed on a
// matrix compare it with the pointer, which a new matrix may reuse)
uint64_t qh_next_mat_uid();

struct qh_mat {
  uint64_t uid = qh_next_mat_uid();
  int64_t rows = 0, nnz = 0;
  int32_t *rowptr = nullptr;   // rows+1
  int32_t *col = nullptr;      // nnz
  double2 *val = nullptr;      // nnz complex128
  float2 *val32 = nullptr;     // nnz complex64 (FP32 path), made on demand
  int lanes64 = 0, lanes32 = 0;  // SpMV variant: lanes per row of the CSR kernel, or QH_KSELL
  int lanes_forced = 0;          // set by qh_mat_set_lanes (not by the autotuner)
  int32_t max_row = 0;         // longest row
  // slot-aligned sliced ELL (built on demand from the CSR)
  int ks_state = 0;            // 0 not built, 1 built, -1 not applicable (a row longer than 64)
  int32_t nslices = 0;
  int32_t *ks_moff = nullptr;  // nslices+1 slot offsets
  int32_t *ks_umid = nullptr;  // nslices: end of the slice's uniform slots (they come first)
  int32_t *ks_cmid = nullptr;  // nslices: end of the slice's scalar slots (the complex ones follow)
  qh_kmeta *ks_meta = nullptr;
  double2 *ks_val = nullptr;   // values of the complex slots, slice/slot/lane order
  float2 *ks_val32 = nullptr;
  double *ks_sval = nullptr;   // values of the scalar slots, slice/slot/lane order
  float *ks_sval32 = nullptr;
  qh_kmeta32 *ks_meta32 = nullptr;  // FP32 metadata (ks_slots + QH_KS_PAD entries)
  int64_t ks_slots = 0;
  int64_t ks_nval = 0;         // stored complex values
  int64_t ks_nsval = 0;        // stored scalar values
  // FP32 row-pair sliced layout (built on demand from the CSR)
  int ks2_state = 0;           // 0 not built, 1 built, -1 not applicable
  int32_t ks2_nslices = 0;
  int32_t *ks2_moff = nullptr;  // nslices+1 slot offsets
  int32_t *ks2_wmid = nullptr;  // end of each slice's in-order uniform slots inside its 2048-row window (first)
  int32_t *ks2_smid = nullptr;  // end of each slice's uniform slots with in-order halves
  int32_t *ks2_umid = nullptr;  // end of each slice's uniform slots
  int32_t *ks2_emid = nullptr;  // end of each slice's scalar slots with a value per row
  int32_t *ks2_cmid = nullptr;  // end of each slice's scalar slots (equal-scalar slots included)
  qh_kmeta2 *ks2_meta = nullptr;
  float4 *ks2_val = nullptr;    // complex slots: (row 2l, row 2l+1) complex64 pairs
  float2 *ks2_sval = nullptr;   // scalar slots: (row 2l, row 2l+1) real pairs
  float *ks2_sval1 = nullptr;   // equal-scalar slots: one real value for both rows of a lane
  int32_t *ks2_xmid = nullptr;  // end of each slice's complex slots (separable-diagonal slots follow)
  qh_sepblk *ks2_sep = nullptr; // blocks with a separable diagonal
  double2 *ks2_gh = nullptr;    // their g and h vectors
  int64_t ks2_slots = 0, ks2_nval = 0, ks2_nsval = 0, ks2_nsval1 = 0, ks2_nsep = 0, ks2_ngh = 0, ks2_nsepblk = 0;
  std::vector<int64_t> blk_off;  // row offsets of the blocks of a block-diagonal matrix (and its row count)
  qh_td_cache *td_cache = nullptr;  // built by the first time-dependent RHS on this matrix
};
// variant codes selecting the sliced-ELL kernel: 128 / 129 / 130 process 2 / 4 / 8
// slots per pass; 131 / 132 process 4 / 8 slots per pass and read the uniform
// slot values with vector loads; 133 / 134 process 2 / 4 slots per pass with the
// loads of the next pass in flight during the products of the current one
#define QH_KSELL 128
// 135 / 136: 129 / 131 with the streamed data under the non-temporal cache policy
#define QH_IS_KSELL(g) ((g) >= 128 && (g) <= 136)
// 140 / 141: the FP32 row-pair sliced layout, 2 / 4 slots per pass (pipelined);
// 142: 2 slots per pass, the slots inside each 2048-row window read from LDS;
// 143 / 144: 140 / 142 with the streamed data under the non-temporal cache policy
#define QH_IS_KS2(g) ((g) >= 140 && (g) <= 144)

struct qh_vec {
  int64_t n = 0;
  int prec = QH_FP64;
  void *d = nullptr;           // double2* (FP64) or float2* high part (FP32)
  float2 *lo = nullptr;        // FP32 compensation word (low part), zero-initialised
  int view = 0;                // 1: d points into another vector's storage (not freed here)
};

// FP32 row-pair sliced layout (qh_ks2.hip): build once per matrix, launch one
// stage (mode 0 SpMV, 1 stages 1-3, 2 stage 4) with variant G = 140 .. 144; td
// (may be null): the merged A_k of a time-dependent RHS summed in the same kernel
// (variants 140, 142, 143)
struct qh_td_ell;
int ks2_build(qh_mat *A);
int ks2_launch(int G, int mode, hipStream_t s, const qh_mat *A, const void *xin, void *kout, void *ynext, float cn,
               const void *x, const void *k1, const void *k2, const void *k3, void *xout, float h, float2 *xlo,
               const qh_td_ell *td = nullptr);

// the Bogacki-Shampine trial step of qh_rk3bs_step with the merged
// time-dependent entries summed in A's stage kernels (td4: the tables of the
// stage times t, t + h/2, t + 3h/4, t + h), where A's FP64 variant is one of the
// one-row sliced variants 128, 134, 136 (*done = 1); otherwise nothing is
// launched (*done = 0)
int qh_rk3bs_td_sliced(qh_mat *A, const qh_vec *x, qh_vec *xn, qh_vec *k1, qh_vec *k2, qh_vec *k3, qh_vec *k4,
                       qh_vec *y, double h, double atol, double rtol, int have_k1, double *err,
                       const qh_td_ell *td4, int *done);

// longest row of A into A->max_row (computed once per matrix)
int qh_mat_max_row(qh_mat *A);

// the merged A_k entries of a time-dependent RHS in their sliced form (slices of
// 64 rows, entry i of row 64 s + l at off[s] + 64 i + l, w[s] entries per row)
// with the stage's coefficient table cf (s_k(t) v per dictionary index)
struct qh_td_ell {
  const int64_t *off;
  const int32_t *w;
  const int32_t *col;
  const uint16_t *idx;
  const double2 *cf;
};

// A's part of the time-dependent RK stages on A's autotuned variant when that is
// a sliced layout (qh_spmv.hip): *G = the variant, or 0 where A runs on its CSR
// (a generator of at most 800 rows, or a CSR variant chosen)
int qh_td_sliced_prepare(qh_mat *A, int prec, int *G);
// one stage of y' = A y on variant G: mode 0 kout = A xin; 1 also ynext = x +
// cn kout; 2 xout = x + h/6 (k1 + 2 k2 + 2 k3 + A xin) (x and the low word xlo of
// an FP32 state may be updated in place, xout = x)
// td (optional, may be null): the merged A_k to add inside A's stage kernel,
// done where the variant is one of the one-row sliced variants 128, 134, 136 or
// the row-pair variants 140, 142, 143 (*fused = 1); otherwise A's part alone
// (*fused = 0)
int qh_td_sliced_stage(int G, const qh_mat *A, int prec, int mode, const void *xin, void *kout, void *ynext,
                       double cn, const void *x, const void *k1, const void *k2, const void *k3, void *xout,
                       double h, float2 *xlo, const qh_td_ell *td, int *fused);

// device allocation helpers (return QH_* codes)
int qh_alloc(void **p, size_t bytes);
void qh_free_ptr(void *p);
// persistent scratch of the synchronous reductions: a device buffer of at least
// dbytes and a pinned host buffer of at least hbytes (either pointer may be
// null); valid until the next qh_scratch call, which may reallocate them
int qh_scratch(size_t dbytes, void **d, size_t hbytes, void **h);
// device address (on the current device) of the fixed scratch array of
// qh_expect_dm_noalloc, resolved once per device (qh_obs.hip)
int qh_noalloc_setup(char **out);

template <typename T>
static inline int qh_alloc_t(T **p, size_t count) {
  return qh_alloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

// --------------------------------------------------------- complex helpers
// Products are written out exactly as C99 complex multiplication, and the
// assembly kernels are compiled with contraction off so that every bit of the
// assembled values matches the host-side C arithmetic.
struct cxd {
  double re, im;
};
__host__ __device__ static inline cxd cxd_make(double r, double i) {
  cxd z;
  z.re = r;
  z.im = i;
  return z;
}

static inline unsigned qh_blocks(int64_t work, int per_block) {
  int64_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Remap a linear block index so that each of the 8 XCDs of an MI355X (which
// receive blocks round-robin) works on one contiguous range of the index space:
// neighbouring rows share an L2 slice, which keeps the near-diagonal gathers of
// x local to one XCD.
__device__ static inline unsigned qh_xcd_block(unsigned b, unsigned nb) {
  const unsigned nx = 8;
  if (nb < nx * 2) return b;
  unsigned per = nb / nx, rem = nb % nx;
  unsigned x = b % nx, k = b / nx;
  // XCD x owns logical blocks [x*per + min(x,rem), ...) of size per + (x<rem)
  unsigned base = x * per + (x < rem ? x : rem);
  unsigned size = per + (x < rem ? 1u : 0u);
  if (k < size) return base + k;
  return b;  // unreachable for a full round-robin distribution
}
