// Bucket accumulation, merge and window reduction kernels (templated on the coordinate field),
// shared by kernels_msm.hip (G1) and kernels_msm_g2.hip (G2). See kernels_msm.hip for the
// pipeline.
#pragma once
#include <algorithm>
#include <string>

#include "bn254.cuh"
#include "internal.h"

namespace g16 {

inline size_t accum_align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Entries per accumulation segment: 32, and 40 for a G1 MSM over per-window buckets, whose
// merges are one or two additions per bucket, so fewer and longer segments pay there (the
// 2^20 G1 MSM: 2.13 against 2.25-2.30 ms; G2 and the prove's merged MSMs were slower).
// Constant-work MSMs over window tables (the prove's) take 64: their accumulation runs the
// same number of rounds of waves at the prove's sizes and the fold starts from half the slots
// (47.8-47.9 against 47.4-47.5 proofs/s). The length never depends on the batch: a wave's
// scratch is sized for its cap of proofs, and every smaller wave of the circuit must cut its
// entries the same way to fit it.
inline uint32_t accum_seg(const MsmPlan& p, size_t point_bytes) {
  if (p.constant && p.merged) return 64u;
  return !p.merged && point_bytes == 128 ? 40u : 32u;
}

template <int = 0>
__global__ void k_clear_cnt(uint32_t* cnt) {
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
}
// Segments a merge thread folds itself, as a chain of dependent additions; a bucket past it
// goes to the heavy path's parallel trees. At 32 a few dense buckets (the narrow top window of
// the standalone G2 MSM) kept the merge running ~450 us with most SIMDs idle.
static constexpr int HEAVY_SPAN = 8;

__device__ __forceinline__ Fr ld_fr(const uint4* p) {
  uint4 a = p[0], b = p[1];
  return {{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}};
}
__device__ __forceinline__ void st_fr(uint4* p, const Fr& x) {
  p[0] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  p[1] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}
__device__ __forceinline__ Fq ld_fq(const uint4* p) {
  uint4 a = p[0], b = p[1];
  return {{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}};
}
__device__ __forceinline__ void st_fq(uint4* p, const Fq& x) {
  p[0] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  p[1] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}

// Field element loads in units of 32 bytes (2 x uint4).
template <class C>
__device__ __forceinline__ void ld_f(Fp<C>& x, const uint4* p) {
  uint4 a = p[0], b = p[1];
  x = {{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}};
}
template <bool INL, class B>
__device__ __forceinline__ void ld_f(Fq2T<INL, B>& x, const uint4* p) {
  ld_f(x.c0, p);
  ld_f(x.c1, p + 2);
}
template <class C>
__device__ __forceinline__ void st_f(uint4* p, const Fp<C>& x) {
  p[0] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  p[1] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}
template <bool INL, class B>
__device__ __forceinline__ void st_f(uint4* p, const Fq2T<INL, B>& x) {
  st_f(p, x.c0);
  st_f(p + 2, x.c1);
}

template <class F> struct Words { static constexpr int N = sizeof(F) / 16; };  // uint4s per F

template <class F>
__device__ __forceinline__ Aff<F> ld_aff(const uint4* base, uint32_t i) {
  constexpr int W = Words<F>::N;
  const uint4* p = base + (size_t)i * 2 * W;
  Aff<F> a;
  ld_f(a.x, p);
  ld_f(a.y, p + W);
  return a;
}

template <class F>
__device__ __forceinline__ void st_aff(uint4* base, size_t i, const Aff<F>& a) {
  constexpr int W = Words<F>::N;
  uint4* p = base + i * 2 * W;
  st_f(p, a.x);
  st_f(p + W, a.y);
}

template <class F>
__device__ __forceinline__ Xyzz<F> ld_xyzz(const uint4* base, size_t i) {
  constexpr int W = Words<F>::N;
  const uint4* p = base + i * 4 * W;
  Xyzz<F> r;
  ld_f(r.X, p);
  ld_f(r.Y, p + W);
  ld_f(r.ZZ, p + 2 * W);
  ld_f(r.ZZZ, p + 3 * W);
  return r;
}

template <class F>
__device__ __forceinline__ void st_xyzz(uint4* base, size_t i, const Xyzz<F>& r) {
  constexpr int W = Words<F>::N;
  uint4* p = base + i * 4 * W;
  st_f(p, r.X);
  st_f(p + W, r.Y);
  st_f(p + 2 * W, r.ZZ);
  st_f(p + 3 * W, r.ZZZ);
}

// How the merge and reduction kernels see a point: XYZZ with the exceptional-case shortcuts
// (XOps), or, for constant-work MSMs, (X : Y : Z) with the complete formulas (POps in words;
// COps below runs them on limbs), stored in an XYZZ-sized slot as (X, Y, Z, 0). `out` is the
// form the host reads (XYZZ).
template <class F>
struct XOps {
  using P = Xyzz<F>;
  static __device__ __forceinline__ P inf() { return xyzz_inf<F>(); }
  static __device__ __forceinline__ P add(const P& a, const P& b) { return xyzz_add(a, b); }
  static __device__ __forceinline__ P load(const uint4* base, size_t i) { return ld_xyzz<F>(base, i); }
  static __device__ __forceinline__ void store(uint4* base, size_t i, const P& p) { st_xyzz(base, i, p); }
  static __device__ __forceinline__ Xyzz<F> out(const P& p) { return p; }
};
template <class F>
struct POps {
  using P = Proj<F>;
  static __device__ __forceinline__ P inf() { return proj_inf<F>(); }
  static __device__ __forceinline__ P add(const P& a, const P& b) { return proj_add(a, b); }
  static __device__ __forceinline__ P load(const uint4* base, size_t i) {
    const Xyzz<F> x = ld_xyzz<F>(base, i);
    return {x.X, x.Y, x.ZZ};
  }
  static __device__ __forceinline__ void store(uint4* base, size_t i, const P& p) {
    st_xyzz(base, i, Xyzz<F>{p.X, p.Y, p.Z, f_zero<F>()});
  }
  static __device__ __forceinline__ Xyzz<F> out(const P& p) { return proj_to_xyzz(p); }
  static __device__ __forceinline__ P sel(bool c, const P& a, const P& b) { return proj_sel(c, a, b); }
  static __device__ __forceinline__ P add_aff(const P& p, const Aff<F>& a) { return proj_add_aff(p, a); }
};

// The constant-work kernels' arithmetic: the complete formulas on 29-bit limbs (P1L for G1, P2L
// for G2; POps for any other field), loaded from and stored to the same (X, Y, Z, 0) slots
// (words below 2q).
template <class F>
struct COps : POps<F> {};
template <>
struct COps<Fm2i> {
  using P = P2L;
  static __device__ __forceinline__ P inf() { return p2l_inf(); }
  static __device__ __forceinline__ P add(const P& a, const P& b) { return p2l_add(a, b); }
  static __device__ __forceinline__ P load(const uint4* base, size_t i) {
    return p2l_from(POps<Fm2i>::load(base, i));
  }
  static __device__ __forceinline__ void store(uint4* base, size_t i, const P& p) {
    POps<Fm2i>::store(base, i, p2l_out(p));
  }
  static __device__ __forceinline__ Xyzz<Fm2i> out(const P& p) {
    return proj_to_xyzz(p2l_out(p));
  }
  static __device__ __forceinline__ P sel(bool c, const P& a, const P& b) { return p2l_sel(c, a, b); }
  static __device__ __forceinline__ P add_aff(const P& p, const Aff<Fm2i>& a) {
    F2L x, y;
    f2l_of(a.x, x);
    f2l_of(a.y, y);
    return p2l_add_aff(p, x, y);
  }
};
template <>
struct COps<Fm> {
  using P = P1L;
  static __device__ __forceinline__ P inf() { return p1l_inf(); }
  static __device__ __forceinline__ P add(const P& a, const P& b) { return p1l_add(a, b); }
  static __device__ __forceinline__ P load(const uint4* base, size_t i) {
    return p1l_from(POps<Fm>::load(base, i));
  }
  static __device__ __forceinline__ void store(uint4* base, size_t i, const P& p) {
    POps<Fm>::store(base, i, p1l_out(p));
  }
  static __device__ __forceinline__ Xyzz<Fm> out(const P& p) { return proj_to_xyzz(p1l_out(p)); }
  static __device__ __forceinline__ P sel(bool c, const P& a, const P& b) { return p1l_sel(c, a, b); }
  static __device__ __forceinline__ P add_aff(const P& p, const Aff<Fm>& a) {
    uint32_t x[9], y[9];
    limbs29_of(a.x.v, x);
    limbs29_of(a.y.v, y);
    return p1l_add_aff(p, x, y);
  }
};

// The tails' additions (merges, heavy buckets, reduction): XOps, or for G1 and G2 the limb
// forms (LOps): points are loaded into limbs, added there, and packed back to the XYZZ words
// (X reduced below 2q) when stored.
template <class F>
struct LOps : XOps<F> {};
template <>
struct LOps<Fm> {
  using P = G1L;
  static __device__ __forceinline__ P inf() { return g1l_inf(); }
  static __device__ __forceinline__ P add(const P& a, const P& b) { return g1l_add(a, b); }
  static __device__ __forceinline__ P load(const uint4* base, size_t i) {
    return g1l_from(ld_xyzz<Fm>(base, i));
  }
  static __device__ __forceinline__ void store(uint4* base, size_t i, const P& p) {
    st_xyzz(base, i, g1l_out(p));
  }
  static __device__ __forceinline__ Xyzz<Fm> out(const P& p) { return g1l_out(p); }
};
template <>
struct LOps<Fm2i> {
  using P = G2L;
  static __device__ __forceinline__ P inf() { return g2l_inf(); }
  static __device__ __forceinline__ P add(const P& a, const P& b) { return g2l_add(a, b); }
  static __device__ __forceinline__ P load(const uint4* base, size_t i) {
    return g2l_from(ld_xyzz<Fm2i>(base, i));
  }
  static __device__ __forceinline__ void store(uint4* base, size_t i, const P& p) {
    st_xyzz(base, i, g2l_out(p));
  }
  static __device__ __forceinline__ Xyzz<Fm2i> out(const P& p) { return g2l_out(p); }
};

// Window tables for merged MSMs (msm_window_tables): entry w n + i = 2^(c w) P_i, affine, from
// the previous window's point by c doublings and one inversion; infinity stays (0, 0).
template <class F>
__global__ void __launch_bounds__(128) k_window_tables(uint4* table, size_t n, int c, int W) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Aff<F> P = ld_aff<F>(table, (uint32_t)i);
  for (int w = 1; w < W; ++w) {
    if (!aff_is_inf(P)) {
      Xyzz<F> X = xyzz_dbl_aff(P);
      for (int k = 1; k < c; ++k) X = xyzz_dbl(X);
      P = xyzz_to_aff(X);
    }
    st_aff(table, (size_t)w * n + i, P);
  }
}

// ---------------------------------------------------------------------------------------
// bucket accumulation
// ---------------------------------------------------------------------------------------
// The bucket loop's running sum. Generic: an XYZZ point. G1 (Fm) and G2 (Fm2i): the limb
// forms G1L and G2L, whose mixed additions never leave 29-bit limbs (bn254.cuh); the value and
// every output are the same.
template <class F>
struct Acc {
  using T = Xyzz<F>;
  static __device__ __forceinline__ T inf() { return xyzz_inf<F>(); }
  static __device__ __forceinline__ T madd(const T& p, const Aff<F>& a, bool negate,
                                           const uint4*, uint32_t) {
    return xyzz_add_aff(p, negate ? aff_neg(a) : a);
  }
  static __device__ __forceinline__ Xyzz<F> out(const T& p) { return p; }
};

template <>
struct Acc<Fm> {
  using T = G1L;
  static __device__ __forceinline__ T inf() { return g1l_inf(); }
  static __device__ __forceinline__ T madd(const T& p, const Aff<Fm>& a, bool negate,
                                           const uint4*, uint32_t) {
    return g1l_madd(p, a, negate);
  }
  static __device__ __forceinline__ Xyzz<Fm> out(const T& p) { return g1l_out(p); }
};

template <>
struct Acc<Fm2i> {
  using T = G2L;
  static __device__ __forceinline__ T inf() { return g2l_inf(); }
  // The doubling case reloads the base (entry i of `bases`) instead of keeping it live.
  static __device__ __forceinline__ T madd(const T& p, const Aff<Fm2i>& a, bool negate,
                                           const uint4* bases, uint32_t i) {
    if (aff_is_inf(a)) return p;
    if (f2l_zero(p.ZZ)) return g2l_from_aff(a, negate);
    G2L r;
    const int k = g2l_madd_core(p, a, negate, r);
    if (k == 1) {
      const Aff<Fm2i> b = ld_aff<Fm2i>(bases, i);
      return g2l_from(xyzz_dbl_aff(negate ? aff_neg(b) : b));
    }
    if (k == 2) return g2l_inf();
    return r;
  }
  static __device__ __forceinline__ Xyzz<Fm2i> out(const T& p) { return g2l_out(p); }
};

// Segment t of the sorted entries: its runs added into XYZZ accumulators, a run whole inside
// the segment stored as its bucket, the end pieces to pfirst / plast.
template <class F>
__device__ __forceinline__ void accumulate_segment(size_t t, const uint2* __restrict__ ent,
                                                   const uint32_t* total, size_t nseg,
                                                   uint32_t seg,
                                                   const uint4* __restrict__ bases,
                                                   uint4* buckets, uint4* pfirst, uint4* plast) {
  if (t >= nseg) return;
  uint32_t T = total[0];
  size_t b = t * seg;
  if (b >= T) return;
  size_t e = b + seg < T ? b + seg : T;
  uint32_t cur = ent[b].x;
  bool starts = (b == 0) || ent[b - 1].x != cur;
  size_t rs = b;
  using A = Acc<F>;
  typename A::T acc = A::inf();
  // The next entry's point is loaded before this one is added, so the load latency of a
  // random base hides behind a mixed addition.
  uint32_t v = ent[b].y;
  Aff<F> a = ld_aff<F>(bases, v >> 1);  // merged: v >> 1 = w n + i, a window-table entry
  for (size_t k = b; k < e; ++k) {
    uint32_t key = ent[k].x;
    if (key != cur) {
      // Run [rs, k) of bucket `cur` ends inside the segment.
      if (starts) st_xyzz(buckets, cur, A::out(acc));
      else st_xyzz(pfirst, t, A::out(acc));  // rs == b: the run began in an earlier segment
      acc = A::inf();
      cur = key;
      starts = true;
      rs = k;
    }
    uint32_t v_next = 0;
    Aff<F> a_next;
    if (k + 1 < e) {
      v_next = ent[k + 1].y;
      a_next = ld_aff<F>(bases, v_next >> 1);
    }
    acc = A::madd(acc, a, v & 1, bases, v >> 1);
    v = v_next;
    a = a_next;
  }
  bool ends = (e == T) || ent[e].x != cur;
  const Xyzz<F> sum = A::out(acc);
  if (starts && ends) {
    st_xyzz(buckets, cur, sum);
  } else {
    if (rs == b) st_xyzz(pfirst, t, sum);
    st_xyzz(plast, t, sum);
  }
}

template <class F>
__global__ void __launch_bounds__(128) k_accumulate(const uint2* __restrict__ ent,
                                                    const uint32_t* total, size_t nseg,
                                                    uint32_t seg,
                                                    const uint4* __restrict__ bases,
                                                    uint4* buckets, uint4* pfirst,
                                                    uint4* plast) {
  accumulate_segment<F>((size_t)blockIdx.x * blockDim.x + threadIdx.x, ent, total, nseg, seg,
                        bases, buckets, pfirst, plast);
}

// ---------------------------------------------------------------------------------------
// Constant work (opt in; crates/cuda/src/msm.rs:143-164): every (scalar, window) has an entry
// (a zero digit's is flagged discard), and from the accumulation to the reduction no decision
// is a branch on the data: points are (X : Y : Z) added by the complete formulas (bn254.cuh),
// loops run a count the plan fixes, and what the data decide is picked by selects after the
// work is done. What the data still decide is where stores land (a bucket's address), as
// with the counting sort before it. Not constant time.
//
// Accumulation: thread t runs exactly `seg` iterations over segment t (a short last segment
// repeats its last entry, added and dropped). Every iteration stores the running sum once:
// to its bucket when a run that began inside the segment ends, to slot 2t (head) when the
// segment's first run ends, else to slot 2t + 1 (tail), which the end overwrites with the
// last run. Slot rows hold the bucket key, or FOLD_NONE.
// ---------------------------------------------------------------------------------------
static constexpr uint32_t FOLD_NONE = 0xFFFFFFFFu;
// Slots a fold thread walks. A group emits at most two slots (its first and last run), so a
// level shrinks the slot count only for FOLD_LEN > 2; 8 measured best (G1 MSM 2^20 3.08 ms,
// G2 2^18 4.30 ms, against 3.13 / 4.38 at 4 and 3.19 / 4.59 at 16).
static constexpr uint32_t FOLD_LEN = 8;
static_assert(FOLD_LEN > 2, "each fold level must shrink the slot count");

template <class F>
__global__ void __launch_bounds__(128) k_accumulate_const(const uint2* __restrict__ ent,
                                                          const uint32_t* total, size_t nseg,
                                                          uint32_t seg,
                                                          const uint4* __restrict__ bases,
                                                          uint4* buckets, uint4* slots,
                                                          uint32_t* rows) {
  using O = COps<F>;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nseg) return;
  const size_t b = t * seg, T = total[0];  // the plan's entry count: b < T for t < nseg
  if (b >= T) {  // not reached on a full entry list; keeps every read inside it regardless
    rows[2 * t] = FOLD_NONE;
    rows[2 * t + 1] = FOLD_NONE;
    return;
  }
  const uint32_t len = (uint32_t)(b + seg < T ? seg : T - b);
  uint32_t cur = ent[b].x;
  bool first = true;
  const typename O::P inf = O::inf();
  typename O::P acc = inf;
  for (uint32_t i = 0; i < seg; ++i) {
    const bool live = i < len;
    const uint2 e = ent[b + (live ? i : len - 1)];
    const uint32_t key = live ? e.x : cur;
    const bool boundary = key != cur;
    const bool to_bucket = boundary && !first;
    const size_t slot = 2 * t + ((boundary && first) ? 0 : 1);
    O::store(to_bucket ? buckets : slots, to_bucket ? cur : slot, acc);
    rows[slot] = cur;
    first = first && !boundary;
    acc = O::sel(boundary, inf, acc);
    cur = key;
    Aff<F> a = ld_aff<F>(bases, e.y >> 2);
    const bool keep = !live || (e.y & 2u) || aff_is_inf(a);
    a.y = f_sel((e.y & 1u) != 0, neg(a.y), a.y);
    const typename O::P sum = O::add_aff(acc, a);
    acc = O::sel(keep, acc, sum);
  }
  const size_t slot = 2 * t + (first ? 0 : 1);
  O::store(slots, slot, acc);
  rows[slot] = cur;
  rows[2 * t + 1] = first ? FOLD_NONE : cur;
}

// One level of the fold: thread g walks slots [g FOLD_LEN, (g + 1) FOLD_LEN) of the level's m
// (exactly FOLD_LEN iterations; past m it repeats the last slot as empty), adding runs of equal
// rows. A run that ends inside the group, other than its first, is complete (all of a bucket's
// pieces are consecutive) and goes to its bucket; the group's first and last runs go to slots
// 2g and 2g + 1 of the next level (k_scan_const finishes the last few hundred slots).
template <class F>
__global__ void __launch_bounds__(128) k_fold_const(const uint4* in, const uint32_t* in_rows,
                                                    uint32_t m, uint4* out, uint32_t* out_rows,
                                                    uint4* buckets) {
  using O = COps<F>;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t groups = (m + FOLD_LEN - 1) / FOLD_LEN;
  if (g >= groups) return;
  const uint32_t lo = g * FOLD_LEN, hi = lo + FOLD_LEN < m ? lo + FOLD_LEN : m;
  uint32_t cur = FOLD_NONE;
  bool first = true;
  const typename O::P inf = O::inf();
  typename O::P acc = inf;
  for (uint32_t i = lo; i < lo + FOLD_LEN; ++i) {
    const bool live = i < hi;
    const uint32_t j = live ? i : hi - 1;
    const uint32_t r = live ? in_rows[j] : FOLD_NONE;
    const bool empty = r == FOLD_NONE;
    const bool boundary = !empty && r != cur;
    const bool finished = boundary && cur != FOLD_NONE;
    const bool to_head = finished && first;
    const bool to_bucket = finished && !to_head;
    const size_t slot = 2 * (size_t)g + (to_head ? 0 : 1);
    O::store(to_bucket ? buckets : out, to_bucket ? cur : slot, acc);
    out_rows[slot] = cur;
    first = first && !finished;
    acc = O::sel(boundary, inf, acc);
    cur = boundary ? r : cur;
    const typename O::P q = O::sel(empty, inf, O::load(in, j));
    acc = O::add(acc, q);
  }
  const bool has_run = cur != FOLD_NONE;
  const size_t slot = 2 * (size_t)g + (first ? 0 : 1);
  O::store(out, slot, acc);
  out_rows[slot] = has_run ? cur : FOLD_NONE;
  out_rows[2 * (size_t)g + 1] = (has_run && !first) ? cur : FOLD_NONE;
}

// The fold's end, for at most BLOCK slots: one workgroup, a slot per thread. The non-empty
// slots are packed to the front in order (their rows stay sorted, so each bucket's pieces are
// adjacent), then a segmented inclusive scan in LDS runs one addition per level for every
// thread: at distance d a thread adds the value d slots back where that slot has its row (the
// slots between belong to the same run) and keeps its own otherwise. The last slot of a run
// then holds the run's sum and stores it to its bucket. The level count depends on m only.
// BLOCK: slots per scan, as many threads as the point form's registers allow in one workgroup.
template <class F>
struct ScanSlots {
  static constexpr uint32_t N = sizeof(typename COps<F>::P) > 128 ? 256 : 512;
};

template <class F, uint32_t BLOCK>
__global__ void __launch_bounds__(BLOCK) k_scan_const(const uint4* in, const uint32_t* in_rows,
                                                      uint32_t m, uint4* buckets) {
  using O = COps<F>;
  using P = typename O::P;
  __shared__ P sh[BLOCK];
  __shared__ uint32_t rows[BLOCK];
  __shared__ uint32_t wave_count[BLOCK / 64];
  const uint32_t i = threadIdx.x, lane = i & 63u, wave = i >> 6;
  const bool live = i < m;
  const uint32_t j = live ? i : m - 1;  // m >= 1
  uint32_t r = live ? in_rows[j] : FOLD_NONE;
  const bool has = r != FOLD_NONE;
  P v = O::sel(has, O::load(in, j), O::inf());
  // Pack: position = non-empty slots before this one.
  const uint64_t ball = __ballot(has);
  const uint32_t before = (uint32_t)__popcll(ball & ((1ull << lane) - 1ull));
  if (lane == 0) wave_count[wave] = (uint32_t)__popcll(ball);
  __syncthreads();
  uint32_t pos = before, total = 0;
#pragma unroll
  for (uint32_t w = 0; w < BLOCK / 64; ++w) {
    pos += w < wave ? wave_count[w] : 0u;
    total += wave_count[w];
  }
  if (has) {
    sh[pos] = v;
    rows[pos] = r;
  }
  if (i >= total) {  // positions no slot packs into
    sh[i] = O::inf();
    rows[i] = FOLD_NONE;
  }
  __syncthreads();
  v = sh[i];
  r = rows[i];
  for (uint32_t d = 1; d < m; d <<= 1) {
    __syncthreads();  // every read of the previous level is done
    sh[i] = v;
    __syncthreads();
    const uint32_t k = i >= d ? i - d : i;
    const bool take = i >= d && rows[k] == r;
    const P sum = O::add(sh[k], v);
    v = O::sel(take, sum, v);
  }
  const bool end = r != FOLD_NONE && (i + 1 == BLOCK || rows[i + 1] != r);
  if (end) O::store(buckets, r, v);
}

// Every bucket starts as the identity (a bucket no entry names is never stored).
template <class F>
__global__ void k_fill_proj_inf(uint4* buckets, size_t K) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < K) POps<F>::store(buckets, i, POps<F>::inf());
}

// Several MSMs' bucket accumulations as one persistent launch over a device-side work queue:
// the grid is sized to what the device holds at once, and each workgroup takes tiles of 128
// segments from a shared counter until every job's segments are done. Tile t belongs to job
// d with start[d] <= t < start[d + 1]; segments past a job's device-side entry count end at once.
static constexpr int QUEUE_JOBS = 4;
static constexpr uint32_t QUEUE_TILE = 128;  // segments per tile = threads per workgroup

struct AccQueueJob {
  const uint2* ent;
  const uint32_t* total;
  const uint4* bases;
  uint4 *buckets, *pfirst, *plast;
  uint64_t nseg;
  uint32_t seg;
};

struct AccQueue {
  AccQueueJob job[QUEUE_JOBS];
  uint32_t start[QUEUE_JOBS + 1];
  int n;
};

template <class F>
__global__ void __launch_bounds__(QUEUE_TILE) k_accumulate_queue(AccQueue q, uint32_t* counter) {
  __shared__ uint32_t tile;
  const uint32_t tiles = q.start[q.n];
  for (;;) {
    if (threadIdx.x == 0) tile = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t t = tile;
    __syncthreads();  // every lane has read `tile` before lane 0 overwrites it
    if (t >= tiles) return;
    int d = 0;
    while (d + 1 < q.n && t >= q.start[d + 1]) ++d;
    const AccQueueJob& j = q.job[d];
    accumulate_segment<F>((size_t)(t - q.start[d]) * QUEUE_TILE + threadIdx.x, j.ent, j.total,
                          j.nseg, j.seg, j.bases, j.buckets, j.pfirst, j.plast);
  }
}

// Buckets that span segments: one thread folds a short span; a heavy one (e.g. every scalar
// equal to 1, in bucket 0 of window 0) is cut into tasks of HEAVY_TASK interior segments
// that a two-level tree sums (k_heavy_tasks, k_heavy_finish): depth log2 instead of a chain.
// cnt[0] = heavy buckets, cnt[1] = tasks; heavy[h] = {bucket, first task, task count};
// tasks[t] = {first segment, end segment}.
static constexpr int HEAVY_TASK = 256;

template <class F>
__global__ void __launch_bounds__(256) k_merge(const uint32_t* offset, uint32_t K, uint4* buckets,
                                               const uint4* pfirst, const uint4* plast,
                                               uint32_t* cnt, uint32_t* heavy, uint32_t* tasks,
                                               uint32_t seg) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  uint32_t s = offset[k], e = offset[k + 1];
  if (s == e) {
    st_xyzz(buckets, k, xyzz_inf<F>());
    return;
  }
  uint32_t sa = s / seg, sb = (e - 1) / seg;
  if (sa == sb) return;  // stored whole by the accumulation
  if (sb - sa > HEAVY_SPAN) {
    uint32_t interior = sb - sa - 1;
    uint32_t ntask = (interior + HEAVY_TASK - 1) / HEAVY_TASK;
    uint32_t slot = atomicAdd(&cnt[0], 1u);
    uint32_t tb = atomicAdd(&cnt[1], ntask);
    heavy[3 * slot] = k;
    heavy[3 * slot + 1] = tb;
    heavy[3 * slot + 2] = ntask;
    for (uint32_t t = 0; t < ntask; ++t) {
      uint32_t lo = sa + 1 + t * HEAVY_TASK;
      tasks[2 * (tb + t)] = lo;
      tasks[2 * (tb + t) + 1] = lo + HEAVY_TASK < sb ? lo + HEAVY_TASK : sb;
    }
    return;
  }
  // One addition site: a G2 addition is over 100 KB of code.
  using O = LOps<F>;
  typename O::P acc = O::load(plast, sa);
  for (uint32_t t = sa + 1; t <= sb; ++t) acc = O::add(acc, O::load(pfirst, t));
  O::store(buckets, k, acc);
}

// Merged windows: every bucket holds many entries and spans many segments (11-32 at the
// 1.1M-constraint size), so a bucket gets MERGE_LANES lanes: each sums every MERGE_LANES-th of
// the bucket's segment pieces (plast of its first segment, pfirst of the rest), then the lanes
// of a bucket add up through LDS. 256-lane workgroups take 256 / MERGE_LANES buckets. A bucket
// past WIDE_SPAN segments (the ones of a boolean witness pile into one; a skewed witness's
// small values would make more) is handed to the heavy path instead, as k_merge does, so no
// lane folds more than WIDE_SPAN / MERGE_LANES pieces in a row. Four lanes measured best
// inside the prove (the GPU is full there, so the tree's extra additions cost more than the
// shorter chains save): 16 lanes 73, 8 lanes 76, 4 lanes 78, 2 lanes 77 proofs/s at the
// 1.1M-constraint size.
static constexpr int MERGE_LANES = 4;
static constexpr uint32_t WIDE_SPAN = 64;

template <class F>
__global__ void __launch_bounds__(256) k_merge_wide(const uint32_t* offset, uint32_t K,
                                                    uint4* buckets, const uint4* pfirst,
                                                    const uint4* plast, uint32_t* cnt,
                                                    uint32_t* heavy, uint32_t* tasks,
                                                    uint32_t seg) {
  using O = LOps<F>;
  __shared__ typename O::P sh[256];
  const uint32_t lane = threadIdx.x % MERGE_LANES;
  const uint32_t k = blockIdx.x * (256 / MERGE_LANES) + threadIdx.x / MERGE_LANES;
  typename O::P acc = O::inf();
  bool store = false;  // the bucket's sum is merged here
  uint32_t s = 0, e = 0;
  if (k < K) {
    s = offset[k];
    e = offset[k + 1];
    if (s != e) {
      const uint32_t sa = s / seg, sb = (e - 1) / seg;
      if (sb - sa > WIDE_SPAN) {
        if (lane == 0) {
          uint32_t interior = sb - sa - 1;
          uint32_t ntask = (interior + HEAVY_TASK - 1) / HEAVY_TASK;
          uint32_t slot = atomicAdd(&cnt[0], 1u);
          uint32_t tb = atomicAdd(&cnt[1], ntask);
          heavy[3 * slot] = k;
          heavy[3 * slot + 1] = tb;
          heavy[3 * slot + 2] = ntask;
          for (uint32_t t = 0; t < ntask; ++t) {
            uint32_t lo = sa + 1 + t * HEAVY_TASK;
            tasks[2 * (tb + t)] = lo;
            tasks[2 * (tb + t) + 1] = lo + HEAVY_TASK < sb ? lo + HEAVY_TASK : sb;
          }
        }
      } else if (sa != sb) {
        store = true;
        for (uint32_t t = sa + lane; t <= sb; t += MERGE_LANES)
          acc = O::add(acc, O::load(t == sa ? plast : pfirst, t));
      }
    }
  }
  sh[threadIdx.x] = acc;
  for (int off = MERGE_LANES / 2; off > 0; off >>= 1) {
    __syncthreads();
    if (lane < (uint32_t)off) sh[threadIdx.x] = O::add(sh[threadIdx.x], sh[threadIdx.x + off]);
  }
  __syncthreads();
  if (k < K && lane == 0) {
    if (s == e) st_xyzz(buckets, k, xyzz_inf<F>());
    else if (store) O::store(buckets, k, sh[threadIdx.x]);
  }
}

// Sum of BLOCK per-thread values in LDS, left in sh[0].
template <class Ops, int BLOCK>
__device__ __forceinline__ void block_tree(typename Ops::P* sh) {
  for (int off = BLOCK / 2; off > 0; off >>= 1) {
    __syncthreads();
    if ((int)threadIdx.x < off) sh[threadIdx.x] = Ops::add(sh[threadIdx.x], sh[threadIdx.x + off]);
  }
  __syncthreads();
}

// Level 1: each task's interior segment partials (one per thread) summed by an LDS tree.
template <class F, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_heavy_tasks(const uint32_t* cnt, const uint32_t* tasks,
                                                       const uint4* pfirst, uint4* task_sum) {
  static_assert(BLOCK == HEAVY_TASK, "one segment partial per thread");
  using O = LOps<F>;
  __shared__ typename O::P sh[BLOCK];
  uint32_t nt = cnt[1];
  for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
    uint32_t lo = tasks[2 * t], hi = tasks[2 * t + 1];
    uint32_t seg = lo + threadIdx.x;
    sh[threadIdx.x] = seg < hi ? O::load(pfirst, seg) : O::inf();
    block_tree<O, BLOCK>(sh);
    if (threadIdx.x == 0) O::store(task_sum, t, sh[0]);
    __syncthreads();
  }
}

// Level 2: a heavy bucket's task sums and its two end pieces (item nt: plast of the first
// segment, nt + 1: pfirst of the last), strided over the block, then a tree.
template <class F, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_heavy_finish(const uint32_t* offset,
                                                        const uint32_t* cnt,
                                                        const uint32_t* heavy,
                                                        const uint4* task_sum, uint4* buckets,
                                                        const uint4* pfirst, const uint4* plast,
                                                        uint32_t seg) {
  using O = LOps<F>;
  __shared__ typename O::P sh[BLOCK];
  uint32_t nh = cnt[0];
  for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
    uint32_t k = heavy[3 * h], tb = heavy[3 * h + 1], nt = heavy[3 * h + 2];
    uint32_t sa = offset[k] / seg, sb = (offset[k + 1] - 1) / seg;
    typename O::P acc = O::inf();
    for (uint32_t t = threadIdx.x; t < nt + 2; t += BLOCK) {
      const uint4* src = t < nt ? task_sum : (t == nt ? plast : pfirst);
      const size_t i = t < nt ? (size_t)tb + t : (t == nt ? sa : sb);
      acc = O::add(acc, O::load(src, i));
    }
    sh[threadIdx.x] = acc;
    block_tree<O, BLOCK>(sh);
    if (threadIdx.x == 0) O::store(buckets, k, sh[0]);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------
// window reduction: S_w = sum_j (j+1) B_{w,j}
// ---------------------------------------------------------------------------------------
// The buckets of a window fall in G groups of B = 2^lg. With T_g = sum_i B_{gB+i} and
// U_g = sum_i (i+1) B_{gB+i},
//   S_w = sum_g U_g + B sum_g g T_g = sum_g U_g + sum_k 2^(k+lg) Q_k,
//   Q_k = sum of T_g over the g whose bit k is set.
// Level 1 (one thread per group, 2B additions) makes T and U; level 2 is a plain sum per
// (window, row) — row 0 sums U, row k+1 is Q_k. A row's items are cut into S slices (S a
// power of two, at most 64, dividing every row's count), one workgroup each: strided
// per-thread sums and an LDS tree; with S > 1 the slice sums go to `out` = a scratch array
// and k_reduce_slices adds each row's S of them, so a long row (one bucket set of 2^15
// buckets: 8192 groups) costs two short trees rather than one long strided chain. Every
// chain is short; the doublings of the bit weights happen once, in the host's Horner pass
// over all windows.
template <class F, class Ops = LOps<F>>
__global__ void __launch_bounds__(128) k_reduce_groups(const uint4* buckets, uint32_t nb,
                                                       int W, uint32_t B, uint4* T,
                                                       uint4* U) {
  uint32_t G = nb / B;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint32_t)W * G) return;
  size_t base = (size_t)t * B;  // = w * nb + g * B
  typename Ops::P run = Ops::inf(), acc = Ops::inf();
  for (int i = (int)B - 1; i >= 0; --i) {
    run = Ops::add(run, Ops::load(buckets, base + i));
    acc = Ops::add(acc, run);
  }
  Ops::store(T, t, run);
  Ops::store(U, t, acc);
}

// Workgroup x = ((w rows + r) S + slice); writes out[x] (the host's XYZZ when S = 1, else a
// slice sum in the policy's form for k_reduce_slices).
template <class F, int BLOCK, class Ops = LOps<F>>
__global__ void __launch_bounds__(BLOCK) k_reduce_rows(const uint4* T, const uint4* U,
                                                       uint32_t G, uint32_t rows, uint32_t S,
                                                       uint4* out) {
  __shared__ typename Ops::P sh[BLOCK];
  const uint32_t sl = blockIdx.x % S, wr = blockIdx.x / S;
  const uint32_t w = wr / rows, r = wr % rows;
  size_t base = (size_t)w * G;
  typename Ops::P acc = Ops::inf();
  const uint4* src = r == 0 ? U : T;
  const uint32_t k = r == 0 ? 0 : r - 1, low = (1u << k) - 1u, count = r == 0 ? G : G / 2;
  const uint32_t per = count / S, m0 = sl * per;
  for (uint32_t m = m0 + threadIdx.x; m < m0 + per; m += BLOCK) {
    // row 0: every g; row k + 1: m with a 1 inserted at bit k
    const uint32_t g = r == 0 ? m : ((m & ~low) << 1) | (1u << k) | (m & low);
    acc = Ops::add(acc, Ops::load(src, base + g));
  }
  sh[threadIdx.x] = acc;
  block_tree<Ops, BLOCK>(sh);
  if (threadIdx.x == 0) {
    if (S == 1) st_xyzz(out, blockIdx.x, Ops::out(sh[0]));
    else Ops::store(out, blockIdx.x, sh[0]);
    __threadfence_system();  // `out` may be pinned host memory the host reads next
  }
}

// Row x = w rows + r: out[x] = sum of parts[x S .. x S + S), in the host's XYZZ.
template <class F, int BLOCK, class Ops = LOps<F>>
__global__ void __launch_bounds__(BLOCK) k_reduce_slices(const uint4* parts, uint32_t S,
                                                         uint4* out) {
  __shared__ typename Ops::P sh[BLOCK];
  typename Ops::P acc = Ops::inf();
  for (uint32_t i = threadIdx.x; i < S; i += BLOCK)
    acc = Ops::add(acc, Ops::load(parts, (size_t)blockIdx.x * S + i));
  sh[threadIdx.x] = acc;
  block_tree<Ops, BLOCK>(sh);
  if (threadIdx.x == 0) {
    st_xyzz(out, blockIdx.x, Ops::out(sh[0]));
    __threadfence_system();  // `out` may be pinned host memory the host reads next
  }
}

inline uint32_t reduce_group_size(const MsmPlan& p) { return p.nb < 4 ? p.nb : 4; }
inline int ilog2_u32(uint32_t x) {
  int l = 0;
  while ((2u << l) <= x) ++l;
  return l;
}

// Slices per row of G groups over `sets` bucket sets: rows hold G or G / 2 items, at least
// 512 and 256 per slice, and no more than 256 workgroups of four waves (one wave per SIMD):
// past that, waves share SIMDs and every tree level slows by as much as slicing saves. In
// practice only a single bucket set (merged windows) is sliced.
// Past 8192 groups (bucket sets of 2^15 and more: widths above 16) a row is cut into slices of
// about 2048 groups, up to 256 per row, so that the row sums run across the device instead of
// in a few long chains.
inline uint32_t reduce_slices(uint32_t G, uint32_t sets) {
  const uint32_t rows = 1 + (uint32_t)ilog2_u32(G);
  uint32_t S = 1;
  if (G > 8192) {
    while (S < 256 && S * 2048 < G) S *= 2;
    return S;
  }
  while (S < 64 && 2 * S * 512 <= G && (uint64_t)2 * S * rows * sets <= 256) S *= 2;
  return S;
}

// Scratch of one accumulation (one base vector over one sorted entry list): buckets, the
// segment end partials, heavy-bucket bookkeeping, task sums, reduction groups, counters.
struct AccumLayout {
  size_t buckets, pfirst, plast, heavy, tasks, task_sum, gT, gU, parts, cnt;
  size_t fold0, frows0, fold1, frows1;  // constant work: the fold's two levels of slots
  size_t total;
};

// Slots the constant-work fold needs: 2 per segment, then 2 per group of FOLD_LEN.
inline size_t fold_slots0(size_t nseg) { return 2 * nseg; }
inline size_t fold_slots1(size_t nseg) { return 2 * ((2 * nseg + FOLD_LEN - 1) / FOLD_LEN) + 2; }

inline AccumLayout accum_layout(const MsmPlan& p, size_t pt) {
  size_t K = (size_t)msm_bucket_sets(p) * p.nb;
  const uint32_t seg = accum_seg(p, pt);
  size_t nseg = (p.max_entries + seg - 1) / seg;
  size_t max_heavy = nseg / HEAVY_SPAN + 2;
  size_t max_tasks = nseg / HEAVY_TASK + max_heavy + 2;
  uint32_t B = reduce_group_size(p);
  size_t G = p.nb / B;
  AccumLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t r = off;
    off += accum_align_up(bytes);
    return r;
  };
  L.buckets = take(pt * K);
  L.pfirst = take(pt * (nseg + 1));
  L.plast = take(pt * (nseg + 1));
  L.heavy = take(12 * max_heavy);
  L.tasks = take(8 * max_tasks);
  L.task_sum = take(pt * max_tasks);
  L.gT = take(pt * G * msm_bucket_sets(p));
  L.gU = take(pt * G * msm_bucket_sets(p));
  const uint32_t S = reduce_slices((uint32_t)G, (uint32_t)msm_bucket_sets(p));
  L.parts = take(S > 1 ? pt * S * (1 + ilog2_u32((uint32_t)G)) * msm_bucket_sets(p) : 0);
  L.cnt = take(16);
  L.fold0 = take(p.constant ? pt * fold_slots0(nseg) : 0);
  L.frows0 = take(p.constant ? 4 * fold_slots0(nseg) : 0);
  L.fold1 = take(p.constant ? pt * fold_slots1(nseg) : 0);
  L.frows1 = take(p.constant ? 4 * fold_slots1(nseg) : 0);
  L.total = off;
  return L;
}

inline size_t msm_accum_scratch_bytes_impl(const MsmPlan& p, int group) {
  return accum_layout(p, group == 1 ? 128 : 256).total;
}

// Pointers into one accumulation's scratch.
struct AccumPtrs {
  uint4 *buckets, *pfirst, *plast, *task_sum, *gT, *gU, *parts;
  uint32_t *heavy, *tasks, *cnt;
  uint4 *fold0, *fold1;
  uint32_t *frows0, *frows1;
  uint32_t seg;  // entries per segment
  size_t nseg;
};

// Launchers of the merge and reduction kernels over field F. The G1 ones are defined here;
// the G2 ones (F = Fm2i, products expanded in place so the latency-bound tail does not spill)
// live in kernels_msm_g2_tail.hip and kernels_msm_g2_reduce.hip, so the slow G2 kernels
// compile in parallel translation units.
template <class F>
struct Tail {
  static void merge(const MsmPlan& p, const AccumPtrs& a, hipStream_t s);
  static void reduce(const MsmPlan& p, const AccumPtrs& a, uint32_t B, void* window_out,
                     hipStream_t s);
};

template <class F>
inline void tail_fold_const(const AccumPtrs& a, hipStream_t s) {
  // Fold levels (2 nseg slots, then 2 per group of FOLD_LEN) until one workgroup's scan holds
  // the rest.
  constexpr uint32_t SCAN = ScanSlots<F>::N;
  size_t m = fold_slots0(a.nseg);
  if (m == 0) return;
  const uint4* in = a.fold0;
  const uint32_t* in_rows = a.frows0;
  bool to1 = true;
  while (m > SCAN) {
    const size_t groups = (m + FOLD_LEN - 1) / FOLD_LEN;
    uint4* out = to1 ? a.fold1 : a.fold0;
    uint32_t* out_rows = to1 ? a.frows1 : a.frows0;
    hipLaunchKernelGGL(k_fold_const<F>, dim3((unsigned)((groups + 127) / 128)), dim3(128), 0, s,
                       in, in_rows, (uint32_t)m, out, out_rows, a.buckets);
    G16_LAUNCH_CHECK("msm fold (constant work)");
    in = out;
    in_rows = out_rows;
    m = 2 * groups;
    to1 = !to1;
  }
  hipLaunchKernelGGL((k_scan_const<F, SCAN>), dim3(1), dim3(SCAN), 0, s, in, in_rows,
                     (uint32_t)m, a.buckets);
  G16_LAUNCH_CHECK("msm fold scan (constant work)");
}

template <class F>
inline void tail_merge_impl(const MsmPlan& p, const AccumPtrs& a, hipStream_t s) {
  if (p.constant) {
    tail_fold_const<F>(a, s);
    return;
  }
  size_t K = (size_t)msm_bucket_sets(p) * p.nb;
  hipLaunchKernelGGL(k_clear_cnt<0>, dim3(1), dim3(64), 0, s, a.cnt);
  G16_LAUNCH_CHECK("msm clear counters");
  // Merged windows give a bucket many segments in a large MSM (11-32 at the 1.1M-constraint
  // size: MERGE_LANES lanes each); a wave of small proofs gives most buckets one segment or
  // none, where lanes that mostly exit at once only cost launches: one thread per bucket then.
  const size_t avg_segs = p.max_entries / ((size_t)a.seg * (K ? K : 1));
  if (p.merged && avg_segs >= 4) {
    constexpr unsigned per_block = 256 / MERGE_LANES;
    hipLaunchKernelGGL(k_merge_wide<F>, dim3((unsigned)((K + per_block - 1) / per_block)),
                       dim3(256), 0, s, p.offset, (uint32_t)K, a.buckets, a.pfirst, a.plast,
                       a.cnt, a.heavy, a.tasks, a.seg);
    G16_LAUNCH_CHECK("msm merge (wide)");
  } else {
    hipLaunchKernelGGL(k_merge<F>, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, s, p.offset,
                       (uint32_t)K, a.buckets, a.pfirst, a.plast, a.cnt, a.heavy, a.tasks, a.seg);
    G16_LAUNCH_CHECK("msm merge");
  }
  hipLaunchKernelGGL((k_heavy_tasks<F, HEAVY_TASK>), dim3(1024), dim3(HEAVY_TASK), 0, s, a.cnt,
                     a.tasks, a.pfirst, a.task_sum);
  G16_LAUNCH_CHECK("msm heavy tasks");
  hipLaunchKernelGGL((k_heavy_finish<F, 64>), dim3(256), dim3(64), 0, s, p.offset, a.cnt, a.heavy,
                     a.task_sum, a.buckets, a.pfirst, a.plast, a.seg);
  G16_LAUNCH_CHECK("msm heavy finish");
}

template <class F, class Ops>
inline void tail_reduce_ops(const MsmPlan& p, const AccumPtrs& a, uint32_t B, void* window_out,
                            hipStream_t s) {
  uint32_t G = p.nb / B;
  uint32_t rows = 1 + (uint32_t)ilog2_u32(G);
  const int sets = msm_bucket_sets(p);
  uint32_t nthreads = (uint32_t)sets * G;
  hipLaunchKernelGGL((k_reduce_groups<F, Ops>), dim3((nthreads + 127) / 128), dim3(128), 0, s,
                     a.buckets, p.nb, sets, B, a.gT, a.gU);
  G16_LAUNCH_CHECK("msm reduce groups");
  const uint32_t S = reduce_slices(G, (uint32_t)sets);
  uint4* wout = static_cast<uint4*>(window_out);
  hipLaunchKernelGGL((k_reduce_rows<F, 256, Ops>), dim3((unsigned)sets * rows * S), dim3(256), 0,
                     s, a.gT, a.gU, G, rows, S, S > 1 ? a.parts : wout);
  G16_LAUNCH_CHECK("msm reduce rows");
  if (S > 1) {
    hipLaunchKernelGGL((k_reduce_slices<F, 64, Ops>), dim3((unsigned)sets * rows), dim3(64), 0, s,
                       a.parts, S, wout);
    G16_LAUNCH_CHECK("msm reduce slices");
  }
}

// Constant-work plans reduce their projective buckets with the complete formulas.
template <class F>
inline void tail_reduce_impl(const MsmPlan& p, const AccumPtrs& a, uint32_t B, void* window_out,
                             hipStream_t s) {
  if (p.constant) tail_reduce_ops<F, COps<F>>(p, a, B, window_out, s);
  else tail_reduce_ops<F, LOps<F>>(p, a, B, window_out, s);
}

// F: the coordinate field in the MSM's Montgomery form R' = 2^261 (Fm, or Fm2i for G2); the
// bases must already be in that form (msm_bases_to_msm_form).
template <class F>
inline AccumPtrs accum_ptrs(const MsmPlan& p, void* scratch) {
  constexpr size_t PT = sizeof(Xyzz<F>);
  AccumLayout L = accum_layout(p, PT);
  char* q = static_cast<char*>(scratch);
  AccumPtrs a;
  a.buckets = reinterpret_cast<uint4*>(q + L.buckets);
  a.pfirst = reinterpret_cast<uint4*>(q + L.pfirst);
  a.plast = reinterpret_cast<uint4*>(q + L.plast);
  a.heavy = reinterpret_cast<uint32_t*>(q + L.heavy);
  a.tasks = reinterpret_cast<uint32_t*>(q + L.tasks);
  a.task_sum = reinterpret_cast<uint4*>(q + L.task_sum);
  a.gT = reinterpret_cast<uint4*>(q + L.gT);
  a.gU = reinterpret_cast<uint4*>(q + L.gU);
  a.parts = reinterpret_cast<uint4*>(q + L.parts);
  a.cnt = reinterpret_cast<uint32_t*>(q + L.cnt);
  a.fold0 = reinterpret_cast<uint4*>(q + L.fold0);
  a.frows0 = reinterpret_cast<uint32_t*>(q + L.frows0);
  a.fold1 = reinterpret_cast<uint4*>(q + L.fold1);
  a.frows1 = reinterpret_cast<uint32_t*>(q + L.frows1);
  a.seg = accum_seg(p, PT);
  a.nseg = (p.max_entries + a.seg - 1) / a.seg;
  return a;
}

// The G1 bucket accumulations of up to QUEUE_JOBS MSMs (each: plan, bases, scratch) as one
// persistent launch, then each job's merges and reduction into its window_out. `counter` is a
// device word the launch takes tiles from (cleared here by a kernel first).
template <class F>
inline void accumulate_queue_impl(const MsmPlan* const* plans, const void* const* bases,
                                  void* const* scratch, void* const* window_out, int n,
                                  uint32_t* counter, int grid, hipStream_t s) {
  if (n < 1 || n > QUEUE_JOBS) throw ArgError{"accumulation queue of " + std::to_string(n) + " jobs"};
  for (int d = 0; d < n; ++d)
    if (plans[d]->constant) throw ArgError{"constant-work MSMs do not go through the queue"};
  AccQueue q{};
  q.n = n;
  q.start[0] = 0;
  AccumPtrs a[QUEUE_JOBS];
  for (int d = 0; d < n; ++d) {
    a[d] = accum_ptrs<F>(*plans[d], scratch[d]);
    const size_t nseg = (plans[d]->max_entries + a[d].seg - 1) / a[d].seg;
    q.job[d] = AccQueueJob{plans[d]->ent, plans[d]->total, static_cast<const uint4*>(bases[d]),
                           a[d].buckets, a[d].pfirst, a[d].plast, nseg, a[d].seg};
    const size_t tiles = (nseg + QUEUE_TILE - 1) / QUEUE_TILE;
    if (q.start[d] + tiles >= (1ull << 32)) throw ArgError{"accumulation queue too long"};
    q.start[d + 1] = q.start[d] + (uint32_t)tiles;
  }
  hipLaunchKernelGGL(k_clear_cnt<0>, dim3(1), dim3(64), 0, s, counter);
  G16_LAUNCH_CHECK("msm queue counter");
  if (q.start[n]) {
    const unsigned blocks = (unsigned)std::min<size_t>((size_t)grid, q.start[n]);
    hipLaunchKernelGGL(k_accumulate_queue<F>, dim3(blocks), dim3(QUEUE_TILE), 0, s, q, counter);
    G16_LAUNCH_CHECK("msm accumulate (queue)");
  }
  for (int d = 0; d < n; ++d) {
    Tail<F>::merge(*plans[d], a[d], s);
    Tail<F>::reduce(*plans[d], a[d], reduce_group_size(*plans[d]), window_out[d], s);
  }
}

template <class F>
inline void accumulate_impl(const MsmPlan& p, const void* bases, void* scratch, void* window_out,
                            hipStream_t s, hipEvent_t* ev) {
  const AccumPtrs a = accum_ptrs<F>(p, scratch);
  const uint32_t B = reduce_group_size(p);
  auto bp = static_cast<const uint4*>(bases);
  const unsigned grid = (unsigned)((a.nseg + 127) / 128);
  if (ev) G16_CHECK(hipEventRecord(ev[0], s));
  if (p.constant) {
    const size_t K = (size_t)msm_bucket_sets(p) * p.nb;
    hipLaunchKernelGGL(k_fill_proj_inf<F>, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, s,
                       a.buckets, K);
    G16_LAUNCH_CHECK("msm buckets (constant work)");
    if (a.nseg) {
      hipLaunchKernelGGL(k_accumulate_const<F>, dim3(grid), dim3(128), 0, s, p.ent, p.total,
                         a.nseg, a.seg, bp, a.buckets, a.fold0, a.frows0);
      G16_LAUNCH_CHECK("msm accumulate (constant work)");
    }
  } else if (a.nseg) {
    hipLaunchKernelGGL(k_accumulate<F>, dim3(grid), dim3(128), 0, s, p.ent, p.total, a.nseg,
                       a.seg, bp, a.buckets, a.pfirst, a.plast);
    G16_LAUNCH_CHECK("msm accumulate");
  }
  if (ev) G16_CHECK(hipEventRecord(ev[1], s));
  Tail<F>::merge(p, a, s);
  if (ev) G16_CHECK(hipEventRecord(ev[2], s));
  Tail<F>::reduce(p, a, B, window_out, s);
  if (ev) G16_CHECK(hipEventRecord(ev[3], s));
}

}  // namespace g16
