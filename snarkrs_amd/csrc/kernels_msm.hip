// Pippenger MSM for G1 and G2 (stages 5-9, crates/groth16/src/cpu.rs:477-632).
//
// Pipeline per scalar vector (shared by every base vector multiplied by the same scalars,
// so the witness MSMs A, B1, B2 and L sort once):
//   1. digits:  Montgomery -> standard scalar, carry-free signed digits per window
//               (crates/msm/src/lib.rs:275-331), stored window-major (16-bit up to c = 15);
//   2. sort:    a two-level counting sort of (set, bucket) keys, coarse bins then fine bins,
//               whose writes all land in runs (see "The sort has two levels" below);
//   3. accumulate: one thread per fixed-length segment of the sorted entries adds its runs
//               into XYZZ accumulators (madd-2008-s); a run that fills its bucket inside
//               the segment is stored as the bucket, the partial runs at a segment's ends go
//               to side arrays and a merge pass folds them (heavy buckets, e.g. every 1
//               scalar in bucket 0 of window 0, are cut into tasks summed by LDS trees);
//   4. reduce:  per window S_w = sum_j (j+1) B_j, as group sums and weighted group sums of
//               8 buckets, then one LDS tree per (window, bit of the group index); the host
//               finishes with one Horner pass over every window and bit.
// Zero scalars produce no entries and a scalar equal to 1 produces one (window 0, bucket 0),
// which is the reference prescan's effect (crates/msm/src/lib.rs:534-690) without a pass of
// its own; bases at infinity are skipped where they are loaded.
#include "msm_accum.cuh"

namespace g16 {

static constexpr int RECODE_BITS = 255;

// ---------------------------------------------------------------------------------------
// digits and the counting sort
// ---------------------------------------------------------------------------------------
// s[idx] by selects (idx is data-dependent; indexing the register array would put it in
// scratch memory). Limbs past the top read as zero.
__device__ __forceinline__ uint32_t limb(const uint32_t s[8], int idx) {
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) r = idx == k ? s[k] : r;
  return r;
}

__device__ __forceinline__ uint32_t read_bits(const uint32_t s[8], int off, int width) {
  int idx = off >> 5;
  int sh = off & 31;
  uint64_t buf = limb(s, idx) | ((uint64_t)limb(s, idx + 1) << 32);
  return (uint32_t)(buf >> sh) & ((1u << width) - 1);
}

// Digit i of the width-c signed recoding, in [-2^(c-1), 2^(c-1)].
__device__ __forceinline__ int signed_digit(const uint32_t s[8], int i, int c) {
  int off = i * c;
  int b = (int)read_bits(s, off, c);
  int d = b - ((b >> (c - 1)) << c);
  int carry = off == 0 ? 0 : (int)read_bits(s, off - 1, 1);
  return d + carry;
}

// Digit codes, one per (window, scalar), window-major. Two widths:
//   32-bit (c >= 16, or constant work): (|d| - 1) | sign << 24, DIGIT_ZERO for d = 0; under
//     constant work a zero digit is (i mod nb) | DIGIT_DISCARD instead, an entry whose addition
//     is computed and dropped;
//   16-bit (c <= 15, variable work): (|d| - 1) | sign << 15, 0x7FFF for d = 0 (|d| - 1 is at
//     most 2^14 - 1 there), half the digit traffic.
static constexpr uint32_t DIGIT_SIGN = 1u << 24;
static constexpr uint32_t DIGIT_ZERO = 1u << 25;
static constexpr uint32_t DIGIT_DISCARD = 1u << 26;

template <class D> struct Digit;
template <> struct Digit<uint32_t> {
  static __device__ __forceinline__ uint32_t make(int d, uint32_t zero) {
    return d == 0 ? zero : ((uint32_t)(d < 0 ? -d : d) - 1u) | (d < 0 ? DIGIT_SIGN : 0u);
  }
  static __device__ __forceinline__ bool zero(uint32_t e) { return (e & DIGIT_ZERO) != 0; }
  static __device__ __forceinline__ uint32_t none() { return DIGIT_ZERO; }
  static __device__ __forceinline__ uint32_t bucket(uint32_t e) { return e & (DIGIT_SIGN - 1u); }
  static __device__ __forceinline__ uint32_t sign(uint32_t e) { return (e >> 24) & 1u; }
  static __device__ __forceinline__ uint32_t discard(uint32_t e) {
    return (e & DIGIT_DISCARD) ? 2u : 0u;
  }
};
template <> struct Digit<uint16_t> {
  static __device__ __forceinline__ uint16_t make(int d, uint32_t) {
    return (uint16_t)(d == 0 ? 0x7FFFu
                             : ((uint32_t)(d < 0 ? -d : d) - 1u) | (d < 0 ? 1u << 15 : 0u));
  }
  static __device__ __forceinline__ bool zero(uint16_t e) { return e == 0x7FFFu; }
  static __device__ __forceinline__ uint16_t none() { return (uint16_t)0x7FFFu; }
  static __device__ __forceinline__ uint32_t bucket(uint16_t e) { return e & 0x7FFFu; }
  static __device__ __forceinline__ uint32_t sign(uint16_t e) { return (uint32_t)e >> 15; }
  static __device__ __forceinline__ uint32_t discard(uint16_t) { return 0; }
};

// Whether a plan's digits are 16-bit.
static bool short_digits(const MsmPlan& p) { return p.c <= 15 && !p.constant; }

// The sort has two levels, so that every write lands in a run instead of at a random place:
//   1. coarse: each (chunk of SORT_CHUNK scalars, digit row) workgroup counts its digits per
//      coarse bin (bucket >> shift) in LDS; one exclusive scan over all the counts, laid out
//      group-major ([set][coarse bin][row of the set][chunk]), gives every (chunk, row, bin) its
//      run in a staging array and every coarse group its range; the scatter writes each entry
//      (fine bits, value) into its run;
//   2. fine: each coarse group is cut into slices of FINE_TILE entries; a workgroup per slice
//      counts its entries per fine bin (bucket & mask), a workgroup per group scans its
//      slices' counts into positions and bucket offsets, and a workgroup per slice writes its
//      entries with their full keys, in order of bucket, inside the group's range.
// A set is a digit row (per-window buckets) or a vector's W rows (merged windows). Entries of a
// bucket come out in any order; its sum does not depend on it.
static constexpr uint32_t SORT_CHUNK = 65536;  // scalars per coarse workgroup (at least)
static constexpr int SORT_THREADS = 1024;
static constexpr int SORT_ITEMS = 8;           // digits per thread per round (loads first)
static constexpr int FINE_THREADS = 256;
static constexpr int FINE_ITEMS = 16;
static constexpr uint32_t FINE_TILE = FINE_THREADS * FINE_ITEMS;  // entries per fine slice
static constexpr uint32_t SCAN_BLOCK = 4096;   // elements per scan workgroup (256 x 16)
static constexpr uint32_t MAX_CBINS = 512;     // coarse bins per set (LDS counters)
static constexpr int MAX_FINE_SHIFT = 12;      // fine bins per coarse group <= 4096

// ctr[key] += 1 from every active lane, returning each lane's old value plus its rank among
// the lanes that share its key. When at least a quarter of the wave shares the first active
// lane's key (a heavy bucket: the ones of a boolean witness all land in bucket 0), those lanes
// take one atomic together instead of 16+ serialised ones on the same LDS word.
__device__ __forceinline__ uint32_t lds_take(uint32_t* ctr, uint32_t key) {
  const uint32_t lane = __lane_id();
  const uint64_t active = __ballot(1);
  const int first = __ffsll((unsigned long long)active) - 1;
  const uint32_t k0 = __shfl(key, first);
  const uint64_t same = __ballot(key == k0);
  if (__popcll(same) >= 16) {
    uint32_t base = 0;
    if ((int)lane == first) base = atomicAdd(&ctr[k0], (uint32_t)__popcll(same));
    base = __shfl(base, first);
    if (key == k0) return base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
  }
  return atomicAdd(&ctr[key], 1u);
}

// Block-wide exclusive prefix of one value per thread (blockDim.x = T, a multiple of 64);
// returns the thread's prefix and writes the block total to *sum.
template <int T>
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* sh, uint32_t* sum) {
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < T / 64; ++w) {
    const uint32_t t = sh[w];
    before += (uint32_t)w < wave ? t : 0u;
    all += t;
  }
  __syncthreads();  // sh may be reused by the caller
  *sum = all;
  return before + x - v;
}

// sprefix[g] = slices before group g, sprefix[groups] = all slices (one workgroup).
template <class D>
__global__ void __launch_bounds__(256) k_digits(const uint4* scalars, uint32_t n, uint32_t total,
                                                int c, int W, D* digits, int constant) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  // scalar i is index j of vector k: its digits go to rows (k W + w) of n
  const uint32_t k = i / n, j = i - k * n;
  Fr s = to_std(ld_fr(scalars + 2 * (size_t)i));
  const uint32_t zero = constant ? (j & ((1u << (c - 1)) - 1u)) | DIGIT_DISCARD : DIGIT_ZERO;
  D* row = digits + (size_t)k * W * n + j;
  for (int w = 0; w < W; ++w) row[(size_t)w * n] = Digit<D>::make(signed_digit(s.v, w, c), zero);
}

template <class D>
__global__ void __launch_bounds__(SORT_THREADS) k_coarse_count(const D* digits, uint32_t n,
                                                               uint32_t cs, int shift,
                                                               uint32_t cbins,
                                                               uint32_t rows_per_set,
                                                               uint32_t chunks,
                                                               uint32_t* ccount) {
  extern __shared__ uint32_t hist[];  // cbins words
  const uint32_t p = blockIdx.x, row = blockIdx.y;
  for (uint32_t j = threadIdx.x; j < cbins; j += SORT_THREADS) hist[j] = 0;
  __syncthreads();
  const size_t lo = (size_t)p * cs;
  const size_t hi = lo + cs < n ? lo + cs : n;
  const D* d = digits + (size_t)row * n;
  for (size_t base = lo; base < hi; base += SORT_THREADS * SORT_ITEMS) {
    D e[SORT_ITEMS];
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
      const size_t i = base + k * SORT_THREADS + threadIdx.x;
      e[k] = i < hi ? d[i] : Digit<D>::none();
    }
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k)
      if (!Digit<D>::zero(e[k])) lds_take(hist, Digit<D>::bucket(e[k]) >> shift);
  }
  __syncthreads();
  const uint32_t set = row / rows_per_set, r = row - set * rows_per_set;
  for (uint32_t j = threadIdx.x; j < cbins; j += SORT_THREADS)
    ccount[(((size_t)set * cbins + j) * rows_per_set + r) * chunks + p] = hist[j];
}

// Each entry of (chunk p, row) into its coarse bin's run of the staging array:
// tmp[pos] = {fine bits of the bucket, value}. The chunk goes through LDS in tiles of
// SORT_THREADS x SORT_ITEMS digits: a tile's entries are counted and ordered by coarse bin in
// LDS first, then written out with consecutive lanes on consecutive positions of each bin's
// run (a wave's stores cover a few runs instead of up to 64 scattered words).
// LDS: cur, lcnt, lstart (MAX_CBINS words each) and the tile's staged keys and values: 70 KB,
// two workgroups per CU.
static constexpr uint32_t SCATTER_TILE = SORT_THREADS * SORT_ITEMS;

template <class D>
__global__ void __launch_bounds__(SORT_THREADS) k_coarse_scatter(
    const D* digits, uint32_t n, uint32_t cs, int shift, uint32_t cbins, uint32_t rows_per_set,
    uint32_t chunks, int W, const uint32_t* cpos, uint2* tmp, int merged, int constant) {
  __shared__ uint32_t cur[MAX_CBINS];     // next global position of each bin's run
  __shared__ uint32_t lcnt[MAX_CBINS];    // the tile's entries per bin
  __shared__ uint32_t lstart[MAX_CBINS];  // their exclusive prefix
  __shared__ uint32_t skey[SCATTER_TILE]; // staged (bin << 12 | fine bits), in bin order
  __shared__ uint32_t sval[SCATTER_TILE];
  __shared__ uint32_t scan_sh[SORT_THREADS / 64];
  const uint32_t p = blockIdx.x, row = blockIdx.y;
  const uint32_t set = row / rows_per_set, r = row - set * rows_per_set;
  for (uint32_t j = threadIdx.x; j < cbins; j += SORT_THREADS)
    cur[j] = cpos[(((size_t)set * cbins + j) * rows_per_set + r) * chunks + p];
  const uint32_t win = row % (uint32_t)W;  // window of this row (rows are vector-major)
  const uint32_t mask = (1u << shift) - 1u;
  const size_t lo = (size_t)p * cs;
  const size_t hi = lo + cs < n ? lo + cs : n;
  const D* d = digits + (size_t)row * n;
  const uint32_t per = (cbins + SORT_THREADS - 1) / SORT_THREADS;  // bins per thread (0 or 1)
  for (size_t base = lo; base < hi; base += SCATTER_TILE) {
    for (uint32_t j = threadIdx.x; j < cbins; j += SORT_THREADS) lcnt[j] = 0;
    __syncthreads();
    D e[SORT_ITEMS];
    uint32_t rank[SORT_ITEMS];
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
      const size_t i = base + k * SORT_THREADS + threadIdx.x;
      e[k] = i < hi ? d[i] : Digit<D>::none();
    }
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k)
      rank[k] = Digit<D>::zero(e[k]) ? 0u : lds_take(lcnt, Digit<D>::bucket(e[k]) >> shift);
    __syncthreads();
    // lstart = exclusive prefix of lcnt over the bins, `per` consecutive bins per thread.
    const uint32_t j0 = threadIdx.x * per;
    uint32_t loc = 0;
    for (uint32_t j = j0; j < j0 + per && j < cbins; ++j) loc += lcnt[j];
    uint32_t all;
    uint32_t run = block_exclusive<SORT_THREADS>(loc, scan_sh, &all);
    for (uint32_t j = j0; j < j0 + per && j < cbins; ++j) {
      lstart[j] = run;
      run += lcnt[j];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
      if (Digit<D>::zero(e[k])) continue;
      const uint32_t i = (uint32_t)(base + k * SORT_THREADS + threadIdx.x);
      const uint32_t b = Digit<D>::bucket(e[k]), bin = b >> shift;
      const uint32_t idx = merged ? win * n + i : i;
      const uint32_t sign = Digit<D>::sign(e[k]);
      const uint32_t pos = lstart[bin] + rank[k];
      skey[pos] = (bin << 12) | (b & mask);
      sval[pos] = constant ? (idx << 2) | Digit<D>::discard(e[k]) | sign : (idx << 1) | sign;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < all; i += SORT_THREADS) {
      const uint32_t key = skey[i], bin = key >> 12;
      tmp[cur[bin] + (i - lstart[bin])] = make_uint2(key & 0xFFFu, sval[i]);
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < cbins; j += SORT_THREADS) cur[j] += lcnt[j];
    // the next tile's clearing of lcnt follows this thread's own reads of it; the barrier at
    // the top of the loop orders everyone else's
    __syncthreads();
  }
}

// The fine level works on slices of at most FINE_TILE entries, so that a heavy coarse group
// (the ones of a boolean witness all land in one) is spread over many workgroups: coarse group
// g = set cbins + bin has max(1, ceil(size / FINE_TILE)) slices, slice s of it covers entries
// [start + s FINE_TILE, min(start + (s + 1) FINE_TILE, end)) of the staging array, and its
// fine-bin counts sit at fcnt[sprefix[g] fb + f ns + s] (ns = the group's slice count): laid
// out [group][fine bin][slice], one exclusive scan per group turns them into positions.

__device__ __forceinline__ uint32_t slices_of(uint32_t size) {
  return size ? (size + FINE_TILE - 1) / FINE_TILE : 1u;
}

__global__ void __launch_bounds__(1024) k_slice_scan(const uint32_t* cpos, uint32_t per_group,
                                                     uint32_t groups, uint32_t* sprefix) {
  __shared__ uint32_t sh[16];
  const uint32_t per = (groups + 1023) / 1024;
  const uint32_t g0 = threadIdx.x * per < groups ? threadIdx.x * per : groups;
  const uint32_t g1 = g0 + per < groups ? g0 + per : groups;
  // Group boundaries eight groups at a time, loaded before they are used (cpos[groups *
  // per_group] is the total, so boundary g1 is always in range).
  uint32_t loc = 0;
  for (uint32_t g = g0; g < g1; g += 8) {
    uint32_t a[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = g + k <= g1 ? cpos[(size_t)(g + k) * per_group] : 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (g + k < g1) loc += slices_of(a[k + 1] - a[k]);
  }
  uint32_t all;
  uint32_t run = block_exclusive<1024>(loc, sh, &all);
  for (uint32_t g = g0; g < g1; g += 8) {
    uint32_t a[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = g + k <= g1 ? cpos[(size_t)(g + k) * per_group] : 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (g + k < g1) {
        sprefix[g + k] = run;
        run += slices_of(a[k + 1] - a[k]);
      }
  }
  if (threadIdx.x == 0) sprefix[groups] = all;
}

struct Slice {
  uint32_t g, s, ns, start, end;  // group, slice, group's slices, slice's entry range
};

// The slice of workgroup b, or ns = 0 past the last slice.
__device__ __forceinline__ Slice find_slice(uint32_t b, const uint32_t* sprefix, uint32_t groups,
                                            const uint32_t* cpos, uint32_t per_group) {
  Slice r{0, 0, 0, 0, 0};
  if (b >= sprefix[groups]) return r;
  uint32_t lo = 0, hi = groups;  // sprefix strictly increases: the last g with sprefix[g] <= b
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (sprefix[mid] <= b) lo = mid;
    else hi = mid;
  }
  r.g = lo;
  r.s = b - sprefix[lo];
  r.ns = sprefix[lo + 1] - sprefix[lo];
  const uint32_t start = cpos[(size_t)lo * per_group], end = cpos[(size_t)(lo + 1) * per_group];
  r.start = start + r.s * FINE_TILE;
  r.end = r.start + FINE_TILE < end ? r.start + FINE_TILE : end;
  return r;
}

__global__ void __launch_bounds__(FINE_THREADS) k_fine_count(const uint2* tmp, const uint32_t* cpos,
                                                             uint32_t per_group, uint32_t groups,
                                                             const uint32_t* sprefix, int shift,
                                                             uint32_t* fcnt) {
  __shared__ uint32_t cnt[1u << MAX_FINE_SHIFT];
  const Slice sl = find_slice(blockIdx.x, sprefix, groups, cpos, per_group);
  if (!sl.ns) return;
  const uint32_t fb = 1u << shift;
  for (uint32_t j = threadIdx.x; j < fb; j += FINE_THREADS) cnt[j] = 0;
  __syncthreads();
  uint32_t f[FINE_ITEMS];
#pragma unroll
  for (int k = 0; k < FINE_ITEMS; ++k) {
    const uint32_t i = sl.start + k * FINE_THREADS + threadIdx.x;
    f[k] = i < sl.end ? tmp[i].x : ~0u;
  }
#pragma unroll
  for (int k = 0; k < FINE_ITEMS; ++k)
    if (f[k] != ~0u) lds_take(cnt, f[k]);
  __syncthreads();
  uint32_t* out = fcnt + (size_t)sprefix[sl.g] * fb + sl.s;
  for (uint32_t j = threadIdx.x; j < fb; j += FINE_THREADS) out[(size_t)j * sl.ns] = cnt[j];
}

// Group g's counts [fine bin][slice] into positions (exclusive prefix from the group's start),
// in place, and the offset of each of its buckets (the position of its slice 0). The last
// group also writes offset[K] and the total.
__global__ void __launch_bounds__(FINE_THREADS) k_fine_prefix(const uint32_t* cpos,
                                                              uint32_t per_group, uint32_t groups,
                                                              const uint32_t* sprefix,
                                                              uint32_t cbins, int shift,
                                                              uint32_t nb, uint32_t* fcnt,
                                                              uint32_t* offset, uint32_t* total) {
  __shared__ uint32_t sh[FINE_THREADS / 64];
  const uint32_t g = blockIdx.x;
  const uint32_t fb = 1u << shift;
  const uint32_t ns = sprefix[g + 1] - sprefix[g];
  const uint32_t m = fb * ns;
  uint32_t* c = fcnt + (size_t)sprefix[g] * fb;
  const uint32_t start = cpos[(size_t)g * per_group];
  const uint32_t set = g / cbins, bin = g - set * cbins;
  const uint32_t key0 = set * nb + (bin << shift);
  const uint32_t per = (m + FINE_THREADS - 1) / FINE_THREADS;
  const uint32_t j0 = threadIdx.x * per;
  uint32_t loc = 0;
  for (uint32_t j = j0; j < j0 + per && j < m; ++j) loc += c[j];
  uint32_t all;
  uint32_t run = start + block_exclusive<FINE_THREADS>(loc, sh, &all);
  for (uint32_t j = j0; j < j0 + per && j < m; ++j) {
    const uint32_t v = c[j];
    c[j] = run;
    const uint32_t f = j / ns;
    if (j == f * ns) offset[key0 + f] = run;
    run += v;
  }
  if (g + 1 == groups && threadIdx.x == 0) {  // key0 + fb = sets nb = K
    offset[key0 + fb] = start + all;
    total[0] = start + all;
  }
}

// Each slice's entries ordered by bucket into ent, keys set nb + (bin << shift) + fine: the
// slice is first ordered by fine bin in LDS, then written with consecutive lanes on
// consecutive positions of each bin's run (as the coarse scatter does).
__global__ void __launch_bounds__(FINE_THREADS) k_fine_scatter(const uint2* tmp,
                                                               const uint32_t* cpos,
                                                               uint32_t per_group,
                                                               uint32_t groups,
                                                               const uint32_t* sprefix,
                                                               uint32_t cbins, int shift,
                                                               uint32_t nb, const uint32_t* fcnt,
                                                               uint2* ent) {
  constexpr uint32_t FB = 1u << MAX_FINE_SHIFT;
  __shared__ uint32_t cur[FB];   // the slice's first position in each fine bin's run
  __shared__ uint32_t lcnt[FB];  // the slice's entries per fine bin, then their exclusive prefix
  __shared__ uint32_t sf[FINE_TILE], sv[FINE_TILE];  // staged fine bits and values
  __shared__ uint32_t scan_sh[FINE_THREADS / 64];
  const Slice sl = find_slice(blockIdx.x, sprefix, groups, cpos, per_group);
  if (!sl.ns) return;  // the same for the whole workgroup
  const uint32_t fb = 1u << shift;
  const uint32_t* pos = fcnt + (size_t)sprefix[sl.g] * fb + sl.s;
  for (uint32_t j = threadIdx.x; j < fb; j += FINE_THREADS) {
    cur[j] = pos[(size_t)j * sl.ns];
    lcnt[j] = 0;
  }
  const uint32_t set = sl.g / cbins, bin = sl.g - set * cbins;
  const uint32_t key0 = set * nb + (bin << shift);
  uint2 t[FINE_ITEMS];
  uint32_t rank[FINE_ITEMS];
#pragma unroll
  for (int k = 0; k < FINE_ITEMS; ++k) {
    const uint32_t i = sl.start + k * FINE_THREADS + threadIdx.x;
    t[k] = i < sl.end ? tmp[i] : make_uint2(~0u, 0u);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < FINE_ITEMS; ++k) rank[k] = t[k].x != ~0u ? lds_take(lcnt, t[k].x) : 0u;
  __syncthreads();
  const uint32_t per = (fb + FINE_THREADS - 1) / FINE_THREADS;
  const uint32_t j0 = threadIdx.x * per;
  uint32_t loc = 0;
  for (uint32_t j = j0; j < j0 + per && j < fb; ++j) loc += lcnt[j];
  uint32_t all;
  uint32_t run = block_exclusive<FINE_THREADS>(loc, scan_sh, &all);
  for (uint32_t j = j0; j < j0 + per && j < fb; ++j) {  // each thread rewrites only its bins
    const uint32_t c = lcnt[j];
    lcnt[j] = run;
    run += c;
  }
  __syncthreads();
  const uint32_t* lstart = lcnt;
#pragma unroll
  for (int k = 0; k < FINE_ITEMS; ++k) {
    if (t[k].x == ~0u) continue;
    const uint32_t q = lstart[t[k].x] + rank[k];
    sf[q] = t[k].x;
    sv[q] = t[k].y;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < all; i += FINE_THREADS) {
    const uint32_t f = sf[i];
    ent[cur[f] + (i - lstart[f])] = make_uint2(key0 + f, sv[i]);
  }
}

// Exclusive scan of count[0..K) into offset[0..K], offset[K] = total (K <= 1024 SCAN_BLOCK).
// Three launches: per-block scans, a scan of the block sums, the fix-up.
__global__ void __launch_bounds__(256) k_scan_blocks(const uint32_t* count, uint32_t K,
                                                     uint32_t* offset, uint32_t* bsum) {
  constexpr int PER = SCAN_BLOCK / 256;
  __shared__ uint32_t sh[256];
  const uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * PER;
  uint32_t v[PER], loc = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    v[k] = base + k < K ? count[base + k] : 0;
    loc += v[k];
  }
  sh[threadIdx.x] = loc;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    uint32_t x = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = sh[threadIdx.x] - loc;  // exclusive prefix within the block
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    if (base + k < K) offset[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = sh[255];
}

__global__ void __launch_bounds__(1024) k_scan_top(uint32_t* bsum, uint32_t nblk) {
  __shared__ uint32_t sh[1024];
  uint32_t t = threadIdx.x;
  uint32_t x = t < nblk ? bsum[t] : 0;
  sh[t] = x;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    uint32_t y = t >= (unsigned)off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += y;
    __syncthreads();
  }
  if (t < nblk) bsum[t] = sh[t] - x;   // exclusive
  if (t == 0) bsum[1024] = sh[1023];   // grand total
}

template <int = 0>
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_scan_fix(uint32_t* offset, const uint32_t* bsum, uint32_t K) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < K) offset[i] += bsum[i / SCAN_BLOCK];
  if (i == 0) offset[K] = bsum[1024];
}

int msm_choose_c(uint32_t n, int g1, int g2) {
  // The reference's cost model (crates/msm/src/lib.rs:160-172) prices a window at
  // n + 2 * buckets points. Here the weights are fitted to this build's measurements on gfx950
  // (sweeps of c at 2^20 and 2^21 points and of the prove's two widths): a G2 entry takes
  // about 3 G1 entries' time, and a bucket's share of the merges and the window reduction
  // about 6 entries of its group (G1) or 8 (G2). Scalars sorted once for g1 G1 and g2 G2 MSMs
  // share one width; it is capped at 16 (the LDS counting sort).
  int best = 3;
  double best_cost = 1e300;
  for (int c = 3; c <= 16; ++c) {
    double windows = (RECODE_BITS + c - 1) / c;
    double nb = (double)(1u << (c - 1));
    double cost = windows * (g1 * ((double)n + 6.0 * nb) + 3.0 * g2 * ((double)n + 8.0 * nb));
    if (cost < best_cost) {
      best_cost = cost;
      best = c;
    }
  }
  return best;
}

int msm_default_c(uint32_t n) { return msm_choose_c(n, 1, 0); }

// The same weights for MSMs over window tables: every window's digits fall into one bucket
// set, so the buckets' merges and reduction are paid once, not once per window (measured on
// gfx950: the 1.1M-constraint prove 78.7 -> 79.8 proofs/s, the 64-proof batch 417-431 ->
// 445-455 proofs/s against the per-window pricing).
int msm_choose_c_merged(uint32_t n, int g1, int g2) {
  int best = 3;
  double best_cost = 1e300;
  for (int c = 3; c <= 16; ++c) {
    double windows = (RECODE_BITS + c - 1) / c;
    double nb = (double)(1u << (c - 1));
    double cost = windows * (g1 + 3.0 * g2) * (double)n + (g1 * 6.0 + 3.0 * g2 * 8.0) * nb;
    if (cost < best_cost) {
      best_cost = cost;
      best = c;
    }
  }
  return best;
}

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// The sort's geometry: coarse bins per set (fine bins per coarse bin at most 2^MAX_FINE_SHIFT,
// coarse bins at most MAX_CBINS), and chunks of at
// least SORT_CHUNK scalars, doubled until the count array fits the scan (1024 SCAN_BLOCK).
struct SortGeom {
  int shift = 0;
  uint32_t cbins = 1, rows_per_set = 1, sets = 1, rows = 1, cs = SORT_CHUNK, chunks = 1;
  size_t q = 0;  // counts: rows * cbins * chunks
  uint32_t groups = 1, max_slices = 1;  // coarse groups, and slices at most (one per group
                                        // plus one per FINE_TILE entries)
  size_t fine_words() const { return ((size_t)1 << shift) * max_slices; }
};

static SortGeom sort_geom(uint32_t n, int c, uint32_t batch, bool merged) {
  SortGeom g;
  const int W = (RECODE_BITS + c - 1) / c;
  const uint32_t nb = 1u << (c - 1);
  g.rows_per_set = merged ? (uint32_t)W : 1u;
  g.rows = batch * (uint32_t)W;
  g.sets = g.rows / g.rows_per_set;
  const size_t per_set = (size_t)n * g.rows_per_set;
  uint32_t lo = nb >> MAX_FINE_SHIFT;
  if (lo < 1) lo = 1;
  // Coarse bins about 2 sqrt(nb): the coarse scatter's runs (a chunk's entries per bin) and the
  // fine scatter's (a slice's entries per fine bin) come out of similar length (measured on
  // gfx950 against 4x fewer and 4x more bins: the 2^20 sort 0.290 against 0.310 and 0.310 ms),
  // and at most one bin per 1024 entries of a set.
  uint32_t cb = 1;
  while ((size_t)cb * cb < (size_t)4 * nb && cb * 2 <= nb && cb * 2 <= MAX_CBINS) cb *= 2;
  while (cb > 1 && (size_t)cb * 1024 > per_set) cb /= 2;
  if (cb < lo) cb = lo;
  const size_t cap = (size_t)1024 * SCAN_BLOCK - 1;
  uint32_t cs = SORT_CHUNK;
  auto chunks_of = [&](uint32_t x) { return n ? (uint32_t)(((size_t)n + x - 1) / x) : 1u; };
  while ((size_t)g.rows * cb * chunks_of(cs) > cap && cs < (1u << 30)) cs *= 2;
  while ((size_t)g.rows * cb * chunks_of(cs) > cap && cb > lo) cb /= 2;
  if ((size_t)g.rows * cb * chunks_of(cs) > cap) throw ArgError{"MSM too large for the sort's scan"};
  g.cbins = cb;
  g.shift = 0;
  while ((cb << g.shift) < nb) ++g.shift;
  g.cs = cs;
  g.chunks = chunks_of(cs);
  g.q = (size_t)g.rows * cb * g.chunks;
  g.groups = g.sets * cb;
  const size_t slices = g.groups + ((size_t)batch * n * W + FINE_TILE - 1) / FINE_TILE;
  if (slices >= (1u << 31)) throw ArgError{"MSM too large for the sort's slices"};
  g.max_slices = (uint32_t)slices;
  return g;
}

size_t msm_sort_scratch_bytes(uint32_t n, int c, uint32_t batch) {
  const int W = (RECODE_BITS + c - 1) / c;
  const size_t nb = (size_t)1 << (c - 1);
  const size_t E = (size_t)batch * n * W;
  // Sized for per-window sets (the larger key count) and the geometry of either layout.
  const size_t K = (size_t)batch * W * nb;
  const SortGeom a = sort_geom(n, c, batch, false), b = sort_geom(n, c, batch, true);
  const size_t q = std::max(a.q, b.q);
  const size_t groups = std::max(a.groups, b.groups);
  const size_t fine = std::max(a.fine_words(), b.fine_words());
  // offset, block sums, counts, positions, entries, staging, slice prefix, fine counts,
  // digits (32-bit: the widest), total
  return align_up(4 * (K + 1)) + align_up(4 * 1025) + 2 * align_up(4 * (q + 1)) +
         2 * align_up(8 * E) + align_up(4 * (groups + 1)) + align_up(4 * fine) +
         align_up(4 * E) + align_up(16);
}

void msm_plan_bind(MsmPlan& p, void* scratch, uint32_t n, int c, bool merged, bool constant,
                   uint32_t batch) {
  if (c < 2 || c > MSM_MAX_C) throw ArgError{"MSM window width must be in [2, 20]"};
  if (batch == 0) throw ArgError{"MSM batch of zero vectors"};
  p.merged = merged;
  p.constant = constant;
  p.c = c;
  p.windows = (RECODE_BITS + c - 1) / c;
  p.n = n;
  p.nb = 1u << (c - 1);
  p.batch = batch;
  p.max_entries = (size_t)batch * n * p.windows;
  if ((size_t)n * batch >= (1u << 31)) throw ArgError{"MSM larger than 2^31 points"};
  if (p.max_entries >= (1ull << 31)) throw ArgError{"MSM entry count exceeds 2^31"};
  if (constant && (merged ? (size_t)n * p.windows : (size_t)n) >= (1ull << 30))
    throw ArgError{"constant-work MSM: base indices must stay below 2^30"};
  const SortGeom g = sort_geom(n, c, batch, merged);
  p.shift = g.shift;
  p.cbins = g.cbins;
  p.rows_per_set = g.rows_per_set;
  p.chunk = g.cs;
  p.chunks = g.chunks;
  p.q = g.q;
  p.groups = g.groups;
  p.max_slices = g.max_slices;
  const size_t K = (size_t)msm_bucket_sets(p) * p.nb;
  char* q = static_cast<char*>(scratch);
  auto take = [&](size_t bytes) {
    char* r = q;
    q += align_up(bytes);
    return r;
  };
  p.offset = reinterpret_cast<uint32_t*>(take(4 * (K + 1)));
  p.block_sums = reinterpret_cast<uint32_t*>(take(4 * 1025));
  p.ccount = reinterpret_cast<uint32_t*>(take(4 * (p.q + 1)));
  p.cpos = reinterpret_cast<uint32_t*>(take(4 * (p.q + 1)));
  p.ent = reinterpret_cast<uint2*>(take(8 * p.max_entries));
  p.tmp = reinterpret_cast<uint2*>(take(8 * p.max_entries));
  p.sprefix = reinterpret_cast<uint32_t*>(take(4 * ((size_t)p.groups + 1)));
  p.fcnt = reinterpret_cast<uint32_t*>(take(4 * g.fine_words()));
  p.digits = reinterpret_cast<uint32_t*>(take(4 * p.max_entries));
  p.total = reinterpret_cast<uint32_t*>(take(16));
}

template <class D>
static void sort_coarse_count(MsmPlan& p, const void* scalars, hipStream_t s) {
  D* digits = reinterpret_cast<D*>(p.digits);
  const uint32_t total = p.batch * p.n;
  hipLaunchKernelGGL(k_digits<D>, dim3((total + 255) / 256), dim3(256), 0, s,
                     static_cast<const uint4*>(scalars), p.n, total, p.c, p.windows, digits,
                     (int)p.constant);
  G16_LAUNCH_CHECK("msm digits");
  hipLaunchKernelGGL(k_coarse_count<D>, dim3(p.chunks, p.batch * p.windows), dim3(SORT_THREADS),
                     4 * p.cbins, s, digits, p.n, p.chunk, p.shift, p.cbins, p.rows_per_set,
                     p.chunks, p.ccount);
  G16_LAUNCH_CHECK("msm coarse count");
}

template <class D>
static void sort_coarse_scatter(MsmPlan& p, hipStream_t s) {
  hipLaunchKernelGGL(k_coarse_scatter<D>, dim3(p.chunks, p.batch * p.windows), dim3(SORT_THREADS),
                     0, s, reinterpret_cast<const D*>(p.digits), p.n, p.chunk, p.shift,
                     p.cbins, p.rows_per_set, p.chunks, p.windows, p.cpos, p.tmp, (int)p.merged,
                     (int)p.constant);
  G16_LAUNCH_CHECK("msm coarse scatter");
}

void msm_sort(MsmPlan& p, const void* scalars, hipStream_t s) {
  const uint32_t Q = (uint32_t)p.q;
  if (p.n) {
    if (short_digits(p)) sort_coarse_count<uint16_t>(p, scalars, s);
    else sort_coarse_count<uint32_t>(p, scalars, s);
  } else {
    hipLaunchKernelGGL(k_fill_u32<0>, dim3((Q + 255) / 256), dim3(256), 0, s, p.ccount,
                       (size_t)Q, 0u);
    G16_LAUNCH_CHECK("msm clear counts");
  }
  const uint32_t nblk = (Q + SCAN_BLOCK - 1) / SCAN_BLOCK;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nblk), dim3(256), 0, s, p.ccount, Q, p.cpos,
                     p.block_sums);
  G16_LAUNCH_CHECK("msm scan");
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, p.block_sums, nblk);
  G16_LAUNCH_CHECK("msm scan top");
  hipLaunchKernelGGL(k_scan_fix, dim3((Q + 255) / 256), dim3(256), 0, s, p.cpos, p.block_sums, Q);
  G16_LAUNCH_CHECK("msm scan fix");
  if (p.n) {
    if (short_digits(p)) sort_coarse_scatter<uint16_t>(p, s);
    else sort_coarse_scatter<uint32_t>(p, s);
  }
  const uint32_t per_group = p.rows_per_set * p.chunks;
  hipLaunchKernelGGL(k_slice_scan, dim3(1), dim3(1024), 0, s, p.cpos, per_group, p.groups,
                     p.sprefix);
  G16_LAUNCH_CHECK("msm slice scan");
  hipLaunchKernelGGL(k_fine_count, dim3(p.max_slices), dim3(FINE_THREADS), 0, s, p.tmp, p.cpos,
                     per_group, p.groups, p.sprefix, p.shift, p.fcnt);
  G16_LAUNCH_CHECK("msm fine count");
  hipLaunchKernelGGL(k_fine_prefix, dim3(p.groups), dim3(FINE_THREADS), 0, s, p.cpos, per_group,
                     p.groups, p.sprefix, p.cbins, p.shift, p.nb, p.fcnt, p.offset, p.total);
  G16_LAUNCH_CHECK("msm fine prefix");
  hipLaunchKernelGGL(k_fine_scatter, dim3(p.max_slices), dim3(FINE_THREADS), 0, s, p.tmp, p.cpos,
                     per_group, p.groups, p.sprefix, p.cbins, p.shift, p.nb, p.fcnt, p.ent);
  G16_LAUNCH_CHECK("msm fine scatter");
}

void msm_reduce_shape(const MsmPlan& p, int* rows, int* lg) {
  uint32_t B = reduce_group_size(p);
  *lg = ilog2_u32(B);
  *rows = 1 + ilog2_u32(p.nb / B);
}

size_t msm_out_points(const MsmPlan& p) {
  int rows, lg;
  msm_reduce_shape(p, &rows, &lg);
  return (size_t)msm_bucket_sets(p) * rows;
}

size_t msm_accum_scratch_bytes(const MsmPlan& p, int group) {
  return msm_accum_scratch_bytes_impl(p, group);
}

template <>
void Tail<Fm>::merge(const MsmPlan& p, const AccumPtrs& a, hipStream_t s) {
  tail_merge_impl<Fm>(p, a, s);
}

template <>
void Tail<Fm>::reduce(const MsmPlan& p, const AccumPtrs& a, uint32_t B, void* window_out,
                      hipStream_t s) {
  tail_reduce_impl<Fm>(p, a, B, window_out, s);
}

// Window tables: entry w n + i = 2^(c w) P_i from P_i at entry i, one thread per point.
void msm_window_tables(int group, void* table, size_t n, int c, int windows, hipStream_t s) {
  if (group == 2) {
    msm_window_tables_g2(table, n, c, windows, s);
    return;
  }
  if (!n || windows < 2) return;
  hipLaunchKernelGGL(k_window_tables<Fm>, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s,
                     static_cast<uint4*>(table), n, c, windows);
  G16_LAUNCH_CHECK("msm window tables");
}

// x -> 32 x mod q (five modular doublings): the R = 2^256 Montgomery form to R' = 2^261.
__global__ void __launch_bounds__(256) k_to_msm_form(const uint4* src, uint4* dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fq x = ld_fq(src + 2 * i);
#pragma unroll
  for (int k = 0; k < 5; ++k) x = add(x, x);
  st_fq(dst + 2 * i, x);
}

void msm_to_msm_form(const void* src, void* dst, size_t n_fq, hipStream_t s) {
  if (!n_fq) return;
  hipLaunchKernelGGL(k_to_msm_form, dim3((unsigned)((n_fq + 255) / 256)), dim3(256), 0, s,
                     static_cast<const uint4*>(src), static_cast<uint4*>(dst), n_fq);
  G16_LAUNCH_CHECK("msm bases to R' form");
}

// G2 lives in kernels_msm_g2*.hip, compiled on their own (the slow ones to compile).
void msm_accumulate_g2(const MsmPlan& p, const void* bases, void* scratch, void* window_out,
                       hipStream_t s, hipEvent_t* ev);

// Workgroups of k_accumulate_queue<Fm> the device holds at once (every CU's share), measured
// once per process.
static int queue_grid() {
  static const int grid = [] {
    int dev = 0, cus = 0, per = 0;
    G16_CHECK(hipGetDevice(&dev));
    G16_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    G16_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_accumulate_queue<Fm>,
                                                           QUEUE_TILE, 0));
    return std::max(1, cus) * std::max(1, per);
  }();
  return grid;
}

void msm_accumulate_g1_queue(const MsmPlan* const* plans, const void* const* bases,
                             void* const* scratch, void* const* window_out, int n,
                             uint32_t* counter, hipStream_t s) {
  accumulate_queue_impl<Fm>(plans, bases, scratch, window_out, n, counter, queue_grid(), s);
}

void msm_accumulate(const MsmPlan& p, int group, const void* bases, void* scratch,
                    void* window_out, hipStream_t s, hipEvent_t* ev) {
  if (group == 1) accumulate_impl<Fm>(p, bases, scratch, window_out, s, ev);
  else msm_accumulate_g2(p, bases, scratch, window_out, s, ev);
}

}  // namespace g16
