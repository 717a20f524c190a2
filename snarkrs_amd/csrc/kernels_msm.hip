// Pippenger MSM for G1 and G2 (stages 5-9, crates/groth16/src/cpu.rs:477-632).
//
// Pipeline per scalar vector (shared by every base vector multiplied by the same scalars,
// so the witness MSMs A, B1, B2 and L sort once):
//   1. digits:  Montgomery -> standard scalar, carry-free signed digits per window
//               (crates/msm/src/lib.rs:275-331), stored window-major;
//   2. sort:    a counting sort of (window, |digit|) keys without global atomics: each
//               workgroup counts one chunk of one window in LDS, a column scan over the
//               chunks and a scan over the buckets place every (bucket, chunk) pair, and the
//               scatter ranks a chunk's entries with LDS atomics;
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
//   32-bit (c = 16, or constant work): (|d| - 1) | sign << 16, DIGIT_ZERO for d = 0; under
//     constant work a zero digit is (i mod nb) | DIGIT_DISCARD instead, an entry whose addition
//     is computed and dropped;
//   16-bit (c <= 15, variable work): (|d| - 1) | sign << 15, 0x7FFF for d = 0 (|d| - 1 is at
//     most 2^14 - 1 there), half the digit traffic.
static constexpr uint32_t DIGIT_ZERO = 1u << 17;
static constexpr uint32_t DIGIT_DISCARD = 1u << 18;

template <class D> struct Digit;
template <> struct Digit<uint32_t> {
  static __device__ __forceinline__ uint32_t make(int d, uint32_t zero) {
    return d == 0 ? zero : ((uint32_t)(d < 0 ? -d : d) - 1u) | (d < 0 ? 1u << 16 : 0u);
  }
  static __device__ __forceinline__ bool zero(uint32_t e) { return (e & DIGIT_ZERO) != 0; }
  static __device__ __forceinline__ uint32_t bucket(uint32_t e) { return e & 0xFFFFu; }
  static __device__ __forceinline__ uint32_t sign(uint32_t e) { return (e >> 16) & 1u; }
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
  static __device__ __forceinline__ uint32_t bucket(uint16_t e) { return e & 0x7FFFu; }
  static __device__ __forceinline__ uint32_t sign(uint16_t e) { return (uint32_t)e >> 15; }
  static __device__ __forceinline__ uint32_t discard(uint16_t) { return 0; }
};

// Whether a plan's digits are 16-bit.
static bool short_digits(const MsmPlan& p) { return p.c <= 15 && !p.constant; }

// Scalars per counting-sort workgroup: a bucket's count within a chunk fits 16 bits, so the
// count kernel's LDS histogram packs two buckets per word (64 KB at c = 16).
static constexpr uint32_t SORT_CHUNK = 32768;
static constexpr int SORT_THREADS = 1024;
static constexpr uint32_t SCATTER_SPAN = 16384;  // buckets per scatter workgroup (64 KB LDS)

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

// Entries per (window, bucket): chunk blockIdx.x of window blockIdx.y counts its digits in an
// LDS histogram and adds each nonzero bin to count[w nb + b] (cleared beforehand).
template <class D>
__global__ void __launch_bounds__(SORT_THREADS) k_chunk_count(const D* digits, uint32_t n,
                                                              uint32_t nb, uint32_t* count) {
  extern __shared__ uint32_t hist[];  // nb / 2 words
  uint32_t p = blockIdx.x, w = blockIdx.y;
  uint32_t words = nb / 2;
  for (uint32_t j = threadIdx.x; j < words; j += SORT_THREADS) hist[j] = 0;
  __syncthreads();
  size_t lo = (size_t)p * SORT_CHUNK;
  size_t hi = lo + SORT_CHUNK < n ? lo + SORT_CHUNK : n;
  const D* d = digits + (size_t)w * n;
  for (size_t i = lo + threadIdx.x; i < hi; i += SORT_THREADS) {
    const D e = d[i];
    if (Digit<D>::zero(e)) continue;
    const uint32_t b = Digit<D>::bucket(e);
    atomicAdd(&hist[b >> 1], 1u << ((b & 1u) << 4));
  }
  __syncthreads();
  uint32_t* out = count + (size_t)w * nb;
  for (uint32_t j = threadIdx.x; j < words; j += SORT_THREADS) {
    const uint32_t v = hist[j];
    if (v & 0xFFFFu) atomicAdd(&out[2 * j], v & 0xFFFFu);
    if (v >> 16) atomicAdd(&out[2 * j + 1], v >> 16);
  }
}

template <int = 0>
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// The scatter's global cursors: the first position of key t = (vector k, window w) nb + b,
// from the bucket offsets and, merged (wpre given), the entries of the vector's earlier windows
// in that bucket (whose offsets are per (k, b)).
__global__ void k_cursor_init(uint32_t* cursor, const uint32_t* offset, const uint32_t* wpre,
                              uint32_t nb, int W, uint32_t K) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K) return;
  const uint32_t vw = t / nb, b = t - vw * nb;
  cursor[t] = wpre ? offset[(vw / (uint32_t)W) * nb + b] + wpre[t] : offset[t];
}

// Merged windows: per (vector k, bucket b), count[(k W + w) nb + b] becomes the entries of the
// vector's windows before w and btotal[k nb + b] the bucket's total over all its windows.
__global__ void __launch_bounds__(256) k_window_scan(uint32_t* count, uint32_t nb, int W,
                                                     uint32_t sets, uint32_t* btotal) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= sets * nb) return;
  const uint32_t k = t / nb, b = t - k * nb;
  uint32_t* col = count + (size_t)k * W * nb + b;
  uint32_t run = 0;
  for (int w = 0; w < W; ++w) {
    uint32_t c = col[(size_t)w * nb];
    col[(size_t)w * nb] = run;
    run += c;
  }
  btotal[t] = run;
}

// Entries of chunk blockIdx.x of digit row blockIdx.y = (vector k, window w) whose bucket lies
// in [blockIdx.z * span, + span), in three steps over one LDS array: a histogram of the chunk's
// digits; each nonzero bin reserves its run of positions with one atomic on the global cursor
// of its key (the runs of a bucket's chunks land in any order; the bucket's sum does not depend
// on it); then each entry takes the next position of its bin. Merged: a vector's windows share
// its one bucket set (key k nb + b) and an entry's value indexes the window table (w n + i).
template <class D>
__global__ void __launch_bounds__(SORT_THREADS) k_chunk_scatter(
    const D* digits, uint32_t n, uint32_t nb, int W, uint32_t span, uint32_t* cursor, int merged,
    uint2* ent, int constant) {
  extern __shared__ uint32_t cur[];  // span words
  const uint32_t p = blockIdx.x, w = blockIdx.y, b0 = blockIdx.z * span;  // w: digit row
  const uint32_t k = w / (uint32_t)W, win = w - k * (uint32_t)W;
  for (uint32_t j = threadIdx.x; j < span; j += SORT_THREADS) cur[j] = 0;
  __syncthreads();
  size_t lo = (size_t)p * SORT_CHUNK;
  size_t hi = lo + SORT_CHUNK < n ? lo + SORT_CHUNK : n;
  const D* d = digits + (size_t)w * n;
  for (size_t i = lo + threadIdx.x; i < hi; i += SORT_THREADS) {
    const D e = d[i];
    if (Digit<D>::zero(e)) continue;
    const uint32_t b = Digit<D>::bucket(e) - b0;
    if (b < span) atomicAdd(&cur[b], 1u);
  }
  __syncthreads();
  uint32_t* gc = cursor + (size_t)w * nb + b0;
  for (uint32_t j = threadIdx.x; j < span; j += SORT_THREADS) {
    const uint32_t c = cur[j];
    if (c) cur[j] = atomicAdd(&gc[j], c);
  }
  __syncthreads();
  for (size_t i = lo + threadIdx.x; i < hi; i += SORT_THREADS) {
    const D e = d[i];
    if (Digit<D>::zero(e)) continue;
    const uint32_t b = Digit<D>::bucket(e) - b0;
    if (b >= span) continue;
    const uint32_t pos = atomicAdd(&cur[b], 1u);
    const uint32_t key = (merged ? k : w) * nb + b0 + b;
    const uint32_t idx = merged ? win * n + (uint32_t)i : (uint32_t)i;
    const uint32_t sign = Digit<D>::sign(e);
    ent[pos] = make_uint2(key, constant ? (idx << 2) | Digit<D>::discard(e) | sign
                                        : (idx << 1) | sign);
  }
}

// Exclusive scan of count[0..K) into offset[0..K], offset[K] = total.
// Three launches: per-1024-element block scans, a scan of the block sums, the fix-up.
__global__ void __launch_bounds__(256) k_scan_blocks(const uint32_t* count, uint32_t K,
                                                     uint32_t* offset, uint32_t* bsum) {
  __shared__ uint32_t sh[256];
  uint32_t base = blockIdx.x * 1024 + threadIdx.x * 4;
  uint32_t v[4], loc = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
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
  for (int k = 0; k < 4; ++k) {
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

__global__ void k_scan_fix(uint32_t* offset, const uint32_t* bsum, uint32_t K,
                           uint32_t* total) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < K) offset[i] += bsum[i / 1024];
  if (i == 0) {
    offset[K] = bsum[1024];
    total[0] = bsum[1024];
  }
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

static uint32_t sort_chunks(uint32_t n) { return (n + SORT_CHUNK - 1) / SORT_CHUNK; }

size_t msm_sort_scratch_bytes(uint32_t n, int c, uint32_t batch) {
  int W = (RECODE_BITS + c - 1) / c;
  size_t nb = (size_t)1 << (c - 1);
  size_t K = (size_t)batch * W * nb;
  size_t E = (size_t)batch * n * W;
  // count, offset, block sums, entries, digits (32-bit: the widest), cursors, btotal, total
  return align_up(4 * (K + 1)) * 2 + align_up(4 * 1025) + align_up(8 * E) + align_up(4 * E) +
         align_up(4 * K) + align_up(4 * (batch * nb + 1)) + align_up(16);
}

void msm_plan_bind(MsmPlan& p, void* scratch, uint32_t n, int c, bool merged, bool constant,
                   uint32_t batch) {
  if (c < 2 || c > 16) throw ArgError{"MSM window width must be in [2, 16]"};
  if (batch == 0) throw ArgError{"MSM batch of zero vectors"};
  p.merged = merged;
  p.constant = constant;
  p.c = c;
  p.windows = (RECODE_BITS + c - 1) / c;
  p.n = n;
  p.nb = 1u << (c - 1);
  p.batch = batch;
  p.max_entries = (size_t)batch * n * p.windows;
  p.chunks = sort_chunks(n);
  size_t K = (size_t)batch * p.windows * p.nb;
  const size_t KS = (size_t)msm_bucket_sets(p) * p.nb;  // keys the offsets scan covers
  if ((KS + 1023) / 1024 > 1024) throw ArgError{"MSM window too wide for the scan"};
  if ((size_t)n * batch >= (1u << 31)) throw ArgError{"MSM larger than 2^31 points"};
  if (p.max_entries >= (1ull << 31)) throw ArgError{"MSM entry count exceeds 2^31"};
  if (constant && (merged ? (size_t)n * p.windows : (size_t)n) >= (1ull << 30))
    throw ArgError{"constant-work MSM: base indices must stay below 2^30"};
  char* q = static_cast<char*>(scratch);
  auto take = [&](size_t bytes) {
    char* r = q;
    q += align_up(bytes);
    return r;
  };
  p.count = reinterpret_cast<uint32_t*>(take(4 * (K + 1)));
  p.offset = reinterpret_cast<uint32_t*>(take(4 * (K + 1)));
  p.block_sums = reinterpret_cast<uint32_t*>(take(4 * 1025));
  p.ent = reinterpret_cast<uint2*>(take(8 * p.max_entries));
  p.digits = reinterpret_cast<uint32_t*>(take(4 * p.max_entries));
  p.cursor = reinterpret_cast<uint32_t*>(take(4 * K));
  p.btotal = reinterpret_cast<uint32_t*>(take(4 * ((size_t)batch * p.nb + 1)));
  p.total = reinterpret_cast<uint32_t*>(take(16));
}

template <class D>
static void sort_count(MsmPlan& p, const void* scalars, hipStream_t s) {
  D* digits = reinterpret_cast<D*>(p.digits);
  const uint32_t total = p.batch * p.n;
  hipLaunchKernelGGL(k_digits<D>, dim3((total + 255) / 256), dim3(256), 0, s,
                     static_cast<const uint4*>(scalars), p.n, total, p.c, p.windows, digits,
                     (int)p.constant);
  G16_LAUNCH_CHECK("msm digits");
  hipLaunchKernelGGL(k_chunk_count<D>, dim3(p.chunks, p.batch * p.windows), dim3(SORT_THREADS),
                     4 * (p.nb / 2), s, digits, p.n, p.nb, p.count);
  G16_LAUNCH_CHECK("msm chunk count");
}

template <class D>
static void sort_scatter(MsmPlan& p, hipStream_t s) {
  const uint32_t span = p.nb < SCATTER_SPAN ? p.nb : SCATTER_SPAN;
  hipLaunchKernelGGL(k_chunk_scatter<D>, dim3(p.chunks, p.batch * p.windows, p.nb / span),
                     dim3(SORT_THREADS), 4 * span, s, reinterpret_cast<const D*>(p.digits), p.n,
                     p.nb, p.windows, span, p.cursor, (int)p.merged, p.ent, (int)p.constant);
  G16_LAUNCH_CHECK("msm chunk scatter");
}

void msm_sort(MsmPlan& p, const void* scalars, hipStream_t s) {
  const uint32_t K = p.batch * (uint32_t)p.windows * p.nb;
  hipLaunchKernelGGL(k_fill_u32<0>, dim3((K + 255) / 256), dim3(256), 0, s, p.count, (size_t)K,
                     0u);
  G16_LAUNCH_CHECK("msm clear counts");
  if (p.n) {
    if (short_digits(p)) sort_count<uint16_t>(p, scalars, s);
    else sort_count<uint32_t>(p, scalars, s);
  }
  // The bucket totals to scan: per (window, bucket), or per bucket over all windows.
  const uint32_t* totals = p.count;
  uint32_t KS = K;
  if (p.merged) {
    const uint32_t KB = p.batch * p.nb;
    hipLaunchKernelGGL(k_window_scan, dim3((KB + 255) / 256), dim3(256), 0, s, p.count, p.nb,
                       p.windows, p.batch, p.btotal);
    G16_LAUNCH_CHECK("msm window scan");
    totals = p.btotal;
    KS = KB;
  }
  uint32_t nblk = (KS + 1023) / 1024;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nblk), dim3(256), 0, s, totals, KS, p.offset,
                     p.block_sums);
  G16_LAUNCH_CHECK("msm scan");
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, p.block_sums, nblk);
  G16_LAUNCH_CHECK("msm scan top");
  hipLaunchKernelGGL(k_scan_fix, dim3((KS + 255) / 256), dim3(256), 0, s, p.offset,
                     p.block_sums, KS, p.total);
  G16_LAUNCH_CHECK("msm scan fix");
  if (p.n) {
    hipLaunchKernelGGL(k_cursor_init, dim3((K + 255) / 256), dim3(256), 0, s, p.cursor,
                       p.offset, p.merged ? p.count : nullptr, p.nb, p.windows, K);
    G16_LAUNCH_CHECK("msm cursors");
    if (short_digits(p)) sort_scatter<uint16_t>(p, s);
    else sort_scatter<uint32_t>(p, s);
  }
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
