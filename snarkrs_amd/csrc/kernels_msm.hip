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

// Digit encoding: (|d| - 1) | sign << 16, or DIGIT_ZERO for d = 0; under constant work a zero
// digit is (i mod nb) | DIGIT_DISCARD instead, an entry whose addition is computed and dropped.
static constexpr uint32_t DIGIT_ZERO = 1u << 17;
static constexpr uint32_t DIGIT_DISCARD = 1u << 18;
// Scalars per counting-sort workgroup: a bucket's count within a chunk fits 16 bits, so the
// LDS histogram packs two buckets per word (64 KB at c = 16).
static constexpr uint32_t SORT_CHUNK = 32768;
static constexpr int SORT_THREADS = 1024;
static constexpr uint32_t SCATTER_SPAN = 16384;  // buckets per scatter workgroup (64 KB LDS)

__global__ void __launch_bounds__(256) k_digits(const uint4* scalars, uint32_t n, int c, int W,
                                                uint32_t* digits, int constant) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fr s = to_std(ld_fr(scalars + 2 * (size_t)i));
  const uint32_t zero = constant ? (i & ((1u << (c - 1)) - 1u)) | DIGIT_DISCARD : DIGIT_ZERO;
  for (int w = 0; w < W; ++w) {
    int d = signed_digit(s.v, w, c);
    uint32_t e = d == 0 ? zero
                        : ((uint32_t)(d < 0 ? -d : d) - 1u) | (d < 0 ? 1u << 16 : 0u);
    digits[(size_t)w * n + i] = e;
  }
}

// Counts of chunk blockIdx.x of window blockIdx.y: chunk_count[w][p][b].
__global__ void __launch_bounds__(SORT_THREADS) k_chunk_count(const uint32_t* digits,
                                                              uint32_t n, uint32_t nb,
                                                              uint32_t chunks,
                                                              uint32_t* chunk_count) {
  extern __shared__ uint32_t hist[];  // nb / 2 words
  uint32_t p = blockIdx.x, w = blockIdx.y;
  uint32_t words = nb / 2;
  for (uint32_t j = threadIdx.x; j < words; j += SORT_THREADS) hist[j] = 0;
  __syncthreads();
  size_t lo = (size_t)p * SORT_CHUNK;
  size_t hi = lo + SORT_CHUNK < n ? lo + SORT_CHUNK : n;
  const uint32_t* d = digits + (size_t)w * n;
  for (size_t i = lo + threadIdx.x; i < hi; i += SORT_THREADS) {
    uint32_t e = d[i];
    if (e & DIGIT_ZERO) continue;
    uint32_t b = e & 0xFFFFu;
    atomicAdd(&hist[b >> 1], 1u << ((b & 1u) << 4));
  }
  __syncthreads();
  uint2* out = reinterpret_cast<uint2*>(chunk_count + ((size_t)w * chunks + p) * nb);
  for (uint32_t j = threadIdx.x; j < words; j += SORT_THREADS) {
    uint32_t v = hist[j];
    out[j] = make_uint2(v & 0xFFFFu, v >> 16);
  }
}

// Per bucket (w, b): chunk_count[w][p][b] becomes the bucket's entries in chunks before p,
// count[w * nb + b] its total.
__global__ void __launch_bounds__(256) k_chunk_scan(uint32_t* chunk_count, uint32_t nb,
                                                    uint32_t chunks, uint32_t K,
                                                    uint32_t* count) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K) return;
  uint32_t w = t / nb, b = t % nb;
  uint32_t* col = chunk_count + (size_t)w * chunks * nb + b;
  uint32_t run = 0;
  uint32_t p = 0;
  for (; p + 4 <= chunks; p += 4) {
    uint32_t v0 = col[(size_t)p * nb], v1 = col[(size_t)(p + 1) * nb];
    uint32_t v2 = col[(size_t)(p + 2) * nb], v3 = col[(size_t)(p + 3) * nb];
    col[(size_t)p * nb] = run;
    col[(size_t)(p + 1) * nb] = run + v0;
    col[(size_t)(p + 2) * nb] = run + v0 + v1;
    col[(size_t)(p + 3) * nb] = run + v0 + v1 + v2;
    run += v0 + v1 + v2 + v3;
  }
  for (; p < chunks; ++p) {
    uint32_t v = col[(size_t)p * nb];
    col[(size_t)p * nb] = run;
    run += v;
  }
  count[t] = run;
}

// Merged windows: per bucket b, count[w * nb + b] becomes the entries of windows before w and
// btotal[b] the bucket's total over all windows.
__global__ void __launch_bounds__(256) k_window_scan(uint32_t* count, uint32_t nb, int W,
                                                     uint32_t* btotal) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  uint32_t run = 0;
  for (int w = 0; w < W; ++w) {
    uint32_t c = count[(size_t)w * nb + b];
    count[(size_t)w * nb + b] = run;
    run += c;
  }
  btotal[b] = run;
}

// Entries of chunk blockIdx.x of window blockIdx.y whose bucket lies in
// [blockIdx.z * span, + span): positions from the bucket offset, the chunk prefix and an LDS
// cursor. Merged (wpre given): buckets are shared by every window, offsets are per bucket,
// wpre[w * nb + b] places window w after the earlier windows, and an entry's value indexes the
// window table (w n + i).
__global__ void __launch_bounds__(SORT_THREADS) k_chunk_scatter(
    const uint32_t* digits, uint32_t n, uint32_t nb, uint32_t chunks, uint32_t span,
    const uint32_t* chunk_count, const uint32_t* offset, const uint32_t* wpre, uint2* ent,
    int constant) {
  extern __shared__ uint32_t cur[];  // span words
  uint32_t p = blockIdx.x, w = blockIdx.y, b0 = blockIdx.z * span;
  const uint32_t* cc = chunk_count + ((size_t)w * chunks + p) * nb + b0;
  const uint32_t* off = offset + (wpre ? 0 : (size_t)w * nb) + b0;
  const uint32_t* wp = wpre ? wpre + (size_t)w * nb + b0 : nullptr;
  for (uint32_t j = threadIdx.x; j < span; j += SORT_THREADS)
    cur[j] = off[j] + cc[j] + (wp ? wp[j] : 0u);
  __syncthreads();
  size_t lo = (size_t)p * SORT_CHUNK;
  size_t hi = lo + SORT_CHUNK < n ? lo + SORT_CHUNK : n;
  const uint32_t* d = digits + (size_t)w * n;
  for (size_t i = lo + threadIdx.x; i < hi; i += SORT_THREADS) {
    uint32_t e = d[i];
    if (e & DIGIT_ZERO) continue;
    uint32_t b = (e & 0xFFFFu) - b0;
    if (b >= span) continue;
    uint32_t pos = atomicAdd(&cur[b], 1u);
    const uint32_t key = (wpre ? 0u : w * nb) + b0 + b;
    const uint32_t idx = wpre ? w * n + (uint32_t)i : (uint32_t)i;
    const uint32_t sign = (e >> 16) & 1u;
    ent[pos] = make_uint2(key, constant ? (idx << 2) | ((e & DIGIT_DISCARD) ? 2u : 0u) | sign
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

size_t msm_sort_scratch_bytes(uint32_t n, int c) {
  int W = (RECODE_BITS + c - 1) / c;
  size_t nb = (size_t)1 << (c - 1);
  size_t K = (size_t)W * nb;
  size_t E = (size_t)n * W;
  return align_up(4 * (K + 1)) * 2 + align_up(4 * 1025) + align_up(8 * E) + align_up(4 * E) +
         align_up(4 * K * sort_chunks(n)) + align_up(4 * (nb + 1)) + align_up(16);
}

void msm_plan_bind(MsmPlan& p, void* scratch, uint32_t n, int c, bool merged, bool constant) {
  if (c < 2 || c > 16) throw ArgError{"MSM window width must be in [2, 16]"};
  p.merged = merged;
  p.constant = constant;
  p.c = c;
  p.windows = (RECODE_BITS + c - 1) / c;
  p.n = n;
  p.nb = 1u << (c - 1);
  p.max_entries = (size_t)n * p.windows;
  p.chunks = sort_chunks(n);
  size_t K = (size_t)p.windows * p.nb;
  if ((K + 1023) / 1024 > 1024) throw ArgError{"MSM window too wide for the scan"};
  if ((size_t)n >= (1u << 31)) throw ArgError{"MSM larger than 2^31 points"};
  if (p.max_entries >= (1ull << 31)) throw ArgError{"MSM entry count exceeds 2^31"};
  if (constant && (merged ? p.max_entries : (size_t)n) >= (1ull << 30))
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
  p.chunk_count = reinterpret_cast<uint32_t*>(take(4 * K * p.chunks));
  p.btotal = reinterpret_cast<uint32_t*>(take(4 * ((size_t)p.nb + 1)));
  p.total = reinterpret_cast<uint32_t*>(take(16));
}

void msm_sort(MsmPlan& p, const void* scalars, hipStream_t s) {
  uint32_t K = (uint32_t)p.windows * p.nb;
  if (p.n) {
    hipLaunchKernelGGL(k_digits, dim3((p.n + 255) / 256), dim3(256), 0, s,
                       static_cast<const uint4*>(scalars), p.n, p.c, p.windows, p.digits,
                       (int)p.constant);
    G16_LAUNCH_CHECK("msm digits");
    hipLaunchKernelGGL(k_chunk_count, dim3(p.chunks, p.windows), dim3(SORT_THREADS),
                       4 * (p.nb / 2), s, p.digits, p.n, p.nb, p.chunks, p.chunk_count);
    G16_LAUNCH_CHECK("msm chunk count");
  }
  hipLaunchKernelGGL(k_chunk_scan, dim3((K + 255) / 256), dim3(256), 0, s, p.chunk_count, p.nb,
                     p.chunks, K, p.count);
  G16_LAUNCH_CHECK("msm chunk scan");
  // The bucket totals to scan: per (window, bucket), or per bucket over all windows.
  const uint32_t* totals = p.count;
  uint32_t KS = K;
  if (p.merged) {
    hipLaunchKernelGGL(k_window_scan, dim3((p.nb + 255) / 256), dim3(256), 0, s, p.count, p.nb,
                       p.windows, p.btotal);
    G16_LAUNCH_CHECK("msm window scan");
    totals = p.btotal;
    KS = p.nb;
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
    uint32_t span = p.nb < SCATTER_SPAN ? p.nb : SCATTER_SPAN;
    hipLaunchKernelGGL(k_chunk_scatter, dim3(p.chunks, p.windows, p.nb / span),
                       dim3(SORT_THREADS), 4 * span, s, p.digits, p.n, p.nb, p.chunks, span,
                       p.chunk_count, p.offset, p.merged ? p.count : nullptr, p.ent,
                       (int)p.constant);
    G16_LAUNCH_CHECK("msm chunk scatter");
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

void msm_accumulate(const MsmPlan& p, int group, const void* bases, void* scratch,
                    void* window_out, hipStream_t s, hipEvent_t* ev) {
  if (group == 1) accumulate_impl<Fm>(p, bases, scratch, window_out, s, ev);
  else msm_accumulate_g2(p, bases, scratch, window_out, s, ev);
}

}  // namespace g16
