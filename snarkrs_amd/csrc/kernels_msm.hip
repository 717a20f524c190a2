// Pippenger MSM for G1 and G2 (stages 5-9, crates/groth16/src/cpu.rs:477-632).
//
// Pipeline per scalar vector (shared by every base vector multiplied by the same scalars,
// so the witness MSMs A, B1, B2 and L sort once):
//   1. digits:  Montgomery -> standard scalar, carry-free signed digits per window
//               (crates/msm/src/lib.rs:275-331), histogram of (window, |digit|) keys;
//   2. sort:    exclusive scan of the histogram and a scatter of (key, index|sign) entries
//               into key order (a one-pass counting sort with an 18-bit digit: the on-GPU
//               radix sort of the digits into buckets);
//   3. accumulate: one thread per fixed-length segment of the sorted entries adds its runs
//               into XYZZ accumulators (madd-2008-s); a run that fills its bucket inside
//               the segment is stored as the bucket, the partial runs at a segment's ends go
//               to side arrays and a merge pass folds them (heavy buckets, e.g. every 1
//               scalar in bucket 0 of window 0, get a whole workgroup and an LDS tree);
//   4. reduce:  per window S_w = sum_j (j+1) B_j: groups of buckets by running sums, then a
//               per-window workgroup tree; the host finishes sum_w 2^(cw) S_w.
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

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

static void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  G16_LAUNCH_CHECK("fill");
}

__device__ __forceinline__ bool scalar_is_one(const uint32_t s[8]) {
  return (s[0] == 1u) && ((s[1] | s[2] | s[3] | s[4] | s[5] | s[6] | s[7]) == 0u);
}

// A scalar equal to 1 has one nonzero digit, window 0 digit 1: key 0. Witnesses are full of
// them (every set bit), so their counter is taken once per wavefront (ballot + popcount)
// rather than once per lane, which would serialise a quarter of a million atomics on one
// address.
__global__ void __launch_bounds__(256) k_digits_count(const uint4* scalars, uint32_t n, int c,
                                                      int W, uint32_t nb, uint4* std_out,
                                                      uint32_t* count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = i < n;
  Fr s = fp_zero<FrCfg>();
  if (live) {
    s = to_std(ld_fr(scalars + 2 * (size_t)i));
    st_fr(std_out + 2 * (size_t)i, s);
  }
  bool one = live && scalar_is_one(s.v);
  uint64_t ones = __ballot(one);
  if (ones && __lane_id() == (uint32_t)(__ffsll((unsigned long long)ones) - 1))
    atomicAdd(&count[0], (uint32_t)__popcll(ones));
  if (!live || one) return;
  for (int w = 0; w < W; ++w) {
    int d = signed_digit(s.v, w, c);
    if (d == 0) continue;
    uint32_t key = (uint32_t)w * nb + (uint32_t)(d < 0 ? -d : d) - 1;
    atomicAdd(&count[key], 1u);
  }
}

__global__ void __launch_bounds__(256) k_scatter(const uint4* std_in, uint32_t n, int c, int W,
                                                 uint32_t nb, uint32_t* cursor, uint32_t* keys,
                                                 uint32_t* vals) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = i < n;
  Fr s = fp_zero<FrCfg>();
  if (live) s = ld_fr(std_in + 2 * (size_t)i);
  bool one = live && scalar_is_one(s.v);
  uint64_t ones = __ballot(one);
  if (ones) {
    uint32_t leader = (uint32_t)(__ffsll((unsigned long long)ones) - 1);
    uint32_t base = 0;
    if (__lane_id() == leader) base = atomicAdd(&cursor[0], (uint32_t)__popcll(ones));
    base = __shfl(base, (int)leader);
    if (one) {
      uint64_t below = ones & ((1ull << __lane_id()) - 1ull);
      uint32_t pos = base + (uint32_t)__popcll(below);
      keys[pos] = 0;
      vals[pos] = i << 1;
    }
  }
  if (!live || one) return;
  for (int w = 0; w < W; ++w) {
    int d = signed_digit(s.v, w, c);
    if (d == 0) continue;
    uint32_t key = (uint32_t)w * nb + (uint32_t)(d < 0 ? -d : d) - 1;
    uint32_t pos = atomicAdd(&cursor[key], 1u);
    keys[pos] = key;
    vals[pos] = (i << 1) | (d < 0 ? 1u : 0u);
  }
}

// Exclusive scan of count[0..K) into offset[0..K], offset[K] = total; cursor = offset.
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

__global__ void k_scan_fix(uint32_t* offset, uint32_t* cursor, const uint32_t* bsum,
                           uint32_t K, uint32_t* total) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < K) {
    uint32_t o = offset[i] + bsum[i / 1024];
    offset[i] = o;
    cursor[i] = o;
  }
  if (i == 0) {
    offset[K] = bsum[1024];
    total[0] = bsum[1024];
  }
}

int msm_default_c(uint32_t n) {
  // The reference's cost model (crates/msm/src/lib.rs:160-172) weighs a bucket twice a
  // point; on the GPU the bucket reduction is latency-bound, so buckets are priced higher
  // (4 per point) and the window is capped at 16.
  int best = 3;
  double best_cost = 1e300;
  for (int c = 3; c <= 16; ++c) {
    double windows = (RECODE_BITS + c - 1) / c;
    double cost = windows * ((double)n + 4.0 * (double)(1u << (c - 1)));
    if (cost < best_cost) {
      best_cost = cost;
      best = c;
    }
  }
  return best;
}

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

size_t msm_sort_scratch_bytes(uint32_t n, int c) {
  int W = (RECODE_BITS + c - 1) / c;
  size_t K = (size_t)W << (c - 1);
  size_t E = (size_t)n * W;
  return align_up(4 * (K + 1)) * 3 + align_up(4 * 1025) + align_up(4 * E) * 2 +
         align_up((size_t)n * 32) + align_up(16);
}

void msm_plan_bind(MsmPlan& p, void* scratch, uint32_t n, int c) {
  p.c = c;
  p.windows = (RECODE_BITS + c - 1) / c;
  p.n = n;
  p.nb = 1u << (c - 1);
  p.seg = SEG;
  p.max_entries = (size_t)n * p.windows;
  size_t K = (size_t)p.windows * p.nb;
  if ((K + 1023) / 1024 > 1024) throw ArgError{"MSM window too wide for the scan"};
  if ((size_t)n >= (1u << 31)) throw ArgError{"MSM larger than 2^31 points"};
  char* q = static_cast<char*>(scratch);
  auto take = [&](size_t bytes) {
    char* r = q;
    q += align_up(bytes);
    return r;
  };
  p.count = reinterpret_cast<uint32_t*>(take(4 * (K + 1)));
  p.offset = reinterpret_cast<uint32_t*>(take(4 * (K + 1)));
  p.cursor = reinterpret_cast<uint32_t*>(take(4 * (K + 1)));
  p.block_sums = reinterpret_cast<uint32_t*>(take(4 * 1025));
  p.keys = reinterpret_cast<uint32_t*>(take(4 * p.max_entries));
  p.vals = reinterpret_cast<uint32_t*>(take(4 * p.max_entries));
  p.scalars_std = reinterpret_cast<uint32_t*>(take((size_t)n * 32));
  p.total = reinterpret_cast<uint32_t*>(take(16));
}

void msm_sort(MsmPlan& p, const void* scalars, hipStream_t s) {
  uint32_t K = (uint32_t)p.windows * p.nb;
  fill_u32(p.count, K + 1, 0, s);
  if (p.n) {
    hipLaunchKernelGGL(k_digits_count, dim3((p.n + 255) / 256), dim3(256), 0, s,
                       static_cast<const uint4*>(scalars), p.n, p.c, p.windows, p.nb,
                       reinterpret_cast<uint4*>(p.scalars_std), p.count);
    G16_LAUNCH_CHECK("msm digits");
  }
  uint32_t nblk = (K + 1023) / 1024;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nblk), dim3(256), 0, s, p.count, K, p.offset,
                     p.block_sums);
  G16_LAUNCH_CHECK("msm scan");
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, p.block_sums, nblk);
  G16_LAUNCH_CHECK("msm scan top");
  hipLaunchKernelGGL(k_scan_fix, dim3((K + 255) / 256), dim3(256), 0, s, p.offset, p.cursor,
                     p.block_sums, K, p.total);
  G16_LAUNCH_CHECK("msm scan fix");
  if (p.n) {
    hipLaunchKernelGGL(k_scatter, dim3((p.n + 255) / 256), dim3(256), 0, s,
                       reinterpret_cast<const uint4*>(p.scalars_std), p.n, p.c, p.windows, p.nb,
                       p.cursor, p.keys, p.vals);
    G16_LAUNCH_CHECK("msm scatter");
  }
}

size_t msm_accum_scratch_bytes(const MsmPlan& p, int group) {
  return msm_accum_scratch_bytes_impl(p, group);
}

template <>
void Tail<Fq>::merge(const MsmPlan& p, const AccumPtrs& a, hipStream_t s) {
  tail_merge_impl<Fq>(p, a, s);
}

template <>
void Tail<Fq>::reduce(const MsmPlan& p, const AccumPtrs& a, uint32_t B, void* window_out,
                      hipStream_t s) {
  tail_reduce_impl<Fq>(p, a, B, window_out, s);
}

// G2 lives in kernels_msm_g2*.hip, compiled on their own (the slow ones to compile).
void msm_accumulate_g2(const MsmPlan& p, const void* bases, void* scratch, void* window_out,
                       hipStream_t s, hipEvent_t* ev);

void msm_accumulate(const MsmPlan& p, int group, const void* bases, void* scratch,
                    void* window_out, hipStream_t s, hipEvent_t* ev) {
  if (group == 1) accumulate_impl<Fq>(p, bases, scratch, window_out, s, ev);
  else msm_accumulate_g2(p, bases, scratch, window_out, s, ev);
}

}  // namespace g16
