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
#include "bn254.cuh"
#include "internal.h"

namespace g16 {

static constexpr int SEG = 32;          // entries per accumulation segment
static constexpr int HEAVY_SPAN = 32;   // segments a merge thread folds itself
static constexpr int RECODE_BITS = 255;

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
__device__ __forceinline__ void ld_f(Fq& x, const uint4* p) { x = ld_fq(p); }
__device__ __forceinline__ void ld_f(Fq2& x, const uint4* p) {
  x.c0 = ld_fq(p);
  x.c1 = ld_fq(p + 2);
}
__device__ __forceinline__ void st_f(uint4* p, const Fq& x) { st_fq(p, x); }
__device__ __forceinline__ void st_f(uint4* p, const Fq2& x) {
  st_fq(p, x.c0);
  st_fq(p + 2, x.c1);
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

// ---------------------------------------------------------------------------------------
// digits and the counting sort
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t read_bits(const uint32_t s[8], int off, int width) {
  int idx = off >> 5;
  if (idx >= 8) return 0;
  int sh = off & 31;
  uint64_t buf = s[idx];
  if (idx + 1 < 8) buf |= (uint64_t)s[idx + 1] << 32;
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

__global__ void __launch_bounds__(256) k_digits_count(const uint4* scalars, uint32_t n, int c,
                                                      int W, uint32_t nb, uint4* std_out,
                                                      uint32_t* count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fr s = to_std(ld_fr(scalars + 2 * (size_t)i));
  st_fr(std_out + 2 * (size_t)i, s);
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
  if (i >= n) return;
  Fr s = ld_fr(std_in + 2 * (size_t)i);
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
    total[1] = 0;  // heavy bucket count
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
         align_up((size_t)n * 32) + align_up(16) + align_up(4 * K);
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
  p.heavy = reinterpret_cast<uint32_t*>(take(4 * K));
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

// ---------------------------------------------------------------------------------------
// bucket accumulation
// ---------------------------------------------------------------------------------------
template <class F>
__global__ void __launch_bounds__(128) k_accumulate(const uint32_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ vals,
                                                    const uint32_t* total, size_t nseg,
                                                    const uint4* __restrict__ bases,
                                                    uint4* buckets, uint4* pfirst,
                                                    uint4* plast) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nseg) return;
  uint32_t T = total[0];
  size_t b = t * SEG;
  if (b >= T) return;
  size_t e = b + SEG < T ? b + SEG : T;
  uint32_t cur = keys[b];
  bool starts = (b == 0) || keys[b - 1] != cur;
  size_t rs = b;
  Xyzz<F> acc = xyzz_inf<F>();
  for (size_t k = b; k < e; ++k) {
    uint32_t key = keys[k];
    if (key != cur) {
      // Run [rs, k) of bucket `cur` ends inside the segment.
      if (starts) st_xyzz(buckets, cur, acc);
      else st_xyzz(pfirst, t, acc);  // rs == b: the run began in an earlier segment
      acc = xyzz_inf<F>();
      cur = key;
      starts = true;
      rs = k;
    }
    uint32_t v = vals[k];
    Aff<F> a = ld_aff<F>(bases, v >> 1);
    if (v & 1) a = aff_neg(a);
    acc = xyzz_add_aff(acc, a);
  }
  bool ends = (e == T) || keys[e] != cur;
  if (starts && ends) {
    st_xyzz(buckets, cur, acc);
  } else {
    if (rs == b) st_xyzz(pfirst, t, acc);
    st_xyzz(plast, t, acc);
  }
}

// Buckets that span segments: one thread folds a short span, a heavy one is queued.
template <class F>
__global__ void k_merge(const uint32_t* offset, uint32_t K, uint4* buckets, const uint4* pfirst,
                        const uint4* plast, uint32_t* total, uint32_t* heavy) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  uint32_t s = offset[k], e = offset[k + 1];
  if (s == e) {
    st_xyzz(buckets, k, xyzz_inf<F>());
    return;
  }
  uint32_t sa = s / SEG, sb = (e - 1) / SEG;
  if (sa == sb) return;  // stored whole by the accumulation
  if (sb - sa > HEAVY_SPAN) {
    uint32_t slot = atomicAdd(&total[1], 1u);
    heavy[slot] = k;
    return;
  }
  Xyzz<F> acc = ld_xyzz<F>(plast, sa);
  for (uint32_t t = sa + 1; t < sb; ++t) acc = xyzz_add(acc, ld_xyzz<F>(pfirst, t));
  acc = xyzz_add(acc, ld_xyzz<F>(pfirst, sb));
  st_xyzz(buckets, k, acc);
}

// One workgroup per heavy bucket: threads stride over the interior segments, then an LDS
// tree, then the two end pieces.
template <class F, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_heavy(const uint32_t* offset, const uint32_t* total,
                                                 const uint32_t* heavy, uint4* buckets,
                                                 const uint4* pfirst, const uint4* plast) {
  __shared__ Xyzz<F> sh[BLOCK];
  uint32_t nh = total[1];
  for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
    uint32_t k = heavy[h];
    uint32_t s = offset[k], e = offset[k + 1];
    uint32_t sa = s / SEG, sb = (e - 1) / SEG;
    Xyzz<F> acc = xyzz_inf<F>();
    for (uint32_t t = sa + 1 + threadIdx.x; t < sb; t += BLOCK)
      acc = xyzz_add(acc, ld_xyzz<F>(pfirst, t));
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) sh[threadIdx.x] = xyzz_add(sh[threadIdx.x], sh[threadIdx.x + off]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      Xyzz<F> r = xyzz_add(ld_xyzz<F>(plast, sa), sh[0]);
      r = xyzz_add(r, ld_xyzz<F>(pfirst, sb));
      st_xyzz(buckets, k, r);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------
// window reduction: S_w = sum_j (j+1) B_{w,j}
// ---------------------------------------------------------------------------------------
// Group g of window w covers buckets [g*B, (g+1)*B): X = sum_i (i+1) B_{gB+i} + gB * sum_i B.
template <class F>
__global__ void __launch_bounds__(128) k_reduce_groups(const uint4* buckets, uint32_t nb,
                                                       int W, uint32_t B, uint4* out) {
  uint32_t G = nb / B;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint32_t)W * G) return;
  uint32_t w = t / G, g = t % G;
  size_t base = (size_t)w * nb + (size_t)g * B;
  Xyzz<F> run = xyzz_inf<F>(), acc = xyzz_inf<F>();
  for (int i = (int)B - 1; i >= 0; --i) {
    run = xyzz_add(run, ld_xyzz<F>(buckets, base + i));
    acc = xyzz_add(acc, run);
  }
  acc = xyzz_add(acc, xyzz_mul_small(run, g * B));
  st_xyzz(out, t, acc);
}

template <class F, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_reduce_windows(const uint4* groups, uint32_t G,
                                                          uint4* out) {
  __shared__ Xyzz<F> sh[BLOCK];
  uint32_t w = blockIdx.x;
  Xyzz<F> acc = xyzz_inf<F>();
  for (uint32_t g = threadIdx.x; g < G; g += BLOCK)
    acc = xyzz_add(acc, ld_xyzz<F>(groups, (size_t)w * G + g));
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int off = BLOCK / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] = xyzz_add(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) st_xyzz(out, w, sh[0]);
}

static uint32_t reduce_group_size(const MsmPlan& p) {
  // About 32k group threads over all windows, at least 2 buckets each.
  uint64_t total = (uint64_t)p.windows * p.nb;
  uint32_t B = 2;
  while ((total / B) > 32768 && B < p.nb) B <<= 1;
  return B > p.nb ? p.nb : B;
}

size_t msm_accum_scratch_bytes(const MsmPlan& p, int group) {
  size_t pt = group == 1 ? 128 : 256;
  size_t K = (size_t)p.windows * p.nb;
  size_t nseg = (p.max_entries + SEG - 1) / SEG;
  uint32_t B = reduce_group_size(p);
  size_t G = p.nb / B;
  return align_up(pt * K) + 2 * align_up(pt * (nseg + 1)) + align_up(pt * G * p.windows);
}

template <class F>
static void accumulate_impl(const MsmPlan& p, const void* bases, void* scratch, void* window_out,
                            hipStream_t s, hipEvent_t* ev) {
  constexpr size_t PT = sizeof(Xyzz<F>);
  size_t K = (size_t)p.windows * p.nb;
  size_t nseg = (p.max_entries + SEG - 1) / SEG;
  uint32_t B = reduce_group_size(p);
  uint32_t G = p.nb / B;
  char* q = static_cast<char*>(scratch);
  uint4* buckets = reinterpret_cast<uint4*>(q);
  q += align_up(PT * K);
  uint4* pfirst = reinterpret_cast<uint4*>(q);
  q += align_up(PT * (nseg + 1));
  uint4* plast = reinterpret_cast<uint4*>(q);
  q += align_up(PT * (nseg + 1));
  uint4* groups = reinterpret_cast<uint4*>(q);
  auto bp = static_cast<const uint4*>(bases);

  if (ev) G16_CHECK(hipEventRecord(ev[0], s));
  if (nseg) {
    hipLaunchKernelGGL(k_accumulate<F>, dim3((unsigned)((nseg + 127) / 128)), dim3(128), 0, s,
                       p.keys, p.vals, p.total, nseg, bp, buckets, pfirst, plast);
    G16_LAUNCH_CHECK("msm accumulate");
  }
  if (ev) G16_CHECK(hipEventRecord(ev[1], s));
  hipLaunchKernelGGL(k_merge<F>, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, s, p.offset,
                     (uint32_t)K, buckets, pfirst, plast, p.total, p.heavy);
  G16_LAUNCH_CHECK("msm merge");
  constexpr int HB = sizeof(F) == sizeof(Fq) ? 128 : 64;
  hipLaunchKernelGGL((k_heavy<F, HB>), dim3(256), dim3(HB), 0, s, p.offset, p.total, p.heavy,
                     buckets, pfirst, plast);
  G16_LAUNCH_CHECK("msm heavy merge");
  if (ev) G16_CHECK(hipEventRecord(ev[2], s));
  uint32_t nthreads = (uint32_t)p.windows * G;
  hipLaunchKernelGGL(k_reduce_groups<F>, dim3((nthreads + 127) / 128), dim3(128), 0, s, buckets,
                     p.nb, p.windows, B, groups);
  G16_LAUNCH_CHECK("msm reduce groups");
  constexpr int RB = sizeof(F) == sizeof(Fq) ? 128 : 64;
  hipLaunchKernelGGL((k_reduce_windows<F, RB>), dim3(p.windows), dim3(RB), 0, s, groups, G,
                     static_cast<uint4*>(window_out));
  G16_LAUNCH_CHECK("msm reduce windows");
  if (ev) G16_CHECK(hipEventRecord(ev[3], s));
}

void msm_accumulate(const MsmPlan& p, int group, const void* bases, void* scratch,
                    void* window_out, hipStream_t s, hipEvent_t* ev) {
  if (group == 1) accumulate_impl<Fq>(p, bases, scratch, window_out, s, ev);
  else accumulate_impl<Fq2>(p, bases, scratch, window_out, s, ev);
}

}  // namespace g16
