// Bucket accumulation, merge and window reduction kernels (templated on the coordinate field),
// shared by kernels_msm.hip (G1) and kernels_msm_g2.hip (G2). See kernels_msm.hip for the
// pipeline.
#pragma once
#include "bn254.cuh"
#include "internal.h"

namespace g16 {

inline size_t accum_align_up(size_t x) { return (x + 255) & ~(size_t)255; }

static constexpr int SEG = 32;          // entries per accumulation segment
static constexpr int HEAVY_SPAN = 32;   // segments a merge thread folds itself

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
template <bool INL>
__device__ __forceinline__ void ld_f(Fq2T<INL>& x, const uint4* p) {
  x.c0 = ld_fq(p);
  x.c1 = ld_fq(p + 2);
}
__device__ __forceinline__ void st_f(uint4* p, const Fq& x) { st_fq(p, x); }
template <bool INL>
__device__ __forceinline__ void st_f(uint4* p, const Fq2T<INL>& x) {
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
__global__ void __launch_bounds__(256) k_merge(const uint32_t* offset, uint32_t K, uint4* buckets, const uint4* pfirst,
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

inline uint32_t reduce_group_size(const MsmPlan& p) {
  // About 32k group threads over all windows, at least 2 buckets each.
  uint64_t total = (uint64_t)p.windows * p.nb;
  uint32_t B = 2;
  while ((total / B) > 32768 && B < p.nb) B <<= 1;
  return B > p.nb ? p.nb : B;
}

inline size_t msm_accum_scratch_bytes_impl(const MsmPlan& p, int group) {
  size_t pt = group == 1 ? 128 : 256;
  size_t K = (size_t)p.windows * p.nb;
  size_t nseg = (p.max_entries + SEG - 1) / SEG;
  uint32_t B = reduce_group_size(p);
  size_t G = p.nb / B;
  return accum_align_up(pt * K) + 2 * accum_align_up(pt * (nseg + 1)) + accum_align_up(pt * G * p.windows);
}

// F: the coordinate field of the merge and reduction kernels; FHot: the same field as the
// bucket-accumulation loop computes in (Fq2T<true> for G2: products expanded in place).
template <class F, class FHot = F>
inline void accumulate_impl(const MsmPlan& p, const void* bases, void* scratch, void* window_out,
                            hipStream_t s, hipEvent_t* ev) {
  static_assert(sizeof(F) == sizeof(FHot), "one layout");
  constexpr size_t PT = sizeof(Xyzz<F>);
  size_t K = (size_t)p.windows * p.nb;
  size_t nseg = (p.max_entries + SEG - 1) / SEG;
  uint32_t B = reduce_group_size(p);
  uint32_t G = p.nb / B;
  char* q = static_cast<char*>(scratch);
  uint4* buckets = reinterpret_cast<uint4*>(q);
  q += accum_align_up(PT * K);
  uint4* pfirst = reinterpret_cast<uint4*>(q);
  q += accum_align_up(PT * (nseg + 1));
  uint4* plast = reinterpret_cast<uint4*>(q);
  q += accum_align_up(PT * (nseg + 1));
  uint4* groups = reinterpret_cast<uint4*>(q);
  auto bp = static_cast<const uint4*>(bases);

  if (ev) G16_CHECK(hipEventRecord(ev[0], s));
  if (nseg) {
    hipLaunchKernelGGL(k_accumulate<FHot>, dim3((unsigned)((nseg + 127) / 128)), dim3(128), 0, s,
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


}  // namespace g16
