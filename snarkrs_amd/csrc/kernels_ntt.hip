// Number-theoretic transforms over Fr for stages 1-3 (crates/ntt/src/lib.rs:429-550).
//
// Layout: each vector is n = 2^L Fr elements, 32 bytes each (8 Montgomery u32 limbs),
// several vectors back to back. A launch performs up to R = 4 radix-2 passes on
// independent 2^R-point sub-transforms held in registers: for passes s0 .. s0+R-1 the
// element indices i = hi * 2^(s0+R) + k * 2^s0 + lo (k < 2^R) interact only among
// themselves, so thread (hi, lo) loads its 2^R elements, runs the butterflies and stores
// them back. Consecutive threads take consecutive lo, so each of the 2^R loads of a wave is
// a contiguous run whenever 2^s0 >= 64. Twiddles come from a table of w^i (i < n/2) in HBM.
//
//   DIF (natural -> bit-reversed), passes from s = L-1 down to 0:
//       (u, v) -> (u + v, (u - v) * T[j << (L-1-s)])
//   DIT (bit-reversed -> natural), passes from s = 0 up to L-1:
//       (u, v) -> (u + v * T[j << (L-1-s)], u - v * T[...])
// With T = inverse twiddles the DIF is the unscaled iNTT; with forward twiddles the DIT is
// the forward NTT. The prover's pipeline iNTT -> coset shift -> NTT needs no permutation:
// the coset table is stored in bit-reversed order and multiplied in as the forward DIT loads
// its first group (crates/ntt/src/lib.rs:419-470 uses the same cancellation).
#include "bn254.cuh"
#include "internal.h"

namespace g16 {

__device__ __forceinline__ Fr ldfr(const uint4* p) {
  uint4 a = p[0], b = p[1];
  return {{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}};
}

__device__ __forceinline__ void stfr(uint4* p, const Fr& x) {
  p[0] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  p[1] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}

template <int R, bool DIF>
__global__ void __launch_bounds__(256) k_ntt_group(uint4* data, int L, int s0,
                                                   const uint4* tw, const uint4* prescale,
                                                   uint32_t per_vec) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t vec = blockIdx.y;
  if (t >= per_vec) return;
  uint32_t lo = t & ((1u << s0) - 1);
  uint32_t hi = t >> s0;
  size_t base = ((size_t)hi << (s0 + R)) + lo;
  uint4* v = data + (((size_t)vec << L) * 2);
  Fr x[1 << R];
#pragma unroll
  for (int k = 0; k < (1 << R); ++k) {
    size_t idx = base + ((size_t)k << s0);
    x[k] = ldfr(v + 2 * idx);
    if (!DIF && prescale) x[k] = mul(x[k], ldfr(prescale + 2 * idx));
  }
  if (DIF) {
#pragma unroll
    for (int m = R - 1; m >= 0; --m) {
      int s = s0 + m;
#pragma unroll
      for (int k = 0; k < (1 << R); ++k) {
        if (k & (1 << m)) continue;
        int k1 = k | (1 << m);
        uint32_t j = ((uint32_t)(k & ((1 << m) - 1)) << s0) + lo;
        Fr w = ldfr(tw + 2 * ((size_t)j << (L - 1 - s)));
        Fr u = x[k], d = x[k1];
        x[k] = add(u, d);
        x[k1] = mul(sub(u, d), w);
      }
    }
  } else {
#pragma unroll
    for (int m = 0; m < R; ++m) {
      int s = s0 + m;
#pragma unroll
      for (int k = 0; k < (1 << R); ++k) {
        if (k & (1 << m)) continue;
        int k1 = k | (1 << m);
        uint32_t j = ((uint32_t)(k & ((1 << m) - 1)) << s0) + lo;
        Fr w = ldfr(tw + 2 * ((size_t)j << (L - 1 - s)));
        Fr u = x[k], d = mul(x[k1], w);
        x[k] = add(u, d);
        x[k1] = sub(u, d);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < (1 << R); ++k) stfr(v + 2 * (base + ((size_t)k << s0)), x[k]);
}

template <bool DIF>
static void launch_group(void* data, int nvec, int L, int s0, int R, const void* tw,
                         const void* prescale, hipStream_t s) {
  uint32_t per_vec = 1u << (L - R);
  dim3 grid((per_vec + 255) / 256, nvec);
  auto d = static_cast<uint4*>(data);
  auto t = static_cast<const uint4*>(tw);
  auto p = static_cast<const uint4*>(prescale);
  switch (R) {
    case 1: hipLaunchKernelGGL((k_ntt_group<1, DIF>), grid, dim3(256), 0, s, d, L, s0, t, p, per_vec); break;
    case 2: hipLaunchKernelGGL((k_ntt_group<2, DIF>), grid, dim3(256), 0, s, d, L, s0, t, p, per_vec); break;
    case 3: hipLaunchKernelGGL((k_ntt_group<3, DIF>), grid, dim3(256), 0, s, d, L, s0, t, p, per_vec); break;
    default: hipLaunchKernelGGL((k_ntt_group<4, DIF>), grid, dim3(256), 0, s, d, L, s0, t, p, per_vec); break;
  }
  G16_LAUNCH_CHECK("ntt group");
}

static constexpr int RMAX = 4;

void ntt_dif_inverse(void* data, int nvec, int L, const void* tw, hipStream_t s) {
  if (L == 0) return;
  // Groups from the top: the last (lowest) group takes the remainder.
  int top = L;
  while (top > 0) {
    int R = top >= RMAX ? RMAX : top;
    int s0 = top - R;
    launch_group<true>(data, nvec, L, s0, R, tw, nullptr, s);
    top = s0;
  }
}

void ntt_dit_forward(void* data, int nvec, int L, const void* tw, const void* prescale,
                     hipStream_t s) {
  if (L == 0) {
    if (prescale) launch_scale(data, (size_t)nvec, prescale, s);  // n == 1: one factor
    return;
  }
  int s0 = 0;
  while (s0 < L) {
    int R = (L - s0) >= RMAX ? RMAX : (L - s0);
    launch_group<false>(data, nvec, L, s0, R, tw, s0 == 0 ? prescale : nullptr, s);
    s0 += R;
  }
}

// out[i] = base^i, 64 consecutive elements per thread re-entering the ladder at its start.
__global__ void k_powers(uint4* out, const uint4* basep, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t lo = t * 64;
  if (lo >= n) return;
  Fr base = ldfr(basep);
  Fr acc = fp_one<FrCfg>();
  // base^lo by square-and-multiply.
  Fr b = base;
  for (size_t e = lo; e; e >>= 1) {
    if (e & 1) acc = mul(acc, b);
    b = sqr(b);
  }
  size_t hi = lo + 64 < n ? lo + 64 : n;
  for (size_t i = lo; i < hi; ++i) {
    stfr(out + 2 * i, acc);
    acc = mul(acc, base);
  }
}

void launch_powers(void* out, const void* base, size_t n, hipStream_t s) {
  if (n == 0) return;
  size_t threads = (n + 63) / 64;
  hipLaunchKernelGGL(k_powers, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s,
                     static_cast<uint4*>(out), static_cast<const uint4*>(base), n);
  G16_LAUNCH_CHECK("powers");
}

__device__ __forceinline__ uint32_t bitrev(uint32_t x, int L) {
  return L == 0 ? 0 : (__builtin_bitreverse32(x) >> (32 - L));
}

// out[p] = scale * shift^bitrev(p).
__global__ void k_coset_table(uint4* out, const uint4* shiftp, const uint4* scalep, int L) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (1u << L)) return;
  uint32_t e = bitrev(p, L);
  Fr acc = ldfr(scalep), b = ldfr(shiftp);
  for (; e; e >>= 1) {
    if (e & 1) acc = mul(acc, b);
    b = sqr(b);
  }
  stfr(out + 2 * (size_t)p, acc);
}

void launch_coset_table(void* out, const void* shift, const void* scale, int L, hipStream_t s) {
  uint32_t n = 1u << L;
  hipLaunchKernelGGL(k_coset_table, dim3((n + 255) / 256), dim3(256), 0, s,
                     static_cast<uint4*>(out), static_cast<const uint4*>(shift),
                     static_cast<const uint4*>(scale), L);
  G16_LAUNCH_CHECK("coset table");
}

__global__ void k_bitrev_permute(uint4* dst, const uint4* src, int L) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << L)) return;
  uint32_t j = bitrev(i, L);
  dst[2 * (size_t)j] = src[2 * (size_t)i];
  dst[2 * (size_t)j + 1] = src[2 * (size_t)i + 1];
}

void launch_bitrev_permute(void* dst, const void* src, int L, hipStream_t s) {
  uint32_t n = 1u << L;
  hipLaunchKernelGGL(k_bitrev_permute, dim3((n + 255) / 256), dim3(256), 0, s,
                     static_cast<uint4*>(dst), static_cast<const uint4*>(src), L);
  G16_LAUNCH_CHECK("bitrev permute");
}

__global__ void k_scale(uint4* data, size_t n, const uint4* fp) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  stfr(data + 2 * i, mul(ldfr(data + 2 * i), ldfr(fp)));
}

void launch_scale(void* data, size_t n, const void* factor, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     static_cast<uint4*>(data), n, static_cast<const uint4*>(factor));
  G16_LAUNCH_CHECK("scale");
}

}  // namespace g16
