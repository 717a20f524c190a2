// Number-theoretic transforms over Fr for stages 1-3 (crates/ntt/src/lib.rs:429-550).
//
// Layout: each vector is n = 2^L Fr elements, 32 bytes each (8 Montgomery u32 limbs),
// several vectors back to back. A launch performs up to R = 4 radix-2 passes on
// independent 2^R-point sub-transforms held in registers: for passes s0 .. s0+R-1 the
// element indices i = hi * 2^(s0+R) + k * 2^s0 + lo (k < 2^R) interact only among
// themselves, so thread (hi, lo) loads its 2^R elements, runs the butterflies and stores
// them back. Consecutive threads take consecutive lo, so each of the 2^R loads of a wave is
// a contiguous run whenever 2^s0 >= 64. Twiddles come from per-pass tables in HBM: pass s
// reads T_s[j] = w^(j 2^(L-1-s)) for j < 2^s, stored contiguously at offset 2^s - 1, so the
// twiddle loads of a wave are as contiguous as its data loads.
//
//   DIF (natural -> bit-reversed), passes from s = L-1 down to 0:
//       (u, v) -> (u + v, (u - v) * T_s[j])
//   DIT (bit-reversed -> natural), passes from s = 0 up to L-1:
//       (u, v) -> (u + v * T_s[j], u - v * T_s[j])
// The tables (twiddles and the coset table) are kept in the Montgomery form R' = 2^261, and
// data are multiplied by them with the 29-bit product (mul_rp): data stay in the R = 2^256
// form and the products need no carry flags.
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

// The 2^R elements of a thread stay in registers only while every index into x[] is a
// compile-time constant; full unrolling of loops over field products exceeds the compiler's
// unroll budget (the array then lands in scratch memory), so the butterflies, loads and stores
// below are expanded by template recursion instead of loops.
template <int R, int M, int K, bool DIF>
__device__ __forceinline__ void ntt_pass(Fr (&x)[1 << R], const uint4* tw, int s0, uint32_t lo) {
  if constexpr (K < (1 << R)) {
    if constexpr ((K & (1 << M)) == 0) {
      constexpr int K1 = K | (1 << M);
      const int s = s0 + M;
      const uint32_t j = ((uint32_t)(K & ((1 << M) - 1)) << s0) + lo;
      const Fr w = ldfr(tw + 2 * (((size_t)1 << s) - 1 + j));
      if constexpr (DIF) {
        const Fr u = x[K], d = x[K1];
        x[K] = add(u, d);
        x[K1] = mul_rp(sub(u, d), w);
      } else {
        const Fr u = x[K], d = mul_rp(x[K1], w);
        x[K] = add(u, d);
        x[K1] = sub(u, d);
      }
    }
    ntt_pass<R, M, K + 1, DIF>(x, tw, s0, lo);
  }
}

// DIF runs passes M = R-1 down to 0, DIT 0 up to R-1.
template <int R, int M, bool DIF>
__device__ __forceinline__ void ntt_passes(Fr (&x)[1 << R], const uint4* tw, int s0,
                                           uint32_t lo) {
  if constexpr (M >= 0 && M < R) {
    ntt_pass<R, M, 0, DIF>(x, tw, s0, lo);
    ntt_passes<R, DIF ? M - 1 : M + 1, DIF>(x, tw, s0, lo);
  }
}

template <int R, int K>
__device__ __forceinline__ void ntt_load(Fr (&x)[1 << R], const uint4* v, size_t base, int s0,
                                         const uint4* prescale) {
  if constexpr (K < (1 << R)) {
    const size_t idx = base + ((size_t)K << s0);
    x[K] = ldfr(v + 2 * idx);
    if (prescale) x[K] = mul_rp(x[K], ldfr(prescale + 2 * idx));
    ntt_load<R, K + 1>(x, v, base, s0, prescale);
  }
}

template <int R, int K>
__device__ __forceinline__ void ntt_store(const Fr (&x)[1 << R], uint4* v, size_t base, int s0) {
  if constexpr (K < (1 << R)) {
    stfr(v + 2 * (base + ((size_t)K << s0)), x[K]);
    ntt_store<R, K + 1>(x, v, base, s0);
  }
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
  ntt_load<R, 0>(x, v, base, s0, DIF ? nullptr : prescale);
  ntt_passes<R, DIF ? R - 1 : 0, DIF>(x, tw, s0, lo);
  ntt_store<R, 0>(x, v, base, s0);
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
    if (prescale) launch_scale_rp(data, (size_t)nvec, prescale, s);  // n == 1: one factor
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

__global__ void k_scale_rp(uint4* data, size_t n, const uint4* fp) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  stfr(data + 2 * i, mul_rp(ldfr(data + 2 * i), ldfr(fp)));
}

void launch_scale_rp(void* data, size_t n, const void* factor_rp, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_scale_rp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     static_cast<uint4*>(data), n, static_cast<const uint4*>(factor_rp));
  G16_LAUNCH_CHECK("scale");
}

// x -> 32 x mod r: the R = 2^256 form to R' = 2^261.
__device__ __forceinline__ Fr fr_to_rp(Fr x) {
#pragma unroll
  for (int k = 0; k < 5; ++k) x = add(x, x);
  return x;
}

__global__ void k_fr_to_rp(uint4* data, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) stfr(data + 2 * i, fr_to_rp(ldfr(data + 2 * i)));
}

void launch_fr_to_rp(void* data, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fr_to_rp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     static_cast<uint4*>(data), n);
  G16_LAUNCH_CHECK("to R' form");
}

// Per-pass twiddle tables from the powers w^i (i < n/2): entry 2^s - 1 + j = w^(j 2^(L-1-s)),
// j < 2^s, for s < L, in the R' form.
__global__ void k_stage_tables(uint4* out, const uint4* powers, int L) {
  size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g + 1 >= ((size_t)1 << L)) return;
  int s = 63 - __builtin_clzll((unsigned long long)(g + 1));
  size_t j = g + 1 - ((size_t)1 << s);
  stfr(out + 2 * g, fr_to_rp(ldfr(powers + 2 * (j << (L - 1 - s)))));
}

void launch_stage_tables(void* out, const void* powers, int L, hipStream_t s) {
  size_t n = ((size_t)1 << L) - 1;
  if (n == 0) return;
  hipLaunchKernelGGL(k_stage_tables, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     static_cast<uint4*>(out), static_cast<const uint4*>(powers), L);
  G16_LAUNCH_CHECK("twiddle tables");
}

}  // namespace g16
