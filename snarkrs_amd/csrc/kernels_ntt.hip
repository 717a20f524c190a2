// Number-theoretic transforms over Fr for stages 1-3 (crates/ntt/src/lib.rs:429-550).
//
// Layout: each vector is n = 2^L Fr elements, 32 bytes each (8 Montgomery u32 limbs),
// several vectors back to back, transformed in place by LDS-tiled passes of several radix-2
// stages each (k_ntt_pass below). Stage s reads twiddles T_s[j] = w^(j 2^(L-1-s)), j < 2^s,
// from per-pass tables stored contiguously at offset 2^s - 1.
//
//   DIF (natural -> bit-reversed), stages from s = L-1 down to 0:
//       (u, v) -> (u + v, (u - v) * T_s[j])
//   DIT (bit-reversed -> natural), stages from s = 0 up to L-1:
//       (u, v) -> (u + v * T_s[j], u - v * T_s[j])
// The tables (twiddles and the coset table) are kept in the Montgomery form R' = 2^261, and
// data are multiplied by them with the 29-bit product (mul_rp): data stay in the R = 2^256
// form and the products need no carry flags.
// With T = inverse twiddles the DIF is the unscaled iNTT; with forward twiddles the DIT is
// the forward NTT. The prover's pipeline iNTT -> coset shift -> NTT needs no permutation:
// the coset table is stored in bit-reversed order and multiplied in as the forward DIT loads
// first pass (crates/ntt/src/lib.rs:419-470 uses the same cancellation).
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

// One pass = B consecutive radix-2 stages s0 .. s0+B-1, done in LDS. For those stages the
// elements i = hi 2^(s0+B) + k 2^s0 + lo (k < 2^B) interact only among themselves; a
// workgroup takes one hi and TLO consecutive lo (2^B TLO <= 2048 elements, 64 KB of LDS held
// as eight 32-bit planes, so neighbouring lanes hit neighbouring banks), loads them with
// consecutive lanes on consecutive lo, runs the B stages with a barrier between stages, and
// stores them back. The pass over the lowest stages (s0 = 0) uses TLO = 1 and reads 2^11
// contiguous elements; higher passes use B <= 6, so every load is a run of >= 32 elements.
static constexpr int NTT_TILE = 2048;  // elements per workgroup
static constexpr int NTT_THREADS = 256;

__device__ __forceinline__ void lds_put(uint32_t* lds, int e, const Fr& x) {
#pragma unroll
  for (int l = 0; l < 8; ++l) lds[l * NTT_TILE + e] = x.v[l];
}
__device__ __forceinline__ Fr lds_get(const uint32_t* lds, int e) {
  Fr x;
#pragma unroll
  for (int l = 0; l < 8; ++l) x.v[l] = lds[l * NTT_TILE + e];
  return x;
}

template <bool DIF>
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_pass(uint4* data, int L, int s0, int B,
                                                          int tlo, const uint4* tw,
                                                          const uint4* prescale) {
  __shared__ uint32_t lds[8 * NTT_TILE];
  const int E = tlo << B;  // elements in this tile
  const uint32_t tiles_lo = (1u << s0) / (uint32_t)tlo;
  const uint32_t hi = blockIdx.x / tiles_lo;
  const uint32_t lo0 = (blockIdx.x % tiles_lo) * (uint32_t)tlo;
  uint4* v = data + (((size_t)blockIdx.y << L) * 2);
  const size_t base = ((size_t)hi << (s0 + B)) + lo0;
  for (int e = threadIdx.x; e < E; e += NTT_THREADS) {
    const int k = e / tlo, l = e % tlo;
    const size_t idx = base + ((size_t)k << s0) + l;
    Fr x = ldfr(v + 2 * idx);
    if (!DIF && prescale) x = mul_rp(x, ldfr(prescale + 2 * idx));
    lds_put(lds, e, x);
  }
  __syncthreads();
  for (int step = 0; step < B; ++step) {
    const int m = DIF ? B - 1 - step : step;  // stage s0 + m
    const int s = s0 + m;
    const uint4* T = tw + 2 * (((size_t)1 << s) - 1);
    for (int u = threadIdx.x; u < E / 2; u += NTT_THREADS) {
      const int q = u / tlo, l = u % tlo;
      const int k = ((q >> m) << (m + 1)) | (q & ((1 << m) - 1));  // bit m of k clear
      const int e0 = k * tlo + l, e1 = e0 + (tlo << m);
      const uint32_t j = ((uint32_t)(k & ((1 << m) - 1)) << s0) + lo0 + (uint32_t)l;
      const Fr w = ldfr(T + 2 * (size_t)j);
      const Fr a = lds_get(lds, e0), c = lds_get(lds, e1);
      if (DIF) {
        lds_put(lds, e0, add(a, c));
        lds_put(lds, e1, mul_rp(sub(a, c), w));
      } else {
        const Fr d = mul_rp(c, w);
        lds_put(lds, e0, add(a, d));
        lds_put(lds, e1, sub(a, d));
      }
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < E; e += NTT_THREADS) {
    const int k = e / tlo, l = e % tlo;
    stfr(v + 2 * (base + ((size_t)k << s0) + l), lds_get(lds, e));
  }
}

template <bool DIF>
static void launch_pass(void* data, int nvec, int L, int s0, int B, const void* tw,
                        const void* prescale, hipStream_t s) {
  int tlo = NTT_TILE >> B;
  if (tlo > (1 << s0)) tlo = 1 << s0;
  const uint32_t tiles = (1u << (L - B)) / (uint32_t)tlo;
  hipLaunchKernelGGL(k_ntt_pass<DIF>, dim3(tiles, nvec), dim3(NTT_THREADS), 0, s,
                     static_cast<uint4*>(data), L, s0, B, tlo, static_cast<const uint4*>(tw),
                     static_cast<const uint4*>(prescale));
  G16_LAUNCH_CHECK("ntt pass");
}

// Pass plan: the lowest 11 stages in one pass (contiguous tiles), the rest in passes of at
// most 6 stages.
static constexpr int LOW_PASS = 11, HIGH_PASS = 6;

void ntt_dif_inverse(void* data, int nvec, int L, const void* tw, hipStream_t s) {
  if (L == 0) return;
  const int low = L < LOW_PASS ? L : LOW_PASS;
  int top = L;
  while (top > low) {
    const int rest = top - low;
    const int B = rest < HIGH_PASS ? rest : HIGH_PASS;
    launch_pass<true>(data, nvec, L, top - B, B, tw, nullptr, s);
    top -= B;
  }
  launch_pass<true>(data, nvec, L, 0, low, tw, nullptr, s);
}

void ntt_dit_forward(void* data, int nvec, int L, const void* tw, const void* prescale,
                     hipStream_t s) {
  if (L == 0) {
    if (prescale) launch_scale_rp(data, (size_t)nvec, prescale, s);  // n == 1: one factor
    return;
  }
  const int low = L < LOW_PASS ? L : LOW_PASS;
  launch_pass<false>(data, nvec, L, 0, low, tw, prescale, s);
  int s0 = low;
  while (s0 < L) {
    const int B = (L - s0) < HIGH_PASS ? (L - s0) : HIGH_PASS;
    launch_pass<false>(data, nvec, L, s0, B, tw, nullptr, s);
    s0 += B;
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

// data[i] *= table[i] (table in the R' form).
__global__ void k_mul_table_rp(uint4* data, size_t n, const uint4* table) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  stfr(data + 2 * i, mul_rp(ldfr(data + 2 * i), ldfr(table + 2 * i)));
}

void launch_mul_table_rp(void* data, size_t n, const void* table_rp, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mul_table_rp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     static_cast<uint4*>(data), n, static_cast<const uint4*>(table_rp));
  G16_LAUNCH_CHECK("table product");
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
