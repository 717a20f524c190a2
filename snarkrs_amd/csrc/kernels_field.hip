// Field probe (device arithmetic checked element by element against the oracle) and the
// batched fixed-base multiplication the synthetic setup uses.
#include <type_traits>

#include "bn254.cuh"
#include "internal.h"

namespace g16 {

template <class C>
__device__ Fp<C> ld(const uint8_t* p, size_t i) {
  const uint4* q = reinterpret_cast<const uint4*>(p + 32 * i);
  uint4 a = q[0], b = q[1];
  return {{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}};
}

template <class C>
__device__ void st(uint8_t* p, size_t i, const Fp<C>& x) {
  uint4* q = reinterpret_cast<uint4*>(p + 32 * i);
  q[0] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  q[1] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}

template <class C>
__global__ void k_probe_fp(int op, const uint8_t* a, const uint8_t* b, uint8_t* out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fp<C> x = ld<C>(a, i);
  Fp<C> y = b ? ld<C>(b, i) : fp_zero<C>();
  Fp<C> r;
  switch (op) {
    case 0: r = add(x, y); break;
    case 1: r = sub(x, y); break;
    case 2: r = mul(x, y); break;
    case 3: r = sqr(x); break;
    case 4: r = inv(x); break;
    case 5: r = neg(x); break;
    case 6: r = from_std(x); break;
    default: r = to_std(x); break;
  }
  // The MSM form keeps values in [0, 2q); the probe reports them fully reduced.
  if constexpr (std::is_same<C, FqmCfg>::value) reduce_once<FqCfg>(r.v);
  st<C>(out, i, r);
}

__global__ void k_probe_fq2(int op, const uint8_t* a, const uint8_t* b, uint8_t* out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fq2 x = {ld<FqCfg>(a, 2 * i), ld<FqCfg>(a, 2 * i + 1)};
  Fq2 y = b ? Fq2{ld<FqCfg>(b, 2 * i), ld<FqCfg>(b, 2 * i + 1)} : fq2_zero();
  Fq2 r;
  switch (op) {
    case 0: r = add(x, y); break;
    case 1: r = sub(x, y); break;
    case 2: r = mul(x, y); break;
    case 3: r = sqr(x); break;
    case 4: r = inv(x); break;
    case 5: r = neg(x); break;
    case 6: r = {from_std(x.c0), from_std(x.c1)}; break;
    default: r = {to_std(x.c0), to_std(x.c1)}; break;
  }
  st<FqCfg>(out, 2 * i, r.c0);
  st<FqCfg>(out, 2 * i + 1, r.c1);
}

void launch_field_probe(int field, int op, const void* a, const void* b, void* out, size_t n,
                        hipStream_t s) {
  if (n == 0) return;
  unsigned grid = (unsigned)((n + 255) / 256);
  auto A = static_cast<const uint8_t*>(a);
  auto B = static_cast<const uint8_t*>(b);
  auto O = static_cast<uint8_t*>(out);
  if (field == 0)
    hipLaunchKernelGGL(k_probe_fp<FrCfg>, dim3(grid), dim3(256), 0, s, op, A, B, O, n);
  else if (field == 1)
    hipLaunchKernelGGL(k_probe_fp<FqCfg>, dim3(grid), dim3(256), 0, s, op, A, B, O, n);
  else if (field == 2)
    hipLaunchKernelGGL(k_probe_fq2, dim3(grid), dim3(256), 0, s, op, A, B, O, n);
  else
    hipLaunchKernelGGL(k_probe_fp<FqmCfg>, dim3(grid), dim3(256), 0, s, op, A, B, O, n);
  G16_LAUNCH_CHECK("field probe");
}

// ---- fixed base: out[i] = k_i * G, affine -------------------------------------------------
__device__ void st_aff(uint8_t* p, size_t i, const G1Aff& a) {
  st<FqCfg>(p, 2 * i, a.x);
  st<FqCfg>(p, 2 * i + 1, a.y);
}
__device__ void st_aff(uint8_t* p, size_t i, const G2Aff& a) {
  st<FqCfg>(p, 4 * i, a.x.c0);
  st<FqCfg>(p, 4 * i + 1, a.x.c1);
  st<FqCfg>(p, 4 * i + 2, a.y.c0);
  st<FqCfg>(p, 4 * i + 3, a.y.c1);
}

template <class F>
__global__ void __launch_bounds__(64) k_fixed_base(const uint8_t* k, uint8_t* out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fr s = to_std(ld<FrCfg>(k, i));
  Aff<F> g;
  if constexpr (sizeof(F) == sizeof(Fq)) g = g1_generator();
  else g = g2_generator();
  Xyzz<F> r = xyzz_mul_plain(g, s.v);
  st_aff(out, i, xyzz_to_aff(r));
}

void launch_fixed_base(int group, const void* k, size_t n, void* out, hipStream_t s) {
  if (n == 0) return;
  unsigned grid = (unsigned)((n + 63) / 64);
  auto K = static_cast<const uint8_t*>(k);
  auto O = static_cast<uint8_t*>(out);
  if (group == 1)
    hipLaunchKernelGGL(k_fixed_base<Fq>, dim3(grid), dim3(64), 0, s, K, O, n);
  else
    hipLaunchKernelGGL(k_fixed_base<Fq2>, dim3(grid), dim3(64), 0, s, K, O, n);
  G16_LAUNCH_CHECK("fixed base");
}

}  // namespace g16

// ---- VALU ceilings (the method of the reference's bench/results/device-microbench.md:7-10:
// dependent-free chains, long enough that launch cost vanishes) ---------------------------
namespace g16 {

// 8 independent 32x32+64 multiply-add chains per lane.
__global__ void __launch_bounds__(256) k_mad_peak(const uint32_t* in, uint64_t* out, int iters) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a[8];
  uint64_t acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = in[(t + k) & 1023];
    acc[k] = a[k];
  }
  uint32_t b = in[t & 1023] | 1u;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = (uint64_t)a[k] * b + acc[k];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = (uint32_t)acc[k];
  }
  uint64_t x = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) x ^= acc[k];
  out[t] = x;
}

// 4 independent chains of Fq Montgomery products per lane.
__global__ void __launch_bounds__(256) k_fqmul_peak(const uint32_t* in, uint32_t* out, int iters) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  Fq x[4], y;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int j = 0; j < 8; ++j) x[k].v[j] = in[(t * 8 + k * 3 + j) & 1023] & 0x0fffffffu;
#pragma unroll
  for (int j = 0; j < 8; ++j) y.v[j] = in[(t + j) & 1023] & 0x0fffffffu;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = mul(x[k], y);
  }
  uint32_t h = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int j = 0; j < 8; ++j) h ^= x[k].v[j];
  out[t] = h;
}

void launch_valu_peak(int kind, const void* in, void* out, unsigned blocks, int iters,
                      hipStream_t s) {
  if (kind == 0)
    hipLaunchKernelGGL(k_mad_peak, dim3(blocks), dim3(256), 0, s,
                       static_cast<const uint32_t*>(in), static_cast<uint64_t*>(out), iters);
  else
    hipLaunchKernelGGL(k_fqmul_peak, dim3(blocks), dim3(256), 0, s,
                       static_cast<const uint32_t*>(in), static_cast<uint32_t*>(out), iters);
  G16_LAUNCH_CHECK("valu peak");
}

}  // namespace g16
