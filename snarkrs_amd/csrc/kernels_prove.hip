// Stage 0 (CSR gather, crates/groth16/src/cpu.rs:140-174) and the pointwise stages
// (C = A o B, crates/groth16/src/cpu.rs:331-337; H = A*B - C, :230-245).
#include "bn254.cuh"
#include "internal.h"

namespace g16 {

__device__ __forceinline__ Fr ldfr4(const uint4* p) {
  uint4 a = p[0], b = p[1];
  return {{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}};
}

__device__ __forceinline__ void stfr4(uint4* p, const Fr& x) {
  p[0] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  p[1] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}

// One thread per row: out[c] = sum_k values[value_id[k]] * w[signal[k]], zero witness values
// skipped and ones added without a multiply (both exact shortcuts, as the reference's).
__global__ void __launch_bounds__(256) k_gather(const uint32_t* __restrict__ row_ptr,
                                                const uint32_t* __restrict__ signal,
                                                const uint32_t* __restrict__ value_id,
                                                const uint4* __restrict__ values,
                                                const uint4* __restrict__ w, uint4* out,
                                                uint32_t n_rows) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_rows) return;
  uint32_t lo = row_ptr[c], hi = row_ptr[c + 1];
  Fr acc = fp_zero<FrCfg>();
  for (uint32_t k = lo; k < hi; ++k) {
    Fr x = ldfr4(w + 2 * (size_t)signal[k]);
    if (is_zero(x)) continue;
    Fr v = ldfr4(values + 2 * (size_t)value_id[k]);
    acc = add(acc, is_one(x) ? v : mul(v, x));
  }
  stfr4(out + 2 * (size_t)c, acc);
}

void launch_gather(const uint32_t* row_ptr, const uint32_t* signal, const uint32_t* value_id,
                   const void* values, const void* witness, void* out, uint32_t n_rows,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3((n_rows + 255) / 256), dim3(256), 0, s, row_ptr, signal,
                     value_id, static_cast<const uint4*>(values),
                     static_cast<const uint4*>(witness), static_cast<uint4*>(out), n_rows);
  G16_LAUNCH_CHECK("gather");
}

// Row c of A.w and of B.w (the shortcuts of k_gather), and C = A o B, in one pass: stage 0
// as the reference's fused gather (crates/cuda/src/kernels/gather.cu:83-129) computes it.
__device__ __forceinline__ Fr gather_row(const uint32_t* __restrict__ signal,
                                         const uint32_t* __restrict__ value_id,
                                         const uint4* __restrict__ values,
                                         const uint4* __restrict__ w, uint32_t lo, uint32_t hi) {
  Fr acc = fp_zero<FrCfg>();
  for (uint32_t k = lo; k < hi; ++k) {
    Fr x = ldfr4(w + 2 * (size_t)signal[k]);
    if (is_zero(x)) continue;
    Fr v = ldfr4(values + 2 * (size_t)value_id[k]);
    acc = add(acc, is_one(x) ? v : mul(v, x));
  }
  return acc;
}

// Grid row y is proof y of a batch: witness y at w + y nv, its (A, B, C) at abc + 3 y n_rows.
__global__ void __launch_bounds__(256) k_gather_abc(const uint32_t* __restrict__ rp_a,
                                                    const uint32_t* __restrict__ sig_a,
                                                    const uint32_t* __restrict__ vid_a,
                                                    const uint32_t* __restrict__ rp_b,
                                                    const uint32_t* __restrict__ sig_b,
                                                    const uint32_t* __restrict__ vid_b,
                                                    const uint4* __restrict__ values,
                                                    const uint4* __restrict__ w, size_t nv,
                                                    uint4* abc, uint32_t n_rows) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  w += 2 * nv * blockIdx.y;
  uint4* a = abc + 2 * (size_t)n_rows * 3 * blockIdx.y;
  uint4* b = a + 2 * (size_t)n_rows;
  uint4* c = b + 2 * (size_t)n_rows;
  const Fr x = gather_row(sig_a, vid_a, values, w, rp_a[r], rp_a[r + 1]);
  const Fr y = gather_row(sig_b, vid_b, values, w, rp_b[r], rp_b[r + 1]);
  stfr4(a + 2 * (size_t)r, x);
  stfr4(b + 2 * (size_t)r, y);
  stfr4(c + 2 * (size_t)r, mul(x, y));
}

void launch_gather_abc(const uint32_t* const row_ptr[2], const uint32_t* const signal[2],
                       const uint32_t* const value_id[2], const void* values,
                       const void* witness, size_t nv, void* abc, uint32_t n_rows,
                       uint32_t batch, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_abc, dim3((n_rows + 255) / 256, batch), dim3(256), 0, s,
                     row_ptr[0], signal[0], value_id[0], row_ptr[1], signal[1], value_id[1],
                     static_cast<const uint4*>(values), static_cast<const uint4*>(witness), nv,
                     static_cast<uint4*>(abc), n_rows);
  G16_LAUNCH_CHECK("gather A, B, C");
}

__global__ void k_mul_into(uint4* out, const uint4* a, const uint4* b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  stfr4(out + 2 * i, mul(ldfr4(a + 2 * i), ldfr4(b + 2 * i)));
}

void launch_mul_into(void* out, const void* a, const void* b, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mul_into, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     static_cast<uint4*>(out), static_cast<const uint4*>(a),
                     static_cast<const uint4*>(b), n);
  G16_LAUNCH_CHECK("pointwise product");
}

// Stage 0 from the caller's rows (h_and_msms_from_rows, crates/groth16/src/cpu.rs:387-424):
// a[0..rows) and b[0..rows) hold A.w and B.w already; C = A*B there, and all three are zero
// on the padding rows (written every proof: the transforms run in place on the slot's buffer).
__global__ void k_rows_abc(uint4* a, uint4* b, uint4* c, size_t rows, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i < rows) {
    stfr4(c + 2 * i, mul(ldfr4(a + 2 * i), ldfr4(b + 2 * i)));
  } else {
    Fr z = fp_zero<FrCfg>();
    stfr4(a + 2 * i, z);
    stfr4(b + 2 * i, z);
    stfr4(c + 2 * i, z);
  }
}

void launch_rows_abc(void* a, void* b, void* c, size_t rows, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_rows_abc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     static_cast<uint4*>(a), static_cast<uint4*>(b), static_cast<uint4*>(c),
                     rows, n);
  G16_LAUNCH_CHECK("rows: C = A*B");
}

// H = A*B - C on the odd coset; no division by Z (folded into the H bases at setup).
__global__ void k_h(uint4* h, const uint4* a, const uint4* b, const uint4* c, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  stfr4(h + 2 * i, sub(mul(ldfr4(a + 2 * i), ldfr4(b + 2 * i)), ldfr4(c + 2 * i)));
}

void launch_h(void* h, const void* a, const void* b, const void* c, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_h, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     static_cast<uint4*>(h), static_cast<const uint4*>(a),
                     static_cast<const uint4*>(b), static_cast<const uint4*>(c), n);
  G16_LAUNCH_CHECK("H = A*B - C");
}

}  // namespace g16
