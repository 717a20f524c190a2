// BN254 field and curve arithmetic for gfx950, usable from host and device code.
//
// Representation (the packed wire format of crates/gpu-layout/src/lib.rs:75-185):
//   * Fr, Fq: 8 little-endian u32 limbs, Montgomery form with R = 2^256, fully reduced;
//   * Fq2 = Fq[u]/(u^2 + 1), c0 then c1;
//   * affine points x, y with the all-zero encoding for infinity;
//   * accumulators in XYZZ coordinates (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2), infinity = ZZ 0.
//
// The multiply is CIOS over 32-bit limbs: every partial product is a 32x32->64 VALU
// multiply-add (v_mad_u64_u32), carries ride 64-bit adds (v_add_co/v_addc); there is no
// MFMA here because a Montgomery product is not a dense contraction. Constants are pinned by
// tests/test_gpu_field.py against Python big integers and the CPU oracle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HD __host__ __device__ __forceinline__

namespace g16 {

struct FrCfg {
  // r = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001
  static HD uint32_t p(int i) {
    constexpr uint32_t P[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                               0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
    return P[i];
  }
  static HD uint32_t one(int i) {  // R mod r
    constexpr uint32_t R1[8] = {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
                                0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
    return R1[i];
  }
  static HD uint32_t r2(int i) {  // R^2 mod r
    constexpr uint32_t R2[8] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                                0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};
    return R2[i];
  }
  static constexpr uint32_t N0 = 0xefffffffu;  // -r^-1 mod 2^32
};

struct FqCfg {
  // q = 0x30644e72e131a029b85045b68181585d97816a916871ca8d3c208c16d87cfd47
  static HD uint32_t p(int i) {
    constexpr uint32_t P[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                               0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
    return P[i];
  }
  static HD uint32_t one(int i) {
    constexpr uint32_t R1[8] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
    return R1[i];
  }
  static HD uint32_t r2(int i) {
    constexpr uint32_t R2[8] = {0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
                                0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u};
    return R2[i];
  }
  static constexpr uint32_t N0 = 0xe4866389u;
};

// Fq for the MSM kernels: the same modulus in Montgomery form with R' = 2^261, whose product
// runs over nine 29-bit limbs (mul below). x_R' = 32 x_R mod q; the MSM converts its bases
// once (to_msm_form) and its window rows back on the host (from_msm_form).
struct FqmCfg {
  static HD uint32_t p(int i) { return FqCfg::p(i); }
  static HD uint32_t one(int i) {  // 2^261 mod q
    constexpr uint32_t R1[8] = {0x157ccc21u, 0x4e8384ebu, 0x0ce148c3u, 0xfb90a602u,
                                0x819caa36u, 0x5301fa84u, 0x563d4475u, 0x0dc83629u};
    return R1[i];
  }
  static HD uint32_t r2(int i) {  // 2^522 mod q
    constexpr uint32_t R2[8] = {0x659bac10u, 0xe1a2a074u, 0x5406005au, 0x63985586u,
                                0x2d3e2632u, 0xff54c580u, 0x34ea65a6u, 0x2a11a68cu};
    return R2[i];
  }
  static constexpr uint32_t N0 = FqCfg::N0;
};

template <class C>
struct Fp {
  uint32_t v[8];
};

typedef Fp<FrCfg> Fr;
typedef Fp<FqCfg> Fq;
typedef Fp<FqmCfg> Fm;

template <class C>
HD Fp<C> fp_zero() {
  Fp<C> r;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = 0;
  return r;
}

template <class C>
HD Fp<C> fp_one() {
  Fp<C> r;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = C::one(i);
  return r;
}

template <class C>
HD bool is_zero(const Fp<C>& a) {
  uint32_t x = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) x |= a.v[i];
  return x == 0;
}

template <class C>
HD bool eq(const Fp<C>& a, const Fp<C>& b) {
  uint32_t x = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) x |= a.v[i] ^ b.v[i];
  return x == 0;
}

template <class C>
HD bool is_one(const Fp<C>& a) {
  uint32_t x = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) x |= a.v[i] ^ C::one(i);
  return x == 0;
}

// 32-bit add and subtract with carry: one v_add_co / v_addc_co (v_sub_co / v_subb_co) per
// limb on the device. The same chains written with 64-bit intermediates compile to 64-bit
// shifts and register moves, about three times the instructions.
HD uint32_t addc32(uint32_t x, uint32_t y, uint32_t& carry) {
  unsigned c;
  uint32_t r = __builtin_addc(x, y, carry, &c);
  carry = c;
  return r;
}
HD uint32_t subb32(uint32_t x, uint32_t y, uint32_t& borrow) {
  unsigned b;
  uint32_t r = __builtin_subc(x, y, borrow, &b);
  borrow = b;
  return r;
}

// a - p if a >= p, computed as a subtraction with borrow and a select.
template <class C>
HD void reduce_once(uint32_t a[8]) {
  uint32_t d[8];
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) d[i] = subb32(a[i], C::p(i), borrow);
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = borrow ? a[i] : d[i];
}

template <class C>
HD Fp<C> add(const Fp<C>& a, const Fp<C>& b) {
  Fp<C> r;
  uint32_t carry = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = addc32(a.v[i], b.v[i], carry);
  reduce_once<C>(r.v);  // p < 2^254: the sum never carries out of 256 bits
  return r;
}

template <class C>
HD Fp<C> sub(const Fp<C>& a, const Fp<C>& b) {
  Fp<C> r;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = subb32(a.v[i], b.v[i], borrow);
  // Add p back when the difference went negative: mask the modulus with the borrow.
  uint32_t mask = 0u - borrow;
  uint32_t carry = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = addc32(r.v[i], C::p(i) & mask, carry);
  return r;
}

template <class C>
HD Fp<C> neg(const Fp<C>& a) {
  return is_zero(a) ? a : sub(fp_zero<C>(), a);
}

template <class C>
HD Fp<C> dbl(const Fp<C>& a) {
  return add(a, a);
}

// Montgomery product a*b*2^-256 mod p, CIOS (Koc, Acar, Kaliski 1996) over 32-bit limbs.
// Each step is one 32x32+32 -> 64 multiply-add (the partial product plus the running word)
// and a 32-bit add-with-carry of the previous step's high word: the carry chain stays in
// 32-bit registers and the carry flag instead of 64-bit register pairs.
template <class C>
HD Fp<C> mul(const Fp<C>& a, const Fp<C>& b) {
  uint32_t t[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) t[i] = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    uint32_t c = 0, co;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      uint64_t x = (uint64_t)a.v[j] * b.v[i] + (uint64_t)t[j];
      t[j] = __builtin_addc((uint32_t)x, c, 0u, &co);
      c = (uint32_t)(x >> 32) + co;  // < 2^32: x's high word is at most 2^32 - 2
    }
    t[8] = __builtin_addc(t[8], c, 0u, &co);
    t[9] = co;
    uint32_t m = t[0] * C::N0;
    uint64_t x = (uint64_t)m * C::p(0) + (uint64_t)t[0];
    c = (uint32_t)(x >> 32);
#pragma unroll
    for (int j = 1; j < 8; ++j) {
      x = (uint64_t)m * C::p(j) + (uint64_t)t[j];
      t[j - 1] = __builtin_addc((uint32_t)x, c, 0u, &co);
      c = (uint32_t)(x >> 32) + co;
    }
    t[7] = __builtin_addc(t[8], c, 0u, &co);
    t[8] = t[9] + co;
  }
  Fp<C> r;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = t[i];
  reduce_once<C>(r.v);  // the CIOS result is below 2p < 2^256
  return r;
}

// Montgomery product for R' = 2^261: CIOS over nine 29-bit limbs with 64-bit column
// accumulators. Every partial product is one 29x29+64 multiply-add; the columns take the
// carries lazily (each stays below 2^63: at most 18 products of 2^58 land in one column), so
// the loop has no carry flags at all, which the 32-bit form above needs on every step. On
// gfx950 this runs 1.42x the 32-bit form's products per second. Inputs and output are the
// usual 8 x 32-bit limbs, the output fully reduced. M29 gives the modulus in 29-bit limbs and
// -modulus^-1 mod 2^29.
struct Fq29 {
  using Cfg = FqCfg;
  static HD uint32_t p29(int j) {
    constexpr uint32_t P[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                               0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
    return P[j];
  }
  static constexpr uint32_t N29 = 0x04866389u;
};

struct Fr29 {
  using Cfg = FrCfg;
  static HD uint32_t p29(int j) {
    constexpr uint32_t P[9] = {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u,
                               0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
    return P[j];
  }
  static constexpr uint32_t N29 = 0x0fffffffu;
};

// The 29-bit limbs of an 8 x 32-bit value, and back (a value below 2^256 in limbs < 2^29).
// Every index is a template constant, so no word array is ever indexed at run time (where a
// large kernel left the loop rolled, a run-time index put the words in scratch memory).
template <int K>
HD uint32_t limb29_at(const uint32_t v[8]) {  // bits [29K, 29K + 29)
  constexpr int o = 29 * K, w = o >> 5, sh = o & 31;
  constexpr uint32_t MASK = (1u << 29) - 1u;
  if constexpr (sh == 0) return v[w] & MASK;
  else if constexpr (w + 1 < 8) return ((v[w] >> sh) | (v[w + 1] << (32 - sh))) & MASK;
  else return (v[w] >> sh) & MASK;
}
template <int I>
HD uint32_t word29_at(const uint32_t R[9]) {  // bits [32I, 32I + 32): two limbs (32I mod 29 < 27)
  constexpr int j = (32 * I) / 29, s = 32 * I - 29 * j;
  static_assert(s < 27, "a word spans at most two limbs");
  return (R[j] >> s) | (R[j + 1] << (29 - s));
}
HD void limbs29_of(const uint32_t v[8], uint32_t A[9]) {
  A[0] = limb29_at<0>(v);
  A[1] = limb29_at<1>(v);
  A[2] = limb29_at<2>(v);
  A[3] = limb29_at<3>(v);
  A[4] = limb29_at<4>(v);
  A[5] = limb29_at<5>(v);
  A[6] = limb29_at<6>(v);
  A[7] = limb29_at<7>(v);
  A[8] = limb29_at<8>(v);
}
HD void pack29(const uint32_t R[9], uint32_t out[8]) {
  out[0] = word29_at<0>(R);
  out[1] = word29_at<1>(R);
  out[2] = word29_at<2>(R);
  out[3] = word29_at<3>(R);
  out[4] = word29_at<4>(R);
  out[5] = word29_at<5>(R);
  out[6] = word29_at<6>(R);
  out[7] = word29_at<7>(R);
}

// The product on limbs: A B 2^-261 mod the modulus, in [0, 2 modulus) (any inputs below
// 2^256), as nine 29-bit limbs.
template <class M29>
HD void mont29_limbs(const uint32_t A[9], const uint32_t B[9], uint32_t R[9]) {
  constexpr uint32_t MASK = (1u << 29) - 1u;
  uint64_t t[10];
#pragma unroll
  for (int j = 0; j < 10; ++j) t[j] = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
#pragma unroll
    for (int j = 0; j < 9; ++j) t[j] += (uint64_t)A[j] * B[i];
    // m = t0 * N29 mod 2^29; the low 32 bits of the product suffice, unmasked t0 included.
    const uint32_t m = ((uint32_t)t[0] * M29::N29) & MASK;
#pragma unroll
    for (int j = 0; j < 9; ++j) t[j] += (uint64_t)m * M29::p29(j);
    t[1] += t[0] >> 29;  // t[0] is now a multiple of 2^29
#pragma unroll
    for (int j = 0; j < 9; ++j) t[j] = t[j + 1];
    t[9] = 0;
  }
  // Carry the columns into 29-bit limbs (the value is below 2 modulus < 2^255).
  uint64_t c = 0;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const uint64_t x = t[j] + c;
    c = x >> 29;
    R[j] = (uint32_t)x & MASK;
  }
}

// Lazily reduced sums of products. A wide value is 17 columns t[k] of weight 2^(29 k), kept
// as 64-bit two's complement words (uint64_t arithmetic wraps, so a column may pass through
// values outside int64 as long as its final value fits); wide_mac adds A B, wide_msub
// subtracts it (one signed 32 x 32 multiply-add per partial product: limbs below 2^30). redc29 then divides the sum by 2^261 modulo the modulus once for all its
// products.
HD void wide_mac(uint64_t t[17], const uint32_t A[9], const uint32_t B[9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = 0; j < 9; ++j) t[i + j] += (uint64_t)A[i] * B[j];
}
HD void wide_msub(uint64_t t[17], const uint32_t A[9], const uint32_t B[9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int32_t na = -(int32_t)A[i];
#pragma unroll
    for (int j = 0; j < 9; ++j) t[i + j] += (uint64_t)((int64_t)na * (int64_t)(int32_t)B[j]);
  }
}

// t 2^-261 mod the modulus as nine 29-bit limbs in [0, 2 modulus), for a wide value
// V = sum_k t[k] 2^(29 k) with -2^261 modulus < V < 2^261 modulus whose columns each hold at
// most 18 partial products of limbs below 2^29 (|t[k]| <= 18 2^58; the reduction's nine
// products and the carries keep every column below 27 2^58 + 2^34 < 2^63). The quotient
// (V + M modulus) / 2^261, M < 2^261, lies in (-modulus, 2 modulus); a negative one gets the
// modulus added (NEG = false: V >= 0, nothing to add).
template <class M29, bool NEG = true>
HD void redc29(uint64_t t[17], uint32_t R[9]) {
  constexpr uint32_t MASK = (1u << 29) - 1u;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const uint32_t m = ((uint32_t)t[i] * M29::N29) & MASK;
#pragma unroll
    for (int j = 0; j < 9; ++j) t[i + j] += (uint64_t)m * M29::p29(j);
    if (i < 8) t[i + 1] += (uint64_t)((int64_t)t[i] >> 29);  // t[i] is now 0 mod 2^29
  }
  // The quotient's columns are t[9..16] plus the carry out of t[8].
  int64_t c = (int64_t)t[8] >> 29;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const int64_t x = (j < 8 ? (int64_t)t[9 + j] : 0) + c;
    R[j] = (uint32_t)x & MASK;
    c = x >> 29;
  }
  // c is 0, or -1 for a negative quotient: then add the modulus (the carry out cancels c).
  if (!NEG) return;
  const uint32_t neg = (uint32_t)c;
  uint32_t k = 0;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const uint32_t x = R[j] + (M29::p29(j) & neg) + k;
    R[j] = x & MASK;
    k = x >> 29;
  }
}

// With FULL = false the result is left in [0, 2 modulus) (any 256-bit inputs): the MSM form
// keeps its values there (see the FqmCfg operations below).
template <class M29, bool FULL = true>
HD void mont29(const uint32_t a[8], const uint32_t b[8], uint32_t out[8]) {
  uint32_t A[9], B[9], R[9];
  limbs29_of(a, A);
  limbs29_of(b, B);
  mont29_limbs<M29>(A, B, R);
  pack29(R, out);
  if (FULL) reduce_once<typename M29::Cfg>(out);
}

// The MSM form's values live in [0, 2q), not reduced below q: the product needs no final
// subtraction (its result is below 2q for any 256-bit inputs), additions subtract 2q and
// subtractions add 2q, and zero tests accept 0 and q. Values leave the MSM kernels only
// through the host, which reduces them fully.
template <>
HD Fp<FqmCfg> mul<FqmCfg>(const Fp<FqmCfg>& a, const Fp<FqmCfg>& b) {
  Fp<FqmCfg> r;
  mont29<Fq29, false>(a.v, b.v, r.v);
  return r;
}

HD uint32_t fqm_2q(int i) {
  constexpr uint32_t Q2[8] = {0xb0f9fa8eu, 0x7841182du, 0xd0e3951au, 0x2f02d522u,
                              0x0302b0bbu, 0x70a08b6du, 0xc2634053u, 0x60c89ce5u};
  return Q2[i];
}

template <>
HD Fp<FqmCfg> add<FqmCfg>(const Fp<FqmCfg>& a, const Fp<FqmCfg>& b) {
  Fp<FqmCfg> r, d;
  uint32_t carry = 0, borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = addc32(a.v[i], b.v[i], carry);  // a + b < 4q < 2^256
#pragma unroll
  for (int i = 0; i < 8; ++i) d.v[i] = subb32(r.v[i], fqm_2q(i), borrow);
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = borrow ? r.v[i] : d.v[i];
  return r;
}

template <>
HD Fp<FqmCfg> sub<FqmCfg>(const Fp<FqmCfg>& a, const Fp<FqmCfg>& b) {
  Fp<FqmCfg> r;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = subb32(a.v[i], b.v[i], borrow);
  uint32_t mask = 0u - borrow, carry = 0;  // negative: add 2q
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = addc32(r.v[i], fqm_2q(i) & mask, carry);
  return r;
}

template <>
HD bool is_zero<FqmCfg>(const Fp<FqmCfg>& a) {
  uint32_t z = 0, q = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    z |= a.v[i];
    q |= a.v[i] ^ FqCfg::p(i);
  }
  return z == 0 || q == 0;
}

template <>
HD Fp<FqmCfg> neg<FqmCfg>(const Fp<FqmCfg>& a) {
  return sub(fp_zero<FqmCfg>(), a);
}

// x * w for x in the usual R = 2^256 form and w stored in the R' = 2^261 form (w 2^261 mod r):
// the 29-bit product divides by 2^261, so the result is x w in the R form. The NTT keeps its
// twiddle and coset tables in the R' form and multiplies data by them this way.
HD Fr mul_rp(const Fr& x, const Fr& w_rp) {
  Fr r;
  mont29<Fr29>(x.v, w_rp.v, r.v);
  return r;
}

template <class C>
HD Fp<C> sqr(const Fp<C>& a) {
  return mul(a, a);
}

template <class C>
HD Fp<C> from_std(const Fp<C>& plain) {
  Fp<C> r2;
#pragma unroll
  for (int i = 0; i < 8; ++i) r2.v[i] = C::r2(i);
  return mul(plain, r2);
}

template <class C>
HD Fp<C> to_std(const Fp<C>& a) {
  Fp<C> one_plain = fp_zero<C>();
  one_plain.v[0] = 1;
  return mul(a, one_plain);
}

// a^(p-2) (Fermat), square-and-multiply over the exponent's bits; zero maps to zero.
template <class C>
HD Fp<C> inv(const Fp<C>& a) {
  Fp<C> acc = fp_one<C>();
  for (int i = 255; i >= 0; --i) {
    acc = sqr(acc);
    uint32_t e = C::p(i >> 5) - ((i >> 5) == 0 ? 2u : 0u);
    if ((e >> (i & 31)) & 1) acc = mul(acc, a);
  }
  return acc;
}

// ---------------------------------------------------------------------------------------
// Fq2 = Fq[u]/(u^2 + 1)
// ---------------------------------------------------------------------------------------
// Fq2T<true> expands its three Fq products in place; Fq2T<false> (= Fq2) calls one
// out-of-line copy of the Fq multiply. A G2 point formula with every product expanded is
// ~17k instructions, past the 64 KB instruction cache a CU pair shares, and costs the
// compiler minutes per kernel, so only the G2 bucket-accumulation loop uses the expanded form.
// B: the base field type, Fq (R = 2^256) or Fm (the MSM's R' = 2^261 form).
template <bool INL, class B = Fq>
struct Fq2T {
  B c0, c1;
};
typedef Fq2T<false> Fq2;
typedef Fq2T<true> Fq2i;
typedef Fq2T<true, Fm> Fm2i;

template <class B>
__host__ __device__ __attribute__((noinline)) B fq_mul_ool(B a, B b) {
  return mul(a, b);
}

template <bool INL, class B>
HD B fq2_part_mul(const B& a, const B& b) {
  if constexpr (INL) return mul(a, b);
  else return fq_mul_ool(a, b);
}

template <bool INL = false, class B = Fq> HD Fq2T<INL, B> fq2_zero() {
  return {B{}, B{}};
}
template <bool INL = false, class C = FqCfg> HD Fq2T<INL, Fp<C>> fq2_one() {
  return {fp_one<C>(), fp_zero<C>()};
}
template <bool INL, class B> HD bool is_zero(const Fq2T<INL, B>& a) {
  return is_zero(a.c0) && is_zero(a.c1);
}
template <bool INL, class B> HD bool eq(const Fq2T<INL, B>& a, const Fq2T<INL, B>& b) {
  return eq(a.c0, b.c0) && eq(a.c1, b.c1);
}
template <bool INL, class B> HD Fq2T<INL, B> add(const Fq2T<INL, B>& a, const Fq2T<INL, B>& b) {
  return {add(a.c0, b.c0), add(a.c1, b.c1)};
}
template <bool INL, class B> HD Fq2T<INL, B> sub(const Fq2T<INL, B>& a, const Fq2T<INL, B>& b) {
  return {sub(a.c0, b.c0), sub(a.c1, b.c1)};
}
template <bool INL, class B> HD Fq2T<INL, B> neg(const Fq2T<INL, B>& a) {
  return {neg(a.c0), neg(a.c1)};
}
template <bool INL, class B> HD Fq2T<INL, B> dbl(const Fq2T<INL, B>& a) { return add(a, a); }

// Karatsuba: 3 base-field products.
// A sum that only feeds a product: the MSM form's product takes any 256-bit operand, so there
// a + b (< 4q < 2^256) needs no reduction; the other fields reduce as usual.
template <class B> HD B add_into_mul(const B& a, const B& b) { return add(a, b); }
template <> HD Fp<FqmCfg> add_into_mul<Fp<FqmCfg>>(const Fp<FqmCfg>& a, const Fp<FqmCfg>& b) {
  Fp<FqmCfg> r;
  uint32_t carry = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = addc32(a.v[i], b.v[i], carry);
  return r;
}

template <bool INL, class B> HD Fq2T<INL, B> mul(const Fq2T<INL, B>& a, const Fq2T<INL, B>& b) {
  B v0 = fq2_part_mul<INL>(a.c0, b.c0);
  B v1 = fq2_part_mul<INL>(a.c1, b.c1);
  B s = fq2_part_mul<INL>(add_into_mul(a.c0, a.c1), add_into_mul(b.c0, b.c1));
  return {sub(v0, v1), sub(sub(s, v0), v1)};
}

// In the MSM form (the G2 bucket accumulation) the Karatsuba products stay wide and each
// coordinate is reduced once: c0 = a0 b0 - a1 b1 (|c0| < 4 q^2), c1 = (a0 + a1)(b0 + b1) -
// a0 b0 - a1 b1 = a0 b1 + a1 b0 (each column a sum of at most 18 limb products): two
// reductions instead of three, and each operand goes to limbs once.
HD Fq2T<true, Fm> mul(const Fq2T<true, Fm>& a, const Fq2T<true, Fm>& b) {
  uint32_t A0[9], A1[9], B0[9], B1[9], R[9];
  limbs29_of(a.c0.v, A0);
  limbs29_of(a.c1.v, A1);
  limbs29_of(b.c0.v, B0);
  limbs29_of(b.c1.v, B1);
  uint64_t t0[17], t1[17];
#pragma unroll
  for (int k = 0; k < 17; ++k) t0[k] = t1[k] = 0;
  wide_mac(t0, A0, B0);
  wide_mac(t1, A1, B1);
#pragma unroll
  for (int k = 0; k < 17; ++k) {
    const uint64_t x = t0[k], y = t1[k];
    t0[k] = x - y;
    t1[k] = 0 - x - y;
  }
  Fq2T<true, Fm> r;
  redc29<Fq29>(t0, R);
  pack29(R, r.c0.v);
#pragma unroll
  for (int j = 0; j < 9; ++j) {  // limbs below 2^30: the columns' final values stay in range
    A0[j] += A1[j];
    B0[j] += B1[j];
  }
  wide_mac(t1, A0, B0);
  redc29<Fq29, false>(t1, R);
  pack29(R, r.c1.v);
  return r;
}

// a b - c d, by default two products and a subtraction.
template <class F> HD F mul_sub(const F& a, const F& b, const F& c, const F& d) {
  return sub(mul(a, b), mul(c, d));
}
// In the MSM form over Fq2: both Karatsuba products wide in the same two accumulators, one
// reduction per coordinate (two instead of four). c0 = a0 b0 - a1 b1 - c0 d0 + c1 d1 and
// c1 = a0 b1 + a1 b0 - c0 d1 - c1 d0: each column at most 18 limb products in magnitude, and
// |c0|, |c1| < 8 q^2.
HD Fq2T<true, Fm> mul_sub(const Fq2T<true, Fm>& a, const Fq2T<true, Fm>& b,
                          const Fq2T<true, Fm>& c, const Fq2T<true, Fm>& d) {
  uint32_t X0[9], X1[9], Y0[9], Y1[9], R[9];
  uint64_t t0[17], t1[17], w[17];
  limbs29_of(a.c0.v, X0);
  limbs29_of(a.c1.v, X1);
  limbs29_of(b.c0.v, Y0);
  limbs29_of(b.c1.v, Y1);
#pragma unroll
  for (int k = 0; k < 17; ++k) t0[k] = w[k] = 0;
  wide_mac(t0, X0, Y0);
  wide_mac(w, X1, Y1);
#pragma unroll
  for (int k = 0; k < 17; ++k) {
    const uint64_t x = t0[k];
    t0[k] = x - w[k];
    t1[k] = 0 - x - w[k];
  }
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    X0[j] += X1[j];
    Y0[j] += Y1[j];
  }
  wide_mac(t1, X0, Y0);  // t1 = a0 b1 + a1 b0
  limbs29_of(c.c0.v, X0);
  limbs29_of(c.c1.v, X1);
  limbs29_of(d.c0.v, Y0);
  limbs29_of(d.c1.v, Y1);
#pragma unroll
  for (int k = 0; k < 17; ++k) w[k] = 0;
  wide_mac(w, X0, Y0);
#pragma unroll
  for (int k = 0; k < 17; ++k) {
    t0[k] -= w[k];
    t1[k] += w[k];
    w[k] = 0;
  }
  wide_mac(w, X1, Y1);
#pragma unroll
  for (int k = 0; k < 17; ++k) {
    t0[k] += w[k];
    t1[k] += w[k];
  }
#pragma unroll
  for (int j = 0; j < 9; ++j) {  // below 2^30: the signed products stay inside int64
    X0[j] += X1[j];
    Y0[j] += Y1[j];
  }
  wide_msub(t1, X0, Y0);  // t1 -= (c0 + c1)(d0 + d1): now a0 b1 + a1 b0 - c0 d1 - c1 d0
  Fq2T<true, Fm> r;
  redc29<Fq29>(t0, R);
  pack29(R, r.c0.v);
  redc29<Fq29>(t1, R);
  pack29(R, r.c1.v);
  return r;
}

// (a0 + a1 u)^2 = (a0 + a1)(a0 - a1) + 2 a0 a1 u: 2 base-field products.
template <bool INL, class B> HD Fq2T<INL, B> sqr(const Fq2T<INL, B>& a) {
  B m = fq2_part_mul<INL>(add_into_mul(a.c0, a.c1), sub(a.c0, a.c1));
  B t = fq2_part_mul<INL>(a.c0, a.c1);
  return {m, add(t, t)};
}

template <bool INL, class B> HD Fq2T<INL, B> inv(const Fq2T<INL, B>& a) {
  B t = inv(add(sqr(a.c0), sqr(a.c1)));
  return {mul(a.c0, t), neg(mul(a.c1, t))};
}

// ---------------------------------------------------------------------------------------
// Points
// ---------------------------------------------------------------------------------------
template <class F>
struct Aff {
  F x, y;
};

template <class F>
struct Xyzz {
  F X, Y, ZZ, ZZZ;
};

typedef Aff<Fq> G1Aff;
typedef Aff<Fq2> G2Aff;
typedef Xyzz<Fq> G1X;
typedef Xyzz<Fq2> G2X;

template <class F> struct FieldOps;
template <class C> struct FieldOps<Fp<C>> {
  static HD Fp<C> zero() { return fp_zero<C>(); }
  static HD Fp<C> one() { return fp_one<C>(); }
};
template <bool INL, class C> struct FieldOps<Fq2T<INL, Fp<C>>> {
  static HD Fq2T<INL, Fp<C>> zero() { return {fp_zero<C>(), fp_zero<C>()}; }
  static HD Fq2T<INL, Fp<C>> one() { return {fp_one<C>(), fp_zero<C>()}; }
};
template <class F> HD F f_zero() { return FieldOps<F>::zero(); }
template <class F> HD F f_one() { return FieldOps<F>::one(); }

template <class F>
HD bool aff_is_inf(const Aff<F>& p) {
  return is_zero(p.x) && is_zero(p.y);
}

template <class F>
HD Xyzz<F> xyzz_inf() {
  return {f_zero<F>(), f_zero<F>(), f_zero<F>(), f_zero<F>()};
}

template <class F>
HD bool xyzz_is_inf(const Xyzz<F>& p) {
  return is_zero(p.ZZ);
}

template <class F>
HD Xyzz<F> xyzz_from_aff(const Aff<F>& a) {
  if (aff_is_inf(a)) return xyzz_inf<F>();
  return {a.x, a.y, f_one<F>(), f_one<F>()};
}

// The point formulas below are ordered to keep few values live (each product consumed as
// early as the formula allows) and, on the device, fenced with scheduling barriers (SB) so
// the compiler does not run several products side by side: a G2 addition otherwise holds more
// than 256 registers per lane and takes a whole SIMD.
#if defined(__HIP_DEVICE_COMPILE__)
#define SB() __builtin_amdgcn_sched_barrier(0)
#else
#define SB() ((void)0)
#endif

// dbl-2008-s-1 (a = 0).
template <class F>
HD Xyzz<F> xyzz_dbl(const Xyzz<F>& p) {
  if (xyzz_is_inf(p)) return p;
  F U = dbl(p.Y);
  F V = sqr(U);
  SB();
  F W = mul(U, V);
  SB();
  F S = mul(p.X, V);
  SB();
  Xyzz<F> r;
  r.ZZ = mul(V, p.ZZ);
  SB();
  F X2 = sqr(p.X);
  SB();
  F M = add(dbl(X2), X2);
  F WY = mul(W, p.Y);
  SB();
  r.ZZZ = mul(W, p.ZZZ);
  SB();
  r.X = sub(sqr(M), dbl(S));
  SB();
  r.Y = sub(mul(M, sub(S, r.X)), WY);
  return r;
}

// mdbl-2008-s-1: doubling of an affine point.
template <class F>
HD Xyzz<F> xyzz_dbl_aff(const Aff<F>& a) {
  F U = dbl(a.y);
  F V = sqr(U);
  SB();
  F W = mul(U, V);
  SB();
  F S = mul(a.x, V);
  SB();
  F X2 = sqr(a.x);
  SB();
  F M = add(dbl(X2), X2);
  F WY = mul(W, a.y);
  SB();
  Xyzz<F> r;
  r.X = sub(sqr(M), dbl(S));
  SB();
  r.Y = sub(mul(M, sub(S, r.X)), WY);
  r.ZZ = V;
  r.ZZZ = W;
  return r;
}

// madd-2008-s: p + affine a, with the doubling and cancellation cases resolved.
template <class F>
HD Xyzz<F> xyzz_add_aff(const Xyzz<F>& p, const Aff<F>& a) {
  if (aff_is_inf(a)) return p;
  if (xyzz_is_inf(p)) return {a.x, a.y, f_one<F>(), f_one<F>()};
  F P = sub(mul(a.x, p.ZZ), p.X);
  SB();
  F R = sub(mul(a.y, p.ZZZ), p.Y);
  SB();
  if (is_zero(P)) {
    if (is_zero(R)) return xyzz_dbl_aff(a);
    return xyzz_inf<F>();
  }
  F PP = sqr(P);
  SB();
  F PPP = mul(P, PP);
  SB();
  Xyzz<F> r;
  r.ZZ = mul(p.ZZ, PP);
  SB();
  F Q = mul(p.X, PP);
  SB();
  r.ZZZ = mul(p.ZZZ, PPP);
  SB();
  r.X = sub(sub(sqr(R), PPP), dbl(Q));
  SB();
  r.Y = mul_sub(R, sub(Q, r.X), p.Y, PPP);
  return r;
}

// add-2008-s: p + q, general. The inputs die after the first six products: the doubling case
// (p = q) doubles (U1, S1, ZZ1 ZZ2, ZZZ1 ZZZ2), which is p in other coordinates.
template <class F>
HD Xyzz<F> xyzz_add(const Xyzz<F>& p, const Xyzz<F>& q) {
  if (xyzz_is_inf(p)) return q;
  if (xyzz_is_inf(q)) return p;
  F U1 = mul(p.X, q.ZZ);
  SB();
  F P = sub(mul(q.X, p.ZZ), U1);
  SB();
  F S1 = mul(p.Y, q.ZZZ);
  SB();
  F R = sub(mul(q.Y, p.ZZZ), S1);
  SB();
  F ZZ12 = mul(p.ZZ, q.ZZ);
  SB();
  F ZZZ12 = mul(p.ZZZ, q.ZZZ);
  SB();
  if (is_zero(P)) {
    if (is_zero(R)) return xyzz_dbl(Xyzz<F>{U1, S1, ZZ12, ZZZ12});
    return xyzz_inf<F>();
  }
  F PP = sqr(P);
  SB();
  F PPP = mul(P, PP);
  SB();
  Xyzz<F> r;
  r.ZZ = mul(ZZ12, PP);
  SB();
  F Q = mul(U1, PP);
  SB();
  r.ZZZ = mul(ZZZ12, PPP);
  SB();
  F SP = mul(S1, PPP);
  SB();
  r.X = sub(sub(sqr(R), PPP), dbl(Q));
  SB();
  r.Y = sub(mul(R, sub(Q, r.X)), SP);
  return r;
}

template <class F>
HD Xyzz<F> xyzz_neg(const Xyzz<F>& p) {
  Xyzz<F> r = p;
  r.Y = neg(p.Y);
  return r;
}

template <class F>
HD Aff<F> aff_neg(const Aff<F>& a) {
  if (aff_is_inf(a)) return a;
  return {a.x, neg(a.y)};
}

template <class F>
HD Aff<F> xyzz_to_aff(const Xyzz<F>& p) {
  if (xyzz_is_inf(p)) return {f_zero<F>(), f_zero<F>()};
  // x = X / ZZ, y = Y / ZZZ: one inversion of ZZZ gives both (ZZ^-1 = ZZZ^-1 * ZZZ / ZZ, and
  // ZZZ / ZZ = ZZZ^2 / (ZZ * ZZZ) ... kept simple: two inversions through one).
  F izzz = inv(p.ZZZ);
  F t = mul(izzz, p.ZZ);       // ZZ / ZZZ = 1 / Z
  F iz2 = sqr(t);              // 1 / ZZ
  return {mul(p.X, iz2), mul(p.Y, izzz)};
}

// ---------------------------------------------------------------------------------------
// G1 bucket sums in limbs
// ---------------------------------------------------------------------------------------
// The G1 bucket loop keeps its running XYZZ sum as nine 29-bit limbs per coordinate and does
// the mixed addition without leaving them: no word/limb conversion per product, no carry
// chains. A difference a - b is taken limb by limb as a + k q - b, with k q written in raised
// limbs (every limb below the top one increased by b' 2^29 - b', the top one lowered by b',
// the value unchanged) so that no limb goes negative, then carried once, all limbs in
// parallel (norm29). Values are bounded rather than reduced:
//   products (mont29_limbs, redc29): exact limbs, below 2q;
//   X < 9q, limbs below 2^29 + 8 (X3 = R^2 + 7q - PPP - 2Q); Y, ZZ, ZZZ < 2q exact.
// Each product stays inside mont29_limbs' bound (operands' product below 2^261 q): the largest
// is P^2 with P = U2 + 10q - X1 < 12q, 144 q^2 < 2^261 q. Every limb sum stays in [0, 2^32).
// Leaving the loop, X is reduced below 2q by a product with one (g1l_out).
struct G1L {
  uint32_t X[9], Y[9], ZZ[9], ZZZ[9];
};

// k q in raised limbs (the comments give k and the raise b').
struct Kq3 {  // 3q, b' = 2: limbs >= 2^30 - 2
  static HD uint32_t k(int j) {
    constexpr uint32_t K[9] = {0x4976f7d5u, 0x430d2222u, 0x5557e9ebu, 0x487f6870u, 0x40918c66u,
                               0x4891c240u, 0x41f4a3c1u, 0x4b14729au, 0x00912ce9u};
    return K[j];
  }
};
struct Kq7 {  // 7q, b' = 4: limbs >= 2^31 - 4, above a sum of three exact limbs
  static HD uint32_t k(int j) {
    constexpr uint32_t K[9] = {0x8b6aecf1u, 0x871ea4fbu, 0x87227725u, 0x93d3f3b2u, 0x96a8f244u,
                               0x93fec540u, 0x849028c3u, 0x84850b68u, 0x0152be21u};
    return K[j];
  }
};
struct Kq10 {  // 10q, b' = 2
  static HD uint32_t k(int j) {
    constexpr uint32_t K[9] = {0x54e1e4c6u, 0x4a2bc721u, 0x5c7a6114u, 0x5c535c26u, 0x573a7eaeu,
                               0x5c908784u, 0x4684cc88u, 0x4f997e06u, 0x01e3eb0eu};
    return K[j];
  }
};
// 2^261 mod q (the MSM form's one) in limbs.
HD uint32_t one29(int j) {
  constexpr uint32_t O[9] = {0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u,
                             0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};
  return O[j];
}

// Carries every limb's bits above 29 into the next limb, all at once: the value is unchanged
// (below 2^261, so the top limb has nothing to carry out) and limbs below 2^32 end below
// 2^29 + 8.
HD void norm29(uint32_t x[9]) {
  constexpr uint32_t MASK = (1u << 29) - 1u;
  uint32_t c[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) c[j] = x[j] >> 29;
  x[0] &= MASK;
#pragma unroll
  for (int j = 1; j < 8; ++j) x[j] = (x[j] & MASK) + c[j - 1];
  x[8] += c[7];
}

// r = a + K - b limb by limb, then carried.
template <class K>
HD void sub29(const uint32_t a[9], const uint32_t b[9], uint32_t r[9]) {
#pragma unroll
  for (int j = 0; j < 9; ++j) r[j] = a[j] + K::k(j) - b[j];
  norm29(r);
}

// Exact limbs of a value below 2q that is 0 mod q: 0 or q.
HD bool zero29(const uint32_t a[9]) {
  uint32_t z = 0, q = 0;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    z |= a[j];
    q |= a[j] ^ Fq29::p29(j);
  }
  return z == 0 || q == 0;
}

HD G1L g1l_from(const Xyzz<Fm>& p) {
  G1L r;
  limbs29_of(p.X.v, r.X);
  limbs29_of(p.Y.v, r.Y);
  limbs29_of(p.ZZ.v, r.ZZ);
  limbs29_of(p.ZZZ.v, r.ZZZ);
  return r;
}

HD G1L g1l_inf() {
  G1L r;
#pragma unroll
  for (int j = 0; j < 9; ++j) r.X[j] = r.Y[j] = r.ZZ[j] = r.ZZZ[j] = 0;
  return r;
}

HD Xyzz<Fm> g1l_out(const G1L& p) {
  uint32_t one[9], X[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) one[j] = one29(j);
  mont29_limbs<Fq29>(p.X, one, X);  // X mod q, below 2q
  Xyzz<Fm> r;
  pack29(X, r.X.v);
  pack29(p.Y, r.Y.v);
  pack29(p.ZZ, r.ZZ.v);
  pack29(p.ZZZ, r.ZZZ.v);
  return r;
}

// madd-2008-s, p + (x, negate ? -y : y), with the doubling and cancellation cases resolved
// (the same result as xyzz_add_aff up to the representation). P = 0 mod q is read off
// PP = P^2 (below 2q, so 0 or q), R = 0 off R^2.
HD G1L g1l_madd(const G1L& p, const Aff<Fm>& a, bool negate) {
  if (aff_is_inf(a)) return p;
  if (zero29(p.ZZ)) {
    G1L r;
    limbs29_of(a.x.v, r.X);
    limbs29_of((negate ? neg(a.y) : a.y).v, r.Y);
#pragma unroll
    for (int j = 0; j < 9; ++j) r.ZZ[j] = r.ZZZ[j] = one29(j);
    return r;
  }
  uint32_t x[9], y[9], P[9], R[9];
  limbs29_of(a.x.v, x);
  limbs29_of(a.y.v, y);
#pragma unroll
  for (int j = 0; j < 9; ++j) y[j] = negate ? Kq3::k(j) - y[j] : y[j];  // 3q - y, below 3q
  mont29_limbs<Fq29>(x, p.ZZ, P);  // U2
  sub29<Kq10>(P, p.X, P);          // P = U2 - X1 + 10q < 12q
  SB();
  mont29_limbs<Fq29>(y, p.ZZZ, R);  // S2
  sub29<Kq3>(R, p.Y, R);            // R = S2 - Y1 + 3q < 5q
  SB();
  uint32_t PP[9];
  mont29_limbs<Fq29>(P, P, PP);
  if (zero29(PP)) {
    uint32_t R2[9];
    mont29_limbs<Fq29>(R, R, R2);
    if (zero29(R2)) return g1l_from(xyzz_dbl_aff(negate ? aff_neg(a) : a));
    return g1l_inf();
  }
  SB();
  uint32_t PPP[9];
  mont29_limbs<Fq29>(P, PP, PPP);
  SB();
  G1L r;
  mont29_limbs<Fq29>(p.ZZ, PP, r.ZZ);
  SB();
  uint32_t Q[9];
  mont29_limbs<Fq29>(p.X, PP, Q);
  SB();
  mont29_limbs<Fq29>(p.ZZZ, PPP, r.ZZZ);
  SB();
  mont29_limbs<Fq29>(R, R, r.X);
#pragma unroll
  for (int j = 0; j < 9; ++j) r.X[j] = r.X[j] + Kq7::k(j) - PPP[j] - Q[j] - Q[j];
  norm29(r.X);  // X3 = R^2 - PPP - 2Q + 7q < 9q
  SB();
  sub29<Kq10>(Q, r.X, Q);  // Q - X3 + 10q < 12q
  // Y3 = R (Q - X3) - Y1 PPP as one wide value (|.| < 60 q^2) and one reduction.
  uint64_t t[17];
#pragma unroll
  for (int k = 0; k < 17; ++k) t[k] = 0;
  wide_mac(t, R, Q);
  wide_msub(t, p.Y, PPP);
  redc29<Fq29>(t, r.Y);
  return r;
}

// add-2008-s on the limb form, p + q (both in the bounds above); the doubling case goes
// through the word form.
HD G1L g1l_add(const G1L& p, const G1L& q) {
  if (zero29(p.ZZ)) return q;
  if (zero29(q.ZZ)) return p;
  uint32_t U1[9], P[9], S1[9], R[9];
  mont29_limbs<Fq29>(p.X, q.ZZ, U1);
  mont29_limbs<Fq29>(q.X, p.ZZ, P);  // U2
  sub29<Kq3>(P, U1, P);              // P = U2 - U1 + 3q < 5q
  SB();
  mont29_limbs<Fq29>(p.Y, q.ZZZ, S1);
  mont29_limbs<Fq29>(q.Y, p.ZZZ, R);  // S2
  sub29<Kq3>(R, S1, R);               // R = S2 - S1 + 3q < 5q
  SB();
  uint32_t PP[9];
  mont29_limbs<Fq29>(P, P, PP);
  if (zero29(PP)) {
    uint32_t R2[9];
    mont29_limbs<Fq29>(R, R, R2);
    if (zero29(R2)) return g1l_from(xyzz_dbl(g1l_out(p)));
    return g1l_inf();
  }
  SB();
  uint32_t PPP[9], Z[9];
  mont29_limbs<Fq29>(P, PP, PPP);
  SB();
  G1L r;
  mont29_limbs<Fq29>(p.ZZ, q.ZZ, Z);
  mont29_limbs<Fq29>(Z, PP, r.ZZ);
  SB();
  mont29_limbs<Fq29>(p.ZZZ, q.ZZZ, Z);
  mont29_limbs<Fq29>(Z, PPP, r.ZZZ);
  SB();
  uint32_t Q[9];
  mont29_limbs<Fq29>(U1, PP, Q);
  SB();
  mont29_limbs<Fq29>(R, R, r.X);
#pragma unroll
  for (int j = 0; j < 9; ++j) r.X[j] = r.X[j] + Kq7::k(j) - PPP[j] - Q[j] - Q[j];
  norm29(r.X);  // X3 = R^2 - PPP - 2Q + 7q < 9q
  SB();
  sub29<Kq10>(Q, r.X, Q);  // Q - X3 + 10q < 12q
  uint64_t t[17];
#pragma unroll
  for (int k = 0; k < 17; ++k) t[k] = 0;
  wide_mac(t, R, Q);
  wide_msub(t, S1, PPP);
  redc29<Fq29>(t, r.Y);  // Y3 = R (Q - X3) - S1 PPP, |.| < 60 q^2
  return r;
}

// ---------------------------------------------------------------------------------------
// G2 bucket sums in limbs
// ---------------------------------------------------------------------------------------
// The same for G2 over Fq2 = Fq[u]/(u^2 + 1), each coordinate two limb vectors. An Fq2 product
// is the schoolbook one, c0 = a0 b0 - a1 b1 and c1 = a0 b1 + a1 b0, accumulated wide with
// signed multiply-adds and reduced once per component: the fourth product costs less than the
// 64-bit column subtractions (and their carry hazards) Karatsuba would need. A reduction takes
// a wide value below 2^261 q in magnitude: with components below A q and B q, |c0| < A B q^2
// and c1 < 2 A B q^2, so 2 A B must stay below 2^261 / q = 169.3. Bounds:
//   X < 5q, exact limbs (X3 is formed below 9q and 4q subtracted when it is at least 4q);
//   Y, ZZ, ZZZ < 2q exact; P = U2 + 6q - X1 < 8q; R = +-S2 + 5q - Y1 < 7q;
//   T = Q + 6q - X3 < 8q; the largest product is P^2 (c1 = 2 p0 p1 < 128 q^2), and
//   Y3 = R T - Y1 PPP has |c1| < 112 q^2.
// Every column holds at most 18 limb products on each side of its sign, as redc29 requires.
struct F2L {
  uint32_t c0[9], c1[9];
};
struct G2L {
  F2L X, Y, ZZ, ZZZ;
};

struct Kq5 {  // 5q, b' = 4: limbs >= 2^31 - 4, above a sum of two exact limbs
  static HD uint32_t k(int j) {
    constexpr uint32_t K[9] = {0x9a70f263u, 0x8515e38du, 0x8e3d3087u, 0x8e29ae10u, 0x8b9d3f54u,
                               0x8e4843bfu, 0x83426641u, 0x87ccbf00u, 0x00f1f584u};
    return K[j];
  }
};
struct Kq6 {  // 6q, b' = 2
  static HD uint32_t k(int j) {
    constexpr uint32_t K[9] = {0x52edefaau, 0x461a4446u, 0x4aafd3d8u, 0x50fed0e3u, 0x412318ceu,
                               0x51238482u, 0x43e94784u, 0x5628e536u, 0x012259d4u};
    return K[j];
  }
};
HD uint32_t q4_29(int j) {  // 4q, exact limbs
  constexpr uint32_t K[9] = {0x01f3f51cu, 0x041182dbu, 0x11ca8d3cu, 0x0b548b43u, 0x161765e0u,
                             0x0b6d0302u, 0x029b8504u, 0x197098d0u, 0x00c19139u};
  return K[j];
}

// A value below 9q in limbs below 2^32, to exact limbs and below 5q.
HD void below5q(uint32_t x[9]) {
  constexpr uint32_t MASK = (1u << 29) - 1u;
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t v = x[j] + c;
    x[j] = v & MASK;
    c = v >> 29;
  }
  x[8] += c;
  uint32_t d[9];
  int32_t b = 0;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const int32_t v = (int32_t)x[j] - (int32_t)q4_29(j) + b;
    d[j] = (uint32_t)v & MASK;
    b = v >> 29;  // 0 or -1
  }
#pragma unroll
  for (int j = 0; j < 9; ++j) x[j] = b < 0 ? x[j] : d[j];
}

HD void wide_zero(uint64_t t[17]) {
#pragma unroll
  for (int k = 0; k < 17; ++k) t[k] = 0;
}

// r = a b (r must not alias a or b). Both components' sums are formed before either is
// reduced, so the two reductions' serial chains can interleave.
HD void f2l_mul(const F2L& a, const F2L& b, F2L& r) {
  uint64_t t[17], u[17];
  wide_zero(t);
  wide_zero(u);
  wide_mac(t, a.c0, b.c0);
  wide_mac(u, a.c0, b.c1);
  wide_msub(t, a.c1, b.c1);
  wide_mac(u, a.c1, b.c0);
  redc29<Fq29>(t, r.c0);
  redc29<Fq29, false>(u, r.c1);
}

// r = a^2 = (a0 + a1)(a0 - a1) + 2 a0 a1 u (r must not alias a).
HD void f2l_sqr(const F2L& a, F2L& r) {
  uint32_t s[9];
  int32_t d[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    s[j] = a.c0[j] + a.c1[j];
    d[j] = (int32_t)a.c0[j] - (int32_t)a.c1[j];
  }
  uint64_t t[17];
  wide_zero(t);
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = 0; j < 9; ++j) t[i + j] += (uint64_t)((int64_t)(int32_t)s[i] * (int64_t)d[j]);
  redc29<Fq29>(t, r.c0);
#pragma unroll
  for (int j = 0; j < 9; ++j) s[j] = a.c0[j] + a.c0[j];
  wide_zero(t);
  wide_mac(t, s, a.c1);
  redc29<Fq29, false>(t, r.c1);
}

HD bool f2l_zero(const F2L& a) { return zero29(a.c0) && zero29(a.c1); }

HD void f2l_of(const Fm2i& a, F2L& r) {
  limbs29_of(a.c0.v, r.c0);
  limbs29_of(a.c1.v, r.c1);
}

HD G2L g2l_from(const Xyzz<Fm2i>& p) {
  G2L r;
  f2l_of(p.X, r.X);
  f2l_of(p.Y, r.Y);
  f2l_of(p.ZZ, r.ZZ);
  f2l_of(p.ZZZ, r.ZZZ);
  return r;
}

// (x, negate ? -y : y) with ZZ = ZZZ = 1.
HD G2L g2l_from_aff(const Aff<Fm2i>& a, bool negate) {
  G2L r;
  f2l_of(a.x, r.X);
  f2l_of(negate ? neg(a.y) : a.y, r.Y);
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    r.ZZ.c0[j] = r.ZZZ.c0[j] = one29(j);
    r.ZZ.c1[j] = r.ZZZ.c1[j] = 0;
  }
  return r;
}

HD G2L g2l_inf() {
  G2L r;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    r.X.c0[j] = r.X.c1[j] = r.Y.c0[j] = r.Y.c1[j] = 0;
    r.ZZ.c0[j] = r.ZZ.c1[j] = r.ZZZ.c0[j] = r.ZZZ.c1[j] = 0;
  }
  return r;
}

HD Xyzz<Fm2i> g2l_out(const G2L& p) {
  uint32_t one[9], X[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) one[j] = one29(j);
  Xyzz<Fm2i> r;
  mont29_limbs<Fq29>(p.X.c0, one, X);  // below 2q
  pack29(X, r.X.c0.v);
  mont29_limbs<Fq29>(p.X.c1, one, X);
  pack29(X, r.X.c1.v);
  pack29(p.Y.c0, r.Y.c0.v);
  pack29(p.Y.c1, r.Y.c1.v);
  pack29(p.ZZ.c0, r.ZZ.c0.v);
  pack29(p.ZZ.c1, r.ZZ.c1.v);
  pack29(p.ZZZ.c0, r.ZZZ.c0.v);
  pack29(p.ZZZ.c1, r.ZZZ.c1.v);
  return r;
}

// madd-2008-s over Fq2 for p not infinity: r = p + (x, negate ? -y : y) and 0, or 1 when the
// sum is a doubling (p = that point, r unset) or 2 when it is infinity. `a` is dead after the
// first two products, so a caller that can reload it keeps it out of the registers of the rest.
HD int g2l_madd_core(const G2L& p, const Aff<Fm2i>& a, bool negate, G2L& r) {
  F2L v, P, R;
  f2l_of(a.x, v);
  f2l_mul(v, p.ZZ, P);  // U2
  sub29<Kq6>(P.c0, p.X.c0, P.c0);
  sub29<Kq6>(P.c1, p.X.c1, P.c1);  // P = U2 - X1 + 6q < 8q
  SB();
  f2l_of(a.y, v);
  f2l_mul(v, p.ZZZ, R);  // S2
  const uint32_t m = 0u - (uint32_t)negate;
#pragma unroll
  for (int j = 0; j < 9; ++j) {  // R = +-S2 - Y1 + 5q < 7q
    R.c0[j] = ((R.c0[j] ^ m) - m) + Kq5::k(j) - p.Y.c0[j];
    R.c1[j] = ((R.c1[j] ^ m) - m) + Kq5::k(j) - p.Y.c1[j];
  }
  norm29(R.c0);
  norm29(R.c1);
  SB();
  F2L PP;
  f2l_sqr(P, PP);
  if (f2l_zero(PP)) {
    F2L R2;
    f2l_sqr(R, R2);
    return f2l_zero(R2) ? 1 : 2;
  }
  SB();
  F2L PPP;
  f2l_mul(P, PP, PPP);
  SB();
  f2l_mul(p.ZZ, PP, r.ZZ);
  SB();
  F2L Q;
  f2l_mul(p.X, PP, Q);
  SB();
  f2l_mul(p.ZZZ, PPP, r.ZZZ);
  SB();
  f2l_sqr(R, r.X);
#pragma unroll
  for (int j = 0; j < 9; ++j) {  // X3 = R^2 - PPP - 2Q + 7q < 9q
    r.X.c0[j] = r.X.c0[j] + Kq7::k(j) - PPP.c0[j] - Q.c0[j] - Q.c0[j];
    r.X.c1[j] = r.X.c1[j] + Kq7::k(j) - PPP.c1[j] - Q.c1[j] - Q.c1[j];
  }
  below5q(r.X.c0);
  below5q(r.X.c1);
  SB();
  sub29<Kq6>(Q.c0, r.X.c0, Q.c0);
  sub29<Kq6>(Q.c1, r.X.c1, Q.c1);  // T = Q - X3 + 6q < 8q
  // Y3 = R T - Y1 PPP, one wide value and one reduction per component.
  uint64_t t[17];
  wide_zero(t);
  wide_mac(t, R.c0, Q.c0);
  wide_msub(t, R.c1, Q.c1);
  wide_msub(t, p.Y.c0, PPP.c0);
  wide_mac(t, p.Y.c1, PPP.c1);
  redc29<Fq29>(t, r.Y.c0);
  SB();
  wide_zero(t);
  wide_mac(t, R.c0, Q.c1);
  wide_mac(t, R.c1, Q.c0);
  wide_msub(t, p.Y.c0, PPP.c1);
  wide_msub(t, p.Y.c1, PPP.c0);
  redc29<Fq29>(t, r.Y.c1);
  return 0;
}

// madd-2008-s over Fq2, p + (x, negate ? -y : y), with the doubling and cancellation cases
// resolved (as g1l_madd).
HD G2L g2l_madd(const G2L& p, const Aff<Fm2i>& a, bool negate) {
  if (aff_is_inf(a)) return p;
  if (f2l_zero(p.ZZ)) return g2l_from_aff(a, negate);
  G2L r;
  const int k = g2l_madd_core(p, a, negate, r);
  if (k == 1) return g2l_from(xyzz_dbl_aff(negate ? aff_neg(a) : a));
  if (k == 2) return g2l_inf();
  return r;
}

// add-2008-s over Fq2 on the limb form, p + q; the doubling case goes through the word form.
// Bounds as g2l_madd's with U1 < 2q: P, R < 5q, T < 8q.
HD G2L g2l_add(const G2L& p, const G2L& q) {
  if (f2l_zero(p.ZZ)) return q;
  if (f2l_zero(q.ZZ)) return p;
  F2L U1, P, S1, R;
  f2l_mul(p.X, q.ZZ, U1);
  f2l_mul(q.X, p.ZZ, P);  // U2
  sub29<Kq3>(P.c0, U1.c0, P.c0);
  sub29<Kq3>(P.c1, U1.c1, P.c1);  // P = U2 - U1 + 3q < 5q
  SB();
  f2l_mul(p.Y, q.ZZZ, S1);
  f2l_mul(q.Y, p.ZZZ, R);  // S2
  sub29<Kq3>(R.c0, S1.c0, R.c0);
  sub29<Kq3>(R.c1, S1.c1, R.c1);  // R = S2 - S1 + 3q < 5q
  SB();
  F2L PP;
  f2l_sqr(P, PP);
  if (f2l_zero(PP)) {
    F2L R2;
    f2l_sqr(R, R2);
    if (f2l_zero(R2)) return g2l_from(xyzz_dbl(g2l_out(p)));
    return g2l_inf();
  }
  SB();
  F2L PPP, Z;
  f2l_mul(P, PP, PPP);
  SB();
  G2L r;
  f2l_mul(p.ZZ, q.ZZ, Z);
  f2l_mul(Z, PP, r.ZZ);
  SB();
  f2l_mul(p.ZZZ, q.ZZZ, Z);
  f2l_mul(Z, PPP, r.ZZZ);
  SB();
  F2L Q;
  f2l_mul(U1, PP, Q);
  SB();
  f2l_sqr(R, r.X);
#pragma unroll
  for (int j = 0; j < 9; ++j) {  // X3 = R^2 - PPP - 2Q + 7q < 9q
    r.X.c0[j] = r.X.c0[j] + Kq7::k(j) - PPP.c0[j] - Q.c0[j] - Q.c0[j];
    r.X.c1[j] = r.X.c1[j] + Kq7::k(j) - PPP.c1[j] - Q.c1[j] - Q.c1[j];
  }
  below5q(r.X.c0);
  below5q(r.X.c1);
  SB();
  sub29<Kq6>(Q.c0, r.X.c0, Q.c0);
  sub29<Kq6>(Q.c1, r.X.c1, Q.c1);  // T = Q - X3 + 6q < 8q
  uint64_t t[17];
  wide_zero(t);
  wide_mac(t, R.c0, Q.c0);
  wide_msub(t, R.c1, Q.c1);
  wide_msub(t, S1.c0, PPP.c0);
  wide_mac(t, S1.c1, PPP.c1);
  redc29<Fq29>(t, r.Y.c0);
  SB();
  wide_zero(t);
  wide_mac(t, R.c0, Q.c1);
  wide_mac(t, R.c1, Q.c0);
  wide_msub(t, S1.c0, PPP.c1);
  wide_msub(t, S1.c1, PPP.c0);
  redc29<Fq29>(t, r.Y.c1);
  return r;
}

// k * a for a plain (standard-form) 256-bit scalar, double-and-add from the top bit.
template <class F>
HD Xyzz<F> xyzz_mul_plain(const Aff<F>& a, const uint32_t k[8]) {
  Xyzz<F> acc = xyzz_inf<F>();
  for (int i = 255; i >= 0; --i) {
    acc = xyzz_dbl(acc);
    if ((k[i >> 5] >> (i & 31)) & 1) acc = xyzz_add_aff(acc, a);
  }
  return acc;
}

// k * p for an XYZZ point and a plain 256-bit scalar.
template <class F>
HD Xyzz<F> xyzz_mul(const Xyzz<F>& p, const uint32_t k[8]) {
  Xyzz<F> acc = xyzz_inf<F>();
  for (int i = 255; i >= 0; --i) {
    acc = xyzz_dbl(acc);
    if ((k[i >> 5] >> (i & 31)) & 1) acc = xyzz_add(acc, p);
  }
  return acc;
}

// Scalar multiplication by a small non-negative integer.
template <class F>
HD Xyzz<F> xyzz_mul_small(const Xyzz<F>& p, uint32_t k) {
  Xyzz<F> acc = xyzz_inf<F>();
  for (int i = 31; i >= 0; --i) {
    acc = xyzz_dbl(acc);
    if ((k >> i) & 1) acc = xyzz_add(acc, p);
  }
  return acc;
}

// Standard generators.
HD G1Aff g1_generator() {
  Fq x = fp_zero<FqCfg>(), y = fp_zero<FqCfg>();
  x.v[0] = 1;
  y.v[0] = 2;
  return {from_std(x), from_std(y)};
}

HD G2Aff g2_generator() {
  Fq x0, x1, y0, y1;
  const uint32_t X0[8] = {0xd992f6edu, 0x46debd5cu, 0xf75edaddu, 0x674322d4u,
                          0x5e5c4479u, 0x426a0066u, 0x121f1e76u, 0x1800deefu};
  const uint32_t X1[8] = {0xaef312c2u, 0x97e485b7u, 0x35a9e712u, 0xf1aa4933u,
                          0x31fb5d25u, 0x7260bfb7u, 0x920d483au, 0x198e9393u};
  const uint32_t Y0[8] = {0x66fa7daau, 0x4ce6cc01u, 0x0c43d37bu, 0xe3d1e769u,
                          0x8dcb408fu, 0x4aab7180u, 0xdb8c6debu, 0x12c85ea5u};
  const uint32_t Y1[8] = {0xd122975bu, 0x55acdadcu, 0x70b38ef3u, 0xbc4b3133u,
                          0x690c3395u, 0xec9e99adu, 0x585ff075u, 0x090689d0u};
  for (int i = 0; i < 8; ++i) {
    x0.v[i] = X0[i];
    x1.v[i] = X1[i];
    y0.v[i] = Y0[i];
    y1.v[i] = Y1[i];
  }
  return {{from_std(x0), from_std(x1)}, {from_std(y0), from_std(y1)}};
}

// ---------------------------------------------------------------------------------------
// Complete formulas for constant-work MSMs (Renes, Costello, Batina, "Complete addition
// formulas for prime order elliptic curves", 2016, Algorithms 7 and 8 for a = 0). A point is
// (X : Y : Z) with x = X / Z, y = Y / Z, and the identity (0 : 1 : 0); the formulas have no
// exceptional case (doubling, opposite points and the identity come out of the same
// products), so an addition's instruction stream does not depend on its operands. b3 = 3 b:
// 9 on G1 (y^2 = x^3 + 3), 9 / (9 + u) = (81 - 9 u) / 82 on the G2 twist
// (y^2 = x^3 + 3 / (9 + u)); the MSM form's constant below is that value times 2^261 mod q.
// ---------------------------------------------------------------------------------------
template <class F>
struct Proj {
  F X, Y, Z;
};

template <class C>
HD Fp<C> f_sel(bool c, const Fp<C>& a, const Fp<C>& b) {
  Fp<C> r;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = c ? a.v[i] : b.v[i];
  return r;
}
template <bool INL, class B>
HD Fq2T<INL, B> f_sel(bool c, const Fq2T<INL, B>& a, const Fq2T<INL, B>& b) {
  return {f_sel(c, a.c0, b.c0), f_sel(c, a.c1, b.c1)};
}
template <class F>
HD Proj<F> proj_sel(bool c, const Proj<F>& a, const Proj<F>& b) {
  return {f_sel(c, a.X, b.X), f_sel(c, a.Y, b.Y), f_sel(c, a.Z, b.Z)};
}
template <class F>
HD Proj<F> proj_inf() {
  return {f_zero<F>(), f_one<F>(), f_zero<F>()};
}

template <class F> struct CurveB3;
template <> struct CurveB3<Fm> {
  static HD Fm times(const Fm& x) {  // 9 x
    const Fm x8 = dbl(dbl(dbl(x)));
    return add(x8, x);
  }
};
template <> struct CurveB3<Fm2i> {
  static HD Fm2i k() {  // components below q
    constexpr uint32_t C0[8] = {0x295cc5c1u, 0x582d62c9u, 0x8acc80b0u, 0x9042cb3du,
                                0x9e1cb150u, 0xf6155910u, 0x1a987d08u, 0x1b2ff417u};
    constexpr uint32_t C1[8] = {0x01e5cc12u, 0xdeb7b4a1u, 0x9c61623au, 0x0b4079f4u,
                                0xfd7d8f2bu, 0x4b0dc77fu, 0x24d1bcddu, 0x17dd1076u};
    Fm2i r;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      r.c0.v[i] = C0[i];
      r.c1.v[i] = C1[i];
    }
    return r;
  }
  static HD Fm2i times(const Fm2i& x) { return mul(x, k()); }
};

// Algorithm 7: P + Q for any two points (12 products and 2 by b3).
template <class F>
HD Proj<F> proj_add(const Proj<F>& p, const Proj<F>& q) {
  F t0 = mul(p.X, q.X);
  F t1 = mul(p.Y, q.Y);
  F t2 = mul(p.Z, q.Z);
  SB();
  F t3 = sub(mul(add(p.X, p.Y), add(q.X, q.Y)), add(t0, t1));  // X1 Y2 + X2 Y1
  SB();
  F t4 = sub(mul(add(p.Y, p.Z), add(q.Y, q.Z)), add(t1, t2));  // Y1 Z2 + Y2 Z1
  SB();
  F t5 = sub(mul(add(p.X, p.Z), add(q.X, q.Z)), add(t0, t2));  // X1 Z2 + X2 Z1
  SB();
  t0 = add(dbl(t0), t0);                // 3 X1 X2
  t2 = CurveB3<F>::times(t2);           // b3 Z1 Z2
  const F s = add(t1, t2), d = sub(t1, t2);
  t5 = CurveB3<F>::times(t5);           // b3 (X1 Z2 + X2 Z1)
  Proj<F> r;
  r.X = mul_sub(t3, d, t4, t5);
  SB();
  r.Y = add(mul(s, d), mul(t5, t0));
  SB();
  r.Z = add(mul(t4, s), mul(t0, t3));
  return r;
}

// Algorithm 8: P + A for any point P and an affine A (not the identity: the affine encoding
// has none; a caller selects around it). 11 products and 2 by b3.
template <class F>
HD Proj<F> proj_add_aff(const Proj<F>& p, const Aff<F>& a) {
  F t0 = mul(p.X, a.x);
  F t1 = mul(p.Y, a.y);
  SB();
  F t3 = sub(mul(add(a.x, a.y), add(p.X, p.Y)), add(t0, t1));  // X1 Y2 + X2 Y1
  SB();
  F t4 = add(mul(a.y, p.Z), p.Y);  // Y1 + Y2 Z1
  SB();
  F t5 = add(mul(a.x, p.Z), p.X);  // X1 + X2 Z1
  SB();
  t0 = add(dbl(t0), t0);           // 3 X1 X2
  const F t2 = CurveB3<F>::times(p.Z);
  const F s = add(t1, t2), d = sub(t1, t2);
  t5 = CurveB3<F>::times(t5);
  Proj<F> r;
  r.X = mul_sub(t3, d, t4, t5);
  SB();
  r.Y = add(mul(s, d), mul(t5, t0));
  SB();
  r.Z = add(mul(t4, s), mul(t0, t3));
  return r;
}

// G1's complete formulas on 29-bit limbs (constant work): (X : Y : Z) with every coordinate
// below 2q in exact limbs, as the products leave them; the same algorithms, values mod q and
// outputs as proj_add / proj_add_aff. b3 = 9 is nine limb multiplies and a subtraction of a
// multiple of q (mul_b3_29), below 2q. Bounds, for inputs below 2q: t0, t1, t2 < 2q
// (3 t0 < 6q); the cross terms (cross29) t3 and, in Algorithm 7, t4, t5 lie in (q, 7q);
// Algorithm 8's t4, t5 (a product plus a coordinate) < 4q; s = t1 + b3 t2 < 4q,
// d = t1 - b3 t2 + 3q < 5q. The output sums stay below 70 q^2 (a reduction takes
// 169 q^2 = 2^261 q) and every factor is carried to limbs below 2^29 + 8 first, so a column
// holds at most 18 such products.
struct P1L {
  uint32_t X[9], Y[9], Z[9];
};

// 9 a mod q for a below 7q (limbs below 2^29 + 8), in exact limbs below 2q: v = 9 a exactly
// (below 63q < 2^261), less k q with k = floor(T / (qt + 1)), T and qt the top limbs of v and
// q. Then k <= T 2^232 / q <= v / q, and v / q - k < (T + qt + 1) / (qt (qt + 1)) + 1, which
// for T < 64 qt is below 1 + 2^-15.
HD void mul_b3_29(const uint32_t a[9], uint32_t r[9]) {
  constexpr uint32_t MASK = (1u << 29) - 1u;
  uint64_t c = 0;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const uint64_t v = 9ull * a[j] + c;
    r[j] = (uint32_t)v & MASK;
    c = v >> 29;
  }
  const uint32_t k = r[8] / (Fq29::p29(8) + 1u);
  int64_t d = 0;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const int64_t x = (int64_t)r[j] - (int64_t)k * Fq29::p29(j) + d;
    r[j] = (uint32_t)x & MASK;
    d = x >> 29;
  }
}

// a0 b1 + a1 b0 as (a0 + a1)(b0 + b1) - p0 - p1 + 5q, for a0, a1, b0, b1 below 2q and
// p0 = a0 b0, p1 = a1 b1 in exact limbs (below 2q): the value lies in (q, 7q).
HD void cross29(const uint32_t a0[9], const uint32_t a1[9], const uint32_t b0[9],
                const uint32_t b1[9], const uint32_t p0[9], const uint32_t p1[9],
                uint32_t r[9]) {
  uint32_t u[9], v[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    u[j] = a0[j] + a1[j];
    v[j] = b0[j] + b1[j];
  }
  norm29(u);
  norm29(v);
  mont29_limbs<Fq29>(u, v, r);  // 16 q^2 < 2^261 q: below 2q
#pragma unroll
  for (int j = 0; j < 9; ++j) r[j] = r[j] + Kq5::k(j) - p0[j] - p1[j];
  norm29(r);
}

HD P1L p1l_inf() {
  P1L r;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    r.X[j] = r.Z[j] = 0;
    r.Y[j] = one29(j);
  }
  return r;
}

HD P1L p1l_sel(bool c, const P1L& a, const P1L& b) {
  P1L r;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    r.X[j] = c ? a.X[j] : b.X[j];
    r.Y[j] = c ? a.Y[j] : b.Y[j];
    r.Z[j] = c ? a.Z[j] : b.Z[j];
  }
  return r;
}

HD P1L p1l_from(const Proj<Fm>& p) {
  P1L r;
  limbs29_of(p.X.v, r.X);
  limbs29_of(p.Y.v, r.Y);
  limbs29_of(p.Z.v, r.Z);
  return r;
}

HD Proj<Fm> p1l_out(const P1L& p) {
  Proj<Fm> r;
  pack29(p.X, r.X.v);
  pack29(p.Y, r.Y.v);
  pack29(p.Z, r.Z.v);
  return r;
}

// Both algorithms' common end: from t0 = X1 X2, t1 = Y1 Y2, t2 = Z1 Z2 (Algorithm 8: Z1) and
// the carried t3, t4, t5 it forms 3 t0, b3 t2, b3 t5 and the three output sums.
HD P1L p1l_finish(uint32_t t0[9], const uint32_t t1[9], const uint32_t t2[9],
                  const uint32_t t3[9], const uint32_t t4[9], const uint32_t t5[9]) {
  uint32_t b[9], s[9], d[9], b5[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) t0[j] *= 3u;
  norm29(t0);         // 3 t0 < 6q
  mul_b3_29(t2, b);   // b3 t2 < 2q
#pragma unroll
  for (int j = 0; j < 9; ++j) s[j] = t1[j] + b[j];
  norm29(s);             // s = t1 + b3 t2 < 4q
  sub29<Kq3>(t1, b, d);  // d = t1 - b3 t2 + 3q < 5q
  mul_b3_29(t5, b5);     // b3 t5 < 2q
  SB();
  P1L r;
  uint64_t w[17];
  wide_zero(w);
  wide_mac(w, t3, d);
  wide_msub(w, t4, b5);
  redc29<Fq29>(w, r.X);  // X3 = t3 d - t4 b3 t5, |.| < 49 q^2
  SB();
  wide_zero(w);
  wide_mac(w, s, d);
  wide_mac(w, b5, t0);
  redc29<Fq29, false>(w, r.Y);  // Y3 = s d + b3 t5 3 t0 < 32 q^2
  SB();
  wide_zero(w);
  wide_mac(w, t4, s);
  wide_mac(w, t0, t3);
  redc29<Fq29, false>(w, r.Z);  // Z3 = t4 s + 3 t0 t3 < 70 q^2
  return r;
}

// Algorithm 8: p + (x2, y2), an affine point (never the identity) with limbs below 2q.
HD P1L p1l_add_aff(const P1L& p, const uint32_t x2[9], const uint32_t y2[9]) {
  uint32_t t0[9], t1[9], t3[9], t4[9], t5[9];
  mont29_limbs<Fq29>(p.X, x2, t0);
  mont29_limbs<Fq29>(p.Y, y2, t1);
  SB();
  cross29(p.X, p.Y, x2, y2, t0, t1, t3);  // X1 y2 + Y1 x2
  SB();
  mont29_limbs<Fq29>(y2, p.Z, t4);
  mont29_limbs<Fq29>(x2, p.Z, t5);
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    t4[j] += p.Y[j];  // Y1 + y2 Z1 < 4q
    t5[j] += p.X[j];  // X1 + x2 Z1 < 4q
  }
  norm29(t4);
  norm29(t5);
  SB();
  return p1l_finish(t0, t1, p.Z, t3, t4, t5);
}

// Algorithm 7: p + q for any two points.
HD P1L p1l_add(const P1L& p, const P1L& q) {
  uint32_t t0[9], t1[9], t2[9], t3[9], t4[9], t5[9];
  mont29_limbs<Fq29>(p.X, q.X, t0);
  mont29_limbs<Fq29>(p.Y, q.Y, t1);
  mont29_limbs<Fq29>(p.Z, q.Z, t2);
  SB();
  cross29(p.X, p.Y, q.X, q.Y, t0, t1, t3);  // X1 Y2 + Y1 X2
  SB();
  cross29(p.Y, p.Z, q.Y, q.Z, t1, t2, t4);  // Y1 Z2 + Z1 Y2
  SB();
  cross29(p.X, p.Z, q.X, q.Z, t0, t2, t5);  // X1 Z2 + Z1 X2
  SB();
  return p1l_finish(t0, t1, t2, t3, t4, t5);
}

// G2's complete formulas on limbs (P2L), the same steps over Fq2 with the bounds above:
// components below 2q, the cross terms Karatsuba's (q, 7q), b3 t2 and b3 t5 Fq2 products with
// the constant (components below q; 2 7 1 < 169: below 2q), and each output component one wide
// sum of four limb products. A component with two products of each sign stays within 18
// products per column and sign (redc29); the all-positive second components (Y3: 64 q^2,
// Z3: 140 q^2 in Algorithm 7) hold 36 products per column and reduce unsigned (redc29u).
struct P2L {
  F2L X, Y, Z;
};

// redc29 for a wide value whose columns are all sums of products (no subtraction): up to 54
// limb products a column, so that with the reduction's nine they stay below 2^64, and the
// carries shift unsigned. The value is below 2^261 q, the result below 2q.
template <class M29>
HD void redc29u(uint64_t t[17], uint32_t R[9]) {
  constexpr uint32_t MASK = (1u << 29) - 1u;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const uint32_t m = ((uint32_t)t[i] * M29::N29) & MASK;
#pragma unroll
    for (int j = 0; j < 9; ++j) t[i + j] += (uint64_t)m * M29::p29(j);
    if (i < 8) t[i + 1] += t[i] >> 29;
  }
  uint64_t c = t[8] >> 29;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const uint64_t x = (j < 8 ? t[9 + j] : 0) + c;
    R[j] = (uint32_t)x & MASK;
    c = x >> 29;
  }
}

HD void f2l_b3(F2L& k) { f2l_of(CurveB3<Fm2i>::k(), k); }

HD P2L p2l_inf() {
  P2L r;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    r.X.c0[j] = r.X.c1[j] = r.Z.c0[j] = r.Z.c1[j] = r.Y.c1[j] = 0;
    r.Y.c0[j] = one29(j);
  }
  return r;
}

HD void f2l_sel(bool c, const F2L& a, const F2L& b, F2L& r) {
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    r.c0[j] = c ? a.c0[j] : b.c0[j];
    r.c1[j] = c ? a.c1[j] : b.c1[j];
  }
}

HD P2L p2l_sel(bool c, const P2L& a, const P2L& b) {
  P2L r;
  f2l_sel(c, a.X, b.X, r.X);
  f2l_sel(c, a.Y, b.Y, r.Y);
  f2l_sel(c, a.Z, b.Z, r.Z);
  return r;
}

HD P2L p2l_from(const Proj<Fm2i>& p) {
  P2L r;
  f2l_of(p.X, r.X);
  f2l_of(p.Y, r.Y);
  f2l_of(p.Z, r.Z);
  return r;
}

HD void f2l_pack(const F2L& a, Fm2i& r) {
  pack29(a.c0, r.c0.v);
  pack29(a.c1, r.c1.v);
}

HD Proj<Fm2i> p2l_out(const P2L& p) {
  Proj<Fm2i> r;
  f2l_pack(p.X, r.X);
  f2l_pack(p.Y, r.Y);
  f2l_pack(p.Z, r.Z);
  return r;
}

// cross29 over Fq2: a0 b1 + a1 b0 = (a0 + a1)(b0 + b1) - p0 - p1 (+ 5q per component).
HD void f2l_cross(const F2L& a0, const F2L& a1, const F2L& b0, const F2L& b1, const F2L& p0,
                  const F2L& p1, F2L& r) {
  F2L u, v;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    u.c0[j] = a0.c0[j] + a1.c0[j];
    u.c1[j] = a0.c1[j] + a1.c1[j];
    v.c0[j] = b0.c0[j] + b1.c0[j];
    v.c1[j] = b0.c1[j] + b1.c1[j];
  }
  norm29(u.c0);
  norm29(u.c1);
  norm29(v.c0);
  norm29(v.c1);
  f2l_mul(u, v, r);  // components below 4q: 2 16 q^2 < 2^261 q, below 2q
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    r.c0[j] = r.c0[j] + Kq5::k(j) - p0.c0[j] - p1.c0[j];
    r.c1[j] = r.c1[j] + Kq5::k(j) - p0.c1[j] - p1.c1[j];
  }
  norm29(r.c0);
  norm29(r.c1);
}

// r = a b + c d (wide, one reduction per component).
HD void f2l_mul2(const F2L& a, const F2L& b, const F2L& c, const F2L& d, F2L& r) {
  uint64_t t[17], u[17];
  wide_zero(t);
  wide_zero(u);
  wide_mac(t, a.c0, b.c0);
  wide_msub(t, a.c1, b.c1);
  wide_mac(t, c.c0, d.c0);
  wide_msub(t, c.c1, d.c1);
  wide_mac(u, a.c0, b.c1);
  wide_mac(u, a.c1, b.c0);
  wide_mac(u, c.c0, d.c1);
  wide_mac(u, c.c1, d.c0);
  redc29<Fq29>(t, r.c0);
  redc29u<Fq29>(u, r.c1);
}

HD P2L p2l_finish(F2L& t0, const F2L& t1, const F2L& t2, const F2L& t3, const F2L& t4,
                  const F2L& t5) {
  F2L k, b, s, d, b5;
  f2l_b3(k);
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    t0.c0[j] *= 3u;
    t0.c1[j] *= 3u;
  }
  norm29(t0.c0);
  norm29(t0.c1);  // 3 t0 < 6q
  f2l_mul(t2, k, b);  // b3 t2 < 2q
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    s.c0[j] = t1.c0[j] + b.c0[j];
    s.c1[j] = t1.c1[j] + b.c1[j];
  }
  norm29(s.c0);
  norm29(s.c1);               // s < 4q
  sub29<Kq3>(t1.c0, b.c0, d.c0);
  sub29<Kq3>(t1.c1, b.c1, d.c1);  // d < 5q
  f2l_mul(t5, k, b5);         // b3 t5 < 2q
  SB();
  P2L r;
  {  // X3 = t3 d - t4 b3 t5: |c0| < 49 q^2, |c1| < 70 q^2
    uint64_t t[17], u[17];
    wide_zero(t);
    wide_zero(u);
    wide_mac(t, t3.c0, d.c0);
    wide_msub(t, t3.c1, d.c1);
    wide_msub(t, t4.c0, b5.c0);
    wide_mac(t, t4.c1, b5.c1);
    wide_mac(u, t3.c0, d.c1);
    wide_mac(u, t3.c1, d.c0);
    wide_msub(u, t4.c0, b5.c1);
    wide_msub(u, t4.c1, b5.c0);
    redc29<Fq29>(t, r.X.c0);
    redc29<Fq29>(u, r.X.c1);
  }
  SB();
  f2l_mul2(s, d, b5, t0, r.Y);  // Y3 = s d + b3 t5 3 t0: |c0| < 32 q^2, c1 < 64 q^2
  SB();
  f2l_mul2(t4, s, t0, t3, r.Z);  // Z3 = t4 s + 3 t0 t3: |c0| < 70 q^2, c1 < 140 q^2
  return r;
}

// Algorithm 8 over Fq2: p + (x2, y2), components below 2q (never the identity).
HD P2L p2l_add_aff(const P2L& p, const F2L& x2, const F2L& y2) {
  F2L t0, t1, t3, t4, t5;
  f2l_mul(p.X, x2, t0);
  f2l_mul(p.Y, y2, t1);
  SB();
  f2l_cross(p.X, p.Y, x2, y2, t0, t1, t3);  // X1 y2 + Y1 x2
  SB();
  f2l_mul(y2, p.Z, t4);
  f2l_mul(x2, p.Z, t5);
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    t4.c0[j] += p.Y.c0[j];
    t4.c1[j] += p.Y.c1[j];
    t5.c0[j] += p.X.c0[j];
    t5.c1[j] += p.X.c1[j];
  }
  norm29(t4.c0);
  norm29(t4.c1);  // Y1 + y2 Z1 < 4q
  norm29(t5.c0);
  norm29(t5.c1);  // X1 + x2 Z1 < 4q
  SB();
  return p2l_finish(t0, t1, p.Z, t3, t4, t5);
}

// Algorithm 7 over Fq2.
HD P2L p2l_add(const P2L& p, const P2L& q) {
  F2L t0, t1, t2, t3, t4, t5;
  f2l_mul(p.X, q.X, t0);
  f2l_mul(p.Y, q.Y, t1);
  f2l_mul(p.Z, q.Z, t2);
  SB();
  f2l_cross(p.X, p.Y, q.X, q.Y, t0, t1, t3);
  SB();
  f2l_cross(p.Y, p.Z, q.Y, q.Z, t1, t2, t4);
  SB();
  f2l_cross(p.X, p.Z, q.X, q.Z, t0, t2, t5);
  SB();
  return p2l_finish(t0, t1, t2, t3, t4, t5);
}

// (X : Y : Z) as XYZZ (X Z, Y Z^2, Z^2, Z^3): the identity (Z = 0) becomes XYZZ's (ZZ = 0).
template <class F>
HD Xyzz<F> proj_to_xyzz(const Proj<F>& p) {
  Xyzz<F> r;
  r.ZZ = sqr(p.Z);
  r.ZZZ = mul(r.ZZ, p.Z);
  r.X = mul(p.X, p.Z);
  r.Y = mul(p.Y, r.ZZ);
  return r;
}

}  // namespace g16
