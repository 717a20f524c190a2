// Host field arithmetic on four 64-bit limbs (Montgomery form, R = 2^256, so the boundary's
// 8 x u32 little-endian limbs are these limbs unchanged): Fq, Fr and Fq2 = Fq[u]/(u^2 + 1).
// Used by the host verifier and by the host's finish of the device MSMs, where the 32-bit
// limb code the kernels use would run at a third of the speed.
#pragma once
#include <cstdint>
#include <cstring>

namespace g16 {
namespace h64 {

typedef unsigned __int128 u128;

struct QTag {  // base field q
  static constexpr uint64_t P[4] = {0x3c208c16d87cfd47ull, 0x97816a916871ca8dull,
                                    0xb85045b68181585dull, 0x30644e72e131a029ull};
  static constexpr uint64_t INV = 0x87d20782e4866389ull;  // -q^-1 mod 2^64
};
struct RTag {  // scalar field r
  static constexpr uint64_t P[4] = {0x43e1f593f0000001ull, 0x2833e84879b97091ull,
                                    0xb85045b68181585dull, 0x30644e72e131a029ull};
  static constexpr uint64_t INV = 0xc2e1f593efffffffull;
};
static_assert(QTag::P[0] * QTag::INV + 1 == 0, "q Montgomery constant");
static_assert(RTag::P[0] * RTag::INV + 1 == 0, "r Montgomery constant");

template <class T>
struct Fp {
  uint64_t v[4];
};
typedef Fp<QTag> Fq;
typedef Fp<RTag> Fr;

template <class T>
inline bool lt_p(const uint64_t a[4]) {
  for (int i = 3; i >= 0; --i)
    if (a[i] != T::P[i]) return a[i] < T::P[i];
  return false;
}

template <class T>
inline void sub_p(uint64_t a[4]) {
  uint64_t borrow = 0;
  for (int i = 0; i < 4; ++i) {
    u128 t = (u128)a[i] - T::P[i] - borrow;
    a[i] = (uint64_t)t;
    borrow = (uint64_t)(t >> 64) ? 1 : 0;
  }
}

template <class T>
inline Fp<T> f_zero() {
  return Fp<T>{{0, 0, 0, 0}};
}

template <class T>
inline bool f_is_zero(const Fp<T>& a) {
  return !(a.v[0] | a.v[1] | a.v[2] | a.v[3]);
}

template <class T>
inline bool f_eq(const Fp<T>& a, const Fp<T>& b) {
  return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2] && a.v[3] == b.v[3];
}

// a - p when a >= p, else a, without branches (a < 2p < 2^256).
template <class T>
inline void reduce_once(uint64_t a[4]) {
  uint64_t d[4], borrow = 0;
  for (int i = 0; i < 4; ++i) {
    u128 t = (u128)a[i] - T::P[i] - borrow;
    d[i] = (uint64_t)t;
    borrow = (uint64_t)(t >> 64) & 1;
  }
  uint64_t keep = 0 - borrow;  // all ones when a < p
  for (int i = 0; i < 4; ++i) a[i] = (a[i] & keep) | (d[i] & ~keep);
}

template <class T>
inline Fp<T> f_add(const Fp<T>& a, const Fp<T>& b) {
  Fp<T> o;
  uint64_t carry = 0;
  for (int i = 0; i < 4; ++i) {
    u128 t = (u128)a.v[i] + b.v[i] + carry;
    o.v[i] = (uint64_t)t;
    carry = (uint64_t)(t >> 64);
  }
  reduce_once<T>(o.v);  // both moduli are below 2^254: no carry out
  return o;
}

template <class T>
inline Fp<T> f_sub(const Fp<T>& a, const Fp<T>& b) {
  Fp<T> o;
  uint64_t borrow = 0;
  for (int i = 0; i < 4; ++i) {
    u128 t = (u128)a.v[i] - b.v[i] - borrow;
    o.v[i] = (uint64_t)t;
    borrow = (uint64_t)(t >> 64) & 1;
  }
  uint64_t mask = 0 - borrow, carry = 0;
  for (int i = 0; i < 4; ++i) {
    u128 t = (u128)o.v[i] + (T::P[i] & mask) + carry;
    o.v[i] = (uint64_t)t;
    carry = (uint64_t)(t >> 64);
  }
  return o;
}

template <class T>
inline Fp<T> f_neg(const Fp<T>& a) {
  return f_is_zero(a) ? a : f_sub(f_zero<T>(), a);
}

// a b / 2^256 mod p: CIOS with the carry words dropped, which is exact when the modulus'
// top limb is below 2^62 (true of both moduli here), then one conditional subtraction.
template <class T>
inline Fp<T> f_mul(const Fp<T>& a, const Fp<T>& b) {
  static_assert(T::P[3] < (1ull << 62), "the carry-free product needs p < 2^254");
  uint64_t t[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    u128 s = (u128)a.v[0] * b.v[i] + t[0];
    uint64_t A = (uint64_t)(s >> 64);
    t[0] = (uint64_t)s;
    uint64_t m = t[0] * T::INV;
    u128 u = (u128)m * T::P[0] + t[0];
    uint64_t C = (uint64_t)(u >> 64);
    for (int j = 1; j < 4; ++j) {
      s = (u128)a.v[j] * b.v[i] + t[j] + A;
      A = (uint64_t)(s >> 64);
      u = (u128)m * T::P[j] + (uint64_t)s + C;
      C = (uint64_t)(u >> 64);
      t[j - 1] = (uint64_t)u;
    }
    t[3] = C + A;
  }
  Fp<T> o{{t[0], t[1], t[2], t[3]}};
  reduce_once<T>(o.v);
  return o;
}

template <class T>
inline Fp<T> f_sqr(const Fp<T>& a) {
  return f_mul(a, a);
}

template <class T>
inline Fp<T> f_dbl(const Fp<T>& a) {
  return f_add(a, a);
}

// 2^256 mod p and 2^512 mod p by doubling, computed once.
template <class T>
struct Consts {
  Fp<T> one, r2;
  Consts() {
    Fp<T> x{{1, 0, 0, 0}};
    for (int i = 0; i < 256; ++i) x = f_add(x, x);
    one = x;
    for (int i = 0; i < 256; ++i) x = f_add(x, x);
    r2 = x;
  }
};
template <class T>
const Consts<T>& consts() {
  static const Consts<T> c;
  return c;
}

template <class T>
inline Fp<T> f_one() {
  return consts<T>().one;
}

template <class T>
inline bool f_is_one(const Fp<T>& a) {
  return f_eq(a, f_one<T>());
}

template <class T>
inline Fp<T> f_from_u64(uint64_t k) {
  return f_mul(Fp<T>{{k, 0, 0, 0}}, consts<T>().r2);
}

// Montgomery form -> the integer, and back.
template <class T>
inline Fp<T> f_to_int(const Fp<T>& a) {
  return f_mul(a, Fp<T>{{1, 0, 0, 0}});
}
template <class T>
inline Fp<T> f_from_int(const Fp<T>& a) {
  return f_mul(a, consts<T>().r2);
}

struct Fq2 {
  Fq c0, c1;
};
inline Fq2 e2_zero() { return {f_zero<QTag>(), f_zero<QTag>()}; }
inline Fq2 e2_one() { return {f_one<QTag>(), f_zero<QTag>()}; }
inline bool e2_is_zero(const Fq2& a) { return f_is_zero(a.c0) && f_is_zero(a.c1); }
inline bool e2_eq(const Fq2& a, const Fq2& b) { return f_eq(a.c0, b.c0) && f_eq(a.c1, b.c1); }
inline Fq2 e2_add(const Fq2& a, const Fq2& b) { return {f_add(a.c0, b.c0), f_add(a.c1, b.c1)}; }
inline Fq2 e2_sub(const Fq2& a, const Fq2& b) { return {f_sub(a.c0, b.c0), f_sub(a.c1, b.c1)}; }
inline Fq2 e2_neg(const Fq2& a) { return {f_neg(a.c0), f_neg(a.c1)}; }
inline Fq2 e2_dbl(const Fq2& a) { return e2_add(a, a); }
inline Fq2 e2_mul(const Fq2& a, const Fq2& b) {
  Fq t0 = f_mul(a.c0, b.c0), t1 = f_mul(a.c1, b.c1);
  Fq m = f_mul(f_add(a.c0, a.c1), f_add(b.c0, b.c1));
  return {f_sub(t0, t1), f_sub(f_sub(m, t0), t1)};
}
inline Fq2 e2_sqr(const Fq2& a) {  // (a0 + a1)(a0 - a1) + 2 a0 a1 u
  Fq m = f_mul(a.c0, a.c1);
  return {f_mul(f_add(a.c0, a.c1), f_sub(a.c0, a.c1)), f_add(m, m)};
}

}  // namespace h64
}  // namespace g16
