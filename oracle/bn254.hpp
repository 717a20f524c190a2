// CPU oracle arithmetic for BN254 — TEST INFRASTRUCTURE ONLY.
//
// This header is the checker that the HIP path is compared against. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the library built from it.
// It is a plain restatement of the field and curve arithmetic the reference takes from
// arkworks 0.5.0 (ark-ff / ark-ec / ark-bn254; not vendored in the reference, called from
// crates/field/src/lib.rs:9-11), written independently of the device code:
//   * 4 x 64-bit limbs, Montgomery form with R = 2^256 (the radix arkworks uses, which is
//     why packed limbs cross the boundary unchanged: crates/gpu-layout/src/lib.rs:30-47);
//   * CIOS multiplication with unsigned __int128;
//   * Fq2 = Fq[u]/(u^2+1) multiplied schoolbook (4M), not Karatsuba as on the device;
//   * points in Jacobian coordinates (the device uses XYZZ), infinity = Z == 0.
// Constants are pinned by tests/test_oracle_field.py against Python big integers.
#pragma once
#include <cstdint>
#include <cstring>

namespace orc {

typedef unsigned __int128 u128;

struct FrParams {
  // r = 21888242871839275222246405745257275088548364400416034343698204186575808495617
  static constexpr uint64_t P[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL,
                                    0xb85045b68181585dULL, 0x30644e72e131a029ULL};
  static constexpr uint64_t N0 = 0xc2e1f593efffffffULL;  // -r^-1 mod 2^64
  static constexpr uint64_t R1[4] = {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL,
                                     0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL};
  static constexpr uint64_t R2[4] = {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL,
                                     0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL};
};

struct FqParams {
  // q = 21888242871839275222246405745257275088696311157297823662689037894645226208583
  static constexpr uint64_t P[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL,
                                    0xb85045b68181585dULL, 0x30644e72e131a029ULL};
  static constexpr uint64_t N0 = 0x87d20782e4866389ULL;
  static constexpr uint64_t R1[4] = {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL,
                                     0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL};
  static constexpr uint64_t R2[4] = {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL,
                                     0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL};
};

static inline bool geq4(const uint64_t a[4], const uint64_t b[4]) {
  for (int i = 3; i >= 0; --i) {
    if (a[i] != b[i]) return a[i] > b[i];
  }
  return true;
}

static inline uint64_t sub4(uint64_t out[4], const uint64_t a[4], const uint64_t b[4]) {
  uint64_t borrow = 0;
  for (int i = 0; i < 4; ++i) {
    u128 d = (u128)a[i] - b[i] - borrow;
    out[i] = (uint64_t)d;
    borrow = (uint64_t)(d >> 127) & 1;  // 1 when the difference went negative
  }
  return borrow;
}

static inline uint64_t add4(uint64_t out[4], const uint64_t a[4], const uint64_t b[4]) {
  uint64_t carry = 0;
  for (int i = 0; i < 4; ++i) {
    u128 s = (u128)a[i] + b[i] + carry;
    out[i] = (uint64_t)s;
    carry = (uint64_t)(s >> 64);
  }
  return carry;
}

template <class PP>
struct Fp {
  uint64_t v[4];  // Montgomery representative x*R mod p, fully reduced

  static Fp zero() { Fp r; std::memset(r.v, 0, 32); return r; }
  static Fp one() { Fp r; std::memcpy(r.v, PP::R1, 32); return r; }
  bool is_zero() const { return (v[0] | v[1] | v[2] | v[3]) == 0; }
  bool operator==(const Fp& o) const { return std::memcmp(v, o.v, 32) == 0; }
  bool operator!=(const Fp& o) const { return !(*this == o); }

  static Fp mont_mul(const uint64_t a[4], const uint64_t b[4]) {
    // CIOS (Koc, Acar, Kaliski 1996), one outer iteration per limb of b.
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
      uint64_t c = 0;
      for (int j = 0; j < 4; ++j) {
        u128 x = (u128)a[j] * b[i] + t[j] + c;
        t[j] = (uint64_t)x;
        c = (uint64_t)(x >> 64);
      }
      u128 x4 = (u128)t[4] + c;
      t[4] = (uint64_t)x4;
      t[5] = (uint64_t)(x4 >> 64);
      uint64_t m = t[0] * PP::N0;
      u128 y = (u128)m * PP::P[0] + t[0];
      c = (uint64_t)(y >> 64);
      for (int j = 1; j < 4; ++j) {
        y = (u128)m * PP::P[j] + t[j] + c;
        t[j - 1] = (uint64_t)y;
        c = (uint64_t)(y >> 64);
      }
      y = (u128)t[4] + c;
      t[3] = (uint64_t)y;
      t[4] = t[5] + (uint64_t)(y >> 64);
    }
    Fp r;
    std::memcpy(r.v, t, 32);
    if (t[4] != 0 || geq4(r.v, PP::P)) sub4(r.v, r.v, PP::P);
    return r;
  }

  Fp operator*(const Fp& o) const { return mont_mul(v, o.v); }
  Fp& operator*=(const Fp& o) { *this = mont_mul(v, o.v); return *this; }
  Fp sqr() const { return mont_mul(v, v); }
  Fp operator+(const Fp& o) const {
    Fp r;
    add4(r.v, v, o.v);  // p < 2^254, no carry out of 256 bits
    if (geq4(r.v, PP::P)) sub4(r.v, r.v, PP::P);
    return r;
  }
  Fp operator-(const Fp& o) const {
    Fp r;
    if (sub4(r.v, v, o.v)) add4(r.v, r.v, PP::P);
    return r;
  }
  Fp& operator+=(const Fp& o) { *this = *this + o; return *this; }
  Fp& operator-=(const Fp& o) { *this = *this - o; return *this; }
  Fp neg() const { return is_zero() ? *this : zero() - *this; }
  Fp dbl() const { return *this + *this; }

  // Standard-form integer <-> Montgomery representative.
  static Fp from_std(const uint64_t s[4]) { return mont_mul(s, PP::R2); }
  static Fp from_u64(uint64_t x) { uint64_t s[4] = {x, 0, 0, 0}; return from_std(s); }
  void to_std(uint64_t out[4]) const {
    static const uint64_t ONE[4] = {1, 0, 0, 0};
    Fp r = mont_mul(v, ONE);
    std::memcpy(out, r.v, 32);
  }
  bool is_one() const { return std::memcmp(v, PP::R1, 32) == 0; }

  // x^e for a little-endian 4-limb exponent, square-and-multiply from the top bit.
  Fp pow(const uint64_t e[4]) const {
    Fp acc = one();
    for (int i = 255; i >= 0; --i) {
      acc = acc.sqr();
      if ((e[i >> 6] >> (i & 63)) & 1) acc *= *this;
    }
    return acc;
  }
  Fp pow_u64(uint64_t e) const { uint64_t x[4] = {e, 0, 0, 0}; return pow(x); }
  // Fermat: x^(p-2). Zero maps to zero.
  Fp inv() const {
    uint64_t e[4];
    std::memcpy(e, PP::P, 32);
    e[0] -= 2;  // P[0] is odd and > 2, no borrow
    return pow(e);
  }
};

typedef Fp<FrParams> Fr;
typedef Fp<FqParams> Fq;

struct Fq2 {
  Fq c0, c1;
  static Fq2 zero() { return {Fq::zero(), Fq::zero()}; }
  static Fq2 one() { return {Fq::one(), Fq::zero()}; }
  bool is_zero() const { return c0.is_zero() && c1.is_zero(); }
  bool operator==(const Fq2& o) const { return c0 == o.c0 && c1 == o.c1; }
  bool operator!=(const Fq2& o) const { return !(*this == o); }
  Fq2 operator+(const Fq2& o) const { return {c0 + o.c0, c1 + o.c1}; }
  Fq2 operator-(const Fq2& o) const { return {c0 - o.c0, c1 - o.c1}; }
  Fq2& operator+=(const Fq2& o) { *this = *this + o; return *this; }
  Fq2& operator-=(const Fq2& o) { *this = *this - o; return *this; }
  // (a0 + a1 u)(b0 + b1 u) with u^2 = -1, schoolbook.
  Fq2 operator*(const Fq2& o) const {
    return {c0 * o.c0 - c1 * o.c1, c0 * o.c1 + c1 * o.c0};
  }
  Fq2& operator*=(const Fq2& o) { *this = *this * o; return *this; }
  Fq2 sqr() const { return *this * *this; }
  Fq2 neg() const { return {c0.neg(), c1.neg()}; }
  Fq2 dbl() const { return *this + *this; }
  Fq2 inv() const {
    Fq t = (c0.sqr() + c1.sqr()).inv();
    return {c0 * t, (c1 * t).neg()};
  }
  Fq2 scale(const Fq& k) const { return {c0 * k, c1 * k}; }
};

// Curve constants: G1: y^2 = x^3 + 3;  G2: y^2 = x^3 + 3/(9+u).
struct G1Curve {
  typedef Fq F;
  static F b() { return Fq::from_u64(3); }
};
struct G2Curve {
  typedef Fq2 F;
  static F b() {
    // 3/(9+u) = 27/82 - (3/82) u, standard-form limbs (tests recompute it from 3*(9+u)^-1).
    static const uint64_t B0[4] = {0x3267e6dc24a138e5ULL, 0xb5b4c5e559dbefa3ULL,
                                   0x81be18991be06ac3ULL, 0x2b149d40ceb8aaaeULL};
    static const uint64_t B1[4] = {0xe4a2bd0685c315d2ULL, 0xa74fa084e52d1852ULL,
                                   0xcd2cafadeed8fdf4ULL, 0x009713b03af0fed4ULL};
    return {Fq::from_std(B0), Fq::from_std(B1)};
  }
};

template <class C>
struct Affine {
  typedef typename C::F F;
  F x, y;
  bool inf;
  static Affine infinity() { return {F::zero(), F::zero(), true}; }
  bool on_curve() const {
    if (inf) return true;
    return y.sqr() == x.sqr() * x + C::b();
  }
  Affine neg() const { return {x, y.neg(), inf}; }
  bool operator==(const Affine& o) const {
    if (inf || o.inf) return inf == o.inf;
    return x == o.x && y == o.y;
  }
};

// Jacobian: (X, Y, Z) represents (X/Z^2, Y/Z^3); Z == 0 is the point at infinity.
template <class C>
struct Jac {
  typedef typename C::F F;
  F X, Y, Z;
  static Jac infinity() { return {F::one(), F::one(), F::zero()}; }
  static Jac from_affine(const Affine<C>& a) {
    if (a.inf) return infinity();
    return {a.x, a.y, F::one()};
  }
  bool is_inf() const { return Z.is_zero(); }

  // dbl-2009-l (a = 0).
  Jac dbl() const {
    if (is_inf()) return *this;
    F A = X.sqr(), B = Y.sqr(), Cc = B.sqr();
    F D = ((X + B).sqr() - A - Cc).dbl();
    F E = A.dbl() + A;
    F Fv = E.sqr();
    Jac r;
    r.X = Fv - D.dbl();
    F c8 = Cc.dbl().dbl().dbl();
    r.Y = E * (D - r.X) - c8;
    r.Z = (Y * Z).dbl();
    return r;
  }

  // add-2007-bl with the equal / opposite cases resolved.
  Jac add(const Jac& o) const {
    if (is_inf()) return o;
    if (o.is_inf()) return *this;
    F Z1Z1 = Z.sqr(), Z2Z2 = o.Z.sqr();
    F U1 = X * Z2Z2, U2 = o.X * Z1Z1;
    F S1 = Y * o.Z * Z2Z2, S2 = o.Y * Z * Z1Z1;
    F H = U2 - U1;
    F rr = (S2 - S1).dbl();
    if (H.is_zero()) {
      if (rr.is_zero()) return dbl();
      return infinity();
    }
    F I = H.dbl().sqr();
    F J = H * I;
    F V = U1 * I;
    Jac r;
    r.X = rr.sqr() - J - V.dbl();
    r.Y = rr * (V - r.X) - (S1 * J).dbl();
    r.Z = ((Z + o.Z).sqr() - Z1Z1 - Z2Z2) * H;
    return r;
  }

  // madd-2007-bl: this + affine point.
  Jac add_affine(const Affine<C>& a) const {
    if (a.inf) return *this;
    if (is_inf()) return from_affine(a);
    F Z1Z1 = Z.sqr();
    F U2 = a.x * Z1Z1;
    F S2 = a.y * Z * Z1Z1;
    F H = U2 - X;
    F rr = (S2 - Y).dbl();
    if (H.is_zero()) {
      if (rr.is_zero()) return dbl();
      return infinity();
    }
    F HH = H.sqr();
    F I = HH.dbl().dbl();
    F J = H * I;
    F V = X * I;
    Jac r;
    r.X = rr.sqr() - J - V.dbl();
    r.Y = rr * (V - r.X) - (Y * J).dbl();
    r.Z = (Z + H).sqr() - Z1Z1 - HH;
    return r;
  }

  Jac neg() const { return {X, Y.neg(), Z}; }

  Affine<C> to_affine() const {
    if (is_inf()) return Affine<C>::infinity();
    F zi = Z.inv();
    F zi2 = zi.sqr();
    return {X * zi2, Y * zi2 * zi, false};
  }

  // k * this, k a standard-form little-endian 256-bit integer.
  Jac mul(const uint64_t k[4]) const {
    Jac acc = infinity();
    for (int i = 255; i >= 0; --i) {
      acc = acc.dbl();
      if ((k[i >> 6] >> (i & 63)) & 1) acc = acc.add(*this);
    }
    return acc;
  }
};

typedef Affine<G1Curve> G1A;
typedef Affine<G2Curve> G2A;
typedef Jac<G1Curve> G1J;
typedef Jac<G2Curve> G2J;

// Standard generators (arkworks ark-bn254 G1 (1, 2); G2 as in EIP-197).
inline G1A g1_generator() {
  return {Fq::from_u64(1), Fq::from_u64(2), false};
}
inline G2A g2_generator() {
  static const uint64_t X0[4] = {0x46debd5cd992f6edULL, 0x674322d4f75edaddULL,
                                 0x426a00665e5c4479ULL, 0x1800deef121f1e76ULL};
  static const uint64_t X1[4] = {0x97e485b7aef312c2ULL, 0xf1aa493335a9e712ULL,
                                 0x7260bfb731fb5d25ULL, 0x198e9393920d483aULL};
  static const uint64_t Y0[4] = {0x4ce6cc0166fa7daaULL, 0xe3d1e7690c43d37bULL,
                                 0x4aab71808dcb408fULL, 0x12c85ea5db8c6debULL};
  static const uint64_t Y1[4] = {0x55acdadcd122975bULL, 0xbc4b313370b38ef3ULL,
                                 0xec9e99ad690c3395ULL, 0x090689d0585ff075ULL};
  return {{Fq::from_std(X0), Fq::from_std(X1)}, {Fq::from_std(Y0), Fq::from_std(Y1)}, false};
}

}  // namespace orc
