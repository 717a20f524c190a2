// Host check of the limb-form point arithmetic (G1L, G2L, P1L, P2L in snarkrs_amd/csrc/bn254.cuh)
// against the word-form XYZZ formulas of the same header: chains of mixed additions with
// random signs, repeated points (doublings) and opposite points (cancellations), general
// additions of such sums (including p + p and p + (-p)), and every value's bound after every
// step (G1 X < 9q, G2 X < 5q, the rest < 2q, limbs as the header states); the complete
// formulas on limbs (P1L, P2L) against proj_add_aff / proj_add, coordinate by coordinate,
// including every coordinate at 2q - 1, and G1's b3 step on its own. Points are random
// field elements, not curve points: the formulas are compared as algebra, not the group law.
// Prints "ok <steps>" and exits 0, or the first mismatches and exits 1.
#include <cstdio>
#include <random>
#include <vector>

#include "../../snarkrs_amd/csrc/bn254.cuh"

using namespace g16;

static std::mt19937_64 rng(2024);

static void rnd_below_2q(Fm& v) {
  for (;;) {
    for (int i = 0; i < 8; ++i) v.v[i] = (uint32_t)rng();
    v.v[7] &= 0x7fffffffu;
    for (int i = 7; i >= 0; --i)
      if (v.v[i] != fqm_2q(i)) {
        if (v.v[i] < fqm_2q(i)) return;
        break;
      }
  }
}

// value(x) < k q for limbs below 2^32 (any carries)
static bool below_kq(const uint32_t x[9], int k) {
  int64_t kq[10], xs[10];
  uint64_t c = 0;
  for (int j = 0; j < 9; ++j) {
    const uint64_t v = (uint64_t)Fq29::p29(j) * (uint64_t)k + c;
    kq[j] = (int64_t)(v & 0x1fffffffu);
    c = v >> 29;
  }
  kq[9] = (int64_t)c;
  int64_t cc = 0;
  for (int j = 0; j < 9; ++j) {
    const int64_t v = (int64_t)x[j] + cc;
    xs[j] = v & 0x1fffffff;
    cc = v >> 29;
  }
  xs[9] = cc;
  for (int j = 9; j >= 0; --j)
    if (xs[j] != kq[j]) return xs[j] < kq[j];
  return false;
}

static bool limbs_below(const uint32_t x[9], uint32_t lim) {
  for (int j = 0; j < 9; ++j)
    if (x[j] >= lim) return false;
  return true;
}

template <class F>
static bool same(const Xyzz<F>& p, const Xyzz<F>& q) {
  const bool pi = is_zero(p.ZZ), qi = is_zero(q.ZZ);
  if (pi || qi) return pi == qi;
  return is_zero(sub(mul(p.X, q.ZZ), mul(q.X, p.ZZ))) &&
         is_zero(sub(mul(p.Y, q.ZZZ), mul(q.Y, p.ZZZ)));
}

static long bad = 0, steps = 0;

static void report(const char* what, int chain, int step) {
  if (++bad <= 8) std::printf("%s: chain %d step %d\n", what, chain, step);
}

static bool g1_bounds(const G1L& a) {
  return below_kq(a.X, 9) && below_kq(a.Y, 2) && below_kq(a.ZZ, 2) && below_kq(a.ZZZ, 2) &&
         limbs_below(a.X, (1u << 29) + 8) && limbs_below(a.Y, 1u << 29) &&
         limbs_below(a.ZZ, 1u << 29) && limbs_below(a.ZZZ, 1u << 29);
}

static bool g2_bounds(const G2L& a) {
  const F2L* c[4] = {&a.X, &a.Y, &a.ZZ, &a.ZZZ};
  const int k[4] = {5, 2, 2, 2};
  for (int i = 0; i < 4; ++i)
    for (const uint32_t* v : {c[i]->c0, c[i]->c1})
      if (!below_kq(v, k[i]) || !limbs_below(v, 1u << 29)) return false;
  return true;
}

// One chain of mixed additions in both forms; returns the limb sum and the word-form sum.
template <class F, class L, class Madd, class Out, class Bounds, class Inf>
static void chains(const char* name, const std::vector<Aff<F>>& B, int n_chains, Madd madd,
                   Out out, Bounds bounds, Inf inf, std::vector<std::pair<L, Xyzz<F>>>* keep) {
  for (int ch = 0; ch < n_chains; ++ch) {
    Xyzz<F> ref = xyzz_inf<F>();
    L acc = inf();
    const int len = 1 + (int)(rng() % 40);
    int prev = -1;
    bool prev_neg = false;
    for (int s = 0; s < len; ++s) {
      int i = (int)(rng() % B.size());
      bool ng = rng() & 1;
      const int mode = (int)(rng() % 10);
      if (mode == 0 && prev >= 0) i = prev, ng = prev_neg;                // doubling when acc = it
      if (mode == 1 && prev >= 0) i = prev, ng = !prev_neg;               // cancellation
      if (mode == 2) acc = inf(), ref = xyzz_inf<F>();                     // a fresh run
      ref = xyzz_add_aff(ref, ng ? aff_neg(B[i]) : B[i]);
      acc = madd(acc, B[i], ng);
      ++steps;
      if (!same(out(acc), ref)) report(name, ch, s);
      if (!bounds(acc)) report("bounds", ch, s);
      prev = i;
      prev_neg = ng;
    }
    if (keep) keep->push_back({acc, ref});
  }
}

template <class F, class L, class Add, class Out, class Bounds>
static void sums(const char* name, const std::vector<std::pair<L, Xyzz<F>>>& pts, Add add,
                 Out out, Bounds bounds) {
  for (size_t i = 0; i + 1 < pts.size(); ++i) {
    const auto& p = pts[i];
    const auto& q = pts[i + 1];
    const L r = add(p.first, q.first);
    ++steps;
    if (!same(out(r), xyzz_add(p.second, q.second))) report(name, (int)i, 0);
    if (!bounds(r)) report("bounds (add)", (int)i, 0);
    const L d = add(p.first, p.first);  // the doubling case
    if (!same(out(d), xyzz_add(p.second, p.second))) report(name, (int)i, 1);
    if (!bounds(d)) report("bounds (add)", (int)i, 1);
  }
}

// G1's complete formulas (P1L) against proj_add_aff / proj_add: the same algorithms, so every
// coordinate agrees mod q, not only the point.
static bool same_proj(const Proj<Fm>& a, const Proj<Fm>& b) {
  return is_zero(sub(a.X, b.X)) && is_zero(sub(a.Y, b.Y)) && is_zero(sub(a.Z, b.Z));
}

static bool p1_bounds(const P1L& a) {
  return below_kq(a.X, 2) && below_kq(a.Y, 2) && below_kq(a.Z, 2) &&
         limbs_below(a.X, 1u << 29) && limbs_below(a.Y, 1u << 29) && limbs_below(a.Z, 1u << 29);
}

static P1L p1_add_aff(const P1L& p, const Aff<Fm>& a) {
  uint32_t x[9], y[9];
  limbs29_of(a.x.v, x);
  limbs29_of(a.y.v, y);
  return p1l_add_aff(p, x, y);
}

// G1's b3 = 9 (mul_b3_29) against the product with 9 2^261 mod q, for values below 7q: random
// ones and those whose 9 a has a top limb next to a multiple of (q's top limb + 1), where the
// quotient estimate changes.
static void b3_g1() {
  uint32_t nine[9];
  limbs29_of(CurveB3<Fm>::times(f_one<Fm>()).v, nine);
  const uint32_t qt1 = Fq29::p29(8) + 1u;
  for (int it = 0; it < 200000; ++it) {
    uint32_t a[9];
    for (int j = 0; j < 8; ++j) a[j] = (uint32_t)rng() & 0x1fffffffu;
    if (it & 1) {
      // 9 a's top limb lands near m (qt + 1): a's top limb near m (qt + 1) / 9
      const uint32_t m = 1 + (uint32_t)(rng() % 62);
      const int64_t t = (int64_t)m * qt1 / 9 + (int64_t)(rng() % 5) - 2;
      a[8] = (uint32_t)(t < 0 ? 0 : t);
    } else {
      a[8] = (uint32_t)(rng() % (7ull * Fq29::p29(8) + 7));
    }
    if (!below_kq(a, 7)) continue;
    uint32_t r[9], ref[9];
    mul_b3_29(a, r);
    mont29_limbs<Fq29>(a, nine, ref);
    Fm x, y;
    pack29(r, x.v);
    pack29(ref, y.v);
    ++steps;
    if (!below_kq(r, 2) || !limbs_below(r, 1u << 29) || !is_zero(sub(x, y)))
      report("G1 b3", it, 0);
  }
}

static void complete_g1(const std::vector<Aff<Fm>>& B) {
  std::vector<std::pair<P1L, Proj<Fm>>> keep;
  for (int ch = 0; ch < 400; ++ch) {
    Proj<Fm> ref = proj_inf<Fm>();
    P1L acc = p1l_inf();
    const int len = 1 + (int)(rng() % 40);
    int prev = -1;
    bool prev_neg = false;
    for (int s = 0; s < len; ++s) {
      int i = (int)(rng() % B.size());
      bool ng = rng() & 1;
      const int mode = (int)(rng() % 10);
      if (mode == 0 && prev >= 0) i = prev, ng = prev_neg;
      if (mode == 1 && prev >= 0) i = prev, ng = !prev_neg;
      const Aff<Fm> a = ng ? aff_neg(B[i]) : B[i];
      ref = proj_add_aff(ref, a);
      acc = p1_add_aff(acc, a);
      ++steps;
      if (!same_proj(p1l_out(acc), ref)) report("G1 complete add_aff", ch, s);
      if (!p1_bounds(acc)) report("bounds (complete)", ch, s);
      prev = i;
      prev_neg = ng;
    }
    keep.push_back({acc, ref});
  }
  for (size_t i = 0; i + 1 < keep.size(); ++i) {
    const auto& p = keep[i];
    const auto& q = keep[i + 1];
    for (int k = 0; k < 2; ++k) {
      const auto& r = k ? p : q;
      const P1L s = p1l_add(p.first, r.first);
      ++steps;
      if (!same_proj(p1l_out(s), proj_add(p.second, r.second))) report("G1 complete add", (int)i, k);
      if (!p1_bounds(s)) report("bounds (complete add)", (int)i, k);
    }
  }
  // Every coordinate at 2q - 1, the largest value the form holds.
  Fm big;
  for (int i = 0; i < 8; ++i) big.v[i] = fqm_2q(i);
  big.v[0] -= 1;
  const Proj<Fm> pb{big, big, big};
  const P1L lb = p1l_from(pb);
  steps += 2;
  if (!same_proj(p1l_out(p1_add_aff(lb, Aff<Fm>{big, big})), proj_add_aff(pb, Aff<Fm>{big, big})))
    report("G1 complete add_aff (2q - 1)", 0, 0);
  if (!same_proj(p1l_out(p1l_add(lb, lb)), proj_add(pb, pb)))
    report("G1 complete add (2q - 1)", 0, 0);
}

static bool same_proj(const Proj<Fm2i>& a, const Proj<Fm2i>& b) {
  return is_zero(sub(a.X, b.X)) && is_zero(sub(a.Y, b.Y)) && is_zero(sub(a.Z, b.Z));
}

static bool p2_bounds(const P2L& a) {
  for (const F2L* c : {&a.X, &a.Y, &a.Z})
    for (const uint32_t* v : {c->c0, c->c1})
      if (!below_kq(v, 2) || !limbs_below(v, 1u << 29)) return false;
  return true;
}

static P2L p2_add_aff(const P2L& p, const Aff<Fm2i>& a) {
  F2L x, y;
  f2l_of(a.x, x);
  f2l_of(a.y, y);
  return p2l_add_aff(p, x, y);
}

static void complete_g2(const std::vector<Aff<Fm2i>>& B) {
  std::vector<std::pair<P2L, Proj<Fm2i>>> keep;
  for (int ch = 0; ch < 200; ++ch) {
    Proj<Fm2i> ref = proj_inf<Fm2i>();
    P2L acc = p2l_inf();
    const int len = 1 + (int)(rng() % 30);
    int prev = -1;
    bool prev_neg = false;
    for (int s = 0; s < len; ++s) {
      int i = (int)(rng() % B.size());
      bool ng = rng() & 1;
      const int mode = (int)(rng() % 10);
      if (mode == 0 && prev >= 0) i = prev, ng = prev_neg;
      if (mode == 1 && prev >= 0) i = prev, ng = !prev_neg;
      const Aff<Fm2i> a = ng ? aff_neg(B[i]) : B[i];
      ref = proj_add_aff(ref, a);
      acc = p2_add_aff(acc, a);
      ++steps;
      if (!same_proj(p2l_out(acc), ref)) report("G2 complete add_aff", ch, s);
      if (!p2_bounds(acc)) report("bounds (complete G2)", ch, s);
      prev = i;
      prev_neg = ng;
    }
    keep.push_back({acc, ref});
  }
  for (size_t i = 0; i + 1 < keep.size(); ++i) {
    const auto& p = keep[i];
    const auto& q = keep[i + 1];
    for (int k = 0; k < 2; ++k) {
      const auto& r = k ? p : q;
      const P2L s = p2l_add(p.first, r.first);
      ++steps;
      if (!same_proj(p2l_out(s), proj_add(p.second, r.second))) report("G2 complete add", (int)i, k);
      if (!p2_bounds(s)) report("bounds (complete G2 add)", (int)i, k);
    }
  }
  Fm big;
  for (int i = 0; i < 8; ++i) big.v[i] = fqm_2q(i);
  big.v[0] -= 1;
  const Fm2i b2{big, big};
  const Proj<Fm2i> pb{b2, b2, b2};
  const P2L lb = p2l_from(pb);
  steps += 2;
  if (!same_proj(p2l_out(p2_add_aff(lb, Aff<Fm2i>{b2, b2})), proj_add_aff(pb, Aff<Fm2i>{b2, b2})))
    report("G2 complete add_aff (2q - 1)", 0, 0);
  if (!same_proj(p2l_out(p2l_add(lb, lb)), proj_add(pb, pb)))
    report("G2 complete add (2q - 1)", 0, 0);
}

int main() {
  std::vector<Aff<Fm>> B1(24);
  for (auto& b : B1) rnd_below_2q(b.x), rnd_below_2q(b.y);
  std::vector<std::pair<G1L, Xyzz<Fm>>> k1;
  chains<Fm, G1L>("G1 madd", B1, 600, g1l_madd, g1l_out, g1_bounds, g1l_inf, &k1);
  // p + (-p) through the general addition: negate Y (word form) and convert back.
  for (size_t i = 0; i < 40 && i < k1.size(); ++i) {
    Xyzz<Fm> w = g1l_out(k1[i].first);
    w.Y = neg(w.Y);
    const G1L r = g1l_add(k1[i].first, g1l_from(w));
    ++steps;
    if (!is_zero(g1l_out(r).ZZ) && !is_zero(k1[i].second.ZZ)) report("G1 p + (-p)", (int)i, 0);
  }
  sums<Fm, G1L>("G1 add", k1, g1l_add, g1l_out, g1_bounds);
  b3_g1();
  complete_g1(B1);

  std::vector<Aff<Fm2i>> B2(24);
  for (auto& b : B2) rnd_below_2q(b.x.c0), rnd_below_2q(b.x.c1), rnd_below_2q(b.y.c0),
      rnd_below_2q(b.y.c1);
  std::vector<std::pair<G2L, Xyzz<Fm2i>>> k2;
  chains<Fm2i, G2L>("G2 madd", B2, 300, g2l_madd, g2l_out, g2_bounds, g2l_inf, &k2);
  for (size_t i = 0; i < 20 && i < k2.size(); ++i) {
    Xyzz<Fm2i> w = g2l_out(k2[i].first);
    w.Y = neg(w.Y);
    const G2L r = g2l_add(k2[i].first, g2l_from(w));
    ++steps;
    if (!is_zero(g2l_out(r).ZZ) && !is_zero(k2[i].second.ZZ)) report("G2 p + (-p)", (int)i, 0);
  }
  sums<Fm2i, G2L>("G2 add", k2, g2l_add, g2l_out, g2_bounds);
  complete_g2(B2);

  if (bad) {
    std::printf("%ld mismatches in %ld steps\n", bad, steps);
    return 1;
  }
  std::printf("ok %ld\n", steps);
  return 0;
}
