// CPU oracle for the Groth16 prove path — TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load liboracle.so.
// It restates the reference CPU back end (a Rust crate that cannot be built here: no cargo)
// in plain C++ so the HIP path can be checked element by element:
//
//   stage 0  gather              crates/groth16/src/cpu.rs:140-174, C = A o B :331-337
//   stages 1-3 iNTT/coset/NTT    crates/ntt/src/lib.rs:508-550 (same result, natural-order
//                                radix-2 here), coset shift crates/field/src/lib.rs:124-130
//   stage 4  H = A*B - C         crates/groth16/src/cpu.rs:230-245 (no division by Z)
//   stages 5-9 five MSMs         crates/groth16/src/cpu.rs:477-632
//   Pippenger                    crates/msm/src/lib.rs:160-172 (window_size),
//                                :275-331 (carry-free signed digits), :534-690 (prescan)
//   stage 11 assembly            crates/groth16/src/prove.rs:391-430
//   Domain                       crates/field/src/lib.rs:48-82
//
// Every function is parity-pinned by tests/ (golden vectors, closed forms, brute force);
// the header of each section names the test that pins it.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <thread>
#include <vector>

#include "bn254.hpp"

using namespace orc;

// ---------------------------------------------------------------------------------------
// threads
// ---------------------------------------------------------------------------------------
static int g_threads = 0;

static int n_threads() {
  if (g_threads > 0) return g_threads;
  const char* e = std::getenv("ORC_THREADS");
  int t = e ? std::atoi(e) : 0;
  if (t <= 0) t = (int)std::thread::hardware_concurrency();
  return t > 0 ? t : 1;
}

// Runs fn(lo, hi) over [0, n) split into contiguous chunks, one per thread.
static void parallel_for(size_t n, const std::function<void(size_t, size_t)>& fn,
                         size_t min_chunk = 1024) {
  int t = n_threads();
  size_t chunks = std::min<size_t>((size_t)t, (n + min_chunk - 1) / min_chunk);
  if (chunks <= 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> pool;
  size_t step = (n + chunks - 1) / chunks;
  for (size_t c = 0; c < chunks; ++c) {
    size_t lo = c * step, hi = std::min(n, lo + step);
    if (lo >= hi) break;
    pool.emplace_back(fn, lo, hi);
  }
  for (auto& th : pool) th.join();
}

// Runs fn(i) for i in [0, n) on a pool pulling indices from an atomic counter.
static void parallel_tasks(size_t n, const std::function<void(size_t)>& fn) {
  int t = std::min<int>(n_threads(), (int)n);
  if (t <= 1) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (int k = 0; k < t; ++k)
    pool.emplace_back([&] {
      for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
    });
  for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------------------
// wire format: 32-byte little-endian limbs, Montgomery for field elements and point
// coordinates, the all-zero encoding for the point at infinity
// (crates/gpu-layout/src/lib.rs:120-185).
// ---------------------------------------------------------------------------------------
template <class F>
static F load_f(const uint8_t* p) {
  F x;
  std::memcpy(x.v, p, 32);
  return x;
}
template <class F>
static void store_f(uint8_t* p, const F& x) { std::memcpy(p, x.v, 32); }

static Fq2 load_fq2(const uint8_t* p) { return {load_f<Fq>(p), load_f<Fq>(p + 32)}; }
static void store_fq2(uint8_t* p, const Fq2& x) {
  store_f(p, x.c0);
  store_f(p + 32, x.c1);
}

static G1A load_g1(const uint8_t* p) {
  G1A a{load_f<Fq>(p), load_f<Fq>(p + 32), false};
  a.inf = a.x.is_zero() && a.y.is_zero();
  return a;
}
static void store_g1(uint8_t* p, const G1A& a) {
  if (a.inf) {
    std::memset(p, 0, 64);
    return;
  }
  store_f(p, a.x);
  store_f(p + 32, a.y);
}
static G2A load_g2(const uint8_t* p) {
  G2A a{load_fq2(p), load_fq2(p + 64), false};
  a.inf = a.x.is_zero() && a.y.is_zero();
  return a;
}
static void store_g2(uint8_t* p, const G2A& a) {
  if (a.inf) {
    std::memset(p, 0, 128);
    return;
  }
  store_fq2(p, a.x);
  store_fq2(p + 64, a.y);
}

template <class C> struct PointIO;
template <> struct PointIO<G1Curve> {
  static constexpr size_t BYTES = 64;
  static G1A load(const uint8_t* p) { return load_g1(p); }
  static void store(uint8_t* p, const G1A& a) { store_g1(p, a); }
};
template <> struct PointIO<G2Curve> {
  static constexpr size_t BYTES = 128;
  static G2A load(const uint8_t* p) { return load_g2(p); }
  static void store(uint8_t* p, const G2A& a) { store_g2(p, a); }
};

// ---------------------------------------------------------------------------------------
// Domain (crates/field/src/lib.rs:48-82, :124-130).  Pinned by test_oracle_field.py
// (ffjavascript roots for 2^28 and 2^4, crates/groth16/src/prove.rs:549-563) and by the
// coset identities of crates/groth16/src/cpu.rs:915-927.
// ---------------------------------------------------------------------------------------
static Fr two_adic_root() {
  // 5^((r-1)/2^28), the ffjavascript / arkworks root of order 2^28.
  static const uint64_t W[4] = {0x9bd61b6e725b19f0ULL, 0x402d111e41112ed4ULL,
                                0x00e0a7eb8ef62abcULL, 0x2a3c09f0a58a7e85ULL};
  return Fr::from_std(W);
}

static Fr domain_root(int log_n) {
  Fr g = two_adic_root();
  for (int i = log_n; i < 28; ++i) g = g.sqr();
  return g;
}

// ---------------------------------------------------------------------------------------
// NTT, natural order in and out (pinned by the naive DFT, impulse and convolution tests in
// test_oracle_ntt.py, the pattern of crates/ntt/src/lib.rs:1298-1380).
// ---------------------------------------------------------------------------------------
static void bit_reverse(Fr* a, int log_n) {
  size_t n = (size_t)1 << log_n;
  for (size_t i = 0; i < n; ++i) {
    size_t j = 0;
    for (int b = 0; b < log_n; ++b) j |= ((i >> b) & 1) << (log_n - 1 - b);
    if (i < j) std::swap(a[i], a[j]);
  }
}

// a_k <- sum_j a_j * root^(jk).  Iterative radix-2 DIT after a bit-reversal.
static void ntt_with_root(Fr* a, int log_n, const Fr& root) {
  if (log_n == 0) return;
  size_t n = (size_t)1 << log_n;
  bit_reverse(a, log_n);
  std::vector<Fr> tw(n / 2);
  tw[0] = Fr::one();
  for (size_t i = 1; i < n / 2; ++i) tw[i] = tw[i - 1] * root;
  for (int s = 0; s < log_n; ++s) {
    size_t half = (size_t)1 << s;
    size_t stride = n / (2 * half);
    parallel_for(n / 2, [&](size_t lo, size_t hi) {
      for (size_t b = lo; b < hi; ++b) {
        size_t blk = b / half, j = b % half;
        size_t i0 = blk * 2 * half + j, i1 = i0 + half;
        Fr t = a[i1] * tw[j * stride];
        Fr u = a[i0];
        a[i0] = u + t;
        a[i1] = u - t;
      }
    }, 4096);
  }
}

static void ntt_forward(Fr* a, int log_n) { ntt_with_root(a, log_n, domain_root(log_n)); }

static void ntt_inverse(Fr* a, int log_n) {
  ntt_with_root(a, log_n, domain_root(log_n).inv());
  Fr ninv = Fr::from_u64((uint64_t)1 << log_n).inv();
  size_t n = (size_t)1 << log_n;
  parallel_for(n, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) a[i] *= ninv;
  });
}

// Stages 1-3 for one vector: coefficients of the evaluations, times shift^j, re-evaluated
// on the domain: out[k] = P(shift * w^k), shift = primitive 2n-th root (snarkjs' odd coset,
// crates/field/src/lib.rs:124-130, crates/ntt/src/lib.rs:508-550).
static void coset_pipeline(Fr* a, int log_n) {
  size_t n = (size_t)1 << log_n;
  ntt_inverse(a, log_n);
  Fr shift = domain_root(log_n + 1);
  parallel_for(n, [&](size_t lo, size_t hi) {
    Fr acc = shift.pow_u64(lo);
    for (size_t i = lo; i < hi; ++i) {
      a[i] *= acc;
      acc *= shift;
    }
  });
  ntt_forward(a, log_n);
}

// ---------------------------------------------------------------------------------------
// Key (the same layout libg16hip takes across its C ABI).
// ---------------------------------------------------------------------------------------
extern "C" {
typedef struct {
  uint32_t n_vars, n_public, domain_size, n_values;
  const uint8_t *alpha_g1, *beta_g1, *beta_g2, *gamma_g2, *delta_g1, *delta_g2;
  const uint8_t* ic;
  const uint32_t* row_ptr[2];
  const uint32_t* signal[2];
  const uint32_t* value_id[2];
  const uint8_t* values;
  const uint8_t *a_query, *b_g1_query, *b_g2_query, *l_query, *h_query;
  uint32_t h_basis;  // 0 snarkjs' odd coset (domain_size H bases), 1 gnark's (domain_size - 1)
} orc_key;
}

static int log2_exact(uint32_t n) {
  if (n == 0 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1u << l) < n) ++l;
  return l;
}

// Stage 0 for one matrix, crates/groth16/src/cpu.rs:140-174 (zero witness skipped, one
// added without a multiply; both bit-identical to the plain sum).
static void gather(const orc_key* k, int m, const Fr* w, Fr* out) {
  const uint32_t* rp = k->row_ptr[m];
  parallel_for(k->domain_size, [&](size_t lo, size_t hi) {
    for (size_t c = lo; c < hi; ++c) {
      Fr acc = Fr::zero();
      for (uint32_t e = rp[c]; e < rp[c + 1]; ++e) {
        const Fr& x = w[k->signal[m][e]];
        if (x.is_zero()) continue;
        Fr v = load_f<Fr>(k->values + 32 * (size_t)k->value_id[m][e]);
        acc += x.is_one() ? v : v * x;
      }
      out[c] = acc;
    }
  });
}

// Stages 1-4 on stage 0's vectors (crates/groth16/src/cpu.rs:191-245): H = A*B - C on the
// odd coset.
static void coset_h(std::vector<Fr>& a, std::vector<Fr>& b, std::vector<Fr>& c, int log_n, Fr* h) {
  size_t n = a.size();
  coset_pipeline(a.data(), log_n);
  coset_pipeline(b.data(), log_n);
  coset_pipeline(c.data(), log_n);
  for (size_t i = 0; i < n; ++i) h[i] = a[i] * b[i] - c[i];
}

// Stages 0-4 (crates/groth16/src/cpu.rs:316-341, :191-245).
static void compute_h(const orc_key* k, const Fr* w, Fr* h) {
  size_t n = k->domain_size;
  int log_n = log2_exact(k->domain_size);
  std::vector<Fr> a(n), b(n), c(n);
  gather(k, 0, w, a.data());
  gather(k, 1, w, b.data());
  for (size_t i = 0; i < n; ++i) c[i] = a[i] * b[i];
  coset_h(a, b, c, log_n, h);
}

// Stage 0 from the caller's rows (crates/groth16/src/cpu.rs:387-424): A.w and B.w padded
// with zero rows to the domain, C = A*B on the rows.
static void rows_h(size_t n, const Fr* a_rows, const Fr* b_rows, size_t rows, Fr* h) {
  int log_n = log2_exact((uint32_t)n);
  std::vector<Fr> a(n, Fr::zero()), b(n, Fr::zero()), c(n, Fr::zero());
  for (size_t i = 0; i < rows; ++i) {
    a[i] = a_rows[i];
    b[i] = b_rows[i];
    c[i] = a[i] * b[i];
  }
  coset_h(a, b, c, log_n, h);
}

// H on the odd coset to gnark's basis (crates/groth16/src/cpu.rs:250-286): the coefficients
// of V(shift x) by the inverse transform, V's coefficient j = g_j shift^-j, placed at
// bitrev(j). These are H's times -2 (t = x^n - 1 = -2 on the odd coset); the caller scales the
// MSM by -1/2, as the reference does. Returns false if coefficient n - 1 is not zero.
static bool to_gnark_coefficients(Fr* h, int log_n) {
  size_t n = (size_t)1 << log_n;
  ntt_inverse(h, log_n);
  Fr shift_inv = domain_root(log_n + 1).inv();
  std::vector<Fr> out(n);
  Fr acc = Fr::one();
  for (size_t j = 0; j < n; ++j) {
    size_t p = 0;
    for (int bit = 0; bit < log_n; ++bit) p |= ((j >> bit) & 1) << (log_n - 1 - bit);
    out[p] = h[j] * acc;
    acc *= shift_inv;
  }
  std::copy(out.begin(), out.end(), h);
  return out[n - 1].is_zero();
}

// ---------------------------------------------------------------------------------------
// MSM: signed-digit Pippenger (crates/msm/src/lib.rs) and a naive double-and-add sum.
// Pinned by test_oracle_msm.py: naive sums, the closed form (sum k_i s_i)*G, digit
// reconstruction (crates/msm/src/lib.rs:1974-2062), window_size KATs (:1969-1970) and the
// adversarial scalar/base classes of crates/cuda/tests/adversarial_gpu.rs:224-281.
// ---------------------------------------------------------------------------------------
static const int RECODE_BITS = 255;  // crates/msm/src/lib.rs:94

// crates/msm/src/lib.rs:160-172: minimise ceil(255/c) * (n + 3 * 2^(c-1)) over c in [3,16].
static int window_size(uint64_t n) {
  int best = 3;
  unsigned __int128 best_cost = ~(unsigned __int128)0;
  for (int c = 3; c <= 16; ++c) {
    unsigned __int128 windows = (RECODE_BITS + c - 1) / c;
    unsigned __int128 cost = windows * ((unsigned __int128)n + 3 * ((unsigned __int128)1 << (c - 1)));
    if (cost < best_cost) {
      best_cost = cost;
      best = c;
    }
  }
  return best;
}

static uint64_t read_bits(const uint64_t s[4], int off, int width) {
  int idx = off / 64;
  if (idx >= 4) return 0;
  int sh = off % 64;
  uint64_t buf = s[idx] >> sh;
  if (sh + width > 64 && idx + 1 < 4) buf |= s[idx + 1] << (64 - sh);
  return buf & ((1ULL << width) - 1);
}

// crates/msm/src/lib.rs:314-331: window i's digit in [-2^(c-1), 2^(c-1)], carry-free.
static int64_t signed_digit(const uint64_t s[4], int i, int c) {
  int off = i * c;
  int64_t b = (int64_t)read_bits(s, off, c);
  int64_t d = b - ((b >> (c - 1)) << c);
  int64_t carry = off == 0 ? 0 : (int64_t)read_bits(s, off - 1, 1);
  return d + carry;
}

struct Scalar4 {
  uint64_t v[4];
};

template <class C>
static Jac<C> msm_pippenger(const uint8_t* bases, const Fr* scalars, size_t n, int force_c) {
  typedef PointIO<C> IO;
  // Prescan (crates/msm/src/lib.rs:534-690): drop infinity bases and zero scalars, fold
  // the ones into one running sum, and keep the rest for the bucket loop.
  std::vector<uint32_t> idx;
  std::vector<Scalar4> big;
  std::vector<uint32_t> ones;
  for (size_t i = 0; i < n; ++i) {
    Affine<C> p = IO::load(bases + IO::BYTES * i);
    if (p.inf) continue;
    Scalar4 s;
    scalars[i].to_std(s.v);
    if ((s.v[0] | s.v[1] | s.v[2] | s.v[3]) == 0) continue;
    if (s.v[0] == 1 && (s.v[1] | s.v[2] | s.v[3]) == 0) {
      ones.push_back((uint32_t)i);
      continue;
    }
    idx.push_back((uint32_t)i);
    big.push_back(s);
  }
  size_t m = idx.size();
  Jac<C> ones_sum = Jac<C>::infinity();
  for (uint32_t i : ones) ones_sum = ones_sum.add_affine(IO::load(bases + IO::BYTES * i));
  if (m == 0) return ones_sum;

  int c = force_c > 0 ? force_c : window_size(m);
  int W = (RECODE_BITS + c - 1) / c;
  size_t nb = (size_t)1 << (c - 1);
  std::vector<Jac<C>> sums(W);
  parallel_tasks((size_t)W, [&](size_t w) {
    std::vector<Jac<C>> buckets(nb, Jac<C>::infinity());
    for (size_t k = 0; k < m; ++k) {
      int64_t d = signed_digit(big[k].v, (int)w, c);
      if (d == 0) continue;
      Affine<C> p = IO::load(bases + IO::BYTES * idx[k]);
      if (d > 0)
        buckets[d - 1] = buckets[d - 1].add_affine(p);
      else
        buckets[-d - 1] = buckets[-d - 1].add_affine(p.neg());
    }
    // Running sum: sum_j (j+1) B_j.
    Jac<C> run = Jac<C>::infinity(), acc = Jac<C>::infinity();
    for (size_t j = nb; j-- > 0;) {
      run = run.add(buckets[j]);
      acc = acc.add(run);
    }
    sums[w] = acc;
  });
  Jac<C> total = sums[W - 1];
  for (int w = W - 2; w >= 0; --w) {
    for (int i = 0; i < c; ++i) total = total.dbl();
    total = total.add(sums[w]);
  }
  return total.add(ones_sum);
}

template <class C>
static Jac<C> msm_naive(const uint8_t* bases, const Fr* scalars, size_t n) {
  typedef PointIO<C> IO;
  int T = n_threads();
  std::vector<Jac<C>> part(T, Jac<C>::infinity());
  parallel_tasks((size_t)T, [&](size_t t) {
    for (size_t i = t; i < n; i += T) {
      Scalar4 s;
      scalars[i].to_std(s.v);
      part[t] = part[t].add(Jac<C>::from_affine(IO::load(bases + IO::BYTES * i)).mul(s.v));
    }
  });
  Jac<C> acc = Jac<C>::infinity();
  for (auto& p : part) acc = acc.add(p);
  return acc;
}

// ---------------------------------------------------------------------------------------
// Fixed-base batch multiplication k_i * G for the setup, 8-bit windows over a table of
// d * 2^(8w) * G, batch-normalised (one inversion per chunk).  Pinned against the plain
// double-and-add `mul` in test_oracle_curve.py.
// ---------------------------------------------------------------------------------------
template <class C>
struct FixedBase {
  std::vector<Affine<C>> table;  // [32][256]
  explicit FixedBase(const Affine<C>& g) : table(32 * 256) {
    Jac<C> base = Jac<C>::from_affine(g);
    for (int w = 0; w < 32; ++w) {
      Jac<C> acc = Jac<C>::infinity();
      for (int d = 0; d < 256; ++d) {
        table[w * 256 + d] = acc.to_affine();
        acc = acc.add(base);
      }
      for (int i = 0; i < 8; ++i) base = base.dbl();
    }
  }
  Jac<C> mul(const uint64_t k[4]) const {
    Jac<C> acc = Jac<C>::infinity();
    for (int w = 0; w < 32; ++w) {
      int d = (int)((k[w / 8] >> (8 * (w % 8))) & 0xff);
      if (d) acc = acc.add_affine(table[w * 256 + d]);
    }
    return acc;
  }
};

template <class F>
static void batch_invert(std::vector<F>& xs) {
  // Montgomery's trick; zeros stay zero.
  size_t n = xs.size();
  std::vector<F> pre(n);
  F acc = F::one();
  for (size_t i = 0; i < n; ++i) {
    pre[i] = acc;
    if (!xs[i].is_zero()) acc = acc * xs[i];
  }
  F inv = acc.inv();
  for (size_t i = n; i-- > 0;) {
    if (xs[i].is_zero()) continue;
    F t = inv * pre[i];
    inv = inv * xs[i];
    xs[i] = t;
  }
}

template <class C>
static void fixed_base_batch(const Affine<C>& g, const Fr* k_mont, size_t n, uint8_t* out) {
  typedef PointIO<C> IO;
  typedef typename C::F F;
  FixedBase<C> fb(g);
  parallel_for(n, [&](size_t lo, size_t hi) {
    std::vector<Jac<C>> pts(hi - lo);
    std::vector<F> zs(hi - lo);
    for (size_t i = lo; i < hi; ++i) {
      uint64_t s[4];
      k_mont[i].to_std(s);
      pts[i - lo] = fb.mul(s);
      zs[i - lo] = pts[i - lo].Z;
    }
    batch_invert(zs);
    for (size_t i = lo; i < hi; ++i) {
      const Jac<C>& p = pts[i - lo];
      if (p.is_inf()) {
        IO::store(out + IO::BYTES * i, Affine<C>::infinity());
        continue;
      }
      F zi = zs[i - lo], zi2 = zi.sqr();
      IO::store(out + IO::BYTES * i, Affine<C>{p.X * zi2, p.Y * zi2 * zi, false});
    }
  }, 256);
}

// ---------------------------------------------------------------------------------------
// SplitMix64 (Steele et al.; the test RNG of crates/gpu-layout/src/lib.rs:702-727) and its
// Fr sampler: 256 random bits, little-endian, reduced mod r.
// ---------------------------------------------------------------------------------------
struct SplitMix64 {
  uint64_t s;
  uint64_t next_u64() {
    s += 0x9E3779B97F4A7C15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  Fr next_fr() {
    uint64_t x[4];
    for (int i = 0; i < 4; ++i) x[i] = next_u64();
    while (geq4(x, FrParams::P)) sub4(x, x, FrParams::P);
    return Fr::from_std(x);
  }
};

// ---------------------------------------------------------------------------------------
// Synthetic circuit, witness and toxic-waste setup.  The construction (the workload of the
// plain-prove and 1.1M-constraint configurations) is specified step by step in DESIGN.md;
// snarkrs_amd/csrc/synth.cpp implements the same specification for the product side and
// test_synth.py holds the two equal.  Setup follows the snarkjs key semantics:
// crates/ceremony/src/setup.rs:879-1000 (public tail rows A[nC+s][s] = 1),
// crates/formats/src/lib.rs:35-75 (query meanings), crates/groth16/src/cpu.rs:230-245 (H
// bases = odd Lagrange basis of the 2n domain, divided by delta).
// ---------------------------------------------------------------------------------------
extern "C" {
typedef struct {
  uint32_t n_constraints;
  uint32_t n_public;        // public inputs (wires 1..=n_public)
  uint32_t n_private_in;    // private inputs after them
  uint32_t bool_permille;   // share of constraints whose output wire is a bit
  uint32_t input_bool_permille;
  uint32_t terms;           // nonzeros per A and B row
  uint64_t seed;            // circuit structure
  uint64_t setup_seed;      // toxic waste
  uint64_t witness_seed;    // witness values
} orc_synth_params;
}

struct Synth {
  uint32_t n_vars = 0, n_public = 0, domain = 0, n_constraints = 0;
  uint32_t n_in = 0, input_bool_permille = 0;
  std::vector<uint8_t> bit_row;
  std::vector<uint32_t> rp[3], sig[3];
  std::vector<Fr> val[3];
  std::vector<Fr> w;
  Fr tau, alpha, beta, gamma, delta;
  std::vector<Fr> u, v, ic_s, l_s, h_s;
};

static uint32_t floor_log2(uint64_t x) {
  uint32_t l = 0;
  while (x >>= 1) ++l;
  return l;
}

// Witness for seed `ws`: inputs (bits with probability input_bool_permille/1000, else
// uniform), then per constraint a drawn bit (bit rows) or the product of its A and B rows.
static void synth_witness(const Synth* S, uint64_t ws, std::vector<Fr>& w) {
  SplitMix64 rng{ws};
  w.assign(S->n_vars, Fr::zero());
  w[0] = Fr::one();
  for (uint32_t i = 1; i <= S->n_in; ++i) {
    if (rng.next_u64() % 1000 < S->input_bool_permille)
      w[i] = (rng.next_u64() & 1) ? Fr::one() : Fr::zero();
    else
      w[i] = rng.next_fr();
  }
  for (uint32_t j = 0; j < S->n_constraints; ++j) {
    uint32_t cur = 1 + S->n_in + j;
    if (S->bit_row[j]) {
      w[cur] = (rng.next_u64() & 1) ? Fr::one() : Fr::zero();
      continue;
    }
    Fr ab[2];
    for (int m = 0; m < 2; ++m) {
      Fr acc = Fr::zero();
      for (uint32_t e = S->rp[m][j]; e < S->rp[m][j + 1]; ++e) acc += S->val[m][e] * w[S->sig[m][e]];
      ab[m] = acc;
    }
    w[cur] = ab[0] * ab[1];
  }
}

static Synth* synth_build(const orc_synth_params* p) {
  Synth* S = new Synth();
  SplitMix64 rng{p->seed};
  uint32_t N = p->n_constraints, T = p->terms;
  uint32_t n_in = p->n_public + p->n_private_in;
  S->n_constraints = N;
  S->n_public = p->n_public;
  S->n_in = n_in;
  S->input_bool_permille = p->input_bool_permille;
  S->n_vars = 1 + n_in + N;
  S->domain = 1u << (floor_log2((uint64_t)N + p->n_public) + 1);
  uint32_t n = S->domain;
  for (int m = 0; m < 3; ++m) S->rp[m].assign(n + 1, 0);
  S->bit_row.assign(N, 0);
  uint32_t cur = 1 + n_in;
  for (uint32_t j = 0; j < N; ++j) {
    // Bit row: b * b = b on the fresh wire, A and B padded to T terms with k x - k x pairs
    // (and a zero term when T - 1 is odd). Product row: T seeded terms each in A and B.
    bool bit = rng.next_u64() % 1000 < p->bool_permille;
    S->bit_row[j] = bit;
    for (int m = 0; m < 2; ++m) {
      auto put = [&](uint32_t sg, const Fr& v) {
        S->sig[m].push_back(sg);
        S->val[m].push_back(v);
      };
      if (bit) {
        put(cur, Fr::one());
        uint32_t pads = T - 1;
        while (pads >= 2) {
          uint32_t x = (uint32_t)(rng.next_u64() % cur);
          Fr k = rng.next_fr();
          put(x, k);
          put(x, Fr::zero() - k);
          pads -= 2;
        }
        if (pads) put((uint32_t)(rng.next_u64() % cur), Fr::zero());
      } else {
        for (uint32_t t = 0; t < T; ++t) {
          uint32_t sg = (uint32_t)(rng.next_u64() % cur);
          put(sg, rng.next_fr());
        }
      }
      S->rp[m][j + 1] = (uint32_t)S->sig[m].size();
    }
    S->sig[2].push_back(cur);
    S->val[2].push_back(Fr::one());
    S->rp[2][j + 1] = (uint32_t)S->sig[2].size();
    ++cur;
  }
  // Public tail rows: A[N + s][s] = 1 for s = 0..=n_public (crates/ceremony/src/setup.rs:987-1001).
  for (uint32_t s = 0; s <= p->n_public; ++s) {
    S->sig[0].push_back(s);
    S->val[0].push_back(Fr::one());
    S->rp[0][N + s + 1] = (uint32_t)S->sig[0].size();
  }
  for (int m = 0; m < 3; ++m)
    for (uint32_t c = 1; c <= n; ++c)
      if (S->rp[m][c] < S->rp[m][c - 1]) S->rp[m][c] = S->rp[m][c - 1];
  synth_witness(S, p->witness_seed, S->w);

  // Toxic waste.
  SplitMix64 trng{p->setup_seed};
  S->tau = trng.next_fr();
  S->alpha = trng.next_fr();
  S->beta = trng.next_fr();
  S->gamma = trng.next_fr();
  S->delta = trng.next_fr();

  // Lagrange basis of the n domain at tau: L_c = w^c (tau^n - 1) / (n (tau - w^c)).
  int log_n = (int)floor_log2(n);
  Fr om = domain_root(log_n);
  Fr zt = S->tau.pow_u64(n) - Fr::one();
  Fr zn = zt * Fr::from_u64(n).inv();
  std::vector<Fr> L(n), den(n);
  Fr wc = Fr::one();
  for (uint32_t c = 0; c < n; ++c) {
    L[c] = wc;
    den[c] = S->tau - wc;
    wc *= om;
  }
  batch_invert(den);
  for (uint32_t c = 0; c < n; ++c) L[c] = L[c] * zn * den[c];

  std::vector<Fr> uu(S->n_vars, Fr::zero()), vv(S->n_vars, Fr::zero()), ww(S->n_vars, Fr::zero());
  std::vector<Fr>* tgt[3] = {&uu, &vv, &ww};
  for (int m = 0; m < 3; ++m)
    for (uint32_t c = 0; c < n; ++c)
      for (uint32_t e = S->rp[m][c]; e < S->rp[m][c + 1]; ++e)
        (*tgt[m])[S->sig[m][e]] += S->val[m][e] * L[c];
  S->u = uu;
  S->v = vv;
  Fr ginv = S->gamma.inv(), dinv = S->delta.inv();
  for (uint32_t i = 0; i < S->n_vars; ++i) {
    Fr t = S->beta * uu[i] + S->alpha * vv[i] + ww[i];
    if (i <= S->n_public)
      S->ic_s.push_back(t * ginv);
    else
      S->l_s.push_back(t * dinv);
  }
  // H bases: odd Lagrange basis of the 2n domain at tau, over delta.
  Fr om2 = domain_root(log_n + 1);
  Fr z2 = (S->tau.pow_u64(2ull * n) - Fr::one()) * Fr::from_u64(2ull * n).inv() * dinv;
  std::vector<Fr> hd(n);
  Fr x = om2, om2sq = om2.sqr();
  S->h_s.resize(n);
  for (uint32_t j = 0; j < n; ++j) {
    S->h_s[j] = x;
    hd[j] = S->tau - x;
    x *= om2sq;
  }
  batch_invert(hd);
  for (uint32_t j = 0; j < n; ++j) S->h_s[j] = S->h_s[j] * z2 * hd[j];
  return S;
}

// ---------------------------------------------------------------------------------------
// C ABI (test-only; called through oracle/oracle.py)
// ---------------------------------------------------------------------------------------
extern "C" {

void orc_set_threads(int t) { g_threads = t; }
int orc_threads(void) { return n_threads(); }
int orc_window_size(uint64_t n) { return window_size(n); }

int64_t orc_signed_digit(const uint64_t* s_std, int i, int c) { return signed_digit(s_std, i, c); }

// Field vector ops on Montgomery values. field: 0 = Fr, 1 = Fq, 2 = Fq2.
// op: 0 add, 1 sub, 2 mul, 3 sqr, 4 inv, 5 neg, 6 std->mont, 7 mont->std
int orc_field_op(int field, int op, const uint8_t* a, const uint8_t* b, uint8_t* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    if (field == 2) {
      Fq2 x = load_fq2(a + 64 * i), y = b ? load_fq2(b + 64 * i) : Fq2::zero(), r;
      switch (op) {
        case 0: r = x + y; break;
        case 1: r = x - y; break;
        case 2: r = x * y; break;
        case 3: r = x.sqr(); break;
        case 4: r = x.inv(); break;
        case 5: r = x.neg(); break;
        case 6: {
          uint64_t s0[4], s1[4];
          std::memcpy(s0, a + 64 * i, 32);
          std::memcpy(s1, a + 64 * i + 32, 32);
          r = {Fq::from_std(s0), Fq::from_std(s1)};
          break;
        }
        case 7: {
          uint64_t s0[4], s1[4];
          x.c0.to_std(s0);
          x.c1.to_std(s1);
          std::memcpy(out + 64 * i, s0, 32);
          std::memcpy(out + 64 * i + 32, s1, 32);
          continue;
        }
        default: return -1;
      }
      store_fq2(out + 64 * i, r);
      continue;
    }
    auto run = [&](auto tag) -> int {
      typedef decltype(tag) F;
      F x = load_f<F>(a + 32 * i), y = b ? load_f<F>(b + 32 * i) : F::zero(), r;
      switch (op) {
        case 0: r = x + y; break;
        case 1: r = x - y; break;
        case 2: r = x * y; break;
        case 3: r = x.sqr(); break;
        case 4: r = x.inv(); break;
        case 5: r = x.neg(); break;
        case 6: {
          uint64_t s[4];
          std::memcpy(s, a + 32 * i, 32);
          r = F::from_std(s);
          break;
        }
        case 7: {
          uint64_t s[4];
          x.to_std(s);
          std::memcpy(out + 32 * i, s, 32);
          return 0;
        }
        default: return -1;
      }
      store_f(out + 32 * i, r);
      return 0;
    };
    int rc = field == 0 ? run(Fr()) : run(Fq());
    if (rc) return rc;
  }
  return 0;
}

void orc_domain_root(int log_n, uint8_t* out) { store_f(out, domain_root(log_n)); }

int orc_ntt(uint8_t* data, int log_n, int inverse) {
  Fr* a = reinterpret_cast<Fr*>(data);
  if (inverse) ntt_inverse(a, log_n); else ntt_forward(a, log_n);
  return 0;
}

int orc_coset_pipeline(uint8_t* data, int log_n) {
  coset_pipeline(reinterpret_cast<Fr*>(data), log_n);
  return 0;
}

void orc_g1_generator(uint8_t* out) { store_g1(out, g1_generator()); }
void orc_g2_generator(uint8_t* out) { store_g2(out, g2_generator()); }

int orc_on_curve(int group, const uint8_t* pts, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    bool ok = group == 1 ? load_g1(pts + 64 * i).on_curve() : load_g2(pts + 128 * i).on_curve();
    if (!ok) return 0;
  }
  return 1;
}

// out[i] = k_i * P_i (k Montgomery), double-and-add.
int orc_scalar_mul(int group, const uint8_t* pts, const uint8_t* k, size_t n, uint8_t* out) {
  parallel_for(n, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      uint64_t s[4];
      load_f<Fr>(k + 32 * i).to_std(s);
      if (group == 1)
        store_g1(out + 64 * i, G1J::from_affine(load_g1(pts + 64 * i)).mul(s).to_affine());
      else
        store_g2(out + 128 * i, G2J::from_affine(load_g2(pts + 128 * i)).mul(s).to_affine());
    }
  }, 16);
  return 0;
}

// out[i] = P_i + Q_i.
int orc_point_add(int group, const uint8_t* p, const uint8_t* q, size_t n, uint8_t* out) {
  for (size_t i = 0; i < n; ++i) {
    if (group == 1)
      store_g1(out + 64 * i, G1J::from_affine(load_g1(p + 64 * i))
                                 .add(G1J::from_affine(load_g1(q + 64 * i))).to_affine());
    else
      store_g2(out + 128 * i, G2J::from_affine(load_g2(p + 128 * i))
                                  .add(G2J::from_affine(load_g2(q + 128 * i))).to_affine());
  }
  return 0;
}

// k_i * G for the standard generator of `group`, 8-bit fixed-base windows.
int orc_fixed_base(int group, const uint8_t* k, size_t n, uint8_t* out) {
  const Fr* kk = reinterpret_cast<const Fr*>(k);
  if (group == 1) fixed_base_batch<G1Curve>(g1_generator(), kk, n, out);
  else fixed_base_batch<G2Curve>(g2_generator(), kk, n, out);
  return 0;
}

// MSM; mode 0 = Pippenger (force_c 0 = reference window model), 1 = naive.
int orc_msm(int group, const uint8_t* bases, const uint8_t* scalars, size_t n, int mode,
            int force_c, uint8_t* out) {
  const Fr* s = reinterpret_cast<const Fr*>(scalars);
  if (group == 1) {
    G1J r = mode ? msm_naive<G1Curve>(bases, s, n) : msm_pippenger<G1Curve>(bases, s, n, force_c);
    store_g1(out, r.to_affine());
  } else {
    G2J r = mode ? msm_naive<G2Curve>(bases, s, n) : msm_pippenger<G2Curve>(bases, s, n, force_c);
    store_g2(out, r.to_affine());
  }
  return 0;
}

int orc_gather(const orc_key* k, int m, const uint8_t* w, uint8_t* out) {
  gather(k, m, reinterpret_cast<const Fr*>(w), reinterpret_cast<Fr*>(out));
  return 0;
}

int orc_compute_h(const orc_key* k, const uint8_t* w, uint8_t* h) {
  if (log2_exact(k->domain_size) < 0) return -1;
  compute_h(k, reinterpret_cast<const Fr*>(w), reinterpret_cast<Fr*>(h));
  return 0;
}

// Stages 0-11 with fixed blinders (crates/groth16/src/prove.rs:175-229, :391-430).
// msm_out: A G1 | B G2 | B G1 | L G1 | H G1 (affine, 64+128+64+64+64 bytes);
// proof_out: A G1 | B G2 | C G1 (256 bytes).
static int prove_tail(const orc_key* k, const Fr* w, std::vector<Fr>& h, const uint8_t* r_m,
                      const uint8_t* s_m, uint8_t* msm_out, uint8_t* proof_out);

int orc_prove(const orc_key* k, const uint8_t* wit, const uint8_t* r_m, const uint8_t* s_m,
              uint8_t* msm_out, uint8_t* proof_out) {
  if (log2_exact(k->domain_size) < 0 || k->h_basis != 0) return -1;
  const Fr* w = reinterpret_cast<const Fr*>(wit);
  std::vector<Fr> h(k->domain_size);
  compute_h(k, w, h.data());
  return prove_tail(k, w, h, r_m, s_m, msm_out, proof_out);
}

// prove_from_rows with fixed blinders (crates/groth16/src/cpu.rs:387-470,
// crates/groth16/src/prove.rs:93-115 without the check): stage 0 from the rows, H in the key's
// basis. -2: more rows than the domain; -3: gnark's coefficient n - 1 is not zero.
int orc_prove_rows(const orc_key* k, const uint8_t* wit, const uint8_t* a_rows,
                   const uint8_t* b_rows, size_t rows, const uint8_t* r_m, const uint8_t* s_m,
                   uint8_t* msm_out, uint8_t* proof_out) {
  int log_n = log2_exact(k->domain_size);
  if (log_n < 0 || k->h_basis > 1) return -1;
  if (rows > k->domain_size) return -2;
  const Fr* w = reinterpret_cast<const Fr*>(wit);
  std::vector<Fr> h(k->domain_size);
  rows_h(k->domain_size, reinterpret_cast<const Fr*>(a_rows), reinterpret_cast<const Fr*>(b_rows),
         rows, h.data());
  if (k->h_basis == 1 && !to_gnark_coefficients(h.data(), log_n)) return -3;
  return prove_tail(k, w, h, r_m, s_m, msm_out, proof_out);
}

// Stages 5-11 on a computed H (prove.rs:391-430).
static int prove_tail(const orc_key* k, const Fr* w, std::vector<Fr>& h, const uint8_t* r_m,
                      const uint8_t* s_m, uint8_t* msm_out, uint8_t* proof_out) {
  size_t n = k->domain_size;
  uint32_t nl = k->n_vars - k->n_public - 1;
  G1J a = msm_pippenger<G1Curve>(k->a_query, w, k->n_vars, 0);
  G2J b2 = msm_pippenger<G2Curve>(k->b_g2_query, w, k->n_vars, 0);
  G1J b1 = msm_pippenger<G1Curve>(k->b_g1_query, w, k->n_vars, 0);
  G1J l = msm_pippenger<G1Curve>(k->l_query, w + k->n_public + 1, nl, 0);
  G1J hh = G1J::infinity();
  if (k->h_basis == 1) {
    // The MSM over gnark's n - 1 bases, then times -1/2 = 1/t on the odd coset.
    hh = msm_pippenger<G1Curve>(k->h_query, h.data(), n - 1, 0);
    uint64_t half[4];
    (Fr::zero() - Fr::from_u64(2).inv()).to_std(half);
    hh = hh.mul(half);
  } else {
    hh = msm_pippenger<G1Curve>(k->h_query, h.data(), n, 0);
  }
  if (msm_out) {
    store_g1(msm_out, a.to_affine());
    store_g2(msm_out + 64, b2.to_affine());
    store_g1(msm_out + 192, b1.to_affine());
    store_g1(msm_out + 256, l.to_affine());
    store_g1(msm_out + 320, hh.to_affine());
  }
  Fr r = load_f<Fr>(r_m), s = load_f<Fr>(s_m);
  uint64_t rs[4], ss[4], rss[4];
  r.to_std(rs);
  s.to_std(ss);
  (r * s).to_std(rss);
  G1J d1 = G1J::from_affine(load_g1(k->delta_g1));
  G2J d2 = G2J::from_affine(load_g2(k->delta_g2));
  G1J pi_a = a.add_affine(load_g1(k->alpha_g1)).add(d1.mul(rs));
  G2J pi_b = b2.add_affine(load_g2(k->beta_g2)).add(d2.mul(ss));
  G1J pib1 = b1.add_affine(load_g1(k->beta_g1)).add(d1.mul(ss));
  G1J pi_c = l.add(hh).add(pi_a.mul(ss)).add(pib1.mul(rs)).add(d1.mul(rss).neg());
  store_g1(proof_out, pi_a.to_affine());
  store_g2(proof_out + 64, pi_b.to_affine());
  store_g1(proof_out + 192, pi_c.to_affine());
  return 0;
}

// Synthetic circuit handle.
void* orc_synth_new(const orc_synth_params* p) { return synth_build(p); }
void orc_synth_free(void* h) { delete static_cast<Synth*>(h); }

void orc_synth_witness(void* h, uint64_t ws, uint8_t* out) {
  std::vector<Fr> w;
  synth_witness(static_cast<Synth*>(h), ws, w);
  std::memcpy(out, w.data(), 32 * w.size());
}

// dims: n_vars, n_public, domain, n_constraints, nnz A, nnz B, nnz C
void orc_synth_dims(void* h, uint32_t* d) {
  Synth* S = static_cast<Synth*>(h);
  d[0] = S->n_vars;
  d[1] = S->n_public;
  d[2] = S->domain;
  d[3] = S->n_constraints;
  for (int m = 0; m < 3; ++m) d[4 + m] = (uint32_t)S->sig[m].size();
}

// what: 0..2 row_ptr[m], 3..5 signal[m], 6..8 values[m], 9 witness, 10 toxic waste
// (tau, alpha, beta, gamma, delta), 11 u, 12 v, 13 ic scalars, 14 l scalars, 15 h scalars.
int orc_synth_get(void* h, int what, void* dst) {
  Synth* S = static_cast<Synth*>(h);
  auto cp_u32 = [&](const std::vector<uint32_t>& x) { std::memcpy(dst, x.data(), 4 * x.size()); };
  auto cp_fr = [&](const std::vector<Fr>& x) { std::memcpy(dst, x.data(), 32 * x.size()); };
  if (what >= 0 && what < 3) cp_u32(S->rp[what]);
  else if (what < 6) cp_u32(S->sig[what - 3]);
  else if (what < 9) cp_fr(S->val[what - 6]);
  else if (what == 9) cp_fr(S->w);
  else if (what == 10) {
    Fr t[5] = {S->tau, S->alpha, S->beta, S->gamma, S->delta};
    std::memcpy(dst, t, sizeof t);
  } else if (what == 11) cp_fr(S->u);
  else if (what == 12) cp_fr(S->v);
  else if (what == 13) cp_fr(S->ic_s);
  else if (what == 14) cp_fr(S->l_s);
  else if (what == 15) cp_fr(S->h_s);
  else return -1;
  return 0;
}

// Closed-form "shadow" of a proof (test_prove.py): with the toxic waste known, every proof
// point is a known multiple of the generator, computed here in O(n + nnz) scalar
// arithmetic with no NTT and no MSM:
//   A(tau) = sum_c a_c L_c(tau), B likewise, ABt = sum_c a_c b_c L_c(tau)
//   msm_A = A(tau), msm_B = B(tau), msm_L = sum_{i>l} w_i (beta u_i + alpha v_i + w_i(tau))/delta
//   msm_H = (A(tau) B(tau) - ABt) / delta
// out (Fr Montgomery): msm_A, msm_B, msm_L, msm_H, pi_a, pi_b, pi_c, and the two sides of
// the verification equation  a*b  and  alpha*beta + sum_{i<=l} w_i(...) + delta*c.
int orc_shadow(void* h, const uint8_t* r_m, const uint8_t* s_m, uint8_t* out) {
  Synth* S = static_cast<Synth*>(h);
  uint32_t n = S->domain;
  int log_n = (int)floor_log2(n);
  Fr om = domain_root(log_n);
  Fr zn = (S->tau.pow_u64(n) - Fr::one()) * Fr::from_u64(n).inv();
  // Lagrange values recomputed one by one (no batch inversion), a different route from
  // the setup's.
  std::vector<Fr> L(n);
  parallel_for(n, [&](size_t lo, size_t hi) {
    Fr wc = om.pow_u64(lo);
    for (size_t c = lo; c < hi; ++c) {
      L[c] = wc * zn * (S->tau - wc).inv();
      wc *= om;
    }
  }, 256);
  std::vector<Fr> ev[3];
  for (int m = 0; m < 3; ++m) {
    ev[m].assign(n, Fr::zero());
    for (uint32_t c = 0; c < n; ++c)
      for (uint32_t e = S->rp[m][c]; e < S->rp[m][c + 1]; ++e)
        ev[m][c] += S->val[m][e] * S->w[S->sig[m][e]];
  }
  Fr At = Fr::zero(), Bt = Fr::zero(), ABt = Fr::zero(), Ct = Fr::zero();
  for (uint32_t c = 0; c < n; ++c) {
    At += ev[0][c] * L[c];
    Bt += ev[1][c] * L[c];
    ABt += ev[0][c] * ev[1][c] * L[c];
    Ct += ev[2][c] * L[c];
  }
  // Per-wire u_i, v_i, w_i(tau) for the L and IC sums.
  std::vector<Fr> pw[3];
  for (int m = 0; m < 3; ++m) {
    pw[m].assign(S->n_vars, Fr::zero());
    for (uint32_t c = 0; c < n; ++c)
      for (uint32_t e = S->rp[m][c]; e < S->rp[m][c + 1]; ++e)
        pw[m][S->sig[m][e]] += S->val[m][e] * L[c];
  }
  Fr dinv = S->delta.inv();
  Fr Lsum = Fr::zero(), Psum = Fr::zero();
  for (uint32_t i = 0; i < S->n_vars; ++i) {
    Fr t = (S->beta * pw[0][i] + S->alpha * pw[1][i] + pw[2][i]) * S->w[i];
    if (i <= S->n_public) Psum += t; else Lsum += t;
  }
  Fr r = load_f<Fr>(r_m), s = load_f<Fr>(s_m);
  Fr mA = At, mB = Bt, mL = Lsum * dinv, mH = (At * Bt - ABt) * dinv;
  Fr pa = S->alpha + mA + r * S->delta;
  Fr pb = S->beta + mB + s * S->delta;
  Fr pb1 = S->beta + mB + s * S->delta;
  Fr pc = mL + mH + s * pa + r * pb1 - r * s * S->delta;
  Fr lhs = pa * pb;
  Fr rhs = S->alpha * S->beta + Psum + S->delta * pc;
  (void)Ct;
  Fr res[9] = {mA, mB, mL, mH, pa, pb, pc, lhs, rhs};
  std::memcpy(out, res, sizeof res);
  return 0;
}

}  // extern "C"
