// Synthetic workloads (include/g16synth.h): a seeded R1CS with a forward-computed witness,
// seeded toxic waste, and the query scalars of the snarkjs key semantics — public tail rows
// A[nC + s][s] = 1 (crates/ceremony/src/setup.rs:987-1001), u_i/v_i/w_i(tau) over the
// Lagrange basis of the domain, IC and L = (beta u + alpha v + w) / gamma or delta
// (crates/formats/src/lib.rs:35-75), H = the odd Lagrange basis of the 2n domain over delta
// (crates/groth16/src/cpu.rs:230-245). Host code on the bn254.cuh arithmetic.
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/g16synth.h"
#include "bn254.cuh"

using namespace g16;

namespace {

// SplitMix64 (Steele, Lea, Flood 2014).
struct Rng {
  uint64_t s;
  uint64_t u64() {
    s += 0x9E3779B97F4A7C15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  // 256 random bits (little-endian u64 words) reduced mod r, returned in Montgomery form.
  Fr fr() {
    Fr x;
    for (int i = 0; i < 4; ++i) {
      uint64_t w = u64();
      x.v[2 * i] = (uint32_t)w;
      x.v[2 * i + 1] = (uint32_t)(w >> 32);
    }
    for (int k = 0; k < 6; ++k) reduce_once<FrCfg>(x.v);  // 2^256 < 6r
    return from_std(x);
  }
};

Fr fr_u64(uint64_t x) {
  Fr r = fp_zero<FrCfg>();
  r.v[0] = (uint32_t)x;
  r.v[1] = (uint32_t)(x >> 32);
  return from_std(r);
}

Fr fr_pow(Fr b, uint64_t e) {
  Fr acc = fp_one<FrCfg>();
  for (; e; e >>= 1) {
    if (e & 1) acc = mul(acc, b);
    b = sqr(b);
  }
  return acc;
}

Fr root_of_unity(int L) {
  Fr w;
  const uint32_t W[8] = {0x725b19f0u, 0x9bd61b6eu, 0x41112ed4u, 0x402d111eu,
                         0x8ef62abcu, 0x00e0a7ebu, 0xa58a7e85u, 0x2a3c09f0u};
  std::memcpy(w.v, W, 32);
  w = from_std(w);
  for (int i = L; i < 28; ++i) w = sqr(w);
  return w;
}

unsigned n_workers() {
  unsigned t = std::thread::hardware_concurrency();
  return t ? (t > 16 ? 16 : t) : 1;
}

template <class Fn>
void parallel_chunks(size_t n, Fn fn) {
  unsigned T = n_workers();
  if (n < 4096) T = 1;
  std::vector<std::thread> pool;
  size_t step = (n + T - 1) / T;
  for (unsigned t = 0; t < T; ++t) {
    size_t lo = t * step, hi = lo + step < n ? lo + step : n;
    if (lo >= hi) break;
    pool.emplace_back(fn, lo, hi);
  }
  for (auto& th : pool) th.join();
}

// x[i] <- 1 / x[i] in place over [lo, hi) (Montgomery's trick; inputs nonzero).
void invert_range(Fr* x, size_t lo, size_t hi) {
  std::vector<Fr> pre(hi - lo);
  Fr acc = fp_one<FrCfg>();
  for (size_t i = lo; i < hi; ++i) {
    pre[i - lo] = acc;
    acc = mul(acc, x[i]);
  }
  Fr iv = inv(acc);
  for (size_t i = hi; i-- > lo;) {
    Fr t = mul(iv, pre[i - lo]);
    iv = mul(iv, x[i]);
    x[i] = t;
  }
}

}  // namespace

struct g16synth {
  uint32_t n_vars = 0, n_public = 0, domain = 0, n_constraints = 0;
  uint32_t n_in = 0, input_bool_permille = 0;
  std::vector<uint32_t> rp[3], sig[3];
  std::vector<Fr> val[3];
  std::vector<uint8_t> bit_row;  // constraint j's output wire is a bit (b * b = b)
  std::vector<Fr> w;
  Fr toxic[5];  // tau, alpha, beta, gamma, delta
  std::vector<Fr> u, v, ic, l, h;
};

// The witness for witness seed `ws`: inputs first (a bit with probability
// input_bool_permille / 1000, else uniform), then each constraint's output wire in order — a
// drawn bit on a bit row, else (A_j . w) * (B_j . w). Any seed satisfies the circuit.
static void fill_witness(const g16synth* S, uint64_t ws, std::vector<Fr>& w) {
  Rng rng{ws};
  const Fr one = fp_one<FrCfg>(), zero = fp_zero<FrCfg>();
  w.assign(S->n_vars, zero);
  w[0] = one;
  for (uint32_t i = 1; i <= S->n_in; ++i) {
    if (rng.u64() % 1000 < S->input_bool_permille)
      w[i] = (rng.u64() & 1) ? one : zero;
    else
      w[i] = rng.fr();
  }
  for (uint32_t j = 0; j < S->n_constraints; ++j) {
    uint32_t wire = 1 + S->n_in + j;
    if (S->bit_row[j]) {
      w[wire] = (rng.u64() & 1) ? one : zero;
      continue;
    }
    Fr lin[2];
    for (int m = 0; m < 2; ++m) {
      Fr acc = zero;
      for (uint32_t e = S->rp[m][j]; e < S->rp[m][j + 1]; ++e)
        acc = add(acc, mul(S->val[m][e], w[S->sig[m][e]]));
      lin[m] = acc;
    }
    w[wire] = mul(lin[0], lin[1]);
  }
}

static void build(g16synth* S, const g16synth_params* p) {
  Rng rng{p->seed};
  const uint32_t N = p->n_constraints, T = p->terms;
  const uint32_t n_in = p->n_public + p->n_private_in;
  S->n_constraints = N;
  S->n_public = p->n_public;
  S->n_in = n_in;
  S->input_bool_permille = p->input_bool_permille;
  S->n_vars = 1 + n_in + N;
  uint32_t lg = 0;
  for (uint64_t x = (uint64_t)N + p->n_public; x >>= 1;) ++lg;
  S->domain = 1u << (lg + 1);  // snarkjs: 2^(floor(log2(nC + nPub)) + 1)
  const uint32_t n = S->domain;
  const Fr one = fp_one<FrCfg>(), zero = fp_zero<FrCfg>();

  // Structure, one constraint (and one fresh wire) at a time. A product row has T seeded
  // terms over earlier wires in A and in B, and C = [wire]. A bit row is b * b = b on the
  // fresh wire b, padded to T terms per A and B row with pairs k x - k x on earlier wires
  // (and one zero term when T - 1 is odd), so every A and B row has T nonzeros and the
  // constraint holds whatever bit the witness picks.
  for (int m = 0; m < 3; ++m) {
    S->rp[m].assign((size_t)n + 1, 0);
    size_t cap = m < 2 ? (size_t)N * T : (size_t)N;
    S->sig[m].reserve(cap + (m == 0 ? p->n_public + 1 : 0));
    S->val[m].reserve(cap + (m == 0 ? p->n_public + 1 : 0));
  }
  S->bit_row.assign(N, 0);
  uint32_t wire = 1 + n_in;
  for (uint32_t j = 0; j < N; ++j) {
    bool bit = rng.u64() % 1000 < p->bool_permille;
    S->bit_row[j] = bit;
    for (int m = 0; m < 2; ++m) {
      if (bit) {
        S->sig[m].push_back(wire);
        S->val[m].push_back(one);
        uint32_t left = T - 1;
        for (; left >= 2; left -= 2) {
          uint32_t x = (uint32_t)(rng.u64() % wire);
          Fr k = rng.fr();
          S->sig[m].push_back(x);
          S->val[m].push_back(k);
          S->sig[m].push_back(x);
          S->val[m].push_back(neg(k));
        }
        if (left) {
          S->sig[m].push_back((uint32_t)(rng.u64() % wire));
          S->val[m].push_back(zero);
        }
      } else {
        for (uint32_t t = 0; t < T; ++t) {
          S->sig[m].push_back((uint32_t)(rng.u64() % wire));
          S->val[m].push_back(rng.fr());
        }
      }
      S->rp[m][j + 1] = (uint32_t)S->sig[m].size();
    }
    S->sig[2].push_back(wire);
    S->val[2].push_back(one);
    S->rp[2][j + 1] = (uint32_t)S->sig[2].size();
    ++wire;
  }
  for (uint32_t s = 0; s <= p->n_public; ++s) {
    S->sig[0].push_back(s);
    S->val[0].push_back(one);
    S->rp[0][(size_t)N + s + 1] = (uint32_t)S->sig[0].size();
  }
  for (int m = 0; m < 3; ++m)
    for (size_t c = 1; c <= n; ++c)
      if (S->rp[m][c] < S->rp[m][c - 1]) S->rp[m][c] = S->rp[m][c - 1];
  fill_witness(S, p->witness_seed, S->w);

  Rng trng{p->setup_seed};
  for (auto& x : S->toxic) x = trng.fr();
  const Fr tau = S->toxic[0], alpha = S->toxic[1], beta = S->toxic[2];
  const Fr gamma_inv = inv(S->toxic[3]), delta_inv = inv(S->toxic[4]);

  // Lagrange basis of the n domain at tau: w^c (tau^n - 1) / (n (tau - w^c)).
  const Fr om = root_of_unity((int)lg + 1);
  const Fr zfac = mul(sub(fr_pow(tau, n), one), inv(fr_u64(n)));
  std::vector<Fr> L(n), den(n);
  parallel_chunks(n, [&](size_t lo, size_t hi) {
    Fr x = fr_pow(om, lo);
    for (size_t c = lo; c < hi; ++c) {
      L[c] = x;
      den[c] = sub(tau, x);
      x = mul(x, om);
    }
    invert_range(den.data(), lo, hi);
    for (size_t c = lo; c < hi; ++c) L[c] = mul(mul(L[c], zfac), den[c]);
  });

  // Per-wire u, v, w(tau): sum over each matrix's nonzeros of coefficient * L_row.
  std::vector<Fr> acc[3];
  for (int m = 0; m < 3; ++m) {
    acc[m].assign(S->n_vars, zero);
    for (size_t c = 0; c < n; ++c)
      for (uint32_t e = S->rp[m][c]; e < S->rp[m][c + 1]; ++e)
        acc[m][S->sig[m][e]] = add(acc[m][S->sig[m][e]], mul(S->val[m][e], L[c]));
  }
  S->u = acc[0];
  S->v = acc[1];
  S->ic.resize((size_t)p->n_public + 1);
  S->l.resize(S->n_vars - S->ic.size());
  parallel_chunks(S->n_vars, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      Fr t = add(add(mul(beta, acc[0][i]), mul(alpha, acc[1][i])), acc[2][i]);
      if (i <= p->n_public) S->ic[i] = mul(t, gamma_inv);
      else S->l[i - S->ic.size()] = mul(t, delta_inv);
    }
  });

  // H bases: odd Lagrange basis of the 2n domain at tau, over delta.
  const Fr om2 = root_of_unity((int)lg + 2);
  const Fr om2sq = sqr(om2);
  const Fr hfac = mul(mul(sub(fr_pow(tau, 2ull * n), one), inv(fr_u64(2ull * n))), delta_inv);
  S->h.resize(n);
  std::vector<Fr> hd(n);
  parallel_chunks(n, [&](size_t lo, size_t hi) {
    Fr x = mul(om2, fr_pow(om2sq, lo));
    for (size_t j = lo; j < hi; ++j) {
      S->h[j] = x;
      hd[j] = sub(tau, x);
      x = mul(x, om2sq);
    }
    invert_range(hd.data(), lo, hi);
    for (size_t j = lo; j < hi; ++j) S->h[j] = mul(mul(S->h[j], hfac), hd[j]);
  });
}

extern "C" {

g16synth* g16synth_new(const g16synth_params* p) {
  if (!p || p->terms == 0) return nullptr;
  if ((uint64_t)p->n_constraints + p->n_public >= (1ull << 27)) return nullptr;
  auto* s = new g16synth();
  build(s, p);
  return s;
}

void g16synth_free(g16synth* s) { delete s; }

void g16synth_dims(const g16synth* s, uint32_t* d) {
  d[0] = s->n_vars;
  d[1] = s->n_public;
  d[2] = s->domain;
  d[3] = s->n_constraints;
  for (int m = 0; m < 3; ++m) d[4 + m] = (uint32_t)s->sig[m].size();
}

int g16synth_get(const g16synth* s, int what, void* dst) {
  auto u32 = [&](const std::vector<uint32_t>& x) { std::memcpy(dst, x.data(), 4 * x.size()); };
  auto fr = [&](const std::vector<Fr>& x) { std::memcpy(dst, x.data(), 32 * x.size()); };
  switch (what) {
    case 0: case 1: case 2: u32(s->rp[what]); return 0;
    case 3: case 4: case 5: u32(s->sig[what - 3]); return 0;
    case 6: case 7: case 8: fr(s->val[what - 6]); return 0;
    case 9: fr(s->w); return 0;
    case 10: std::memcpy(dst, s->toxic, sizeof s->toxic); return 0;
    case 11: fr(s->u); return 0;
    case 12: fr(s->v); return 0;
    case 13: fr(s->ic); return 0;
    case 14: fr(s->l); return 0;
    case 15: fr(s->h); return 0;
    default: return -1;
  }
}

int g16synth_witness(const g16synth* s, uint64_t witness_seed, uint8_t* out) {
  if (!s || !out) return -1;
  std::vector<Fr> w;
  fill_witness(s, witness_seed, w);
  std::memcpy(out, w.data(), 32 * w.size());
  return 0;
}

void g16synth_scalars(uint64_t seed, size_t n, uint32_t zero_one_permille, uint8_t* out) {
  Rng rng{seed};
  const Fr one = fp_one<FrCfg>(), zero = fp_zero<FrCfg>();
  for (size_t i = 0; i < n; ++i) {
    Fr x;
    if (rng.u64() % 1000 < zero_one_permille)
      x = (rng.u64() & 1) ? one : zero;
    else
      x = rng.fr();
    std::memcpy(out + 32 * i, x.v, 32);
  }
}

}  // extern "C"
