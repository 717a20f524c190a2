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
#include <type_traits>

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

// Inside a transform, values lie in [0, 2r) (4r < 2^256): a twiddle product skips its final
// subtraction (the 29-bit product returns below 2r for any 256-bit operands), a sum subtracts
// 2r once it reaches 2r, a difference adds 2r when negative. A pass that stores a transform's
// result reduces it below r, so every vector leaves the transforms fully reduced.
__device__ __forceinline__ uint32_t fr_2r(int i) {
  constexpr uint32_t R2[8] = {0xe0000002u, 0x87c3eb27u, 0xf372e122u, 0x5067d090u,
                              0x0302b0bau, 0x70a08b6du, 0xc2634053u, 0x60c89ce5u};
  return R2[i];
}
__device__ __forceinline__ Fr tmul(const Fr& x, const Fr& w_rp) {
  Fr r;
  mont29<Fr29, false>(x.v, w_rp.v, r.v);
  return r;
}
__device__ __forceinline__ Fr tadd(const Fr& a, const Fr& b) {
  Fr r, d;
  uint32_t carry = 0, borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = addc32(a.v[i], b.v[i], carry);
#pragma unroll
  for (int i = 0; i < 8; ++i) d.v[i] = subb32(r.v[i], fr_2r(i), borrow);
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = borrow ? r.v[i] : d.v[i];
  return r;
}
__device__ __forceinline__ Fr tsub(const Fr& a, const Fr& b) {
  Fr r;
  uint32_t borrow = 0, carry = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = subb32(a.v[i], b.v[i], borrow);
  const uint32_t mask = 0u - borrow;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = addc32(r.v[i], fr_2r(i) & mask, carry);
  return r;
}
__device__ __forceinline__ Fr tfull(Fr x) {  // [0, 2r) -> [0, r)
  reduce_once<FrCfg>(x.v);
  return x;
}

__device__ __forceinline__ uint32_t bitrev(uint32_t x, int L) {
  return L == 0 ? 0 : (__builtin_bitreverse32(x) >> (32 - L));
}

// One pass = B consecutive radix-2 stages s0 .. s0+B-1, done in LDS. For those stages the
// elements i = hi 2^(s0+B) + k 2^s0 + lo (k < 2^B) interact only among themselves; a
// workgroup takes one hi and TLO consecutive lo (2^B TLO = NTT_TILE elements), loads them
// with consecutive lanes on consecutive lo, runs the B stages and stores them back. The tile is
// 4096 elements (128 KB of LDS, one workgroup of 512 threads per CU on gfx950's 160 KB), so a
// 2^21 transform is two passes: the lowest 12 stages over contiguous tiles, then 9 stages over
// runs of 8 consecutive elements (256 bytes). The LDS holds the tile as two planes of 16-byte
// halves, so consecutive lanes read consecutive 16-byte words.
//
// Stages go two at a time (radix 4): a thread loads 4 elements, does the 2 x 2 butterflies of
// both stages in registers and stores them, one barrier per pair; an odd stage count ends (DIT)
// or starts (DIF) with a radix-2 step. The products are the same as radix 2's.
//
// The prover's iNTT -> coset shift -> NTT runs as three passes: the inverse's high stages, one
// fused pass that does the inverse's low stages, multiplies by the coset table and does the
// forward's low stages on the same tile in LDS, then the forward's high stages.
static constexpr int NTT_LOG_TILE = 12;
static constexpr int NTT_TILE = 1 << NTT_LOG_TILE;  // elements per workgroup
static constexpr int NTT_THREADS = 512;
static constexpr int NTT_HIGH_PASS = 10;  // stages per pass above the low pass: runs >= 4

struct NttLds {
  uint4 lo[NTT_TILE];  // words 0-3 of each element
  uint4 hi[NTT_TILE];  // words 4-7
};

__device__ __forceinline__ void lds_put(NttLds& t, int e, const Fr& x) {
  t.lo[e] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  t.hi[e] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}
__device__ __forceinline__ Fr lds_get(const NttLds& t, int e) {
  const uint4 a = t.lo[e], b = t.hi[e];
  return {{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}};
}

// Stage s0 + m's twiddle for tile-local k (bits above m ignored) and lo = lo0 + l.
__device__ __forceinline__ Fr twiddle(const uint4* tw, int s0, int m, uint32_t k, uint32_t lo) {
  const uint4* T = tw + 2 * (((size_t)1 << (s0 + m)) - 1);
  const uint32_t j = ((k & ((1u << m) - 1u)) << s0) + lo;
  return ldfr(T + 2 * (size_t)j);
}

// The B stages of one tile in LDS (element e = k tlo + l). DIT: stages ascending,
// (a, b) -> (a + b w, a - b w). DIF: descending, (a, b) -> (a + b, (a - b) w).
template <bool DIF>
__device__ __forceinline__ void tile_stages(NttLds& t, int lt, int s0, int B, uint32_t lo0,
                                            const uint4* tw) {
  const int tlo = 1 << lt, E = tlo << B;
  auto radix2 = [&](int m) {
    for (int u = threadIdx.x; u < E / 2; u += NTT_THREADS) {
      const int q = u >> lt, l = u & (tlo - 1);
      const int k = ((q >> m) << (m + 1)) | (q & ((1 << m) - 1));  // bit m of k clear
      const int e0 = (k << lt) + l, e1 = e0 + (tlo << m);
      const Fr w = twiddle(tw, s0, m, (uint32_t)k, lo0 + (uint32_t)l);
      const Fr a = lds_get(t, e0), c = lds_get(t, e1);
      if (DIF) {
        lds_put(t, e0, tadd(a, c));
        lds_put(t, e1, tmul(tsub(a, c), w));
      } else {
        const Fr d = tmul(c, w);
        lds_put(t, e0, tadd(a, d));
        lds_put(t, e1, tsub(a, d));
      }
    }
    __syncthreads();
  };
  // Stages m and m + 1 on e0..e3 = k0, k0 + 2^m, k0 + 2^(m+1), k0 + 2^m + 2^(m+1) (times tlo,
  // plus l): stage m pairs (e0, e1), (e2, e3) with one twiddle, stage m + 1 pairs (e0, e2) with
  // the same low bits and (e1, e3) with bit m set. FIRST (stages 0 and 1 of the transform:
  // s0 = m = 0, lo = 0): T_0[0] = T_1[0] = 1 and T_1[1] is the same fourth root for every
  // element, so one product remains of four.
  auto radix4 = [&](int m, auto first) {
    constexpr bool FIRST = decltype(first)::value;
    const Fr w4 = FIRST ? twiddle(tw, 0, 1, 1u, 0u) : Fr{};
    for (int u = threadIdx.x; u < E / 4; u += NTT_THREADS) {
      const int q = u >> lt, l = u & (tlo - 1);
      const int k0 = ((q >> m) << (m + 2)) | (q & ((1 << m) - 1));
      const int e0 = (k0 << lt) + l, e1 = e0 + (tlo << m), e2 = e0 + (tlo << (m + 1)),
                e3 = e2 + (tlo << m);
      const uint32_t lo = lo0 + (uint32_t)l;
      Fr x0 = lds_get(t, e0), x1 = lds_get(t, e1), x2 = lds_get(t, e2), x3 = lds_get(t, e3);
      if (FIRST) {
        if (DIF) {  // stage 1 (twiddles 1, w4), then stage 0 (twiddle 1)
          const Fr y0 = tadd(x0, x2), y2 = tsub(x0, x2);
          const Fr y1 = tadd(x1, x3), y3 = tmul(tsub(x1, x3), w4);
          x0 = tadd(y0, y1);
          x1 = tsub(y0, y1);
          x2 = tadd(y2, y3);
          x3 = tsub(y2, y3);
        } else {  // stage 0 (twiddle 1), then stage 1 (twiddles 1, w4)
          const Fr y0 = tadd(x0, x1), y1 = tsub(x0, x1), y2 = tadd(x2, x3);
          const Fr f3 = tmul(tsub(x2, x3), w4);
          x0 = tadd(y0, y2);
          x2 = tsub(y0, y2);
          x1 = tadd(y1, f3);
          x3 = tsub(y1, f3);
        }
      } else {
        const Fr wa = twiddle(tw, s0, m, (uint32_t)k0, lo);
        const Fr wb = twiddle(tw, s0, m + 1, (uint32_t)k0, lo);
        const Fr wc = twiddle(tw, s0, m + 1, (uint32_t)k0 | (1u << m), lo);
        if (DIF) {  // stage m + 1, then m
          Fr y0 = tadd(x0, x2), y2 = tmul(tsub(x0, x2), wb);
          Fr y1 = tadd(x1, x3), y3 = tmul(tsub(x1, x3), wc);
          x0 = tadd(y0, y1);
          x1 = tmul(tsub(y0, y1), wa);
          x2 = tadd(y2, y3);
          x3 = tmul(tsub(y2, y3), wa);
        } else {  // stage m, then m + 1
          const Fr d1 = tmul(x1, wa), d3 = tmul(x3, wa);
          Fr y0 = tadd(x0, d1), y1 = tsub(x0, d1), y2 = tadd(x2, d3), y3 = tsub(x2, d3);
          const Fr f2 = tmul(y2, wb), f3 = tmul(y3, wc);
          x0 = tadd(y0, f2);
          x2 = tsub(y0, f2);
          x1 = tadd(y1, f3);
          x3 = tsub(y1, f3);
        }
      }
      lds_put(t, e0, x0);
      lds_put(t, e1, x1);
      lds_put(t, e2, x2);
      lds_put(t, e3, x3);
    }
    __syncthreads();
  };
  using Gen = std::false_type;
  using First = std::true_type;
  if (DIF) {
    int m = B - 1;
    if (B & 1) radix2(m--);
    for (; m > 0; m -= 2) {
      if (m == 1 && s0 == 0) radix4(0, First{});
      else radix4(m - 1, Gen{});
    }
  } else {
    int m = 0;
    if (s0 == 0 && B >= 2) {
      radix4(0, First{});
      m = 2;
    }
    for (; m + 1 < B; m += 2) radix4(m, Gen{});
    if (m < B) radix2(m);
  }
}

// Which vectors a launch covers: grid row y is vector (y / g) 3 + y % g + o of the buffer, so
// (g, o) = (3, 0) is every vector, (2, 0) the A and B of each proof's (A, B, C) triple and
// (1, 2) the C. The last forward pass over C can write H = A B - C to h (vector y) instead of C,
// with A and B already final.
struct NttVecs {
  int g = 3, o = 0;
  uint4* h = nullptr;
};

// MODE 0: DIF stages (inverse twiddles tw). MODE 1: DIT stages (forward twiddles tw), each
// element first multiplied by prescale[idx] when given. MODE 2 (s0 = 0 only): DIF stages with
// tw, the prescale product, then DIT stages with tw2, on the same tile. MODE 3: DIT stages, then
// h[idx] = a[idx] b[idx] - x stored instead of x (H = A B - C, no division by Z:
// crates/groth16/src/cpu.rs:230-245, fused as crates/cuda/src/kernels/ntt.cu:268-275 does).
// Where a pass reads and what it stores: src (nullptr: the pass works in place on data) read at
// bit-reversed indices when brev is set (the first pass of a natural-order forward transform:
// the permutation folded into its loads, out of place), and full = the stored values are a
// transform's result (reduced below r).
struct NttIo {
  const uint4* src = nullptr;
  int brev = 0;
  int full = 0;
};

template <int MODE>
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_pass(uint4* data, int L, int s0, int B,
                                                          int lt, const uint4* tw,
                                                          const uint4* prescale,
                                                          const uint4* tw2, NttVecs vs,
                                                          NttIo io) {
  __shared__ NttLds t;
  const int tlo = 1 << lt, E = tlo << B;  // elements in this tile
  const uint32_t tiles_lo = 1u << (s0 - lt);
  const uint32_t hi = blockIdx.x >> (s0 - lt);
  const uint32_t lo0 = (blockIdx.x & (tiles_lo - 1)) << lt;
  const uint32_t y = blockIdx.y;
  const size_t voff = ((size_t)((y / vs.g) * 3 + y % vs.g + vs.o) << L) * 2;
  uint4* v = data + voff;
  const uint4* in = io.src ? io.src + voff : v;
  const size_t base = ((size_t)hi << (s0 + B)) + lo0;
  for (int e = threadIdx.x; e < E; e += NTT_THREADS) {
    const int k = e >> lt, l = e & (tlo - 1);
    const size_t idx = base + ((size_t)k << s0) + l;
    const size_t from = io.brev ? (size_t)bitrev((uint32_t)idx, L) : idx;
    Fr x = ldfr(in + 2 * from);
    if (MODE == 1 && prescale) x = tmul(x, ldfr(prescale + 2 * idx));
    lds_put(t, e, x);
  }
  __syncthreads();
  if (MODE == 0) tile_stages<true>(t, lt, s0, B, lo0, tw);
  if (MODE == 1 || MODE == 3) tile_stages<false>(t, lt, s0, B, lo0, tw);
  if (MODE == 2) {
    tile_stages<true>(t, lt, s0, B, lo0, tw);
    for (int e = threadIdx.x; e < E; e += NTT_THREADS) {
      const size_t idx = base + ((size_t)(e >> lt) << s0) + (e & (tlo - 1));
      lds_put(t, e, tmul(lds_get(t, e), ldfr(prescale + 2 * idx)));
    }
    __syncthreads();
    tile_stages<false>(t, lt, s0, B, lo0, tw2);
  }
  for (int e = threadIdx.x; e < E; e += NTT_THREADS) {
    const int k = e >> lt, l = e & (tlo - 1);
    const size_t idx = base + ((size_t)k << s0) + l;
    if (MODE == 3) {
      // A and B as their last passes left them (below 2r: the Fr product takes them and
      // returns below r); C reduced first, so H comes out below r.
      const uint4* a = data + (((size_t)(3 * y) << L) * 2);
      const uint4* b = a + (((size_t)1 << L) * 2);
      uint4* h = vs.h + (((size_t)y << L) * 2);
      stfr(h + 2 * idx, sub(mul(ldfr(a + 2 * idx), ldfr(b + 2 * idx)), tfull(lds_get(t, e))));
    } else {
      const Fr x = lds_get(t, e);
      stfr(v + 2 * idx, io.full ? tfull(x) : x);
    }
  }
}

template <int MODE>
static void launch_pass(void* data, int nvec, int L, int s0, int B, const void* tw,
                        const void* prescale, const void* tw2, hipStream_t s,
                        NttVecs vs = NttVecs{}, NttIo io = NttIo{}) {
  int lt = NTT_LOG_TILE - B;  // log2 of the consecutive lo per tile
  if (lt > s0) lt = s0;
  if (lt < 0) throw ArgError{"NTT pass of more stages than a tile holds"};
  const uint32_t tiles = 1u << (L - B - lt);
  hipLaunchKernelGGL(k_ntt_pass<MODE>, dim3(tiles, nvec), dim3(NTT_THREADS), 0, s,
                     static_cast<uint4*>(data), L, s0, B, lt, static_cast<const uint4*>(tw),
                     static_cast<const uint4*>(prescale), static_cast<const uint4*>(tw2), vs, io);
  G16_LAUNCH_CHECK("ntt pass");
}

// Pass plan: the lowest min(L, 12) stages in one pass over contiguous tiles; the stages above
// in passes of at most NTT_HIGH_PASS, as equal as they can be.
static int high_passes(int rest) { return (rest + NTT_HIGH_PASS - 1) / NTT_HIGH_PASS; }

static int low_stages(int L) { return L < NTT_LOG_TILE ? L : NTT_LOG_TILE; }

// Stages [low, L) in ascending passes: pass p covers [start(p), start(p + 1)).
static int high_start(int L, int p) {
  const int low = low_stages(L), rest = L - low, np = high_passes(rest);
  return low + (rest * p) / np;
}

void ntt_dif_inverse(void* data, int nvec, int L, const void* tw, hipStream_t s) {
  if (L == 0) return;
  const int np = high_passes(L - low_stages(L));
  for (int p = np - 1; p >= 0; --p) {
    const int a = high_start(L, p), b = high_start(L, p + 1);
    launch_pass<0>(data, nvec, L, a, b - a, tw, nullptr, nullptr, s);
  }
  launch_pass<0>(data, nvec, L, 0, low_stages(L), tw, nullptr, nullptr, s, NttVecs{},
                 NttIo{nullptr, 0, 1});
}

void ntt_dit_forward(void* data, int nvec, int L, const void* tw, const void* prescale,
                     hipStream_t s, const void* natural_src, void* tmp) {
  if (L == 0) {
    if (natural_src && natural_src != data)
      G16_CHECK(hipMemcpyAsync(data, natural_src, 32 * (size_t)nvec, hipMemcpyDeviceToDevice, s));
    if (prescale) launch_scale_rp(data, (size_t)nvec, prescale, s);  // n == 1: one factor
    return;
  }
  const int np = high_passes(L - low_stages(L));
  // With natural_src, the first pass gathers it at bit-reversed indices into tmp (np > 0) or
  // into data (np == 0), and the high passes go from tmp back to data.
  void* first = natural_src ? (np > 0 ? tmp : data) : data;
  launch_pass<1>(first, nvec, L, 0, low_stages(L), tw, prescale, nullptr, s, NttVecs{},
                 NttIo{static_cast<const uint4*>(natural_src), natural_src ? 1 : 0,
                       np == 0 ? 1 : 0});
  for (int p = 0; p < np; ++p) {
    const int a = high_start(L, p), b = high_start(L, p + 1);
    const bool from_tmp = natural_src && p == 0;
    launch_pass<1>(data, nvec, L, a, b - a, tw, nullptr, nullptr, s, NttVecs{},
                   NttIo{from_tmp ? static_cast<const uint4*>(tmp) : nullptr, 0,
                         p == np - 1 ? 1 : 0});
  }
}

void ntt_coset_pipeline(void* data, int nvec, int L, const void* tw_inv, const void* tw_fwd,
                        const void* coset, hipStream_t s, void* h_out) {
  const size_t n = (size_t)1 << L;
  if (h_out && nvec % 3) throw ArgError{"H from the transforms needs (A, B, C) triples"};
  uint4* v = static_cast<uint4*>(data);
  const int np = L == 0 ? 0 : high_passes(L - low_stages(L));
  if (L == 0) {
    launch_scale_rp(data, (size_t)nvec, coset, s);
  } else {
    for (int p = np - 1; p >= 0; --p) {
      const int a = high_start(L, p), b = high_start(L, p + 1);
      launch_pass<0>(data, nvec, L, a, b - a, tw_inv, nullptr, nullptr, s);
    }
    // Without high passes the fused pass is the last: its output is the result (and with
    // h_out the H below reads it fully reduced).
    launch_pass<2>(data, nvec, L, 0, low_stages(L), tw_inv, coset, tw_fwd, s, NttVecs{},
                   NttIo{nullptr, 0, np == 0 ? 1 : 0});
    for (int p = 0; p < np; ++p) {
      const int a = high_start(L, p), b = high_start(L, p + 1);
      if (h_out && p == np - 1) {
        // Every A and B first (left below 2r: only the H pass reads them), then the C passes
        // store H from the final A and B.
        const int m = nvec / 3;
        launch_pass<1>(data, 2 * m, L, a, b - a, tw_fwd, nullptr, nullptr, s, NttVecs{2, 0});
        launch_pass<3>(data, m, L, a, b - a, tw_fwd, nullptr, nullptr, s,
                       NttVecs{1, 2, static_cast<uint4*>(h_out)});
      } else {
        launch_pass<1>(data, nvec, L, a, b - a, tw_fwd, nullptr, nullptr, s, NttVecs{},
                       NttIo{nullptr, 0, p == np - 1 ? 1 : 0});
      }
    }
  }
  if (h_out && np == 0)
    for (int k = 0; k < nvec / 3; ++k)
      launch_h(static_cast<uint4*>(h_out) + 2 * n * k, v + 6 * n * k, v + 6 * n * k + 2 * n,
               v + 6 * n * k + 4 * n, n, s);
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
