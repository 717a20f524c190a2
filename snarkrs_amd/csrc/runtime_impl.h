// Host-side types shared by runtime.hip (devices, standalone entry points) and prove.hip
// (prepared circuits, proofs, batches, the server).
#pragma once
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/g16hip.h"
#include "bn254.cuh"
#include "host64.h"
#include "internal.h"

namespace g16 {

// ---- errors ---------------------------------------------------------------------------
struct StatusError {
  g16hip_status st;
  std::string msg;
};

inline g16hip_status fail(g16hip_status st, const std::string& msg) {
  set_last_error(msg);
  return st;
}

template <class F>
g16hip_status guarded(F&& f) {
  try {
    set_last_error("");
    return f();
  } catch (const HipError& e) {
    // Sizes are checked against the device budget before anything is allocated (a refusal
    // there is a StatusError); an allocation the driver still refuses afterwards (memory
    // taken by another process in between) is reported as a memory refusal too.
    return fail(e.code == hipErrorOutOfMemory ? G16HIP_DEVICE_MEMORY : G16HIP_DEVICE, e.msg);
  } catch (const ArgError& e) {
    return fail(G16HIP_BACKEND, e.msg);
  } catch (const StatusError& e) {
    return fail(e.st, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(G16HIP_BACKEND, "host allocation failed");
  }
}

// ---- device memory ---------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    if (n <= bytes && p) return;
    release();
    if (n == 0) return;
    G16_CHECK(hipMalloc(&p, n));
    bytes = n;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

inline void upload(DevBuf& b, const void* src, size_t n, hipStream_t s) {
  b.alloc(n);
  if (n) G16_CHECK(hipMemcpyAsync(b.p, src, n, hipMemcpyHostToDevice, s));
}

// ---- host-side arithmetic helpers (bn254.cuh compiled for the CPU) ----------------------
inline Fr fr_load(const uint8_t* p) {
  Fr x;
  std::memcpy(x.v, p, 32);
  return x;
}
inline G1Aff g1_load(const uint8_t* p) {
  G1Aff a;
  std::memcpy(a.x.v, p, 32);
  std::memcpy(a.y.v, p + 32, 32);
  return a;
}
inline G2Aff g2_load(const uint8_t* p) {
  G2Aff a;
  std::memcpy(a.x.c0.v, p, 32);
  std::memcpy(a.x.c1.v, p + 32, 32);
  std::memcpy(a.y.c0.v, p + 64, 32);
  std::memcpy(a.y.c1.v, p + 96, 32);
  return a;
}
inline void g1_store(uint8_t* p, const G1Aff& a) {
  std::memcpy(p, a.x.v, 32);
  std::memcpy(p + 32, a.y.v, 32);
}
inline void g2_store(uint8_t* p, const G2Aff& a) {
  std::memcpy(p, a.x.c0.v, 32);
  std::memcpy(p + 32, a.x.c1.v, 32);
  std::memcpy(p + 64, a.y.c0.v, 32);
  std::memcpy(p + 96, a.y.c1.v, 32);
}

inline bool fq_canonical(const Fq& x) {
  for (int i = 7; i >= 0; --i) {
    if (x.v[i] != FqCfg::p(i)) return x.v[i] < FqCfg::p(i);
  }
  return false;
}

// Domain root of order 2^L: 5^((r-1)/2^28) squared 28 - L times (crates/field/src/lib.rs:48-82).
inline Fr domain_root(int L) {
  Fr w;
  const uint32_t W[8] = {0x725b19f0u, 0x9bd61b6eu, 0x41112ed4u, 0x402d111eu,
                         0x8ef62abcu, 0x00e0a7ebu, 0xa58a7e85u, 0x2a3c09f0u};
  std::memcpy(w.v, W, 32);
  w = from_std(w);
  for (int i = L; i < 28; ++i) w = sqr(w);
  return w;
}

inline Fr fr_from_u64(uint64_t x) {
  Fr r = fp_zero<FrCfg>();
  r.v[0] = (uint32_t)x;
  r.v[1] = (uint32_t)(x >> 32);
  return from_std(r);
}

// The host's finish of a windowed MSM: a few hundred point operations over the reduction's
// rows, on the 64-bit-limb host field code (the 32-bit-limb field code the kernels share
// runs at a third of the speed on the host). Both keep field elements fully reduced in the
// R = 2^256 Montgomery form, so a value means the same bytes in either.
namespace hx {
using h64::Fq;
using h64::Fq2;
inline Fq mul(const Fq& a, const Fq& b) { return h64::f_mul(a, b); }
inline Fq sqr(const Fq& a) { return h64::f_mul(a, a); }
inline Fq add(const Fq& a, const Fq& b) { return h64::f_add(a, b); }
inline Fq sub(const Fq& a, const Fq& b) { return h64::f_sub(a, b); }
inline bool is_zero(const Fq& a) { return h64::f_is_zero(a); }
inline Fq2 mul(const Fq2& a, const Fq2& b) { return h64::e2_mul(a, b); }
inline Fq2 sqr(const Fq2& a) { return h64::e2_sqr(a); }
inline Fq2 add(const Fq2& a, const Fq2& b) { return h64::e2_add(a, b); }
inline Fq2 sub(const Fq2& a, const Fq2& b) { return h64::e2_sub(a, b); }
inline bool is_zero(const Fq2& a) { return h64::e2_is_zero(a); }
template <class F>
F dbl(const F& a) {
  return add(a, a);
}

template <class F>
struct Pt {  // XYZZ, ZZ = 0 at infinity
  F X, Y, ZZ, ZZZ;
};

// dbl-2008-s-1 (a = 0), as xyzz_dbl.
template <class F>
Pt<F> pt_dbl(const Pt<F>& p) {
  if (is_zero(p.ZZ)) return p;
  F U = dbl(p.Y), V = sqr(U), W = mul(U, V), S = mul(p.X, V);
  F X2 = sqr(p.X), M = add(dbl(X2), X2);
  Pt<F> r;
  r.ZZ = mul(V, p.ZZ);
  r.ZZZ = mul(W, p.ZZZ);
  r.X = sub(sqr(M), dbl(S));
  r.Y = sub(mul(M, sub(S, r.X)), mul(W, p.Y));
  return r;
}

// add-2008-s with the doubling and cancellation cases, as xyzz_add.
template <class F>
Pt<F> pt_add(const Pt<F>& p, const Pt<F>& q) {
  if (is_zero(p.ZZ)) return q;
  if (is_zero(q.ZZ)) return p;
  F U1 = mul(p.X, q.ZZ), S1 = mul(p.Y, q.ZZZ);
  F P = sub(mul(q.X, p.ZZ), U1), R = sub(mul(q.Y, p.ZZZ), S1);
  F ZZ12 = mul(p.ZZ, q.ZZ), ZZZ12 = mul(p.ZZZ, q.ZZZ);
  if (is_zero(P)) {
    if (is_zero(R)) return pt_dbl(Pt<F>{U1, S1, ZZ12, ZZZ12});
    return Pt<F>{};
  }
  F PP = sqr(P), PPP = mul(P, PP), Q = mul(U1, PP);
  Pt<F> r;
  r.ZZ = mul(ZZ12, PP);
  r.ZZZ = mul(ZZZ12, PPP);
  r.X = sub(sub(sqr(R), PPP), dbl(Q));
  r.Y = sub(mul(R, sub(Q, r.X)), mul(S1, PPP));
  return r;
}

// Rows come off the MSM kernels in the R' = 2^261 Montgomery form, below 2q: x_R = x_R' / 32,
// one R-form product by 2^251 (= 2^256 / 32) after a conditional subtraction of q.
inline Fq from_msm_form(Fq x) {
  if (!h64::lt_p<h64::QTag>(x.v)) h64::sub_p<h64::QTag>(x.v);
  return h64::f_mul(x, Fq{{0, 0, 0, 0x0800000000000000ull}});
}
inline Fq2 from_msm_form(const Fq2& x) {
  return {from_msm_form(x.c0), from_msm_form(x.c1)};
}

template <class F> struct Of;
template <> struct Of<g16::Fq> { typedef Fq T; };
template <> struct Of<g16::Fq2> { typedef Fq2 T; };

// a^(q - 2) (Fermat); 0 maps to 0.
inline Fq inv(const Fq& a) {
  const uint64_t* P = h64::QTag::P;
  const uint64_t E[4] = {P[0] - 2, P[1], P[2], P[3]};
  Fq acc = h64::f_one<h64::QTag>();
  for (int i = 255; i >= 0; --i) {
    acc = sqr(acc);
    if ((E[i >> 6] >> (i & 63)) & 1) acc = mul(acc, a);
  }
  return acc;
}
// 1 / (a0 + a1 u) = (a0 - a1 u) / (a0^2 + a1^2).
inline Fq2 inv(const Fq2& a) {
  const Fq t = inv(add(sqr(a.c0), sqr(a.c1)));
  return {mul(a.c0, t), h64::f_neg(mul(a.c1, t))};
}

template <class G>
Pt<typename Of<G>::T> in(const Xyzz<G>& p) {
  typedef typename Of<G>::T H;
  static_assert(sizeof(H) == sizeof(G), "host field layouts");
  Pt<H> q;
  std::memcpy(&q.X, &p.X, sizeof(H));
  std::memcpy(&q.Y, &p.Y, sizeof(H));
  std::memcpy(&q.ZZ, &p.ZZ, sizeof(H));
  std::memcpy(&q.ZZZ, &p.ZZZ, sizeof(H));
  return q;
}
template <class G>
Xyzz<G> out(const Pt<typename Of<G>::T>& q) {
  typedef typename Of<G>::T H;
  Xyzz<G> p;
  std::memcpy(&p.X, &q.X, sizeof(H));
  std::memcpy(&p.Y, &q.Y, sizeof(H));
  std::memcpy(&p.ZZ, &q.ZZ, sizeof(H));
  std::memcpy(&p.ZZZ, &q.ZZZ, sizeof(H));
  return p;
}
}  // namespace hx

// Point operations for the host's part of an MSM or a proof (finishing sums, blinding
// products, stage 11, affine output), on the 64-bit field code and the kernels' types.
template <class G>
Xyzz<G> host_add(const Xyzz<G>& a, const Xyzz<G>& b) {
  return hx::out<G>(hx::pt_add(hx::in(a), hx::in(b)));
}

// k * p for a plain (standard-form) 256-bit scalar, double-and-add from the top bit.
template <class G>
Xyzz<G> host_mul(const Xyzz<G>& p, const uint32_t k[8]) {
  const auto q = hx::in(p);
  decltype(hx::in(p)) acc{};
  for (int i = 255; i >= 0; --i) {
    acc = hx::pt_dbl(acc);
    if ((k[i >> 5] >> (i & 31)) & 1) acc = hx::pt_add(acc, q);
  }
  return hx::out<G>(acc);
}

// x = X / ZZ, y = Y / ZZZ with one inversion: ZZ / ZZZ = 1 / Z, squared 1 / ZZ.
template <class G>
Aff<G> host_to_aff(const Xyzz<G>& p) {
  typedef typename hx::Of<G>::T H;
  const auto q = hx::in(p);
  Aff<G> a;
  if (hx::is_zero(q.ZZ)) {
    std::memset(&a, 0, sizeof(a));
    return a;
  }
  const H izzz = hx::inv(q.ZZZ);
  const H iz2 = hx::sqr(hx::mul(izzz, q.ZZ));
  const H x = hx::mul(q.X, iz2), y = hx::mul(q.Y, izzz);
  std::memcpy(&a.x, &x, sizeof(H));
  std::memcpy(&a.y, &y, sizeof(H));
  return a;
}

// The MSM from the reduction's rows (msm_accumulate, in the kernels' R' form): sum over
// windows w and rows j of 2^e P[w][j], e = c w for row 0 and c w + lg + j - 1 for row
// j >= 1 — one Horner pass over the exponents, one doubling each (the reference finishes
// sum_w 2^(cw) S_w on the host the same way, crates/cuda/src/msm.rs:3976-4010).
template <class F>
Xyzz<F> msm_combine(const Xyzz<F>* P, const MsmPlan& p) {
  typedef typename hx::Of<F>::T H;
  static_assert(sizeof(H) == sizeof(F), "host field layouts");
  int rows, lg;
  msm_reduce_shape(p, &rows, &lg);
  const int c = p.c, W = msm_bucket_sets(p);  // one set when the windows are merged
  int emax = c * (W - 1) + (rows > 1 ? lg + rows - 2 : 0);
  std::vector<hx::Pt<H>> E((size_t)emax + 1, hx::Pt<H>{});
  for (int w = 0; w < W; ++w)
    for (int j = 0; j < rows; ++j) {
      int e = c * w + (j == 0 ? 0 : lg + j - 1);
      hx::Pt<H> q;
      const Xyzz<F>& s = P[(size_t)w * rows + j];
      std::memcpy(&q.X, &s.X, sizeof(H));
      std::memcpy(&q.Y, &s.Y, sizeof(H));
      std::memcpy(&q.ZZ, &s.ZZ, sizeof(H));
      std::memcpy(&q.ZZZ, &s.ZZZ, sizeof(H));
      q = {hx::from_msm_form(q.X), hx::from_msm_form(q.Y), hx::from_msm_form(q.ZZ),
           hx::from_msm_form(q.ZZZ)};
      E[e] = hx::pt_add(E[e], q);
    }
  hx::Pt<H> acc = E[emax];
  for (int e = emax - 1; e >= 0; --e) acc = hx::pt_add(hx::pt_dbl(acc), E[e]);
  Xyzz<F> out;
  std::memcpy(&out.X, &acc.X, sizeof(H));
  std::memcpy(&out.Y, &acc.Y, sizeof(H));
  std::memcpy(&out.ZZ, &acc.ZZ, sizeof(H));
  std::memcpy(&out.ZZZ, &acc.ZZZ, sizeof(H));
  return out;
}

inline float ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  G16_CHECK(hipEventElapsedTime(&ms, a, b));
  return ms;
}

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  uint64_t us() const {
    return (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::steady_clock::now() - t0).count();
  }
  double ms() const { return us() / 1000.0; }
};

// Runs fn on a pooled host thread. Threads are kept and reused; a task never waits for a busy
// one (the pool starts another thread when none is idle), so tasks may wait for each other.
void host_pool_run(std::function<void()> fn);

// Runs fn() on a pooled thread and keeps its exception for the joiner.
struct HostTask {
  std::exception_ptr err;
  template <class Fn>
  explicit HostTask(Fn fn) : st_(std::make_shared<State>()) {
    std::shared_ptr<State> st = st_;
    host_pool_run([st, fn]() mutable {
      try {
        fn();
      } catch (...) {
        st->err = std::current_exception();
      }
      std::lock_guard<std::mutex> lk(st->m);
      st->done = true;
      st->cv.notify_all();
    });
  }
  HostTask(const HostTask&) = delete;
  HostTask& operator=(const HostTask&) = delete;
  ~HostTask() { join(); }
  void join() {
    std::unique_lock<std::mutex> lk(st_->m);
    st_->cv.wait(lk, [&] { return st_->done; });
    err = st_->err;
  }

 private:
  struct State {
    std::mutex m;
    std::condition_variable cv;
    bool done = false;
    std::exception_ptr err;
  };
  std::shared_ptr<State> st_;
};

// ---- devices ---------------------------------------------------------------------------
struct NttTables {
  // fwd, inv: per-pass twiddle tables (2^L - 1 entries); coset: shift^bitrev(p) / n; both in
  // the R' = 2^261 form. scalars (R form): [w, w^-1, shift, 1/n, 1, 0, shift^-1, -1/(2n)].
  // gnark (built for gnark-basis keys only): -shift^-bitrev(p) / (2n), R' form.
  DevBuf fwd, inv, coset, scalars, gnark;
};
void build_ntt_tables(NttTables& t, int L, hipStream_t s);
void build_gnark_h_table(NttTables& t, int L, hipStream_t s);

}  // namespace g16

struct g16hip_device {
  int ordinal = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // serialises the standalone entry points' scratch
  g16::DevBuf scratch_a, scratch_b, scratch_c, scratch_d, scratch_e;
  g16::DevBuf msm_bases;  // the standalone MSM's bases in the MSM kernels' R' form
  std::map<int, std::unique_ptr<g16::NttTables>> ntt;
  hipEvent_t ev[6] = {};
  g16hip_msm_stats stats = {};
  std::mutex stats_mu;
  g16hip_kernel_stats kstats = {};
  bool constant_work = false;  // MsmPlan::constant for standalone MSMs and new circuits
  // Cap on the device memory a circuit prepared afterwards (its key and first slot) and a
  // further slot may plan against; 0 = none (g16hip_set_vram_limit, G16HIP_VRAM_LIMIT).
  std::atomic<size_t> vram_limit{0};
  size_t vram_reserve = 0;  // bytes left to other users (G16HIP_VRAM_RESERVE)
};

namespace g16 {

// Bytes an allocation may take now: free device memory less a headroom for the runtime and
// the reserve, capped by the device's limit (the reference's device_budget,
// crates/cuda/src/msm.rs:1057-1091).
constexpr size_t VRAM_HEADROOM = 256ull << 20;
inline size_t device_budget(const g16hip_device* d) {
  size_t free_b = 0, total_b = 0;
  G16_CHECK(hipMemGetInfo(&free_b, &total_b));
  size_t budget = free_b > VRAM_HEADROOM + d->vram_reserve ? free_b - VRAM_HEADROOM - d->vram_reserve : 0;
  const size_t limit = d->vram_limit.load();
  return limit && limit < budget ? limit : budget;
}

// Device allocations are rounded up to the pages the driver maps them in when planned.
inline size_t dev_pages(size_t bytes) {
  constexpr size_t PAGE = 2ull << 20;
  return (bytes + PAGE - 1) / PAGE * PAGE;
}

inline StatusError too_big(size_t need, size_t budget, const std::string& what) {
  return StatusError{G16HIP_DEVICE_MEMORY, what + " needs " + std::to_string(need) +
                                               " bytes of device memory and " +
                                               std::to_string(budget) + " are available to it"};
}

inline void note_accumulate(g16hip_device* d, int group, uint64_t entries, double ms) {
  std::lock_guard<std::mutex> lk(d->stats_mu);
  d->kstats.launches[group - 1] += 1;
  d->kstats.entries[group - 1] += entries;
  d->kstats.accumulate_ms[group - 1] += ms;
}

inline void set_device(const g16hip_device* d) { G16_CHECK(hipSetDevice(d->ordinal)); }

// Device scratch of one standalone MSM: the sort's, the accumulation's (window sums behind
// it) and the bases in the MSM kernels' form.
struct MsmScratch {
  void* sort = nullptr;
  void* accum = nullptr;
  void* bases = nullptr;
};

// Bytes of each piece for n points at width c, each rounded up to 256 bytes.
struct MsmScratchSizes {
  size_t sort = 0, accum = 0, bases = 0;
  size_t total() const { return sort + accum + bases; }
};
MsmScratchSizes msm_scratch_sizes(int group, uint32_t n, int c, bool constant);
// c, or the library's width for n when c <= 0; throws outside [2, 16].
int msm_width(uint32_t n, int c);
// One device MSM (group 1 or 2) on caller-provided scratch, result affine in the wire
// encoding; the caller holds d->mu with d's device current.
void msm_device_on_scratch(g16hip_device* d, int group, const void* bases, const void* scalars,
                           uint32_t n, int c, const MsmScratch& sc, uint8_t* out);

}  // namespace g16
