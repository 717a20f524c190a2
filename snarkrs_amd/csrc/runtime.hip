// The C ABI of include/g16hip.h: device handles, prepared circuits, the prove pipeline
// (stages 0-9 on the GPU, stage 11 on the host) and the standalone entry points.
#include <algorithm>
#include <chrono>
#include <exception>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/g16hip.h"
#include "bn254.cuh"
#include "internal.h"

using namespace g16;

namespace g16 {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw HipError{std::string(what) + ": " + hipGetErrorString(e)};
  }
}

}  // namespace g16

// ---------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------
static thread_local std::string t_last_error;

namespace g16 {
void set_last_error(const std::string& msg) { t_last_error = msg; }
}  // namespace g16

static g16hip_status fail(g16hip_status st, const std::string& msg) {
  t_last_error = msg;
  return st;
}

struct StatusError {
  g16hip_status st;
  std::string msg;
};

template <class F>
static g16hip_status guarded(F&& f) {
  try {
    t_last_error.clear();
    return f();
  } catch (const HipError& e) {
    bool oom = e.msg.find("out of memory") != std::string::npos;
    return fail(oom ? G16HIP_DEVICE_MEMORY : G16HIP_DEVICE, e.msg);
  } catch (const ArgError& e) {
    return fail(G16HIP_BACKEND, e.msg);
  } catch (const StatusError& e) {
    return fail(e.st, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(G16HIP_BACKEND, "host allocation failed");
  }
}

// ---------------------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------------------
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

static void upload(DevBuf& b, const void* src, size_t n, hipStream_t s) {
  b.alloc(n);
  if (n) G16_CHECK(hipMemcpyAsync(b.p, src, n, hipMemcpyHostToDevice, s));
}

// ---------------------------------------------------------------------------------------
// host-side helpers (the same bn254.cuh arithmetic, compiled for the CPU)
// ---------------------------------------------------------------------------------------
static Fr fr_load(const uint8_t* p) {
  Fr x;
  std::memcpy(x.v, p, 32);
  return x;
}
static G1Aff g1_load(const uint8_t* p) {
  G1Aff a;
  std::memcpy(a.x.v, p, 32);
  std::memcpy(a.y.v, p + 32, 32);
  return a;
}
static G2Aff g2_load(const uint8_t* p) {
  G2Aff a;
  std::memcpy(a.x.c0.v, p, 32);
  std::memcpy(a.x.c1.v, p + 32, 32);
  std::memcpy(a.y.c0.v, p + 64, 32);
  std::memcpy(a.y.c1.v, p + 96, 32);
  return a;
}
static void g1_store(uint8_t* p, const G1Aff& a) {
  std::memcpy(p, a.x.v, 32);
  std::memcpy(p + 32, a.y.v, 32);
}
static void g2_store(uint8_t* p, const G2Aff& a) {
  std::memcpy(p, a.x.c0.v, 32);
  std::memcpy(p + 32, a.x.c1.v, 32);
  std::memcpy(p + 64, a.y.c0.v, 32);
  std::memcpy(p + 96, a.y.c1.v, 32);
}

static bool fq_canonical(const Fq& x) {
  for (int i = 7; i >= 0; --i) {
    if (x.v[i] != FqCfg::p(i)) return x.v[i] < FqCfg::p(i);
  }
  return false;
}

// Domain root of order 2^L: 5^((r-1)/2^28) squared 28 - L times (crates/field/src/lib.rs:48-82).
static Fr domain_root(int L) {
  Fr w;
  const uint32_t W[8] = {0x725b19f0u, 0x9bd61b6eu, 0x41112ed4u, 0x402d111eu,
                         0x8ef62abcu, 0x00e0a7ebu, 0xa58a7e85u, 0x2a3c09f0u};
  std::memcpy(w.v, W, 32);
  w = from_std(w);
  for (int i = L; i < 28; ++i) w = sqr(w);
  return w;
}

static Fr fr_from_u64(uint64_t x) {
  Fr r = fp_zero<FrCfg>();
  r.v[0] = (uint32_t)x;
  r.v[1] = (uint32_t)(x >> 32);
  return from_std(r);
}

// sum_w 2^(c*w) S_w on the host (the reference finishes the same way, crates/cuda/src/msm.rs:3976-4010).
template <class F>
static Xyzz<F> horner(const std::vector<Xyzz<F>>& S, int c) {
  Xyzz<F> acc = xyzz_inf<F>();
  for (int w = (int)S.size() - 1; w >= 0; --w) {
    for (int i = 0; i < c && w != (int)S.size() - 1; ++i) acc = xyzz_dbl(acc);
    acc = xyzz_add(acc, S[w]);
  }
  return acc;
}

static float ev_ms(hipEvent_t a, hipEvent_t b) {
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
};

// ---------------------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------------------
struct NttTables {
  DevBuf fwd, inv, coset, scalars;  // scalars: [w, w^-1, shift, 1/n]
};

struct g16hip_device {
  int ordinal = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // serialises the standalone entry points' scratch
  DevBuf scratch_a, scratch_b, scratch_c, scratch_d, scratch_e;
  std::map<int, std::unique_ptr<NttTables>> ntt;
  hipEvent_t ev[6] = {};
  g16hip_msm_stats stats = {};
  std::mutex stats_mu;
  g16hip_kernel_stats kstats = {};
};

static void note_accumulate(g16hip_device* d, int group, uint64_t entries, double ms) {
  std::lock_guard<std::mutex> lk(d->stats_mu);
  d->kstats.launches[group - 1] += 1;
  d->kstats.entries[group - 1] += entries;
  d->kstats.accumulate_ms[group - 1] += ms;
}

static void set_device(const g16hip_device* d) { G16_CHECK(hipSetDevice(d->ordinal)); }

static void build_ntt_tables(NttTables& t, int L, hipStream_t s) {
  size_t n = (size_t)1 << L;
  size_t half = n / 2 ? n / 2 : 1;
  Fr host[6];
  host[0] = domain_root(L);
  host[1] = inv(host[0]);
  host[2] = domain_root(L + 1);  // coset shift: primitive 2n-th root (snarkjs' odd coset)
  host[3] = inv(fr_from_u64(n));
  host[4] = fp_one<FrCfg>();
  host[5] = fp_zero<FrCfg>();
  upload(t.scalars, host, sizeof host, s);
  t.fwd.alloc(32 * half);
  t.inv.alloc(32 * half);
  t.coset.alloc(32 * n);
  Fr* sc = t.scalars.as<Fr>();
  launch_powers(t.fwd.p, sc + 0, half, s);
  launch_powers(t.inv.p, sc + 1, half, s);
  launch_coset_table(t.coset.p, sc + 2, sc + 3, L, s);
}

static NttTables& ntt_tables(g16hip_device* d, int L) {
  auto it = d->ntt.find(L);
  if (it != d->ntt.end()) return *it->second;
  auto t = std::make_unique<NttTables>();
  build_ntt_tables(*t, L, d->stream);
  NttTables& ref = *t;
  d->ntt[L] = std::move(t);
  return ref;
}

// One MSM on device-resident bases and scalars with scratch from `d`.
template <class F>
static Xyzz<F> run_msm(g16hip_device* d, int group, const void* bases, const void* scalars,
                       uint32_t n, int c) {
  if (c <= 0) c = msm_default_c(n);
  if (c < 2 || c > 20) throw ArgError{"window width out of range"};
  MsmPlan plan;
  d->scratch_d.alloc(msm_sort_scratch_bytes(n, c));
  msm_plan_bind(plan, d->scratch_d.p, n, c);
  d->scratch_e.alloc(msm_accum_scratch_bytes(plan, group) + sizeof(Xyzz<F>) * plan.windows);
  G16_CHECK(hipEventRecord(d->ev[0], d->stream));
  msm_sort(plan, scalars, d->stream);
  char* wout = static_cast<char*>(d->scratch_e.p) + msm_accum_scratch_bytes(plan, group);
  msm_accumulate(plan, group, bases, d->scratch_e.p, wout, d->stream, d->ev + 1);
  std::vector<Xyzz<F>> S(plan.windows);
  uint32_t entries = 0;
  G16_CHECK(hipMemcpyAsync(S.data(), wout, sizeof(Xyzz<F>) * plan.windows,
                           hipMemcpyDeviceToHost, d->stream));
  G16_CHECK(hipMemcpyAsync(&entries, plan.total, 4, hipMemcpyDeviceToHost, d->stream));
  G16_CHECK(hipEventRecord(d->ev[5], d->stream));
  G16_CHECK(hipStreamSynchronize(d->stream));
  g16hip_msm_stats& st = d->stats;
  st.sort_ms = ev_ms(d->ev[0], d->ev[1]);
  st.accumulate_ms = ev_ms(d->ev[1], d->ev[2]);
  st.merge_ms = ev_ms(d->ev[2], d->ev[3]);
  st.reduce_ms = ev_ms(d->ev[3], d->ev[4]);
  st.total_ms = ev_ms(d->ev[0], d->ev[5]);
  st.entries = entries;
  st.c = (uint32_t)c;
  st.windows = (uint32_t)plan.windows;
  note_accumulate(d, group, entries, st.accumulate_ms);
  return horner(S, c);
}

// ---------------------------------------------------------------------------------------
// prepared circuit
// ---------------------------------------------------------------------------------------
// A proof runs on three streams of its circuit: stream (0) stages 0-4, the H sort and the H
// MSM; side[0] the witness sort and the A, B1 and L MSMs; side[1] the B2 (G2) MSM. Stages 5-8
// read only the witness, so they start as soon as it is on the device and overlap the
// transforms, and the latency-bound tail of each MSM (merges, window reduction) overlaps the
// throughput-bound accumulation of another.
struct g16hip_circuit {
  g16hip_device* dev = nullptr;
  std::mutex mu;
  hipStream_t stream = nullptr;
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t ev[8];
  hipEvent_t sync_ev[4];  // witness ready, witness sorted, side[0] done, side[1] done
  hipEvent_t done_ev[5];  // window sums of A, B1, L, B2, H copied to host_wout
  uint8_t* host_wout = nullptr;  // pinned: window sums + the two entry totals
  hipEvent_t mev[5][4];  // per MSM of a proof: around accumulation, merge, reduction
  uint32_t n_vars = 0, n_public = 0, domain = 0;
  int log_n = 0;
  G1Aff alpha_g1, beta_g1, delta_g1;
  G2Aff beta_g2, delta_g2;
  DevBuf row_ptr[2], signal[2], value_id[2], values;
  DevBuf a_q, b1_q, b2_q, l_q, h_q;
  NttTables tables;
  DevBuf witness, abc, hbuf;
  DevBuf sort_w, sort_h, accum_w1, accum_w2, accum_h, wout;
  MsmPlan plan_w, plan_h;
  size_t resident = 0;
};

struct g16hip_h {
  g16hip_circuit* owner = nullptr;
  DevBuf buf;
};

static void validate_key(const g16hip_key* k) {
  if (!k) throw ArgError{"null key"};
  uint32_t n = k->domain_size;
  if (n == 0 || (n & (n - 1))) throw ArgError{"domain size " + std::to_string(n) + " is not a power of two"};
  if (n > (1u << 27)) throw ArgError{"domain size exceeds 2^27"};
  if ((uint64_t)k->n_public + 1 > k->n_vars)
    throw ArgError{"n_public " + std::to_string(k->n_public) + " exceeds n_vars " + std::to_string(k->n_vars)};
  for (int m = 0; m < 2; ++m) {
    const uint32_t* rp = k->row_ptr[m];
    if (!rp) throw ArgError{"missing row_ptr"};
    if (rp[0] != 0) throw ArgError{"matrix " + std::to_string(m) + " row_ptr starts at " + std::to_string(rp[0])};
    for (uint32_t c = 0; c < n; ++c)
      if (rp[c + 1] < rp[c]) throw ArgError{"matrix " + std::to_string(m) + " row_ptr is not monotone"};
    uint32_t nnz = rp[n];
    for (uint32_t e = 0; e < nnz; ++e) {
      if (k->signal[m][e] >= k->n_vars) throw ArgError{"signal index out of range"};
      if (k->value_id[m][e] >= k->n_values) throw ArgError{"value index out of range"};
    }
  }
  const uint8_t* pts[] = {k->alpha_g1, k->beta_g1, k->delta_g1};
  for (auto p : pts) {
    G1Aff a = g1_load(p);
    if (!fq_canonical(a.x) || !fq_canonical(a.y)) throw ArgError{"non-canonical key point"};
    if (aff_is_inf(a)) throw ArgError{"a key point is the point at infinity, which no honest setup produces"};
  }
  if (aff_is_inf(g2_load(k->beta_g2)) || aff_is_inf(g2_load(k->delta_g2)))
    throw ArgError{"a key point is the point at infinity, which no honest setup produces"};
}

static void prepare_circuit(g16hip_circuit* C, const g16hip_key* k) {
  hipStream_t s = C->stream;
  uint32_t n = k->domain_size;
  C->n_vars = k->n_vars;
  C->n_public = k->n_public;
  C->domain = n;
  C->log_n = 0;
  while ((1u << C->log_n) < n) ++C->log_n;
  C->alpha_g1 = g1_load(k->alpha_g1);
  C->beta_g1 = g1_load(k->beta_g1);
  C->delta_g1 = g1_load(k->delta_g1);
  C->beta_g2 = g2_load(k->beta_g2);
  C->delta_g2 = g2_load(k->delta_g2);
  for (int m = 0; m < 2; ++m) {
    uint32_t nnz = k->row_ptr[m][n];
    upload(C->row_ptr[m], k->row_ptr[m], 4 * (size_t)(n + 1), s);
    upload(C->signal[m], k->signal[m], 4 * (size_t)nnz, s);
    upload(C->value_id[m], k->value_id[m], 4 * (size_t)nnz, s);
  }
  upload(C->values, k->values, 32 * (size_t)k->n_values, s);
  size_t nv = k->n_vars;
  upload(C->a_q, k->a_query, 64 * nv, s);
  upload(C->b1_q, k->b_g1_query, 64 * nv, s);
  upload(C->b2_q, k->b_g2_query, 128 * nv, s);
  // L covers the private wires only; it is stored behind n_public + 1 points at infinity so
  // that it multiplies the full witness and shares the witness sort with A, B1 and B2.
  size_t pub = (size_t)k->n_public + 1;
  C->l_q.alloc(64 * nv);
  G16_CHECK(hipMemsetAsync(C->l_q.p, 0, 64 * pub, s));
  if (nv > pub)
    G16_CHECK(hipMemcpyAsync(static_cast<char*>(C->l_q.p) + 64 * pub, k->l_query, 64 * (nv - pub),
                             hipMemcpyHostToDevice, s));
  upload(C->h_q, k->h_query, 64 * (size_t)n, s);
  build_ntt_tables(C->tables, C->log_n, s);
  // Per-proof scratch, sized once here so a proof allocates nothing.
  C->witness.alloc(32 * nv);
  C->abc.alloc(32 * 3 * (size_t)n);
  C->hbuf.alloc(32 * (size_t)n);
  int cw = msm_default_c((uint32_t)nv), ch = msm_default_c(n);
  C->sort_w.alloc(msm_sort_scratch_bytes((uint32_t)nv, cw));
  C->sort_h.alloc(msm_sort_scratch_bytes(n, ch));
  msm_plan_bind(C->plan_w, C->sort_w.p, (uint32_t)nv, cw);
  msm_plan_bind(C->plan_h, C->sort_h.p, n, ch);
  C->accum_w1.alloc(msm_accum_scratch_bytes(C->plan_w, 1));
  C->accum_w2.alloc(msm_accum_scratch_bytes(C->plan_w, 2));
  C->accum_h.alloc(msm_accum_scratch_bytes(C->plan_h, 1));
  C->wout.alloc(256 * (size_t)(4 * C->plan_w.windows + C->plan_h.windows));
  G16_CHECK(hipHostMalloc((void**)&C->host_wout,
                          256 * (size_t)(4 * C->plan_w.windows + C->plan_h.windows) + 64,
                          hipHostMallocDefault));
  G16_CHECK(hipStreamSynchronize(s));
  size_t tot = 0;
  DevBuf* bufs[] = {&C->row_ptr[0], &C->row_ptr[1], &C->signal[0], &C->signal[1], &C->value_id[0],
                    &C->value_id[1], &C->values, &C->a_q, &C->b1_q, &C->b2_q, &C->l_q, &C->h_q,
                    &C->tables.fwd, &C->tables.inv, &C->tables.coset, &C->witness, &C->abc,
                    &C->hbuf, &C->sort_w, &C->sort_h, &C->accum_w1, &C->accum_w2,
                    &C->accum_h, &C->wout};
  for (auto b : bufs) tot += b->bytes;
  C->resident = tot;
}

static void check_witness_len(g16hip_circuit* C, size_t n) {
  if (n != C->n_vars)
    throw StatusError{G16HIP_WITNESS_LENGTH, "witness has " + std::to_string(n) +
                                                 " entries, proving key expects " +
                                                 std::to_string(C->n_vars)};
}

// Stages 0-4 from the circuit's device witness buffer into `h_out` (device).
static void stages_0_4(g16hip_circuit* C, void* h_out) {
  hipStream_t s = C->stream;
  size_t n = C->domain;
  char* abc = C->abc.as<char>();
  void* A = abc;
  void* B = abc + 32 * n;
  void* Cv = abc + 64 * n;
  G16_CHECK(hipEventRecord(C->ev[0], s));
  launch_gather(C->row_ptr[0].as<uint32_t>(), C->signal[0].as<uint32_t>(),
                C->value_id[0].as<uint32_t>(), C->values.p, C->witness.p, A, C->domain, s);
  launch_gather(C->row_ptr[1].as<uint32_t>(), C->signal[1].as<uint32_t>(),
                C->value_id[1].as<uint32_t>(), C->values.p, C->witness.p, B, C->domain, s);
  launch_mul_into(Cv, A, B, n, s);
  G16_CHECK(hipEventRecord(C->ev[1], s));
  ntt_dif_inverse(abc, 3, C->log_n, C->tables.inv.p, s);
  ntt_dit_forward(abc, 3, C->log_n, C->tables.fwd.p, C->tables.coset.p, s);
  G16_CHECK(hipEventRecord(C->ev[2], s));
  launch_h(h_out, A, B, Cv, n, s);
  G16_CHECK(hipEventRecord(C->ev[3], s));
}

// Host work that rides on the MSM results: with `blind`, s*A and r*B1 (the two stage-11
// products that need an MSM output) are computed as soon as A and B1 are done, while the GPU
// is still on the H MSM.
struct BlindProducts {
  bool want = false;
  Fr r_std, s_std;
  G1X sA, rB1;
};

// Runs fn() on a thread and keeps the first exception for the joiner.
struct HostTask {
  std::exception_ptr err;  // declared first: the thread may set it as soon as it starts
  std::thread th;
  template <class Fn>
  explicit HostTask(Fn fn) : th([this, fn] {
    try {
      fn();
    } catch (...) {
      err = std::current_exception();
    }
  }) {}
  void join() {
    if (th.joinable()) th.join();
  }
};

// Stages 5-9 (and, with `with_h`, stages 0-4 first on the main stream): the witness is
// already being copied to C->witness on the main stream when this is called; H is read from
// `h` (device), which stages 0-4 write when `with_h`. Each MSM's window sums go to pinned
// host memory on its own stream, and a host thread finishes that MSM (Horner, affine) as soon
// as they land.
static void run_stages(g16hip_circuit* C, bool with_h, void* h, g16hip_msm_out* out,
                       g16hip_timings* t, BlindProducts* blind = nullptr) {
  hipStream_t s0 = C->stream, s1 = C->side[0], s2 = C->side[1];
  MsmPlan& pw = C->plan_w;
  MsmPlan& ph = C->plan_h;
  char* wout = C->wout.as<char>();
  uint8_t* hw = C->host_wout;
  const size_t Ww = pw.windows, Wh = ph.windows;
  const size_t off[5] = {0, 128 * Ww, 256 * Ww, 384 * Ww, 640 * Ww};
  const size_t len[5] = {128 * Ww, 128 * Ww, 128 * Ww, 256 * Ww, 128 * Wh};
  uint32_t* totals = reinterpret_cast<uint32_t*>(hw + 640 * Ww + 128 * Wh);
  auto finish_on = [&](int k, hipStream_t st) {
    G16_CHECK(hipMemcpyAsync(hw + off[k], wout + off[k], len[k], hipMemcpyDeviceToHost, st));
    G16_CHECK(hipEventRecord(C->done_ev[k], st));
  };
  G16_CHECK(hipEventRecord(C->sync_ev[0], s0));  // witness on the device
  G16_CHECK(hipEventRecord(C->ev[4], s0));
  // side[0]: witness sort, then A, B1, L.
  G16_CHECK(hipStreamWaitEvent(s1, C->sync_ev[0], 0));
  msm_sort(pw, C->witness.p, s1);
  G16_CHECK(hipEventRecord(C->sync_ev[1], s1));
  G16_CHECK(hipMemcpyAsync(&totals[0], pw.total, 4, hipMemcpyDeviceToHost, s1));
  msm_accumulate(pw, 1, C->a_q.p, C->accum_w1.p, wout + off[0], s1, C->mev[0]);
  finish_on(0, s1);
  msm_accumulate(pw, 1, C->b1_q.p, C->accum_w1.p, wout + off[1], s1, C->mev[1]);
  finish_on(1, s1);
  msm_accumulate(pw, 1, C->l_q.p, C->accum_w1.p, wout + off[2], s1, C->mev[2]);
  finish_on(2, s1);
  G16_CHECK(hipEventRecord(C->sync_ev[2], s1));
  // side[1]: B2 on the same sorted entries.
  G16_CHECK(hipStreamWaitEvent(s2, C->sync_ev[1], 0));
  msm_accumulate(pw, 2, C->b2_q.p, C->accum_w2.p, wout + off[3], s2, C->mev[3]);
  finish_on(3, s2);
  G16_CHECK(hipEventRecord(C->sync_ev[3], s2));
  // main: stages 0-4, then the H sort and MSM.
  if (with_h) stages_0_4(C, h);
  msm_sort(ph, h, s0);
  G16_CHECK(hipMemcpyAsync(&totals[1], ph.total, 4, hipMemcpyDeviceToHost, s0));
  msm_accumulate(ph, 1, C->h_q.p, C->accum_h.p, wout + off[4], s0, C->mev[4]);
  finish_on(4, s0);
  G16_CHECK(hipStreamWaitEvent(s0, C->sync_ev[2], 0));
  G16_CHECK(hipStreamWaitEvent(s0, C->sync_ev[3], 0));
  G16_CHECK(hipEventRecord(C->ev[5], s0));

  auto g1_sum = [&](int k, int c) {
    G16_CHECK(hipEventSynchronize(C->done_ev[k]));
    std::vector<G1X> S(len[k] / sizeof(G1X));
    std::memcpy(S.data(), hw + off[k], len[k]);
    return horner(S, c);
  };
  G1X A, B1;
  HostTask ta([&] {
    A = g1_sum(0, pw.c);
    g1_store(out->a_g1, xyzz_to_aff(A));
    if (blind && blind->want) blind->sA = xyzz_mul(A, blind->s_std.v);
  });
  HostTask tb([&] {
    B1 = g1_sum(1, pw.c);
    g1_store(out->b_g1, xyzz_to_aff(B1));
    if (blind && blind->want) blind->rB1 = xyzz_mul(B1, blind->r_std.v);
  });
  HostTask tl([&] { g1_store(out->l_g1, xyzz_to_aff(g1_sum(2, pw.c))); });
  HostTask t2([&] {
    G16_CHECK(hipEventSynchronize(C->done_ev[3]));
    std::vector<G2X> S(Ww);
    std::memcpy(S.data(), hw + off[3], len[3]);
    g2_store(out->b_g2, xyzz_to_aff(horner(S, pw.c)));
  });
  std::exception_ptr err;
  try {
    g1_store(out->h_g1, xyzz_to_aff(g1_sum(4, ph.c)));
    G16_CHECK(hipStreamSynchronize(s0));
  } catch (...) {
    err = std::current_exception();
  }
  for (HostTask* task : {&ta, &tb, &tl, &t2}) {
    task->join();
    if (!err && task->err) err = task->err;
  }
  if (err) {
    (void)hipStreamSynchronize(s1);
    (void)hipStreamSynchronize(s2);
    std::rethrow_exception(err);
  }
  for (int k = 0; k < 5; ++k)
    note_accumulate(C->dev, k == 3 ? 2 : 1, k == 4 ? totals[1] : totals[0],
                    ev_ms(C->mev[k][0], C->mev[k][1]));
  static const bool timeline = std::getenv("G16HIP_TIMELINE") != nullptr;
  if (timeline) {
    // Milliseconds from the witness arriving on the device, per stream (the
    // G16_CUDA_MEMORY_TRACE kind of switch, crates/cuda/src/backend.rs:89-93).
    static const char* names[5] = {"A", "B1", "L", "B2", "H"};
    std::fprintf(stderr, "[g16hip timeline]");
    if (with_h)
      std::fprintf(stderr, " ntt %.2f-%.2f h_done %.2f |", ev_ms(C->ev[4], C->ev[1]),
                   ev_ms(C->ev[4], C->ev[2]), ev_ms(C->ev[4], C->ev[3]));
    for (int k = 0; k < 5; ++k)
      std::fprintf(stderr, " %s acc %.2f-%.2f merge %.2f reduce %.2f |", names[k],
                   ev_ms(C->ev[4], C->mev[k][0]), ev_ms(C->ev[4], C->mev[k][1]),
                   ev_ms(C->ev[4], C->mev[k][2]), ev_ms(C->ev[4], C->mev[k][3]));
    std::fprintf(stderr, " end %.2f\n", ev_ms(C->ev[4], C->ev[5]));
  }
  if (t) {
    if (with_h) {
      t->gather_us += (uint64_t)(ev_ms(C->ev[0], C->ev[1]) * 1000.0f);
      t->ntt_us += (uint64_t)(ev_ms(C->ev[1], C->ev[2]) * 1000.0f);
      t->pointwise_us += (uint64_t)(ev_ms(C->ev[2], C->ev[3]) * 1000.0f);
    }
    // Stages 5-9 overlap stages 0-4: msm_us is the wall time from the witness arriving to the
    // last MSM finishing, as the reference reports overlapped MSMs (crates/groth16/src/lib.rs:333-335).
    t->msm_us += (uint64_t)(ev_ms(C->ev[4], C->ev[5]) * 1000.0f);
  }
}

static void add_stage_timings(g16hip_circuit* C, g16hip_timings* t) {
  if (!t) return;
  t->gather_us += (uint64_t)(ev_ms(C->ev[0], C->ev[1]) * 1000.0f);
  t->ntt_us += (uint64_t)(ev_ms(C->ev[1], C->ev[2]) * 1000.0f);
  t->pointwise_us += (uint64_t)(ev_ms(C->ev[2], C->ev[3]) * 1000.0f);
}

static void load_witness(g16hip_circuit* C, const void* w, size_t n, bool on_device) {
  check_witness_len(C, n);
  G16_CHECK(hipMemcpyAsync(C->witness.p, w, 32 * n,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, C->stream));
}

static void h_and_msms_locked(g16hip_circuit* C, const void* w, size_t n, bool on_device,
                              g16hip_msm_out* out, g16hip_timings* t) {
  load_witness(C, w, n, on_device);
  run_stages(C, true, C->hbuf.p, out, t);
}

// Stage 11 (crates/groth16/src/prove.rs:391-430), on the host:
//   pi_a = A + alpha1 + r delta1,  pi_b = B2 + beta2 + s delta2,
//   pi_c = L + H + s pi_a + r (B1 + beta1 + s delta1) - r s delta1
//        = L + H + s A + r B1 + (s alpha1 + r beta1 + r s delta1).
// Everything that needs only the key and (r, s) is computed beside the GPU work; s A and r B1
// as soon as those MSMs finish (BlindProducts). Group elements are canonical, so the proof is
// the same point whatever order the terms are added in.
struct BlinderTerms {
  G1X d1r, k1;  // r delta1;  s alpha1 + r beta1 + r s delta1
  G2X d2s;      // s delta2
};

static BlinderTerms blinder_terms(const g16hip_circuit* C, const Fr& r_std, const Fr& s_std,
                                  const Fr& rs_std) {
  BlinderTerms b;
  b.d1r = xyzz_mul_plain(C->delta_g1, r_std.v);
  b.k1 = xyzz_add(xyzz_add(xyzz_mul_plain(C->alpha_g1, s_std.v),
                           xyzz_mul_plain(C->beta_g1, r_std.v)),
                  xyzz_mul_plain(C->delta_g1, rs_std.v));
  b.d2s = xyzz_mul_plain(C->delta_g2, s_std.v);
  return b;
}

static void assemble(const g16hip_circuit* C, const g16hip_msm_out& m, const BlinderTerms& b,
                     const BlindProducts& bp, g16hip_proof* proof) {
  G1X pi_a = xyzz_add(xyzz_add_aff(xyzz_from_aff(g1_load(m.a_g1)), C->alpha_g1), b.d1r);
  G2X pi_b = xyzz_add(xyzz_add_aff(xyzz_from_aff(g2_load(m.b_g2)), C->beta_g2), b.d2s);
  G1X pi_c = xyzz_add_aff(xyzz_from_aff(g1_load(m.l_g1)), g1_load(m.h_g1));
  pi_c = xyzz_add(pi_c, bp.sA);
  pi_c = xyzz_add(pi_c, bp.rB1);
  pi_c = xyzz_add(pi_c, b.k1);
  g1_store(proof->a, xyzz_to_aff(pi_a));
  g2_store(proof->b, xyzz_to_aff(pi_b));
  g1_store(proof->c, xyzz_to_aff(pi_c));
}

static g16hip_status prove_impl(g16hip_circuit* C, const void* w, size_t n, bool on_device,
                                const uint8_t* r8, const uint8_t* s8, g16hip_proof* proof,
                                g16hip_timings* t) {
  return guarded([&] {
    if (!C || !r8 || !s8 || !proof) throw ArgError{"null argument"};
    check_witness_len(C, n);
    // w[0] is the constant-one wire (crates/groth16/src/prove.rs:272-283).
    Fr w0;
    if (on_device) {
      G16_CHECK(hipSetDevice(C->dev->ordinal));
      G16_CHECK(hipMemcpy(w0.v, w, 32, hipMemcpyDeviceToHost));
    } else {
      std::memcpy(w0.v, w, 32);
    }
    if (!is_one(w0)) throw StatusError{G16HIP_CONSTANT_WIRE, "w[0] is the constant-one wire and must be 1"};
    Fr r = fr_load(r8), s = fr_load(s8);
    if (is_zero(r) || is_zero(s)) throw StatusError{G16HIP_ZERO_BLINDER, "a blinder is zero"};
    BlindProducts bp;
    bp.want = true;
    bp.r_std = to_std(r);
    bp.s_std = to_std(s);
    Fr rs_std = to_std(mul(r, s));
    BlinderTerms terms;
    HostTask th([&] { terms = blinder_terms(C, bp.r_std, bp.s_std, rs_std); });
    g16hip_msm_out m;
    try {
      std::lock_guard<std::mutex> lk(C->mu);
      G16_CHECK(hipSetDevice(C->dev->ordinal));
      load_witness(C, w, n, on_device);
      run_stages(C, true, C->hbuf.p, &m, t, &bp);
    } catch (...) {
      th.join();
      throw;
    }
    th.join();
    if (th.err) std::rethrow_exception(th.err);
    Timer ta;
    assemble(C, m, terms, bp, proof);
    if (t) t->assemble_us += ta.us();
    return G16HIP_OK;
  });
}

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
extern "C" {

const char* g16hip_last_error(void) { return t_last_error.c_str(); }

g16hip_status g16hip_open(int ordinal, g16hip_device** out) {
  return guarded([&] {
    if (!out) throw ArgError{"null output handle"};
    int n = 0;
    G16_CHECK(hipGetDeviceCount(&n));
    if (ordinal < 0 || ordinal >= n)
      throw ArgError{"device ordinal " + std::to_string(ordinal) + " out of range (" +
                     std::to_string(n) + " devices)"};
    auto* d = new g16hip_device();
    d->ordinal = ordinal;
    G16_CHECK(hipSetDevice(ordinal));
    G16_CHECK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    for (auto& e : d->ev) G16_CHECK(hipEventCreate(&e));
    *out = d;
    return G16HIP_OK;
  });
}

void g16hip_close(g16hip_device* d) {
  if (!d) return;
  (void)hipSetDevice(d->ordinal);
  (void)hipStreamSynchronize(d->stream);
  d->ntt.clear();
  d->scratch_a.release();
  d->scratch_b.release();
  d->scratch_c.release();
  d->scratch_d.release();
  d->scratch_e.release();
  for (auto& e : d->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(d->stream);
  delete d;
}

g16hip_status g16hip_mem_info(g16hip_device* d, size_t* free_b, size_t* total_b) {
  return guarded([&] {
    set_device(d);
    G16_CHECK(hipMemGetInfo(free_b, total_b));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_field_probe(g16hip_device* d, int field, int op, const uint8_t* a,
                                 const uint8_t* b, uint8_t* out, size_t n) {
  return guarded([&] {
    if (field < 0 || field > 2 || op < 0 || op > 7) throw ArgError{"bad field/op"};
    if (n == 0) return G16HIP_OK;
    std::lock_guard<std::mutex> lk(d->mu);
    set_device(d);
    size_t w = field == 2 ? 64 : 32;
    upload(d->scratch_a, a, n * w, d->stream);
    if (b) upload(d->scratch_b, b, n * w, d->stream);
    d->scratch_c.alloc(n * w);
    launch_field_probe(field, op, d->scratch_a.p, b ? d->scratch_b.p : nullptr, d->scratch_c.p,
                       n, d->stream);
    G16_CHECK(hipMemcpyAsync(out, d->scratch_c.p, n * w, hipMemcpyDeviceToHost, d->stream));
    G16_CHECK(hipStreamSynchronize(d->stream));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_fixed_base(g16hip_device* d, int group, const uint8_t* k, size_t n,
                                uint8_t* out) {
  return guarded([&] {
    if (group != 1 && group != 2) throw ArgError{"group must be 1 or 2"};
    if (n == 0) return G16HIP_OK;
    std::lock_guard<std::mutex> lk(d->mu);
    set_device(d);
    size_t w = group == 1 ? 64 : 128;
    upload(d->scratch_a, k, n * 32, d->stream);
    d->scratch_c.alloc(n * w);
    launch_fixed_base(group, d->scratch_a.p, n, d->scratch_c.p, d->stream);
    G16_CHECK(hipMemcpyAsync(out, d->scratch_c.p, n * w, hipMemcpyDeviceToHost, d->stream));
    G16_CHECK(hipStreamSynchronize(d->stream));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_valu_probe(g16hip_device* d, int kind, double* per_second) {
  return guarded([&] {
    if (!d || !per_second || kind < 0 || kind > 1) throw ArgError{"bad argument"};
    std::lock_guard<std::mutex> lk(d->mu);
    set_device(d);
    std::vector<uint32_t> seed(1024);
    for (uint32_t i = 0; i < 1024; ++i) seed[i] = 0x9E3779B9u * (i + 1);
    upload(d->scratch_a, seed.data(), 4096, d->stream);
    const unsigned blocks = 256 * 16;  // 16 workgroups of 256 per CU
    const int iters = kind == 0 ? 4096 : 256;
    d->scratch_b.alloc((size_t)blocks * 256 * 8);
    launch_valu_peak(kind, d->scratch_a.p, d->scratch_b.p, blocks, 8, d->stream);  // warm
    G16_CHECK(hipEventRecord(d->ev[0], d->stream));
    launch_valu_peak(kind, d->scratch_a.p, d->scratch_b.p, blocks, iters, d->stream);
    G16_CHECK(hipEventRecord(d->ev[1], d->stream));
    G16_CHECK(hipStreamSynchronize(d->stream));
    double sec = ev_ms(d->ev[0], d->ev[1]) / 1e3;
    double ops = (double)blocks * 256 * iters * (kind == 0 ? 8 : 4);
    *per_second = ops / sec;
    return G16HIP_OK;
  });
}

static g16hip_status msm_host(g16hip_device* d, int group, const uint8_t* bases,
                              const uint8_t* scalars, size_t n, uint8_t* out) {
  return guarded([&] {
    if (!d || !out) throw ArgError{"null argument"};
    size_t pw = group == 1 ? 64 : 128;
    if (n == 0) {
      std::memset(out, 0, pw);
      return G16HIP_OK;
    }
    if (n >= (1u << 31)) throw ArgError{"MSM larger than 2^31 points"};
    std::lock_guard<std::mutex> lk(d->mu);
    set_device(d);
    upload(d->scratch_a, bases, n * pw, d->stream);
    upload(d->scratch_b, scalars, n * 32, d->stream);
    if (group == 1)
      g1_store(out, xyzz_to_aff(run_msm<Fq>(d, 1, d->scratch_a.p, d->scratch_b.p, (uint32_t)n, 0)));
    else
      g2_store(out, xyzz_to_aff(run_msm<Fq2>(d, 2, d->scratch_a.p, d->scratch_b.p, (uint32_t)n, 0)));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_msm_g1(g16hip_device* d, const uint8_t* b, const uint8_t* s, size_t n,
                            uint8_t* out) {
  return msm_host(d, 1, b, s, n, out);
}

g16hip_status g16hip_msm_g2(g16hip_device* d, const uint8_t* b, const uint8_t* s, size_t n,
                            uint8_t* out) {
  return msm_host(d, 2, b, s, n, out);
}

static g16hip_status msm_dev(g16hip_device* d, int group, const void* bases, const void* scalars,
                             size_t n, int c, uint8_t* out) {
  return guarded([&] {
    if (!d || !out) throw ArgError{"null argument"};
    size_t pw = group == 1 ? 64 : 128;
    if (n == 0) {
      std::memset(out, 0, pw);
      return G16HIP_OK;
    }
    if (n >= (1u << 31)) throw ArgError{"MSM larger than 2^31 points"};
    std::lock_guard<std::mutex> lk(d->mu);
    set_device(d);
    if (group == 1)
      g1_store(out, xyzz_to_aff(run_msm<Fq>(d, 1, bases, scalars, (uint32_t)n, c)));
    else
      g2_store(out, xyzz_to_aff(run_msm<Fq2>(d, 2, bases, scalars, (uint32_t)n, c)));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_msm_g1_device(g16hip_device* d, const void* b, const void* s, size_t n, int c,
                                   uint8_t* out) {
  return msm_dev(d, 1, b, s, n, c, out);
}

g16hip_status g16hip_msm_g2_device(g16hip_device* d, const void* b, const void* s, size_t n, int c,
                                   uint8_t* out) {
  return msm_dev(d, 2, b, s, n, c, out);
}

g16hip_status g16hip_msm_last_stats(g16hip_device* d, g16hip_msm_stats* out) {
  return guarded([&] {
    if (!d || !out) throw ArgError{"null argument"};
    std::lock_guard<std::mutex> lk(d->mu);
    *out = d->stats;
    return G16HIP_OK;
  });
}

g16hip_status g16hip_kernel_stats_get(g16hip_device* d, g16hip_kernel_stats* out) {
  return guarded([&] {
    if (!d || !out) throw ArgError{"null argument"};
    std::lock_guard<std::mutex> lk(d->stats_mu);
    *out = d->kstats;
    return G16HIP_OK;
  });
}

g16hip_status g16hip_kernel_stats_reset(g16hip_device* d) {
  return guarded([&] {
    if (!d) throw ArgError{"null argument"};
    std::lock_guard<std::mutex> lk(d->stats_mu);
    d->kstats = {};
    return G16HIP_OK;
  });
}

static void ntt_on_device(g16hip_device* d, void* data, int L, int kind) {
  NttTables& t = ntt_tables(d, L);
  size_t n = (size_t)1 << L;
  hipStream_t s = d->stream;
  Fr* sc = t.scalars.as<Fr>();
  if (kind == 0) {
    // natural -> natural: permute, then the bit-reversed-input DIT.
    d->scratch_c.alloc(32 * n);
    launch_bitrev_permute(d->scratch_c.p, data, L, s);
    G16_CHECK(hipMemcpyAsync(data, d->scratch_c.p, 32 * n, hipMemcpyDeviceToDevice, s));
    ntt_dit_forward(data, 1, L, t.fwd.p, nullptr, s);
  } else if (kind == 1) {
    ntt_dif_inverse(data, 1, L, t.inv.p, s);
    d->scratch_c.alloc(32 * n);
    launch_bitrev_permute(d->scratch_c.p, data, L, s);
    G16_CHECK(hipMemcpyAsync(data, d->scratch_c.p, 32 * n, hipMemcpyDeviceToDevice, s));
    launch_scale(data, n, sc + 3, s);
  } else {
    ntt_dif_inverse(data, 1, L, t.inv.p, s);
    ntt_dit_forward(data, 1, L, t.fwd.p, t.coset.p, s);
  }
}

g16hip_status g16hip_ntt(g16hip_device* d, uint8_t* data, int log_n, int kind) {
  return guarded([&] {
    if (!d || !data) throw ArgError{"null argument"};
    if (log_n < 0 || log_n > 27 || kind < 0 || kind > 2) throw ArgError{"bad NTT size or kind"};
    std::lock_guard<std::mutex> lk(d->mu);
    set_device(d);
    size_t n = (size_t)1 << log_n;
    upload(d->scratch_a, data, 32 * n, d->stream);
    ntt_on_device(d, d->scratch_a.p, log_n, kind);
    G16_CHECK(hipMemcpyAsync(data, d->scratch_a.p, 32 * n, hipMemcpyDeviceToHost, d->stream));
    G16_CHECK(hipStreamSynchronize(d->stream));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_ntt_device(g16hip_device* d, void* data, int log_n, int kind) {
  return guarded([&] {
    if (!d || !data) throw ArgError{"null argument"};
    if (log_n < 0 || log_n > 27 || kind < 0 || kind > 2) throw ArgError{"bad NTT size or kind"};
    std::lock_guard<std::mutex> lk(d->mu);
    set_device(d);
    ntt_on_device(d, data, log_n, kind);
    G16_CHECK(hipStreamSynchronize(d->stream));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_prepare(g16hip_device* d, const g16hip_key* key, g16hip_circuit** out) {
  return guarded([&] {
    if (!d || !out) throw ArgError{"null argument"};
    validate_key(key);
    set_device(d);
    auto C = std::make_unique<g16hip_circuit>();
    C->dev = d;
    G16_CHECK(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
    for (auto& st : C->side) G16_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (auto& e : C->ev) G16_CHECK(hipEventCreate(&e));
    for (auto& e : C->sync_ev) G16_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : C->done_ev) G16_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& row : C->mev)
      for (auto& e : row) G16_CHECK(hipEventCreate(&e));
    try {
      prepare_circuit(C.get(), key);
    } catch (...) {
      (void)hipDeviceSynchronize();
      for (auto& e : C->ev) (void)hipEventDestroy(e);
      for (auto& e : C->sync_ev) (void)hipEventDestroy(e);
      for (auto& e : C->done_ev) (void)hipEventDestroy(e);
      if (C->host_wout) (void)hipHostFree(C->host_wout);
      for (auto& row : C->mev)
        for (auto& e : row) (void)hipEventDestroy(e);
      for (auto& st : C->side) (void)hipStreamDestroy(st);
      (void)hipStreamDestroy(C->stream);
      throw;
    }
    *out = C.release();
    return G16HIP_OK;
  });
}

void g16hip_circuit_release(g16hip_circuit* C) {
  if (!C) return;
  (void)hipSetDevice(C->dev->ordinal);
  (void)hipStreamSynchronize(C->stream);
  for (auto& st : C->side) (void)hipStreamSynchronize(st);
  for (auto& e : C->ev) (void)hipEventDestroy(e);
  for (auto& e : C->sync_ev) (void)hipEventDestroy(e);
  for (auto& e : C->done_ev) (void)hipEventDestroy(e);
  if (C->host_wout) (void)hipHostFree(C->host_wout);
  for (auto& row : C->mev)
    for (auto& e : row) (void)hipEventDestroy(e);
  for (auto& st : C->side) (void)hipStreamDestroy(st);
  (void)hipStreamDestroy(C->stream);
  delete C;
}

g16hip_status g16hip_circuit_dims(const g16hip_circuit* C, uint32_t* nv, uint32_t* np,
                                  uint32_t* dom) {
  return guarded([&] {
    if (!C) throw ArgError{"null circuit"};
    if (nv) *nv = C->n_vars;
    if (np) *np = C->n_public;
    if (dom) *dom = C->domain;
    return G16HIP_OK;
  });
}

size_t g16hip_circuit_resident_bytes(const g16hip_circuit* C) { return C ? C->resident : 0; }

g16hip_status g16hip_compute_h(g16hip_circuit* C, const uint8_t* w, size_t n, g16hip_h** h_out,
                               g16hip_timings* t) {
  return guarded([&] {
    if (!C || !w || !h_out) throw ArgError{"null argument"};
    std::lock_guard<std::mutex> lk(C->mu);
    G16_CHECK(hipSetDevice(C->dev->ordinal));
    load_witness(C, w, n, false);
    auto H = std::make_unique<g16hip_h>();
    H->owner = C;
    H->buf.alloc(32 * (size_t)C->domain);
    stages_0_4(C, H->buf.p);
    G16_CHECK(hipStreamSynchronize(C->stream));
    add_stage_timings(C, t);
    *h_out = H.release();
    return G16HIP_OK;
  });
}

g16hip_status g16hip_h_to_host(g16hip_circuit* C, const g16hip_h* h, uint8_t* out) {
  return guarded([&] {
    if (!C || !h || !out) throw ArgError{"null argument"};
    if (h->owner != C) throw ArgError{"H handle came from another circuit's compute_h"};
    std::lock_guard<std::mutex> lk(C->mu);
    G16_CHECK(hipSetDevice(C->dev->ordinal));
    G16_CHECK(hipMemcpy(out, h->buf.p, 32 * (size_t)C->domain, hipMemcpyDeviceToHost));
    return G16HIP_OK;
  });
}

void g16hip_h_release(g16hip_h* h) {
  if (!h) return;
  (void)hipSetDevice(h->owner->dev->ordinal);
  delete h;
}

g16hip_status g16hip_msms(g16hip_circuit* C, const uint8_t* w, size_t n, const g16hip_h* h,
                          g16hip_msm_out* out, g16hip_timings* t) {
  return guarded([&] {
    if (!C || !w || !h || !out) throw ArgError{"null argument"};
    if (h->owner != C)
      throw ArgError{"H is a device handle from another circuit's compute_h, which this circuit cannot read"};
    std::lock_guard<std::mutex> lk(C->mu);
    G16_CHECK(hipSetDevice(C->dev->ordinal));
    load_witness(C, w, n, false);
    run_stages(C, false, h->buf.p, out, t);
    return G16HIP_OK;
  });
}

g16hip_status g16hip_h_and_msms(g16hip_circuit* C, const uint8_t* w, size_t n,
                                g16hip_msm_out* out, g16hip_timings* t) {
  return guarded([&] {
    if (!C || !w || !out) throw ArgError{"null argument"};
    std::lock_guard<std::mutex> lk(C->mu);
    G16_CHECK(hipSetDevice(C->dev->ordinal));
    h_and_msms_locked(C, w, n, false, out, t);
    return G16HIP_OK;
  });
}

g16hip_status g16hip_h_and_msms_device(g16hip_circuit* C, const void* w, size_t n,
                                       g16hip_msm_out* out, g16hip_timings* t) {
  return guarded([&] {
    if (!C || !w || !out) throw ArgError{"null argument"};
    std::lock_guard<std::mutex> lk(C->mu);
    G16_CHECK(hipSetDevice(C->dev->ordinal));
    h_and_msms_locked(C, w, n, true, out, t);
    return G16HIP_OK;
  });
}

g16hip_status g16hip_prove_with_blinders(g16hip_circuit* C, const uint8_t* w, size_t n,
                                         const uint8_t* r, const uint8_t* s, g16hip_proof* proof,
                                         g16hip_timings* t) {
  if (!w) return fail(G16HIP_BACKEND, "null witness");
  return prove_impl(C, w, n, false, r, s, proof, t);
}

g16hip_status g16hip_prove_with_blinders_device(g16hip_circuit* C, const void* w, size_t n,
                                                const uint8_t* r, const uint8_t* s,
                                                g16hip_proof* proof, g16hip_timings* t) {
  if (!w) return fail(G16HIP_BACKEND, "null witness");
  return prove_impl(C, w, n, true, r, s, proof, t);
}

}  // extern "C"
