// The C ABI of include/g16hip.h, part 1: errors, devices and the standalone entry points
// (field probe, fixed base, MSMs, NTTs, statistics). Prepared circuits and proofs are in
// prove.hip.
#include <cstdio>
#include <cstdlib>
#include <deque>

#include "runtime_impl.h"

namespace g16 {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    // HIP keeps the last error a runtime call returned and hipGetLastError reports it; read it
    // here so that a failure this library caught (an allocation refused, then handled) is not
    // reported again by the next launch check on this thread.
    (void)hipGetLastError();
    throw HipError{std::string(what) + ": " + hipGetErrorString(e), e};
  }
}

static thread_local std::string t_last_error;

void set_last_error(const std::string& msg) { t_last_error = msg; }

// The host thread pool. It lives as long as the process (never destroyed, so no thread waits
// on a destroyed mutex at exit); each thread takes queued tasks until it has none, then waits.
namespace {
struct HostPool {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::function<void()>> tasks;
  size_t idle = 0;

  void loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      while (tasks.empty()) {
        ++idle;
        cv.wait(lk);
        --idle;
      }
      std::function<void()> fn = std::move(tasks.front());
      tasks.pop_front();
      lk.unlock();
      fn();
      lk.lock();
    }
  }
};
HostPool& host_pool() {
  static HostPool* pool = new HostPool;
  return *pool;
}
}  // namespace

void host_pool_run(std::function<void()> fn) {
  HostPool& p = host_pool();
  std::unique_lock<std::mutex> lk(p.mu);
  p.tasks.push_back(std::move(fn));
  // Idle threads take queued tasks in turn; when there are fewer of them than tasks queued,
  // one more thread starts, so a task never waits behind a running one.
  if (p.idle >= p.tasks.size()) {
    p.cv.notify_one();
    return;
  }
  lk.unlock();
  std::thread([&p] { p.loop(); }).detach();
}

void build_ntt_tables(NttTables& t, int L, hipStream_t s) {
  size_t n = (size_t)1 << L;
  size_t half = n / 2 ? n / 2 : 1;
  Fr host[8];
  host[0] = domain_root(L);
  host[1] = inv(host[0]);
  host[2] = domain_root(L + 1);  // coset shift: primitive 2n-th root (snarkjs' odd coset)
  host[3] = inv(fr_from_u64(n));
  host[4] = fp_one<FrCfg>();
  host[5] = fp_zero<FrCfg>();
  host[6] = inv(host[2]);
  host[7] = neg(inv(fr_from_u64(2 * n)));
  upload(t.scalars, host, sizeof host, s);
  t.fwd.alloc(32 * (n > 1 ? n - 1 : 1));
  t.inv.alloc(32 * (n > 1 ? n - 1 : 1));
  t.coset.alloc(32 * n);
  Fr* sc = t.scalars.as<Fr>();
  DevBuf powers;
  powers.alloc(32 * half);
  launch_powers(powers.p, sc + 0, half, s);
  launch_stage_tables(t.fwd.p, powers.p, L, s);
  launch_powers(powers.p, sc + 1, half, s);
  launch_stage_tables(t.inv.p, powers.p, L, s);
  launch_coset_table(t.coset.p, sc + 2, sc + 3, L, s);
  launch_fr_to_rp(t.coset.p, n, s);
  G16_CHECK(hipStreamSynchronize(s));  // before `powers` is freed
}

// H's odd-coset evaluations to gnark's basis (crates/cuda/src/kernels/rows.cu:37-68): the
// unscaled DIF inverse leaves n g_bitrev(p) at position p, where g are the coefficients of
// V(shift x), V = A*B - C = t H, and t = x^n - 1 = -2 on the odd coset; so H's coefficient
// bitrev(p) is that times shift^-bitrev(p) / n * (-1/2), which this table holds (R' form).
void build_gnark_h_table(NttTables& t, int L, hipStream_t s) {
  size_t n = (size_t)1 << L;
  t.gnark.alloc(32 * n);
  Fr* sc = t.scalars.as<Fr>();
  launch_coset_table(t.gnark.p, sc + 6, sc + 7, L, s);
  launch_fr_to_rp(t.gnark.p, n, s);
}

NttTables& ntt_tables_for(g16hip_device* d, int L) {
  auto it = d->ntt.find(L);
  if (it != d->ntt.end()) return *it->second;
  auto t = std::make_unique<NttTables>();
  build_ntt_tables(*t, L, d->stream);
  NttTables& ref = *t;
  d->ntt[L] = std::move(t);
  return ref;
}

static size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }

MsmScratchSizes msm_scratch_sizes(int group, uint32_t n, int c, bool constant) {
  MsmPlan plan;
  static char anchor;  // the plan is bound only to size its pieces
  msm_plan_bind(plan, &anchor, n, c, false, constant);
  const size_t pt = group == 1 ? 128 : 256;  // XYZZ point
  MsmScratchSizes z;
  z.sort = round256(msm_sort_scratch_bytes(n, c));
  z.accum = round256(msm_accum_scratch_bytes(plan, group) + pt * msm_out_points(plan));
  z.bases = round256(32 * (size_t)n * (group == 1 ? 2 : 4));
  return z;
}

int msm_width(uint32_t n, int c) {
  if (c <= 0) c = msm_default_c(n);
  if (c < 2 || c > 16) throw ArgError{"window width " + std::to_string(c) + " is outside [2, 16]"};
  return c;
}

// One MSM on device-resident bases and scalars with the given scratch (d's stream, events,
// statistics).
template <class F>
static Xyzz<F> run_msm_on(g16hip_device* d, int group, const void* bases, const void* scalars,
                          uint32_t n, int c, const MsmScratch& sc) {
  MsmPlan plan;
  msm_plan_bind(plan, sc.sort, n, c, false, d->constant_work);
  const size_t npts = msm_out_points(plan);
  const size_t coords = (size_t)n * (group == 1 ? 2 : 4);
  G16_CHECK(hipEventRecord(d->ev[0], d->stream));
  msm_to_msm_form(bases, sc.bases, coords, d->stream);
  msm_sort(plan, scalars, d->stream);
  char* wout = static_cast<char*>(sc.accum) + msm_accum_scratch_bytes(plan, group);
  msm_accumulate(plan, group, sc.bases, sc.accum, wout, d->stream, d->ev + 1);
  std::vector<Xyzz<F>> S(npts);
  uint32_t entries = 0;
  G16_CHECK(hipMemcpyAsync(S.data(), wout, sizeof(Xyzz<F>) * npts, hipMemcpyDeviceToHost,
                           d->stream));
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
  return msm_combine(S.data(), plan);
}

// The same with scratch the device owns and grows as needed.
template <class F>
static Xyzz<F> run_msm(g16hip_device* d, int group, const void* bases, const void* scalars,
                       uint32_t n, int c) {
  c = msm_width(n, c);
  const MsmScratchSizes z = msm_scratch_sizes(group, n, c, d->constant_work);
  d->scratch_d.alloc(z.sort);
  d->scratch_e.alloc(z.accum);
  d->msm_bases.alloc(z.bases);
  return run_msm_on<F>(d, group, bases, scalars, n, c,
                       MsmScratch{d->scratch_d.p, d->scratch_e.p, d->msm_bases.p});
}

// One device MSM on caller-provided scratch, result affine in the wire encoding (the caller
// holds d->mu and has made d's device current).
void msm_device_on_scratch(g16hip_device* d, int group, const void* bases, const void* scalars,
                           uint32_t n, int c, const MsmScratch& sc, uint8_t* out) {
  if (group == 1) g1_store(out, host_to_aff(run_msm_on<Fq>(d, 1, bases, scalars, n, c, sc)));
  else g2_store(out, host_to_aff(run_msm_on<Fq2>(d, 2, bases, scalars, n, c, sc)));
}


// Device bytes one standalone MSM of n points at width c takes: the uploaded bases and
// scalars, the bases in the MSM form, the sort scratch and the accumulation scratch (with
// the fold's slots under constant work).
static size_t msm_run_bytes(int group, uint32_t n, int c, bool constant) {
  size_t pw = group == 1 ? 64 : 128;
  MsmPlan plan;
  static char anchor;  // the plan is bound only to size its pieces
  msm_plan_bind(plan, &anchor, n, c, false, constant);
  return 2 * pw * (size_t)n + 32 * (size_t)n + msm_sort_scratch_bytes(n, c) +
         msm_accum_scratch_bytes(plan, group) + 256 * msm_out_points(plan);
}

// One MSM over host slices: upload, run, result in XYZZ (the caller holds d->mu).
template <class F>
static Xyzz<F> msm_from_host(g16hip_device* d, int group, const uint8_t* bases,
                             const uint8_t* scalars, size_t n) {
  size_t pw = group == 1 ? 64 : 128;
  upload(d->scratch_a, bases, n * pw, d->stream);
  upload(d->scratch_b, scalars, n * 32, d->stream);
  return run_msm<F>(d, group, d->scratch_a.p, d->scratch_b.p, (uint32_t)n, 0);
}

static Fq fq_small(uint32_t k) {
  Fq x = fp_zero<FqCfg>();
  x.v[0] = k;
  return from_std(x);
}

template <class F>
static bool xyzz_on_curve(const Xyzz<F>& p);

template <>
bool xyzz_on_curve<Fq>(const G1X& p) {
  G1Aff a = host_to_aff(p);
  if (aff_is_inf(a)) return true;
  Fq b = fq_small(3);
  return eq(sqr(a.y), add(mul(sqr(a.x), a.x), b));
}

template <>
bool xyzz_on_curve<Fq2>(const G2X& p) {
  G2Aff a = host_to_aff(p);
  if (aff_is_inf(a)) return true;
  // b' = 3 / (9 + u)
  Fq2 xi{fq_small(9), fq_small(1)};
  Fq2 three{fq_small(3), fp_zero<FqCfg>()};
  Fq2 b = mul(three, inv(xi));
  return eq(sqr(a.y), add(mul(sqr(a.x), a.x), b));
}

}  // namespace g16

// The ceremony's MSM seam (crates/cuda/src/ceremony.rs:65-312): MSMs over host slices for
// `groth16 setup` and `zkey verify`'s recompute, cut into runs of at most max_run points.
struct g16hip_ceremony {
  g16hip_device* dev = nullptr;
  size_t max_run = 1u << 21;
  size_t vram_limit = 0;  // 0: whatever hipMalloc grants
  std::mutex one;         // one call on the device at a time
  uint64_t last_runs = 0;
  g16hip_ceremony_stats tally = {};
};

namespace g16 {

constexpr size_t CEREMONY_MIN_RUN = 1u << 10;
constexpr int CEREMONY_ATTEMPTS = 8;

// One run, re-run while its sum is off the curve (CEREMONY_ATTEMPTS in all); a memory
// refusal comes straight out, before anything is allocated when a limit is set.
template <class F>
static Xyzz<F> ceremony_run(g16hip_ceremony* cer, int group, const uint8_t* bases,
                            const uint8_t* scalars, size_t n) {
  g16hip_device* d = cer->dev;
  if (cer->vram_limit) {
    size_t need = msm_run_bytes(group, (uint32_t)n, msm_default_c((uint32_t)n), d->constant_work);
    if (need > cer->vram_limit)
      throw StatusError{G16HIP_DEVICE_MEMORY,
                        "a run of " + std::to_string(n) + " points needs " + std::to_string(need) +
                            " bytes, over the limit of " + std::to_string(cer->vram_limit)};
  }
  for (int attempt = 1;; ++attempt) {
    Xyzz<F> p;
    {
      std::lock_guard<std::mutex> lk(d->mu);
      set_device(d);
      p = msm_from_host<F>(d, group, bases, scalars, n);
    }
    if (xyzz_on_curve(p)) return p;
    cer->tally.retries += 1;
    if (attempt >= CEREMONY_ATTEMPTS)
      throw StatusError{G16HIP_DEVICE, "ceremony MSM: the sum is not a point on the curve after " +
                                           std::to_string(attempt) + " attempts"};
    std::fprintf(stderr, "g16hip: ceremony MSM attempt %d of %d gave a sum off the curve; re-running\n",
                 attempt, CEREMONY_ATTEMPTS);
  }
}

template <class F>
static Xyzz<F> ceremony_msm(g16hip_ceremony* cer, int group, const uint8_t* bases,
                            const uint8_t* scalars, size_t n) {
  size_t pw = group == 1 ? 64 : 128;
  Timer t;
  Xyzz<F> sum = xyzz_inf<F>();
  size_t at = 0, len = std::max(cer->max_run, CEREMONY_MIN_RUN);
  uint64_t runs = 0;
  while (at < n) {
    size_t m = std::min(len, n - at);
    try {
      Xyzz<F> p = ceremony_run<F>(cer, group, bases + pw * at, scalars + 32 * at, m);
      sum = host_add(sum, p);
      at += m;
      ++runs;
    } catch (const StatusError& e) {
      if (e.st != G16HIP_DEVICE_MEMORY || m <= CEREMONY_MIN_RUN) throw;
      len = std::max(m / 2, CEREMONY_MIN_RUN);
    } catch (const HipError& e) {
      if (e.code != hipErrorOutOfMemory || m <= CEREMONY_MIN_RUN) throw;
      len = std::max(m / 2, CEREMONY_MIN_RUN);
    }
  }
  cer->last_runs = runs;
  cer->tally.calls += 1;
  cer->tally.runs += runs;
  cer->tally.points += n;
  cer->tally.ms += t.ms();
  return sum;
}

}  // namespace g16

using namespace g16;

extern "C" {

g16hip_status g16hip_ceremony_open(g16hip_device* d, size_t max_run, size_t vram_limit,
                                   g16hip_ceremony** out) {
  return guarded([&] {
    if (!d || !out) throw ArgError{"null argument"};
    auto* c = new g16hip_ceremony();
    c->dev = d;
    if (max_run) c->max_run = max_run;
    c->vram_limit = vram_limit;
    *out = c;
    return G16HIP_OK;
  });
}

void g16hip_ceremony_close(g16hip_ceremony* c) { delete c; }

g16hip_status g16hip_ceremony_msm(g16hip_ceremony* c, int group, const uint8_t* bases,
                                  const uint8_t* scalars, size_t n, uint8_t* out) {
  return guarded([&] {
    if (!c || !out || (n && (!bases || !scalars))) throw ArgError{"null argument"};
    if (group != 1 && group != 2) throw ArgError{"group must be 1 or 2"};
    size_t pw = group == 1 ? 64 : 128;
    std::lock_guard<std::mutex> lk(c->one);
    if (n == 0) {
      c->last_runs = 0;
      std::memset(out, 0, pw);
      return G16HIP_OK;
    }
    if (group == 1) g1_store(out, host_to_aff(ceremony_msm<Fq>(c, 1, bases, scalars, n)));
    else g2_store(out, host_to_aff(ceremony_msm<Fq2>(c, 2, bases, scalars, n)));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_set_constant_work(g16hip_device* d, int on) {
  return guarded([&] {
    if (!d) throw ArgError{"null device"};
    std::lock_guard<std::mutex> lk(d->mu);
    d->constant_work = on != 0;
    return G16HIP_OK;
  });
}

g16hip_status g16hip_msm_scratch_bytes(int group, size_t n, size_t* bytes) {
  return guarded([&] {
    if (!bytes || (group != 1 && group != 2)) throw ArgError{"bad argument"};
    if (n >= (1u << 31)) throw ArgError{"MSM larger than 2^31 points"};
    *bytes = n ? msm_run_bytes(group, (uint32_t)n, msm_default_c((uint32_t)n), false) : 0;
    return G16HIP_OK;
  });
}

g16hip_status g16hip_ceremony_last_runs(g16hip_ceremony* c, uint64_t* runs) {
  return guarded([&] {
    if (!c || !runs) throw ArgError{"null argument"};
    std::lock_guard<std::mutex> lk(c->one);
    *runs = c->last_runs;
    return G16HIP_OK;
  });
}

g16hip_status g16hip_ceremony_stats_get(g16hip_ceremony* c, g16hip_ceremony_stats* out) {
  return guarded([&] {
    if (!c || !out) throw ArgError{"null argument"};
    std::lock_guard<std::mutex> lk(c->one);
    *out = c->tally;
    return G16HIP_OK;
  });
}

const char* g16hip_last_error(void) { return t_last_error.c_str(); }

g16hip_status g16hip_open(int ordinal, g16hip_device** out) {
  return guarded([&] {
    if (!out) throw ArgError{"null output handle"};
    int n = 0;
    G16_CHECK(hipGetDeviceCount(&n));
    if (ordinal < 0 || ordinal >= n)
      throw ArgError{"device ordinal " + std::to_string(ordinal) + " out of range (" +
                     std::to_string(n) + " devices)"};
    auto bytes_env = [](const char* name) -> size_t {
      const char* v = std::getenv(name);
      if (!v || !*v) return 0;
      char* end = nullptr;
      unsigned long long x = std::strtoull(v, &end, 10);
      if (*end) throw ArgError{std::string(name) + "=" + v + " is not a byte count"};
      return (size_t)x;
    };
    const size_t limit = bytes_env("G16HIP_VRAM_LIMIT"), reserve = bytes_env("G16HIP_VRAM_RESERVE");
    auto* d = new g16hip_device();
    d->ordinal = ordinal;
    d->vram_limit = limit;
    d->vram_reserve = reserve;
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
  d->msm_bases.release();
  d->scratch_d.release();
  d->scratch_e.release();
  for (auto& e : d->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(d->stream);
  delete d;
}

g16hip_status g16hip_mem_info(g16hip_device* d, size_t* free_b, size_t* total_b) {
  return guarded([&] {
    if (!d || !free_b || !total_b) throw ArgError{"null argument"};
    set_device(d);
    G16_CHECK(hipMemGetInfo(free_b, total_b));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_set_vram_limit(g16hip_device* d, size_t bytes) {
  return guarded([&] {
    if (!d) throw ArgError{"null device"};
    d->vram_limit = bytes;
    return G16HIP_OK;
  });
}

g16hip_status g16hip_device_budget(g16hip_device* d, size_t* bytes) {
  return guarded([&] {
    if (!d || !bytes) throw ArgError{"null argument"};
    set_device(d);
    *bytes = device_budget(d);
    return G16HIP_OK;
  });
}

g16hip_status g16hip_field_probe(g16hip_device* d, int field, int op, const uint8_t* a,
                                 const uint8_t* b, uint8_t* out, size_t n) {
  return guarded([&] {
    if (field < 0 || field > 3 || op < 0 || op > 7) throw ArgError{"bad field/op"};
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
    if (group == 1) g1_store(out, host_to_aff(msm_from_host<Fq>(d, 1, bases, scalars, n)));
    else g2_store(out, host_to_aff(msm_from_host<Fq2>(d, 2, bases, scalars, n)));
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
      g1_store(out, host_to_aff(run_msm<Fq>(d, 1, bases, scalars, (uint32_t)n, c)));
    else
      g2_store(out, host_to_aff(run_msm<Fq2>(d, 2, bases, scalars, (uint32_t)n, c)));
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
  NttTables& t = ntt_tables_for(d, L);
  size_t n = (size_t)1 << L;
  hipStream_t s = d->stream;
  Fr* sc = t.scalars.as<Fr>();
  if (kind == 0) {
    // natural -> natural: the bit-reversed-input DIT reading `data` at bit-reversed indices
    // (its first pass out of place into scratch, the next back into `data`).
    d->scratch_c.alloc(32 * n);
    ntt_dit_forward(data, 1, L, t.fwd.p, nullptr, s, data, d->scratch_c.p);
  } else if (kind == 1) {
    ntt_dif_inverse(data, 1, L, t.inv.p, s);
    d->scratch_c.alloc(32 * n);
    launch_bitrev_permute(d->scratch_c.p, data, L, s);
    G16_CHECK(hipMemcpyAsync(data, d->scratch_c.p, 32 * n, hipMemcpyDeviceToDevice, s));
    launch_scale(data, n, sc + 3, s);
  } else {
    ntt_coset_pipeline(data, 1, L, t.inv.p, t.fwd.p, t.coset.p, s);
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


}  // extern "C"
