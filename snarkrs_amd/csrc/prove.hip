// The C ABI of include/g16hip.h, part 2: prepared circuits (Backend::prepare /
// PreparedCircuit, crates/groth16/src/lib.rs:371-455), the prove pipeline (stages 0-9 on the
// GPU, stage 11 on the host, crates/groth16/src/prove.rs:175-229, :391-430), batched proving
// and the serving loop.
//
// A circuit owns its resident key and one or more proof slots. A slot is everything one
// proof in flight needs: three streams, events, and the per-proof device scratch (witness,
// A/B/C, H, sort and accumulation buffers, window sums) plus pinned host memory for the
// window sums. Single proofs run on slot 0 with per-stage events (timings, host work started
// per MSM); batches and the server run many slots at once, each replaying its proof pipeline
// captured once as a hipGraph.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <deque>
#include <unordered_map>

#include "runtime_impl.h"
#include "verify.h"

using namespace g16;

namespace g16 {

struct Slot {
  hipStream_t stream = nullptr;
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t ev[8] = {};
  hipEvent_t sync_ev[4] = {};  // witness ready, witness sorted, side[0] done, side[1] done
  hipEvent_t done_ev[5] = {};  // reduction rows of A, B1, L, B2, H written to host_wout
  hipEvent_t mev[5][4] = {};   // per MSM: before accumulation, after it, after merges, after reduction
  DevBuf witness, abc, hbuf, sort_w, sort_h, accum_w1, accum_w2, accum_h;
  MsmPlan plan_w, plan_h;
  // Pinned, device-mapped host memory: the reduction kernels write each MSM's rows straight
  // here (no copy on the stream, which would wait for CUs behind other streams' kernels).
  uint8_t* host_wout = nullptr;
  uint8_t* host_wout_dev = nullptr;
  // Page-locked staging for a host witness: the caller's (pageable) witness is copied here in
  // chunks and each chunk goes up by DMA as soon as it is in (the reference's pinned witness
  // staging, crates/cuda/src/stages.rs:946-985).
  uint8_t* host_witness = nullptr;
  size_t host_witness_bytes = 0;
  double witness_host_ms = -1.0;  // the last host witness's staging time (timeline only)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t bytes = 0;

  ~Slot() {
    for (hipStream_t st : {stream, side[0], side[1]})
      if (st) (void)hipStreamSynchronize(st);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : sync_ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : done_ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& row : mev)
      for (auto& e : row)
        if (e) (void)hipEventDestroy(e);
    if (host_wout) (void)hipHostFree(host_wout);
    if (host_witness) (void)hipHostFree(host_witness);
    for (hipStream_t st : side)
      if (st && st != stream) (void)hipStreamDestroy(st);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// A wave: up to `cap` proofs of one circuit proved together. Every stage is one launch over all
// of them (the fused gather over cap witnesses, the transforms of 3 cap vectors, each MSM's
// sort, accumulation, merge and reduction over all cap scalar vectors with a bucket set per
// proof), on one stream, so a batch keeps one hardware queue per circuit. Host buffers are
// double-buffered: the host finishes wave i (Horner passes, blinders, assembly) while the
// device runs wave i + 1.
struct Wave {
  uint32_t cap = 0;
  hipStream_t stream = nullptr;
  hipEvent_t done[2] = {};
  DevBuf witness, abc, hbuf, sort_w, sort_h;
  DevBuf accum_g1[4], accum_w2;    // A, B1, L, H (one accumulation queue), B2
  DevBuf queue;                    // the accumulation queue's counter
  uint8_t* host_wout[2] = {};      // reduction rows, pinned and device-mapped
  uint8_t* host_wout_dev[2] = {};
  uint8_t* staging[2] = {};        // page-locked witnesses, cap * 32 n_vars bytes each
  size_t wout_bytes = 0, bytes = 0;
  // The device part of a wave of m proofs on buffer b, captured once and replayed (graph
  // replay on): key 2 m + b.
  std::map<uint32_t, std::pair<hipGraph_t, hipGraphExec_t>> graphs;

  ~Wave() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& g : graphs) {
      if (g.second.second) (void)hipGraphExecDestroy(g.second.second);
      if (g.second.first) (void)hipGraphDestroy(g.second.first);
    }
    for (auto& e : done)
      if (e) (void)hipEventDestroy(e);
    for (int b = 0; b < 2; ++b) {
      if (host_wout[b]) (void)hipHostFree(host_wout[b]);
      if (staging[b]) (void)hipHostFree(staging[b]);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace g16

struct g16hip_circuit {
  g16hip_device* dev = nullptr;
  uint32_t n_vars = 0, n_public = 0, domain = 0;
  int log_n = 0;
  int cw = 0, ch = 0;  // window widths of the witness and H MSMs
  bool merged = false; // bases stored as window tables (one bucket set per MSM)
  G1Aff alpha_g1, beta_g1, delta_g1;
  G2Aff beta_g2, delta_g2;
  bool has_matrices = true;  // false: proves from A.w and B.w rows only
  uint32_t h_basis = 0;      // 0 snarkjs' odd coset, 1 gnark's bit-reversed coefficients
  uint32_t n_h = 0;          // H bases: domain_size, or domain_size - 1 for gnark's
  bool constant = false;     // constant-work MSMs (the device's setting at prepare)
  VkBytes vk;  // the verifying key, for the check of a proof's result
  DevBuf row_ptr[2], signal[2], value_id[2], values;
  DevBuf a_q, b1_q, b2_q, l_q, h_q;
  NttTables tables;
  size_t key_bytes = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::vector<std::unique_ptr<Slot>> slots;
  std::vector<bool> busy;
  uint32_t growing = 0;  // slots being added outside the lock
  // Batched proving: the wave (built on first use, rebuilt for another cap) and the proofs per
  // wave g16hip_prove_batch uses (0: the default; g16hip_batch_autotune measures one).
  std::mutex wave_mu;
  std::unique_ptr<g16::Wave> wave;
  std::atomic<uint32_t> wave_cap{0};
};

struct g16hip_h {
  g16hip_circuit* owner = nullptr;
  DevBuf buf;
};

namespace g16 {

// ---------------------------------------------------------------------------------------
// key validation (crates/groth16/src/cpu.rs:73-122) and upload
// ---------------------------------------------------------------------------------------
static void validate_key(const g16hip_key* k) {
  if (!k) throw ArgError{"null key"};
  uint32_t n = k->domain_size;
  if (n == 0 || (n & (n - 1))) throw ArgError{"domain size " + std::to_string(n) + " is not a power of two"};
  if (n > (1u << 27)) throw ArgError{"domain size exceeds 2^27"};
  if ((uint64_t)k->n_public + 1 > k->n_vars)
    throw ArgError{"n_public " + std::to_string(k->n_public) + " exceeds n_vars " + std::to_string(k->n_vars)};
  if (k->h_basis > 1) throw ArgError{"unknown H basis " + std::to_string(k->h_basis)};
  if (k->h_basis == 1 && n < 2) throw ArgError{"a gnark-basis key needs a domain of at least 2"};
  if (!k->a_query || !k->b_g1_query || !k->b_g2_query || !k->h_query ||
      (!k->l_query && k->n_vars > k->n_public + 1))
    throw ArgError{"missing query points"};
  const bool matrices = k->row_ptr[0] || k->row_ptr[1];
  for (int m = 0; m < 2 && matrices; ++m) {
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
  if (!k->gamma_g2 || !k->ic) throw ArgError{"the key has no verifying-key points (gamma_g2, ic)"};
}

// Window widths, window tables, and the device bytes the key and one proof slot take, fixed
// before anything is allocated (the reference's resident_bytes check at prepare,
// crates/cuda/src/backend.rs:367-381, :626-643).
struct KeyPlan {
  int cw = 0, ch = 0;
  bool merged = false;
  size_t key_bytes = 0, slot_bytes = 0;
  size_t total() const { return key_bytes + slot_bytes; }
};

static size_t slot_bytes_for(uint32_t nv, uint32_t n, uint32_t nh, int cw, int ch, bool merged,
                             bool constant);

static KeyPlan plan_key(const g16hip_key* k, bool merged, bool constant) {
  const size_t n = k->domain_size, nv = k->n_vars, nh = k->h_basis == 1 ? n - 1 : n;
  KeyPlan P;
  P.merged = merged;
  // The witness digits feed A, B1, L (G1) and B2 (G2); H feeds one G1 MSM.
  // G16HIP_C_WITNESS / G16HIP_C_H override the widths (2..16, and H up to 20 over window
  // tables).
  P.cw = merged ? msm_choose_c_merged((uint32_t)nv, 3, 1) : msm_choose_c((uint32_t)nv, 3, 1);
  P.ch = merged ? msm_choose_c_merged((uint32_t)nh, 1, 0) : msm_default_c((uint32_t)nh);
  if (const char* e = std::getenv("G16HIP_C_WITNESS")) P.cw = std::max(2, std::min(16, std::atoi(e)));
  if (const char* e = std::getenv("G16HIP_C_H"))
    P.ch = std::max(2, std::min(merged ? MSM_MAX_C : 16, std::atoi(e)));
  const size_t tw = merged ? (255 + P.cw - 1) / P.cw : 1, th = merged ? (255 + P.ch - 1) / P.ch : 1;
  size_t b = 0;
  if (k->row_ptr[0]) {
    for (int m = 0; m < 2; ++m) {
      const size_t nnz = k->row_ptr[m][n];
      b += dev_pages(4 * (n + 1)) + 2 * dev_pages(4 * nnz);
    }
    b += dev_pages(32 * (size_t)k->n_values);
  }
  b += 3 * dev_pages(64 * nv * tw) + dev_pages(128 * nv * tw) + dev_pages(64 * nh * th);
  // NTT tables: two twiddle tables, the coset table, the scalars, gnark's table, and the
  // powers built on the way.
  b += 2 * dev_pages(32 * (n > 1 ? n - 1 : 1)) + dev_pages(32 * n) + dev_pages(256) +
       (k->h_basis == 1 ? dev_pages(32 * n) : 0) + dev_pages(32 * (n / 2 ? n / 2 : 1));
  P.key_bytes = b;
  P.slot_bytes = slot_bytes_for((uint32_t)nv, (uint32_t)n, (uint32_t)nh, P.cw, P.ch, merged,
                                constant);
  return P;
}

// Window tables take W times the bases' memory; they are used when the device keeps 8 GiB of
// its budget free after the key and a slot (G16HIP_WINDOW_TABLES=0 turns them off). A key
// that does not fit even without them is refused before anything is allocated.
static KeyPlan choose_plan(const g16hip_device* d, const g16hip_key* k) {
  const size_t budget = device_budget(d);
  const char* wt = std::getenv("G16HIP_WINDOW_TABLES");
  if (!(wt && wt[0] == '0')) {
    KeyPlan P = plan_key(k, true, d->constant_work);
    if (P.total() + (8ull << 30) <= budget) return P;
  }
  KeyPlan P = plan_key(k, false, d->constant_work);
  if (P.total() > budget)
    throw too_big(P.total(), budget, "this key's resident data and one proof's stage scratch");
  return P;
}

static void upload_key(g16hip_circuit* C, const g16hip_key* k, const KeyPlan& plan, hipStream_t s) {
  uint32_t n = k->domain_size;
  C->n_vars = k->n_vars;
  C->n_public = k->n_public;
  C->domain = n;
  C->log_n = 0;
  while ((1u << C->log_n) < n) ++C->log_n;
  C->has_matrices = k->row_ptr[0] != nullptr;
  C->h_basis = k->h_basis;
  C->n_h = k->h_basis == 1 ? n - 1 : n;
  C->alpha_g1 = g1_load(k->alpha_g1);
  C->beta_g1 = g1_load(k->beta_g1);
  C->delta_g1 = g1_load(k->delta_g1);
  C->beta_g2 = g2_load(k->beta_g2);
  C->delta_g2 = g2_load(k->delta_g2);
  std::memcpy(C->vk.alpha_g1, k->alpha_g1, 64);
  std::memcpy(C->vk.beta_g2, k->beta_g2, 128);
  std::memcpy(C->vk.gamma_g2, k->gamma_g2, 128);
  std::memcpy(C->vk.delta_g2, k->delta_g2, 128);
  C->vk.ic.assign(k->ic, k->ic + 64 * ((size_t)k->n_public + 1));
  for (int m = 0; m < 2 && C->has_matrices; ++m) {
    uint32_t nnz = k->row_ptr[m][n];
    upload(C->row_ptr[m], k->row_ptr[m], 4 * (size_t)(n + 1), s);
    upload(C->signal[m], k->signal[m], 4 * (size_t)nnz, s);
    upload(C->value_id[m], k->value_id[m], 4 * (size_t)nnz, s);
  }
  if (C->has_matrices) upload(C->values, k->values, 32 * (size_t)k->n_values, s);
  const size_t nh = C->n_h;
  size_t nv = k->n_vars;
  // Window tables: with entry w n + i = 2^(c w) P_i precomputed (the key's bases are fixed),
  // an MSM accumulates the digits of every window into one set of buckets, so the merges and
  // the window reduction run once instead of once per window.
  C->cw = plan.cw;
  C->ch = plan.ch;
  C->merged = plan.merged;
  const int Ww = (255 + C->cw - 1) / C->cw, Wh = (255 + C->ch - 1) / C->ch;
  const size_t tw = C->merged ? (size_t)Ww : 1, th = C->merged ? (size_t)Wh : 1;
  auto put = [&](DevBuf& b, const void* src, size_t npts, size_t pt, size_t windows) {
    b.alloc(pt * npts * windows);
    if (npts) G16_CHECK(hipMemcpyAsync(b.p, src, pt * npts, hipMemcpyHostToDevice, s));
  };
  put(C->a_q, k->a_query, nv, 64, tw);
  put(C->b1_q, k->b_g1_query, nv, 64, tw);
  put(C->b2_q, k->b_g2_query, nv, 128, tw);
  // L covers the private wires only; it is stored behind n_public + 1 points at infinity so
  // that it multiplies the full witness and shares the witness sort with A, B1 and B2.
  size_t pub = (size_t)k->n_public + 1;
  C->l_q.alloc(64 * nv * tw);
  G16_CHECK(hipMemsetAsync(C->l_q.p, 0, 64 * pub, s));
  if (nv > pub)
    G16_CHECK(hipMemcpyAsync(static_cast<char*>(C->l_q.p) + 64 * pub, k->l_query, 64 * (nv - pub),
                             hipMemcpyHostToDevice, s));
  put(C->h_q, k->h_query, nh, 64, th);
  // The bases serve only the MSMs: keep them in the MSM kernels' R' = 2^261 form.
  msm_to_msm_form(C->a_q.p, C->a_q.p, 2 * nv, s);
  msm_to_msm_form(C->b1_q.p, C->b1_q.p, 2 * nv, s);
  msm_to_msm_form(C->l_q.p, C->l_q.p, 2 * nv, s);
  msm_to_msm_form(C->h_q.p, C->h_q.p, 2 * nh, s);
  msm_to_msm_form(C->b2_q.p, C->b2_q.p, 4 * nv, s);
  if (C->merged) {
    msm_window_tables(1, C->a_q.p, nv, C->cw, Ww, s);
    msm_window_tables(1, C->b1_q.p, nv, C->cw, Ww, s);
    msm_window_tables(1, C->l_q.p, nv, C->cw, Ww, s);
    msm_window_tables(2, C->b2_q.p, nv, C->cw, Ww, s);
    msm_window_tables(1, C->h_q.p, nh, C->ch, Wh, s);
  }
  build_ntt_tables(C->tables, C->log_n, s);
  if (C->h_basis == 1) build_gnark_h_table(C->tables, C->log_n, s);
  G16_CHECK(hipStreamSynchronize(s));
  DevBuf* bufs[] = {&C->row_ptr[0], &C->row_ptr[1], &C->signal[0], &C->signal[1],
                    &C->value_id[0], &C->value_id[1], &C->values, &C->a_q, &C->b1_q, &C->b2_q,
                    &C->l_q, &C->h_q, &C->tables.fwd, &C->tables.inv, &C->tables.coset,
                    &C->tables.gnark};
  C->key_bytes = 0;
  for (auto b : bufs) C->key_bytes += b->bytes;
}

// A slot's streams and events. HIP spreads a process's streams over a few hardware queues
// (GPU_MAX_HW_QUEUES, 4 by default) in creation order, and streams that share a queue run
// one after the other; a slot's three streams are therefore created back to back, before
// anything else of the circuit, so that they land on different queues.
// G16HIP_SERIAL=1 runs a slot's three streams' work on its main stream, one kernel at a time
// (for traces whose kernel durations are each kernel's alone).
static bool serial_slots() {
  static const bool on = std::getenv("G16HIP_SERIAL") != nullptr;
  return on;
}

static std::unique_ptr<Slot> new_slot() {
  auto S = std::make_unique<Slot>();
  G16_CHECK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
  if (serial_slots())
    S->side[0] = S->side[1] = S->stream;
  else
    for (auto& st : S->side) G16_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  for (auto& e : S->ev) G16_CHECK(hipEventCreate(&e));
  for (auto& e : S->sync_ev) G16_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (auto& e : S->done_ev) G16_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (auto& row : S->mev)
    for (auto& e : row) G16_CHECK(hipEventCreate(&e));
  return S;
}

// Per-proof scratch for one slot, sized once so a proof allocates nothing.
static void size_slot(const g16hip_circuit* C, Slot* S) {
  size_t nv = C->n_vars, n = C->domain;
  S->witness.alloc(32 * nv);
  S->abc.alloc(32 * 3 * n);
  S->hbuf.alloc(32 * n);
  S->sort_w.alloc(msm_sort_scratch_bytes((uint32_t)nv, C->cw));
  S->sort_h.alloc(msm_sort_scratch_bytes(C->n_h, C->ch));
  msm_plan_bind(S->plan_w, S->sort_w.p, (uint32_t)nv, C->cw, C->merged, C->constant);
  msm_plan_bind(S->plan_h, S->sort_h.p, C->n_h, C->ch, C->merged, C->constant);
  S->accum_w1.alloc(msm_accum_scratch_bytes(S->plan_w, 1));
  S->accum_w2.alloc(msm_accum_scratch_bytes(S->plan_w, 2));
  S->accum_h.alloc(msm_accum_scratch_bytes(S->plan_h, 1));
  size_t wbytes = 256 * (4 * msm_out_points(S->plan_w) + msm_out_points(S->plan_h));
  G16_CHECK(hipHostMalloc((void**)&S->host_wout, wbytes + 64,
                          hipHostMallocMapped | hipHostMallocCoherent));
  G16_CHECK(hipHostGetDevicePointer((void**)&S->host_wout_dev, S->host_wout, 0));
  S->host_witness_bytes = 32 * nv;
  G16_CHECK(hipHostMalloc((void**)&S->host_witness, S->host_witness_bytes ? S->host_witness_bytes : 32,
                          hipHostMallocDefault));
  DevBuf* bufs[] ={&S->witness, &S->abc, &S->hbuf, &S->sort_w, &S->sort_h, &S->accum_w1,
                    &S->accum_w2, &S->accum_h};
  S->bytes = 0;
  for (auto b : bufs) S->bytes += b->bytes;
}

static std::unique_ptr<Slot> make_slot(const g16hip_circuit* C) {
  auto S = new_slot();
  size_slot(C, S.get());
  return S;
}

// Device bytes of one slot (what size_slot allocates), rounded up to pages.
static size_t slot_bytes_for(uint32_t nv, uint32_t n, uint32_t nh, int cw, int ch, bool merged,
                             bool constant) {
  MsmPlan pw, ph;
  static char anchor;  // the plans are bound only to size their pieces
  msm_plan_bind(pw, &anchor, nv, cw, merged, constant);
  msm_plan_bind(ph, &anchor, nh, ch, merged, constant);
  return dev_pages(32 * (size_t)nv) + dev_pages(32 * 3 * (size_t)n) + dev_pages(32 * (size_t)n) +
         dev_pages(msm_sort_scratch_bytes(nv, cw)) + dev_pages(msm_sort_scratch_bytes(nh, ch)) +
         dev_pages(msm_accum_scratch_bytes(pw, 1)) + dev_pages(msm_accum_scratch_bytes(pw, 2)) +
         dev_pages(msm_accum_scratch_bytes(ph, 1));
}

static size_t slot_bytes_estimate(const g16hip_circuit* C) {
  return slot_bytes_for(C->n_vars, C->domain, C->n_h, C->cw, C->ch, C->merged, C->constant);
}

static bool try_add_slot(g16hip_circuit* C, size_t reserve);

// Exclusive use of one of the circuit's slots. Waits for a free one; with `cap` > 0, first
// adds a slot (up to `cap` in all, while the device keeps `reserve` bytes free) when every
// slot is busy.
struct SlotLease {
  g16hip_circuit* C;
  size_t idx;
  Slot* S;
  explicit SlotLease(g16hip_circuit* c, uint32_t cap = 0, size_t reserve = 0) : C(c) {
    std::unique_lock<std::mutex> lk(C->mu);
    for (;;) {
      for (idx = 0; idx < C->busy.size() && C->busy[idx]; ++idx) {
      }
      if (idx < C->busy.size()) break;
      if (C->slots.size() + C->growing < cap) {
        C->growing++;
        lk.unlock();
        bool added = false;
        try {
          added = try_add_slot(C, reserve);
        } catch (...) {
          lk.lock();
          C->growing--;
          throw;
        }
        lk.lock();
        C->growing--;
        if (added) continue;
        cap = 0;  // no memory for another: wait for a slot to free up
      }
      C->cv.wait(lk);
    }
    C->busy[idx] = true;
    S = C->slots[idx].get();
  }
  ~SlotLease() {
    {
      std::lock_guard<std::mutex> lk(C->mu);
      C->busy[idx] = false;
    }
    C->cv.notify_all();
  }
};

// ---------------------------------------------------------------------------------------
// the pipeline
// ---------------------------------------------------------------------------------------
static void check_witness_len(const g16hip_circuit* C, size_t n) {
  if (n != C->n_vars)
    throw StatusError{G16HIP_WITNESS_LENGTH, "witness has " + std::to_string(n) +
                                                 " entries, proving key expects " +
                                                 std::to_string(C->n_vars)};
}

// A host witness goes up through the slot's page-locked staging: the calling thread and
// helper threads copy it in, in order, in pieces of 512 KiB, and the calling thread enqueues
// each DMA chunk (1 MiB first, then 8 MiB) as soon as all its pieces are in (copying pieces
// itself while it waits), so the host copy overlaps the transfer. Measured on the MI355X box:
// one 38 MB DMA from page-locked memory runs at 57 GB/s, 8 MiB chunks at 54, 1 MiB chunks at
// 38 (each DMA costs ~9 us); eight threads copy into page-locked memory at ~69 GB/s. The
// staging is free at the start of a proof: a slot is leased by one proof at a time, and a
// proof returns only after its streams have finished.
static constexpr size_t kStagingHelpers = 7;
static void stage_witness(Slot* S, const uint8_t* w, size_t bytes) {
  constexpr size_t PIECE = 512u << 10, FIRST = 1u << 20, CHUNK = 8u << 20;
  uint8_t* dst = S->host_witness;
  char* dev = S->witness.as<char>();
  const size_t np = (bytes + PIECE - 1) / PIECE;
  if (np <= 2) {
    std::memcpy(dst, w, bytes);
    G16_CHECK(hipMemcpyAsync(dev, dst, bytes, hipMemcpyHostToDevice, S->stream));
    return;
  }
  std::atomic<size_t> next{0};
  std::unique_ptr<std::atomic<uint8_t>[]> ready(new std::atomic<uint8_t>[np]);
  for (size_t c = 0; c < np; ++c) ready[c].store(0, std::memory_order_relaxed);
  auto copy_one = [&]() -> bool {
    const size_t c = next.fetch_add(1, std::memory_order_relaxed);
    if (c >= np) return false;
    std::memcpy(dst + c * PIECE, w + c * PIECE, std::min(PIECE, bytes - c * PIECE));
    ready[c].store(1, std::memory_order_release);
    return true;
  };
  // Pooled helpers (no thread is started per proof once the pool has them).
  std::vector<std::unique_ptr<HostTask>> helpers;
  const size_t nh = std::min<size_t>(kStagingHelpers, np - 1);
  for (size_t h = 0; h < nh; ++h)
    helpers.push_back(std::make_unique<HostTask>([&] {
      while (copy_one()) {
      }
    }));
  std::exception_ptr err;
  try {
    size_t off = 0, p = 0;  // next byte to send, next piece to wait for
    while (off < bytes) {
      const size_t len = std::min(off == 0 ? FIRST : CHUNK, bytes - off);
      for (; p * PIECE < off + len; ++p)
        while (!ready[p].load(std::memory_order_acquire))
          if (!copy_one()) std::this_thread::yield();
      G16_CHECK(hipMemcpyAsync(dev + off, dst + off, len, hipMemcpyHostToDevice, S->stream));
      off += len;
    }
  } catch (...) {
    err = std::current_exception();
  }
  for (auto& t : helpers) t->join();
  if (err) std::rethrow_exception(err);
}

static bool timeline_on() {
  static const bool on = std::getenv("G16HIP_TIMELINE") != nullptr;
  return on;
}

static void load_witness(const g16hip_circuit* C, Slot* S, const void* w, size_t n,
                         bool on_device) {
  check_witness_len(C, n);
  if (timeline_on()) {
    G16_CHECK(hipEventRecord(S->ev[6], S->stream));
    S->witness_host_ms = -1.0;
  }
  if (!on_device && timeline_on()) {
    const auto t0 = std::chrono::steady_clock::now();
    stage_witness(S, static_cast<const uint8_t*>(w), 32 * n);
    S->witness_host_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  } else if (on_device)
    G16_CHECK(hipMemcpyAsync(S->witness.p, w, 32 * n, hipMemcpyDeviceToDevice, S->stream));
  else
    stage_witness(S, static_cast<const uint8_t*>(w), 32 * n);
}

// Stage 0 handed over by the caller: constraint i's A.w and B.w (host memory), i < n.
struct Rows {
  const uint8_t* a;
  const uint8_t* b;
  size_t n;
};

// The gather path needs the key's matrices and snarkjs' H basis; a key without them proves
// from rows (crates/groth16/src/cpu.rs:176-189).
static void require_gather(const g16hip_circuit* C) {
  if (C->h_basis == 1)
    throw ArgError{"a gnark key carries no constraint matrices; prove it from A.w and B.w with "
                   "h_and_msms_from_rows"};
  if (!C->has_matrices)
    throw ArgError{"the key carries no constraint matrices; prove it from A.w and B.w with "
                   "h_and_msms_from_rows"};
}

// Stages 0-4 on the slot's main stream into `h_out` (device): stage 0 by gather or from
// `rows`, then H on the odd coset, taken to gnark's coefficient basis for a gnark key.
static void stages_0_4(const g16hip_circuit* C, Slot* S, void* h_out, bool timed,
                       const Rows* rows = nullptr) {
  hipStream_t s = S->stream;
  size_t n = C->domain;
  char* abc = S->abc.as<char>();
  void* A = abc;
  void* B = abc + 32 * n;
  void* Cv = abc + 64 * n;
  if (!rows) require_gather(C);
  if (timed) G16_CHECK(hipEventRecord(S->ev[0], s));
  if (rows) {
    if (rows->n) {
      G16_CHECK(hipMemcpyAsync(A, rows->a, 32 * rows->n, hipMemcpyHostToDevice, s));
      G16_CHECK(hipMemcpyAsync(B, rows->b, 32 * rows->n, hipMemcpyHostToDevice, s));
    }
    launch_rows_abc(A, B, Cv, rows->n, n, s);
  } else {
    const uint32_t* rp[2] = {C->row_ptr[0].as<uint32_t>(), C->row_ptr[1].as<uint32_t>()};
    const uint32_t* sg[2] = {C->signal[0].as<uint32_t>(), C->signal[1].as<uint32_t>()};
    const uint32_t* vi[2] = {C->value_id[0].as<uint32_t>(), C->value_id[1].as<uint32_t>()};
    launch_gather_abc(rp, sg, vi, C->values.p, S->witness.p, C->n_vars, abc, C->domain, 1, s);
  }
  if (timed) G16_CHECK(hipEventRecord(S->ev[1], s));
  // The transforms, with H = A*B - C written by the last pass.
  ntt_coset_pipeline(abc, 3, C->log_n, C->tables.inv.p, C->tables.fwd.p, C->tables.coset.p, s,
                     h_out);
  if (timed) G16_CHECK(hipEventRecord(S->ev[2], s));
  if (C->h_basis == 1) {
    // Coefficients in bit-reversed order, the order gnark's computeH leaves and pk.G1.Z
    // expects (crates/groth16/src/cpu.rs:250-286).
    ntt_dif_inverse(h_out, 1, C->log_n, C->tables.inv.p, s);
    launch_mul_table_rp(h_out, n, C->tables.gnark.p, s);
  }
  if (timed) G16_CHECK(hipEventRecord(S->ev[3], s));
}

// H's coefficient n - 1 (position n - 1 in the bit-reversed order) is zero whenever the
// arithmetic is right, since A*B - C vanishes on the domain; a gnark-basis proof checks it.
static void check_gnark_h(const g16hip_circuit* C, const Slot* S) {
  if (C->h_basis != 1) return;
  uint32_t last[8];
  G16_CHECK(hipMemcpy(last, S->hbuf.as<char>() + 32 * ((size_t)C->domain - 1), 32,
                      hipMemcpyDeviceToHost));
  for (uint32_t v : last)
    if (v)
      throw StatusError{G16HIP_BACKEND,
                        "H has a nonzero coefficient " + std::to_string(C->domain - 1) +
                            ", so the quotient by the vanishing polynomial is not exact"};
}

static const size_t kMsmCount = 5;  // A, B1, L, B2, H

struct WinLayout {
  size_t off[5], len[5], totals;
};

// Reduction rows of each MSM in the slot's window-sum buffer (device and pinned host).
static WinLayout win_layout(const Slot* S) {
  size_t Ww = msm_out_points(S->plan_w), Wh = msm_out_points(S->plan_h);
  return {{0, 128 * Ww, 256 * Ww, 384 * Ww, 640 * Ww},
          {128 * Ww, 128 * Ww, 128 * Ww, 256 * Ww, 128 * Wh},
          640 * Ww + 128 * Wh};
}

// Enqueues stages (0-4 if `with_h`) 5-9 on the slot's three streams: side[0] sorts the
// witness digits once and accumulates A, B1 and L; side[1] accumulates B2 on the same
// entries; the main stream runs stages 0-4 and then the H sort and MSM. Each MSM's reduction
// rows are written by its kernel into pinned host memory. With `timed`, events bracket every
// phase (done_ev[k] marks MSM k's rows on the host) and the entry counts are copied out at
// the end for the kernel statistics; a graph capture passes timed = false.
static void enqueue_stages(const g16hip_circuit* C, Slot* S, bool with_h, void* h, bool timed,
                           const Rows* rows = nullptr) {
  hipStream_t s0 = S->stream, s1 = S->side[0], s2 = S->side[1];
  MsmPlan& pw = S->plan_w;
  MsmPlan& ph = S->plan_h;
  uint8_t* wout = S->host_wout_dev;
  WinLayout L = win_layout(S);
  uint32_t* totals = reinterpret_cast<uint32_t*>(S->host_wout + L.totals);
  auto rows_done = [&](int k, hipStream_t st) {
    if (timed) G16_CHECK(hipEventRecord(S->done_ev[k], st));
  };
  hipEvent_t* mev0 = timed ? S->mev[0] : nullptr;
  hipEvent_t* mev1 = timed ? S->mev[1] : nullptr;
  hipEvent_t* mev2 = timed ? S->mev[2] : nullptr;
  hipEvent_t* mev3 = timed ? S->mev[3] : nullptr;
  hipEvent_t* mev4 = timed ? S->mev[4] : nullptr;
  G16_CHECK(hipEventRecord(S->sync_ev[0], s0));  // the witness is on the device
  if (timed) G16_CHECK(hipEventRecord(S->ev[4], s0));
  G16_CHECK(hipStreamWaitEvent(s1, S->sync_ev[0], 0));
  msm_sort(pw, S->witness.p, s1);
  G16_CHECK(hipEventRecord(S->sync_ev[1], s1));
  // Entry totals for the kernel statistics, copied while the MSMs run rather than after them.
  if (timed) G16_CHECK(hipMemcpyAsync(&totals[0], pw.total, 4, hipMemcpyDeviceToHost, s1));
  // B2 is the longest chain: it is submitted first, so it also starts first where streams
  // end up sharing a hardware queue.
  G16_CHECK(hipStreamWaitEvent(s2, S->sync_ev[1], 0));
  msm_accumulate(pw, 2, C->b2_q.p, S->accum_w2.p, wout + L.off[3], s2, mev3);
  rows_done(3, s2);
  G16_CHECK(hipEventRecord(S->sync_ev[3], s2));
  msm_accumulate(pw, 1, C->a_q.p, S->accum_w1.p, wout + L.off[0], s1, mev0);
  rows_done(0, s1);
  msm_accumulate(pw, 1, C->b1_q.p, S->accum_w1.p, wout + L.off[1], s1, mev1);
  rows_done(1, s1);
  msm_accumulate(pw, 1, C->l_q.p, S->accum_w1.p, wout + L.off[2], s1, mev2);
  rows_done(2, s1);
  G16_CHECK(hipEventRecord(S->sync_ev[2], s1));
  if (with_h) stages_0_4(C, S, h, timed, rows);
  msm_sort(ph, h, s0);
  if (timed) G16_CHECK(hipMemcpyAsync(&totals[1], ph.total, 4, hipMemcpyDeviceToHost, s0));
  msm_accumulate(ph, 1, C->h_q.p, S->accum_h.p, wout + L.off[4], s0, mev4);
  rows_done(4, s0);
  G16_CHECK(hipStreamWaitEvent(s0, S->sync_ev[2], 0));
  G16_CHECK(hipStreamWaitEvent(s0, S->sync_ev[3], 0));
  if (timed) G16_CHECK(hipEventRecord(S->ev[5], s0));
}

// The stage-11 products that need an MSM output: s*A and r*B1 (prove.rs:391-430 groups the
// rest with the key, see BlinderTerms).
struct BlindProducts {
  bool want = false;
  Fr r_std, s_std;
  G1X sA, rB1;
};

// Finishes MSM k on the host: Horner over its reduction rows, affine output, and the blinding
// product where one rides on it. `wait` blocks until the sums are on the host.
static void finish_msms(const g16hip_circuit* C, Slot* S, bool per_msm_events,
                        g16hip_msm_out* out, BlindProducts* blind) {
  WinLayout L = win_layout(S);
  const uint8_t* hw = S->host_wout;
  const MsmPlan& pw = S->plan_w;
  const MsmPlan& ph = S->plan_h;
  auto wait = [&](int k) {
    if (per_msm_events) G16_CHECK(hipEventSynchronize(S->done_ev[k]));
  };
  auto g1_sum = [&](int k, const MsmPlan& p) {
    wait(k);
    std::vector<G1X> v(L.len[k] / sizeof(G1X));
    std::memcpy(v.data(), hw + L.off[k], L.len[k]);
    return msm_combine(v.data(), p);
  };
  bool want = blind && blind->want;
  HostTask ta([&] {
    G1X A = g1_sum(0, pw);
    g1_store(out->a_g1, host_to_aff(A));
    if (want) blind->sA = host_mul(A, blind->s_std.v);
  });
  HostTask tb([&] {
    G1X B1 = g1_sum(1, pw);
    g1_store(out->b_g1, host_to_aff(B1));
    if (want) blind->rB1 = host_mul(B1, blind->r_std.v);
  });
  HostTask tl([&] { g1_store(out->l_g1, host_to_aff(g1_sum(2, pw))); });
  HostTask t2([&] {
    wait(3);
    std::vector<G2X> v(L.len[3] / sizeof(G2X));
    std::memcpy(v.data(), hw + L.off[3], L.len[3]);
    g2_store(out->b_g2, host_to_aff(msm_combine(v.data(), pw)));
  });
  std::exception_ptr err;
  try {
    g1_store(out->h_g1, host_to_aff(g1_sum(4, ph)));
  } catch (...) {
    err = std::current_exception();
  }
  for (HostTask* t : {&ta, &tb, &tl, &t2}) {
    t->join();
    if (!err && t->err) err = t->err;
  }
  if (err) std::rethrow_exception(err);
}

// Eager run with events: timings, kernel statistics, the optional per-proof timeline.
// `beside` runs on the calling thread once the GPU work is enqueued (host work that overlaps
// it without a thread of its own).
static void run_timed(g16hip_circuit* C, Slot* S, bool with_h, void* h, g16hip_msm_out* out,
                      g16hip_timings* t, BlindProducts* blind, const Rows* rows = nullptr,
                      const std::function<void()>& beside = nullptr) {
  enqueue_stages(C, S, with_h, h, true, rows);
  std::exception_ptr err;
  try {
    if (beside) beside();
    finish_msms(C, S, true, out, blind);
  } catch (...) {
    err = std::current_exception();
  }
  // ev[5] follows every stream's work (the main stream waited for both side streams before
  // recording it): one event wait instead of three stream synchronisations.
  G16_CHECK(hipEventSynchronize(S->ev[5]));
  if (err) std::rethrow_exception(err);
  if (rows) check_gnark_h(C, S);
  WinLayout L = win_layout(S);
  const uint32_t* totals = reinterpret_cast<const uint32_t*>(S->host_wout + L.totals);
  for (size_t k = 0; k < kMsmCount; ++k)
    note_accumulate(C->dev, k == 3 ? 2 : 1, k == 4 ? totals[1] : totals[0],
                    ev_ms(S->mev[k][0], S->mev[k][1]));
  if (timeline_on()) {
    // Milliseconds from the witness arriving on the device, per stream (the
    // G16_CUDA_MEMORY_TRACE kind of switch, crates/cuda/src/backend.rs:89-93).
    static const char* names[5] = {"A", "B1", "L", "B2", "H"};
    std::fprintf(stderr, "[g16hip timeline]");
    if (S->witness_host_ms >= 0)
      std::fprintf(stderr, " witness: staged on the host in %.2f, on the device %.2f after the "
                   "first copy |", S->witness_host_ms, ev_ms(S->ev[6], S->ev[4]));
    else
      std::fprintf(stderr, " witness: on the device %.2f after the copy |",
                   ev_ms(S->ev[6], S->ev[4]));
    if (with_h)
      std::fprintf(stderr, " ntt %.2f-%.2f h_done %.2f |", ev_ms(S->ev[4], S->ev[1]),
                   ev_ms(S->ev[4], S->ev[2]), ev_ms(S->ev[4], S->ev[3]));
    for (size_t k = 0; k < kMsmCount; ++k)
      std::fprintf(stderr, " %s acc %.2f-%.2f merge %.2f reduce %.2f |", names[k],
                   ev_ms(S->ev[4], S->mev[k][0]), ev_ms(S->ev[4], S->mev[k][1]),
                   ev_ms(S->ev[4], S->mev[k][2]), ev_ms(S->ev[4], S->mev[k][3]));
    std::fprintf(stderr, " end %.2f\n", ev_ms(S->ev[4], S->ev[5]));
  }
  if (t) {
    if (with_h) {
      t->gather_us += (uint64_t)(ev_ms(S->ev[0], S->ev[1]) * 1000.0f);
      t->ntt_us += (uint64_t)(ev_ms(S->ev[1], S->ev[2]) * 1000.0f);
      t->pointwise_us += (uint64_t)(ev_ms(S->ev[2], S->ev[3]) * 1000.0f);
    }
    // Stages 5-9 overlap stages 0-4: msm_us is the wall time from the witness arriving to the
    // last MSM finishing, as the reference reports overlapped MSMs
    // (crates/groth16/src/lib.rs:333-335).
    t->msm_us += (uint64_t)(ev_ms(S->ev[4], S->ev[5]) * 1000.0f);
  }
}

// Graph run: stages 0-9 captured once per slot (the fork/join of the three streams is part of
// the graph; there are no memset nodes: every clear is a fill kernel), then replayed.
// Replays go on the slot's own main stream, which lives as long as the graph: launching a
// slot's graph on a stream that is destroyed before the graph's next launch is not done.
static void run_graph(g16hip_circuit* C, Slot* S, g16hip_msm_out* out, BlindProducts* blind,
                      const std::function<void()>& beside = nullptr) {
  if (!S->exec) {
    G16_CHECK(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_stages(C, S, true, S->hbuf.p, false);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(S->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    G16_CHECK(hipStreamEndCapture(S->stream, &S->graph));
    G16_CHECK(hipGraphInstantiate(&S->exec, S->graph, nullptr, nullptr, 0));
  }
  G16_CHECK(hipGraphLaunch(S->exec, S->stream));
  std::exception_ptr err;
  try {
    if (beside) beside();
  } catch (...) {
    err = std::current_exception();
  }
  G16_CHECK(hipStreamSynchronize(S->stream));
  if (err) std::rethrow_exception(err);
  finish_msms(C, S, false, out, blind);
}

// ---------------------------------------------------------------------------------------
// stage 11 (crates/groth16/src/prove.rs:391-430), on the host:
//   pi_a = A + alpha1 + r delta1,  pi_b = B2 + beta2 + s delta2,
//   pi_c = L + H + s pi_a + r (B1 + beta1 + s delta1) - r s delta1
//        = L + H + s A + r B1 + (s alpha1 + r beta1 + r s delta1).
// The key-only terms are computed beside the GPU work; s A and r B1 as soon as those MSMs
// finish. Group elements are canonical, so the proof is the same point whatever order the
// terms are added in.
// ---------------------------------------------------------------------------------------
struct BlinderTerms {
  G1X d1r, k1;  // r delta1;  s alpha1 + r beta1 + r s delta1
  G2X d2s;      // s delta2
};

static BlinderTerms blinder_terms(const g16hip_circuit* C, const Fr& r_std, const Fr& s_std,
                                  const Fr& rs_std) {
  BlinderTerms b;
  b.d1r = host_mul(xyzz_from_aff(C->delta_g1), r_std.v);
  b.k1 = host_add(host_add(host_mul(xyzz_from_aff(C->alpha_g1), s_std.v),
                           host_mul(xyzz_from_aff(C->beta_g1), r_std.v)),
                  host_mul(xyzz_from_aff(C->delta_g1), rs_std.v));
  b.d2s = host_mul(xyzz_from_aff(C->delta_g2), s_std.v);
  return b;
}

static void assemble(const g16hip_circuit* C, const g16hip_msm_out& m, const BlinderTerms& b,
                     const BlindProducts& bp, g16hip_proof* proof) {
  G1X pi_a = host_add(host_add(xyzz_from_aff(g1_load(m.a_g1)), xyzz_from_aff(C->alpha_g1)),
                      b.d1r);
  G2X pi_b = host_add(host_add(xyzz_from_aff(g2_load(m.b_g2)), xyzz_from_aff(C->beta_g2)),
                      b.d2s);
  G1X pi_c = host_add(xyzz_from_aff(g1_load(m.l_g1)), xyzz_from_aff(g1_load(m.h_g1)));
  pi_c = host_add(pi_c, bp.sA);
  pi_c = host_add(pi_c, bp.rB1);
  pi_c = host_add(pi_c, b.k1);
  g1_store(proof->a, host_to_aff(pi_a));
  g2_store(proof->b, host_to_aff(pi_b));
  g1_store(proof->c, host_to_aff(pi_c));
}

// Stages 0-11 for one witness on a leased slot: eager with timings (`graph` false) or as a
// graph replay.
static void prove_on_slot(g16hip_circuit* C, Slot* S, const void* w, size_t n, bool on_device,
                          const uint8_t* r8, const uint8_t* s8, g16hip_proof* proof,
                          g16hip_timings* t, bool graph, const Rows* rows = nullptr) {
  check_witness_len(C, n);
  // w[0] is the constant-one wire (crates/groth16/src/prove.rs:272-283). A device witness's
  // w[0] comes back with the proof (pinned bytes after the MSM totals) instead of through a
  // synchronous copy that would hold the GPU idle before the proof starts; the call still
  // fails with CONSTANT_WIRE and writes no proof.
  Fr w0;
  uint8_t* w0_pinned = S->host_wout + win_layout(S).totals + 16;
  if (on_device) {
    G16_CHECK(hipMemcpyAsync(w0_pinned, w, 32, hipMemcpyDeviceToHost, S->stream));
  } else {
    std::memcpy(w0.v, w, 32);
    if (!is_one(w0)) throw StatusError{G16HIP_CONSTANT_WIRE, "w[0] is the constant-one wire and must be 1"};
  }
  Fr r = fr_load(r8), s = fr_load(s8);
  if (is_zero(r) || is_zero(s)) throw StatusError{G16HIP_ZERO_BLINDER, "a blinder is zero"};
  BlindProducts bp;
  bp.want = true;
  bp.r_std = to_std(r);
  bp.s_std = to_std(s);
  Fr rs_std = to_std(mul(r, s));
  BlinderTerms terms;
  const std::function<void()> beside = [&] { terms = blinder_terms(C, bp.r_std, bp.s_std, rs_std); };
  g16hip_msm_out m;
  load_witness(C, S, w, n, on_device);
  if (graph && !rows) run_graph(C, S, &m, &bp, beside);
  else run_timed(C, S, true, S->hbuf.p, &m, t, &bp, rows, beside);
  if (on_device) {  // the slot's streams are synchronised: the copy has landed
    std::memcpy(w0.v, w0_pinned, 32);
    if (!is_one(w0)) throw StatusError{G16HIP_CONSTANT_WIRE, "w[0] is the constant-one wire and must be 1"};
  }
  Timer ta;
  assemble(C, m, terms, bp, proof);
  if (t) t->assemble_us += ta.us();
}

// Waves (and the server's slot proofs) launch their kernels one by one (the eager path) or
// replay the pipeline captured once as a hipGraph: g16hip_set_graph_replay, or G16HIP_GRAPHS=1
// at start-up. Eager is the default.
static std::atomic<int> g_graph_replay{-1};

static bool replay_graphs() {
  int v = g_graph_replay.load(std::memory_order_relaxed);
  if (v < 0) {
    const char* e = std::getenv("G16HIP_GRAPHS");
    v = (e && e[0] == '1') ? 1 : 0;
    int expected = -1;
    g_graph_replay.compare_exchange_strong(expected, v);
    v = g_graph_replay.load(std::memory_order_relaxed);
  }
  return v == 1;
}

// ---------------------------------------------------------------------------------------
// batched proving: waves
// ---------------------------------------------------------------------------------------
// Waves need the gather path and snarkjs' H basis; other keys prove item by item on a slot.
// Waves take gather-path keys with snarkjs' H basis and variable-work MSMs (a constant-work
// circuit proves on slots: its accumulation is not in the persistent queue).
static bool can_wave(const g16hip_circuit* C) {
  return C->has_matrices && C->h_basis == 0 && !C->constant;
}

struct WavePlans {
  MsmPlan w, h;      // bound to the wave's scratch for m proofs
  MsmPlan vw, vh;    // one proof's view (its rows)
  size_t rw = 0, rh = 0;  // reduction rows per proof
};

static WavePlans wave_plans(const g16hip_circuit* C, void* sort_w, void* sort_h, uint32_t m) {
  WavePlans P;
  msm_plan_bind(P.w, sort_w, C->n_vars, C->cw, C->merged, C->constant, m);
  msm_plan_bind(P.h, sort_h, C->n_h, C->ch, C->merged, C->constant, m);
  P.vw = msm_vector_view(P.w);
  P.vh = msm_vector_view(P.h);
  P.rw = msm_out_points(P.vw);
  P.rh = msm_out_points(P.vh);
  return P;
}

// Rows of MSM j (A, B1, L, B2, H) for proof k of a wave of cap proofs, in bytes from the start
// of the rows buffer: every MSM's block holds cap proofs' rows.
static size_t wave_rows_offset(const WavePlans& P, uint32_t cap, int j, uint32_t k) {
  const size_t g1 = 128 * P.rw * cap;
  const size_t base[5] = {0, g1, 2 * g1, 3 * g1, 3 * g1 + 256 * P.rw * cap};
  const size_t per[5] = {128 * P.rw, 128 * P.rw, 128 * P.rw, 256 * P.rw, 128 * P.rh};
  return base[j] + per[j] * k;
}

static size_t wave_rows_bytes(const WavePlans& P, uint32_t cap) {
  return wave_rows_offset(P, cap, 4, cap);
}

static size_t wave_bytes_for(const g16hip_circuit* C, uint32_t cap) {
  static char anchor;  // bound only to size the pieces
  const WavePlans P = wave_plans(C, &anchor, &anchor, cap);
  const size_t nv = C->n_vars, n = C->domain;
  return dev_pages(32 * nv * cap) + dev_pages(96 * n * cap) + dev_pages(32 * n * cap) +
         dev_pages(msm_sort_scratch_bytes((uint32_t)nv, C->cw, cap)) +
         dev_pages(msm_sort_scratch_bytes(C->n_h, C->ch, cap)) +
         3 * dev_pages(msm_accum_scratch_bytes(P.w, 1)) + dev_pages(msm_accum_scratch_bytes(P.w, 2)) +
         dev_pages(msm_accum_scratch_bytes(P.h, 1)) + dev_pages(256);
}

// Proofs per wave when none was measured: enough to give a transform about 2^21 elements per
// vector set (small circuits fill the device together), at most 16.
static uint32_t default_wave_cap(const g16hip_circuit* C) {
  uint32_t cap = (1u << 21) / std::max<uint32_t>(C->domain, 1);
  return std::max(1u, std::min(16u, cap));
}

static uint32_t wave_cap_of(const g16hip_circuit* C) {
  const uint32_t cap = C->wave_cap.load();
  return cap ? cap : default_wave_cap(C);
}

// Builds a wave for cap proofs, refused before allocating when the device budget cannot hold it.
static std::unique_ptr<Wave> make_wave(g16hip_circuit* C, uint32_t cap) {
  const size_t need = wave_bytes_for(C, cap);
  const size_t budget = device_budget(C->dev);
  if (need > budget) throw too_big(need, budget, "a wave of " + std::to_string(cap) + " proofs");
  auto W = std::make_unique<Wave>();
  W->cap = cap;
  G16_CHECK(hipStreamCreateWithFlags(&W->stream, hipStreamNonBlocking));
  for (auto& e : W->done) G16_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const size_t nv = C->n_vars, n = C->domain;
  W->witness.alloc(32 * nv * cap);
  W->abc.alloc(96 * n * cap);
  W->hbuf.alloc(32 * n * cap);
  W->sort_w.alloc(msm_sort_scratch_bytes((uint32_t)nv, C->cw, cap));
  W->sort_h.alloc(msm_sort_scratch_bytes(C->n_h, C->ch, cap));
  const WavePlans P = wave_plans(C, W->sort_w.p, W->sort_h.p, cap);
  for (int j = 0; j < 3; ++j) W->accum_g1[j].alloc(msm_accum_scratch_bytes(P.w, 1));
  W->accum_g1[3].alloc(msm_accum_scratch_bytes(P.h, 1));
  W->accum_w2.alloc(msm_accum_scratch_bytes(P.w, 2));
  W->queue.alloc(256);
  W->wout_bytes = wave_rows_bytes(P, cap);
  for (int b = 0; b < 2; ++b) {
    G16_CHECK(hipHostMalloc((void**)&W->host_wout[b], W->wout_bytes + 64,
                            hipHostMallocMapped | hipHostMallocCoherent));
    G16_CHECK(hipHostGetDevicePointer((void**)&W->host_wout_dev[b], W->host_wout[b], 0));
    G16_CHECK(hipHostMalloc((void**)&W->staging[b], std::max<size_t>(32 * nv * cap, 32),
                            hipHostMallocDefault));
  }
  W->bytes = need;
  return W;
}

struct WaveItem {
  const uint8_t* witness;
  Fr r, s;  // Montgomery
  size_t index;  // in the caller's batch
};

// The device part of a wave of m proofs on buffer b (everything after the witnesses are in
// staging), on the wave's stream: the upload, stages 0-4, both sorts and the five MSMs.
static void wave_device(g16hip_circuit* C, Wave* W, uint32_t m, int b) {
  const size_t nv = C->n_vars, wb = 32 * nv;
  hipStream_t s = W->stream;
  G16_CHECK(hipMemcpyAsync(W->witness.p, W->staging[b], wb * m, hipMemcpyHostToDevice, s));
  WavePlans P = wave_plans(C, W->sort_w.p, W->sort_h.p, m);
  // The buffers are sized for the wave's cap; a plan for m <= cap proofs must fit them.
  if (msm_accum_scratch_bytes(P.w, 1) > W->accum_g1[0].bytes ||
      msm_accum_scratch_bytes(P.h, 1) > W->accum_g1[3].bytes ||
      msm_accum_scratch_bytes(P.w, 2) > W->accum_w2.bytes)
    throw StatusError{G16HIP_BACKEND, "a wave of " + std::to_string(m) +
                                          " proofs needs more accumulation scratch than its "
                                          "cap's buffers hold"};
  const uint32_t* rp[2] = {C->row_ptr[0].as<uint32_t>(), C->row_ptr[1].as<uint32_t>()};
  const uint32_t* sg[2] = {C->signal[0].as<uint32_t>(), C->signal[1].as<uint32_t>()};
  const uint32_t* vi[2] = {C->value_id[0].as<uint32_t>(), C->value_id[1].as<uint32_t>()};
  launch_gather_abc(rp, sg, vi, C->values.p, W->witness.p, nv, W->abc.p, C->domain, m, s);
  ntt_coset_pipeline(W->abc.p, 3 * (int)m, C->log_n, C->tables.inv.p, C->tables.fwd.p,
                     C->tables.coset.p, s, W->hbuf.p);
  uint8_t* rows = W->host_wout_dev[b];
  const uint32_t cap = W->cap;
  msm_sort(P.w, W->witness.p, s);
  msm_sort(P.h, W->hbuf.p, s);
  // The four G1 MSMs (A, B1, L over the witness entries, H over its own) in one accumulation
  // queue, then B2 (G2).
  const MsmPlan* plans[4] = {&P.w, &P.w, &P.w, &P.h};
  const void* bases[4] = {C->a_q.p, C->b1_q.p, C->l_q.p, C->h_q.p};
  void* scratch[4] = {W->accum_g1[0].p, W->accum_g1[1].p, W->accum_g1[2].p, W->accum_g1[3].p};
  void* outs[4] = {rows + wave_rows_offset(P, cap, 0, 0), rows + wave_rows_offset(P, cap, 1, 0),
                   rows + wave_rows_offset(P, cap, 2, 0), rows + wave_rows_offset(P, cap, 4, 0)};
  msm_accumulate_g1_queue(plans, bases, scratch, outs, 4, W->queue.as<uint32_t>(), s);
  msm_accumulate(P.w, 2, C->b2_q.p, W->accum_w2.p, rows + wave_rows_offset(P, cap, 3, 0), s);
}

// Copies the wave's witnesses into page-locked staging with a few host threads and enqueues one
// upload, then stages 0-9 for all m proofs on the wave's stream; done[b] follows the last rows.
static void enqueue_wave(g16hip_circuit* C, Wave* W, const std::vector<WaveItem>& it, int b) {
  const uint32_t m = (uint32_t)it.size();
  const size_t nv = C->n_vars, wb = 32 * nv;
  hipStream_t s = W->stream;
  uint8_t* stage = W->staging[b];
  {
    std::atomic<uint32_t> next{0};
    auto copy = [&] {
      for (uint32_t k; (k = next.fetch_add(1)) < m;) std::memcpy(stage + wb * k, it[k].witness, wb);
    };
    std::vector<std::unique_ptr<HostTask>> th;
    for (uint32_t h = 1; h < std::min<uint32_t>(m, 4); ++h)
      th.push_back(std::make_unique<HostTask>(copy));
    copy();
    for (auto& t : th) t->join();
  }
  if (replay_graphs()) {
    auto gi = W->graphs.find(2 * m + (uint32_t)b);
    if (gi == W->graphs.end()) {
      hipGraph_t g = nullptr;
      hipGraphExec_t x = nullptr;
      G16_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        wave_device(C, W, m, b);
      } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      G16_CHECK(hipStreamEndCapture(s, &g));
      const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      if (e != hipSuccess) (void)hipGraphDestroy(g);
      G16_CHECK(e);
      gi = W->graphs.emplace(2 * m + (uint32_t)b, std::make_pair(g, x)).first;
    }
    G16_CHECK(hipGraphLaunch(gi->second.second, s));
  } else {
    wave_device(C, W, m, b);
  }
  G16_CHECK(hipEventRecord(W->done[b], s));
}

// Waits for wave b and finishes its proofs on the host: each proof's Horner passes, blinding
// products and stage 11, a few proofs per host thread.
static void finish_wave(g16hip_circuit* C, Wave* W, const std::vector<WaveItem>& it, int b,
                        g16hip_proof* proofs, g16hip_status* st, std::string* msg) {
  G16_CHECK(hipEventSynchronize(W->done[b]));
  const uint32_t m = (uint32_t)it.size(), cap = W->cap;
  static char anchor;
  const WavePlans P = wave_plans(C, &anchor, &anchor, m);
  const uint8_t* rows = W->host_wout[b];
  auto g1 = [&](int j, uint32_t k, const MsmPlan& v) {
    return msm_combine(reinterpret_cast<const G1X*>(rows + wave_rows_offset(P, cap, j, k)), v);
  };
  auto one = [&](uint32_t k) {
    const WaveItem& w = it[k];
    st[w.index] = guarded([&] {
      g16hip_msm_out mo;
      BlindProducts bp;
      bp.want = true;
      bp.r_std = to_std(w.r);
      bp.s_std = to_std(w.s);
      const G1X A = g1(0, k, P.vw), B1 = g1(1, k, P.vw);
      g1_store(mo.a_g1, host_to_aff(A));
      g1_store(mo.b_g1, host_to_aff(B1));
      g1_store(mo.l_g1, host_to_aff(g1(2, k, P.vw)));
      g2_store(mo.b_g2, host_to_aff(msm_combine(
                            reinterpret_cast<const G2X*>(rows + wave_rows_offset(P, cap, 3, k)),
                            P.vw)));
      g1_store(mo.h_g1, host_to_aff(g1(4, k, P.vh)));
      bp.sA = host_mul(A, bp.s_std.v);
      bp.rB1 = host_mul(B1, bp.r_std.v);
      const BlinderTerms terms = blinder_terms(C, bp.r_std, bp.s_std, to_std(mul(w.r, w.s)));
      assemble(C, mo, terms, bp, &proofs[w.index]);
      return G16HIP_OK;
    });
    if (st[w.index] != G16HIP_OK) msg[w.index] = g16hip_last_error();
  };
  std::atomic<uint32_t> next{0};
  auto work = [&] {
    for (uint32_t k; (k = next.fetch_add(1)) < m;) one(k);
  };
  std::vector<std::unique_ptr<HostTask>> th;
  for (uint32_t h = 1; h < std::min<uint32_t>(m, 8); ++h)
    th.push_back(std::make_unique<HostTask>(work));
  work();
  for (auto& t : th) t->join();
}

// Proves the items of one circuit in waves of the circuit's cap, the host finishing wave i while
// the device runs wave i + 1. Items that fail the witness checks get their status and no proof.
static void prove_waves(g16hip_circuit* C, const g16hip_batch_item* items,
                        const std::vector<size_t>& idx, g16hip_proof* proofs, g16hip_status* st,
                        std::string* msg) {
  std::vector<WaveItem> ok;
  for (size_t i : idx) {
    const g16hip_status v = guarded([&] {
      const g16hip_batch_item& x = items[i];
      if (!x.witness || !x.r || !x.s) throw ArgError{"null argument"};
      check_witness_len(C, x.n_witness);
      if (!is_one(fr_load(x.witness)))
        throw StatusError{G16HIP_CONSTANT_WIRE, "w[0] is the constant-one wire and must be 1"};
      const Fr r = fr_load(x.r), s = fr_load(x.s);
      if (is_zero(r) || is_zero(s)) throw StatusError{G16HIP_ZERO_BLINDER, "a blinder is zero"};
      ok.push_back(WaveItem{x.witness, r, s, i});
      return G16HIP_OK;
    });
    if (v != G16HIP_OK) {
      st[i] = v;
      msg[i] = g16hip_last_error();
    }
  }
  if (ok.empty()) return;
  std::lock_guard<std::mutex> lk(C->wave_mu);
  set_device(C->dev);
  const uint32_t cap = wave_cap_of(C);
  if (!C->wave || C->wave->cap != cap) {
    C->wave.reset();
    C->wave = make_wave(C, cap);
  }
  Wave* W = C->wave.get();
  std::vector<std::vector<WaveItem>> waves;
  for (size_t a = 0; a < ok.size(); a += cap)
    waves.emplace_back(ok.begin() + a, ok.begin() + std::min(ok.size(), a + cap));
  for (size_t i = 0; i < waves.size(); ++i) {
    enqueue_wave(C, W, waves[i], (int)(i & 1));
    if (i) finish_wave(C, W, waves[i - 1], (int)((i - 1) & 1), proofs, st, msg);
  }
  finish_wave(C, W, waves.back(), (int)((waves.size() - 1) & 1), proofs, st, msg);
}

}  // namespace g16

// ---------------------------------------------------------------------------------------
// server (continuous batching)
// ---------------------------------------------------------------------------------------
namespace g16 {

struct Request {
  g16hip_circuit* circuit = nullptr;
  const uint8_t* witness = nullptr;  // caller-owned until the ticket is waited for
  size_t n_witness = 0;
  uint8_t r[32], s[32];
  Timer submitted;
  g16hip_proof proof{};
  g16hip_status status = G16HIP_OK;
  std::string error;
  double latency_ms = 0;
  bool done = false;
};

}  // namespace g16

struct g16hip_server {
  g16hip_device* dev = nullptr;
  size_t reserve = 0;
  uint32_t lanes = 1;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<g16::Request>> queue;  // first come, first served
  std::unordered_map<uint64_t, std::shared_ptr<g16::Request>> tickets;
  uint64_t next_ticket = 1;
  std::vector<std::thread> workers;
  bool closing = false;
};

namespace g16 {

// Adds one slot to `C` if the device keeps `reserve` bytes free after it (admission by free
// device memory); returns whether it did.
static bool try_add_slot(g16hip_circuit* C, size_t reserve) {
  if (device_budget(C->dev) < slot_bytes_estimate(C) + reserve) return false;
  auto S = make_slot(C);
  std::lock_guard<std::mutex> lk(C->mu);
  C->slots.push_back(std::move(S));
  C->busy.push_back(false);
  C->cv.notify_all();
  return true;
}

// A lane: takes the oldest request and every queued request of the same circuit, up to the
// circuit's wave cap, and proves them as one wave (continuous batching: requests that arrive
// while other waves run join the next one). Keys that cannot wave, or a wave the device budget
// cannot hold, prove request by request on a slot, leasing (or adding) one as memory allows.
static void server_lane(g16hip_server* srv) {
  (void)hipSetDevice(srv->dev->ordinal);
  for (;;) {
    std::vector<std::shared_ptr<Request>> group;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      srv->cv_work.wait(lk, [&] { return srv->closing || !srv->queue.empty(); });
      if (srv->queue.empty()) return;  // closing, nothing left
      g16hip_circuit* C = srv->queue.front()->circuit;
      const uint32_t cap = can_wave(C) ? wave_cap_of(C) : 1;
      for (auto it = srv->queue.begin(); it != srv->queue.end() && group.size() < cap;) {
        if ((*it)->circuit == C) {
          group.push_back(*it);
          it = srv->queue.erase(it);
        } else {
          ++it;
        }
      }
    }
    g16hip_circuit* C = group[0]->circuit;
    const size_t m = group.size();
    std::vector<g16hip_batch_item> items(m);
    std::vector<g16hip_proof> proofs(m);
    std::vector<g16hip_status> st(m, G16HIP_DEVICE);
    std::vector<std::string> msg(m);
    std::vector<size_t> idx(m);
    for (size_t k = 0; k < m; ++k) {
      items[k] = g16hip_batch_item{C, group[k]->witness, group[k]->n_witness, group[k]->r,
                                   group[k]->s};
      idx[k] = k;
    }
    bool waved = false;
    if (can_wave(C)) {
      const g16hip_status w = guarded([&] {
        prove_waves(C, items.data(), idx, proofs.data(), st.data(), msg.data());
        return G16HIP_OK;
      });
      waved = w == G16HIP_OK || w != G16HIP_DEVICE_MEMORY;
      if (!waved) std::fill(st.begin(), st.end(), G16HIP_DEVICE);
      else if (w != G16HIP_OK)
        for (size_t k = 0; k < m; ++k)
          if (st[k] == G16HIP_DEVICE && msg[k].empty()) {
            st[k] = w;
            msg[k] = g16hip_last_error();
          }
    }
    if (!waved) {
      for (size_t k = 0; k < m; ++k) {
        st[k] = guarded([&] {
          SlotLease lease(C, srv->lanes, srv->reserve);
          prove_on_slot(C, lease.S, items[k].witness, items[k].n_witness, false, items[k].r,
                        items[k].s, &proofs[k], nullptr, replay_graphs());
          return G16HIP_OK;
        });
        if (st[k] != G16HIP_OK) msg[k] = g16hip_last_error();
      }
    }
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      for (size_t k = 0; k < m; ++k) {
        Request& req = *group[k];
        req.status = st[k];
        req.error = msg[k];
        if (st[k] == G16HIP_OK) req.proof = proofs[k];
        req.latency_ms = req.submitted.ms();
        req.done = true;
      }
    }
    srv->cv_done.notify_all();
  }
}

}  // namespace g16

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
extern "C" {

g16hip_status g16hip_prepare(g16hip_device* d, const g16hip_key* key, g16hip_circuit** out) {
  return guarded([&] {
    if (!d || !out) throw ArgError{"null argument"};
    validate_key(key);
    set_device(d);
    auto C = std::make_unique<g16hip_circuit>();
    C->dev = d;
    {
      std::lock_guard<std::mutex> lk(d->mu);
      C->constant = d->constant_work;
    }
    // Refused for device memory before anything is allocated (streams included).
    const KeyPlan plan = choose_plan(d, key);
    // The key goes up on slot 0's main stream (a failure destroys the slot, which waits for
    // its streams first).
    auto S = new_slot();
    upload_key(C.get(), key, plan, S->stream);
    size_slot(C.get(), S.get());
    C->slots.push_back(std::move(S));
    C->busy.push_back(false);
    *out = C.release();
    return G16HIP_OK;
  });
}

g16hip_status g16hip_key_resident_bytes(const g16hip_key* key, int window_tables, size_t* bytes) {
  return guarded([&] {
    if (!bytes) throw ArgError{"null argument"};
    validate_key(key);
    *bytes = plan_key(key, window_tables != 0, false).total();
    return G16HIP_OK;
  });
}

void g16hip_circuit_release(g16hip_circuit* C) {
  if (!C) return;
  (void)hipSetDevice(C->dev->ordinal);
  C->slots.clear();
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

size_t g16hip_circuit_resident_bytes(const g16hip_circuit* C) {
  if (!C) return 0;
  std::lock_guard<std::mutex> lk(const_cast<g16hip_circuit*>(C)->mu);
  size_t tot = C->key_bytes;
  for (auto& S : C->slots) tot += S->bytes;
  return tot;
}

g16hip_status g16hip_circuit_add_slots(g16hip_circuit* C, uint32_t count, size_t reserve_bytes,
                                       uint32_t* total_slots) {
  return guarded([&] {
    if (!C) throw ArgError{"null circuit"};
    set_device(C->dev);
    for (uint32_t i = 0; i < count; ++i)
      if (!try_add_slot(C, reserve_bytes)) break;
    if (total_slots) {
      std::lock_guard<std::mutex> lk(C->mu);
      *total_slots = (uint32_t)C->slots.size();
    }
    return G16HIP_OK;
  });
}

g16hip_status g16hip_compute_h(g16hip_circuit* C, const uint8_t* w, size_t n, g16hip_h** h_out,
                               g16hip_timings* t) {
  return guarded([&] {
    if (!C || !w || !h_out) throw ArgError{"null argument"};
    set_device(C->dev);
    SlotLease lease(C);
    Slot* S = lease.S;
    load_witness(C, S, w, n, false);
    auto H = std::make_unique<g16hip_h>();
    H->owner = C;
    H->buf.alloc(32 * (size_t)C->domain);
    stages_0_4(C, S, H->buf.p, true);
    G16_CHECK(hipStreamSynchronize(S->stream));
    if (t) {
      t->gather_us += (uint64_t)(ev_ms(S->ev[0], S->ev[1]) * 1000.0f);
      t->ntt_us += (uint64_t)(ev_ms(S->ev[1], S->ev[2]) * 1000.0f);
      t->pointwise_us += (uint64_t)(ev_ms(S->ev[2], S->ev[3]) * 1000.0f);
    }
    *h_out = H.release();
    return G16HIP_OK;
  });
}

g16hip_status g16hip_h_to_host(g16hip_circuit* C, const g16hip_h* h, uint8_t* out) {
  return guarded([&] {
    if (!C || !h || !out) throw ArgError{"null argument"};
    if (h->owner != C) throw ArgError{"H handle came from another circuit's compute_h"};
    set_device(C->dev);
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
    set_device(C->dev);
    SlotLease lease(C);
    load_witness(C, lease.S, w, n, false);
    run_timed(C, lease.S, false, h->buf.p, out, t, nullptr);
    return G16HIP_OK;
  });
}

static g16hip_status h_and_msms_impl(g16hip_circuit* C, const void* w, size_t n, bool on_device,
                                     g16hip_msm_out* out, g16hip_timings* t) {
  return guarded([&] {
    if (!C || !w || !out) throw ArgError{"null argument"};
    set_device(C->dev);
    SlotLease lease(C);
    load_witness(C, lease.S, w, n, on_device);
    run_timed(C, lease.S, true, lease.S->hbuf.p, out, t, nullptr);
    return G16HIP_OK;
  });
}

g16hip_status g16hip_h_and_msms(g16hip_circuit* C, const uint8_t* w, size_t n,
                                g16hip_msm_out* out, g16hip_timings* t) {
  return h_and_msms_impl(C, w, n, false, out, t);
}

g16hip_status g16hip_h_and_msms_device(g16hip_circuit* C, const void* w, size_t n,
                                       g16hip_msm_out* out, g16hip_timings* t) {
  return h_and_msms_impl(C, w, n, true, out, t);
}

// The rows' shape (crates/groth16/src/cpu.rs:394-406).
static Rows check_rows(const g16hip_circuit* C, const uint8_t* a, const uint8_t* b, size_t rows) {
  if (rows > C->domain)
    throw ArgError{"A.w and B.w have " + std::to_string(rows) + " rows; both must be one per "
                   "constraint, at most the domain size " + std::to_string(C->domain)};
  if (rows && (!a || !b)) throw ArgError{"null rows"};
  return Rows{a, b, rows};
}

g16hip_status g16hip_h_and_msms_from_rows(g16hip_circuit* C, const uint8_t* w, size_t n,
                                          const uint8_t* a_rows, const uint8_t* b_rows,
                                          size_t n_rows, g16hip_msm_out* out, g16hip_timings* t) {
  return guarded([&] {
    if (!C || !w || !out) throw ArgError{"null argument"};
    Rows rows = check_rows(C, a_rows, b_rows, n_rows);
    set_device(C->dev);
    SlotLease lease(C);
    load_witness(C, lease.S, w, n, false);
    run_timed(C, lease.S, true, lease.S->hbuf.p, out, t, nullptr, &rows);
    return G16HIP_OK;
  });
}

static g16hip_status prove_impl(g16hip_circuit* C, const void* w, size_t n, bool on_device,
                                const uint8_t* r, const uint8_t* s, g16hip_proof* proof,
                                g16hip_timings* t) {
  return guarded([&] {
    if (!C || !w || !r || !s || !proof) throw ArgError{"null argument"};
    set_device(C->dev);
    SlotLease lease(C);
    prove_on_slot(C, lease.S, w, n, on_device, r, s, proof, t, false);
    return G16HIP_OK;
  });
}

g16hip_status g16hip_prove_with_blinders(g16hip_circuit* C, const uint8_t* w, size_t n,
                                         const uint8_t* r, const uint8_t* s, g16hip_proof* proof,
                                         g16hip_timings* t) {
  return prove_impl(C, w, n, false, r, s, proof, t);
}

// prove then self_verify (crates/groth16/src/prove.rs:66-91): the public inputs are
// witness[1..=n_public], checked against the circuit's own verifying key.
g16hip_status g16hip_prove(g16hip_circuit* C, const uint8_t* w, size_t n, const uint8_t* r,
                           const uint8_t* s, g16hip_proof* proof, g16hip_timings* t) {
  g16hip_status st = prove_impl(C, w, n, false, r, s, proof, t);
  if (st != G16HIP_OK) return st;
  Timer tv;
  std::string why;
  st = groth16_check(C->vk, w + 32, C->n_public, *proof, VerifyMode::Checked, &why);
  if (t) t->verify_us += tv.us();
  if (st != G16HIP_OK) return fail(G16HIP_SELF_VERIFY, "the proof failed its check: " + why);
  return G16HIP_OK;
}

// prove_from_rows (crates/groth16/src/prove.rs:93-115): checked like prove.
g16hip_status g16hip_prove_from_rows(g16hip_circuit* C, const uint8_t* w, size_t n,
                                     const uint8_t* a_rows, const uint8_t* b_rows, size_t n_rows,
                                     const uint8_t* r, const uint8_t* s, g16hip_proof* proof,
                                     g16hip_timings* t) {
  g16hip_status st = guarded([&] {
    if (!C || !w || !r || !s || !proof) throw ArgError{"null argument"};
    Rows rows = check_rows(C, a_rows, b_rows, n_rows);
    set_device(C->dev);
    SlotLease lease(C);
    prove_on_slot(C, lease.S, w, n, false, r, s, proof, t, false, &rows);
    return G16HIP_OK;
  });
  if (st != G16HIP_OK) return st;
  Timer tv;
  std::string why;
  st = groth16_check(C->vk, w + 32, C->n_public, *proof, VerifyMode::Checked, &why);
  if (t) t->verify_us += tv.us();
  if (st != G16HIP_OK) return fail(G16HIP_SELF_VERIFY, "the proof failed its check: " + why);
  return G16HIP_OK;
}

g16hip_status g16hip_circuit_vk(const g16hip_circuit* C, uint8_t* out) {
  return guarded([&] {
    if (!C || !out) throw ArgError{"null argument"};
    std::memcpy(out, C->vk.alpha_g1, 64);
    std::memcpy(out + 64, C->vk.beta_g2, 128);
    std::memcpy(out + 192, C->vk.gamma_g2, 128);
    std::memcpy(out + 320, C->vk.delta_g2, 128);
    std::memcpy(out + 448, C->vk.ic.data(), C->vk.ic.size());
    return G16HIP_OK;
  });
}

g16hip_status g16hip_prove_with_blinders_device(g16hip_circuit* C, const void* w, size_t n,
                                                const uint8_t* r, const uint8_t* s,
                                                g16hip_proof* proof, g16hip_timings* t) {
  return prove_impl(C, w, n, true, r, s, proof, t);
}

void g16hip_set_graph_replay(int on) { g_graph_replay.store(on ? 1 : 0); }

g16hip_status g16hip_prove_batch(const g16hip_batch_item* items, size_t n, g16hip_proof* proofs,
                                 g16hip_status* statuses) {
  return guarded([&] {
    if (n && (!items || !proofs)) throw ArgError{"null argument"};
    // One worker per slot of each circuit in the batch; each takes that circuit's items in
    // order and replays its slot's captured pipeline.
    std::unordered_map<g16hip_circuit*, std::vector<size_t>> by_circuit;
    for (size_t i = 0; i < n; ++i) {
      if (!items[i].circuit) throw ArgError{"null circuit in batch item " + std::to_string(i)};
      by_circuit[items[i].circuit].push_back(i);
    }
    // Every item starts as not proved; a circuit's worker overwrites each status it settles.
    std::vector<g16hip_status> st(n, G16HIP_DEVICE);
    std::vector<std::string> msg(n);
    std::vector<std::unique_ptr<HostTask>> workers;
    for (auto& kv : by_circuit) {
      g16hip_circuit* C = kv.first;
      const std::vector<size_t>* list = &kv.second;
      workers.push_back(std::make_unique<HostTask>([=, &st, &msg] {
        (void)hipSetDevice(C->dev->ordinal);
        if (can_wave(C)) {
          // One worker per circuit: its waves, on the circuit's own stream.
          const g16hip_status w = guarded([&] {
            prove_waves(C, items, *list, proofs, st.data(), msg.data());
            return G16HIP_OK;
          });
          if (w != G16HIP_OK) {
            const std::string why = g16hip_last_error();
            for (size_t i : *list)
              if (st[i] == G16HIP_DEVICE && msg[i].empty()) {
                st[i] = w;
                msg[i] = why;
              }
          }
          return;
        }
        for (size_t i : *list) {  // rows-only and gnark-basis keys: proof by proof on a slot
          st[i] = guarded([&] {
            SlotLease lease(C);
            prove_on_slot(C, lease.S, items[i].witness, items[i].n_witness, false, items[i].r,
                          items[i].s, &proofs[i], nullptr, false);
            return G16HIP_OK;
          });
          if (st[i] != G16HIP_OK) msg[i] = g16hip_last_error();
        }
      }));
    }
    for (auto& w : workers) w->join();
    g16hip_status first = G16HIP_OK;
    std::string first_msg;
    for (size_t i = 0; i < n; ++i) {
      if (statuses) statuses[i] = st[i];
      if (first == G16HIP_OK && st[i] != G16HIP_OK) {
        first = st[i];
        first_msg = "batch item " + std::to_string(i) + ": " + msg[i];
      }
    }
    if (first != G16HIP_OK) throw StatusError{first, first_msg};
    return G16HIP_OK;
  });
}

g16hip_status g16hip_set_batch_cap(g16hip_circuit* C, uint32_t cap) {
  return guarded([&] {
    if (!C) throw ArgError{"null circuit"};
    if (cap > 64) throw ArgError{"at most 64 proofs per wave"};
    std::lock_guard<std::mutex> lk(C->wave_mu);
    C->wave_cap = cap;
    return G16HIP_OK;
  });
}

g16hip_status g16hip_batch_autotune(g16hip_circuit* C, const uint8_t* w, size_t n,
                                    uint32_t max_cap, uint32_t* caps, double* rates,
                                    size_t len, size_t* tried, uint32_t* best) {
  return guarded([&] {
    if (!C || !w || !best) throw ArgError{"null argument"};
    if (!can_wave(C)) throw ArgError{"batched proving needs the key's matrices and snarkjs' H basis"};
    check_witness_len(C, n);
    if (!is_one(fr_load(w))) throw StatusError{G16HIP_CONSTANT_WIRE, "w[0] is the constant-one wire and must be 1"};
    set_device(C->dev);
    max_cap = std::max(1u, std::min(max_cap ? max_cap : 32u, 64u));
    uint8_t r8[32] = {}, s8[32] = {};
    const Fr r = fr_from_u64(0x5eed1), s = fr_from_u64(0x5eed2);
    std::memcpy(r8, r.v, 32);
    std::memcpy(s8, s.v, 32);
    uint32_t best_cap = 1;
    double best_rate = 0;
    size_t k = 0;
    int worse = 0;
    for (uint32_t cap = 1; cap <= max_cap; cap *= 2) {
      {
        std::lock_guard<std::mutex> lk(C->wave_mu);
        if (wave_bytes_for(C, cap) > device_budget(C->dev)) break;
        C->wave_cap = cap;
      }
      // one warm wave, then two timed ones, all of the caller's witness
      std::vector<g16hip_batch_item> items(2 * cap, g16hip_batch_item{C, w, n, r8, s8});
      std::vector<g16hip_proof> pr(2 * cap);
      std::vector<g16hip_status> stv(2 * cap);
      if (g16hip_prove_batch(items.data(), cap, pr.data(), stv.data()) != G16HIP_OK)
        throw StatusError{stv[0], g16hip_last_error()};
      Timer t;
      if (g16hip_prove_batch(items.data(), 2 * cap, pr.data(), stv.data()) != G16HIP_OK)
        throw StatusError{stv[0], g16hip_last_error()};
      const double rate = 2.0 * cap / (t.us() * 1e-6);
      if (caps && rates && k < len) {
        caps[k] = cap;
        rates[k] = rate;
      }
      ++k;
      if (rate > best_rate * 1.02) {
        best_rate = rate;
        best_cap = cap;
        worse = 0;
      } else if (++worse >= 2) {
        break;
      }
    }
    {
      std::lock_guard<std::mutex> lk(C->wave_mu);
      C->wave_cap = best_cap;
    }
    if (tried) *tried = k;
    *best = best_cap;
    return G16HIP_OK;
  });
}

g16hip_status g16hip_server_open(g16hip_device* d, size_t reserve_bytes, uint32_t lanes,
                                 g16hip_server** out) {
  return guarded([&] {
    if (!d || !out) throw ArgError{"null argument"};
    if (lanes == 0 || lanes > 64) throw ArgError{"lanes must be in [1, 64]"};
    auto srv = std::make_unique<g16hip_server>();
    srv->dev = d;
    srv->reserve = reserve_bytes;
    srv->lanes = lanes;
    for (uint32_t i = 0; i < lanes; ++i) srv->workers.emplace_back(server_lane, srv.get());
    *out = srv.release();
    return G16HIP_OK;
  });
}

g16hip_status g16hip_server_submit(g16hip_server* srv, g16hip_circuit* C, const uint8_t* w,
                                   size_t n, const uint8_t* r, const uint8_t* s,
                                   uint64_t* ticket) {
  return guarded([&] {
    if (!srv || !C || !w || !r || !s || !ticket) throw ArgError{"null argument"};
    if (C->dev != srv->dev) throw ArgError{"circuit belongs to another device"};
    check_witness_len(C, n);
    auto req = std::make_shared<Request>();
    req->circuit = C;
    req->witness = w;
    req->n_witness = n;
    std::memcpy(req->r, r, 32);
    std::memcpy(req->s, s, 32);
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      if (srv->closing) throw ArgError{"server is closing"};
      srv->queue.push_back(req);
      *ticket = srv->next_ticket++;
      srv->tickets[*ticket] = req;
    }
    srv->cv_work.notify_one();
    return G16HIP_OK;
  });
}

g16hip_status g16hip_server_wait(g16hip_server* srv, uint64_t ticket, g16hip_proof* proof,
                                 double* latency_ms) {
  return guarded([&] {
    if (!srv || !proof) throw ArgError{"null argument"};
    std::shared_ptr<Request> req;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      auto it = srv->tickets.find(ticket);
      if (it == srv->tickets.end()) throw ArgError{"unknown ticket"};
      req = it->second;
      srv->cv_done.wait(lk, [&] { return req->done; });
      srv->tickets.erase(ticket);
    }
    if (latency_ms) *latency_ms = req->latency_ms;
    if (req->status != G16HIP_OK) throw StatusError{req->status, req->error};
    *proof = req->proof;
    return G16HIP_OK;
  });
}

void g16hip_server_close(g16hip_server* srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->closing = true;
  }
  srv->cv_work.notify_all();
  for (auto& t : srv->workers)
    if (t.joinable()) t.join();
  delete srv;
}

}  // extern "C"
