// Internal declarations shared by the kernel translation units and the runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace g16 {

// Throws on a HIP error; the C ABI layer turns it into G16HIP_DEVICE + message.
struct HipError {
  std::string msg;
  hipError_t code = hipErrorUnknown;
};
struct ArgError {
  std::string msg;
};

void hip_check(hipError_t e, const char* what);
// The message g16hip_last_error() returns on this thread.
void set_last_error(const std::string& msg);
#define G16_CHECK(call) ::g16::hip_check((call), #call)
#define G16_LAUNCH_CHECK(what) ::g16::hip_check(hipGetLastError(), what)

// ---- field / setup kernels (kernels_field.hip) ----
void launch_field_probe(int field, int op, const void* a, const void* b, void* out, size_t n,
                        hipStream_t s);
void launch_fixed_base(int group, const void* k, size_t n, void* out, hipStream_t s);
// kind 0: 8 chains of v_mad_u64_u32 per lane, kind 1: 4 chains of Fq products per lane.
void launch_valu_peak(int kind, const void* in, void* out, unsigned blocks, int iters,
                      hipStream_t s);

// ---- NTT kernels (kernels_ntt.hip) ----
// out[i] = base^i for i < n (Montgomery Fr).
void launch_powers(void* out, const void* base, size_t n, hipStream_t s);
// Per-pass twiddle tables from powers w^i (i < n/2): entry 2^s - 1 + j = w^(j 2^(L-1-s)),
// in the R' = 2^261 form the NTT multiplies with (bn254.cuh mul_rp).
void launch_stage_tables(void* out, const void* powers, int log_n, hipStream_t s);
void launch_coset_table(void* out, const void* shift, const void* scale, int log_n,
                        hipStream_t s);
void launch_fr_to_rp(void* data, size_t n, hipStream_t s);  // R form -> R' form, in place
// Transforms of `nvec` vectors stored back to back (stride n elements), in place, with
// per-pass twiddle tables (launch_stage_tables).
// dif_inverse: natural order in -> bit-reversed out, unscaled (w^-1 twiddles).
// dit_forward: bit-reversed in -> natural out; if `prescale` is given, element p is first
// multiplied by prescale[p] (the coset table in bit-reversed order, R' form). With
// `natural_src` the input is read from there in natural order (the bit reversal folded into
// the first pass's loads; natural_src may be data) and `tmp` (n elements per vector, distinct
// from data and natural_src) holds the first pass's output.
void ntt_dif_inverse(void* data, int nvec, int log_n, const void* tw_inv, hipStream_t s);
void ntt_dit_forward(void* data, int nvec, int log_n, const void* tw_fwd, const void* prescale,
                     hipStream_t s, const void* natural_src = nullptr, void* tmp = nullptr);
// ntt_dif_inverse then ntt_dit_forward with `coset` as the prescale, the inverse's lowest pass
// and the forward's first pass fused into one (the same tiles, in LDS). With h_out (nvec = 3:
// A, B, C back to back), H = A B - C goes to h_out from the last pass and C is not stored.
void ntt_coset_pipeline(void* data, int nvec, int log_n, const void* tw_inv, const void* tw_fwd,
                        const void* coset, hipStream_t s, void* h_out = nullptr);
void launch_bitrev_permute(void* dst, const void* src, int log_n, hipStream_t s);
void launch_scale(void* data, size_t n, const void* factor, hipStream_t s);
void launch_scale_rp(void* data, size_t n, const void* factor_rp, hipStream_t s);
void launch_mul_table_rp(void* data, size_t n, const void* table_rp, hipStream_t s);

// ---- prove stages (kernels_prove.hip) ----
void launch_gather(const uint32_t* row_ptr, const uint32_t* signal, const uint32_t* value_id,
                   const void* values, const void* witness, void* out, uint32_t n_rows,
                   hipStream_t s);
// A = A.w, B = B.w and C = A o B row by row in one kernel (matrices 0 = A, 1 = B), for
// `batch` witnesses of nv elements back to back; proof k's (A, B, C) go to abc + 3 k n_rows.
void launch_gather_abc(const uint32_t* const row_ptr[2], const uint32_t* const signal[2],
                       const uint32_t* const value_id[2], const void* values,
                       const void* witness, size_t nv, void* abc, uint32_t n_rows,
                       uint32_t batch, hipStream_t s);
void launch_mul_into(void* out, const void* a, const void* b, size_t n, hipStream_t s);
void launch_h(void* h, const void* a, const void* b, const void* c, size_t n, hipStream_t s);
void launch_rows_abc(void* a, void* b, void* c, size_t rows, size_t n, hipStream_t s);

// ---- MSM (kernels_msm.hip) ----
// Window widths the kernels take: up to 20 (buckets of 2^19 per set; the 32-bit digit code has
// 24 bits for |d| - 1). The standalone entry points keep to 16.
constexpr int MSM_MAX_C = 20;
struct MsmPlan {
  int c = 0;              // window width
  int windows = 0;        // ceil(255 / c)
  uint32_t n = 0;         // scalars per vector
  uint32_t nb = 0;        // buckets per window, 2^(c-1)
  // Scalar vectors sorted and accumulated together against the same bases (the proofs of one
  // batch): vector k is scalars [k n, (k + 1) n), digits are laid out per (k, window), and
  // every vector has bucket sets of its own (set index k, or k windows + w per window).
  uint32_t batch = 1;
  size_t max_entries = 0; // batch * n * windows
  // The two-level sort: coarse bins (bucket >> shift) per set, digit rows per set (windows
  // when merged, else 1), scalars per chunk, chunks, and the count array's length.
  // The fine level: coarse groups (sets * cbins) and an upper bound on their slices.
  int shift = 0;
  uint32_t cbins = 1, rows_per_set = 1, chunk = 0, chunks = 0;
  size_t q = 0;
  uint32_t groups = 0, max_slices = 0;
  // Merged windows: the bases come as window tables (entry w n + i = 2^(c w) P_i), so the
  // digits of every window fall into one set of nb buckets and the window sum is the MSM.
  bool merged = false;
  // Constant work (opt in): every (scalar, window) emits an entry, so the entry count is
  // n * windows whatever the scalars; a zero digit goes to bucket (i mod nb) flagged as
  // discarded, and the accumulation computes its addition and keeps the old sum by a select.
  // Entry values are then index << 2 | discard << 1 | sign (index < 2^30).
  bool constant = false;
  // device scratch (owned by the caller)
  uint32_t* offset = nullptr;   // [K + 1] first entry of each bucket (K = sets * nb); [K] = total
  uint32_t* block_sums = nullptr;
  uint32_t* ccount = nullptr;   // [q + 1] entries per (set, coarse bin, row of the set, chunk)
  uint32_t* cpos = nullptr;     // [q + 1] their exclusive prefix: each run's first position
  uint2* ent = nullptr;         // [max_entries] sorted entries: {key = set * nb + bucket,
                                //                 value = index << 1 | sign}
  uint2* tmp = nullptr;         // [max_entries] coarse-sorted entries {fine bits, value}
  uint32_t* sprefix = nullptr;  // [groups + 1] slices before each coarse group
  uint32_t* fcnt = nullptr;     // [2^shift * max_slices] fine counts, then positions
  uint32_t* digits = nullptr;   // [rows * n] encoded signed digits (16 or 32-bit), row-major
  uint32_t* total = nullptr;    // [4] entry count
};
// Bucket sets the accumulation and reduction work on: one per window, or one when merged, for
// each scalar vector.
inline int msm_bucket_sets(const MsmPlan& p) {
  return (int)p.batch * (p.merged ? 1 : p.windows);
}
// The plan as seen by one vector's host finish (its rows are msm_out_points of this view).
inline MsmPlan msm_vector_view(const MsmPlan& p) {
  MsmPlan v = p;
  v.batch = 1;
  v.max_entries = (size_t)p.n * p.windows;
  return v;
}
size_t msm_sort_scratch_bytes(uint32_t n, int c, uint32_t batch = 1);
void msm_plan_bind(MsmPlan& p, void* scratch, uint32_t n, int c, bool merged = false,
                   bool constant = false, uint32_t batch = 1);
// Window tables for merged MSMs: `table` holds n affine bases (MSM form) at [0, n); fills
// entries w n + i = 2^(c w) P_i for 1 <= w < windows. group 1 = G1, 2 = G2.
void msm_window_tables(int group, void* table, size_t n, int c, int windows, hipStream_t s);
void msm_window_tables_g2(void* table, size_t n, int c, int windows, hipStream_t s);
// Digits of `n` Montgomery scalars into (window, bucket)-sorted entries.
void msm_sort(MsmPlan& p, const void* scalars_mont, hipStream_t s);
size_t msm_accum_scratch_bytes(const MsmPlan& p, int group);
// Bucket accumulation + window reduction for one base vector over the sorted entries;
// writes msm_out_points(p) XYZZ points to `window_out` (device memory): `rows` per window,
// row 0 = sum_g U_g and row k + 1 = sum of T_g over the groups g with bit k set, where the
// window's buckets fall in groups of 2^lg, T_g is a group's sum and U_g its sum weighted
// 1, 2, ..., 2^lg. The window sum is then S_w = row 0 + sum_k 2^(k + lg) row k+1
// (msm_combine on the host).
// If `ev` is given, ev[0..3] are recorded before the accumulation kernel, after it, after the
// merges and after the window reduction.
void msm_accumulate(const MsmPlan& p, int group, const void* bases, void* scratch,
                    void* window_out, hipStream_t s, hipEvent_t* ev = nullptr);
// The G1 bucket accumulations of n <= 4 MSMs (plans[k] sorted, bases[k] in the MSM form,
// scratch[k] of msm_accum_scratch_bytes) as ONE persistent launch whose workgroups take tiles
// of segments of every job from a device-side counter (a device word of the caller, 16 bytes,
// cleared by a kernel here), then each MSM's merges and reduction into window_out[k].
void msm_accumulate_g1_queue(const MsmPlan* const* plans, const void* const* bases,
                             void* const* scratch, void* const* window_out, int n,
                             uint32_t* counter, hipStream_t s);
void msm_reduce_shape(const MsmPlan& p, int* rows, int* lg);
// The MSM kernels keep coordinates in Montgomery form with R' = 2^261 (bn254.cuh Fm): bases go
// in as 32 x mod q. dst may equal src; n_fq counts 32-byte coordinates.
void msm_to_msm_form(const void* src, void* dst, size_t n_fq, hipStream_t s);
size_t msm_out_points(const MsmPlan& p);
int msm_default_c(uint32_t n);
// Width for scalars sorted once for g1 G1 MSMs and g2 G2 MSMs (per-window buckets), and the
// same over window tables (one bucket set).
int msm_choose_c(uint32_t n, int g1, int g2);
int msm_choose_c_merged(uint32_t n, int g1, int g2);

}  // namespace g16
