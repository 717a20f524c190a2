/*
 * g16hip — the C ABI of the MI355X (gfx950) Groth16 prover back end.
 *
 * This is the boundary a thin Rust crate `crates/hip` would bind to implement the
 * reference's back-end contract, `snarkrs_groth16::{Backend, PreparedCircuit}`
 * (crates/groth16/src/lib.rs:371-455), and its standalone MSM plugin API,
 * `snarkrs_msm::MsmBackend` (crates/msm/src/lib.rs:52-82). Every entry point below names
 * the reference interface it replaces.
 *
 * Wire format (crates/gpu-layout/src/lib.rs:75-185, crates/formats/src/binfile.rs:10-25):
 *   - a field element is 32 bytes, 8 little-endian u32 limbs. Fr values (witness, r, s,
 *     key coefficients, H) are MONTGOMERY form with R = 2^256 (arkworks' own limbs, so a
 *     Rust caller passes `Fr.0.0` unchanged);
 *   - a G1 affine point is x then y (64 bytes), a G2 affine point x then y with each Fq2
 *     stored c0 then c1 (128 bytes); coordinates are Montgomery; the all-zero encoding is
 *     the point at infinity;
 *   - every output point is affine in the same encoding.
 *
 * Errors: every call returns a g16hip_status that maps 1:1 onto ProveError
 * (crates/groth16/src/lib.rs:79-154); g16hip_last_error() returns the message of the last
 * failure on the calling thread.
 *
 * Threading: a prepared circuit may be used from many threads at once
 * (crates/groth16/src/lib.rs:380-381); calls on one circuit are serialised internally on
 * that circuit's own HIP stream and scratch space. Different circuits run concurrently.
 */
#ifndef G16HIP_H
#define G16HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ProveError, crates/groth16/src/lib.rs:79-154. */
typedef enum {
  G16HIP_OK = 0,
  G16HIP_WITNESS_LENGTH = 1, /* ProveError::WitnessLength { got, want } */
  G16HIP_BACKEND = 2,        /* ProveError::Backend: malformed key/arguments, deterministic */
  G16HIP_DEVICE = 3,         /* ProveError::Device: a device fault, may succeed on retry */
  G16HIP_DEVICE_MEMORY = 4,  /* ProveError::DeviceMemory: refused before allocating */
  G16HIP_CONSTANT_WIRE = 5,  /* ProveError::ConstantWire: witness[0] != 1 */
  G16HIP_ZERO_BLINDER = 6,   /* ProveError::ZeroBlinder: r or s is zero */
  G16HIP_SELF_VERIFY = 7,    /* ProveError::SelfVerify: the proof failed the check of its result */
  /* VerifyError, crates/groth16/src/verify.rs:16-31 */
  G16HIP_PUBLIC_COUNT = 8,   /* VerifyError::PublicInputCount { got, want } */
  G16HIP_EMPTY_KEY = 9,      /* VerifyError::EmptyVerifyingKey */
  G16HIP_PAIRING_FAILED = 10,/* VerifyError::PairingFailed */
  G16HIP_INVALID_PROOF = 11, /* VerifyError::InvalidProof: off the curve or subgroup, or infinity */
  G16HIP_UNSAFE_KEY = 12,    /* VerifyError::UnsafeKey (ZkeyError::UnsafeVerifyingKey) */
  G16HIP_BAD_JSON = 13       /* JsonError (crates/groth16/src/json.rs:39-41), ZkeyError::BadJson */
} g16hip_status;

typedef struct g16hip_device g16hip_device;
typedef struct g16hip_circuit g16hip_circuit;
/* HPoly::Device (crates/groth16/src/lib.rs:194-218): H left on the device by compute_h. */
typedef struct g16hip_h g16hip_h;

/*
 * A parsed proving key, host memory, borrowed for the duration of g16hip_prepare.
 * Field meanings follow `ProvingKey` (crates/formats/src/lib.rs:35-75) with section 4 in
 * CSR form (`Coefficients`, crates/formats/src/lib.rs:118-150): for matrix m (0 = A,
 * 1 = B), row c's nonzeros are k in [row_ptr[m][c], row_ptr[m][c+1]); each is
 * values[value_id[m][k]] * witness[signal[m][k]]; values are Fr Montgomery (the section 4
 * double-Montgomery bytes already divided by R, crates/formats/src/binfile.rs:488-503).
 */
typedef struct {
  uint32_t n_vars;
  uint32_t n_public;
  uint32_t domain_size; /* power of two, <= 2^27 */
  uint32_t n_values;
  const uint8_t* alpha_g1; /* G1 */
  const uint8_t* beta_g1;  /* G1 */
  const uint8_t* beta_g2;  /* G2 */
  const uint8_t* gamma_g2; /* G2 */
  const uint8_t* delta_g1; /* G1 */
  const uint8_t* delta_g2; /* G2 */
  const uint8_t* ic;       /* G1[n_public + 1] */
  const uint32_t* row_ptr[2];  /* [domain_size + 1] each */
  const uint32_t* signal[2];   /* [row_ptr[m][domain_size]] */
  const uint32_t* value_id[2]; /* [row_ptr[m][domain_size]] */
  const uint8_t* values;       /* Fr[n_values] */
  const uint8_t* a_query;    /* section 5, G1[n_vars] */
  const uint8_t* b_g1_query; /* section 6, G1[n_vars] */
  const uint8_t* b_g2_query; /* section 7, G2[n_vars] */
  const uint8_t* l_query;    /* section 8, G1[n_vars - n_public - 1] */
  const uint8_t* h_query;    /* section 9, G1[domain_size]; gnark: G1[domain_size - 1] */
  /* HBasis (crates/formats/src/lib.rs:66-89): 0 = snarkjs' odd coset (every zkey), 1 = gnark's
   * pk.G1.Z, [tau^bitrev(j) t(tau) / delta]_1 for j < domain_size - 1. A key may leave the
   * matrices (row_ptr, signal, value_id, values) NULL: it then proves from A.w and B.w rows
   * only (g16hip_h_and_msms_from_rows), as a gnark key, which carries no matrices, must. */
  uint32_t h_basis;
} g16hip_key;

/* StageTimings, crates/groth16/src/lib.rs:322-368 (microseconds; calls add to them). */
typedef struct {
  uint64_t gather_us;
  uint64_t ntt_us;
  uint64_t pointwise_us;
  uint64_t msm_us;
  uint64_t assemble_us;
  uint64_t verify_us;
} g16hip_timings;

/* MsmOutputs, crates/groth16/src/lib.rs:311-318, each point affine. */
typedef struct {
  uint8_t a_g1[64];
  uint8_t b_g2[128];
  uint8_t b_g1[64];
  uint8_t l_g1[64];
  uint8_t h_g1[64];
} g16hip_msm_out;

/* Proof, crates/groth16/src/lib.rs:188-192. */
typedef struct {
  uint8_t a[64];
  uint8_t b[128];
  uint8_t c[64];
} g16hip_proof;

/* VerifyingKey (crates/formats/src/lib.rs:353-360): points in the wire encoding above,
 * borrowed for the call. ic holds n_ic = n_public + 1 G1 points, the constant wire's first. */
typedef struct {
  const uint8_t* alpha_g1; /* G1 */
  const uint8_t* beta_g2;  /* G2 */
  const uint8_t* gamma_g2; /* G2 */
  const uint8_t* delta_g2; /* G2 */
  const uint8_t* ic;       /* G1[n_ic] */
  uint32_t n_ic;
} g16hip_vk;

/* Message of the last failed call on this thread ("" if none). */
const char* g16hip_last_error(void);

/* Opens GPU `ordinal`; replaces CudaBackend::new (crates/cuda/src/backend.rs:152-184). */
g16hip_status g16hip_open(int ordinal, g16hip_device** out);
void g16hip_close(g16hip_device* dev);
/* Free device memory in bytes, the admission budget (crates/cuda/src/backend.rs:626-643). */
g16hip_status g16hip_mem_info(g16hip_device* dev, size_t* free_bytes, size_t* total_bytes);
/* CudaBackend::with_vram_limit (crates/cuda/src/backend.rs:120-130): caps the device memory a
 * circuit prepared after this call (its resident key and one proof slot) and each further slot
 * may plan against, below what the driver reports free (0 = no cap). A key over the budget is
 * refused by g16hip_prepare with G16HIP_DEVICE_MEMORY before anything is allocated
 * (crates/cuda/src/backend.rs:367-381). G16HIP_VRAM_LIMIT (bytes) sets the same at
 * g16hip_open; G16HIP_VRAM_RESERVE (bytes) leaves that much of the free memory to others. */
g16hip_status g16hip_set_vram_limit(g16hip_device* dev, size_t bytes);
/* The budget an allocation may plan against now (device_budget, crates/cuda/src/msm.rs:1057-1091):
 * free memory less a 256 MiB headroom and the reserve, capped by the limit. */
g16hip_status g16hip_device_budget(g16hip_device* dev, size_t* bytes);
/* resident_bytes (crates/cuda/src/backend.rs:626-643): device bytes `key` and one proof slot take
 * once prepared, with (window_tables != 0) or without the per-window base tables, each buffer
 * rounded up to 2 MiB pages. g16hip_prepare uses the tables only when the budget keeps 8 GiB
 * free after them. */
g16hip_status g16hip_key_resident_bytes(const g16hip_key* key, int window_tables, size_t* bytes);

/* Backend::prepare (crates/groth16/src/lib.rs:371-378): validates the key
 * (crates/groth16/src/cpu.rs:73-122), uploads CSR, twiddles, coset table and the five base
 * vectors, and keeps them resident. The key is only read. A key whose resident data and first
 * proof slot exceed the device budget (g16hip_device_budget) is G16HIP_DEVICE_MEMORY with
 * nothing allocated. */
g16hip_status g16hip_prepare(g16hip_device* dev, const g16hip_key* key, g16hip_circuit** out);
void g16hip_circuit_release(g16hip_circuit* c);
/* PreparedCircuit::{n_vars, n_public, domain_size} (crates/groth16/src/lib.rs:382-387). */
g16hip_status g16hip_circuit_dims(const g16hip_circuit* c, uint32_t* n_vars, uint32_t* n_public,
                                  uint32_t* domain_size);
/* Device bytes this circuit holds resident (key + per-proof scratch). */
size_t g16hip_circuit_resident_bytes(const g16hip_circuit* c);

/* PreparedCircuit::compute_h (crates/groth16/src/lib.rs:398-399): stages 0-4. `witness` is
 * Fr[n_witness] Montgomery in host memory; H stays on the device in *h_out. */
g16hip_status g16hip_compute_h(g16hip_circuit* c, const uint8_t* witness, size_t n_witness,
                               g16hip_h** h_out, g16hip_timings* t);
/* PreparedCircuit::h_to_host (crates/groth16/src/lib.rs:446-454): H (Fr[domain_size],
 * Montgomery) copied down. Debug/test only. */
g16hip_status g16hip_h_to_host(g16hip_circuit* c, const g16hip_h* h, uint8_t* out);
void g16hip_h_release(g16hip_h* h);

/* PreparedCircuit::msms (crates/groth16/src/lib.rs:401-410): stages 5-9. `h` must come from
 * g16hip_compute_h on this circuit; another circuit's handle is refused (require_handle,
 * crates/groth16/src/lib.rs:285-307). */
g16hip_status g16hip_msms(g16hip_circuit* c, const uint8_t* witness, size_t n_witness,
                          const g16hip_h* h, g16hip_msm_out* out, g16hip_timings* t);

/* PreparedCircuit::h_and_msms (crates/groth16/src/lib.rs:412-423): stages 0-9. */
g16hip_status g16hip_h_and_msms(g16hip_circuit* c, const uint8_t* witness, size_t n_witness,
                                g16hip_msm_out* out, g16hip_timings* t);
/* The same with the witness already in device memory (Fr[n_witness] Montgomery). */
g16hip_status g16hip_h_and_msms_device(g16hip_circuit* c, const void* d_witness,
                                       size_t n_witness, g16hip_msm_out* out, g16hip_timings* t);
/* PreparedCircuit::h_and_msms_from_rows (crates/groth16/src/lib.rs:425-444): stages 1-9 with
 * stage 0 done by the caller. a_rows[i], b_rows[i] are constraint i's A.w and B.w (Fr
 * Montgomery, host memory), n_rows <= domain_size of them, every later row zero; C = A*B on
 * the rows. A gnark-basis key takes H's coefficients in bit-reversed order (the conversion
 * runs on the device); a nonzero coefficient n - 1 is refused as the reference refuses it
 * (crates/groth16/src/cpu.rs:265-286). */
g16hip_status g16hip_h_and_msms_from_rows(g16hip_circuit* c, const uint8_t* witness,
                                          size_t n_witness, const uint8_t* a_rows,
                                          const uint8_t* b_rows, size_t n_rows,
                                          g16hip_msm_out* out, g16hip_timings* t);
/* prove_from_rows (crates/groth16/src/prove.rs:93-115): the proof from rows with caller r, s,
 * checked against the circuit's verifying key like g16hip_prove (G16HIP_SELF_VERIFY). */
g16hip_status g16hip_prove_from_rows(g16hip_circuit* c, const uint8_t* witness, size_t n_witness,
                                     const uint8_t* a_rows, const uint8_t* b_rows, size_t n_rows,
                                     const uint8_t* r, const uint8_t* s, g16hip_proof* proof,
                                     g16hip_timings* t);

/* prove_with_blinders (crates/groth16/src/prove.rs:175-229): stages 0-11 with caller r, s
 * (Fr Montgomery), checking the witness shape first (prove.rs:272-283). Assembly (stage 11)
 * runs on the host like the reference's, its blinder terms overlapped with the GPU work. */
g16hip_status g16hip_prove_with_blinders(g16hip_circuit* c, const uint8_t* witness,
                                         size_t n_witness, const uint8_t* r, const uint8_t* s,
                                         g16hip_proof* proof, g16hip_timings* t);
g16hip_status g16hip_prove_with_blinders_device(g16hip_circuit* c, const void* d_witness,
                                                size_t n_witness, const uint8_t* r,
                                                const uint8_t* s, g16hip_proof* proof,
                                                g16hip_timings* t);

/* prove (crates/groth16/src/prove.rs:12-41, :79-91): g16hip_prove_with_blinders, then the
 * check of the result against the circuit's own verifying key (g16hip_verify below); a proof
 * that fails it is G16HIP_SELF_VERIFY with the verifier's message. r and s come from the
 * caller's CSPRNG (the reference's sample_blinders, prove.rs:138-146); zero is refused. The
 * check's time goes to t->verify_us. */
g16hip_status g16hip_prove(g16hip_circuit* c, const uint8_t* witness, size_t n_witness,
                           const uint8_t* r, const uint8_t* s, g16hip_proof* proof,
                           g16hip_timings* t);
/* PreparedCircuit::key().vk: alpha_g1 | beta_g2 | gamma_g2 | delta_g2 | ic[n_public + 1], in
 * the wire encoding (64 + 3 * 128 + 64 * (n_public + 1) bytes). */
g16hip_status g16hip_circuit_vk(const g16hip_circuit* c, uint8_t* out);

/* ---- Verification and snarkjs files (host code; no device needed). ---- */

/* verify (crates/groth16/src/verify.rs:49-52): the key's structure checked
 * (crates/formats/src/lib.rs:1326-1350, gamma_g2 != delta_g2 included), A and C on the curve,
 * B on the curve and in the prime-order subgroup, none of them infinity, then
 * e(-A, B) e(alpha, beta) e(L_bar, gamma) e(C, delta) = 1 with
 * L_bar = IC[0] + sum_i public[i] IC[i + 1]. public_inputs: Fr[n_public] Montgomery.
 * G16HIP_OK, or the VerifyError status. */
g16hip_status g16hip_verify(const g16hip_vk* vk, const uint8_t* public_inputs, size_t n_public,
                            const g16hip_proof* proof);
/* verify_allowing_unblinded (verify.rs:66-76): every check but gamma_g2 != delta_g2. */
g16hip_status g16hip_verify_allowing_unblinded(const g16hip_vk* vk, const uint8_t* public_inputs,
                                               size_t n_public, const g16hip_proof* proof);
/* verify_unchecked (verify.rs:104-134): the bare pairing check. */
g16hip_status g16hip_verify_unchecked(const g16hip_vk* vk, const uint8_t* public_inputs,
                                      size_t n_public, const g16hip_proof* proof);

/* The optimal ate pairing (ark-bn254's Bn254::pairing, the reference's verifier arithmetic):
 * *is_one = whether prod_i e(g1[i], g2[i]) = 1 (Bn254::multi_pairing). Points must be valid. */
g16hip_status g16hip_pairing_check(const uint8_t* g1, const uint8_t* g2, size_t n, int* is_one);
/* e(g1, g2) in Fq12 = Fq6[w]/(w^2 - v), Fq6 = Fq2[v]/(v^3 - (9 + u)): 12 Fq Montgomery
 * coordinates, c0.{a0,a1,a2} then c1.{a0,a1,a2}, each Fq2 c0 then c1 (384 bytes). */
g16hip_status g16hip_pairing(const uint8_t* g1, const uint8_t* g2, uint8_t* out);
/* *valid = the point decodes, is on the curve and (group 2) in the prime-order subgroup. */
g16hip_status g16hip_point_check(int group, const uint8_t* point, int* valid);

/* proof_to_string / public_to_string (crates/groth16/src/json.rs:97-114): snarkjs'
 * proof.json and public.json text, NUL-terminated. With buf NULL only *len (the text's length
 * without the NUL) is set; otherwise cap must exceed it. */
g16hip_status g16hip_proof_to_json(const g16hip_proof* proof, char* buf, size_t cap, size_t* len);
g16hip_status g16hip_public_to_json(const uint8_t* public_inputs, size_t n, char* buf, size_t cap,
                                    size_t* len);
/* proof_from_str / public_from_str (json.rs:150-243): strict readers; coordinates must be
 * below their modulus, z exactly 1 (or ffjavascript's infinity), points on the curve and in
 * the subgroup. public: *n receives the count; with out NULL nothing else is written, else out
 * holds cap scalars (Fr Montgomery). G16HIP_BAD_JSON with the reason on a malformed file. */
g16hip_status g16hip_proof_from_json(const char* text, size_t len, g16hip_proof* out);
g16hip_status g16hip_public_from_json(const char* text, size_t len, uint8_t* out, size_t cap,
                                      size_t* n);
/* VerifyingKey::from_json_str (check_structure != 0) / from_json_str_unchecked
 * (crates/formats/src/lib.rs:1371-1430): snarkjs' verification_key.json. *n_public receives
 * nPublic; out (cap bytes, NULL to ask for n_public only) receives the g16hip_circuit_vk
 * layout. */
g16hip_status g16hip_vk_from_json(const char* text, size_t len, int check_structure, uint8_t* out,
                                  size_t cap, uint32_t* n_public);

/* ---- Batched proving and serving (no counterpart in the reference, whose provers run one
 * proof per call; these are for callers that keep many proofs in flight on one GPU). ---- */

/* Adds up to `count` proof slots to `c` (each: streams + per-proof scratch, so one more proof
 * can be in flight), stopping when the device would keep less than `reserve_bytes` free:
 * admission by free device memory. *total_slots receives the slot count. A circuit starts
 * with one slot. */
g16hip_status g16hip_circuit_add_slots(g16hip_circuit* c, uint32_t count, size_t reserve_bytes,
                                       uint32_t* total_slots);

typedef struct {
  g16hip_circuit* circuit;
  const uint8_t* witness; /* Fr[n_witness] Montgomery, host memory */
  size_t n_witness;
  const uint8_t* r; /* Fr Montgomery */
  const uint8_t* s;
} g16hip_batch_item;

/* Batches and the server run each proof on a slot's three streams, kernel by kernel (the
 * default), or, with graph replay on, replay the slot's proof pipeline captured once as a
 * hipGraph (stages 0-9 with their fork/join; no memset nodes). Process-wide; the default can
 * also be set with G16HIP_GRAPHS=1. Results are bit-identical either way. */
void g16hip_set_graph_replay(int on);

/* Proves every item (any mix of circuits). Each circuit's items run in waves of up to its
 * batch cap: a wave is up to cap proofs of that circuit proved together, every stage one launch
 * over all of them (their witnesses staged and uploaded together, one fused gather, the
 * transforms of all their vectors, each MSM's sort / accumulation / merge / reduction over all
 * their scalar vectors with a bucket set per proof), on one stream per circuit, the host
 * finishing wave i while the device runs wave i + 1. Circuits run concurrently. Keys without
 * matrices or with gnark's H basis prove item by item. proofs[i] is bit-identical to
 * g16hip_prove_with_blinders on item i. statuses (optional) receives each item's status; the
 * return value is the first failure's. A wave larger than the device budget is refused as
 * G16HIP_DEVICE_MEMORY before anything is allocated. */
g16hip_status g16hip_prove_batch(const g16hip_batch_item* items, size_t n, g16hip_proof* proofs,
                                 g16hip_status* statuses);
/* The batch cap (proofs per wave) of circuit c; 0 = the library's default (about 2^21 transform
 * elements per wave, at most 16 proofs). At most 64. */
g16hip_status g16hip_set_batch_cap(g16hip_circuit* c, uint32_t cap);
/* Measures batched throughput of circuit c on the given witness (Fr[n_witness] Montgomery,
 * host memory) for caps 1, 2, 4, ... up to max_cap (0 = 32) while the device budget holds a wave,
 * stopping after two caps that do not improve on the best by 2%, and sets the best as c's cap
 * (*best). caps[i] / rates[i] (optional, len entries) receive each cap tried and its proofs/s;
 * *tried (optional) the number tried. The reference has no batching; this replaces its fixed
 * launch shapes (crates/cuda/src/msm.rs:744-853) for batches by a measurement per (n, batch). */
g16hip_status g16hip_batch_autotune(g16hip_circuit* c, const uint8_t* witness, size_t n_witness,
                                    uint32_t max_cap, uint32_t* caps, double* rates, size_t len,
                                    size_t* tried, uint32_t* best);

/* A serving loop with continuous batching. `lanes` proofs are in flight at once (a host
 * thread each; 1..64): a free lane takes the oldest request of any circuit, leases a slot of
 * that circuit — adding one, up to `lanes` per circuit, while the device keeps
 * `reserve_bytes` free (admission by free device memory), else waiting for one — and proves on
 * it. submit returns at once; the witness stays owned by the
 * caller and must stay valid until wait returns for its ticket. wait blocks for a ticket's
 * proof and reports its latency (submit to completion). Slots' streams spread over the
 * process's hardware queues (GPU_MAX_HW_QUEUES, 4 by default) and streams that share a queue
 * run one after the other, so more queues let more proofs overlap. */
typedef struct g16hip_server g16hip_server;
g16hip_status g16hip_server_open(g16hip_device* dev, size_t reserve_bytes, uint32_t lanes,
                                 g16hip_server** out);
g16hip_status g16hip_server_submit(g16hip_server* srv, g16hip_circuit* c, const uint8_t* witness,
                                   size_t n_witness, const uint8_t* r, const uint8_t* s,
                                   uint64_t* ticket);
g16hip_status g16hip_server_wait(g16hip_server* srv, uint64_t ticket, g16hip_proof* proof,
                                 double* latency_ms);
/* Finishes the queued requests, then stops the workers. */
void g16hip_server_close(g16hip_server* srv);

/* MsmBackend::try_msm_g1 / try_msm_g2 (crates/msm/src/lib.rs:61-81): bases affine, scalars
 * Fr Montgomery (as `&[Fr]` is), host memory; result affine. Empty input is infinity. */
g16hip_status g16hip_msm_g1(g16hip_device* dev, const uint8_t* bases, const uint8_t* scalars,
                            size_t n, uint8_t* out);
g16hip_status g16hip_msm_g2(g16hip_device* dev, const uint8_t* bases, const uint8_t* scalars,
                            size_t n, uint8_t* out);
/* The same with bases and scalars already in device memory; window width c (0 = the
 * library's choice). Used for the resident-input MSM throughput figure. */
g16hip_status g16hip_msm_g1_device(g16hip_device* dev, const void* d_bases, const void* d_scalars,
                                   size_t n, int c, uint8_t* out);
g16hip_status g16hip_msm_g2_device(g16hip_device* dev, const void* d_bases, const void* d_scalars,
                                   size_t n, int c, uint8_t* out);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc). The device MSM of
 * g16hip_msm_g1_device / g16hip_msm_g2_device (group 1 or 2, same arguments, same result bit
 * for bit) with every piece of its scratch (the sort's digits, counts and entries, the
 * accumulation's buckets and segment pieces, the bases in the MSM form, the window sums) in a
 * static device array of G16HIP_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds
 * until the process ends; it calls no hipMalloc or hipFree. An MSM whose scratch exceeds the
 * array is G16HIP_DEVICE_MEMORY (2^17 G1 or 2^16 G2 points at the library's width fit). Calls
 * through one device handle serialise; calls that overlap on one GPU through different handles
 * share the array and race. The array keeps the last call's contents; nothing in it needs to
 * start at zero. */
#define G16HIP_NOALLOC_SCRATCH_BYTES ((size_t)96 << 20)
g16hip_status g16hip_msm_device_noalloc(g16hip_device* dev, int group, const void* d_bases,
                                        const void* d_scalars, size_t n, int c, uint8_t* out);

/* Work::Constant (crates/cuda/src/msm.rs:143-164, CudaBackend::with_constant_work,
 * crates/cuda/src/backend.rs:100-118), opt in: for standalone MSMs on dev and circuits
 * prepared on dev after the call, every (scalar, window) emits an entry, zero digits into
 * their window's buckets (i mod 2^(c-1)) as entries whose addition is computed and discarded,
 * ones as ordinary digits, so the entry count, the sort and the accumulation grid are
 * n * windows whatever the scalars. Results are bit-identical. Not constant time: bucket
 * sizes, the merge of buckets that span segments and the point formulas' special cases still
 * follow the scalars. Off by default. */
g16hip_status g16hip_set_constant_work(g16hip_device* dev, int on);

/* Phase timings of the last standalone MSM on `dev` (HIP events on the stream its kernels
 * ran on), the per-stage split the reference's StageTimings gives for a proof
 * (crates/groth16/src/lib.rs:322-368) at kernel granularity. */
typedef struct {
  float sort_ms;       /* digits + histogram + scan + scatter */
  float accumulate_ms; /* the bucket-accumulation kernel alone */
  float merge_ms;      /* spanning and heavy buckets */
  float reduce_ms;     /* window reduction */
  float total_ms;      /* first kernel start to window sums on the host */
  uint64_t entries;    /* sorted (window, bucket) entries = bucket additions */
  uint32_t c;          /* window width used */
  uint32_t windows;
} g16hip_msm_stats;
g16hip_status g16hip_msm_last_stats(g16hip_device* dev, g16hip_msm_stats* out);

/* Running totals over every bucket-accumulation kernel launched on `dev` (standalone MSMs and
 * proofs alike) since the last reset, index 0 = G1, 1 = G2: the per-launch figures a roofline
 * needs (work = entries, time = HIP events on the launch stream). */
typedef struct {
  uint64_t launches[2];
  uint64_t entries[2];
  double accumulate_ms[2];
} g16hip_kernel_stats;
g16hip_status g16hip_kernel_stats_get(g16hip_device* dev, g16hip_kernel_stats* out);
g16hip_status g16hip_kernel_stats_reset(g16hip_device* dev);

/* The ceremony's MsmBackend (crates/cuda/src/ceremony.rs:65-312, the CudaMsmBackend behind
 * `groth16 setup` and `zkey verify`'s recompute): MSMs over host slices (bases affine,
 * scalars Fr Montgomery), each call cut into runs of at most max_run points (0 = 2^21) whose
 * sums are added on the host. A run refused for device memory is halved, down to 2^10
 * points; with vram_limit != 0 a run whose scratch would exceed it is refused before anything
 * is allocated (the reference's with_vram_limit). A run whose sum is not on the curve is
 * re-run, 8 attempts in all, then G16HIP_DEVICE. Calls on one handle run one at a time; every
 * call goes to the device (the reference sends calls under 2^12 points to its CPU MSM; this
 * library has no CPU path). */
typedef struct g16hip_ceremony g16hip_ceremony;
typedef struct {
  uint64_t calls, runs, points, retries;
  double ms;
} g16hip_ceremony_stats;
g16hip_status g16hip_ceremony_open(g16hip_device* dev, size_t max_run, size_t vram_limit,
                                   g16hip_ceremony** out);
void g16hip_ceremony_close(g16hip_ceremony* c);
/* group 1 or 2; out: the affine sum (64 or 128 bytes). Empty input is infinity. */
g16hip_status g16hip_ceremony_msm(g16hip_ceremony* c, int group, const uint8_t* bases,
                                  const uint8_t* scalars, size_t n, uint8_t* out);
/* Device bytes a standalone MSM of n points takes (uploaded inputs, the bases' MSM form,
 * sort and accumulation scratch): what vram_limit is compared with. */
g16hip_status g16hip_msm_scratch_bytes(int group, size_t n, size_t* bytes);
/* Runs the last call took (CudaMsmBackend::last_runs), and the running tally. */
g16hip_status g16hip_ceremony_last_runs(g16hip_ceremony* c, uint64_t* runs);
g16hip_status g16hip_ceremony_stats_get(g16hip_ceremony* c, g16hip_ceremony_stats* out);

/* Standalone NTT over Fr, in place on host memory (Fr[2^log_n] Montgomery, natural order):
 * kind 0 = forward, 1 = inverse (with 1/n), 2 = the prover's coset pipeline
 * iNTT -> x shift^j -> NTT (CpuNtt::{ntt, intt_coset_ntt}, crates/ntt/src/lib.rs:415, :508). */
g16hip_status g16hip_ntt(g16hip_device* dev, uint8_t* data, int log_n, int kind);
g16hip_status g16hip_ntt_device(g16hip_device* dev, void* d_data, int log_n, int kind);

/* Field arithmetic probe (the role of crates/cuda/src/kernels/field_probe.cu:16-112):
 * field 0 = Fr, 1 = Fq, 2 = Fq2, 3 = Fq in the MSM kernels' Montgomery form R' = 2^261 (no
 * reference counterpart); op 0 add, 1 sub, 2 mul, 3 sqr, 4 inv, 5 neg, 6 to Montgomery,
 * 7 from Montgomery. n elements, host memory. */
g16hip_status g16hip_field_probe(g16hip_device* dev, int field, int op, const uint8_t* a,
                                 const uint8_t* b, uint8_t* out, size_t n);

/* VALU ceilings for the roofline, measured the way the reference's device microbenchmark is
 * (bench/results/device-microbench.md:7-10): kind 0 = 32x32+64 multiply-adds per second
 * (independent chains), kind 1 = Fq Montgomery products per second. */
g16hip_status g16hip_valu_probe(g16hip_device* dev, int kind, double* per_second);

/* Batched fixed-base multiplication k_i * G (G the standard generator of `group`, 1 or 2),
 * k Fr Montgomery, output affine: the device half of the synthetic setup that builds the
 * proving keys the benchmarks use (the role of crates/ceremony's setup MSM seam). */
g16hip_status g16hip_fixed_base(g16hip_device* dev, int group, const uint8_t* k, size_t n,
                                uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif /* G16HIP_H */
