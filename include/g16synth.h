/*
 * g16synth — synthetic Groth16 workloads for benchmarks and tests of libg16hip.
 *
 * No counterpart in the reference's proving API: this is the host half of the "repo's own
 * setup path" that produces the benchmark keys (the reference uses crates/ceremony's
 * `groth16 setup` on real circuits). It builds a seeded R1CS whose witness is computed
 * forward, a seeded toxic waste (tau, alpha, beta, gamma, delta) and every query scalar of
 * the snarkjs key semantics; the points are then k * G through g16hip_fixed_base on the GPU.
 * The construction is specified step by step in DESIGN.md and is bit-identical to the CPU
 * oracle's (tests/test_synth.py).
 */
#ifndef G16SYNTH_H
#define G16SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t n_constraints;
  uint32_t n_public;           /* public inputs: wires 1..=n_public */
  uint32_t n_private_in;       /* private inputs after them */
  uint32_t bool_permille;      /* share of constraints whose output wire is a bit */
  uint32_t input_bool_permille;/* share of input wires that are bits */
  uint32_t terms;              /* nonzeros per A and B row */
  uint64_t seed;               /* circuit structure */
  uint64_t setup_seed;         /* toxic waste */
  uint64_t witness_seed;       /* witness values (any seed satisfies the circuit) */
} g16synth_params;

typedef struct g16synth g16synth;

/* Builds the circuit, witness, toxic waste and query scalars on the host (multithreaded). */
g16synth* g16synth_new(const g16synth_params* p);
void g16synth_free(g16synth* s);
/* out[7]: n_vars, n_public, domain_size, n_constraints, nnz(A), nnz(B), nnz(C). */
void g16synth_dims(const g16synth* s, uint32_t* out);
/* Copies one array out. what: 0..2 row_ptr of A, B, C (u32[domain+1]); 3..5 signal (u32[nnz]);
 * 6..8 values (Fr[nnz], Montgomery); 9 witness (Fr[n_vars]); 10 toxic waste
 * (Fr[5]: tau, alpha, beta, gamma, delta); 11 u_i(tau) (Fr[n_vars]); 12 v_i(tau) (Fr[n_vars]);
 * 13 IC scalars (Fr[n_public+1]); 14 L scalars (Fr[n_vars-n_public-1]); 15 H scalars
 * (Fr[domain]). Returns 0, or -1 for an unknown `what`. */
int g16synth_get(const g16synth* s, int what, void* dst);

/* Another satisfying witness of the same circuit, for witness seed `witness_seed`
 * (Fr[n_vars], Montgomery): the keys stay the same, so batches of independent proofs can share
 * one prepared circuit. Returns 0, or -1 for a null argument. */
int g16synth_witness(const g16synth* s, uint64_t witness_seed, uint8_t* out);

/* Seeded scalar vectors (Fr Montgomery) for the standalone MSM workloads: each scalar is, with
 * probability zero_one_permille/1000, 0 or 1 (equally likely), else uniform in [0, r). */
void g16synth_scalars(uint64_t seed, size_t n, uint32_t zero_one_permille, uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif /* G16SYNTH_H */
