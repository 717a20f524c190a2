/*
 * g16zkey — snarkjs .zkey / .wtns files to and from the g16hip_key wire format.
 *
 * Decoding replaces ProvingKey::load and Witness::load (crates/formats/src/lib.rs:430,
 * :505-760, :905-1130; crates/formats/src/wtns.rs:17-63): the iden3 binfile container
 * ("zkey" v1 / "wtns" v2; 4 magic bytes, u32 version, u32 nSections, then u32 id + u64 length +
 * payload per section, crates/formats/src/binfile.rs:6-8), Montgomery point coordinates with
 * (0,0) as infinity, section 4's double-Montgomery coefficients sorted into CSR by a counting
 * sort with values interned, and the plain-integer witness converted to Montgomery form.
 * The checks follow the reference's: protocol, moduli, power-of-two domain <= 2^27, section
 * lengths, canonical coordinates, record indices in range, no header point at infinity, and
 * witness[0] == 1.
 */
#ifndef G16ZKEY_H
#define G16ZKEY_H

#include <stddef.h>
#include <stdint.h>

#include "g16hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct g16zkey g16zkey;

/* Parses `path`; on success *out owns every array g16zkey_key points into. Errors are
 * G16HIP_BACKEND with the reason in g16hip_last_error(). */
g16hip_status g16zkey_load(const char* path, g16zkey** out);
/* The parsed key as the C ABI takes it (valid until g16zkey_free). */
const g16hip_key* g16zkey_key(const g16zkey* z);
void g16zkey_free(g16zkey* z);

/* Writes `key` as a snarkjs .zkey (sections in snarkjs' order 1,2,4,3,9,8,5,6,7,10; section 10
 * an empty contribution list). */
g16hip_status g16zkey_write(const char* path, const g16hip_key* key);

/* Reads a .wtns; *out = Fr[*n] in Montgomery form (free with g16wtns_free). */
g16hip_status g16wtns_load(const char* path, uint8_t** out, size_t* n);
void g16wtns_free(uint8_t* w);
/* Writes Fr[n] Montgomery values as a .wtns (plain little-endian integers). */
g16hip_status g16wtns_write(const char* path, const uint8_t* witness, size_t n);

#ifdef __cplusplus
}
#endif
#endif /* G16ZKEY_H */
