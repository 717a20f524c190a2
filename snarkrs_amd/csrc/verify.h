// Host-side Groth16 verification over BN254 and the snarkjs JSON encodings, shared by
// verify.cpp (the C ABI of the checks) and prove.hip (the check of a proof's result).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/g16hip.h"

namespace g16 {

// A verifying key in the boundary's wire encoding (points affine, Montgomery, 0 = infinity).
struct VkBytes {
  uint8_t alpha_g1[64];
  uint8_t beta_g2[128];
  uint8_t gamma_g2[128];
  uint8_t delta_g2[128];
  std::vector<uint8_t> ic;  // G1[n_public + 1]
};

enum class VerifyMode { Checked, AllowingUnblinded, Unchecked };

// verify / verify_allowing_unblinded / verify_unchecked (crates/groth16/src/verify.rs:49-134):
// G16HIP_OK, or the VerifyError status with its message in *why. `pub` is Fr[n] Montgomery.
g16hip_status groth16_check(const VkBytes& vk, const uint8_t* pub, size_t n, const g16hip_proof& p,
                            VerifyMode mode, std::string* why);

}  // namespace g16
