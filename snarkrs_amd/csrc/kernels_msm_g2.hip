// G2 bucket accumulation (msm_accum.cuh) on Fq2 with its products expanded in place (Fm2i),
// in a translation unit of its own so it compiles in parallel with the rest; the merge and
// reduction launchers are in kernels_msm_g2_tail.hip and kernels_msm_g2_reduce.hip.
#include "msm_accum.cuh"

namespace g16 {

void msm_window_tables_g2(void* table, size_t n, int c, int windows, hipStream_t s) {
  if (!n || windows < 2) return;
  hipLaunchKernelGGL(k_window_tables<Fm2i>, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s,
                     static_cast<uint4*>(table), n, c, windows);
  G16_LAUNCH_CHECK("msm window tables (G2)");
}

void msm_accumulate_g2(const MsmPlan& p, const void* bases, void* scratch, void* window_out,
                       hipStream_t s, hipEvent_t* ev) {
  accumulate_impl<Fm2i>(p, bases, scratch, window_out, s, ev);
}

}  // namespace g16
