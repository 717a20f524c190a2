// G2 instantiation of the MSM accumulation kernels (msm_accum.cuh), in a translation unit of
// its own so it compiles in parallel with the rest: the bucket-accumulation loop runs on Fq2
// with its products expanded in place (Fq2i), the merge and reduction kernels on the compact
// out-of-line form.
#include "msm_accum.cuh"

namespace g16 {

void msm_accumulate_g2(const MsmPlan& p, const void* bases, void* scratch, void* window_out,
                       hipStream_t s, hipEvent_t* ev) {
  accumulate_impl<Fq2, Fq2i>(p, bases, scratch, window_out, s, ev);
}

}  // namespace g16
