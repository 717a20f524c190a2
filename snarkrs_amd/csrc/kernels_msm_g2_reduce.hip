// G2 window reduction on Fm2i: see msm_accum.cuh.
#include "msm_accum.cuh"

namespace g16 {

template <>
void Tail<Fm2i>::reduce(const MsmPlan& p, const AccumPtrs& a, uint32_t B, void* window_out,
                        hipStream_t s) {
  tail_reduce_impl<Fm2i>(p, a, B, window_out, s);
}

}  // namespace g16
