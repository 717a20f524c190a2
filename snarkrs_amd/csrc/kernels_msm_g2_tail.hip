// G2 bucket merges (spanning and heavy buckets) on Fm2i: see msm_accum.cuh.
#include "msm_accum.cuh"

namespace g16 {

template <>
void Tail<Fm2i>::merge(const MsmPlan& p, const AccumPtrs& a, hipStream_t s) {
  tail_merge_impl<Fm2i>(p, a, s);
}

}  // namespace g16
