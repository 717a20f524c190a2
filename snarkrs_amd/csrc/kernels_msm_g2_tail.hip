// G2 bucket merges (spanning and heavy buckets) on Fq2i: see msm_accum.cuh.
#include "msm_accum.cuh"

namespace g16 {

template <>
void Tail<Fq2i>::merge(const MsmPlan& p, const AccumPtrs& a, hipStream_t s) {
  tail_merge_impl<Fq2i>(p, a, s);
}

}  // namespace g16
