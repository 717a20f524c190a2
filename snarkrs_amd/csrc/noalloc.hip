// g16hip_msm_device_noalloc: the device MSM (g16hip_msm_g1_device / _g2_device) with every
// piece of its scratch in a static device array instead of memory the library allocates, for
// callers that cannot allocate device memory (code under graph capture, tight serving loops).
// It has no counterpart in the reference.
#include <map>

#include "runtime_impl.h"

using namespace g16;

// One copy per GPU, held until the process ends. Host code never uses this name as an
// address: each device's copy is resolved once with hipGetSymbolAddress (device_scratch).
static __device__ __align__(256) uint8_t scratch[G16HIP_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;  // device ordinal -> that device's copy of `scratch`

// The device address of `scratch` on d's GPU, looked up on the first call there with that
// device current.
void* device_scratch(const g16hip_device* d) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(d->ordinal);
  if (it != g_scratch_ptr.end()) return it->second;
  void* p = nullptr;
  G16_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw StatusError{G16HIP_DEVICE, "hipGetSymbolAddress gave no address for the static scratch"};
  g_scratch_ptr[d->ordinal] = p;
  return p;
}

}  // namespace

extern "C" g16hip_status g16hip_msm_device_noalloc(g16hip_device* d, int group, const void* bases,
                                                   const void* scalars, size_t n, int c,
                                                   uint8_t* out) {
  return guarded([&] {
    if (!d || !out) throw ArgError{"null argument"};
    if (group != 1 && group != 2) throw ArgError{"group must be 1 or 2"};
    const size_t pw = group == 1 ? 64 : 128;
    if (n == 0) {
      std::memset(out, 0, pw);
      return G16HIP_OK;
    }
    if (!bases || !scalars) throw ArgError{"null argument"};
    if (n >= (1u << 31)) throw ArgError{"MSM larger than 2^31 points"};
    std::lock_guard<std::mutex> lk(d->mu);
    set_device(d);
    char* base = static_cast<char*>(device_scratch(d));
    const int w = msm_width((uint32_t)n, c);
    const MsmScratchSizes z = msm_scratch_sizes(group, (uint32_t)n, w, d->constant_work);
    if (z.total() > G16HIP_NOALLOC_SCRATCH_BYTES)
      throw StatusError{G16HIP_DEVICE_MEMORY,
                        "an MSM of " + std::to_string(n) + " points at window width " +
                            std::to_string(w) + " needs " + std::to_string(z.total()) +
                            " bytes of scratch and the static scratch holds " +
                            std::to_string((size_t)G16HIP_NOALLOC_SCRATCH_BYTES)};
    msm_device_on_scratch(d, group, bases, scalars, (uint32_t)n, w,
                          MsmScratch{base, base + z.sort, base + z.sort + z.accum}, out);
    return G16HIP_OK;
  });
}
