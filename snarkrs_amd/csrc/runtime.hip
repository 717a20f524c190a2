// The C ABI of include/g16hip.h: device handles, prepared circuits, the prove pipeline
// (stages 0-9 on the GPU, stage 11 on the host) and the standalone entry points.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <future>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/g16hip.h"
#include "bn254.cuh"
#include "internal.h"

using namespace g16;

namespace g16 {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw HipError{std::string(what) + ": " + hipGetErrorString(e)};
  }
}

}  // namespace g16

// ---------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------
static thread_local std::string t_last_error;

static g16hip_status fail(g16hip_status st, const std::string& msg) {
  t_last_error = msg;
  return st;
}

template <class F>
static g16hip_status guarded(F&& f) {
  try {
    t_last_error.clear();
    return f();
  } catch (const HipError& e) {
    bool oom = e.msg.find("out of memory") != std::string::npos;
    return fail(oom ? G16HIP_DEVICE_MEMORY : G16HIP_DEVICE, e.msg);
  } catch (const ArgError& e) {
    return fail(G16HIP_BACKEND, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(G16HIP_BACKEND, "host allocation failed");
  }
}

// ---------------------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    if (n <= bytes && p) return;
    release();
    if (n == 0) return;
    G16_CHECK(hipMalloc(&p, n));
    bytes = n;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

static void upload(DevBuf& b, const void* src, size_t n, hipStream_t s) {
  b.alloc(n);
  if (n) G16_CHECK(hipMemcpyAsync(b.p, src, n, hipMemcpyHostToDevice, s));
}

struct g16hip_device {
  int ordinal = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // serialises the standalone entry points' scratch
  DevBuf scratch_a, scratch_b, scratch_c, scratch_d;
};

static void set_device(const g16hip_device* d) { G16_CHECK(hipSetDevice(d->ordinal)); }

// ---------------------------------------------------------------------------------------
// C ABI: device, field probe, fixed base
// ---------------------------------------------------------------------------------------
extern "C" {

const char* g16hip_last_error(void) { return t_last_error.c_str(); }

g16hip_status g16hip_open(int ordinal, g16hip_device** out) {
  return guarded([&] {
    if (!out) throw ArgError{"null output handle"};
    int n = 0;
    G16_CHECK(hipGetDeviceCount(&n));
    if (ordinal < 0 || ordinal >= n)
      throw ArgError{"device ordinal " + std::to_string(ordinal) + " out of range (" +
                     std::to_string(n) + " devices)"};
    auto* d = new g16hip_device();
    d->ordinal = ordinal;
    G16_CHECK(hipSetDevice(ordinal));
    G16_CHECK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    *out = d;
    return G16HIP_OK;
  });
}

void g16hip_close(g16hip_device* d) {
  if (!d) return;
  (void)hipSetDevice(d->ordinal);
  (void)hipStreamSynchronize(d->stream);
  d->scratch_a.release();
  d->scratch_b.release();
  d->scratch_c.release();
  d->scratch_d.release();
  (void)hipStreamDestroy(d->stream);
  delete d;
}

g16hip_status g16hip_mem_info(g16hip_device* d, size_t* free_b, size_t* total_b) {
  return guarded([&] {
    set_device(d);
    G16_CHECK(hipMemGetInfo(free_b, total_b));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_field_probe(g16hip_device* d, int field, int op, const uint8_t* a,
                                 const uint8_t* b, uint8_t* out, size_t n) {
  return guarded([&] {
    if (field < 0 || field > 2 || op < 0 || op > 7) throw ArgError{"bad field/op"};
    if (n == 0) return G16HIP_OK;
    std::lock_guard<std::mutex> lk(d->mu);
    set_device(d);
    size_t w = field == 2 ? 64 : 32;
    upload(d->scratch_a, a, n * w, d->stream);
    if (b) upload(d->scratch_b, b, n * w, d->stream);
    d->scratch_c.alloc(n * w);
    launch_field_probe(field, op, d->scratch_a.p, b ? d->scratch_b.p : nullptr, d->scratch_c.p,
                       n, d->stream);
    G16_CHECK(hipMemcpyAsync(out, d->scratch_c.p, n * w, hipMemcpyDeviceToHost, d->stream));
    G16_CHECK(hipStreamSynchronize(d->stream));
    return G16HIP_OK;
  });
}

g16hip_status g16hip_fixed_base(g16hip_device* d, int group, const uint8_t* k, size_t n,
                                uint8_t* out) {
  return guarded([&] {
    if (group != 1 && group != 2) throw ArgError{"group must be 1 or 2"};
    if (n == 0) return G16HIP_OK;
    std::lock_guard<std::mutex> lk(d->mu);
    set_device(d);
    size_t w = group == 1 ? 64 : 128;
    upload(d->scratch_a, k, n * 32, d->stream);
    d->scratch_c.alloc(n * w);
    launch_fixed_base(group, d->scratch_a.p, n, d->scratch_c.p, d->stream);
    G16_CHECK(hipMemcpyAsync(out, d->scratch_c.p, n * w, hipMemcpyDeviceToHost, d->stream));
    G16_CHECK(hipStreamSynchronize(d->stream));
    return G16HIP_OK;
  });
}

}  // extern "C"
