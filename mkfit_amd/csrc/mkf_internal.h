// mkf_internal.h -- context, error reporting, pointer checks and buffer
// addressing shared by the fit (mkf_fit.hip) and finding (mkf_find.hip)
// translation units of libmkfit_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "../../include/mkfit_amd.h"
#include "mkf_track_math.h"

// ============================================================================
// Context and error reporting
// ============================================================================

struct GraphKey {
  const void* ptrs[12];
  int n, slots, pflags, fflags, block, bpc, variant, pca_pflags, tile;
  bool operator<(const GraphKey& o) const {
    int c = std::memcmp(ptrs, o.ptrs, sizeof(ptrs));
    if (c) return c < 0;
    return std::tie(n, slots, pflags, fflags, block, bpc, variant, pca_pflags, tile) <
           std::tie(o.n, o.slots, o.pflags, o.fflags, o.block, o.bpc, o.variant, o.pca_pflags, o.tile);
  }
};

struct mkf_context {
  int device = 0;
  int num_cus = 256;
  float bfield = 3.8112f;
  float* d_rl = nullptr;
  float* d_xi = nullptr;
  mkf_launch_shape shape{256, 0, 0, 0};
  mkf_find_shape find_shape[2] = {{0, 2}, {0, 2}};   // clone engine, standard finding
  hipStream_t capture_stream = nullptr;
  // Instantiated batched-fit graphs by batch (pointers, sizes, flags, shape),
  // at most kMaxGraphs; the least recently launched one is destroyed to make
  // room for a new one.
  struct GraphEntry {
    hipGraphExec_t exec;
    unsigned long long last_use;
  };
  static constexpr size_t kMaxGraphs = 16;
  std::map<GraphKey, GraphEntry> graphs;
  unsigned long long graph_tick = 0;
  std::mutex mu;
  void* adj_scratch = nullptr;   // adjoint checkpoints and per-track gradient rows
  size_t adj_bytes = 0;
  // Recorded on the stream of the last mkf_fit_adjoint call after its last
  // kernel that touches adj_scratch; the next call's stream waits on it, so
  // calls on different streams take the scratch one after the other.
  hipEvent_t adj_done = nullptr;
  // Track order of the batched fit (sort tiles), n_tracks ints; the graphs
  // capture its address, so growing it drops every cached graph.  fit_done
  // orders its users across streams as adj_done does for adj_scratch.
  int* fit_order = nullptr;
  size_t fit_order_cap = 0;
  hipEvent_t fit_done = nullptr;
  // This device's fixed scratch array of mkf_fit_adjoint_noalloc (device address from hipGetSymbolAddress at
  // context creation); shared by every context on the device.
  void* adj_static = nullptr;
  size_t adj_static_bytes = 0;
};

namespace mkf_detail {

inline thread_local std::string g_last_error;

inline int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return fail(MKF_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

// A device pointer must be non-null, device (or managed) memory, and the
// allocation it lies in must extend `bytes` past it.
inline int check_dev(const void* p, size_t bytes, const char* name) {
  if (!p) return fail(MKF_ERR_INVALID_ARG, "%s is NULL", name);
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return fail(MKF_ERR_INVALID_ARG, "%s is not a HIP allocation", name);
  }
  if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged)
    return fail(MKF_ERR_INVALID_ARG, "%s is not device memory", name);
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
    (void)hipGetLastError();
    return fail(MKF_ERR_INVALID_ARG, "%s: cannot query its allocation", name);
  }
  const char* b = (const char*)base;
  const char* q = (const char*)p;
  if (q < b || (size_t)(q - b) + bytes > size)
    return fail(MKF_ERR_BOUNDS, "%s: %zu bytes needed past the pointer, allocation has %zu", name, bytes,
                size - (size_t)(q - b));
  return MKF_OK;
}

#define CHECK(expr)            \
  do {                         \
    int rc_ = (expr);          \
    if (rc_) return rc_;       \
  } while (0)

inline int set_device(mkf_context* c) {
  if (!c) return fail(MKF_ERR_INVALID_ARG, "context is NULL");
  HIP_TRY(hipSetDevice(c->device));
  return MKF_OK;
}

inline mkf::Material mat_of(const mkf_context* c) { return mkf::Material{c->d_rl, c->d_xi}; }

// Orders a call's use of a shared scratch buffer after the previous call's:
// `st` waits for `done` (when a previous call recorded it).  The caller
// records `done` on `st` after its last kernel that uses the buffer
// (scratch_release).  Both run under the owner's mutex.
inline int scratch_acquire(hipEvent_t& done, hipStream_t st) {
  if (!done) HIP_TRY(hipEventCreateWithFlags(&done, hipEventDisableTiming));
  else HIP_TRY(hipStreamWaitEvent(st, done, 0));
  return MKF_OK;
}
inline int scratch_release(hipEvent_t done, hipStream_t st) {
  HIP_TRY(hipEventRecord(done, st));
  return MKF_OK;
}

// Sets mkf_context::adj_static for the context's device (mkf_adjoint.hip); the device must be current.
int adjoint_noalloc_setup(mkf_context* c);

}  // namespace mkf_detail

namespace mkf {

// Buffer addressing for the arrays read every step.  Each gets one buffer
// resource (base, byte size) held in SGPRs; element e of track t is read with
// the per-lane byte offset t*4 in a VGPR and the element stride e*n*4 in an
// SGPR soffset, so no per-lane address arithmetic is spent on the 9 hit
// elements.  The byte size doubles as a hardware bounds check: an access past
// it returns 0 instead of touching other memory.
constexpr int kBufFlags = 0x00020000;  // gfx9-family raw buffer: 32-bit dword data

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_buf(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, kBufFlags);
}
__device__ __forceinline__ float bld_f(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ int bld_i(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  return (int)__builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0);
}
__device__ __forceinline__ unsigned bld_u8(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  return (unsigned)__builtin_amdgcn_raw_buffer_load_b8(r, (int)vo, (int)so, 0);
}

// Workgroups are dealt to the 8 XCDs in turn.  This renumbers block b so that
// XCD x walks the x-th contiguous eighth of [0, gridDim.x): XCD x takes
// gridDim.x / 8 + (x < gridDim.x % 8) blocks, so the map is a bijection.  Work
// ordered by layer, event or sort tile then shares one L2 per XCD instead of
// all eight.
__device__ __forceinline__ int xcd_major_block() {
  const int nb = gridDim.x, r8 = nb % 8, q8 = nb / 8, x = blockIdx.x % 8;
  return x * q8 + min(x, r8) + blockIdx.x / 8;
}

}  // namespace mkf
