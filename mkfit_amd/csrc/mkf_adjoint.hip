// mkf_adjoint.hip -- reverse-mode gradient of the fit's total chi2 with respect
// to per-layer material (x/X0) and per-layer alignment (r-phi and z offsets).
//
// One lane per track.  Forward: the fit of the alignment model (hits moved by
// their layer's offsets, material from the layer's x/X0; mkf_adjoint_math.h)
// in FP32 (or, with fp64 set, the same model in FP64), saving the 27-entry
// state every `checkpoint_every` layers.
// Reverse: the adjoint lambda of the state (27 floats, zero after the last
// hit) is carried back layer by layer.  At layer l the state entering the
// layer is recomputed from the segment's checkpoint, and the step's
// vector-Jacobian product
//     v = lambda_out^T d(state_out)/d(in) + d(chi2_l)/d(in),
//     in = (state_in, theta_l)  (30 directions)
// is evaluated by the hand-written reverse sweep through the step
// (mkf_adjoint_rev.h) -- or, as a cross-check, from forward tangents, Dual<D>
// carrying D input directions per pass (10 passes).  v[0:27] is the new
// lambda, v[27:30] the track's gradient for layer l.
// Per-track gradient rows go to HBM and a second kernel sums them over tracks
// in double, in a fixed order (deterministic result).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/mkfit_amd.h"
#include "mkf_adjoint_math.h"
#include "mkf_adjoint_rev.h"
#include "mkf_internal.h"

using namespace mkf;
using namespace mkf_adj;
using namespace mkf_detail;

namespace {

constexpr int kDirs = 30;       // 27 state entries + 3 layer parameters
constexpr int kBlock = 128;
// Tangents per Dual pass: what fits in registers (a double Dual is twice as wide).
template <class T> constexpr int kTanD = 3;
template <> constexpr int kTanD<double> = 1;

struct AdjArgs {
  int n, n_layers, ck, pflags;
  float xi_per_x0, bfield;
  const float* __restrict__ seed_par;
  const float* __restrict__ seed_err;
  const int* __restrict__ seed_chg;
  const float* __restrict__ hit_pos;
  const float* __restrict__ hit_err;
  const int* __restrict__ n_hits;
  const float* __restrict__ theta;
  float* __restrict__ chi2;
  void* __restrict__ ckpt;      // [n_ck][27][n] of the model's scalar type
  void* __restrict__ gtrack;    // [n_layers * 3][n] of the model's scalar type
};

// Hit l of track t and the reciprocal of its transverse radius in the model's
// scalar type (the FP32 model rounds the double root to float first).
__device__ __forceinline__ void load_hit(const AdjArgs& a, int l, int t, float (&hp)[3], float (&he)[6],
                                         float& rh_inv) {
  const size_t n = (size_t)a.n;
#pragma unroll
  for (int e = 0; e < 3; ++e) hp[e] = a.hit_pos[((size_t)l * 3 + e) * n + t];
#pragma unroll
  for (int e = 0; e < 6; ++e) he[e] = a.hit_err[((size_t)l * 6 + e) * n + t];
  const double x = hp[0], y = hp[1];
  rh_inv = 1.f / (float)__builtin_sqrt(x * x + y * y);
}
__device__ __forceinline__ void load_hit(const AdjArgs& a, int l, int t, float (&hp)[3], float (&he)[6],
                                         double& rh_inv) {
  float unused;
  load_hit(a, l, t, hp, he, unused);
  const double x = hp[0], y = hp[1];
  rh_inv = 1.0 / __builtin_sqrt(x * x + y * y);
}

// TAN: the step's product from forward tangents instead of the reverse sweep.
// The FP32 reverse sweep fits two waves per SIMD (256 VGPRs).
template <class T, bool TAN>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu((sizeof(T) == 4 && !TAN) ? 2 : 1, 8)))
k_adjoint(AdjArgs a) {
  __shared__ T vbuf[TAN ? kBlock : 1][kDirs + 1];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  const size_t n = (size_t)a.n;
  T* __restrict__ ckpt = static_cast<T*>(a.ckpt);
  T* __restrict__ gtrack = static_cast<T*>(a.gtrack);
  const int nh = min(a.n_hits[t], a.n_layers);
  const int chg = a.seed_chg[t];
  T par[6], err[21];
  // ---- forward, with checkpoints
#pragma unroll
  for (int e = 0; e < 6; ++e) par[e] = a.seed_par[e * n + t];
#pragma unroll
  for (int e = 0; e < 21; ++e) err[e] = a.seed_err[e * n + t];
  T chi2 = T(0);
  for (int l = 0; l < nh; ++l) {
    if (l % a.ck == 0) {
      T* ck = ckpt + (size_t)(l / a.ck) * 27 * n + t;
#pragma unroll
      for (int e = 0; e < 6; ++e) ck[e * n] = par[e];
#pragma unroll
      for (int e = 0; e < 21; ++e) ck[(6 + e) * n] = err[e];
    }
    float hp[3], he[6];
    T rhi;
    load_hit(a, l, t, hp, he, rhi);
    chi2 += align_step<T>(par, err, chg, hp, he, T(a.theta[3 * l]), T(a.theta[3 * l + 1]), T(a.theta[3 * l + 2]),
                          a.xi_per_x0, a.bfield, a.pflags, rhi);
  }
  a.chi2[t] = (float)chi2;
  for (int l = nh; l < a.n_layers; ++l)
    for (int q = 0; q < 3; ++q) gtrack[((size_t)l * 3 + q) * n + t] = T(0);
  // ---- reverse sweep
  T lam[27];
#pragma unroll
  for (int e = 0; e < 27; ++e) lam[e] = T(0);
  for (int l = nh - 1; l >= 0; --l) {
    const int c0 = (l / a.ck) * a.ck;
    const T* ck = ckpt + (size_t)(l / a.ck) * 27 * n + t;
#pragma unroll
    for (int e = 0; e < 6; ++e) par[e] = ck[e * n];
#pragma unroll
    for (int e = 0; e < 21; ++e) err[e] = ck[(6 + e) * n];
    for (int j = c0; j < l; ++j) {                      // recompute the state entering layer l
      float hp[3], he[6];
      T rhi;
      load_hit(a, j, t, hp, he, rhi);
      (void)align_step<T>(par, err, chg, hp, he, T(a.theta[3 * j]), T(a.theta[3 * j + 1]), T(a.theta[3 * j + 2]),
                          a.xi_per_x0, a.bfield, a.pflags, rhi);
    }
    float hp[3], he[6];
    T rhi;
    load_hit(a, l, t, hp, he, rhi);
    const T th0 = a.theta[3 * l], th1 = a.theta[3 * l + 1], th2 = a.theta[3 * l + 2];
    if constexpr (!TAN) {
      T gp[6], ge[21], gth[3];
      (void)align_step_vjp<T>(par, err, chg, hp, he, th0, th1, th2, a.xi_per_x0, a.bfield, a.pflags, rhi,
                              *reinterpret_cast<const T(*)[6]>(lam), *reinterpret_cast<const T(*)[21]>(lam + 6), gp,
                              ge, gth);
#pragma unroll
      for (int e = 0; e < 6; ++e) lam[e] = gp[e];
#pragma unroll
      for (int e = 0; e < 21; ++e) lam[6 + e] = ge[e];
      for (int q = 0; q < 3; ++q) gtrack[((size_t)l * 3 + q) * n + t] = gth[q];
    } else {
      T* vb = vbuf[threadIdx.x];
      constexpr int TD = kTanD<T>;
      for (int d0 = 0; d0 < kDirs; d0 += TD) {
        using DV = Dual<T, TD>;
        DV dp[6], de[21], m, dr, dz;
#pragma unroll
        for (int e = 0; e < 6; ++e) {
          dp[e].v = par[e];
#pragma unroll
          for (int k = 0; k < TD; ++k) dp[e].d[k] = (e == d0 + k) ? T(1) : T(0);
        }
#pragma unroll
        for (int e = 0; e < 21; ++e) {
          de[e].v = err[e];
#pragma unroll
          for (int k = 0; k < TD; ++k) de[e].d[k] = (6 + e == d0 + k) ? T(1) : T(0);
        }
        m.v = th0; dr.v = th1; dz.v = th2;
#pragma unroll
        for (int k = 0; k < TD; ++k) {
          m.d[k] = (27 == d0 + k) ? T(1) : T(0);
          dr.d[k] = (28 == d0 + k) ? T(1) : T(0);
          dz.d[k] = (29 == d0 + k) ? T(1) : T(0);
        }
        const DV c2 = align_step<DV>(dp, de, chg, hp, he, m, dr, dz, a.xi_per_x0, a.bfield, a.pflags, rhi);
#pragma unroll
        for (int k = 0; k < TD; ++k) {
          T v = c2.d[k];
#pragma unroll
          for (int e = 0; e < 6; ++e) v += lam[e] * dp[e].d[k];
#pragma unroll
          for (int e = 0; e < 21; ++e) v += lam[6 + e] * de[e].d[k];
          if (d0 + k < kDirs) vb[d0 + k] = v;
        }
      }
#pragma unroll
      for (int e = 0; e < 27; ++e) lam[e] = vb[e];
      for (int q = 0; q < 3; ++q) gtrack[((size_t)l * 3 + q) * n + t] = vb[27 + q];
    }
  }
}

// Sum over tracks of each parameter's row, in double and in a fixed order:
// block (row, part) sums its fixed slice of the row into partial[row][part],
// then one thread per row adds its kSumParts partials in order.  The result
// does not depend on timing.
constexpr int kSumParts = 64;
template <class T>
__global__ void __launch_bounds__(256) k_sum_parts(const T* __restrict__ rows, int n, double* __restrict__ partial) {
  __shared__ double part[256];
  const int chunk = (n + kSumParts - 1) / kSumParts;
  const int lo = blockIdx.y * chunk, hi = min(n, lo + chunk);
  const T* r = rows + (size_t)blockIdx.x * n;
  double s = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += 256) s += (double)r[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)blockIdx.x * kSumParts + blockIdx.y] = part[0];
}

__global__ void __launch_bounds__(256) k_sum_rows(const double* __restrict__ partial, int n_rows, float* __restrict__ out) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_rows) return;
  double s = 0.0;
  if (partial)
    for (int p = 0; p < kSumParts; ++p) s += partial[(size_t)row * kSumParts + p];
  out[row] = (float)s;
}

}  // namespace

// Device memory of mkf_fit_adjoint_noalloc: the same layout as the context's adjoint scratch (sum partials,
// checkpoints, per-track gradient rows), in a fixed array each device holds for the life of the process.  Host
// code uses the address hipGetSymbolAddress returns for it on each device (mkf_context::adj_static), never the
// array's name.  Every call on one device shares it; calls that overlap on one device race.
constexpr size_t kAdjNoallocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kAdjNoallocBytes];

namespace {

// Checks the batch; returns through `need` the scratch bytes the call takes and through `sum_bytes` the
// 256-byte aligned size of the sums' partials at its front.
int adjoint_check(const mkf_adjoint_batch* b, size_t& need, size_t& sum_bytes) {
  if (!b) return fail(MKF_ERR_INVALID_ARG, "batch is NULL");
  if (b->n_tracks < 0 || b->n_layers <= 0 || b->n_layers > 1024)
    return fail(MKF_ERR_INVALID_ARG, "need n_tracks >= 0 and 1 <= n_layers <= 1024");
  if (b->checkpoint_every < 1) return fail(MKF_ERR_INVALID_ARG, "checkpoint_every must be >= 1");
  if (b->pflags & MKF_PF_USE_PARAM_B_FIELD)
    return fail(MKF_ERR_INVALID_ARG, "the adjoint supports the uniform field only");
  const size_t N = b->n_tracks, L = b->n_layers;
  CHECK(check_dev(b->theta, L * 3 * 4, "theta"));
  CHECK(check_dev(b->grad, L * 3 * 4, "grad"));
  need = 0;
  sum_bytes = 0;
  if (N == 0) return MKF_OK;
  if (L * 6 * N * 4 >= ((size_t)1 << 32)) return fail(MKF_ERR_INVALID_ARG, "batch too large");
  CHECK(check_dev(b->seed_par, 6 * N * 4, "seed_par"));
  CHECK(check_dev(b->seed_err, 21 * N * 4, "seed_err"));
  CHECK(check_dev(b->seed_chg, N * 4, "seed_chg"));
  CHECK(check_dev(b->hit_pos, L * 3 * N * 4, "hit_pos"));
  CHECK(check_dev(b->hit_err, L * 6 * N * 4, "hit_err"));
  CHECK(check_dev(b->n_hits, N * 4, "n_hits"));
  CHECK(check_dev(b->chi2, N * 4, "chi2"));
  const size_t n_ck = (L + b->checkpoint_every - 1) / b->checkpoint_every;
  const size_t word = b->fp64 ? 8 : 4;
  // scratch: the sums' partials (256-byte aligned size), checkpoints, per-track gradient rows
  sum_bytes = (L * 3 * kSumParts * sizeof(double) + 255) / 256 * 256;
  need = sum_bytes + (n_ck * 27 + L * 3) * N * word;
  return MKF_OK;
}

// The forward fit, reverse sweep and sums of a checked batch, with `scratch` holding at least the bytes
// adjoint_check asked for.  Every byte of the scratch that a kernel reads is written earlier in the same call.
int adjoint_launch(const mkf_context* c, const mkf_adjoint_batch* b, void* scr, size_t sum_bytes, hipStream_t st) {
  const size_t N = b->n_tracks, L = b->n_layers;
  if (N == 0) {
    hipLaunchKernelGGL(k_sum_rows, dim3((unsigned)((L * 3 + 255) / 256)), dim3(256), 0, st, (const double*)nullptr,
                       (int)(L * 3), b->grad);
    HIP_TRY(hipGetLastError());
    return MKF_OK;
  }
  const size_t n_ck = (L + b->checkpoint_every - 1) / b->checkpoint_every;
  const size_t word = b->fp64 ? 8 : 4;
  AdjArgs a;
  a.n = b->n_tracks;
  a.n_layers = b->n_layers;
  a.ck = b->checkpoint_every;
  a.pflags = b->pflags;
  a.xi_per_x0 = b->xi_per_x0;
  a.bfield = c->bfield;
  a.seed_par = b->seed_par; a.seed_err = b->seed_err; a.seed_chg = b->seed_chg;
  a.hit_pos = b->hit_pos; a.hit_err = b->hit_err; a.n_hits = b->n_hits;
  a.theta = b->theta;
  a.chi2 = b->chi2;
  double* partial = static_cast<double*>(scr);
  a.ckpt = static_cast<char*>(scr) + sum_bytes;
  a.gtrack = static_cast<char*>(a.ckpt) + n_ck * 27 * N * word;
  const dim3 sum_grid((unsigned)(L * 3), (unsigned)kSumParts);
  const dim3 fin_grid((unsigned)((L * 3 + 255) / 256));
  const dim3 grid((unsigned)((N + kBlock - 1) / kBlock));
  if (b->fp64) {
    if (b->tangent)
      hipLaunchKernelGGL((k_adjoint<double, true>), grid, dim3(kBlock), 0, st, a);
    else
      hipLaunchKernelGGL((k_adjoint<double, false>), grid, dim3(kBlock), 0, st, a);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sum_parts<double>, sum_grid, dim3(256), 0, st, static_cast<const double*>(a.gtrack), a.n,
                       partial);
  } else {
    if (b->tangent)
      hipLaunchKernelGGL((k_adjoint<float, true>), grid, dim3(kBlock), 0, st, a);
    else
      hipLaunchKernelGGL((k_adjoint<float, false>), grid, dim3(kBlock), 0, st, a);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sum_parts<float>, sum_grid, dim3(256), 0, st, static_cast<const float*>(a.gtrack), a.n,
                       partial);
  }
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_sum_rows, fin_grid, dim3(256), 0, st, (const double*)partial, (int)(L * 3), b->grad);
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

}  // namespace

namespace mkf_detail {

// Called by mkf_context_create with the context's device current: the device address of this device's copy
// of `scratch`.
int adjoint_noalloc_setup(mkf_context* c) {
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return fail(MKF_ERR_HIP, "no-allocation scratch: %s", hipGetErrorString(e));
  c->adj_static = p;
  c->adj_static_bytes = kAdjNoallocBytes;
  return MKF_OK;
}

}  // namespace mkf_detail

extern "C" int mkf_fit_adjoint(mkf_context* c, const mkf_adjoint_batch* b, void* stream) {
  CHECK(set_device(c));
  size_t need = 0, sum_bytes = 0;
  CHECK(adjoint_check(b, need, sum_bytes));
  hipStream_t st = (hipStream_t)stream;
  if (b->n_tracks == 0) return adjoint_launch(c, b, nullptr, 0, st);
  std::lock_guard<std::mutex> lk(c->mu);
  CHECK(scratch_acquire(c->adj_done, st));
  if (need > c->adj_bytes) {
    HIP_TRY(hipDeviceSynchronize());
    if (c->adj_scratch) HIP_TRY(hipFree(c->adj_scratch));
    c->adj_scratch = nullptr;
    c->adj_bytes = 0;
    HIP_TRY(hipMalloc(&c->adj_scratch, need));
    c->adj_bytes = need;
  }
  CHECK(adjoint_launch(c, b, c->adj_scratch, sum_bytes, st));
  return scratch_release(c->adj_done, st);
}

extern "C" int mkf_fit_adjoint_noalloc(mkf_context* c, const mkf_adjoint_batch* b, void* stream) {
  CHECK(set_device(c));
  size_t need = 0, sum_bytes = 0;
  CHECK(adjoint_check(b, need, sum_bytes));
  if (!c->adj_static) return fail(MKF_ERR_HIP, "no-allocation scratch not set up for this context");
  if (need > c->adj_static_bytes)
    return fail(MKF_ERR_INVALID_ARG,
                "batch needs %zu bytes of adjoint scratch; mkf_fit_adjoint_noalloc has %zu (fewer tracks, or a "
                "larger checkpoint_every)", need, c->adj_static_bytes);
  return adjoint_launch(c, b, c->adj_static, sum_bytes, (hipStream_t)stream);
}
