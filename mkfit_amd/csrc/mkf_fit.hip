// mkf_fit.hip -- propagation / Kalman / fit kernels and the C ABI (include/mkfit_amd.h).
//
// Layout in HBM: every per-track quantity is element-major, [element][track],
// so lane t of a wavefront reads a[e*n + t] and each element load is one
// fully coalesced 256-byte wave transaction.  Per-hit arrays are
// [slot][element][track].
//
// Two fit forms share the same per-lane step (mkf_track_math.h):
//   plain   -- k_fit launched once per hit slot; the 27-float state makes a
//              round trip through HBM per slot (~270 B per track-slot).
//   batched -- k_fit launched once over all slots on a persistent grid: the
//              state is loaded once, stays in VGPRs across the track's slots,
//              and only the hits stream (~37 B per track-slot).  It is
//              launched from a cached hipGraph; launch shape and register
//              budget come from mkf_autotune_batched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mkfit_amd.h"
#include "mkf_internal.h"
#include "mkf_track_math.h"

using namespace mkf;
using namespace mkf_detail;

// ============================================================================
// Kernels: single operations (the MPlex free functions)
// ============================================================================

__global__ void __launch_bounds__(256)
k_propagate_r(int n, const float* __restrict__ in_err, const float* __restrict__ in_par,
              const int* __restrict__ in_chg, const float* __restrict__ ms_rad, float* __restrict__ out_err,
              float* __restrict__ out_par, int pflags, const int* __restrict__ no_mat, int* __restrict__ out_fail,
              float bfield, Material mat) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float par[6], err[21];
#pragma unroll
  for (int e = 0; e < 6; ++e) par[e] = in_par[e * n + t];
#pragma unroll
  for (int e = 0; e < 21; ++e) err[e] = in_err[e * n + t];
  const bool nm = no_mat ? no_mat[t] != 0 : false;
  const bool failed = propagate_to_r(par, err, in_chg[t], ms_rad[t], pflags, nm, bfield, mat);
#pragma unroll
  for (int e = 0; e < 6; ++e) out_par[e * n + t] = par[e];
#pragma unroll
  for (int e = 0; e < 21; ++e) out_err[e * n + t] = err[e];
  if (out_fail) out_fail[t] = failed ? 1 : 0;
}

__global__ void __launch_bounds__(256)
k_propagate_z(int n, const float* __restrict__ in_err, const float* __restrict__ in_par,
              const int* __restrict__ in_chg, const float* __restrict__ ms_z, float* __restrict__ out_err,
              float* __restrict__ out_par, int pflags, const int* __restrict__ no_mat, float bfield, Material mat) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float par[6], err[21];
#pragma unroll
  for (int e = 0; e < 6; ++e) par[e] = in_par[e * n + t];
#pragma unroll
  for (int e = 0; e < 21; ++e) err[e] = in_err[e * n + t];
  const bool nm = no_mat ? no_mat[t] != 0 : false;
  propagate_to_z(par, err, in_chg[t], ms_z[t], pflags, nm, bfield, mat);
#pragma unroll
  for (int e = 0; e < 6; ++e) out_par[e * n + t] = par[e];
#pragma unroll
  for (int e = 0; e < 21; ++e) out_err[e * n + t] = err[e];
}

template <bool ENDCAP>
__global__ void __launch_bounds__(256)
k_kalman(int op, int n, const float* __restrict__ ps_err, const float* __restrict__ ps_par,
         const float* __restrict__ ms_err, const float* __restrict__ ms_par, float* __restrict__ out_err,
         float* __restrict__ out_par, float* __restrict__ out_chi2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float par[6], err[21], mp[3], me[6];
#pragma unroll
  for (int e = 0; e < 6; ++e) par[e] = ps_par[e * n + t];
#pragma unroll
  for (int e = 0; e < 21; ++e) err[e] = ps_err[e * n + t];
#pragma unroll
  for (int e = 0; e < 3; ++e) mp[e] = ms_par[e * n + t];
#pragma unroll
  for (int e = 0; e < 6; ++e) me[e] = ms_err[e * n + t];
  float chi2;
  if (op & MKF_KFO_UPDATE_PARAMS)
    chi2 = ENDCAP ? kalman_endcap<true, true>(par, err, mp, me) : kalman_barrel<true, true>(par, err, mp, me);
  else
    chi2 = ENDCAP ? kalman_endcap<true, false>(par, err, mp, me) : kalman_barrel<true, false>(par, err, mp, me);
  if ((op & MKF_KFO_CALCULATE_CHI2) && out_chi2) out_chi2[t] = chi2;
  if (op & MKF_KFO_UPDATE_PARAMS) {
#pragma unroll
    for (int e = 0; e < 6; ++e) out_par[e * n + t] = par[e];
#pragma unroll
    for (int e = 0; e < 21; ++e) out_err[e * n + t] = err[e];
  }
}

// ============================================================================
// Fit kernels
// ============================================================================

struct FitArgs {
  int n, n_slots;
  float* __restrict__ par;
  float* __restrict__ err;
  int* __restrict__ chg;
  float* __restrict__ chi2;
  const float* __restrict__ hit_pos;
  const float* __restrict__ hit_err;
  const int* __restrict__ n_hits;
  const unsigned char* __restrict__ hit_flags;
  const float* __restrict__ seed_par;
  const float* __restrict__ seed_err;
  const int* __restrict__ seed_chg;
  int pflags, fit_flags, pca_pflags;
  float bfield;
  Material mat;
  // Batched form with a sort tile: position i of the launch fits track
  // order[i] (k_fit_order); NULL: position i fits track i.
  const int* __restrict__ order;
};

// One hit slot of one track: the body of the reference's per-layer loop
// (MkFinder::BkFitFitTracks, MkFinder.cc:1612-1740; FitTracksSteered,
// MkFitter.cc:477-500).
// Buffer resources for the arrays read on every step (uniform across the grid);
// the once-per-track state goes through ordinary global loads so that only
// four descriptors (16 SGPRs) stay live in the slot loop.
struct FitBufs {
  __amdgpu_buffer_rsrc_t hit_pos, hit_err, n_hits, hit_flags;
  uint32_t nb;  // n * 4: byte stride between elements
};

__device__ __forceinline__ FitBufs make_bufs(const FitArgs& a) {
  FitBufs b;
  const uint32_t n = (uint32_t)a.n, nb = n * 4u, S = (uint32_t)a.n_slots;
  b.nb = nb;
  b.hit_pos = make_buf(a.hit_pos, 3u * nb * S);
  b.hit_err = make_buf(a.hit_err, 6u * nb * S);
  b.n_hits = make_buf(a.n_hits, nb);
  b.hit_flags = make_buf(a.hit_flags, a.hit_flags ? n * S : 0u);
  return b;
}

// One hit slot as read from HBM.
struct HitIn {
  float mp[3], me[6];
  unsigned flags;
};

// Loads slot s of track tb/4.  Slots at or past n_slots fall outside the
// buffers and read as zeros, so the next-slot prefetch needs no guard.
__device__ __forceinline__ void load_hit(const FitArgs& a, const FitBufs& b, int s, uint32_t tb, HitIn& h) {
  h.flags = a.hit_flags ? bld_u8(b.hit_flags, tb >> 2, (uint32_t)s * (uint32_t)a.n) : 0u;
  const uint32_t sp = (uint32_t)s * 3u * b.nb, se = (uint32_t)s * 6u * b.nb;
#pragma unroll
  for (int e = 0; e < 3; ++e) h.mp[e] = bld_f(b.hit_pos, tb, sp + (uint32_t)e * b.nb);
#pragma unroll
  for (int e = 0; e < 6; ++e) h.me[e] = bld_f(b.hit_err, tb, se + (uint32_t)e * b.nb);
}

// FP64 fallback of the covariance update (MKF_FIT_FP64_FALLBACK).  The test
// flags a negative or NaN variance or a covariance beyond what the variances
// allow by more than rounding, C_ij^2 > (1 + 1e-5) C_ii C_jj, with the
// products formed in double (exact for float factors).
__device__ __forceinline__ bool ill_conditioned(const float (&C)[21]) {
  bool bad = false;
#pragma unroll
  for (int i = 0; i < 6; ++i) bad |= !(C[S(i, i)] >= 0.f);
#pragma unroll
  for (int i = 1; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < i; ++j) {
      const double cij = C[S(i, j)];
      bad |= cij * cij > (double)C[S(i, i)] * (double)C[S(j, j)] * 1.00001;
    }
  return bad;
}

// The updated covariance in the Joseph form (I - KH) C (I - KH)^T + K V K^T,
// evaluated in double from the FP32 pre-update covariance C and rounded to
// float; H is the barrel (tangent-plane rotation and z) or endcap (x, y)
// projection, two rows h0 and h1, and V the hit covariance projected the same
// way.  Symmetric and positive semi-definite by construction for a positive
// semi-definite C.  With CH = C H^T (two columns), K = CH S^-1 and
// s_ab = h_a . CH_b, the product expands to rank-2 corrections of C:
//   AC   = C - K0 CH0^T - K1 CH1^T                       (A = I - K H)
//   AC h0 = CH0 - K0 s00 - K1 s01,  AC h1 = CH1 - K0 s01 - K1 s11
//   C'   = AC - (AC h0) K0^T - (AC h1) K1^T + K V K^T
// so no 6x6 temporary is formed and the routine stays inline at a modest
// register cost (about 50 VGPRs of doubles).
__device__ __forceinline__ void joseph_fp64(const float (&pre)[21], float px, float py, const float (&mp)[3],
                                            const float (&me)[6], bool endcap, float (&out)[21]) {
  const double vxx = me[0], vyx = me[1], vyy = me[2], vzx = me[3], vzy = me[4], vzz = me[5];
  double a00 = 1.0, a01 = 0.0, V00, V01, V11;
  if (endcap) {
    V00 = vxx; V01 = vyx; V11 = vyy;
  } else {
    const double mx = mp[0], my = mp[1];
    const double r = __builtin_sqrt(mx * mx + my * my);
    a00 = -(my + (double)py) / (2 * r);
    a01 = (mx + (double)px) / (2 * r);
    V00 = a00 * a00 * vxx + 2 * a00 * a01 * vyx + a01 * a01 * vyy;
    V01 = a00 * vzx + a01 * vzy;
    V11 = vzz;
  }
  // CH0 = C h0, CH1 = C h1: barrel h0 = (a00, a01, 0..), h1 = e2; endcap h0 = e0, h1 = e1.
  double CH0[6], CH1[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    CH0[i] = endcap ? (double)pre[S(i, 0)] : a00 * (double)pre[S(i, 0)] + a01 * (double)pre[S(i, 1)];
    CH1[i] = endcap ? (double)pre[S(i, 1)] : (double)pre[S(i, 2)];
  }
  const double s00 = endcap ? CH0[0] : a00 * CH0[0] + a01 * CH0[1];
  const double s01 = endcap ? CH1[0] : a00 * CH1[0] + a01 * CH1[1];
  const double s11 = endcap ? CH1[1] : CH1[2];
  const double S00 = s00 + V00, S01 = s01 + V01, S11 = s11 + V11;
  const double det = S00 * S11 - S01 * S01;
  const double i00 = S11 / det, i01 = -S01 / det, i11 = S00 / det;
  double K0[6], K1[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    K0[i] = CH0[i] * i00 + CH1[i] * i01;
    K1[i] = CH0[i] * i01 + CH1[i] * i11;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const double g0 = CH0[i] - K0[i] * s00 - K1[i] * s01;    // (AC h0)_i
    const double g1 = CH1[i] - K0[i] * s01 - K1[i] * s11;    // (AC h1)_i
    const double kv0 = K0[i] * V00 + K1[i] * V01, kv1 = K0[i] * V01 + K1[i] * V11;
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      const double ac = (double)pre[S(i, j)] - K0[i] * CH0[j] - K1[i] * CH1[j];
      out[S(i, j)] = (float)(ac - g0 * K0[j] - g1 * K1[j] + kv0 * K0[j] + kv1 * K1[j]);
    }
  }
}

template <bool FB>
__device__ __forceinline__ void fit_step(float (&par)[6], float (&err)[21], int& chg, float& chi2, const FitArgs& a,
                                         const HitIn& h) {
  const bool missing = h.flags & MKF_HIT_MISSING;
  float mp[3], me[6];
  if (missing) {
    mp[0] = par[0]; mp[1] = par[1]; mp[2] = par[2];
    me[0] = 666.f; me[1] = 0.f; me[2] = 666.f; me[3] = 0.f; me[4] = 0.f; me[5] = 666.f;
  } else {
#pragma unroll
    for (int e = 0; e < 3; ++e) mp[e] = h.mp[e];
#pragma unroll
    for (int e = 0; e < 6; ++e) me[e] = h.me[e];
  }
  float c2;
  const bool endcap = h.flags & MKF_HIT_ENDCAP;
  float pre[FB ? 21 : 1], px = 0.f, py = 0.f;
  if (!endcap) {
    const float r = fhypot(mp[0], mp[1]);                                  // MkBase.h:55
    propagate_to_r(par, err, chg, r, a.pflags, missing, a.bfield, a.mat);
  } else {
    propagate_to_z(par, err, chg, mp[2], a.pflags, missing, a.bfield, a.mat);
  }
  if (FB) {
#pragma unroll
    for (int e = 0; e < 21; ++e) pre[e] = err[e];
    px = par[0];
    py = par[1];
  }
  c2 = endcap ? kalman_endcap<true, true>(par, err, mp, me) : kalman_barrel<true, true>(par, err, mp, me);
  if constexpr (FB) {
    if (ill_conditioned(err)) joseph_fp64(pre, px, py, mp, me, endcap, err);
  }
  if (a.fit_flags & MKF_FIT_CHI2) chi2 += c2;
  if ((a.fit_flags & MKF_FIT_CHARGE_FLIP) && par[3] < 0.f) {
    chg = -chg;
    par[3] = -par[3];
  }
}

// MkBase::PropagateTracksToPCAZ (MkBase.h:91-104): the z of the point of
// closest approach to the origin of the track's straight-line r-z projection,
// then propagateHelixToZ there.  std::hypot of the reference is formed in double
// and rounded, as glibc's hypotf does.
__device__ __forceinline__ void propagate_to_pca(float (&par)[6], float (&err)[21], int chg, const FitArgs& a) {
  const float slope = tanf(par[5]);
  const double xd = par[0], yd = par[1];
  const float r = (float)__builtin_sqrt(xd * xd + yd * yd);
  const float z = (slope * (slope * par[2] - r)) / (1 + slope * slope);
  propagate_to_z(par, err, chg, z, a.pca_pflags, false, a.bfield, a.mat);
}

// from_seed: read the seed arrays (first touch of an out-of-place fit).
__device__ __forceinline__ void load_state(const FitArgs& a, int t, bool from_seed, float (&par)[6], float (&err)[21],
                                           int& chg, float& chi2) {
  const size_t n = (size_t)a.n;
  const float* P = (from_seed ? a.seed_par : a.par) + t;
  const float* E = (from_seed ? a.seed_err : a.err) + t;
#pragma unroll
  for (int e = 0; e < 6; ++e) par[e] = P[e * n];
#pragma unroll
  for (int e = 0; e < 21; ++e) err[e] = E[e * n];
  chg = from_seed ? a.seed_chg[t] : a.chg[t];
  chi2 = ((a.fit_flags & MKF_FIT_CHI2) && !from_seed) ? a.chi2[t] : 0.f;
}

__device__ __forceinline__ void store_state(const FitArgs& a, int t, const float (&par)[6], const float (&err)[21],
                                            int chg, float chi2) {
  const size_t n = (size_t)a.n;
  float* P = a.par + t;
  float* E = a.err + t;
#pragma unroll
  for (int e = 0; e < 6; ++e) P[e * n] = par[e];
#pragma unroll
  for (int e = 0; e < 21; ++e) E[e * n] = err[e];
  if ((a.fit_flags & MKF_FIT_CHARGE_FLIP) || a.seed_par) a.chg[t] = chg;
  if (a.fit_flags & MKF_FIT_CHI2) a.chi2[t] = chi2;
}

// One kernel body serves both fit forms, so they execute identical machine
// code and agree bit for bit:
//   plain   -- launched once per slot s with [s_begin, s_end) = [s, s+1) and one
//              thread per track; the state round-trips through HBM per launch.
//   batched -- launched once with [0, n_slots) on a persistent grid; each
//              thread keeps its track in VGPRs across all of its slots.
// With a seed input, the launch that starts at slot 0 reads the seeds and also
// carries tracks without hits over to the output.  MINW = minimum waves per
// SIMD requested from the register allocator (the autotuner's occupancy knob).
// FORM only names the instantiation (0 = plain, 1 = batched) so profiles keep
// the two forms apart; the body does not depend on it.
// PF: software-pipeline the hit loads (slot s+1 in flight while slot s
// computes) at the cost of 9 more live VGPRs; the loads do not change the
// arithmetic, so both settings give the same bits.
// With a track order (sort tiles) the blocks walk the positions in XCD-major
// order, so the wavefronts of one tile run on one XCD and the cache lines of
// the tile's hits, which its wavefronts share, are fetched into one L2.
template <int MINW, int FORM, bool FB, bool PF = true>
__global__ void __attribute__((amdgpu_flat_work_group_size(64, 512), amdgpu_waves_per_eu(MINW, 8)))
k_fit(FitArgs a, int s_begin, int s_end) {
  const int stride = gridDim.x * blockDim.x;
  const bool from_seed = s_begin == 0 && a.seed_par != nullptr;
  const FitBufs b = make_bufs(a);
  // The launch that ends at the last slot also takes every track to its PCA.
  const bool to_pca = (a.fit_flags & MKF_FIT_PROP_TO_PCA) && s_end >= a.n_slots;
  const int blk = a.order ? xcd_major_block() : (int)blockIdx.x;
  for (int i = blk * blockDim.x + threadIdx.x; i < a.n; i += stride) {
    const int t = a.order ? a.order[i] : i;
    const uint32_t tb = (uint32_t)t * 4u;  // lane byte offset in a 4-byte element array
    const int s1 = min(min(bld_i(b.n_hits, tb, 0), a.n_slots), s_end);
    if (s_begin >= s1 && !from_seed && !to_pca) continue;
    float par[6], err[21], chi2;
    int chg;
    load_state(a, t, from_seed, par, err, chg, chi2);
    if (PF) {
      HitIn cur, nxt;
      if (s_begin < s1) load_hit(a, b, s_begin, tb, nxt);
      for (int s = s_begin; s < s1; ++s) {
        cur = nxt;
        if (s + 1 < s1) load_hit(a, b, s + 1, tb, nxt);
        fit_step<FB>(par, err, chg, chi2, a, cur);
      }
    } else {
      for (int s = s_begin; s < s1; ++s) {
        HitIn cur;
        load_hit(a, b, s, tb, cur);
        fit_step<FB>(par, err, chg, chi2, a, cur);
      }
    }
    if (to_pca) propagate_to_pca(par, err, chg, a);
    store_state(a, t, par, err, chg, chi2);
  }
}

// Track order of the batched fit.  One block per tile of TILE consecutive
// tracks: each track gets the key (hit count, endcap flags of its slots,
// position in the tile) -- slot s of the flags in bit s mod 40 -- and the
// block sorts the tile's keys in LDS (bitonic network, 1024 threads), then
// writes the tracks in key order.  Tracks of one length and one barrel/endcap
// pattern then share wavefronts, so a wavefront's slot loop lasts as long as
// its tracks and takes one branch per slot; the permutation stays inside the
// tile, so the lines of a tile's hits are shared by its own wavefronts.  The
// fit of each track is the same whatever its position, so the order changes
// no result.
template <int TILE>
__global__ void __launch_bounds__(1024) k_fit_order(const int* __restrict__ n_hits,
                                                     const unsigned char* __restrict__ hit_flags, int n, int n_slots,
                                                     int* __restrict__ order) {
  static_assert(TILE >= 1024 && (TILE & (TILE - 1)) == 0 && TILE <= 16384, "tile: power of two, 1024..16384");
  __shared__ uint64_t key[TILE];
  const int base = blockIdx.x * TILE;
  const int m = min(TILE, n - base);
  for (int j = threadIdx.x; j < TILE; j += 1024) {
    uint64_t k = ~0ull;  // padding sorts last
    if (j < m) {
      const int t = base + j;
      const int nh = min(max(n_hits[t], 0), n_slots);
      uint64_t mask = 0;
      if (hit_flags)
        for (int s = 0; s < nh; ++s)
          if (hit_flags[(size_t)s * n + t] & MKF_HIT_ENDCAP) mask |= 1ull << (s % 40);
      k = ((uint64_t)min(nh, 1023) << 54) | (mask << 14) | (uint64_t)j;
    }
    key[j] = k;
  }
  for (int size = 2; size <= TILE; size <<= 1)
    for (int half = size >> 1; half > 0; half >>= 1) {
      __syncthreads();
      for (int p = threadIdx.x; p < TILE / 2; p += 1024) {
        const int lo = 2 * half * (p / half) + (p % half), hi = lo + half;
        const uint64_t x = key[lo], y = key[hi];
        const bool ascending = (lo & size) == 0;
        if ((x > y) == ascending) {
          key[lo] = y;
          key[hi] = x;
        }
      }
    }
  __syncthreads();
  for (int j = threadIdx.x; j < m; j += 1024) order[base + j] = base + (int)(key[j] & 0x3fff);
}

using OrderKernel = void (*)(const int*, const unsigned char*, int, int, int*);
static int order_kernel(int tile, OrderKernel* k) {
  switch (tile) {
    case 1024: *k = k_fit_order<1024>; return MKF_OK;
    case 4096: *k = k_fit_order<4096>; return MKF_OK;
    case 16384: *k = k_fit_order<16384>; return MKF_OK;
    default: return fail(MKF_ERR_INVALID_ARG, "sort_tile must be 0, 1024, 4096 or 16384");
  }
}

using FitKernel = void (*)(FitArgs, int, int);
// Register budgets below what the kernel needs (5 or 6 waves: 96 or 80
// VGPRs) make the compiler spill to scratch, and the rematerialised
// expressions then round differently; the variants stop at 4 waves so every
// choice gives identical results.
// [fp64 fallback][variant]
static const FitKernel kFusedVariants[2][6] = {
    {k_fit<1, 1, false>, k_fit<2, 1, false>, k_fit<4, 1, false>,
     k_fit<1, 1, false, false>, k_fit<2, 1, false, false>, k_fit<4, 1, false, false>},
    {k_fit<1, 1, true>, k_fit<2, 1, true>, k_fit<4, 1, true>,
     k_fit<1, 1, true, false>, k_fit<2, 1, true, false>, k_fit<4, 1, true, false>}};
static const FitKernel kPlainKernel[2] = {k_fit<1, 0, false>, k_fit<1, 0, true>};
static constexpr int kNumFusedVariants = 6;

// ============================================================================
// C ABI
// ============================================================================

namespace {

int validate_batch(const mkf_fit_batch* b, FitArgs& a, const mkf_context* c) {
  if (!b) return fail(MKF_ERR_INVALID_ARG, "batch is NULL");
  if (b->n_tracks < 0 || b->n_slots < 0) return fail(MKF_ERR_INVALID_ARG, "negative batch size");
  const size_t n = (size_t)b->n_tracks, S = (size_t)b->n_slots;
  // Buffer addressing: every array of the batch must stay below 4 GiB.
  if (21 * n * 4 >= ((size_t)1 << 32) || 6 * n * 4 * (S > 0 ? S : 1) >= ((size_t)1 << 32))
    return fail(MKF_ERR_INVALID_ARG, "batch too large: split it so every array stays below 4 GiB");
  a.n = b->n_tracks;
  a.n_slots = b->n_slots;
  a.par = b->par;
  a.err = b->err;
  a.chg = b->chg;
  a.chi2 = b->chi2;
  a.hit_pos = b->hit_pos;
  a.hit_err = b->hit_err;
  a.n_hits = b->n_hits;
  a.hit_flags = b->hit_flags;
  a.seed_par = b->seed_par;
  a.seed_err = b->seed_err;
  a.seed_chg = b->seed_chg;
  a.pflags = b->pflags;
  a.fit_flags = b->fit_flags;
  a.pca_pflags = b->pca_pflags;
  a.bfield = c->bfield;
  a.mat = mat_of(c);
  a.order = nullptr;
  if (n == 0) return MKF_OK;
  CHECK(check_dev(b->par, 6 * n * 4, "par"));
  CHECK(check_dev(b->err, 21 * n * 4, "err"));
  CHECK(check_dev(b->chg, n * 4, "chg"));
  if (b->fit_flags & MKF_FIT_CHI2) CHECK(check_dev(b->chi2, n * 4, "chi2"));
  if (b->seed_par) {
    CHECK(check_dev(b->seed_par, 6 * n * 4, "seed_par"));
    CHECK(check_dev(b->seed_err, 21 * n * 4, "seed_err"));
    CHECK(check_dev(b->seed_chg, n * 4, "seed_chg"));
  } else if (b->seed_err || b->seed_chg) {
    return fail(MKF_ERR_INVALID_ARG, "seed_err/seed_chg set without seed_par");
  }
  CHECK(check_dev(b->n_hits, n * 4, "n_hits"));
  if (S > 0) {
    CHECK(check_dev(b->hit_pos, 3 * n * S * 4, "hit_pos"));
    CHECK(check_dev(b->hit_err, 6 * n * S * 4, "hit_err"));
    if (b->hit_flags) CHECK(check_dev(b->hit_flags, n * S, "hit_flags"));
  }
  if (((b->pflags | ((b->fit_flags & MKF_FIT_PROP_TO_PCA) ? b->pca_pflags : 0)) & MKF_PF_APPLY_MATERIAL) && !c->d_rl)
    return fail(MKF_ERR_INVALID_ARG, "material requested but the context has no material grid");
  return MKF_OK;
}

int grid_for(int n, int block) { return (n + block - 1) / block; }

// The batched fit: with a sort tile, k_fit_order writes the track order to
// `order` (n ints of device memory) and the fused kernel follows it.
int launch_fused(mkf_context* c, FitArgs a, const mkf_launch_shape& sh, int* order, hipStream_t st) {
  if (a.n == 0) return MKF_OK;
  a.order = nullptr;
  if (sh.sort_tile > 0) {
    OrderKernel ok;
    CHECK(order_kernel(sh.sort_tile, &ok));
    if (!order) return fail(MKF_ERR_INVALID_ARG, "no track-order buffer");
    hipLaunchKernelGGL(ok, dim3((a.n + sh.sort_tile - 1) / sh.sort_tile), dim3(1024), 0, st, a.n_hits, a.hit_flags,
                       a.n, a.n_slots, order);
    HIP_TRY(hipGetLastError());
    a.order = order;
  }
  const int block = sh.block;
  int grid = sh.blocks_per_cu > 0 ? sh.blocks_per_cu * c->num_cus : grid_for(a.n, block);
  grid = std::max(1, std::min(grid, grid_for(a.n, block)));
  const int fb = (a.fit_flags & MKF_FIT_FP64_FALLBACK) ? 1 : 0;
  hipLaunchKernelGGL(kFusedVariants[fb][sh.variant], dim3(grid), dim3(block), 0, st, a, 0, a.n_slots);
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

bool shape_ok(const mkf_launch_shape& s) {
  return (s.block == 64 || s.block == 128 || s.block == 256 || s.block == 512) && s.blocks_per_cu >= 0 &&
         s.blocks_per_cu <= 64 && s.variant >= 0 && s.variant < kNumFusedVariants &&
         (s.sort_tile == 0 || s.sort_tile == 1024 || s.sort_tile == 4096 || s.sort_tile == 16384);
}

// Grows the context's track-order buffer to n ints.  Cached graphs hold the
// old address, so they are dropped (after the device drains) when it moves.
// Called under c->mu.
int ensure_fit_order(mkf_context* c, int n) {
  if ((size_t)n <= c->fit_order_cap) return MKF_OK;
  HIP_TRY(hipDeviceSynchronize());
  for (auto& kv : c->graphs) HIP_TRY(hipGraphExecDestroy(kv.second.exec));
  c->graphs.clear();
  if (c->fit_order) HIP_TRY(hipFree(c->fit_order));
  c->fit_order = nullptr;
  c->fit_order_cap = 0;
  HIP_TRY(hipMalloc(&c->fit_order, (size_t)n * sizeof(int)));
  c->fit_order_cap = (size_t)n;
  return MKF_OK;
}

}  // namespace

extern "C" {

const char* mkf_last_error(void) { return g_last_error.c_str(); }

int mkf_num_batched_variants(void) { return kNumFusedVariants; }

int mkf_context_create(const mkf_config* cfg, mkf_context** out) {
  if (!cfg || !out) return fail(MKF_ERR_INVALID_ARG, "cfg/out is NULL");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return fail(MKF_ERR_NO_DEVICE, "no HIP device");
  }
  if (cfg->device < 0 || cfg->device >= ndev) return fail(MKF_ERR_INVALID_ARG, "device %d out of range", cfg->device);
  if ((cfg->rl_grid == nullptr) != (cfg->xi_grid == nullptr))
    return fail(MKF_ERR_INVALID_ARG, "rl_grid and xi_grid must both be set or both be NULL");
  auto* c = new mkf_context();
  c->device = cfg->device;
  c->bfield = cfg->bfield;
  hipError_t e = hipSetDevice(c->device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, c->device);
  if (e == hipSuccess && cfg->rl_grid) {
    const size_t bytes = sizeof(float) * kNBinsZ * kNBinsR;
    e = hipMalloc(&c->d_rl, bytes);
    if (e == hipSuccess) e = hipMalloc(&c->d_xi, bytes);
    if (e == hipSuccess) e = hipMemcpy(c->d_rl, cfg->rl_grid, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_xi, cfg->xi_grid, bytes, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    int rc = fail(MKF_ERR_HIP, "context setup: %s", hipGetErrorString(e));
    mkf_context_destroy(c);
    return rc;
  }
  if (int rc = adjoint_noalloc_setup(c)) {
    mkf_context_destroy(c);
    return rc;
  }
  *out = c;
  return MKF_OK;
}

int mkf_context_destroy(mkf_context* c) {
  if (!c) return MKF_OK;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
  if (c->capture_stream) (void)hipStreamDestroy(c->capture_stream);
  if (c->adj_scratch) (void)hipFree(c->adj_scratch);
  if (c->adj_done) (void)hipEventDestroy(c->adj_done);
  if (c->fit_order) (void)hipFree(c->fit_order);
  if (c->fit_done) (void)hipEventDestroy(c->fit_done);
  if (c->d_rl) (void)hipFree(c->d_rl);
  if (c->d_xi) (void)hipFree(c->d_xi);
  delete c;
  return MKF_OK;
}

int mkf_propagate_helix_to_r(mkf_context* c, int n, const float* in_err, const float* in_par, const int* in_chg,
                             const float* ms_rad, float* out_err, float* out_par, int pflags, const int* no_mat,
                             int* out_fail, void* stream) {
  CHECK(set_device(c));
  if (n < 0) return fail(MKF_ERR_INVALID_ARG, "n < 0");
  if (n == 0) return MKF_OK;
  const size_t N = n;
  CHECK(check_dev(in_err, 21 * N * 4, "in_err"));
  CHECK(check_dev(in_par, 6 * N * 4, "in_par"));
  CHECK(check_dev(in_chg, N * 4, "in_chg"));
  CHECK(check_dev(ms_rad, N * 4, "ms_rad"));
  CHECK(check_dev(out_err, 21 * N * 4, "out_err"));
  CHECK(check_dev(out_par, 6 * N * 4, "out_par"));
  if (no_mat) CHECK(check_dev(no_mat, N * 4, "no_mat"));
  if (out_fail) CHECK(check_dev(out_fail, N * 4, "out_fail"));
  if ((pflags & MKF_PF_APPLY_MATERIAL) && !c->d_rl) return fail(MKF_ERR_INVALID_ARG, "no material grid");
  hipLaunchKernelGGL(k_propagate_r, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, in_err, in_par,
                     in_chg, ms_rad, out_err, out_par, pflags, no_mat, out_fail, c->bfield, mat_of(c));
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

int mkf_propagate_helix_to_z(mkf_context* c, int n, const float* in_err, const float* in_par, const int* in_chg,
                             const float* ms_z, float* out_err, float* out_par, int pflags, const int* no_mat,
                             void* stream) {
  CHECK(set_device(c));
  if (n < 0) return fail(MKF_ERR_INVALID_ARG, "n < 0");
  if (n == 0) return MKF_OK;
  const size_t N = n;
  CHECK(check_dev(in_err, 21 * N * 4, "in_err"));
  CHECK(check_dev(in_par, 6 * N * 4, "in_par"));
  CHECK(check_dev(in_chg, N * 4, "in_chg"));
  CHECK(check_dev(ms_z, N * 4, "ms_z"));
  CHECK(check_dev(out_err, 21 * N * 4, "out_err"));
  CHECK(check_dev(out_par, 6 * N * 4, "out_par"));
  if (no_mat) CHECK(check_dev(no_mat, N * 4, "no_mat"));
  if ((pflags & MKF_PF_APPLY_MATERIAL) && !c->d_rl) return fail(MKF_ERR_INVALID_ARG, "no material grid");
  hipLaunchKernelGGL(k_propagate_z, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, in_err, in_par,
                     in_chg, ms_z, out_err, out_par, pflags, no_mat, c->bfield, mat_of(c));
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

static int kalman_common(mkf_context* c, int op, int n, const float* ps_err, const float* ps_par,
                         const float* ms_err, const float* ms_par, float* out_err, float* out_par, float* out_chi2,
                         void* stream, bool endcap) {
  CHECK(set_device(c));
  if (n < 0) return fail(MKF_ERR_INVALID_ARG, "n < 0");
  if (!(op & (MKF_KFO_CALCULATE_CHI2 | MKF_KFO_UPDATE_PARAMS)) || (op & ~3))
    return fail(MKF_ERR_INVALID_ARG, "kf_op must be a nonzero combination of CALCULATE_CHI2 and UPDATE_PARAMS");
  if (n == 0) return MKF_OK;
  const size_t N = n;
  CHECK(check_dev(ps_err, 21 * N * 4, "ps_err"));
  CHECK(check_dev(ps_par, 6 * N * 4, "ps_par"));
  CHECK(check_dev(ms_err, 6 * N * 4, "ms_err"));
  CHECK(check_dev(ms_par, 3 * N * 4, "ms_par"));
  if (op & MKF_KFO_UPDATE_PARAMS) {
    CHECK(check_dev(out_err, 21 * N * 4, "out_err"));
    CHECK(check_dev(out_par, 6 * N * 4, "out_par"));
  }
  if ((op & MKF_KFO_CALCULATE_CHI2) && out_chi2) CHECK(check_dev(out_chi2, N * 4, "out_chi2"));
  if (endcap)
    hipLaunchKernelGGL(k_kalman<true>, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, op, n, ps_err,
                       ps_par, ms_err, ms_par, out_err, out_par, out_chi2);
  else
    hipLaunchKernelGGL(k_kalman<false>, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, op, n, ps_err,
                       ps_par, ms_err, ms_par, out_err, out_par, out_chi2);
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

int mkf_kalman_operation(mkf_context* c, int op, int n, const float* ps_err, const float* ps_par,
                         const float* ms_err, const float* ms_par, float* out_err, float* out_par, float* out_chi2,
                         void* stream) {
  return kalman_common(c, op, n, ps_err, ps_par, ms_err, ms_par, out_err, out_par, out_chi2, stream, false);
}

int mkf_kalman_operation_endcap(mkf_context* c, int op, int n, const float* ps_err, const float* ps_par,
                                const float* ms_err, const float* ms_par, float* out_err, float* out_par,
                                float* out_chi2, void* stream) {
  return kalman_common(c, op, n, ps_err, ps_par, ms_err, ms_par, out_err, out_par, out_chi2, stream, true);
}

int mkf_fit_plain(mkf_context* c, const mkf_fit_batch* b, void* stream) {
  CHECK(set_device(c));
  FitArgs a;
  CHECK(validate_batch(b, a, c));
  if (a.n == 0) return MKF_OK;
  const int launches = std::max(a.n_slots, (a.seed_par || (a.fit_flags & MKF_FIT_PROP_TO_PCA)) ? 1 : 0);
  for (int s = 0; s < launches; ++s) {
    hipLaunchKernelGGL(kPlainKernel[(a.fit_flags & MKF_FIT_FP64_FALLBACK) ? 1 : 0], dim3(grid_for(a.n, 256)),
                       dim3(256), 0, (hipStream_t)stream, a, s, s + 1);
    HIP_TRY(hipGetLastError());
  }
  return MKF_OK;
}

int mkf_fit_batched(mkf_context* c, const mkf_fit_batch* b, void* stream) {
  CHECK(set_device(c));
  FitArgs a;
  CHECK(validate_batch(b, a, c));
  if (a.n == 0) return MKF_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  const bool sorted = c->shape.sort_tile > 0;
  if (sorted) CHECK(ensure_fit_order(c, a.n));
  GraphKey key;
  const void* ptrs[12] = {b->par,    b->err,       b->chg,      b->chi2,     b->hit_pos,  b->hit_err,
                          b->n_hits, b->hit_flags, b->seed_par, b->seed_err, b->seed_chg, nullptr};
  std::memcpy(key.ptrs, ptrs, sizeof(ptrs));
  key.n = a.n; key.slots = a.n_slots; key.pflags = a.pflags; key.fflags = a.fit_flags;
  key.block = c->shape.block; key.bpc = c->shape.blocks_per_cu; key.variant = c->shape.variant;
  key.pca_pflags = a.pca_pflags;
  key.tile = c->shape.sort_tile;
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    if (c->graphs.size() >= mkf_context::kMaxGraphs) {
      // Evict the least recently launched graph; a launch of it may still be
      // in flight, so the device drains first (rare: only callers that cycle
      // through more than kMaxGraphs distinct batches get here).
      auto lru = c->graphs.begin();
      for (auto j = c->graphs.begin(); j != c->graphs.end(); ++j)
        if (j->second.last_use < lru->second.last_use) lru = j;
      HIP_TRY(hipDeviceSynchronize());
      HIP_TRY(hipGraphExecDestroy(lru->second.exec));
      c->graphs.erase(lru);
    }
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->capture_stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_fused(c, a, c->shape, c->fit_order, c->capture_stream);
    hipError_t e = hipStreamEndCapture(c->capture_stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(MKF_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(MKF_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    it = c->graphs.emplace(key, mkf_context::GraphEntry{ge, 0}).first;
  }
  it->second.last_use = ++c->graph_tick;
  // The track-order buffer is shared by every call on this context: a call on
  // another stream waits for the previous call's graph to finish with it.
  if (sorted) CHECK(scratch_acquire(c->fit_done, (hipStream_t)stream));
  HIP_TRY(hipGraphLaunch(it->second.exec, (hipStream_t)stream));
  if (sorted) CHECK(scratch_release(c->fit_done, (hipStream_t)stream));
  return MKF_OK;
}

int mkf_fit_order(mkf_context* c, const mkf_fit_batch* b, int sort_tile, int* order, void* stream) {
  CHECK(set_device(c));
  FitArgs a;
  CHECK(validate_batch(b, a, c));
  if (a.n == 0) return MKF_OK;
  OrderKernel ok;
  CHECK(order_kernel(sort_tile, &ok));
  CHECK(check_dev(order, (size_t)a.n * sizeof(int), "order"));
  hipLaunchKernelGGL(ok, dim3((a.n + sort_tile - 1) / sort_tile), dim3(1024), 0, (hipStream_t)stream, a.n_hits,
                     a.hit_flags, a.n, a.n_slots, order);
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

int mkf_set_batched_shape(mkf_context* c, const mkf_launch_shape* s) {
  if (!c || !s) return fail(MKF_ERR_INVALID_ARG, "NULL argument");
  if (!shape_ok(*s)) return fail(MKF_ERR_INVALID_ARG, "invalid launch shape");
  std::lock_guard<std::mutex> lk(c->mu);
  c->shape = *s;
  return MKF_OK;
}

int mkf_get_batched_shape(mkf_context* c, mkf_launch_shape* s) {
  if (!c || !s) return fail(MKF_ERR_INVALID_ARG, "NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  *s = c->shape;
  return MKF_OK;
}

int mkf_autotune_batched(mkf_context* c, const mkf_fit_batch* b, float* scratch, int reps, mkf_launch_shape* best,
                         float* best_ms) {
  CHECK(set_device(c));
  FitArgs a;
  CHECK(validate_batch(b, a, c));
  if (a.n == 0) return fail(MKF_ERR_INVALID_ARG, "empty batch");
  if (reps < 1) reps = 1;
  const size_t N = a.n;
  CHECK(check_dev(scratch, 29 * N * 4, "scratch_state"));
  hipStream_t st = c->capture_stream;
  // Save the state the fit mutates: par, err, chi2, chg.
  auto save = [&](bool restore) -> hipError_t {
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    float* sp = scratch;
    hipError_t e;
    if (restore) {
      if ((e = hipMemcpyAsync(a.par, sp, 6 * N * 4, k, st))) return e;
      if ((e = hipMemcpyAsync(a.err, sp + 6 * N, 21 * N * 4, k, st))) return e;
      if ((e = hipMemcpyAsync(a.chg, sp + 27 * N, N * 4, k, st))) return e;
      if (a.fit_flags & MKF_FIT_CHI2) if ((e = hipMemcpyAsync(a.chi2, sp + 28 * N, N * 4, k, st))) return e;
    } else {
      if ((e = hipMemcpyAsync(sp, a.par, 6 * N * 4, k, st))) return e;
      if ((e = hipMemcpyAsync(sp + 6 * N, a.err, 21 * N * 4, k, st))) return e;
      if ((e = hipMemcpyAsync(sp + 27 * N, a.chg, N * 4, k, st))) return e;
      if (a.fit_flags & MKF_FIT_CHI2) if ((e = hipMemcpyAsync(sp + 28 * N, a.chi2, N * 4, k, st))) return e;
    }
    return hipSuccess;
  };
  HIP_TRY(save(false));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const int blocks[] = {64, 128, 256, 512};
  const int bpcs[] = {0, 1, 2, 4, 8, 16};
  // Sort tiles (k_fit_order) stay out of the search: the fused kernel's loads through a tile's permutation
  // cost more than the ordering saves (2M uneven tracks: 1.70 ms at best against 1.61 ms unordered, and
  // 0.80 ms for the same batch packed in fit order); set_batched_shape still selects them.
  const int tiles[] = {0};
  mkf_launch_shape bs = c->shape;
  float bt = 1e30f;
  int rc = MKF_OK;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    rc = ensure_fit_order(c, a.n);
  }
  int* order = c->fit_order;
  if (!rc) rc = scratch_acquire(c->fit_done, st);
  for (int tile : tiles)
  for (int v = 0; v < kNumFusedVariants && !rc; ++v)
    for (int bl : blocks)
      for (int bp : bpcs) {
        if (rc) break;
        mkf_launch_shape sh{bl, bp, v, tile};
        float tot = 0.f;
        for (int r = 0; r < reps + 1 && !rc; ++r) {  // first run is a warm-up
          if (save(true) != hipSuccess) { rc = fail(MKF_ERR_HIP, "autotune restore failed"); break; }
          (void)hipEventRecord(e0, st);
          rc = launch_fused(c, a, sh, order, st);
          (void)hipEventRecord(e1, st);
          if (hipEventSynchronize(e1) != hipSuccess) { rc = fail(MKF_ERR_HIP, "autotune sync failed"); break; }
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (r > 0) tot += ms;
        }
        if (!rc && tot / reps < bt) { bt = tot / reps; bs = sh; }
      }
  if (!rc && save(true) != hipSuccess) rc = fail(MKF_ERR_HIP, "autotune final restore failed");
  if (c->fit_done) (void)scratch_release(c->fit_done, st);
  (void)hipStreamSynchronize(st);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    c->shape = bs;
  }
  if (best) *best = bs;
  if (best_ms) *best_ms = bt;
  return MKF_OK;
}

}  // extern "C"
