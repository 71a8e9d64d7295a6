// oracle_core.h -- CPU restatement of mkFit's Matriplex Kalman hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
// compared against, and the CPU baseline timed beside it.  Nothing in the
// product (mkfit_amd/) includes, links or calls it.
//
// It restates, operation for operation, the reference CPU code:
//   helixAtRFromIterativeCCS_impl    mkFit/PropagationMPlex.icc:6-246
//   propagateHelixToRMPlex           mkFit/PropagationMPlex.cc:408-505
//   propagateHelixToZMPlex/helixAtZ  mkFit/PropagationMPlex.cc:510-712
//   applyMaterialEffects             mkFit/PropagationMPlex.cc:716-767
//   MultHelixProp[Transp][Endcap]    mkFit/PropagationMPlex.cc:86-140 with the
//                                    J sparsity of mkFit/GenMPlexOps.pl:254-324
//   kalmanOperation[Endcap]          mkFit/KalmanUtilsMPlex.cc:548-681,774-888
//                                    helpers KalmanUtilsMPlex.cc:14-394, .icc
//   2x2 Cramer inverse (double det)  Matriplex/MatriplexSym.h:422-444
//   sincos4, hipo                    Matrix.h:38-56
//   constants                        Config.h:77-85,123-126,230-248,418-421
//
// Layout mirrors Matriplex (Matriplex/Matriplex.h:65): a block holds NN lanes,
// element e of lane n lives at a[e*NN + n]; symmetric 6x6 is packed lower
// triangle (MatriplexSym.h:15-22).  Loops over lanes carry `omp simd`, as the
// reference's do, so the float instantiation vectorises like the reference
// when built with its compiler flags.  T = float is the parity path, T = double
// is the FP64 oracle (same algorithm, same float-valued constants).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#ifndef ORC_NN
#define ORC_NN 16
#endif

namespace orc {

constexpr int NN = ORC_NN;

// ---- constants (Config.h) ---------------------------------------------------
constexpr float kPI      = 3.14159265358979323846;   // Config.h:77
constexpr float kTwoPI   = 6.28318530717958647692;   // Config.h:78
constexpr float kPIOver2 = kPI / 2.0f;               // Config.h:79
constexpr float kSol     = 0.299792458;              // Config.h:85
constexpr int   kNiter   = 5;                        // Config.h:230
constexpr float kMagC1   = 3.8114;                   // Config.h:245-248
constexpr float kMagB0   = -3.94991e-06;
constexpr float kMagB1   = 7.53701e-06;
constexpr float kMagA    = 2.43878e-11;
constexpr float kRangeZME = 300.;                    // Config.h:123-126
constexpr int   kNBinsZME = 300;
constexpr float kRangeRME = 120.;
constexpr int   kNBinsRME = 120;

enum { PF_use_param_b_field = 0x1, PF_apply_material = 0x2 };   // Config.h:17-23
enum { KFO_Calculate_Chi2 = 1, KFO_Update_Params = 2 };          // KalmanUtilsMPlex.h:11-15

// Default of Ctx::tie_shift for every context created afterwards (orc_set_tie_shift; checker only).
inline float g_tie_shift = 0.f;

struct Ctx {
  float bfield;            // uniform field in T (Config::Bfield)
  const float* rl_grid;    // [kNBinsZME][kNBinsRME] radiation lengths (Config::RlgridME)
  const float* xi_grid;    // [kNBinsZME][kNBinsRME] (Config::XigridME)
  // Optional per-lane records [NN] of how close the helix step came to its two
  // thresholds: min over iterations of | |r - r0| - 1e-4 | in units of the
  // float spacing at r (the switch of the derivative terms) and of
  // |oodotp - 5| (the failure test).  A checker uses them to tell threshold
  // ties from rounding noise.
  float* tie_r = nullptr;
  float* tie_dot = nullptr;
  // min over lookups of the distance, in float spacings of the coordinate, of
  // the material lookup point to a bin edge across which the grid changes
  float* tie_mat = nullptr;
  // Moves the helix step's 1e-4 cm threshold by this many float spacings of r
  // (0 = the reference's test): the checker's "other side of the tie" result.
  float tie_shift = g_tie_shift;
};

// Symmetric packed index, MatriplexSym.h:21.
constexpr int sym(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

template <class T> inline T hipo(T x, T y) { return std::sqrt(x * x + y * y); }   // Matrix.h:38-41

template <class T> inline void sincos4(const T x, T& s, T& c) {                    // Matrix.h:48-56
  const T x2 = x * x;
  c = T(1) - T(0.5f) * x2 + T(0.04166667f) * x2 * x2;
  s = x - T(0.16666667f) * x * x2;
}

template <class T> inline T bfield_from_zr(T z, T r) {                          // Config.h:418-421
  return (T(kMagB0) * z * z + T(kMagB1) * z + T(kMagC1)) * (T(kMagA) * r * r + T(1.f));
}

// Jacobian sparsity patterns: 0 = structural zero, 1 = structural one, 2 = general.
// Barrel: GenMPlexOps.pl:257-264.  Endcap: GenMPlexOps.pl:303-310.
constexpr int kJPatBarrel[36] = {2, 2, 0, 2, 2, 0,
                                 2, 2, 0, 2, 2, 0,
                                 2, 2, 1, 2, 2, 2,
                                 2, 2, 0, 2, 2, 0,
                                 2, 2, 0, 2, 2, 0,
                                 0, 0, 0, 0, 0, 1};
constexpr int kJPatEndcap[36] = {1, 0, 2, 2, 2, 2,
                                 0, 1, 2, 2, 2, 2,
                                 0, 0, 0, 0, 0, 0,
                                 0, 0, 0, 1, 0, 0,
                                 0, 0, 2, 2, 1, 2,
                                 0, 0, 0, 0, 0, 1};

// temp = J * C (C symmetric), full 6x6 result; sum over k ascending, skipping
// structural zeros and folding structural ones, as the generated .ah code does
// (Matriplex/GenMul.pm:482-533).  The lane loop is innermost so each term is
// one vectorisable pass over the block.
template <class T, const int* PAT>
inline void mult_helix_prop(const T* __restrict__ J, const T* __restrict__ C, T* __restrict__ out) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      T* o = out + (i * 6 + j) * NN;
      bool first = true;
      for (int k = 0; k < 6; ++k) {
        const int p = PAT[i * 6 + k];
        if (p == 0) continue;
        const T* c = C + sym(k, j) * NN;
        const T* a = J + (i * 6 + k) * NN;
        if (first) {
          if (p == 1) {
#pragma omp simd
            for (int n = 0; n < NN; ++n) o[n] = c[n];
          } else {
#pragma omp simd
            for (int n = 0; n < NN; ++n) o[n] = a[n] * c[n];
          }
          first = false;
        } else {
          if (p == 1) {
#pragma omp simd
            for (int n = 0; n < NN; ++n) o[n] = o[n] + c[n];
          } else {
#pragma omp simd
            for (int n = 0; n < NN; ++n) o[n] = o[n] + a[n] * c[n];
          }
        }
      }
      if (first) {
#pragma omp simd
        for (int n = 0; n < NN; ++n) o[n] = 0;
      }
    }
}

// C = temp * J^T, symmetric result (lower triangle only).
template <class T, const int* PAT>
inline void mult_helix_prop_transp(const T* __restrict__ J, const T* __restrict__ temp, T* __restrict__ C) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      T* o = C + sym(i, j) * NN;
      bool first = true;
      for (int k = 0; k < 6; ++k) {
        const int p = PAT[j * 6 + k];
        if (p == 0) continue;
        const T* b = temp + (i * 6 + k) * NN;
        const T* a = J + (j * 6 + k) * NN;
        if (first) {
          if (p == 1) {
#pragma omp simd
            for (int n = 0; n < NN; ++n) o[n] = b[n];
          } else {
#pragma omp simd
            for (int n = 0; n < NN; ++n) o[n] = b[n] * a[n];
          }
          first = false;
        } else {
          if (p == 1) {
#pragma omp simd
            for (int n = 0; n < NN; ++n) o[n] = o[n] + b[n];
          } else {
#pragma omp simd
            for (int n = 0; n < NN; ++n) o[n] = o[n] + b[n] * a[n];
          }
        }
      }
      if (first) {
#pragma omp simd
        for (int n = 0; n < NN; ++n) o[n] = 0;
      }
    }
}

template <class T> inline void squash_phi(T* par) {                                // PropagationMPlex.h:8-15
#pragma omp simd
  for (int n = 0; n < NN; ++n) {
    if (par[4 * NN + n] >= T(kPI)) par[4 * NN + n] -= T(kTwoPI);
    if (par[4 * NN + n] < -T(kPI)) par[4 * NN + n] += T(kTwoPI);
  }
}

// ---- helixAtRFromIterativeCCS_impl (PropagationMPlex.icc:6-246) --------------
// outPar must hold inPar on entry (PropagationMPlex.cc:420); errorProp zeroed
// and failFlag cleared here (PropagationMPlex.cc:396-405).
template <class T>
inline void helix_at_r(const T* inPar, const int* inChg, const T* msRad, T* outPar, T* errorProp,
                       int* failFlag, int pflags, const Ctx& ctx) {
  std::memset(errorProp, 0, sizeof(T) * 36 * NN);
  for (int n = 0; n < NN; ++n) failFlag[n] = 0;
#pragma omp simd
  for (int n = 0; n < NN; ++n) {
    T* E = errorProp + n;
    auto J = [&](int i, int j) -> T& { return E[(i * 6 + j) * NN]; };
    J(0, 0) = 1; J(1, 1) = 1; J(2, 2) = 1; J(3, 3) = 1; J(4, 4) = 1; J(5, 5) = 1;

    T r0 = hipo(inPar[0 * NN + n], inPar[1 * NN + n]);
    const T bf = (pflags & PF_use_param_b_field) ? bfield_from_zr(inPar[2 * NN + n], r0) : T(ctx.bfield);
    const T k = T(inChg[n]) * T(100.f) / (-T(kSol) * bf);
    const T r = msRad[n];

    const T xin = inPar[0 * NN + n];
    const T yin = inPar[1 * NN + n];
    const T ipt = inPar[3 * NN + n];
    const T phiin = inPar[4 * NN + n];
    const T theta = inPar[5 * NN + n];

    const T kinv = T(1) / k;
    const T pt = T(1) / ipt;

    T D = 0, cosa = 0, sina = 0, cosah = 0, sinah = 0, id = 0;
    T cosPorT = std::cos(phiin), sinPorT = std::sin(phiin);
    T pxin = cosPorT * pt;
    T pyin = sinPorT * pt;

    T dDdx = r0 > T(0) ? -xin / r0 : T(0);
    T dDdy = r0 > T(0) ? -yin / r0 : T(0);
    T dDdipt = 0;
    T dDdphi = 0;

    for (int i = 0; i < kNiter; ++i) {
      r0 = hipo(outPar[0 * NN + n], outPar[1 * NN + n]);
      const T oodotp = r0 * pt / (pxin * outPar[0 * NN + n] + pyin * outPar[1 * NN + n]);
      if (ctx.tie_dot) ctx.tie_dot[n] = std::min(ctx.tie_dot[n], float(std::abs(oodotp - T(5.0f))));
      if (oodotp > T(5.0f) || oodotp < T(0)) {
        id = 0;
        failFlag[n] = 1;
      } else {
        id = (r - r0) * oodotp;
      }
      D += id;

      sincos4(id * ipt * kinv * T(0.5f), sinah, cosah);
      cosa = T(1) - T(2) * sinah * sinah;
      sina = T(2) * sinah * cosah;

      if (i + 1 != kNiter) {
        const T x = outPar[0 * NN + n];
        const T y = outPar[1 * NN + n];
        T thr = T(0.0001f);
        if (ctx.tie_r || ctx.tie_shift != 0.f) {
          const float rf = float(r), ulp = std::nextafter(rf, 2 * rf + 1) - rf;
          if (ctx.tie_r) ctx.tie_r[n] = std::min(ctx.tie_r[n], float(std::abs(std::abs(r - r0) - thr)) / ulp);
          thr += T(ctx.tie_shift * ulp);
        }
        const T oor0 = (r0 > T(0) && std::abs(r - r0) < thr) ? T(1) / r0 : T(0);
        const T dadipt = id * kinv;
        const T dadx = -x * ipt * kinv * oor0;
        const T dady = -y * ipt * kinv * oor0;
        const T pxca = pxin * cosa;
        const T pxsa = pxin * sina;
        const T pyca = pyin * cosa;
        const T pysa = pyin * sina;
        T tmp;
        tmp = k * dadx;
        dDdx -= (x * (T(1) + tmp * (pxca - pysa)) + y * tmp * (pyca + pxsa)) * oor0;
        tmp = k * dady;
        dDdy -= (x * tmp * (pxca - pysa) + y * (T(1) + tmp * (pyca + pxsa))) * oor0;
        tmp = dadipt * ipt;
        dDdipt -= k * (x * (pxca * tmp - pysa * tmp - pyca - pxsa + pyin) +
                       y * (pyca * tmp + pxsa * tmp - pysa + pxca - pxin)) * pt * oor0;
        dDdphi += k * (x * (pysa - pxin + pxca) - y * (pxsa - pyin + pyca)) * oor0;
      }

      outPar[0 * NN + n] = outPar[0 * NN + n] + T(2) * k * sinah * (pxin * cosah - pyin * sinah);
      outPar[1 * NN + n] = outPar[1 * NN + n] + T(2) * k * sinah * (pyin * cosah + pxin * sinah);
      const T pxinold = pxin;
      pxin = pxin * cosa - pyin * sina;
      pyin = pyin * cosa + pxinold * sina;
    }

    const T alpha = D * ipt * kinv;
    const T dadx = dDdx * ipt * kinv;
    const T dady = dDdy * ipt * kinv;
    const T dadipt = (ipt * dDdipt + D) * kinv;
    const T dadphi = dDdphi * ipt * kinv;

    sincos4(alpha, sina, cosa);

    J(0, 0) = T(1) + k * dadx * (cosPorT * cosa - sinPorT * sina) * pt;
    J(0, 1) = k * dady * (cosPorT * cosa - sinPorT * sina) * pt;
    J(0, 2) = 0;
    J(0, 3) = k * (cosPorT * (ipt * dadipt * cosa - sina) + sinPorT * ((T(1) - cosa) - ipt * dadipt * sina)) * pt * pt;
    J(0, 4) = k * (cosPorT * dadphi * cosa - sinPorT * dadphi * sina - sinPorT * sina + cosPorT * cosa - cosPorT) * pt;
    J(0, 5) = 0;

    J(1, 0) = k * dadx * (sinPorT * cosa + cosPorT * sina) * pt;
    J(1, 1) = T(1) + k * dady * (sinPorT * cosa + cosPorT * sina) * pt;
    J(1, 2) = 0;
    J(1, 3) = k * (sinPorT * (ipt * dadipt * cosa - sina) + cosPorT * (ipt * dadipt * sina - (T(1) - cosa))) * pt * pt;
    J(1, 4) = k * (sinPorT * dadphi * cosa + cosPorT * dadphi * sina + sinPorT * cosa + cosPorT * sina - sinPorT) * pt;
    J(1, 5) = 0;

    cosPorT = std::cos(theta);
    sinPorT = std::sin(theta);
    sinPorT = T(1) / sinPorT;

    outPar[2 * NN + n] = inPar[2 * NN + n] + k * alpha * cosPorT * pt * sinPorT;

    J(2, 0) = k * cosPorT * dadx * pt * sinPorT;
    J(2, 1) = k * cosPorT * dady * pt * sinPorT;
    J(2, 2) = 1;
    J(2, 3) = k * cosPorT * (ipt * dadipt - alpha) * pt * pt * sinPorT;
    J(2, 4) = k * dadphi * cosPorT * pt * sinPorT;
    J(2, 5) = -k * alpha * pt * sinPorT * sinPorT;

    outPar[3 * NN + n] = ipt;
    J(3, 0) = 0; J(3, 1) = 0; J(3, 2) = 0; J(3, 3) = 1; J(3, 4) = 0; J(3, 5) = 0;

    outPar[4 * NN + n] = inPar[4 * NN + n] + alpha;
    J(4, 0) = dadx;
    J(4, 1) = dady;
    J(4, 2) = 0;
    J(4, 3) = dadipt;
    J(4, 4) = T(1) + dadphi;
    J(4, 5) = 0;

    outPar[5 * NN + n] = theta;
    J(5, 0) = 0; J(5, 1) = 0; J(5, 2) = 0; J(5, 3) = 0; J(5, 4) = 0; J(5, 5) = 1;
  }
}

// ---- helixAtZ (PropagationMPlex.cc:604-712) ---------------------------------
template <class T>
inline void helix_at_z(const T* inPar, const int* inChg, const T* msZ, T* outPar, T* errorProp,
                       int pflags, const Ctx& ctx) {
  std::memset(errorProp, 0, sizeof(T) * 36 * NN);
#pragma omp simd
  for (int n = 0; n < NN; ++n) {
    T* E = errorProp + n;
    auto J = [&](int i, int j) -> T& { return E[(i * 6 + j) * NN]; };
    J(0, 0) = 1; J(1, 1) = 1; J(3, 3) = 1; J(4, 4) = 1; J(5, 5) = 1;

    const T zout = msZ[n];
    const T zin = inPar[2 * NN + n];
    const T ipt = inPar[3 * NN + n];
    const T phiin = inPar[4 * NN + n];
    const T theta = inPar[5 * NN + n];

    const T bf = (pflags & PF_use_param_b_field)
                     ? bfield_from_zr(zin, hipo(inPar[0 * NN + n], inPar[1 * NN + n]))
                     : T(ctx.bfield);
    const T k = T(inChg[n]) * T(100.f) / (-T(kSol) * bf);
    const T kinv = T(1) / k;
    const T pt = T(1) / ipt;

    T cosahTmp = 0, sinahTmp = 0;
    const T cosP = std::cos(phiin), sinP = std::sin(phiin);
    const T cosT = std::cos(theta), sinT = std::sin(theta);
    const T tanT = sinT / cosT;
    const T icos2T = T(1) / (cosT * cosT);
    const T pxin = cosP * pt;
    const T pyin = sinP * pt;

    const T deltaZ = zout - zin;
    const T alpha = deltaZ * tanT * ipt * kinv;

    sincos4(alpha * T(0.5f), sinahTmp, cosahTmp);
    const T cosah = cosahTmp;
    const T sinah = sinahTmp;
    const T cosa = T(1) - T(2) * sinah * sinah;
    const T sina = T(2) * sinah * cosah;

    outPar[0 * NN + n] = outPar[0 * NN + n] + T(2) * k * sinah * (pxin * cosah - pyin * sinah);
    outPar[1 * NN + n] = outPar[1 * NN + n] + T(2) * k * sinah * (pyin * cosah + pxin * sinah);
    outPar[2 * NN + n] = zout;
    outPar[4 * NN + n] = phiin + alpha;

    const T pxcaMpysa = pxin * cosa - pyin * sina;
    J(0, 2) = -tanT * ipt * pxcaMpysa;
    J(0, 3) = k * pt * pt * (cosP * (alpha * cosa - sina) + sinP * T(2) * sinah * (sinah - alpha * cosah));
    J(0, 4) = -T(2) * k * pt * sinah * (sinP * cosah + cosP * sinah);
    J(0, 5) = deltaZ * ipt * pxcaMpysa * icos2T;

    const T pycaPpxsa = pyin * cosa + pxin * sina;
    J(1, 2) = -tanT * ipt * pycaPpxsa;
    J(1, 3) = k * pt * pt * (sinP * (alpha * cosa - sina) - cosP * T(2) * sinah * (sinah - alpha * cosah));
    J(1, 4) = T(2) * k * pt * sinah * (cosP * cosah - sinP * sinah);
    J(1, 5) = deltaZ * ipt * pycaPpxsa * icos2T;

    J(4, 2) = -ipt * tanT * kinv;
    J(4, 3) = tanT * deltaZ * kinv;
    J(4, 5) = ipt * deltaZ * kinv * icos2T;
  }
}

// ---- applyMaterialEffects (PropagationMPlex.cc:716-767) ---------------------
template <class T>
inline void apply_material_effects(const T* hitsRl, const T* hitsXi, const T* propSign, T* outErr, T* outPar,
                                   bool isBarrel) {
#pragma omp simd
  for (int n = 0; n < NN; ++n) {
    T radL = hitsRl[n];
    if (radL < T(1e-13f)) continue;
    const T theta = outPar[5 * NN + n];
    const T pt = T(1) / outPar[3 * NN + n];
    const T p = pt / std::sin(theta);
    const T p2 = p * p;
    constexpr float mpi = 0.140;
    constexpr float mpi2 = mpi * mpi;
    const T beta2 = p2 / (p2 + T(mpi2));
    const T beta = std::sqrt(beta2);
    const T invCos = isBarrel ? p / pt : T(1) / std::abs(std::cos(theta));
    radL = radL * invCos;
    if (radL < T(1e-13f)) continue;
    const T thetaMSC = T(0.0136f) * (T(1) + T(0.038f) * std::log(radL)) / (beta * p);
    const T thetaMSC2 = thetaMSC * thetaMSC * radL;
    outErr[sym(4, 4) * NN + n] += thetaMSC2;
    outErr[sym(5, 5) * NN + n] += thetaMSC2;
    const T gamma2 = (p2 + T(mpi2)) / T(mpi2);
    const T gamma = std::sqrt(gamma2);
    constexpr float me = 0.0005;
    const T wmax = T(2) * T(me) * beta2 * gamma2 / (T(1) + T(2) * gamma * T(me) / T(mpi) + T(me) * T(me) / (T(mpi) * T(mpi)));
    constexpr float I = 16.0e-9 * 10.75;
    const T deltahalf = std::log(T(28.816e-9f) * std::sqrt(T(2.33f) * T(0.498f)) / T(I)) + std::log(beta * gamma) - T(0.5f);
    const T dEdx = beta < T(1)
                       ? (T(2) * (hitsXi[n] * invCos * (T(0.5f) * std::log(T(2) * T(me) * beta2 * gamma2 * wmax / (T(I) * T(I))) - beta2 - deltahalf) / beta2))
                       : T(0);
    const T dP = propSign[n] * dEdx / beta;
    outPar[3 * NN + n] = p / (std::max(p + dP, T(0.001f)) * pt);
    outErr[sym(3, 3) * NN + n] += dP * dP / (p2 * pt * pt);
  }
}

// Material lookup, MaterialEffects.h:8-11 and PropagationMPlex.cc:446-468/543-565.
// The bin arithmetic is float in the reference; other T are rounded to float first.
template <class T> inline int zbin_me(T z) {
  const float zf = (float)static_cast<double>(z);
  return (std::abs(zf) * kNBinsZME) / (kRangeZME);
}
template <class T> inline int rbin_me(T r) {
  const float rf = (float)static_cast<double>(r);
  return (rf * kNBinsRME) / (kRangeRME);
}

template <class T> inline void lookup_material(const Ctx& ctx, T zq, T rq, T& rl, T& xi) {
  const int zb = zbin_me(zq), rb = rbin_me(rq);
  const bool ok = zb >= 0 && zb < kNBinsZME && rb >= 0 && rb < kNBinsRME && ctx.rl_grid;
  rl = ok ? T(ctx.rl_grid[zb * kNBinsRME + rb]) : T(0);
  xi = ok ? T(ctx.xi_grid[zb * kNBinsRME + rb]) : T(0);
}

// Distance of a lookup point to the nearest bin edge across which the material grid changes, in float
// spacings of the coordinate crossing it (for Ctx::tie_mat).
inline float material_edge_margin(const Ctx& ctx, float zq, float rq) {
  if (!ctx.rl_grid) return 1e30f;
  auto val = [&](int zb, int rb) -> float {
    return (zb >= 0 && zb < kNBinsZME && rb >= 0 && rb < kNBinsRME) ? ctx.rl_grid[zb * kNBinsRME + rb] : 0.f;
  };
  const float zf = std::abs(zq), rf = rq;
  const int zb = zbin_me(zq), rb = rbin_me(rq);
  float m = 1e30f;
  const float xz = zf * kNBinsZME / kRangeZME, xr = rf * kNBinsRME / kRangeRME;
  const int ez = (int)std::lround(xz), er = (int)std::lround(xr);
  if (val(ez - 1, rb) != val(ez, rb))
    m = std::min(m, std::abs(xz - ez) / ((std::nextafter(zf, 2 * zf + 1) - zf) * kNBinsZME / kRangeZME));
  if (val(zb, er - 1) != val(zb, er))
    m = std::min(m, std::abs(xr - er) / ((std::nextafter(rf, 2 * rf + 1) - rf) * kNBinsRME / kRangeRME));
  return m;
}

// ---- propagateHelixToRMPlex (PropagationMPlex.cc:408-505) -------------------
template <class T>
// matRl / matXi, when given, replace the grid lookup with per-lane material
// (the per-layer x/X0 of the alignment model).
inline void propagate_helix_to_r(const T* inErr, const T* inPar, const int* inChg, const T* msRad, T* outErr,
                                 T* outPar, int N_proc, int pflags, const int* noMatEff, int* failOut,
                                 const Ctx& ctx, const T* matRl = nullptr, const T* matXi = nullptr) {
  std::memcpy(outErr, inErr, sizeof(T) * 21 * NN);
  std::memcpy(outPar, inPar, sizeof(T) * 6 * NN);
  alignas(64) T errorProp[36 * NN];
  alignas(64) int failFlag[NN];
  helix_at_r(inPar, inChg, msRad, outPar, errorProp, failFlag, pflags, ctx);

  if (pflags & PF_apply_material) {
    alignas(64) T hitsRl[NN], hitsXi[NN], propSign[NN];
    for (int n = 0; n < NN; ++n) {
      hitsRl[n] = 0; hitsXi[n] = 0; propSign[n] = 1;
      if (n >= N_proc) continue;
      if (!(failFlag[n] || (noMatEff && noMatEff[n]))) {
        if (matRl) {
          hitsRl[n] = matRl[n];
          hitsXi[n] = matXi[n];
        } else {
          lookup_material(ctx, outPar[2 * NN + n], msRad[n], hitsRl[n], hitsXi[n]);
          if (ctx.tie_mat)
            ctx.tie_mat[n] = std::min(ctx.tie_mat[n], material_edge_margin(ctx, float(outPar[2 * NN + n]), float(msRad[n])));
        }
      }
      const T r0 = hipo(inPar[0 * NN + n], inPar[1 * NN + n]);
      const T r = msRad[n];
      propSign[n] = (r > r0 ? T(1) : T(-1));
    }
    apply_material_effects(hitsRl, hitsXi, propSign, outErr, outPar, true);
  }

  squash_phi(outPar);

  alignas(64) T temp[36 * NN];
  mult_helix_prop<T, kJPatBarrel>(errorProp, outErr, temp);
  mult_helix_prop_transp<T, kJPatBarrel>(errorProp, temp, outErr);

  for (int n = 0; n < N_proc; ++n) {
    if (failFlag[n]) {
      for (int i = 0; i < 6; ++i) outPar[i * NN + n] = inPar[i * NN + n];
      for (int i = 0; i < 21; ++i) outErr[i * NN + n] = inErr[i * NN + n];
    }
  }
  if (failOut)
    for (int n = 0; n < NN; ++n) failOut[n] = failFlag[n];
}

// ---- propagateHelixToZMPlex (PropagationMPlex.cc:510-601) -------------------
template <class T>
inline void propagate_helix_to_z(const T* inErr, const T* inPar, const int* inChg, const T* msZ, T* outErr,
                                 T* outPar, int N_proc, int pflags, const int* noMatEff, const Ctx& ctx) {
  std::memcpy(outErr, inErr, sizeof(T) * 21 * NN);
  std::memcpy(outPar, inPar, sizeof(T) * 6 * NN);
  alignas(64) T errorProp[36 * NN];
  helix_at_z(inPar, inChg, msZ, outPar, errorProp, pflags, ctx);

  if (pflags & PF_apply_material) {
    alignas(64) T hitsRl[NN], hitsXi[NN], propSign[NN];
    for (int n = 0; n < NN; ++n) {
      hitsRl[n] = 0; hitsXi[n] = 0; propSign[n] = 1;
      if (n >= N_proc) continue;
      if (!(noMatEff && noMatEff[n])) {
        const T rq = T(std::hypot(outPar[0 * NN + n], outPar[1 * NN + n]));
        lookup_material(ctx, msZ[n], rq, hitsRl[n], hitsXi[n]);
        if (ctx.tie_mat) ctx.tie_mat[n] = std::min(ctx.tie_mat[n], material_edge_margin(ctx, float(msZ[n]), float(rq)));
      }
      const T zout = msZ[n];
      const T zin = inPar[2 * NN + n];
      propSign[n] = (std::abs(zout) > std::abs(zin) ? T(1) : T(-1));
    }
    apply_material_effects(hitsRl, hitsXi, propSign, outErr, outPar, false);
  }

  squash_phi(outPar);

  alignas(64) T temp[36 * NN];
  mult_helix_prop<T, kJPatEndcap>(errorProp, outErr, temp);
  mult_helix_prop_transp<T, kJPatEndcap>(errorProp, temp, outErr);
}

// ---- 2x2 symmetric Cramer inverse, determinant in double (MatriplexSym.h:427-444)
template <class T> inline void invert_cramer_sym2(T& a0, T& a1, T& a2) {
  const double det = (double)a0 * a2 - (double)a1 * a1;
  const T s = T(1 / det);
  const T tmp = s * a2;
  a1 *= -s;
  a2 = s * a0;
  a0 = tmp;
}

// ---- kalmanOperation, barrel (KalmanUtilsMPlex.cc:548-681) ------------------
template <class T>
inline void kalman_operation(int kfOp, const T* psErr, const T* psPar, const T* msErr, const T* msPar, T* outErr,
                             T* outPar, T* outChi2) {
#pragma omp simd
  for (int n = 0; n < NN; ++n) {
    auto C = [&](int i, int j) { return psErr[sym(i, j) * NN + n]; };
    const T mx = msPar[0 * NN + n], my = msPar[1 * NN + n], mz = msPar[2 * NN + n];
    const T px = psPar[0 * NN + n], py = psPar[1 * NN + n], pz = psPar[2 * NN + n];
    const T r = std::hypot(mx, my);
    const T a00 = -(my + py) / (2 * r);                 // rotT00
    const T a01 = (mx + px) / (2 * r);                  // rotT01

    const T rg0 = mx - px, rg1 = my - py, rg2 = mz - pz;              // SubtractFirst3
    T b[6];                                                             // AddIntoUpperLeft3x3
    for (int i = 0; i < 6; ++i) b[i] = psErr[i * NN + n] + msErr[i * NN + n];

    const T l0 = a00 * rg0 + a01 * rg1;                                 // RotateResidulsOnTangentPlane
    const T l1 = rg2;

    T c[9];                                                             // ProjectResErr
    c[0] = a00 * b[0] + a01 * b[1];
    c[1] = a00 * b[1] + a01 * b[2];
    c[2] = a00 * b[3] + a01 * b[4];
    c[3] = b[3];
    c[4] = b[4];
    c[5] = b[5];
    c[6] = a01 * b[0] - a00 * b[1];
    c[7] = a01 * b[1] - a00 * b[2];
    c[8] = a01 * b[3] - a00 * b[4];
    T s0 = c[0] * a00 + c[1] * a01;                                     // ProjectResErrTransp
    T s1 = c[3] * a00 + c[4] * a01;
    T s2 = c[5];

    invert_cramer_sym2(s0, s1, s2);

    if (kfOp & KFO_Calculate_Chi2) {                                    // Chi2Similarity
      outChi2[n] = s0 * l0 * l0 + s2 * l1 * l1 + 2 * (s1 * l1 * l0);
    }

    if (kfOp & KFO_Update_Params) {
      T h[9];                                                           // KalmanHTG
      h[0] = a00 * s0; h[1] = a00 * s1; h[2] = 0;
      h[3] = a01 * s0; h[4] = a01 * s1; h[5] = 0;
      h[6] = s1;       h[7] = s2;       h[8] = 0;
      T K[6][2];                                                        // KalmanGain (6x3, third column 0)
      for (int i = 0; i < 6; ++i) {
        K[i][0] = C(i, 0) * h[0] + C(i, 1) * h[3] + C(i, 2) * h[6];
        K[i][1] = C(i, 0) * h[1] + C(i, 1) * h[4] + C(i, 2) * h[7];
      }
      for (int i = 0; i < 6; ++i)                                       // MultResidualsAdd
        outPar[i * NN + n] = psPar[i * NN + n] + K[i][0] * l0 + K[i][1] * l1;
      if (outPar[4 * NN + n] >= T(kPI)) outPar[4 * NN + n] -= T(kTwoPI);   // squashPhiMPlex
      if (outPar[4 * NN + n] < -T(kPI)) outPar[4 * NN + n] += T(kTwoPI);
      T KH[6][3];                                                       // KHMult
      for (int i = 0; i < 6; ++i) {
        KH[i][0] = K[i][0] * a00;
        KH[i][1] = K[i][0] * a01;
        KH[i][2] = K[i][1];
      }
      for (int i = 0; i < 6; ++i)                                       // KHC.ah, then Subtract
        for (int j = 0; j <= i; ++j) {
          const T t = KH[i][0] * C(0, j) + KH[i][1] * C(1, j) + KH[i][2] * C(2, j);
          outErr[sym(i, j) * NN + n] = C(i, j) - t;
        }
    }
  }
}

// ---- kalmanOperationEndcap (KalmanUtilsMPlex.cc:774-888) --------------------
template <class T>
inline void kalman_operation_endcap(int kfOp, const T* psErr, const T* psPar, const T* msErr, const T* msPar,
                                    T* outErr, T* outPar, T* outChi2) {
#pragma omp simd
  for (int n = 0; n < NN; ++n) {
    auto C = [&](int i, int j) { return psErr[sym(i, j) * NN + n]; };
    const T r0 = msPar[0 * NN + n] - psPar[0 * NN + n];               // SubtractFirst2
    const T r1 = msPar[1 * NN + n] - psPar[1 * NN + n];
    T s0 = psErr[0 * NN + n] + msErr[0 * NN + n];                      // AddIntoUpperLeft2x2
    T s1 = psErr[1 * NN + n] + msErr[1 * NN + n];
    T s2 = psErr[2 * NN + n] + msErr[2 * NN + n];

    invert_cramer_sym2(s0, s1, s2);

    if (kfOp & KFO_Calculate_Chi2) outChi2[n] = s0 * r0 * r0 + s2 * r1 * r1 + 2 * (s1 * r1 * r0);

    if (kfOp & KFO_Update_Params) {
      T K[6][2];                                                        // KalmanGain62.ah
      for (int i = 0; i < 6; ++i) {
        K[i][0] = C(i, 0) * s0 + C(i, 1) * s1;
        K[i][1] = C(i, 0) * s1 + C(i, 1) * s2;
      }
      for (int i = 0; i < 6; ++i)
        outPar[i * NN + n] = psPar[i * NN + n] + K[i][0] * r0 + K[i][1] * r1;
      if (outPar[4 * NN + n] >= T(kPI)) outPar[4 * NN + n] -= T(kTwoPI);
      if (outPar[4 * NN + n] < -T(kPI)) outPar[4 * NN + n] += T(kTwoPI);
      for (int i = 0; i < 6; ++i)                                       // K62HC.ah, then Subtract
        for (int j = 0; j <= i; ++j) {
          const T t = K[i][0] * C(0, j) + K[i][1] * C(1, j);
          outErr[sym(i, j) * NN + n] = C(i, j) - t;
        }
    }
  }
}

}  // namespace orc
