// mkf_track_math.h -- per-lane helix propagation and Kalman update for gfx950.
//
// One wavefront lane owns one track.  The state (6 parameters, 21 packed
// covariance entries, charge) lives in VGPRs; the propagation Jacobian is kept
// only in its non-structural entries (17 barrel, 15 endcap) and the similarity
// J*C*J^T is evaluated row by row so no 6x6 temporary is ever materialised.
// All arithmetic is FP32 VALU (the 6x6 products are far too small for MFMA);
// the only FP64 is the 2x2 determinant, as in the reference.
//
// The arithmetic follows the reference's expressions so results agree to FP32
// rounding.  Where a step is regrouped to save instructions (the sine and
// cosine pairs, the material terms, the dDdx / dDdy increments of the helix
// iteration), the comment at the step gives the algebra;
// the cancelling parts (the dDdipt / dDdphi increments, the Jacobian's small
// differences, the Kalman gain and covariance update) keep the reference's
// term order.  Sources:
//   helix step      mkFit/PropagationMPlex.icc:6-246
//   to-R driver     mkFit/PropagationMPlex.cc:408-505
//   helixAtZ        mkFit/PropagationMPlex.cc:604-712, driver :510-601
//   material        mkFit/PropagationMPlex.cc:716-767, lookup :446-468, :543-565
//   similarity      term order of the GenMul products (Matriplex/GenMul.pm:482-533)
//                   over the J patterns of mkFit/GenMPlexOps.pl:254-324
//   Kalman update   mkFit/KalmanUtilsMPlex.cc:548-681 (barrel), :774-888 (endcap)
//   2x2 inverse     Matriplex/MatriplexSym.h:427-444
#pragma once

#include <hip/hip_runtime.h>

namespace mkf {

constexpr float kPI = 3.14159265358979323846;   // Config.h:77
constexpr float kTwoPI = 6.28318530717958647692;
constexpr float kSol = 0.299792458;             // Config.h:85
constexpr int kNiter = 5;                       // Config.h:230
constexpr float kMagC1 = 3.8114, kMagB0 = -3.94991e-06, kMagB1 = 7.53701e-06, kMagA = 2.43878e-11;
constexpr int kNBinsZ = 300, kNBinsR = 120;     // Config.h:123-126
constexpr float kRangeZ = 300., kRangeR = 120.;

constexpr int PF_PARAM_B = 0x1;
constexpr int PF_MATERIAL = 0x2;

struct Material {
  const float* __restrict__ rl;  // [300][120] or nullptr
  const float* __restrict__ xi;
};

__host__ __device__ constexpr int S(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// FP32 divide, reciprocal and square root as hardware approximations refined by
// one FMA residual step: for operands in the normal range (all the fit's
// operands) the results are the correctly rounded ones or within one ulp, at
// 4 instructions instead of the ~10 of the IEEE sequences with full
// special-case handling.
__device__ __forceinline__ float frcp(float b) {
  float r = __builtin_amdgcn_rcpf(b);
  return __builtin_fmaf(__builtin_fmaf(-b, r, 1.f), r, r);
}
__device__ __forceinline__ float fdiv(float a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  const float q = a * r;
  return __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
}
__device__ __forceinline__ float fsqrt(float x) {
  const float y = __builtin_amdgcn_sqrtf(x);
  return __builtin_fmaf(__builtin_fmaf(-y, y, x), 0.5f * __builtin_amdgcn_rcpf(y), y);
}

__device__ __forceinline__ float hipo(float x, float y) { return fsqrt(x * x + y * y); }
// std::hypot of the reference (MkBase.h:55, KalmanUtilsMPlex.cc:585): the sum of
// squares with one rounding, then the refined square root.
__device__ __forceinline__ float fhypot(float x, float y) { return fsqrt(__builtin_fmaf(x, x, y * y)); }

// ln x as the hardware base-2 logarithm (v_log_f32, 1 ulp) times ln 2, for
// positive normal arguments away from 1 (the material terms' arguments), where
// it stays within a few ulp of the correctly rounded value.
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

__device__ __forceinline__ void sincos4(float x, float& s, float& c) {   // Matrix.h:48-56
  const float x2 = x * x;
  c = 1.f - 0.5f * x2 + 0.04166667f * x2 * x2;
  s = x - 0.16666667f * x * x2;
}

__device__ __forceinline__ float bfield_from_zr(float z, float r) {   // Config.h:418-421
  return (kMagB0 * z * z + kMagB1 * z + kMagC1) * (kMagA * r * r + 1.f);
}

__device__ __forceinline__ void squash_phi(float& phi) {   // PropagationMPlex.h:8-15
  if (phi >= kPI) phi -= kTwoPI;
  if (phi < -kPI) phi += kTwoPI;
}

// Barrel Jacobian, non-structural entries (rows 3 and 5 are unit rows, (2,2) = 1).
struct JacR {
  float j00, j01, j03, j04;
  float j10, j11, j13, j14;
  float j20, j21, j23, j24, j25;
  float j40, j41, j43, j44;
  float isin_t, cos_t;   // 1 / sin(theta) and cos(theta), reused by the material step
};

// Endcap Jacobian (PropagationMPlex.cc:618-624 initial pattern; rows 2 and 3 structural).
struct JacZ {
  float j02, j03, j04, j05;
  float j12, j13, j14, j15;
  float j42, j43, j45;
  float isin_t, cos_t;   // 1 / sin(theta) and cos(theta), reused by the material step
};

// ---- helix to radius r ------------------------------------------------------
// par in/out (out only meaningful when the return value is false = no failure).
//
// The iteration and the Jacobian are the reference's (PropagationMPlex.icc:6-246)
// with some products regrouped.  Per iteration, with (pxn, pyn) the rotated
// momentum that becomes the next (pxin, pyin), g = x pxn + y pyn and
// w = ipt / r0 (using k * kinv = 1): k dadx = -x w and k dady = -y w, so the
// dDdx / dDdy increments are x (1 - w g) / r0 and y (1 - w g) / r0.  The dDdipt
// and dDdphi brackets, and the small differences in the Jacobian, keep the
// reference's term order: their terms nearly cancel once the step has
// converged, and the order fixes the rounding.
// trig, when given: {sin, cos} of the input phi and theta, as sincosf gives them.
__device__ __forceinline__ bool helix_at_r(float (&par)[6], int chg, float r, float bfield, bool param_b, JacR& J,
                                           const float* trig = nullptr) {
  const float xin = par[0], yin = par[1], zin = par[2];
  const float ipt = par[3], phiin = par[4], theta = par[5];
  float r0 = hipo(xin, yin);
  const float bf = param_b ? bfield_from_zr(zin, r0) : bfield;
  const float k = fdiv(float(chg) * 100.f, -kSol * bf);
  const float kinv = frcp(k);
  const float pt = frcp(ipt);
  const float iptkinv = ipt * kinv;

  float D = 0.f, cosa = 0.f, sina = 0.f, cosah = 0.f, sinah = 0.f, id = 0.f;
  float cosP, sinP;
  if (trig) {
    sinP = trig[0];
    cosP = trig[1];
  } else {
    sincosf(phiin, &sinP, &cosP);
  }
  float pxin = cosP * pt;
  float pyin = sinP * pt;

  float dDdx = r0 > 0.f ? fdiv(-xin, r0) : 0.f;
  float dDdy = r0 > 0.f ? fdiv(-yin, r0) : 0.f;
  float dDdipt = 0.f;
  float dDdphi = 0.f;
  bool fail = false;
  float x = xin, y = yin;

#pragma unroll
  for (int i = 0; i < kNiter; ++i) {
    r0 = hipo(x, y);
    const float oodotp = fdiv(r0 * pt, pxin * x + pyin * y);
    if (oodotp > 5.0f || oodotp < 0) {
      id = 0.0f;
      fail = true;
    } else {
      id = (r - r0) * oodotp;
    }
    D += id;

    sincos4(id * ipt * kinv * 0.5f, sinah, cosah);
    cosa = 1.f - 2.f * sinah * sinah;
    sina = 2.f * sinah * cosah;
    const float pxn = pxin * cosa - pyin * sina;
    const float pyn = pyin * cosa + pxin * sina;

    const float oor0 = (r0 > 0.f && fabsf(r - r0) < 0.0001f) ? frcp(r0) : 0.f;
    // Every derivative term is scaled by oor0: with oor0 == 0 the sums keep
    // their values, so the block is skipped (a whole wavefront skips it on the
    // first iterations, before the step has converged to within 1 um).
    if (i + 1 != kNiter && oor0 != 0.f) {
      const float g = x * pxn + y * pyn;
      const float q = oor0 * (1.f - ipt * oor0 * g);
      dDdx -= x * q;
      dDdy -= y * q;
      const float pxca = pxin * cosa;
      const float pxsa = pxin * sina;
      const float pyca = pyin * cosa;
      const float pysa = pyin * sina;
      const float tmp = id * kinv * ipt;
      dDdipt -= k * (x * (pxca * tmp - pysa * tmp - pyca - pxsa + pyin) +
                     y * (pyca * tmp + pxsa * tmp - pysa + pxca - pxin)) * pt * oor0;
      dDdphi += k * (x * (pysa - pxin + pxca) - y * (pxsa - pyin + pyca)) * oor0;
    }

    x = x + 2.f * k * sinah * (pxin * cosah - pyin * sinah);
    y = y + 2.f * k * sinah * (pyin * cosah + pxin * sinah);
    pxin = pxn;
    pyin = pyn;
  }

  const float alpha = D * ipt * kinv;
  const float dadx = dDdx * ipt * kinv;
  const float dady = dDdy * ipt * kinv;
  const float dadipt = (ipt * dDdipt + D) * kinv;
  const float dadphi = dDdphi * ipt * kinv;

  sincos4(alpha, sina, cosa);
  // (cosP cosa - sinP sina) and (sinP cosa + cosP sina) are cos / sin(phi + alpha);
  // the small differences (u cosa - sina), (1 - cosa) keep the reference's form,
  // which forms them from alpha-sized quantities.
  const float cc = cosP * cosa - sinP * sina;
  const float ss = sinP * cosa + cosP * sina;
  const float kp = k * pt, kpp = kp * pt;
  const float u = ipt * dadipt;
  const float ucs = u * cosa - sina, omc = 1.f - cosa, us = u * sina;

  J.j00 = 1.f + k * dadx * cc * pt;
  J.j01 = k * dady * cc * pt;
  J.j03 = kpp * (cosP * ucs + sinP * (omc - us));
  J.j04 = kp * (cosP * dadphi * cosa - sinP * dadphi * sina - sinP * sina + cosP * cosa - cosP);
  J.j10 = k * dadx * ss * pt;
  J.j11 = 1.f + k * dady * ss * pt;
  J.j13 = kpp * (sinP * ucs + cosP * (us - omc));
  J.j14 = kp * (sinP * dadphi * cosa + cosP * dadphi * sina + sinP * cosa + cosP * sina - sinP);

  float cosT, sinT;
  if (trig) {
    sinT = trig[2];
    cosT = trig[3];
  } else {
    sincosf(theta, &sinT, &cosT);
  }
  const float isinT = frcp(sinT);
  J.isin_t = isinT;
  J.cos_t = cosT;
  const float kc = kp * cosT * isinT;   // k pt cot(theta)

  par[0] = x;
  par[1] = y;
  par[2] = zin + k * alpha * cosT * pt * isinT;
  J.j20 = kc * dadx;
  J.j21 = kc * dady;
  J.j23 = kc * pt * (u - alpha);
  J.j24 = kc * dadphi;
  J.j25 = -kp * alpha * isinT * isinT;
  par[4] = phiin + alpha;
  J.j40 = dadx;
  J.j41 = dady;
  J.j43 = dadipt;
  J.j44 = 1.f + dadphi;
  return fail;
}

// ---- helix to plane z (helixAtZ) --------------------------------------------
__device__ __forceinline__ void helix_at_z(float (&par)[6], int chg, float zout, float bfield, bool param_b, JacZ& J,
                                           const float* trig = nullptr) {
  const float zin = par[2], ipt = par[3], phiin = par[4], theta = par[5];
  const float bf = param_b ? bfield_from_zr(zin, hipo(par[0], par[1])) : bfield;
  const float k = fdiv(float(chg) * 100.f, -kSol * bf);
  const float kinv = frcp(k);
  const float pt = frcp(ipt);
  float cosP, sinP, cosT, sinT;
  if (trig) {
    sinP = trig[0];
    cosP = trig[1];
    sinT = trig[2];
    cosT = trig[3];
  } else {
    sincosf(phiin, &sinP, &cosP);
    sincosf(theta, &sinT, &cosT);
  }
  J.isin_t = frcp(sinT);
  J.cos_t = cosT;
  const float tanT = fdiv(sinT, cosT);
  const float icos2T = frcp(cosT * cosT);
  const float pxin = cosP * pt;
  const float pyin = sinP * pt;
  const float deltaZ = zout - zin;
  const float alpha = deltaZ * tanT * ipt * kinv;
  float sinah, cosah;
  sincos4(alpha * 0.5f, sinah, cosah);
  const float cosa = 1.f - 2.f * sinah * sinah;
  const float sina = 2.f * sinah * cosah;

  par[0] = par[0] + 2.f * k * sinah * (pxin * cosah - pyin * sinah);
  par[1] = par[1] + 2.f * k * sinah * (pyin * cosah + pxin * sinah);
  par[2] = zout;
  par[4] = phiin + alpha;

  const float pxcaMpysa = pxin * cosa - pyin * sina;
  J.j02 = -tanT * ipt * pxcaMpysa;
  J.j03 = k * pt * pt * (cosP * (alpha * cosa - sina) + sinP * 2.f * sinah * (sinah - alpha * cosah));
  J.j04 = -2 * k * pt * sinah * (sinP * cosah + cosP * sinah);
  J.j05 = deltaZ * ipt * pxcaMpysa * icos2T;
  const float pycaPpxsa = pyin * cosa + pxin * sina;
  J.j12 = -tanT * ipt * pycaPpxsa;
  J.j13 = k * pt * pt * (sinP * (alpha * cosa - sina) - cosP * 2.f * sinah * (sinah - alpha * cosah));
  J.j14 = 2 * k * pt * sinah * (cosP * cosah - sinP * sinah);
  J.j15 = deltaZ * ipt * pycaPpxsa * icos2T;
  J.j42 = -ipt * tanT * kinv;
  J.j43 = tanT * deltaZ * kinv;
  J.j45 = ipt * deltaZ * kinv * icos2T;
}

// ---- applyMaterialEffects, one lane -----------------------------------------
// The reference's formulas (PropagationMPlex.cc:716-767) with its ten divisions
// and three logarithms regrouped around three reciprocals and two logarithms:
//   p = pt / sin(theta);  E^2 = p^2 + m^2;  beta^2 = p^2 / E^2;  gamma^2 = E^2 / m^2
//   thetaMSC = 0.0136 (1 + 0.038 ln radL) / (beta p)
//   wmax = 2 me beta^2 gamma^2 / (1 + 2 gamma me / m + me^2 / m^2)
//   0.5 ln(2 me beta^2 gamma^2 wmax / I^2) - deltahalf
//     = 0.5 ln(wmax) + 0.5 ln(2 me) - ln(28.816e-9 sqrt(2.33 * 0.498)) + 0.5
// since deltahalf carries ln(beta gamma) = 0.5 ln(beta^2 gamma^2) and ln(I / I).  The terms
// agree with the reference's to a few FP32 ulp.  sin_t / cos_t are the sine and
// cosine of par[5], which the helix step has already formed; isin_t = 1 / sin_t.
__device__ __forceinline__ void material_effects(float radL, float xi, float propSign, float (&par)[6],
                                                 float (&err)[21], bool is_barrel, float isin_t, float cos_t) {
  if (radL < 1e-13f) return;
  constexpr float mpi = 0.140;
  constexpr float mpi2 = mpi * mpi;
  constexpr float me = 0.0005;
  constexpr float I = 16.0e-9 * 10.75;
  const float pt = frcp(par[3]);
  const float p = pt * isin_t;
  const float p2 = p * p;
  const float e2 = p2 + mpi2;
  const float beta2 = p2 * frcp(e2);
  const float beta = fsqrt(beta2);
  const float invCos = is_barrel ? isin_t : frcp(fabsf(cos_t));
  radL = radL * invCos;
  if (radL < 1e-13f) return;
  const float thetaMSC = 0.0136f * (1.f + 0.038f * flog(radL)) * frcp(beta * p);
  const float thetaMSC2 = thetaMSC * thetaMSC * radL;
  err[S(4, 4)] += thetaMSC2;
  err[S(5, 5)] += thetaMSC2;
  const float gamma = fsqrt(e2 * (1.f / mpi2));
  const float wmax = (2.f * me / mpi2) * p2 * frcp(1.f + gamma * (2.f * me / mpi) + me * me / (mpi * mpi));
  // 0.5 ln(2 me) - ln(28.816e-9 sqrt(2.33 * 0.498)) + 0.5 (the ln I terms cancel)
  constexpr float kLogTerm = 14.3341009f;
  const float dEdx = beta < 1.f ? 2.f * xi * invCos * (0.5f * flog(wmax) + kLogTerm - beta2) * frcp(beta2) : 0.f;
  const float dP = propSign * dEdx * frcp(beta);
  par[3] = fdiv(p, fmaxf(p + dP, 0.001f) * pt);
  err[S(3, 3)] += dP * dP * frcp(p2 * pt * pt);
}

__device__ __forceinline__ void lookup_material(const Material& m, float zq, float rq, float& rl, float& xi) {
  const int zb = (int)fdiv(fabsf(zq) * kNBinsZ, kRangeZ);   // MaterialEffects.h:8-9
  const int rb = (int)fdiv(rq * kNBinsR, kRangeR);
  if (m.rl && zb >= 0 && zb < kNBinsZ && rb >= 0 && rb < kNBinsR) {
    rl = m.rl[zb * kNBinsR + rb];
    xi = m.xi[zb * kNBinsR + rb];
  } else {
    rl = 0.f;
    xi = 0.f;
  }
}

// ---- similarity C' = J C J^T, barrel pattern --------------------------------
__device__ __forceinline__ void similarity_r(const JacR& J, float (&C)[21]) {
  float out[21];
  float t[6];
  // row i of temp = J(i,:) * C, then out(i, j<=i) = sum_k temp(i,k) J(j,k)
  auto out_row = [&](int i) {
    // j = 0, 1, 4: k in {0,1,3,4};  j = 2: k in {0,1,2(one),3,4,5};  j = 3: temp(i,3);  j = 5: temp(i,5)
    if (0 <= i) out[S(i, 0)] = t[0] * J.j00 + t[1] * J.j01 + t[3] * J.j03 + t[4] * J.j04;
    if (1 <= i) out[S(i, 1)] = t[0] * J.j10 + t[1] * J.j11 + t[3] * J.j13 + t[4] * J.j14;
    if (2 <= i) out[S(i, 2)] = t[0] * J.j20 + t[1] * J.j21 + t[2] + t[3] * J.j23 + t[4] * J.j24 + t[5] * J.j25;
    if (3 <= i) out[S(i, 3)] = t[3];
    if (4 <= i) out[S(i, 4)] = t[0] * J.j40 + t[1] * J.j41 + t[3] * J.j43 + t[4] * J.j44;
    if (5 <= i) out[S(i, 5)] = t[5];
  };
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if (i == 3 || i == 5) {
#pragma unroll
      for (int j = 0; j < 6; ++j) t[j] = C[S(i, j)];
    } else {
      float a0, a1, a3, a4;
      if (i == 0) { a0 = J.j00; a1 = J.j01; a3 = J.j03; a4 = J.j04; }
      else if (i == 1) { a0 = J.j10; a1 = J.j11; a3 = J.j13; a4 = J.j14; }
      else if (i == 2) { a0 = J.j20; a1 = J.j21; a3 = J.j23; a4 = J.j24; }
      else { a0 = J.j40; a1 = J.j41; a3 = J.j43; a4 = J.j44; }
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        if (i == 2)
          t[j] = a0 * C[S(0, j)] + a1 * C[S(1, j)] + C[S(2, j)] + a3 * C[S(3, j)] + a4 * C[S(4, j)] + J.j25 * C[S(5, j)];
        else
          t[j] = a0 * C[S(0, j)] + a1 * C[S(1, j)] + a3 * C[S(3, j)] + a4 * C[S(4, j)];
      }
    }
    out_row(i);
  }
#pragma unroll
  for (int e = 0; e < 21; ++e) C[e] = out[e];
}

// ---- similarity, endcap pattern ---------------------------------------------
// J rows: 0 = (1,0,j02,j03,j04,j05), 1 = (0,1,j12,j13,j14,j15), 2 = 0,
//         3 = e3, 4 = (0,0,j42,j43,1,j45), 5 = e5.
__device__ __forceinline__ void similarity_z(const JacZ& J, float (&C)[21]) {
  float out[21];
  float t[6];
  auto out_row = [&](int i) {
    if (0 <= i) out[S(i, 0)] = t[0] + t[2] * J.j02 + t[3] * J.j03 + t[4] * J.j04 + t[5] * J.j05;
    if (1 <= i) out[S(i, 1)] = t[1] + t[2] * J.j12 + t[3] * J.j13 + t[4] * J.j14 + t[5] * J.j15;
    if (2 <= i) out[S(i, 2)] = 0.f;
    if (3 <= i) out[S(i, 3)] = t[3];
    if (4 <= i) out[S(i, 4)] = t[2] * J.j42 + t[3] * J.j43 + t[4] + t[5] * J.j45;
    if (5 <= i) out[S(i, 5)] = t[5];
  };
#pragma unroll
  for (int i = 0; i < 6; ++i) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      if (i == 0) t[j] = C[S(0, j)] + J.j02 * C[S(2, j)] + J.j03 * C[S(3, j)] + J.j04 * C[S(4, j)] + J.j05 * C[S(5, j)];
      else if (i == 1) t[j] = C[S(1, j)] + J.j12 * C[S(2, j)] + J.j13 * C[S(3, j)] + J.j14 * C[S(4, j)] + J.j15 * C[S(5, j)];
      else if (i == 2) t[j] = 0.f;
      else if (i == 3) t[j] = C[S(3, j)];
      else if (i == 4) t[j] = J.j42 * C[S(2, j)] + J.j43 * C[S(3, j)] + C[S(4, j)] + J.j45 * C[S(5, j)];
      else t[j] = C[S(5, j)];
    }
    out_row(i);
  }
#pragma unroll
  for (int e = 0; e < 21; ++e) C[e] = out[e];
}

// ---- propagateHelixToRMPlex, one lane ---------------------------------------
// On failure the state is left as it was (the reference resets output to input).
__device__ __forceinline__ bool propagate_to_r(float (&par)[6], float (&err)[21], int chg, float r, int pflags,
                                               bool no_mat, float bfield, const Material& mat,
                                               const float* trig = nullptr) {
  float out[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) out[i] = par[i];
  JacR J;
  const bool fail = helix_at_r(out, chg, r, bfield, pflags & PF_PARAM_B, J, trig);
  if (fail) return true;
  if (pflags & PF_MATERIAL) {
    float rl = 0.f, xi = 0.f;
    if (!no_mat) lookup_material(mat, out[2], r, rl, xi);
    const float r0 = hipo(par[0], par[1]);
    const float propSign = r > r0 ? 1.f : -1.f;
    material_effects(rl, xi, propSign, out, err, true, J.isin_t, J.cos_t);
  }
  squash_phi(out[4]);
  similarity_r(J, err);
#pragma unroll
  for (int i = 0; i < 6; ++i) par[i] = out[i];
  return false;
}

// ---- propagateHelixToZMPlex, one lane ---------------------------------------
__device__ __forceinline__ void propagate_to_z(float (&par)[6], float (&err)[21], int chg, float zout, int pflags,
                                               bool no_mat, float bfield, const Material& mat,
                                               const float* trig = nullptr) {
  const float zin = par[2];
  JacZ J;
  helix_at_z(par, chg, zout, bfield, pflags & PF_PARAM_B, J, trig);
  if (pflags & PF_MATERIAL) {
    float rl = 0.f, xi = 0.f;
    if (!no_mat) lookup_material(mat, zout, fhypot(par[0], par[1]), rl, xi);
    const float propSign = fabsf(zout) > fabsf(zin) ? 1.f : -1.f;
    material_effects(rl, xi, propSign, par, err, false, J.isin_t, J.cos_t);
  }
  squash_phi(par[4]);
  similarity_z(J, err);
}

// ---- 2x2 inverse with a double determinant ----------------------------------
// As the reference (MatriplexSym.h:427-444): the products of two floats are
// exact in double, so a0 a2 - a1^2 is rounded once; a strong update can cancel
// it to a few ulp of its terms, and this keeps the reference's value there.
__device__ __forceinline__ void invert_sym2(float& a0, float& a1, float& a2) {
  const double det = (double)a0 * a2 - (double)a1 * a1;
  // 1/det to well below float resolution: hardware FP64 reciprocal refined by
  // one Newton step, then rounded to float as the reference's TT(1)/det is.
  double rd = __builtin_amdgcn_rcp(det);
  rd = __builtin_fma(__builtin_fma(-det, rd, 1.0), rd, rd);
  const float s = (float)rd;
  const float tmp = s * a2;
  a1 *= -s;
  a2 = s * a0;
  a0 = tmp;
}

// ---- kalmanOperation (barrel), one lane --------------------------------------
// ps state in par/err is replaced by the updated state when `update`.  The
// gain and the covariance update keep the reference's products and their order
// (KalmanHTG, KalmanGain, KHMult, KHC): a strong update cancels most of C, and
// the order fixes the rounding of what remains.
template <bool CHI2, bool UPDATE>
__device__ __forceinline__ float kalman_barrel(float (&par)[6], float (&err)[21], const float (&mp)[3],
                                               const float (&me)[6]) {
  const float r = fhypot(mp[0], mp[1]);
  const float a00 = fdiv(-(mp[1] + par[1]), 2 * r);
  const float a01 = fdiv(mp[0] + par[0], 2 * r);
  const float rg0 = mp[0] - par[0], rg1 = mp[1] - par[1], rg2 = mp[2] - par[2];
  float b[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) b[i] = err[i] + me[i];
  const float l0 = a00 * rg0 + a01 * rg1;
  const float l1 = rg2;
  const float c0 = a00 * b[0] + a01 * b[1];
  const float c1 = a00 * b[1] + a01 * b[2];
  float s0 = c0 * a00 + c1 * a01;
  float s1 = b[3] * a00 + b[4] * a01;
  float s2 = b[5];
  invert_sym2(s0, s1, s2);
  float chi2 = 0.f;
  if (CHI2) chi2 = s0 * l0 * l0 + s2 * l1 * l1 + 2 * (s1 * l1 * l0);
  if (UPDATE) {
    const float h0 = a00 * s0, h1 = a00 * s1, h3 = a01 * s0, h4 = a01 * s1, h6 = s1, h7 = s2;
    float K0[6], K1[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      K0[i] = err[S(i, 0)] * h0 + err[S(i, 1)] * h3 + err[S(i, 2)] * h6;
      K1[i] = err[S(i, 0)] * h1 + err[S(i, 1)] * h4 + err[S(i, 2)] * h7;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) par[i] = par[i] + K0[i] * l0 + K1[i] * l1;
    squash_phi(par[4]);
    float out[21];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const float kh0 = K0[i] * a00, kh1 = K0[i] * a01, kh2 = K1[i];
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        const float t = kh0 * err[S(0, j)] + kh1 * err[S(1, j)] + kh2 * err[S(2, j)];
        out[S(i, j)] = err[S(i, j)] - t;
      }
    }
#pragma unroll
    for (int e = 0; e < 21; ++e) err[e] = out[e];
  }
  return chi2;
}

// ---- kalmanOperationEndcap, one lane ----------------------------------------
template <bool CHI2, bool UPDATE>
__device__ __forceinline__ float kalman_endcap(float (&par)[6], float (&err)[21], const float (&mp)[3],
                                               const float (&me)[6]) {
  const float r0 = mp[0] - par[0];
  const float r1 = mp[1] - par[1];
  float s0 = err[0] + me[0];
  float s1 = err[1] + me[1];
  float s2 = err[2] + me[2];
  invert_sym2(s0, s1, s2);
  float chi2 = 0.f;
  if (CHI2) chi2 = s0 * r0 * r0 + s2 * r1 * r1 + 2 * (s1 * r1 * r0);
  if (UPDATE) {
    float K0[6], K1[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      K0[i] = err[S(i, 0)] * s0 + err[S(i, 1)] * s1;
      K1[i] = err[S(i, 0)] * s1 + err[S(i, 1)] * s2;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) par[i] = par[i] + K0[i] * r0 + K1[i] * r1;
    squash_phi(par[4]);
    float out[21];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        const float t = K0[i] * err[S(0, j)] + K1[i] * err[S(1, j)];
        out[S(i, j)] = err[S(i, j)] - t;
      }
#pragma unroll
    for (int e = 0; e < 21; ++e) err[e] = out[e];
  }
  return chi2;
}

}  // namespace mkf
