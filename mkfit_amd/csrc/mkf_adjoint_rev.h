// mkf_adjoint_rev.h -- reverse-mode derivative of one fit step of the alignment
// model (mkf_adjoint_math.h), written out by hand.
//
// align_step_vjp recomputes the step's forward values (recording the helix
// iterations' entering position and momentum), then sweeps back through the
// step's stages -- Kalman update and chi2, similarity J C J^T, material,
// helix -- carrying adjoints, and returns
//     g = lambda_par^T d(par_out)/d(in) + lambda_err^T d(err_out)/d(in)
//         + d(chi2)/d(in),        in = (par_in, err_in, theta_l)
// in one pass: the cost of a few forward steps instead of one tangent pass per
// input direction.  The branches are the forward's (helix failure, the
// iteration's convergence test, material cut-offs, the momentum floor); the
// derivative of each stage is that of its exact formula evaluated at the
// forward's values, so in FP64 it equals the tangent derivative of
// mkf_adjoint_math.h to rounding.  Covariances are 21 packed entries, each an
// independent input, as in the tangent form.
#pragma once

#include "mkf_adjoint_math.h"

namespace mkf_adj {

// ---- Kalman update (barrel) and chi2: adjoint ----------------------------------
// In: the state before the update (par, err), the hit (mp, me) and the adjoints
// of the updated state (lpo, leo) and of chi2 (lc2).  Out: adjoints of par, err
// and mp (accumulated into lp, le, lmp).
template <class T>
__device__ __forceinline__ void kalman_rev(const T (&par)[6], const T (&err)[21], const T (&mp)[3],
                                           const float (&me)[6], const T (&lpo)[6], const T (&leo)[21], T lc2,
                                           T (&lp)[6], T (&le)[21], T (&lmp)[3]) {
  // forward values, as kalman_t
  const T rk = vsqrt(mp[0] * mp[0] + mp[1] * mp[1]);
  const T a00 = vdiv(-(mp[1] + par[1]), T(2.f) * rk);
  const T a01 = vdiv(mp[0] + par[0], T(2.f) * rk);
  const T rg0 = mp[0] - par[0], rg1 = mp[1] - par[1], rg2 = mp[2] - par[2];
  T b[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) b[i] = err[i] + me[i];
  const T l0 = a00 * rg0 + a01 * rg1;
  const T l1 = rg2;
  const T c0 = a00 * b[0] + a01 * b[1];
  const T c1 = a00 * b[1] + a01 * b[2];
  const T s0 = c0 * a00 + c1 * a01;
  const T s1 = b[3] * a00 + b[4] * a01;
  const T s2 = b[5];
  T g0 = s0, g1 = s1, g2 = s2;
  invert_sym2_t(g0, g1, g2);
  // s = 1 / det as the tangent form takes it (the float model's determinant in double)
  const T s = T(1.0 / ((double)s0 * (double)s2 - (double)s1 * (double)s1));
  const T h0 = a00 * g0, h1 = a00 * g1, h3 = a01 * g0, h4 = a01 * g1, h6 = g1, h7 = g2;
  T K0[6], K1[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    K0[i] = err[S(i, 0)] * h0 + err[S(i, 1)] * h3 + err[S(i, 2)] * h6;
    K1[i] = err[S(i, 0)] * h1 + err[S(i, 1)] * h4 + err[S(i, 2)] * h7;
  }
  // out_ij = C_ij - (kh0_i C_0j + kh1_i C_1j + kh2_i C_2j), kh0 = K0 a00, kh1 = K0 a01, kh2 = K1
  T lK0[6], lK1[6];
  T la00 = T(0), la01 = T(0);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    T lkh0 = T(0), lkh1 = T(0), lkh2 = T(0);
    const T kh0 = K0[i] * a00, kh1 = K0[i] * a01, kh2 = K1[i];
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      const T L = leo[S(i, j)];
      le[S(i, j)] += L;
      lkh0 -= L * err[S(0, j)];
      lkh1 -= L * err[S(1, j)];
      lkh2 -= L * err[S(2, j)];
      le[S(0, j)] -= L * kh0;
      le[S(1, j)] -= L * kh1;
      le[S(2, j)] -= L * kh2;
    }
    lK0[i] = lkh0 * a00 + lkh1 * a01;
    lK1[i] = lkh2;
    la00 += lkh0 * K0[i];
    la01 += lkh1 * K0[i];
  }
  // par'_i = par_i + K0_i l0 + K1_i l1 (the phi squash has derivative 1)
  T ll0 = T(0), ll1 = T(0);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    lp[i] += lpo[i];
    lK0[i] += lpo[i] * l0;
    lK1[i] += lpo[i] * l1;
    ll0 += lpo[i] * K0[i];
    ll1 += lpo[i] * K1[i];
  }
  // K0_i = C_i0 h0 + C_i1 h3 + C_i2 h6;  K1_i = C_i0 h1 + C_i1 h4 + C_i2 h7
  T lh0 = T(0), lh1 = T(0), lh3 = T(0), lh4 = T(0), lh6 = T(0), lh7 = T(0);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    le[S(i, 0)] += lK0[i] * h0 + lK1[i] * h1;
    le[S(i, 1)] += lK0[i] * h3 + lK1[i] * h4;
    le[S(i, 2)] += lK0[i] * h6 + lK1[i] * h7;
    lh0 += lK0[i] * err[S(i, 0)];
    lh3 += lK0[i] * err[S(i, 1)];
    lh6 += lK0[i] * err[S(i, 2)];
    lh1 += lK1[i] * err[S(i, 0)];
    lh4 += lK1[i] * err[S(i, 1)];
    lh7 += lK1[i] * err[S(i, 2)];
  }
  T lg0 = lh0 * a00 + lh3 * a01;
  T lg1 = lh1 * a00 + lh4 * a01 + lh6;
  T lg2 = lh7;
  la00 += lh0 * g0 + lh1 * g1;
  la01 += lh3 * g0 + lh4 * g1;
  // chi2 = g0 l0^2 + g2 l1^2 + 2 g1 l1 l0
  lg0 += lc2 * l0 * l0;
  lg2 += lc2 * l1 * l1;
  lg1 += lc2 * T(2) * l1 * l0;
  ll0 += lc2 * (T(2) * g0 * l0 + T(2) * g1 * l1);
  ll1 += lc2 * (T(2) * g2 * l1 + T(2) * g1 * l0);
  // (g0, g1, g2) = s (s2, -s1, s0), s = 1 / (s0 s2 - s1^2)
  const T ls = lg0 * s2 - lg1 * s1 + lg2 * s0;
  T ls0 = lg2 * s, ls1 = -lg1 * s, ls2 = lg0 * s;
  const T ldet = -ls * s * s;
  ls0 += ldet * s2;
  ls2 += ldet * s0;
  ls1 -= T(2) * ldet * s1;
  // s0 = c0 a00 + c1 a01, s1 = b3 a00 + b4 a01, s2 = b5
  T lb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  const T lc0 = ls0 * a00, lc1 = ls0 * a01;
  la00 += ls0 * c0 + ls1 * b[3];
  la01 += ls0 * c1 + ls1 * b[4];
  lb[3] += ls1 * a00;
  lb[4] += ls1 * a01;
  lb[5] += ls2;
  // c0 = a00 b0 + a01 b1, c1 = a00 b1 + a01 b2
  la00 += lc0 * b[0] + lc1 * b[1];
  la01 += lc0 * b[1] + lc1 * b[2];
  lb[0] += lc0 * a00;
  lb[1] += lc0 * a01 + lc1 * a00;
  lb[2] += lc1 * a01;
  // l0 = a00 rg0 + a01 rg1, l1 = rg2
  la00 += ll0 * rg0;
  la01 += ll0 * rg1;
  const T lrg0 = ll0 * a00, lrg1 = ll0 * a01, lrg2 = ll1;
  // b_i = C_i + me_i
#pragma unroll
  for (int i = 0; i < 6; ++i) le[i] += lb[i];
  // rg = mp - par
  lmp[0] += lrg0;
  lmp[1] += lrg1;
  lmp[2] += lrg2;
  lp[0] -= lrg0;
  lp[1] -= lrg1;
  lp[2] -= lrg2;
  // a00 = -(mp1 + p1) / (2 rk), a01 = (mp0 + p0) / (2 rk)
  const T i2r = T(1) / (T(2) * rk);
  lmp[1] -= la00 * i2r;
  lp[1] -= la00 * i2r;
  lmp[0] += la01 * i2r;
  lp[0] += la01 * i2r;
  const T lrk = -(la00 * a00 + la01 * a01) / rk;
  lmp[0] += lrk * mp[0] / rk;
  lmp[1] += lrk * mp[1] / rk;
}

// ---- similarity C' = J C J^T: adjoint -------------------------------------------
// lJ must start at zero; lC accumulates.
template <class T>
__device__ __forceinline__ void similarity_rev(const JacRT<T>& J, const T (&Cm)[21], const T (&lo)[21], JacRT<T>& lJ,
                                               T (&lC)[21]) {
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    T t[6];
    T a0 = T(0), a1 = T(0), a3 = T(0), a4 = T(0);
    if (i == 3 || i == 5) {
#pragma unroll
      for (int j = 0; j < 6; ++j) t[j] = Cm[S(i, j)];
    } else {
      a0 = i == 0 ? J.j00 : i == 1 ? J.j10 : i == 2 ? J.j20 : J.j40;
      a1 = i == 0 ? J.j01 : i == 1 ? J.j11 : i == 2 ? J.j21 : J.j41;
      a3 = i == 0 ? J.j03 : i == 1 ? J.j13 : i == 2 ? J.j23 : J.j43;
      a4 = i == 0 ? J.j04 : i == 1 ? J.j14 : i == 2 ? J.j24 : J.j44;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        t[j] = a0 * Cm[S(0, j)] + a1 * Cm[S(1, j)] + a3 * Cm[S(3, j)] + a4 * Cm[S(4, j)];
        if (i == 2) t[j] = t[j] + Cm[S(2, j)] + J.j25 * Cm[S(5, j)];
      }
    }
    T lt[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    {  // out(i, 0) = t0 j00 + t1 j01 + t3 j03 + t4 j04
      const T L = lo[S(i, 0)];
      lt[0] += L * J.j00; lt[1] += L * J.j01; lt[3] += L * J.j03; lt[4] += L * J.j04;
      lJ.j00 += L * t[0]; lJ.j01 += L * t[1]; lJ.j03 += L * t[3]; lJ.j04 += L * t[4];
    }
    if (i >= 1) {
      const T L = lo[S(i, 1)];
      lt[0] += L * J.j10; lt[1] += L * J.j11; lt[3] += L * J.j13; lt[4] += L * J.j14;
      lJ.j10 += L * t[0]; lJ.j11 += L * t[1]; lJ.j13 += L * t[3]; lJ.j14 += L * t[4];
    }
    if (i >= 2) {
      const T L = lo[S(i, 2)];
      lt[0] += L * J.j20; lt[1] += L * J.j21; lt[2] += L; lt[3] += L * J.j23; lt[4] += L * J.j24;
      lt[5] += L * J.j25;
      lJ.j20 += L * t[0]; lJ.j21 += L * t[1]; lJ.j23 += L * t[3]; lJ.j24 += L * t[4]; lJ.j25 += L * t[5];
    }
    if (i >= 3) lt[3] += lo[S(i, 3)];
    if (i >= 4) {
      const T L = lo[S(i, 4)];
      lt[0] += L * J.j40; lt[1] += L * J.j41; lt[3] += L * J.j43; lt[4] += L * J.j44;
      lJ.j40 += L * t[0]; lJ.j41 += L * t[1]; lJ.j43 += L * t[3]; lJ.j44 += L * t[4];
    }
    if (i >= 5) lt[5] += lo[S(i, 5)];
    if (i == 3 || i == 5) {
#pragma unroll
      for (int j = 0; j < 6; ++j) lC[S(i, j)] += lt[j];
    } else {
      T la0 = T(0), la1 = T(0), la3 = T(0), la4 = T(0);
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        lC[S(0, j)] += lt[j] * a0;
        lC[S(1, j)] += lt[j] * a1;
        lC[S(3, j)] += lt[j] * a3;
        lC[S(4, j)] += lt[j] * a4;
        la0 += lt[j] * Cm[S(0, j)];
        la1 += lt[j] * Cm[S(1, j)];
        la3 += lt[j] * Cm[S(3, j)];
        la4 += lt[j] * Cm[S(4, j)];
        if (i == 2) {
          lC[S(2, j)] += lt[j];
          lC[S(5, j)] += lt[j] * J.j25;
          lJ.j25 += lt[j] * Cm[S(5, j)];
        }
      }
      if (i == 0) { lJ.j00 += la0; lJ.j01 += la1; lJ.j03 += la3; lJ.j04 += la4; }
      if (i == 1) { lJ.j10 += la0; lJ.j11 += la1; lJ.j13 += la3; lJ.j14 += la4; }
      if (i == 2) { lJ.j20 += la0; lJ.j21 += la1; lJ.j23 += la3; lJ.j24 += la4; }
      if (i == 4) { lJ.j40 += la0; lJ.j41 += la1; lJ.j43 += la3; lJ.j44 += la4; }
    }
  }
}

// ---- material (barrel): adjoint -------------------------------------------------
// par: the state entering the material step.  lp3 holds the adjoint of the
// momentum after the step and returns that of the momentum before it; the
// adjoints of the added variances come in as le33, le44, le55; lp5 and lm
// accumulate.  The covariance entries themselves pass through unchanged.
template <class T>
__device__ __forceinline__ void material_rev(T m, float xi_per_x0, float propSign, const T (&par)[6], T& lp3, T& lp5,
                                             T le33, T le44, T le55, T& lm) {
  if (m < T(1e-13f)) return;
  constexpr float mpi = 0.140;
  constexpr float mpi2 = mpi * mpi;
  constexpr float me = 0.0005;
  constexpr float I = 16.0e-9 * 10.75;
  const T xi = m * T(xi_per_x0);
  const T theta = par[5];
  const T pt = vrcp(par[3]);
  const T sT = vsin(theta), cT = vcos(theta);
  const T p = vdiv(pt, sT);
  const T p2 = p * p;
  const T e2 = p2 + T(mpi2);
  const T beta2 = vdiv(p2, e2);
  const T beta = vsqrt(beta2);
  const T invCos = vdiv(p, pt);
  const T radL = m * invCos;
  if (radL < T(1e-13f)) return;
  const T U = T(0.0136f) * (T(1.f) + T(0.038f) * vlog(radL));
  const T bp = beta * p;
  const T tms = vdiv(U, bp);
  const T gamma2 = vdiv(e2, T(mpi2));
  const T gamma = vsqrt(gamma2);
  const T den = T(1.f) + vdiv(T(2.f) * gamma * T(me), T(mpi)) + T(me) * T(me) / (T(mpi) * T(mpi));
  const T wmax = vdiv(T(2.f) * T(me) * beta2 * gamma2, den);
  const T W = vdiv(T(2.f) * T(me) * beta2 * gamma2 * wmax, T(I * I));
  const T deltahalf = slog(T(28.816e-9f) * ssqrt(T(2.33f) * T(0.498f)) / T(I)) + vlog(beta * gamma) - T(0.5f);
  const bool ok = beta < T(1.f);
  const T A = T(0.5f) * vlog(W) - beta2 - deltahalf;
  const T dEdx = ok ? T(2.f) * vdiv(xi * invCos * A, beta2) : T(0);
  const T dP = vdiv(T(propSign) * dEdx, beta);
  const bool floor_off = p + dP >= T(0.001f);
  const T M = floor_off ? p + dP : T(0.001f);
  const T Q = p2 * pt * pt;

  T l_p = T(0), l_pt = T(0), l_dP = T(0), l_beta = T(0), l_beta2 = T(0), l_p2 = T(0), l_gamma = T(0);
  T l_gamma2 = T(0), l_invCos = T(0), l_radL = T(0), l_xi = T(0);
  // err33 += dP^2 / (p2 pt^2)
  l_dP += le33 * T(2) * dP / Q;
  const T l_Q = -le33 * dP * dP / (Q * Q);
  l_p2 += l_Q * pt * pt;
  l_pt += l_Q * p2 * T(2) * pt;
  // par3' = p / (M pt)
  l_p += lp3 / (M * pt);
  const T l_M = -lp3 * p / (M * M * pt);
  l_pt += -lp3 * p / (M * pt * pt);
  if (floor_off) {
    l_p += l_M;
    l_dP += l_M;
  }
  // dP = propSign dEdx / beta
  const T l_dEdx = l_dP * T(propSign) / beta;
  l_beta -= l_dP * dP / beta;
  if (ok) {
    // dEdx = 2 xi invCos A / beta2, A = 0.5 ln W - beta2 - deltahalf,
    // W = (2 me / I^2) beta2 gamma2 wmax, deltahalf = c + ln(beta gamma) - 0.5
    const T l_N = T(2) * l_dEdx / beta2;
    l_beta2 -= l_dEdx * dEdx / beta2;
    l_xi += l_N * invCos * A;
    l_invCos += l_N * xi * A;
    const T l_A = l_N * xi * invCos;
    l_beta2 += l_A * (T(0.5f) / beta2 - T(1));
    l_gamma2 += l_A * T(0.5f) / gamma2;
    const T l_wmax = l_A * T(0.5f) / wmax;
    l_beta -= l_A / beta;
    l_gamma -= l_A / gamma;
    // wmax = 2 me beta2 gamma2 / den, den = 1 + 2 gamma me / mpi + (me / mpi)^2
    const T l_num = l_wmax / den;
    const T l_den = -l_wmax * wmax / den;
    l_beta2 += l_num * T(2.f) * T(me) * gamma2;
    l_gamma2 += l_num * T(2.f) * T(me) * beta2;
    l_gamma += l_den * T(2.f) * T(me) / T(mpi);
  }
  // gamma = sqrt(gamma2), gamma2 = e2 / mpi2
  l_gamma2 += l_gamma * T(0.5f) / gamma;
  T l_e2 = l_gamma2 / T(mpi2);
  // thetaMSC^2 radL added to err44 and err55; thetaMSC = U / (beta p), U = 0.0136 (1 + 0.038 ln radL)
  const T l_tms2 = le44 + le55;
  const T l_tms = l_tms2 * T(2) * tms * radL;
  l_radL += l_tms2 * tms * tms;
  const T l_bp = -l_tms * tms / bp;
  l_radL += (l_tms / bp) * T(0.0136f) * T(0.038f) / radL;
  l_beta += l_bp * p;
  l_p += l_bp * beta;
  // radL = m invCos, invCos = p / pt
  lm += l_radL * invCos;
  l_invCos += l_radL * m;
  l_p += l_invCos / pt;
  l_pt -= l_invCos * invCos / pt;
  // beta = sqrt(beta2), beta2 = p2 / e2, e2 = p2 + mpi2, p2 = p^2
  l_beta2 += l_beta * T(0.5f) / beta;
  l_p2 += l_beta2 / e2;
  l_e2 -= l_beta2 * beta2 / e2;
  l_p2 += l_e2;
  l_p += l_p2 * T(2) * p;
  // p = pt / sin(theta), xi = m xi_per_x0, pt = 1 / par3
  l_pt += l_p / sT;
  lp5 -= l_p * p * cT / sT;
  lm += l_xi * T(xi_per_x0);
  lp3 = -l_pt * pt * pt;
}

// ---- helix to radius r: adjoint ---------------------------------------------------
// par: the state entering the helix step, tp its recorded iterations.  lo: the
// adjoints of the output parameters, lJ those of the Jacobian entries.  Returns
// the adjoints of par (lp) and of the target radius (lr).
template <class T>
__device__ __forceinline__ void helix_rev(const T (&par)[6], int chg, T r, float bf, const HelixTape<T>& tp,
                                          const T (&lo)[6], const JacRT<T>& lJ, T (&lp)[6], T& lr) {
  const T xin = par[0], yin = par[1], ipt = par[3], phiin = par[4], theta = par[5];
  const T k = vdiv(T(chg) * T(100.f), -T(mkf::kSol) * T(bf));
  const T kinv = vrcp(k);
  const T pt = vrcp(ipt);
  const T cP = vcos(phiin), sP = vsin(phiin);
  const T D = tp.D;
  const T alpha = D * ipt * kinv;
  const T dadx = tp.dDdx * ipt * kinv, dady = tp.dDdy * ipt * kinv;
  const T dadipt = (ipt * tp.dDdipt + D) * kinv, dadphi = tp.dDdphi * ipt * kinv;
  T sa, ca;
  sincos4_t(alpha, sa, ca);
  const T CC = cP * ca - sP * sa, SS = sP * ca + cP * sa;
  const T U = ipt * dadipt;
  const T cT = vcos(theta), sT = vsin(theta), isT = vrcp(sT);

  T l_pt = T(0), l_ipt = lo[3], l_phiin = lo[4], l_theta = lo[5], l_cP = T(0), l_sP = T(0);
  T l_alpha = lo[4];                                       // phi_out = phi_in + alpha
  T l_dadx = lJ.j40, l_dady = lJ.j41, l_dadipt = lJ.j43, l_dadphi = lJ.j44;
  T l_U = T(0), l_CC = T(0), l_SS = T(0), l_ca = T(0), l_sa = T(0);
  // rows 0 and 1: j00 = 1 + k dadx CC pt, j01 = k dady CC pt, j10 = k dadx SS pt, j11 = 1 + k dady SS pt
  l_dadx += (lJ.j00 * CC + lJ.j10 * SS) * k * pt;
  l_dady += (lJ.j01 * CC + lJ.j11 * SS) * k * pt;
  l_CC += (lJ.j00 * dadx + lJ.j01 * dady) * k * pt;
  l_SS += (lJ.j10 * dadx + lJ.j11 * dady) * k * pt;
  l_pt += (lJ.j00 * dadx + lJ.j01 * dady) * k * CC + (lJ.j10 * dadx + lJ.j11 * dady) * k * SS;
  {  // j03 = k pt^2 (cP (U ca - sa) + sP (1 - ca - U sa)), j13 = k pt^2 (sP (U ca - sa) + cP (U sa - 1 + ca))
    const T kp2 = k * pt * pt;
    const T in3 = cP * (U * ca - sa) + sP * (T(1) - ca - U * sa);
    const T in13 = sP * (U * ca - sa) + cP * (U * sa - T(1) + ca);
    l_pt += (lJ.j03 * in3 + lJ.j13 * in13) * T(2) * k * pt;
    const T a3 = lJ.j03 * kp2, a13 = lJ.j13 * kp2;
    l_U += a3 * CC + a13 * SS;
    l_ca += a3 * (cP * U - sP) + a13 * (sP * U + cP);
    l_sa += a3 * (-cP - sP * U) + a13 * (cP * U - sP);
    l_cP += a3 * (U * ca - sa) + a13 * (U * sa - T(1) + ca);
    l_sP += a3 * (T(1) - ca - U * sa) + a13 * (U * ca - sa);
  }
  {  // j04 = k pt (dadphi CC + CC - cP), j14 = k pt (dadphi SS + SS - sP)
    const T kpt = k * pt;
    l_dadphi += (lJ.j04 * CC + lJ.j14 * SS) * kpt;
    l_CC += lJ.j04 * kpt * (dadphi + T(1));
    l_SS += lJ.j14 * kpt * (dadphi + T(1));
    l_cP -= lJ.j04 * kpt;
    l_sP -= lJ.j14 * kpt;
    l_pt += lJ.j04 * k * (dadphi * CC + CC - cP) + lJ.j14 * k * (dadphi * SS + SS - sP);
  }
  // CC = cP ca - sP sa, SS = sP ca + cP sa
  l_cP += l_CC * ca + l_SS * sa;
  l_sP += -l_CC * sa + l_SS * ca;
  l_ca += l_CC * cP + l_SS * sP;
  l_sa += -l_CC * sP + l_SS * cP;
  {  // row 2 and z with KT = k cT pt / sin(theta): j20 = KT dadx, j21 = KT dady, j24 = KT dadphi,
     // j23 = KT pt (U - alpha), j25 = -k alpha pt / sin^2(theta), z = z_in + KT alpha
    const T KT = k * cT * pt * isT;
    const T l_KT = lJ.j20 * dadx + lJ.j21 * dady + lJ.j24 * dadphi + lJ.j23 * pt * (U - alpha) + lo[2] * alpha;
    l_dadx += lJ.j20 * KT;
    l_dady += lJ.j21 * KT;
    l_dadphi += lJ.j24 * KT;
    l_U += lJ.j23 * KT * pt;
    l_alpha += -lJ.j23 * KT * pt + lo[2] * KT - lJ.j25 * k * pt * isT * isT;
    l_pt += lJ.j23 * KT * (U - alpha) - lJ.j25 * k * alpha * isT * isT + l_KT * k * cT * isT;
    const T l_isT = -lJ.j25 * T(2) * k * alpha * pt * isT + l_KT * k * cT * pt;
    const T l_cT = l_KT * k * pt * isT;
    l_theta += -l_cT * sT - l_isT * cT * isT * isT;
  }
  // U = ipt dadipt
  l_ipt += l_U * dadipt;
  l_dadipt += l_U * ipt;
  // sincos4(alpha): s = a - c3 a^3, c = 1 - a^2 / 2 + c4 a^4
  l_alpha += l_sa * (T(1) - T(3) * T(0.16666667f) * alpha * alpha) +
             l_ca * (-alpha + T(4) * T(0.04166667f) * alpha * alpha * alpha);
  // alpha = D ipt kinv; dadx = dDdx ipt kinv; dady, dadphi alike; dadipt = (ipt dDdipt + D) kinv
  T l_D = l_alpha * ipt * kinv + l_dadipt * kinv;
  l_ipt += (l_alpha * D + l_dadx * tp.dDdx + l_dady * tp.dDdy + l_dadphi * tp.dDdphi + l_dadipt * tp.dDdipt) * kinv;
  const T l_dDdx = l_dadx * ipt * kinv, l_dDdy = l_dady * ipt * kinv, l_dDdphi = l_dadphi * ipt * kinv;
  const T l_dDdipt = l_dadipt * kinv * ipt;

  // the iterations, last first; l_x .. l_py are the adjoints of the values leaving iteration i
  T l_x = lo[0], l_y = lo[1], l_px = T(0), l_py = T(0), l_r = T(0);
#pragma unroll
  for (int i = mkf::kNiter - 1; i >= 0; --i) {
    const T x = tp.x[i], y = tp.y[i], px = tp.px[i], py = tp.py[i];
    const T r0 = vsqrt(x * x + y * y);
    const T dot = px * x + py * y;
    const T oodotp = vdiv(r0 * pt, dot);
    const bool fail = oodotp > T(5.0f) || oodotp < T(0);
    const T id = fail ? T(0) : (r - r0) * oodotp;
    const T u = id * ipt * kinv * T(0.5f);
    T sh, ch;
    sincos4_t(u, sh, ch);
    const T cai = T(1.f) - T(2.f) * sh * sh;
    const T sai = T(2.f) * sh * ch;
    const bool conv = i + 1 != mkf::kNiter && r0 > T(0) && vabs(r - r0) < T(0.0001f);
    T l_cai = T(0), l_sai = T(0), l_sh = T(0), l_ch = T(0), l_id = l_D, l_r0 = T(0);
    {  // px' = px ca - py sa, py' = py ca + px sa
      const T lpx1 = l_px, lpy1 = l_py;
      l_cai += lpx1 * px + lpy1 * py;
      l_sai += -lpx1 * py + lpy1 * px;
      l_px = lpx1 * cai + lpy1 * sai;
      l_py = -lpx1 * sai + lpy1 * cai;
    }
    {  // x' = x + 2k sh (px ch - py sh), y' = y + 2k sh (py ch + px sh)
      const T A = px * ch - py * sh, B = py * ch + px * sh, k2 = T(2) * k;
      l_sh += l_x * k2 * (A - sh * py) + l_y * k2 * (B + sh * px);
      l_ch += (l_x * px + l_y * py) * k2 * sh;
      l_px += l_x * k2 * sh * ch + l_y * k2 * sh * sh;
      l_py += -l_x * k2 * sh * sh + l_y * k2 * sh * ch;
    }
    if (conv) {
      // The increments with o = 1 / r0, P = px', Q = py', w = k ipt kinv o, R = x P + y Q:
      //   dDdx -= o x (1 - w R),  dDdy -= o y (1 - w R),
      //   dDdipt -= k pt o (tmp R + (y P - x Q) + (x py - y px)),  tmp = id kinv ipt,
      //   dDdphi += k o (x (px ca + py sa - px) - y (Q - py)).
      const T o = vrcp(r0);
      const T la = -l_dDdx, lb = -l_dDdy, lc = -l_dDdipt, ld = l_dDdphi;
      const T P = px * cai - py * sai, Q = py * cai + px * sai;
      const T w = k * ipt * kinv * o;
      const T R = x * P + y * Q;
      const T Z = T(1) - w * R;
      const T Sab = la * x + lb * y;
      const T tmp = id * kinv * ipt;
      const T Bi = tmp * R + (y * P - x * Q) + (x * py - y * px);
      const T Bp = x * (px * cai + py * sai - px) - y * (Q - py);
      const T kpo = k * pt * o, ko = k * o;
      const T gP = -o * Sab * w * x + lc * kpo * (tmp * x + y);
      const T gQ = -o * Sab * w * y + lc * kpo * (tmp * y - x) - ld * ko * y;
      l_x += o * (Z * la - w * P * Sab) + lc * kpo * (tmp * P - Q + py) + ld * ko * (px * cai + py * sai - px);
      l_y += o * (Z * lb - w * Q * Sab) + lc * kpo * (tmp * Q + P - px) - ld * ko * (Q - py);
      l_px += -lc * kpo * y + ld * ko * x * (cai - T(1)) + gP * cai + gQ * sai;
      l_py += lc * kpo * x + ld * ko * (x * sai + y) - gP * sai + gQ * cai;
      l_cai += ld * ko * x * px + gP * px + gQ * py;
      l_sai += ld * ko * x * py - gP * py + gQ * px;
      const T gtmp = lc * kpo * R;
      l_id += gtmp * kinv * ipt;
      l_ipt += gtmp * id * kinv - o * Sab * R * k * kinv * o;
      l_pt += lc * k * o * Bi;
      const T go = Sab * (T(1) - T(2) * w * R) + lc * k * pt * Bi + ld * k * Bp;
      l_r0 -= go * o * o;
    }
    // ca = 1 - 2 sh^2, sa = 2 sh ch
    l_sh += -l_cai * T(4) * sh + l_sai * T(2) * ch;
    l_ch += l_sai * T(2) * sh;
    // sincos4(u)
    const T l_u = l_sh * (T(1) - T(3) * T(0.16666667f) * u * u) + l_ch * (-u + T(4) * T(0.04166667f) * u * u * u);
    // u = id ipt kinv / 2
    l_id += l_u * ipt * kinv * T(0.5f);
    l_ipt += l_u * id * kinv * T(0.5f);
    if (!fail) {  // id = (r - r0) oodotp, oodotp = r0 pt / dot, dot = px x + py y
      l_r += l_id * oodotp;
      l_r0 -= l_id * oodotp;
      const T l_oo = l_id * (r - r0);
      l_r0 += l_oo * pt / dot;
      l_pt += l_oo * r0 / dot;
      const T l_dot = -l_oo * oodotp / dot;
      l_px += l_dot * x;
      l_py += l_dot * y;
      l_x += l_dot * px;
      l_y += l_dot * py;
    }
    if (r0 > T(0)) {  // r0 = sqrt(x^2 + y^2)
      l_x += l_r0 * x / r0;
      l_y += l_r0 * y / r0;
    }
  }
  // entering values: x = xin, y = yin, px = cP pt, py = sP pt, dDdx = -xin / r0, dDdy = -yin / r0
  T l_xin = l_x, l_yin = l_y;
  l_cP += l_px * pt;
  l_sP += l_py * pt;
  l_pt += l_px * cP + l_py * sP;
  const T r0i = vsqrt(xin * xin + yin * yin);
  if (r0i > T(0)) {
    l_xin -= l_dDdx / r0i;
    l_yin -= l_dDdy / r0i;
    const T l_r0i = (l_dDdx * xin + l_dDdy * yin) / (r0i * r0i);
    l_xin += l_r0i * xin / r0i;
    l_yin += l_r0i * yin / r0i;
  }
  l_phiin += -l_cP * sP + l_sP * cP;
  l_ipt -= l_pt * pt * pt;
  lp[0] = l_xin;
  lp[1] = l_yin;
  lp[2] = lo[2];
  lp[3] = l_ipt;
  lp[4] = l_phiin;
  lp[5] = l_theta;
  lr = l_r;
}

// ---- one step: value and vector-Jacobian product ---------------------------------
// Returns chi2 of the step; g_par, g_err, g_th receive
//   lam_par^T d(par_out)/d(.) + lam_err^T d(err_out)/d(.) + d(chi2)/d(.)
// for the entering state and the layer's (x/X0, r-phi offset, z offset).
template <class T>
__device__ __forceinline__ T align_step_vjp(const T (&par)[6], const T (&err)[21], int chg, const float (&hp)[3],
                                            const float (&he)[6], T m, T drphi, T dz, float xi_per_x0, float bfield,
                                            int pflags, T rh_inv, const T (&lam_par)[6], const T (&lam_err)[21],
                                            T (&g_par)[6], T (&g_err)[21], T (&g_th)[3]) {
  // ---- forward, as align_step
  T mp[3];
  mp[0] = hp[0] - drphi * (hp[1] * rh_inv);
  mp[1] = hp[1] + drphi * (hp[0] * rh_inv);
  mp[2] = hp[2] + dz;
  const T r = vsqrt(mp[0] * mp[0] + mp[1] * mp[1]);
  T ph[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) ph[i] = par[i];
  JacRT<T> J;
  HelixTape<T> tp;
  const bool fail = helix_at_r_t(ph, chg, r, bfield, J, &tp);
  const bool mat = !fail && (pflags & mkf::PF_MATERIAL);
  const T r0 = ssqrt(par[0] * par[0] + par[1] * par[1]);
  const float propSign = r > r0 ? 1.f : -1.f;
  T pk[6], em[21], ek[21];          // state entering the update; covariance after material
#pragma unroll
  for (int i = 0; i < 6; ++i) pk[i] = fail ? par[i] : ph[i];
#pragma unroll
  for (int e = 0; e < 21; ++e) em[e] = err[e];
  if (mat) material_t(m, m * T(xi_per_x0), propSign, pk, em);
  if (!fail) {
    if (pk[4] >= T(mkf::kPI)) pk[4] -= T(mkf::kTwoPI);
    if (pk[4] < T(-mkf::kPI)) pk[4] += T(mkf::kTwoPI);
  }
#pragma unroll
  for (int e = 0; e < 21; ++e) ek[e] = em[e];
  if (!fail) similarity_t(J, ek);
  T pu[6], eu[21];
#pragma unroll
  for (int i = 0; i < 6; ++i) pu[i] = pk[i];
#pragma unroll
  for (int e = 0; e < 21; ++e) eu[e] = ek[e];
  const T chi2 = kalman_t(pu, eu, mp, he);

  // ---- reverse
  T lpk[6], lek[21], lmp[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int i = 0; i < 6; ++i) lpk[i] = T(0);
#pragma unroll
  for (int e = 0; e < 21; ++e) lek[e] = T(0);
  kalman_rev(pk, ek, mp, he, lam_par, lam_err, T(1), lpk, lek, lmp);
  T lm = T(0), lr = T(0);
  if (fail) {
#pragma unroll
    for (int i = 0; i < 6; ++i) g_par[i] = lpk[i];
#pragma unroll
    for (int e = 0; e < 21; ++e) g_err[e] = lek[e];
  } else {
    JacRT<T> lJ = {};
#pragma unroll
    for (int e = 0; e < 21; ++e) g_err[e] = T(0);
    similarity_rev(J, em, lek, lJ, g_err);
    T lph[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) lph[i] = lpk[i];
    if (mat) material_rev(m, xi_per_x0, propSign, ph, lph[3], lph[5], g_err[S(3, 3)], g_err[S(4, 4)], g_err[S(5, 5)], lm);
    helix_rev(par, chg, r, bfield, tp, lph, lJ, g_par, lr);
  }
  // r = |mp_xy|; mp = hit moved by (drphi, dz)
  lmp[0] += lr * mp[0] / r;
  lmp[1] += lr * mp[1] / r;
  g_th[0] = lm;
  g_th[1] = -lmp[0] * (hp[1] * rh_inv) + lmp[1] * (hp[0] * rh_inv);
  g_th[2] = lmp[2];
  return chi2;
}

}  // namespace mkf_adj
