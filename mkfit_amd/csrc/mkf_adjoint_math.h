// mkf_adjoint_math.h -- the barrel fit step written once for every model type:
// float or double (the forward pass and the recomputation from checkpoints)
// and Dual<T, D>, a value with D forward tangents, which gives D columns of the
// step's Jacobian per evaluation (the check of the reverse-mode derivative in
// mkf_adjoint_rev.h).  The arithmetic follows the reference's order of
// operations as cited in mkf_track_math.h (the material step and the 2x2
// determinant in their reference form); only the scalar type differs.
#pragma once

#include <hip/hip_runtime.h>

#include "mkf_track_math.h"

namespace mkf_adj {

using mkf::S;

// A value with D forward tangents over the scalar type T (float: the FP32 model;
// double: the same model evaluated in FP64).
template <class T, int D>
struct Dual {
  T v;
  T d[D];
};

// Scalar type of a model value (float, double, or a Dual's).
template <class V> struct ScalarOf { using type = V; };
template <class T, int D> struct ScalarOf<Dual<T, D>> { using type = T; };
template <class V> using SOf = typename ScalarOf<V>::type;
// Keeps a scalar operand out of template deduction, so a float literal combines
// with a double Dual.
template <class T> struct Id { using type = T; };
template <class T> using NoDed = typename Id<T>::type;

// ---- Dual arithmetic ----------------------------------------------------------
template <class T, int D> __device__ __forceinline__ Dual<T, D> cst(T x) {
  Dual<T, D> r;
  r.v = x;
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = T(0);
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator+(const Dual<T, D>& a, const Dual<T, D>& b) {
  Dual<T, D> r; r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator-(const Dual<T, D>& a, const Dual<T, D>& b) {
  Dual<T, D> r; r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator-(const Dual<T, D>& a) {
  Dual<T, D> r; r.v = -a.v;
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = -a.d[i];
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator*(const Dual<T, D>& a, const Dual<T, D>& b) {
  Dual<T, D> r; r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator/(const Dual<T, D>& a, const Dual<T, D>& b) {
  Dual<T, D> r;
  const T ib = T(1) / b.v;
  r.v = a.v / b.v;
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator+(const Dual<T, D>& a, NoDed<T> b) {
  Dual<T, D> r = a; r.v = a.v + b; return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator+(NoDed<T> a, const Dual<T, D>& b) { return b + a; }
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator-(const Dual<T, D>& a, NoDed<T> b) {
  Dual<T, D> r = a; r.v = a.v - b; return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator-(NoDed<T> a, const Dual<T, D>& b) {
  Dual<T, D> r; r.v = a - b.v;
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = -b.d[i];
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator*(const Dual<T, D>& a, NoDed<T> b) {
  Dual<T, D> r; r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator*(NoDed<T> a, const Dual<T, D>& b) { return b * a; }
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator/(const Dual<T, D>& a, NoDed<T> b) {
  Dual<T, D> r; r.v = a.v / b;
  const T ib = T(1) / b;
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = a.d[i] * ib;
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> operator/(NoDed<T> a, const Dual<T, D>& b) {
  return cst<T, D>(a) / b;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D>& operator+=(Dual<T, D>& a, const Dual<T, D>& b) { a = a + b; return a; }
template <class T, int D> __device__ __forceinline__ Dual<T, D>& operator-=(Dual<T, D>& a, const Dual<T, D>& b) { a = a - b; return a; }
template <class T, int D> __device__ __forceinline__ Dual<T, D>& operator-=(Dual<T, D>& a, NoDed<T> b) { a.v -= b; return a; }
template <class T, int D> __device__ __forceinline__ Dual<T, D>& operator+=(Dual<T, D>& a, NoDed<T> b) { a.v += b; return a; }

// ---- scalar functions over all model types ----------------------------------
// float: the fit's own refined hardware divide / reciprocal / root (mkf_track_math.h);
// double: IEEE operations.
__device__ __forceinline__ float val(float x) { return x; }
__device__ __forceinline__ double val(double x) { return x; }
template <class T, int D> __device__ __forceinline__ T val(const Dual<T, D>& x) { return x.v; }

__device__ __forceinline__ float vrcp(float x) { return mkf::frcp(x); }
__device__ __forceinline__ float vdiv(float a, float b) { return mkf::fdiv(a, b); }
__device__ __forceinline__ float vsqrt(float x) { return mkf::fsqrt(x); }
__device__ __forceinline__ float vsin(float x) { return sinf(x); }
__device__ __forceinline__ float vcos(float x) { return cosf(x); }
__device__ __forceinline__ float vlog(float x) { return logf(x); }
__device__ __forceinline__ float vabs(float x) { return fabsf(x); }
__device__ __forceinline__ double vrcp(double x) { return 1.0 / x; }
__device__ __forceinline__ double vdiv(double a, double b) { return a / b; }
__device__ __forceinline__ double vsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double vsin(double x) { return sin(x); }
__device__ __forceinline__ double vcos(double x) { return cos(x); }
__device__ __forceinline__ double vlog(double x) { return log(x); }
__device__ __forceinline__ double vabs(double x) { return fabs(x); }
// plain-math versions for the value part of a Dual
__device__ __forceinline__ float ssqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double ssqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float ssin(float x) { return sinf(x); }
__device__ __forceinline__ double ssin(double x) { return sin(x); }
__device__ __forceinline__ float scos(float x) { return cosf(x); }
__device__ __forceinline__ double scos(double x) { return cos(x); }
__device__ __forceinline__ float slog(float x) { return logf(x); }
__device__ __forceinline__ double slog(double x) { return log(x); }

template <class T, int D> __device__ __forceinline__ Dual<T, D> vrcp(const Dual<T, D>& x) { return T(1) / x; }
template <class T, int D> __device__ __forceinline__ Dual<T, D> vdiv(const Dual<T, D>& a, const Dual<T, D>& b) {
  return a / b;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> vsqrt(const Dual<T, D>& x) {
  Dual<T, D> r; r.v = ssqrt(x.v);
  const T h = r.v > T(0) ? T(0.5) / r.v : T(0);
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = x.d[i] * h;
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> vsin(const Dual<T, D>& x) {
  Dual<T, D> r; r.v = ssin(x.v);
  const T c = scos(x.v);
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = x.d[i] * c;
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> vcos(const Dual<T, D>& x) {
  Dual<T, D> r; r.v = scos(x.v);
  const T s = -ssin(x.v);
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = x.d[i] * s;
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> vlog(const Dual<T, D>& x) {
  Dual<T, D> r; r.v = slog(x.v);
  const T ix = T(1) / x.v;
#pragma unroll
  for (int i = 0; i < D; ++i) r.d[i] = x.d[i] * ix;
  return r;
}
template <class T, int D> __device__ __forceinline__ Dual<T, D> vabs(const Dual<T, D>& x) { return x.v < T(0) ? -x : x; }

template <class V> __device__ __forceinline__ V vmax(const V& a, const V& b) { return val(a) >= val(b) ? a : b; }
template <class V> struct MkC { static __device__ __forceinline__ V c(SOf<V> x) { return x; } };
template <class T, int D> struct MkC<Dual<T, D>> {
  static __device__ __forceinline__ Dual<T, D> c(T x) { return cst<T, D>(x); }
};
template <class V> __device__ __forceinline__ V C(SOf<V> x) { return MkC<V>::c(x); }

// ---- helix to radius r (barrel), approximate Jacobian -------------------------
template <class V>
struct JacRT {
  V j00, j01, j03, j04, j10, j11, j13, j14, j20, j21, j23, j24, j25, j40, j41, j43, j44;
};

// Forward values of the helix step that its reverse needs: the position and
// momentum entering each iteration and the accumulated path and derivatives.
template <class V>
struct HelixTape {
  V x[mkf::kNiter], y[mkf::kNiter], px[mkf::kNiter], py[mkf::kNiter];
  V D, dDdx, dDdy, dDdipt, dDdphi;
};

template <class V>
__device__ __forceinline__ void sincos4_t(const V& x, V& s, V& c) {   // Matrix.h:48-56
  const V x2 = x * x;
  c = 1.f - 0.5f * x2 + 0.04166667f * x2 * x2;
  s = x - 0.16666667f * x * x2;
}

template <class V>
__device__ __forceinline__ bool helix_at_r_t(V (&par)[6], int chg, const V& r, float bf, JacRT<V>& J,
                                             HelixTape<V>* tape = nullptr) {
  const V xin = par[0], yin = par[1], zin = par[2];
  const V ipt = par[3], phiin = par[4], theta = par[5];
  V r0 = vsqrt(xin * xin + yin * yin);
  using T = SOf<V>;
  const T k = vdiv(T(chg) * T(100.f), -T(mkf::kSol) * T(bf));
  const T kinv = vrcp(k);
  const V pt = vrcp(ipt);
  V D = C<V>(0.f), cosa = C<V>(0.f), sina = C<V>(0.f), cosah = C<V>(0.f), sinah = C<V>(0.f), id = C<V>(0.f);
  V cosPorT = vcos(phiin), sinPorT = vsin(phiin);
  V pxin = cosPorT * pt;
  V pyin = sinPorT * pt;
  V dDdx = val(r0) > 0.f ? vdiv(-xin, r0) : C<V>(0.f);
  V dDdy = val(r0) > 0.f ? vdiv(-yin, r0) : C<V>(0.f);
  V dDdipt = C<V>(0.f), dDdphi = C<V>(0.f);
  bool fail = false;
  V x = xin, y = yin;
#pragma unroll
  for (int i = 0; i < mkf::kNiter; ++i) {
    if (tape) {
      tape->x[i] = x;
      tape->y[i] = y;
      tape->px[i] = pxin;
      tape->py[i] = pyin;
    }
    r0 = vsqrt(x * x + y * y);
    const V oodotp = vdiv(r0 * pt, pxin * x + pyin * y);
    if (val(oodotp) > 5.0f || val(oodotp) < 0) {
      id = C<V>(0.f);
      fail = true;
    } else {
      id = (r - r0) * oodotp;
    }
    D += id;
    sincos4_t(id * ipt * kinv * 0.5f, sinah, cosah);
    cosa = 1.f - 2.f * sinah * sinah;
    sina = 2.f * sinah * cosah;
    if (i + 1 != mkf::kNiter) {
      const V oor0 = (val(r0) > T(0) && vabs(val(r) - val(r0)) < T(0.0001f)) ? vrcp(r0) : C<V>(0.f);
      const V dadipt = id * kinv;
      const V dadx = -x * ipt * kinv * oor0;
      const V dady = -y * ipt * kinv * oor0;
      const V pxca = pxin * cosa, pxsa = pxin * sina, pyca = pyin * cosa, pysa = pyin * sina;
      V tmp;
      tmp = k * dadx;
      dDdx -= (x * (1.f + tmp * (pxca - pysa)) + y * tmp * (pyca + pxsa)) * oor0;
      tmp = k * dady;
      dDdy -= (x * tmp * (pxca - pysa) + y * (1.f + tmp * (pyca + pxsa))) * oor0;
      tmp = dadipt * ipt;
      dDdipt -= k * (x * (pxca * tmp - pysa * tmp - pyca - pxsa + pyin) +
                     y * (pyca * tmp + pxsa * tmp - pysa + pxca - pxin)) * pt * oor0;
      dDdphi += k * (x * (pysa - pxin + pxca) - y * (pxsa - pyin + pyca)) * oor0;
    }
    x = x + 2.f * k * sinah * (pxin * cosah - pyin * sinah);
    y = y + 2.f * k * sinah * (pyin * cosah + pxin * sinah);
    const V pxinold = pxin;
    pxin = pxin * cosa - pyin * sina;
    pyin = pyin * cosa + pxinold * sina;
  }
  if (tape) {
    tape->D = D;
    tape->dDdx = dDdx;
    tape->dDdy = dDdy;
    tape->dDdipt = dDdipt;
    tape->dDdphi = dDdphi;
  }
  const V alpha = D * ipt * kinv;
  const V dadx = dDdx * ipt * kinv;
  const V dady = dDdy * ipt * kinv;
  const V dadipt = (ipt * dDdipt + D) * kinv;
  const V dadphi = dDdphi * ipt * kinv;
  sincos4_t(alpha, sina, cosa);
  J.j00 = 1.f + k * dadx * (cosPorT * cosa - sinPorT * sina) * pt;
  J.j01 = k * dady * (cosPorT * cosa - sinPorT * sina) * pt;
  J.j03 = k * (cosPorT * (ipt * dadipt * cosa - sina) + sinPorT * ((1.f - cosa) - ipt * dadipt * sina)) * pt * pt;
  J.j04 = k * (cosPorT * dadphi * cosa - sinPorT * dadphi * sina - sinPorT * sina + cosPorT * cosa - cosPorT) * pt;
  J.j10 = k * dadx * (sinPorT * cosa + cosPorT * sina) * pt;
  J.j11 = 1.f + k * dady * (sinPorT * cosa + cosPorT * sina) * pt;
  J.j13 = k * (sinPorT * (ipt * dadipt * cosa - sina) + cosPorT * (ipt * dadipt * sina - (1.f - cosa))) * pt * pt;
  J.j14 = k * (sinPorT * dadphi * cosa + cosPorT * dadphi * sina + sinPorT * cosa + cosPorT * sina - sinPorT) * pt;
  cosPorT = vcos(theta);
  sinPorT = vrcp(vsin(theta));
  par[0] = x;
  par[1] = y;
  par[2] = zin + k * alpha * cosPorT * pt * sinPorT;
  J.j20 = k * cosPorT * dadx * pt * sinPorT;
  J.j21 = k * cosPorT * dady * pt * sinPorT;
  J.j23 = k * cosPorT * (ipt * dadipt - alpha) * pt * pt * sinPorT;
  J.j24 = k * dadphi * cosPorT * pt * sinPorT;
  J.j25 = -k * alpha * pt * sinPorT * sinPorT;
  par[4] = phiin + alpha;
  J.j40 = dadx;
  J.j41 = dady;
  J.j43 = dadipt;
  J.j44 = 1.f + dadphi;
  return fail;
}

// ---- applyMaterialEffects (barrel), material given per layer -----------------
template <class V>
__device__ __forceinline__ void material_t(V radL, const V& xi, float propSign, V (&par)[6], V (&err)[21]) {
  if (val(radL) < 1e-13f) return;
  const V theta = par[5];
  const V pt = vrcp(par[3]);
  const V p = vdiv(pt, vsin(theta));
  const V p2 = p * p;
  constexpr float mpi = 0.140;
  constexpr float mpi2 = mpi * mpi;
  const V beta2 = vdiv(p2, p2 + mpi2);
  const V beta = vsqrt(beta2);
  const V invCos = vdiv(p, pt);
  radL = radL * invCos;
  if (val(radL) < 1e-13f) return;
  const V thetaMSC = vdiv(0.0136f * (1.f + 0.038f * vlog(radL)), beta * p);
  const V thetaMSC2 = thetaMSC * thetaMSC * radL;
  err[S(4, 4)] += thetaMSC2;
  err[S(5, 5)] += thetaMSC2;
  const V gamma2 = vdiv(p2 + mpi2, C<V>(mpi2));
  const V gamma = vsqrt(gamma2);
  constexpr float me = 0.0005;
  const V wmax = vdiv(2.f * me * beta2 * gamma2,
                      1.f + vdiv(2.f * gamma * me, C<V>(mpi)) + SOf<V>(me) * SOf<V>(me) / (SOf<V>(mpi) * SOf<V>(mpi)));
  constexpr float I = 16.0e-9 * 10.75;
  using T = SOf<V>;
  const V deltahalf = slog(T(28.816e-9f) * ssqrt(T(2.33f) * T(0.498f)) / T(I)) + vlog(beta * gamma) - 0.5f;
  const V dEdx = val(beta) < 1.f
                     ? (2.f * vdiv(xi * invCos * (0.5f * vlog(vdiv(2.f * me * beta2 * gamma2 * wmax, C<V>(I * I))) -
                                                  beta2 - deltahalf),
                                   beta2))
                     : C<V>(0.f);
  const V dP = vdiv(propSign * dEdx, beta);
  par[3] = vdiv(p, vmax(p + dP, C<V>(0.001f)) * pt);
  err[S(3, 3)] += vdiv(dP * dP, p2 * pt * pt);
}

// ---- similarity C' = J C J^T (barrel pattern) --------------------------------
template <class V>
__device__ __forceinline__ void similarity_t(const JacRT<V>& J, V (&Cm)[21]) {
  V out[21];
  V t[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if (i == 3 || i == 5) {
#pragma unroll
      for (int j = 0; j < 6; ++j) t[j] = Cm[S(i, j)];
    } else {
      const V& a0 = i == 0 ? J.j00 : i == 1 ? J.j10 : i == 2 ? J.j20 : J.j40;
      const V& a1 = i == 0 ? J.j01 : i == 1 ? J.j11 : i == 2 ? J.j21 : J.j41;
      const V& a3 = i == 0 ? J.j03 : i == 1 ? J.j13 : i == 2 ? J.j23 : J.j43;
      const V& a4 = i == 0 ? J.j04 : i == 1 ? J.j14 : i == 2 ? J.j24 : J.j44;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        if (i == 2)
          t[j] = a0 * Cm[S(0, j)] + a1 * Cm[S(1, j)] + Cm[S(2, j)] + a3 * Cm[S(3, j)] + a4 * Cm[S(4, j)] +
                 J.j25 * Cm[S(5, j)];
        else
          t[j] = a0 * Cm[S(0, j)] + a1 * Cm[S(1, j)] + a3 * Cm[S(3, j)] + a4 * Cm[S(4, j)];
      }
    }
    if (0 <= i) out[S(i, 0)] = t[0] * J.j00 + t[1] * J.j01 + t[3] * J.j03 + t[4] * J.j04;
    if (1 <= i) out[S(i, 1)] = t[0] * J.j10 + t[1] * J.j11 + t[3] * J.j13 + t[4] * J.j14;
    if (2 <= i) out[S(i, 2)] = t[0] * J.j20 + t[1] * J.j21 + t[2] + t[3] * J.j23 + t[4] * J.j24 + t[5] * J.j25;
    if (3 <= i) out[S(i, 3)] = t[3];
    if (4 <= i) out[S(i, 4)] = t[0] * J.j40 + t[1] * J.j41 + t[3] * J.j43 + t[4] * J.j44;
    if (5 <= i) out[S(i, 5)] = t[5];
  }
#pragma unroll
  for (int e = 0; e < 21; ++e) Cm[e] = out[e];
}

// ---- 2x2 inverse: FP32 values as in the fit (double determinant), tangents of 1/det;
// FP64: the plain inverse.
__device__ __forceinline__ void invert_sym2_t(float& a0, float& a1, float& a2) { mkf::invert_sym2(a0, a1, a2); }
template <int D>
__device__ __forceinline__ void invert_sym2_t(Dual<float, D>& a0, Dual<float, D>& a1, Dual<float, D>& a2) {
  float v0 = a0.v, v1 = a1.v, v2 = a2.v;
  mkf::invert_sym2(v0, v1, v2);                       // values exactly as the float path
  const float s = (float)(1.0 / ((double)a0.v * a2.v - (double)a1.v * a1.v));
  Dual<float, D> n0, n1, n2;
  n0.v = v0; n1.v = v1; n2.v = v2;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const float ddet = a0.d[i] * a2.v + a0.v * a2.d[i] - 2.f * a1.v * a1.d[i];
    const float ds = -s * s * ddet;
    n0.d[i] = ds * a2.v + s * a2.d[i];
    n1.d[i] = -(ds * a1.v + s * a1.d[i]);
    n2.d[i] = ds * a0.v + s * a0.d[i];
  }
  a0 = n0; a1 = n1; a2 = n2;
}
template <class V>
__device__ __forceinline__ void invert_sym2_t(V& a0, V& a1, V& a2) {   // double and Dual<double>
  const V s = vrcp(a0 * a2 - a1 * a1);
  const V tmp = s * a2;
  a1 = -(a1 * s);
  a2 = s * a0;
  a0 = tmp;
}

// ---- kalmanOperation (barrel): chi2 and update --------------------------------
template <class V>
__device__ __forceinline__ V kalman_t(V (&par)[6], V (&err)[21], const V (&mp)[3], const float (&me)[6]) {
  const V r = vsqrt(mp[0] * mp[0] + mp[1] * mp[1]);
  const V a00 = vdiv(-(mp[1] + par[1]), 2.f * r);
  const V a01 = vdiv(mp[0] + par[0], 2.f * r);
  const V rg0 = mp[0] - par[0], rg1 = mp[1] - par[1], rg2 = mp[2] - par[2];
  V b[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) b[i] = err[i] + me[i];
  const V l0 = a00 * rg0 + a01 * rg1;
  const V l1 = rg2;
  const V c0 = a00 * b[0] + a01 * b[1];
  const V c1 = a00 * b[1] + a01 * b[2];
  V s0 = c0 * a00 + c1 * a01;
  V s1 = b[3] * a00 + b[4] * a01;
  V s2 = b[5];
  invert_sym2_t(s0, s1, s2);
  const V chi2 = s0 * l0 * l0 + s2 * l1 * l1 + 2.f * (s1 * l1 * l0);
  const V h0 = a00 * s0, h1 = a00 * s1, h3 = a01 * s0, h4 = a01 * s1, h6 = s1, h7 = s2;
  V K0[6], K1[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    K0[i] = err[S(i, 0)] * h0 + err[S(i, 1)] * h3 + err[S(i, 2)] * h6;
    K1[i] = err[S(i, 0)] * h1 + err[S(i, 1)] * h4 + err[S(i, 2)] * h7;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) par[i] = par[i] + K0[i] * l0 + K1[i] * l1;
  if (val(par[4]) >= mkf::kPI) par[4] -= mkf::kTwoPI;
  if (val(par[4]) < -mkf::kPI) par[4] += mkf::kTwoPI;
  V out[21];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const V kh0 = K0[i] * a00, kh1 = K0[i] * a01, kh2 = K1[i];
#pragma unroll
    for (int j = 0; j <= i; ++j) out[S(i, j)] = err[S(i, j)] - (kh0 * err[S(0, j)] + kh1 * err[S(1, j)] + kh2 * err[S(2, j)]);
  }
#pragma unroll
  for (int e = 0; e < 21; ++e) err[e] = out[e];
  return chi2;
}

// ---- one fit step of the alignment model: the hit moved by the layer's
// alignment, propagation to its radius with the layer's material, update.
template <class V>
__device__ __forceinline__ V align_step(V (&par)[6], V (&err)[21], int chg, const float (&hp)[3], const float (&he)[6],
                                        const V& m, const V& drphi, const V& dz, float xi_per_x0, float bfield,
                                        int pflags, SOf<V> rh_inv) {
  using T = SOf<V>;
  V mp[3];
  mp[0] = hp[0] - drphi * (hp[1] * rh_inv);
  mp[1] = hp[1] + drphi * (hp[0] * rh_inv);
  mp[2] = hp[2] + dz;
  const V r = vsqrt(mp[0] * mp[0] + mp[1] * mp[1]);
  V out[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) out[i] = par[i];
  JacRT<V> J;
  const bool fail = helix_at_r_t(out, chg, r, bfield, J);
  if (!fail) {
    if (pflags & mkf::PF_MATERIAL) {
      const T r0 = ssqrt(val(par[0]) * val(par[0]) + val(par[1]) * val(par[1]));
      material_t(m, m * T(xi_per_x0), val(r) > r0 ? 1.f : -1.f, out, err);
    }
    if (val(out[4]) >= mkf::kPI) out[4] -= mkf::kTwoPI;
    if (val(out[4]) < -mkf::kPI) out[4] += mkf::kTwoPI;
    similarity_t(J, err);
#pragma unroll
    for (int i = 0; i < 6; ++i) par[i] = out[i];
  }
  return kalman_t(par, err, mp, he);
}

}  // namespace mkf_adj
