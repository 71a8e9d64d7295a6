// oracle.cc -- C entry points over the oracle_core.h restatement.
//
// TEST INFRASTRUCTURE ONLY (checker and CPU baseline); the product library
// never links it.
//
// The fit driver restates the reference's harness and fit loops:
//   runFittingTestPlex            mkFit/fittestMPlex.cc:45-110  (blocks of NN,
//                                 grain max(1, 32/NN) blocks per task)
//   MkFitter::InputTracksForFit   mkFit/MkFitter.cc:286-321     (packing)
//   MkFitter::FitTracksSteered    mkFit/MkFitter.cc:473-503     (barrel/endcap)
//   MkFinder::BkFitFitTracks      mkFit/MkFinder.cc:1573-1742   (chi2 accumulation,
//                                 missing-hit fake hit {666,0,666,0,0,666} at the
//                                 current track position with no material, charge flip)
//   MkBase::PropagateTracksToHit{R,Z} mkFit/MkBase.h:48-89      (target = hypot / z of hit)
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "oracle_core.h"

using namespace orc;

namespace {

Ctx make_ctx(const orc_ctx* c) {
  Ctx x;
  x.bfield = c ? c->bfield : 3.8112f;
  x.rl_grid = c ? c->rl_grid : nullptr;
  x.xi_grid = c ? c->xi_grid : nullptr;
  return x;
}

// Gather lanes [t0, t0+cnt) of an element-major batch array into a block;
// padding lanes replicate lane t0 so every lane computes on valid numbers.
template <class Td, class Ts>
inline void pack(Td* blk, const Ts* src, int ne, int n, int t0, int cnt) {
  for (int e = 0; e < ne; ++e)
    for (int l = 0; l < NN; ++l) blk[e * NN + l] = Td(src[(size_t)e * n + t0 + (l < cnt ? l : 0)]);
}
template <class Td, class Ts>
inline void unpack(Td* dst, const Ts* blk, int ne, int n, int t0, int cnt) {
  for (int e = 0; e < ne; ++e)
    for (int l = 0; l < cnt; ++l) dst[(size_t)e * n + t0 + l] = Td(blk[e * NN + l]);
}

template <class F>
void for_each_block(int n, int n_threads, F&& body) {
  const int nblk = (n + NN - 1) / NN;
  const int grain = std::max(1, 32 / NN);                 // fittestMPlex.cc:67, Config::numSeedsPerTask
  if (n_threads <= 1) {
    for (int b = 0; b < nblk; ++b) body(b);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < n_threads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        const int b0 = next.fetch_add(grain);
        if (b0 >= nblk) break;
        for (int b = b0; b < std::min(nblk, b0 + grain); ++b) body(b);
      }
    });
  for (auto& th : pool) th.join();
}

template <class T>
int propagate_r_batch(const orc_ctx* c, int n, const T* in_err, const T* in_par, const int* in_chg, const T* ms_rad,
                      T* out_err, T* out_par, int pflags, const int* no_mat, int* fail_out) {
  const Ctx ctx = make_ctx(c);
  for_each_block(n, 1, [&](int b) {
    const int t0 = b * NN, cnt = std::min(NN, n - t0);
    alignas(64) T ie[21 * NN], ip[6 * NN], oe[21 * NN], op[6 * NN], r[NN];
    alignas(64) int ch[NN], nm[NN], fl[NN];
    pack(ie, in_err, 21, n, t0, cnt);
    pack(ip, in_par, 6, n, t0, cnt);
    pack(r, ms_rad, 1, n, t0, cnt);
    pack(ch, in_chg, 1, n, t0, cnt);
    if (no_mat) pack(nm, no_mat, 1, n, t0, cnt);
    propagate_helix_to_r(ie, ip, ch, r, oe, op, cnt, pflags, no_mat ? nm : nullptr, fl, ctx);
    unpack(out_err, oe, 21, n, t0, cnt);
    unpack(out_par, op, 6, n, t0, cnt);
    if (fail_out) unpack(fail_out, fl, 1, n, t0, cnt);
  });
  return 0;
}

template <class T>
int propagate_z_batch(const orc_ctx* c, int n, const T* in_err, const T* in_par, const int* in_chg, const T* ms_z,
                      T* out_err, T* out_par, int pflags, const int* no_mat) {
  const Ctx ctx = make_ctx(c);
  for_each_block(n, 1, [&](int b) {
    const int t0 = b * NN, cnt = std::min(NN, n - t0);
    alignas(64) T ie[21 * NN], ip[6 * NN], oe[21 * NN], op[6 * NN], z[NN];
    alignas(64) int ch[NN], nm[NN];
    pack(ie, in_err, 21, n, t0, cnt);
    pack(ip, in_par, 6, n, t0, cnt);
    pack(z, ms_z, 1, n, t0, cnt);
    pack(ch, in_chg, 1, n, t0, cnt);
    if (no_mat) pack(nm, no_mat, 1, n, t0, cnt);
    propagate_helix_to_z(ie, ip, ch, z, oe, op, cnt, pflags, no_mat ? nm : nullptr, ctx);
    unpack(out_err, oe, 21, n, t0, cnt);
    unpack(out_par, op, 6, n, t0, cnt);
  });
  return 0;
}

template <class T>
int helix_r_batch(const orc_ctx* c, int n, const T* in_par, const int* in_chg, const T* ms_rad, T* out_par, T* out_jac,
                  int* fail_out, int pflags) {
  const Ctx ctx = make_ctx(c);
  for_each_block(n, 1, [&](int b) {
    const int t0 = b * NN, cnt = std::min(NN, n - t0);
    alignas(64) T ip[6 * NN], op[6 * NN], r[NN], J[36 * NN];
    alignas(64) int ch[NN], fl[NN];
    pack(ip, in_par, 6, n, t0, cnt);
    pack(r, ms_rad, 1, n, t0, cnt);
    pack(ch, in_chg, 1, n, t0, cnt);
    std::memcpy(op, ip, sizeof(op));
    helix_at_r(ip, ch, r, op, J, fl, pflags, ctx);
    unpack(out_par, op, 6, n, t0, cnt);
    if (out_jac) unpack(out_jac, J, 36, n, t0, cnt);
    if (fail_out) unpack(fail_out, fl, 1, n, t0, cnt);
  });
  return 0;
}

template <class T>
int helix_z_batch(const orc_ctx* c, int n, const T* in_par, const int* in_chg, const T* ms_z, T* out_par, T* out_jac,
                  int pflags) {
  const Ctx ctx = make_ctx(c);
  for_each_block(n, 1, [&](int b) {
    const int t0 = b * NN, cnt = std::min(NN, n - t0);
    alignas(64) T ip[6 * NN], op[6 * NN], z[NN], J[36 * NN];
    alignas(64) int ch[NN];
    pack(ip, in_par, 6, n, t0, cnt);
    pack(z, ms_z, 1, n, t0, cnt);
    pack(ch, in_chg, 1, n, t0, cnt);
    std::memcpy(op, ip, sizeof(op));
    helix_at_z(ip, ch, z, op, J, pflags, ctx);
    unpack(out_par, op, 6, n, t0, cnt);
    if (out_jac) unpack(out_jac, J, 36, n, t0, cnt);
  });
  return 0;
}

template <class T>
int kalman_batch(int op, int endcap, int n, const T* ps_err, const T* ps_par, const T* ms_err, const T* ms_par,
                 T* out_err, T* out_par, T* out_chi2) {
  for_each_block(n, 1, [&](int b) {
    const int t0 = b * NN, cnt = std::min(NN, n - t0);
    alignas(64) T pe[21 * NN], pp[6 * NN], me[6 * NN], mp[3 * NN], oe[21 * NN], opar[6 * NN], c2[NN];
    pack(pe, ps_err, 21, n, t0, cnt);
    pack(pp, ps_par, 6, n, t0, cnt);
    pack(me, ms_err, 6, n, t0, cnt);
    pack(mp, ms_par, 3, n, t0, cnt);
    if (endcap)
      kalman_operation_endcap(op, pe, pp, me, mp, oe, opar, c2);
    else
      kalman_operation(op, pe, pp, me, mp, oe, opar, c2);
    if (op & KFO_Update_Params) {
      if (out_err) unpack(out_err, oe, 21, n, t0, cnt);
      if (out_par) unpack(out_par, opar, 6, n, t0, cnt);
    }
    if ((op & KFO_Calculate_Chi2) && out_chi2) unpack(out_chi2, c2, 1, n, t0, cnt);
  });
  return 0;
}

// Conditioning test of an updated covariance (packed 6x6, one lane): a negative
// or NaN variance, or a covariance beyond what the variances allow by more
// than rounding, C_ij^2 > (1 + 1e-5) C_ii C_jj (this also catches a nonzero
// covariance of a zero variance; a zero row, as the z row after an endcap step,
// passes).  Healthy track covariances reach correlations of 1 - 1e-7, which
// FP32 rounding alone can push to 1 + 3e-7; the margin keeps those out.  The
// products are formed in double, where the product of two floats is exact, so
// the decision is the same on every platform for the same input bits.
template <class T>
bool ill_conditioned(const T* C, int lane) {
  for (int i = 0; i < 6; ++i)
    if (!(C[sym(i, i) * NN + lane] >= T(0))) return true;
  for (int i = 1; i < 6; ++i)
    for (int j = 0; j < i; ++j) {
      const double cij = (double)C[sym(i, j) * NN + lane];
      const double lim = (double)C[sym(i, i) * NN + lane] * (double)C[sym(j, j) * NN + lane];
      if (cij * cij > lim * 1.00001) return true;
    }
  return false;
}

// The FP64 fallback: the updated covariance in the Joseph form
//   C' = (I - K H) C (I - K H)^T + K V K^T,   K = C H^T (H C H^T + V)^-1,
// which is symmetric and positive semi-definite by construction, evaluated in
// double from the pre-update state and rounded to float.  H is the 2x6
// measurement projection of the barrel (rotation onto the tangent plane,
// KalmanUtilsMPlex.cc:585-591, and z) or the endcap (x, y) update; V the hit
// covariance projected the same way.  No counterpart in the reference: an
// extension for covariances the FP32 form (I - KH)C leaves non-positive.
template <class T>
void joseph_fp64(const T* psErr, const T* psPar, const T* msErr, const T* msPar, bool endcap, T* outErr, int lane) {
  double C[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) C[i][j] = (double)psErr[sym(i, j) * NN + lane];
  const double vxx = msErr[0 * NN + lane], vyx = msErr[1 * NN + lane], vyy = msErr[2 * NN + lane];
  const double vzx = msErr[3 * NN + lane], vzy = msErr[4 * NN + lane], vzz = msErr[5 * NN + lane];
  double h0[6] = {0, 0, 0, 0, 0, 0}, h1[6] = {0, 0, 0, 0, 0, 0}, V00, V01, V11;
  if (endcap) {
    h0[0] = 1; h1[1] = 1;
    V00 = vxx; V01 = vyx; V11 = vyy;
  } else {
    const double mx = msPar[0 * NN + lane], my = msPar[1 * NN + lane];
    const double px = psPar[0 * NN + lane], py = psPar[1 * NN + lane];
    const double r = std::sqrt(mx * mx + my * my);
    const double a00 = -(my + py) / (2 * r), a01 = (mx + px) / (2 * r);
    h0[0] = a00; h0[1] = a01; h1[2] = 1;
    V00 = a00 * a00 * vxx + 2 * a00 * a01 * vyx + a01 * a01 * vyy;
    V01 = a00 * vzx + a01 * vzy;
    V11 = vzz;
  }
  double CH0[6], CH1[6];
  for (int i = 0; i < 6; ++i) {
    CH0[i] = 0; CH1[i] = 0;
    for (int k = 0; k < 6; ++k) { CH0[i] += C[i][k] * h0[k]; CH1[i] += C[i][k] * h1[k]; }
  }
  double S00 = V00, S01 = V01, S11 = V11;
  for (int k = 0; k < 6; ++k) { S00 += h0[k] * CH0[k]; S01 += h0[k] * CH1[k]; S11 += h1[k] * CH1[k]; }
  const double det = S00 * S11 - S01 * S01;
  const double i00 = S11 / det, i01 = -S01 / det, i11 = S00 / det;
  double K0[6], K1[6], A[6][6], AC[6][6];
  for (int i = 0; i < 6; ++i) {
    K0[i] = CH0[i] * i00 + CH1[i] * i01;
    K1[i] = CH0[i] * i01 + CH1[i] * i11;
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) A[i][j] = (i == j ? 1.0 : 0.0) - K0[i] * h0[j] - K1[i] * h1[j];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      AC[i][j] = 0;
      for (int k = 0; k < 6; ++k) AC[i][j] += A[i][k] * C[k][j];
    }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      double v = K0[i] * (V00 * K0[j] + V01 * K1[j]) + K1[i] * (V01 * K0[j] + V11 * K1[j]);
      for (int k = 0; k < 6; ++k) v += AC[i][k] * A[j][k];
      outErr[sym(i, j) * NN + lane] = T(v);
    }
}

// One NN block of the fit: BkFitFitTracks-style per-layer loop with per-lane
// hit counts, per-lane barrel/endcap choice (FitTracksSteered generalised to a
// per-lane pattern: both operations run where a block mixes them, each lane
// keeps its own) and the fake-hit rule for missing layers.
template <class T, class B>
void fit_block(const Ctx& ctx_in, const B& bt, int t0, int cnt, float* tie_out = nullptr) {
  const int n = bt.n_tracks;
  Ctx ctx = ctx_in;
  alignas(64) float tie_r[NN], tie_dot[NN], tie_mat[NN];
  if (tie_out) {
    for (int l = 0; l < NN; ++l) tie_r[l] = tie_dot[l] = tie_mat[l] = 1e30f;
    ctx.tie_r = tie_r;
    ctx.tie_dot = tie_dot;
    ctx.tie_mat = tie_mat;
  }
  alignas(64) T err[21 * NN], par[6 * NN], chi2[NN];
  alignas(64) T errP[21 * NN], parP[6 * NN], errU[21 * NN], parU[6 * NN], c2[NN];
  alignas(64) T errE[21 * NN], parE[6 * NN], c2E[NN], errPE[21 * NN], parPE[6 * NN];
  alignas(64) T msPar[3 * NN], msErr[6 * NN], tgt[NN];
  alignas(64) int chg[NN], nh[NN], nomat[NN], fail[NN];
  unsigned char fl[NN];
  pack(err, bt.err, 21, n, t0, cnt);
  pack(par, bt.par, 6, n, t0, cnt);
  pack(chi2, bt.chi2, 1, n, t0, cnt);
  pack(chg, bt.chg, 1, n, t0, cnt);
  int maxh = 0;
  for (int l = 0; l < NN; ++l) {
    nh[l] = l < cnt ? bt.n_hits[t0 + l] : 0;
    maxh = std::max(maxh, nh[l]);
  }
  maxh = std::min(maxh, bt.n_slots);
  for (int s = 0; s < maxh; ++s) {
    const float* hp = bt.hit_pos + (size_t)s * 3 * n;
    const float* he = bt.hit_err + (size_t)s * 6 * n;
    bool any_b = false, any_e = false;
    for (int l = 0; l < NN; ++l) {
      const bool act = l < cnt && s < nh[l];
      const int t = t0 + (l < cnt ? l : 0);
      fl[l] = bt.hit_flags ? bt.hit_flags[(size_t)s * n + t] : 0;
      if (!act) fl[l] |= 0x80;
      if ((fl[l] & ORC_HIT_MISSING) || !act) {            // MkFinder.cc:1635-1647
        nomat[l] = 1;
        for (int i = 0; i < 3; ++i) msPar[i * NN + l] = par[i * NN + l];
        const T fake[6] = {666, 0, 666, 0, 0, 666};
        for (int i = 0; i < 6; ++i) msErr[i * NN + l] = fake[i];
      } else {
        nomat[l] = 0;
        for (int i = 0; i < 3; ++i) msPar[i * NN + l] = T(hp[(size_t)i * n + t]);
        for (int i = 0; i < 6; ++i) msErr[i * NN + l] = T(he[(size_t)i * n + t]);
      }
      if (act) {
        if (fl[l] & ORC_HIT_ENDCAP) any_e = true; else any_b = true;
      }
    }
    const int kfop = KFO_Calculate_Chi2 | KFO_Update_Params;
    if (any_b) {
      for (int l = 0; l < NN; ++l) tgt[l] = T(std::hypot(msPar[0 * NN + l], msPar[1 * NN + l]));   // MkBase.h:55
      propagate_helix_to_r(err, par, chg, tgt, errP, parP, cnt, bt.pflags, nomat, fail, ctx);
      kalman_operation(kfop, errP, parP, msErr, msPar, errU, parU, c2);
    }
    if (any_e) {
      for (int l = 0; l < NN; ++l) tgt[l] = msPar[2 * NN + l];                                      // MkBase.h:84
      propagate_helix_to_z(err, par, chg, tgt, errPE, parPE, cnt, bt.pflags, nomat, ctx);
      kalman_operation_endcap(kfop, errPE, parPE, msErr, msPar, errE, parE, c2E);
    }
    for (int l = 0; l < cnt; ++l) {
      if (fl[l] & 0x80) continue;
      const bool ec = fl[l] & ORC_HIT_ENDCAP;
      T* se = ec ? errE : errU;
      const T* sp = ec ? parE : parU;
      if ((bt.fit_flags & ORC_FIT_FP64_FALLBACK) && ill_conditioned(se, l))
        joseph_fp64(ec ? errPE : errP, ec ? parPE : parP, msErr, msPar, ec, se, l);
      for (int i = 0; i < 21; ++i) err[i * NN + l] = se[i * NN + l];
      for (int i = 0; i < 6; ++i) par[i * NN + l] = sp[i * NN + l];
      if (bt.fit_flags & ORC_FIT_CHI2) chi2[l] += ec ? c2E[l] : c2[l];                              // MkFinder.cc:1740
      if ((bt.fit_flags & ORC_FIT_CHARGE_FLIP) && par[3 * NN + l] < 0) {                             // MkFinder.cc:1671-1678
        chg[l] = -chg[l];
        par[3 * NN + l] = -par[3 * NN + l];
      }
    }
  }
  if (bt.fit_flags & ORC_FIT_PROP_TO_PCA) {                                   // MkBase.h:91-104
    for (int l = 0; l < NN; ++l) {
      const T slope = std::tan(par[5 * NN + l]);
      tgt[l] = (slope * (slope * par[2 * NN + l] - std::hypot(par[0 * NN + l], par[1 * NN + l]))) / (1 + slope * slope);
    }
    propagate_helix_to_z(err, par, chg, tgt, errP, parP, cnt, bt.pca_pflags, nullptr, ctx);
    std::memcpy(err, errP, sizeof(err));
    std::memcpy(par, parP, sizeof(par));
  }
  unpack(bt.err, err, 21, n, t0, cnt);
  unpack(bt.par, par, 6, n, t0, cnt);
  unpack(bt.chi2, chi2, 1, n, t0, cnt);
  unpack(bt.chg, chg, 1, n, t0, cnt);
  if (tie_out) {
    unpack(tie_out, tie_r, 1, n, t0, cnt);
    unpack(tie_out + n, tie_dot, 1, n, t0, cnt);
    unpack(tie_out + 2 * (size_t)n, tie_mat, 1, n, t0, cnt);
  }
}

template <class T, class B>
int fit_batch(const orc_ctx* c, const B* b, int n_threads, float* tie_out = nullptr, float tie_shift = 0.f) {
  if (!b || b->n_tracks < 0 || b->n_slots < 0) return 1;
  Ctx ctx = make_ctx(c);
  if (tie_shift != 0.f) ctx.tie_shift = tie_shift;
  for_each_block(b->n_tracks, n_threads, [&](int blk) {
    const int t0 = blk * NN, cnt = std::min(NN, b->n_tracks - t0);
    fit_block<T>(ctx, *b, t0, cnt, tie_out);
  });
  return 0;
}

}  // namespace

extern "C" {

int orc_simd_width(void) { return NN; }

int orc_propagate_to_r_f32(const orc_ctx* c, int n, const float* ie, const float* ip, const int* ch, const float* r,
                           float* oe, float* op, int pf, const int* nm, int* fo) {
  return propagate_r_batch<float>(c, n, ie, ip, ch, r, oe, op, pf, nm, fo);
}
int orc_propagate_to_r_f64(const orc_ctx* c, int n, const double* ie, const double* ip, const int* ch,
                           const double* r, double* oe, double* op, int pf, const int* nm, int* fo) {
  return propagate_r_batch<double>(c, n, ie, ip, ch, r, oe, op, pf, nm, fo);
}
int orc_propagate_to_z_f32(const orc_ctx* c, int n, const float* ie, const float* ip, const int* ch, const float* z,
                           float* oe, float* op, int pf, const int* nm) {
  return propagate_z_batch<float>(c, n, ie, ip, ch, z, oe, op, pf, nm);
}
int orc_propagate_to_z_f64(const orc_ctx* c, int n, const double* ie, const double* ip, const int* ch,
                           const double* z, double* oe, double* op, int pf, const int* nm) {
  return propagate_z_batch<double>(c, n, ie, ip, ch, z, oe, op, pf, nm);
}
int orc_helix_at_r_f32(const orc_ctx* c, int n, const float* ip, const int* ch, const float* r, float* op, float* J,
                       int* fo, int pf) {
  return helix_r_batch<float>(c, n, ip, ch, r, op, J, fo, pf);
}
int orc_helix_at_r_f64(const orc_ctx* c, int n, const double* ip, const int* ch, const double* r, double* op,
                       double* J, int* fo, int pf) {
  return helix_r_batch<double>(c, n, ip, ch, r, op, J, fo, pf);
}
int orc_helix_at_z_f32(const orc_ctx* c, int n, const float* ip, const int* ch, const float* z, float* op, float* J,
                       int pf) {
  return helix_z_batch<float>(c, n, ip, ch, z, op, J, pf);
}
int orc_helix_at_z_f64(const orc_ctx* c, int n, const double* ip, const int* ch, const double* z, double* op,
                       double* J, int pf) {
  return helix_z_batch<double>(c, n, ip, ch, z, op, J, pf);
}
int orc_kalman_op_f32(int op, int ec, int n, const float* pe, const float* pp, const float* me, const float* mp,
                      float* oe, float* opar, float* c2) {
  return kalman_batch<float>(op, ec, n, pe, pp, me, mp, oe, opar, c2);
}
int orc_kalman_op_f64(int op, int ec, int n, const double* pe, const double* pp, const double* me, const double* mp,
                      double* oe, double* opar, double* c2) {
  return kalman_batch<double>(op, ec, n, pe, pp, me, mp, oe, opar, c2);
}
int orc_fit_f32(const orc_ctx* c, const orc_fit_batch_f32* b, int nt) { return fit_batch<float>(c, b, nt); }
int orc_fit_f64(const orc_ctx* c, const orc_fit_batch_f64* b, int nt) { return fit_batch<double>(c, b, nt); }
// The FP32 fit, also returning per track how close its steps came to their thresholds: ties [3][n_tracks] =
// Ctx::tie_r, Ctx::tie_dot, Ctx::tie_mat (1e30 where a track never met one).
int orc_fit_ties_f32(const orc_ctx* c, const orc_fit_batch_f32* b, int nt, float* ties) {
  return ties ? fit_batch<float>(c, b, nt, ties) : 1;
}
// The FP32 fit with the helix step's 1e-4 cm threshold moved by tie_shift float spacings of r: the result on
// the other side of every threshold tie within that distance (Ctx::tie_shift).
int orc_fit_tie_shift_f32(const orc_ctx* c, const orc_fit_batch_f32* b, int nt, float tie_shift) {
  return fit_batch<float>(c, b, nt, nullptr, tie_shift);
}
// Sets the threshold shift of every later call (fits and finders; 0 restores the reference's test).  Not
// thread-safe: the checker sets it between calls.
int orc_set_tie_shift(float tie_shift) {
  g_tie_shift = tie_shift;
  return 0;
}

}  // extern "C"
