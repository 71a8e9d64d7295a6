// oracle_align.cc -- the fit as a function of per-layer material and alignment
// (the model whose gradient the adjoint kernel computes).
//
// TEST INFRASTRUCTURE ONLY (checker).  The fit is the reference's chi2-
// accumulating barrel fit (MkFinder::BkFitFitTracks, MkFinder.cc:1612-1740,
// restated in oracle.cc), one hit per layer in layer order, with two inputs
// made explicit parameters theta[l] = (x/X0, d_rphi, d_z) of layer l:
//   * material: the propagation to layer l's hit uses radL = x/X0 and
//     xi = xi_per_x0 * x/X0 in place of the grid lookup (PropagationMPlex.cc:446-468);
//   * alignment: layer l's hits are moved by d_rphi along the local r-phi
//     direction of the hit and by d_z along z before the propagation to the
//     hit radius and the update:
//       x' = x - d_rphi * y / r_h,  y' = y + d_rphi * x / r_h,  z' = z + d_z,
//     with r_h = hypot(x, y) of the stored hit.
// The loss is the sum over tracks and layers of the update chi2.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"
#include "oracle_core.h"

using namespace orc;

namespace {

template <class T>
void align_block(const Ctx& ctx, const orc_align_batch& b, const T* theta, int t0, int cnt, double* chi2_out) {
  const int n = b.n_tracks;
  alignas(64) T err[21 * NN], par[6 * NN], errP[21 * NN], parP[6 * NN], errU[21 * NN], parU[6 * NN];
  alignas(64) T msPar[3 * NN], msErr[6 * NN], rad[NN], rl[NN], xi[NN], c2[NN];
  alignas(64) int chg[NN], fail[NN], nh[NN];
  double chi2[NN];
  for (int l = 0; l < NN; ++l) {
    const int t = t0 + (l < cnt ? l : 0);
    for (int e = 0; e < 6; ++e) par[e * NN + l] = T(b.seed_par[(size_t)e * n + t]);
    for (int e = 0; e < 21; ++e) err[e * NN + l] = T(b.seed_err[(size_t)e * n + t]);
    chg[l] = b.seed_chg[t];
    nh[l] = l < cnt ? std::min(b.n_hits[t], b.n_layers) : 0;
    chi2[l] = 0;
  }
  int maxh = 0;
  for (int l = 0; l < NN; ++l) maxh = std::max(maxh, nh[l]);
  for (int s = 0; s < maxh; ++s) {
    const float* hp = b.hit_pos + (size_t)s * 3 * n;
    const float* he = b.hit_err + (size_t)s * 6 * n;
    const T m = theta[3 * s + 0], drphi = theta[3 * s + 1], dz = theta[3 * s + 2];
    for (int l = 0; l < NN; ++l) {
      const int t = t0 + (l < cnt ? l : 0);
      const T x = T(hp[t]), y = T(hp[(size_t)n + t]), z = T(hp[(size_t)2 * n + t]);
      const T rh = std::hypot(x, y);
      msPar[0 * NN + l] = x - drphi * (y / rh);
      msPar[1 * NN + l] = y + drphi * (x / rh);
      msPar[2 * NN + l] = z + dz;
      for (int e = 0; e < 6; ++e) msErr[e * NN + l] = T(he[(size_t)e * n + t]);
      rad[l] = std::hypot(msPar[0 * NN + l], msPar[1 * NN + l]);
      rl[l] = m;
      xi[l] = T(b.xi_per_x0) * m;
    }
    propagate_helix_to_r(err, par, chg, rad, errP, parP, cnt, b.pflags, nullptr, fail, ctx, rl, xi);
    kalman_operation(KFO_Calculate_Chi2 | KFO_Update_Params, errP, parP, msErr, msPar, errU, parU, c2);
    for (int l = 0; l < cnt; ++l) {
      if (s >= nh[l]) continue;
      for (int e = 0; e < 21; ++e) err[e * NN + l] = errU[e * NN + l];
      for (int e = 0; e < 6; ++e) par[e * NN + l] = parU[e * NN + l];
      chi2[l] += (double)c2[l];
    }
  }
  for (int l = 0; l < cnt; ++l) chi2_out[t0 + l] = chi2[l];
}

template <class T>
int align_chi2(const orc_ctx* c, const orc_align_batch* b, const T* theta, double* chi2_out, int n_threads) {
  if (!b || !theta || !chi2_out || b->n_tracks < 0 || b->n_layers <= 0) return 1;
  Ctx ctx;
  ctx.bfield = c ? c->bfield : 3.8112f;
  ctx.rl_grid = nullptr;
  ctx.xi_grid = nullptr;
  const int n = b->n_tracks;
  const int nblk = (n + NN - 1) / NN;
  std::atomic<int> next{0};
  auto work = [&]() {
    for (;;) {
      const int blk = next.fetch_add(1);
      if (blk >= nblk) break;
      align_block<T>(ctx, *b, theta, blk * NN, std::min(NN, n - blk * NN), chi2_out);
    }
  };
  n_threads = std::max(1, n_threads);
  std::vector<std::thread> pool;
  for (int i = 1; i < n_threads; ++i) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  return 0;
}

}  // namespace

extern "C" int orc_align_chi2_f32(const orc_ctx* c, const orc_align_batch* b, const float* theta, double* chi2,
                                  int n_threads) {
  return align_chi2<float>(c, b, theta, chi2, n_threads);
}
extern "C" int orc_align_chi2_f64(const orc_ctx* c, const orc_align_batch* b, const double* theta, double* chi2,
                                  int n_threads) {
  return align_chi2<double>(c, b, theta, chi2, n_threads);
}
