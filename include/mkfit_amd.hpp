// mkfit_amd.hpp -- C++ host interface over the C ABI (mkfit_amd.h), shaped like
// the reference's MkFitter / MkFinder surface so a caller's code keeps its
// structure: pack tracks, fit or find, unpack.
//
//   MkFitter::InputTracksForFit / FitTracks / OutputFittedTracks  (MkFitter.h:57-88)
//   EventOfHits::SuckInHits                                       (HitStructures.h:280-319)
//   MkFinder::FindTracksBestHit, FindTracksCloneEngine,
//            OutputTracksAndHitIdx                                 (MkBuilder.cc:1362-2060)
//   MkFinder per-layer steps: PropagateTracksToR/Z, SelectHitIndices,
//            AddBestHit, FindCandidatesCloneEngine, UpdateWithLastHit
//                                                                  (MkBase.h:14-106,
//                                                                   MkFinder.h:126-228)
//
// Records mirror the reference's: TrackState (Track.h:74-143: x, y, z, 1/pT,
// phi, theta; packed 6x6 covariance; charge), Hit (Hit.h:165-183: position and
// packed 3x3 covariance) and HitOnTrack (Hit.h:314-327).  Where the reference
// works on NN-track Matriplex blocks, these classes hold the whole batch in
// device memory in the same element-major order.  Errors throw mkfit_amd::Error
// carrying mkf_last_error().  Header-only; link libmkfit_amd.so and the HIP
// runtime (libamdhip64).
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mkfit_amd.h"

namespace mkfit_amd {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(int rc) {
  if (rc != MKF_OK) throw Error(std::string("mkfit_amd: ") + mkf_last_error());
}
inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

struct TrackState {
  float par[6] = {0, 0, 0, 0, 0, 0};
  float err[21] = {};
  int charge = 0;
};
struct Hit {
  float pos[3] = {0, 0, 0};
  float err[6] = {0, 0, 0, 0, 0, 0};
};
struct HitOnTrack {
  int index = -1;
  int layer = -1;
};
struct Track {
  TrackState state;
  float chi2 = 0.f;
  float score = 0.f;
  int n_found = 0;
  std::vector<HitOnTrack> hots;
};

// Owning device array.
template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  ~DeviceArray() { release(); }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) check_hip(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
    n_ = n;
  }
  void upload(const std::vector<T>& v, hipStream_t s) {
    resize(v.size());
    if (n_) check_hip(hipMemcpyAsync(p_, v.data(), n_ * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
  void download(std::vector<T>& v, hipStream_t s) const {
    v.resize(n_);
    if (n_) check_hip(hipMemcpyAsync(v.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost, s), "download");
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  void swap(DeviceArray& o) {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
  }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
};

class Context {
 public:
  // rl_grid / xi_grid: host [300][120] material grids (Config::RlgridME / XigridME) or nullptr.
  explicit Context(float bfield = 3.8112f, const float* rl_grid = nullptr, const float* xi_grid = nullptr,
                   int device = 0) {
    mkf_config cfg{device, bfield, rl_grid, xi_grid};
    check(mkf_context_create(&cfg, &h_));
    check_hip(hipSetDevice(device), "hipSetDevice");
    check_hip(hipStreamCreate(&stream_), "hipStreamCreate");
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  ~Context() {
    if (stream_) (void)hipStreamDestroy(stream_);
    mkf_context_destroy(h_);
  }
  mkf_context* get() const { return h_; }
  hipStream_t stream() const { return stream_; }
  void sync() const { check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

 private:
  mkf_context* h_ = nullptr;
  hipStream_t stream_ = nullptr;
};

// ---- the fit ------------------------------------------------------------------

class MkFitter {
 public:
  explicit MkFitter(Context& c) : c_(c) {}

  // Tracks [beg, end) with, per track, its hits in fit order.  A hit with
  // index < 0 in hots is missing (fake hit, MKF_HIT_MISSING); endcap[l] marks
  // layers fitted with the endcap update (LayerInfo::is_barrel() false).
  void InputTracksForFit(const std::vector<Track>& tracks, const std::vector<std::vector<Hit>>& layer_hits,
                         const std::vector<bool>& endcap, int beg, int end) {
    n_ = end - beg;
    slots_ = 0;
    for (int i = beg; i < end; ++i) slots_ = std::max<int>(slots_, (int)tracks[i].hots.size());
    const size_t n = n_, S = slots_;
    std::vector<float> par(6 * n), err(21 * n), hp(3 * n * S, 0.f), he(6 * n * S, 0.f);
    std::vector<int> chg(n), nh(n);
    std::vector<unsigned char> fl(n * S, 0);
    for (size_t t = 0; t < n; ++t) {
      const Track& tr = tracks[beg + t];
      for (int e = 0; e < 6; ++e) par[e * n + t] = tr.state.par[e];
      for (int e = 0; e < 21; ++e) err[e * n + t] = tr.state.err[e];
      chg[t] = tr.state.charge;
      nh[t] = (int)tr.hots.size();
      for (size_t s = 0; s < tr.hots.size(); ++s) {
        const HitOnTrack& h = tr.hots[s];
        unsigned char f = (h.layer >= 0 && h.layer < (int)endcap.size() && endcap[h.layer]) ? MKF_HIT_ENDCAP : 0;
        if (h.index < 0 || h.layer < 0) {
          f |= MKF_HIT_MISSING;
        } else {
          const Hit& hit = layer_hits[h.layer][h.index];
          for (int e = 0; e < 3; ++e) hp[(s * 3 + e) * n + t] = hit.pos[e];
          for (int e = 0; e < 6; ++e) he[(s * 6 + e) * n + t] = hit.err[e];
        }
        fl[s * n + t] = f;
      }
    }
    hipStream_t st = c_.stream();
    seed_par_.upload(par, st);
    seed_err_.upload(err, st);
    seed_chg_.upload(chg, st);
    hit_pos_.upload(hp, st);
    hit_err_.upload(he, st);
    n_hits_.upload(nh, st);
    hit_flags_.upload(fl, st);
    par_.resize(6 * n);
    err_.resize(21 * n);
    chg_.resize(n);
    chi2_.resize(n);
    c_.sync();
  }

  // batched = one persistent kernel (mkf_fit_batched); otherwise one launch per
  // hit slot as the reference's per-layer loop (mkf_fit_plain).  Same results.
  void FitTracks(int pflags, int fit_flags = MKF_FIT_CHI2, bool batched = true) {
    mkf_fit_batch b{};
    b.n_tracks = n_;
    b.n_slots = slots_;
    b.par = par_.get(); b.err = err_.get(); b.chg = chg_.get(); b.chi2 = chi2_.get();
    b.hit_pos = hit_pos_.get(); b.hit_err = hit_err_.get(); b.n_hits = n_hits_.get(); b.hit_flags = hit_flags_.get();
    b.pflags = pflags;
    b.fit_flags = fit_flags;
    b.seed_par = seed_par_.get(); b.seed_err = seed_err_.get(); b.seed_chg = seed_chg_.get();
    check(batched ? mkf_fit_batched(c_.get(), &b, c_.stream()) : mkf_fit_plain(c_.get(), &b, c_.stream()));
  }

  void OutputFittedTracks(std::vector<Track>& tracks, int beg, int end) {
    std::vector<float> par, err, chi2;
    std::vector<int> chg;
    par_.download(par, c_.stream());
    err_.download(err, c_.stream());
    chg_.download(chg, c_.stream());
    chi2_.download(chi2, c_.stream());
    c_.sync();
    const size_t n = n_;
    for (size_t t = 0; t < n && (int)t < end - beg; ++t) {
      Track& tr = tracks[beg + t];
      for (int e = 0; e < 6; ++e) tr.state.par[e] = par[e * n + t];
      for (int e = 0; e < 21; ++e) tr.state.err[e] = err[e * n + t];
      tr.state.charge = chg[t];
      tr.chi2 = chi2[t];
    }
  }

 private:
  Context& c_;
  int n_ = 0, slots_ = 0;
  DeviceArray<float> seed_par_, seed_err_, hit_pos_, hit_err_, par_, err_, chi2_;
  DeviceArray<int> seed_chg_, n_hits_, chg_;
  DeviceArray<unsigned char> hit_flags_;
};

// ---- finding --------------------------------------------------------------------

class EventOfHits {
 public:
  EventOfHits(Context& c, const std::vector<mkf_layer_desc>& layers) : c_(c), n_layers_((int)layers.size()) {
    check(mkf_event_of_hits_create(c.get(), n_layers_, layers.data(), &h_));
  }
  EventOfHits(const EventOfHits&) = delete;
  EventOfHits& operator=(const EventOfHits&) = delete;
  ~EventOfHits() { mkf_event_of_hits_destroy(h_); }

  // LayerOfHits::SuckInHits for every layer.
  void SuckInHits(const std::vector<std::vector<Hit>>& layer_hits) {
    std::vector<int> off(n_layers_ + 1, 0);
    for (int l = 0; l < n_layers_; ++l) off[l + 1] = off[l] + (int)layer_hits[l].size();
    std::vector<float> flat((size_t)off[n_layers_] * 9);
    for (int l = 0; l < n_layers_; ++l)
      for (size_t i = 0; i < layer_hits[l].size(); ++i) {
        float* d = &flat[((size_t)off[l] + i) * 9];
        std::memcpy(d, layer_hits[l][i].pos, 3 * sizeof(float));
        std::memcpy(d + 3, layer_hits[l][i].err, 6 * sizeof(float));
      }
    hits_.upload(flat, c_.stream());
    check(mkf_event_of_hits_load(h_, off.data(), hits_.get(), c_.stream()));
  }
  mkf_event_of_hits* get() const { return h_; }
  int n_layers() const { return n_layers_; }

 private:
  Context& c_;
  int n_layers_;
  mkf_event_of_hits* h_ = nullptr;
  DeviceArray<float> hits_;
};

class MkFinder {
 public:
  explicit MkFinder(Context& c) : c_(c) {}

  // Seeds [beg, end): state and the seed hits on layers < first_layer.
  void InputTracksAndHitIdx(const std::vector<Track>& seeds, int beg, int end, int first_layer, int n_layers) {
    n_ = end - beg;
    first_ = first_layer;
    L_ = n_layers;
    const size_t n = n_;
    std::vector<float> par(6 * n), err(21 * n);
    std::vector<int> chg(n), hot((size_t)L_ * n, -9);
    for (size_t t = 0; t < n; ++t) {
      const Track& s = seeds[beg + t];
      for (int e = 0; e < 6; ++e) par[e * n + t] = s.state.par[e];
      for (int e = 0; e < 21; ++e) err[e * n + t] = s.state.err[e];
      chg[t] = s.state.charge;
      for (const HitOnTrack& h : s.hots)
        if (h.layer >= 0 && h.layer < first_) hot[(size_t)h.layer * n + t] = h.index;
    }
    hipStream_t st = c_.stream();
    par_.upload(par, st);
    err_.upload(err, st);
    chg_.upload(chg, st);
    hot_seed_.upload(hot, st);
    seed_hits_.resize((size_t)std::max(first_, 1) * n);
    if (n && first_ > 0)
      check_hip(hipMemcpyAsync(seed_hits_.get(), hot_seed_.get(), (size_t)first_ * n * sizeof(int),
                               hipMemcpyDeviceToDevice, st), "copy seed hits");
    c_.sync();
  }

  // MkBuilder::FindTracksBestHit: one track per seed.
  void FindTracksBestHit(EventOfHits& eoh, float chi2_cut, int pflags = 0) {
    const size_t n = n_;
    out_par_.resize(6 * n); out_err_.resize(21 * n); out_chg_.resize(n); out_chi2_.resize(n);
    out_nf_.resize(n); out_hot_.resize((size_t)L_ * n);
    hipStream_t st = c_.stream();
    if (n) {
      check_hip(hipMemcpyAsync(out_par_.get(), par_.get(), 6 * n * 4, hipMemcpyDeviceToDevice, st), "copy");
      check_hip(hipMemcpyAsync(out_err_.get(), err_.get(), 21 * n * 4, hipMemcpyDeviceToDevice, st), "copy");
      check_hip(hipMemcpyAsync(out_chg_.get(), chg_.get(), n * 4, hipMemcpyDeviceToDevice, st), "copy");
      check_hip(hipMemsetAsync(out_chi2_.get(), 0, n * 4, st), "memset");
      check_hip(hipMemcpyAsync(out_hot_.get(), hot_seed_.get(), (size_t)L_ * n * 4, hipMemcpyDeviceToDevice, st),
                "copy");
    }
    mkf_find_batch b{};
    b.n_tracks = n_;
    b.par = out_par_.get(); b.err = out_err_.get(); b.chg = out_chg_.get(); b.chi2 = out_chi2_.get();
    b.hot_idx = out_hot_.get(); b.n_found = out_nf_.get();
    b.first_layer = first_; b.chi2_cut = chi2_cut; b.pflags = pflags;
    check(mkf_find_tracks_best_hit(c_.get(), eoh.get(), &b, st));
    K_ = 1;
  }

  // MkBuilder::FindTracksCloneEngine: up to max_cands candidates per seed.
  void FindTracksCloneEngine(EventOfHits& eoh, float chi2_cut, int max_cands = 6, int pflags = 0,
                             int intra_pflags = 0, int max_holes = 4, int max_consec_holes = 1) {
    const size_t n = n_, K = max_cands;
    out_nc_.resize(n); out_par_.resize(K * 6 * n); out_err_.resize(K * 21 * n); out_chg_.resize(K * n);
    out_chi2_.resize(K * n); out_score_.resize(K * n); out_nf_.resize(K * n); out_hot_.resize(K * L_ * n);
    mkf_ce_batch b{};
    b.n_seeds = n_;
    b.seed_par = par_.get(); b.seed_err = err_.get(); b.seed_chg = chg_.get(); b.seed_hits = seed_hits_.get();
    b.first_layer = first_; b.chi2_cut = chi2_cut; b.pflags = pflags; b.intra_pflags = intra_pflags;
    b.prop_to_hit = 1; b.max_cands = max_cands; b.max_holes = max_holes; b.max_consec_holes = max_consec_holes;
    b.n_cands = out_nc_.get(); b.par = out_par_.get(); b.err = out_err_.get(); b.chg = out_chg_.get();
    b.chi2 = out_chi2_.get(); b.score = out_score_.get(); b.n_found = out_nf_.get(); b.hot_idx = out_hot_.get();
    check(mkf_find_tracks_clone_engine(c_.get(), eoh.get(), &b, c_.stream()));
    K_ = max_cands;
  }

  // ---- per-layer steps, for callers that run the layer loop themselves --------
  // The lanes are the seeds of InputTracksAndHitIdx.  As in the reference,
  // Par[iC] is a lane's current state and Par[iP] its state propagated to the
  // layer (MkBase.h:14-106).

  // Lanes <- seeds: Par[iC] = seed state, chi2 0, NFoundHits = seed hits found.
  void BeginLayerSteps() {
    const size_t n = n_;
    hipStream_t st = c_.stream();
    std::vector<int> hot, nf(n, 0);
    hot_seed_.download(hot, st);
    c_.sync();
    lane_hots_.assign(n, {});
    for (size_t t = 0; t < n; ++t)
      for (int l = 0; l < first_; ++l) {
        const int h = hot[(size_t)l * n + t];
        lane_hots_[t].push_back({h, l});
        nf[t] += h >= 0;
      }
    cur_par_.resize(6 * n); cur_err_.resize(21 * n); prop_par_.resize(6 * n); prop_err_.resize(21 * n);
    lane_chg_.resize(n); lane_chi2_.resize(n);
    if (n) {
      check_hip(hipMemcpyAsync(cur_par_.get(), par_.get(), 6 * n * 4, hipMemcpyDeviceToDevice, st), "copy");
      check_hip(hipMemcpyAsync(cur_err_.get(), err_.get(), 21 * n * 4, hipMemcpyDeviceToDevice, st), "copy");
      check_hip(hipMemcpyAsync(lane_chg_.get(), chg_.get(), n * 4, hipMemcpyDeviceToDevice, st), "copy");
    }
    lane_chi2_.upload(std::vector<float>(n, 0.f), st);
    lane_nf_.upload(nf, st);
    lane_ninside_.upload(std::vector<int>(n, 0), st);
    lane_ntail_.upload(std::vector<int>(n, 0), st);
    xsize_.resize(n); xarr_.resize((size_t)MKF_HIT_IDX_MAX * n); wsr_.resize(n); code_.resize(n);
    c_.sync();
  }

  // MkBase::PropagateTracksToR / ToZ with one target for all lanes: Par[iC] -> Par[iP].
  void PropagateTracksToR(float r, int pflags = 0) {
    target_.upload(std::vector<float>(n_, r), c_.stream());
    check(mkf_propagate_helix_to_r(c_.get(), n_, cur_err_.get(), cur_par_.get(), lane_chg_.get(), target_.get(),
                                   prop_err_.get(), prop_par_.get(), pflags, nullptr, nullptr, c_.stream()));
  }
  void PropagateTracksToZ(float z, int pflags = 0) {
    target_.upload(std::vector<float>(n_, z), c_.stream());
    check(mkf_propagate_helix_to_z(c_.get(), n_, cur_err_.get(), cur_par_.get(), lane_chg_.get(), target_.get(),
                                   prop_err_.get(), prop_par_.get(), pflags, nullptr, c_.stream()));
  }

  // Par[iP] of every lane, element-major [6][n].
  void PropagatedParameters(std::vector<float>& par) {
    prop_par_.download(par, c_.stream());
    c_.sync();
  }

  // Hit candidates of every lane on `layer` from Par[iP].
  void SelectHitIndices(EventOfHits& eoh, int layer) {
    check(mkf_select_hit_indices(eoh.get(), layer, n_, prop_par_.get(), prop_err_.get(), xsize_.get(), xarr_.get(),
                                 wsr_.get(), c_.stream()));
  }

  // Par[iP] -> Par[iC] with the best hit or a fake hit; returns each lane's
  // HitOnTrack index for the layer (-9 when none was added).
  const std::vector<int>& AddBestHit(EventOfHits& eoh, int layer, float chi2_cut, bool prop_to_hit = false,
                                     int intra_pflags = 0) {
    mkf_finder_lanes ln = lanes(true);
    check(mkf_add_best_hit(c_.get(), eoh.get(), layer, &ln, xsize_.get(), xarr_.get(), wsr_.get(), chi2_cut,
                           prop_to_hit, intra_pflags, code_.get(), c_.stream()));
    code_.download(codes_, c_.stream());
    c_.sync();
    for (int t = 0; t < n_; ++t)
      if (codes_[t] != -9) lane_hots_[t].push_back({codes_[t], layer});
    return codes_;
  }

  // IdxChi2List records of every lane for the layer, from Par[iP].
  void FindCandidatesCloneEngine(EventOfHits& eoh, int layer, float chi2_cut,
                                 std::vector<std::vector<mkf_idx_chi2>>& records, int max_holes = 4,
                                 int max_consec_holes = 1, bool prop_to_hit = true, int intra_pflags = 0) {
    const size_t n = n_;
    n_entries_.resize(n);
    entries_.resize(n * MKF_CE_ENTRIES_PER_LANE);
    mkf_finder_lanes ln = lanes(false);
    check(mkf_find_candidates_clone_engine(c_.get(), eoh.get(), layer, &ln, xsize_.get(), xarr_.get(), wsr_.get(),
                                           chi2_cut, prop_to_hit, intra_pflags, max_holes, max_consec_holes,
                                           n_entries_.get(), entries_.get(), c_.stream()));
    std::vector<int> ne;
    std::vector<mkf_idx_chi2> all;
    n_entries_.download(ne, c_.stream());
    entries_.download(all, c_.stream());
    c_.sync();
    records.assign(n, {});
    for (size_t t = 0; t < n; ++t)
      records[t].assign(all.begin() + t * MKF_CE_ENTRIES_PER_LANE,
                        all.begin() + t * MKF_CE_ENTRIES_PER_LANE + ne[t]);
  }

  // Par[iC] updated with hit hit_idx[lane] of the layer (lanes with < 0 keep theirs).
  void UpdateWithLastHit(EventOfHits& eoh, int layer, const std::vector<int>& hit_idx, bool prop_to_hit = true,
                         int pflags = 0) {
    target_idx_.upload(hit_idx, c_.stream());
    mkf_finder_lanes ln = lanes(false);
    ln.prop_par = cur_par_.get(); ln.prop_err = cur_err_.get();
    ln.par = prop_par_.get(); ln.err = prop_err_.get();
    check(mkf_update_with_last_hit(c_.get(), eoh.get(), layer, &ln, target_idx_.get(), prop_to_hit, pflags,
                                   c_.stream()));
    cur_par_.swap(prop_par_);
    cur_err_.swap(prop_err_);
  }

  // Each lane's current state, chi2, found-hit count and HitOnTrack list.
  void OutputLanes(std::vector<Track>& tracks) {
    std::vector<float> par, err, chi2;
    std::vector<int> chg, nf;
    hipStream_t st = c_.stream();
    cur_par_.download(par, st); cur_err_.download(err, st); lane_chg_.download(chg, st);
    lane_chi2_.download(chi2, st); lane_nf_.download(nf, st);
    c_.sync();
    const size_t n = n_;
    tracks.assign(n, Track());
    for (size_t t = 0; t < n; ++t) {
      Track& tr = tracks[t];
      for (int e = 0; e < 6; ++e) tr.state.par[e] = par[e * n + t];
      for (int e = 0; e < 21; ++e) tr.state.err[e] = err[e * n + t];
      tr.state.charge = chg[t];
      tr.chi2 = chi2[t];
      tr.n_found = nf[t];
      tr.hots = lane_hots_[t];
    }
  }

  // The best candidate of every seed (OutputTracksAndHitIdx / the first
  // candidate after MergeCandsAndBestShortOne), hits in layer order.
  void OutputTracksAndHitIdx(std::vector<Track>& tracks) {
    std::vector<float> par, err, chi2, score;
    std::vector<int> chg, nf, hot;
    hipStream_t st = c_.stream();
    out_par_.download(par, st); out_err_.download(err, st); out_chg_.download(chg, st);
    out_chi2_.download(chi2, st); out_nf_.download(nf, st); out_hot_.download(hot, st);
    if (K_ > 1) out_score_.download(score, st);
    c_.sync();
    const size_t n = n_;
    tracks.assign(n, Track());
    for (size_t t = 0; t < n; ++t) {
      Track& tr = tracks[t];
      for (int e = 0; e < 6; ++e) tr.state.par[e] = par[e * n + t];
      for (int e = 0; e < 21; ++e) tr.state.err[e] = err[e * n + t];
      tr.state.charge = chg[t];
      tr.chi2 = chi2[t];
      tr.n_found = nf[t];
      tr.score = K_ > 1 ? score[t] : 0.f;
      for (int l = 0; l < L_; ++l) {
        const int h = hot[(size_t)l * n + t];
        if (h != -9) tr.hots.push_back({h, l});
      }
    }
  }

 private:
  mkf_finder_lanes lanes(bool writes_current) {
    mkf_finder_lanes ln{};
    ln.n = n_;
    ln.prop_par = prop_par_.get(); ln.prop_err = prop_err_.get();
    ln.par = writes_current ? cur_par_.get() : nullptr;
    ln.err = writes_current ? cur_err_.get() : nullptr;
    ln.chg = lane_chg_.get(); ln.chi2 = lane_chi2_.get(); ln.n_found = lane_nf_.get();
    ln.n_inside_minus_one = lane_ninside_.get(); ln.n_tail_minus_one = lane_ntail_.get();
    return ln;
  }

  Context& c_;
  int n_ = 0, first_ = 0, L_ = 0, K_ = 1;
  DeviceArray<float> par_, err_, out_par_, out_err_, out_chi2_, out_score_;
  DeviceArray<int> chg_, hot_seed_, seed_hits_, out_chg_, out_nf_, out_hot_, out_nc_;
  // per-layer step state
  DeviceArray<float> cur_par_, cur_err_, prop_par_, prop_err_, lane_chi2_, target_;
  DeviceArray<int> lane_chg_, lane_nf_, lane_ninside_, lane_ntail_, xsize_, xarr_, wsr_, code_, n_entries_,
      target_idx_;
  DeviceArray<mkf_idx_chi2> entries_;
  std::vector<int> codes_;
  std::vector<std::vector<HitOnTrack>> lane_hots_;
};

}  // namespace mkfit_amd
