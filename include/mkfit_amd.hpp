// mkfit_amd.hpp -- C++ host interface over the C ABI (mkfit_amd.h) with the
// reference's records and call signatures, so a reference caller links
// against it with its own code unchanged where the signatures below match:
//
//   records    Track, TrackState, Hit, HitOnTrack: the x86-64 memory images of
//              the reference's classes (Track.h:74-143, 360-367, 619-620;
//              Hit.h:147-160, 181-182, 251-260, 301-327), checked below with
//              static_assert, with the reference's accessors
//   MkFitter   InputTracksAndHits / InputTracksForFit / FitTracksWithInterSlurp /
//              FitTracks / FitTracksSteered / OutputTracks / OutputFittedTracks
//              (MkFitter.h:60-87)
//   MkFinder   Setup, InputTracksAndHitIdx / OutputTracksAndHitIdx
//              (MkFinder.h:128-150), PropagateTracksToR / ToZ (MkBase.h:35-76),
//              SelectHitIndices, AddBestHit, FindCandidatesCloneEngine,
//              UpdateWithLastHit, BkFitInputTracks / BkFitFitTracks[BH] /
//              BkFitPropTracksToPCA / BkFitOutputTracks (MkFinder.h:183-228)
//   MkBuilder  find_tracks_load_seeds, FindTracksBestHit / Standard /
//              CloneEngine, BackwardFitBH / BackwardFit, export_tracks /
//              export_best_comb_cands (MkBuilder.h:155-237)
//   Config     the propagation-flag globals of Config.cc:52-58 and includePCA
//
// Where the reference packs NN tracks per Matriplex and a caller loops over
// blocks of NN, these classes take any range [beg, end) and hold it in device
// memory in the same element-major order; one call with the whole range is the
// fast way to use them, a loop over NN-blocks gives the same results.  The
// event's hits live in EventOfHits (device LayerOfHits); the fitter reads the
// hits of an Event's layerHits_.  Errors throw mkfit_amd::Error carrying
// mkf_last_error().  Header-only; link libmkfit_amd.so and the HIP runtime
// (libamdhip64).
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
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
inline void require(bool ok, const char* what) {
  if (!ok) throw Error(std::string("mkfit_amd: ") + what);
}

// ---- the reference's small vector / symmetric matrix storage ---------------------
// SVector<float, D> and SMatrix<float, D, D, MatRepSym> keep D floats and the
// lower triangle row by row, (0,0), (1,0), (1,1), (2,0), ...

template <int D>
struct SVector {
  float fArray[D] = {};
  float operator[](int i) const { return fArray[i]; }
  float& operator[](int i) { return fArray[i]; }
  float At(int i) const { return fArray[i]; }
  float& At(int i) { return fArray[i]; }
  const float* Array() const { return fArray; }
  float* Array() { return fArray; }
};

template <int D>
struct SMatrixSym {
  static constexpr int kSize = D * (D + 1) / 2;
  float fRep[kSize] = {};
  static constexpr int offset(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }
  float operator()(int i, int j) const { return fRep[offset(i, j)]; }
  float& operator()(int i, int j) { return fRep[offset(i, j)]; }
  float At(int i, int j) const { return fRep[offset(i, j)]; }
  float& At(int i, int j) { return fRep[offset(i, j)]; }
  const float* Array() const { return fRep; }
  float* Array() { return fRep; }
};

using SVector3 = SVector<3>;
using SVector6 = SVector<6>;
using SMatrixSym33 = SMatrixSym<3>;
using SMatrixSym66 = SMatrixSym<6>;

// ---- records -------------------------------------------------------------------

// Hit.h:314-327: index in the low 24 bits, layer in the high 8, both signed.
struct HitOnTrack {
  int index : 24;
  int layer : 8;
  HitOnTrack() : index(-1), layer(-1) {}
  HitOnTrack(int i, int l) : index(i), layer(l) {}
  bool operator<(const HitOnTrack o) const { return layer != o.layer ? layer < o.layer : index < o.index; }
};
using HoTVec = std::vector<HitOnTrack>;

// Hit.h:165-303: MeasurementState (position, packed 3x3 covariance), mcHitID,
// PackedData.  Public members so that their offsets can be checked.
class Hit {
 public:
  Hit() = default;
  Hit(const SVector3& position, const SMatrixSym33& error, int mcHitID = -1)
      : pos_(position), err_(error), mcHitID_(mcHitID) {}

  const float* posArray() const { return pos_.Array(); }
  const float* errArray() const { return err_.Array(); }
  const SVector3& position() const { return pos_; }
  const SVector3& parameters() const { return pos_; }
  const SMatrixSym33& error() const { return err_; }
  float x() const { return pos_[0]; }
  float y() const { return pos_[1]; }
  float z() const { return pos_[2]; }
  float r() const { return std::hypot(pos_[0], pos_[1]); }
  float phi() const { return std::atan2(pos_[1], pos_[0]); }
  float exx() const { return std::sqrt(err_(0, 0)); }
  float eyy() const { return std::sqrt(err_(1, 1)); }
  float ezz() const { return std::sqrt(err_(2, 2)); }
  int mcHitID() const { return mcHitID_; }
  void setMCHitID(int id) { mcHitID_ = id; }
  unsigned packedData() const { return pdata_; }
  void setPackedData(unsigned d) { pdata_ = d; }
  // Hit::PackedData (Hit.h:251-282): detid_in_layer bits 0-11, the charge code bits 12-19.
  unsigned detIDinLayer() const { return pdata_ & 0xfffu; }
  unsigned chargeCode() const { return (pdata_ >> 12) & 0xffu; }
  unsigned chargePerCM() const { return chargeCode() == 0 ? 0u : ((chargeCode() - 1u) << 3) + 1620u; }
  static unsigned minChargePerCM() { return 1620; }
  static unsigned maxChargePerCM() { return 1620 + (0xfe << 3); }

  SVector3 pos_;
  SMatrixSym33 err_;
  int mcHitID_ = -1;
  unsigned pdata_ = 0;
};
using HitVec = std::vector<Hit>;

// Hit.h:333-338: a dead-module region of a layer, in phi and q (z on barrel layers, r on disks).
struct DeadRegion {
  float phi1, phi2, q1, q2;
  DeadRegion(float a1, float a2, float b1, float b2) : phi1(a1), phi2(a2), q1(b1), q2(b2) {}
};
using DeadVec = std::vector<DeadRegion>;

// Track.h:74-143: x, y, z, 1/pT, phi, theta; packed 6x6 covariance; charge.
struct TrackState {
  TrackState() = default;
  TrackState(const SVector6& par, const SMatrixSym66& err, int chg) : parameters(par), errors(err), charge(chg) {}
  float x() const { return parameters[0]; }
  float y() const { return parameters[1]; }
  float z() const { return parameters[2]; }
  float posR() const { return std::hypot(parameters[0], parameters[1]); }
  float invpT() const { return parameters[3]; }
  float momPhi() const { return parameters[4]; }
  float theta() const { return parameters[5]; }
  float pT() const { return std::abs(1.f / parameters[3]); }

  SVector6 parameters;
  SMatrixSym66 errors;
  short charge = 0;
  bool valid = true;
};

// Track.h:145-367 (TrackBase) and 384-620 (Track).
class Track {
 public:
  Track() = default;
  Track(const TrackState& state, float chi2, int label, int nHits, const HitOnTrack* hits)
      : state_(state), chi2_(chi2), label_(label) {
    reserveHits(nHits);
    for (int h = 0; h < nHits; ++h) addHitIdx(hits[h].index, hits[h].layer, 0.f);
  }

  const TrackState& state() const { return state_; }
  void setState(const TrackState& s) { state_ = s; }
  const SVector6& parameters() const { return state_.parameters; }
  const SMatrixSym66& errors() const { return state_.errors; }
  SVector6& parameters_nc() { return state_.parameters; }
  SMatrixSym66& errors_nc() { return state_.errors; }
  TrackState& state_nc() { return state_; }
  const float* posArray() const { return state_.parameters.Array(); }
  const float* errArray() const { return state_.errors.Array(); }

  float x() const { return state_.x(); }
  float y() const { return state_.y(); }
  float z() const { return state_.z(); }
  float posR() const { return state_.posR(); }
  float invpT() const { return state_.invpT(); }
  float momPhi() const { return state_.momPhi(); }
  float theta() const { return state_.theta(); }
  float pT() const { return state_.pT(); }

  int charge() const { return state_.charge; }
  float chi2() const { return chi2_; }
  float score() const { return score_; }
  int label() const { return label_; }
  void setCharge(int chg) { state_.charge = (short)chg; }
  void setChi2(float c) { chi2_ = c; }
  void setScore(float s) { score_ = s; }
  void setLabel(int l) { label_ = l; }
  unsigned getRawStatus() const { return status_; }
  void setRawStatus(unsigned s) { status_ = s; }

  void reserveHits(int n) { hitsOnTrk_.reserve(n); }
  void resetHits() {
    lastHitIdx_ = -1;
    nFoundHits_ = 0;
    hitsOnTrk_.clear();
  }
  void resizeHits(int nHits, int nFoundHits) {
    hitsOnTrk_.resize(nHits);
    lastHitIdx_ = (short)(nHits - 1);
    nFoundHits_ = (short)nFoundHits;
  }
  // Track.h:470-479: an index >= 0 or -9 counts as found and adds its chi2.
  void addHitIdx(int hitIdx, int hitLyr, float chi2) {
    hitsOnTrk_.push_back(HitOnTrack(hitIdx, hitLyr));
    ++lastHitIdx_;
    if (hitIdx >= 0 || hitIdx == -9) {
      ++nFoundHits_;
      chi2_ += chi2;
    }
  }
  void addHitIdx(const HitOnTrack& hot, float chi2) { addHitIdx(hot.index, hot.layer, chi2); }
  HitOnTrack getHitOnTrack(int i) const { return hitsOnTrk_[i]; }
  int getHitIdx(int i) const { return hitsOnTrk_[i].index; }
  int getHitLyr(int i) const { return hitsOnTrk_[i].layer; }
  HitOnTrack getLastHitOnTrack() const { return hitsOnTrk_[lastHitIdx_]; }
  int getLastHitIdx() const { return hitsOnTrk_[lastHitIdx_].index; }
  int getLastHitLyr() const { return hitsOnTrk_[lastHitIdx_].layer; }
  const HitOnTrack* getHitsOnTrackArray() const { return hitsOnTrk_.data(); }
  const HitOnTrack* BeginHitsOnTrack() const { return hitsOnTrk_.data(); }
  const HitOnTrack* EndHitsOnTrack() const { return hitsOnTrk_.data() + (lastHitIdx_ + 1); }
  void setHitIdxLyr(int i, int idx, int lyr) { hitsOnTrk_[i] = HitOnTrack(idx, lyr); }
  int nFoundHits() const { return nFoundHits_; }
  int nTotalHits() const { return lastHitIdx_ + 1; }
  void countAndSetNFoundHits() {
    nFoundHits_ = 0;
    for (int i = 0; i <= lastHitIdx_; ++i) nFoundHits_ += hitsOnTrk_[i].index >= 0 || hitsOnTrk_[i].index == -9;
  }
  // Track.h:553-574: -1 hits after the first found hit (counting from the
  // last), and -1 hits after the last found hit.
  int nInsideMinusOneHits() const {
    int n = 0;
    bool inside = false;
    for (int i = lastHitIdx_; i >= 0; --i) {
      if (hitsOnTrk_[i].index >= 0) inside = true;
      if (inside && hitsOnTrk_[i].index == -1) ++n;
    }
    return n;
  }
  int nTailMinusOneHits() const {
    int n = 0;
    for (int i = lastHitIdx_; i >= 0; --i) {
      if (hitsOnTrk_[i].index >= 0) return n;
      if (hitsOnTrk_[i].index == -1) ++n;
    }
    return n;
  }

  TrackState state_;
  float chi2_ = 0.f;
  float score_ = 0.f;
  short lastHitIdx_ = -1;
  short nFoundHits_ = 0;
  unsigned status_ = 0;
  int label_ = -1;
  HoTVec hitsOnTrk_;
};
using TrackVec = std::vector<Track>;

// The byte images the reference's event files and callers rely on.
static_assert(sizeof(HitOnTrack) == 4, "HitOnTrack is 4 bytes");
static_assert(sizeof(Hit) == 44 && offsetof(Hit, err_) == 12 && offsetof(Hit, mcHitID_) == 36 &&
                  offsetof(Hit, pdata_) == 40, "Hit byte image");
static_assert(sizeof(TrackState) == 112 && offsetof(TrackState, errors) == 24 &&
                  offsetof(TrackState, charge) == 108 && offsetof(TrackState, valid) == 110,
              "TrackState byte image");
static_assert(std::is_standard_layout<Track>::value, "Track is standard-layout");
static_assert(sizeof(Track) == 160 && offsetof(Track, chi2_) == 112 && offsetof(Track, score_) == 116 &&
                  offsetof(Track, lastHitIdx_) == 120 && offsetof(Track, nFoundHits_) == 122 &&
                  offsetof(Track, status_) == 124 && offsetof(Track, label_) == 128 &&
                  offsetof(Track, hitsOnTrk_) == 136,
              "Track byte image");

// ---- configuration ---------------------------------------------------------------

enum PropagationFlagsEnum { PF_none = 0, PF_use_param_b_field = MKF_PF_USE_PARAM_B_FIELD,
                            PF_apply_material = MKF_PF_APPLY_MATERIAL };

// Config.h:25-50.
struct PropagationFlags {
  bool use_param_b_field = false;
  bool apply_material = false;
  PropagationFlags() = default;
  PropagationFlags(int pfe) : use_param_b_field(pfe & PF_use_param_b_field), apply_material(pfe & PF_apply_material) {}
  int raw() const { return (use_param_b_field ? PF_use_param_b_field : 0) | (apply_material ? PF_apply_material : 0); }
};

// The globals of Config.cc:52-58 and Config.cc:109 the hot path reads
// (default-initialised as there; a geometry sets them, e.g.
// Geoms/CylCowWLids.cc:225-231).
namespace Config {
inline bool finding_requires_propagation_to_hit_pos = false;
inline PropagationFlags finding_inter_layer_pflags;
inline PropagationFlags finding_intra_layer_pflags;
inline PropagationFlags backward_fit_pflags;
inline PropagationFlags forward_fit_pflags;
inline PropagationFlags seed_fit_pflags;
inline PropagationFlags pca_prop_pflags;
inline bool includePCA = false;
}  // namespace Config

// IterationConfig.h:93-104: the finding parameters of one iteration.
struct IterationParams {
  int nlayers_per_seed = 3;
  int maxCandsPerSeed = 5;
  int maxHolesPerCand = 4;
  int maxConsecHoles = 1;
  float chi2Cut_min = 15.0f;
  float chi2CutOverlap = 3.5f;
  float pTCutOverlap = 1.0f;
};

// SteeringParams.h:90-163: the layer plan of a region.  The device finders walk
// the layers of an EventOfHits in their order, so a plan here must list those
// layers in order (fill_plan(0, n - 1)); m_bkw_fit_last is the lowest plan
// index the backward fit visits.
struct SteeringParams {
  std::vector<int> m_layer_plan;
  int m_region = 0;
  int m_fwd_search_pickup = 0;
  int m_bkw_fit_last = 0;
  int m_bkw_search_pickup = -1;
  void reserve_plan(int n) { m_layer_plan.reserve(n); }
  void append_plan(int layer) { m_layer_plan.push_back(layer); }
  void fill_plan(int first, int last) {
    for (int i = first; i <= last; ++i) append_plan(i);
  }
  void set_iterator_limits(int fwd_search_pu, int bkw_fit_last, int bkw_search_pu = -1) {
    m_fwd_search_pickup = fwd_search_pu;
    m_bkw_fit_last = bkw_fit_last;
    m_bkw_search_pickup = bkw_search_pu;
  }
};

// Event.h:57-62: the parts of an event the hot path reads.
struct Event {
  std::vector<HitVec> layerHits_;
  TrackVec simTracks_, seedTracks_, candidateTracks_, fitTracks_;
};

// ---- device memory and the context ---------------------------------------------------

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
  void download(std::vector<T>& v, hipStream_t s) const { download(v, n_, s); }
  void download(std::vector<T>& v, size_t count, hipStream_t s) const {
    v.resize(count);
    if (count) check_hip(hipMemcpyAsync(v.data(), p_, count * sizeof(T), hipMemcpyDeviceToHost, s), "download");
  }
  void copy_from(const DeviceArray& o, size_t count, hipStream_t s) {
    if (count)
      check_hip(hipMemcpyAsync(p_, o.p_, count * sizeof(T), hipMemcpyDeviceToDevice, s), "device copy");
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

// Host element-major packing of track states: [6][n], [21][n], [n].
struct PackedStates {
  std::vector<float> par, err;
  std::vector<int> chg;
  void pack(const TrackVec& tracks, int beg, int end) {
    const size_t n = end - beg;
    par.resize(6 * n);
    err.resize(21 * n);
    chg.resize(n);
    for (size_t t = 0; t < n; ++t) {
      const Track& tr = tracks[beg + t];
      for (int e = 0; e < 6; ++e) par[e * n + t] = tr.parameters()[e];
      for (int e = 0; e < 21; ++e) err[e * n + t] = tr.errors().fRep[e];
      chg[t] = tr.charge();
    }
  }
};

inline void unpack_state(Track& tr, const std::vector<float>& par, const std::vector<float>& err, size_t n,
                         size_t t) {
  for (int e = 0; e < 6; ++e) tr.parameters_nc()[e] = par[e * n + t];
  for (int e = 0; e < 21; ++e) tr.errors_nc().fRep[e] = err[e * n + t];
}

// ---- the fit ------------------------------------------------------------------

// MkFitter (MkFitter.h:21-87).  The hit of slot hi of a track is its hi-th
// HitOnTrack, looked up in the layer that HitOnTrack names (the reference reads
// layer hi, MkFitter.cc:79-89, 335-349, and assumes the two agree).  A
// HitOnTrack with index < 0 is fitted as the backward fit's fake hit
// {666, 0, 666, 0, 0, 666} at the track position (MKF_HIT_MISSING); the
// reference's forward fit has no rule for it (MkFitter.cc:343-348).  Chi2 and
// Label pass through as in the reference, whose fit does not accumulate chi2
// (MkFitter.cc:434-436); SetFitFlags(MKF_FIT_CHI2) adds the fit's chi2 sum.
class MkFitter {
 public:
  explicit MkFitter(Context& c, int n_hits = 0) : c_(c), Nhits_(n_hits) {}

  void SetNhits(int nhits) { Nhits_ = nhits; }
  int Nhits() const { return Nhits_; }
  void SetFitFlags(int fit_flags) { fit_flags_ = fit_flags; }
  // true (default): one persistent kernel (mkf_fit_batched); false: one launch
  // per hit slot like the reference's per-layer loop (mkf_fit_plain).  Same results.
  void SetBatched(bool b) { batched_ = b; }

  // MkFitter::InputTracksAndHits (MkFitter.cc:56-91): states and hits.
  void InputTracksAndHits(const TrackVec& tracks, const std::vector<HitVec>& layerHits, int beg, int end) {
    InputTracksForFit(tracks, beg, end);
    pack_hits(layerHits);
  }

  // MkFitter::InputTracksForFit (MkFitter.cc:286-321): states and HitOnTrack
  // lists; the hits are read by the fit call.
  void InputTracksForFit(const TrackVec& tracks, int beg, int end) {
    require(beg >= 0 && beg <= end && end <= (int)tracks.size(), "InputTracksForFit: bad range");
    n_ = end - beg;
    int most = 0;
    for (int i = beg; i < end; ++i) most = std::max(most, tracks[i].nTotalHits());
    slots_ = Nhits_ > 0 ? Nhits_ : most;
    const size_t n = n_;
    in_.pack(tracks, beg, end);
    chi2_in_.resize(n);
    label_.resize(n);
    hots_.assign((size_t)slots_ * n, HitOnTrack());
    n_hits_h_.assign(n, 0);
    for (size_t t = 0; t < n; ++t) {
      const Track& tr = tracks[beg + t];
      chi2_in_[t] = tr.chi2();
      label_[t] = tr.label();
      n_hits_h_[t] = std::min(slots_, tr.nTotalHits());
      for (int s = 0; s < n_hits_h_[t]; ++s) hots_[(size_t)s * n + t] = tr.getHitOnTrack(s);
    }
    hipStream_t st = c_.stream();
    seed_par_.upload(in_.par, st);
    seed_err_.upload(in_.err, st);
    seed_chg_.upload(in_.chg, st);
    n_hits_.upload(n_hits_h_, st);
    par_.resize(6 * n);
    err_.resize(21 * n);
    chg_.resize(n);
    chi2_.resize(n);
    hits_packed_ = false;
    c_.sync();
  }

  // MkFitter::FitTracksWithInterSlurp (MkFitter.cc:323-363): the hits of
  // layersohits, barrel updates, Config::forward_fit_pflags.
  void FitTracksWithInterSlurp(const std::vector<HitVec>& layersohits, int N_proc) {
    pack_hits(layersohits);
    fit(N_proc, nullptr, Config::forward_fit_pflags);
  }

  // MkFitter::FitTracks (MkFitter.cc:429-445): propagate to each hit's r and
  // update (barrel).  The hits are those of InputTracksAndHits, else ev->layerHits_.
  void FitTracks(int N_proc, const Event* ev, PropagationFlags pflags) { fit(N_proc, ev, pflags, nullptr); }

  // MkFitter::FitTracksSteered (MkFitter.cc:470-503): slot hi is a barrel
  // (propagation to r) or an endcap (propagation to z) update by is_barrel[hi].
  void FitTracksSteered(const bool is_barrel[], int N_proc, const Event* ev, PropagationFlags pflags) {
    fit(N_proc, ev, pflags, is_barrel);
  }

  // MkFitter::OutputTracks (MkFitter.cc:505-522) with iCP = iC: parameters,
  // errors, charge, chi2, label of tracks [beg, end), which must exist.
  static constexpr int iC = 0;
  void OutputTracks(TrackVec& tracks, int beg, int end, int iCP) const {
    require(iCP == iC, "OutputTracks: only the updated state (iC) is kept");
    require(beg >= 0 && end - beg == n_ && end <= (int)tracks.size(), "OutputTracks: range differs from the input");
    std::vector<float> par, err, chi2;
    std::vector<int> chg;
    par_.download(par, c_.stream());
    err_.download(err, c_.stream());
    chg_.download(chg, c_.stream());
    chi2_.download(chi2, c_.stream());
    c_.sync();
    const size_t n = n_;
    for (size_t t = 0; t < n; ++t) {
      Track& tr = tracks[beg + t];
      unpack_state(tr, par, err, n, t);
      tr.setCharge(chg[t]);
      tr.setChi2((fit_flags_ & MKF_FIT_CHI2) ? chi2_in_[t] + chi2[t] : chi2_in_[t]);
      tr.setLabel(label_[t]);
    }
  }
  void OutputFittedTracks(TrackVec& tracks, int beg, int end) const { OutputTracks(tracks, beg, end, iC); }

 private:
  void pack_hits(const std::vector<HitVec>& layerHits) {
    const size_t n = n_, S = slots_;
    hp_.assign(3 * n * S, 0.f);
    he_.assign(6 * n * S, 0.f);
    fl_.assign(n * S, 0);
    for (size_t t = 0; t < n; ++t)
      for (int s = 0; s < n_hits_h_[t]; ++s) {
        const HitOnTrack h = hots_[(size_t)s * n + t];
        if (h.index < 0 || h.layer < 0) {
          fl_[s * n + t] = MKF_HIT_MISSING;
          continue;
        }
        require(h.layer < (int)layerHits.size() && h.index < (int)layerHits[h.layer].size(),
                "fit: HitOnTrack names a hit the layer does not have");
        const Hit& hit = layerHits[h.layer][h.index];
        for (int e = 0; e < 3; ++e) hp_[(s * 3 + e) * n + t] = hit.posArray()[e];
        for (int e = 0; e < 6; ++e) he_[(s * 6 + e) * n + t] = hit.errArray()[e];
      }
    hit_pos_.upload(hp_, c_.stream());
    hit_err_.upload(he_, c_.stream());
    c_.sync();
    hits_packed_ = true;
  }

  void fit(int N_proc, const Event* ev, PropagationFlags pflags, const bool* is_barrel = nullptr) {
    require(N_proc == n_, "fit: N_proc must equal the number of tracks input");
    if (!hits_packed_) {
      require(ev != nullptr, "fit: no hits (InputTracksAndHits, or an Event with layerHits_)");
      pack_hits(ev->layerHits_);
    }
    const size_t n = n_;
    std::vector<unsigned char> fl = fl_;
    if (is_barrel)
      for (int s = 0; s < slots_; ++s)
        if (!is_barrel[s])
          for (size_t t = 0; t < n; ++t) fl[s * n + t] |= MKF_HIT_ENDCAP;
    hit_flags_.upload(fl, c_.stream());
    mkf_fit_batch b{};
    b.n_tracks = n_;
    b.n_slots = slots_;
    b.par = par_.get(); b.err = err_.get(); b.chg = chg_.get(); b.chi2 = chi2_.get();
    b.hit_pos = hit_pos_.get(); b.hit_err = hit_err_.get(); b.n_hits = n_hits_.get(); b.hit_flags = hit_flags_.get();
    b.pflags = pflags.raw();
    b.fit_flags = fit_flags_;
    b.seed_par = seed_par_.get(); b.seed_err = seed_err_.get(); b.seed_chg = seed_chg_.get();
    check(batched_ ? mkf_fit_batched(c_.get(), &b, c_.stream()) : mkf_fit_plain(c_.get(), &b, c_.stream()));
    c_.sync();
  }

  Context& c_;
  int Nhits_ = 0, n_ = 0, slots_ = 0, fit_flags_ = 0;
  bool batched_ = true, hits_packed_ = false;
  PackedStates in_;
  std::vector<float> chi2_in_, hp_, he_;
  std::vector<int> label_, n_hits_h_;
  std::vector<unsigned char> fl_;
  std::vector<HitOnTrack> hots_;   // [slot][n]
  DeviceArray<float> seed_par_, seed_err_, hit_pos_, hit_err_, par_, err_, chi2_;
  DeviceArray<int> seed_chg_, n_hits_, chg_;
  DeviceArray<unsigned char> hit_flags_;
};

// ---- event of hits --------------------------------------------------------------------

class EventOfHits;

// LayerOfHits (HitStructures.h:100-276): a handle on one layer of an EventOfHits.
class LayerOfHits {
 public:
  LayerOfHits(EventOfHits* eoh, int layer, const mkf_layer_desc& info) : eoh_(eoh), layer_(layer), info_(info) {}
  int layer_id() const { return layer_; }
  bool is_barrel() const { return info_.is_barrel != 0; }
  const mkf_layer_desc& layer_info() const { return info_; }
  EventOfHits& event_of_hits() const { return *eoh_; }

 private:
  EventOfHits* eoh_;
  int layer_;
  mkf_layer_desc info_;
};

// EventOfHits (HitStructures.h:280-319).
class EventOfHits {
 public:
  EventOfHits(Context& c, const std::vector<mkf_layer_desc>& layers) : c_(c), n_layers_((int)layers.size()) {
    check(mkf_event_of_hits_create(c.get(), n_layers_, layers.data(), &h_));
    for (int l = 0; l < n_layers_; ++l) m_layers_of_hits.emplace_back(this, l, layers[l]);
  }
  EventOfHits(const EventOfHits&) = delete;
  EventOfHits& operator=(const EventOfHits&) = delete;
  ~EventOfHits() { mkf_event_of_hits_destroy(h_); }

  // LayerOfHits::SuckInHits of every layer (HitStructures.cc:77-183).
  void SuckInHits(const std::vector<HitVec>& layer_hits) {
    require((int)layer_hits.size() == n_layers_, "SuckInHits: one HitVec per layer");
    std::vector<int> off(n_layers_ + 1, 0);
    for (int l = 0; l < n_layers_; ++l) off[l + 1] = off[l] + (int)layer_hits[l].size();
    std::vector<float> flat((size_t)off[n_layers_] * 9);
    for (int l = 0; l < n_layers_; ++l)
      for (size_t i = 0; i < layer_hits[l].size(); ++i) {
        float* d = &flat[((size_t)off[l] + i) * 9];
        std::memcpy(d, layer_hits[l][i].posArray(), 3 * sizeof(float));
        std::memcpy(d + 3, layer_hits[l][i].errArray(), 6 * sizeof(float));
      }
    codes_.clear();
    for (int l = 0; l < n_layers_; ++l)
      for (const Hit& h : layer_hits[l]) codes_.push_back((unsigned char)h.chargeCode());
    hits_.upload(flat, c_.stream());
    check(mkf_event_of_hits_load(h_, off.data(), hits_.get(), c_.stream()));
    n_hits_.assign(off.begin(), off.end());
    c_.sync();
  }
  // The iteration's hit mask (MkJob::get_mask_for_layer, MkFinder.cc:476-482):
  // per layer, true for a hit an earlier iteration took, which hit selection
  // then passes over; an empty vector clears the mask.  Set after SuckInHits
  // (which clears it).
  void SetIterationHitMasks(const std::vector<std::vector<bool>>& layer_masks) {
    if (layer_masks.empty()) {
      check(mkf_event_of_hits_set_hit_mask(h_, nullptr, c_.stream()));
      return;
    }
    require((int)layer_masks.size() == n_layers_, "SetIterationHitMasks: one mask per layer");
    std::vector<unsigned char> flat;
    for (int l = 0; l < n_layers_; ++l) {
      require((int)layer_masks[l].size() == n_hits(l), "SetIterationHitMasks: one flag per hit of the layer");
      for (bool b : layer_masks[l]) flat.push_back(b ? 1 : 0);
    }
    mask_.upload(flat, c_.stream());
    check(mkf_event_of_hits_set_hit_mask(h_, mask_.get(), c_.stream()));
    c_.sync();
  }
  // The strip charge cut of FindCandidates[CloneEngine] (passStripChargePCMfromTrack,
  // MkFinder.cc:863-896) over the charge codes the sucked-in hits carry in their
  // packed data (setupAsStrip / setupAsPixel, Hit.h:285-299); off, every hit
  // passes.  Set after SuckInHits (which turns it off).
  void UseStripCharges(bool on) {
    if (!on || codes_.empty()) {
      check(mkf_event_of_hits_set_hit_charge(h_, nullptr, c_.stream()));
      return;
    }
    charge_.upload(codes_, c_.stream());
    check(mkf_event_of_hits_set_hit_charge(h_, charge_.get(), c_.stream()));
    c_.sync();
  }
  // EventOfHits::SuckInDeads (HitStructures.h:310-313): the layer's dead
  // modules, replacing its earlier ones; kept across SuckInHits.
  void SuckInDeads(int layer, const DeadVec& deadv) {
    require(layer >= 0 && layer < n_layers_, "SuckInDeads: layer out of range");
    deads_.resize(n_layers_);
    deads_[layer] = deadv;
    std::vector<float> flat;
    for (int l = 0; l < n_layers_; ++l)
      for (const DeadRegion& d : deads_[l]) flat.insert(flat.end(), {(float)l, d.phi1, d.phi2, d.q1, d.q2});
    check(mkf_event_of_hits_set_dead_regions(h_, (int)(flat.size() / 5), flat.empty() ? nullptr : flat.data(),
                                             c_.stream()));
  }
  mkf_event_of_hits* get() const { return h_; }
  int n_layers() const { return n_layers_; }
  int n_hits(int layer) const { return n_hits_.empty() ? 0 : n_hits_[layer + 1] - n_hits_[layer]; }

  std::vector<LayerOfHits> m_layers_of_hits;

 private:
  Context& c_;
  int n_layers_;
  mkf_event_of_hits* h_ = nullptr;
  DeviceArray<float> hits_;
  DeviceArray<unsigned char> mask_, charge_;
  std::vector<unsigned char> codes_;
  std::vector<int> n_hits_;
  std::vector<DeadVec> deads_;
};

using IdxChi2List = mkf_idx_chi2;

// Per-seed candidate lists of a combinatorial search, device-resident (the
// EventOfCombCandidates of the reference).
struct CandidateStore {
  int n = 0, K = 0, L = 0;
  DeviceArray<float> par, err, chi2, score;
  DeviceArray<int> chg, n_found, hot, n_cands;
  void resize(int n_, int K_, int L_) {
    n = n_; K = K_; L = L_;
    const size_t N = n, KK = K;
    n_cands.resize(N); par.resize(KK * 6 * N); err.resize(KK * 21 * N); chg.resize(KK * N);
    chi2.resize(KK * N); score.resize(KK * N); n_found.resize(KK * N); hot.resize(KK * L * N);
  }
};

// ---- finding ------------------------------------------------------------------------

// MkFinder (MkFinder.h:22-359): the lanes are the tracks of the last
// InputTracksAndHitIdx / BkFitInputTracks call.  As in the reference, Par[iC]
// is a lane's current state and Par[iP] its state propagated to the layer;
// N_proc must equal the number of lanes.  Iteration parameters come from
// Setup, propagation settings from Config.
class MkFinder {
 public:
  static constexpr int iC = 0, iP = 1;
  // The reference's Matriplex width NN (Config.h:424-433: 16 with AVX-512),
  // the block of lanes the backward fit steps together.
  static constexpr int kDefaultNN = 16;

  explicit MkFinder(Context& c) : c_(c) {}

  void Setup(const IterationParams& ip) { ip_ = ip; }
  void Release() {}
  void SetMatriplexWidth(int nn) {
    require(nn >= 1 && nn <= 64 && (nn & (nn - 1)) == 0, "SetMatriplexWidth: a power of two, 1..64");
    nn_ = nn;
  }

  // MkFinder::InputTracksAndHitIdx (MkFinder.cc:41-56, copy_in MkFinder.h:238-254).
  void InputTracksAndHitIdx(const TrackVec& tracks, int beg, int end, bool inputProp) {
    require(beg >= 0 && beg <= end && end <= (int)tracks.size(), "InputTracksAndHitIdx: bad range");
    n_ = end - beg;
    const size_t n = n_;
    PackedStates s;
    s.pack(tracks, beg, end);
    std::vector<float> chi2(n);
    std::vector<int> nf(n), nin(n), ntail(n);
    label_.resize(n);
    status_.resize(n);
    lane_hots_.assign(n, {});
    for (size_t t = 0; t < n; ++t) {
      const Track& tr = tracks[beg + t];
      chi2[t] = tr.chi2();
      nf[t] = tr.nFoundHits();
      nin[t] = tr.nInsideMinusOneHits();
      ntail[t] = tr.nTailMinusOneHits();
      label_[t] = tr.label();
      status_[t] = tr.getRawStatus();
      lane_hots_[t].assign(tr.BeginHitsOnTrack(), tr.EndHitsOnTrack());
    }
    hipStream_t st = c_.stream();
    for (auto* a : {&cur_par_, &prop_par_}) a->resize(6 * n);
    for (auto* a : {&cur_err_, &prop_err_}) a->resize(21 * n);
    (inputProp ? prop_par_ : cur_par_).upload(s.par, st);
    (inputProp ? prop_err_ : cur_err_).upload(s.err, st);
    lane_chg_.upload(s.chg, st);
    lane_chi2_.upload(chi2, st);
    lane_nf_.upload(nf, st);
    lane_ninside_.upload(nin, st);
    lane_ntail_.upload(ntail, st);
    xsize_.resize(n); xarr_.resize((size_t)MKF_HIT_IDX_MAX * n); wsr_.resize(n); code_.resize(n);
    bk_loaded_ = false;
    c_.sync();
  }

  // MkFinder::OutputTracksAndHitIdx (MkFinder.cc:87-101, copy_out MkFinder.h:256-267):
  // tracks [beg, end), which must exist.
  void OutputTracksAndHitIdx(TrackVec& tracks, int beg, int end, bool outputProp) const {
    require(beg >= 0 && end - beg == n_ && end <= (int)tracks.size(), "OutputTracksAndHitIdx: range differs");
    std::vector<float> par, err, chi2;
    std::vector<int> chg, nf;
    hipStream_t st = c_.stream();
    (outputProp ? prop_par_ : cur_par_).download(par, st);
    (outputProp ? prop_err_ : cur_err_).download(err, st);
    lane_chg_.download(chg, st); lane_chi2_.download(chi2, st); lane_nf_.download(nf, st);
    c_.sync();
    const size_t n = n_;
    for (size_t t = 0; t < n; ++t) {
      Track& tr = tracks[beg + t];
      unpack_state(tr, par, err, n, t);
      tr.setCharge(chg[t]);
      tr.setChi2(chi2[t]);
      tr.setLabel(label_[t]);
      tr.resizeHits((int)lane_hots_[t].size(), nf[t]);
      std::copy(lane_hots_[t].begin(), lane_hots_[t].end(), tr.hitsOnTrk_.begin());
    }
  }

  // MkBase::PropagateTracksToR / ToZ (MkBase.h:35-76): Par[iC] -> Par[iP].
  void PropagateTracksToR(float r, int N_proc, PropagationFlags pf) {
    lanes_check(N_proc);
    target_.upload(std::vector<float>(n_, r), c_.stream());
    check(mkf_propagate_helix_to_r(c_.get(), n_, cur_err_.get(), cur_par_.get(), lane_chg_.get(), target_.get(),
                                   prop_err_.get(), prop_par_.get(), pf.raw(), nullptr, nullptr, c_.stream()));
  }
  void PropagateTracksToZ(float z, int N_proc, PropagationFlags pf) {
    lanes_check(N_proc);
    target_.upload(std::vector<float>(n_, z), c_.stream());
    check(mkf_propagate_helix_to_z(c_.get(), n_, cur_err_.get(), cur_par_.get(), lane_chg_.get(), target_.get(),
                                   prop_err_.get(), prop_par_.get(), pf.raw(), nullptr, c_.stream()));
  }

  // MkFinder::SelectHitIndices (MkFinder.cc:227-684) from Par[iP].
  void SelectHitIndices(const LayerOfHits& layer_of_hits, int N_proc) {
    lanes_check(N_proc);
    check(mkf_select_hit_indices(layer_of_hits.event_of_hits().get(), layer_of_hits.layer_id(), n_,
                                 prop_par_.get(), prop_err_.get(), xsize_.get(), xarr_.get(), wsr_.get(),
                                 c_.stream()));
  }

  // MkFinder::AddBestHit (MkFinder.cc:691-832): Par[iP] -> Par[iC] with the
  // best hit or a fake hit, which is appended to each lane's HitOnTrack list
  // (none when the window lies outside the layer).
  void AddBestHit(const LayerOfHits& layer_of_hits, int N_proc) {
    lanes_check(N_proc);
    mkf_finder_lanes ln = lanes(true);
    check(mkf_add_best_hit(c_.get(), layer_of_hits.event_of_hits().get(), layer_of_hits.layer_id(), &ln,
                           xsize_.get(), xarr_.get(), wsr_.get(), ip_.chi2Cut_min,
                           Config::finding_requires_propagation_to_hit_pos,
                           Config::finding_intra_layer_pflags.raw(), code_.get(), c_.stream()));
    std::vector<int> codes;
    code_.download(codes, c_.stream());
    c_.sync();
    for (int t = 0; t < n_; ++t)
      if (codes[t] != -9) lane_hots_[t].push_back(HitOnTrack(codes[t], layer_of_hits.layer_id()));
  }

  // MkFinder::FindCandidatesCloneEngine (MkFinder.cc:1114-1284) from Par[iP]:
  // per lane the IdxChi2List records CandCloner::add_cand receives, in the
  // order it receives them; trk_idx is the lane (the reference keys a record
  // by its seed and stores the candidate index, both the caller's bookkeeping
  // of the lane).
  void FindCandidatesCloneEngine(const LayerOfHits& layer_of_hits, std::vector<std::vector<IdxChi2List>>& records,
                                 int N_proc) {
    lanes_check(N_proc);
    const size_t n = n_;
    n_entries_.resize(n);
    entries_.resize(n * MKF_CE_ENTRIES_PER_LANE);
    mkf_finder_lanes ln = lanes(false);
    check(mkf_find_candidates_clone_engine(c_.get(), layer_of_hits.event_of_hits().get(), layer_of_hits.layer_id(),
                                           &ln, xsize_.get(), xarr_.get(), wsr_.get(), ip_.chi2Cut_min,
                                           Config::finding_requires_propagation_to_hit_pos,
                                           Config::finding_intra_layer_pflags.raw(), ip_.maxHolesPerCand,
                                           ip_.maxConsecHoles, n_entries_.get(), entries_.get(), c_.stream()));
    std::vector<int> ne;
    std::vector<IdxChi2List> all;
    n_entries_.download(ne, c_.stream());
    entries_.download(all, c_.stream());
    c_.sync();
    records.assign(n, {});
    for (size_t t = 0; t < n; ++t)
      records[t].assign(all.begin() + t * MKF_CE_ENTRIES_PER_LANE, all.begin() + t * MKF_CE_ENTRIES_PER_LANE + ne[t]);
  }

  // MkFinder::UpdateWithLastHit (MkFinder.cc:1290-1308): Par[iP] updated with
  // each lane's last HitOnTrack into Par[iC] (Config::finding_inter_layer_pflags);
  // lanes whose last HitOnTrack is no hit keep Par[iP].
  void UpdateWithLastHit(const LayerOfHits& layer_of_hits, int N_proc) {
    lanes_check(N_proc);
    std::vector<int> idx(n_, -1);
    for (int t = 0; t < n_; ++t)
      if (!lane_hots_[t].empty()) idx[t] = lane_hots_[t].back().index;
    target_idx_.upload(idx, c_.stream());
    mkf_finder_lanes ln = lanes(true);
    check(mkf_update_with_last_hit(c_.get(), layer_of_hits.event_of_hits().get(), layer_of_hits.layer_id(), &ln,
                                   target_idx_.get(), Config::finding_requires_propagation_to_hit_pos,
                                   Config::finding_inter_layer_pflags.raw(), c_.stream()));
  }

  // ---- backward fit (MkFinder.cc:1342-1749) ----
  // BkFitInputTracks (MkFinder.cc:1342-1366): state, charge and HitOnTrack list
  // of cands [beg, end); the covariance is scaled by 100 and chi2 starts at 0
  // in the fit.
  void BkFitInputTracks(TrackVec& cands, int beg, int end) {
    require(beg >= 0 && beg <= end && end <= (int)cands.size(), "BkFitInputTracks: bad range");
    n_ = end - beg;
    const size_t n = n_;
    bk_in_.pack(cands, beg, end);
    bk_hots_.assign(n, {});
    for (size_t t = 0; t < n; ++t) {
      const Track& tr = cands[beg + t];
      bk_hots_[t].assign(tr.BeginHitsOnTrack(), tr.EndHitsOnTrack());
    }
    bk_loaded_ = true;
    bk_fitted_ = false;
  }

  // BkFitFitTracks (MkFinder.cc:1573-1742, the combinatorial candidates' fit)
  // and BkFitFitTracksBH (:1455-1568, the best-hit tracks' fit: fake hits with
  // material).  Walks st_par's plan in reverse down to m_bkw_fit_last over
  // the lanes in blocks of the Matriplex width.
  void BkFitFitTracks(const EventOfHits& eoh, const SteeringParams& st_par, int N_proc, bool chiDebug = false) {
    (void)chiDebug;
    bk_fit(eoh, st_par, N_proc, false);
  }
  void BkFitFitTracksBH(const EventOfHits& eoh, const SteeringParams& st_par, int N_proc, bool chiDebug = false) {
    (void)chiDebug;
    bk_fit(eoh, st_par, N_proc, true);
  }
  // BkFitPropTracksToPCA (MkFinder.cc:1746-1749) with Config::pca_prop_pflags:
  // Par[iP] becomes the state at the point of closest approach.
  void BkFitPropTracksToPCA(int N_proc) {
    require(bk_fitted_ && N_proc == n_, "BkFitPropTracksToPCA: after BkFitFitTracks over the same lanes");
    bk_pca_ = true;
    bk_run();
  }
  // BkFitOutputTracks (MkFinder.cc:1403-1421): parameters, errors (Par[iP]
  // with outputProp, i.e. after BkFitPropTracksToPCA, else Par[iC]), chi2 and,
  // where chi2 is not NaN, the score getScoreCand.  As there, the charge the
  // fit may have flipped is not copied out.
  void BkFitOutputTracks(TrackVec& cands, int beg, int end, bool outputProp) {
    require(bk_fitted_ && beg >= 0 && end - beg == n_ && end <= (int)cands.size(),
            "BkFitOutputTracks: after BkFitFitTracks, over the same range");
    require(!outputProp || bk_pca_, "BkFitOutputTracks: Par[iP] is set by BkFitPropTracksToPCA");
    if (outputProp != bk_pca_) {   // Par[iC] wanted after the PCA step: the fit without it
      bk_pca_ = outputProp;
      bk_run();
    }
    std::vector<float> par, err, chi2, score;
    hipStream_t st = c_.stream();
    bk_par_.download(par, st); bk_err_.download(err, st); bk_chi2_.download(chi2, st);
    bk_score_.download(score, st);
    c_.sync();
    const size_t n = n_;
    for (size_t t = 0; t < n; ++t) {
      Track& tr = cands[beg + t];
      unpack_state(tr, par, err, n, t);
      tr.setChi2(chi2[t]);
      if (!std::isnan(chi2[t])) tr.setScore(score[t]);
    }
  }

 private:
  friend class MkBuilder;

  void lanes_check(int N_proc) const { require(N_proc == n_, "N_proc must equal the number of lanes input"); }

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

  void bk_fit(const EventOfHits& eoh, const SteeringParams& st_par, int N_proc, bool bh) {
    require(bk_loaded_ && N_proc == n_, "BkFitFitTracks: after BkFitInputTracks, N_proc lanes");
    const int L = (int)st_par.m_layer_plan.size();
    require(L >= 1 && L <= MKF_MAX_FIND_LAYERS && L <= eoh.n_layers(), "BkFitFitTracks: plan size");
    for (int i = 0; i < L; ++i)
      require(st_par.m_layer_plan[i] == i, "BkFitFitTracks: the plan must list the event's layers in order");
    require(st_par.m_bkw_fit_last >= 0 && st_par.m_bkw_fit_last < L, "BkFitFitTracks: m_bkw_fit_last");
    // The HitOnTrack list as one entry per plan layer: on a layer the list
    // visits several times, the earliest found hit (the one the reverse walk
    // ends on), else the last code.
    const size_t n = n_;
    std::vector<int> hot((size_t)L * n, -9);
    for (size_t t = 0; t < n; ++t)
      for (const HitOnTrack& h : bk_hots_[t]) {
        require(h.layer >= 0 && h.layer < L, "BkFitFitTracks: a HitOnTrack layer outside the plan");
        int& slot = hot[(size_t)h.layer * n + t];
        if (slot < 0) slot = h.index;
      }
    bk_eoh_ = eoh.get();
    bk_L_ = L;
    bk_last_ = st_par.m_bkw_fit_last;
    bk_bh_ = bh;
    bk_pca_ = false;
    hipStream_t st = c_.stream();
    bk_hot_.upload(hot, st);
    bk_in_par_.upload(bk_in_.par, st);
    bk_in_err_.upload(bk_in_.err, st);
    bk_in_chg_.upload(bk_in_.chg, st);
    bk_par_.resize(6 * n); bk_err_.resize(21 * n); bk_chg_.resize(n); bk_chi2_.resize(n); bk_score_.resize(n);
    bk_run();
    bk_fitted_ = true;
  }

  void bk_run() {
    mkf_bkfit_batch b{};
    b.n_cands = n_;
    b.in_par = bk_in_par_.get(); b.in_err = bk_in_err_.get(); b.in_chg = bk_in_chg_.get();
    b.hot_idx = bk_hot_.get();
    b.n_layers = bk_L_;
    b.last_layer = bk_last_;
    b.group = std::min(nn_, 64);
    b.pflags = Config::backward_fit_pflags.raw();
    b.pca_pflags = Config::pca_prop_pflags.raw();
    b.flags = (bk_pca_ ? MKF_BKFIT_PCA : 0) | (bk_bh_ ? MKF_BKFIT_FAKE_MATERIAL : 0);
    b.par = bk_par_.get(); b.err = bk_err_.get(); b.chg = bk_chg_.get(); b.chi2 = bk_chi2_.get();
    b.score = bk_score_.get();
    check(mkf_backward_fit(c_.get(), bk_eoh_, &b, c_.stream()));
    c_.sync();
  }

  Context& c_;
  IterationParams ip_;
  int n_ = 0, nn_ = kDefaultNN;
  std::vector<int> label_;
  std::vector<unsigned> status_;
  std::vector<HoTVec> lane_hots_;
  DeviceArray<float> cur_par_, cur_err_, prop_par_, prop_err_, lane_chi2_, target_;
  DeviceArray<int> lane_chg_, lane_nf_, lane_ninside_, lane_ntail_, xsize_, xarr_, wsr_, code_, n_entries_,
      target_idx_;
  DeviceArray<IdxChi2List> entries_;
  // backward fit
  PackedStates bk_in_;
  std::vector<HoTVec> bk_hots_;
  bool bk_loaded_ = false, bk_fitted_ = false, bk_pca_ = false, bk_bh_ = false;
  mkf_event_of_hits* bk_eoh_ = nullptr;
  int bk_L_ = 0, bk_last_ = 0;
  DeviceArray<float> bk_in_par_, bk_in_err_, bk_par_, bk_err_, bk_chi2_, bk_score_;
  DeviceArray<int> bk_in_chg_, bk_hot_, bk_chg_;
};

// MkBuilder (MkBuilder.h:114-237): the event-level searches, each one device
// launch over every seed.  Seeds carry their hits on the first
// IterationParams::nlayers_per_seed layers of the EventOfHits; the search runs
// over the remaining layers in order.
class MkBuilder {
 public:
  explicit MkBuilder(Context& c) : c_(c) {}

  // MkBuilder::begin_event (MkBuilder.cc:201-230) for an event whose hits an
  // EventOfHits holds.
  void begin_event(EventOfHits& eoh, const IterationParams& ip = IterationParams()) {
    eoh_ = &eoh;
    ip_ = ip;
    require(ip.nlayers_per_seed >= 1 && ip.nlayers_per_seed < eoh.n_layers(), "begin_event: nlayers_per_seed");
    require(eoh.n_layers() <= MKF_MAX_FIND_LAYERS, "begin_event: too many layers for one search");
  }
  void end_event() { eoh_ = nullptr; }

  // MkBuilder::find_tracks_load_seeds[_BH] (MkBuilder.cc:1310-1335).
  void find_tracks_load_seeds_BH(const TrackVec& in_seeds) { find_tracks_load_seeds(in_seeds); }
  void find_tracks_load_seeds(const TrackVec& in_seeds) {
    require(eoh_ != nullptr, "find_tracks_load_seeds: begin_event first");
    n_ = (int)in_seeds.size();
    const size_t n = n_;
    const int first = ip_.nlayers_per_seed, L = eoh_->n_layers();
    seeds_ = in_seeds;
    PackedStates s;
    s.pack(in_seeds, 0, n_);
    std::vector<int> hot((size_t)L * n, -9);
    std::vector<float> chi2(n);
    for (size_t t = 0; t < n; ++t) {
      for (const HitOnTrack* h = in_seeds[t].BeginHitsOnTrack(); h != in_seeds[t].EndHitsOnTrack(); ++h)
        if (h->layer >= 0 && h->layer < first) hot[(size_t)h->layer * n + t] = h->index;
      chi2[t] = in_seeds[t].chi2();
    }
    hipStream_t st = c_.stream();
    seed_par_.upload(s.par, st);
    seed_err_.upload(s.err, st);
    seed_chg_.upload(s.chg, st);
    seed_chi2_.upload(chi2, st);
    seed_hot_.upload(hot, st);
    kind_ = None;
    c_.sync();
  }

  // MkBuilder::FindTracksBestHit (MkBuilder.cc:1362-1480).
  void FindTracksBestHit() {
    ready();
    const size_t n = n_;
    const int L = eoh_->n_layers();
    bh_par_.resize(6 * n); bh_err_.resize(21 * n); bh_chg_.resize(n); bh_chi2_.resize(n); bh_nf_.resize(n);
    bh_hot_.resize((size_t)L * n);
    hipStream_t st = c_.stream();
    bh_score_.upload(std::vector<float>(n, 0.f), st);
    bh_par_.copy_from(seed_par_, 6 * n, st);
    bh_err_.copy_from(seed_err_, 21 * n, st);
    bh_chg_.copy_from(seed_chg_, n, st);
    bh_chi2_.copy_from(seed_chi2_, n, st);
    bh_hot_.copy_from(seed_hot_, (size_t)L * n, st);
    mkf_find_batch b{};
    b.n_tracks = n_;
    b.par = bh_par_.get(); b.err = bh_err_.get(); b.chg = bh_chg_.get(); b.chi2 = bh_chi2_.get();
    b.hot_idx = bh_hot_.get(); b.n_found = bh_nf_.get();
    b.first_layer = ip_.nlayers_per_seed; b.chi2_cut = ip_.chi2Cut_min;
    b.pflags = Config::finding_inter_layer_pflags.raw();
    b.prop_to_hit = Config::finding_requires_propagation_to_hit_pos;
    b.intra_pflags = Config::finding_intra_layer_pflags.raw();
    check(mkf_find_tracks_best_hit(c_.get(), eoh_->get(), &b, st));
    c_.sync();
    kind_ = BestHit;
  }

  // MkBuilder::FindTracksCloneEngine (MkBuilder.cc:1847-2060) and
  // FindTracksStandard (:1643-1845).
  void FindTracksCloneEngine() { combinatorial(false); }
  void FindTracksStandard() { combinatorial(true); }

  // MkBuilder::BackwardFitBH (MkBuilder.cc:2069-2122) over the best-hit tracks
  // and BackwardFit (:2124-2253) over the best candidate of every seed, with
  // Config::backward_fit_pflags, includePCA and pca_prop_pflags, in blocks of
  // the Matriplex width.  Parameters, errors, chi2 and score are replaced; the
  // charge is left as found (BkFitOutputTracks, MkFinder.cc:1403-1446).
  void BackwardFitBH() {
    require(kind_ == BestHit, "BackwardFitBH: after FindTracksBestHit");
    back_fit(bh_par_, bh_err_, bh_chg_, bh_hot_, bh_chi2_, bh_score_, true);
  }
  void BackwardFit() {
    require(kind_ == CloneEngine || kind_ == Standard, "BackwardFit: after FindTracksCloneEngine / Standard");
    back_fit(cands_.par, cands_.err, cands_.chg, cands_.hot, cands_.chi2, cands_.score, false);
  }
  void SetMatriplexWidth(int nn) {
    require(nn >= 1 && nn <= 64 && (nn & (nn - 1)) == 0, "SetMatriplexWidth: a power of two, 1..64");
    nn_ = nn;
  }

  // MkBuilder::export_tracks (best-hit) / export_best_comb_cands
  // (MkBuilder.cc:416-455): one track per seed, its HitOnTrack list in layer
  // order, label and status from the seed.  Seeds without a candidate give no
  // track in export_best_comb_cands.
  void export_tracks(TrackVec& out_vec) {
    require(kind_ == BestHit, "export_tracks: after FindTracksBestHit");
    std::vector<float> par, err, chi2, score;
    std::vector<int> chg, nf, hot;
    hipStream_t st = c_.stream();
    bh_par_.download(par, st); bh_err_.download(err, st); bh_chg_.download(chg, st);
    bh_chi2_.download(chi2, st); bh_nf_.download(nf, st); bh_hot_.download(hot, st); bh_score_.download(score, st);
    c_.sync();
    const size_t n = n_;
    out_vec.clear();
    out_vec.reserve(n);
    for (size_t t = 0; t < n; ++t) {
      Track tr;
      fill(tr, t, par, err, chg[t], chi2[t], score[t], nf[t], &hot[0], n);
      out_vec.push_back(std::move(tr));
    }
  }
  void export_best_comb_cands(TrackVec& out_vec) {
    require(kind_ == CloneEngine || kind_ == Standard, "export_best_comb_cands: after a combinatorial search");
    std::vector<float> par, err, chi2, score;
    std::vector<int> chg, nf, hot, nc;
    const size_t n = n_;
    const int L = eoh_->n_layers();
    hipStream_t st = c_.stream();
    cands_.par.download(par, 6 * n, st); cands_.err.download(err, 21 * n, st); cands_.chg.download(chg, n, st);
    cands_.chi2.download(chi2, n, st); cands_.score.download(score, n, st); cands_.n_found.download(nf, n, st);
    cands_.hot.download(hot, (size_t)L * n, st); cands_.n_cands.download(nc, st);
    c_.sync();
    out_vec.clear();
    out_vec.reserve(n);
    for (size_t t = 0; t < n; ++t) {
      if (nc[t] == 0) continue;
      Track tr;
      fill(tr, t, par, err, chg[t], chi2[t], score[t], nf[t], &hot[0], n);
      out_vec.push_back(std::move(tr));
    }
  }
  // Every candidate of every seed, best first (the EventOfCombCandidates).
  void export_comb_cands(std::vector<TrackVec>& out) {
    require(kind_ == CloneEngine || kind_ == Standard, "export_comb_cands: after a combinatorial search");
    std::vector<float> par, err, chi2, score;
    std::vector<int> chg, nf, hot, nc;
    hipStream_t st = c_.stream();
    cands_.par.download(par, st); cands_.err.download(err, st); cands_.chg.download(chg, st);
    cands_.chi2.download(chi2, st); cands_.score.download(score, st); cands_.n_found.download(nf, st);
    cands_.hot.download(hot, st); cands_.n_cands.download(nc, st);
    c_.sync();
    const size_t n = n_, L = eoh_->n_layers();
    out.assign(n, {});
    for (size_t t = 0; t < n; ++t)
      for (int k = 0; k < nc[t]; ++k) {
        Track tr;
        const size_t s = k * n + t;
        std::vector<float> p(par.begin() + k * 6 * n, par.begin() + (k + 1) * 6 * n);
        std::vector<float> e(err.begin() + k * 21 * n, err.begin() + (k + 1) * 21 * n);
        fill(tr, t, p, e, chg[s], chi2[s], score[s], nf[s], &hot[k * L * n], n);
        out[t].push_back(std::move(tr));
      }
  }
  int total_cands() const {
    if (kind_ == BestHit) return n_;
    if (kind_ != CloneEngine && kind_ != Standard) return 0;
    std::vector<int> nc;
    cands_.n_cands.download(nc, c_.stream());
    c_.sync();
    int s = 0;
    for (int v : nc) s += v;
    return s;
  }

 private:
  enum Kind { None, BestHit, CloneEngine, Standard };

  void ready() const {
    require(eoh_ != nullptr, "begin_event first");
    require(!seeds_.empty() || n_ == 0, "find_tracks_load_seeds first");
  }

  void fill(Track& tr, size_t t, const std::vector<float>& par, const std::vector<float>& err, int chg, float chi2,
            float score, int nf, const int* hot, size_t n) const {
    unpack_state(tr, par, err, n, t);
    tr.setCharge(chg);
    tr.setChi2(chi2);
    tr.setScore(score);
    tr.setLabel(seeds_[t].label());
    tr.setRawStatus(seeds_[t].getRawStatus());
    const int L = eoh_->n_layers();
    int cnt = 0;
    for (int l = 0; l < L; ++l) cnt += hot[(size_t)l * n + t] != -9;
    tr.resizeHits(cnt, nf);
    int i = 0;
    for (int l = 0; l < L; ++l) {
      const int h = hot[(size_t)l * n + t];
      if (h != -9) tr.setHitIdxLyr(i++, h, l);
    }
  }

  void combinatorial(bool standard) {
    ready();
    const int K = ip_.maxCandsPerSeed, L = eoh_->n_layers();
    require(K >= 1 && K <= MKF_CE_MAX_CANDS, "maxCandsPerSeed: 1..MKF_CE_MAX_CANDS");
    cands_.resize(n_, K, L);
    mkf_ce_batch b{};
    b.n_seeds = n_;
    b.seed_par = seed_par_.get(); b.seed_err = seed_err_.get(); b.seed_chg = seed_chg_.get();
    b.seed_hits = seed_hot_.get();
    b.first_layer = ip_.nlayers_per_seed; b.chi2_cut = ip_.chi2Cut_min;
    b.pflags = Config::finding_inter_layer_pflags.raw();
    b.intra_pflags = Config::finding_intra_layer_pflags.raw();
    b.prop_to_hit = Config::finding_requires_propagation_to_hit_pos;
    b.max_cands = K; b.max_holes = ip_.maxHolesPerCand; b.max_consec_holes = ip_.maxConsecHoles;
    b.n_cands = cands_.n_cands.get(); b.par = cands_.par.get(); b.err = cands_.err.get(); b.chg = cands_.chg.get();
    b.chi2 = cands_.chi2.get(); b.score = cands_.score.get(); b.n_found = cands_.n_found.get();
    b.hot_idx = cands_.hot.get();
    check(standard ? mkf_find_tracks_standard(c_.get(), eoh_->get(), &b, c_.stream())
                   : mkf_find_tracks_clone_engine(c_.get(), eoh_->get(), &b, c_.stream()));
    c_.sync();
    kind_ = standard ? Standard : CloneEngine;
  }

  // The first n entries of each array are the tracks to fit (slot 0 of a
  // candidate store); results go back in place, charge excepted.
  void back_fit(DeviceArray<float>& par, DeviceArray<float>& err, DeviceArray<int>& chg, DeviceArray<int>& hot,
                DeviceArray<float>& chi2, DeviceArray<float>& score, bool bh) {
    chg_scratch_.resize(n_);
    mkf_bkfit_batch b{};
    b.n_cands = n_;
    b.in_par = par.get(); b.in_err = err.get(); b.in_chg = chg.get();
    b.hot_idx = hot.get();
    b.n_layers = eoh_->n_layers();
    b.last_layer = 0;
    b.group = nn_;
    b.pflags = Config::backward_fit_pflags.raw();
    b.pca_pflags = Config::pca_prop_pflags.raw();
    b.flags = (Config::includePCA ? MKF_BKFIT_PCA : 0) | (bh ? MKF_BKFIT_FAKE_MATERIAL : 0);
    b.par = par.get(); b.err = err.get(); b.chg = chg_scratch_.get(); b.chi2 = chi2.get(); b.score = score.get();
    check(mkf_backward_fit(c_.get(), eoh_->get(), &b, c_.stream()));
    c_.sync();
  }

  Context& c_;
  EventOfHits* eoh_ = nullptr;
  IterationParams ip_;
  TrackVec seeds_;
  int n_ = 0, nn_ = MkFinder::kDefaultNN;
  Kind kind_ = None;
  DeviceArray<float> seed_par_, seed_err_, seed_chi2_, bh_par_, bh_err_, bh_chi2_, bh_score_;
  DeviceArray<int> seed_chg_, seed_hot_, bh_chg_, bh_nf_, bh_hot_, chg_scratch_;
  CandidateStore cands_;
};

}  // namespace mkfit_amd
