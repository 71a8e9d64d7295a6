/* mkfit_amd.h -- C ABI of the MI355X (gfx950) Kalman fit and hit-selection library.
 *
 * Drop-in boundary for mkFit's Matriplex hot path.  Each entry point replaces a
 * reference C++ interface (cited per function) that operates on caller-owned
 * NN-wide Matriplex blocks; here the "block" is the whole batch resident in GPU
 * memory, with the same element-major order (Matriplex/Matriplex.h:65):
 * element e of track t is at a[e*n + t].
 *
 *   parameters  [6][n]   CCS form x, y, z, 1/pT, phi, theta (MPlexLV, Matrix.h:79)
 *   covariance  [21][n]  packed lower triangle (MPlexLS, MatriplexSym.h:15-22)
 *   charge      [n]      int (MPlexQI)
 *   hit pos     [3][n]   (MPlexHV), hit cov [6][n] xx yx yy zx zy zz (MPlexHS)
 *
 * All pointers passed to compute entry points are device pointers (hipMalloc'd
 * memory); they are checked against their allocation's extent before any
 * launch.  `stream` is a hipStream_t (NULL = the default stream).  Functions
 * return 0 on success and a nonzero MKF_ERR_* code otherwise, with a message in
 * mkf_last_error().  Work is enqueued asynchronously on `stream`.
 */
#ifndef MKFIT_AMD_H
#define MKFIT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define MKF_OK              0
#define MKF_ERR_INVALID_ARG 1
#define MKF_ERR_HIP         2
#define MKF_ERR_NO_DEVICE   3
#define MKF_ERR_BOUNDS      4

/* PropagationFlags (Config.h:17-47). */
#define MKF_PF_USE_PARAM_B_FIELD 0x1
#define MKF_PF_APPLY_MATERIAL    0x2

/* KalmanFilterOperation (mkFit/KalmanUtilsMPlex.h:11-15). */
#define MKF_KFO_CALCULATE_CHI2 1
#define MKF_KFO_UPDATE_PARAMS  2

/* Fit flags. */
#define MKF_FIT_CHI2        1 /* accumulate chi2 of every hit into chi2[] (MkFinder.cc:1740) */
#define MKF_FIT_CHARGE_FLIP 2 /* flip charge and 1/pT when 1/pT < 0 after an update (MkFinder.cc:1671-1678) */
#define MKF_FIT_PROP_TO_PCA 8 /* after the last hit, propagate to the point of closest approach to the origin in
                                 the r-z plane (MkBase::PropagateTracksToPCAZ via BkFitPropTracksToPCA,
                                 MkBase.h:91-104, MkFinder.cc:1746-1749) with pca_pflags */
#define MKF_FIT_FP64_FALLBACK 4 /* no counterpart in the reference: where an updated covariance fails the
                                   conditioning test (negative variance, or C_ij^2 > (1+1e-5) C_ii C_jj), recompute
                                   it in FP64 in the Joseph form (I-KH)C(I-KH)^T + KVK^T */

/* Per-slot hit flags (hit_flags[slot][track]). */
#define MKF_HIT_ENDCAP  1 /* propagate to the hit z, endcap update (MkFitter.cc:488-497) */
#define MKF_HIT_MISSING 2 /* no hit: fake hit {666,0,666,0,0,666} at the track position, no material (MkFinder.cc:1635-1647) */

typedef struct mkf_context mkf_context;

typedef struct mkf_config {
  int device;            /* HIP device ordinal */
  float bfield;          /* uniform Bz in T (Config::Bfield, Config.h:244) */
  const float* rl_grid;  /* host [300][120] radiation-length grid (Config::RlgridME) or NULL */
  const float* xi_grid;  /* host [300][120] energy-loss grid (Config::XigridME) or NULL */
} mkf_config;

/* Context: uploads the material grid to the device once; replaces the
 * Config:: globals read by the hot path (Config.cc:42-58, Config.h:164-165). */
int mkf_context_create(const mkf_config* cfg, mkf_context** out);
int mkf_context_destroy(mkf_context* ctx);
const char* mkf_last_error(void);

/* propagateHelixToRMPlex (mkFit/PropagationMPlex.h:30-34, .cc:408-505).
 * no_mat [n] may be NULL (noMatEffPtr); out_fail [n] may be NULL. */
int mkf_propagate_helix_to_r(mkf_context* ctx, int n, const float* in_err, const float* in_par, const int* in_chg,
                             const float* ms_rad, float* out_err, float* out_par, int pflags, const int* no_mat,
                             int* out_fail, void* stream);

/* propagateHelixToZMPlex (mkFit/PropagationMPlex.h:45-49, .cc:510-601). */
int mkf_propagate_helix_to_z(mkf_context* ctx, int n, const float* in_err, const float* in_par, const int* in_chg,
                             const float* ms_z, float* out_err, float* out_par, int pflags, const int* no_mat,
                             void* stream);

/* kalmanOperation (mkFit/KalmanUtilsMPlex.h:42-46, .cc:548-681).  out_err/out_par
 * are written when kf_op has UPDATE_PARAMS, out_chi2 when it has CALCULATE_CHI2
 * (unused outputs may be NULL). */
int mkf_kalman_operation(mkf_context* ctx, int kf_op, int n, const float* ps_err, const float* ps_par,
                         const float* ms_err, const float* ms_par, float* out_err, float* out_par, float* out_chi2,
                         void* stream);

/* kalmanOperationEndcap (mkFit/KalmanUtilsMPlex.h:73-77, .cc:774-888). */
int mkf_kalman_operation_endcap(mkf_context* ctx, int kf_op, int n, const float* ps_err, const float* ps_par,
                                const float* ms_err, const float* ms_par, float* out_err, float* out_par,
                                float* out_chi2, void* stream);

/* A batch of tracks to fit, all arrays device-resident.  Slot s of track t is
 * the s-th hit along the track; a track uses slots [0, n_hits[t]). */
typedef struct mkf_fit_batch {
  int n_tracks;
  int n_slots;
  float* par;                     /* [6][n]  in: seed state, out: fitted state */
  float* err;                     /* [21][n] */
  int* chg;                       /* [n] */
  float* chi2;                    /* [n] accumulated in place when MKF_FIT_CHI2 */
  const float* hit_pos;           /* [n_slots][3][n] */
  const float* hit_err;           /* [n_slots][6][n] */
  const int* n_hits;              /* [n] */
  const unsigned char* hit_flags; /* [n_slots][n] MKF_HIT_*, or NULL: all barrel, none missing */
  int pflags;                     /* MKF_PF_* used for every propagation (forward_fit_pflags) */
  int fit_flags;                  /* MKF_FIT_* */
  /* Optional out-of-place input (MkFitter::InputTracksForFit, MkFitter.cc:286-321):
   * when seed_par is non-NULL the fit reads the seed state from seed_par/seed_err/
   * seed_chg, writes the result to par/err/chg, and starts chi2 from 0; the seed
   * arrays are not modified.  When NULL, par/err/chg/chi2 are updated in place. */
  const float* seed_par;          /* [6][n] or NULL */
  const float* seed_err;          /* [21][n] */
  const int* seed_chg;            /* [n] */
  int pca_pflags;                 /* MKF_PF_* of the propagation to the PCA (Config::pca_prop_pflags) */
} mkf_fit_batch;

/* Plain fit: one kernel launch per hit slot, state round-trips through HBM
 * between launches.  Replaces the per-layer loop of MkFitter::FitTracksWithInterSlurp
 * / FitTracksSteered (mkFit/MkFitter.cc:323-363,473-503) driven over all NN
 * blocks by runFittingTestPlex (mkFit/fittestMPlex.cc:45-110). */
int mkf_fit_plain(mkf_context* ctx, const mkf_fit_batch* batch, void* stream);

/* Batched fit: one persistent fused propagate+update kernel over the whole
 * batch, state kept in registers across slots, replayed from a cached hipGraph
 * when the same batch is fitted again.  Same results as mkf_fit_plain. */
int mkf_fit_batched(mkf_context* ctx, const mkf_fit_batch* batch, void* stream);

/* Launch shape used by mkf_fit_batched (set by the autotuner or by hand). */
typedef struct mkf_launch_shape {
  int block;          /* threads per block: 64, 128, 256 or 512 */
  int blocks_per_cu;  /* persistent grid = blocks_per_cu * CU count; 0: one thread per track */
  int variant;        /* kernel variant index (occupancy / register bound) */
  int sort_tile;      /* 0, or 1024 / 4096 / 16384: before the fit, the tracks of each tile of
                         sort_tile consecutive tracks are ordered by (hit count, endcap pattern of
                         their slots), so a wavefront runs tracks of one length and one branch per
                         slot (the reference groups its blocks the same way: one is_barrel pattern
                         per block, MkFitter.cc:473-503).  Results do not depend on it. */
} mkf_launch_shape;

/* The track order mkf_fit_batched follows with a sort tile: order[i] is the
 * track fitted at position i (a permutation of [0, n_tracks) that stays within
 * each tile).  No counterpart in the reference, whose drivers group tracks
 * into blocks of one layer signature on the host (MkBuilder.cc:594-625,
 * MkFitter.cc:473-503); exposed for callers that fit the same batch in
 * several passes and for timing the ordering on its own. */
int mkf_fit_order(mkf_context* ctx, const mkf_fit_batch* batch, int sort_tile, int* order, void* stream);

int mkf_set_batched_shape(mkf_context* ctx, const mkf_launch_shape* shape);
int mkf_get_batched_shape(mkf_context* ctx, mkf_launch_shape* shape);
/* Times every launch shape on `batch` (restoring the batch's state between
 * trials) and keeps the fastest for subsequent mkf_fit_batched calls.
 * scratch_state must hold a copy area of 29*n_tracks floats (device). */
int mkf_autotune_batched(mkf_context* ctx, const mkf_fit_batch* batch, float* scratch_state, int reps,
                         mkf_launch_shape* best, float* best_ms);

/* Number of variants mkf_set_batched_shape accepts. */
int mkf_num_batched_variants(void);

/* ---- Gradient of the fit ------------------------------------------------------
 *
 * No counterpart in the reference (whose fit has no derivative code): the
 * reverse-mode gradient of the fit's total chi2 with respect to per-layer
 * material and alignment, for calibration.  The fit is the barrel chi2 fit of
 * mkf_fit_batched (one hit per layer, slot l = layer l, uniform field) with
 * two inputs made parameters theta[l] = (x/X0, d_rphi, d_z) of layer l:
 *   material   the propagation to layer l uses radL = x/X0 and
 *              xi = xi_per_x0 * x/X0 instead of the material grid;
 *   alignment  layer l's hits are moved by d_rphi (cm) along their local r-phi
 *              direction and d_z (cm) along z before the propagation and update:
 *              x' = x - d_rphi y/r, y' = y + d_rphi x/r, z' = z + d_z.
 * The forward states are kept only every `checkpoint_every` layers and are
 * recomputed within each segment on the way back. */
typedef struct mkf_adjoint_batch {
  int n_tracks, n_layers;
  const float* seed_par;  /* [6][n] */
  const float* seed_err;  /* [21][n] */
  const int* seed_chg;    /* [n] */
  const float* hit_pos;   /* [n_layers][3][n] */
  const float* hit_err;   /* [n_layers][6][n] */
  const int* n_hits;      /* [n] hits on layers 0 .. n_hits-1 */
  const float* theta;     /* [n_layers][3] device */
  int pflags;             /* MKF_PF_APPLY_MATERIAL: material from theta (MKF_PF_USE_PARAM_B_FIELD unsupported) */
  float xi_per_x0;        /* energy-loss scale per unit x/X0 (GeV) */
  int checkpoint_every;   /* layers per recomputed segment, >= 1 */
  float* chi2;            /* [n] out: the track's total chi2 */
  float* grad;            /* [n_layers][3] out: d(sum over tracks of chi2)/d theta */
  int fp64;               /* 0: the FP32 fit model (as mkf_fit_batched computes it); 1: the same model in FP64,
                             forward and reverse (checkpoints twice the size) */
  int tangent;            /* 0: each step's derivative by the hand-written reverse sweep; 1: from forward
                             tangents, 30 input directions (a cross-check, several times slower) */
} mkf_adjoint_batch;

int mkf_fit_adjoint(mkf_context* ctx, const mkf_adjoint_batch* batch, void* stream);

/* mkf_fit_adjoint without any device allocation: no counterpart in the reference.  For callers that cannot
 * allocate device memory -- code under HIP graph capture, tight fit-and-descend loops, harnesses that forbid
 * hipMalloc.  Same kernels and results as mkf_fit_adjoint; the checkpoints, gradient rows and sum partials
 * live in a fixed 64 MiB device array that each GPU holds for the life of the process, so a batch may need at
 * most 64 MiB: 64*3*8*n_layers (256-aligned) + (ceil(n_layers/checkpoint_every)*27 + 3*n_layers) * n_tracks *
 * (fp64 ? 8 : 4) bytes, e.g. 79,000 tracks on 25 layers every 5 in FP32; larger batches return
 * MKF_ERR_INVALID_ARG.  Every call on one GPU shares that array: calls that overlap on one GPU (other streams,
 * other contexts, other threads) race and must be ordered by the caller. */
int mkf_fit_adjoint_noalloc(mkf_context* ctx, const mkf_adjoint_batch* batch, void* stream);

/* ---- Best-hit track finding -------------------------------------------------
 *
 * Hits are the reference's Hit (Hit.h:165-183) as 9 floats: x, y, z, then the
 * packed 3x3 covariance xx, yx, yy, zx, zy, zz.  Layer l of an event owns the
 * hits [hit_offset[l], hit_offset[l+1]) of one device array [total][9]. */

/* One layer: LayerInfo geometry (TrackerInfo.h:35-98: m_rin, m_rout, m_zmin,
 * m_zmax, m_propagate_to, m_q_bin, is_barrel) and the IterationLayerConfig
 * selection settings (IterationConfig.h:44-90: min/max dphi and dq windows,
 * dynamic-window and chi2-cut coefficients c_dp_*, c_dq_*, c_c2_*). */
typedef struct mkf_layer_desc {
  int is_barrel;
  float rin, rout, zmin, zmax;
  float propagate_to;
  float q_bin; /* > 0: bin width in cm; < 0: number of bins (LayerOfHits::setup_bins) */
  float min_dphi, max_dphi, min_dq, max_dq;
  float c_dp_sf, c_dp_0, c_dp_1, c_dp_2;
  float c_dq_sf, c_dq_0, c_dq_1, c_dq_2;
  float c_c2_sf, c_c2_0, c_c2_1, c_c2_2;
} mkf_layer_desc;

#define MKF_MAX_FIND_LAYERS 64    /* layers one track is searched through */
#define MKF_MAX_EOH_LAYERS 2048   /* layers of one mkf_event_of_hits (several events may share it) */
#define MKF_HIT_IDX_MAX 16 /* candidates per track and layer (MPlexHitIdxMax, MkFitter.h:18) */

/* Window-selection result per track (WSR_Result, TrackerInfo.h:17-31). */
#define MKF_WSR_INSIDE  0
#define MKF_WSR_EDGE    1
#define MKF_WSR_OUTSIDE 2

typedef struct mkf_event_of_hits mkf_event_of_hits;

/* EventOfHits (HitStructures.h:280-319): one LayerOfHits per layer, in the
 * order the finding visits them.  Device storage grows on demand at load.
 * Several events may share one: event e's layers are then a contiguous run
 * of layers (e.g. e * L .. e * L + L - 1), so that one load builds every
 * event's LayerOfHits and one finder launch serves seeds of all of them (a
 * seed names its event's first layer in the batch's layer_base).  At most
 * MKF_MAX_EOH_LAYERS layers, and 1024 q bins per layer. */
int mkf_event_of_hits_create(mkf_context* ctx, int n_layers, const mkf_layer_desc* layers, mkf_event_of_hits** out);
int mkf_event_of_hits_destroy(mkf_event_of_hits* eoh);

/* LayerOfHits::SuckInHits for every layer (HitStructures.cc:77-183): bins each
 * hit by (q, phi) -- q = z on barrel layers, r on endcaps -- sorts the layer's
 * hits stably by (q bin, fine phi bin) and records each bin's [first, end)
 * range.  `hit_offset` is a host array [n_layers + 1]; `hits` a device array
 * [hit_offset[n_layers]][9].  The hits are copied in sorted order, so `hits`
 * may be reused once the stream has passed this call. */
int mkf_event_of_hits_load(mkf_event_of_hits* eoh, const int* hit_offset, const float* hits, void* stream);

/* The iteration's hit mask (MkJob::get_mask_for_layer, MkFinder.cc:476-482): mask
 * is a device array with one byte per hit of the last load, in the order the
 * hits were given to it; a nonzero byte marks a hit taken by an earlier
 * iteration, which hit selection then passes over (it does not count toward the
 * MKF_HIT_IDX_MAX hits of a window).  NULL clears the mask; every load clears it.
 * The mask is copied on `stream`, so `mask` may be reused once the stream has
 * passed this call. */
int mkf_event_of_hits_set_hit_mask(mkf_event_of_hits* eoh, const unsigned char* mask, void* stream);

/* Strip cluster charge (Hit::PackedData::charge_pcm, Hit.h:251-282): code is a
 * device array of one byte per loaded hit, in load order, holding each hit's
 * 8-bit charge code (0: below 1620 per cm; 0xff: overflow or pixel-like).  The
 * combinatorial finders and mkf_find_candidates_clone_engine then take a hit
 * only if its charge per cm times the track's path-length factor at the hit
 * exceeds 1620 (passStripChargePCMfromTrack, MkFinder.cc:863-896, applied on
 * every layer: the layer descriptors carry no pixel flag).  Best-hit finding
 * does not apply it, as the reference's AddBestHit does not.  NULL clears the
 * codes (every hit passes, as code 0xff); a load clears them.  The codes are
 * copied on `stream`, so `code` may be reused once the stream has passed this
 * call. */
int mkf_event_of_hits_set_hit_charge(mkf_event_of_hits* eoh, const unsigned char* code, void* stream);

/* Dead modules (DeadRegion, Hit.h:333-337; LayerOfHits::SuckInDeads,
 * HitStructures.cc:186-212): regions is a host array [n][5] of {layer, phi1,
 * phi2, q1, q2}, each marking the (q bin, phi bin) rectangle of its layer (the
 * phi range wrapping when phi1's bin lies after phi2's).  The combinatorial
 * finders then give a candidate whose window's central q row crosses a marked
 * bin, and that takes no hit on the layer, the invalid hit -7, which is no hole
 * (MkFinder.cc:451-462, 1092-1096, 1263-1267); best-hit finding and the
 * per-layer entry points do not model it, as the reference's AddBestHit does
 * not.  Replaces earlier regions; n = 0 clears them; loads keep them. */
int mkf_event_of_hits_set_dead_regions(mkf_event_of_hits* eoh, int n, const float* regions, void* stream);

/* Device views of a loaded layer (valid until the next load or destroy):
 * ranks [n] sorted position -> hit index within the layer (LayerOfHits
 * m_hit_ranks), bins [nq*256][2] first/end per (q bin, phi bin) (m_phi_bin_infos;
 * a bin without hits is empty, and in phi bins 1..255 holds {p, p} with p the
 * position where its hits would start, so those bins tile their q row),
 * phis/qs [n] per sorted hit (m_hit_phis, m_hit_qs). */
int mkf_event_of_hits_layer(mkf_event_of_hits* eoh, int layer, int* n_hits, int* nq, const int** ranks,
                            const int** bins, const float** phis, const float** qs);

/* MkFinder::SelectHitIndices (MkFinder.cc:227-684) for n tracks whose state
 * (par, err; device, element-major) is already at `layer`: xsize [n] (-1 when
 * the window lies outside the layer), xarr [MKF_HIT_IDX_MAX][n] hit indices
 * within the layer in scan order (-1 padded), wsr [n] MKF_WSR_*. */
int mkf_select_hit_indices(mkf_event_of_hits* eoh, int layer, int n, const float* par, const float* err,
                           int* xsize, int* xarr, int* wsr, void* stream);

/* ---- Per-layer finding steps --------------------------------------------
 * The MkFinder operations a caller's own layer loop applies to one layer
 * (MkBuilder.cc:1395-1478, 1926-2060), on the per-lane MkFinder members
 * (MkFinder.h:238-300) held as device arrays, element-major.  xsize / xarr /
 * wsr are what mkf_select_hit_indices wrote for the same lanes and layer;
 * xarr entries outside [0, hits on the layer) are skipped. */
typedef struct mkf_finder_lanes {
  int n;
  const float* prop_par;   /* [6][n]  Par[iP]: the state propagated to the layer */
  const float* prop_err;   /* [21][n] Err[iP] */
  float* par;              /* [6][n]  Par[iC]: out, the state updated on the layer */
  float* err;              /* [21][n] Err[iC] */
  int* chg;                /* [n] Chg, in/out: flipped with 1/pT when an update leaves 1/pT < 0 */
  float* chi2;             /* [n] Chi2 */
  int* n_found;            /* [n] NFoundHits */
  int* n_inside_minus_one; /* [n] NInsideMinusOneHits */
  int* n_tail_minus_one;   /* [n] NTailMinusOneHits */
} mkf_finder_lanes;

/* MkFinder::AddBestHit (MkFinder.cc:691-832) with add_hit (MkFinder.h:312-354):
 * per lane the smallest |chi2| below getHitSelDynamicChi2Cut among its xarr
 * hits, else a fake hit (-3 window on the layer edge, -2 after an earlier -1,
 * else -1); chi2 and the hit counters advance; then the Kalman update
 * (m_update_param_foo = kalmanPropagateAndUpdate[Endcap], KalmanUtilsMPlex.cc:
 * 471-504, 697-730) of prop_* with the chosen or fake hit into par / err, with
 * the charge flip.  hit_code [n] out: the HitOnTrack index added for the layer,
 * or -9 when the window lies outside the layer (no HitOnTrack; the update
 * still uses the fake hit).  prop_to_hit / intra_pflags: as in mkf_find_batch. */
int mkf_add_best_hit(mkf_context* ctx, mkf_event_of_hits* eoh, int layer, const mkf_finder_lanes* lanes,
                     const int* xsize, const int* xarr, const int* wsr, float chi2_cut, int prop_to_hit,
                     int intra_pflags, int* hit_code, void* stream);

/* IdxChi2List (Track.h:28-42), plus the entry's position in its lane's xarr. */
typedef struct mkf_idx_chi2 {
  int trk_idx;      /* lane */
  int hit_idx;      /* hit index within the layer, or the fake code -1 / -2 / -3 */
  unsigned module;  /* 0 for hits (no module ids), 0xffffffff for the fake entry */
  int nhits, ntailholes, noverlaps, nholes;
  float pt, chi2, chi2_hit, score;
  int xhit_pos;     /* position in xarr; -1 for the fake entry */
} mkf_idx_chi2;

#define MKF_CE_ENTRIES_PER_LANE (MKF_HIT_IDX_MAX + 1)

/* MkFinder::FindCandidatesCloneEngine (MkFinder.cc:1114-1284): per lane the
 * IdxChi2List records its hits pass (|chi2| below the dynamic cut and
 * isStripQCompatible), in xarr order, then the invalid-hit record (-1, or -2
 * past max_holes / max_consec_holes, -3 on the layer edge), none when the
 * window lies outside the layer.  Reads prop_*, chg, chi2 and the counters of
 * `lanes`; writes n_entries [n] and entries [n][MKF_CE_ENTRIES_PER_LANE].
 * The reference hands each record to CandCloner::add_cand in the order
 * (xarr position, lane) for hits, then lane order for invalid records. */
int mkf_find_candidates_clone_engine(mkf_context* ctx, mkf_event_of_hits* eoh, int layer,
                                     const mkf_finder_lanes* lanes, const int* xsize, const int* xarr,
                                     const int* wsr, float chi2_cut, int prop_to_hit, int intra_pflags,
                                     int max_holes, int max_consec_holes, int* n_entries, mkf_idx_chi2* entries,
                                     void* stream);

/* MkFinder::UpdateWithLastHit (MkFinder.cc:1290-1308): per lane the Kalman
 * update of prop_* with hit hit_idx[t] of the layer (kalmanPropagateAndUpdate
 * with the inter-layer flags `pflags`, charge flip) into par / err.  Lanes with
 * hit_idx < 0 or past the layer's hits copy prop_* unchanged. */
int mkf_update_with_last_hit(mkf_context* ctx, mkf_event_of_hits* eoh, int layer, const mkf_finder_lanes* lanes,
                             const int* hit_idx, int prop_to_hit, int pflags, void* stream);

typedef struct mkf_find_batch {
  int n_tracks;
  float* par;     /* [6][n]  in: seed state; out: state when the track stopped or after the last layer */
  float* err;     /* [21][n] */
  int* chg;       /* [n] */
  float* chi2;    /* [n] in: seed chi2; out: plus the chi2 of every added hit */
  int* hot_idx;   /* [n_layers][n] in: seed hit indices on layers < first_layer; out: per layer the hit
                     index within the layer, -1 (no hit, within the tail), -2 (no hit, stop: too many
                     holes), -3 (no hit, window on the layer edge), or -9 (layer not reached) */
  int* n_found;   /* [n] out: hits found, seed hits included */
  int first_layer;/* layers before it hold the seed hits */
  float chi2_cut; /* IterationParams::chi2Cut_min (IterationConfig.h:101) */
  int pflags;     /* Config::finding_inter_layer_pflags: MKF_PF_* for the layer-to-layer propagation */
  int prop_to_hit;  /* Config::finding_requires_propagation_to_hit_pos (Config.cc:52): chi2 and update after
                       propagating to each hit (kalmanPropagateAndComputeChi2 / ...AndUpdate) */
  int intra_pflags; /* Config::finding_intra_layer_pflags for that propagation */
  int n_layers;      /* layers per event searched (hot_idx rows); 0: all layers of the event of hits */
  const int* layer_base; /* [n] device: first layer (in the event of hits) of each seed's event, or NULL: 0.
                            A seed whose event would run past the last layer is left untouched. */
} mkf_find_batch;

/* MkBuilder::FindTracksBestHit (MkBuilder.cc:1362-1480) over the layers of
 * `eoh` from first_layer on: per layer, propagate to LayerInfo::m_propagate_to
 * (PropagateTracksToR / ToZ), stop tracks that cannot reach a barrel layer,
 * SelectHitIndices, AddBestHit (MkFinder.cc:691-832: the smallest |chi2| below
 * the dynamic cut among at most MKF_HIT_IDX_MAX candidates, else a fake hit
 * code), Kalman update with the chosen or fake hit and charge flip
 * (kalmanPropagateAndUpdate, KalmanUtilsMPlex.cc:471-504, without the
 * propagation to the hit), and stop after a -2.  One GPU lane carries one seed
 * through all layers. */
int mkf_find_tracks_best_hit(mkf_context* ctx, mkf_event_of_hits* eoh, const mkf_find_batch* batch, void* stream);

#define MKF_CE_MAX_CANDS 8

typedef struct mkf_ce_batch {
  int n_seeds;
  const float* seed_par;  /* [6][n] */
  const float* seed_err;  /* [21][n] */
  const int* seed_chg;    /* [n] */
  const int* seed_hits;   /* [first_layer][n] hit index of the seed on each layer before first_layer */
  int first_layer;        /* >= 1 */
  float chi2_cut;         /* IterationParams::chi2Cut_min */
  int pflags;             /* finding_inter_layer_pflags: layer-to-layer and update propagation */
  int intra_pflags;       /* finding_intra_layer_pflags: propagation to a hit for its chi2 */
  int prop_to_hit;        /* finding_requires_propagation_to_hit_pos (Config.cc:52) */
  int max_cands;          /* IterationParams::maxCandsPerSeed, 1..MKF_CE_MAX_CANDS (IterationConfig.h:97) */
  int max_holes;          /* maxHolesPerCand */
  int max_consec_holes;   /* maxConsecHoles */
  /* out: per seed its candidates after the final merge, best first (K = max_cands) */
  int* n_cands;           /* [n] */
  float* par;             /* [K][6][n] */
  float* err;             /* [K][21][n] */
  int* chg;               /* [K][n] */
  float* chi2;            /* [K][n] */
  float* score;           /* [K][n]  getScoreCand of each candidate; -1e16 for empty slots */
  int* n_found;           /* [K][n] */
  int* hot_idx;           /* [K][n_layers][n] hit index or fake code per layer, -9 none */
  int n_layers;           /* layers per event searched; 0: all layers of the event of hits */
  const int* layer_base;  /* [n] device: first layer of each seed's event, or NULL: 0.  A seed whose event would
                             run past the last layer gets no candidates: n_cands 0, hot_idx -9, score -1e16,
                             n_found 0 (its par/err/chg/chi2 slots are not written) */
} mkf_ce_batch;

/* MkBuilder::FindTracksCloneEngine / find_tracks_in_layers (MkBuilder.cc:1847-2060)
 * over the layers of `eoh` from first_layer on: up to max_cands candidates per
 * seed; per layer every candidate is propagated, its hits selected and scored
 * (FindCandidatesCloneEngine, MkFinder.cc:1114-1284), the best max_cands
 * (hit, candidate) combinations per seed become the next candidates
 * (CandCloner::ProcessSeedRange, CandCloner.cc:37-191) and are updated with
 * their new hit (UpdateWithLastHit, MkFinder.cc:1290-1308); the final order is
 * CombCandidate::MergeCandsAndBestShortOne (HitStructures.cc:511-552).  Hit
 * cluster charge and overlap (second-module) hits are not modelled: hits carry
 * no charge or module id.  Returns MKF_ERR_INVALID_ARG if a layer of `eoh`
 * holds 2^28 hits or more (hit records carry 28-bit positions in the layer). */
int mkf_find_tracks_clone_engine(mkf_context* ctx, mkf_event_of_hits* eoh, const mkf_ce_batch* batch, void* stream);

/* MkBuilder::FindTracksStandard (MkBuilder.cc:1643-1845) with MkFinder::FindCandidates (MkFinder.cc:902-1109),
 * same batch and outputs as mkf_find_tracks_clone_engine: per layer every (candidate, accepted hit) pair
 * becomes a new candidate -- the candidate's state at the layer propagated to the hit with intra_pflags and
 * updated with it, scored by getScoreCand with its updated pT and tail holes -- and every candidate inside
 * the layer also becomes one with the invalid hit and its state at the layer; these and the held-back
 * candidates are sorted by score, the best max_cands without a -2 hit are kept, and the first -2 one in that
 * order competes for the best short candidate.  Hit charge and overlap hits as in the clone engine. */
int mkf_find_tracks_standard(mkf_context* ctx, mkf_event_of_hits* eoh, const mkf_ce_batch* batch, void* stream);

/* Launch shape of the combinatorial finders (mkf_find_tracks_clone_engine,
 * mkf_find_tracks_standard), per finder: no counterpart in the reference, whose
 * Matriplex width and thread pool are fixed at build time; the shapes give the
 * same results.  slots: candidate slots per seed in a 64-lane wavefront, 6 (ten
 * seeds per wavefront) or 8 (eight), at least max_cands; 0 (the default): 6
 * when max_cands <= 6, else 8.  waves_per_simd: 2 (the default) or 1, the
 * occupancy the kernel's registers are allotted for. */
typedef struct mkf_find_shape {
  int slots;
  int waves_per_simd;
} mkf_find_shape;
int mkf_set_find_shape(mkf_context* ctx, int standard, const mkf_find_shape* shape);
int mkf_get_find_shape(mkf_context* ctx, int standard, mkf_find_shape* shape);
/* The finders' autotuner (as mkf_autotune_batched is the fit's): times every
 * shape valid for the batch's max_cands on `batch` (one warm-up launch and
 * `reps` timed launches each, on `stream`, synchronising it), keeps the fastest
 * for later calls of that finder on this context and returns it with its time
 * per launch.  The batch's outputs are those of a finding run (the same for
 * every shape). */
int mkf_autotune_find(mkf_context* ctx, mkf_event_of_hits* eoh, const mkf_ce_batch* batch, int standard, int reps,
                      mkf_find_shape* best, float* best_ms, void* stream);

/* ---- Backward fit of found candidates -------------------------------------- */

#define MKF_BKFIT_PCA           1 /* Config::includePCA: after the last layer, propagate to the PCA with pca_pflags
                                     (BkFitPropTracksToPCA) and output that state (BkFitOutputTracks' Par[iP]) */
#define MKF_BKFIT_FAKE_MATERIAL 2 /* fake hits take material: MkFinder::BkFitFitTracksBH (MkFinder.cc:1455-1570),
                                     the best-hit driver's form; without it MkFinder::BkFitFitTracks */

typedef struct mkf_bkfit_batch {
  int n_cands;
  const float* in_par;   /* [6][n]  the candidate's state after finding (a clone-engine slot: mkf_ce_batch par) */
  const float* in_err;   /* [21][n] scaled by 100 before the fit (BkFitInputTracks) */
  const int* in_chg;     /* [n] */
  const int* hot_idx;    /* [n_layers][n] the candidate's HitOnTrack per plan layer: hit index within the layer,
                            a fake-hit code < 0 (skipped), or -9 (no node); an index past the layer's hits counts
                            as a fake-hit code */
  int n_layers;          /* plan layers per candidate (hot_idx rows), 1..MKF_MAX_FIND_LAYERS */
  int last_layer;        /* SteeringParams::m_bkw_fit_last: the lowest plan layer visited (0: all) */
  const int* layer_base; /* [n] device: first layer of each candidate's event in eoh, or NULL: 0.  A candidate
                            whose plan would run past the last layer of eoh is skipped: none of its outputs is written */
  int group;             /* candidates fitted as one block, 1, 2, 4, ..., 64: the reference's NN-wide MkFinder
                            (MkBuilder::fit_cands, blocks of NN from the first candidate).  A block visits every
                            plan layer on which one of its candidates has a hit; the others of the block take a
                            fake hit {666,0,666,0,0,666} at their position there (MkFinder.cc:1635-1647) */
  int pflags;            /* Config::backward_fit_pflags */
  int pca_pflags;        /* Config::pca_prop_pflags */
  int flags;             /* MKF_BKFIT_* */
  float* par;            /* [6][n] out (may alias in_par) */
  float* err;            /* [21][n] out (may alias in_err) */
  int* chg;              /* [n] out (may alias in_chg) */
  float* chi2;           /* [n] out: sum of the update chi2 of every visited layer, fake hits included */
  float* score;          /* [n] or NULL; out: getScoreCand (HitStructures.h:574-585) with the HitOnTrack counts of
                            hot_idx (TrackCand::addHitIdx) and the fitted pT and chi2; left as it was where chi2 is
                            NaN */
} mkf_bkfit_batch;

/* MkBuilder::BackwardFit / fit_cands (MkBuilder.cc:2124-2253), or BackwardFitBH / fit_cands_BH (:2069-2122)
 * with MKF_BKFIT_FAKE_MATERIAL, over one candidate per seed: BkFitInputTracks (covariance x100, chi2 0;
 * MkFinder.cc:1368-1399), BkFitFitTracks over the plan layers in reverse (the layer's HitOnTrack hit, or a
 * fake hit; propagation to the hit's r or z with pflags, Kalman update, charge flip; MkFinder.cc:1573-1742),
 * optionally the PCA step (:1746-1749), BkFitOutputTracks (:1425-1446).  One GPU lane per candidate; the
 * lanes of a block share their per-layer decision through a wavefront ballot. */
int mkf_backward_fit(mkf_context* ctx, mkf_event_of_hits* eoh, const mkf_bkfit_batch* batch, void* stream);

#ifdef __cplusplus
}
#endif
#endif
