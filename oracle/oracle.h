/* oracle.h -- C entry points of the CPU oracle (test infrastructure only).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this library, and only as the checker / CPU baseline.  The product library
 * (mkfit_amd) never links it.
 *
 * All batch arrays are element-major with leading dimension n (the track
 * count): element e of track t is at a[e*n + t].  Covariances are the packed
 * lower triangle of the symmetric 6x6 (21 floats); hit covariances the packed
 * lower triangle of the 3x3 (6 floats: xx, yx, yy, zx, zy, zz).
 */
#ifndef MKFIT_ORACLE_H
#define MKFIT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx {
  float bfield;          /* uniform Bz in T */
  const float* rl_grid;  /* [300][120] or NULL: no material anywhere */
  const float* xi_grid;  /* [300][120] */
} orc_ctx;

/* Fit flags. */
#define ORC_FIT_CHI2        1 /* accumulate chi2 per hit (MkFinder::BkFitFitTracks) */
#define ORC_FIT_CHARGE_FLIP 2 /* flip charge when 1/pT turns negative after update */
#define ORC_FIT_FP64_FALLBACK 4 /* recompute an ill-conditioned updated covariance in FP64 (Joseph form) */
#define ORC_FIT_PROP_TO_PCA 8 /* MkBase::PropagateTracksToPCAZ after the last hit (BkFitPropTracksToPCA) */

/* Hit-slot flags. */
#define ORC_HIT_ENDCAP  1 /* propagate to the hit z and use the endcap update */
#define ORC_HIT_MISSING 2 /* no hit on this layer: fake hit at the track position */

typedef struct orc_fit_batch_f32 {
  int n_tracks, n_slots;
  float* par;  /* [6][n]  in: seed state, out: fitted state */
  float* err;  /* [21][n] */
  int* chg;    /* [n] */
  float* chi2; /* [n] accumulated (in/out) */
  const float* hit_pos;           /* [n_slots][3][n] */
  const float* hit_err;           /* [n_slots][6][n] */
  const int* n_hits;              /* [n] */
  const unsigned char* hit_flags; /* [n_slots][n] or NULL */
  int pflags;                     /* 1 = parametrised B, 2 = material */
  int fit_flags;                  /* ORC_FIT_* */
  int pca_pflags;                 /* flags of the propagation to the PCA (ORC_FIT_PROP_TO_PCA) */
} orc_fit_batch_f32;

typedef struct orc_fit_batch_f64 {
  int n_tracks, n_slots;
  double* par;
  double* err;
  int* chg;
  double* chi2;
  const float* hit_pos;
  const float* hit_err;
  const int* n_hits;
  const unsigned char* hit_flags;
  int pflags;
  int fit_flags;
  int pca_pflags;
} orc_fit_batch_f64;

int orc_simd_width(void);

int orc_propagate_to_r_f32(const orc_ctx* ctx, int n, const float* in_err, const float* in_par, const int* in_chg,
                           const float* ms_rad, float* out_err, float* out_par, int pflags, const int* no_mat,
                           int* fail_out);
int orc_propagate_to_r_f64(const orc_ctx* ctx, int n, const double* in_err, const double* in_par, const int* in_chg,
                           const double* ms_rad, double* out_err, double* out_par, int pflags, const int* no_mat,
                           int* fail_out);
int orc_propagate_to_z_f32(const orc_ctx* ctx, int n, const float* in_err, const float* in_par, const int* in_chg,
                           const float* ms_z, float* out_err, float* out_par, int pflags, const int* no_mat);
int orc_propagate_to_z_f64(const orc_ctx* ctx, int n, const double* in_err, const double* in_par, const int* in_chg,
                           const double* ms_z, double* out_err, double* out_par, int pflags, const int* no_mat);
/* Helix step only: parameters and the 6x6 Jacobian [36][n] (row-major element). */
int orc_helix_at_r_f32(const orc_ctx* ctx, int n, const float* in_par, const int* in_chg, const float* ms_rad,
                       float* out_par, float* out_jac, int* fail_out, int pflags);
int orc_helix_at_r_f64(const orc_ctx* ctx, int n, const double* in_par, const int* in_chg, const double* ms_rad,
                       double* out_par, double* out_jac, int* fail_out, int pflags);
int orc_helix_at_z_f32(const orc_ctx* ctx, int n, const float* in_par, const int* in_chg, const float* ms_z,
                       float* out_par, float* out_jac, int pflags);
int orc_helix_at_z_f64(const orc_ctx* ctx, int n, const double* in_par, const int* in_chg, const double* ms_z,
                       double* out_par, double* out_jac, int pflags);
int orc_kalman_op_f32(int op, int endcap, int n, const float* ps_err, const float* ps_par, const float* ms_err,
                      const float* ms_par, float* out_err, float* out_par, float* out_chi2);
int orc_kalman_op_f64(int op, int endcap, int n, const double* ps_err, const double* ps_par, const double* ms_err,
                      const double* ms_par, double* out_err, double* out_par, double* out_chi2);

/* Whole fit, one thread per chunk of two NN blocks (the reference's TBB grain). */
int orc_fit_f32(const orc_ctx* ctx, const orc_fit_batch_f32* b, int n_threads);
int orc_fit_f64(const orc_ctx* ctx, const orc_fit_batch_f64* b, int n_threads);
int orc_fit_ties_f32(const orc_ctx* c, const orc_fit_batch_f32* b, int n_threads, float* ties);
int orc_fit_tie_shift_f32(const orc_ctx* c, const orc_fit_batch_f32* b, int n_threads, float tie_shift);
int orc_set_tie_shift(float tie_shift);

/* ---- Fit as a function of per-layer material and alignment (oracle_align.cc) */

typedef struct orc_align_batch {
  int n_tracks, n_layers;  /* slot l of every track is its hit on barrel layer l */
  const float* seed_par;   /* [6][n] */
  const float* seed_err;   /* [21][n] */
  const int* seed_chg;     /* [n] */
  const float* hit_pos;    /* [n_layers][3][n] */
  const float* hit_err;    /* [n_layers][6][n] */
  const int* n_hits;       /* [n] */
  int pflags;              /* ORC PF flags; material from theta when PF_apply_material */
  float xi_per_x0;         /* xi = xi_per_x0 * x/X0 */
} orc_align_batch;

/* theta [n_layers][3] = (x/X0, d_rphi, d_z); chi2 [n] = sum of the update chi2 of each track. */
int orc_align_chi2_f32(const orc_ctx* ctx, const orc_align_batch* b, const float* theta, double* chi2, int n_threads);
int orc_align_chi2_f64(const orc_ctx* ctx, const orc_align_batch* b, const double* theta, double* chi2,
                       int n_threads);

/* ---- Best-hit track finding (oracle_find.cc) ------------------------------ */

typedef struct orc_layer_desc {
  int is_barrel;
  float rin, rout, zmin, zmax; /* LayerInfo limits */
  float propagate_to;          /* LayerInfo::m_propagate_to */
  float q_bin;                 /* > 0 bin width, < 0 number of bins */
  float min_dphi, max_dphi, min_dq, max_dq;          /* IterationLayerConfig selection limits */
  float c_dp_sf, c_dp_0, c_dp_1, c_dp_2;             /* dynamic windows (IterationConfig.h:70-86) */
  float c_dq_sf, c_dq_0, c_dq_1, c_dq_2;
  float c_c2_sf, c_c2_0, c_c2_1, c_c2_2;
} orc_layer_desc;

typedef struct orc_find_event {
  int n_layers;
  const orc_layer_desc* layers;  /* [n_layers] */
  const int* hit_offset;         /* [n_layers + 1] */
  const float* hits;             /* [total][9] x y z xx yx yy zx zy zz */
  int first_layer;               /* first layer searched (the seed hits fill the ones before) */
  float chi2_cut;                /* IterationParams::chi2Cut_min */
  int pflags;                    /* finding_inter_layer_pflags */
  int prop_to_hit;               /* best hit: finding_requires_propagation_to_hit_pos (clone engine: see orc_ce_params) */
  int intra_pflags;              /* best hit: finding_intra_layer_pflags */
  const unsigned char* hit_mask; /* [total] per hit in hit order, nonzero: taken by an earlier iteration, never
                                    selected (MkFinder.cc:476-482); NULL: none */
  int n_dead;                    /* dead-module regions (DeadRegion, Hit.h:333-337) marked in their layers' */
  const float* dead_regions;     /* bins (LayerOfHits::SuckInDeads): [n_dead][5] layer, phi1, phi2, q1, q2 */
  const unsigned char* hit_charge; /* [total] per hit in hit order, Hit::PackedData's 8-bit charge code (Hit.h:251-274):
                                      the combinatorial finders' strip charge cut (MkFinder.cc:867-896) applies;
                                      NULL: every hit passes it (as code 0xff, pixel-like) */
} orc_find_event;

typedef struct orc_find_tracks {
  int n_tracks;
  float* par;     /* [6][n] in: seed state, out: state at the end */
  float* err;     /* [21][n] */
  int* chg;       /* [n] */
  float* chi2;    /* [n] */
  int* hot_idx;   /* [n_layers][n] in: seed hit indices on layers < first_layer; out: hit index per layer,
                     -1/-2/-3 fake-hit codes (index-desc.txt), -9 layer not reached */
  int* n_found;   /* [n] out */
  float* margins; /* optional out [3][n_layers][n], the checker's decision margins per layer searched: nearest
                     window edge to a hit of the visited bins (cm), nearest bin edge to a window limit or a
                     visited hit's phi (fraction of a bin), smallest relative chi2 gap (best vs other, any vs
                     cut); 1e30 where the layer was not searched */
  int* work;      /* optional out [3][n]: per track the algorithmic work units of the search -- propagations to a
                     layer, window hits evaluated (chi2, after a propagation to the hit with prop_to_hit), Kalman
                     updates -- counted over the layers where the track is live */
} orc_find_tracks;

int orc_loh_geometry(const orc_layer_desc* d, int* nq, float* qmin, float* qmax, float* fq);
/* LayerOfHits of one layer: ranks [n] (sorted -> original), bins [nq*256][2], phis/qs [n] sorted. */
int orc_loh_build(const orc_layer_desc* d, int n_hits, const float* hits, int* ranks, int* bins, float* phis,
                  float* qs);
/* SelectHitIndices on states already at the layer: xsize [n], xarr [16][n] (-1 padded), wsr [n]. */
/* mask: [n_hits] or NULL, as orc_find_event's hit_mask for the layer */
int orc_select_hit_indices_masked(const orc_layer_desc* d, int n_hits, const float* hits, const unsigned char* mask,
                                  int n, const float* par, const float* err, int* xsize, int* xarr, int* wsr);
int orc_select_hit_indices(const orc_layer_desc* d, int n_hits, const float* hits, int n, const float* par,
                           const float* err, int* xsize, int* xarr, int* wsr);
int orc_find_best_hit(const orc_ctx* ctx, const orc_find_event* ev, orc_find_tracks* tracks, int n_threads);

/* Per-layer steps on MkFinder lane state (host arrays, element-major [k][n]). */
typedef struct orc_finder_lanes {
  int n;
  const float* prop_par;   /* [6][n]  state at the layer (Par[iP]) */
  const float* prop_err;   /* [21][n] */
  float* par;              /* [6][n]  out: updated state (Par[iC]) */
  float* err;              /* [21][n] */
  int* chg;                /* [n] in/out */
  float* chi2;             /* [n] in/out */
  int* n_found;            /* [n] in/out */
  int* n_inside_minus_one; /* [n] in/out */
  int* n_tail_minus_one;   /* [n] in/out */
} orc_finder_lanes;

typedef struct orc_idx_chi2 {  /* IdxChi2List (Track.h:28-42) + position in the lane's XHitArr (-1: invalid hit) */
  int trk_idx, hit_idx;
  unsigned module;
  int nhits, ntailholes, noverlaps, nholes;
  float pt, chi2, chi2_hit, score;
  int xhit_pos;
} orc_idx_chi2;

/* AddBestHit on `layer` of ev (hits, chi2_cut, prop_to_hit, intra_pflags) for lanes with select results
 * xsize [n], xarr [16][n], wsr [n]; hit_code [n] = HitOnTrack index added, -9 none (window outside). */
int orc_add_best_hit(const orc_ctx* ctx, const orc_find_event* ev, int layer, const orc_finder_lanes* lanes,
                     const int* xsize, const int* xarr, const int* wsr, int* hit_code);

/* ---- Clone-engine track finding (oracle_ce.cc) ---------------------------- */

typedef struct orc_ce_params {
  int max_cands;        /* IterationParams::maxCandsPerSeed (<= 16) */
  int max_holes;        /* maxHolesPerCand */
  int max_consec_holes; /* maxConsecHoles */
  int prop_to_hit;      /* Config::finding_requires_propagation_to_hit_pos */
  int intra_pflags;     /* Config::finding_intra_layer_pflags (chi2 propagation to the hit) */
} orc_ce_params;

typedef struct orc_ce_tracks {
  int n_seeds;
  const float* seed_par; /* [6][n] */
  const float* seed_err; /* [21][n] */
  const int* seed_chg;   /* [n] */
  const int* seed_hits;  /* [first_layer][n] */
  /* out, per seed the candidates after the final merge, best first (K = max_cands) */
  int* n_cands;          /* [n] */
  float* par;            /* [K][6][n] */
  float* err;            /* [K][21][n] */
  int* chg;              /* [K][n] */
  float* chi2;           /* [K][n] */
  float* score;          /* [K][n] */
  int* n_found;          /* [K][n] */
  int* hot_idx;          /* [K][n_layers][n]: hit index per layer, fake codes, -9 none */
  float* margins;        /* optional out [4][n_layers][n], the checker's decision margins per layer: window
                            edge (cm), bin edge (bin fraction), chi2 against the cut (relative), smallest
                            gap of the scores the selection compares relative to their chi2 sums; 1e30
                            where none */
  float* merge_gap;      /* [n] with margins: the same for the final merge */
  int* work;             /* optional out [3][n]: per seed the algorithmic work units -- candidates propagated to a
                            layer, window hits evaluated (propagation to the hit and chi2), candidates updated
                            with their new hit */
} orc_ce_tracks;

int orc_find_clone_engine(const orc_ctx* ctx, const orc_find_event* ev, const orc_ce_params* p, orc_ce_tracks* t,
                          int n_threads);
/* Standard combinatorial finding (MkBuilder::FindTracksStandard, MkFinder::FindCandidates) with the same inputs
 * and outputs; p->intra_pflags also serves the update of a candidate with its new hit. */
int orc_find_standard(const orc_ctx* ctx, const orc_find_event* ev, const orc_ce_params* p, orc_ce_tracks* t,
                      int n_threads);

/* FindCandidatesCloneEngine on `layer` (reads prop_*, chg, chi2 and the counters of `lanes`; the chi2 cut
 * from ev, the rest from p): n_entries [n], entries [n][17] per lane in XHitArr order, invalid record last. */
int orc_find_candidates_ce(const orc_ctx* ctx, const orc_find_event* ev, const orc_ce_params* p, int layer,
                           const orc_finder_lanes* lanes, const int* xsize, const int* xarr, const int* wsr,
                           int* n_entries, orc_idx_chi2* entries);
/* UpdateWithLastHit on `layer`: update prop_* with hit hit_idx[t] (ev->pflags for the propagation to the hit
 * when p->prop_to_hit) into par / err / chg; lanes with hit_idx outside the layer copy prop_* unchanged. */
/* Checker hook: passStripChargePCMfromTrack (MkFinder.cc:867-896) for one state par[6] and hit covariance herr[6]
   given the hit's 8-bit charge code: 1 when the hit passes; *qcorr the corrected charge per cm. */
int orc_pass_strip_charge(int barrel, unsigned code, const float* par, const float* herr, float* qcorr);
int orc_update_with_last_hit(const orc_ctx* ctx, const orc_find_event* ev, const orc_ce_params* p, int layer,
                             const orc_finder_lanes* lanes, const int* hit_idx);

/* ---- Backward fit of found candidates (oracle_bkfit.cc) ------------------- */

typedef struct orc_bkfit_params {
  int group;         /* candidates fitted as one block (the reference's NN-wide MkFinder), >= 1 */
  int pflags;        /* Config::backward_fit_pflags */
  int pca_pflags;    /* Config::pca_prop_pflags */
  int include_pca;   /* Config::includePCA: propagate to the PCA and output that state */
  int fake_material; /* 0: BkFitFitTracks (fake hits without material); 1: BkFitFitTracksBH (with) */
  int last_layer;    /* SteeringParams::m_bkw_fit_last: the lowest plan index visited */
} orc_bkfit_params;

typedef struct orc_bkfit_tracks_f32 {
  int n, n_layers;
  const int* hot_idx;    /* [n_layers][n] HitOnTrack index per plan layer (hit index in the layer, fake code < 0,
                            -9 no node); indices past the layer's hits count as fake codes */
  const int* layer_base; /* [n] first layer of each candidate's event in ev, or NULL: 0; candidates whose plan
                            runs past ev's layers are left untouched */
  float* par;            /* [6][n]  in: the candidate's state; out: fitted (at the PCA with include_pca) */
  float* err;            /* [21][n] */
  int* chg;              /* [n] */
  float* chi2;           /* [n] out */
  float* score;          /* [n] or NULL: out getScoreCand unless chi2 is NaN */
} orc_bkfit_tracks_f32;

typedef struct orc_bkfit_tracks_f64 {
  int n, n_layers;
  const int* hot_idx;
  const int* layer_base;
  double* par;
  double* err;
  int* chg;
  double* chi2;
  double* score;
} orc_bkfit_tracks_f64;

/* ties: NULL or [3][n] out, the threshold margins of orc_fit_ties_f32; tie_shift: as orc_fit_tie_shift_f32. */
int orc_backward_fit_f32(const orc_ctx* ctx, const orc_find_event* ev, const orc_bkfit_params* p,
                         const orc_bkfit_tracks_f32* t, int n_threads, float* ties, float tie_shift);
int orc_backward_fit_f64(const orc_ctx* ctx, const orc_find_event* ev, const orc_bkfit_params* p,
                         const orc_bkfit_tracks_f64* t, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
