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
  int blocks_per_cu;  /* persistent grid = blocks_per_cu * CU count */
  int variant;        /* kernel variant index (occupancy / register bound) */
} mkf_launch_shape;

int mkf_set_batched_shape(mkf_context* ctx, const mkf_launch_shape* shape);
int mkf_get_batched_shape(mkf_context* ctx, mkf_launch_shape* shape);
/* Times every launch shape on `batch` (restoring the batch's state between
 * trials) and keeps the fastest for subsequent mkf_fit_batched calls.
 * scratch_state must hold a copy area of 29*n_tracks floats (device). */
int mkf_autotune_batched(mkf_context* ctx, const mkf_fit_batch* batch, float* scratch_state, int reps,
                         mkf_launch_shape* best, float* best_ms);

/* Number of variants mkf_set_batched_shape accepts. */
int mkf_num_batched_variants(void);

#ifdef __cplusplus
}
#endif
#endif
