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

#ifdef __cplusplus
}
#endif
#endif
