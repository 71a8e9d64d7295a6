/* mkfit_toymc.h -- seeded toy Monte Carlo for the Kalman fit and hit finding.
 *
 * The reference's own simulator was removed (README.md:433: "currently defunct
 * Simulation"), so this library generates its inputs: tracks from a helix in a
 * uniform Bz through concentric barrel cylinders and endcap disks, with
 * multiple scattering and energy loss in thin layers, Gaussian hit smearing,
 * and seeds built from the truth state at the first crossed layer.
 *
 * Output arrays follow the fit's element-major layout: element e of track t at
 * a[e*n + t]; per-hit arrays are [slot][element][track].  Track parameters are
 * the reference's CCS form (x, y, z, 1/pT, phi, theta) (mkFit/PropagationMPlex.icc:38-42),
 * covariances packed lower-triangle (Matriplex/MatriplexSym.h:15-22), hit
 * covariances (xx, yx, yy, zx, zy, zz) as in Hit.h:165-183.
 *
 * Host-only C code; deterministic for a given seed and independent of the
 * thread count (one counter-based stream per track).
 */
#ifndef MKFIT_TOYMC_H
#define MKFIT_TOYMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MKF_MAX_LAYERS 64

/* Layer ids: barrel 0..n_barrel-1, then +z disks, then -z disks. */
typedef struct mkf_geometry {
  int n_barrel;
  float barrel_r[MKF_MAX_LAYERS];
  float barrel_zmax[MKF_MAX_LAYERS];
  int n_disk; /* disks per side */
  float disk_z[MKF_MAX_LAYERS]; /* |z| of each disk */
  float disk_rmin[MKF_MAX_LAYERS];
  float disk_rmax[MKF_MAX_LAYERS];
  float x_over_x0[MKF_MAX_LAYERS]; /* per layer id */
  float xi[MKF_MAX_LAYERS];        /* per layer id, GeV (energy-loss scale as in Config::Xi) */
} mkf_geometry;

typedef struct mkf_toymc_params {
  uint64_t seed;
  float bfield;          /* T */
  float pt_min, pt_max;  /* GeV, uniform */
  float eta_max;         /* |eta| uniform below this */
  float vtx_sxy, vtx_sz; /* cm, Gaussian vertex */
  float sig_rphi, sig_z; /* barrel hit resolution, cm */
  float sig_r_endcap;    /* endcap radial resolution, cm (r-phi uses sig_rphi) */
  int simulate_material; /* multiple scattering + energy loss in the truth */
  int hits_min, hits_max;/* if hits_max > 0: keep U{hits_min..hits_max} crossed layers */
  float missing_frac;    /* probability a crossed layer (not the first) has no hit */
  float seed_sig_pos, seed_sig_z, seed_rel_ipt, seed_sig_ang; /* seed smearing and covariance */
} mkf_toymc_params;

typedef struct mkf_toymc_output {
  int n_tracks, n_slots;
  float* seed_par;         /* [6][n] */
  float* seed_err;         /* [21][n] */
  int* chg;                /* [n] */
  float* hit_pos;          /* [n_slots][3][n] */
  float* hit_err;          /* [n_slots][6][n] */
  int* n_hits;             /* [n] */
  unsigned char* hit_flags;/* [n_slots][n] or NULL: bit0 endcap, bit1 missing */
  int* hit_layer;          /* [n_slots][n] or NULL */
  float* truth_par;        /* [n_slots][6][n] or NULL: true CCS state at each crossing */
} mkf_toymc_output;

/* One event for track finding: every layer's hits (signal hits of the
 * generated tracks plus noise hits uniform on the layer surface) in the
 * reference's Hit order -- pos x, y, z then the packed 3x3 covariance
 * (Hit.h:165-183) -- and one seed per track built from its hits on the first
 * `seed_layers` layers (state: truth at the last seed hit, smeared as for the
 * fit seeds).  Tracks that do not cross all seed layers get no seed.
 * Barrel-only geometries (layers = barrel layers) or disk-only ones (layers =
 * the +z disks; preset 4). */
typedef struct mkf_toymc_find_output {
  int n_tracks;            /* generated tracks (capacity of the seed arrays) */
  int layer_capacity;      /* hits per layer the caller allocated */
  float* hits;             /* [n_layers][layer_capacity][9] */
  int* n_layer_hits;       /* [n_layers] filled */
  int* hit_mc;             /* [n_layers][layer_capacity]: generating track, -1 for noise */
  int n_seeds;             /* out */
  float* seed_par;         /* [6][n_tracks] (first n_seeds used) */
  float* seed_err;         /* [21][n_tracks] */
  int* seed_chg;           /* [n_tracks] */
  int* seed_hits;          /* [seed_layers][n_tracks]: hit index on each seed layer */
  int* seed_mc;            /* [n_tracks]: generating track of each seed */
} mkf_toymc_find_output;

int mkf_toymc_find_event(const mkf_geometry* geo, const mkf_toymc_params* p, float noise_per_cm2, int seed_layers,
                         mkf_toymc_find_output* out, int n_threads);
/* Noise hits that mkf_toymc_find_event puts on barrel layer `layer`. */
int mkf_toymc_noise_count(const mkf_geometry* geo, int layer, float noise_per_cm2);

/* Presets: 1 = 17 barrel layers r = 4 + 6i cm, |z| <= 120;
 *          2 = 25 barrel layers r = 3 + (107/24)i cm, |z| <= 120;
 *          3 = the 25 barrel layers with staged lengths plus 2 x 9 disks at |z| = 30..270 cm;
 *          4 = 2 x 9 disks only, |z| = 30..270 cm, r from |z|/sinh(2.5) to 110 cm.
 * x_over_x0 = 0.02 and xi = 0.02 * 2.2e-3 GeV on every layer. */
int mkf_toymc_geometry_preset(int which, mkf_geometry* geo);
void mkf_toymc_default_params(mkf_toymc_params* p);
int mkf_toymc_generate(const mkf_geometry* geo, const mkf_toymc_params* p, mkf_toymc_output* out, int n_threads);
/* Fills a [300][120] (|z| bins of 1 cm, r bins of 1 cm) material grid with each
 * layer's x/X0 and xi, the form the fit's material lookup reads
 * (Config.h:123-126,164-165; MaterialEffects.h:8-11). */
int mkf_toymc_material_grid(const mkf_geometry* geo, float* rl_grid, float* xi_grid);

#ifdef __cplusplus
}
#endif
#endif
