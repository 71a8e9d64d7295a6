// toymc.cc -- seeded toy Monte Carlo (host C++), see include/mkfit_toymc.h.
//
// Truth tracks are exact helices in a uniform Bz, stepped analytically from
// surface to surface (cylinder: circle-circle intersection; disk: linear in z),
// with a Highland multiple-scattering kick and a mean energy loss applied at
// every crossed layer.  All arithmetic is double precision.
#include "../../include/mkfit_toymc.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kC = 0.0029979245800;  // GeV / (T cm)
constexpr double kPi = 3.14159265358979323846;
constexpr double kMinCosCrossing = 0.3;

// splitmix64 + xoshiro256**: one stream per track, keyed by (seed, track).
inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
struct Rng {
  uint64_t s[4];
  Rng(uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull);
    for (auto& v : s) v = splitmix64(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double gauss() {                                                           // Box-Muller
    double u1 = uniform();
    while (u1 <= 0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * kPi * u2);
  }
};

struct State {
  double x, y, z, px, py, pz;
  int q;
};

struct Crossing {
  int layer;
  bool endcap;
  State s;
};

inline double wrap_2pi(double a) {
  a = std::fmod(a, 2 * kPi);
  return a < 0 ? a + 2 * kPi : a;
}

// Transverse path length to the first crossing of the cylinder radius r ahead of s
// (moving outward), or -1.
double path_to_cylinder(const State& s, double r, double kappa) {
  const double pt = std::hypot(s.px, s.py);
  const double psi = std::atan2(s.py, s.px);
  const double cx = s.x - std::sin(psi) / kappa, cy = s.y + std::cos(psi) / kappa;
  const double rho = 1.0 / std::fabs(kappa);
  const double d = std::hypot(cx, cy);
  if (d < 1e-12 || d + rho < r || std::fabs(rho - r) > d) return -1;
  const double a = (r * r - rho * rho + d * d) / (2 * d);
  const double h2 = r * r - a * a;
  const double h = h2 > 0 ? std::sqrt(h2) : 0;
  const double ux = cx / d, uy = cy / d;
  const double beta0 = std::atan2(s.y - cy, s.x - cx);
  double best = -1;
  for (int sgn = -1; sgn <= 1; sgn += 2) {
    const double Px = a * ux - sgn * h * uy, Py = a * uy + sgn * h * ux;
    const double beta = std::atan2(Py - cy, Px - cx);
    const double db = kappa > 0 ? wrap_2pi(beta - beta0) : wrap_2pi(beta0 - beta);
    const double l = db * rho;
    if (l <= 1e-9) continue;
    // direction at the crossing: must be moving outward
    const double psiP = psi + kappa * l;
    if (Px * std::cos(psiP) + Py * std::sin(psiP) <= 0) continue;
    if (best < 0 || l < best) best = l;
  }
  (void)pt;
  return best;
}

State advance(const State& s, double l, double kappa) {
  const double pt = std::hypot(s.px, s.py);
  const double psi = std::atan2(s.py, s.px);
  State o = s;
  o.x = s.x + (std::sin(psi + kappa * l) - std::sin(psi)) / kappa;
  o.y = s.y + (std::cos(psi) - std::cos(psi + kappa * l)) / kappa;
  o.z = s.z + l * s.pz / pt;
  o.px = pt * std::cos(psi + kappa * l);
  o.py = pt * std::sin(psi + kappa * l);
  return o;
}

// Highland scattering + mean energy loss (Bethe form of mkFit/PropagationMPlex.cc:716-767,
// with the physical sign: momentum decreases along the flight).
void apply_material(State& s, double xx0, double xi, const double n_hat[3], Rng& rng) {
  if (xx0 <= 0) return;
  const double p = std::sqrt(s.px * s.px + s.py * s.py + s.pz * s.pz);
  const double ux = s.px / p, uy = s.py / p, uz = s.pz / p;
  const double cosinc = std::fabs(ux * n_hat[0] + uy * n_hat[1] + uz * n_hat[2]);
  const double t = xx0 / std::max(cosinc, 1e-3);
  const double m = 0.140, me = 0.0005, I = 16.0e-9 * 10.75;
  const double E = std::sqrt(p * p + m * m);
  const double beta = p / E, beta2 = beta * beta, gamma = E / m, gamma2 = gamma * gamma;
  const double th0 = 0.0136 / (beta * p) * std::sqrt(t) * (1 + 0.038 * std::log(t));
  // two orthogonal directions perpendicular to u
  double ax = -uy, ay = ux, az = 0;
  double an = std::sqrt(ax * ax + ay * ay);
  if (an < 1e-9) { ax = 1; ay = 0; an = 1; }
  ax /= an; ay /= an;
  const double bx = uy * az - uz * ay, by = uz * ax - ux * az, bz = ux * ay - uy * ax;
  const double g1 = rng.gauss() * th0, g2 = rng.gauss() * th0;
  double vx = ux + g1 * ax + g2 * bx, vy = uy + g1 * ay + g2 * by, vz = uz + g1 * az + g2 * bz;
  const double vn = std::sqrt(vx * vx + vy * vy + vz * vz);
  vx /= vn; vy /= vn; vz /= vn;
  double pn = p;
  if (xi > 0) {
    const double wmax = 2 * me * beta2 * gamma2 / (1 + 2 * gamma * me / m + me * me / (m * m));
    const double deltahalf = std::log(28.816e-9 * std::sqrt(2.33 * 0.498) / I) + std::log(beta * gamma) - 0.5;
    const double dEdx = 2 * (xi / std::max(cosinc, 1e-3)) *
                        (0.5 * std::log(2 * me * beta2 * gamma2 * wmax / (I * I)) - beta2 - deltahalf) / beta2;
    pn = std::max(p - dEdx / beta, 0.05);
  }
  s.px = pn * vx; s.py = pn * vy; s.pz = pn * vz;
}

void ccs_of(const State& s, double out[6]) {
  const double pt = std::hypot(s.px, s.py);
  out[0] = s.x; out[1] = s.y; out[2] = s.z;
  out[3] = 1.0 / pt;
  out[4] = std::atan2(s.py, s.px);
  out[5] = std::atan2(pt, s.pz);
}

void gen_track(const mkf_geometry& g, const mkf_toymc_params& p, const mkf_toymc_output& o, int t) {
  const int n = o.n_tracks, S = o.n_slots;
  Rng rng(p.seed, (uint64_t)t);
  State s;
  s.q = (rng.next() >> 63) ? 1 : -1;
  const double pt = p.pt_min + (p.pt_max - p.pt_min) * rng.uniform();
  const double eta = -p.eta_max + 2 * p.eta_max * rng.uniform();
  const double phi = 2 * kPi * rng.uniform();
  s.x = rng.gauss() * p.vtx_sxy;
  s.y = rng.gauss() * p.vtx_sxy;
  s.z = rng.gauss() * p.vtx_sz;
  const double theta = 2 * std::atan(std::exp(-eta));
  s.px = pt * std::cos(phi);
  s.py = pt * std::sin(phi);
  s.pz = pt / std::tan(theta);

  const int nb = g.n_barrel, nd = g.n_disk;
  std::vector<Crossing> xs;
  std::vector<char> used(nb + 2 * nd, 0);
  const int cap = std::min(S, MKF_MAX_LAYERS * 2);
  while ((int)xs.size() < cap) {
    const double ptc = std::hypot(s.px, s.py);
    const double kappa = -s.q * kC * p.bfield / ptc;
    double best = -1;
    int bl = -1;
    const double rc = std::hypot(s.x, s.y);
    for (int i = 0; i < nb; ++i) {
      if (used[i] || g.barrel_r[i] <= rc + 1e-7) continue;
      const double l = path_to_cylinder(s, g.barrel_r[i], kappa);
      if (l < 0) continue;
      const double zz = s.z + l * s.pz / ptc;
      if (std::fabs(zz) > g.barrel_zmax[i]) continue;
      if (best < 0 || l < best) { best = l; bl = i; }
    }
    for (int i = 0; i < 2 * nd; ++i) {
      const int lid = nb + i;
      if (used[lid]) continue;
      const double zd = (i < nd ? 1.0 : -1.0) * g.disk_z[i % nd];
      if (s.pz == 0) continue;
      const double l = (zd - s.z) * ptc / s.pz;
      if (l <= 1e-9) continue;
      // only on the outgoing half-turn: no curlers reaching a disk after looping
      if (std::fabs(kappa * l) >= kPi) continue;
      const State a = advance(s, l, kappa);
      const double ra = std::hypot(a.x, a.y);
      if (ra < g.disk_rmin[i % nd] || ra > g.disk_rmax[i % nd]) continue;
      if (a.x * a.px + a.y * a.py < kMinCosCrossing * ra * std::hypot(a.px, a.py)) continue;
      if (best < 0 || l < best) { best = l; bl = lid; }
    }
    if (bl < 0) break;
    // Stop once the track has turned back inward (the fit follows the outgoing branch).
    State a = advance(s, best, kappa);
    // Stop near the apex of the transverse circle: the reference propagator
    // refuses steps whose 1/cos(position, momentum) exceeds 5
    // (mkFit/PropagationMPlex.icc:84-92); keep a margin below that.
    if (bl < nb && a.x * a.px + a.y * a.py < kMinCosCrossing * std::hypot(a.x, a.y) * std::hypot(a.px, a.py)) break;
    used[bl] = 1;
    Crossing c;
    c.layer = bl;
    c.endcap = bl >= nb;
    c.s = a;
    xs.push_back(c);
    if (p.simulate_material) {
      double nh[3];
      if (c.endcap) { nh[0] = 0; nh[1] = 0; nh[2] = 1; }
      else { const double r = std::hypot(a.x, a.y); nh[0] = a.x / r; nh[1] = a.y / r; nh[2] = 0; }
      apply_material(a, g.x_over_x0[bl], g.xi[bl], nh, rng);
    }
    s = a;
  }

  int nh = (int)xs.size();
  if (p.hits_max > 0) {
    const int lo = std::max(1, p.hits_min), hi = std::max(lo, p.hits_max);
    const int target = lo + (int)(rng.uniform() * (hi - lo + 1));
    nh = std::min(nh, target);
  }
  nh = std::min(nh, S);
  o.n_hits[t] = nh;
  o.chg[t] = s.q;

  for (int k = 0; k < S; ++k) {
    unsigned char fl = 0;
    float pos[3] = {0, 0, 0}, err[6] = {0, 0, 0, 0, 0, 0};
    int lyr = -1;
    double tr[6] = {0, 0, 0, 0, 0, 0};
    if (k < nh) {
      const Crossing& c = xs[k];
      lyr = c.layer;
      ccs_of(c.s, tr);
      const double phih = std::atan2(c.s.y, c.s.x);
      const double sp = std::sin(phih), cp = std::cos(phih);
      if (!c.endcap) {
        const double r = g.barrel_r[c.layer];
        const double ph = phih + rng.gauss() * p.sig_rphi / r;
        pos[0] = float(r * std::cos(ph));
        pos[1] = float(r * std::sin(ph));
        pos[2] = float(c.s.z + rng.gauss() * p.sig_z);
        const double v = double(p.sig_rphi) * p.sig_rphi;
        err[0] = float(v * sp * sp);
        err[1] = float(-v * sp * cp);
        err[2] = float(v * cp * cp);
        err[5] = float(double(p.sig_z) * p.sig_z);
      } else {
        fl |= 1;
        const double rh = std::hypot(c.s.x, c.s.y);
        const double r = rh + rng.gauss() * p.sig_r_endcap;
        const double ph = phih + rng.gauss() * p.sig_rphi / rh;
        pos[0] = float(r * std::cos(ph));
        pos[1] = float(r * std::sin(ph));
        pos[2] = float(c.s.z);
        const double vr = double(p.sig_r_endcap) * p.sig_r_endcap, vp = double(p.sig_rphi) * p.sig_rphi;
        err[0] = float(vr * cp * cp + vp * sp * sp);
        err[1] = float((vr - vp) * sp * cp);
        err[2] = float(vr * sp * sp + vp * cp * cp);
        err[5] = float(double(p.sig_z) * p.sig_z);
      }
      if (k > 0 && p.missing_frac > 0 && rng.uniform() < p.missing_frac) {
        fl |= 2;
        for (float& v : pos) v = 0;
        for (float& v : err) v = 0;
      }
    }
    for (int e = 0; e < 3; ++e) o.hit_pos[((size_t)k * 3 + e) * n + t] = pos[e];
    for (int e = 0; e < 6; ++e) o.hit_err[((size_t)k * 6 + e) * n + t] = err[e];
    if (o.hit_flags) o.hit_flags[(size_t)k * n + t] = fl;
    if (o.hit_layer) o.hit_layer[(size_t)k * n + t] = lyr;
    if (o.truth_par)
      for (int e = 0; e < 6; ++e) o.truth_par[((size_t)k * 6 + e) * n + t] = float(tr[e]);
  }

  // Seed: truth at the first crossing (or the vertex), smeared, with a diagonal covariance.
  double sd[6];
  if (!xs.empty()) ccs_of(xs[0].s, sd); else ccs_of(s, sd);
  const double sig[6] = {p.seed_sig_pos, p.seed_sig_pos, p.seed_sig_z, p.seed_rel_ipt * sd[3], p.seed_sig_ang,
                         p.seed_sig_ang};
  for (int e = 0; e < 6; ++e) {
    double v = sd[e] + rng.gauss() * sig[e];
    if (e == 3) v = std::fabs(v);
    if (e == 4) v = std::atan2(std::sin(v), std::cos(v));
    o.seed_par[(size_t)e * n + t] = float(v);
  }
  for (int e = 0; e < 21; ++e) o.seed_err[(size_t)e * n + t] = 0;
  for (int i = 0; i < 6; ++i) o.seed_err[(size_t)(i * (i + 1) / 2 + i) * n + t] = float(sig[i] * sig[i]);
}

}  // namespace

extern "C" {

void mkf_toymc_default_params(mkf_toymc_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->seed = 12345;
  p->bfield = 3.8f;
  p->pt_min = 0.5f;
  p->pt_max = 10.f;
  p->eta_max = 1.f;
  p->vtx_sxy = 0.01f;
  p->vtx_sz = 5.f;
  p->sig_rphi = 0.01f;
  p->sig_z = 0.1f;
  p->sig_r_endcap = 0.1f;
  p->simulate_material = 1;
  p->hits_min = 0;
  p->hits_max = 0;
  p->missing_frac = 0.f;
  p->seed_sig_pos = 0.01f;
  p->seed_sig_z = 0.1f;
  p->seed_rel_ipt = 0.1f;
  p->seed_sig_ang = 0.005f;
}

int mkf_toymc_geometry_preset(int which, mkf_geometry* g) {
  if (!g) return 1;
  std::memset(g, 0, sizeof(*g));
  if (which == 1) {
    g->n_barrel = 17;
    for (int i = 0; i < 17; ++i) { g->barrel_r[i] = 4.f + 6.f * i; g->barrel_zmax[i] = 120.f; }
  } else if (which == 2 || which == 3) {
    g->n_barrel = 25;
    for (int i = 0; i < 25; ++i) {
      g->barrel_r[i] = float(3.0 + 107.0 / 24.0 * i);
      g->barrel_zmax[i] = 120.f;
    }
    if (which == 3) {
      // Staged barrel: inner layers (r < 16 cm) are 27 cm long, outer ones 115 cm;
      // disks at |z| = 30..270 cm cover r from |z|/sinh(2.5) up to 16 cm (|z| <= 90)
      // or from 16 cm to 110 cm (|z| >= 120), so no surface overlaps another.
      for (int i = 0; i < 25; ++i) g->barrel_zmax[i] = g->barrel_r[i] < 16.f ? 27.f : 115.f;
      g->n_disk = 9;
      for (int i = 0; i < 9; ++i) {
        const float z = 30.f * (i + 1);
        g->disk_z[i] = z;
        const float rlo = float(z / std::sinh(2.5));
        if (z <= 90.f) { g->disk_rmin[i] = rlo; g->disk_rmax[i] = 16.f; }
        else { g->disk_rmin[i] = std::max(16.f, rlo); g->disk_rmax[i] = 110.f; }
      }
    }
  } else if (which == 4) {
    // Forward disks only: 9 disks at z = +-(30..270) cm, r from |z|/sinh(2.5) to 110 cm.
    g->n_disk = 9;
    for (int i = 0; i < 9; ++i) {
      const float z = 30.f * (i + 1);
      g->disk_z[i] = z;
      g->disk_rmin[i] = float(z / std::sinh(2.5));
      g->disk_rmax[i] = 110.f;
    }
  } else {
    return 2;
  }
  const int nl = g->n_barrel + 2 * g->n_disk;
  for (int l = 0; l < nl; ++l) { g->x_over_x0[l] = 0.02f; g->xi[l] = 0.02f * 2.2e-3f; }
  return 0;
}

int mkf_toymc_generate(const mkf_geometry* geo, const mkf_toymc_params* p, mkf_toymc_output* out, int n_threads) {
  if (!geo || !p || !out || out->n_tracks < 0 || out->n_slots < 1) return 1;
  if (geo->n_barrel + 2 * geo->n_disk > MKF_MAX_LAYERS) return 2;
  const int n = out->n_tracks;
  if (n_threads < 1) n_threads = 1;
  std::atomic<int> next{0};
  auto work = [&]() {
    for (;;) {
      const int t0 = next.fetch_add(256);
      if (t0 >= n) break;
      for (int t = t0; t < std::min(n, t0 + 256); ++t) gen_track(*geo, *p, *out, t);
    }
  };
  std::vector<std::thread> pool;
  for (int i = 1; i < n_threads; ++i) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  return 0;
}

// Finding events use either a barrel-only geometry (layers = the barrel layers)
// or a disk-only one (layers = the +z disks).
static int find_layers(const mkf_geometry* g) {
  if (g->n_barrel > 0 && g->n_disk == 0) return g->n_barrel;
  if (g->n_barrel == 0 && g->n_disk > 0) return g->n_disk;
  return 0;
}

int mkf_toymc_noise_count(const mkf_geometry* g, int layer, float noise_per_cm2) {
  if (!g || layer < 0 || layer >= find_layers(g) || noise_per_cm2 <= 0) return 0;
  const double area = g->n_barrel > 0 ? 2 * kPi * g->barrel_r[layer] * 2 * g->barrel_zmax[layer]
                                      : kPi * (double(g->disk_rmax[layer]) * g->disk_rmax[layer] -
                                               double(g->disk_rmin[layer]) * g->disk_rmin[layer]);
  return (int)std::llround(area * noise_per_cm2);
}

int mkf_toymc_find_event(const mkf_geometry* g, const mkf_toymc_params* p, float noise_per_cm2, int seed_layers,
                         mkf_toymc_find_output* o, int n_threads) {
  if (!g || !p || !o || o->n_tracks < 0 || find_layers(g) == 0 || seed_layers < 1 || seed_layers > find_layers(g))
    return 1;
  const int n = o->n_tracks, L = find_layers(g), cap = o->layer_capacity;
  const bool disks = g->n_barrel == 0;
  const int n_gen_layers = g->n_barrel + 2 * g->n_disk;
  for (int l = 0; l < L; ++l)
    if (n + mkf_toymc_noise_count(g, l, noise_per_cm2) > cap) return 2;
  // Tracks through the fit generator (one hit per crossed layer, in layer order).
  std::vector<float> seed_par(6 * (size_t)n), seed_err(21 * (size_t)n), hit_pos((size_t)n_gen_layers * 3 * n),
      hit_err((size_t)n_gen_layers * 6 * n), truth((size_t)n_gen_layers * 6 * n);
  std::vector<int> chg(n), nh(n), hit_layer((size_t)n_gen_layers * n);
  std::vector<unsigned char> flags((size_t)n_gen_layers * n);
  mkf_toymc_output to;
  to.n_tracks = n;
  to.n_slots = n_gen_layers;
  to.seed_par = seed_par.data();
  to.seed_err = seed_err.data();
  to.chg = chg.data();
  to.hit_pos = hit_pos.data();
  to.hit_err = hit_err.data();
  to.n_hits = nh.data();
  to.hit_flags = flags.data();
  to.hit_layer = hit_layer.data();
  to.truth_par = truth.data();
  mkf_toymc_params pp = *p;
  pp.hits_min = pp.hits_max = 0;
  pp.missing_frac = 0;
  int rc = mkf_toymc_generate(g, &pp, &to, n_threads);
  if (rc) return rc;
  // Signal hits per layer, in track order; then noise.
  std::vector<std::vector<int>> sig_index(L, std::vector<int>(n, -1));
  for (int l = 0; l < L; ++l) o->n_layer_hits[l] = 0;
  for (int t = 0; t < n; ++t)
    for (int k = 0; k < nh[t]; ++k) {
      const int l = hit_layer[(size_t)k * n + t];   // disk mode: +z disks are layers 0..n_disk-1
      if (l < 0 || l >= L) continue;
      const int i = o->n_layer_hits[l]++;
      float* h = o->hits + ((size_t)l * cap + i) * 9;
      for (int e = 0; e < 3; ++e) h[e] = hit_pos[((size_t)k * 3 + e) * n + t];
      for (int e = 0; e < 6; ++e) h[3 + e] = hit_err[((size_t)k * 6 + e) * n + t];
      o->hit_mc[(size_t)l * cap + i] = t;
      sig_index[l][t] = i;
    }
  for (int l = 0; l < L; ++l) {
    Rng rng(p->seed ^ 0x5bd1e995u, (uint64_t)(1u << 30) + l);
    const int nn = mkf_toymc_noise_count(g, l, noise_per_cm2);
    const double vr = double(p->sig_rphi) * p->sig_rphi, vz = double(p->sig_z) * p->sig_z;
    for (int k = 0; k < nn; ++k) {
      const double ph = -kPi + 2 * kPi * rng.uniform(), u = rng.uniform();
      double r, z;
      if (disks) {   // uniform on the annulus
        const double r0 = g->disk_rmin[l], r1 = g->disk_rmax[l];
        r = std::sqrt(r0 * r0 + u * (r1 * r1 - r0 * r0));
        z = g->disk_z[l];
      } else {
        r = g->barrel_r[l];
        z = -g->barrel_zmax[l] + 2 * g->barrel_zmax[l] * u;
      }
      const int i = o->n_layer_hits[l]++;
      float* h = o->hits + ((size_t)l * cap + i) * 9;
      const double sp = std::sin(ph), cp = std::cos(ph);
      h[0] = float(r * cp); h[1] = float(r * sp); h[2] = float(z);
      if (disks) {   // radial and r-phi resolution in the disk plane, as the signal hits
        const double vrr = double(p->sig_r_endcap) * p->sig_r_endcap;
        h[3] = float(vrr * cp * cp + vr * sp * sp); h[4] = float((vrr - vr) * sp * cp);
        h[5] = float(vrr * sp * sp + vr * cp * cp);
      } else {
        h[3] = float(vr * sp * sp); h[4] = float(-vr * sp * cp); h[5] = float(vr * cp * cp);
      }
      h[6] = 0.f; h[7] = 0.f; h[8] = float(vz);
      o->hit_mc[(size_t)l * cap + i] = -1;
    }
  }
  // Seeds: tracks with hits on all seed layers; state = truth at the last seed
  // hit, smeared with the fit-seed widths.
  int ns = 0;
  for (int t = 0; t < n; ++t) {
    bool ok = true;
    for (int l = 0; l < seed_layers; ++l) ok = ok && sig_index[l][t] >= 0;
    if (!ok) continue;
    Rng rng(p->seed ^ 0x27d4eb2fu, (uint64_t)t);
    int k = -1;   // slot of the hit on the last seed layer
    for (int s = 0; s < nh[t]; ++s)
      if (hit_layer[(size_t)s * n + t] == seed_layers - 1) k = s;
    double sd[6];
    for (int e = 0; e < 6; ++e) sd[e] = truth[((size_t)k * 6 + e) * n + t];
    const double sig[6] = {p->seed_sig_pos, p->seed_sig_pos, p->seed_sig_z, p->seed_rel_ipt * sd[3], p->seed_sig_ang,
                           p->seed_sig_ang};
    for (int e = 0; e < 6; ++e) {
      double v = sd[e] + rng.gauss() * sig[e];
      if (e == 3) v = std::fabs(v);
      if (e == 4) v = std::atan2(std::sin(v), std::cos(v));
      o->seed_par[(size_t)e * n + ns] = float(v);
    }
    for (int e = 0; e < 21; ++e) o->seed_err[(size_t)e * n + ns] = 0.f;
    for (int i = 0; i < 6; ++i) o->seed_err[(size_t)(i * (i + 1) / 2 + i) * n + ns] = float(sig[i] * sig[i]);
    o->seed_chg[ns] = chg[t];
    for (int l = 0; l < seed_layers; ++l) o->seed_hits[(size_t)l * n + ns] = sig_index[l][t];
    o->seed_mc[ns] = t;
    ++ns;
  }
  o->n_seeds = ns;
  return 0;
}

int mkf_toymc_material_grid(const mkf_geometry* g, float* rl, float* xi) {
  if (!g || !rl || !xi) return 1;
  const int NZ = 300, NR = 120;
  std::fill(rl, rl + NZ * NR, 0.f);
  std::fill(xi, xi + NZ * NR, 0.f);
  for (int i = 0; i < g->n_barrel; ++i) {
    const float r = g->barrel_r[i];
    const int rb0 = std::max(0, (int)std::floor(r - 0.5f)), rb1 = std::min(NR - 1, (int)std::floor(r + 0.5f));
    const int zb1 = std::min(NZ - 1, (int)std::floor(g->barrel_zmax[i]));
    for (int zb = 0; zb <= zb1; ++zb)
      for (int rb = rb0; rb <= rb1; ++rb) { rl[zb * NR + rb] = g->x_over_x0[i]; xi[zb * NR + rb] = g->xi[i]; }
  }
  for (int i = 0; i < g->n_disk; ++i) {
    const int lid = g->n_barrel + i;  // +z and -z disks share |z| bins and values
    const float z = g->disk_z[i];
    const int zb0 = std::max(0, (int)std::floor(z - 0.5f)), zb1 = std::min(NZ - 1, (int)std::floor(z + 0.5f));
    const int rb0 = std::max(0, (int)std::floor(g->disk_rmin[i])), rb1 = std::min(NR - 1, (int)std::floor(g->disk_rmax[i]));
    for (int zb = zb0; zb <= zb1; ++zb)
      for (int rb = rb0; rb <= rb1; ++rb) { rl[zb * NR + rb] = g->x_over_x0[lid]; xi[zb * NR + rb] = g->xi[lid]; }
  }
  return 0;
}

}  // extern "C"
