// event_file.cc -- reader and writer of mkFit's binary event files (the
// layout is described in include/mkfit_event_file.h).  Host code: an event is
// read whole (its byte size leads it) and parsed with bounds checks, so a
// truncated or foreign file gives an error instead of a wild read.
#include "../../include/mkfit_event_file.h"

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int kMagic = 0xBEEF;
constexpr int kVersion = 6, kMinVersion = 4;
constexpr int kTrackBytes = 160, kHitBytes = 44, kHotBytes = 4, kStateBytes = 112, kMcInfoBytes = 16;
constexpr int kBeamSpotFloats = 8;
// Record offsets (bytes) within Track / TrackState / Hit.
constexpr int kTsPar = 0, kTsErr = 24, kTsChg = 108, kTsValid = 110;
constexpr int kTrChi2 = 112, kTrScore = 116, kTrLastHit = 120, kTrNFound = 122, kTrStatus = 124, kTrLabel = 128;
constexpr int kHitPos = 0, kHitErr = 12, kHitMc = 36, kHitPacked = 40;
const float kBeamSpotDefault[kBeamSpotFloats] = {0.f, 0.f, 0.f, 5.f, 5e-4f, 5e-4f, 0.f, 0.f};

thread_local std::string g_err;

int fail(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return 1;
}

template <class T>
T get(const unsigned char* p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}
template <class T>
void put(unsigned char* p, T v) {
  std::memcpy(p, &v, sizeof(T));
}

// HitOnTrack: index in the low 24 bits, layer in the high 8, both signed.
void unpack_hot(uint32_t w, int& index, int& layer) {
  index = (int)(w << 8) >> 8;
  layer = (int)w >> 24;
}
uint32_t pack_hot(int index, int layer) { return ((uint32_t)index & 0xffffffu) | ((uint32_t)layer << 24); }

struct TrackSet {
  int n = 0;
  std::vector<float> par, err, chi2, score;
  std::vector<int> chg, label, n_found, hot_offset, hot_index, hot_layer;
  std::vector<unsigned> status;
};

struct Cursor {
  const unsigned char* p;
  size_t left;
  bool take(size_t n, const unsigned char*& out) {
    if (n > left) return false;
    out = p;
    p += n;
    left -= n;
    return true;
  }
  bool i32(int& v) {
    const unsigned char* q;
    if (!take(4, q)) return false;
    v = get<int>(q);
    return true;
  }
};

}  // namespace

struct mkf_event_file {
  std::FILE* fp = nullptr;
  bool writing = false;
  mkf_event_file_header hdr{};
  long pos = 0;
  int n_written = 0;
  // last event read
  std::vector<int> hit_offset, hit_mc;
  std::vector<float> hits;
  std::vector<unsigned> hit_packed;
  std::vector<uint64_t> masks;
  TrackSet sets[3];
  std::vector<float> ts_par, ts_err;
  std::vector<int> ts_chg, mc_info;
  float beam_spot[kBeamSpotFloats] = {};
  bool has_event = false, has_masks = false, has_bs = false;
  int n_states = 0;
};

namespace {

int read_tracks(Cursor& c, TrackSet& s) {
  int n, data_size;
  if (!c.i32(n) || !c.i32(data_size)) return fail("truncated track section header");
  if (n < 0 || data_size < 8) return fail("bad track section (%d tracks, %d bytes)", n, data_size);
  const unsigned char* rec;
  if (!c.take((size_t)n * kTrackBytes, rec)) return fail("truncated track records");
  s = TrackSet();
  s.n = n;
  s.par.resize(6 * (size_t)n); s.err.resize(21 * (size_t)n); s.chi2.resize(n); s.score.resize(n);
  s.chg.resize(n); s.label.resize(n); s.n_found.resize(n); s.status.resize(n); s.hot_offset.resize(n + 1, 0);
  // HitOnTrack offsets are summed in 64 bits: each track may claim up to 32768
  // entries, so an int sum over many tracks could wrap.  The total must fit an
  // int and the bytes left in the event.
  int64_t hot_sum = 0;
  const int64_t hot_cap = std::min<int64_t>(INT_MAX, (int64_t)(c.left / kHotBytes));
  for (int t = 0; t < n; ++t) {
    const unsigned char* r = rec + (size_t)t * kTrackBytes;
    for (int e = 0; e < 6; ++e) s.par[(size_t)e * n + t] = get<float>(r + kTsPar + 4 * e);
    for (int e = 0; e < 21; ++e) s.err[(size_t)e * n + t] = get<float>(r + kTsErr + 4 * e);
    s.chg[t] = get<int16_t>(r + kTsChg);
    s.chi2[t] = get<float>(r + kTrChi2);
    s.score[t] = get<float>(r + kTrScore);
    const int n_hots = get<int16_t>(r + kTrLastHit) + 1;
    if (n_hots < 0) return fail("track %d: negative hit count", t);
    s.n_found[t] = get<int16_t>(r + kTrNFound);
    s.status[t] = get<uint32_t>(r + kTrStatus);
    s.label[t] = get<int32_t>(r + kTrLabel);
    hot_sum += n_hots;
    if (hot_sum > hot_cap) return fail("track %d: HitOnTrack count beyond the bytes left in the event", t);
    s.hot_offset[t + 1] = (int)hot_sum;
  }
  const int total = s.hot_offset[n];
  if ((long)data_size != 8L + (long)n * kTrackBytes + (long)total * kHotBytes)
    return fail("track section size %d does not match its %d tracks and %d hits", data_size, n, total);
  const unsigned char* h;
  if (!c.take((size_t)total * kHotBytes, h)) return fail("truncated HitOnTrack lists");
  s.hot_index.resize(total);
  s.hot_layer.resize(total);
  for (int k = 0; k < total; ++k) unpack_hot(get<uint32_t>(h + 4 * (size_t)k), s.hot_index[k], s.hot_layer[k]);
  return 0;
}

void write_bytes(std::FILE* fp, const void* p, size_t n, bool& ok) {
  if (n && std::fwrite(p, 1, n, fp) != n) ok = false;
}

void track_record(unsigned char* r, int t, int n, const float* par, const float* err, const int* chg,
                  const float* chi2, const float* score, const int* label, const unsigned* status, int n_hots,
                  int n_found) {
  std::memset(r, 0, kTrackBytes);
  for (int e = 0; e < 6; ++e) put<float>(r + kTsPar + 4 * e, par[(size_t)e * n + t]);
  for (int e = 0; e < 21; ++e) put<float>(r + kTsErr + 4 * e, err[(size_t)e * n + t]);
  put<int16_t>(r + kTsChg, (int16_t)chg[t]);
  r[kTsValid] = 1;
  put<float>(r + kTrChi2, chi2 ? chi2[t] : 0.f);
  put<float>(r + kTrScore, score ? score[t] : 0.f);
  put<int16_t>(r + kTrLastHit, (int16_t)(n_hots - 1));
  put<int16_t>(r + kTrNFound, (int16_t)n_found);
  put<uint32_t>(r + kTrStatus, status ? status[t] : 0u);
  put<int32_t>(r + kTrLabel, label ? label[t] : -1);
}

int write_tracks(std::FILE* fp, const mkf_event_data* ev, int w, long& bytes) {
  const int n = ev->n_tracks[w];
  if (n < 0) return fail("negative track count");
  if (n && (!ev->par[w] || !ev->err[w] || !ev->chg[w] || !ev->hot_offset[w]))
    return fail("track set %d: par, err, chg and hot_offset are required", w);
  const int total = n ? ev->hot_offset[w][n] : 0;
  if (total && (!ev->hot_index[w] || !ev->hot_layer[w])) return fail("track set %d: hot_index/hot_layer required", w);
  std::vector<unsigned char> rec((size_t)n * kTrackBytes);
  for (int t = 0; t < n; ++t) {
    const int b = ev->hot_offset[w][t], e = ev->hot_offset[w][t + 1];
    if (e < b || e - b > 32767) return fail("track %d: bad HitOnTrack range", t);
    int nf = 0;
    for (int k = b; k < e; ++k) nf += ev->hot_index[w][k] >= 0;
    track_record(&rec[(size_t)t * kTrackBytes], t, n, ev->par[w], ev->err[w], ev->chg[w], ev->chi2[w], ev->score[w],
                 ev->label[w], ev->status[w], e - b, nf);
  }
  std::vector<uint32_t> hots(total);
  for (int k = 0; k < total; ++k) hots[k] = pack_hot(ev->hot_index[w][k], ev->hot_layer[w][k]);
  const int data_size = 8 + n * kTrackBytes + total * kHotBytes;
  bool ok = true;
  write_bytes(fp, &n, 4, ok);
  write_bytes(fp, &data_size, 4, ok);
  write_bytes(fp, rec.data(), rec.size(), ok);
  write_bytes(fp, hots.data(), hots.size() * 4, ok);
  if (!ok) return fail("write failed");
  bytes += data_size;
  return 0;
}

}  // namespace

extern "C" {

const char* mkf_event_file_last_error(void) { return g_err.c_str(); }

int mkf_event_file_open_read(const char* path, mkf_event_file** out, mkf_event_file_header* hdr) {
  if (!path || !out) return fail("NULL argument");
  *out = nullptr;
  std::FILE* fp = std::fopen(path, "rb");
  if (!fp) return fail("cannot open '%s'", path);
  mkf_event_file_header h;
  if (std::fread(&h, sizeof(h), 1, fp) != 1) {
    std::fclose(fp);
    return fail("'%s': no header", path);
  }
  const char* why = nullptr;
  if (h.magic != kMagic) why = "wrong magic";
  else if (h.format_version < kMinVersion || h.format_version > kVersion) why = "unsupported format version";
  else if (h.sizeof_track != kTrackBytes) why = "sizeof(Track) differs from 160";
  else if (h.sizeof_hit != kHitBytes) why = "sizeof(Hit) differs from 44";
  else if (h.sizeof_hot != kHotBytes) why = "sizeof(HitOnTrack) differs from 4";
  if (why) {
    std::fclose(fp);
    return fail("'%s': %s", path, why);
  }
  auto* f = new mkf_event_file();
  f->fp = fp;
  f->hdr = h;
  f->pos = sizeof(h);
  if (hdr) *hdr = h;
  *out = f;
  return 0;
}

int mkf_event_file_open_write(const char* path, int n_layers, int n_events, int extra_sections,
                              mkf_event_file** out) {
  if (!path || !out) return fail("NULL argument");
  *out = nullptr;
  if (n_layers < 0 || n_events < 0) return fail("negative n_layers or n_events");
  std::FILE* fp = std::fopen(path, "wb");
  if (!fp) return fail("cannot create '%s'", path);
  auto* f = new mkf_event_file();
  f->fp = fp;
  f->writing = true;
  f->hdr = {kMagic, kVersion, kTrackBytes, kHitBytes, kHotBytes, n_layers, n_events, extra_sections};
  if (std::fwrite(&f->hdr, sizeof(f->hdr), 1, fp) != 1) {
    std::fclose(fp);
    delete f;
    return fail("cannot write the header of '%s'", path);
  }
  *out = f;
  return 0;
}

int mkf_event_file_close(mkf_event_file* f) {
  if (!f) return 0;
  int rc = 0;
  if (f->writing && f->hdr.n_events != f->n_written) {   // DataFile::CloseWrite
    f->hdr.n_events = f->n_written;
    if (std::fseek(f->fp, 0, SEEK_SET) != 0 || std::fwrite(&f->hdr, sizeof(f->hdr), 1, f->fp) != 1)
      rc = fail("cannot update the event count");
  }
  if (f->fp && std::fclose(f->fp) != 0) rc = fail("close failed");
  delete f;
  return rc;
}

int mkf_event_file_read_event(mkf_event_file* f, mkf_event_counts* counts) {
  if (!f || !counts) return fail("NULL argument");
  if (f->writing) return fail("file is open for writing");
  if (std::fseek(f->fp, f->pos, SEEK_SET) != 0) return fail("seek failed");
  int evsize;
  if (std::fread(&evsize, 4, 1, f->fp) != 1) return MKF_EF_END;
  if (evsize < 4) return fail("bad event size %d", evsize);
  std::vector<unsigned char> buf((size_t)evsize - 4);
  if (!buf.empty() && std::fread(buf.data(), 1, buf.size(), f->fp) != buf.size()) return fail("truncated event");
  Cursor c{buf.data(), buf.size()};
  const int ex = f->hdr.extra_sections;
  f->has_event = false;
  if (read_tracks(c, f->sets[0])) return 1;
  f->n_states = 0;
  if (ex & MKF_EF_SIM_TRACK_STATES) {
    int n;
    const unsigned char* r;
    if (!c.i32(n) || n < 0 || !c.take((size_t)n * kStateBytes, r)) return fail("bad sim track states");
    f->n_states = n;
    f->ts_par.resize(6 * (size_t)n); f->ts_err.resize(21 * (size_t)n); f->ts_chg.resize(n);
    for (int t = 0; t < n; ++t) {
      const unsigned char* q = r + (size_t)t * kStateBytes;
      for (int e = 0; e < 6; ++e) f->ts_par[(size_t)e * n + t] = get<float>(q + kTsPar + 4 * e);
      for (int e = 0; e < 21; ++e) f->ts_err[(size_t)e * n + t] = get<float>(q + kTsErr + 4 * e);
      f->ts_chg[t] = get<int16_t>(q + kTsChg);
    }
  }
  int nl;
  if (!c.i32(nl) || nl < 0 || nl > 4096) return fail("bad layer count");
  f->hit_offset.assign(nl + 1, 0);
  std::vector<const unsigned char*> recs(nl);
  for (int l = 0; l < nl; ++l) {
    int nh;
    if (!c.i32(nh) || nh < 0 || !c.take((size_t)nh * kHitBytes, recs[l])) return fail("bad hits of layer %d", l);
    f->hit_offset[l + 1] = f->hit_offset[l] + nh;
  }
  const int total = f->hit_offset[nl];
  f->hits.resize((size_t)total * 9);
  f->hit_mc.resize(total);
  f->hit_packed.resize(total);
  for (int l = 0; l < nl; ++l)
    for (int i = 0; i < f->hit_offset[l + 1] - f->hit_offset[l]; ++i) {
      const unsigned char* q = recs[l] + (size_t)i * kHitBytes;
      const size_t g = (size_t)f->hit_offset[l] + i;
      for (int e = 0; e < 3; ++e) f->hits[g * 9 + e] = get<float>(q + kHitPos + 4 * e);
      for (int e = 0; e < 6; ++e) f->hits[g * 9 + 3 + e] = get<float>(q + kHitErr + 4 * e);
      f->hit_mc[g] = get<int32_t>(q + kHitMc);
      f->hit_packed[g] = get<uint32_t>(q + kHitPacked);
    }
  f->has_masks = (ex & MKF_EF_HIT_ITER_MASKS) != 0;
  f->masks.assign(f->has_masks ? total : 0, 0);
  if (f->has_masks) {
    const unsigned char* q;
    if (!c.take((size_t)total * 8, q)) return fail("truncated hit masks");
    std::memcpy(f->masks.data(), q, (size_t)total * 8);
  }
  int nm;
  const unsigned char* q;
  if (!c.i32(nm) || nm < 0 || !c.take((size_t)nm * kMcInfoBytes, q)) return fail("bad MC hit info");
  f->mc_info.resize(4 * (size_t)nm);
  if (nm) std::memcpy(f->mc_info.data(), q, (size_t)nm * kMcInfoBytes);
  f->sets[1] = TrackSet();
  f->sets[2] = TrackSet();
  if ((ex & MKF_EF_SEEDS) && read_tracks(c, f->sets[1])) return 1;
  if ((ex & MKF_EF_CMSSW_TRACKS) && read_tracks(c, f->sets[2])) return 1;
  f->has_bs = (ex & MKF_EF_BEAM_SPOT) != 0;
  if (f->has_bs) {
    if (!c.take(4 * kBeamSpotFloats, q)) return fail("truncated beam spot");
    std::memcpy(f->beam_spot, q, 4 * kBeamSpotFloats);
  }
  if (c.left != 0) return fail("event has %zu bytes past its last section", c.left);
  f->pos += evsize;
  f->has_event = true;
  for (int w = 0; w < 3; ++w) {
    counts->n_tracks[w] = f->sets[w].n;
    counts->n_hots[w] = f->sets[w].n ? f->sets[w].hot_offset[f->sets[w].n] : 0;
  }
  counts->n_track_states = f->n_states;
  counts->n_layers = nl;
  counts->n_hits = total;
  counts->n_mc_hit_info = nm;
  counts->has_masks = f->has_masks;
  counts->has_beam_spot = f->has_bs;
  return 0;
}

int mkf_event_file_get_hits(mkf_event_file* f, int* hit_offset, float* hits, int* hit_mc, unsigned* hit_packed,
                            uint64_t* hit_masks) {
  if (!f || !f->has_event) return fail("no event read");
  if (hit_offset) std::memcpy(hit_offset, f->hit_offset.data(), f->hit_offset.size() * 4);
  if (hits) std::memcpy(hits, f->hits.data(), f->hits.size() * 4);
  if (hit_mc) std::memcpy(hit_mc, f->hit_mc.data(), f->hit_mc.size() * 4);
  if (hit_packed) std::memcpy(hit_packed, f->hit_packed.data(), f->hit_packed.size() * 4);
  if (hit_masks) {
    if (!f->has_masks) return fail("the file has no hit masks");
    std::memcpy(hit_masks, f->masks.data(), f->masks.size() * 8);
  }
  return 0;
}

int mkf_event_file_get_tracks(mkf_event_file* f, int which, const mkf_event_tracks* o) {
  if (!f || !f->has_event) return fail("no event read");
  if (which < 0 || which > 2 || !o) return fail("bad track set");
  const TrackSet& s = f->sets[which];
  auto cp = [](void* dst, const void* src, size_t bytes) {
    if (dst && bytes) std::memcpy(dst, src, bytes);
  };
  cp(o->par, s.par.data(), s.par.size() * 4);
  cp(o->err, s.err.data(), s.err.size() * 4);
  cp(o->chg, s.chg.data(), s.chg.size() * 4);
  cp(o->chi2, s.chi2.data(), s.chi2.size() * 4);
  cp(o->score, s.score.data(), s.score.size() * 4);
  cp(o->label, s.label.data(), s.label.size() * 4);
  cp(o->n_found, s.n_found.data(), s.n_found.size() * 4);
  cp(o->status, s.status.data(), s.status.size() * 4);
  if (o->hot_offset) {
    if (s.n) std::memcpy(o->hot_offset, s.hot_offset.data(), s.hot_offset.size() * 4);
    else o->hot_offset[0] = 0;
  }
  cp(o->hot_index, s.hot_index.data(), s.hot_index.size() * 4);
  cp(o->hot_layer, s.hot_layer.data(), s.hot_layer.size() * 4);
  return 0;
}

int mkf_event_file_get_track_states(mkf_event_file* f, float* par, float* err, int* chg) {
  if (!f || !f->has_event) return fail("no event read");
  if (par) std::memcpy(par, f->ts_par.data(), f->ts_par.size() * 4);
  if (err) std::memcpy(err, f->ts_err.data(), f->ts_err.size() * 4);
  if (chg) std::memcpy(chg, f->ts_chg.data(), f->ts_chg.size() * 4);
  return 0;
}

int mkf_event_file_get_mc_hit_info(mkf_event_file* f, int* info) {
  if (!f || !f->has_event) return fail("no event read");
  if (info && !f->mc_info.empty()) std::memcpy(info, f->mc_info.data(), f->mc_info.size() * 4);
  return 0;
}

int mkf_event_file_get_beam_spot(mkf_event_file* f, float* bs) {
  if (!f || !f->has_event) return fail("no event read");
  if (!f->has_bs) return fail("the file has no beam spot");
  if (bs) std::memcpy(bs, f->beam_spot, sizeof(f->beam_spot));
  return 0;
}

int mkf_event_file_write_event(mkf_event_file* f, const mkf_event_data* ev) {
  if (!f || !ev) return fail("NULL argument");
  if (!f->writing) return fail("file is open for reading");
  const int ex = f->hdr.extra_sections;
  const int nl = ev->n_layers;
  if (nl < 0 || (nl && !ev->hit_offset)) return fail("bad layers");
  const int total = nl ? ev->hit_offset[nl] : 0;
  if (total && !ev->hits) return fail("hits required");
  if ((ex & MKF_EF_HIT_ITER_MASKS) && total && !ev->hit_masks) return fail("the file has hit masks: hit_masks required");
  std::FILE* fp = f->fp;
  if (std::fseek(fp, 0, SEEK_END) != 0) return fail("seek failed");
  const long start = std::ftell(fp);
  long evsize = 4;
  bool ok = true;
  write_bytes(fp, &evsize, 4, ok);                        // fixed up below
  if (write_tracks(fp, ev, 0, evsize)) return 1;
  if (ex & MKF_EF_SIM_TRACK_STATES) {
    const int n = ev->n_track_states;
    if (n < 0 || (n && (!ev->ts_par || !ev->ts_err || !ev->ts_chg))) return fail("bad sim track states");
    std::vector<unsigned char> rec((size_t)n * kStateBytes, 0);
    for (int t = 0; t < n; ++t) {
      unsigned char* r = &rec[(size_t)t * kStateBytes];
      for (int e = 0; e < 6; ++e) put<float>(r + kTsPar + 4 * e, ev->ts_par[(size_t)e * n + t]);
      for (int e = 0; e < 21; ++e) put<float>(r + kTsErr + 4 * e, ev->ts_err[(size_t)e * n + t]);
      put<int16_t>(r + kTsChg, (int16_t)ev->ts_chg[t]);
      r[kTsValid] = 1;
    }
    write_bytes(fp, &n, 4, ok);
    write_bytes(fp, rec.data(), rec.size(), ok);
    evsize += 4 + (long)n * kStateBytes;
  }
  write_bytes(fp, &nl, 4, ok);
  evsize += 4;
  for (int l = 0; l < nl; ++l) {
    const int b = ev->hit_offset[l], nh = ev->hit_offset[l + 1] - b;
    if (nh < 0) return fail("hit_offset decreases at layer %d", l);
    std::vector<unsigned char> rec((size_t)nh * kHitBytes, 0);
    for (int i = 0; i < nh; ++i) {
      unsigned char* r = &rec[(size_t)i * kHitBytes];
      const size_t g = (size_t)b + i;
      for (int e = 0; e < 3; ++e) put<float>(r + kHitPos + 4 * e, ev->hits[g * 9 + e]);
      for (int e = 0; e < 6; ++e) put<float>(r + kHitErr + 4 * e, ev->hits[g * 9 + 3 + e]);
      put<int32_t>(r + kHitMc, ev->hit_mc ? ev->hit_mc[g] : -1);
      put<uint32_t>(r + kHitPacked, ev->hit_packed ? ev->hit_packed[g] : 0u);
    }
    write_bytes(fp, &nh, 4, ok);
    write_bytes(fp, rec.data(), rec.size(), ok);
    evsize += 4 + (long)nh * kHitBytes;
  }
  if (ex & MKF_EF_HIT_ITER_MASKS) {
    write_bytes(fp, ev->hit_masks, (size_t)total * 8, ok);
    evsize += (long)total * 8;
  }
  const int nm = ev->n_mc_hit_info;
  if (nm < 0 || (nm && !ev->mc_hit_info)) return fail("bad MC hit info");
  write_bytes(fp, &nm, 4, ok);
  write_bytes(fp, ev->mc_hit_info, (size_t)nm * kMcInfoBytes, ok);
  evsize += 4 + (long)nm * kMcInfoBytes;
  if ((ex & MKF_EF_SEEDS) && write_tracks(fp, ev, 1, evsize)) return 1;
  if ((ex & MKF_EF_CMSSW_TRACKS) && write_tracks(fp, ev, 2, evsize)) return 1;
  if (ex & MKF_EF_BEAM_SPOT) {
    write_bytes(fp, ev->beam_spot ? ev->beam_spot : kBeamSpotDefault, 4 * kBeamSpotFloats, ok);
    evsize += 4 * kBeamSpotFloats;
  }
  if (!ok) return fail("write failed");
  if (evsize > 0x7fffffffL) return fail("event larger than 2 GiB");
  const int es = (int)evsize;
  if (std::fseek(fp, start, SEEK_SET) != 0 || std::fwrite(&es, 4, 1, fp) != 1 || std::fseek(fp, 0, SEEK_END) != 0)
    return fail("cannot write the event size");
  ++f->n_written;
  return 0;
}

}  // extern "C"
