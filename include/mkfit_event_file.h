/* mkfit_event_file.h -- C ABI of the reader and writer of mkFit's binary
 * event files ("memory files"), so events written by the reference's
 * Event::write_out can feed the GPU finders and fitters.
 *
 * Format (Event.h:74-127 DataFileHeader / DataFile, Event.cc:123-200 write_out,
 * :213-330 read_in, :400-440 write_tracks / read_tracks): a header of eight
 * int32 (magic 0xBEEF, format version 4..6, sizeof(Track), sizeof(Hit),
 * sizeof(HitOnTrack), number of layers, number of events, extra-section bits),
 * then per event an int32 byte size followed by
 *   sim tracks          int32 n, int32 section bytes, n Track records, then each
 *                       track's HitOnTrack list (lastHitIdx + 1 entries)
 *   [sim track states]  int32 n, n TrackState records       (extra bit 0x1)
 *   layers              int32 n_layers, per layer int32 n and n Hit records
 *   [hit masks]         per layer n uint64                    (extra bit 0x8)
 *   MC hit info         int32 n, n records of 4 int32
 *   [seeds]             as sim tracks                          (extra bit 0x2)
 *   [CMSSW tracks]      as sim tracks                          (extra bit 0x4)
 *   [beam spot]         8 float32                              (extra bit 0x10)
 * Records are the x86-64 memory images of the reference's classes:
 *   TrackState 112 B: parameters f32[6] @0, errors f32[21] @24 (packed lower
 *                     triangle, SMatrixSym66), charge int16 @108, valid u8 @110
 *   Track      160 B: TrackState @0, chi2 f32 @112, score f32 @116,
 *                     lastHitIdx int16 @120, nFoundHits int16 @122,
 *                     status u32 @124, label int32 @128, the hit vector's
 *                     three pointers @136 (meaningless on file)
 *   Hit         44 B: position f32[3] @0, packed 3x3 covariance f32[6] @12,
 *                     mcHitID int32 @36, packed data u32 @40
 *   HitOnTrack   4 B: index = signed low 24 bits, layer = signed high 8 bits
 * (Track.h:74-143, 360-367, 619-620; Hit.h:147-160, 181-182, 251-260, 301-303,
 * 314-320, 340-346.)  All arrays below are host memory. */
#ifndef MKFIT_EVENT_FILE_H
#define MKFIT_EVENT_FILE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MKF_EF_SIM_TRACK_STATES 0x1
#define MKF_EF_SEEDS            0x2
#define MKF_EF_CMSSW_TRACKS     0x4
#define MKF_EF_HIT_ITER_MASKS   0x8
#define MKF_EF_BEAM_SPOT        0x10

#define MKF_EF_TRACKS_SIM   0
#define MKF_EF_TRACKS_SEED  1
#define MKF_EF_TRACKS_CMSSW 2

typedef struct mkf_event_file_header {
  int magic, format_version, sizeof_track, sizeof_hit, sizeof_hot, n_layers, n_events, extra_sections;
} mkf_event_file_header;

/* Sizes of the event last read. */
typedef struct mkf_event_counts {
  int n_tracks[3];   /* sim, seed, CMSSW tracks (0 when the section is absent) */
  int n_hots[3];     /* HitOnTrack entries of each track set */
  int n_track_states;
  int n_layers;
  int n_hits;        /* all layers */
  int n_mc_hit_info;
  int has_masks, has_beam_spot;
} mkf_event_counts;

/* One set of tracks, element-major like the library's batches. */
typedef struct mkf_event_tracks {
  float* par;        /* [6][n] */
  float* err;        /* [21][n] */
  int* chg;          /* [n] */
  float* chi2;       /* [n] */
  float* score;      /* [n] */
  int* label;        /* [n] */
  int* n_found;      /* [n] */
  unsigned* status;  /* [n] Track::Status raw bits */
  int* hot_offset;   /* [n + 1]: track t's HitOnTrack entries are [hot_offset[t], hot_offset[t+1]) */
  int* hot_index;    /* [n_hots] */
  int* hot_layer;    /* [n_hots] */
} mkf_event_tracks;

/* One event to write; a NULL section pointer with a nonzero count is an error. */
typedef struct mkf_event_data {
  int n_layers;
  const int* hit_offset;        /* [n_layers + 1] */
  const float* hits;            /* [total][9] x y z xx yx yy zx zy zz */
  const int* hit_mc;            /* [total] mcHitID, NULL: -1 */
  const unsigned* hit_packed;   /* [total] Hit::PackedData raw, NULL: 0 */
  const uint64_t* hit_masks;    /* [total], written when the file has MKF_EF_HIT_ITER_MASKS */
  int n_tracks[3];
  const float* par[3];          /* [6][n] */
  const float* err[3];          /* [21][n] */
  const int* chg[3];
  const float* chi2[3];         /* NULL: 0 */
  const float* score[3];        /* NULL: 0 */
  const int* label[3];          /* NULL: -1 */
  const unsigned* status[3];    /* NULL: 0 */
  const int* hot_offset[3];     /* [n + 1] */
  const int* hot_index[3];
  const int* hot_layer[3];
  int n_track_states;           /* written when the file has MKF_EF_SIM_TRACK_STATES */
  const float* ts_par;          /* [6][n] */
  const float* ts_err;          /* [21][n] */
  const int* ts_chg;
  int n_mc_hit_info;
  const int* mc_hit_info;       /* [n][4]: mcTrackID, layer, ithLayerHit, mcHitID */
  const float* beam_spot;       /* [8], written when the file has MKF_EF_BEAM_SPOT (NULL: the defaults) */
} mkf_event_data;

typedef struct mkf_event_file mkf_event_file;

/* DataFile::OpenRead (Event.cc:866-938): checks the magic, a format version in
 * 4..6 and the three record sizes; `hdr` may be NULL. */
int mkf_event_file_open_read(const char* path, mkf_event_file** out, mkf_event_file_header* hdr);
/* DataFile::OpenWrite (Event.cc:940-949): a header with n_events (fixed up
 * at close to the number written), n_layers and the extra sections. */
int mkf_event_file_open_write(const char* path, int n_layers, int n_events, int extra_sections,
                              mkf_event_file** out);
int mkf_event_file_close(mkf_event_file* f);

/* Event::read_in (Event.cc:213-330) of the next event (DataFile::
 * AdvancePosToNextEvent); fills `counts` and keeps the event for the getters.
 * Returns MKF_EF_END (and leaves counts untouched) past the last event. */
#define MKF_EF_END 100
int mkf_event_file_read_event(mkf_event_file* f, mkf_event_counts* counts);
/* The last event read: hit_offset [n_layers + 1], hits [n_hits][9]; hit_mc,
 * hit_packed, hit_masks [n_hits] may be NULL. */
int mkf_event_file_get_hits(mkf_event_file* f, int* hit_offset, float* hits, int* hit_mc, unsigned* hit_packed,
                            uint64_t* hit_masks);
int mkf_event_file_get_tracks(mkf_event_file* f, int which, const mkf_event_tracks* out);
/* par [6][n], err [21][n], chg [n] of the sim track states. */
int mkf_event_file_get_track_states(mkf_event_file* f, float* par, float* err, int* chg);
int mkf_event_file_get_mc_hit_info(mkf_event_file* f, int* info);
int mkf_event_file_get_beam_spot(mkf_event_file* f, float* bs);

/* Event::write_out (Event.cc:123-200) of one event. */
int mkf_event_file_write_event(mkf_event_file* f, const mkf_event_data* ev);

const char* mkf_event_file_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
