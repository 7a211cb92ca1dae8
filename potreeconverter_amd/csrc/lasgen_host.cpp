// Host side of the synthetic LAS generator: fills records on the CPU and
// writes a LAS 1.4 file whose header bounds are the exact bounds of the
// quantized points (the same formula laszip uses to decode a point,
// scale * X + offset).
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "lasfile.h"
#include "lasgen.h"

extern "C" {

// Fills `params` for one of the named workloads; returns 0 on success.
//   0: 500 m tile, offset 0, scale 0.001 (CPU-runnable config)
//   1: 5 km tile in UTM-like coordinates (offset 500000, 5000000, 0)
//   2: precision scene: 819.2 m tile, scale 0.0001, offset (1e6, 1e6, 0),
//      10% of the points in near-duplicate clusters, 5% on a lattice whose
//      pitch (0.4 m) is the node spacing at level 4 (819.2 / 128 / 2^4)
int pcg_lasgen_preset(int32_t preset, int64_t n, uint64_t seed, LasGenParams* p) {
  memset(p, 0, sizeof(*p));
  p->seed = seed;
  p->n = n;
  p->amplitude = 40.0;
  p->num_buildings = 20;
  p->tile_points = 4096;
  p->wall_fraction = 0.05;
  p->extra_bytes = 0;
  p->point_format = 2;
  double scale = 0.001;
  if (preset == 0) {
    p->extent = 500.0;
  } else if (preset == 1) {
    p->extent = 5000.0;
    p->origin[0] = p->offset[0] = 500000.0;
    p->origin[1] = p->offset[1] = 5000000.0;
  } else if (preset == 2) {
    scale = 0.0001;
    p->extent = 819.2;
    p->origin[0] = p->offset[0] = 1000000.0;
    p->origin[1] = p->offset[1] = 1000000.0;
    p->span_units = 8192000;
    p->dup_fraction = 0.10;
    p->lattice_fraction = 0.05;
    p->lattice_pitch_units = 4000;
  } else {
    return 1;
  }
  for (int i = 0; i < 3; i++) p->inv_scale[i] = 1.0 / scale;
  lasgen_prepare(*p);
  return 0;
}

// Sets the record format of a preset's points (0-7) and whether 7 extra
// bytes follow; returns 0 on success.
int pcg_lasgen_configure(LasGenParams* p, int32_t point_format, int32_t extra_bytes) {
  if (point_format < 0 || point_format > 7) return 1;
  p->point_format = point_format;
  p->extra_bytes = extra_bytes ? 1 : 0;
  return 0;
}

int32_t pcg_lasgen_record_length(const LasGenParams* p) { return lasgen_record_length(*p); }

// Generates records [first, first + count) into `out`.
void pcg_lasgen_fill(const LasGenParams* p, int64_t first, int64_t count, uint8_t* out) {
  LasGenBuilding bld[64];
  int nb = p->num_buildings < 64 ? p->num_buildings : 64;
  for (int b = 0; b < nb; b++) bld[b] = lasgen_building(*p, b);
  int rl = lasgen_record_length(*p);
  for (int64_t i = 0; i < count; i++) lasgen_point(*p, bld, first + i, out + i * rl);
}

// Header bounds from integer extents (decoding formula scale * X + offset).
void pcg_lasgen_bounds(const LasGenParams* p, const int32_t qmin[3], const int32_t qmax[3], double mn[3],
                       double mx[3]) {
  for (int i = 0; i < 3; i++) {
    double s = 1.0 / p->inv_scale[i];
    mn[i] = s * qmin[i] + p->offset[i];
    mx[i] = s * qmax[i] + p->offset[i];
  }
}

// Generates the whole point set and writes a LAS file.
int pcg_lasgen_write(const LasGenParams* p, const char* path) {
  int rl = pcg_lasgen_record_length(p);
  std::vector<uint8_t> recs((size_t)p->n * rl);
  pcg_lasgen_fill(p, 0, p->n, recs.data());
  int32_t qmin[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, qmax[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (int64_t i = 0; i < p->n; i++) {
    for (int k = 0; k < 3; k++) {
      int32_t v;
      memcpy(&v, &recs[(size_t)i * rl + 4 * k], 4);
      qmin[k] = std::min(qmin[k], v);
      qmax[k] = std::max(qmax[k], v);
    }
  }
  lasfile::Header h;
  h.point_format = (uint8_t)p->point_format;
  h.record_length = (uint16_t)rl;
  h.num_points = (uint64_t)p->n;
  for (int i = 0; i < 3; i++) {
    h.scale[i] = 1.0 / p->inv_scale[i];
    h.offset[i] = p->offset[i];
  }
  pcg_lasgen_bounds(p, qmin, qmax, h.min, h.max);
  if (p->extra_bytes) {
    lasfile::ExtraBytesDescriptor a, b, c;
    a.data_type = 9; a.name = "gauss"; a.description = "float32 N(0,1)";
    b.data_type = 3; b.name = "uniform16"; b.description = "uint16 uniform";
    c.data_type = 2; c.name = "uniform8"; c.description = "int8 uniform";
    h.extra = {a, b, c};
  }
  return lasfile::write_file(path, h, recs.data()) ? 0 : 1;
}

}  // extern "C"
