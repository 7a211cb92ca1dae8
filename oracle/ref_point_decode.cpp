// Test infrastructure: decodes uncompressed LAS point records with the
// reference's vendored LASzip raw item readers
// (Converter/libs/laszip/src/lasreaditemraw.hpp), compiled from the
// reference checkout by oracle/build_ref.sh, so that the oracle's restated
// decode (orc_decode_points) can be checked against laszip itself.
//
// usage: ref_point_decode <point format 0-7> <record length> <records.bin> <out.bin>
// out.bin holds one 32-byte entry per record, the fields the chunker's
// attribute handlers read from laszip's point struct:
//   0 u16 intensity      2 u8 return_number    3 u8 number_of_returns
//   4 u8 classification (the 5-bit field)      5 i8 scan_angle_rank
//   6 u8 user_data       7 u8 extended_classification_flags
//   8 u16 point_source_ID                       10 i16 extended_scan_angle
//   12 u8 extended_classification              14 u16 rgb[3]
//   24 f64 gps_time
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "bytestreamin_array.hpp"
#include "lasreaditemraw.hpp"

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s format record_length records.bin out.bin\n", argv[0]);
    return 2;
  }
  const int fmt = atoi(argv[1]);
  const int rl = atoi(argv[2]);
  FILE* f = fopen(argv[3], "rb");
  if (!f || fmt < 0 || fmt > 7) return 2;
  std::vector<U8> data;
  U8 buf[1 << 16];
  size_t got;
  while ((got = fread(buf, 1, sizeof(buf), f)) > 0) data.insert(data.end(), buf, buf + got);
  fclose(f);
  const size_t n = data.size() / rl;
  // the item sequence laszip uses for an uncompressed record of each format
  LASreadItemRaw_POINT10_LE p10;
  LASreadItemRaw_GPSTIME11_LE gps11;
  LASreadItemRaw_RGB12_LE rgb12;
  LASreadItemRaw_WAVEPACKET13_LE wave13;
  LASreadItemRaw_POINT14_LE p14;
  // laszip's point struct starts zeroed and keeps what earlier reads left
  LAStempReadPoint10 point;
  memset(&point, 0, sizeof(point));
  U16 rgb[4] = {0, 0, 0, 0};
  F64 gps = 0;
  U8 wave[29];
  std::vector<U8> out(n * 32, 0);
  U32 context = 0;
  for (size_t i = 0; i < n; i++) {
    ByteStreamInArrayLE in(data.data() + i * rl, rl);
    p10.init(&in); gps11.init(&in); rgb12.init(&in); wave13.init(&in); p14.init(&in);
    if (fmt <= 5) {
      p10.read((U8*)&point, context);
      if (fmt == 1 || fmt == 3 || fmt == 4 || fmt == 5) gps11.read((U8*)&gps, context);
      if (fmt == 2 || fmt == 3 || fmt == 5) rgb12.read((U8*)rgb, context);
      if (fmt == 4 || fmt == 5) wave13.read(wave, context);
    } else {
      p14.read((U8*)&point, context);
      gps = point.gps_time;
      if (fmt == 7) rgb12.read((U8*)rgb, context);  // RGB14 raw: the same 6-byte copy
    }
    U8* o = &out[i * 32];
    memcpy(o + 0, &point.intensity, 2);
    o[2] = point.return_number;
    o[3] = point.number_of_returns;
    o[4] = point.classification & 31;  // laszip_point.classification is a 5-bit field over this byte
    memcpy(o + 5, &point.scan_angle_rank, 1);
    o[6] = point.user_data;
    o[7] = point.extended_classification_flags;
    memcpy(o + 8, &point.point_source_ID, 2);
    memcpy(o + 10, &point.extended_scan_angle, 2);
    o[12] = point.extended_classification;
    memcpy(o + 14, rgb, 6);
    memcpy(o + 24, &gps, 8);
  }
  FILE* g = fopen(argv[4], "wb");
  if (!g) return 2;
  fwrite(out.data(), 1, out.size(), g);
  fclose(g);
  return 0;
}
