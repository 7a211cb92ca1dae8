// Minimal uncompressed LAS 1.2/1.4 header reader and writer (host only).
//
// Covers what the counting-sort chunker needs from an input file: record
// format and length, point count (legacy 32-bit and LAS 1.4 64-bit fields,
// the larger wins as in Converter/modules/LasLoader/LasLoader.cpp:103),
// scale/offset/bounds (LasLoader.cpp:87-101) and the extra-bytes VLR
// (record id 4, 192-byte descriptors; PotreeConverter.h:80-133).
// LAZ is not handled: entropy decoding stays outside this path.
#pragma once

#include <stdint.h>
#include <string.h>

#include <fstream>
#include <string>
#include <vector>

namespace lasfile {

struct ExtraBytesDescriptor {
  uint8_t data_type = 0;
  uint8_t options = 0;
  std::string name;
  std::string description;
  double scale[3] = {1.0, 1.0, 1.0};
  double offset[3] = {0.0, 0.0, 0.0};
};

struct Header {
  uint8_t version_major = 1;
  uint8_t version_minor = 4;
  uint16_t header_size = 375;
  uint32_t offset_to_points = 375;
  uint32_t num_vlrs = 0;
  uint8_t point_format = 2;
  uint16_t record_length = 26;
  uint64_t num_points = 0;
  double scale[3] = {0.001, 0.001, 0.001};
  double offset[3] = {0.0, 0.0, 0.0};
  double min[3] = {0, 0, 0};
  double max[3] = {0, 0, 0};
  std::vector<ExtraBytesDescriptor> extra;
};

template <class T>
inline T rd(const uint8_t* p) {
  T v;
  memcpy(&v, p, sizeof(T));
  return v;
}
template <class T>
inline void wr(uint8_t* p, T v) {
  memcpy(p, &v, sizeof(T));
}

inline std::string cstr(const uint8_t* p, size_t n) {
  size_t k = 0;
  while (k < n && p[k] != 0) k++;
  return std::string((const char*)p, k);
}

// Parses the header and VLRs; returns false on a malformed file.
inline bool read_header(const std::string& path, Header& h, std::string* err = nullptr) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    if (err) *err = "cannot open " + path;
    return false;
  }
  uint8_t b[375];
  memset(b, 0, sizeof(b));
  f.read((char*)b, 227);
  if (f.gcount() < 227 || memcmp(b, "LASF", 4) != 0) {
    if (err) *err = "not a LAS file: " + path;
    return false;
  }
  h.version_major = b[24];
  h.version_minor = b[25];
  h.header_size = rd<uint16_t>(b + 94);
  h.offset_to_points = rd<uint32_t>(b + 96);
  h.num_vlrs = rd<uint32_t>(b + 100);
  h.point_format = b[104] & 0x3F;
  if (b[104] & 0xC0) {
    if (err) *err = "compressed (LAZ) point data is not supported: " + path;
    return false;
  }
  h.record_length = rd<uint16_t>(b + 105);
  uint64_t legacy = rd<uint32_t>(b + 107);
  for (int i = 0; i < 3; i++) h.scale[i] = rd<double>(b + 131 + 8 * i);
  for (int i = 0; i < 3; i++) h.offset[i] = rd<double>(b + 155 + 8 * i);
  h.max[0] = rd<double>(b + 179); h.min[0] = rd<double>(b + 187);
  h.max[1] = rd<double>(b + 195); h.min[1] = rd<double>(b + 203);
  h.max[2] = rd<double>(b + 211); h.min[2] = rd<double>(b + 219);
  uint64_t extended = 0;
  if (h.header_size >= 375) {
    f.seekg(0);
    f.read((char*)b, 375);
    extended = rd<uint64_t>(b + 247);
  }
  h.num_points = legacy > extended ? legacy : extended;

  f.seekg(h.header_size);
  for (uint32_t v = 0; v < h.num_vlrs; v++) {
    uint8_t vh[54];
    f.read((char*)vh, 54);
    if (f.gcount() < 54) break;
    uint16_t record_id = rd<uint16_t>(vh + 18);
    uint16_t len = rd<uint16_t>(vh + 20);
    std::vector<uint8_t> data(len);
    if (len) f.read((char*)data.data(), len);
    if (record_id == 4) {
      for (size_t o = 0; o + 192 <= data.size(); o += 192) {
        ExtraBytesDescriptor d;
        d.data_type = data[o + 2];
        d.options = data[o + 3];
        d.name = cstr(&data[o + 4], 32);
        if (d.options & 0x08) memcpy(d.scale, &data[o + 112], 24);
        if (d.options & 0x10) memcpy(d.offset, &data[o + 136], 24);
        d.description = cstr(&data[o + 160], 32);
        h.extra.push_back(d);
      }
    }
  }
  return true;
}

// Reads `count` records starting at record `first` into `dst`.
inline bool read_points(const std::string& path, const Header& h, uint64_t first, uint64_t count, uint8_t* dst) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.seekg((std::streamoff)(h.offset_to_points + first * h.record_length));
  f.read((char*)dst, (std::streamsize)(count * h.record_length));
  return (uint64_t)f.gcount() == count * h.record_length;
}

// Writes a LAS 1.4 file with the given header fields and raw records.
inline bool write_file(const std::string& path, const Header& h, const uint8_t* records) {
  std::vector<uint8_t> vlr;
  if (!h.extra.empty()) {
    vlr.assign(54 + 192 * h.extra.size(), 0);
    memcpy(&vlr[2], "LASF_Spec", 9);
    wr<uint16_t>(&vlr[18], 4);
    wr<uint16_t>(&vlr[20], (uint16_t)(192 * h.extra.size()));
    memcpy(&vlr[22], "extra bytes", 11);
    for (size_t i = 0; i < h.extra.size(); i++) {
      uint8_t* d = &vlr[54 + 192 * i];
      const auto& e = h.extra[i];
      d[2] = e.data_type;
      d[3] = e.options;
      memcpy(d + 4, e.name.data(), e.name.size() < 32 ? e.name.size() : 32);
      memcpy(d + 112, e.scale, 24);
      memcpy(d + 136, e.offset, 24);
      memcpy(d + 160, e.description.data(), e.description.size() < 32 ? e.description.size() : 32);
    }
  }
  uint8_t b[375];
  memset(b, 0, sizeof(b));
  memcpy(b, "LASF", 4);
  b[24] = 1;
  b[25] = 4;
  memcpy(b + 26, "potreeconverter_amd", 19);
  memcpy(b + 58, "lasgen", 6);
  wr<uint16_t>(b + 94, 375);
  wr<uint32_t>(b + 96, (uint32_t)(375 + vlr.size()));
  wr<uint32_t>(b + 100, h.extra.empty() ? 0u : 1u);
  b[104] = h.point_format;
  wr<uint16_t>(b + 105, h.record_length);
  // legacy point count: 0 for formats 6-10 and above 2^32 - 1 points (LAS 1.4)
  wr<uint32_t>(b + 107, h.num_points <= 0xFFFFFFFFull && h.point_format < 6 ? (uint32_t)h.num_points : 0u);
  for (int i = 0; i < 3; i++) wr<double>(b + 131 + 8 * i, h.scale[i]);
  for (int i = 0; i < 3; i++) wr<double>(b + 155 + 8 * i, h.offset[i]);
  wr<double>(b + 179, h.max[0]); wr<double>(b + 187, h.min[0]);
  wr<double>(b + 195, h.max[1]); wr<double>(b + 203, h.min[1]);
  wr<double>(b + 211, h.max[2]); wr<double>(b + 219, h.min[2]);
  wr<uint64_t>(b + 247, h.num_points);
  std::ofstream f(path, std::ios::binary);
  if (!f) return false;
  f.write((const char*)b, 375);
  if (!vlr.empty()) f.write((const char*)vlr.data(), (std::streamsize)vlr.size());
  f.write((const char*)records, (std::streamsize)(h.num_points * h.record_length));
  return (bool)f;
}

}  // namespace lasfile
