// Point attribute schema of the converter output (host side).
//
// Mirrors Converter/include/Attributes.h:16-186 (AttributeType codes, the
// Attribute fields written to metadata.json, Attributes::getOffset/get) and
// the per-format attribute lists of Converter/include/PotreeConverter.h:136-193.
#pragma once
#include <stdint.h>

#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "lasfile.h"
#include "pcg_common.h"

namespace pcg {

enum class AttributeType : int32_t {
  INT8 = 0, INT16 = 1, INT32 = 2, INT64 = 3,
  UINT8 = 10, UINT16 = 11, UINT32 = 12, UINT64 = 13,
  FLOAT = 20, DOUBLE = 21, UNDEFINED = 123456,
};

const char* attribute_type_name(AttributeType t);
AttributeType attribute_type_from_name(const std::string& s);
int attribute_type_size(AttributeType t);

struct Vec3 {
  double x = 0, y = 0, z = 0;
  Vec3() {}
  Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
  Vec3 operator-(const Vec3& r) const { return Vec3(x - r.x, y - r.y, z - r.z); }
  Vec3 operator+(const Vec3& r) const { return Vec3(x + r.x, y + r.y, z + r.z); }
  Vec3 operator+(double s) const { return Vec3(x + s, y + s, z + s); }
  Vec3 operator*(double s) const { return Vec3(x * s, y * s, z * s); }
  // std::max(std::max(x, y), z) as in Vector3.h:236-239
  double max() const {
    double m = (x < y) ? y : x;
    return (m < z) ? z : m;
  }
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

struct Attribute {
  std::string name, description;
  int32_t size = 0, numElements = 0, elementSize = 0;
  AttributeType type = AttributeType::UNDEFINED;
  Vec3 min = Vec3(INFINITY, INFINITY, INFINITY);
  Vec3 max = Vec3(-INFINITY, -INFINITY, -INFINITY);
  Vec3 scale = Vec3(1, 1, 1);
  Vec3 offset = Vec3(0, 0, 0);
  std::vector<int64_t> histogram = std::vector<int64_t>(256, 0);
  Attribute() {}
  Attribute(std::string n, int s, int ne, int es, AttributeType t)
      : name(n), size(s), numElements(ne), elementSize(es), type(t) {}
};

struct Attributes {
  std::vector<Attribute> list;
  int32_t bytes = 0;
  Vec3 posScale = Vec3(1, 1, 1);
  Vec3 posOffset = Vec3(0, 0, 0);
  Attributes() {}
  explicit Attributes(const std::vector<Attribute>& l) : list(l) { for (auto& a : list) bytes += a.size; }
  int getOffset(const std::string& name) const {
    int o = 0;
    for (auto& a : list) { if (a.name == name) return o; o += a.size; }
    return -1;
  }
  Attribute* get(const std::string& name) {
    for (auto& a : list) if (a.name == name) return &a;
    return nullptr;
  }
};

// Output attributes, scale and offset of one LAS source
// (PotreeConverter.h:29-76, :136-193, :195-315) and the GPU transcode/stat
// tables that produce those records.  Returns false with set_error() for
// unsupported point formats.
bool output_attributes_for(const lasfile::Header& h, Attributes& out, ChunkerParams& params);

// The output attributes of several sources with an optional filter
// (computeOutputAttributes(sources, requestedAttributes),
// PotreeConverter.h:195-315) and each source's transcode / statistics tables
// into that layout; statTargets[s][k] names the output attribute (empty:
// not in the output) and component that statistic k of source s folds.
bool output_attributes_for_sources(const std::vector<lasfile::Header>& hs, const std::vector<std::string>& requested,
                                   Attributes& out, std::vector<ChunkerParams>& params,
                                   std::vector<std::vector<std::pair<std::string, int>>>& statTargets);

// chunks/metadata.json text (chunker_countsort_laszip.cpp:1005-1068).
// whether the chunker folds min/max for the attribute (the reference has a
// handler for it)
bool has_statistics(const Attribute& a);
std::string chunk_metadata_json(const Vec3& min, const Vec3& max, const Attributes& a);
bool parse_chunk_metadata(const std::string& text, Vec3& min, Vec3& max, Attributes& a);

// %.17g, the formatting of std::setprecision(max_digits10)
std::string fmt17(double v);

}  // namespace pcg
