// Attribute schema, GPU transcode tables and chunk metadata text.
#include "attributes.h"
#include "json_dump.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <sstream>

namespace pcg {

const char* attribute_type_name(AttributeType t) {
  switch (t) {
    case AttributeType::INT8: return "int8";
    case AttributeType::INT16: return "int16";
    case AttributeType::INT32: return "int32";
    case AttributeType::INT64: return "int64";
    case AttributeType::UINT8: return "uint8";
    case AttributeType::UINT16: return "uint16";
    case AttributeType::UINT32: return "uint32";
    case AttributeType::UINT64: return "uint64";
    case AttributeType::FLOAT: return "float";
    case AttributeType::DOUBLE: return "double";
    case AttributeType::UNDEFINED: return "undefined";
  }
  return "error";
}

AttributeType attribute_type_from_name(const std::string& s) {
  const AttributeType all[] = {AttributeType::INT8, AttributeType::INT16, AttributeType::INT32, AttributeType::INT64,
                               AttributeType::UINT8, AttributeType::UINT16, AttributeType::UINT32,
                               AttributeType::UINT64, AttributeType::FLOAT, AttributeType::DOUBLE};
  for (auto t : all) if (s == attribute_type_name(t)) return t;
  return AttributeType::UNDEFINED;
}

int attribute_type_size(AttributeType t) {
  switch (t) {
    case AttributeType::INT8: case AttributeType::UINT8: return 1;
    case AttributeType::INT16: case AttributeType::UINT16: return 2;
    case AttributeType::INT32: case AttributeType::UINT32: case AttributeType::FLOAT: return 4;
    case AttributeType::INT64: case AttributeType::UINT64: case AttributeType::DOUBLE: return 8;
    default: return 0;
  }
}

static StatKind stat_kind_of(AttributeType t) {
  switch (t) {
    case AttributeType::INT8: return ST_I8;
    case AttributeType::INT16: return ST_I16;
    case AttributeType::INT32: return ST_I32;
    case AttributeType::INT64: return ST_I64;
    case AttributeType::UINT8: return ST_U8;
    case AttributeType::UINT16: return ST_U16;
    case AttributeType::UINT32: return ST_U32;
    case AttributeType::UINT64: return ST_U64;
    case AttributeType::FLOAT: return ST_F32;
    case AttributeType::DOUBLE: return ST_F64;
    default: return ST_NONE;
  }
}

bool has_statistics(const Attribute& a) {
  if (a.type == AttributeType::UNDEFINED || a.size == 0) return false;
  // the wave packet of formats 4-5 has no handler in
  // chunker_countsort_laszip.cpp:534-546
  static const char* const no_handler[] = {"wave packet descriptor index", "byte offset to waveform data",
                                           "waveform packet size", "return point waveform location", "XYZ(t)"};
  for (const char* n : no_handler)
    if (a.name == n) return false;
  return true;
}

bool output_attributes_for(const lasfile::Header& h, Attributes& out, ChunkerParams& P) {
  using T = AttributeType;
  Attribute xyz("position", 12, 3, 4, T::INT32), intensity("intensity", 2, 1, 2, T::UINT16);
  Attribute returnNumber("return number", 1, 1, 1, T::UINT8), numberOfReturns("number of returns", 1, 1, 1, T::UINT8);
  Attribute classification("classification", 1, 1, 1, T::UINT8), scanAngleRank("scan angle rank", 1, 1, 1, T::UINT8);
  Attribute userData("user data", 1, 1, 1, T::UINT8), pointSourceId("point source id", 2, 1, 2, T::UINT16);
  Attribute gpsTime("gps-time", 8, 1, 8, T::DOUBLE), rgb("rgb", 6, 3, 2, T::UINT16);
  Attribute wavePacketDescriptorIndex("wave packet descriptor index", 1, 1, 1, T::UINT8);
  Attribute byteOffsetToWaveformData("byte offset to waveform data", 8, 1, 8, T::UINT64);
  Attribute waveformPacketSize("waveform packet size", 4, 1, 4, T::UINT32);
  Attribute returnPointWaveformLocation("return point waveform location", 4, 1, 4, T::FLOAT);
  Attribute XYZt("XYZ(t)", 12, 3, 4, T::FLOAT), classificationFlags("classification flags", 1, 1, 1, T::UINT8);
  Attribute scanAngle("scan angle", 2, 1, 2, T::INT16);
  std::vector<Attribute> list;
  int fmt = h.point_format;
  switch (fmt) {
    case 0: list = {xyz, intensity, returnNumber, numberOfReturns, classification, scanAngleRank, userData, pointSourceId}; break;
    case 1: list = {xyz, intensity, returnNumber, numberOfReturns, classification, scanAngleRank, userData, pointSourceId, gpsTime}; break;
    case 2: list = {xyz, intensity, returnNumber, numberOfReturns, classification, scanAngleRank, userData, pointSourceId, rgb}; break;
    case 3: list = {xyz, intensity, returnNumber, numberOfReturns, classification, scanAngleRank, userData, pointSourceId, gpsTime, rgb}; break;
    case 4:
      list = {xyz, intensity, returnNumber, numberOfReturns, classification, scanAngleRank, userData, pointSourceId, gpsTime,
              wavePacketDescriptorIndex, byteOffsetToWaveformData, waveformPacketSize, returnPointWaveformLocation, XYZt};
      break;
    case 5:
      list = {xyz, intensity, returnNumber, numberOfReturns, classification, scanAngleRank, userData, pointSourceId, gpsTime,
              rgb, wavePacketDescriptorIndex, byteOffsetToWaveformData, waveformPacketSize, returnPointWaveformLocation,
              XYZt};
      break;
    case 6:
      list = {xyz, intensity, returnNumber, numberOfReturns, classificationFlags, classification, userData, scanAngle,
              pointSourceId, gpsTime};
      break;
    case 7:
      list = {xyz, intensity, returnNumber, numberOfReturns, classificationFlags, classification, userData, scanAngle,
              pointSourceId, gpsTime, rgb};
      break;
    default:
      // PotreeConverter.h:181-185 (formats above 7 end the conversion)
      set_error("currently unsupported LAS format: " + std::to_string(fmt));
      return false;
  }
  const size_t numStandard = list.size();
  // extra bytes (PotreeConverter.h:80-133, LasLoader.cpp:7-69)
  for (auto& e : h.extra) {
    static const T base[] = {T::UNDEFINED, T::UINT8, T::INT8, T::UINT16, T::INT16, T::UINT32, T::INT32, T::UINT64,
                             T::INT64, T::FLOAT, T::DOUBLE};
    T t = T::UNDEFINED;
    int ne = 1;
    if (e.data_type > 0 && e.data_type <= 30) {
      t = base[(e.data_type - 1) % 10 + 1];
      ne = e.data_type <= 10 ? 1 : (e.data_type <= 20 ? 2 : 3);
    }
    int es = attribute_type_size(t);
    Attribute a(e.name, ne * es, ne, es, t);
    a.description = e.description;
    a.scale = Vec3(e.scale[0], e.scale[1], e.scale[2]);
    a.offset = Vec3(e.offset[0], e.offset[1], e.offset[2]);
    list.push_back(a);
  }
  out = Attributes(list);
  Vec3 min(h.min[0], h.min[1], h.min[2]), max(h.max[0], h.max[1], h.max[2]);
  Vec3 scale(h.scale[0], h.scale[1], h.scale[2]);
  Vec3 size = max - min;
  scale.x = std::max(scale.x, size.x / pow(2.0, 30.0));
  scale.y = std::max(scale.y, size.y / pow(2.0, 30.0));
  scale.z = std::max(scale.z, size.z / pow(2.0, 30.0));
  out.posScale = scale;
  out.posOffset = min;

  // transcode + statistics tables (chunker_countsort_laszip.cpp:391-651)
  memset(&P, 0, sizeof(P));
  P.bpp_in = h.record_length;
  P.bpp_out = out.bytes;
  auto op = [&](int src, int dst, int len, int mode) {
    P.ops[P.num_ops++] = TranscodeOp{src, dst, len, mode};
  };
  auto st = [&](int src, int kind) { P.stats[P.num_stats++] = StatOp{src, kind}; };
  // the statistics start with std_stat(fmt, 0 .. std_stat_count(fmt)), the
  // order the count kernel folds them in; the attributes without a handler
  // in the reference (the wave packet of formats 4-5) get neither an op nor
  // a statistic, so their bytes stay 0 and their min/max stay +-inf
  P.point_format = fmt;
  for (int i = 0; i < std_stat_count(fmt); i++) {
    const StdStat ss = std_stat(fmt, i);
    st(ss.src, ss.kind);
  }
  op(12, out.getOffset("intensity"), 2, TC_COPY);
  if (fmt <= 5) {
    op(14, out.getOffset("return number"), 1, TC_LOW3);
    op(14, out.getOffset("number of returns"), 1, TC_MID3);
    op(15, out.getOffset("classification"), 1, TC_LOW5);
    op(16, out.getOffset("scan angle rank"), 1, TC_COPY);
    op(17, out.getOffset("user data"), 1, TC_COPY);
    op(18, out.getOffset("point source id"), 2, TC_COPY);
    if (fmt == 1 || fmt == 3 || fmt == 4 || fmt == 5) op(20, out.getOffset("gps-time"), 8, TC_COPY);
    if (fmt == 2 || fmt == 3 || fmt == 5) op(fmt == 2 ? 20 : 28, out.getOffset("rgb"), 6, TC_COPY);
    P.class_src = 15;
    P.class_mask = 31;
  } else {
    // laszip's point-14 reader: legacy return fields from the 4-bit ones,
    // classification flags, the full classification byte (chunker :467-479
    // takes extended_classification when it is above 31, the 5-bit legacy
    // field - the same byte - otherwise), scan angle as stored
    op(14, out.getOffset("return number"), 1, TC_RET14);
    op(14, out.getOffset("number of returns"), 1, TC_NRET14);
    op(15, out.getOffset("classification flags"), 1, TC_LOW4);
    op(16, out.getOffset("classification"), 1, TC_COPY);
    op(17, out.getOffset("user data"), 1, TC_COPY);
    op(18, out.getOffset("scan angle"), 2, TC_COPY);
    op(20, out.getOffset("point source id"), 2, TC_COPY);
    op(22, out.getOffset("gps-time"), 8, TC_COPY);
    if (fmt == 7) op(30, out.getOffset("rgb"), 6, TC_COPY);
    P.class_src = 16;
    P.class_mask = 255;
  }
  static const int core[] = {20, 28, 26, 34, 57, 63, 30, 36};
  int srcOff = core[fmt];
  int dstOff = 0;
  for (size_t i = 0; i < numStandard; i++) dstOff += out.list[i].size;
  int extraLen = 0;
  for (size_t i = numStandard; i < out.list.size(); i++) {
    const Attribute& a = out.list[i];
    StatKind k = stat_kind_of(a.type);
    for (int e = 0; e < a.numElements && e < 3 && k != ST_NONE; e++) {
      if (P.num_stats >= kMaxStats) { set_error("too many attribute components"); return false; }
      st(srcOff + extraLen + e * a.elementSize, k);
    }
    extraLen += a.size;
  }
  if (extraLen > 0) op(srcOff, dstOff, extraLen, TC_COPY);
  if (srcOff + extraLen > h.record_length) {
    set_error("LAS record shorter than its declared attributes");
    return false;
  }
  return true;
}

bool output_attributes_for_sources(const std::vector<lasfile::Header>& hs, const std::vector<std::string>& requested,
                                   Attributes& out, std::vector<ChunkerParams>& params,
                                   std::vector<std::vector<std::pair<std::string, int>>>& statTargets) {
  if (hs.empty()) {
    set_error("no sources");
    return false;
  }
  std::vector<Attributes> own(hs.size());
  params.assign(hs.size(), ChunkerParams());
  statTargets.assign(hs.size(), {});
  // PotreeConverter.h:195-258: the union of the sources' attribute lists by
  // name (first definition wins; sources in the caller's order), the smallest
  // LAS scale and the bounds of all headers
  std::vector<Attribute> full;
  Vec3 scaleMin(INFINITY, INFINITY, INFINITY), mn(INFINITY, INFINITY, INFINITY), mx(-INFINITY, -INFINITY, -INFINITY);
  for (size_t s = 0; s < hs.size(); s++) {
    if (!output_attributes_for(hs[s], own[s], params[s])) return false;
    for (const Attribute& a : own[s].list) {
      bool seen = false;
      for (const Attribute& b : full) seen = seen || b.name == a.name;
      if (!seen) full.push_back(a);
    }
    for (int k = 0; k < 3; k++) {
      scaleMin[k] = std::min(scaleMin[k], hs[s].scale[k]);
      mn[k] = std::min(mn[k], hs[s].min[k]);
      mx[k] = std::max(mx[k], hs[s].max[k]);
    }
  }
  // :260-283: position first, then the requested names that exist, in the
  // order requested (a name requested twice is listed twice)
  std::vector<Attribute> list = full;
  if (!requested.empty()) {
    std::vector<std::string> should{"position"};
    should.insert(should.end(), requested.begin(), requested.end());
    list.clear();
    for (const std::string& nm : should)
      for (const Attribute& a : full)
        if (a.name == nm) {
          list.push_back(a);
          break;
        }
  }
  out = Attributes(list);
  // computeScaleOffset (:29-76)
  Vec3 size = mx - mn, scale = scaleMin;
  for (int k = 0; k < 3; k++) scale[k] = std::max(scale[k], size[k] / pow(2.0, 30.0));
  out.posScale = scale;
  out.posOffset = mn;
  for (size_t s = 0; s < hs.size(); s++) {
    const Attributes& A = own[s];
    ChunkerParams& P = params[s];
    // the source's own layout -> the output layout, by attribute name
    // (createAttributeHandlers, chunker_countsort_laszip.cpp:380-546): ops of
    // attributes the output lacks are dropped
    auto owner = [&](int dst, std::string& name, int& base) {
      int o = 0;
      for (const Attribute& a : A.list) {
        if (dst >= o && dst < o + a.size) {
          name = a.name;
          base = o;
          return true;
        }
        o += a.size;
      }
      return false;
    };
    size_t numStandard = 0;
    {
      ChunkerParams tmp;
      Attributes a2;
      lasfile::Header h0 = hs[s];
      h0.extra.clear();
      output_attributes_for(h0, a2, tmp);
      numStandard = a2.list.size();
    }
    int extraStart = 0;
    for (size_t i = 0; i < numStandard; i++) extraStart += A.list[i].size;
    std::vector<TranscodeOp> ops;
    for (int i = 0; i < P.num_ops; i++) {
      TranscodeOp op = P.ops[i];
      if (op.dst >= extraStart) continue;  // the extra bytes, below
      std::string name;
      int base = 0;
      if (!owner(op.dst, name, base)) continue;
      const int o = out.getOffset(name);
      if (o < 0) continue;
      op.dst = o + (op.dst - base);
      ops.push_back(op);
    }
    // extra attributes (:548-651): the reference writes each included one at
    // the offset it has in the SOURCE's layout (the sizes of the source's
    // standard attributes plus those of the included extra attributes before
    // it), not at the output attribute's offset; a layout where the two
    // differ would write outside the attribute, so it is refused
    static const int core[] = {20, 28, 26, 34, 57, 63, 30, 36};
    int src = core[hs[s].point_format], refDst = extraStart;
    for (size_t i = numStandard; i < A.list.size(); i++) {
      const Attribute& a = A.list[i];
      const int o = out.getOffset(a.name);
      if (o >= 0) {
        if (o != refDst) {
          set_error("extra attribute \"" + a.name + "\" of source " + std::to_string(s) + " would land at byte " +
                    std::to_string(refDst) + " of the output record instead of its attribute's byte " +
                    std::to_string(o) + " (chunker_countsort_laszip.cpp:579-582 writes it at its source-layout "
                    "offset): attribute filters and source mixes that move extra bytes are not supported");
          return false;
        }
        ops.push_back(TranscodeOp{src, o, a.size, TC_COPY});
        refDst += a.size;
      }
      src += a.size;
    }
    if ((int)ops.size() > kMaxOps) {
      set_error("too many attribute copies for one source");
      return false;
    }
    P.num_ops = (int32_t)ops.size();
    for (size_t i = 0; i < ops.size(); i++) P.ops[i] = ops[i];
    P.bpp_out = out.bytes;
    if (!out.get("classification")) P.class_src = -1;
    // which output attribute component each statistic of this source folds
    // (the count kernel's order: position, then every attribute with a
    // handler, one per component)
    auto& st = statTargets[s];
    for (int k = 0; k < 3; k++) st.emplace_back("position", k);
    for (size_t i = 1; i < A.list.size(); i++) {
      const Attribute& a = A.list[i];
      if (!has_statistics(a)) continue;
      const int comps = a.numElements < 3 ? a.numElements : 3;
      for (int e = 0; e < comps; e++) st.emplace_back(out.get(a.name) ? a.name : std::string(), e);
    }
  }
  return true;
}

std::string fmt17(double v) {
  char buf[64];
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  if (std::isnan(v)) return "nan";
  snprintf(buf, sizeof(buf), "%.17g", v);
  return buf;
}

// chunks/metadata.json as the reference's writeMetadata writes it
// (chunker_countsort_laszip.cpp:1005-1068): an nlohmann::json object dumped
// with dump(4), so keys in lexicographic order, one array element per line,
// numbers in the library's form (json_dump.h).  min/max/scale/offset of an
// attribute exist only for 1-3 elements; the histogram only for one-byte
// attributes that counted something.
std::string chunk_metadata_json(const Vec3& min, const Vec3& max, const Attributes& a) {
  namespace J = jsondump;
  auto vec = [](const Vec3& v, int n, int indent) {
    std::vector<std::string> e{J::number(v.x)};
    if (n >= 2) e.push_back(J::number(v.y));
    if (n >= 3) e.push_back(J::number(v.z));
    return J::array(e, indent);
  };
  const std::string p(12, ' ');
  std::vector<std::string> attrs;
  for (const Attribute& t : a.list) {
    std::string o = "{\n";
    o += p + "\"description\": " + J::string(t.description) + ",\n";
    o += p + "\"elementSize\": " + std::to_string(t.elementSize) + ",\n";
    const bool empty = std::all_of(t.histogram.begin(), t.histogram.end(), [](int64_t h) { return h == 0; });
    if (t.size == 1 && !empty) {
      std::vector<std::string> h;
      for (int64_t c : t.histogram) h.push_back(std::to_string(c));
      o += p + "\"histogram\": " + J::array(h, 12) + ",\n";
    }
    const bool ranged = t.numElements >= 1 && t.numElements <= 3;
    if (ranged) {
      o += p + "\"max\": " + vec(t.max, t.numElements, 12) + ",\n";
      o += p + "\"min\": " + vec(t.min, t.numElements, 12) + ",\n";
    }
    o += p + "\"name\": " + J::string(t.name) + ",\n";
    o += p + "\"numElements\": " + std::to_string(t.numElements) + ",\n";
    if (ranged) {
      o += p + "\"offset\": " + vec(t.offset, t.numElements, 12) + ",\n";
      o += p + "\"scale\": " + vec(t.scale, t.numElements, 12) + ",\n";
    }
    o += p + "\"size\": " + std::to_string(t.size) + ",\n";
    o += p + "\"type\": " + J::string(attribute_type_name(t.type)) + "\n        }";
    attrs.push_back(o);
  }
  std::string s = "{\n    \"attributes\": " + J::array(attrs, 4) + ",\n";
  s += "    \"max\": " + vec(max, 3, 4) + ",\n";
  s += "    \"min\": " + vec(min, 3, 4) + ",\n";
  s += "    \"offset\": " + vec(a.posOffset, 3, 4) + ",\n";
  s += "    \"scale\": " + vec(a.posScale, 3, 4) + "\n}";
  return s;
}

// ---- minimal JSON reader for chunks/metadata.json ----
namespace {
struct J {
  enum K { NUL, NUM, STR, ARR, OBJ } k = NUL;
  double num = 0;
  std::string str;
  std::vector<J> arr;
  std::vector<std::pair<std::string, J>> obj;
  const J* at(const std::string& key) const {
    for (auto& kv : obj) if (kv.first == key) return &kv.second;
    return nullptr;
  }
};
void ws(const char*& p) { while (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r') p++; }
bool parse(const char*& p, J& v) {
  ws(p);
  if (*p == '{') {
    v.k = J::OBJ; p++; ws(p);
    if (*p == '}') { p++; return true; }
    while (*p) {
      J key, val;
      if (!parse(p, key) || key.k != J::STR) return false;
      ws(p);
      if (*p++ != ':') return false;
      if (!parse(p, val)) return false;
      v.obj.push_back({key.str, val});
      ws(p);
      if (*p == ',') { p++; continue; }
      if (*p == '}') { p++; return true; }
      return false;
    }
    return false;
  }
  if (*p == '[') {
    v.k = J::ARR; p++; ws(p);
    if (*p == ']') { p++; return true; }
    while (*p) {
      J e;
      if (!parse(p, e)) return false;
      v.arr.push_back(e);
      ws(p);
      if (*p == ',') { p++; continue; }
      if (*p == ']') { p++; return true; }
      return false;
    }
    return false;
  }
  if (*p == '"') {
    v.k = J::STR; p++;
    while (*p && *p != '"') v.str += *p++;
    if (*p != '"') return false;
    p++;
    return true;
  }
  if (!strncmp(p, "null", 4)) { v.k = J::NUL; p += 4; return true; }
  char* e = nullptr;
  v.k = J::NUM;
  v.num = strtod(p, &e);
  if (e == p) return false;
  p = e;
  return true;
}
}  // namespace

bool parse_chunk_metadata(const std::string& text, Vec3& min, Vec3& max, Attributes& a) {
  const char* p = text.c_str();
  J js;
  if (!parse(p, js) || js.k != J::OBJ) { set_error("malformed chunks/metadata.json"); return false; }
  auto v3 = [](const J* x, Vec3& out) {
    if (!x || x->arr.size() < 3) return false;
    out = Vec3(x->arr[0].num, x->arr[1].num, x->arr[2].num);
    return true;
  };
  Vec3 sc, of;
  if (!v3(js.at("min"), min) || !v3(js.at("max"), max) || !v3(js.at("scale"), sc) || !v3(js.at("offset"), of)) {
    set_error("chunks/metadata.json lacks min/max/scale/offset");
    return false;
  }
  std::vector<Attribute> list;
  const J* ja = js.at("attributes");
  if (!ja) { set_error("chunks/metadata.json lacks attributes"); return false; }
  for (auto& o : ja->arr) {
    Attribute t(o.at("name")->str, (int)o.at("size")->num, (int)o.at("numElements")->num,
                (int)o.at("elementSize")->num, attribute_type_from_name(o.at("type")->str));
    t.description = o.at("description") ? o.at("description")->str : "";
    if (const J* h = o.at("histogram"))
      for (size_t i = 0; i < h->arr.size() && i < 256; i++) t.histogram[i] = (int64_t)h->arr[i].num;
    // indexer.cpp:106-123: a missing (null) min or max reads back as +inf
    auto comp = [](const J* arr, int i, double dflt) {
      if (!arr || (int)arr->arr.size() <= i || arr->arr[i].k == J::NUL) return dflt;
      return arr->arr[i].num;
    };
    for (int i = 0; i < t.numElements && i < 3; i++) {
      t.min[i] = comp(o.at("min"), i, INFINITY);
      t.max[i] = comp(o.at("max"), i, INFINITY);
      t.scale[i] = comp(o.at("scale"), i, 1.0);
      t.offset[i] = comp(o.at("offset"), i, 0.0);
    }
    list.push_back(t);
  }
  a = Attributes(list);
  a.posScale = sc;
  a.posOffset = of;
  return true;
}

}  // namespace pcg
