// JSON text as nlohmann::json 3.7.3 writes it with dump(4): the layout of
// chunks/metadata.json (Converter/src/chunker_countsort_laszip.cpp:1005-1068,
// `js.dump(4)`; the vendored library is Converter/libs/json/json.hpp).
//
// Numbers: the library prints a double with the Grisu2 algorithm
// (F. Loitsch, "Printing Floating-Point Numbers Quickly and Accurately with
// Integers", PLDI 2010): digits of the upper boundary of the value's
// rounding interval, scaled by a cached power of ten into a 64-bit
// fixed-point window, generated until they fall inside the (conservatively
// narrowed) interval, then the last digit moved toward the value.  The
// digit string is usually, not always, the shortest that round-trips; to
// reproduce the library byte for byte the same steps are taken here: the
// same boundaries, the same choice of cached power (the first power 10^k,
// k = -300 + 8i, at or above the one that places the product in the window
// [-60, -32]), 64-bit products rounded half up, and the same exponent form
// (plain decimal for decimal exponents in (-4, 15], d.ddde+XX otherwise,
// integral values with a trailing ".0").  Non-finite values print "null".
//
// kPow10 holds 10^k rounded to the nearest 64-bit significand f with
// 2^63 <= f < 2^64 (10^k ~= f * 2^e), for k = -300, -292, ..., 324; the
// table is recomputed from exact rationals in tests/test_json_dump.py.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

namespace pcg {
namespace jsondump {

struct Pow10 {
  uint64_t f;
  int e;
  int k;
};

static const Pow10 kPow10[] = {
    {0xAB70FE17C79AC6CAull, -1060, -300}, {0xFF77B1FCBEBCDC4Full, -1034, -292},
    {0xBE5691EF416BD60Cull, -1007, -284}, {0x8DD01FAD907FFC3Cull, -980, -276},
    {0xD3515C2831559A83ull, -954, -268}, {0x9D71AC8FADA6C9B5ull, -927, -260},
    {0xEA9C227723EE8BCBull, -901, -252}, {0xAECC49914078536Dull, -874, -244},
    {0x823C12795DB6CE57ull, -847, -236}, {0xC21094364DFB5637ull, -821, -228},
    {0x9096EA6F3848984Full, -794, -220}, {0xD77485CB25823AC7ull, -768, -212},
    {0xA086CFCD97BF97F4ull, -741, -204}, {0xEF340A98172AACE5ull, -715, -196},
    {0xB23867FB2A35B28Eull, -688, -188}, {0x84C8D4DFD2C63F3Bull, -661, -180},
    {0xC5DD44271AD3CDBAull, -635, -172}, {0x936B9FCEBB25C996ull, -608, -164},
    {0xDBAC6C247D62A584ull, -582, -156}, {0xA3AB66580D5FDAF6ull, -555, -148},
    {0xF3E2F893DEC3F126ull, -529, -140}, {0xB5B5ADA8AAFF80B8ull, -502, -132},
    {0x87625F056C7C4A8Bull, -475, -124}, {0xC9BCFF6034C13053ull, -449, -116},
    {0x964E858C91BA2655ull, -422, -108}, {0xDFF9772470297EBDull, -396, -100},
    {0xA6DFBD9FB8E5B88Full, -369, -92}, {0xF8A95FCF88747D94ull, -343, -84},
    {0xB94470938FA89BCFull, -316, -76}, {0x8A08F0F8BF0F156Bull, -289, -68},
    {0xCDB02555653131B6ull, -263, -60}, {0x993FE2C6D07B7FACull, -236, -52},
    {0xE45C10C42A2B3B06ull, -210, -44}, {0xAA242499697392D3ull, -183, -36},
    {0xFD87B5F28300CA0Eull, -157, -28}, {0xBCE5086492111AEBull, -130, -20},
    {0x8CBCCC096F5088CCull, -103, -12}, {0xD1B71758E219652Cull, -77, -4},
    {0x9C40000000000000ull, -50, 4}, {0xE8D4A51000000000ull, -24, 12},
    {0xAD78EBC5AC620000ull, 3, 20}, {0x813F3978F8940984ull, 30, 28},
    {0xC097CE7BC90715B3ull, 56, 36}, {0x8F7E32CE7BEA5C70ull, 83, 44},
    {0xD5D238A4ABE98068ull, 109, 52}, {0x9F4F2726179A2245ull, 136, 60},
    {0xED63A231D4C4FB27ull, 162, 68}, {0xB0DE65388CC8ADA8ull, 189, 76},
    {0x83C7088E1AAB65DBull, 216, 84}, {0xC45D1DF942711D9Aull, 242, 92},
    {0x924D692CA61BE758ull, 269, 100}, {0xDA01EE641A708DEAull, 295, 108},
    {0xA26DA3999AEF774Aull, 322, 116}, {0xF209787BB47D6B85ull, 348, 124},
    {0xB454E4A179DD1877ull, 375, 132}, {0x865B86925B9BC5C2ull, 402, 140},
    {0xC83553C5C8965D3Dull, 428, 148}, {0x952AB45CFA97A0B3ull, 455, 156},
    {0xDE469FBD99A05FE3ull, 481, 164}, {0xA59BC234DB398C25ull, 508, 172},
    {0xF6C69A72A3989F5Cull, 534, 180}, {0xB7DCBF5354E9BECEull, 561, 188},
    {0x88FCF317F22241E2ull, 588, 196}, {0xCC20CE9BD35C78A5ull, 614, 204},
    {0x98165AF37B2153DFull, 641, 212}, {0xE2A0B5DC971F303Aull, 667, 220},
    {0xA8D9D1535CE3B396ull, 694, 228}, {0xFB9B7CD9A4A7443Cull, 720, 236},
    {0xBB764C4CA7A44410ull, 747, 244}, {0x8BAB8EEFB6409C1Aull, 774, 252},
    {0xD01FEF10A657842Cull, 800, 260}, {0x9B10A4E5E9913129ull, 827, 268},
    {0xE7109BFBA19C0C9Dull, 853, 276}, {0xAC2820D9623BF429ull, 880, 284},
    {0x80444B5E7AA7CF85ull, 907, 292}, {0xBF21E44003ACDD2Dull, 933, 300},
    {0x8E679C2F5E44FF8Full, 960, 308}, {0xD433179D9C8CB841ull, 986, 316},
    {0x9E19DB92B4E31BA9ull, 1013, 324},
};

struct Fp {
  uint64_t f;
  int e;
};

inline Fp fp_mul(Fp a, Fp b) {
  const unsigned __int128 p = (unsigned __int128)a.f * b.f + ((unsigned __int128)1 << 63);
  return Fp{(uint64_t)(p >> 64), a.e + b.e + 64};
}

inline Fp fp_normalize(Fp x) {
  while ((x.f >> 63) == 0) { x.f <<= 1; x.e--; }
  return x;
}

// digits of a positive finite double into buf; returns the digit count and
// sets the decimal exponent (value = digits * 10^dexp)
inline int grisu2(double v, char* buf, int& dexp) {
  uint64_t bits;
  memcpy(&bits, &v, 8);
  const uint64_t F = bits & ((1ull << 52) - 1);
  const int E = (int)(bits >> 52);
  Fp w = E == 0 ? Fp{F, 1 - 1075} : Fp{F + (1ull << 52), E - 1075};
  // rounding interval: halfway to the neighbours (the lower neighbour is
  // closer when the significand is a power of two above the smallest binade)
  const bool lowerCloser = F == 0 && E > 1;
  Fp mPlus = fp_normalize(Fp{2 * w.f + 1, w.e - 1});
  Fp mMinus = lowerCloser ? Fp{4 * w.f - 1, w.e - 2} : Fp{2 * w.f - 1, w.e - 1};
  mMinus.f <<= (mMinus.e - mPlus.e);
  mMinus.e = mPlus.e;
  w = fp_normalize(w);
  // cached power: c.e + w.e + 64 in [-60, -32]
  const int need = -60 - mPlus.e - 1;
  const int kMin = need * 78913 / (1 << 18) + (need > 0 ? 1 : 0);  // ceil(need * log10(2))
  const int idx = (300 + kMin + 7) / 8;
  const Pow10& c = kPow10[idx];
  const Fp cp{c.f, c.e};
  const Fp W = fp_mul(w, cp);
  Fp Lo = fp_mul(mMinus, cp), Hi = fp_mul(mPlus, cp);
  Lo.f += 1;
  Hi.f -= 1;
  dexp = -c.k;
  const int sh = -Hi.e;
  const uint64_t one = 1ull << sh;
  uint64_t delta = Hi.f - Lo.f;
  uint64_t dist = Hi.f - W.f;
  uint32_t p1 = (uint32_t)(Hi.f >> sh);
  uint64_t p2 = Hi.f & (one - 1);
  int len = 0;
  // move the last digit toward W while that stays inside the interval and
  // brings the digits closer to W
  auto round_toward = [&](uint64_t rest, uint64_t ten) {
    while (rest < dist && delta - rest >= ten && (rest + ten < dist || dist - rest > rest + ten - dist)) {
      buf[len - 1]--;
      rest += ten;
    }
  };
  int n = 1;
  uint32_t pw = 1;
  while (n < 10 && p1 >= pw * 10u) { pw *= 10; n++; }
  while (n > 0) {
    buf[len++] = (char)('0' + p1 / pw);
    p1 %= pw;
    n--;
    const uint64_t rest = ((uint64_t)p1 << sh) + p2;
    if (rest <= delta) {
      dexp += n;
      round_toward(rest, (uint64_t)pw << sh);
      return len;
    }
    pw /= 10;
  }
  int m = 0;
  for (;;) {
    p2 *= 10;
    buf[len++] = (char)('0' + (p2 >> sh));
    p2 &= one - 1;
    m++;
    delta *= 10;
    dist *= 10;
    if (p2 <= delta) break;
  }
  dexp -= m;
  round_toward(p2, one);
  return len;
}

inline std::string number(double v) {
  if (!std::isfinite(v)) return "null";
  std::string out;
  if (std::signbit(v)) {
    out += '-';
    v = -v;
  }
  if (v == 0) return out + "0.0";
  char d[32];
  int dexp = 0;
  const int k = grisu2(v, d, dexp);
  const int n = k + dexp;  // position of the decimal point
  if (k <= n && n <= 15) {
    out.append(d, k);
    out.append((size_t)(n - k), '0');
    return out + ".0";
  }
  if (0 < n && n <= 15) {
    out.append(d, n);
    out += '.';
    out.append(d + n, k - n);
    return out;
  }
  if (-4 < n && n <= 0) {
    out += "0.";
    out.append((size_t)-n, '0');
    out.append(d, k);
    return out;
  }
  out += d[0];
  if (k > 1) {
    out += '.';
    out.append(d + 1, k - 1);
  }
  int x = n - 1;
  out += 'e';
  out += x < 0 ? '-' : '+';
  if (x < 0) x = -x;
  if (x < 10) out += '0';
  out += std::to_string(x);
  return out;
}

// a JSON string (UTF-8 passed through, control characters escaped)
inline std::string string(const std::string& s) {
  std::string out = "\"";
  for (unsigned char ch : s) {
    switch (ch) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\b': out += "\\b"; break;
      case '\f': out += "\\f"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      default:
        if (ch < 0x20) {
          char u[8];
          snprintf(u, sizeof(u), "\\u%04x", ch);
          out += u;
        } else {
          out += (char)ch;
        }
    }
  }
  return out + "\"";
}

// a non-empty array as dump(4) lays it out at nesting `indent` (spaces):
// one element per line, four spaces deeper; an empty one as []
inline std::string array(const std::vector<std::string>& items, int indent) {
  if (items.empty()) return "[]";
  const std::string pad((size_t)indent + 4, ' ');
  std::string out = "[\n";
  for (size_t i = 0; i < items.size(); i++) {
    out += pad;
    out += items[i];
    out += i + 1 < items.size() ? ",\n" : "\n";
  }
  return out + std::string((size_t)indent, ' ') + "]";
}

}  // namespace jsondump
}  // namespace pcg
