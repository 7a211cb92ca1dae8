"""Writes the fixtures that pin chunks/metadata.json to nlohmann::json.

The reference writes chunks/metadata.json with nlohmann::json 3.7.3's
dump(4) (Converter/src/chunker_countsort_laszip.cpp:1005-1068; the library
is vendored at Converter/libs/json/json.hpp).  This script compiles two small
drivers against that header, straight from the reference checkout:

  json_numbers.json      json(x).dump() of 5865 doubles: random bit patterns,
                         decimal coordinates, integers, powers of ten and two
                         and their neighbours, subnormals, extremes
  chunks_metadata_*.json the oracle's chunks/metadata.json of a case, read by
                         nlohmann::json, the position and attribute
                         min/max/scale/offset entries made doubles (as
                         writeMetadata builds them), written with dump(4)

Run from the repository root:  python tests/golden/make_json_golden.py
"""
import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from helpers import make_las, oracle_convert  # noqa: E402

JSON_HPP_DIR = "/root/reference/Converter/libs/json"

NUMBERS_DRIVER = r'''
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include "json.hpp"
int main() {
  std::string line;
  while (std::getline(std::cin, line)) {
    uint64_t b = std::stoull(line, nullptr, 16);
    double v;
    std::memcpy(&v, &b, 8);
    std::cout << nlohmann::json(v).dump() << "\n";
  }
}
'''

METADATA_DRIVER = r'''
#include <iostream>
#include <iterator>
#include <string>
#include "json.hpp"
using nlohmann::json;
static json doubles(const json& a) {
  json r = json::array();
  for (auto& v : a) r.push_back(v.is_null() ? json(nullptr) : json(v.get<double>()));
  return r;
}
int main() {
  std::string text((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
  json j = json::parse(text);
  for (const char* k : {"min", "max", "scale", "offset"}) j[k] = doubles(j[k]);
  for (auto& a : j["attributes"])
    for (const char* k : {"min", "max", "scale", "offset"}) a[k] = doubles(a[k]);
  std::cout << j.dump(4);
}
'''

METADATA_CASES = [
    # name, preset, points, seed, extra bytes, point format
    ("chunks_metadata_c0_2m", 0, 2_000_000, 0, False, 2),
    ("chunks_metadata_p1_pdrf7_extra", 1, 300_000, 9, True, 7),
    ("chunks_metadata_p2_pdrf3", 2, 300_000, 5, False, 3),
]


def build(td, name, src):
    cpp = os.path.join(td, name + ".cpp")
    with open(cpp, "w") as f:
        f.write(src)
    exe = os.path.join(td, name)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", JSON_HPP_DIR, cpp, "-o", exe], check=True)
    return exe


def number_inputs():
    rng = np.random.default_rng(20260)
    vals = []
    bits = rng.integers(0, 2**63 - 1, size=1500, dtype=np.int64).astype(np.uint64)
    vals += [struct.unpack("<d", struct.pack("<Q", int(b)))[0] for b in bits]
    vals += list(rng.integers(-10**9, 10**9, size=800) * 0.001)
    vals += list(rng.integers(0, 10**8, size=400) * 0.0001 + 1e6)
    vals += list(rng.integers(0, 10**10, size=400) * 0.001 + 5e6)
    vals += [float(v) for v in rng.integers(-2**53, 2**53, size=300)]
    vals += list(rng.random(400) * 10.0 ** rng.integers(-12, 20, size=400))
    for e in range(-323, 309):
        vals.append(float("1e%d" % e))
    vals += [5e-324 * k for k in range(1, 8)]
    for e in range(-1074, 1024, 7):
        v = 2.0 ** e
        vals += [v, np.nextafter(v, 0), np.nextafter(v, np.inf)]
    vals += [5e-324, 2.2250738585072014e-308, 2.225073858507201e-308, 1.7976931348623157e308, 0.1, 0.2, 0.3,
             1 / 3, 2 / 3, 0.001, 0.0001, 500000.0, 5000000.0, 1e6, 65535.0, 1e15, 1e16, 1e-4, 1e-5, 123456789012345.6,
             9007199254740993.0, 0.5, 1.5]
    vals = [float(v) for v in vals if np.isfinite(v)]
    vals += [-v for v in vals[:500]]
    vals += [0.0, -0.0]
    return vals[:6000] + vals[-2:]


def main():
    assert os.path.isdir(JSON_HPP_DIR), "reference checkout with Converter/libs/json needed"
    with tempfile.TemporaryDirectory() as td:
        nd = build(td, "numbers", NUMBERS_DRIVER)
        vals = number_inputs()
        hexes = ["%016x" % struct.unpack("<Q", struct.pack("<d", v))[0] for v in vals]
        out = subprocess.run([nd], input="\n".join(hexes) + "\n", capture_output=True, text=True, check=True).stdout
        texts = out.splitlines()
        assert len(texts) == len(hexes)
        with open(os.path.join(HERE, "json_numbers.json"), "w") as f:
            json.dump({"producer": "nlohmann::json 3.7.3 (Converter/libs/json/json.hpp) json(x).dump()",
                       "bits": hexes, "text": texts}, f, indent=0)
        print("json_numbers", len(texts))
        md = build(td, "metadata", METADATA_DRIVER)
        for name, preset, n, seed, eb, fmt in METADATA_CASES:
            las = os.path.join(td, name + ".las")
            make_las(las, preset, n, seed, eb, fmt)
            work = os.path.join(td, name)
            oracle_convert(las, work, "poisson", keep_chunks=True, chunk_only=True)
            with open(os.path.join(work, "chunks", "metadata.json")) as f:
                mine = f.read()
            text = subprocess.run([md], input=mine, capture_output=True, text=True, check=True).stdout
            case = {"preset": preset, "points": n, "seed": seed, "extra_bytes": eb, "point_format": fmt}
            with open(os.path.join(HERE, name + ".json"), "w") as f:
                json.dump({"case": case, "producer": "nlohmann::json 3.7.3 dump(4) of the document writeMetadata "
                           "builds (chunker_countsort_laszip.cpp:1005-1068)", "text": text}, f, indent=0)
            print(name, len(text), "bytes")


if __name__ == "__main__":
    main()
