"""chunks/metadata.json text against nlohmann::json 3.7.3's dump(4).

potreeconverter_amd/csrc/json_dump.h restates the library's number printing
(Grisu2) and string escaping; the fixtures were written by the vendored
json.hpp itself (tests/golden/make_json_golden.py)."""
import json
import os
import re
import struct
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from helpers import make_las, oracle_convert

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "potreeconverter_amd", "csrc")
GOLDEN = os.path.join(ROOT, "tests", "golden")
JSON_HPP_DIR = "/root/reference/Converter/libs/json"

HARNESS = r'''
#include <cstdint>
#include <cstring>
#include <iostream>
#include <string>
#include "json_dump.h"
int main() {
  std::string line;
  while (std::getline(std::cin, line)) {
    uint64_t b = std::stoull(line, nullptr, 16);
    double v;
    std::memcpy(&v, &b, 8);
    std::cout << pcg::jsondump::number(v) << "\n";
  }
}
'''


@pytest.fixture(scope="module")
def formatter(tmp_path_factory):
    td = tmp_path_factory.mktemp("jd")
    src = td / "h.cpp"
    src.write_text(HARNESS)
    exe = td / "h"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True)

    def run(hexes):
        out = subprocess.run([str(exe)], input="\n".join(hexes) + "\n", capture_output=True, text=True, check=True)
        return out.stdout.splitlines()
    return run


def test_power_table_is_ten_to_the_k():
    """kPow10 = 10^k rounded to the nearest 64-bit significand, k = -300 .. 324 step 8."""
    text = open(os.path.join(CSRC, "json_dump.h")).read()
    rows = re.findall(r"\{0x([0-9A-F]+)ull, (-?\d+), (-?\d+)\}", text)
    assert len(rows) == 79
    for i, (f, e, k) in enumerate(rows):
        f, e, k = int(f, 16), int(e), int(k)
        assert k == -300 + 8 * i
        exact = Fraction(10) ** k / Fraction(2) ** e
        assert 2**63 <= f < 2**64
        assert abs(exact - f) <= Fraction(1, 2)


def test_numbers_equal_nlohmann_fixture(formatter):
    fx = json.load(open(os.path.join(GOLDEN, "json_numbers.json")))
    got = formatter(fx["bits"])
    bad = [(b, w, g) for b, w, g in zip(fx["bits"], fx["text"], got) if w != g]
    assert len(got) == len(fx["text"]) and not bad, bad[:10]


def test_numbers_round_trip(formatter):
    rng = np.random.default_rng(5)
    bits = ["%016x" % int(b) for b in rng.integers(0, 0x7FEFFFFFFFFFFFFF, size=20000, dtype=np.int64)]
    for b, t in zip(bits, formatter(bits)):
        assert struct.unpack("<d", struct.pack("<Q", int(b, 16)))[0] == float(t)


@pytest.mark.skipif(not os.path.isdir(JSON_HPP_DIR), reason="reference checkout not present")
def test_numbers_equal_live_nlohmann(formatter, tmp_path):
    """200 000 more random doubles against the vendored json.hpp itself."""
    src = tmp_path / "n.cpp"
    src.write_text(r'''
#include <cstdint>
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
''')
    exe = tmp_path / "n"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", JSON_HPP_DIR, str(src), "-o", str(exe)], check=True)
    rng = np.random.default_rng(11)
    raw = rng.integers(0, 2**63 - 1, size=150000, dtype=np.int64)
    dec = (rng.integers(-10**12, 10**12, size=50000) * 0.001).view(np.int64)
    bits = ["%016x" % (int(b) & 0x7FFFFFFFFFFFFFFF) for b in np.concatenate([raw, dec])]
    bits = [b for b in bits if int(b, 16) < 0x7FF0000000000000]
    want = subprocess.run([str(exe)], input="\n".join(bits) + "\n", capture_output=True, text=True,
                          check=True).stdout.splitlines()
    got = formatter(bits)
    bad = [(b, w, g) for b, w, g in zip(bits, want, got) if w != g]
    assert not bad, bad[:10]


CASES = sorted(p for p in os.listdir(GOLDEN) if p.startswith("chunks_metadata_"))


@pytest.mark.parametrize("name", CASES)
def test_oracle_chunk_metadata_equals_nlohmann_dump(name, tmp_path):
    fx = json.load(open(os.path.join(GOLDEN, name)))
    c = fx["case"]
    las = tmp_path / "in.las"
    make_las(las, c["preset"], c["points"], c["seed"], c["extra_bytes"], c["point_format"])
    out = tmp_path / "out"
    oracle_convert(las, out, "poisson", keep_chunks=True, chunk_only=True)
    assert open(out / "chunks" / "metadata.json").read() == fx["text"]
