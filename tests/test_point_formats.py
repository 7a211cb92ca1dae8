"""LAS point formats 0-7: record decode, output layout and transcode.

The reference reads points through LASzip; for uncompressed files that is
LASzip's raw item readers (Converter/libs/laszip/src/lasreaditemraw.hpp),
which oracle/build_ref.sh compiles from the reference checkout into
oracle/_ref/ref_point_decode.  The oracle's restated decode must give the
same laszip point fields for every format, including the point-14 mapping
of 4-bit return fields onto the 3-bit legacy ones and of classification
bytes above 31.  The chunker's attribute handlers
(chunker_countsort_laszip.cpp:380-658) are then checked on the oracle's
chunk files against those decoded fields.
"""
import ctypes
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import potreeconverter_amd as pc
from helpers import REF_POINT_DECODE, make_las, oracle, oracle_convert, read

CORE = {0: 20, 1: 28, 2: 26, 3: 34, 4: 57, 5: 63, 6: 30, 7: 36}
# output record bytes of each format (PotreeConverter.h:136-186 attribute sizes)
OUT = {0: 21, 1: 29, 2: 27, 3: 35, 4: 58, 5: 64, 6: 31, 7: 37}


def generated_records(fmt, n, seed, extra):
    g = pc.LasGen(0, n, seed=seed, extra_bytes=extra, point_format=fmt)
    rl = g.record_length
    assert rl == CORE[fmt] + (7 if extra else 0)
    a = np.frombuffer(g.records(), dtype=np.uint8).reshape(n, rl).copy()
    return a, rl


def with_edge_bytes(a, fmt):
    """Every value of the returns byte, the flags byte and the class byte."""
    n = a.shape[0]
    k = np.arange(n)
    a[:, 14] = (k % 256).astype(np.uint8)
    a[:, 15] = ((k // 256) % 256).astype(np.uint8)
    a[:, 16 if fmt >= 6 else 15] = ((k * 7) % 256).astype(np.uint8)
    return a


def oracle_decode(fmt, a, rl):
    out = np.zeros((a.shape[0], 32), dtype=np.uint8)
    buf = np.ascontiguousarray(a)
    rc = oracle().orc_decode_points(fmt, rl, buf.ctypes.data_as(ctypes.c_void_p), a.shape[0],
                                    out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    return out


def laszip_decode(fmt, a, rl, tmp_path):
    src, dst = tmp_path / f"r{fmt}.bin", tmp_path / f"d{fmt}.bin"
    np.ascontiguousarray(a).tofile(src)
    subprocess.run([REF_POINT_DECODE, str(fmt), str(rl), str(src), str(dst)], check=True)
    return np.fromfile(dst, dtype=np.uint8).reshape(a.shape[0], 32)


@pytest.mark.skipif(not os.path.exists(REF_POINT_DECODE), reason="oracle/_ref not built (bash oracle/build_ref.sh)")
@pytest.mark.parametrize("fmt", range(8))
def test_oracle_decode_equals_laszip_raw_readers(fmt, tmp_path):
    a, rl = generated_records(fmt, 20_000, 3, fmt % 2 == 1)
    a = np.concatenate([a, with_edge_bytes(a[:65536 % len(a)].copy(), fmt)])
    got = oracle_decode(fmt, a, rl)
    ref = laszip_decode(fmt, a, rl, tmp_path)
    bad = np.nonzero((got != ref).any(axis=1))[0]
    assert len(bad) == 0, (fmt, bad[:5], got[bad[:2]], ref[bad[:2]])


def test_point14_legacy_returns_cases():
    """The 4-bit -> 3-bit return mapping of laszip's point-14 reader on the
    cases that matter: up to 7 returns copies, more than 7 clamps the count
    to 7 and maps late returns to 6 or (the last) 7."""
    a, rl = generated_records(6, 64, 1, False)
    cases = [(1, 1, 1, 1), (3, 5, 3, 5), (7, 7, 7, 7), (9, 5, 1, 5), (2, 9, 2, 7), (6, 12, 6, 7),
             (7, 12, 6, 7), (11, 12, 6, 7), (12, 12, 7, 7), (15, 12, 7, 7), (15, 15, 7, 7), (0, 0, 0, 0)]
    for i, (r, nr, _, _) in enumerate(cases):
        a[i, 14] = r | (nr << 4)
    d = oracle_decode(6, a[:len(cases)], rl)
    for i, (_, _, er, enr) in enumerate(cases):
        assert (d[i, 2], d[i, 3]) == (er, enr), cases[i]


def chunk_records(out, bpo):
    files = sorted(f for f in os.listdir(out / "chunks") if f.endswith(".bin"))
    return np.concatenate([np.frombuffer(read(out / "chunks" / f), dtype=np.uint8).reshape(-1, bpo) for f in files])


@pytest.mark.parametrize("fmt", range(8))
def test_oracle_chunks_apply_the_handlers(fmt, tmp_path):
    """Every output record holds the handler values of its laszip fields:
    the attribute list of the format, classification = extended class if
    above 31 else the 5-bit class, no bytes for attributes without a handler
    (the wave packet), extra bytes copied; metadata min/max are those of the
    handler values."""
    n, extra = 30_000, fmt in (1, 4, 6)
    las = tmp_path / "a.las"
    make_las(las, 0, n, 5, extra, fmt)
    out = tmp_path / "o"
    oracle_convert(las, out, "poisson", chunk_only=True)
    meta = json.load(open(out / "chunks" / "metadata.json"))
    attrs = meta["attributes"]
    bpo = sum(a["size"] for a in attrs)
    assert bpo == OUT[fmt] + (7 if extra else 0)
    recs = chunk_records(out, bpo)
    assert len(recs) == n
    raw = read(las)
    off = struct.unpack_from("<I", raw, 96)[0]
    rl = struct.unpack_from("<H", raw, 105)[0]
    src = np.frombuffer(raw[off:off + n * rl], dtype=np.uint8).reshape(n, rl)
    dec = oracle_decode(fmt, src, rl)
    # match output records to inputs by the (requantised) position + intensity
    # + gps/user fields: compare the multisets of per-attribute value columns
    offs, o = {}, 0
    for a in attrs:
        offs[a["name"]] = (o, a["size"])
        o += a["size"]

    def col(arr, name):
        s, z = offs[name]
        return np.sort(arr[:, s:s + z].copy().view(f"<u{z}" if z in (1, 2, 4, 8) else np.uint8).ravel())

    exp = {
        "intensity": dec[:, 0:2], "return number": dec[:, 2:3], "number of returns": dec[:, 3:4],
        "user data": dec[:, 6:7], "point source id": dec[:, 8:10],
    }
    cls = np.where(dec[:, 12] > 31, dec[:, 12], dec[:, 4]).astype(np.uint8)[:, None]
    exp["classification"] = cls
    if fmt <= 5:
        exp["scan angle rank"] = dec[:, 5:6]
    else:
        exp["scan angle"] = dec[:, 10:12]
        exp["classification flags"] = dec[:, 7:8]
    if fmt in (1, 3, 4, 5, 6, 7):
        exp["gps-time"] = dec[:, 24:32]
    for name, e in exp.items():
        z = e.shape[1]
        assert np.array_equal(col(recs, name), np.sort(np.ascontiguousarray(e).view(f"<u{z}").ravel())), name
    if fmt in (2, 3, 5, 7):
        s, _ = offs["rgb"]
        assert np.array_equal(np.sort(recs[:, s:s + 6].copy().view("<u2").reshape(-1, 3), axis=0),
                              np.sort(dec[:, 14:20].copy().view("<u2").reshape(-1, 3), axis=0))
    for a in attrs:
        name = a["name"]
        if name in ("wave packet descriptor index", "byte offset to waveform data", "waveform packet size",
                    "return point waveform location", "XYZ(t)"):
            s, z = offs[name]
            assert not recs[:, s:s + z].any(), name
            assert all(v is None for v in a["min"] + a["max"]), name
    # classification histogram and min/max of the handler values
    c = next(a for a in attrs if a["name"] == "classification")
    assert c["histogram"] == np.bincount(cls.ravel(), minlength=256).tolist()
    assert c["min"] == [float(cls.min())] and c["max"] == [float(cls.max())]
    if extra:
        s, _ = offs["gauss"]
        assert np.array_equal(np.sort(recs[:, s:s + 7].copy().view(np.uint8).reshape(-1, 7), axis=0),
                              np.sort(src[:, CORE[fmt]:CORE[fmt] + 7], axis=0))
