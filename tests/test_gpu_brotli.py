"""--encoding BROTLI on the gfx950 path against the CPU oracle, byte for byte.

The oracle's BROTLI encoding is pinned to reference-indexer output by
tests/test_oracle_golden.py (the *_brotli fixtures).  The GPU lays every node
out (attribute arrays, position and rgb as Morton codes, points in
position-code order with std::sort's order for equal codes) and compresses
the nodes with the same brotli release on host threads, so octree.bin,
hierarchy.bin and metadata.json must equal the oracle's.

At sizes where the oracle's compression would take long, the BROTLI output
is checked against the same conversion's DEFAULT output instead: every node
decompresses (system libbrotlidec) to the arrays that its DEFAULT records
define, in ascending position-code order.
"""
import ctypes
import os
import shutil

import numpy as np
import pytest

import potreeconverter_amd as pc
from helpers import make_las, oracle_convert
from potree_format import canonical_nodes

pytestmark = pytest.mark.gpu

FILES = ["octree.bin", "hierarchy.bin", "metadata.json"]

CASES = [
    # preset, points, seed, extra bytes, point format, method
    (0, 150_000, 2, False, 2, "poisson"),
    (0, 400_000, 5, True, 2, "poisson_average"),
    (1, 1_000_000, 1, False, 2, "poisson"),
    (2, 300_000, 11, True, 2, "poisson"),          # near-duplicates: equal position codes
    (2, 1_000_000, 3, False, 2, "poisson_average"),
    (0, 300_000, 26, True, 6, "poisson"),          # no rgb
    (0, 300_000, 27, False, 7, "poisson_average"),
]


def _files_equal(a, b):
    out = {}
    for n in FILES:
        with open(os.path.join(a, n), "rb") as fa, open(os.path.join(b, n), "rb") as fb:
            out[n] = fa.read() == fb.read()
    return out


@pytest.mark.parametrize("preset,n,seed,eb,fmt,method", CASES,
                         ids=[f"p{c[0]}_{c[1]}_s{c[2]}_pdrf{c[4]}_{c[5]}{'_eb' if c[3] else ''}" for c in CASES])
def test_brotli_convert_matches_oracle(preset, n, seed, eb, fmt, method, tmp_path):
    print(f"brotli case preset={preset} n={n} seed={seed} extra={eb} pdrf={fmt} method={method}", flush=True)
    las = tmp_path / "in.las"
    make_las(las, preset, n, seed, eb, fmt)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert(las, cpu, method, encoding="BROTLI")
    pc.convert(las, gpu, method, encoding="BROTLI")
    res = _files_equal(cpu, gpu)
    assert all(res.values()), res


def test_other_encoding_names_write_records_unchanged(tmp_path):
    """indexer.cpp:1462: only "BROTLI" compresses; the name is echoed."""
    las = tmp_path / "in.las"
    make_las(las, 0, 200_000, 9)
    cpu, gpu, plain = tmp_path / "cpu", tmp_path / "gpu", tmp_path / "plain"
    oracle_convert(las, cpu, "poisson", encoding="UNCOMPRESSED")
    pc.convert(las, gpu, "poisson", encoding="UNCOMPRESSED")
    assert all(_files_equal(cpu, gpu).values())
    pc.convert(las, plain, "poisson")
    for n in ("octree.bin", "hierarchy.bin"):
        assert (gpu / n).read_bytes() == (plain / n).read_bytes()
    assert '"encoding": "UNCOMPRESSED"' in (gpu / "metadata.json").read_text()


def test_brotli_indexing_of_reference_format_chunks(tmp_path):
    las = tmp_path / "in.las"
    make_las(las, 2, 400_000, 13, True)
    cpu = tmp_path / "cpu"
    oracle_convert(las, cpu, "poisson_average", keep_chunks=True, encoding="BROTLI")
    gpu = tmp_path / "gpu"
    os.makedirs(gpu)
    shutil.copytree(cpu / "chunks", gpu / "chunks")
    pc.do_indexing(gpu, "poisson_average", "in", 400_000, encoding="BROTLI")
    assert all(_files_equal(cpu, gpu).values())


# ------------------------------------------------------------ decode check

def _brotli_decoder():
    lib = ctypes.CDLL("libbrotlidec.so.1")
    lib.BrotliDecoderDecompress.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t),
                                            ctypes.c_void_p]
    lib.BrotliDecoderDecompress.restype = ctypes.c_int

    def decode(data, size):
        out = np.empty(size, np.uint8)
        n = ctypes.c_size_t(size)
        src = np.frombuffer(data, np.uint8)
        ok = lib.BrotliDecoderDecompress(len(data), src.ctypes.data, ctypes.byref(n), out.ctypes.data)
        assert ok == 1 and n.value == size
        return out
    return decode


def _spread21(v):
    v = v.astype(np.uint64)
    r = np.zeros_like(v)
    for i in range(21):
        r |= ((v >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i)
    return r


def _morton(x, y, z):
    return _spread21(x) | (_spread21(y) << np.uint64(1)) | (_spread21(z) << np.uint64(2))


def test_brotli_nodes_decode_to_the_default_records(tmp_path):
    """5M points: BROTLI output against the same points' DEFAULT output."""
    import json
    las = tmp_path / "in.las"
    make_las(las, 2, 5_000_000, 21, True)
    plain, enc = tmp_path / "plain", tmp_path / "enc"
    pc.convert(las, plain, "poisson_average")
    pc.convert(las, enc, "poisson_average", encoding="BROTLI")
    meta = json.loads((enc / "metadata.json").read_text())
    attrs = meta["attributes"]
    bpp = sum(a["size"] for a in attrs)
    cols = [(a["name"], a["size"]) for a in attrs]
    soa = sum(16 if nm == "position" else 8 if nm == "rgb" else sz for nm, sz in cols)
    a_nodes = canonical_nodes(str(plain))
    b_nodes = canonical_nodes(str(enc))
    assert a_nodes.keys() == b_nodes.keys()
    decode = _brotli_decoder()
    checked = 0
    for name, (typ, mask, npts, size, data) in a_nodes.items():
        btyp, bmask, bnpts, bsize, bdata = b_nodes[name]
        assert (typ, mask, npts) == (btyp, bmask, bnpts), name
        if npts == 0:
            continue
        rec = np.frombuffer(data, np.uint8).reshape(npts, bpp)
        arr = decode(bdata, npts * soa)
        xyz = rec[:, :12].copy().view(np.int32).reshape(npts, 3).astype(np.int64)
        # the minimum starts at int64 max converted to int32, i.e. -1
        # (indexer.cpp:1164-1167), so it never exceeds -1
        m = (xyz - np.minimum(xyz.min(axis=0), -1)).astype(np.uint64)
        lo = _morton(m[:, 0] & np.uint64(0xFFFF), m[:, 1] & np.uint64(0xFFFF), m[:, 2] & np.uint64(0xFFFF))
        up = _morton(m[:, 0] >> np.uint64(16), m[:, 1] >> np.uint64(16), m[:, 2] >> np.uint64(16))
        order = np.lexsort((lo, up))  # a stable order; equal codes are compared as sets below
        pos = arr[:16 * npts].view(np.uint64).reshape(npts, 2)
        assert np.array_equal(pos[:, 0], up[order]) and np.array_equal(pos[:, 1], lo[order]), name
        # every column holds the records' values in the layout's point order:
        # rows of (codes, column bytes) agree as multisets
        rows_enc, rows_ref = [pos.view(np.uint8).reshape(npts, 16)], [np.stack([up, lo], 1).view(np.uint8)]
        o_soa, o_rec = 16 * npts, 0
        for nm, sz in cols:
            if nm == "position":
                o_rec += sz
                continue
            w = 8 if nm == "rgb" else sz
            col = arr[o_soa:o_soa + w * npts].reshape(npts, w)
            if nm == "rgb":
                c = rec[:, o_rec:o_rec + 6].copy().view(np.uint16).reshape(npts, 3).astype(np.uint64)
                ref = _morton(c[:, 0], c[:, 1], c[:, 2]).view(np.uint8).reshape(npts, 8)
            else:
                ref = rec[:, o_rec:o_rec + sz]
            rows_enc.append(col)
            rows_ref.append(ref)
            o_soa += w * npts
            o_rec += sz
        e = np.ascontiguousarray(np.concatenate(rows_enc, 1))
        r = np.ascontiguousarray(np.concatenate(rows_ref, 1))
        e = e.view(np.dtype((np.void, e.shape[1]))).ravel()
        r = r.view(np.dtype((np.void, r.shape[1]))).ravel()
        assert np.array_equal(np.sort(e), np.sort(r)), name
        checked += 1
    assert checked > 100
