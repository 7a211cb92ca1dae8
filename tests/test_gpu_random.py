"""-m random (SamplerRandom, Converter/include/sampler_random.h) on the gfx950
path against the CPU oracle, byte for byte, through the C ABI.

The reference seeds every leaf shuffle from the clock; both sides here run
the seeded variant (the library's "random_seed" option, the oracle's
orc_set_random_seed), which the oracle fixtures pin to the reference indexer
with its clock fixed at the seed (tests/golden/*random*.json).
"""
import os
import shutil

import pytest

import potreeconverter_amd as pc
from helpers import make_las, oracle_convert

pytestmark = pytest.mark.gpu

FILES = ["octree.bin", "hierarchy.bin", "metadata.json"]


def _set(key, value):
    rc = pc.lib().pcg_set_option(key.encode(), value)
    assert rc == 0, pc.lib().pcg_last_error()


def _files_equal(a, b):
    out = {}
    for n in FILES:
        with open(os.path.join(a, n), "rb") as fa, open(os.path.join(b, n), "rb") as fb:
            out[n] = fa.read() == fb.read()
    return out


CASES = [
    # preset, points, seed, extra bytes, point format, random seed
    (0, 150_000, 2, False, 2, 7),
    (0, 2_000_000, 0, False, 2, 1),
    (1, 1_000_000, 1, False, 2, 4_000_000_000),
    (2, 1_000_000, 3, False, 2, 123_456_789),   # near-duplicates, lattice on cell borders
    (2, 300_000, 11, True, 2, 0),
    (0, 400_000, 24, True, 6, 99),
]


@pytest.mark.parametrize("preset,n,seed,eb,fmt,rs", CASES,
                         ids=[f"p{c[0]}_{c[1]}_s{c[2]}_pdrf{c[4]}_r{c[5]}{'_eb' if c[3] else ''}" for c in CASES])
def test_random_convert_matches_oracle(preset, n, seed, eb, fmt, rs, tmp_path):
    print(f"random case preset={preset} n={n} seed={seed} extra={eb} pdrf={fmt} random_seed={rs}", flush=True)
    las = tmp_path / "in.las"
    make_las(las, preset, n, seed, eb, fmt)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert(las, cpu, "random", random_seed=rs)
    try:
        _set("random_seed", rs)
        pc.convert(las, gpu, "random")
    finally:
        _set("random_seed", -1)
    res = _files_equal(cpu, gpu)
    assert all(res.values()), res


def test_random_groups_plain_and_brotli(tmp_path):
    """Sampling groups, per-node launches and BROTLI output on -m random."""
    las = tmp_path / "in.las"
    make_las(las, 0, 2_000_000, 7, False)
    cpu, cpub = tmp_path / "cpu", tmp_path / "cpub"
    oracle_convert(las, cpu, "random", random_seed=5)
    oracle_convert(las, cpub, "random", encoding="BROTLI", random_seed=5)
    try:
        _set("random_seed", 5)
        for g in (1, 3):
            print(f"random groups case {g}", flush=True)
            _set("sample_groups", g)
            out = tmp_path / f"g{g}"
            pc.convert(las, out, "random")
            assert all(_files_equal(cpu, out).values()), g
        _set("sample_groups", 0)
        print("random plain case", flush=True)
        _set("plain_sampling", 1)
        out = tmp_path / "plain"
        pc.convert(las, out, "random")
        assert all(_files_equal(cpu, out).values())
        _set("plain_sampling", 0)
        print("random brotli case", flush=True)
        out = tmp_path / "br"
        pc.convert(las, out, "random", encoding="BROTLI")
        assert all(_files_equal(cpub, out).values())
    finally:
        _set("random_seed", -1)
        _set("sample_groups", 0)
        _set("plain_sampling", 0)


def test_random_indexing_of_reference_format_chunks(tmp_path):
    las = tmp_path / "in.las"
    make_las(las, 1, 600_000, 31)
    cpu = tmp_path / "cpu"
    oracle_convert(las, cpu, "random", keep_chunks=True, random_seed=31)
    gpu = tmp_path / "gpu"
    os.makedirs(gpu)
    shutil.copytree(cpu / "chunks", gpu / "chunks")
    try:
        _set("random_seed", 31)
        pc.do_indexing(gpu, "random", "in", 600_000)
    finally:
        _set("random_seed", -1)
    assert all(_files_equal(cpu, gpu).values())


def test_random_without_seed_conserves_points(tmp_path):
    """The reference's own behaviour (clock seed): every input point is
    written exactly once, whatever the seed."""
    import numpy as np
    from potree_format import canonical_nodes
    las = tmp_path / "in.las"
    g = make_las(las, 0, 300_000, 3)
    out = tmp_path / "out"
    pc.convert(las, out, "random")
    total = sum(v[2] for v in canonical_nodes(str(out)).values())
    assert total == 300_000
