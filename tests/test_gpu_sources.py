"""Several sources and the --attributes filter on the GPU against the oracle.

pcg_convert_sources / pcg_do_chunking_sources (doChunking(sources, targetDir,
min, max, state, outputAttributes, monitor) with the output attributes of
computeOutputAttributes(sources, requestedAttributes)) must write the
oracle's chunk files, chunks/metadata.json, octree.bin, hierarchy.bin and
metadata.json; the oracle's multi-source rules are pinned in
tests/test_oracle_sources.py.
"""
import filecmp
import os

import numpy as np
import pytest

import potreeconverter_amd as pc
from helpers import make_las, oracle_convert_sources

pytestmark = pytest.mark.gpu


def same_tree(a, b, names):
    return {n: filecmp.cmp(os.path.join(a, n), os.path.join(b, n), shallow=False) for n in names}


CASES = [
    # sources (preset, points, seed, extra bytes, point format), attributes, method
    ([(0, 300_000, 1, False, 2), (0, 250_000, 2, False, 2), (0, 200_000, 3, False, 2)], None, "poisson"),
    ([(0, 300_000, 4, True, 2), (0, 200_000, 5, False, 3)], None, "poisson_average"),
    ([(0, 300_000, 4, True, 2), (2, 200_000, 5, False, 3)], ["rgb", "intensity", "nothing", "gps-time"], "poisson"),
    ([(1, 400_000, 6, False, 7), (1, 150_000, 7, True, 6)], ["classification", "rgb"], "poisson_average"),
]


@pytest.mark.parametrize("sources,attributes,method", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_sources_match_oracle(sources, attributes, method, tmp_path):
    print(f"sources case {sources} {attributes} {method}", flush=True)
    paths = []
    for i, (preset, n, seed, eb, fmt) in enumerate(sources):
        p = tmp_path / f"s{i}.las"
        make_las(p, preset, n, seed, eb, fmt)
        paths.append(p)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert_sources(paths, cpu, method, attributes=attributes, keep_chunks=True)
    pc.convert_sources(paths, gpu, method, attributes=attributes, keep_chunks=True)
    res = same_tree(cpu, gpu, ["octree.bin", "hierarchy.bin", "metadata.json"])
    assert all(res.values()), res
    a, b = sorted(os.listdir(cpu / "chunks")), sorted(os.listdir(gpu / "chunks"))
    assert a == b
    assert all(same_tree(cpu / "chunks", gpu / "chunks", a).values())


def test_chunking_with_the_callers_box(tmp_path):
    paths = [tmp_path / "a.las", tmp_path / "b.las"]
    make_las(paths[0], 0, 200_000, 8)
    make_las(paths[1], 0, 150_000, 9)
    mn = [-20.0, -30.0, -100.0]
    mx = [m + 700.0 for m in mn]
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert_sources(paths, cpu, "poisson", cube_min=mn, cube_max=mx, chunk_only=True)
    pc.do_chunking_sources(paths, gpu, cube_min=mn, cube_max=mx)
    a, b = sorted(os.listdir(cpu / "chunks")), sorted(os.listdir(gpu / "chunks"))
    assert a == b and all(same_tree(cpu / "chunks", gpu / "chunks", a).values())


def test_orphaned_hierarchy_chunk_fails_like_the_oracle(tmp_path):
    """Two tiles 1e6 m apart in one cube (presets 0 and 2): Poisson-average
    unlinks level-4 node r0000 without writing it while nodes below it were
    written, so hierarchy.bin has no proxy for their chunk.  The reference
    indexer crashes there (HierarchyBuilder.h:317-326 dereferences the
    missing node); the library and the oracle stop with status 123."""
    a, b = tmp_path / "a.las", tmp_path / "b.las"
    make_las(a, 0, 300_000, 4, True, 2)
    make_las(b, 2, 200_000, 5, False, 3)
    assert oracle_convert_sources([a, b], tmp_path / "cpu", "poisson_average", check=False) == 123
    with pytest.raises(pc.PotreeGpuError, match="error 123: didn't find proxy node r0000"):
        pc.convert_sources([a, b], tmp_path / "gpu", "poisson_average")


def test_moved_extra_bytes_are_refused_like_the_oracle(tmp_path):
    a, b = tmp_path / "a.las", tmp_path / "b.las"
    make_las(a, 0, 30_000, 1, False, 3)
    make_las(b, 0, 30_000, 2, True, 2)
    assert oracle_convert_sources([a, b], tmp_path / "cpu", "poisson", check=False) != 0
    with pytest.raises(pc.PotreeGpuError, match="extra attribute"):
        pc.convert_sources([a, b], tmp_path / "gpu", "poisson")
    with pytest.raises(pc.PotreeGpuError, match="extra attribute"):
        pc.do_chunking_sources([b], tmp_path / "gpu2", attributes=["gauss"])
