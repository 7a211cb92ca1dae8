"""Graph capture of the chunk-subtree sampling ("sample_graph"): jobs run
with the captured graph, eager jobs between them, and every graph job writes
the eager job's bytes."""
import filecmp
import os

import pytest

import potreeconverter_amd as pc
from helpers import make_las

pytestmark = pytest.mark.gpu

FILES = ["octree.bin", "hierarchy.bin", "metadata.json"]


def same(a, b):
    return {n: filecmp.cmp(os.path.join(a, n), os.path.join(b, n), shallow=False) for n in FILES}


@pytest.mark.parametrize("method", ["poisson", "poisson_average", "random"])
def test_graph_jobs_match_eager(method, tmp_path):
    inputs = []
    for i, (n, seed) in enumerate(((2_000_000, 7), (1_200_000, 3))):
        las = tmp_path / f"in{i}.las"
        make_las(las, 0, n, seed, False)
        inputs.append(las)
    old_seed = pc.get_option("random_seed")
    pc.set_option("random_seed", 11)
    try:
        eager = []
        for i, las in enumerate(inputs):
            out = tmp_path / f"eager{i}"
            pc.convert(las, out, method)
            eager.append(out)
        # graph jobs, each input three times, with eager jobs in between
        for rep in range(3):
            for i, las in enumerate(inputs):
                print(f"graph case {method} input {i} replay {rep}", flush=True)
                pc.set_option("sample_graph", 1)
                out = tmp_path / f"graph{i}_{rep}"
                pc.convert(las, out, method)
                pc.set_option("sample_graph", 0)
                res = same(eager[i], out)
                assert all(res.values()), (i, rep, res)
                between = tmp_path / f"between{i}_{rep}"
                pc.convert(inputs[1 - i], between, method)
                res = same(eager[1 - i], between)
                assert all(res.values()), (i, rep, res)
    finally:
        pc.set_option("sample_graph", 0)
        pc.set_option("random_seed", old_seed)
