"""The subtree fixtures against the live reference samplers.

tests/golden/subtree_*.json were written by oracle/_ref/ref_sampler (the
reference SamplerPoisson / SamplerPoissonAverage / SamplerRandom behind a
subtree driver) on the subtrees of tests/subtrees.py.  Where the reference
build is present, the fixtures must still be what it writes for the
subtrees as generated today (the GPU tests regenerate the same inputs).
"""
import glob
import hashlib
import json
import os
import subprocess

import pytest

from helpers import REF_SAMPLER
from subtrees import make_subtree, parse_reference_output

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "subtree_*.json")))


@pytest.mark.skipif(not os.path.exists(REF_SAMPLER), reason="reference samplers not built (oracle/build_ref.sh)")
@pytest.mark.parametrize("fixture", FIXTURES, ids=[os.path.basename(p)[:-5] for p in FIXTURES])
def test_fixture_is_what_the_reference_writes(fixture, tmp_path):
    fx = json.load(open(fixture))
    case = fx["case"]
    t = make_subtree(case["subtree"])
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    t.input_file(src)
    subprocess.run([REF_SAMPLER, case["method"], str(src), str(dst)], check=True,
                   env=dict(os.environ, REF_CLOCK_NS=str(case["random_seed"])))
    nodes, events = parse_reference_output(dst, len(t.levels))
    assert events == fx["events"]
    assert [[c, ch, hashlib.sha256(r).hexdigest()] for c, ch, sm, r in nodes] == fx["nodes"]
    assert all(sm == 1 for c, ch, sm, r in nodes)


def test_subtrees_are_well_formed():
    for name in ("chunk", "upper", "sparse"):
        t = make_subtree(name)
        parents = {}
        for i, row in enumerate(t.children):
            for c in row:
                if c >= 0:
                    assert c not in parents and c > i and t.levels[c] == t.levels[i] + 1
                    parents[c] = i
            assert (max(row) < 0) or t.counts[i] == 0
        assert set(parents) == set(range(1, len(t.levels)))
