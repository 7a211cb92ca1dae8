"""pcg_sample_subtree (one Sampler::sample call on the GPU) against the
reference samplers.

The fixtures subtree_*.json hold what the reference SamplerPoisson,
SamplerPoissonAverage and SamplerRandom leave after one sample() call on the
subtrees of tests/subtrees.py (oracle/_ref/ref_sampler, written by
tests/golden/make_golden.py): every node's point count, its linked children,
the sha256 of its records, and the onNodeCompleted / onNodeDiscarded calls
in order.
"""
import glob
import hashlib
import json
import os

import pytest

import potreeconverter_amd as pc
from subtrees import BASE_SPACING, BPP, RGB, SCALE, make_subtree

pytestmark = pytest.mark.gpu

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "subtree_*.json")))


@pytest.mark.parametrize("fixture", FIXTURES, ids=[os.path.basename(p)[:-5] for p in FIXTURES])
def test_subtree_sampling_matches_reference(fixture):
    fx = json.load(open(fixture))
    case = fx["case"]
    print(f"subtree case {case}", flush=True)
    t = make_subtree(case["subtree"])
    pc._check(pc.lib().pcg_set_option(b"random_seed", case["random_seed"]))
    try:
        counts, recs, children, events = pc.sample_subtree(
            case["method"], t.children, t.bounds, t.levels, t.counts, b"".join(t.records), BPP, RGB,
            [SCALE] * 3, [0.0] * 3, BASE_SPACING, sampled=t.sampled)
    finally:
        pc.lib().pcg_set_option(b"random_seed", -1)
    assert [list(e) for e in events] == fx["events"]
    completed = {e[0] for e in fx["events"] if e[1] == 0}
    linked = {c for row in fx["nodes"] for c in row[1] if c >= 0} | {0}
    for i, (cnt, ch, digest) in enumerate(fx["nodes"]):
        assert children[i] == ch, i
        assert counts[i] == cnt, i
        # a node unlinked without a callback (Poisson-average) is dropped:
        # its records are unspecified
        if i in completed or i in linked:
            assert hashlib.sha256(recs[i]).hexdigest() == digest, i


def test_subtree_of_one_leaf():
    """sample() on a leaf changes nothing (Poisson) and calls nothing."""
    t = make_subtree("sparse")
    leaf = 1
    counts, recs, children, events = pc.sample_subtree(
        "poisson", [[-1] * 8], [t.bounds[leaf]], [t.levels[leaf]], [t.counts[leaf]], t.records[leaf], BPP, RGB,
        [SCALE] * 3, [0.0] * 3, BASE_SPACING)
    assert counts == [t.counts[leaf]] and recs[0] == t.records[leaf] and events == []


def test_subtree_rejects_malformed_trees():
    t = make_subtree("sparse")
    bad = [row[:] for row in t.children]
    bad[1][0] = 2  # node 2 gets two parents
    with pytest.raises(pc.PotreeGpuError, match="bad child"):
        pc.sample_subtree("poisson", bad, t.bounds, t.levels, t.counts, b"".join(t.records), BPP, RGB,
                          [SCALE] * 3, [0.0] * 3, BASE_SPACING)
    counts = t.counts[:]
    counts[0] = 5  # records at a node with children
    with pytest.raises(pc.PotreeGpuError, match="only nodes without children"):
        pc.sample_subtree("poisson", t.children, t.bounds, t.levels, counts, b"".join(t.records) + bytes(5 * BPP),
                          BPP, RGB, [SCALE] * 3, [0.0] * 3, BASE_SPACING)
