"""Writes the golden fixtures of the indexer path from the reference itself.

For each case: generate the seeded LAS file, chunk it with the CPU oracle
(the reference chunker cannot be built here: laszip needs a CMake-generated
header), index those chunks with the REFERENCE indexer built from the
reference sources (oracle/build_ref.sh -> oracle/_ref/ref_indexer), and store
its metadata.json plus a per-node digest of hierarchy.bin / octree.bin
(type, childMask, numPoints, byteSize, sha256 of the node's records).
The reference writes nodes in thread-completion order, so the digest is
keyed by node name rather than by file offset.

SamplerRandom seeds its leaf shuffles from the clock; the driver fixes the
clock at the case's seed (REF_CLOCK_NS), the seeded variant the oracle and
the GPU path implement.

Edge-case scenes (tests/scenes.py: long equal-distance runs, duplicate
stacks, exact-spacing lattices, the 10 001-check limit) are written the same
way from their scene files.

Pre-chunked sets of the batching workload (potreeconverter_amd.ChunkGen:
level-4 chunks "rABCD" with log-normal point counts and three extra-bytes
attributes) are written as a chunks/ directory by the library's host
generator and indexed by the reference indexer directly.

Subtree cases (tests/subtrees.py) pin one Sampler::sample call: the
reference samplers run on a flattened subtree by oracle/_ref/ref_sampler,
which records the callbacks.

Run from the repository root:  python tests/golden/make_golden.py [case names]
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from helpers import REF_INDEXER, REF_SAMPLER, digest_output, make_las, oracle_convert  # noqa: E402
from scenes import make_scene  # noqa: E402

CASES = [
    # name, preset, points, seed, extra bytes, method
    ("c0_2m_poisson", 0, 2_000_000, 0, False, "poisson"),
    ("c0_2m_poisson_average", 0, 2_000_000, 0, False, "poisson_average"),
    ("c0_400k_extra_poisson_average", 0, 400_000, 5, True, "poisson_average"),
    ("c1_1m_utm_poisson", 1, 1_000_000, 1, False, "poisson"),
    ("c0_150k_poisson", 0, 150_000, 2, False, "poisson"),
    ("c3_1m_precision_poisson", 2, 1_000_000, 3, False, "poisson"),
    ("c3_1m_precision_poisson_average", 2, 1_000_000, 3, False, "poisson_average"),
    ("c3_300k_precision_extra_poisson", 2, 300_000, 11, True, "poisson"),
    # BROTLI node encoding (--encoding BROTLI)
    ("c0_150k_poisson_brotli", 0, 150_000, 2, False, "poisson", "BROTLI"),
    ("c0_400k_extra_poisson_average_brotli", 0, 400_000, 5, True, "poisson_average", "BROTLI"),
    ("c3_300k_precision_extra_poisson_brotli", 2, 300_000, 11, True, "poisson", "BROTLI"),
    # SamplerRandom with the clock fixed at the seed (ref_driver.cpp)
    ("c0_150k_random", 0, 150_000, 2, False, "random", "DEFAULT", 7),
    ("c3_300k_precision_extra_random", 2, 300_000, 11, True, "random", "DEFAULT", 123456789),
    ("c1_1m_utm_random_brotli", 1, 1_000_000, 1, False, "random", "BROTLI", 4000000000),
]

# edge-case scenes of tests/scenes.py: name, scene, method
SCENE_CASES = [
    ("scene_stacks_poisson", "stacks", "poisson"),
    ("scene_stacks_poisson_average", "stacks", "poisson_average"),
    ("scene_dups_poisson", "dups", "poisson"),
    ("scene_dups_poisson_average", "dups", "poisson_average"),
    ("scene_spacing_poisson", "spacing", "poisson"),
    ("scene_spacing_poisson_average", "spacing", "poisson_average"),
    ("scene_sphere_poisson", "sphere", "poisson"),
    ("scene_sphere_poisson_average", "sphere", "poisson_average"),
]

# pre-chunked sets: name, ChunkGen arguments, method
CHUNKGEN_CASES = [
    ("c2_64chunks_poisson_average", dict(num_chunks=64, median=4000, sigma=1.16, min_count=100, max_count=100_000,
                                         seed=3), "poisson_average"),
    ("c2_64chunks_poisson", dict(num_chunks=64, median=4000, sigma=1.16, min_count=100, max_count=100_000,
                                 seed=3), "poisson"),
    ("c2_512chunks_poisson_average", dict(num_chunks=512, median=2000, sigma=1.16, min_count=100,
                                          max_count=200_000, seed=7), "poisson_average"),
]

# one Sampler::sample call on a subtree of tests/subtrees.py: name, subtree,
# method, clock of the random leaf shuffles
SUBTREE_CASES = [
    ("subtree_chunk_poisson", "chunk", "poisson", 0),
    ("subtree_chunk_poisson_average", "chunk", "poisson_average", 0),
    ("subtree_chunk_random", "chunk", "random", 4242),
    ("subtree_upper_poisson", "upper", "poisson", 0),
    ("subtree_upper_poisson_average", "upper", "poisson_average", 0),
    ("subtree_upper_random", "upper", "random", 77),
    ("subtree_sparse_poisson", "sparse", "poisson", 0),
    ("subtree_sparse_poisson_average", "sparse", "poisson_average", 0),
]


def write_subtree_fixture(td, name, subtree, method, clock):
    import hashlib
    from subtrees import make_subtree, parse_reference_output
    t = make_subtree(subtree)
    src, dst = os.path.join(td, name + ".in"), os.path.join(td, name + ".out")
    t.input_file(src)
    subprocess.run([REF_SAMPLER, method, src, dst], check=True, env=dict(os.environ, REF_CLOCK_NS=str(clock)))
    nodes, events = parse_reference_output(dst, len(t.levels))
    fixture = {
        "case": {"subtree": subtree, "method": method, "random_seed": clock},
        "producer": "reference samplers (Converter/include/sampler_*.h), one Sampler::sample call",
        "nodes": [[c, ch, hashlib.sha256(r).hexdigest()] for c, ch, sm, r in nodes],
        "events": events,
    }
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(fixture, f, indent=0, sort_keys=True)
    print(name, len(nodes), "nodes", len(events), "events", sum(e[1] for e in events), "discarded")


def write_fixture(td, name, case, las, n, method, encoding="DEFAULT", random_seed=None, chunkgen=None):
    work = os.path.join(td, name)
    if chunkgen is not None:
        chunkgen.write(work)
    else:
        oracle_convert(las, work, method, keep_chunks=True, chunk_only=True)
    env = dict(os.environ, REF_CLOCK_NS=str(random_seed or 0))
    subprocess.run([REF_INDEXER, work, method, name, str(n), encoding], check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, env=env)
    with open(os.path.join(work, "metadata.json")) as f:
        meta = f.read()
    fixture = {
        "case": case,
        "producer": "reference indexer (Converter/src/indexer.cpp) on oracle-chunked input",
        "metadata_json": meta,
        "nodes": digest_output(work),
    }
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(fixture, f, indent=0, sort_keys=True)
    print(name, len(fixture["nodes"]), "nodes")


def main():
    assert os.path.exists(REF_INDEXER), "build the reference indexer first: bash oracle/build_ref.sh"
    td = tempfile.mkdtemp()
    try:
        only = set(sys.argv[1:])
        for name, preset, n, seed, eb, method, *rest in CASES:
            if only and name not in only:
                continue
            encoding = rest[0] if rest else "DEFAULT"
            random_seed = rest[1] if len(rest) > 1 else None
            las = os.path.join(td, name + ".las")
            make_las(las, preset, n, seed, eb)
            case = {"preset": preset, "points": n, "seed": seed, "extra_bytes": eb, "method": method,
                    "encoding": encoding, **({"random_seed": random_seed} if random_seed is not None else {})}
            write_fixture(td, name, case, las, n, method, encoding, random_seed)
        for name, scene, method in SCENE_CASES:
            if only and name not in only:
                continue
            las = os.path.join(td, name + ".las")
            n = make_scene(scene, las)
            case = {"scene": scene, "points": n, "method": method, "encoding": "DEFAULT"}
            write_fixture(td, name, case, las, n, method)
        from potreeconverter_amd import ChunkGen
        for name, args, method in CHUNKGEN_CASES:
            if only and name not in only:
                continue
            g = ChunkGen(**args)
            case = {"chunkgen": args, "points": g.total, "method": method, "encoding": "DEFAULT"}
            write_fixture(td, name, case, None, g.total, method, chunkgen=g)
        for name, subtree, method, clock in SUBTREE_CASES:
            if only and name not in only:
                continue
            write_subtree_fixture(td, name, subtree, method, clock)
    finally:
        shutil.rmtree(td, ignore_errors=True)


if __name__ == "__main__":
    main()
