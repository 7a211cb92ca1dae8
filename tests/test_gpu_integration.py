"""Callers outside Python, on the GPU.

- tests/c_api/convert.c: a C99 program against include/potree_gpu.h
  converts a generated LAS tile; its output must equal the oracle's.
- oracle/_ref/ref_indexer_gpu: the reference indexer (Converter/src/
  indexer.cpp, built from the reference sources) with the GPU samplers of
  INTEGRATION.md (PotreeGpuSampler.h) plugged in through the Sampler
  interface; on the fixtures' inputs it must write what the reference
  indexer with its own samplers wrote.
"""
import filecmp
import json
import os
import subprocess

import pytest

import potreeconverter_amd as pc
from helpers import digest_output, make_las, oracle_convert

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_GPU = os.path.join(ROOT, "oracle", "_ref", "ref_indexer_gpu")
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("method", ["poisson", "poisson_average"])
def test_c_program_converts_like_the_oracle(method, tmp_path):
    lib = os.path.join(ROOT, "potreeconverter_amd")
    exe = tmp_path / "convert"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), os.path.join(ROOT, "tests", "c_api", "convert.c"),
                    f"-I{ROOT}/include", f"-L{lib}", "-lpotree_gpu", f"-Wl,-rpath,{lib}"], check=True)
    print(f"c program {method}", flush=True)
    r = subprocess.run([str(exe), str(tmp_path), "300000", "4", method], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "missing input:" in r.stdout
    cpu = tmp_path / "cpu"
    oracle_convert(tmp_path / "in.las", cpu, method)
    for f in ("octree.bin", "hierarchy.bin", "metadata.json"):
        assert filecmp.cmp(cpu / f, tmp_path / "out" / f, shallow=False), f
        assert filecmp.cmp(cpu / f, tmp_path / "split" / f, shallow=False), f


CASES = ["c0_150k_poisson", "c0_400k_extra_poisson_average", "c0_150k_random", "c2_64chunks_poisson",
         "c2_64chunks_poisson_average", "c2_512chunks_poisson_average", "scene_dups_poisson"]


@pytest.mark.skipif(not os.path.exists(REF_GPU), reason="oracle/_ref/ref_indexer_gpu not built (oracle/build_ref.sh)")
@pytest.mark.parametrize("name", CASES)
def test_reference_indexer_with_gpu_samplers(name, tmp_path):
    fx = json.load(open(os.path.join(GOLDEN, name + ".json")))
    case = fx["case"]
    print(f"reference indexer + GPU samplers: {name}", flush=True)
    work = tmp_path / "work"
    if "chunkgen" in case:
        pc.ChunkGen(**case["chunkgen"]).write(work)
    else:
        las = tmp_path / "in.las"
        if "scene" in case:
            from scenes import make_scene
            make_scene(case["scene"], las)
        else:
            make_las(las, case["preset"], case["points"], case["seed"], case["extra_bytes"])
        oracle_convert(las, work, case["method"], keep_chunks=True, chunk_only=True)
    env = dict(os.environ, REF_CLOCK_NS=str(case.get("random_seed", 0)))
    r = subprocess.run([REF_GPU, str(work), case["method"], name, str(case["points"]), "DEFAULT"],
                       capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-2000:]
    with open(work / "metadata.json") as f:
        assert f.read() == fx["metadata_json"]
    got = digest_output(str(work))
    assert set(got) == set(fx["nodes"])
    bad = [n for n in got if got[n] != fx["nodes"][n]]
    assert not bad, f"{len(bad)} nodes differ, e.g. {bad[:5]}"
