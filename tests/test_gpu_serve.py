"""The serving loop (potreeconverter_amd.serve) on the GPU: every job of a
short Poisson stream is admitted, converted and accounted for."""
import pytest

pytestmark = pytest.mark.gpu


def test_serving_loop_converts_every_job():
    from potreeconverter_amd import serve
    print("serving loop: 12 jobs", flush=True)
    r = serve.serve(num_jobs=12, rate=40.0, min_points=100_000, max_points=1_000_000, seed=5, method="poisson")
    assert r["jobs"] == 12 and r["points"] > 12 * 100_000
    assert r["octree_bytes_conserved"] is True
    assert 0 < r["latency_p50_s"] <= r["latency_p99_s"] <= r["latency_max_s"]
    assert r["convert_p50_s"] <= r["latency_p50_s"] + 1e-9 or r["convert_p50_s"] <= r["latency_p99_s"]


def test_admission_waits_under_a_small_budget():
    """With a budget that holds one job's working memory, jobs wait for
    admission instead of running out of memory."""
    from potreeconverter_amd import serve
    print("serving loop: small budget", flush=True)
    # one job's working memory plus little more: a burst of arrivals must
    # queue for admission
    budget = 1_600_000 * serve.WORK_BYTES_PER_POINT
    r = serve.serve(num_jobs=8, rate=2000.0, min_points=1_000_000, max_points=1_500_000, seed=6, method="poisson",
                    hbm_budget=budget)
    assert r["admission_waits"] > 0 and r["octree_bytes_conserved"] is True
