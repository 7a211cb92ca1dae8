"""Launch shapes and the autotuner: every value of the sampling launch
options, and whatever the autotuner picks, gives the oracle's bytes."""
import ctypes
import filecmp
import os

import pytest

import potreeconverter_amd as pc
from helpers import make_las, oracle_convert

pytestmark = pytest.mark.gpu

FILES = ["octree.bin", "hierarchy.bin", "metadata.json"]
TUNED = ["sample_groups", "sample_order", "sample_graph", "set_policy", "ring_min_nodes", "ring_blocks_per_cu", "copy_blocks",
         "neighbor_blocks", "avg_blocks", "bucket_min_nodes", "accept_pipe_nodes", "ring_max_cand"]


def same(a, b):
    return {n: filecmp.cmp(os.path.join(a, n), os.path.join(b, n), shallow=False) for n in FILES}


@pytest.fixture
def keep_options():
    saved = {k: pc.get_option(k) for k in TUNED}
    yield
    for k, v in saved.items():
        pc.set_option(k, v)


SHAPES = [
    {"ring_blocks_per_cu": 1, "copy_blocks": 256, "neighbor_blocks": 256, "avg_blocks": 64, "bucket_min_nodes": 1,
     "accept_pipe_nodes": 0},
    {"ring_min_nodes": 1, "sample_order": 1, "copy_blocks": 65536, "neighbor_blocks": 16384, "avg_blocks": 8192,
     "accept_pipe_nodes": 1 << 30, "ring_max_cand": 4096},
    {"sample_groups": 1, "sample_order": 1, "ring_min_nodes": 64, "ring_blocks_per_cu": 2,
     "bucket_min_nodes": 1 << 30},
]


@pytest.mark.parametrize("method", ["poisson", "poisson_average"])
def test_launch_shapes_agree(method, tmp_path, keep_options):
    las = tmp_path / "in.las"
    make_las(las, 0, 2_000_000, 7, False)
    cpu = tmp_path / "cpu"
    oracle_convert(las, cpu, method)
    for i, shape in enumerate(SHAPES):
        print(f"shape case {i} {method} {shape}", flush=True)
        for k, v in shape.items():
            pc.set_option(k, v)
        out = tmp_path / f"s{i}"
        pc.convert(las, out, method)
        res = same(cpu, out)
        assert all(res.values()), (shape, res)
        for k in shape:
            pc.set_option(k, {"ring_blocks_per_cu": 3, "copy_blocks": 8192, "neighbor_blocks": 4096,
                              "avg_blocks": 2048, "ring_min_nodes": 512, "sample_order": 0,
                              "sample_groups": 0, "bucket_min_nodes": 512, "accept_pipe_nodes": 1024,
                              "ring_max_cand": 1 << 30}[k])


def test_autotune_keeps_bytes(tmp_path, keep_options):
    """The autotuner on 4M points in HBM: it returns a value for every option
    it searches, is never slower than the options it started from by more
    than noise, and the conversion with its choice writes the same files."""
    L = pc.lib()
    n = 4_000_000
    g = pc.LasGen(1, n, seed=5)
    d = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(d), n * 26 + 64) == 0
    q = (ctypes.c_int32 * 6)(2**31 - 1, 2**31 - 1, 2**31 - 1, -2**31, -2**31, -2**31)
    dq = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(dq), 64) == 0
    try:
        assert L.pcg_copy(dq, q, 24, 0) == 0
        assert L.pcg_lasgen_device(g.buf, 0, n, d, dq) == 0
        assert L.pcg_copy(q, dq, 24, 1) == 0
        hdr = ctypes.create_string_buffer(375)
        assert L.pcg_lasgen_header(g.buf, q, hdr) == 0
        before = tmp_path / "before"
        print("autotune: default conversion", flush=True)
        assert L.pcg_convert_device(d, n, hdr, b"poisson", str(before).encode(), None) == 0, L.pcg_last_error()
        print("autotune: search", flush=True)
        chosen, best, default = pc.autotune_device(d.value, n, hdr, "poisson", passes=1, repeats=1)
        assert sorted(chosen) == sorted(TUNED)
        for k, v in chosen.items():
            assert pc.get_option(k) == v
        assert 0 < best <= default
        after = tmp_path / "after"
        print(f"autotune: conversion with {chosen}", flush=True)
        assert L.pcg_convert_device(d, n, hdr, b"poisson", str(after).encode(), None) == 0, L.pcg_last_error()
        res = same(before, after)
        assert all(res.values()), (chosen, res)
    finally:
        L.pcg_device_free(d)
        L.pcg_device_free(dq)


@pytest.mark.parametrize("method", ["poisson", "poisson_average"])
def test_pipelined_acceptance_forms_agree(method, tmp_path, keep_options):
    """Every cell-path node through each form of the acceptance: the
    pipelined one with 16-bit positions (by size, the default), with 32-bit
    positions (pipe_form 1), and one wavefront per node with the accepted
    bits in global memory (pipe_form 2), and k_accept on wide launches
    (accept_pipe_nodes 0); all levels on the cell-table path."""
    las = tmp_path / "in.las"
    make_las(las, 0, 2_000_000, 7, False)
    cpu = tmp_path / "cpu"
    oracle_convert(las, cpu, method)
    saved_form = pc.get_option("pipe_form")
    try:
        pc.set_option("ring_min_nodes", 1 << 20)
        for form, pipe_nodes in ((0, 1 << 30), (1, 1 << 30), (2, 1 << 30), (0, 0)):
            print(f"acceptance form {form} accept_pipe_nodes={pipe_nodes} {method}", flush=True)
            pc.set_option("pipe_form", form)
            pc.set_option("accept_pipe_nodes", pipe_nodes)
            out = tmp_path / f"f{form}_{pipe_nodes}"
            pc.convert(las, out, method)
            res = same(cpu, out)
            assert all(res.values()), (form, pipe_nodes, res)
    finally:
        pc.set_option("pipe_form", saved_form)
