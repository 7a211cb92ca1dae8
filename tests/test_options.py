"""Library options through the C ABI (no GPU needed): set, read back, and
the range checks of the launch-shape options."""
import pytest

import potreeconverter_amd as pc


def test_set_get_roundtrip():
    for key, value in (("ring_blocks_per_cu", 2), ("copy_blocks", 4096), ("neighbor_blocks", 2048),
                       ("avg_blocks", 1024), ("sample_order", 1), ("ring_min_nodes", 300),
                       ("accept_pipe_nodes", 0)):
        old = pc.get_option(key)
        try:
            pc.set_option(key, value)
            assert pc.get_option(key) == value
        finally:
            pc.set_option(key, old)


@pytest.mark.parametrize("key,value", [("ring_blocks_per_cu", 0), ("ring_blocks_per_cu", 4), ("copy_blocks", 1),
                                       ("neighbor_blocks", 1 << 21), ("avg_blocks", 0), ("sample_order", 2),
                                       ("accept_pipe_nodes", -1)])
def test_out_of_range_is_refused(key, value):
    old = pc.get_option(key)
    with pytest.raises(pc.PotreeGpuError):
        pc.set_option(key, value)
    assert pc.get_option(key) == old


def test_unknown_option():
    with pytest.raises(pc.PotreeGpuError):
        pc.get_option("no_such_option")


def test_autotune_parameter_list():
    L = pc.lib()
    names = [L.pcg_autotune_param_name(i).decode() for i in range(L.pcg_autotune_params())]
    assert names == ["sample_groups", "sample_order", "sample_graph", "set_policy", "ring_min_nodes", "ring_blocks_per_cu", "copy_blocks",
                     "neighbor_blocks", "avg_blocks", "bucket_min_nodes", "accept_pipe_nodes", "ring_max_cand"]
    assert L.pcg_autotune_param_name(len(names)) is None
    for n in names:
        pc.get_option(n)
