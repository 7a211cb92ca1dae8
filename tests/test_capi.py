"""The C ABI library loads and exports every entry point of include/potree_gpu.h."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "potree_gpu.h")).read()
    return sorted(set(re.findall(r"\b(pcg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "potreeconverter_amd", "libpotree_gpu.so"))
    names = declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_package_refuses_to_run_without_library(tmp_path, monkeypatch):
    import potreeconverter_amd as pc
    monkeypatch.setattr(pc, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(pc, "_lib", None)
    try:
        pc.lib()
        raised = False
    except pc.PotreeGpuError:
        raised = True
    assert raised


def test_unknown_option_is_an_error():
    import potreeconverter_amd as pc

    L = pc.lib()
    assert L.pcg_set_option(b"no_such_option", 1) != 0
    assert b"unknown option" in L.pcg_last_error()
    assert L.pcg_set_option(b"ring_min_nodes", 0) != 0
    assert L.pcg_set_option(b"sample_groups", -1) != 0
    assert L.pcg_set_option(b"sample_groups", 9) != 0
    assert L.pcg_set_option(b"random_seed", -2) != 0
    assert L.pcg_set_option(b"random_seed", 1 << 32) != 0
    assert L.pcg_set_option(b"random_seed", (1 << 32) - 1) == 0
    assert L.pcg_set_option(b"random_seed", -1) == 0
