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


def declarations():
    """name -> (return type, [parameter types]) of every pcg_ function in the header."""
    text = open(os.path.join(ROOT, "include", "potree_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for ret, name, params in re.findall(r"([A-Za-z_][A-Za-z0-9_\s\*]*?)\b(pcg_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        params = params.strip()
        plist = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
        types = []
        for p in plist:
            pointer = "*" in p or "[" in p
            base = re.sub(r"\bconst\b", "", p.split("[")[0]).replace("*", " ").split()[0]
            types.append("pointer" if pointer else base)
        out[name] = (ret.strip().splitlines()[-1].strip(), types)
    return out


SIZES = {"int32_t": 4, "uint32_t": 4, "int64_t": 8, "uint64_t": 8, "int": 4, "double": 8}


def test_every_binding_matches_the_header():
    """Each ctypes binding in potreeconverter_amd declares the header's
    parameter list: a missing argtypes would pass a 64-bit size as a 32-bit
    int (a 5.2 GB allocation request truncated to under 1 GB)."""
    import potreeconverter_amd as pc

    L = pc.lib()
    decls = declarations()
    assert len(decls) >= 20
    bad = []
    for name, (ret, params) in decls.items():
        fn = getattr(L, name)
        if fn.argtypes is None:
            bad.append(f"{name}: no argtypes")
            continue
        if len(fn.argtypes) != len(params):
            bad.append(f"{name}: {len(fn.argtypes)} argtypes for {len(params)} parameters")
            continue
        for i, (want, got) in enumerate(zip(params, fn.argtypes)):
            if want == "pointer":
                ok = got in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(got, ctypes._Pointer)
            else:
                ok = ctypes.sizeof(got) == SIZES[want] and (want != "double") == (got is not ctypes.c_double)
            if not ok:
                bad.append(f"{name}: parameter {i} is {want}, bound as {got.__name__}")
        want_ret = {"int": ctypes.c_int, "int32_t": ctypes.c_int32, "uint64_t": ctypes.c_uint64, "void": None,
                    "const char*": ctypes.c_char_p}[ret]
        if fn.restype is not want_ret and not (want_ret is not None and fn.restype is not None
                                                and ctypes.sizeof(fn.restype) == ctypes.sizeof(want_ret)):
            bad.append(f"{name}: returns {ret}, bound as {fn.restype}")
    assert not bad, bad
