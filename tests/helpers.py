"""Shared helpers for the tests: oracle loading, LAS generation, digests."""
import ctypes
import hashlib
import os

from potree_format import canonical_nodes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_INDEXER = os.path.join(ROOT, "oracle", "_ref", "ref_indexer")
REF_SAMPLER = os.path.join(ROOT, "oracle", "_ref", "ref_sampler")
REF_POINT_DECODE = os.path.join(ROOT, "oracle", "_ref", "ref_point_decode")
_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        O = ctypes.CDLL(ORACLE_LIB)
        O.orc_convert.argtypes = [ctypes.c_char_p] * 3 + [ctypes.c_int32, ctypes.c_int32,
                                                          ctypes.POINTER(ctypes.c_double)]
        O.orc_index.argtypes = [ctypes.c_char_p] * 3 + [ctypes.POINTER(ctypes.c_double)]
        O.orc_last_error.restype = ctypes.c_char_p
        O.orc_decode_points.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_void_p]
        O.orc_set_random_seed.argtypes = [ctypes.c_uint32]
        O.orc_set_chunking.argtypes = [ctypes.c_int64, ctypes.c_int64]
        O.orc_convert_encoded.argtypes = [ctypes.c_char_p] * 4 + [ctypes.c_int32, ctypes.c_int32,
                                                                  ctypes.POINTER(ctypes.c_double)]
        O.orc_count_cells.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32] + \
            [ctypes.POINTER(ctypes.c_double)] * 6 + [ctypes.c_int32, ctypes.c_void_p]
        O.orc_create_lut.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_void_p]
        O.orc_create_lut.restype = ctypes.c_int64
        _oracle = O
    return _oracle


def oracle_convert(las, outdir, method, keep_chunks=False, chunk_only=False, encoding="DEFAULT", random_seed=0):
    t = (ctypes.c_double * 8)()
    oracle().orc_set_random_seed(ctypes.c_uint32(random_seed))
    rc = oracle().orc_convert_encoded(str(las).encode(), str(outdir).encode(), method.encode(), encoding.encode(),
                                      int(keep_chunks), int(chunk_only), t)
    assert rc == 0, oracle().orc_last_error()
    return list(t)


def oracle_convert_sources(paths, outdir, method, attributes=None, cube_min=None, cube_max=None, keep_chunks=False,
                           chunk_only=False, check=True):
    """The oracle's doChunking(sources, ..., min, max, ...) + doIndexing with
    the --attributes filter; returns the status (asserts 0 with check)."""
    O = oracle()
    if not hasattr(O, "_sources_bound"):
        O.orc_convert_sources.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_char_p), ctypes.c_int32, ctypes.c_char_p,
                                          ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_double)]
        O._sources_bound = True
    ps = (ctypes.c_char_p * len(paths))(*[str(p).encode() for p in paths])
    at = list(attributes or [])
    names = (ctypes.c_char_p * max(len(at), 1))(*[a.encode() for a in at])
    mn = (ctypes.c_double * 3)(*cube_min) if cube_min is not None else None
    mx = (ctypes.c_double * 3)(*cube_max) if cube_max is not None else None
    t = (ctypes.c_double * 8)()
    rc = O.orc_convert_sources(ps, len(paths), mn, mx, names if at else None, len(at), str(outdir).encode(),
                               method.encode(), int(keep_chunks), int(chunk_only), t)
    if check:
        assert rc == 0, O.orc_last_error()
    return rc


def oracle_index(outdir, method, name=""):
    """indexer::doIndexing of the oracle on outdir/chunks."""
    t = (ctypes.c_double * 8)()
    rc = oracle().orc_index(str(outdir).encode(), method.encode(), name.encode(), t)
    assert rc == 0, oracle().orc_last_error()
    return list(t)


def make_las(path, preset, n, seed, extra_bytes=False, point_format=2):
    import potreeconverter_amd as pc
    g = pc.LasGen(preset, n, seed=seed, extra_bytes=extra_bytes, point_format=point_format)
    g.write(str(path))
    return g


def digest_output(outdir):
    """name -> [type, childMask, numPoints, byteSize, sha256(node bytes)]"""
    out = {}
    for name, (typ, mask, npts, size, data) in canonical_nodes(outdir).items():
        out[name] = [typ, mask, npts, size, hashlib.sha256(data).hexdigest()]
    return out


def read(path):
    with open(path, "rb") as f:
        return f.read()
