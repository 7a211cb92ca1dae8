"""MI355X (gfx950) counting-sort chunker + octree indexer, PotreeConverter 2.x compatible.

The compute path is libpotree_gpu.so (hand-written HIP kernels plus the C++
drop-ins for doChunking / doIndexing).  This module is a thin ctypes view
of its C ABI (include/potree_gpu.h); it never falls back to a CPU
implementation: a missing library raises at import of `lib()`.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpotree_gpu.so")
_lib = None


class PotreeGpuError(RuntimeError):
    pass


def lib():
    """Loads the native library (built by potreeconverter_amd/build.py)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PotreeGpuError(f"{LIB_PATH} is missing; run potreeconverter_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        c = ctypes
        # every entry point gets its full parameter list (tests/test_capi.py
        # checks them against include/potree_gpu.h): without argtypes ctypes
        # passes Python ints as 32-bit C ints, which truncates 64-bit sizes
        L.pcg_init.argtypes = []
        L.pcg_init.restype = c.c_int
        L.pcg_last_error.argtypes = []
        L.pcg_last_error.restype = c.c_char_p
        L.pcg_convert.argtypes = [c.c_char_p, c.c_char_p, c.c_char_p, c.c_int32, c.POINTER(c.c_double)]
        L.pcg_do_chunking.argtypes = [c.c_char_p, c.c_char_p, c.POINTER(c.c_double)]
        L.pcg_do_indexing.argtypes = [c.c_char_p, c.c_char_p, c.c_char_p, c.c_int64, c.POINTER(c.c_double)]
        L.pcg_count_cells.argtypes = [c.c_void_p, c.c_int64, c.c_int32] + [c.POINTER(c.c_double)] * 6 + [
            c.c_int32, c.c_void_p, c.c_void_p]
        L.pcg_count_cells_noalloc.argtypes = L.pcg_count_cells.argtypes
        L.pcg_convert_device.argtypes = [c.c_void_p, c.c_int64, c.c_void_p, c.c_char_p, c.c_char_p,
                                         c.POINTER(c.c_double)]
        L.pcg_convert_encoded.argtypes = [c.c_char_p, c.c_char_p, c.c_char_p, c.c_char_p, c.c_int32,
                                          c.POINTER(c.c_double)]
        L.pcg_do_indexing_encoded.argtypes = [c.c_char_p, c.c_char_p, c.c_char_p, c.c_char_p, c.c_int64,
                                              c.POINTER(c.c_double)]
        L.pcg_convert_device_encoded.argtypes = [c.c_void_p, c.c_int64, c.c_void_p, c.c_char_p, c.c_char_p,
                                                 c.c_char_p, c.POINTER(c.c_double)]
        L.pcg_device_alloc.argtypes = [c.POINTER(c.c_void_p), c.c_uint64]
        L.pcg_device_free.argtypes = [c.c_void_p]
        L.pcg_copy.argtypes = [c.c_void_p, c.c_void_p, c.c_uint64, c.c_int32]
        L.pcg_synchronize.argtypes = []
        L.pcg_lasgen_preset.argtypes = [c.c_int32, c.c_int64, c.c_uint64, c.c_void_p]
        L.pcg_lasgen_configure.argtypes = [c.c_void_p, c.c_int32, c.c_int32]
        L.pcg_lasgen_record_length.argtypes = [c.c_void_p]
        L.pcg_lasgen_record_length.restype = c.c_int32
        L.pcg_lasgen_fill.argtypes = [c.c_void_p, c.c_int64, c.c_int64, c.c_void_p]
        L.pcg_lasgen_fill.restype = None
        L.pcg_lasgen_write.argtypes = [c.c_void_p, c.c_char_p]
        L.pcg_lasgen_device.argtypes = [c.c_void_p, c.c_int64, c.c_int64, c.c_void_p, c.c_void_p]
        L.pcg_lasgen_header.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p]
        L.pcg_do_indexing_device.argtypes = [c.c_void_p, c.c_int32, c.POINTER(c.c_char_p), c.POINTER(c.c_int64),
                                             c.c_char_p, c.c_char_p, c.c_char_p, c.c_char_p, c.c_char_p,
                                             c.POINTER(c.c_double)]
        L.pcg_chunkgen_params.argtypes = [c.c_int32, c.c_double, c.c_double, c.c_int64, c.c_int64, c.c_uint64,
                                          c.c_void_p]
        L.pcg_chunkgen_counts.argtypes = [c.c_void_p, c.POINTER(c.c_int64), c.c_void_p]
        L.pcg_chunkgen_fill.argtypes = [c.c_void_p, c.c_int32, c.c_int64, c.c_int64, c.c_void_p]
        L.pcg_chunkgen_fill.restype = None
        L.pcg_chunkgen_write.argtypes = [c.c_void_p, c.c_char_p]
        L.pcg_chunkgen_device.argtypes = [c.c_void_p, c.c_void_p, c.c_char_p, c.c_int64]
        L.pcg_sample_subtree.argtypes = [c.c_char_p, c.c_int32, c.c_void_p, c.c_void_p, c.c_void_p, c.c_void_p,
                                         c.c_void_p, c.c_void_p, c.c_int32, c.c_int32, c.POINTER(c.c_double),
                                         c.POINTER(c.c_double), c.c_double, c.c_void_p, c.c_void_p, c.c_int64,
                                         c.c_void_p, c.c_void_p, c.POINTER(c.c_int32)]
        L.pcg_device_memory.argtypes = [c.POINTER(c.c_uint64), c.POINTER(c.c_uint64)]
        L.pcg_workspace_bytes.argtypes = []
        L.pcg_workspace_bytes.restype = c.c_uint64
        L.pcg_do_chunking_sources.argtypes = [c.POINTER(c.c_char_p), c.c_int32, c.c_char_p, c.POINTER(c.c_double),
                                              c.POINTER(c.c_double), c.POINTER(c.c_char_p), c.c_int32,
                                              c.POINTER(c.c_double)]
        L.pcg_convert_sources.argtypes = [c.POINTER(c.c_char_p), c.c_int32, c.c_char_p, c.c_char_p, c.c_char_p,
                                          c.POINTER(c.c_char_p), c.c_int32, c.c_int32, c.POINTER(c.c_double)]
        L.pcg_release.argtypes = []
        L.pcg_set_option.argtypes = [c.c_char_p, c.c_int64]
        L.pcg_get_option.argtypes = [c.c_char_p, c.POINTER(c.c_int64)]
        L.pcg_autotune_params.argtypes = []
        L.pcg_autotune_params.restype = c.c_int32
        L.pcg_autotune_param_name.argtypes = [c.c_int32]
        L.pcg_autotune_param_name.restype = c.c_char_p
        L.pcg_autotune_device.argtypes = [c.c_void_p, c.c_int64, c.c_void_p, c.c_char_p, c.c_int32, c.c_int32,
                                          c.POINTER(c.c_int64), c.POINTER(c.c_double), c.POINTER(c.c_double)]
        for name in ("pcg_set_option", "pcg_release", "pcg_convert", "pcg_do_chunking", "pcg_do_indexing", "pcg_count_cells", "pcg_count_cells_noalloc", "pcg_convert_device",
                     "pcg_convert_encoded", "pcg_do_indexing_encoded", "pcg_convert_device_encoded",
                     "pcg_device_alloc", "pcg_device_free", "pcg_copy", "pcg_synchronize", "pcg_lasgen_preset",
                     "pcg_lasgen_configure",
                     "pcg_lasgen_write", "pcg_lasgen_device", "pcg_lasgen_header", "pcg_init",
                     "pcg_do_indexing_device", "pcg_chunkgen_params", "pcg_chunkgen_counts", "pcg_chunkgen_write",
                     "pcg_chunkgen_device", "pcg_sample_subtree", "pcg_device_memory",
                     "pcg_do_chunking_sources", "pcg_convert_sources", "pcg_get_option", "pcg_autotune_device"):
            getattr(L, name).restype = c.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise PotreeGpuError(f"error {rc}: {lib().pcg_last_error().decode(errors='replace')}")


class LasGen:
    """Parameters of the seeded synthetic LAS generator (preset 0: 500 m tile,
    preset 1: 5 km tile at UTM-like coordinates, preset 2: precision scene with
    scale 0.0001, offset 1e6 m, near-duplicate clusters and a lattice whose
    pitch equals the level-4 node spacing)."""

    def __init__(self, preset, n, seed=0, extra_bytes=False, point_format=2):
        self.buf = ctypes.create_string_buffer(256)
        _check(lib().pcg_lasgen_preset(preset, n, seed, self.buf))
        _check(lib().pcg_lasgen_configure(self.buf, point_format, int(bool(extra_bytes))))
        self.n = n
        self.point_format = point_format

    @property
    def record_length(self):
        return lib().pcg_lasgen_record_length(self.buf)

    def records(self, first=0, count=None):
        count = self.n - first if count is None else count
        out = ctypes.create_string_buffer(count * self.record_length)
        lib().pcg_lasgen_fill(self.buf, first, count, out)
        return out.raw

    def write(self, path):
        _check(lib().pcg_lasgen_write(self.buf, str(path).encode()))


class ChunkGen:
    """Parameters of the seeded synthetic pre-chunked input of phase 2:
    num_chunks level-4 chunks "rABCD" of a 2048 m cube, log-normal point
    counts round(median * exp(sigma * g)) clipped to [min_count, max_count],
    34-byte records (LAS format 2 output layout + float32 / uint16 / int8
    extra bytes).  C2 = ChunkGen(4096, 50_000, C2_SIGMA, 1_000, 5_000_000)."""

    RECORD_BYTES = 34

    def __init__(self, num_chunks, median, sigma, min_count, max_count, seed=0):
        self.buf = ctypes.create_string_buffer(128)
        _check(lib().pcg_chunkgen_params(num_chunks, median, sigma, min_count, max_count, seed, self.buf))
        self.num_chunks = num_chunks
        counts = (ctypes.c_int64 * num_chunks)()
        names = ctypes.create_string_buffer(6 * num_chunks)
        _check(lib().pcg_chunkgen_counts(self.buf, counts, names))
        self.counts = list(counts)
        self.names = [names.raw[6 * i:6 * i + 5].decode() for i in range(num_chunks)]
        self.total = sum(self.counts)

    def records(self, chunk, first=0, count=None):
        count = self.counts[chunk] - first if count is None else count
        out = ctypes.create_string_buffer(count * self.RECORD_BYTES)
        lib().pcg_chunkgen_fill(self.buf, chunk, first, count, out)
        return out.raw

    def write(self, outdir):
        """outdir/chunks/<name>.bin + metadata.json (host generation)."""
        _check(lib().pcg_chunkgen_write(self.buf, str(outdir).encode()))

    def device(self, d_records):
        """Generates the records into d_records (a device pointer of at least
        total * 34 bytes) on the GPU; returns the chunks/metadata.json text."""
        meta = ctypes.create_string_buffer(1 << 20)
        _check(lib().pcg_chunkgen_device(self.buf, ctypes.c_void_p(d_records), meta, len(meta)))
        return meta.value.decode()


def do_indexing_device(d_records, names, counts, metadata, method="poisson", encoding=None, name="",
                       outdir=None):
    """indexer::doIndexing on chunk records resident in HBM (ascending chunk
    names, records back to back); returns the 24 statistics of
    pcg_convert_device."""
    n = len(names)
    cn = (ctypes.c_char_p * n)(*[x.encode() for x in names])
    cc = (ctypes.c_int64 * n)(*counts)
    st = (ctypes.c_double * 24)()
    _check(lib().pcg_do_indexing_device(ctypes.c_void_p(d_records), n, cn, cc, metadata.encode(), method.encode(),
                                        encoding.encode() if encoding else None, name.encode(),
                                        str(outdir).encode() if outdir else None, st))
    return list(st)


def sample_subtree(method, children, bounds, levels, counts, records, bpp, rgb_offset, scale, offset,
                   base_spacing, sampled=None):
    """One Sampler::sample call on a flattened subtree (node 0 = the node
    handed to sample()).  children: 8 ids per node (-1 none); bounds: min
    xyz + max xyz per node; levels; counts and records (bytes, node order) of
    the nodes without children; sampled: 1 for holders sampled before.
    Returns (counts, records per node, children per node, events), events
    as (node, 0 = completed | 1 = discarded) in callback order."""
    n = len(levels)
    I32, I64, F64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    ch = (I32 * (8 * n))(*[x for row in children for x in row])
    bd = (F64 * (6 * n))(*[x for row in bounds for x in row])
    lv = (I32 * n)(*levels)
    sm = (ctypes.c_uint8 * n)(*(sampled or [0] * n))
    cn = (I64 * n)(*counts)
    total = sum(counts)
    # Poisson / random only move records; Poisson-average copies the
    # accepted ones into every sampled node
    height = max(levels) - min(levels) + 1
    cap = total * (height + 1) if method == "poisson_average" else total
    out = ctypes.create_string_buffer(max(cap * bpp, 1))
    oc = (I64 * n)()
    och = (I32 * (8 * n))()
    ev = (I32 * (4 * n))()
    ne = I32()
    _check(lib().pcg_sample_subtree(method.encode(), n, ch, bd, lv, sm, cn, records, bpp, rgb_offset,
                                    (F64 * 3)(*scale), (F64 * 3)(*offset), base_spacing, oc, out, cap, och, ev,
                                    ctypes.byref(ne)))
    recs, o = [], 0
    for i in range(n):
        recs.append(out.raw[o * bpp:(o + oc[i]) * bpp])
        o += oc[i]
    return (list(oc), recs, [list(och[8 * i:8 * i + 8]) for i in range(n)],
            [(ev[2 * e], ev[2 * e + 1]) for e in range(ne.value)])


def convert(las_path, outdir, method="poisson", keep_chunks=False, encoding=None):
    """PotreeConverter <las> -o <outdir> -m <method> [--keep-chunks] [--encoding <encoding>]."""
    t = (ctypes.c_double * 8)()
    flags = 1 if keep_chunks else 0
    if encoding is None:
        _check(lib().pcg_convert(str(las_path).encode(), str(outdir).encode(), method.encode(), flags, t))
    else:
        _check(lib().pcg_convert_encoded(str(las_path).encode(), str(outdir).encode(), method.encode(),
                                         encoding.encode(), flags, t))
    return list(t)


def set_option(key, value):
    _check(lib().pcg_set_option(key.encode(), int(value)))


def get_option(key):
    v = ctypes.c_int64()
    _check(lib().pcg_get_option(key.encode(), ctypes.byref(v)))
    return v.value


def autotune_device(d_las, n, header, method="poisson", passes=2, repeats=2):
    """Searches the sampling launch options on the caller's workload (records
    in HBM) and leaves the fastest set; returns ({option: value}, seconds with
    them, seconds with the options as they were)."""
    L = lib()
    k = L.pcg_autotune_params()
    chosen = (ctypes.c_int64 * k)()
    best, default = ctypes.c_double(), ctypes.c_double()
    _check(L.pcg_autotune_device(ctypes.c_void_p(d_las), n, header, method.encode(), passes, repeats, chosen,
                                 ctypes.byref(best), ctypes.byref(default)))
    names = [L.pcg_autotune_param_name(i).decode() for i in range(k)]
    return dict(zip(names, list(chosen))), best.value, default.value


def _strings(values):
    values = list(values or [])
    return (ctypes.c_char_p * max(len(values), 1))(*[str(v).encode() for v in values]), len(values)


def do_chunking_sources(las_paths, outdir, cube_min=None, cube_max=None, attributes=None):
    """doChunking(sources, targetDir, min, max, ..., outputAttributes from
    the --attributes filter) over several LAS files."""
    paths, n = _strings(las_paths)
    names, na = _strings(attributes)
    t = (ctypes.c_double * 8)()
    mn = (ctypes.c_double * 3)(*cube_min) if cube_min is not None else None
    mx = (ctypes.c_double * 3)(*cube_max) if cube_max is not None else None
    _check(lib().pcg_do_chunking_sources(paths, n, str(outdir).encode(), mn, mx, names if na else None, na, t))
    return list(t)


def convert_sources(las_paths, outdir, method="poisson", attributes=None, encoding=None, keep_chunks=False):
    """PotreeConverter <las files...> -o <outdir> -m <method> [--attributes ...]."""
    paths, n = _strings(las_paths)
    names, na = _strings(attributes)
    t = (ctypes.c_double * 8)()
    _check(lib().pcg_convert_sources(paths, n, str(outdir).encode(), method.encode(),
                                     encoding.encode() if encoding else None, names if na else None, na,
                                     1 if keep_chunks else 0, t))
    return list(t)


def do_chunking(las_path, outdir):
    t = (ctypes.c_double * 8)()
    _check(lib().pcg_do_chunking(str(las_path).encode(), str(outdir).encode(), t))
    return list(t)


def do_indexing(outdir, method="poisson", name="", points_total=0, encoding=None):
    t = (ctypes.c_double * 8)()
    if encoding is None:
        _check(lib().pcg_do_indexing(str(outdir).encode(), method.encode(), name.encode(), points_total, t))
    else:
        _check(lib().pcg_do_indexing_encoded(str(outdir).encode(), method.encode(), encoding.encode(),
                                             name.encode(), points_total, t))
    return list(t)
