"""Readers for the Potree 2.0 output files, used to compare conversions.

`decode_hierarchy` walks hierarchy.bin the way the Potree viewer does
(resources/page_template/libs/potree/potree.js:66502-66587 in the
reference): records are read breadth-first per hierarchy chunk, proxies
(type 2) point at the chunk that continues the subtree, and the chunk's
first record replaces the proxy.  `canonical_nodes` then maps every node
name to its record fields and its octree.bin bytes, so two conversions can
be compared independently of the order in which nodes were written.
"""
import json
import os
import struct

REC = 22


def decode_hierarchy(hier: bytes, first_chunk_size: int):
    nodes = {}

    def parse_chunk(root_name, offset, size, root_was_proxy):
        order = [root_name]
        pending = []
        for i in range(size // REC):
            name = order[i]
            typ, mask, npts, boff, bsize = struct.unpack_from("<BBIqq", hier, offset + REC * i)
            node = nodes.setdefault(name, {})
            if i == 0 and root_was_proxy:
                node.update(numPoints=npts, byteOffset=boff, byteSize=bsize)
            elif typ == 2:
                node.update(numPoints=npts, hierOffset=boff, hierSize=bsize)
            else:
                node.update(numPoints=npts, byteOffset=boff, byteSize=bsize)
            node["type"] = typ
            node["childMask"] = mask
            if typ == 2:
                pending.append(name)
                continue
            for c in range(8):
                if mask & (1 << c):
                    order.append(name + str(c))
        for name in pending:
            n = nodes[name]
            parse_chunk(name, n["hierOffset"], n["hierSize"], True)

    parse_chunk("r", 0, first_chunk_size, False)
    return nodes


def load_output(outdir):
    with open(os.path.join(outdir, "metadata.json"), "rb") as f:
        meta_bytes = f.read()
    meta = json.loads(meta_bytes.decode())
    with open(os.path.join(outdir, "hierarchy.bin"), "rb") as f:
        hier = f.read()
    with open(os.path.join(outdir, "octree.bin"), "rb") as f:
        octree = f.read()
    return meta_bytes, meta, hier, octree


def canonical_nodes(outdir):
    """name -> (type, childMask, numPoints, byteSize, node bytes)."""
    _, meta, hier, octree = load_output(outdir)
    nodes = decode_hierarchy(hier, meta["hierarchy"]["firstChunkSize"])
    out = {}
    for name, n in nodes.items():
        off = n.get("byteOffset", 0)
        size = n.get("byteSize", 0)
        out[name] = (n["type"], n["childMask"], n["numPoints"], size, octree[off:off + size])
    return out


def compare_outputs(dir_a, dir_b, check_metadata=True):
    """Returns a list of human-readable differences (empty when equal)."""
    diffs = []
    ma, _, _, _ = load_output(dir_a)
    mb, _, _, _ = load_output(dir_b)
    if check_metadata and ma != mb:
        diffs.append("metadata.json differs")
    na, nb = canonical_nodes(dir_a), canonical_nodes(dir_b)
    if set(na) != set(nb):
        only_a = sorted(set(na) - set(nb))[:10]
        only_b = sorted(set(nb) - set(na))[:10]
        diffs.append(f"node sets differ: only in a {only_a}, only in b {only_b}")
    for name in sorted(set(na) & set(nb)):
        if na[name] != nb[name]:
            a, b = na[name], nb[name]
            what = [k for k, x, y in zip(("type", "childMask", "numPoints", "byteSize", "bytes"), a, b) if x != y]
            diffs.append(f"node {name}: {what} differ ({a[:4]} vs {b[:4]})")
            if len(diffs) > 20:
                break
    return diffs
