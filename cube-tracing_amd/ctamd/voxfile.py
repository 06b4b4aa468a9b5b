"""Writes MagicaVoxel .vox files (chunk layout of the public VOX format:
"VOX " version, MAIN { SIZE XYZI ... RGBA MATL MATT nTRN nGRP nSHP }) and
procedural models, the inputs of Scene.from_vox in tests and benchmarks."""
import struct


def _chunk(cid, content=b"", children=b""):
    return cid.encode() + struct.pack("<ii", len(content), len(children)) + content + children


def _str(s):
    b = s.encode()
    return struct.pack("<i", len(b)) + b


def _dict(d):
    return struct.pack("<i", len(d)) + b"".join(_str(k) + _str(v) for k, v in d.items())


def write_vox(models, palette=None, matl=None, matt=None, graph=None, version=150):
    """models: [((sx, sy, sz), [(x, y, z, color_index), ...])];
    palette: 256 (r, g, b, a) tuples for RGBA (entry i = color index i+1) or None;
    matl: {index: {key: value}}; matt: {index: (type, weight)};
    graph: None, or [(translation (tx, ty, tz), [model ids], rotation or None)],
    written as nTRN(0) -> nGRP(1) -> nTRN -> nSHP per entry."""
    body = b""
    for (sx, sy, sz), vox in models:
        body += _chunk("SIZE", struct.pack("<iii", sx, sy, sz))
        body += _chunk("XYZI", struct.pack("<i", len(vox)) + b"".join(struct.pack("<4B", *v) for v in vox))
    if graph is not None:
        nid = 2
        kids = []
        nodes = b""
        for t, mids, rot in graph:
            frame = {"_t": f"{t[0]} {t[1]} {t[2]}"}
            if rot is not None:
                frame["_r"] = str(rot)
            kids.append(nid)
            nodes += _chunk("nTRN", struct.pack("<i", nid) + _dict({}) + struct.pack("<iiii", nid + 1, -1, 0, 1)
                            + _dict(frame))
            nodes += _chunk("nSHP", struct.pack("<i", nid + 1) + _dict({}) + struct.pack("<i", len(mids))
                            + b"".join(struct.pack("<i", m) + _dict({}) for m in mids))
            nid += 2
        body += _chunk("nTRN", struct.pack("<i", 0) + _dict({}) + struct.pack("<iiii", 1, -1, -1, 1) + _dict({}))
        body += _chunk("nGRP", struct.pack("<i", 1) + _dict({}) + struct.pack("<i", len(kids))
                       + b"".join(struct.pack("<i", k) for k in kids))
        body += nodes
    if palette is not None:
        body += _chunk("RGBA", b"".join(struct.pack("<4B", *c) for c in palette))
    for k, d in (matl or {}).items():
        body += _chunk("MATL", struct.pack("<i", k) + _dict(d))
    for k, (ty, w) in (matt or {}).items():
        body += _chunk("MATT", struct.pack("<iifi", k, ty, w, 0))
    return b"VOX " + struct.pack("<i", version) + _chunk("MAIN", b"", body)


def monument(n=48, seed=5, emissive_colors=(200, 201), fill=0.3):
    """A procedural single-model file: a stepped tower with noise-carved
    walls over an n x n x n grid, palette colors 1..255, a few emissive
    palette entries (MATL _emit)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    z, y, x = np.mgrid[0:n, 0:n, 0:n]
    r = np.maximum(np.abs(x - n / 2), np.abs(y - n / 2))
    solid = (r < (n / 2) * (1 - z / (1.3 * n))) & (rng.random((n, n, n)) < fill + 0.5 * (z < 3))
    lamp = solid & (rng.random((n, n, n)) < 0.01)
    col = (1 + (z * 7 + (x // 4) * 3 + (y // 4)) % 190).astype(np.uint8)
    if emissive_colors:
        col[lamp] = rng.choice(emissive_colors, size=int(lamp.sum()))
    zz, yy, xx = np.nonzero(solid)
    vox = list(zip(xx.tolist(), yy.tolist(), zz.tolist(), col[zz, yy, xx].tolist()))
    pal = [((i * 37) % 256, (i * 91) % 256, (i * 53) % 256, 255) for i in range(1, 257)]
    matl = {c: {"_type": "_emit", "_emit": "1", "_flux": "2"} for c in emissive_colors}
    return write_vox([((n, n, n), vox)], palette=pal, matl=matl, version=200)
