"""Oracle for the .vox map loader -- test infrastructure only.

A numpy restatement of what the library's ct_scene_from_vox builds from a
MagicaVoxel file: the loader's per-region rules (src/gvox/map_loader.cpp)
over the VOX chunk format, with the choices the library fixes where the
reference defers to gvox (one region per model or per chunk-aligned box,
scene-graph placement v - floor(s/2) + t, MATL `_emit` / MATT type 3 as
emissive).  Only tests use it, as the checker of the library's result.

Pinned by: tests that write .vox files from known voxel lists and check
the closed forms (axis remap table, AABB = g e .. (g+1) e, light centre
g e + e/2, palette entry i <-> color index i+1, per-region material
dedup, z-y-x primitive order, later-entry-wins duplicates, scene-graph
offsets) against this module and the library, and the voxel counts of the
reference's own assets/models/*.vox (XYZI counts, e.g. monu7 226,036).
"""
import struct

import numpy as np

E = np.float32(0.03125)  # VOXEL_EXTENT, shared.inl:45

# map_loader.cpp:226-262: grid (x, y, z) <- placed (p0, p1, p2) by axis_direction
AXIS_PERM = {0: (1, 2, 0), 3: (1, 2, 0),   # X_BOTTOM_TOP, X_TOP_BOTTOM: x<-y, y<-z, z<-x
             1: (2, 0, 1), 4: (2, 0, 1),   # Y_*: x<-z, y<-x, z<-y
             2: (0, 1, 2), 5: (0, 1, 2)}   # Z_*: identity


def default_palette():
    """MagicaVoxel's default palette as (256, 3) uint8 rgb: 0 empty, then the
    6-level cube (ff..00; red outer, blue inner) without black, then 10-step
    red, green, blue and grey ramps."""
    lv = [0xFF, 0xCC, 0x99, 0x66, 0x33, 0x00]
    ramp = [0xEE, 0xDD, 0xBB, 0xAA, 0x88, 0x77, 0x55, 0x44, 0x22, 0x11]
    rows = [(0, 0, 0)]
    rows += [(r, g, b) for r in lv for g in lv for b in lv][:-1]
    rows += [(v, 0, 0) for v in ramp] + [(0, v, 0) for v in ramp] + [(0, 0, v) for v in ramp]
    rows += [(v, v, v) for v in ramp]
    return np.array(rows, np.uint8)


def _dict(b, o):
    n, = struct.unpack_from("<i", b, o)
    o += 4
    d = {}
    for _ in range(n):
        kl, = struct.unpack_from("<i", b, o)
        k = b[o + 4:o + 4 + kl].decode("latin1")
        o += 4 + kl
        vl, = struct.unpack_from("<i", b, o)
        d[k] = b[o + 4:o + 4 + vl].decode("latin1")
        o += 4 + vl
    return d, o


def parse(b):
    """(models [(size, cells[z, y, x] uint8)], palette (256, 3), emissive (256,) bool, nodes)."""
    assert b[:4] == b"VOX "
    assert b[8:12] == b"MAIN"
    nc, nk = struct.unpack_from("<ii", b, 12)
    o, end = 20 + nc, 20 + nc + nk
    models, nodes, size = [], {}, None
    pal = default_palette()
    emissive = np.zeros(256, bool)
    while o < end:
        cid = b[o:o + 4].decode("latin1")
        n, m = struct.unpack_from("<ii", b, o + 4)
        c = o + 12
        if cid == "SIZE":
            size = struct.unpack_from("<iii", b, c)
        elif cid == "XYZI":
            nv, = struct.unpack_from("<i", b, c)
            v = np.frombuffer(b, np.uint8, 4 * nv, c + 4).reshape(nv, 4)
            cells = np.zeros((size[2], size[1], size[0]), np.uint8)
            lin = (v[:, 2].astype(np.int64) * size[1] + v[:, 1]) * size[0] + v[:, 0]
            _, at = np.unique(lin[::-1], return_index=True)
            last = nv - 1 - at  # a cell written twice keeps the later entry
            cells.reshape(-1)[lin[last]] = v[last, 3]
            models.append((size, cells))
        elif cid == "RGBA":
            p = np.frombuffer(b, np.uint8, 1024, c).reshape(256, 4)
            pal = np.zeros((256, 3), np.uint8)
            pal[1:] = p[:255, :3]
        elif cid == "MATL":
            mid, = struct.unpack_from("<i", b, c)
            d, _ = _dict(b, c + 4)
            if 0 < mid < 256 and d.get("_type") == "_emit" and "_emit" in d:
                emissive[mid] = float(d["_emit"]) > 0
        elif cid == "MATT":
            mid, ty, w = struct.unpack_from("<iif", b, c)
            if 0 < mid < 256:
                emissive[mid] = ty == 3 and w > 0
        elif cid in ("nTRN", "nGRP", "nSHP"):
            nid, = struct.unpack_from("<i", b, c)
            _, q = _dict(b, c + 4)
            if cid == "nTRN":
                child, _, _, nf = struct.unpack_from("<iiii", b, q)
                q += 16
                t, rot = (0, 0, 0), False
                for f in range(nf):
                    d, q = _dict(b, q)
                    if f == 0:
                        if "_t" in d:
                            t = tuple(int(x) for x in d["_t"].split()[:3])
                        rot = "_r" in d and int(d["_r"]) != 4
                nodes[nid] = ("T", [child], t, rot)
            elif cid == "nGRP":
                k, = struct.unpack_from("<i", b, q)
                nodes[nid] = ("G", list(struct.unpack_from(f"<{k}i", b, q + 4)), None, False)
            else:
                k, = struct.unpack_from("<i", b, q)
                q += 4
                ms = []
                for _ in range(k):
                    ms.append(struct.unpack_from("<i", b, q)[0])
                    _, q = _dict(b, q + 4)
                nodes[nid] = ("S", ms, None, False)
        o = c + n + m
    return models, pal, emissive, nodes


def placements(models, nodes, model=-1):
    if model >= 0:
        return [(model, (0, 0, 0))]
    if not nodes:
        return [(k, (0, 0, 0)) for k in range(len(models))]
    out = []

    def walk(nid, t):
        kind, ch, tt, rot = nodes[nid]
        if kind == "T":
            if rot:
                raise ValueError("rotated node")
            for c in ch:
                walk(c, tuple(a + d for a, d in zip(t, tt)))
        elif kind == "G":
            for c in ch:
                walk(c, t)
        else:
            for mi in ch:
                s = models[mi][0]
                out.append((mi, tuple(t[a] - s[a] // 2 for a in range(3))))
    walk(0, (0, 0, 0))
    return out


def load(b, axis=0, chunk=(0, 0, 0), model=-1, e=E, flux=np.float32(1.0)):
    """dict of numpy arrays: instances (first, count), aabbs (n, 6) f32,
    prims (n, 2) u32, mats: diffuse (m, 3) / emission (m, 3) f32, lights:
    position / emissive (l, 3) f32, inst_prim (l, 2) u32."""
    e, flux = np.float32(e), np.float32(flux)
    models, pal, emissive, nodes = parse(b)
    perm = AXIS_PERM[axis]
    inst, aabbs, prims, diff, emis, lpos, lemi, lip = [], [], [], [], [], [], [], []
    n_prim = n_mat = n_light = 0
    col = pal.astype(np.float32) / np.float32(255.0)
    lit = emissive & (pal.astype(np.int32).sum(1) > 0)
    for mi, off in placements(models, nodes, model):
        size, cells = models[mi]
        lo = np.array(off)
        hi = lo + np.array(size)
        if any(chunk):
            cs = np.array([c if c else s for c, s in zip(chunk, size)])
            b0 = np.where(np.array(chunk) > 0, np.floor_divide(lo, cs), 0)
            b1 = np.where(np.array(chunk) > 0, np.floor_divide(hi - 1, cs), 0)
        else:
            cs, b0, b1 = np.array(size), np.zeros(3, int), np.zeros(3, int)
        for bz in range(b0[2], b1[2] + 1):
            for by in range(b0[1], b1[1] + 1):
                for bx in range(b0[0], b1[0] + 1):
                    bb = np.array([bx, by, bz])
                    rlo = np.where(np.array(chunk) > 0, np.maximum(lo, bb * cs), lo)
                    rhi = np.where(np.array(chunk) > 0, np.minimum(hi, (bb + 1) * cs), hi)
                    sub = cells[rlo[2] - lo[2]:rhi[2] - lo[2], rlo[1] - lo[1]:rhi[1] - lo[1],
                                rlo[0] - lo[0]:rhi[0] - lo[0]]
                    z, y, x = np.nonzero(sub)   # row-major: z, then y, then x
                    if z.size == 0:
                        continue
                    ci = sub[z, y, x]
                    p = np.stack([x + rlo[0], y + rlo[1], z + rlo[2]], 1)
                    g = p[:, list(perm)].astype(np.float32)
                    iid = len(inst)
                    # per-region palette dedup in order of first appearance (map_loader.cpp:183-220)
                    ids, first_at = np.unique(ci, return_index=True)
                    order = ids[np.argsort(first_at)]
                    local = np.full(256, -1, np.int64)
                    local[order] = np.arange(order.size) + n_mat
                    em = np.where(lit[order, None], col[order] * flux, np.float32(0)).astype(np.float32)
                    diff.append(col[order])
                    emis.append(em)
                    aabbs.append(np.concatenate([g * e, (g + np.float32(1)) * e], 1).astype(np.float32))
                    lidx = np.full(ci.size, 0xFFFFFFFF, np.uint64)
                    is_l = lit[ci]
                    lidx[is_l] = n_light + np.arange(int(is_l.sum()))
                    prims.append(np.stack([local[ci].astype(np.uint64), lidx], 1).astype(np.uint32))
                    lpos.append((g[is_l] * e + e * np.float32(0.5)).astype(np.float32))
                    lemi.append((col[ci[is_l]] * flux).astype(np.float32))
                    lip.append(np.stack([np.full(int(is_l.sum()), iid), np.nonzero(is_l)[0]], 1).astype(np.uint32))
                    inst.append((n_prim, ci.size))
                    n_prim += ci.size
                    n_mat += order.size
                    n_light += int(is_l.sum())
    cat = (lambda a, k: np.concatenate(a) if a else np.zeros((0, k), np.float32))
    return {"instances": np.array(inst, np.uint32).reshape(-1, 2), "aabbs": cat(aabbs, 6),
            "prims": cat(prims, 2).astype(np.uint32), "diffuse": cat(diff, 3), "emission": cat(emis, 3),
            "light_pos": cat(lpos, 3), "light_emissive": cat(lemi, 3), "light_ids": cat(lip, 2).astype(np.uint32)}
