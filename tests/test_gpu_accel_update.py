"""Scene updates through the task queue, on the GPU, against the oracle:
the library and the task-semantics model take the same task sequence and
must hold the same records (instances, AABBs, primitives, lights, the
remapping tables); rebuilt BLASes equal the oracle's LBVH, refitted ones
keep their topology with exact bounds; traversal and whole frames across an
update (remapped temporal reuse, previous transforms in the motion
vectors) are bit-exact.  accel_struct_mngr.cpp:2115-2680."""
import math

import numpy as np
import pytest

import ctamd
import oracle as O
from accel_model import INVALID, AccelModel
from helpers import bits, di_array, oracle_scene_from_accel, pres_array, random_rays, res_array, to_device

pytestmark = pytest.mark.gpu


def _workload(**kw):
    args = dict(width=48, height=32, grid=48, chunk=16, frames=8, max_depth=2)
    args.update(kw)
    return ctamd.hd_di(**args)


def _load(w):
    sc = ctamd.Scene(ctamd.scene_desc(w))
    ni = sc.info.instance_count
    acc = ctamd.Accel(max_instances=ni + 8, max_primitives=sc.info.primitive_count + 32 * ni + 1024,
                      max_cube_lights=sc.info.cube_light_count + 8)
    ctx = ctamd.Ctx(w.width, w.height)
    ctamd.load_scene(sc, acc, ctx)
    return sc, acc, ctx


def _model(acc):
    W = acc.read_world()
    return AccelModel.from_records(acc.read_instances(), W["aabbs"], W["prims"], W["lights"], W["light_count"])


def _firsts(acc):
    return {i: r.first_primitive_index for i, r in enumerate(acc.read_instances())}


def _compare(acc, m, prev_first=None):
    """Library records == model records; remapping tables too."""
    inst = acc.read_instances()
    W = acc.read_world()
    used = {i for i, r in enumerate(inst) if r.primitive_count}
    assert used == set(m.instances)
    for i in used:
        r, mi = inst[i], m.instances[i]
        assert r.primitive_count == len(mi.voxels), i
        assert np.array_equal(bits(r.transform[:]), bits(mi.transform)), i
        f, n = r.first_primitive_index, r.primitive_count
        assert np.array_equal(bits(W["aabbs"][f:f + n]), bits(np.array([v.aabb for v in mi.voxels], np.float32))), i
        assert np.array_equal(W["prims"][f:f + n, 0], [v.material for v in mi.voxels]), i
        assert np.array_equal(W["prims"][f:f + n, 1], [v.light for v in mi.voxels]), i
    assert W["light_count"] == len(m.lights)
    for l, rec in enumerate(m.lights):
        g = W["lights"][l]
        raw = bytearray(bytes(g))
        raw[24:32] = b"\0" * 8
        assert bytes(raw) == rec.data, l
        assert (g.instance_info.instance_id, g.instance_info.primitive_id) == (rec.instance, rec.primitive), l
    if prev_first is not None:
        pf = {i: prev_first.get(i, inst[i].first_primitive_index) for i in range(len(inst))}
        assert np.array_equal(W["remap_prims"], m.remap_slot_table(pf, W["capacity"]))
        assert np.array_equal(W["remap_lights"], m.remap_light_table(len(W["remap_lights"])))
    return inst, W


def _oracle(acc, sc, ctx):
    cfg = ctx.light_config()
    cfg.cube_light_count = acc.read_world()["light_count"]
    osc, inst, W = oracle_scene_from_accel(acc, sc.materials, cfg)
    return osc


def _blas_equal(acc, osc, i, n):
    g, o = acc.read_bvh(i, n), osc.read_bvh(i, n)
    for k in ("keys", "ids", "child", "parent"):
        assert np.array_equal(g[k], o[k]), (i, k)
    assert np.array_equal(bits(g["boxes"]), bits(o["boxes"])), i


def _trace_equal(acc, osc, sc, seed):
    rng = np.random.default_rng(seed)
    o, d = random_rays(rng, 3000, sc.grid_min, sc.grid_max, inside_frac=0.3)
    gt, gid = acc.trace_closest(o, d)
    ot, oid = osc.trace_closest(o, d)
    assert np.array_equal(gid, oid) and np.array_equal(bits(gt), bits(ot))
    assert 0.1 < (gid[:, 0] != INVALID).mean() < 1.0
    dist = np.where(gid[:, 0] != INVALID, gt, 50.0).astype(np.float32)
    assert np.array_equal(acc.trace_any(o, d, dist * 0.5), osc.trace_any(o, d, dist * 0.5))
    gb, ob = acc.trace_bounce(o, d), osc.trace_bounce(o, d)
    for k in ("ids", "pos", "nrm", "dist", "wo"):
        assert np.array_equal(bits(gb[k]) if k != "ids" else gb[k], bits(ob[k]) if k != "ids" else ob[k]), k


def _light_voxels(acc, i):
    r = acc.read_instances()[i]
    W = acc.read_world()
    f = r.first_primitive_index
    return [k for k in range(r.primitive_count) if W["prims"][f + k, 1] != INVALID]


def test_primitive_deletion_and_undo(gpu):
    w = _workload()
    sc, acc, ctx = _load(w)
    assert acc.status() == ctamd.CT_AS_SETTLING
    assert acc.update_scene() and acc.status() == ctamd.CT_AS_IDLE  # the load settles
    assert not acc.update_scene()                                      # nothing queued
    m = _model(acc)
    inst0 = acc.read_instances()
    lit = [i for i in range(len(inst0)) if _light_voxels(acc, i)]
    assert len(lit) >= 3
    a, b, c = lit[:3]
    la = _light_voxels(acc, a)
    rounds = [
        [("delete_primitive", a, la[0]), ("delete_primitive", b, 0)],
        [("delete_primitive", c, inst0[c].primitive_count - 1), ("delete_primitive", a, 3)],
        [("delete_primitive", a, la[-1] if la[-1] < len(m.instances[a].voxels) - 2 else 1)],
    ]
    for tasks in rounds:
        pf = _firsts(acc)
        for t in tasks:
            acc.delete_primitive(t[1], t[2])
        assert acc.update_scene()
        m.update(tasks)
        assert acc.status() == ctamd.CT_AS_SETTLING and acc.remapping_active()
        _compare(acc, m, pf)
        osc = _oracle(acc, sc, ctx)
        for t in tasks:
            _blas_equal(acc, osc, t[1], len(m.instances[t[1]].voxels))
        _trace_equal(acc, osc, sc, 11)
    # the next call settles: tables cleared, status idle
    pf = _firsts(acc)
    assert acc.update_scene()
    m.update([])
    assert acc.status() == ctamd.CT_AS_IDLE
    _compare(acc, m, pf)
    assert not acc.read_world()["remap_prims"].any()
    # undo the last two deletions, one per call
    for _ in range(2):
        pf = _firsts(acc)
        acc.undo()
        acc.update_scene()
        m.update([("undo",)])
        _compare(acc, m, pf)
    osc = _oracle(acc, sc, ctx)
    _blas_equal(acc, osc, a, len(m.instances[a].voxels))
    _trace_equal(acc, osc, sc, 12)


def _rot_y(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    r = np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], np.float32)
    return r.T.reshape(16)


def _union_check(g, leaf_box):
    """Every child box of every node is the exact union of what lies below."""
    n = g["n"]
    child, boxes = g["child"].reshape(-1, 2), g["boxes"].reshape(-1, 12)

    def box_of(ref):
        if ref & 0x80000000:
            return leaf_box[g["ids"][ref & 0x7FFFFFFF]]
        b = boxes[ref]
        return np.concatenate([np.minimum(b[0:3], b[6:9]), np.maximum(b[3:6], b[9:12])])

    for k in range(n - 1):
        for c in range(2):
            assert np.array_equal(bits(boxes[k, 6 * c:6 * c + 6]), bits(box_of(child[k, c]))), (k, c)


def test_update_blas_transform_and_refit(gpu):
    w = _workload()
    sc, acc, ctx = _load(w)
    acc.update_scene()
    m = _model(acc)
    inst = acc.read_instances()
    i = 4
    n = inst[i].primitive_count
    before = acc.read_bvh(i, n)
    # move five voxels by a few voxel widths (the reference's frame-2000 test edits two, main.cpp:1344-1371)
    W = acc.read_world()
    f = inst[i].first_primitive_index
    rng = np.random.default_rng(3)
    edits = []
    for p in rng.choice(n, 5, replace=False):
        shift = rng.integers(-3, 4, 3).astype(np.float32) * ctamd.CT_VOXEL_EXTENT
        box = W["aabbs"][f + p].copy()
        box[:3] += shift
        box[3:] += shift
        edits.append((int(p), box))
    edits.append((n + 7, np.zeros(6, np.float32)))  # out of range: skipped
    tr = _rot_y(0.1)
    pf = _firsts(acc)
    acc.update_blas(i, tr, edits)
    acc.update_blas(2, tr)  # transform only
    acc.update_scene()
    m.update([("update_blas", i, tr, edits), ("update_blas", 2, tr, [])])
    _compare(acc, m, pf)
    after = acc.read_bvh(i, n)
    for k in ("keys", "ids", "child", "parent"):
        assert np.array_equal(before[k], after[k]), k          # topology kept
    assert not np.array_equal(bits(before["boxes"]), bits(after["boxes"]))
    # leaf boxes are the fresh build's, each node the exact union below it
    osc = _oracle(acc, sc, ctx)
    o = osc.read_bvh(i, n)
    leaf_box = {}
    for k, (c0, c1) in enumerate(o["child"].reshape(-1, 2)):
        for c, ref in enumerate((c0, c1)):
            if ref & 0x80000000:
                leaf_box[o["ids"][ref & 0x7FFFFFFF]] = o["boxes"].reshape(-1, 12)[k, 6 * c:6 * c + 6]
    assert len(leaf_box) == n
    _union_check(after, leaf_box)
    ob = o["boxes"].reshape(-1, 12)[0]
    gb = after["boxes"].reshape(-1, 12)[0]
    root = lambda b: np.concatenate([np.minimum(b[0:3], b[6:9]), np.maximum(b[3:6], b[9:12])])
    assert np.array_equal(bits(root(gb)), bits(root(ob)))
    _trace_equal(acc, osc, sc, 21)


def _frames_equal(ctx, ofr, w, f):
    for which in (ctamd.CT_BUF_DI, ctamd.CT_BUF_PREV_DI):
        assert np.array_equal(di_array(ctx.read(which, w.width, w.height)), di_array(ofr.read(which))), (f, which)
    hit = di_array(ofr.read(ctamd.CT_BUF_DI))[:, 3].view(np.float32) > 0
    for which in (ctamd.CT_BUF_INT_RESERVOIR, ctamd.CT_BUF_PREV_RESERVOIR):
        ga, oa = res_array(ctx.read(which, w.width, w.height)), res_array(ofr.read(which))
        assert np.array_equal(ga[hit], oa[hit]), (f, which)
    for which in (ctamd.CT_BUF_PATH_RESERVOIR, ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR):
        assert np.array_equal(pres_array(ctx.read(which, w.width, w.height)), pres_array(ofr.read(which))), (f, which)
    assert np.array_equal(bits(ctx.read(ctamd.CT_BUF_VELOCITY, w.width, w.height)),
                          bits(ofr.read(ctamd.CT_BUF_VELOCITY))), f
    assert np.array_equal(bits(ctx.read(ctamd.CT_BUF_IMAGE, w.width, w.height)), bits(ofr.read(ctamd.CT_BUF_IMAGE))), f
    gs, os_ = ctx.stats(), ofr.stats()
    assert (gs.shadow_rays, gs.bounce_rays) == (os_.shadow_rays, os_.bounce_rays), f


def test_frames_across_an_update_with_remapping(gpu):
    """Frames 0-1, then deletions (hit voxels, a light) and a transform
    update; frame 2 runs with REMAP_BIT and the previous transforms, the
    settle, then frame 3 as usual -- all bit-exact against the oracle."""
    w = _workload(frames=6)
    w.camera = lambda f: ((0.2, 0.35, 2.2), (-0.09, -0.15, -0.98), (0.0, 1.0, 0.0))  # static: reuse everywhere
    sc, acc, ctx = _load(w)
    acc.update_scene()
    osc = _oracle(acc, sc, ctx)
    ofr = O.OFrame(osc, w.width, w.height)
    for f in range(2):
        p = ctamd.frame_params(w, f)
        ctx.render(acc, p)
        ofr.render(p)
        _frames_equal(ctx, ofr, w, f)
    # voxels the previous frame hit: delete a few, and a light
    di = di_array(ctx.read(ctamd.CT_BUF_DI, w.width, w.height))
    hits = di[di[:, 3].view(np.float32) > 0][:, 11:13]
    pairs, seen = [], set()
    for inst, prim in hits[:: max(len(hits) // 6, 1)]:
        if (inst, prim) not in seen and len(pairs) < 4:
            seen.add((int(inst), int(prim)))
            pairs.append((int(inst), int(prim)))
    lit = next(i for i in range(len(acc.read_instances())) if _light_voxels(acc, i) and i not in {q for q, _ in pairs})
    inst_before = acc.read_instances()
    prev_t = np.stack([np.array(r.transform[:], np.float32) for r in inst_before])
    prev_first = np.array([r.first_primitive_index for r in inst_before], np.uint32)
    by_inst = {}
    for inst, prim in pairs:
        by_inst.setdefault(inst, []).append(prim)
    for inst, prims in by_inst.items():
        for prim in sorted(prims, reverse=True):   # higher first: earlier deletions do not shift them
            acc.delete_primitive(inst, prim)
    acc.delete_primitive(lit, _light_voxels(acc, lit)[0])
    moved = (pairs[0][0] + 1) % len(inst_before)
    acc.update_blas(moved, _rot_y(0.5))
    acc.update_scene()
    W = acc.read_world()
    assert (W["remap_prims"] == INVALID).sum() >= len(pairs) + 1
    assert (W["remap_lights"] != 0).any()
    osc2 = _oracle(acc, sc, ctx)
    osc2.set_previous(prev_t, prev_first, W["remap_prims"], W["remap_lights"])
    ofr.set_scene(osc2)
    ofr.set_previous_scene(osc)
    p = ctamd.frame_params(w, 2)
    p.is_active |= ctamd.CT_REMAP_BIT
    ctx.render(acc, p)
    ofr.render(p)
    _frames_equal(ctx, ofr, w, 2)
    # the remapping changed something: frame 2 without it differs
    assert acc.update_scene()  # settle
    osc3 = _oracle(acc, sc, ctx)
    ofr.set_scene(osc3)
    ofr.set_previous_scene(None)
    p = ctamd.frame_params(w, 3)
    ctx.render(acc, p)
    ofr.render(p)
    _frames_equal(ctx, ofr, w, 3)


def test_delete_from_gpu_keeps_the_callers_compaction(gpu):
    w = _workload()
    sc, acc, ctx = _load(w)
    acc.update_scene()
    m = _model(acc)
    inst = acc.read_instances()
    i = 5
    f, n = inst[i].first_primitive_index, inst[i].primitive_count
    W = acc.read_world()
    rng = np.random.default_rng(9)
    # the caller's pass: drop k non-light voxels, keep the order of the rest
    cand = [k for k in range(n) if W["prims"][f + k, 1] == INVALID]
    gone = sorted(rng.choice(cand, 6, replace=False).tolist())
    keep = [k for k in range(n) if k not in gone]
    aab = np.zeros((n, 6), np.float32)
    pri = np.zeros((n, 2), np.uint32)
    aab[: len(keep)] = W["aabbs"][f + np.array(keep)]
    pri[: len(keep)] = W["prims"][f + np.array(keep)]
    rem = np.zeros(n, np.uint32)
    for new, old in enumerate(keep):
        if new != old:
            rem[old] = new
    rem[gone] = INVALID
    world = acc.world()
    to_device(world.aabbs + 24 * f, aab)
    to_device(world.primitives + 8 * f, pri)
    to_device(world.remapped_primitives + 4 * f, rem)
    pf = _firsts(acc)
    acc.delete_primitives_from_gpu(i, len(gone))
    acc.update_scene()
    from accel_model import Voxel
    compacted = [Voxel(tuple(aab[k].tolist()), int(pri[k, 0]), int(pri[k, 1])) for k in range(len(keep))]
    m.update([("delete_from_gpu", i, compacted, len(gone))])
    inst2, W2 = _compare(acc, m)
    assert np.array_equal(W2["remap_prims"][f:f + n], rem)
    osc = _oracle(acc, sc, ctx)
    _blas_equal(acc, osc, i, len(keep))
    _trace_equal(acc, osc, sc, 31)
    acc.update_scene()  # settle clears the caller's entries too
    assert not acc.read_world()["remap_prims"].any()


def test_temporal_visibility_traces_the_previous_set(gpu):
    """get_previous_tlas while SETTLING: an update that deletes occluders
    keeps the scene from before it as the previous set; frame 2 (CT_REMAP_BIT)
    traces temporal reuse's target-at-previous queries there
    (restir_resampling.glsl:369 -> light.glsl:102) and matches the oracle
    given the pre-update scene, bit for bit, while an oracle that traces the
    current scene for them gives other reservoirs.  After the settle the
    previous set is the current one again."""
    w = _workload(frames=4, max_depth=1)
    w.camera = lambda f: ((0.2, 0.35, 2.2), (-0.09, -0.15, -0.98), (0.0, 1.0, 0.0))  # static: reuse everywhere
    sc, acc, ctx = _load(w)
    acc.update_scene()
    osc = _oracle(acc, sc, ctx)
    ofr, ofr_cur = O.OFrame(osc, w.width, w.height), O.OFrame(osc, w.width, w.height)
    for f in range(2):
        p = ctamd.frame_params(w, f)
        ctx.render(acc, p)
        ofr.render(p)
        ofr_cur.render(p)
        _frames_equal(ctx, ofr, w, f)
    t0, w0 = acc.tlas(), acc.world()
    assert acc.tlas(previous=True).nodes == t0.nodes  # idle: previous = current
    # delete a sixth of the voxels the camera does not see: occluders of shadow rays
    di = di_array(ctx.read(ctamd.CT_BUF_DI, w.width, w.height))
    seen = {(int(a), int(b)) for a, b in di[di[:, 3].view(np.float32) > 0][:, 11:13]}
    inst_before = acc.read_instances()
    prev_t = np.stack([np.array(r.transform[:], np.float32) for r in inst_before])
    prev_first = np.array([r.first_primitive_index for r in inst_before], np.uint32)
    rng = np.random.default_rng(4)
    n_del = 0
    for i, r in enumerate(inst_before):
        cand = [k for k in range(r.primitive_count) if (i, k) not in seen and rng.random() < 1 / 6]
        for k in sorted(cand, reverse=True)[: r.primitive_count - 1]:
            acc.delete_primitive(i, k)
            n_del += 1
    assert n_del > 1000
    acc.update_scene()
    assert acc.status() == ctamd.CT_AS_SETTLING
    tp, tc = acc.tlas(previous=True), acc.tlas()
    assert tp.nodes != tc.nodes and tp.generation == t0.generation and tc.generation != t0.generation
    wp = acc.world(previous=True)
    assert wp.instances != acc.world().instances and wp.cube_lights == acc.world().cube_lights
    W = acc.read_world()
    osc2 = _oracle(acc, sc, ctx)
    osc2.set_previous(prev_t, prev_first, W["remap_prims"], W["remap_lights"])
    ofr.set_scene(osc2)
    ofr.set_previous_scene(osc)
    ofr_cur.set_scene(osc2)
    p = ctamd.frame_params(w, 2)
    p.is_active |= ctamd.CT_REMAP_BIT
    ctx.render(acc, p)
    ofr.render(p)
    ofr_cur.render(p)
    _frames_equal(ctx, ofr, w, 2)
    hit = di_array(ofr.read(ctamd.CT_BUF_DI))[:, 3].view(np.float32) > 0
    a = res_array(ofr.read(ctamd.CT_BUF_INT_RESERVOIR))[hit]
    b = res_array(ofr_cur.read(ctamd.CT_BUF_INT_RESERVOIR))[hit]
    assert (a != b).any(axis=1).sum() > 0
    # settle: the previous set is the current one again
    assert acc.update_scene()
    assert acc.tlas(previous=True).nodes == acc.tlas().nodes
    osc3 = _oracle(acc, sc, ctx)
    ofr.set_scene(osc3)
    ofr.set_previous_scene(None)
    p = ctamd.frame_params(w, 3)
    ctx.render(acc, p)
    ofr.render(p)
    _frames_equal(ctx, ofr, w, 3)
