"""GPU voxel deletion: marks from the brush (squared_brush.slang) or from a
list, the device compaction (rearrangement.slang) and the
DELETE_PRIMITIVE_BLAS_FROM_GPU tasks (accel_struct_mngr.cpp:2717-2871),
against the brush oracle and the task-semantics model: records, remapping
entries, rebuilt BLASes and traversal."""
import numpy as np
import pytest

import brush_model as B
import ctamd
from ctamd.voxfile import monument
from helpers import bits, di_array
from test_gpu_accel_update import _blas_equal, _compare, _load, _model, _oracle, _trace_equal, _workload

pytestmark = pytest.mark.gpu


def _vox_scene():
    """A light-free .vox scene cut into 16^3 regions (more than 32 instances)."""
    sc = ctamd.Scene.from_vox(monument(n=64, seed=12, emissive_colors=()), chunk=(16, 16, 16))
    ni = sc.info.instance_count
    acc = ctamd.Accel(max_instances=ni + 8, max_primitives=sc.info.primitive_count + 32 * ni + 1024,
                      max_cube_lights=8)
    ctx = ctamd.Ctx(32, 32)
    ctamd.load_scene(sc, acc, ctx)
    return sc, acc, ctx


def _check_compaction(acc, m, plan, before):
    """Device records right after processing, then the model after update_scene."""
    W = acc.read_world()
    rms = {}
    for i, dele in plan.items():
        f, n = before[i]
        moves, keep, remap = B.rearrange(n, dele)
        exp = B.compacted_records(list(range(n)), n, dele)
        assert np.array_equal(bits(W["aabbs"][f:f + keep]), bits(before["aabbs"][f + np.array(exp, int)])), i
        assert np.array_equal(W["prims"][f:f + keep], before["prims"][f + np.array(exp, int)]), i
        rm = np.zeros(n, np.uint32)
        for k, v in remap.items():
            rm[k] = v
        assert np.array_equal(W["remap_prims"][f:f + n], rm), i
        rms[i] = (f, n, keep, rm)
    acc.update_scene()
    # the compaction's remapping entries survive the update that applies it
    # (a kept instance's slots take no other entries)
    W = acc.read_world()
    for i, (f, n, keep, rm) in rms.items():
        if keep:
            assert np.array_equal(W["remap_prims"][f:f + n], rm), i
    m.update([("delete_from_gpu", i, B.compacted_records(m.instances[i].voxels, before[i][1], plan[i]),
               len(set(plan[i]))) for i in sorted(plan)])
    _compare(acc, m)


def _snapshot(acc):
    inst = acc.read_instances()
    W = acc.read_world()
    snap = {i: (r.first_primitive_index, r.primitive_count) for i, r in enumerate(inst) if r.primitive_count}
    snap["aabbs"], snap["prims"] = W["aabbs"].copy(), W["prims"].copy()
    return snap


def test_marked_voxels_compact_and_rebuild(gpu):
    sc, acc, ctx = _vox_scene()
    acc.update_scene()  # settle: the accel is idle
    m = _model(acc)
    before = _snapshot(acc)
    ids = sorted(k for k in before if isinstance(k, int))
    assert len(ids) > 40
    rng = np.random.default_rng(17)
    big = [i for i in ids if before[i][1] > 40]
    plan = {}
    plan[big[0]] = rng.choice(before[big[0]][1], before[big[0]][1] // 5, replace=False).tolist()   # 20 %
    hi = [i for i in big if i >= 32]
    plan[hi[0]] = [int(rng.integers(0, before[hi[0]][1]))]                                          # one voxel, id >= 32
    n = before[hi[1]][1]
    plan[hi[1]] = list(range(n - 5, n))                                                              # tail only
    plan[hi[2]] = rng.choice(before[hi[2]][1], before[hi[2]][1] // 2, replace=False).tolist()       # half
    plan[ids[-1]] = list(range(before[ids[-1]][1]))                                                  # all: deleted
    marks = [(i, p) for i, d in plan.items() for p in d]
    order = rng.permutation(len(marks))
    marks = [marks[k] for k in order]
    acc.brush_mark(marks[: len(marks) // 2])
    acc.brush_mark(marks[len(marks) // 2:] + marks[:7])  # repeated marks count once
    assert acc.brush_counters() == (len(plan), len(marks))
    assert acc.process_voxel_modifications() == (len(plan), len(marks))
    assert acc.brush_counters() == (0, 0)
    _check_compaction(acc, m, plan, before)
    assert acc.read_instances()[ids[-1]].primitive_count == 0
    osc = _oracle(acc, sc, ctx)
    for i in plan:
        if i != ids[-1]:
            _blas_equal(acc, osc, i, before[i][1] - len(plan[i]))
    _trace_equal(acc, osc, sc, 23)
    acc.update_scene()  # settle clears the remapping entries
    assert not acc.read_world()["remap_prims"].any()
    # The host mirror took the compaction's moves without reading the slots
    # back: CPU deletions in the compacted instances move the records it
    # holds at the ends of the kept ranges (six per instance, index 0 each
    # time, so six tails are read).
    ops = [("delete_primitive", i, 0) for i in sorted(plan) if i != ids[-1] for _ in range(6)]
    for _, i, p in ops:
        acc.delete_primitive(i, p)
    acc.update_scene()
    m.update(ops)
    _compare(acc, m)


def test_brush_stroke_on_a_frame(gpu):
    w = _workload(width=64, height=40)
    sc, acc, ctx = _load(w)
    acc.update_scene()
    ctx.render(acc, ctamd.frame_params(w, 0))
    m = _model(acc)
    before = _snapshot(acc)
    di = di_array(ctx.read(ctamd.CT_BUF_PREV_DI, w.width, w.height))
    dist, hit = di[:, 3].view(np.float32), di[:, 11:13]  # distance; instance_hit (instance, primitive)
    # a cursor away from the edges over a voxel, and one whose window wraps at the left edge
    hits = [(x, y) for y in range(8, 32) for x in range(8, 56) if dist[y * w.width + x] > 0]
    x, y = hits[len(hits) // 2]
    ctamd.brush_apply(ctx, acc, 1, y, 3)
    assert acc.brush_counters() == (0, 0)
    expected = B.brush_marks(dist, hit, w.width, w.height, x, y, 3)
    assert expected
    ctamd.brush_apply(ctx, acc, x, y, 3)
    assert acc.brush_counters() == (1, len(expected))
    plan = {}
    for i, p in expected:
        plan.setdefault(i, []).append(p)
    assert acc.process_voxel_modifications() == (1, len(expected))
    _check_compaction(acc, m, plan, before)
    osc = _oracle(acc, sc, ctx)
    (i,) = plan
    _blas_equal(acc, osc, i, before[i][1] - len(plan[i]))
    _trace_equal(acc, osc, sc, 29)


def test_processing_waits_for_an_idle_accel(gpu):
    sc, acc, ctx = _vox_scene()
    # right after the load the accel is settling: marks stay until it is idle
    acc.brush_mark([(0, 0), (0, 1)])
    assert acc.status() != ctamd.CT_AS_IDLE
    assert acc.process_voxel_modifications() == (0, 0)
    assert acc.brush_counters() == (1, 2)
    acc.update_scene()
    assert acc.process_voxel_modifications() == (1, 2)


def test_gpu_deletion_removes_the_lights_of_deleted_voxels(gpu):
    """Lights of deleted voxels leave the light list (last light into the
    hole, owner relinked), lights of moved voxels follow them, and the light
    remapping maps original indices to final ones -- as the task model does."""
    w = _workload()
    sc, acc, ctx = _load(w)
    acc.update_scene()
    m = _model(acc)
    before = _snapshot(acc)
    W = acc.read_world()
    n_lights = W["light_count"]
    ids = sorted(k for k in before if isinstance(k, int))
    lit = {i: [k for k in range(before[i][1]) if W["prims"][before[i][0] + k, 1] != 0xFFFFFFFF] for i in ids}
    a, b = [i for i in ids if len(lit[i]) >= 3][:2]
    dark_a = [k for k in range(before[a][1]) if k not in lit[a]]
    plan = {a: lit[a][:2] + dark_a[:3] + [before[a][1] - 1], b: [lit[b][-1]]}
    acc.brush_mark([(i, p) for i, d in plan.items() for p in d])
    assert acc.process_voxel_modifications()[0] == 2
    _check_compaction(acc, m, plan, before)
    W2 = acc.read_world()
    gone = {int(W["prims"][before[i][0] + p, 1]) for i, d in plan.items() for p in d} - {0xFFFFFFFF}
    assert W2["light_count"] == n_lights - len(gone) == len(m.lights)
    assert np.array_equal(W2["remap_lights"], m.remap_light_table(len(W2["remap_lights"])))
    # every light's owner points back at it
    inst = acc.read_instances()
    for l in range(W2["light_count"]):
        o = W2["lights"][l].instance_info
        assert W2["prims"][inst[o.instance_id].first_primitive_index + o.primitive_id, 1] == l
    osc = _oracle(acc, sc, ctx)
    _trace_equal(acc, osc, sc, 37)


def test_undo_after_a_gpu_deletion_that_removed_lights(gpu):
    """A CPU deletion of a lit voxel, then a GPU deletion that removes lights,
    then two undos: the first pops the GPU deletion and changes nothing, the
    second is refused with CT_E_STATE because the light list is no longer
    the one the CPU deletion left; records stay equal to the task model's.
    Also: a load while deletions are queued is refused."""
    w = _workload()
    sc, acc, ctx = _load(w)
    acc.update_scene()
    m = _model(acc)
    W = acc.read_world()
    before = _snapshot(acc)
    ids = sorted(k for k in before if isinstance(k, int))
    lit = {i: [k for k in range(before[i][1]) if W["prims"][before[i][0] + k, 1] != 0xFFFFFFFF] for i in ids}
    a, b = [i for i in ids if len(lit[i]) >= 3][:2]
    acc.delete_primitive(b, lit[b][0])
    acc.update_scene()
    m.update([("delete_primitive", b, lit[b][0])])
    _compare(acc, m)
    acc.update_scene()  # settle
    m.update([])
    before = _snapshot(acc)
    plan = {a: lit[a][:2]}
    acc.brush_mark([(a, p) for p in plan[a]])
    assert acc.process_voxel_modifications()[0] == 1
    with pytest.raises(ctamd.CtError) as e:
        ctamd.load_scene(sc, acc, ctx)  # a GPU deletion is queued
    assert e.value.code == ctamd.CT_E_STATE
    _check_compaction(acc, m, plan, before)
    acc.update_scene()  # settle
    m.update([])
    acc.undo()  # pops the GPU deletion
    acc.update_scene()
    assert m.update([("undo",)]) == []
    _compare(acc, m)
    acc.update_scene()
    m.update([])
    acc.undo()
    with pytest.raises(ctamd.CtError) as e:
        acc.update_scene()
    assert e.value.code == ctamd.CT_E_STATE
    refused = m.update([("undo",)])
    assert len(refused) == 1
    _compare(acc, m)
    osc = _oracle(acc, sc, ctx)
    _trace_equal(acc, osc, sc, 41)
