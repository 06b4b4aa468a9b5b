"""The oracle of the accel manager's task semantics against what fixes them
independently of its own code: the reference's swap-with-last rule for
primitive deletion (accel_struct_mngr.cpp:664-669), consistency of every
back-reference between voxels and lights, the remapping tables read as a
permutation of the surviving voxels, exact round trips through undo, and
closed forms of the UPDATE_BLAS transform product."""
import math

import numpy as np
import pytest

from accel_model import INVALID, AccelModel, Instance, LightRec, Voxel, mat4_mul


def _scene(rng, counts=(40, 33, 64), light_frac=0.25):
    m = AccelModel()
    ident = np.eye(4, dtype=np.float32).reshape(16)
    uid = 0
    for i, n in enumerate(counts):
        vox = []
        for k in range(n):
            lo = np.array([uid % 97, uid // 97, i], np.float32)  # unique box per voxel
            uid += 1
            light = INVALID
            if rng.random() < light_frac:
                light = len(m.lights)
                m.lights.append(LightRec(bytes([len(m.lights) % 251]) * 40, i, k))
            vox.append(Voxel(tuple((lo.tolist() + (lo + 1).tolist())), int(rng.integers(0, 8)), light))
        m.instances[i] = Instance(ident.copy(), vox)
        m.prev_transform[i] = ident.copy()
    return m


def _check_links(m):
    """Every light names a voxel that names it back, and vice versa."""
    for l, rec in enumerate(m.lights):
        assert m.instances[rec.instance].voxels[rec.primitive].light == l
    n = 0
    for i, inst in m.instances.items():
        for k, v in enumerate(inst.voxels):
            if v.light != INVALID:
                n += 1
                assert m.lights[v.light].instance == i and m.lights[v.light].primitive == k
    assert n == len(m.lights)


def _snapshot(m):
    return ({i: [(v.aabb, v.material) for v in inst.voxels] for i, inst in m.instances.items()},
            sorted(r.data for r in m.lights))


def test_delete_moves_the_last_voxel_into_the_hole():
    m = _scene(np.random.default_rng(1))
    before = list(m.instances[0].voxels)
    m.update([("delete_primitive", 0, 5)])
    after = list(m.instances[0].voxels)
    assert len(after) == len(before) - 1
    assert after[5] is before[-1]
    assert all(after[k] is before[k] for k in range(len(after)) if k != 5)
    # deleting the last voxel moves nothing
    m.update([("delete_primitive", 0, len(after) - 1)])
    assert m.instances[0].voxels == after[:-1]


@pytest.mark.parametrize("seed", range(6))
def test_random_deletions_keep_links_and_multisets(seed):
    rng = np.random.default_rng(100 + seed)
    m = _scene(rng)
    _check_links(m)
    voxels0 = {i: {v.aabb for v in inst.voxels} for i, inst in m.instances.items()}
    lights0 = sorted(r.data for r in m.lights)
    deleted, dead_lights = {i: set() for i in m.instances}, []
    for step in range(25):
        tasks = []
        for _ in range(int(rng.integers(1, 4))):
            i = int(rng.integers(0, 3))
            n = len(m.instances[i].voxels) - sum(1 for t in tasks if t[1] == i)
            if n <= 1:
                continue
            d = int(rng.integers(0, n))
            tasks.append(("delete_primitive", i, d))
        # record what each task deletes as it is applied
        for t in tasks:
            v = m.instances[t[1]].voxels[t[2]]
            deleted[t[1]].add(v.aabb)
            if v.light != INVALID:
                dead_lights.append(m.lights[v.light].data)
            m._delete_primitive(t[1], t[2])
        m.settling = True
        _check_links(m)
        for i, inst in m.instances.items():
            assert {v.aabb for v in inst.voxels} == voxels0[i] - deleted[i]
            assert len({v.aabb for v in inst.voxels}) == len(inst.voxels)
        left = list(lights0)
        for d in dead_lights:
            left.remove(d)
        assert sorted(r.data for r in m.lights) == left


def test_remapping_tables_follow_every_surviving_voxel():
    """Read the primitive table as the shader does (0 = unchanged, INVALID =
    deleted, else the new index): it must send each voxel's previous index
    to its current one; likewise for lights."""
    rng = np.random.default_rng(7)
    m = _scene(rng)
    for rnd in range(12):
        prev = {i: [v.aabb for v in inst.voxels] for i, inst in m.instances.items()}
        prev_lights = [r.data for r in m.lights]
        tasks = []
        for _ in range(3):
            i = int(rng.integers(0, 3))
            n = len(m.instances[i].voxels) - sum(1 for t in tasks if t[1] == i)
            if n > 1:
                tasks.append(("delete_primitive", i, int(rng.integers(0, n))))
        m.update(tasks)
        for i, inst in m.instances.items():
            now = {v.aabb: k for k, v in enumerate(inst.voxels)}
            for p, box in enumerate(prev[i]):
                e = m.remap_prims.get((i, p), 0)
                if box not in now:
                    assert e == INVALID
                elif e == 0:
                    # unchanged index, or a move to index 0 the 0 encoding cannot tell apart
                    assert now[box] in (p, 0)
                else:
                    assert e == now[box]
        assert len(set(prev_lights)) == len(prev_lights)  # records are unique here
        now_l = {r.data: k for k, r in enumerate(m.lights)}
        for l, data in enumerate(prev_lights):
            e = m.remap_lights.get(l, 0)
            if data not in now_l:
                assert e == INVALID
            elif e == 0:
                assert now_l[data] in (l, 0)
            else:
                assert e == now_l[data]


def test_settle_clears_the_tables():
    m = _scene(np.random.default_rng(3))
    m.update([("delete_primitive", 1, 2)])
    assert m.remap_prims
    m.update([])
    assert not m.remap_prims and not m.remap_lights


@pytest.mark.parametrize("seed", range(4))
def test_undo_restores_exactly(seed):
    rng = np.random.default_rng(200 + seed)
    m = _scene(rng)
    snaps = [_snapshot(m)]
    order = []
    for _ in range(6):
        i = int(rng.integers(0, 3))
        d = int(rng.integers(0, len(m.instances[i].voxels)))
        m.update([("delete_primitive", i, d)])
        snaps.append(_snapshot(m))
        order.append(i)
    for k in range(6):
        prev_layout = {i: [v.aabb for v in inst.voxels] for i, inst in m.instances.items()}
        m.update([("undo",)])
        assert _snapshot(m) == snaps[-2 - k]
        _check_links(m)
        # the undo's remapping sends the moved voxel back
        for (i, p), e in m.remap_prims.items():
            assert m.instances[i].voxels[e].aabb == prev_layout[i][p]
    exact = [(v.aabb, v.material, v.light) for v in m.instances[0].voxels]
    assert len(exact) == len(snaps[0][0][0])
    m.update([("undo",)])  # empty stack: no effect
    assert _snapshot(m) == snaps[0]


def test_undo_of_other_kinds_pops_without_effect():
    m = _scene(np.random.default_rng(4))
    m.update([("delete_primitive", 0, 0)])
    m.update([("update_blas", 1, np.eye(4, dtype=np.float32).reshape(16), [(0, (9, 9, 9, 10, 10, 10))])])
    s = _snapshot(m)
    m.update([("undo",)])        # pops the update: nothing changes
    assert _snapshot(m) == s
    m.update([("undo",)])        # then the deletion is undone
    assert len(m.instances[0].voxels) == 40


def _translation(x, y, z):
    t = np.eye(4, dtype=np.float32)
    t[:3, 3] = (x, y, z)
    return t.T.reshape(16)  # column-major


def _rot_y(a):
    c, s = math.cos(a), math.sin(a)
    r = np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], np.float32)
    return r.T.reshape(16)


def test_update_transform_is_the_column_major_product():
    a = mat4_mul(_translation(1.0, 2.0, 3.0), _translation(0.5, -4.0, 8.0))
    assert np.array_equal(a, _translation(1.5, -2.0, 11.0))       # exact in float32
    r = mat4_mul(_rot_y(0.3), _rot_y(0.5))
    assert np.allclose(r, _rot_y(0.8), atol=1e-6)
    # order: instance transform first, task transform on the right (object space)
    tr = mat4_mul(_translation(5, 0, 0), _rot_y(math.pi / 2))
    p = tr.reshape(4, 4).T @ np.array([1, 0, 0, 1], np.float32)
    assert np.allclose(p[:3], [5, 0, -1], atol=1e-6)
    m = _scene(np.random.default_rng(5))
    m.update([("update_blas", 2, _translation(0, 1, 0), [(3, (7, 7, 7, 8, 8, 8)), (999, (0, 0, 0, 1, 1, 1))])])
    assert m.instances[2].voxels[3].aabb == (7.0, 7.0, 7.0, 8.0, 8.0, 8.0)
    assert np.array_equal(m.prev_transform[2], np.eye(4, dtype=np.float32).reshape(16))
    m.update([])
    assert np.array_equal(m.prev_transform[2], _translation(0, 1, 0))


def _check_light_remap(m, prev_lights):
    now_l = {r.data: k for k, r in enumerate(m.lights)}
    for l, data in enumerate(prev_lights):
        e = m.remap_lights.get(l, 0)
        if data not in now_l:
            assert e == INVALID, l
        elif e == 0:
            assert now_l[data] in (l, 0), l
        else:
            assert e == now_l[data], l


def _check_prim_remap(m, prev):
    for i, inst in m.instances.items():
        now = {v.aabb: k for k, v in enumerate(inst.voxels)}
        for p, box in enumerate(prev.get(i, [])):
            e = m.remap_prims.get((i, p), 0)
            if box not in now:
                assert e == INVALID, (i, p)
            elif e == 0:
                assert now[box] in (p, 0), (i, p)
            else:
                assert e == now[box], (i, p)


def _unique_light_scene(seed):
    m = _scene(np.random.default_rng(seed), light_frac=0.5)
    for k, r in enumerate(m.lights):  # unique records, so a light can be followed by its bytes
        r.data = k.to_bytes(4, "little") * 10
    return m


def test_remapping_composes_repeated_deletions_of_one_slot():
    """Deleting the voxel that the previous deletion of the same update moved
    into the hole: the first moved voxel is gone too, and its remapping entry
    must say so (entries are composed over the update, not written per
    task)."""
    m = _unique_light_scene(11)
    lit = [k for k, v in enumerate(m.instances[2].voxels) if v.light != INVALID]
    for d in (5, lit[0]):
        prev = {i: [v.aabb for v in inst.voxels] for i, inst in m.instances.items()}
        prev_lights = [r.data for r in m.lights]
        m.update([("delete_primitive", 2, d), ("delete_primitive", 2, d), ("delete_primitive", 2, d)])
        _check_links(m)
        _check_prim_remap(m, prev)
        _check_light_remap(m, prev_lights)


def test_remapping_composes_gpu_and_cpu_deletions():
    """GPU deletions of an update are applied first (their compaction already
    ran); CPU deletions in the same update then move lights that the GPU
    removal had moved: every previous light index maps to where its record
    is now."""
    rng = np.random.default_rng(12)
    for trial in range(20):
        m = _unique_light_scene(300 + trial)
        prev_lights = [r.data for r in m.lights]
        vox = m.instances[0].voxels
        keep = [v for k, v in enumerate(vox) if rng.random() < 0.7]
        dropped = len(vox) - len(keep)
        compacted = keep + [v for v in vox if v not in keep]
        tasks = [("delete_from_gpu", 0, compacted, dropped)]
        for _ in range(3):
            i = int(rng.integers(1, 3))
            n = len(m.instances[i].voxels) - sum(1 for t in tasks if t[0] == "delete_primitive" and t[1] == i)
            tasks.append(("delete_primitive", i, int(rng.integers(0, n))))
        m.update(tasks)
        _check_links(m)
        _check_light_remap(m, prev_lights)


def test_undo_refused_after_a_gpu_deletion_removed_lights():
    """A CPU deletion's undo record names light indices; a GPU deletion
    that removes lights in between changes them, so the undo is refused
    (CT_E_STATE in the library) and nothing changes."""
    m = _unique_light_scene(13)
    lit = [k for k, v in enumerate(m.instances[1].voxels) if v.light != INVALID]
    m.update([("delete_primitive", 1, lit[0])])
    vox = m.instances[0].voxels
    lit0 = [v for v in vox if v.light != INVALID]
    keep = [v for v in vox if v is not lit0[0]]
    m.update([("delete_from_gpu", 0, keep + [lit0[0]], 1)])
    before = _snapshot(m)
    refused = m.update([("undo",)])       # pops the GPU deletion: nothing to undo
    assert not refused and _snapshot(m) == before
    refused = m.update([("undo",)])       # the CPU deletion: its light list has changed
    assert len(refused) == 1 and "light list" in str(refused[0])
    assert _snapshot(m) == before
    _check_links(m)


def test_undo_still_applies_when_the_light_list_is_as_left():
    m = _unique_light_scene(14)
    snap = _snapshot(m)
    lit = [k for k, v in enumerate(m.instances[1].voxels) if v.light != INVALID]
    m.update([("delete_primitive", 1, lit[0])])
    m.update([("update_blas", 2, np.eye(4, dtype=np.float32).reshape(16), [])])
    m.update([("undo",)])
    assert m.update([("undo",)]) == []
    assert _snapshot(m) == snap
    _check_links(m)
