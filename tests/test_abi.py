"""The C-ABI library loads, exports every declared entry point, and fails
cleanly without a GPU.  No GPU compute here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import ctamd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "ct_abi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ct_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ctamd.lib()
    names = _declared()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match_reference_sizes():
    for name, (cls, size) in ctamd.STRUCT_SIZES.items():
        assert C.sizeof(cls) == size, name


def test_no_device_is_an_error_not_a_crash():
    if ctamd.device_available():
        return
    h = C.c_void_p()
    d = ctamd.AccelDesc(4, 1024, 4, 0)
    rc = ctamd.lib().ct_accel_create(C.byref(d), C.byref(h))
    assert rc == ctamd.CT_E_NODEVICE and not h.value
    assert b"gfx950" in ctamd.lib().ct_last_error_message()


def test_invalid_arguments_are_rejected():
    h = C.c_void_p()
    assert ctamd.lib().ct_accel_create(None, C.byref(h)) == ctamd.CT_E_INVALID
    d = ctamd.AccelDesc(0, 1024, 4, 0)
    assert ctamd.lib().ct_accel_create(C.byref(d), C.byref(h)) == ctamd.CT_E_INVALID
    assert ctamd.lib().ct_scene_generate(None, C.byref(h)) == ctamd.CT_E_INVALID


def test_camera_matrices():
    iv, ip = ctamd.camera_matrices((0, 0, 2.5), (0, 0, -1), (0, 1, 0), np.radians(60.0), 1.0)
    IV = iv.reshape(4, 4).T  # column-major -> row-major
    assert np.allclose(IV[:3, 3], [0, 0, 2.5])           # camera position
    assert np.allclose(IV[:3, 2], [0, 0, 1], atol=1e-6)   # view z axis = -forward
    IP = ip.reshape(4, 4).T
    # centre pixel direction: inv_proj * (0, 0, 1, 1) points down -z
    v = IP @ np.array([0, 0, 1, 1.0])
    assert v[2] < 0 and abs(v[0]) < 1e-6 and abs(v[1]) < 1e-6
    # y flip: top of NDC (-1) maps to +y in view space (main.cpp:1460)
    vt = IP @ np.array([0, -1, 1, 1.0])
    assert vt[1] / -vt[2] > 0 and np.isclose(vt[1] / -vt[2], np.tan(np.radians(30.0)), rtol=1e-5)
    pvp = ctamd.prev_view_proj(iv, ip)
    M = pvp.reshape(4, 4).T @ IV @ IP
    assert np.allclose(M, np.eye(4), atol=1e-4)


def test_scene_generator_counts_and_layout():
    w = ctamd.small_plain()
    sc = ctamd.Scene(ctamd.scene_desc(w))
    assert sc.info.primitive_count == 31457 == round(0.12 * 64 ** 3)
    assert sc.info.cube_light_count == 8 and sc.info.instance_count == 1
    a = sc.aabbs
    assert np.all(a[:, 3:] - a[:, :3] == np.float32(0.03125))
    g = np.round((a[:, :3] + 1.0) / 0.03125).astype(int)
    assert len({tuple(x) for x in g}) == a.shape[0]          # distinct cells
    assert g.min() >= 0 and g.max() <= 63
    lights = list(sc.lights)
    for L in lights:
        p = L.instance_info.primitive_id
        assert sc.primitives[p, 1] != 0xFFFFFFFF
        assert np.allclose(L.position[:], (a[p, :3] + a[p, 3:]) / 2)
        assert L.emissive[0] == 10.0 and L.size == np.float32(0.03125)
    sc2 = ctamd.Scene(ctamd.scene_desc(w))
    assert np.array_equal(sc.aabbs, sc2.aabbs)                # deterministic


def test_scene_generator_chunks():
    w = ctamd.hd_di(width=16, height=16, grid=96, chunk=32)
    sc = ctamd.Scene(ctamd.scene_desc(w))
    assert sc.info.instance_count == 27
    assert sc.info.primitive_count == round(0.25 * 96 ** 3)
    first = [i.first_primitive_index for i in sc.instances]
    counts = [i.primitive_count for i in sc.instances]
    assert first[0] == 0 and all(first[k] + counts[k] == first[k + 1] for k in range(26))
    for k, inst in enumerate(sc.instances):
        a = sc.aabbs[inst.first_primitive_index: inst.first_primitive_index + inst.primitive_count]
        cell = np.floor((a[:, :3] + 1.5) / 0.03125 / 32 + 1e-6).astype(int)
        assert len({tuple(x) for x in cell}) == 1             # one chunk per instance


def _light_voxel(sc, L):
    inst = sc.instances[L.instance_info.instance_id]
    assert L.instance_info.primitive_id < inst.primitive_count
    return inst.first_primitive_index + L.instance_info.primitive_id


def test_scene_generator_palette():
    """The palette workload's scene at 128^3: 256 shared materials, one entry
    per voxel, 3 emissive entries, and exactly the voxels of those entries
    are cube lights, in voxel order, with the entry's emission."""
    w = ctamd.palette_pt(grid=128, chunk=32, width=16, height=16)
    assert w.palette == (256, 3)
    sc = ctamd.Scene(ctamd.scene_desc(w))
    assert sc.info.primitive_count == round(0.12 * 128 ** 3) and sc.info.material_count == 256
    mats = sc.materials
    emi = np.array([m.emission[:] for m in mats], np.float32)
    dif = np.array([m.diffuse[:] for m in mats], np.float32)
    emissive_entries = np.nonzero(emi.max(axis=1) > 0)[0]
    assert len(emissive_entries) == 3
    assert np.all((dif >= 0.2) & (dif <= 0.8)) and np.all((emi[emissive_entries] >= 1) & (emi[emissive_entries] <= 20))
    entry = sc.primitives[:, 0]
    assert entry.max() < 256 and len(np.unique(entry)) == 256
    lit = np.isin(entry, emissive_entries)
    assert np.array_equal(np.nonzero(lit)[0], np.nonzero(sc.primitives[:, 1] != 0xFFFFFFFF)[0])
    assert sc.info.cube_light_count == lit.sum() > 1000
    a = sc.aabbs
    vox = [_light_voxel(sc, L) for L in sc.lights]
    assert vox == sorted(vox) and np.array_equal(vox, np.nonzero(lit)[0])
    for j, L in enumerate(sc.lights):
        v = vox[j]
        assert sc.primitives[v, 1] == j
        assert np.array_equal(np.float32(L.emissive[:]), emi[entry[v]])
        assert np.allclose(L.position[:], (a[v, :3] + a[v, 3:]) / 2)


def test_scene_generator_light_ids_across_chunks():
    """Per-voxel materials over several chunks: each light names the
    instance and primitive of its voxel."""
    w = ctamd.hd_di(width=16, height=16, grid=96, chunk=32)
    sc = ctamd.Scene(ctamd.scene_desc(w))
    a = sc.aabbs
    ids = {L.instance_info.instance_id for L in sc.lights}
    assert len(ids) > 10
    for j, L in enumerate(sc.lights):
        v = _light_voxel(sc, L)
        assert sc.primitives[v, 1] == j and sc.primitives[v, 0] == v
        assert np.allclose(L.position[:], (a[v, :3] + a[v, 3:]) / 2)


@pytest.mark.parametrize("count", [1, 7, 12 ** 3])
def test_scene_generator_occupancy_extremes(count):
    d = ctamd.scene_desc(ctamd.small_plain())
    d.grid[:] = [12, 12, 12]
    d.chunk[:] = [5, 5, 5]
    d.occupied_count, d.emissive_count = count, 1
    sc = ctamd.Scene(d)
    assert sc.info.primitive_count == count
    g = np.round(sc.aabbs[:, :3] / 0.03125).astype(int)
    assert len({tuple(x) for x in g}) == count


def test_scene_generator_rejects_bad_palettes():
    h = C.c_void_p()
    for size, emissive in ((257, 3), (4, 5)):
        d = ctamd.scene_desc(ctamd.small_plain())
        d.palette_size, d.palette_emissive = size, emissive
        assert ctamd.lib().ct_scene_generate(C.byref(d), C.byref(h)) == ctamd.CT_E_INVALID and not h.value
