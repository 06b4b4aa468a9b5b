"""MagicaVoxel loader (ct_scene_from_vox): closed-form cases written from
known voxel lists, the library against the numpy oracle on procedural and
(when present) the reference's own model files, and malformed input."""
import os

import numpy as np
import pytest

import ctamd
import vox_model as V
from helpers import bits
from ctamd.voxfile import write_vox, monument

E = np.float32(ctamd.CT_VOXEL_EXTENT)
PAL = [((7 * i) % 256, (13 * i + 1) % 256, (29 * i + 2) % 256, 255) for i in range(256)]


def lib_arrays(sc):
    info = sc.info
    inst = np.array([(i.first_primitive_index, i.primitive_count) for i in sc.instances], np.uint32).reshape(-1, 2)
    mats = list(sc.materials)
    lights = list(sc.lights)
    return {"instances": inst, "aabbs": sc.aabbs, "prims": sc.primitives,
            "diffuse": np.array([m.diffuse[:] for m in mats], np.float32).reshape(-1, 3),
            "emission": np.array([m.emission[:] for m in mats], np.float32).reshape(-1, 3),
            "light_pos": np.array([l.position[:] for l in lights], np.float32).reshape(-1, 3),
            "light_emissive": np.array([l.emissive[:] for l in lights], np.float32).reshape(-1, 3),
            "light_ids": np.array([(l.instance_info.instance_id, l.instance_info.primitive_id) for l in lights],
                                  np.uint32).reshape(-1, 2),
            "count": (info.instance_count, info.primitive_count, info.material_count, info.cube_light_count)}


def assert_same(g, o):
    for k in o:
        a, b = np.asarray(g[k]), np.asarray(o[k])
        assert a.shape == b.shape, (k, a.shape, b.shape)
        if a.dtype == np.float32:
            assert np.array_equal(bits(a), bits(b)), k
        else:
            assert np.array_equal(a, b), k


def test_single_voxel_closed_form():
    data = write_vox([((4, 5, 6), [(1, 2, 3, 7)])], palette=PAL)
    # every axis direction, map_loader.cpp:226-262
    expect = {0: (2, 3, 1), 1: (3, 1, 2), 2: (1, 2, 3), 3: (2, 3, 1), 4: (3, 1, 2), 5: (1, 2, 3)}
    for axis, g in expect.items():
        sc = ctamd.Scene.from_vox(data, axis=axis)
        assert (sc.info.instance_count, sc.info.primitive_count, sc.info.material_count,
                sc.info.cube_light_count) == (1, 1, 1, 0)
        gf = np.array(g, np.float32)
        assert np.array_equal(sc.aabbs[0], np.concatenate([gf * E, (gf + 1) * E]))
        m = sc.materials[0]
        # color index 7 is RGBA entry 6
        assert np.array_equal(np.array(m.diffuse[:], np.float32),
                              np.array(PAL[6][:3], np.float32) / np.float32(255))
        assert m.type == 0 and m.emission[:] == [0.0, 0.0, 0.0]
        assert (m.shininess, m.roughness, m.ior, m.dissolve, m.illum) == (1.0, 1.0, 1.0, 1.0, 3)
        assert tuple(sc.primitives[0]) == (0, ctamd.CT_INVALID_INDEX)
        assert tuple(sc.instances[0].transform[:]) == tuple(np.eye(4, dtype=np.float32).reshape(16))


def test_emissive_voxel_is_a_cube_light():
    vox = [(0, 0, 0, 5), (1, 0, 0, 9), (2, 0, 0, 5)]
    data = write_vox([((3, 1, 1), vox)], palette=PAL, matl={5: {"_type": "_emit", "_emit": "0.5"},
                                                            9: {"_type": "_diffuse"}}, version=200)
    sc = ctamd.Scene.from_vox(data, axis=ctamd.CT_AXIS_Z_BOTTOM_TOP)
    assert sc.info.cube_light_count == 2 and sc.info.material_count == 2
    col = np.array(PAL[4][:3], np.float32) / np.float32(255)
    for k, (px, L) in enumerate(zip((0, 2), sc.lights)):
        assert np.array_equal(np.array(L.position[:], np.float32),
                              np.array([px * E + E * np.float32(0.5), E * np.float32(0.5), E * np.float32(0.5)],
                                       np.float32))
        assert np.array_equal(np.array(L.emissive[:], np.float32), col)
        assert L.size == E and L.type == ctamd.CT_LIGHT_CUBE
        assert (L.instance_info.instance_id, L.instance_info.primitive_id) == (0, px)
    assert sc.primitives.tolist() == [[0, 0], [1, ctamd.CT_INVALID_INDEX], [0, 1]]
    assert np.array_equal(np.array(sc.materials[0].emission[:], np.float32), col)
    # flux scales emission (map_loader.cpp:150-156)
    sc2 = ctamd.Scene.from_vox(data, axis=2, flux=4.0)
    assert np.array_equal(np.array(sc2.lights[0].emissive[:], np.float32), col * np.float32(4))


def test_matt_emissive_and_default_palette():
    data = write_vox([((2, 1, 1), [(0, 0, 0, 1), (1, 0, 0, 2)])], matt={2: (3, 0.7)})
    sc = ctamd.Scene.from_vox(data, axis=2)
    # default palette: index 1 = ffffff, index 2 = (ff, ff, cc)
    assert sc.materials[0].diffuse[:] == [1.0, 1.0, 1.0]
    assert np.array_equal(np.array(sc.materials[1].diffuse[:], np.float32),
                          np.array([255, 255, 204], np.float32) / np.float32(255))
    assert sc.info.cube_light_count == 1 and sc.lights[0].instance_info.primitive_id == 1
    pal = V.default_palette()
    assert pal.shape == (256, 3) and tuple(pal[215]) == (0x00, 0x00, 0x33) and tuple(pal[216]) == (0xEE, 0, 0)
    assert tuple(pal[255]) == (0x11, 0x11, 0x11) and len({tuple(c) for c in pal}) == 256


def test_order_dedup_and_regions():
    rng = np.random.default_rng(3)
    cells = [(x, y, z) for z in range(4) for y in range(4) for x in range(4) if (x + y + z) % 2 == 0]
    vox = [(x, y, z, 1 + (x + 2 * z) % 3) for x, y, z in cells]
    order = rng.permutation(len(vox))
    shuffled = [vox[i] for i in order]
    # a cell written twice keeps the later entry
    data = write_vox([((4, 4, 4), [(0, 0, 0, 9)] + shuffled)], palette=PAL)
    sc = ctamd.Scene.from_vox(data, axis=2)
    assert sc.info.instance_count == 1 and sc.info.primitive_count == len(cells)
    lo = (sc.aabbs[:, :3] / E).astype(int)
    assert [tuple(p) for p in lo] == cells  # z, y, x order
    assert sc.info.material_count == 3  # one material per palette index in the region
    # 2^3 chunks of 2^3: one instance per non-empty chunk, materials per chunk
    sc = ctamd.Scene.from_vox(data, axis=2, chunk=(2, 2, 2))
    assert sc.info.instance_count == 8
    inst = [(i.first_primitive_index, i.primitive_count) for i in sc.instances]
    assert inst[0][0] == 0 and all(a[0] + a[1] == b[0] for a, b in zip(inst, inst[1:]))
    assert sum(c for _, c in inst) == len(cells)
    assert sc.info.material_count > 3
    assert_same(lib_arrays(sc), V.load(data, axis=2, chunk=(2, 2, 2)))


def test_scene_graph_translation_and_models():
    a = ((4, 4, 2), [(0, 0, 0, 1), (3, 3, 1, 2)])
    b = ((3, 3, 3), [(1, 1, 1, 3)])
    data = write_vox([a, b], palette=PAL, graph=[((10, 0, -5), [0], None), ((0, 7, 0), [1], 4)], version=200)
    sc = ctamd.Scene.from_vox(data, axis=2)
    assert sc.info.instance_count == 2
    lo = (sc.aabbs[:, :3] / E).astype(int).tolist()
    # v - floor(s/2) + t
    assert lo == [[0 - 2 + 10, 0 - 2, 0 - 1 - 5], [3 - 2 + 10, 3 - 2, 1 - 1 - 5], [1 - 1, 1 - 1 + 7, 1 - 1]]
    assert_same(lib_arrays(sc), V.load(data, axis=2))
    # one model alone sits at the origin
    sc1 = ctamd.Scene.from_vox(data, axis=2, model=1)
    assert (sc1.aabbs[0, :3] / E).astype(int).tolist() == [1, 1, 1]
    # without a scene graph both models sit at the origin, in file order
    sc2 = ctamd.Scene.from_vox(write_vox([a, b], palette=PAL), axis=2)
    assert (sc2.aabbs[:, :3] / E).astype(int).tolist() == [[0, 0, 0], [3, 3, 1], [1, 1, 1]]
    with pytest.raises(ctamd.CtError) as ei:
        ctamd.Scene.from_vox(write_vox([a], palette=PAL, graph=[((0, 0, 0), [0], 17)], version=200))
    assert ei.value.code == ctamd.CT_E_UNSUPPORTED


@pytest.mark.parametrize("axis,chunk", [(0, (0, 0, 0)), (0, (16, 16, 16)), (1, (32, 8, 64)), (2, (7, 11, 13))])
def test_library_equals_oracle_on_procedural_model(axis, chunk):
    data = monument(n=40)
    sc = ctamd.Scene.from_vox(data, axis=axis, chunk=chunk)
    o = V.load(data, axis=axis, chunk=chunk)
    assert sc.info.cube_light_count > 0
    assert_same(lib_arrays(sc), o)
    lo = np.minimum.reduce(sc.aabbs[:, :3])
    hi = np.maximum.reduce(sc.aabbs[:, 3:])
    assert np.array_equal(sc.grid_min, lo) and np.array_equal(sc.grid_max, hi)


def test_malformed_files_are_errors():
    data = monument(n=12)
    for cut in list(range(0, 64)) + list(range(64, len(data), 97)):
        with pytest.raises(ctamd.CtError) as ei:
            ctamd.Scene.from_vox(data[:cut])
        assert ei.value.code == ctamd.CT_E_INVALID, cut
    bad = bytearray(write_vox([((2, 2, 2), [(2, 0, 0, 1)])]))  # x outside the model
    with pytest.raises(ctamd.CtError):
        ctamd.Scene.from_vox(bytes(bad))
    with pytest.raises(ctamd.CtError):
        ctamd.Scene.from_vox(write_vox([((2, 2, 2), [])]))  # no voxels
    with pytest.raises(ctamd.CtError):
        ctamd.Scene.from_vox(data, model=3)


ASSETS = "/root/reference/assets/models"


@pytest.mark.skipif(not os.path.isdir(ASSETS), reason="the reference's model files are not on this machine")
@pytest.mark.parametrize("name,voxels", [("monu7.vox", 226036), ("monu9.vox", 32832), ("room.vox", 82536),
                                         ("chr_sword.vox", 334), ("deer.vox", 1415)])
def test_reference_models(name, voxels):
    with open(os.path.join(ASSETS, name), "rb") as fh:
        data = fh.read()
    sc = ctamd.Scene.from_vox(data)
    assert sc.info.primitive_count == voxels  # the file's XYZI count (no duplicates in these files)
    assert_same(lib_arrays(sc), V.load(data))
    if name == "room.vox":
        # color indices 164, 166 and 248 are MATL _emit: 12 + 42 + 627 voxels
        assert sc.info.cube_light_count == 681
