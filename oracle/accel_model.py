"""Oracle of the acceleration-structure manager's task semantics: a
restatement, on per-instance voxel lists, of what process_task_queue and
its helpers do to the scene records (src/accel_struct_mngr.cpp:2115-2443,
delete_aabb_device_buffer :607-732, update_remapping_buffer :734-773,
update_light_remapping_buffer :775-816, delete_light_device_buffer
:1155-1196, the undo path :853-1083 and the settle :2586-2680).

Test infrastructure only: the tests apply the same task sequence to the
library and to this model and compare the scene records, the remapping
tables and the previous-frame state.  Pure Python over small scenes.

Where the reference's text is inconsistent the model follows the semantics
the library documents in include/ct_abi.h (TASK):
* the light swap runs whether or not the deleted voxel is the instance's
  last (the reference skips the back-reference fix then, :669), and the
  deleted/moved light indices are reset per task (the reference keeps them
  in variables declared outside its task loop, :2137-2140);
* UPDATE_BLAS skips primitive indices >= the instance's count (the
  reference's test is `>`, :532);
* an undo restores with instance-local remapping values (the reference
  writes a slot index there, :1024).
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field

import numpy as np

INVALID = 0xFFFFFFFF


def mat4_mul(m, t):
    """Column-major m * t with daxa_f32mat4x4_mult's float32 operation order
    (math.inl:92-117): each element ((m0 t0 + m1 t1) + m2 t2) + m3 t3."""
    m = np.asarray(m, np.float32).reshape(16)
    t = np.asarray(t, np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    for c in range(4):
        for r in range(4):
            acc = np.float32(m[0 * 4 + r] * t[c * 4 + 0])
            acc = np.float32(acc + np.float32(m[1 * 4 + r] * t[c * 4 + 1]))
            acc = np.float32(acc + np.float32(m[2 * 4 + r] * t[c * 4 + 2]))
            acc = np.float32(acc + np.float32(m[3 * 4 + r] * t[c * 4 + 3]))
            out[c * 4 + r] = acc
    return out


@dataclass
class Voxel:
    aabb: tuple          # 6 float32 values, min then max
    material: int
    light: int           # index in the light list or INVALID


@dataclass
class LightRec:
    data: bytes          # the 40-byte LIGHT record with instance_info zeroed
    instance: int
    primitive: int


@dataclass
class Instance:
    transform: np.ndarray
    voxels: list = field(default_factory=list)


@dataclass
class Done:
    kind: str
    instance: int = 0
    deleted: int = 0
    last: int = 0
    voxel: Voxel | None = None
    light: int = INVALID
    light_moved: int = INVALID
    light_rec: LightRec | None = None
    lights_after: int = 0            # length of the light list the deletion left
    moved_rec: LightRec | None = None  # the record it moved into the hole


class StateError(Exception):
    """A task the library refuses with CT_E_STATE (the update goes on)."""


class AccelModel:
    """Scene records by instance id, the cube light list, the remapping
    tables keyed by (instance, previous local index) and by light index, the
    previous-frame transforms and the undo stack."""

    def __init__(self):
        self.instances: dict[int, Instance] = {}
        self.lights: list[LightRec] = []
        self.remap_prims: dict[tuple, int] = {}
        self.remap_lights: dict[int, int] = {}
        self.prev_transform: dict[int, np.ndarray] = {}
        self.done: list[Done] = []
        self.settling = False
        # moves of the update in progress (see update)
        self._origin, self._final, self._porigin, self._pfinal = [], {}, {}, {}

    # ---------------------------------------------------------- loading
    @classmethod
    def from_records(cls, instances, aabbs, prims, lights, light_count):
        """A committed scene: `instances` by id (ct_instance records,
        primitive_count 0 = unused id), aabbs [slots, 6] and prims [slots, 2]
        by slot, the first light_count LIGHT records."""
        m = cls()
        for i, ins in enumerate(instances):
            if ins.primitive_count == 0:
                continue
            f = ins.first_primitive_index
            vox = [Voxel(tuple(np.asarray(aabbs[f + k], np.float32).tolist()), int(prims[f + k, 0]),
                         int(prims[f + k, 1])) for k in range(ins.primitive_count)]
            m.instances[i] = Instance(np.array(ins.transform[:], np.float32), vox)
            m.prev_transform[i] = m.instances[i].transform.copy()
        for l in lights[:light_count]:
            raw = bytearray(bytes(l))
            inst, prim = l.instance_info.instance_id, l.instance_info.primitive_id
            raw[24:32] = b"\0" * 8   # instance_info lives at bytes 24..31 of LIGHT
            m.lights.append(LightRec(bytes(raw), int(inst), int(prim)))
        return m

    def build(self, voxels_per_instance, transforms, lights):
        """BUILD_BLAS_FROM_CPU of n instances: ids are the lowest free ones
        (uuid.hpp:22-37); `lights` are (record bytes, instance k of this
        batch, primitive) and voxels name lights by position in `lights`."""
        self._settle()
        ids = []
        base = len(self.lights)
        for vox, tr in zip(voxels_per_instance, transforms):
            i = 0
            while i in self.instances:
                i += 1
            self.instances[i] = Instance(np.asarray(tr, np.float32).reshape(16).copy(),
                                         [Voxel(tuple(v.aabb), v.material,
                                                INVALID if v.light == INVALID else base + v.light) for v in vox])
            self.prev_transform[i] = self.instances[i].transform.copy()
            ids.append(i)
        for data, k, p in lights:
            self.lights.append(LightRec(data, ids[k], p))
        self.settling = True
        return ids

    # ---------------------------------------------------------- updates
    def update(self, tasks):
        """One update_scene call: settle the previous update, then apply the
        tasks in order.  Each task is a tuple (kind, args...).  Light moves
        are tracked in the index space of the list as the update found it
        (the previous frame's): origin[l] is the original index of the light
        now at l, and each original index that moved gets one remapping
        entry, its final position.  Returns the StateErrors of refused
        tasks."""
        self._settle()
        self._origin = list(range(len(self.lights)))
        self._final = {}
        self._porigin, self._pfinal = {}, {}
        refused = []
        self._gpu_deletions([t for t in tasks if t[0] == "delete_from_gpu"])
        for t in tasks:
            try:
                getattr(self, "_" + t[0])(*t[1:])
            except StateError as e:
                refused.append(e)
        self.remap_lights.update(self._final)
        for (i, o), f in self._pfinal.items():
            self.remap_prims[(i, o)] = f
        self.settling = True
        return refused

    def _prim_origin(self, i, n):
        if i not in self._porigin:
            self._porigin[i] = list(range(n))
        return self._porigin[i]

    def _prim_removed(self, i, n, l):
        org = self._prim_origin(i, n)
        if l < len(org) and org[l] is not None:
            self._pfinal[(i, org[l])] = INVALID

    def _prim_moved(self, i, n, src, dst):
        org = self._prim_origin(i, n)
        o = org[src] if src < len(org) else None
        while len(org) <= dst:
            org.append(None)
        org[dst] = o
        if o is not None:
            self._pfinal[(i, o)] = dst

    def _prim_restored(self, i, n, l):
        org = self._prim_origin(i, n)
        while len(org) <= l:
            org.append(None)
        org[l] = None

    def _light_removed(self, l):
        o = self._origin[l] if l < len(self._origin) else None
        if o is not None:
            self._final[o] = INVALID

    def _light_moved(self, src, dst):
        o = self._origin[src] if src < len(self._origin) else None
        while len(self._origin) <= dst:
            self._origin.append(None)
        self._origin[dst] = o
        if o is not None:
            self._final[o] = dst

    def _light_restored(self, l):
        while len(self._origin) <= l:
            self._origin.append(None)
        self._origin[l] = None

    def _gpu_deletions(self, tasks):
        """The GPU deletions of one update, before the task loop: each
        instance takes its compacted voxels; lights of surviving voxels get
        the voxel's new index; lights of deleted voxels are removed as a CPU
        deletion removes one (last light into the hole, owner relinked),
        highest index first; the light remapping maps original indices to
        final ones."""
        alive, gone = set(), set()
        for _, i, compacted, count in tasks:
            inst = self.instances.get(i)
            if inst is None:
                continue
            old = [v.light for v in inst.voxels]
            keep = len(old) - min(count, len(old))
            inst.voxels = [copy.copy(v) for v in list(compacted)[:keep]]
            for k, v in enumerate(inst.voxels):
                if v.light != INVALID and v.light < len(self.lights):
                    self.lights[v.light].primitive = k
                    alive.add(v.light)
            gone.update(L for L in old if L != INVALID and L < len(self.lights))
        gone -= alive
        for L in sorted(gone, reverse=True):
            Lx = len(self.lights) - 1
            self._light_removed(L)
            if Lx != L:
                self.lights[L] = copy.deepcopy(self.lights[Lx])
                self._light_moved(Lx, L)
                self._relink(L)
            self.lights.pop()

    def _settle(self):
        if self.settling:
            self.remap_prims.clear()
            self.remap_lights.clear()
            self.prev_transform = {i: inst.transform.copy() for i, inst in self.instances.items()}
            self.settling = False

    def _relink(self, l):
        rec = self.lights[l]
        inst = self.instances.get(rec.instance)
        if inst is not None and rec.primitive < len(inst.voxels):
            inst.voxels[rec.primitive].light = l

    def _delete_primitive(self, i, d):
        inst = self.instances[i]
        vox = inst.voxels
        assert 0 <= d < len(vox) and len(vox) > 1
        last = len(vox) - 1
        rec = Done("delete_primitive", i, d, last, copy.deepcopy(vox[d]))
        L = vox[d].light
        Lx = INVALID
        if L != INVALID:
            rec.light_rec = copy.deepcopy(self.lights[L])
            Lx = len(self.lights) - 1
            if Lx == L:
                Lx = INVALID
            self._light_removed(L)
            if Lx != INVALID:
                self.lights[L] = copy.deepcopy(self.lights[Lx])
                self._light_moved(Lx, L)
                self._relink(L)
            self.lights.pop()          # the list shrinks by one; its last entry is now stale
        if last != d:
            vox[d] = vox[last]
            if vox[d].light != INVALID:
                self.lights[vox[d].light].primitive = d
        vox.pop()
        rec.light, rec.light_moved = L, Lx
        rec.lights_after = len(self.lights)
        if Lx != INVALID:
            rec.moved_rec = copy.deepcopy(self.lights[L])
        self._prim_removed(i, last + 1, d)
        if last != d:
            self._prim_moved(i, last + 1, last, d)
        self.done.append(rec)

    def _lights_as_left(self, rec):
        """The light list is still the one the deletion left: same length
        and the record it moved into the hole still there."""
        if rec.light == INVALID:
            return True
        if len(self.lights) != rec.lights_after:
            return False
        return rec.light_moved == INVALID or self.lights[rec.light] == rec.moved_rec

    def _undo(self):
        if not self.done:
            return
        rec = self.done.pop()
        if rec.kind != "delete_primitive":
            return
        inst = self.instances.get(rec.instance)
        if inst is None or len(inst.voxels) != rec.last:
            raise StateError("UNDO of a primitive deletion whose instance has changed since")
        if not self._lights_as_left(rec):
            raise StateError("UNDO of a primitive deletion whose cube light list has changed since")
        vox = inst.voxels
        d, last = rec.deleted, rec.last
        vox.append(None)
        if last != d:
            vox[last] = vox[d]
            if vox[last].light != INVALID:
                self.lights[vox[last].light].primitive = last
            self._prim_moved(rec.instance, last, d, last)
        self._prim_restored(rec.instance, last, d)
        vox[d] = copy.deepcopy(rec.voxel)
        if rec.light != INVALID:
            self.lights.append(None)
            if rec.light_moved != INVALID:
                self.lights[rec.light_moved] = copy.deepcopy(self.lights[rec.light])
                self._light_moved(rec.light, rec.light_moved)
                self._relink(rec.light_moved)
            self.lights[rec.light] = copy.deepcopy(rec.light_rec)
            self._light_restored(rec.light)

    def _delete_blas(self, i):
        del self.instances[i]
        self.done.append(Done("delete_blas", i))

    def _update_blas(self, i, transform, edits=()):
        inst = self.instances[i]
        inst.transform = mat4_mul(inst.transform, transform)
        for p, box in edits:
            if p < len(inst.voxels):
                inst.voxels[p].aabb = tuple(np.asarray(box, np.float32).tolist())
        self.done.append(Done("update_blas", i))

    def _delete_from_gpu(self, i, compacted, count):
        """The caller's pass left `compacted` (the instance's voxels after
        its own compaction) and drops `count` from the end; the voxels were
        taken in _gpu_deletions, an instance left empty is deleted here."""
        if i in self.instances and not self.instances[i].voxels:
            del self.instances[i]
        self.done.append(Done("delete_from_gpu", i))

    # ---------------------------------------------------------- export
    def remap_slot_table(self, prev_first, capacity):
        """The remapping tables as the library lays them out: by slot
        (previous first index + previous local index) and by light index."""
        rp = np.zeros(capacity, np.uint32)
        for (i, p), v in self.remap_prims.items():
            rp[prev_first[i] + p] = v
        return rp

    def remap_light_table(self, capacity):
        rl = np.zeros(capacity, np.uint32)
        for l, v in self.remap_lights.items():
            rl[l] = v
        return rl
