"""Oracle for the GPU voxel-deletion brush -- test infrastructure only.

Restates squared_brush.slang:79-133 (which voxels a stroke marks) and the
intended result of rearrangement.slang:62-204 (the swap-with-tail
compaction of one instance and its remapping entries), one voxel at a
time in plain Python.  Only tests use it, as the checker of the library.

Pinned by: tests on hand-made cases -- no deletion keeps everything in
place; deleting only tail voxels moves nothing; deleting the whole
instance leaves nothing; the surviving voxel multiset equals the input
minus the deleted ones; survivors below n - d never move; every move goes
from a tail slot to a deleted slot; the remapping table is the inverse of
the moves -- and on the brush side the window clipping table, including
the u32 wrap at the left and top edges (squared_brush.slang:108-109).
"""
import numpy as np

INVALID = 0xFFFFFFFF
U32 = 1 << 32


def brush_window(x, y, window, W, H):
    """(xmin, xmax, ymin, ymax) of the stroke, in u32 arithmetic, or None."""
    xmin = max((x - window) % U32, 0)
    ymin = max((y - window) % U32, 0)
    xmax = min((x + window) % U32, W - 1)
    ymax = min((y + window) % U32, H - 1)
    if xmin > xmax or ymin > ymax:
        return None
    return xmin, xmax, ymin, ymax


def brush_marks(dist, hit, W, H, x, y, window):
    """Voxels (instance, primitive) a stroke at (x, y) marks, from the
    stored DI: dist (H*W,) f32 and hit (H*W, 2) u32 per pixel."""
    c = y * W + x
    if not dist[c] > 0:
        return set()
    win = brush_window(x, y, window, W, H)
    if win is None:
        return set()
    xmin, xmax, ymin, ymax = win
    out = set()
    for py in range(ymin, ymax + 1):
        for px in range(xmin, xmax + 1):
            q = py * W + px
            if dist[q] > 0 and hit[q, 0] == hit[c, 0]:
                out.add((int(hit[q, 0]), int(hit[q, 1])))
    return out


def rearrange(n, deleted):
    """Swap-with-tail compaction of an instance of n voxels whose local
    indices `deleted` are marked.  Returns (moves {from: to}, keep, remap
    {local: value}) with remap[moved from] = to and remap[deleted] = ~0."""
    dele = sorted(set(deleted))
    keep = n - len(dele)
    holes = [i for i in dele if i < keep]
    tail = [s for s in range(keep, n) if s not in set(dele)]
    assert len(holes) == len(tail)
    moves = dict(zip(tail, holes))
    remap = {i: INVALID for i in dele}
    remap.update(moves)
    return moves, keep, remap


def compacted_records(records, n, deleted):
    """The instance's first n - d records after the compaction."""
    moves, keep, _ = rearrange(n, deleted)
    out = list(records[:n])
    for s, i in moves.items():
        out[i] = records[s]
    return out[:keep]
