"""Pins of the brush oracle: compaction invariants and the stroke window."""
import numpy as np
import pytest

import brush_model as B


@pytest.mark.parametrize("n,deleted", [(10, []), (10, [7, 8, 9]), (10, list(range(10))), (1, [0]),
                                       (37, [0, 5, 36]), (70, [1, 2, 3, 64, 65, 69])])
def test_compaction_cases(n, deleted):
    moves, keep, remap = B.rearrange(n, deleted)
    assert keep == n - len(set(deleted))
    recs = list(range(n))
    out = B.compacted_records(recs, n, deleted)
    assert sorted(out) == sorted(set(range(n)) - set(deleted))   # survivors, each once
    for i in range(keep):
        if i not in deleted:
            assert out[i] == i                                     # survivors below keep stay
    for s, i in moves.items():
        assert s >= keep and s not in deleted and i < keep and i in deleted
        assert remap[s] == i and out[i] == s
    assert all(remap[i] == B.INVALID for i in deleted)
    if all(i >= keep for i in deleted):
        assert not moves                                           # tail-only deletion moves nothing


def test_random_compaction_is_a_stable_hole_fill():
    rng = np.random.default_rng(1)
    for _ in range(200):
        n = int(rng.integers(1, 200))
        d = rng.choice(n, int(rng.integers(0, n + 1)), replace=False).tolist()
        moves, keep, remap = B.rearrange(n, d)
        # holes ascending receive tail survivors ascending
        assert list(moves.values()) == sorted(moves.values()) and list(moves) == sorted(moves)
        out = B.compacted_records(list(range(n)), n, d)
        assert len(out) == keep and len(set(out)) == keep and not set(out) & set(d)


def test_stroke_window_wraps_like_the_shader():
    assert B.brush_window(10, 10, 3, 64, 32) == (7, 13, 7, 13)
    assert B.brush_window(2, 10, 3, 64, 32) is None               # x - window wraps in u32
    assert B.brush_window(10, 1, 3, 64, 32) is None
    assert B.brush_window(62, 30, 3, 64, 32) == (59, 63, 27, 31)   # clipped at the far edges
    assert B.brush_window(3, 3, 3, 64, 32) == (0, 6, 0, 6)
    dist = np.ones(64 * 32, np.float32)
    hit = np.zeros((64 * 32, 2), np.uint32)
    hit[:, 1] = np.arange(64 * 32)
    hit[64 * 10 + 11, 0] = 1                                       # another instance
    m = B.brush_marks(dist, hit, 64, 32, 10, 10, 1)
    assert m == {(0, 64 * yy + xx) for yy in (9, 10, 11) for xx in (9, 10, 11)} - {(0, 64 * 10 + 11)}
    dist[64 * 10 + 10] = 0
    assert B.brush_marks(dist, hit, 64, 32, 10, 10, 1) == set()   # sky under the cursor
