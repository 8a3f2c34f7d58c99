"""Scale checks against closed forms (no oracle at this size).

A perfect n^3 lattice of FP32 spheres at unit spacing with radius 0.6 has
AABB overlaps exactly between lattice neighbours (|step| = 1 on each axis:
13 directions) and sphere contacts exactly between face neighbours (distance
1 <= 1.2; sqrt(2) > 1.2).  At n = 210 the shape count (9,261,000) puts the
chunk-slot scans beyond 2,048 tiles and the pair count (~1.2e8) does the same
for the narrowphase scans, so the large-input scan path, multi-pass key sort
and the pair emission all run at full size.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lattice(scenes, n):
    idx = np.arange(n ** 3, dtype=np.int64)
    pos = np.stack([idx % n, (idx // n) % n, idx // (n * n)], axis=1).astype(np.float32)
    del idx
    N = n ** 3
    rot = np.zeros((N, 4), np.float32)
    rot[:, 0] = 1
    B = np.zeros((N, 3), np.float32)
    B[:, 0] = 0.6
    g = max(1, n // 2)
    return scenes._one_shape_per_body(np.zeros(N, np.int32), B, pos, rot, np.float32, (g, g, g), "unit-lattice",
                                      {"side": n})


def test_unit_lattice_closed_forms(cdx, ctx, scenes):
    """Both pair searches at full size: closed-form totals, identical pair lists."""
    n = 210
    s = _lattice(scenes, n)
    m = n - 1
    pairs = 3 * m * n * n + 6 * m * m * n + 4 * m ** 3
    contacts = 3 * m * n * n
    lists = {}
    for mode, name in ((cdx.PAIRS_CELLS, "cells"), (cdx.PAIRS_BINS, "bins")):
        sc = cdx.Scene(ctx, s.dtype).load(s)
        sc.set_pair_search(mode)
        sc.step()
        sc.sync()
        c = sc.counts()
        print(f"n={n} {name}: {c.pairs} pairs (closed form {pairs}), {c.contacts} contacts (closed form {contacts}), "
              f"{c.bin_entries} bin entries, {c.active_bins} active bins", flush=True)
        assert bool(c.flags & cdx.FLAG_CELL_SEARCH) == (name == "cells")
        assert c.pairs == pairs
        assert c.contacts == contacts
        lists[name] = sc.read_pairs()
        if name == "bins":
            assert c.active_bins == cells_active
        else:
            cells_active = c.active_bins  # built on request from the cell search's AABBs
        sc.close()
    assert np.array_equal(lists["cells"], lists["bins"])
    assert np.all(np.diff(lists["cells"]) > 0)
