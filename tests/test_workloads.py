"""Synthetic benchmark layouts are deterministic and consistent with one spacing per level."""
import numpy as np

from granite_amd import workloads


def test_uneven_amr_layout():
    a = workloads.uneven_amr_levels()
    b = workloads.uneven_amr_levels()
    assert a == b
    assert len(a) == 4 and sum(len(bl) for _, bl in a) == 250
    for l, (dx, blocks) in enumerate(a):
        assert dx == (1.0 / 32) * 2 ** (3 - l)
        for n, lo, hi in blocks:
            assert all(v in workloads.EDGES for v in n)
            # GridBlock derives dx = (hi - lo) / n; every block of the level must get dx exactly
            assert all((hi[d] - lo[d]) / n[d] == dx for d in range(3))
    cells = workloads.cells(a)
    assert 10e6 < cells < 15e6
    assert np.isclose(cells / 250, 36.8**3, rtol=0.25)
