"""The config-2 cell set: counts and ranges (host only)."""
import numpy as np

from rsdft_develop_amd import batch


def test_config2_cells_draws():
    cells = batch.config2_cells(64, seed=2024)
    assert len(cells) == 64
    grids = {c.cell.ngrid for c in cells}
    assert grids <= {(n, n, n) for n in (24, 32, 40, 48)} and len(grids) == 4
    nb = np.array([c.cell.nband for c in cells])
    assert nb.min() >= 20 and nb.max() <= 40
    sc = np.array([c.scale for c in cells])
    assert sc.min() >= 0.95 and sc.max() <= 1.10
    for c in cells:
        a = c.cell.aa[0, 0]
        assert np.isclose(a, 10.26 * c.scale)
        assert c.cell.natom == 8
    # same seed, same cells
    again = batch.config2_cells(64, seed=2024)
    assert all(np.array_equal(x.cell.frac, y.cell.frac) for x, y in zip(cells, again))
