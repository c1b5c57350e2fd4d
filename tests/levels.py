"""Comparison of eigenvalue lists that contain degenerate shells."""
import numpy as np


def shell_means(e, gap=1e-4):
    """Means and sizes of the clusters of sorted levels closer than `gap`."""
    e = np.sort(np.asarray(e, dtype=np.float64))
    cuts = np.flatnonzero(np.diff(e) > gap) + 1
    return np.array([s.mean() for s in np.split(e, cuts)]), np.diff(np.r_[0, cuts, e.size])


def assert_levels_match(got, want, shell_tol=1e-8, level_tol=1e-8):
    """Shell means (invariant under rotations inside a degenerate shell) to
    shell_tol, single sorted levels to level_tol, same shell structure."""
    mg, ng = shell_means(got)
    mw, nw = shell_means(want)
    np.testing.assert_array_equal(ng, nw)
    np.testing.assert_allclose(mg, mw, rtol=0, atol=shell_tol)
    np.testing.assert_allclose(np.sort(got), np.sort(want), rtol=0, atol=level_tol)
