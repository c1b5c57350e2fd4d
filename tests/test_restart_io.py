"""Restart files (wf.dat1, vrho.dat1): round trips, and interchange with a
Fortran program compiled here (amdflang) that reads and writes the records
with its own unformatted I/O in the reference writers' sequence
(io_write_wf_simple_module.f90:153-278, io_module.f90:117-224)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from rsdft_develop_amd import restart

HERE = os.path.dirname(os.path.abspath(__file__))
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)


def _sample(ngrid=(5, 4, 3), nb=3, seed=0):
    rng = np.random.default_rng(seed)
    ML = int(np.prod(ngrid))
    bands = rng.standard_normal((ML, nb))
    occ = np.array([2.0, 1.25, 0.5][:nb])
    aa = np.diag([5.0, 4.2, 3.1])
    vr = restart.VrhoData(ngrid, aa, rng.random((2, 3)), np.array([0, 0]), np.array([14]), -0.123,
                          *(rng.standard_normal(ML) for _ in range(4)))
    return bands, occ, aa, vr


def test_wf_round_trip_fp64_and_fp32(tmp_path):
    bands, occ, aa, _ = _sample()
    for oc, tol in ((3, 0.0), (4, 1e-6)):
        p = tmp_path / f"wf{oc}.dat1"
        restart.write_wf(str(p), bands, occ, aa, (5, 4, 3), oc=oc)
        w = restart.read_wf(str(p))
        assert w.ngrid == (5, 4, 3) and w.mb == 3 and w.band_range == (1, 3) and w.oc == oc
        np.testing.assert_array_equal(w.occ, occ)
        np.testing.assert_array_equal(w.aa, aa)
        np.testing.assert_allclose(w.bands, bands, rtol=tol, atol=0)


def test_vrho_round_trip(tmp_path):
    _, _, _, vr = _sample()
    p = tmp_path / "vrho.dat1"
    restart.write_vrho(str(p), vr)
    got = restart.read_vrho(str(p))
    assert got.ngrid == vr.ngrid and got.efermi == vr.efermi
    np.testing.assert_array_equal(got.kind, vr.kind)
    np.testing.assert_array_equal(got.zn, vr.zn)
    np.testing.assert_array_equal(got.frac, vr.frac)
    for f in ("rho", "vloc", "vh", "vxc"):
        np.testing.assert_array_equal(getattr(got, f), getattr(vr, f))


def test_reader_places_points_by_the_grid_map(tmp_path):
    """A file whose points come in another order (another rank layout)
    reads back into the natural order."""
    bands, occ, aa, _ = _sample(nb=1)
    ng = (5, 4, 3)
    LL = restart.grid_map(ng)
    perm = np.random.default_rng(3).permutation(LL.shape[1])
    p = tmp_path / "perm.dat1"
    w = restart.RecordWriter(str(p))
    w.write(np.int32(-1))
    w.write(np.array([60, *ng], dtype="<i4"))
    w.write(np.array([1, 1, 1], dtype="<i4"))
    w.write(np.array([1, 1, 1], dtype="<i4"))
    w.write(np.array([0, 3, 1], dtype="<i4"))
    w.write(np.ascontiguousarray(LL[:, perm]))
    w.write(occ[:1])
    w.write(aa, restart.recip(aa), np.zeros(3))
    for rec in (np.ones(8, "<i4"), np.array([0, 60], "<i4"), np.array([0, 1], "<i4"), np.array([0, 1], "<i4"),
                np.array([0, 1], "<i4")):
        w.write(rec)
    w.write(bands[perm, 0])
    w.close()
    np.testing.assert_array_equal(restart.read_wf(str(p)).bands[:, 0], bands[:, 0])


def test_reciprocal_vectors():
    aa = np.array([[4.0, 0.5, 0.0], [0.0, 3.0, 0.2], [0.1, 0.0, 2.0]])
    bb = restart.recip(aa)
    np.testing.assert_allclose(aa.T @ bb, 2 * np.pi * np.eye(3), atol=1e-12)


@pytest.fixture(scope="module")
def fortran_prog(tmp_path_factory):
    if FC is None:
        pytest.skip("amdflang not available")
    d = tmp_path_factory.mktemp("fio")
    exe = d / "restart_io"
    r = subprocess.run([FC, "-O1", "-o", str(exe), os.path.join(HERE, "fortran_io", "restart_io.f90")],
                       capture_output=True, text=True, cwd=d)
    assert r.returncode == 0, r.stderr
    return exe


def test_fortran_reads_python_files(tmp_path, fortran_prog):
    bands, occ, aa, vr = _sample()
    ng = (5, 4, 3)
    restart.write_wf(str(tmp_path / "wf.dat1"), bands, occ, aa, ng)
    restart.write_vrho(str(tmp_path / "vrho.dat1"), vr)
    out = subprocess.run([str(fortran_prog), "read", str(tmp_path / "wf.dat1"), str(tmp_path / "vrho.dat1")],
                         capture_output=True, text=True, check=True).stdout.split("\n")
    kv = {}
    for line in out:
        tok = line.split()
        if not tok:
            continue
        nkey = 2 if tok[0] in ("LL", "occ") else 1
        kv.setdefault(" ".join(tok[:nkey]), []).append(" ".join(tok[nkey:]))
    assert kv["wf"][0].split() == ["-1", "60", "5", "4", "3", "3", "3"]
    assert kv["LL last"][0].split() == ["4", "3", "2"]
    assert float(kv["occ sum"][0]) == pytest.approx(occ.sum(), rel=1e-15)
    assert float(kv["bb11"][0]) == pytest.approx(2 * np.pi / 5.0, rel=1e-15)
    w = np.arange(1, 61, dtype=np.float64)
    for b in range(3):
        n, val = kv["band"][b].split()
        assert int(n) == b + 1 and float(val) == pytest.approx(float(bands[:, b] @ w), rel=1e-14)
    assert kv["version"][0].strip() == restart.VERSION
    assert kv["atoms"][0].split() == ["1", "2", "1", "14"]
    assert float(kv["efermi"][0]) == vr.efermi
    sums = [float(x) for x in kv["fields"][0].split()]
    np.testing.assert_allclose(sums, [vr.rho.sum(), vr.vloc.sum(), vr.vh.sum(), vr.vxc.sum()], rtol=1e-14)


def test_python_reads_fortran_file(tmp_path, fortran_prog):
    subprocess.run([str(fortran_prog), "write", str(tmp_path / "wf.dat1"), str(tmp_path / "unused")], check=True)
    w = restart.read_wf(str(tmp_path / "wf.dat1"))
    assert w.ngrid == (4, 3, 2) and w.mb == 3 and w.band_range == (1, 3)
    np.testing.assert_array_equal(w.occ, [2.0, 1.5, 0.25])
    np.testing.assert_array_equal(w.aa, np.diag([4.0, 3.0, 2.0]))
    for n in range(3):
        np.testing.assert_array_equal(w.bands[:, n], np.arange(1, 25) + 100.0 * (n + 1))
