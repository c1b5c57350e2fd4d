"""Host logic of the mixed-precision SCF: when the loop leaves FP32.

The SCF loop is driven with scripted per-iteration energy changes instead of
a device context, so these tests run without a GPU."""
import numpy as np

from rsdft_develop_amd import scf


def _run(des, **kw):
    """SCF.run over scripted dE/atom values; returns the precisions used."""
    run = object.__new__(scf.SCF)
    run.p = scf.ScfParams(mixed=True, etol_per_atom=1e-6, max_iter=len(des), **kw)
    run.cell = type("C", (), {"natom": 1})()
    run.occ, run.eewald, run.setup_s = np.zeros(1), 0.0, 0.0
    seq = iter(des)

    def step(it, e_prev, precision=64):
        return scf.ScfIteration(it, 0.0, 0.0, 0.0, next(seq), 0.0, np.zeros(1), {}, precision)

    run.step = step
    run.update_potential = lambda: 0.0
    res = run.run()
    return [r.precision for r in res.iterations], res.converged


def test_switch_below_twice_the_tolerance_by_default():
    precs, conv = _run([1.0, 1e-2, 1e-4, 3e-6, 1.5e-6, 9e-7, 5e-7])
    # 1.5e-6 < 2e-6 switches after that iteration; the first FP64 one is
    # below the tolerance but measured against an FP32 energy, so the second
    # FP64 one ends the loop
    assert precs == [32, 32, 32, 32, 32, 64, 64] and conv


def test_first_fp64_iteration_never_stops():
    precs, conv = _run([1.0, 1e-2, 1.5e-6, 9e-7])
    assert precs == [32, 32, 32, 64] and not conv


def test_explicit_threshold():
    precs, conv = _run([1.0, 1e-2, 5e-5, 1e-5, 9e-7], fp64_below=1e-4)
    assert precs == [32, 32, 32, 64, 64] and conv


def test_fp32_never_ends_the_loop():
    # below the stop tolerance in FP32: two FP64 iterations must confirm it
    precs, conv = _run([1.0, 5e-7, 4e-7, 3e-7])
    assert precs == [32, 32, 64, 64] and conv


def test_stalled_fp32_switches():
    # FP32 floor at 3e-5 (above 2 etol): three iterations without a new
    # lowest dE move the loop to FP64
    precs, conv = _run([1.0, 1e-3, 3e-5, 4e-5, 3.5e-5, 3.2e-5, 1e-5, 8e-7])
    assert precs == [32, 32, 32, 32, 32, 32, 64, 64] and conv
