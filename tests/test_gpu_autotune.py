"""Autotuner (pf_system_autotune / pf_batch_autotune): the chosen schedule is one of the
candidates, every candidate computes the same density! (to the rounding of the transform
form), and the tuned choice then serves later calls.  Schedules are reset at the end so
that later tests run the built-in defaults."""
import numpy as np
import pytest

from helpers import diblock_system

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")


def seeded(s, seed=1):
    rng = np.random.default_rng(seed)
    for m in s.mids:
        s.fields[m] = 0.3 * rng.standard_normal(s.shape)


@pytest.fixture(autouse=True)
def reset_after():
    yield
    P.schedule_reset()


@pytest.mark.parametrize("dims,dim,crystal,precision,tol", [
    ((32, 32, 32), 3, "cubic", "fp64", 1e-12),
    ((64, 64, 64), 3, "cubic", "fp32", 2e-5),
    ((48, 48, 1), 2, "hexagonal", "fp64", 1e-12),
])
def test_system_autotune(dims, dim, crystal, precision, tol):
    s, _, _ = diblock_system(P, dims, dim, crystal, (5.0,), N=100, f=0.35, chi=0.2, Ns=20,
                             precision=precision)
    seeded(s)
    s.density_update()
    ref = {m: s.density[m] for m in s.mids}
    sched, ms = s.autotune(reps=2)
    assert ms > 0
    if dims[2] > 1:
        assert sched["slab"] in ((0, 1, 2, 3) if precision == "fp32" and dims[:2] == (64, 64)
                                 else (0, 1, 2))
        assert sched["pencil"] in ((0, 1, 2, 3, 4, 5, 6) if dims[2] == 64 else (0, 1, 2))
    else:
        assert sched["planar"] in (0, 1, 2, 3)
    seeded(s)
    s.density_update()
    for m in s.mids:
        err = np.max(np.abs(s.density[m] - ref[m])) / np.max(np.abs(ref[m]))
        assert err < tol


def test_batch_autotune_then_scft():
    syss = []
    for dims, dim, crystal in [((32, 32, 1), 2, "hexagonal"), ((16, 16, 16), 3, "cubic"),
                               ((16, 16, 16), 3, "cubic")]:
        s, _, _ = diblock_system(P, dims, dim, crystal, (4.5,), N=100, f=0.35, chi=0.2, Ns=20)
        seeded(s, seed=len(syss))
        syss.append(s)
    batch = P.Batch(syss)
    batch.residuals_update()
    ref = [s.field_error() for s in syss]
    out = batch.autotune(reps=2)
    assert len(out) == 2
    for r in out:
        assert r["bytes"] == 8
        if r["nz"] > 1:
            assert r["split"] in (1, 2)
    for k, s in enumerate(syss):
        seeded(s, seed=k)
    batch.residuals_update()
    for a, s in zip(ref, syss):
        assert s.field_error() == pytest.approx(a, rel=1e-10)


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-5), ("fp64", 1e-12)])
def test_every_schedule_matches_default(precision, tol):
    """Every slab x pencil candidate of a 64^3 grid (generic, compile-time and
    shuffle-exchange slabs; generic, LDS (two rows, one row, one row with 512 threads, half
    a row), register and pair pencils) gives the default
    schedule's densities to the transform rounding."""
    dims = (64, 64, 64)
    s, _, _ = diblock_system(P, dims, 3, "cubic", (8.0,), N=100, f=0.35, chi=0.2, Ns=12,
                             precision=precision)
    seeded(s, 9)
    s.density_update()
    ref = {m: s.density[m] for m in s.mids}
    pencils = (0, 1, 2, 3, 4, 5, 6) if precision == "fp32" else (0, 1, 2, 4, 5, 6)
    slabs = (0, 1, 2, 3) if precision == "fp32" else (0, 1, 2)
    for slab in slabs:
        for pencil in pencils:
            P.schedule_set(precision, dims, slab=slab, pencil=pencil)
            seeded(s, 9)
            s.density_update()
            for m in s.mids:
                err = np.max(np.abs(s.density[m] - ref[m])) / np.max(np.abs(ref[m]))
                assert err < tol, (slab, pencil, err)


def test_fp64_128_compile_time_matches_generic():
    # the FP64 128 x 128 xy-slab and 128-line z-pencil (one workgroup per CU) with 256
    # threads (1 / 1), with 512 (2 / 4, the default), the half-row pencil (2 / 6) and the
    # default schedule against the generic kernels (0 / 0) on a 128^3 cubic diblock
    dims = (128, 128, 128)
    out = []
    for sched in ((0, 0), (1, 1), (2, 4), (2, 6), (-1, -1)):
        P.schedule_set("fp64", dims, slab=sched[0], pencil=sched[1])
        try:
            s, _, c = diblock_system(P, dims, 3, "cubic", (14.0,), N=100, f=0.35, chi=0.2, Ns=8,
                                     precision="fp64")
            seeded(s, 4)
            s.density_update()
            out.append((c.Q, {m: s.density[m] for m in s.mids}))
            del s
        finally:
            P.schedule_reset()
    for k in range(1, len(out)):
        assert out[k][0] == pytest.approx(out[0][0], rel=1e-12), k
        for m in out[0][1]:
            err = np.max(np.abs(out[k][1][m] - out[0][1][m])) / np.max(np.abs(out[0][1][m]))
            assert err < 1e-12, (k, m, err)


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-5), ("fp64", 1e-12)])
def test_every_planar_form_matches_default(precision, tol):
    # the generic planar kernel and the compile-time one with 256, 512 and 1024 threads on
    # a 96^2 hexagonal plane
    dims = (96, 96, 1)
    s, _, _ = diblock_system(P, dims, 2, "hexagonal", (9.0,), N=100, f=0.35, chi=0.2, Ns=24,
                             precision=precision)
    seeded(s, 5)
    s.density_update()
    ref = {m: s.density[m] for m in s.mids}
    for planar in (0, 1, 2, 3):
        P.schedule_set(precision, dims, planar=planar)
        seeded(s, 5)
        s.density_update()
        for m in s.mids:
            err = np.max(np.abs(s.density[m] - ref[m])) / np.max(np.abs(ref[m]))
            assert err < tol, (planar, err)


def test_fp32_128_forms_match():
    # the FP32 128^3 slab (256, 512, 1024 threads) and LDS pencil (256, 512, 1024 threads,
    # half rows)
    dims = (128, 128, 128)
    s, _, _ = diblock_system(P, dims, 3, "cubic", (14.0,), N=100, f=0.35, chi=0.2, Ns=8,
                             precision="fp32")
    seeded(s, 6)
    s.density_update()
    ref = {m: s.density[m] for m in s.mids}
    for slab, pencil in ((1, 1), (2, 4), (4, 5), (4, 6), (0, 0)):
        P.schedule_set("fp32", dims, slab=slab, pencil=pencil)
        seeded(s, 6)
        s.density_update()
        for m in s.mids:
            err = np.max(np.abs(s.density[m] - ref[m])) / np.max(np.abs(ref[m]))
            assert err < 2e-5, (slab, pencil, err)
