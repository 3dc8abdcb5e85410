"""Launches whose lanes have different contour lengths.

Every propagator launch steps all lanes of a system (or of a batch) at the same step
index t; a lane with fewer steps (a homopolymer beside a longer copolymer, a batch of
chains with different Ns) is finished for the remaining t.  The compile-time sweep kernels
then load that lane's buffers and drop its results.  These tests put such lanes beside
live ones on the sizes those kernels serve (32^3 and 64^3, FP32 and FP64) and check the
densities and partition functions against the oracle, the FP64 path and one-at-a-time
evaluation."""
import numpy as np
import pytest

import oracle as O
from helpers import diblock_system

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")


def blend3d(mod, dims, a, precision=None):
    # diblock (Ns 40) + A homopolymer (Ns 16) + B homopolymer (Ns 24): six lanes of three
    # lengths in every launch
    kw = {} if precision is None else {"precision": precision}
    cell = mod.UnitCell(3, "cubic", a)
    A, B = mod.Monomer(1), mod.Monomer(2)
    di = mod.Multiblock([A, B], [30, 70], [1.0, 1.1], 40)
    ha = mod.Homopolymer(A, 40, 1.0, 16)
    hb = mod.Homopolymer(B, 60, 1.1, 24)
    itx = mod.FloryInteraction()
    itx.set_interaction(1, 2, 0.2)
    sys = mod.FieldSystem(dims, cell, monomers=[A, B], **kw)
    sys.add_species(di, 0.7)
    sys.add_species(ha, 0.2)
    sys.add_species(hb, 0.1)
    sys.add_interaction(itx)
    return sys


def rand_fields(shape, seed):
    rng = np.random.default_rng(seed)
    return {m: 0.3 * rng.standard_normal(shape) for m in (1, 2)}


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def test_uneven_lanes_fp64_vs_oracle_32cubed():
    f = rand_fields((32, 32, 32), 3)
    so, sg = blend3d(O, (32, 32, 32), 5.5), blend3d(P, (32, 32, 32), 5.5, "fp64")
    O.fieldinit(so, fields=f)
    P.fieldinit(sg, fields=f)
    for a, b in zip(sg.species, so.species):
        assert a.Q == pytest.approx(b.Q, rel=1e-12)
    for m in (1, 2):
        assert rel(sg.density[m], so.density[m]) < 1e-12
    assert sg.free_energy() == pytest.approx(so.free_energy(), rel=1e-12)


@pytest.mark.parametrize("n", [32, 64])
def test_uneven_lanes_fp32_vs_fp64(n):
    f = rand_fields((n, n, n), 4)
    s64, s32 = blend3d(P, (n, n, n), 5.5, "fp64"), blend3d(P, (n, n, n), 5.5, "fp32")
    P.fieldinit(s64, fields=f)
    P.fieldinit(s32, fields=f)
    for a, b in zip(s32.species, s64.species):
        assert a.Q == pytest.approx(b.Q, rel=2e-5)
    for m in (1, 2):
        assert rel(s32.density[m], s64.density[m]) < 2e-5


@pytest.mark.parametrize("precision,n", [("fp32", 64), ("fp32", 32), ("fp64", 32)])
def test_batch_of_different_ns_matches_single(precision, n):
    # three diblocks with Ns 30, 50 and 70 in one batch (their lanes finish at different
    # steps of the shared launches) against each evaluated alone
    def make(ns):
        return diblock_system(P, (n, n, n), 3, "cubic", (5.0,), f=0.4, Ns=ns,
                              precision=precision)[0]

    nss = (30, 50, 70)
    singles, batched = [make(k) for k in nss], [make(k) for k in nss]
    for k in range(3):
        f = rand_fields((n, n, n), 10 + k)
        for m in (1, 2):
            singles[k].fields[m] = f[m]
            batched[k].fields[m] = f[m]
        singles[k].density_update()
    P.Batch(batched).density_update()
    tol = 1e-6 if precision == "fp32" else 1e-12
    for a, b in zip(singles, batched):
        assert b.species[0].Q == pytest.approx(a.species[0].Q, rel=tol)
        for m in (1, 2):
            assert rel(b.density[m], a.density[m]) < tol
