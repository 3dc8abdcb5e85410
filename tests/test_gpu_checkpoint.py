"""Checkpointed propagator sweeps (pf_system_set_checkpoint) against full stacks.

With checkpoints, a chain keeps q and q+ only at segment boundaries and recomputes each
segment when Simpson, the stress or the reverse sweep needs it.  Densities, Q, free energy,
residuals, the stress and the whole free-energy gradient must match the full-stack results
to rounding (a
homopolymer then runs a real q+ sweep instead of reversing q: the same operator)."""
import numpy as np
import pytest

from helpers import cos_fields, diblock_system

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")


def blend(**kw):
    cell = P.UnitCell(2, "rectangular", 7.0, 5.0)
    A, B, C = P.Monomer(1, 1.0), P.Monomer(2, 1.3), P.Monomer(3, 0.8)
    tri = P.Multiblock([A, B, A], [12, 20, 8], [1.0, 1.1, 1.0], 20)
    hom = P.Homopolymer(B, 30, 0.9, 16)
    pt = P.Point(C)
    fl = P.FloryInteraction()
    fl.set_interaction(1, 2, 0.3)
    fl.set_interaction(2, 3, 0.15)
    fl.set_interaction(1, 3, 0.4)
    sys = P.FieldSystem((12, 8, 1), cell, monomers=[A, B, C], **kw)
    sys.add_species(tri, 0.6)
    sys.add_species(hom, 0.3)
    sys.add_species(pt, 0.1)
    sys.add_interaction(fl)
    return sys


CASES = [
    ("lam64", lambda **k: diblock_system(P, (64, 1, 1), 1, "lamellar", (9.0,), N=100, f=0.35,
                                         chi=0.2, Ns=50, **k)[0]),
    ("hex32", lambda **k: diblock_system(P, (32, 32, 1), 2, "hexagonal", (5.0,), N=100, f=0.3,
                                         chi=0.2, Ns=40, **k)[0]),
    ("cubic16", lambda **k: diblock_system(P, (16, 16, 16), 3, "cubic", (4.5,), N=100, f=0.35,
                                           chi=0.2, Ns=30, **k)[0]),
    ("blend12x8", blend),
]


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


@pytest.mark.parametrize("K", [-1, 3])
@pytest.mark.parametrize("name,build", CASES, ids=[c[0] for c in CASES])
def test_checkpointed_matches_full_fp64(name, build, K):
    full, ck = build(), build()
    ck.set_checkpoint(K)
    rng = np.random.default_rng(21)
    f = {m: 0.5 * rng.standard_normal(full.shape) for m in full.mids}
    for s in (full, ck):
        for m, w in f.items():
            s.fields[m] = w
    gf = P.fhelm_grad(full, "textbook")
    gc = P.fhelm_grad(ck, "textbook")
    assert gc.fhelm == pytest.approx(gf.fhelm, rel=1e-12, abs=1e-14)
    for a, b in zip(ck.species, full.species):
        assert a.Q == pytest.approx(b.Q, rel=1e-12)
    for m in full.mids:
        assert rel(ck.density[m], full.density[m]) < 1e-12
        assert rel(gc.dw[m], gf.dw[m]) < 1e-11
    assert rel(gc.dcell, gf.dcell) < 1e-10
    for a, b in zip(gc.dN, gf.dN):
        for x, y in zip(a, b):
            assert x == pytest.approx(y, rel=1e-10, abs=1e-14)
    assert rel(gc.dphi, gf.dphi) < 1e-11
    # residuals! and field_error through the checkpointed density
    for s in (full, ck):
        for m, w in f.items():
            s.fields[m] = w
        s.residuals_update()
    assert ck.field_error() == pytest.approx(full.field_error(), rel=1e-11)
    # the reference stress formula on segments recomputed from the checkpoints
    sf, sc = full.scfstress(), ck.scfstress()
    assert rel(sc, sf) < 1e-10, (sc, sf)


def test_checkpointed_64cubed_fp32():
    full = diblock_system(P, (64, 64, 64), 3, "cubic", (16.33,), N=100, f=0.35, chi=0.2, Ns=100,
                          b=(1.0, 1.0), precision="fp32")[0]
    ck = diblock_system(P, (64, 64, 64), 3, "cubic", (16.33,), N=100, f=0.35, chi=0.2, Ns=100,
                        b=(1.0, 1.0), precision="fp32")[0]
    ck.set_checkpoint(-1)
    rng = np.random.default_rng(5)
    f = {m: 0.5 * rng.standard_normal(full.shape) for m in full.mids}
    for s in (full, ck):
        for m, w in f.items():
            s.fields[m] = w
    gf = P.fhelm_grad(full, "reference")
    gc = P.fhelm_grad(ck, "reference")
    # FP32 arithmetic, identical operations in a different launch grouping
    assert gc.fhelm == pytest.approx(gf.fhelm, rel=1e-6)
    for m in full.mids:
        assert rel(gc.dw[m], gf.dw[m]) < 1e-5
    assert rel(gc.dcell, gf.dcell) < 1e-4


def test_checkpointed_variable_cell_fp64():
    # DomainUpdater on a checkpointed system: the same fixed point as with full stacks
    # (the forward-difference cell Jacobian amplifies last-bit differences, so the two
    # trajectories agree at the fixed point, to the tolerances that bound it)
    def build():
        return diblock_system(P, (64, 1, 1), 1, "lamellar", (16.0,), chi=0.2)

    out = []
    for K in (0, -1):
        s, c, _ = build()
        s.set_checkpoint(K)
        P.fieldinit(s, fields=cos_fields(s.shape))
        r = P.scft(s, P.AndersonUpdater(nhist=10), domain=P.DomainUpdater(nskip=1, nper=5),
                   opts=P.SCFTOptions(nsteps=600, ftol=1e-9, stol=1e-9))
        assert r.converged
        out.append((r.fhelm, c.params[0]))
    assert out[1][0] == pytest.approx(out[0][0], rel=3e-7)
    assert out[1][1] == pytest.approx(out[0][1], rel=3e-6)
