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


def test_checkpointed_systems_in_a_batch_fp64():
    # checkpointed systems join a batch: each runs its own segment-wise density! inside the
    # batch's launch sequence (eagerly, not from a graph), beside systems with full stacks;
    # densities equal one-at-a-time bit for bit, and a batched Anderson scft! (fixed cell)
    # and a variable-cell one match one-at-a-time
    def make():
        out = []
        for k, (name, build) in enumerate(CASES):
            s = build()
            if k != 1:
                s.set_checkpoint(-1 if k % 2 == 0 else 3)
            rng = np.random.default_rng(40 + k)
            for m in s.mids:
                s.fields[m] = 0.3 * rng.standard_normal(s.shape)
            out.append(s)
        return out

    single, batch_sys = make(), make()
    for s in single:
        s.density_update()
    b = P.Batch(batch_sys)
    b.density_update()
    for a, c in zip(single, batch_sys):
        for m in a.mids:
            np.testing.assert_array_equal(np.asarray(c.density[m]), np.asarray(a.density[m]))
    ref = [P.scft(s, P.AndersonUpdater(nhist=10), opts=P.SCFTOptions(nsteps=60, ftol=1e-7))
           for s in single]
    res = b.scft(nhist=10, nsteps=60, ftol=1e-7)
    for k, (r, q) in enumerate(zip(res, ref)):
        print(f"{CASES[k][0]}: batch {r.nsteps} {r.converged} {r.fhelm:.14g}; single {q.nsteps} "
              f"{q.converged} {q.fhelm:.14g}", flush=True)
        assert r.nsteps == q.nsteps and r.converged == q.converged
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-12)


def test_checkpointed_variable_cell_in_a_batch_fp64():
    # a checkpointed lamellar melt and a full-stack one with a DomainUpdater in one batch:
    # the stress passes recompute the checkpointed segments, as one at a time
    def make():
        out = []
        for K, L in ((-1, 9.0), (0, 11.0)):
            s = diblock_system(P, (64, 1, 1), 1, "lamellar", (L,), N=100, f=0.5, chi=0.2,
                               Ns=100, precision="fp64")[0]
            s.set_checkpoint(K)
            P.fieldinit(s, seed=12345, scale=0.1)
            out.append(s)
        return out

    dom = lambda: P.DomainUpdater(nper=10, nskip=5, lam=0.1)
    single = make()
    ref = []
    for s in single:
        r = P.scft(s, P.AndersonUpdater(nhist=10), domain=dom(),
                   opts=P.SCFTOptions(nsteps=120, ftol=1e-4, stol=1e-4))
        ref.append((r, float(s.cell.params[0])))
    batch_sys = make()
    res = P.Batch(batch_sys).scft(nhist=10, domain=dom(), nsteps=120, ftol=1e-4, stol=1e-4)
    for k, (r, (q, qa)) in enumerate(zip(res, ref)):
        a = float(batch_sys[k].cell.params[0])
        print(f"case {k}: batch {r.nsteps} {r.converged} fhelm {r.fhelm:.12g} L {a:.10g}; "
              f"single {q.nsteps} {q.converged} fhelm {q.fhelm:.12g} L {qa:.10g}", flush=True)
        assert r.nsteps == q.nsteps and r.converged == q.converged
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-8)
        assert a == pytest.approx(qa, rel=1e-6)


@pytest.mark.parametrize("dims,dim,crystal,params", [((64, 1, 1), 1, "lamellar", (9.0,)),
                                                     ((24, 16, 1), 2, "rectangular", (7.0, 5.0)),
                                                     ((128, 128, 1), 2, "square", (12.0,)),
                                                     ((16, 16, 16), 3, "cubic", (6.0,))])
def test_checkpointed_rqm4_matches_full_fp64(dims, dim, crystal, params):
    # RQM4 steps (mde.jl:83-104) recomputed segment by segment from the checkpoints give
    # the full-stack density!, Q, free energy, residuals and stress (the 128^2 FP64 plane
    # runs its sub-steps in one plane of LDS)
    def build():
        return diblock_system(P, dims, dim, crystal, params, N=100, f=0.35, chi=0.2, Ns=40,
                              mde="RQM4", precision="fp64")[0]

    full, ck = build(), build()
    ck.set_checkpoint(-1)
    rng = np.random.default_rng(33)
    f = {m: 0.4 * rng.standard_normal(full.shape) for m in full.mids}
    for s in (full, ck):
        for m, w in f.items():
            s.fields[m] = w
        s.residuals_update()
    assert ck.species[0].Q == pytest.approx(full.species[0].Q, rel=1e-12)
    for m in full.mids:
        assert rel(ck.density[m], full.density[m]) < 1e-12
    assert ck.free_energy() == pytest.approx(full.free_energy(), rel=1e-12)
    assert ck.field_error() == pytest.approx(full.field_error(), rel=1e-11)
    assert rel(ck.scfstress(), full.scfstress()) < 1e-10
