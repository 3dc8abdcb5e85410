"""Batched SCFT with the momentum and Euler updaters (pf_batch_scft_momentum,
pf_batch_scft_euler): every system follows scft!(sys, updater; domain) as the
one-at-a-time path runs it (src/updaters/momentum.jl:47-165, src/updaters/euler.jl:36-66,
src/simulate.jl:60-148).

Fixed cell: 2-D hexagonal melts of different grids with :MOM, :RMS and :ADAM.  Variable
cell: the 64^3 workload's composition and seed protocol on a 16^3 grid (tests/config2.py:
:MOM with lam = 0.1 and DomainUpdater(nskip = 100, nper = 10)), three seeds in one batch.
Point species: blends of a point solvent and two homopolymers (the examples.ipynb blend).
The batch's reductions differ from the one-system path in the last bits, and the
DomainUpdater's forward-difference Jacobian (dx = 1e-8) scales such differences by about
1e8, hence the looser bounds with a variable cell.
"""
import numpy as np
import pytest

from config2 import A_CELL, PROTO, build, seeded_fields
from helpers import diblock_system

P = pytest.importorskip("polyfields_hip")
pytestmark = pytest.mark.gpu


def _hex(n, a, NA, chiN, Ns, seed):
    cell = P.UnitCell(2, "hexagonal", a)
    A, B = P.Monomer(1), P.Monomer(2)
    ch = P.Multiblock([A, B], [NA, 100 - NA], [1.0, 1.0], Ns)
    itx = P.FloryInteraction()
    itx.set_interaction(1, 2, chiN / 100.0)
    s = P.FieldSystem((n, n, 1), cell, monomers=[A, B], precision="fp64")
    s.add_species(ch, 1.0)
    s.add_interaction(itx)
    P.fieldinit(s, seed=seed, scale=0.1)
    return s


HEX = [(32, 4.5, 30, 14.0, 64), (48, 5.5, 35, 18.0, 80), (64, 6.5, 40, 22.0, 100)]


@pytest.mark.parametrize("method,lam", [("MOM", 0.1), ("RMS", 0.01), ("ADAM", 0.01)])
def test_batch_momentum_fixed_cell_matches_one_at_a_time_fp64(method, lam):
    def make():
        return [_hex(*c, seed=200 + k) for k, c in enumerate(HEX)]

    single = make()
    ref = [P.scft(s, P.MomentumUpdater(method=method, lam=lam),
                  opts=P.SCFTOptions(nsteps=80, ftol=1e-5)) for s in single]
    batch_sys = make()
    res = P.Batch(batch_sys).scft(updater=P.MomentumUpdater(method=method, lam=lam),
                                  nsteps=80, ftol=1e-5)
    for k, (r, q) in enumerate(zip(res, ref)):
        print(f"{method} case {k}: batch {r.nsteps} {r.converged} {r.fhelm:.14g} {r.ferr:.6g}; "
              f"single {q.nsteps} {q.converged} {q.fhelm:.14g} {q.ferr:.6g}", flush=True)
        assert r.converged == q.converged
        assert r.nsteps == q.nsteps
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-12)
        assert r.ferr == pytest.approx(q.ferr, rel=1e-9)
        for m in (1, 2):
            np.testing.assert_allclose(np.asarray(batch_sys[k].fields[m]),
                                       np.asarray(single[k].fields[m]), rtol=0, atol=1e-10)


def _cfg2_16(seed):
    s, cell, ch = build(P, grid=16, a=A_CELL, precision="fp64")
    for m, w in seeded_fields(s.shape, seed).items():
        s.fields[m] = w
    return s


def test_batch_momentum_variable_cell_matches_one_at_a_time_fp64():
    seeds = [2024, 7, 11]
    kw = dict(nsteps=400, ftol=PROTO["ftol"])
    single = [_cfg2_16(sd) for sd in seeds]
    ref = []
    for s in single:
        r = P.scft(s, P.MomentumUpdater(method="MOM", lam=PROTO["lam"]),
                   domain=P.DomainUpdater(nskip=PROTO["nskip"], nper=PROTO["nper"]),
                   opts=P.SCFTOptions(stol=PROTO["stol"], **kw))
        ref.append((r, float(s.cell.params[0])))
    batch_sys = [_cfg2_16(sd) for sd in seeds]
    res = P.Batch(batch_sys).scft(updater=P.MomentumUpdater(method="MOM", lam=PROTO["lam"]),
                                  domain=P.DomainUpdater(nskip=PROTO["nskip"],
                                                         nper=PROTO["nper"]),
                                  stol=PROTO["stol"], **kw)
    for k, (r, (q, qa)) in enumerate(zip(res, ref)):
        a = float(batch_sys[k].cell.params[0])
        print(f"seed {seeds[k]}: batch {r.nsteps} {r.converged} fhelm {r.fhelm:.12g} a {a:.10g} "
              f"serr {r.serr:.3g}; single {q.nsteps} {q.converged} fhelm {q.fhelm:.12g} "
              f"a {qa:.10g} serr {q.serr:.3g}", flush=True)
        assert r.converged == q.converged
        assert r.nsteps == q.nsteps
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-8)
        assert a == pytest.approx(qa, rel=1e-6)
        assert a != pytest.approx(A_CELL, rel=1e-3)   # the cell moved


@pytest.mark.parametrize("make_updater", [
    lambda: P.MomentumUpdater(method="NAG", lam=0.1, gamma=0.9),
    lambda: P.EulerUpdater(method="PEC", lam=0.05),
    lambda: P.EulerUpdater(method="PECE", lam=0.05)], ids=["NAG", "PEC", "PECE"])
def test_batch_lookahead_updaters_match_one_at_a_time_fp64(make_updater):
    # NAG (momentum.jl:141-165) and the Euler predictor-corrector (euler.jl:36-66) run
    # residuals! inside their step: at NAG's look-ahead fields, at the Euler predictor
    # (and, for PECE, at the corrected fields); the step after it moves along the
    # residual the previous step left
    def make():
        return [_hex(*c, seed=400 + k) for k, c in enumerate(HEX)]

    single = make()
    ref = [P.scft(s, make_updater(), opts=P.SCFTOptions(nsteps=80, ftol=1e-5)) for s in single]
    batch_sys = make()
    res = P.Batch(batch_sys).scft(updater=make_updater(), nsteps=80, ftol=1e-5)
    for k, (r, q) in enumerate(zip(res, ref)):
        print(f"case {k}: batch {r.nsteps} {r.converged} {r.fhelm:.14g} {r.ferr:.6g}; "
              f"single {q.nsteps} {q.converged} {q.fhelm:.14g} {q.ferr:.6g}", flush=True)
        assert r.converged == q.converged
        assert r.nsteps == q.nsteps
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-12)
        assert r.ferr == pytest.approx(q.ferr, rel=1e-9)
        for m in (1, 2):
            np.testing.assert_allclose(np.asarray(batch_sys[k].fields[m]),
                                       np.asarray(single[k].fields[m]), rtol=0, atol=1e-10)


def test_batch_euler_variable_cell_matches_one_at_a_time_fp64():
    # EulerUpdater(:PECE) with a DomainUpdater on two lamellar melts: the cell update
    # ends with residuals! at the new cell, which the next predictor moves along
    def make():
        out = []
        for L, chi, seed in ((10.0, 0.2, 12345), (12.0, 0.25, 7)):
            s = diblock_system(P, (64, 1, 1), 1, "lamellar", (L,), N=100, f=0.5, chi=chi,
                               Ns=100, precision="fp64")[0]
            P.fieldinit(s, seed=seed, scale=0.1)
            out.append(s)
        return out

    dom = lambda: P.DomainUpdater(nper=10, nskip=5, lam=0.1)
    single = make()
    ref = []
    for s in single:
        r = P.scft(s, P.EulerUpdater(method="PECE", lam=0.1), domain=dom(),
                   opts=P.SCFTOptions(nsteps=150, ftol=1e-4, stol=1e-4))
        ref.append((r, float(s.cell.params[0])))
    batch_sys = make()
    res = P.Batch(batch_sys).scft(updater=P.EulerUpdater(method="PECE", lam=0.1),
                                  domain=dom(), nsteps=150, ftol=1e-4, stol=1e-4)
    for k, (r, (q, qa)) in enumerate(zip(res, ref)):
        a = float(batch_sys[k].cell.params[0])
        print(f"case {k}: batch {r.nsteps} {r.converged} fhelm {r.fhelm:.12g} L {a:.10g}; "
              f"single {q.nsteps} {q.converged} fhelm {q.fhelm:.12g} L {qa:.10g}", flush=True)
        assert r.converged == q.converged
        assert r.nsteps == q.nsteps
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-8)
        assert a == pytest.approx(qa, rel=1e-6)


def _blend(L, chi_ab, seed):
    # point solvent + two homopolymers with v = 2 and a chi triangle (the examples.ipynb
    # blend), 1-D
    cell = P.UnitCell(1, "lamellar", L)
    a, b, c = P.Monomer(1, vol=2.0), P.Monomer(2, vol=2.0), P.Monomer(3, vol=1.0)
    s = P.FieldSystem((128, 1, 1), cell, monomers=[a, b, c], precision="fp64")
    s.add_species(P.Point(c), 0.4)
    s.add_species(P.Homopolymer(a, 100, 1.0), 0.3)
    s.add_species(P.Homopolymer(b, 100, 1.0), 0.3)
    itx = P.FloryInteraction()
    itx.set_interaction(1, 2, chi_ab)
    itx.set_interaction(1, 3, 1.25)
    itx.set_interaction(2, 3, 1.25)
    s.add_interaction(itx)
    rng = np.random.default_rng(seed)
    for m in (1, 2, 3):
        s.fields[m] = 0.05 * rng.standard_normal((1, 1, 128))
    return s


def test_batch_blends_with_point_species_match_one_at_a_time_fp64():
    cases = [(20.0, 0.5, 1), (24.0, 0.6, 2), (18.0, 0.45, 3)]
    single = [_blend(*c) for c in cases]
    ref = [P.scft(s, P.MomentumUpdater(method="MOM", lam=0.01),
                  opts=P.SCFTOptions(nsteps=300, ftol=1e-6)) for s in single]
    batch_sys = [_blend(*c) for c in cases]
    res = P.Batch(batch_sys).scft(updater=P.MomentumUpdater(method="MOM", lam=0.01),
                                  nsteps=300, ftol=1e-6)
    for k, (r, q) in enumerate(zip(res, ref)):
        print(f"blend {k}: batch {r.nsteps} {r.converged} {r.fhelm:.14g}; single {q.nsteps} "
              f"{q.converged} {q.fhelm:.14g}", flush=True)
        assert r.converged == q.converged
        assert r.nsteps == q.nsteps
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-12)
        for m in (1, 2, 3):
            np.testing.assert_allclose(np.asarray(batch_sys[k].density[m]),
                                       np.asarray(single[k].density[m]), rtol=0, atol=1e-11)


def test_batch_sweeps_fill_the_stacks_as_density_does():
    # pf_batch_sweeps runs the propagate! chain alone: the q and q+ slices equal those of
    # a full batched density! of the same fields (FP32 64^3 and FP64 2-D)
    for prec, make in (("fp32", lambda k: build(P, grid=64, a=A_CELL, precision="fp32")[0]),
                       ("fp64", lambda k: _hex(*HEX[k % 3], seed=300 + k))):
        a = [make(k) for k in range(3)]
        b = [make(k) for k in range(3)]
        P.Batch(a).density_update()
        P.Batch(b).sweeps()
        for sa, sb in zip(a, b):
            ca, cb = sa.species[0], sb.species[0]
            for s in (0, 1, ca.Ns // 2, ca.Ns):
                np.testing.assert_array_equal(cb.q[s], ca.q[s])
                np.testing.assert_array_equal(cb.qc[s], ca.qc[s])
        print(f"{prec}: sweeps() stacks equal density_update() stacks", flush=True)


def test_batch_rqm4_matches_one_at_a_time_fp64():
    # RQM4 systems (mde.jl:83-104) in a batch beside an RK2 one: 1-D, 2-D and 3-D grids and
    # a 128 x 128 FP64 plane (one plane of LDS), the densities of one batched density! and
    # an Anderson scft! equal one-at-a-time
    def make():
        out = []
        for k, (dims, dim, cr, prm) in enumerate([((64, 1, 1), 1, "lamellar", (12.0,)),
                                                  ((32, 32, 1), 2, "hexagonal", (8.0,)),
                                                  ((16, 16, 16), 3, "cubic", (7.0,)),
                                                  ((32, 32, 32), 3, "cubic", (8.0,)),
                                                  ((128, 128, 1), 2, "square", (12.0,))]):
            cell = P.UnitCell(dim, cr, *prm)
            A, B = P.Monomer(1), P.Monomer(2)
            ch = P.Multiblock([A, B], [35, 65], [1.0, 1.0], 40)
            itx = P.FloryInteraction()
            itx.set_interaction(1, 2, 0.2)
            s = P.FieldSystem(dims, cell, monomers=[A, B], precision="fp64", mde="RQM4")
            s.add_species(ch, 1.0)
            s.add_interaction(itx)
            P.fieldinit(s, seed=500 + k, scale=0.1)
            out.append(s)
        out.append(_hex(*HEX[0], seed=77))   # an RK2 system in the same batch
        return out

    single, batch_sys = make(), make()
    for s in single:
        s.density_update()
    P.Batch(batch_sys).density_update()
    for a, b in zip(single, batch_sys):
        for m in (1, 2):
            np.testing.assert_array_equal(np.asarray(b.density[m]), np.asarray(a.density[m]))
    ref = [P.scft(s, P.AndersonUpdater(nhist=10), opts=P.SCFTOptions(nsteps=40, ftol=1e-7))
           for s in single]
    res = P.Batch(batch_sys).scft(nhist=10, nsteps=40, ftol=1e-7)
    for k, (r, q) in enumerate(zip(res, ref)):
        print(f"rqm4 case {k}: batch {r.nsteps} {r.fhelm:.14g}; single {q.nsteps} {q.fhelm:.14g}",
              flush=True)
        assert r.nsteps == q.nsteps and r.converged == q.converged
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-12)


def _grand(mu, seed):
    # diblock + homopolymer + point species at fixed chemical potentials (grand ensemble)
    cell = P.UnitCell(2, "rectangular", 6.0, 4.5)
    a, b, c = P.Monomer(1), P.Monomer(2), P.Monomer(3)
    s = P.FieldSystem((16, 12, 1), cell, monomers=[a, b, c], ensemble="grand",
                      precision="fp64")
    for sp, v in zip((P.Diblock(a, b, 40, 0.5, 1.0, 1.0, 20), P.Homopolymer(a, 20, 1.0, 10),
                      P.Point(c)), mu):
        s.add_species(sp, v)
    fl = P.FloryInteraction()
    fl.set_interaction(1, 2, 0.4)
    fl.set_interaction(1, 3, 0.2)
    s.add_interaction(fl)
    rng = np.random.default_rng(seed)
    for m in (1, 2, 3):
        s.fields[m] = 0.1 * rng.standard_normal((1, 12, 16))
    return s


@pytest.mark.parametrize("make_updater", [
    lambda: P.MomentumUpdater(method="MOM", lam=0.05),
    lambda: P.MomentumUpdater(method="NAG", lam=0.05, gamma=0.9),
    lambda: P.EulerUpdater(method="PECE", lam=0.05)], ids=["MOM", "NAG", "PECE"])
def test_batch_grand_ensemble_matches_one_at_a_time_fp64(make_updater):
    # grand-ensemble systems in a batch (momentum and Euler updaters; the reference refuses
    # the grand ensemble for Anderson mixing only, anderson.jl:61-63): the free energy
    # takes each system's species fractions phi = Q e^mu from its last pass
    cases = [((0.1, -0.3, -1.2), 5), ((0.3, -0.5, -1.0), 6), ((-0.2, -0.1, -1.4), 7)]
    single = [_grand(*c) for c in cases]
    ref = [P.scft(s, make_updater(), opts=P.SCFTOptions(nsteps=60, ftol=1e-6)) for s in single]
    batch_sys = [_grand(*c) for c in cases]
    res = P.Batch(batch_sys).scft(updater=make_updater(), nsteps=60, ftol=1e-6)
    for k, (r, q) in enumerate(zip(res, ref)):
        print(f"grand {k}: batch {r.nsteps} {r.converged} {r.fhelm:.14g}; single {q.nsteps} "
              f"{q.converged} {q.fhelm:.14g}", flush=True)
        assert r.converged == q.converged
        assert r.nsteps == q.nsteps
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-12)
        for m in (1, 2, 3):
            np.testing.assert_allclose(np.asarray(batch_sys[k].density[m]),
                                       np.asarray(single[k].density[m]), rtol=0, atol=1e-11)


def test_batch_anderson_refuses_grand_ensemble():
    with pytest.raises(P.PolyFieldsError, match="canonical"):
        P.Batch([_grand((0.1, -0.3, -1.2), 5)]).scft(nhist=5, nsteps=3)
