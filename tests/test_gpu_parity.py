"""HIP path (through the C ABI) against the FP64 CPU oracle on the same inputs.

FP64 ("plain") mode: densities, partition functions, stress, residuals and free energy
agree to rounding (the transforms differ in algorithm, not in precision); bounds are
written per check.  FP32 mode reports its error against FP64 with its own bound."""
import numpy as np
import pytest

import oracle as O
from helpers import cos_fields, diblock_system, load_seed

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")


def both(build, fields, precision="fp64"):
    so = build(O)
    sg = build(P, precision)
    O.fieldinit(so[0], fields=fields)
    P.fieldinit(sg[0], fields=fields)
    return so, sg


def rand_fields(shape, mids=(1, 2), scale=0.3, seed=0):
    rng = np.random.default_rng(seed)
    return {m: scale * rng.standard_normal(shape) for m in mids}


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    ("lam32", (32, 1, 1), 1, "lamellar", (10.0,), dict(f=0.5)),
    ("hex16", (16, 16, 1), 2, "hexagonal", (6.0,), dict(f=0.35)),
    ("rect24x12", (24, 12, 1), 2, "rectangular", (7.0, 4.0), dict(f=0.3)),
    ("cubic8", (8, 8, 8), 3, "cubic", (4.0,), dict(f=0.35)),
    ("ortho12x6x6", (12, 6, 6), 3, "orthorhombic", (5.0, 4.0, 3.5), dict(f=0.4)),
    ("tric8x6x4", (8, 6, 4), 3, "triclinic", (6.0, 5.0, 4.0, 1.3, 1.4, 1.2), dict(f=0.5)),
    # sizes served by the compile-time specialised kernels
    ("lam128", (128, 1, 1), 1, "lamellar", (12.0,), dict(f=0.5)),
    ("hex48", (48, 48, 1), 2, "hexagonal", (7.0,), dict(f=0.35)),
    ("hex96", (96, 96, 1), 2, "hexagonal", (9.0,), dict(f=0.3)),
    ("cubic32", (32, 32, 32), 3, "cubic", (5.0,), dict(f=0.35)),
    ("cubic64", (64, 64, 64), 3, "cubic", (8.0,), dict(f=0.35)),
]


@pytest.mark.parametrize("name,dims,dim,crystal,params,kw", CASES, ids=[c[0] for c in CASES])
def test_density_stress_residuals_fp64(name, dims, dim, crystal, params, kw):
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        return diblock_system(mod, dims, dim, crystal, params, Ns=20 if dims[2] == 64 else 40,
                              **kw, **extra)

    shape = (dims[2], dims[1], dims[0])
    (so, _, co), (sg, _, cg) = both(build, rand_fields(shape))
    assert cg.Q == pytest.approx(co.Q, rel=1e-12)
    for s in (0, 1, co.Ns // 2, co.Ns):
        assert rel(cg.q[s], co.q[s]) < 1e-12
        assert rel(cg.qc[s], co.qc[s]) < 1e-12
    for m in (1, 2):
        assert rel(sg.density[m], so.density[m]) < 1e-12
        assert rel(sg.residuals[m], so.residuals[m]) < 1e-11
        assert rel(sg.fields[m], so.fields[m]) < 1e-14
    assert sg.field_error() == pytest.approx(so.field_error(), rel=1e-11)
    assert sg.free_energy() == pytest.approx(so.free_energy(), rel=1e-12, abs=1e-14)
    np.testing.assert_allclose(sg.scfstress(), so.scfstress(), rtol=1e-10, atol=1e-15)


def test_blend_point_homopolymers_fp64():
    # Point + two homopolymers with v = 2 and a chi triangle (examples.ipynb blend).
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        cell = mod.UnitCell(1, "lamellar", 20.0)
        a, b, c = mod.Monomer(1, vol=2.0), mod.Monomer(2, vol=2.0), mod.Monomer(3, vol=1.0)
        sys = mod.FieldSystem((64, 1, 1), cell, monomers=[a, b, c], **extra)
        sys.add_species(mod.Point(c), 0.4)
        sys.add_species(mod.Homopolymer(a, 100, 1.0), 0.3)
        sys.add_species(mod.Homopolymer(b, 100, 1.0), 0.3)
        itx = mod.FloryInteraction()
        itx.set_interaction(1, 2, 0.5)
        itx.set_interaction(1, 3, 1.25)
        itx.set_interaction(2, 3, 1.25)
        sys.add_interaction(itx)
        return sys, cell, None

    (so, _, _), (sg, _, _) = both(build, rand_fields((1, 1, 64), mids=(1, 2, 3), scale=0.05))
    for m in (1, 2, 3):
        assert rel(sg.density[m], so.density[m]) < 1e-12
        assert rel(sg.residuals[m], so.residuals[m]) < 1e-10
    for i in range(3):
        assert sg.species[i].Q == pytest.approx(so.species[i].Q, rel=1e-12)
    assert sg.free_energy() == pytest.approx(so.free_energy(), rel=1e-12)
    assert sg.free_energy_bulk() == pytest.approx(so.free_energy_bulk(), rel=1e-14)


CONSTRAINTS = [("strict", 0.0), ("compressible", 0.5), ("unconstrained", -1.0)]
CGRIDS = [((32, 1, 1), 1, "lamellar", (10.0,)), ((8, 8, 8), 3, "cubic", (4.0,))]


@pytest.mark.parametrize("label,inv_zeta", CONSTRAINTS, ids=[c[0] for c in CONSTRAINTS])
@pytest.mark.parametrize("dims,dim,crystal,params", CGRIDS, ids=["lam32", "cubic8"])
def test_compressibility_and_edwards_residuals_fp64(label, inv_zeta, dims, dim, crystal,
                                                    params):
    # the three constraint modes of compress.jl:37-61 (eta from the density sum, with
    # 1 / inv_zeta, or none) with Edwards excluded-volume potentials (edwards.jl:73-87) on
    # both monomers beside the Flory pair: potentials, residuals, field error, free energy
    # and a short Anderson run against the oracle
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        sys, cell, ch = diblock_system(mod, dims, dim, crystal, params, f=0.4, Ns=40,
                                       compress=mod.Compressibility(inv_zeta), **extra)
        ed = mod.EdwardsInteraction()
        ed.set_interaction(1, 0.8)
        ed.set_interaction(2, 0.3)
        sys.add_interaction(ed)
        return sys, cell, ch

    shape = (dims[2], dims[1], dims[0])
    (so, _, co), (sg, _, cg) = both(build, rand_fields(shape, scale=0.2, seed=9))
    assert cg.Q == pytest.approx(co.Q, rel=1e-12)
    for m in (1, 2):
        assert rel(sg.density[m], so.density[m]) < 1e-12
        assert rel(sg.potentials[m], so.potentials[m]) < 1e-12
        assert rel(sg.residuals[m], so.residuals[m]) < 1e-11
    assert sg.field_error() == pytest.approx(so.field_error(), rel=1e-11)
    assert sg.free_energy() == pytest.approx(so.free_energy(), rel=1e-12, abs=1e-14)
    assert sg.free_energy_bulk() == pytest.approx(so.free_energy_bulk(), rel=1e-14)
    opts = dict(nsteps=15, ftol=1e-14)
    upd = dict(nhist=4, iter_euler=3, lam_euler=0.05)
    ro = O.scft(so, O.AndersonUpdater(**upd), opts=O.SCFTOptions(**opts))
    rg = P.scft(sg, P.AndersonUpdater(**upd), opts=P.SCFTOptions(**opts))
    assert rg.nsteps == ro.nsteps
    assert rg.fhelm == pytest.approx(ro.fhelm, rel=1e-9)
    assert rg.ferr == pytest.approx(ro.ferr, rel=1e-8)
    for m in (1, 2):
        assert rel(sg.fields[m], so.fields[m]) < 1e-8


def test_anderson_trajectory_fp64():
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        return diblock_system(mod, (32, 32, 1), 2, "hexagonal", (9.0,), f=0.3, chi=0.25,
                              Ns=50, **extra)

    (so, _, _), (sg, _, _) = both(build, rand_fields((1, 32, 32), scale=0.5))
    uo, ug = O.AndersonUpdater(nhist=5), P.AndersonUpdater(nhist=5)
    uo.setup(so)
    ug.setup(sg)
    for it in range(14):
        uo.step()
        ug.step()
        fo, fg = so.field_error(), sg.field_error()
        eo, eg = so.free_energy(), sg.free_energy()
        assert fg == pytest.approx(fo, rel=1e-8), it
        assert eg == pytest.approx(eo, rel=1e-10), it


def test_fixed_cell_converged_state_fp64():
    # Fixed cell: the converged state is a smooth function of the inputs, so plain FP64
    # must reproduce the oracle's converged fhelm and densities to 1e-9.
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        return diblock_system(mod, (64, 1, 1), 1, "lamellar", (16.5,), chi=0.2, **extra)

    (so, _, _), (sg, _, _) = both(build, cos_fields((1, 1, 64)))
    opts = dict(nsteps=400, ftol=1e-10)
    ro = O.scft(so, O.AndersonUpdater(nhist=10), opts=O.SCFTOptions(**opts))
    rg = P.scft(sg, P.AndersonUpdater(nhist=10), opts=P.SCFTOptions(**opts))
    assert rg.converged and ro.converged
    assert rg.fhelm == pytest.approx(ro.fhelm, rel=1e-9)
    for m in (1, 2):
        assert rel(sg.density[m], so.density[m]) < 1e-8


def test_variable_cell_fixed_point_fp64():
    # Variable cell.  The reference's forward-difference Jacobian (dx = 1e-8,
    # domain.jl:101) amplifies last-bit rounding ~1e8 times, and its first Newton
    # iteration differences stresses of two field states, so loose-tolerance trajectories
    # are not reproducible even by the oracle under 1-ulp FFT jitter.  The fixed point is:
    # with serr refreshed every step (nskip = 1) and ftol = stol = 1e-9, jittered oracle
    # runs agree to 9e-8 in fhelm and 6e-7 in L (bounded by stol / J).
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        return diblock_system(mod, (64, 1, 1), 1, "lamellar", (16.0,), chi=0.2, **extra)

    (so, co, _), (sg, cg, _) = both(build, cos_fields((1, 1, 64)))
    opts = dict(nsteps=600, ftol=1e-9, stol=1e-9)
    ro = O.scft(so, O.AndersonUpdater(nhist=10), domain=O.DomainUpdater(nskip=1, nper=5),
                opts=O.SCFTOptions(**opts))
    rg = P.scft(sg, P.AndersonUpdater(nhist=10), domain=P.DomainUpdater(nskip=1, nper=5),
                opts=P.SCFTOptions(**opts))
    assert rg.converged and ro.converged
    assert rg.fhelm == pytest.approx(ro.fhelm, rel=3e-7)
    assert cg.params[0] == pytest.approx(co.params[0], rel=3e-6)


def test_variable_cell_readme_settings_converge():
    # The README's own settings (nskip = 5, nper = 20, ftol = stol = 1e-4) converge on the
    # GPU to a state meeting both tolerances.
    sys, cell, _ = diblock_system(P, (64, 1, 1), 1, "lamellar", (10.0,), chi=0.2)
    P.fieldinit(sys, fields=cos_fields(sys.shape))
    dom = P.DomainUpdater(nskip=5, nper=20)
    r = P.scft(sys, P.AndersonUpdater(nhist=10), domain=dom,
               opts=P.SCFTOptions(nsteps=1500, ftol=1e-4, stol=1e-4))
    assert r.converged and r.ferr < 1e-4 and r.serr <= 1e-4


def test_readme_state_on_gpu():
    sys, cell, ch = diblock_system(P, (64, 1, 1), 1, "lamellar", (16.528709413909947,), chi=0.2)
    P.fieldinit(sys, fields=cos_fields(sys.shape))
    res = P.scft(sys, P.AndersonUpdater(nhist=10), opts=P.SCFTOptions(nsteps=400, ftol=1e-10))
    assert res.converged
    assert res.fhelm == pytest.approx(0.0812381288285528, abs=5e-8)


def test_hex_cylinder_notebook_state_on_gpu():
    d, fields = load_seed("cyl2d")
    sys, cell, ch = diblock_system(P, (64, 64, 1), 2, "hexagonal", (float(d["cell_params"][0]),),
                                   f=0.3, chi=0.3)
    P.fieldinit(sys, fields=fields)
    res = P.scft(sys, P.AndersonUpdater(nhist=10), opts=P.SCFTOptions(nsteps=150, ftol=1e-7))
    assert res.ferr < 2e-7
    assert res.fhelm == pytest.approx(0.10659231121049459, abs=1e-7)


FP32_CASES = [CASES[i] for i in (0, 1, 2, 3, 7, 8, 9, 10)]


@pytest.mark.parametrize("name,dims,dim,crystal,params,kw", FP32_CASES, ids=[c[0] for c in FP32_CASES])
def test_fp32_propagator_error_vs_fp64(name, dims, dim, crystal, params, kw):
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        return diblock_system(mod, dims, dim, crystal, params, Ns=40, **kw, **extra)

    shape = (dims[2], dims[1], dims[0])
    (so, _, co), (sg, _, cg) = both(build, rand_fields(shape), precision="fp32")
    assert cg.Q == pytest.approx(co.Q, rel=2e-5)
    for m in (1, 2):
        assert rel(sg.density[m], so.density[m]) < 2e-5
    assert sg.free_energy() == pytest.approx(so.free_energy(), abs=2e-6)
    np.testing.assert_allclose(sg.scfstress(), so.scfstress(), rtol=1e-3, atol=1e-8)


@pytest.mark.parametrize("kind", ["PEC", "PECE", "MOM", "RMS", "ADAM", "NAG"])
def test_euler_momentum_updaters_fp64(kind):
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        return diblock_system(mod, (32, 1, 1), 1, "lamellar", (10.0,), **extra)

    (so, _, _), (sg, _, _) = both(build, rand_fields((1, 1, 32), scale=0.1))
    if kind in ("PEC", "PECE"):
        uo, ug = O.EulerUpdater(method=kind, lam=0.005), P.EulerUpdater(method=kind, lam=0.005)
    else:
        uo = O.MomentumUpdater(method=kind, lam=0.005)
        ug = P.MomentumUpdater(method=kind, lam=0.005)
    uo.setup(so)
    ug.setup(sg)
    for _ in range(10):
        uo.step()
        ug.step()
    assert sg.free_energy() == pytest.approx(so.free_energy(), rel=1e-10)
    for m in (1, 2):
        assert rel(sg.fields[m], so.fields[m]) < 1e-10


@pytest.mark.parametrize("kind", ["MOM", "RMS", "ADAM", "NAG"])
def test_scft_momentum_matches_oracle_fp64(kind):
    # scft! queues the next step's residuals! early for MOM / RMS / ADAM (their step starts
    # with residuals!), not for NAG; the loop must give the oracle's values either way
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        return diblock_system(mod, (32, 1, 1), 1, "lamellar", (10.0,), **extra)

    (so, _, _), (sg, _, _) = both(build, rand_fields((1, 1, 32), scale=0.1, seed=5))
    opts = dict(nsteps=12, ftol=1e-14)
    ro = O.scft(so, O.MomentumUpdater(method=kind, lam=0.005), opts=O.SCFTOptions(**opts))
    rg = P.scft(sg, P.MomentumUpdater(method=kind, lam=0.005), opts=P.SCFTOptions(**opts))
    assert rg.nsteps == ro.nsteps and rg.converged == ro.converged
    assert rg.fhelm == pytest.approx(ro.fhelm, rel=1e-10)
    assert rg.ferr == pytest.approx(ro.ferr, rel=1e-8)
    for m in (1, 2):
        assert rel(sg.fields[m], so.fields[m]) < 1e-10


@pytest.mark.parametrize("kind,lam", [("MOM", 0.05), ("ADAM", 0.005), ("NAG", 0.05)])
def test_scft_momentum_converges_like_oracle_fp64(kind, lam):
    # runs that reach ftol = 1e-3 before nsteps (176-222 steps on the oracle): the
    # converged exit, with the queued residuals! of MOM / ADAM discarded, must stop at the
    # oracle's step with its fields
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        return diblock_system(mod, (32, 1, 1), 1, "lamellar", (10.0,), **extra)

    (so, _, _), (sg, _, _) = both(build, rand_fields((1, 1, 32), scale=0.1, seed=5))
    opts = dict(nsteps=300, ftol=1e-3)
    ro = O.scft(so, O.MomentumUpdater(method=kind, lam=lam), opts=O.SCFTOptions(**opts))
    rg = P.scft(sg, P.MomentumUpdater(method=kind, lam=lam), opts=P.SCFTOptions(**opts))
    assert ro.converged and ro.nsteps < 300
    assert rg.converged and rg.nsteps == ro.nsteps
    assert rg.fhelm == pytest.approx(ro.fhelm, rel=1e-9)
    assert rg.ferr == pytest.approx(ro.ferr, rel=1e-6)
    for m in (1, 2):
        assert rel(sg.fields[m], so.fields[m]) < 1e-8


def test_errors_are_raised_not_hidden():
    sys, cell, ch = diblock_system(P, (16, 1, 1), 1, "lamellar", (10.0,))
    with pytest.raises(P.PolyFieldsError, match="initialized"):
        sys.validate()
    with pytest.raises(P.PolyFieldsError, match="descending"):
        P.FieldSystem((1, 16, 1), P.UnitCell(1, "lamellar", 5.0))


def _lam_build(L):
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        return diblock_system(mod, (64, 1, 1), 1, "lamellar", (L,), chi=0.2, **extra)
    return build


def test_cell_change_then_density_fp64():
    # Build at L=10, move the cell to 13.7 (update!(cell)), evaluate; the oracle is built
    # at 13.7 directly.
    so, co, _ = _lam_build(13.7)(O)
    sg, cg, _ = _lam_build(10.0)(P, "fp64")
    f = cos_fields((1, 1, 64))
    O.fieldinit(so, fields=f)
    cg.params = [13.7]
    P.fieldinit(sg, fields=f)
    for m in (1, 2):
        assert rel(sg.density[m], so.density[m]) < 1e-12
    np.testing.assert_allclose(sg.scfstress(), so.scfstress(), rtol=1e-10)


def test_domain_step_from_identical_state_fp64():
    (so, co, _), (sg, cg, _) = both(_lam_build(15.0), cos_fields((1, 1, 64)))
    do, dg = O.DomainUpdater(nskip=5, nper=20), P.DomainUpdater(nskip=5, nper=20)
    do.setup(so)
    dg.setup(sg)
    np.testing.assert_allclose(dg.stress, do.stress, rtol=1e-10)
    L_prev = 15.0
    for it in range(3):
        do.step(nsteps=1)
        dg.step(nsteps=1)
        # the forward-difference Jacobian carries ~6e-4 relative noise from last-bit
        # rounding (measured on the oracle under 1-ulp FFT jitter); the step inherits it
        step_o = co.params[0] - L_prev
        assert abs(cg.params[0] - co.params[0]) < 2e-3 * abs(step_o), it
        L_prev = co.params[0]
        cg.params = [co.params[0]]       # re-synchronise the cells for the next step
        P.fieldinit(sg, fields={m: so.fields[m] for m in (1, 2)})
        dg.scfstress_update()
        np.testing.assert_allclose(dg.stress, do.stress, rtol=1e-9, err_msg=str(it))


def _hex_system(mod, n, a, fA_N, chiN, Ns, precision=None):
    extra = {} if precision is None else {"precision": precision}
    cell = mod.UnitCell(2, "hexagonal", a)
    A, B = mod.Monomer(1), mod.Monomer(2)
    ch = mod.Multiblock([A, B], [fA_N, 100 - fA_N], [1.0, 1.0], Ns)
    itx = mod.FloryInteraction()
    itx.set_interaction(1, 2, chiN / 100.0)
    sys = mod.FieldSystem((n, n, 1), cell, monomers=[A, B], **extra)
    sys.add_species(ch, 1.0)
    sys.add_interaction(itx)
    return sys


BATCH = [(32, 4.5, 30, 14.0, 64), (48, 5.5, 35, 18.0, 80), (64, 6.5, 40, 22.0, 100),
         (96, 7.5, 28, 26.0, 128), (32, 5.0, 45, 12.0, 70)]


def test_batch_density_mixed_grids_fp64():
    sg = [_hex_system(P, *c, precision="fp64") for c in BATCH]
    so = [_hex_system(O, *c) for c in BATCH]
    for k, (g, o) in enumerate(zip(sg, so)):
        f = rand_fields((1, BATCH[k][0], BATCH[k][0]), seed=k)
        for m in (1, 2):
            g.fields[m] = f[m]
            o.fields[m] = f[m]
    b = P.Batch(sg)
    b.density_update()
    for g, o in zip(sg, so):
        o.density_update()
        assert g.species[0].Q == pytest.approx(o.species[0].Q, rel=1e-12)
        for m in (1, 2):
            assert rel(g.density[m], o.density[m]) < 1e-12


def test_batch_scft_matches_one_at_a_time_fp64():
    def make():
        out = []
        for k, c in enumerate(BATCH):
            s = _hex_system(P, *c, precision="fp64")
            P.fieldinit(s, seed=100 + k, scale=0.1)
            out.append(s)
        return out

    single = make()
    ref = [P.scft(s, P.AndersonUpdater(nhist=10), opts=P.SCFTOptions(nsteps=60, ftol=1e-6))
           for s in single]
    batch_sys = make()
    res = P.Batch(batch_sys).scft(nhist=10, nsteps=60, ftol=1e-6)
    for r, q in zip(res, ref):
        assert r.converged == q.converged
        assert r.nsteps == q.nsteps
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-13)
        assert r.ferr == pytest.approx(q.ferr, rel=1e-12)


def test_batch_3d_density_fp32_matches_single():
    def make():
        return diblock_system(P, (32, 32, 32), 3, "cubic", (5.0,), f=0.35, Ns=40,
                              precision="fp32")[0]
    singles = [make() for _ in range(3)]
    batched = [make() for _ in range(3)]
    for k in range(3):
        f = rand_fields((32, 32, 32), seed=k)
        for m in (1, 2):
            singles[k].fields[m] = f[m]
            batched[k].fields[m] = f[m]
        singles[k].density_update()
    P.Batch(batched).density_update()
    for a, b in zip(singles, batched):
        assert b.species[0].Q == pytest.approx(a.species[0].Q, rel=1e-6)
        for m in (1, 2):
            assert rel(b.density[m], a.density[m]) < 1e-6


@pytest.mark.parametrize("dims,dim,crystal,params", [((32, 1, 1), 1, "lamellar", (9.0,)),
                                                     ((24, 12, 1), 2, "rectangular", (7.0, 4.0)),
                                                     ((8, 8, 8), 3, "cubic", (6.0,)),
                                                     ((12, 8, 6), 3, "orthorhombic",
                                                      (7.0, 5.0, 4.0)),
                                                     # compile-time 3-D slabs and pencils
                                                     ((32, 32, 32), 3, "cubic", (8.0,)),
                                                     ((64, 64, 64), 3, "cubic", (12.0,)),
                                                     # compile-time planes (1024 and 512
                                                     # threads)
                                                     ((64, 64, 1), 2, "hexagonal", (9.0,)),
                                                     ((96, 96, 1), 2, "square", (10.0,)),
                                                     # FP64 planes too large for two in LDS:
                                                     # one plane, the sub-steps in turn
                                                     # (compile-time 128^2, generic 108x96)
                                                     ((128, 128, 1), 2, "square", (12.0,)),
                                                     ((108, 96, 1), 2, "rectangular",
                                                      (11.0, 10.0)),
                                                     ((128, 128, 4), 3, "tetragonal",
                                                      (12.0, 3.0))])
def test_rqm4_propagator_fp64(dims, dim, crystal, params):
    def build(mod, precision=None):
        extra = {} if precision is None else {"precision": precision}
        return diblock_system(mod, dims, dim, crystal, params, Ns=40, mde="RQM4", **extra)

    (so, _, co), (sg, _, cg) = both(build, rand_fields((dims[2], dims[1], dims[0])))
    assert cg.Q == pytest.approx(co.Q, rel=1e-12)
    for s in (1, co.Ns // 2, co.Ns):
        assert rel(cg.q[s], co.q[s]) < 1e-12
        assert rel(cg.qc[s], co.qc[s]) < 1e-12
    for m in (1, 2):
        assert rel(sg.density[m], so.density[m]) < 1e-12
    assert sg.free_energy() == pytest.approx(so.free_energy(), rel=1e-12)


def test_grand_ensemble_parity_fp64():
    # diblock + homopolymer + point species at fixed chemical potentials
    # (system.jl:248-259, 317-340 without the canonical mean subtraction)
    def make(mod):
        cell = mod.UnitCell(2, "rectangular", 6.0, 4.5)
        a, b, c = mod.Monomer(1), mod.Monomer(2), mod.Monomer(3)
        dib = mod.Diblock(a, b, 40, 0.5, 1.0, 1.0, 20)
        hom = mod.Homopolymer(a, 20, 1.0, 10)
        pt = mod.Point(c)
        fl = mod.FloryInteraction()
        fl.set_interaction(1, 2, 0.4)
        fl.set_interaction(1, 3, 0.2)
        s = mod.FieldSystem((16, 12, 1), cell, monomers=[a, b, c], ensemble="grand")
        for sp, v in zip((dib, hom, pt), (0.1, -0.3, -1.2)):
            s.add_species(sp, v)
        s.add_interaction(fl)
        return s
    f = rand_fields((1, 12, 16), mids=(1, 2, 3))
    so, sg = make(O), make(P)
    O.fieldinit(so, fields=f)
    P.fieldinit(sg, fields=f)
    for isp in range(3):
        q, phi, mu = sg.species_scalars(isp)
        assert q == pytest.approx(so.species[isp].Q, rel=1e-12)
        assert phi == pytest.approx(so.phi_species[isp], rel=1e-12)
        assert mu == pytest.approx(so.mu_species[isp], rel=1e-14)
    for m in (1, 2, 3):
        assert rel(sg.density[m], so.density[m]) < 1e-12
        assert rel(sg.residuals[m], so.residuals[m]) < 1e-10
        assert rel(sg.fields[m], so.fields[m]) < 1e-14
    assert sg.free_energy() == pytest.approx(so.free_energy(), rel=1e-12)


@pytest.mark.parametrize("n", [64, 128])
def test_rqm4_planar_fp32_error_vs_fp64(n):
    # the compile-time RQM4 plane in FP32 against the same sweep in FP64 (128^2: FP32 runs
    # both planes in LDS, FP64 one plane at a time)
    f = rand_fields((1, n, n))
    out = {}
    for prec in ("fp64", "fp32"):
        s, _, c = diblock_system(P, (n, n, 1), 2, "square", (11.0,), Ns=40, mde="RQM4",
                                 precision=prec)
        P.fieldinit(s, fields=f)
        s.density_update()
        out[prec] = (c.Q, {m: s.density[m] for m in (1, 2)})
    assert out["fp32"][0] == pytest.approx(out["fp64"][0], rel=1e-5)
    for m in (1, 2):
        assert rel(out["fp32"][1][m], out["fp64"][1][m]) < 1e-5


@pytest.mark.parametrize("g", [16, 64])
def test_rqm4_3d_fp32_error_vs_fp64(g):
    # the 3-D RQM4 sweep in FP32 against the same sweep in FP64 (64^3: compile-time forms)
    dims = (g, g, g)
    f = rand_fields((g, g, g))
    out = {}
    for prec in ("fp64", "fp32"):
        s, _, c = diblock_system(P, dims, 3, "cubic", (7.0,), Ns=30, mde="RQM4", precision=prec)
        P.fieldinit(s, fields=f)
        s.density_update()
        out[prec] = (c.Q, {m: s.density[m] for m in (1, 2)})
    assert out["fp32"][0] == pytest.approx(out["fp64"][0], rel=1e-5)
    for m in (1, 2):
        assert rel(out["fp32"][1][m], out["fp64"][1][m]) < 1e-5


def test_fieldinit_from_reference_file_and_roundtrip(tmp_path):
    import os
    from helpers import GOLDEN
    path = os.path.join(GOLDEN, "seed_lam1d.json")
    # the lam1d seed is a converged 64-point lamellar state of the chi = 0.25 diblock
    sg, cg, _ = diblock_system(P, (64, 1, 1), 1, "lamellar", (5.0,), chi=0.25)
    P.fieldinit(sg, path)
    # add_cell! installs the file's cell as a new object (src/fileio.jl:58-61)
    assert sg.cell.params[0] == pytest.approx(13.100000015310233)
    so, co, _ = diblock_system(O, (64, 1, 1), 1, "lamellar", (13.100000015310233,), chi=0.25)
    d, f = load_seed("lam1d")
    O.fieldinit(so, fields=f)
    assert sg.field_error() == pytest.approx(so.field_error(), rel=1e-9)
    assert sg.free_energy() == pytest.approx(so.free_energy(), rel=1e-12)
    assert so.field_error() < 1e-5
    out = str(tmp_path / "w.json")
    P.savefields(sg, out)
    back = P.loadfields(out)
    for m in (1, 2):
        np.testing.assert_array_equal(back["fields"][m], sg.fields[m])
    # interpolated seeding onto a finer grid, as examples.ipynb does (64 -> 128)
    sh, ch, _ = diblock_system(P, (128, 1, 1), 1, "lamellar", (5.0,), chi=0.25)
    P.fieldinit(sh, path)
    s2, c2, _ = diblock_system(O, (128, 1, 1), 1, "lamellar", (13.100000015310233,), chi=0.25)
    import oracle.fileio as OF
    O.fieldinit(s2, fields={m: OF.interpolate_grid(f[m], (128, 1, 1)) for m in (1, 2)})
    assert sh.field_error() == pytest.approx(s2.field_error(), rel=1e-9)
    assert sh.free_energy() == pytest.approx(s2.free_energy(), rel=1e-12)


def test_fieldinit_from_densities_fp64():
    # field_type = :rho -> uniform_fields!(sys) (system.jl:348-357), then residuals!
    rng = np.random.default_rng(11)
    rho = {1: 0.5 + 0.2 * rng.standard_normal((1, 16, 16))}
    rho[2] = 1.0 - rho[1]
    sg, _, _ = diblock_system(P, (16, 16, 1), 2, "square", (6.0,), chi=0.3)
    so, _, _ = diblock_system(O, (16, 16, 1), 2, "square", (6.0,), chi=0.3)
    P.fieldinit(sg, rho, field_type="rho")
    for m in (1, 2):
        so.density[m] = rho[m].copy()
    so.uniform_fields()
    so.residuals_update()
    for m in (1, 2):
        assert rel(sg.fields[m], so.fields[m]) < 1e-12
        assert rel(sg.density[m], so.density[m]) < 1e-12


def test_large_batch_matches_one_at_a_time_fp64():
    # 64 systems: the per-system Anderson solves run on several host threads; every
    # system's trajectory must still equal its own one-at-a-time run
    # chi N below the order-disorder transition: the random seeds relax without diverging
    cfgs = [(16, 4.0 + 0.02 * k, 30 + (k % 15), 6.0 + 0.05 * k, 40) for k in range(64)]

    def make(k):
        s = _hex_system(P, *cfgs[k], precision="fp64")
        P.fieldinit(s, seed=500 + k, scale=0.1)
        return s

    batch_sys = [make(k) for k in range(64)]
    res = P.Batch(batch_sys).scft(nhist=10, iter_euler=8, lam_euler=0.05, nsteps=25,
                                  ftol=1e-14)
    for k in (0, 17, 42, 63):
        q = P.scft(make(k), P.AndersonUpdater(nhist=10, iter_euler=8, lam_euler=0.05),
                   opts=P.SCFTOptions(nsteps=25, ftol=1e-14))
        assert np.isfinite(q.fhelm)
        assert res[k].nsteps == q.nsteps
        assert res[k].fhelm == pytest.approx(q.fhelm, rel=1e-13)
        assert res[k].ferr == pytest.approx(q.ferr, rel=1e-12)
