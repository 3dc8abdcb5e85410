"""The oracle against closed forms, invariants and library routines.

Each check is chosen so that a plausible slip in the restatement (a dropped LW factor, a
wrong ds, a transposed basis, a wrong Simpson weight, a wrong sign in dG) fails it."""
import math

import numpy as np
import pytest

import oracle as O
from oracle.cell import NPARAMS
from oracle.updaters import lu_inverse
from helpers import diblock_system

RNG = np.random.default_rng(7)


def _random_fields(sys, scale=0.3):
    for m in sys.mids:
        sys.fields[m] = scale * RNG.standard_normal(sys.shape)


def test_uniform_field_partition_function_closed_form():
    # Uniform w: every step multiplies q by exp(-ds v w), so Q = exp(-sum_b v w_b Nref_b).
    sys, cell, ch = diblock_system(O, (16, 16, 1), 2, "hexagonal", (5.0,), f=0.35)
    vA, vB = 0.3, -0.2
    sys.fields[1][...] = vA
    sys.fields[2][...] = vB
    sys.density_update()
    expect = math.exp(-(vA * ch.Nref_block[0] + vB * ch.Nref_block[1]))
    assert ch.Q == pytest.approx(expect, rel=1e-13)


def test_zero_field_densities_equal_block_fractions():
    sys, cell, ch = diblock_system(O, (8, 8, 8), 3, "cubic", (4.0,), f=0.35)
    sys.density_update()
    assert ch.Q == pytest.approx(1.0, abs=1e-14)
    np.testing.assert_allclose(ch.density[1], ch.f_block[0], rtol=1e-13)
    np.testing.assert_allclose(ch.density[2], ch.f_block[1], rtol=1e-13)


def test_single_mode_decay_exact():
    # w = 0, q = cos(2 pi x / L): one Strang step is exactly exp(-b^2 k^2 ds / 6).
    L, ds, b = 7.3, 0.9, 1.3
    sys, cell, _ = diblock_system(O, (32, 1, 1), 1, "lamellar", (L,))
    x = np.arange(32) / 32
    qin = np.cos(2 * np.pi * 3 * x).reshape(1, 1, 32)
    ops = sys.solver_update(np.zeros(sys.shape), 1.0, b, ds)
    qout = sys.propagate(ops, qin)
    k2 = (2 * math.pi * 3 / L) ** 2
    np.testing.assert_allclose(qout, qin * math.exp(-b * b * k2 * ds / 6), atol=1e-15)


def test_q_qdagger_overlap_is_Q_for_every_s():
    # The Strang operator is symmetric, so <q_s q+_s> = Q at every contour point.
    sys, cell, ch = diblock_system(O, (12, 10, 8), 3, "orthorhombic", (5.0, 4.0, 3.5), f=0.3)
    _random_fields(sys)
    sys.density_update()
    n = sys.ngrid
    for s in range(ch.Ns + 1):
        assert np.sum(ch.q[s] * ch.qc[s]) / n == pytest.approx(ch.Q, rel=1e-12)
    total = sum(np.mean(sys.density[m]) for m in sys.mids)
    assert total == pytest.approx(1.0, rel=1e-12)


def test_homopolymer_reversal_matches_backward_sweep():
    # A homopolymer copies q reversed (multiblock.jl:153-156); an A-A "diblock" of the
    # same monomer propagates q+ for real.  Both must agree.
    cell = O.UnitCell(2, "square", 6.0)
    a = O.Monomer(1, vol=1.5)
    homo = O.Homopolymer(a, 40, 1.1, 40)
    aa = O.Multiblock([a, a], [20, 20], [1.1, 1.1], 40)
    out = []
    for sp in (homo, aa):
        sys = O.FieldSystem((12, 12, 1), cell, monomers=[a])
        sys.add_species(sp, 1.0)
        sys.fields[1] = 0.4 * np.random.default_rng(3).standard_normal(sys.shape)
        sys.density_update()
        out.append((sp.Q, sys.density[1].copy()))
    assert out[0][0] == pytest.approx(out[1][0], rel=1e-13)
    np.testing.assert_allclose(out[0][1], out[1][1], rtol=1e-12)


def test_rk2_second_order_in_ds():
    errs = []
    Qs = []
    for Ns in (80, 160, 320, 640):
        sys, cell, ch = diblock_system(O, (32, 1, 1), 1, "lamellar", (9.0,), Ns=Ns)
        x = np.arange(32) / 32
        sys.fields[1] = (0.03 * np.cos(2 * np.pi * x) + 0.01 * np.sin(4 * np.pi * x)).reshape(1, 1, 32)
        sys.fields[2] = -sys.fields[1]
        sys.density_update()
        Qs.append(ch.Q)
    for i in range(3):
        errs.append(abs(Qs[i + 1] - Qs[i]))
    assert errs[0] / errs[1] == pytest.approx(4.0, rel=0.05)
    assert errs[1] / errs[2] == pytest.approx(4.0, rel=0.05)


@pytest.mark.parametrize("dim,crystal", sorted(NPARAMS))
def test_dksq_matches_finite_difference_of_ksq(dim, crystal):
    params = {1: [7.0], 2: [6.0, 4.5, 1.2], 3: [6.0, 5.0, 4.0, 1.3, 1.4, 1.2]}[dim]
    if crystal == "monoclinic":
        params = [6.0, 5.0, 4.0, 1.3]
    npr = NPARAMS[(dim, crystal)]
    params = params[:npr]
    dims = {1: (8, 1, 1), 2: (8, 6, 1), 3: (8, 6, 4)}[dim]
    cell = O.UnitCell(dim, crystal, *params)
    cell.setup(dims)
    h = 1e-6
    for k in range(npr):
        cp, cm = O.UnitCell(dim, crystal, *params), O.UnitCell(dim, crystal, *params)
        cp.params[k] += h
        cm.params[k] -= h
        for c in (cp, cm):
            c.setup(dims)
            c.update()
        fd = (cp.ksq - cm.ksq) / (2 * h)
        np.testing.assert_allclose(cell.dksq[k], fd, rtol=1e-6, atol=1e-7 * np.abs(cell.ksq).max())


def test_basis_dimensions_roundtrip():
    p = [6.0, 5.0, 4.0, 1.3, 1.4, 1.2]
    R = O.basis_matrix(*p)
    np.testing.assert_allclose(O.basis_dimensions(R), p, rtol=1e-13)
    cell = O.UnitCell(3, "triclinic", *p)
    np.testing.assert_allclose(cell.G.T @ cell.R, 2 * math.pi * np.eye(3), atol=1e-12)


def test_lu_inverse_matches_numpy_and_flags_singular():
    A = RNG.standard_normal((7, 7)) + 3 * np.eye(7)
    np.testing.assert_allclose(lu_inverse(A), np.linalg.inv(A), rtol=1e-12, atol=1e-14)
    with pytest.raises(ZeroDivisionError):
        lu_inverse(np.ones((3, 3)))


def test_stress_matches_finite_difference_of_log_Q():
    # sigma = -phi dQ with dQ ~ (1/(Q Nref)) dQ/dtheta (multiblock.jl:233-295), i.e.
    # sigma ~ -(phi/Nref) d ln Q / dtheta at fixed fields.  The reference formula is the
    # continuum derivative, so it agrees with the discrete FD to O(ds^2); centrosymmetric
    # fields avoid the Re*Re quirk.
    L = 15.0
    x = np.arange(64) / 64
    w = (1.2 * np.cos(2 * np.pi * x) + 0.4 * np.cos(4 * np.pi * x)).reshape(1, 1, 64)

    def logQ(Lv, exact=False, Ns=200):
        sys, cell, ch = diblock_system(O, (64, 1, 1), 1, "lamellar", (Lv,), Ns=Ns)
        sys.fields[1], sys.fields[2] = w.copy(), -w.copy()
        sys.density_update()
        return math.log(ch.Q), sys.scfstress(exact=exact)[0], ch.Nref

    h = 1e-5
    fd = -(logQ(L + h)[0] - logQ(L - h)[0]) / (2 * h) / 100.0
    _, s_ref, _ = logQ(L)
    _, s_exact, _ = logQ(L, exact=True)
    assert s_ref == pytest.approx(fd, rel=0.02)
    assert s_exact == pytest.approx(fd, rel=0.02)
    assert abs(fd) > 1e-4


def test_homogeneous_melt_free_energy():
    # Symmetric diblock, zero fields: the SCF state is homogeneous and
    # f = chi fA fB (translational term phi(phi-1)/N vanishes for phi = 1).
    sys, cell, ch = diblock_system(O, (16, 1, 1), 1, "lamellar", (10.0,), chi=0.2)
    sys.residuals_update()
    assert sys.free_energy() == pytest.approx(0.2 * 0.25, rel=1e-13)
    assert sys.field_error() == pytest.approx(0.0, abs=1e-6)


def test_rqm4_fourth_order_and_symmetric():
    # RQM4 (mde.jl:83-104): Richardson of one full and two half Strang steps is fourth
    # order in ds (successive error ratios climb towards 16; RK2 gives 4), and remains a
    # symmetric operator (<q_s q+_s> = Q for every s).
    Qs = []
    for Ns in (40, 80, 160, 320):
        sys, cell, ch = diblock_system(O, (32, 1, 1), 1, "lamellar", (9.0,), Ns=Ns, mde="RQM4")
        x = np.arange(32) / 32
        sys.fields[1] = (0.03 * np.cos(2 * np.pi * x) + 0.01 * np.sin(4 * np.pi * x)).reshape(1, 1, 32)
        sys.fields[2] = -sys.fields[1]
        sys.density_update()
        Qs.append(ch.Q)
        if Ns == 40:
            for s in range(ch.Ns + 1):
                assert np.sum(ch.q[s] * ch.qc[s]) / sys.ngrid == pytest.approx(ch.Q, rel=1e-12)
    e = [abs(Qs[i + 1] - Qs[i]) for i in range(3)]
    r = [e[0] / e[1], e[1] / e[2]]
    assert 10.0 < r[0] < r[1] < 17.0


def _grand_blend(mod, ensemble, vals, dims=(16, 12, 1)):
    cell = mod.UnitCell(2, "rectangular", 6.0, 4.5)
    a, b, c = mod.Monomer(1), mod.Monomer(2), mod.Monomer(3)
    dib = mod.Diblock(a, b, 40, 0.5, 1.0, 1.0, 20)
    hom = mod.Homopolymer(a, 20, 1.0, 10)
    pt = mod.Point(c)
    fl = mod.FloryInteraction()
    fl.set_interaction(1, 2, 0.4)
    fl.set_interaction(1, 3, 0.2)
    s = mod.FieldSystem(dims, cell, monomers=[a, b, c], ensemble=ensemble)
    for sp, v in zip((dib, hom, pt), vals):
        s.add_species(sp, v)
    s.add_interaction(fl)
    return s


def test_grand_matches_canonical_at_matching_mu():
    # muphi! (system.jl:248-259): the grand ensemble with mu_p = log(phi_p / Q_p) taken from
    # a canonical evaluation at the same fields gives back phi_p and the same densities;
    # residuals differ only by the canonical mean subtraction (system.jl:332-337)
    rng = np.random.default_rng(21)
    f = {m: 0.3 * rng.standard_normal((1, 12, 16)) for m in (1, 2, 3)}
    for m in f:
        f[m] -= f[m].mean()
    can = _grand_blend(O, "canonical", (0.5, 0.3, 0.2))
    O.fieldinit(can, fields={m: w.copy() for m, w in f.items()})
    Q = [sp.Q for sp in can.species]
    mu = [math.log(p / q) for p, q in zip(can.phi_species, Q)]
    gr = _grand_blend(O, "grand", mu)
    O.fieldinit(gr, fields={m: w.copy() for m, w in f.items()})
    for isp in range(3):
        assert gr.phi_species[isp] == pytest.approx(can.phi_species[isp], rel=1e-12)
    for m in (1, 2, 3):
        np.testing.assert_allclose(gr.density[m], can.density[m], rtol=1e-12, atol=1e-15)
        rg = gr.residuals[m] - gr.residuals[m].mean()
        np.testing.assert_allclose(rg, can.residuals[m], rtol=1e-10, atol=1e-13)
    # point species, closed form: Q = <exp(-v w)>, phi = Q exp(mu)
    wq = np.mean(np.exp(-gr.fields[3]))
    assert gr.species[2].Q == pytest.approx(wq, rel=1e-13)
    assert gr.phi_species[2] == pytest.approx(wq * math.exp(mu[2]), rel=1e-13)
