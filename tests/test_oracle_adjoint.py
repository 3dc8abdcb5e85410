"""Pins of the oracle's free-energy derivative (oracle/adjoint.py).

Every component of the reverse-mode derivative is checked against central finite
differences (h = 1e-6, FP64) of the oracle's own forward free energy, which is itself
pinned by the reference's known-answer tests and invariants.  A planted mistake (the
Parseval multiplicity of the kx = 0 plane doubled) must fail the cell check.
"""
import numpy as np
import pytest

import oracle
from oracle import adjoint as adj

H = 1e-6


def _diblock(dims, dim, crystal, params, N=100, f=0.35, chi=0.2, Ns=20, seed=3, scale=0.5):
    cell = oracle.UnitCell(dim, crystal, *params)
    a, b = oracle.Monomer(1), oracle.Monomer(2)
    ch = oracle.Diblock(a, b, N, f, 1.0, 1.2, Ns)
    itx = oracle.FloryInteraction()
    itx.set_interaction(1, 2, chi)
    sys = oracle.FieldSystem(dims, cell, monomers=[a, b])
    sys.add_species(ch, 1.0)
    sys.add_interaction(itx)
    rng = np.random.default_rng(seed)
    for mid in sys.mids:
        sys.fields[mid] = scale * rng.standard_normal(sys.shape)
    return sys


def _blend(seed=5):
    """ABA triblock + B homopolymer + point C with Flory and Edwards terms, rectangular
    2-D cell (two cell parameters)."""
    cell = oracle.UnitCell(2, "rectangular", 7.0, 5.0)
    A, B, C = oracle.Monomer(1, 1.0), oracle.Monomer(2, 1.3), oracle.Monomer(3, 0.8)
    tri = oracle.Multiblock([A, B, A], [12, 20, 8], [1.0, 1.1, 1.0], 20)
    hom = oracle.Homopolymer(B, 30, 0.9, 16)
    pt = oracle.Point(C)
    fl = oracle.FloryInteraction()
    fl.set_interaction(1, 2, 0.3)
    fl.set_interaction(2, 3, 0.15)
    fl.set_interaction(1, 3, 0.4)
    ed = oracle.EdwardsInteraction()
    ed.set_interaction(2, 0.7)
    sys = oracle.FieldSystem((12, 8, 1), cell, monomers=[A, B, C])
    sys.add_species(tri, 0.6)
    sys.add_species(hom, 0.3)
    sys.add_species(pt, 0.1)
    sys.add_interaction(fl)
    sys.add_interaction(ed)
    rng = np.random.default_rng(seed)
    for mid in sys.mids:
        sys.fields[mid] = 0.4 * rng.standard_normal(sys.shape)
    return sys


def _fd(sys, functional, setter, h=H):
    setter(+h)
    fp = adj.free_energy(sys, functional)
    setter(-2 * h)
    fm = adj.free_energy(sys, functional)
    setter(+h)
    return (fp - fm) / (2 * h)


def _field_setter(sys, d):
    def s(h):
        for mid in sys.mids:
            sys.fields[mid] += h * d[mid]
    return s


def _cell_setter(sys, k):
    def s(h):
        sys.cell.params[k] += h
        sys.cell.update()
    return s


def _chi_setter(sys, pair):
    itx = [i for i in sys.interactions if hasattr(i, "chis")][0]

    def s(h):
        itx.chis[pair] += h
    return s


def _u_setter(sys, mid):
    itx = [i for i in sys.interactions if hasattr(i, "uvals")][0]

    def s(h):
        itx.uvals[mid] += h
    return s


def _N_setter(sp, b):
    def s(h):
        N = list(sp.N_block)
        N[b] += h
        adj.set_block_lengths(sp, N)
    return s


def _phi_setter(sys, isp):
    def s(h):
        sys.phi_species[isp] += h
    return s


def _close(a, b, rtol=2e-6, atol=1e-10):
    assert abs(a - b) <= atol + rtol * max(abs(a), abs(b)), (a, b)


def _check_all(sys, functional, ndir=3, seed=11):
    g = adj.fhelm_grad(sys, functional)
    rng = np.random.default_rng(seed)
    for _ in range(ndir):
        d = {mid: rng.standard_normal(sys.shape) for mid in sys.mids}
        lin = sum(float(np.sum(g.dw[m] * d[m])) for m in sys.mids)
        _close(lin, _fd(sys, functional, _field_setter(sys, d)))
    for k in range(sys.cell.nparams):
        _close(g.dcell[k], _fd(sys, functional, _cell_setter(sys, k)))
    for pair, v in g.dchi.items():
        _close(v, _fd(sys, functional, _chi_setter(sys, pair)))
    for mid, v in g.du.items():
        _close(v, _fd(sys, functional, _u_setter(sys, mid)))
    for isp, sp in enumerate(sys.species):
        _close(g.dphi[isp], _fd(sys, functional, _phi_setter(sys, isp)))
        for b, v in enumerate(g.dN[isp]):
            _close(v, _fd(sys, functional, _N_setter(sp, b)))
    return g


@pytest.mark.parametrize("functional", adj.FUNCTIONALS)
def test_adjoint_fd_lamellar(functional):
    sys = _diblock((32, 1, 1), 1, "lamellar", (9.0,))
    _check_all(sys, functional)


@pytest.mark.parametrize("functional", adj.FUNCTIONALS)
def test_adjoint_fd_hexagonal(functional):
    sys = _diblock((16, 16, 1), 2, "hexagonal", (8.0,), Ns=16)
    _check_all(sys, functional)


def test_adjoint_fd_cubic():
    sys = _diblock((8, 8, 8), 3, "cubic", (6.0,), Ns=12)
    _check_all(sys, "textbook", ndir=2)


def test_adjoint_fd_triclinic():
    sys = _diblock((8, 6, 4), 3, "triclinic", (5.0, 4.5, 4.0, 1.4, 1.5, 1.7), Ns=10)
    _check_all(sys, "reference", ndir=1)


@pytest.mark.parametrize("functional", adj.FUNCTIONALS)
def test_adjoint_fd_blend(functional):
    sys = _blend()
    _check_all(sys, functional)


def test_density_vjp_is_dQ_trapezoid():
    """Closed form: with rho_bar = 0 and Q_bar = 1 the adjoint of q is q+/n, so dQ/dw is
    the trapezoid rule -(v/2n) sum_s ds (q_{s+1} q+_{s+1} + q_s q+_s) over the steps."""
    sys = _diblock((16, 16, 1), 2, "hexagonal", (8.0,), Ns=16)
    dw, _ = adj.density_vjp(sys, {m: np.zeros(sys.shape) for m in sys.mids}, [1.0])
    sp = sys.species[0]
    n = sys.ngrid
    exp = {m: np.zeros(sys.shape) for m in sys.mids}
    for b in range(sp.nblocks):
        mid = sp.mid_block[b]
        for s in range(sp.begin_block[b], sp.begin_block[b + 1]):
            exp[mid] += -(sp.ds_block[b] / (2 * n)) * (sp.q[s] * sp.qc[s] +
                                                       sp.q[s - 1] * sp.qc[s - 1])
    for m in sys.mids:
        np.testing.assert_allclose(dw[m], exp[m], rtol=1e-11, atol=1e-16)


def test_dQ_dtheta_converges_to_reference_stress():
    """The reference's analytic stress (multiblock.jl:233-295) is the continuum limit of
    d log Q / d theta: for a centrosymmetric 1-D field (where its Re*Re form is exact) the
    gap to the exact discrete derivative closes at second order in ds, with the
    reference's own formula and with the quirk-free form alike."""
    x = np.arange(32) / 32
    w = 0.8 * np.cos(2 * np.pi * x) + 0.3 * np.cos(4 * np.pi * x)
    gaps = []
    for Ns in (80, 160, 320):
        sys = _diblock((32, 1, 1), 1, "lamellar", (9.0,), Ns=Ns)
        sys.fields[1] = w.reshape(sys.shape).copy()
        sys.fields[2] = -w.reshape(sys.shape).copy()
        sp = sys.species[0]
        _, dcell = adj.density_vjp(sys, {m: np.zeros(sys.shape) for m in sys.mids}, [1.0])
        dlogQ = dcell[0] / sp.Q
        # species_stress is the reference's dQ_theta = d log Q / d theta / Nref
        quirk = sys.species_stress(sp, exact=False)[0] * sp.Nref
        exact = sys.species_stress(sp, exact=True)[0] * sp.Nref
        assert abs(quirk - exact) <= 1e-12 * abs(exact)
        gaps.append(abs(dlogQ - quirk) / abs(dlogQ))
    r1, r2 = gaps[0] / gaps[1], gaps[1] / gaps[2]
    assert 3.5 < r1 < 4.5 and 3.5 < r2 < 4.5, gaps
    assert gaps[-1] < 0.01


def test_planted_multiplicity_mistake_fails():
    """Doubling the kx = 0 plane (as the reference stress does, multiblock.jl:276-282)
    in the adjoint's Parseval weights must break the finite-difference check."""
    sys = _diblock((16, 16, 1), 2, "hexagonal", (8.0,), Ns=16)
    good = adj._mult

    def bad(s):
        m = good(s)
        m[..., 0] = 2.0
        return m
    adj._mult = bad
    try:
        g = adj.fhelm_grad(sys, "reference")
    finally:
        adj._mult = good
    fd = _fd(sys, "reference", _cell_setter(sys, 0))
    assert abs(g.dcell[0] - fd) > 1e-4 * abs(fd)
