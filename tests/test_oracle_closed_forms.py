"""Closed forms that pin oracle functions no reference value covers.

* Stress multiplicity (src/species/multiblock.jl:273-284).  The reference doubles every
  entry of the half spectrum but k = 0, the kx = 0 plane included.  For a field that is
  constant along x (the halved axis), every mode lies in the kx = 0 plane, so the
  reference stress along the varying axis is exactly twice -(1/Nref) d ln Q / d theta,
  up to the O(ds^2) gap between the Simpson-weighted stress and the discrete propagator's
  derivative, and the x component is zero.  A planted mistake (the kx = 0 plane counted
  once) gives 1x instead of 2x and fails.
* Updaters (src/updaters/momentum.jl:63-165, src/updaters/euler.jl:36-66).  A melt of
  point particles has the residual in closed form: with one monomer species in the
  canonical ensemble, strict incompressibility and no interactions, rho = e^-w / <e^-w>
  and the residual is rho - 1.  The first step of every updater then follows from its
  definition alone; in the linear-response limit (small w, r = -w) the Euler
  predictor-corrector multiplies w by 1 - 2 lam + lam^2 / 2 per step (the reference's
  predictor adds 1.5 lam r, euler.jl:45-48), where the textbook Heun step would give
  1 - lam + lam^2 / 2.
* Compressibility (src/system/compress.jl:37-61) and the Edwards potential
  (src/interactions/edwards.jl:61-87) on two point species A, B with chi_AB and u_A: the
  densities, potentials, pressure and residuals in closed form for inv_zeta = 0 (strict),
  > 0 (compressible) and < 0 (unconstrained).
"""
import math

import numpy as np
import pytest

import oracle as O


# ---------------------------------------------------------------------------------------
def _xconst_system(dims, dim, crystal, params, Ns=100):
    cell = O.UnitCell(dim, crystal, *params)
    a, b = O.Monomer(1), O.Monomer(2)
    ch = O.Diblock(a, b, 100, 0.5, 1.0, 1.0, Ns)
    itx = O.FloryInteraction()
    itx.set_interaction(1, 2, 0.2)
    s = O.FieldSystem(dims, cell, monomers=[a, b])
    s.add_species(ch, 1.0)
    s.add_interaction(itx)
    nz, ny, nx = s.shape
    y = np.arange(ny) / ny
    z = np.arange(nz) / nz
    w = (0.8 * np.cos(2 * np.pi * y)[None, :, None] + 0.3 * np.cos(4 * np.pi * y)[None, :, None]
         + 0.5 * np.cos(2 * np.pi * z)[:, None, None])
    w = np.broadcast_to(w, s.shape).copy()       # even about the origin, constant along x
    return s, cell, ch, {1: w, 2: -w}


def _stress_and_fd(dims, dim, crystal, params, k, Ns=200):
    def logq(p):
        s, cell, ch, f = _xconst_system(dims, dim, crystal, p, Ns)
        O.fieldinit(s, fields=f)
        return math.log(ch.Q), s, ch
    h = 1e-5
    pp, pm = list(params), list(params)
    pp[k] += h
    pm[k] -= h
    lq_p, _, ch = logq(pp)
    lq_m, _, _ = logq(pm)
    _, s, ch = logq(list(params))
    fd = -(lq_p - lq_m) / (2 * h) / ch.Nref
    return s, fd


@pytest.mark.parametrize("dims,dim,crystal,params,k", [
    ((24, 16, 1), 2, "rectangular", (5.0, 7.0), 1),
    ((12, 10, 8), 3, "orthorhombic", (4.0, 6.0, 5.5), 1),
    ((12, 10, 8), 3, "orthorhombic", (4.0, 6.0, 5.5), 2)])
def test_stress_counts_the_kx0_plane_twice(dims, dim, crystal, params, k):
    ratio, once = [], []
    for Ns in (100, 200, 400):
        s, fd = _stress_and_fd(dims, dim, crystal, params, k, Ns=Ns)
        assert abs(fd) > 1e-3
        ratio.append(s.scfstress()[k] / fd)
        once.append(s.scfstress(exact=True)[k] / fd)
        assert s.scfstress()[0] == 0.0                 # nothing varies along x
    # the gap to the closed form is O(ds^2): it shrinks 4x per halving of ds, and the
    # Richardson extrapolation of the ratio lands on 2
    g = [r - 2.0 for r in ratio]
    assert 3.5 < g[0] / g[1] < 4.5 and 3.5 < g[1] / g[2] < 4.5
    assert ratio[2] + (ratio[2] - ratio[1]) / 3 == pytest.approx(2.0, abs=1e-4)
    # planted mistake: the kx = 0 plane weighted once extrapolates to 1, not 2
    assert once[2] + (once[2] - once[1]) / 3 == pytest.approx(1.0, abs=1e-4)


# ---------------------------------------------------------------------------------------
def _point_melt(w0):
    cell = O.UnitCell(1, "lamellar", 10.0)
    a = O.Monomer(1)
    s = O.FieldSystem((w0.size, 1, 1), cell, monomers=[a])
    s.add_species(O.Point(a), 1.0)
    O.fieldinit(s, fields={1: w0.reshape(1, 1, -1)})
    return s


def _r(w):
    """Residual of the point melt in closed form (mean zero by construction)."""
    e = np.exp(-w)
    return e / e.mean() - 1.0


def _w0(amp=0.4, n=32, seed=5):
    w = amp * np.random.default_rng(seed).standard_normal(n)
    return w - w.mean()


@pytest.mark.parametrize("method", ["MOM", "RMS", "ADAM", "NAG"])
def test_momentum_first_step_closed_form(method):
    lam, gam, b1, b2 = 0.05, 0.9, 0.8, 0.99
    w0 = _w0()
    s = _point_melt(w0)
    u = O.MomentumUpdater(method=method, lam=lam, gamma=gam, beta1=b1, beta2=b2)
    u.setup(s)
    u.step()
    r0 = _r(w0)
    step = {"MOM": (1 - b1) * r0,
            "RMS": r0 / (math.sqrt(1 - b2) * np.abs(r0) + 1e-8),
            "ADAM": (1 - b1) * r0 / (math.sqrt(1 - b2) * np.abs(r0) + 1e-8),
            "NAG": (1 - gam) * r0}[method]
    np.testing.assert_allclose(s.fields[1].ravel(), w0 + lam * step, rtol=0, atol=1e-13)


@pytest.mark.parametrize("method", ["PEC", "PECE"])
def test_euler_step_closed_form(method):
    lam = 0.07
    w0 = _w0()
    s = _point_melt(w0)
    u = O.EulerUpdater(method=method, lam=lam)
    u.setup(s)
    u.step()
    r0 = _r(w0)
    w1 = w0 + 1.5 * lam * r0 + 0.5 * lam * _r(w0 + lam * r0)
    np.testing.assert_allclose(s.fields[1].ravel(), w1 - (w1.mean() if method == "PECE" else 0),
                               rtol=0, atol=1e-13)


def test_euler_linear_response_factor():
    lam, eps = 0.07, 1e-6
    w0 = _w0(amp=eps)
    s = _point_melt(w0)
    u = O.EulerUpdater(method="PEC", lam=lam)
    u.setup(s)
    u.step()
    w1 = s.fields[1].ravel()
    factor = 1 - 2 * lam + lam ** 2 / 2
    np.testing.assert_allclose(w1, factor * w0, rtol=0, atol=1e-5 * np.abs(w0).max())
    heun = 1 - lam + lam ** 2 / 2                  # the textbook step is far outside
    assert np.abs(w1 - heun * w0).max() > 0.05 * np.abs(w0).max()


# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("inv_zeta", [0.0, 0.5, -1.0])
@pytest.mark.parametrize("ensemble", ["canonical", "grand"])
def test_compressibility_and_edwards_closed_form(inv_zeta, ensemble):
    chi, u, va, vb = 0.7, 0.9, 1.0, 2.0
    vals = (0.4, 0.6) if ensemble == "canonical" else (0.3, -0.2)
    cell = O.UnitCell(2, "square", 6.0)
    A, B = O.Monomer(1, vol=va), O.Monomer(2, vol=vb)
    s = O.FieldSystem((12, 8, 1), cell, monomers=[A, B], ensemble=ensemble,
                      compress=O.Compressibility(inv_zeta))
    s.add_species(O.Point(A), vals[0])
    s.add_species(O.Point(B), vals[1])
    fl, ed = O.FloryInteraction(), O.EdwardsInteraction()
    fl.set_interaction(1, 2, chi)
    ed.set_interaction(1, u)
    s.add_interaction(fl)
    s.add_interaction(ed)
    rng = np.random.default_rng(9)
    w = {m: 0.3 * rng.standard_normal((1, 8, 12)) for m in (1, 2)}
    O.fieldinit(s, fields={m: x.copy() for m, x in w.items()})
    # point species: q = e^{-v w}, Q = <q>; canonical phi fixed, grand phi = Q e^mu
    rho = {}
    for m, v, val in ((1, va, vals[0]), (2, vb, vals[1])):
        q = np.exp(-v * w[m])
        phi = val if ensemble == "canonical" else q.mean() * math.exp(val)
        rho[m] = phi * q / q.mean()
    pot = {1: chi * rho[2] / vb + u * rho[1] / va, 2: chi * rho[1] / va}
    if inv_zeta == 0.0:
        eta = (rho[1] + rho[2] - 1.0 + (w[1] - pot[1]) + (w[2] - pot[2])) / 2
    elif inv_zeta > 0.0:
        eta = (rho[1] + rho[2] - 1.0) / inv_zeta
    else:
        eta = np.zeros_like(rho[1])
    for m in (1, 2):
        res = pot[m] + eta - w[m]
        wm = w[m]
        if ensemble == "canonical":
            res, wm = res - res.mean(), wm - wm.mean()
        np.testing.assert_allclose(s.density[m], rho[m], rtol=1e-13, atol=0)
        np.testing.assert_allclose(s.potentials[m], pot[m], rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(s.residuals[m], res, rtol=0, atol=1e-13)
        np.testing.assert_allclose(s.fields[m], wm, rtol=0, atol=1e-15)
    fexc = (chi * (rho[1] / va) * (rho[2] / vb)).mean() + (u / 2 * (rho[1] / va) ** 2).mean()
    ffield = -(rho[1] * s.fields[1] + rho[2] * s.fields[2]).mean()
    if ensemble == "canonical":
        ftrans = sum(p * (p - 1.0) / v for p, v in zip(vals, (va, vb)))
        assert s.free_energy() == pytest.approx(ftrans + ffield + fexc, rel=1e-13)
