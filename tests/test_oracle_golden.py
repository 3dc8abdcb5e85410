"""The oracle against the converged states PolyFields.jl printed itself.

* README.md:62-79 -- config-1 lamellar diblock (64x1x1, N=100, fA=0.5, chi=0.2)
  converged to fhelm 0.0812381288285528 at L = 16.528709413909947 with ferr 5.09e-6 and
  serr 1.66e-6.  Converging the fields fully at that L must reproduce fhelm to well
  inside what ferr = 5e-6 allows, and the stress must sit near serr.
* examples/examples.ipynb (cylinder cell) -- 2-D hexagonal diblock (fA=0.3, chi=0.3)
  from seeds/cyl2d.json converged to fhelm 0.10659231121049459 at
  a = 18.381681210100236.  This exercises the hexagonal basis and the folded k^2 of
  src/system/cell.jl:273-278; the notebook ran on 128^2, the seed grid is 64^2, and the
  smooth solution is resolution-converged at 64^2 to ~1e-9.
"""
import numpy as np
import pytest

import oracle as O
from helpers import diblock_system, cos_fields, load_seed


def test_readme_lamellar_state():
    L = 16.528709413909947
    sys, cell, ch = diblock_system(O, (64, 1, 1), 1, "lamellar", (L,), chi=0.2)
    O.fieldinit(sys, fields=cos_fields(sys.shape))
    res = O.scft(sys, O.AndersonUpdater(nhist=10), opts=O.SCFTOptions(nsteps=400, ftol=1e-10))
    assert res.converged
    assert res.fhelm == pytest.approx(0.0812381288285528, abs=5e-8)
    assert sys.scfstress()[0] == pytest.approx(1.6582423877606872e-6, abs=4e-7)


def test_readme_variable_cell_run_lands_on_readme_cell():
    # The README's own run (Anderson nhist=10, DomainUpdater nskip=5 nper=20,
    # ftol=stol=1e-4) from a lamellar start: the cell and fhelm land where the README's
    # did, to the precision its tolerances allow.
    sys, cell, ch = diblock_system(O, (64, 1, 1), 1, "lamellar", (10.0,), chi=0.2)
    O.fieldinit(sys, fields=cos_fields(sys.shape))
    res = O.scft(sys, O.AndersonUpdater(nhist=10), domain=O.DomainUpdater(nskip=5, nper=20),
                 opts=O.SCFTOptions(nsteps=1500, ftol=1e-4, stol=1e-4))
    assert res.converged
    assert cell.params[0] == pytest.approx(16.528709413909947, abs=0.1)
    assert res.fhelm == pytest.approx(0.0812381288285528, abs=1e-4)


def test_hex_cylinder_notebook_state():
    d, fields = load_seed("cyl2d")
    a = float(d["cell_params"][0])
    sys, cell, ch = diblock_system(O, (64, 64, 1), 2, "hexagonal", (a,), f=0.3, chi=0.3)
    O.fieldinit(sys, fields=fields)
    res = O.scft(sys, O.AndersonUpdater(nhist=10), opts=O.SCFTOptions(nsteps=150, ftol=1e-7))
    assert res.ferr < 2e-7
    assert res.fhelm == pytest.approx(0.10659231121049459, abs=1e-7)
    # serr is a norm (domain.jl:30), so only |stress| is pinned.
    assert abs(sys.scfstress()[0]) == pytest.approx(2.3585375325984883e-5, abs=5e-6)


def test_hex_pin_rejects_signed_wavevectors(monkeypatch):
    # A planted mistake: the physically signed j instead of the reference's folded |j|
    # moves the hexagonal free energy far outside the pin above.
    import oracle.cell as C
    real = C.folded_indices

    def signed(dims):
        JX, JY, JZ = real(dims)
        nxh, ny, nz = C.ksize(dims)
        iy = np.arange(ny)
        jy = np.where(2 * iy <= ny, iy, iy - ny).astype(float)
        return JX, np.broadcast_to(jy[None, :, None], JX.shape).copy(), JZ

    monkeypatch.setattr(C, "folded_indices", signed)
    d, fields = load_seed("cyl2d")
    sys, cell, ch = diblock_system(O, (64, 64, 1), 2, "hexagonal", (float(d["cell_params"][0]),),
                                   f=0.3, chi=0.3)
    O.fieldinit(sys, fields=fields)
    res = O.scft(sys, O.AndersonUpdater(nhist=10), opts=O.SCFTOptions(nsteps=60, ftol=1e-7))
    assert abs(res.fhelm - 0.10659231121049459) > 1e-4
