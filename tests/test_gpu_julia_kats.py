"""The reference's seeded known answers on the GPU.

fieldinit(sys, seed=12345, scale=0.1) draws the fields as the reference's fieldinit!
does (Julia 1.5's MersenneTwister and randn, Dict order of the monomers;
pf_system_fieldinit_seeded), so the values printed in the reference's test files hold
for the HIP path in plain FP64:
  test/test_system.jl:22-34 and test/test_updaters.jl:17-30.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")


def system_blend(mod, **kw):
    cell = mod.UnitCell(1, "lamellar", 10.0)
    a, b, c = mod.Monomer(1, 1.0), mod.Monomer(2, 1.0), mod.Monomer(3, 1.0)
    point = mod.Point(c)
    chain = mod.Diblock(a, b, 100, 0.5, 1.0, 1.0, 100)
    itx = mod.FloryInteraction()
    itx.set_interaction(1, 2, 0.2)
    sys = mod.FieldSystem((32, 1, 1), cell, monomers=[a, b, c], **kw)
    sys.add_species(point, 0.75)
    sys.add_species(chain, 0.25)
    sys.add_interaction(itx)
    return sys, point, chain


def system_diblock(mod, **kw):
    cell = mod.UnitCell(1, "lamellar", 10.0)
    a, b = mod.Monomer(1, 1.0), mod.Monomer(2, 1.0)
    chain = mod.Diblock(a, b, 100, 0.5, 1.0, 1.0, 100)
    itx = mod.FloryInteraction()
    itx.set_interaction(1, 2, 0.2)
    sys = mod.FieldSystem((32, 1, 1), cell, monomers=[a, b], **kw)
    sys.add_species(chain, 1.0)
    sys.add_interaction(itx)
    return sys


def test_seeded_fields_equal_oracle():
    sg, _, _ = system_blend(P, precision="fp64")
    so, _, _ = system_blend(O)
    P.fieldinit(sg, seed=12345, scale=0.1)
    O.fieldinit(so, seed=12345, scale=0.1)
    for m in (1, 2, 3):
        np.testing.assert_allclose(sg.fields[m], so.fields[m], rtol=0, atol=1e-16)


def test_system_kats_on_gpu():
    sys, point, chain = system_blend(P, precision="fp64")
    P.fieldinit(sys, seed=12345, scale=0.1)
    sys.density_update()      # test_system.jl:25-26 (species densities of the seeded fields)
    assert point.Q == pytest.approx(1.004177455788938, rel=1e-13)
    assert chain.Q == pytest.approx(8.24300700515971, rel=1e-13)
    assert sys.free_energy() == pytest.approx(-0.1681580628744596, rel=1e-13)
    assert sys.free_energy_bulk() == pytest.approx(-0.9686022902416352, rel=1e-14)
    assert sys.scfstress()[0] == pytest.approx(-0.0004272750367130805, rel=1e-12)


def test_updater_kats_on_gpu():
    sys = system_diblock(P, precision="fp64")
    P.fieldinit(sys, seed=12345, scale=0.1)
    u1 = P.EulerUpdater(method="PECE", lam=0.005)
    u1.setup(sys)
    u2 = P.MomentumUpdater(method="NAG", lam=0.005)
    u2.setup(sys)
    u3 = P.AndersonUpdater(nhist=5)
    u3.setup(sys)
    for _ in range(10):
        u1.step()
    assert sys.free_energy() == pytest.approx(0.0626842750325202, rel=1e-12)
    for _ in range(10):
        u2.step()
    assert sys.free_energy() == pytest.approx(0.06172669754976512, rel=1e-12)
    for _ in range(10):
        u3.step()
    # the least-squares solve amplifies last-bit differences of the transforms; the
    # reference's own test accepts sqrt(eps) = 1.5e-8
    assert sys.free_energy() == pytest.approx(0.12798422498635653, rel=1e-9)
