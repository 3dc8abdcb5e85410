"""The reference's seeded known answers, reproduced by the oracle.

fieldinit!(sys; seed, scale) (src/fileio.jl:14-23) draws with Julia 1.5's
MersenneTwister and randn in the iteration order of the Dict of fields; oracle/julia_rng.py
restates them (the generator is Julia's standard library, outside the reference tree).
These tests pin that restatement, and the oracle behind it, on the values the reference's
own test files hold:
  test/test_system.jl:22-34 (point and chain Q, free energy, bulk free energy, stress);
  test/test_updaters.jl:17-30 (free energy after 10 Euler PECE steps and 10 Anderson
  steps; the value after 10 NAG steps is in the file as a comment).
The point species' Q depends only on the C-monomer draws, so it pins the stream and the
Dict order on its own; a planted mistake (ascending monomer order) misses it.
"""
import warnings

import numpy as np
import pytest

import oracle as O
from oracle import julia_rng as J


def system_blend():
    # test/test_system.jl:2-17
    cell = O.UnitCell(1, "lamellar", 10.0)
    a, b, c = O.Monomer(1, 1.0), O.Monomer(2, 1.0), O.Monomer(3, 1.0)
    point = O.Point(c)
    chain = O.Diblock(a, b, 100, 0.5, 1.0, 1.0, 100)
    itx = O.FloryInteraction()
    itx.set_interaction(1, 2, 0.2)
    sys = O.FieldSystem((32, 1, 1), cell, monomers=[a, b, c])
    sys.add_species(point, 0.75)
    sys.add_species(chain, 0.25)
    sys.add_interaction(itx)
    return sys, point, chain


def system_diblock():
    # test/test_updaters.jl:2-14
    cell = O.UnitCell(1, "lamellar", 10.0)
    a, b = O.Monomer(1, 1.0), O.Monomer(2, 1.0)
    chain = O.Diblock(a, b, 100, 0.5, 1.0, 1.0, 100)
    itx = O.FloryInteraction()
    itx.set_interaction(1, 2, 0.2)
    sys = O.FieldSystem((32, 1, 1), cell, monomers=[a, b])
    sys.add_species(chain, 1.0)
    sys.add_interaction(itx)
    return sys


def test_system_kats():
    sys, point, chain = system_blend()
    O.fieldinit(sys, seed=12345, scale=0.1)          # test_system.jl:22
    sys.species_density(point)                        # test_system.jl:25-26
    sys.species_density(chain)
    assert point.Q == pytest.approx(1.004177455788938, rel=1e-14)      # :28
    assert chain.Q == pytest.approx(8.24300700515971, rel=1e-13)       # :29
    assert sys.free_energy() == pytest.approx(-0.1681580628744596, rel=1e-13)   # :31
    assert sys.free_energy_bulk() == pytest.approx(-0.9686022902416352, rel=1e-14)  # :32
    assert sys.scfstress()[0] == pytest.approx(-0.0004272750367130805, rel=1e-12)   # :33


def test_updater_kats():
    sys = system_diblock()
    O.fieldinit(sys, seed=12345, scale=0.1)          # test_updaters.jl:17
    u1 = O.EulerUpdater(method="PECE", lam=0.005)     # :19-21 (each set up at creation)
    u1.setup(sys)
    u2 = O.MomentumUpdater(method="NAG", lam=0.005)
    u2.setup(sys)
    u3 = O.AndersonUpdater(nhist=5)
    u3.setup(sys)
    for _ in range(10):
        u1.step()
    assert sys.free_energy() == pytest.approx(0.0626842750325202, rel=1e-13)     # :24
    for _ in range(10):
        u2.step()
    assert sys.free_energy() == pytest.approx(0.06172669754976512, rel=1e-13)    # :27 (comment)
    for _ in range(10):
        u3.step()
    # Anderson's least-squares solve amplifies last-bit differences of the transforms and
    # the matrix inverse (pocketfft / numpy against FFTW / LAPACK): measured 8e-11; the
    # reference's own test accepts sqrt(eps) = 1.5e-8
    assert sys.free_energy() == pytest.approx(0.12798422498635653, rel=1e-9)     # :30


def test_dict_order_matters():
    # Julia's Dict{Int} visits 1, 3, 2; drawing the grids in ascending id instead moves
    # the C draws and misses the point species' Q
    assert J.dict_order([1, 2, 3]) == [1, 3, 2]
    sys, point, _ = system_blend()
    rng = J.seeded(12345)
    f = {m: 0.1 * np.array(J.randn_array(rng, 32)).reshape(1, 1, 32) for m in (1, 2, 3)}
    O.fieldinit(sys, fields=f)
    sys.species_density(point)
    assert abs(point.Q - 1.004177455788938) > 1e-6


def test_small_arrays_draw_element_by_element():
    # randn!(::MersenneTwister, A) of fewer than 13 elements draws each normal in turn;
    # from 13 up it takes the raw draws first.  The two agree unless a draw is rejected.
    for n in (5, 12, 40):
        a = J.randn_array(J.seeded(7), n, bulk=False)
        b = J.randn_array(J.seeded(7), n, bulk=True)
        assert len(a) == len(b) == n
    # a stream with a rejection in its first 40 draws separates them
    diff = [s for s in range(40) if J.randn_array(J.seeded(s), 40, bulk=False)
            != J.randn_array(J.seeded(s), 40, bulk=True)]
    assert diff, "no seed below 40 hits the ziggurat's slow path in 40 draws"


def test_stream_moments():
    # 20000 draws: mean, variance and tail mass of a standard normal
    x = np.array(J.randn_array(J.seeded(2024), 20000))
    assert abs(x.mean()) < 0.03
    assert abs(x.var() - 1.0) < 0.03
    assert abs(np.mean(np.abs(x) > 1.959964) - 0.05) < 0.006
