"""The oracle against the reference's own reproducible known-answer tests
(test/test_cell.jl, test/test_species.jl, test/test_interactions.jl,
test/test_system.jl:33)."""
import math

import numpy as np
import pytest

import oracle as O

RTOL = math.sqrt(np.finfo(float).eps)   # Julia's default `≈`


@pytest.mark.parametrize("dim,crystal,params,vol", [
    (1, "lamellar", (10.0,), 10.0),                                   # test_cell.jl:3-4
    (2, "square", (10.0,), 100.0),                                    # :7-8
    (2, "rectangular", (10.0, 5.0), 50.0),                            # :10-11
    (2, "hexagonal", (10.0,), 86.60254037844385),                     # :13-14
    (2, "oblique", (10.0, 5.0, math.pi / 4), 35.35533905932737),      # :16-17
    (3, "cubic", (10.0,), 1000.0),                                    # :20-21
    (3, "tetragonal", (10.0, 5.0), 500.0),                            # :23-24
    (3, "orthorhombic", (10.0, 5.0, 2.0), 100.0),                     # :26-27
    (3, "hexagonal", (10.0, 5.0), 433.01270189221924),                # :29-30
    (3, "monoclinic", (10.0, 5.0, 2.0, math.pi / 4), 70.71067811865474),  # :32-33
    (3, "triclinic", (10.0, 5.0, 2.0, math.pi / 2, math.pi / 2, math.pi / 2), 100.0),  # :35-36
])
def test_cell_volumes(dim, crystal, params, vol):
    cell = O.UnitCell(dim, crystal, *params)
    assert cell.volume == pytest.approx(vol, rel=RTOL)


def test_species_constructors():
    amon = O.Monomer(1, vol=2.0)
    bmon = O.Monomer(2, vol=1.0)
    assert O.Point(amon).Nref == pytest.approx(2.0, rel=RTOL)              # test_species.jl:5-6
    assert O.Homopolymer(bmon, 100, 1.0, 200).Nref == pytest.approx(100.0)  # :8-9
    ch = O.Diblock(amon, bmon, 100, 0.25, 1.0, 1.0, 200)                   # :11-13
    assert ch.f_block[0] == pytest.approx(0.4, rel=RTOL)
    assert ch.Nref_block[1] == pytest.approx(75.0, rel=RTOL)
    ch = O.Multiblock([amon, amon, bmon, amon, bmon, bmon], [10] * 6, [1.0] * 6, 60)  # :15-16
    assert ch.monomer_fraction(1) == pytest.approx(2 / 3, rel=RTOL)


def test_interaction_pairs():
    fh = O.FloryInteraction()                                               # test_interactions.jl
    fh.set_interaction(1, 2, 0.5)
    fh.set_interaction(2, 1, 0.5)
    assert len(fh.chis) == 1
    with pytest.raises(AssertionError):
        fh.set_interaction(1, 1, 0.5)


def _blend_system():
    """test/test_system.jl:1-22 (Point + Diblock on 32x1x1, L=10, chi_AB=0.2)."""
    cell = O.UnitCell(1, "lamellar", 10.0)
    a, b, c = O.Monomer(1), O.Monomer(2), O.Monomer(3)
    point = O.Point(c)
    chain = O.Diblock(a, b, 100, 0.5, 1.0, 1.0, 100)
    itx = O.FloryInteraction()
    itx.set_interaction(1, 2, 0.2)
    sys = O.FieldSystem((32, 1, 1), cell, monomers=[a, b, c])
    sys.add_species(point, 0.75)
    sys.add_species(chain, 0.25)
    sys.add_interaction(itx)
    return sys, point, chain


def test_validate_rejects_uninitialised_fields():
    sys, _, _ = _blend_system()
    with pytest.raises(ValueError):                                         # test_system.jl:24
        sys.validate()


def test_free_energy_bulk_kat():
    sys, _, _ = _blend_system()
    O.fieldinit(sys, seed=12345, scale=0.1)
    # test_system.jl:33 -- seed independent, so reproducible without Julia's RNG.
    assert sys.free_energy_bulk() == pytest.approx(-0.9686022902416352, rel=RTOL)


def test_chain_grid_config_shapes():
    # Ns is adjusted 100 -> 102 for the fA=0.35 diblock (species/base.jl:27-56,
    # multiblock.jl:55-59); block starts are odd, block lengths even.
    with pytest.warns(UserWarning):
        ch = O.Diblock(O.Monomer(1), O.Monomer(2), 100, 0.35, 1.0, 1.0, 100)
    assert ch.Ns == 102
    assert ch.begin_block == [1, 37, 103]
    assert ch.ds_block == pytest.approx([35 / 36, 65 / 66])
    ch = O.Diblock(O.Monomer(1), O.Monomer(2), 100, 0.5, 1.0, 1.0, 100)
    assert ch.begin_block == [1, 51, 101] and ch.ds_block == [1.0, 1.0]
    with pytest.raises(ValueError):
        O.Diblock(O.Monomer(1), O.Monomer(2), 100, 0.29, 1.0, 1.0, 100)    # 28 + 71 != 100
