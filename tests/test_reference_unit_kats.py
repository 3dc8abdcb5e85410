"""The reference's unit known answers (test/test_cell.jl, test/test_species.jl,
test/test_interactions.jl) against the drop-in's own host mirror: every cell volume
through the C ABI's geometry helper (pf_cell_geometry), the chain constructors through its
contour grid (pf_chain_grid), and the interaction tables, with Julia's default `≈`
tolerance.  No GPU is needed."""
import math

import numpy as np
import pytest

P = pytest.importorskip("polyfields_hip")

RTOL = math.sqrt(np.finfo(float).eps)


@pytest.fixture(autouse=True)
def _library():
    try:
        P.lib()
    except RuntimeError as e:
        pytest.skip(f"library not built: {e}")


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
    assert P.UnitCell(dim, crystal, *params).volume == pytest.approx(vol, rel=RTOL)


def test_species_constructors():
    amon = P.Monomer(1, vol=2.0)
    bmon = P.Monomer(2, vol=1.0)
    assert P.Point(amon).Nref == pytest.approx(2.0, rel=RTOL)                 # test_species.jl:5-6
    assert P.Homopolymer(bmon, 100, 1.0, 200).Nref == pytest.approx(100.0, rel=RTOL)   # :8-9
    ch = P.Diblock(amon, bmon, 100, 0.25, 1.0, 1.0, 200)                      # :11-13
    assert ch.f_block[0] == pytest.approx(0.4, rel=RTOL)
    assert ch.Nref_block[1] == pytest.approx(75.0, rel=RTOL)
    ch = P.Multiblock([amon, amon, bmon, amon, bmon, bmon], [10] * 6, [1.0] * 6, 60)   # :15-16
    assert ch.monomer_fraction(1) == pytest.approx(2 / 3, rel=RTOL)


def test_interaction_pairs():
    fh = P.FloryInteraction()                                                 # test_interactions.jl
    P.set_interaction(fh, 1, 2, 0.5)
    with pytest.warns(UserWarning):   # the reference warns on replacing a pair (flory.jl:29)
        P.set_interaction(fh, 2, 1, 0.5)
    assert len(fh.chis) == 1
    with pytest.raises(AssertionError):
        P.set_interaction(fh, 1, 1, 0.5)
