"""Field-file format and grid interpolation (host code of the product vs the oracle)."""
import os

import numpy as np
import pytest

import oracle.fileio as OF
from helpers import GOLDEN

P = pytest.importorskip("polyfields_hip")


@pytest.mark.parametrize("inshape,outshape", [((64, 1, 1), (128, 1, 1)), ((48, 1, 1), (20, 1, 1)),
                                              ((16, 12, 1), (32, 24, 1)), ((8, 6, 4), (12, 9, 6)),
                                              ((16, 1, 1), (16, 8, 1)), ((16, 8, 1), (16, 1, 1))])
def test_interpolate_grid_matches_natural_spline(inshape, outshape):
    rng = np.random.default_rng(3)
    g = rng.standard_normal((inshape[2], inshape[1], inshape[0]))
    a = P.interpolate_grid(g, outshape)
    b = OF.interpolate_grid(g, outshape)
    assert a.shape == (outshape[2], outshape[1], outshape[0])
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


def test_interpolate_grid_identity_and_linear():
    rng = np.random.default_rng(4)
    g = rng.standard_normal((4, 6, 8))
    np.testing.assert_allclose(P.interpolate_grid(g, (8, 6, 4)), g, atol=1e-14)
    x = np.linspace(-1.0, 2.0, 10)
    line = (3.0 * x - 0.5).reshape(1, 1, 10)
    out = P.interpolate_grid(line, (37, 1, 1)).ravel()
    xo = np.linspace(-1.0, 2.0, 37)
    np.testing.assert_allclose(out, 3.0 * xo - 0.5, atol=1e-13)


def test_loadfields_reads_reference_seed_file():
    d = P.loadfields(os.path.join(GOLDEN, "seed_lam1d.json"))
    o = OF.loadfields(os.path.join(GOLDEN, "seed_lam1d.json"))
    assert d["grid"] == (64, 1, 1) and d["crystal"] == "lamellar" and d["field_type"] == "omega"
    assert d["cell_params"] == pytest.approx([13.100000015310233])
    for m in (1, 2):
        np.testing.assert_array_equal(d["fields"][m], o["fields"][m])
