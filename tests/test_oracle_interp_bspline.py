"""interpolate_grid (src/utils.jl:64-109) against a restatement of the algorithm its
dependency publishes: Interpolations.jl 0.13.1 (Manifest.toml:234-238),
``interpolate(A, BSpline(Cubic(Line(OnGrid()))))`` evaluated at the reference's sample
points.

Interpolations.jl represents the interpolant as a cubic B-spline with one padding
coefficient on each side of every axis, c_0 .. c_{n+1} for samples y_1 .. y_n on the grid
points x = 1 .. n (OnGrid: knots on the samples).  The prefilter solves
  (c_{j-1} + 4 c_j + c_{j+1}) / 6 = y_j          j = 1 .. n
with the Line boundary rows
  c_0 - 2 c_1 + c_2 = 0,   c_{n-1} - 2 c_n + c_{n+1} = 0,
i.e. a zero second derivative at the end knots.  Evaluation sums c_j beta3(x - j) with the
centred cubic B-spline beta3, axis by axis (a tensor product).  Here the system is solved
densely and the spline summed term by term: an algorithm independent of the oracle's
(scipy's natural CubicSpline applied axis by axis), so agreement pins the oracle to the
reference's interpolant.  A planted mistake (the Flat boundary rows c_0 = c_2, zero slope)
must fail the same comparison.
"""
import numpy as np
import pytest

import oracle.fileio as OF


def beta3(t):
    a = np.abs(t)
    return np.where(a < 1.0, 2.0 / 3.0 - a * a + 0.5 * a ** 3,
                    np.where(a < 2.0, (2.0 - a) ** 3 / 6.0, 0.0))


def prefilter_matrix(n, bc="line"):
    """(n+2) x n map from samples to padded coefficients."""
    M = np.zeros((n + 2, n + 2))
    for j in range(1, n + 1):
        M[j, j - 1:j + 2] = [1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0]
    if bc == "line":
        M[0, 0:3] = [1.0, -2.0, 1.0]
        M[n + 1, n - 1:n + 2] = [1.0, -2.0, 1.0]
    else:   # planted mistake: Flat, c_0 = c_2 and c_{n+1} = c_{n-1}
        M[0, [0, 2]] = [1.0, -1.0]
        M[n + 1, [n - 1, n + 1]] = [-1.0, 1.0]
    rhs = np.zeros((n + 2, n))
    rhs[1:n + 1] = np.eye(n)
    return np.linalg.solve(M, rhs)


def eval_matrix(n, xs):
    """len(xs) x (n+2): beta3(x - j) for coefficient j = 0 .. n+1."""
    j = np.arange(n + 2, dtype=float)
    return beta3(np.asarray(xs, dtype=float)[:, None] - j[None, :])


def bspline_along(arr, axis, xs, bc):
    n = arr.shape[axis]
    A = eval_matrix(n, xs) @ prefilter_matrix(n, bc)        # samples -> values at xs
    return np.moveaxis(np.tensordot(A, np.moveaxis(arr, axis, 0), axes=(1, 0)), 0, axis)


def interpolate_grid_restated(grid, outshape, bc="line"):
    """src/utils.jl:64-109 step by step: squeeze, the interpolant of the squeezed array,
    sample points range(1, stop; length = out) with stop = 1 for an output axis of one
    point, and itp evaluated at the first ndims(itp) of (hx, hy, hz)."""
    g = np.asarray(grid, dtype=float).transpose(2, 1, 0)     # Julia (x, y, z) order
    inshape = g.shape
    keep = [i for i in range(3) if inshape[i] != 1]
    sqz = g.reshape([inshape[i] for i in keep])
    hooks = [np.linspace(1.0, 1.0 if outshape[i] == 1 else inshape[i], outshape[i])
             for i in range(3)]
    out = sqz
    for d in range(len(keep)):
        out = bspline_along(out, d, hooks[d], bc)
    # the evaluated dimensions are (hx, hy, ...)[:ndims]; the remaining output axes copy
    full = out.reshape(out.shape + (1,) * (3 - out.ndim))
    full = np.broadcast_to(full, tuple(outshape[i] if i >= len(keep) else full.shape[i]
                                       for i in range(3)))
    return np.ascontiguousarray(full.transpose(2, 1, 0))


SHAPES = [((64, 1, 1), (128, 1, 1)), ((48, 1, 1), (20, 1, 1)), ((16, 12, 1), (32, 24, 1)),
          ((8, 6, 4), (12, 9, 6)), ((16, 1, 1), (16, 8, 1)), ((16, 8, 1), (16, 1, 1))]


@pytest.mark.parametrize("inshape,outshape", SHAPES)
def test_oracle_is_the_cubic_line_bspline(inshape, outshape):
    rng = np.random.default_rng(11)
    g = rng.standard_normal((inshape[2], inshape[1], inshape[0]))
    ref = interpolate_grid_restated(g, outshape)
    got = OF.interpolate_grid(g, outshape)
    assert got.shape == ref.shape == (outshape[2], outshape[1], outshape[0])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_planted_flat_boundary_fails():
    rng = np.random.default_rng(12)
    g = rng.standard_normal((1, 1, 24))
    wrong = interpolate_grid_restated(g, (61, 1, 1), bc="flat")
    got = OF.interpolate_grid(g, (61, 1, 1))
    assert np.max(np.abs(wrong - got)) > 1e-3


def test_restatement_interpolates_samples_and_cubics_with_zero_end_curvature():
    # on the samples the spline returns the data; a cubic whose second derivative is zero
    # at both end knots (p'' linear through 0 at x = 1 and x = n means p'' = 0, so p is
    # linear) is reproduced, and a parabola is not (its curvature is nonzero at the ends)
    n = 11
    x = np.arange(1, n + 1, dtype=float)
    rng = np.random.default_rng(13)
    y = rng.standard_normal(n)
    np.testing.assert_allclose(bspline_along(y, 0, x, "line"), y, atol=1e-13)
    xs = np.linspace(1.0, n, 41)
    np.testing.assert_allclose(bspline_along(2.0 * x - 1.0, 0, xs, "line"), 2.0 * xs - 1.0,
                               atol=1e-12)
    par = bspline_along(x * x, 0, xs, "line")
    assert np.max(np.abs(par - xs * xs)) > 1e-3
