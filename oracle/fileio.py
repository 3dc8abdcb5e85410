"""Field files and grid interpolation -- restates src/fileio.jl:77-133 and
interpolate_grid (src/utils.jl:64-109).  Test infrastructure only.

interpolate_grid: Interpolations.jl 0.13.1 BSpline(Cubic(Line(OnGrid()))) is the cubic
spline through the samples with zero second derivative at the end knots (the "Line"
boundary condition), i.e. the natural cubic spline; evaluated with scipy's
CubicSpline(bc_type="natural") at range(1, n_in; length = n_out).

Pinned by tests/test_oracle_interp_bspline.py: a restatement of Interpolations.jl
0.13.1's published algorithm for that interpolant (B-spline prefilter with the Line
boundary rows c_0 - 2c_1 + c_2 = 0, tensor-product evaluation, the reference's squeeze and
sample points) agrees with this function to 1e-12 on 1-D, 2-D and 3-D grids; the Flat
boundary rows (a planted mistake) do not.  The reference's only printed use (examples.ipynb,
lam1d.json interpolated 64 -> 128, ferr 9.713161299981836e-6) is not reproduced by the
current source's sampling range (which stretches the profile by 127/126 and gives ferr
~0.1).
"""
import json

import numpy as np
from scipy.interpolate import CubicSpline


def interpolate_grid(grid, outshape):
    g = np.asarray(grid, dtype=float)
    inshape = (g.shape[2], g.shape[1], g.shape[0])
    out = g.transpose(2, 1, 0)
    for ax in range(3):
        nin, nout = inshape[ax], outshape[ax]
        if nout == 1:
            out = np.take(out, [0], axis=ax)
        elif nin == 1:
            out = np.repeat(out, nout, axis=ax)
        else:
            xs = np.arange(1, nin + 1, dtype=float)
            out = CubicSpline(xs, out, axis=ax, bc_type="natural")(np.linspace(1.0, nin, nout))
    return np.ascontiguousarray(out.transpose(2, 1, 0))


def loadfields(path):
    with open(path) as fh:
        raw = json.load(fh)
    nx, ny, nz = raw["grid"]
    return {"grid": tuple(raw["grid"]), "cell_params": list(raw["cell_params"]),
            "crystal": raw["crystal"], "field_type": raw["field_type"],
            "fields": {int(m): np.array(raw["fields"][str(m)], dtype=float).reshape(nz, ny, nx)
                       for m in raw["mids"]}}
