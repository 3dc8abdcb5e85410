"""Field files and grid interpolation -- restates src/fileio.jl:77-133 and
interpolate_grid (src/utils.jl:64-109).  Test infrastructure only.

interpolate_grid: Interpolations.jl 0.13.1 BSpline(Cubic(Line(OnGrid()))) is the cubic
spline through the samples with zero second derivative at the end knots (the "Line"
boundary condition), i.e. the natural cubic spline; evaluated with scipy's
CubicSpline(bc_type="natural") at range(1, n_in; length = n_out).

Parity unpinned: the reference's only printed use (examples.ipynb, lam1d.json
interpolated 64 -> 128, ferr 9.713161299981836e-6) is not reproduced by the current
source's sampling range (which stretches the profile by 127/126 and gives ferr ~0.1); it
is pinned here only by identity, exact reproduction of linear data and scipy's spline.
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
