"""Field files and seeding (host side): restates src/fileio.jl and interpolate_grid
(src/utils.jl:64-109).

The JSON layout is the reference's (fileio.jl:77-100): keys dim, grid, mids, date,
crystal, cell_params, field_type and fields {mid: [z][y][x]}.  Grids are numpy arrays of
shape (Nz, Ny, Nx), the reference's x-fastest order.
"""
import datetime
import json

import numpy as np

from ._lib import call, dptr


def _natural_spline_weights(n):
    """Second-derivative solve of the natural cubic spline on knots 1..n: returns a
    function mapping samples (n, ...) to second derivatives (n, ...).  Natural boundary
    conditions (zero second derivative at both end knots) are what Interpolations.jl's
    BSpline(Cubic(Line(OnGrid()))) imposes."""
    def solve(y):
        y = np.asarray(y, dtype=float)
        if n < 3:
            return np.zeros_like(y)
        # tridiagonal system for the interior second derivatives: m[i-1] + 4 m[i] + m[i+1]
        # = 6 (y[i-1] - 2 y[i] + y[i+1]), m[0] = m[n-1] = 0 (unit knot spacing)
        rhs = 6.0 * (y[:-2] - 2.0 * y[1:-1] + y[2:])
        k = n - 2
        c = np.zeros(k)
        d = np.zeros_like(rhs)
        b = 4.0
        c[0] = 1.0 / b
        d[0] = rhs[0] / b
        for i in range(1, k):
            den = 4.0 - c[i - 1]
            c[i] = 1.0 / den
            d[i] = (rhs[i] - d[i - 1]) / den
        m = np.zeros_like(y)
        m[k] = d[k - 1]
        for i in range(k - 2, -1, -1):
            m[i + 1] = d[i] - c[i] * m[i + 2]
        return m
    return solve


def _spline_eval(y, m, pts):
    """Evaluate the natural cubic spline (samples y, second derivatives m on knots 1..n)
    at 1-based positions pts, along axis 0."""
    n = y.shape[0]
    j = np.clip(np.floor(pts).astype(int), 1, max(n - 1, 1))   # interval [j, j+1]
    t = pts - j
    j0 = j - 1
    j1 = np.minimum(j0 + 1, n - 1)
    a = 1.0 - t
    shape = (-1,) + (1,) * (y.ndim - 1)
    a = a.reshape(shape)
    t = t.reshape(shape)
    return (a * y[j0] + t * y[j1] +
            ((a ** 3 - a) * m[j0] + (t ** 3 - t) * m[j1]) / 6.0)


def interpolate_grid(grid, outshape):
    """``interpolate_grid(grid, outshape)`` (src/utils.jl:64-109): cubic B-spline with
    natural ("Line") boundary conditions through the samples, evaluated at
    ``range(1, n_in; length = n_out)`` along every non-singleton axis.  ``grid`` and the
    result use the reference's (Nx, Ny, Nz) dims order via (Nz, Ny, Nx) arrays;
    ``outshape`` is (Nx, Ny, Nz)."""
    g = np.asarray(grid, dtype=float)
    inshape = (g.shape[2], g.shape[1], g.shape[0])
    out = g.transpose(2, 1, 0)                  # axes x, y, z
    for ax in range(3):
        nin, nout = inshape[ax], outshape[ax]
        if nout == 1:
            out = np.take(out, [0], axis=ax)     # single output point: hook at 1
            continue
        if nin == 1:                             # copy along a singleton input axis
            out = np.repeat(out, nout, axis=ax)
            continue
        pts = np.linspace(1.0, float(nin), nout)
        y = np.moveaxis(out, ax, 0)
        m = _natural_spline_weights(nin)(y)
        out = np.moveaxis(_spline_eval(y, m, pts), 0, ax)
    return np.ascontiguousarray(out.transpose(2, 1, 0))


def loadfields(path):
    """``loadfields(fpath)`` (src/fileio.jl:107-133)."""
    with open(path) as fh:
        raw = json.load(fh)
    d = {
        "dim": int(raw["dim"]),
        "grid": tuple(int(x) for x in raw["grid"]),
        "mids": tuple(int(x) for x in raw["mids"]),
        "crystal": str(raw["crystal"]),
        "cell_params": [float(x) for x in raw["cell_params"]],
        "field_type": str(raw["field_type"]),
    }
    nx, ny, nz = d["grid"]
    d["fields"] = {}
    for mid in d["mids"]:
        arr = np.array(raw["fields"][str(mid)], dtype=float)
        if arr.shape != (nz, ny, nx):
            raise ValueError(f"field {mid} has shape {arr.shape}, grid says {(nz, ny, nx)}")
        d["fields"][mid] = arr
    return d


def savefields(sys, path, field_type="omega"):
    """``savefields(sys, fpath; field_type)`` (src/fileio.jl:77-100)."""
    view = sys.density if field_type == "rho" else sys.fields
    d = {
        "dim": sys.cell.dim,
        "grid": list(sys.dims),
        "mids": sys.mids,
        "date": datetime.datetime.now().strftime("%Y-%m-%d %H:%M:%S"),
        "crystal": sys.cell.crystal,
        "cell_params": [float(x) for x in sys.cell.params],
        "field_type": field_type,
        "fields": {str(m): view[m].tolist() for m in sys.mids},
    }
    with open(path, "w") as fh:
        json.dump(d, fh, indent=4)


def fieldinit_dict(sys, fields, field_type="omega", interpolate=True):
    """``fieldinit!(sys, fields::Dict; field_type, interpolate)`` (src/fileio.jl:25-54)."""
    import warnings
    for mid, arr in fields.items():
        if mid not in sys.monomers:
            warnings.warn(f"Monomer with ID = {mid} not present in system -- skipping.")
            continue
        a = np.asarray(arr, dtype=float)
        if a.shape != sys.shape:
            if not interpolate:
                raise ValueError(f"Input field dimensions {a.shape} does not match system "
                                 f"grid {sys.shape}.")
            a = interpolate_grid(a, sys.dims)
        a = np.ascontiguousarray(a)
        if field_type == "rho":
            call("pf_system_set_density", sys._h, int(mid), dptr(a))
        else:
            sys.fields[mid] = a
    if field_type == "rho":
        call("pf_uniform_fields", sys._h)
    sys.residuals_update()


def fieldinit_file(sys, path, load_cell=True, interpolate=True):
    """``fieldinit!(sys, fpath; load_cell, interpolate)`` (src/fileio.jl:56-64)."""
    d = loadfields(path)
    if load_cell:
        from .api import UnitCell
        sys.add_cell(UnitCell(d["dim"], d["crystal"], *d["cell_params"]))
    fieldinit_dict(sys, d["fields"], field_type=d["field_type"], interpolate=interpolate)
