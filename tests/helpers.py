"""Builders shared by the tests: the same systems on the oracle and on the HIP library."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_seed(name):
    d = np.load(os.path.join(GOLDEN, f"seed_{name}.npz"))
    fields = {int(m): d[f"field_{int(m)}"] for m in d["mids"]}
    return d, fields


def diblock_system(mod, dims, dim, crystal, params, N=100, f=0.5, chi=0.2, Ns=100,
                   b=(1.0, 1.0), **kw):
    """A diblock melt built through the reference-shaped API of ``mod`` (the oracle or
    the HIP host package)."""
    cell = mod.UnitCell(dim, crystal, *params)
    a = mod.Monomer(1)
    bm = mod.Monomer(2)
    ch = mod.Diblock(a, bm, N, f, b[0], b[1], Ns)
    itx = mod.FloryInteraction()
    itx.set_interaction(1, 2, chi)
    sys = mod.FieldSystem(dims, cell, monomers=[a, bm], **kw)
    sys.add_species(ch, 1.0)
    sys.add_interaction(itx)
    return sys, cell, ch


def cos_fields(shape, amp=1.0):
    nz, ny, nx = shape
    x = np.arange(nx) / nx
    w = amp * np.cos(2 * np.pi * x)
    w = np.broadcast_to(w, shape).copy()
    return {1: w, 2: -w}
