"""Anderson mixing with histories deeper than the one-pass Gram kernel serves (the
reference's default AndersonUpdater() keeps nhist = 50, src/updaters/anderson.jl:36;
Gram loop :147-181, mixing :211-224, history shift :209-225).

The GPU runs the tiled Gram for every step of such an updater.  Bounds: the oracle's own
trajectory moves, under one-ulp perturbations of its transforms, by about 5e-12 (ferr)
in the first 30 steps and 1e-7 by step 60 (the least-squares problem over a full
50-deep history amplifies rounding); the bounds below sit two to three orders of
magnitude above that, and far below what a wrong history slot would give.
"""
import numpy as np
import pytest

import oracle as O
from helpers import diblock_system

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")


def _hex(mod, **kw):
    return diblock_system(mod, (32, 32, 1), 2, "hexagonal", (9.0,), f=0.3, chi=0.25, Ns=50,
                          **kw)


def test_anderson_default_depth_trajectory_fp64():
    rng = np.random.default_rng(0)
    fields = {m: 0.5 * rng.standard_normal((1, 32, 32)) for m in (1, 2)}
    (so, _, _), (sg, _, _) = _hex(O), _hex(P, precision="fp64")
    O.fieldinit(so, fields=fields)
    P.fieldinit(sg, fields=fields)
    uo, ug = O.AndersonUpdater(), P.AndersonUpdater()   # the default nhist = 50
    assert uo.nhist == ug.nhist == 50
    uo.setup(so)
    ug.setup(sg)
    for it in range(61):   # iterations 1..61: the history is full from 51 and shifts
        uo.step()
        ug.step()
        fo, fg = so.field_error(), sg.field_error()
        eo, eg = so.free_energy(), sg.free_energy()
        tf, te = (1e-9, 1e-10) if it < 30 else (1e-5, 1e-7)
        assert fg == pytest.approx(fo, rel=tf), it
        assert eg == pytest.approx(eo, rel=te), it


def test_batch_deep_history_matches_one_at_a_time_fp64():
    # the batch's tiled Gram uses the same partition of each system's points as the
    # one-system path, so every trajectory is the same to rounding of the host solve
    cfgs = [(32, 4.5, 30, 14.0, 64), (48, 5.5, 35, 11.0, 80), (32, 5.0, 45, 12.0, 70)]

    def make(k):
        n, a, NA, chiN, Ns = cfgs[k]
        cell = P.UnitCell(2, "hexagonal", a)
        A, B = P.Monomer(1), P.Monomer(2)
        ch = P.Multiblock([A, B], [NA, 100 - NA], [1.0, 1.0], Ns)
        itx = P.FloryInteraction()
        itx.set_interaction(1, 2, chiN / 100.0)
        s = P.FieldSystem((n, n, 1), cell, monomers=[A, B], precision="fp64")
        s.add_species(ch, 1.0)
        s.add_interaction(itx)
        P.fieldinit(s, seed=300 + k, scale=0.1)
        return s

    opts = dict(nsteps=40, ftol=1e-14)
    ref = [P.scft(make(k), P.AndersonUpdater(nhist=24, iter_euler=3, lam_euler=0.05),
                  opts=P.SCFTOptions(**opts)) for k in range(len(cfgs))]
    res = P.Batch([make(k) for k in range(len(cfgs))]).scft(nhist=24, iter_euler=3,
                                                             lam_euler=0.05, **opts)
    for r, q in zip(res, ref):
        assert np.isfinite(q.fhelm)
        assert r.nsteps == q.nsteps
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-13)
        assert r.ferr == pytest.approx(q.ferr, rel=1e-12)
