"""Instances of the 256-system hexagonal sweep (bench.py's batched_scft workload) that the
GPU batch does not solve fail in the reference algorithm too.

The sweep's protocol is scft!(AndersonUpdater(nhist=10, iter_euler=10^9, lam_euler=0.05);
nsteps=1500, ftol=1e-2) (pure Euler mixing, src/updaters/anderson.jl:98-147), then
AndersonUpdater(nhist=20) to ftol=1e-6.  For instances 94 and 235 the fields diverge in
the first phase until the propagator's Q turns negative, where the reference's muphi!
raises a DomainError; the oracle, running the same algorithm on the CPU, fails on both,
and the GPU batch reports both as diverged (tests/test_gpu_sweep_failures.py).
"""
import os
import sys
import warnings

import numpy as np
import pytest

import oracle as O

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def _instance(i):
    n, Ns, NA, chiN, a, fs = bench.sweep_instances(256)[i]
    cell = O.UnitCell(2, "hexagonal", a)
    A, B = O.Monomer(1), O.Monomer(2)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ch = O.Multiblock([A, B], [NA, 100 - NA], [1.0, 1.0], Ns)
    itx = O.FloryInteraction()
    itx.set_interaction(1, 2, chiN / 100.0)
    s = O.FieldSystem((n, n, 1), cell, monomers=[A, B])
    s.add_species(ch, 1.0)
    s.add_interaction(itx)
    rng = np.random.default_rng(fs)
    O.fieldinit(s, fields={m: 0.1 * rng.standard_normal((1, n, n)) for m in (1, 2)})
    return s


def _phase1(s, log=None):
    p = bench.PH1
    return O.scft(s, O.AndersonUpdater(nhist=p["nhist"], iter_euler=p["iter_euler"],
                                       lam_euler=p["lam_euler"]),
                  opts=O.SCFTOptions(nsteps=p["nsteps"], ftol=p["ftol"]), log=log)


@pytest.mark.parametrize("i", [94, 235])
def test_unsolved_sweep_instances_fail_in_the_oracle(i):
    # the reference stops with a DomainError once Q turns negative (log(phi / Q) in
    # muphi!, src/system/system.jl:251), or with "Diverging field updater" once ferr is NaN
    log = []
    with pytest.raises((RuntimeError, ValueError), match="Diverging|DomainError"):
        _phase1(_instance(i), log)
    assert 100 < len(log) <= bench.PH1["nsteps"]
    assert not any(l[1] < bench.PH1["ftol"] for l in log)
