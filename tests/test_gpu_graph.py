"""Graph replay of the batched density!/residuals! launch sequence (pf_batch_set_graph).

One captured graph is replayed several times, with new fields and eager work of the
program (a single-system density! on the same device, host round trips) between the
replays; every replay must equal the eager launch sequence on the same inputs bit for bit.
"""
import numpy as np
import pytest

from helpers import diblock_system

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")


def systems(precision):
    out = []
    for dims, dim, crystal, params, f in [
        ((32, 32, 1), 2, "hexagonal", (4.0,), 0.3),
        ((48, 48, 1), 2, "hexagonal", (5.0,), 0.4),
        ((24, 24, 1), 2, "hexagonal", (4.5,), 0.35),   # generic planar kernel
        ((32, 32, 32), 3, "cubic", (4.0,), 0.35),
        ((32, 32, 32), 3, "cubic", (4.5,), 0.3),
    ]:
        s, _, _ = diblock_system(P, dims, dim, crystal, params, N=100, f=f, chi=0.2, Ns=40,
                                 precision=precision)
        out.append(s)
    return out


def field_sets(syss, k):
    rng = np.random.default_rng(100 + k)
    return [{m: 0.3 * rng.standard_normal(s.shape) for m in s.mids} for s in syss]


def state(syss):
    return [np.stack([s.density[m] for m in s.mids] + [s.residuals[m] for m in s.mids] +
                     [s.fields[m] for m in s.mids]) for s in syss]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_equal_eager(precision):
    syss = systems(precision)
    batch = P.Batch(syss)
    extra, _, _ = diblock_system(P, (16, 16, 1), 2, "hexagonal", (4.0,), precision=precision)
    extra.fields[1] = np.zeros(extra.shape)
    extra.fields[2] = np.zeros(extra.shape)
    results = {}
    for mode in (True, False):
        batch.use_graph(mode)
        results[mode] = []
        for k in range(4):
            for s, f in zip(syss, field_sets(syss, k)):
                for m, w in f.items():
                    s.fields[m] = w
            batch.residuals_update()
            results[mode].append(state(syss))
            extra.density_update()                 # eager work between replays
            extra.synchronize()
            batch.density_update()                 # the other captured sequence
    for k in range(4):
        for a, b in zip(results[True][k], results[False][k]):
            assert np.array_equal(a, b), f"replay {k} differs from the eager launches"


def test_graph_batched_scft_equals_eager():
    res = {}
    for mode in (True, False):
        syss = systems("fp64")[:3]
        for s, f in zip(syss, field_sets(syss, 7)):
            P.fieldinit(s, fields=f)
        batch = P.Batch(syss)
        batch.use_graph(mode)
        r = batch.scft(nhist=10, iter_euler=1000000000, lam_euler=0.1, nsteps=60, ftol=1e-12)
        res[mode] = [(x.nsteps, x.ferr, x.fhelm) for x in r]
    assert res[True] == res[False]


def test_graph_sweep_of_64_systems_equals_eager_fp32():
    # 64 two-dimensional systems on four grid sizes (four lane groups on their own
    # streams inside the graph), 40 replayed FP32 Euler iterations: every system's free
    # energy equals the eager run's bit for bit, in repeated runs
    def run(graph):
        ss = []
        for k in range(64):
            n = (32, 48, 64, 96)[k % 4]
            s, _, _ = diblock_system(P, (n, n, 1), 2, "hexagonal", (4.0 + 0.03 * k,), N=100,
                                     f=(0.3, 0.35, 0.4)[k % 3], chi=0.12 + 0.001 * k, Ns=40 + (k % 9),
                                     precision="fp32")
            rng = np.random.default_rng(700 + k)
            for m in (1, 2):
                s.fields[m] = 0.1 * rng.standard_normal(s.shape)
            ss.append(s)
        b = P.Batch(ss)
        b.use_graph(graph)
        r = b.scft(nhist=10, iter_euler=1000, lam_euler=0.1, nsteps=40, ftol=1e-14)
        return np.array([x.fhelm for x in r])

    ref = run(False)
    for _ in range(3):
        np.testing.assert_array_equal(run(True), ref)
