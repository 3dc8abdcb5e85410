"""Converged variable-cell states, plain FP64 on the GPU against the oracle (scft! with
AndersonUpdater and DomainUpdater, src/simulate.jl:60-148, src/updaters/domain.jl:50-120).

SEEDED is the lamellar configuration of the reference README from seeded 0.1 N(0,1)
fields at L = 10, run to ftol = 1e-12, stol = 1e-14.  Its converged state depends on
last-bit rounding (the reference's dx = 1e-8 forward-difference Jacobian and position-
dependent stress), so the bounds are several times the oracle's own floor under one-ulp
transform perturbations (tests/test_oracle_vcell_floor.py).  COSINE starts from
centrosymmetric fields near the optimal cell; its fixed point is well posed and the GPU
reproduces fhelm, L and the density amplitudes to 1e-9.
"""
import pytest

import oracle as O
from test_oracle_vcell_floor import COSINE_BOUNDS, SEEDED_BOUNDS
from vcell_protocols import COSINE, SEEDED, compare, run

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")


@pytest.mark.parametrize("name,proto,bounds,rho_abs", [
    ("seeded", SEEDED, SEEDED_BOUNDS, 1e-5),
    ("cosine", COSINE, COSINE_BOUNDS, 1e-6)])
def test_variable_cell_converged_state_fp64(name, proto, bounds, rho_abs):
    ref = run(O, proto)
    gpu = run(P, proto, precision="fp64")
    assert ref[0].converged and gpu[0].converged
    assert gpu[0].ferr < proto["ftol"] and gpu[0].serr <= proto["stol"]
    d = compare(gpu, ref)
    print(name, gpu[0].nsteps, ref[0].nsteps, {k: f"{v:.2e}" for k, v in d.items()})
    for k, b in bounds.items():
        assert d[k] < b, (k, d[k])
    # pointwise densities also carry the lamellae's position, a neutral direction of the
    # fixed point (a translation changes neither fhelm nor the amplitudes)
    assert d["rho"] < rho_abs
