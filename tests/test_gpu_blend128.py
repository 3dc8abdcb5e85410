"""Large blend: 128^3 orthorhombic cell holding an ABC linear triblock (N = 200, block
fractions 0.3/0.4/0.3, b = 1.0/1.2/0.9, Ns = 200) and an A homopolymer (N = 100, volume
fraction 0.15); chi_AB N = 25, chi_BC N = 20, chi_AC N = 35 (N = 200); fields
0.1 N(0,1); cell edges a, b, c drawn uniform in [6, 10] Rg (Rg = sqrt(200/6) b).

* the same composition on a 16x12x8 grid: FP64 GPU against the FP64 oracle;
* 128^3: the FP64 GPU path against the FP64 oracle at full size (density, Q, free
  energy, stress; about 20 s of oracle time on the host);
* 128^3: FP32 propagator / FP64 sums against the FP64 GPU path (density, Q, free energy,
  field error, stress), and the FP64 free-energy gradient against central differences
  (h = 1e-6) in the fields and in the three cell parameters.
"""
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")

NREF = 200
RG = math.sqrt(NREF / 6.0)


def cell_params(seed=5):
    rng = np.random.default_rng(seed)
    return tuple(float(x) * RG for x in rng.uniform(6.0, 10.0, 3))


def blend(mod, dims, params, **kw):
    cell = mod.UnitCell(3, "orthorhombic", *params)
    A, B, C = mod.Monomer(1), mod.Monomer(2), mod.Monomer(3)
    tri = mod.Multiblock([A, B, C], [60, 80, 60], [1.0, 1.2, 0.9], 200)
    hom = mod.Homopolymer(A, 100, 1.0, 100)
    itx = mod.FloryInteraction()
    itx.set_interaction(1, 2, 25.0 / NREF)
    itx.set_interaction(2, 3, 20.0 / NREF)
    itx.set_interaction(1, 3, 35.0 / NREF)
    sys = mod.FieldSystem(dims, cell, monomers=[A, B, C], **kw)
    sys.add_species(tri, 0.85)
    sys.add_species(hom, 0.15)
    sys.add_interaction(itx)
    return sys


def seed_fields(sys, seed=11, scale=0.1):
    rng = np.random.default_rng(seed)
    f = {m: scale * rng.standard_normal(sys.shape) for m in sys.mids}
    for m, w in f.items():
        sys.fields[m] = w
    return f


def test_blend_small_fp64_parity():
    # same grid spacing as the 128^3 cell on a 16x12x8 grid
    params = tuple(p * d / 128 for p, d in zip(cell_params(), (16, 12, 8)))
    so = blend(O, (16, 12, 8), params)
    sg = blend(P, (16, 12, 8), params)
    f = seed_fields(so)
    for m, w in f.items():
        sg.fields[m] = w
    O.fieldinit(so, fields=f)
    P.fieldinit(sg, fields=f)
    for isp in range(2):
        assert sg.species[isp].Q == pytest.approx(so.species[isp].Q, rel=1e-12)
    for m in so.mids:
        np.testing.assert_allclose(sg.density[m], so.density[m], rtol=1e-11, atol=1e-14)
        np.testing.assert_allclose(sg.residuals[m], so.residuals[m], rtol=1e-9, atol=1e-12)
    assert sg.free_energy() == pytest.approx(so.free_energy(), rel=1e-11)
    np.testing.assert_allclose(sg.scfstress(), so.scfstress(), rtol=1e-9, atol=1e-14)


@pytest.fixture(scope="module")
def big():
    params = cell_params()
    s64 = blend(P, (128, 128, 128), params)
    f = seed_fields(s64)
    s32 = blend(P, (128, 128, 128), params, precision="fp32")
    for m, w in f.items():
        s32.fields[m] = w
    return s64, s32, f


def test_blend_128_fp64_matches_oracle_full_size(big):
    import os
    import time
    s64, _, f = big
    O.set_fft_workers(max(1, min(16, len(os.sched_getaffinity(0)))))
    so = blend(O, (128, 128, 128), cell_params())
    for m, w in f.items():
        so.fields[m] = w
    t0 = time.perf_counter()
    so.density_update()
    st_o = so.scfstress()
    t_o = time.perf_counter() - t0
    for m, w in f.items():
        s64.fields[m] = w
    s64.density_update()
    for isp in range(2):
        assert s64.species[isp].Q == pytest.approx(so.species[isp].Q, rel=1e-12)
    e_rho = 0.0
    for m in so.mids:
        a, b = np.asarray(s64.density[m]), np.asarray(so.density[m])
        e_rho = max(e_rho, float(np.max(np.abs(a - b)) / np.max(np.abs(b))))
    assert e_rho < 1e-11
    assert s64.free_energy() == pytest.approx(so.free_energy(), rel=1e-11)
    st_g = s64.scfstress()
    np.testing.assert_allclose(st_g, st_o, rtol=1e-9, atol=1e-15)
    print(f"128^3 blend FP64 GPU vs oracle: rho {e_rho:.2e} relative, stress "
          f"{float(np.max(np.abs(st_g - st_o) / np.abs(st_o))):.2e}; oracle {t_o:.1f} s")
    del so


def test_blend_128_fp32_vs_fp64(big):
    import time
    s64, s32, f = big
    out = {}
    for name, s in (("fp64", s64), ("fp32", s32)):
        for m, w in f.items():
            s.fields[m] = w
        s.density_update()
        s.synchronize()
        t0 = time.perf_counter()
        s.density_update()
        s.synchronize()
        dt = time.perf_counter() - t0
        out[name] = dict(t=dt, Q=[sp.Q for sp in s.species], F=s.free_energy(),
                         rho={m: s.density[m] for m in s.mids}, stress=s.scfstress())
    e_rho = max(float(np.max(np.abs(out["fp32"]["rho"][m] - out["fp64"]["rho"][m])))
                for m in s64.mids)
    e_Q = max(abs(a / b - 1) for a, b in zip(out["fp32"]["Q"], out["fp64"]["Q"]))
    e_F = abs(out["fp32"]["F"] - out["fp64"]["F"])
    e_S = float(np.max(np.abs(out["fp32"]["stress"] - out["fp64"]["stress"]))
                / np.max(np.abs(out["fp64"]["stress"])))
    steps = 2 * 200 + 100
    print(f"128^3 blend: density! fp64 {out['fp64']['t'] * 1e3:.1f} ms, fp32 "
          f"{out['fp32']['t'] * 1e3:.1f} ms ({steps} MDE steps); fp32 vs fp64: rho {e_rho:.2e}, "
          f"Q {e_Q:.2e}, fhelm {e_F:.2e}, stress {e_S:.2e}")
    assert e_rho < 1e-4 and e_Q < 1e-4 and e_F < 1e-5 and e_S < 1e-3


def test_blend_128_gradient_fd(big):
    s64, _, f = big
    for m, w in f.items():
        s64.fields[m] = w
    g = P.fhelm_grad(s64, "textbook")

    def F():
        s64.density_update()
        v = s64.free_energy()
        for isp, sp in enumerate(s64.species):
            phi = s64.phi_species[isp]
            v += phi * (math.log(phi / sp.Q) - phi) / sp.Nref
        return v

    rng = np.random.default_rng(3)
    for _ in range(3):
        d = {m: rng.standard_normal(s64.shape) for m in s64.mids}
        lin = sum(float(np.sum(g.dw[m] * d[m])) for m in s64.mids)
        vals = []
        for h in (1e-6, -1e-6):
            for m in s64.mids:
                s64.fields[m] = f[m] + h * d[m]
            vals.append(F())
        fd = (vals[0] - vals[1]) / 2e-6
        assert lin == pytest.approx(fd, rel=1e-5, abs=1e-9)
    for m in s64.mids:
        s64.fields[m] = f[m]
    p0 = np.array(s64.cell.params)
    for k in range(3):
        vals = []
        for h in (1e-6, -1e-6):
            p = p0.copy()
            p[k] += h
            s64.cell.params = p
            vals.append(F())
        s64.cell.params = p0
        fd = (vals[0] - vals[1]) / 2e-6
        print(f"128^3 blend dF/dtheta[{k}] = {g.dcell[k]:.10e}, FD {fd:.10e}")
        assert g.dcell[k] == pytest.approx(fd, rel=1e-5, abs=1e-10)


def test_blend_128_checkpointed_gradient(big):
    """sqrt(Ns) checkpointing on the 128^3 blend: the gradient (cell derivative included)
    from recomputed segments equals the full-stack one.  Stacks held per chain: full
    2 (Ns + 1) slices; checkpointed 2 (Ns/K + 1) + 2 (K + 1) with K = ceil(sqrt(Ns + 1))."""
    s64, _, f = big
    for m, w in f.items():
        s64.fields[m] = w
    gf = P.fhelm_grad(s64, "textbook")
    ck = blend(P, (128, 128, 128), cell_params())
    ck.set_checkpoint(-1)
    for m, w in f.items():
        ck.fields[m] = w
    gc = P.fhelm_grad(ck, "textbook")
    assert gc.fhelm == pytest.approx(gf.fhelm, rel=1e-12)
    np.testing.assert_allclose(gc.dcell, gf.dcell, rtol=1e-9)
    for m in s64.mids:
        err = np.max(np.abs(gc.dw[m] - gf.dw[m])) / np.max(np.abs(gf.dw[m]))
        assert err < 1e-10
    full = sum(2 * (sp.Ns + 1) for sp in s64.species[:1]) + (s64.species[1].Ns + 1)
    kept = 0
    for sp in s64.species:
        K = math.ceil(math.sqrt(sp.Ns + 1))
        nseg = math.ceil(sp.Ns / K)
        kept += 2 * (nseg + 1) + 2 * (K + 1)
    print(f"128^3 blend: slices held {full} (full stacks) vs {kept} (checkpointed), "
          f"{full * 128 ** 3 * 8 / 2 ** 30:.2f} GiB vs {kept * 128 ** 3 * 8 / 2 ** 30:.2f} GiB FP64")


def test_blend_128_variable_cell_scft_checkpointed():
    """The config's variable-cell SCFT on the 128^3 blend (FP64): scft! with a
    DomainUpdater relaxing the three orthorhombic edges, on sqrt(Ns)-checkpointed stacks
    (the stress passes recompute the segments) against the same run on full stacks.  The
    forward-difference cell Jacobian (dx = 1e-8) scales the rounding gap between the two
    Simpson summation orders (1e-15) by about 1e8, hence the bounds."""
    out = []
    for K in (0, -1):
        s = blend(P, (128, 128, 128), cell_params(), precision="fp64")
        s.set_checkpoint(K)
        seed_fields(s)
        start = np.array(s.cell.params, dtype=float)
        r = P.scft(s, P.MomentumUpdater(method="MOM", lam=0.05),
                   domain=P.DomainUpdater(nskip=4, nper=2, lam=0.1),
                   opts=P.SCFTOptions(nsteps=12, ftol=1e-9, stol=1e-9))
        cell = np.array(s.cell.params, dtype=float)
        print(f"K={K}: nsteps {r.nsteps} fhelm {r.fhelm:.12g} ferr {r.ferr:.4g} serr "
              f"{r.serr:.4g} cell {cell}", flush=True)
        out.append((r, cell, start))
    (rf, cf, start), (rc, cc, _) = out
    assert rf.nsteps == rc.nsteps == 12
    assert np.max(np.abs(cf - start)) > 1e-6 * np.max(start)     # the cell moved
    assert rc.fhelm == pytest.approx(rf.fhelm, rel=1e-7)
    np.testing.assert_allclose(cc, cf, rtol=1e-6)
    assert rc.ferr == pytest.approx(rf.ferr, rel=1e-4)
