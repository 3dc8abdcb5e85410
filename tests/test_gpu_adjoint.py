"""Free-energy gradient on the GPU (pf_fhelm_grad / pf_density_vjp through the C ABI).

* Parity: FP64 GPU gradient against the FP64 oracle's reverse sweep (itself pinned by
  central finite differences in tests/test_oracle_adjoint.py) on 1-D, 2-D, 3-D cells and
  a blend with a homopolymer, a point species and Edwards terms.
* The 64^3 cubic diblock (N = 100, fA = 0.35, chi N = 20, Ns = 100 -> 102, fields
  0.5 N(0,1)): gradient with respect to the fields on 64 random directions, the cell
  parameter, chi and fA against central finite differences (h = 1e-6, FP64) of the GPU
  free energy.
* The reference's analytic stress (multiblock.jl:233-295) as the continuum limit of
  d log Q / d theta, and the reference's SCF residual: at a converged state the field
  gradient of the incompressible textbook functional is the Simpson-vs-trapezoid
  quadrature gap.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from oracle import adjoint as OA

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")
from polyfields_hip._lib import call  # noqa: E402


def diblock(mod, dims, dim, crystal, params, N=100, f=0.35, chi=0.2, Ns=20, b=(1.0, 1.2),
            **kw):
    cell = mod.UnitCell(dim, crystal, *params)
    a, bm = mod.Monomer(1), mod.Monomer(2)
    ch = mod.Diblock(a, bm, N, f, b[0], b[1], Ns)
    itx = mod.FloryInteraction()
    itx.set_interaction(1, 2, chi)
    sys = mod.FieldSystem(dims, cell, monomers=[a, bm], **kw)
    sys.add_species(ch, 1.0)
    sys.add_interaction(itx)
    return sys


def blend(mod, **kw):
    cell = mod.UnitCell(2, "rectangular", 7.0, 5.0)
    A, B, Cm = mod.Monomer(1, 1.0), mod.Monomer(2, 1.3), mod.Monomer(3, 0.8)
    tri = mod.Multiblock([A, B, A], [12, 20, 8], [1.0, 1.1, 1.0], 20)
    hom = mod.Homopolymer(B, 30, 0.9, 16)
    pt = mod.Point(Cm)
    fl = mod.FloryInteraction()
    fl.set_interaction(1, 2, 0.3)
    fl.set_interaction(2, 3, 0.15)
    fl.set_interaction(1, 3, 0.4)
    ed = mod.EdwardsInteraction()
    ed.set_interaction(2, 0.7)
    sys = mod.FieldSystem((12, 8, 1), cell, monomers=[A, B, Cm], **kw)
    sys.add_species(tri, 0.6)
    sys.add_species(hom, 0.3)
    sys.add_species(pt, 0.1)
    sys.add_interaction(fl)
    sys.add_interaction(ed)
    return sys


def set_fields(sys, fields):
    for m, w in fields.items():
        sys.fields[m] = np.ascontiguousarray(w)


def rand_fields(sys, scale, seed):
    rng = np.random.default_rng(seed)
    return {m: scale * rng.standard_normal(sys.shape) for m in sys.mids}


def relmax(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


PARITY = [
    ("lam32", lambda m, **k: diblock(m, (32, 1, 1), 1, "lamellar", (9.0,), **k)),
    ("hex16", lambda m, **k: diblock(m, (16, 16, 1), 2, "hexagonal", (8.0,), Ns=16, **k)),
    ("cubic8", lambda m, **k: diblock(m, (8, 8, 8), 3, "cubic", (6.0,), Ns=12, **k)),
    ("tric8x6x4", lambda m, **k: diblock(m, (8, 6, 4), 3, "triclinic",
                                          (5.0, 4.5, 4.0, 1.4, 1.5, 1.7), Ns=10, **k)),
    ("cubic32", lambda m, **k: diblock(m, (32, 32, 32), 3, "cubic", (7.0,), Ns=10, **k)),
    ("blend12x8", blend),
]


@pytest.mark.parametrize("functional", ["reference", "textbook"])
@pytest.mark.parametrize("name,build", PARITY, ids=[p[0] for p in PARITY])
def test_gradient_parity_fp64(name, build, functional):
    so, sg = build(O), build(P)
    fields = rand_fields(so, 0.5, seed=7)
    set_fields(so, fields)
    set_fields(sg, fields)
    go = OA.fhelm_grad(so, functional)
    gg = P.fhelm_grad(sg, functional)
    assert gg.fhelm == pytest.approx(go.fhelm, rel=1e-11, abs=1e-13)
    dwo = np.stack([go.dw[m] for m in so.mids])
    dwg = np.stack([gg.dw[m] for m in sg.mids])
    assert relmax(dwg, dwo) < 1e-10
    assert relmax(gg.dcell, go.dcell) < 1e-9
    for k, v in go.dchi.items():
        assert gg.dchi[k] == pytest.approx(v, rel=1e-11)
    for k, v in go.du.items():
        assert gg.du[k] == pytest.approx(v, rel=1e-11)
    assert relmax(gg.dphi, go.dphi) < 1e-10
    for a, b in zip(gg.dN, go.dN):
        assert len(a) == len(b)
        for x, y in zip(a, b):
            assert x == pytest.approx(y, rel=1e-9, abs=1e-13)


def test_density_vjp_parity_fp64():
    so, sg = PARITY[1][1](O), PARITY[1][1](P)
    fields = rand_fields(so, 0.5, seed=3)
    set_fields(so, fields)
    set_fields(sg, fields)
    rng = np.random.default_rng(4)
    rb = {m: rng.standard_normal(so.shape) / so.ngrid for m in so.mids}
    dwo, dco = OA.density_vjp(so, rb, [0.7])
    dwg, dcg, _ = P.density_vjp(sg, rb, [0.7])
    assert relmax(np.stack([dwg[m] for m in sg.mids]), np.stack([dwo[m] for m in so.mids])) < 1e-10
    assert relmax(dcg, dco) < 1e-9


# --------------------------------------------------------------------------------------
# 64^3 cubic diblock: gradient against central finite differences

A_RG = 4.0 * math.sqrt(100 / 6.0)       # a = 4.0 Rg in units of b, N = 100


def cfg4(precision="fp64"):
    return diblock(P, (64, 64, 64), 3, "cubic", (A_RG,), N=100, f=0.35, chi=0.2, Ns=100,
                   b=(1.0, 1.0), precision=precision)


def free_energy(sys, functional):
    sys.density_update()
    f = sys.free_energy()
    if functional == "textbook":
        for isp, sp in enumerate(sys.species):
            phi = sys.phi_species[isp]
            f += phi * (math.log(phi / sp.Q) - phi) / sp.Nref
    return f


def central(sys, functional, setter, h=1e-6):
    setter(h)
    fp = free_energy(sys, functional)
    setter(-2 * h)
    fm = free_energy(sys, functional)
    setter(h)
    return (fp - fm) / (2 * h)


@pytest.fixture(scope="module")
def cfg4_state():
    sys = cfg4()
    fields = rand_fields(sys, 0.5, seed=2024)
    set_fields(sys, fields)
    return sys, fields


@pytest.mark.parametrize("functional", ["reference", "textbook"])
def test_cfg4_gradient_matches_oracle_full_size(cfg4_state, functional):
    # the config-4 system at its full 64^3 grid: the GPU reverse sweep against the
    # oracle's (about 3 s of oracle time per functional on the host)
    import os
    sys, fields = cfg4_state
    O.set_fft_workers(max(1, min(16, len(os.sched_getaffinity(0)))))
    so = diblock(O, (64, 64, 64), 3, "cubic", (A_RG,), N=100, f=0.35, chi=0.2, Ns=100,
                 b=(1.0, 1.0))
    set_fields(so, fields)
    set_fields(sys, fields)
    go = OA.fhelm_grad(so, functional)
    gg = P.fhelm_grad(sys, functional)
    dwo = np.stack([go.dw[m] for m in so.mids])
    dwg = np.stack([gg.dw[m] for m in sys.mids])
    print(f"64^3 gradient ({functional}) GPU vs oracle: fhelm {abs(gg.fhelm / go.fhelm - 1):.2e}, "
          f"dw {relmax(dwg, dwo):.2e}, da {relmax(gg.dcell, go.dcell):.2e}", flush=True)
    assert gg.fhelm == pytest.approx(go.fhelm, rel=1e-11, abs=1e-13)
    assert relmax(dwg, dwo) < 1e-10
    assert relmax(gg.dcell, go.dcell) < 1e-9
    for k, v in go.dchi.items():
        assert gg.dchi[k] == pytest.approx(v, rel=1e-10)
    for a, b in zip(gg.dN, go.dN):
        for x, y in zip(a, b):
            assert x == pytest.approx(y, rel=1e-9, abs=1e-13)


@pytest.mark.parametrize("functional", ["reference", "textbook"])
def test_cfg4_fields_fd_64_directions(cfg4_state, functional):
    sys, fields = cfg4_state
    set_fields(sys, fields)
    g = P.fhelm_grad(sys, functional)
    rng = np.random.default_rng(99)
    lins, fds = [], []
    for _ in range(64):
        d = {m: rng.standard_normal(sys.shape) for m in sys.mids}
        lins.append(sum(float(np.sum(g.dw[m] * d[m])) for m in sys.mids))

        def setter(h, d=d):
            for m in sys.mids:
                fields[m] += h * d[m]
            set_fields(sys, fields)
        fds.append(central(sys, functional, setter))
    set_fields(sys, fields)
    lins, fds = np.array(lins), np.array(fds)
    err = np.abs(lins - fds)
    scale = float(np.sqrt(np.mean(fds ** 2)))
    rel = err / np.abs(fds)
    print(f"64^3 {functional}: 64 directions, |<g,d> - FD|: max {err.max():.2e}, "
          f"rms(FD) {scale:.2e}, max/rms {err.max() / scale:.2e}, "
          f"relative median {np.median(rel):.2e} max {rel.max():.2e}")
    # central FD with h = 1e-6 carries ~1e-15/h of rounding: each direction must agree to
    # 1e-6 of its own size or of the typical directional derivative, whichever is larger
    assert np.all(err <= 1e-6 * np.maximum(np.abs(fds), scale))
    assert np.median(rel) < 1e-7


def test_cfg4_cell_chi_fA_fd(cfg4_state):
    sys, fields = cfg4_state
    set_fields(sys, fields)
    for functional in ("reference", "textbook"):
        g = P.fhelm_grad(sys, functional)

        def cell_setter(h):
            p = np.array(sys.cell.params)
            p[0] += h
            sys.cell.params = p

        def chi_setter(h):
            itx = sys.interactions[0]
            itx.chis[(1, 2)] += h
            call("pf_system_set_flory", sys._h, 1, 2, itx.chis[(1, 2)])

        sp = sys.species[0]
        N = sp.N

        def f_setter(h):
            # Diblock(A, B, N, f): N_A = f N, N_B = (1 - f) N on the fixed contour grid
            NA, NB = sp.N_block
            P.set_block_lengths(sys, sp, [NA + h * N, NB - h * N])

        fd_cell = central(sys, functional, cell_setter)
        fd_chi = central(sys, functional, chi_setter)
        fd_f = central(sys, functional, f_setter)
        dF_df = N * (g.dN[0][0] - g.dN[0][1])
        print(f"64^3 {functional}: dF/da {g.dcell[0]:.12e} FD {fd_cell:.12e}; "
              f"dF/dchi {g.dchi[(1, 2)]:.12e} FD {fd_chi:.12e}; dF/dfA {dF_df:.12e} FD {fd_f:.12e}")
        assert g.dcell[0] == pytest.approx(fd_cell, rel=1e-6)
        assert g.dchi[(1, 2)] == pytest.approx(fd_chi, rel=1e-7)
        assert dF_df == pytest.approx(fd_f, rel=1e-6)


def test_cfg4_fp32_gradient_error(cfg4_state):
    """FP32 propagator / FP64 sums: error of the FP32 gradient against FP64."""
    sys, fields = cfg4_state
    set_fields(sys, fields)
    g64 = P.fhelm_grad(sys, "textbook")
    s32 = cfg4("fp32")
    set_fields(s32, fields)
    g32 = P.fhelm_grad(s32, "textbook")
    ew = relmax(np.stack([g32.dw[m] for m in s32.mids]), np.stack([g64.dw[m] for m in sys.mids]))
    ec = abs(g32.dcell[0] - g64.dcell[0]) / abs(g64.dcell[0])
    print(f"64^3 FP32 gradient vs FP64: dw {ew:.2e} (max-normalised), dF/da {ec:.2e}, "
          f"fhelm {abs(g32.fhelm - g64.fhelm):.2e}")
    assert ew < 1e-4 and ec < 1e-3


# --------------------------------------------------------------------------------------
# the reference's analytic stress and SCF residual

@pytest.mark.parametrize("case", ["tric32", "hex64x64", "blend32_ckpt"])
def test_compile_time_reverse_kernels_match_generic(case):
    # the reverse sweep through the compile-time kernels (32^2 / 64^2 planes, 32 / 64
    # lines) against the generic kernels of the same system (schedule 0 / 0)
    def make():
        if case == "tric32":
            s = diblock(P, (32, 32, 32), 3, "triclinic", (9.0, 8.5, 8.0, 1.4, 1.5, 1.7), Ns=16)
        elif case == "hex64x64":
            s = diblock(P, (64, 64, 1), 2, "hexagonal", (12.0,), Ns=16)
        else:
            cell = P.UnitCell(3, "orthorhombic", 9.0, 8.0, 7.0)
            A, B = P.Monomer(1), P.Monomer(2)
            tri = P.Multiblock([A, B, A], [12, 20, 8], [1.0, 1.1, 1.0], 20)
            hom = P.Homopolymer(B, 30, 0.9, 16)
            fl = P.FloryInteraction()
            fl.set_interaction(1, 2, 0.3)
            s = P.FieldSystem((32, 32, 32), cell, monomers=[A, B])
            s.add_species(tri, 0.7)
            s.add_species(hom, 0.3)
            s.add_interaction(fl)
            s.set_checkpoint(-1)
        rng = np.random.default_rng(3)
        P.fieldinit(s, fields={m: 0.4 * rng.standard_normal(s.shape) for m in s.mids})
        return s
    dims = {"tric32": (32, 32, 32), "hex64x64": (64, 64, 1), "blend32_ckpt": (32, 32, 32)}[case]
    try:
        P.schedule_set("fp64", dims, slab=0, pencil=0)
        gg = P.fhelm_grad(make(), "textbook")
    finally:
        P.schedule_reset()
    gf = P.fhelm_grad(make(), "textbook")
    assert gf.fhelm == pytest.approx(gg.fhelm, rel=1e-13)
    for m in gg.dw:
        assert np.max(np.abs(gf.dw[m] - gg.dw[m])) <= 1e-11 * np.max(np.abs(gg.dw[m]))
    np.testing.assert_allclose(gf.dcell, gg.dcell, rtol=1e-10, atol=1e-14)
    for k in gg.dchi:
        assert gf.dchi[k] == pytest.approx(gg.dchi[k], rel=1e-10, abs=1e-14)


def test_dlogQ_dtheta_matches_reference_stress_at_second_order():
    """scfstress (the reference's formula, exact for this centrosymmetric 1-D field) is
    the continuum limit of d log Q / d theta from the reverse sweep."""
    x = np.arange(32) / 32
    w = 0.8 * np.cos(2 * np.pi * x) + 0.3 * np.cos(4 * np.pi * x)
    gaps = []
    for Ns in (80, 160, 320):
        sys = diblock(P, (32, 1, 1), 1, "lamellar", (9.0,), Ns=Ns)
        set_fields(sys, {1: w.reshape(sys.shape), 2: -w.reshape(sys.shape)})
        sp = sys.species[0]
        _, dcell, _ = P.density_vjp(sys, {m: np.zeros(sys.shape) for m in sys.mids}, [1.0])
        dlogQ = dcell[0] / sp.Q
        stress = sys.scfstress()[0]            # = -phi dQ_theta, dQ_theta = dlogQ/Nref
        ref = -stress * sp.Nref
        gaps.append(abs(dlogQ - ref) / abs(dlogQ))
    r1, r2 = gaps[0] / gaps[1], gaps[1] / gaps[2]
    print(f"stress gap vs ds: {gaps}, ratios {r1:.3f} {r2:.3f}")
    assert 3.5 < r1 < 4.5 and 3.5 < r2 < 4.5
    assert gaps[-1] < 0.01


def test_gradient_through_reference_residual_at_scf_point():
    """At a converged SCF state (reference scft! loop, fixed cell), with eta the
    pressure field of the reference's residual (res = pot + eta - w = 0):
      dF_tb/dw + J^T eta / n = (rho_trap - rho) / n  (+ J^T res / n -> 0),
    where rho_trap is the trapezoid-rule density (the exact d Q / d w) and rho the
    reference's Simpson density."""
    sys = diblock(P, (64, 1, 1), 1, "lamellar", (16.0,), N=100, f=0.5, chi=0.2, Ns=100,
                  b=(1.0, 1.0))
    x = np.arange(64) / 64
    w = np.cos(2 * np.pi * x).reshape(sys.shape)
    P.fieldinit(sys, fields={1: w, 2: -w})
    upd = P.AndersonUpdater(nhist=10)
    res = P.scft(sys, upd, opts=P.SCFTOptions(nsteps=3000, ftol=1e-10))
    assert res.converged
    n = sys.ngrid
    sys.residuals_update()
    pot = {m: sys.potentials[m] for m in sys.mids}
    wf = {m: sys.fields[m] for m in sys.mids}
    rho = {m: sys.density[m] for m in sys.mids}
    eta = wf[1] - pot[1]
    g = P.fhelm_grad(sys, "textbook")
    h, _, _ = P.density_vjp(sys, {m: eta / n for m in sys.mids}, [0.0])
    sp = sys.species[0]
    t, _, _ = P.density_vjp(sys, {m: np.zeros(sys.shape) for m in sys.mids},
                            [-1.0 / (sp.Nref * sp.Q)])
    scale = max(float(np.max(np.abs(rho[m]))) for m in sys.mids) / n
    lhs = {m: g.dw[m] + h[m] for m in sys.mids}
    gap = {m: t[m] - rho[m] / n for m in sys.mids}
    err = max(float(np.max(np.abs(lhs[m] - gap[m]))) for m in sys.mids) / scale
    size = max(float(np.max(np.abs(gap[m]))) for m in sys.mids) / scale
    print(f"SCF point: |grad L - quadrature gap| / |rho/n| = {err:.2e}; gap itself {size:.2e}")
    assert err < 1e-7
    assert size > 1e3 * err          # the identity is not trivially zero


def test_torch_function_gradcheck():
    torch = pytest.importorskip("torch")
    sys = diblock(P, (8, 8, 1), 2, "square", (5.0,), Ns=12)
    F = P.FhelmFunction()
    rng = np.random.default_rng(1)
    w = torch.tensor(0.4 * rng.standard_normal((2,) + sys.shape), dtype=torch.float64,
                     requires_grad=True)
    cell = torch.tensor([5.0], dtype=torch.float64, requires_grad=True)
    chi = torch.tensor([0.2], dtype=torch.float64, requires_grad=True)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert torch.autograd.gradcheck(lambda a, b, c: F.apply(sys, "textbook", a, b, c),
                                        (w, cell, chi), eps=1e-6, atol=1e-8, rtol=1e-5)
