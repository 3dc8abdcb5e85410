// Reverse-mode derivative of the free energy with respect to the fields, the cell
// parameters, the interaction parameters, the block lengths and the volume fractions.
//
// F = ftrans(Q, phi, Nref) - <sum_a rho_a w_a> + E(rho)   (src/system/system.jl:402-433)
// with rho = density!(sys) (src/system/system.jl:266-289).  The reverse pass:
//   rbar_a = dF/drho_a = (dE/drho_a - w_a) / n              (k_rbar)
//   per species: total dF/dQ = Qbar - phi <rbar, rho_sp> / Q (the 1/Q of the Simpson
//   normalisation, src/species/multiblock.jl:219-221), then the reverse sweeps of the
//   chain (Engine::adjoint) or the closed form of a point species;
//   block lengths at a fixed contour grid: ds_b = v_b N_b / steps_b and Nref = sum v_b N_b.
#include <cmath>
#include <cstring>

#include "pf_host.h"

namespace pf {

static void ensure(double*& p, size_t count) {
    if (!p) p = (double*)dev_alloc(sizeof(double) * count);
}

void fhelm_grad(System& s, int functional, const double* rho_bar, const double* Q_bar,
                GradResult& g) {
    if (functional < -1 || functional > 1) throw Error(-2, "unknown functional");
    if (!s.canonical)
        throw Error(-2, "the free-energy derivative is defined for the canonical ensemble");
    if (s.mde != 0)
        throw Error(-2, "the free-energy derivative is defined for the RK2 propagator");
    if (!s.fields_set) throw Error(-2, "fields are not initialised");
    PF_HIP(hipSetDevice(s.device));
    const size_t n = s.n;
    const int nm = s.nmon(), nsp = (int)s.species.size();
    const int np = (int)s.cell.params.size();
    hipStream_t st = s.stream;
    ensure(s.g_rbar, n * nm);
    ensure(s.g_dw, n * nm);
    ensure(s.g_part, 48 * 512);
    ensure(s.g_out, 64);
    size_t blk_len = 0;
    for (const Species& sp : s.species)
        if (sp.kind == 0) blk_len += n * sp.nblk();
    if (blk_len > s.g_blk_len) {
        dev_free(s.g_blk);
        s.g_blk = (double*)dev_alloc(sizeof(double) * blk_len);
        s.g_blk_len = blk_len;
    }

    s.density();
    s.fetch_scalars();
    const double dn = (double)n;
    // the functional itself
    g.fhelm = 0.0;
    if (functional >= 0) {
        double v[5];
        s.sys_sums(v);
        double ftrans = 0.0;
        for (int i = 0; i < nsp; ++i) {
            const double p = s.phi[i];
            if (p <= 1e-8) continue;
            const double mu = functional == 0 ? p : std::log(p / s.Q[i]);
            ftrans += p * (mu - 1.0) / s.species[i].Nref;
        }
        g.fhelm = ftrans - v[2] / dn + v[3] / dn + v[4] / dn;
        launch_rbar(s.residual_args(), n, s.g_rbar, st);
    } else {
        PF_HIP(hipMemcpyAsync(s.g_rbar, rho_bar, sizeof(double) * n * nm,
                              hipMemcpyHostToDevice, st));
    }

    // per-block densities and the dots <rbar_mon(b), rho_b>, <rbar, rho_point>
    std::vector<double*> blk(nsp, nullptr);
    {
        size_t o = 0;
        for (int i = 0; i < nsp; ++i)
            if (s.species[i].kind == 0) {
                blk[i] = s.g_blk + o;
                o += n * s.species[i].nblk();
            }
    }
    s.eng->block_densities(s, blk);
    const int nparts = std::min(reduce_parts(n), 256);
    auto run_dots = [&](const std::vector<std::pair<const double*, const double*>>& xy) {
        std::vector<double> out;
        for (size_t a = 0; a < xy.size(); a += 48) {
            DotArgs da{};
            for (size_t b = a; b < std::min(xy.size(), a + 48); ++b) {
                da.x[da.k] = xy[b].first;
                da.y[da.k] = xy[b].second;
                ++da.k;
            }
            launch_dots(da, n, s.g_part, nparts, s.g_out, st);
            std::vector<double> h(da.k);
            PF_HIP(hipMemcpyAsync(h.data(), s.g_out, sizeof(double) * da.k,
                                  hipMemcpyDeviceToHost, st));
            PF_HIP(hipStreamSynchronize(st));
            out.insert(out.end(), h.begin(), h.end());
        }
        return out;
    };
    std::vector<std::pair<const double*, const double*>> xy;
    for (int i = 0; i < nsp; ++i) {
        const Species& sp = s.species[i];
        if (sp.kind == 0) {
            for (int b = 0; b < sp.nblk(); ++b)
                xy.push_back({s.g_rbar + (size_t)sp.block_slot[b] * n, blk[i] + (size_t)b * n});
        } else {
            xy.push_back({s.g_rbar + (size_t)sp.slots[0] * n, sp.rho_sp});
        }
    }
    const size_t nsp_dots = xy.size();
    if (functional >= 0) {
        for (int a = 0; a < nm; ++a)
            for (int b = a; b < nm; ++b)
                if ((a != b && s.pair[a * PF_MAXMON + b]) || (a == b && s.u_set[a]))
                    xy.push_back({s.rho + (size_t)a * n, s.rho + (size_t)b * n});
    }
    const std::vector<double> dots = run_dots(xy);

    // per-species scalars
    g.dphi.assign(nsp, 0.0);
    AdjointIO io;
    io.rbar = s.g_rbar;
    io.qtot.assign(nsp, 0.0);
    io.dw = s.g_dw;
    std::vector<std::vector<double>> D(nsp);
    std::vector<double> rr(nsp, 0.0), Nref_bar(nsp, 0.0);
    {
        size_t q = 0;
        for (int i = 0; i < nsp; ++i) {
            const Species& sp = s.species[i];
            const int nd = sp.kind == 0 ? sp.nblk() : 1;
            for (int b = 0; b < nd; ++b) {
                D[i].push_back(dots[q++]);
                rr[i] += D[i].back();
            }
        }
    }
    for (int i = 0; i < nsp; ++i) {
        const Species& sp = s.species[i];
        const double phi = s.phi[i], Q = s.Q[i];
        const bool active = phi > 1e-8;
        double Qbar = 0.0;
        g.dphi[i] = rr[i];
        Nref_bar[i] = -phi * rr[i] / sp.Nref;
        if (functional >= 0 && active) {
            const double mu = functional == 0 ? phi : std::log(phi / Q);
            if (functional == 1) Qbar = -phi / (sp.Nref * Q);
            g.dphi[i] += functional == 0 ? (2.0 * phi - 1.0) / sp.Nref : mu / sp.Nref;
            Nref_bar[i] += -phi * (mu - 1.0) / (sp.Nref * sp.Nref);
        } else if (functional < 0) {
            Qbar = Q_bar[i];
        }
        io.qtot[i] = Qbar - phi * rr[i] / Q;
    }

    // dF/dw: explicit -rho/n, point species, then the chain reverse sweeps
    launch_fill<double>(s.g_dw, 0.0, n * nm, st);
    if (functional >= 0) launch_axpy(s.g_dw, s.rho, -1.0 / dn, n * nm, st);
    for (int i = 0; i < nsp; ++i) {
        const Species& sp = s.species[i];
        if (sp.kind != 1) continue;
        const int slot = sp.slots[0];
        launch_point_grad(s.w + (size_t)slot * n, s.mon_vol[slot], n, s.g_rbar + (size_t)slot * n,
                          s.phi[i] / s.Q[i], io.qtot[i] / dn, s.g_dw + (size_t)slot * n, st);
    }
    s.eng->adjoint(s, io);
    g.dtheta = io.dtheta;
    if ((int)g.dtheta.size() != np) g.dtheta.assign(np, 0.0);

    // block lengths
    g.dN.clear();
    for (int i = 0; i < nsp; ++i) {
        const Species& sp = s.species[i];
        if (sp.kind != 0) continue;
        const double phi = s.phi[i];
        for (int b = 0; b < sp.nblk(); ++b) {
            const double ds = sp.ds_block[b];
            const double dds = io.dds_ld[i][b] + io.wdot[i][b] / ds + phi * D[i][b] / ds;
            const double dNref = dds * ds / sp.Nref_block[b] + Nref_bar[i];
            g.dN.push_back(sp.vol_block[b] * dNref);
        }
    }

    // interaction parameters
    g.dchi.assign((size_t)nm * nm, 0.0);
    g.du.assign(nm, 0.0);
    if (functional >= 0) {
        size_t q = nsp_dots;
        for (int a = 0; a < nm; ++a)
            for (int b = a; b < nm; ++b) {
                if (a != b && s.pair[a * PF_MAXMON + b]) {
                    const double v = dots[q++] / (s.mon_vol[a] * s.mon_vol[b] * dn);
                    g.dchi[a * nm + b] = v;
                    g.dchi[b * nm + a] = v;
                } else if (a == b && s.u_set[a]) {
                    g.du[a] = 0.5 * dots[q++] / (s.mon_vol[a] * s.mon_vol[a] * dn);
                }
            }
    }
    g.dw.resize(n * nm);
    PF_HIP(hipMemcpyAsync(g.dw.data(), s.g_dw, sizeof(double) * n * nm, hipMemcpyDeviceToHost,
                          st));
    PF_HIP(hipStreamSynchronize(st));
}

}  // namespace pf
