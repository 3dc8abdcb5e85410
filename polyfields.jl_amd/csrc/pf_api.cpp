// extern "C" boundary: argument checks, exception capture, handle management.
#include <cstring>

#include "../../include/polyfields_hip.h"
#include "pf_batch.h"
#include "pf_julia_random.h"

struct pf_system {
    pf::System* s;
};
struct pf_updater {
    pf::Updater* u;
};
struct pf_domain {
    pf::Domain* d;
};
struct pf_batch {
    pf::Batch* b;
    std::vector<int> all;
};

namespace {
thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const pf::Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    } catch (...) {
        g_err = "unknown error";
        return -2;
    }
}

void need(bool ok, const char* what) {
    if (!ok) throw pf::Error(-1, std::string("invalid argument: ") + what);
}

pf::System& S(pf_system* p) {
    need(p && p->s, "system handle");
    return *p->s;
}
}  // namespace

namespace {
int chain_blocks(const pf::System& s) {
    int nb = 0;
    for (const auto& sp : s.species)
        if (sp.kind == 0) nb += sp.nblk();
    return nb;
}

void copy_grad(const pf::System& s, const pf::GradResult& g, double* dw, double* dcell,
               int nparams, double* dN, int ndN) {
    if (dw) std::memcpy(dw, g.dw.data(), sizeof(double) * g.dw.size());
    if (dcell) {
        need(nparams == (int)s.cell.params.size(), "nparams");
        for (int k = 0; k < nparams; ++k) dcell[k] = g.dtheta[k];
    }
    if (dN) {
        need(ndN == (int)g.dN.size(), "ndN (total blocks of the chain species)");
        for (int k = 0; k < ndN; ++k) dN[k] = g.dN[k];
    }
}
}  // namespace

extern "C" {

const char* pf_last_error(void) { return g_err.c_str(); }

int pf_version(void) { return 1; }

int pf_device_count(int* count) {
    return guard([&] {
        need(count, "count");
        int c = 0;
        PF_HIP(hipGetDeviceCount(&c));
        *count = c;
    });
}

int pf_cell_geometry(int dim, const char* crystal, const double* params, int nparams,
                     double* R9, double* G9, double* volume, double* dGG) {
    return guard([&] {
        need(crystal && params && nparams > 0 && nparams <= PF_MAXPAR, "cell parameters");
        pf::Cell c;
        c.dim = dim;
        c.crystal = crystal;
        c.params.assign(params, params + nparams);
        c.update();
        if (R9) std::memcpy(R9, c.R, sizeof(c.R));
        if (G9) std::memcpy(G9, c.G, sizeof(c.G));
        if (volume) *volume = c.volume;
        if (dGG) std::memcpy(dGG, c.dGG.data(), sizeof(double) * c.dGG.size());
    });
}

int pf_chain_grid(int nblk, const double* Nref_block, int Ns, int* begin, double* ds_block,
                  double* Nref) {
    return guard([&] {
        need(nblk > 0 && Nref_block && Ns > 0 && begin && ds_block && Nref, "chain grid");
        std::vector<double> nb(Nref_block, Nref_block + nblk);
        std::vector<int> bg;
        std::vector<double> ds;
        pf::chain_grid(nb, Ns, Nref, bg, ds);
        for (int i = 0; i <= nblk; ++i) begin[i] = bg[i];
        for (int i = 0; i < nblk; ++i) ds_block[i] = ds[i];
    });
}

int pf_system_create(pf_system** out, int nx, int ny, int nz, int dim, const char* crystal,
                     const double* params, int nparams, int ensemble, double inv_zeta,
                     int precision, int device) {
    return guard([&] {
        need(out && crystal && params && nparams > 0 && nparams <= PF_MAXPAR, "system create");
        need(ensemble == PF_CANONICAL || ensemble == PF_GRAND, "ensemble");
        std::vector<double> p(params, params + nparams);
        auto* h = new pf_system{nullptr};
        try {
            h->s = new pf::System(nx, ny, nz, dim, crystal, p, precision,
                                  ensemble == PF_CANONICAL, inv_zeta, device);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int pf_system_destroy(pf_system* sys) {
    return guard([&] {
        if (!sys) return;
        delete sys->s;
        delete sys;
    });
}

int pf_system_add_monomer(pf_system* sys, int id, double vol, double charge) {
    return guard([&] { S(sys).add_monomer(id, vol, charge); });
}

int pf_system_add_multiblock(pf_system* sys, int nblk, const int* block_mids, const int* N_block,
                             const double* b_block, int Ns, double phi_or_mu, int* species_index,
                             int* Ns_used) {
    return guard([&] {
        need(nblk > 0 && block_mids && N_block && b_block, "multiblock arrays");
        auto& s = S(sys);
        int i = s.add_chain(std::vector<int>(block_mids, block_mids + nblk),
                            std::vector<int>(N_block, N_block + nblk),
                            std::vector<double>(b_block, b_block + nblk), Ns, phi_or_mu);
        if (species_index) *species_index = i;
        if (Ns_used) *Ns_used = s.species[i].Ns;
    });
}

int pf_system_add_point(pf_system* sys, int mid, double phi_or_mu, int* species_index) {
    return guard([&] {
        int i = S(sys).add_point(mid, phi_or_mu);
        if (species_index) *species_index = i;
    });
}

int pf_system_set_flory(pf_system* sys, int a, int b, double chi) {
    return guard([&] { S(sys).set_flory(a, b, chi); });
}

int pf_system_set_edwards(pf_system* sys, int mid, double u) {
    return guard([&] { S(sys).set_edwards(mid, u); });
}

int pf_system_set_field(pf_system* sys, int mid, const double* src) {
    return guard([&] {
        auto& s = S(sys);
        need(src, "field source");
        const int m = s.slot_of(mid);
        need(m >= 0, "monomer id");
        PF_HIP(hipMemcpyAsync(s.w + (size_t)m * s.n, src, sizeof(double) * s.n,
                              hipMemcpyHostToDevice, s.stream));
        s.sync();
        s.fields_set = true;
    });
}

int pf_system_fieldinit_seeded(pf_system* sys, uint64_t seed, double scale) {
    return guard([&] {
        auto& s = S(sys);
        pf::JuliaMT g(seed);
        std::vector<double> buf(s.n);
        for (int mid : pf::julia_dict_order(s.mon_order)) {
            pf::julia_randn_array(g, buf.data(), (int64_t)s.n);
            for (double& v : buf) v *= scale;
            PF_HIP(hipMemcpyAsync(s.w + (size_t)s.slot_of(mid) * s.n, buf.data(),
                                  sizeof(double) * s.n, hipMemcpyHostToDevice, s.stream));
            s.sync();
        }
        s.fields_set = true;
        s.residuals();
    });
}

int pf_julia_randn(uint64_t seed, const int64_t* counts, int narrays, double* out) {
    return guard([&] {
        need(counts && out && narrays >= 0, "arguments");
        pf::JuliaMT g(seed);
        size_t off = 0;
        for (int k = 0; k < narrays; ++k) {
            need(counts[k] >= 0, "array length");
            pf::julia_randn_array(g, out + off, counts[k]);
            off += (size_t)counts[k];
        }
    });
}

int pf_system_set_density(pf_system* sys, int mid, const double* src) {
    return guard([&] {
        auto& s = S(sys);
        need(src, "density source");
        const int m = s.slot_of(mid);
        need(m >= 0, "monomer id");
        PF_HIP(hipMemcpyAsync(s.rho + (size_t)m * s.n, src, sizeof(double) * s.n,
                              hipMemcpyHostToDevice, s.stream));
        s.sync();
    });
}

int pf_uniform_fields(pf_system* sys) {
    return guard([&] { S(sys).uniform_fields(); });
}

int pf_system_get_grid(pf_system* sys, int which, int mid, double* dst) {
    return guard([&] {
        auto& s = S(sys);
        need(dst, "destination");
        const int m = s.slot_of(mid);
        need(m >= 0, "monomer id");
        const double* base = which == PF_GRID_FIELD      ? s.w
                             : which == PF_GRID_DENSITY  ? s.rho
                             : which == PF_GRID_RESIDUAL ? s.res
                             : which == PF_GRID_POTENTIAL ? s.pot
                                                          : nullptr;
        need(base, "grid kind");
        PF_HIP(hipMemcpyAsync(dst, base + (size_t)m * s.n, sizeof(double) * s.n,
                              hipMemcpyDeviceToHost, s.stream));
        s.sync();
    });
}

int pf_system_get_species_density(pf_system* sys, int species, int mid, double* dst) {
    return guard([&] {
        auto& s = S(sys);
        need(dst && species >= 0 && species < (int)s.species.size(), "species");
        const auto& sp = s.species[species];
        const int m = s.slot_of(mid);
        auto it = std::find(sp.slots.begin(), sp.slots.end(), m);
        need(m >= 0 && it != sp.slots.end(), "monomer of species");
        PF_HIP(hipMemcpyAsync(dst, sp.rho_sp + (size_t)(it - sp.slots.begin()) * s.n,
                              sizeof(double) * s.n, hipMemcpyDeviceToHost, s.stream));
        s.sync();
    });
}

int pf_system_get_propagator(pf_system* sys, int species, int sl, int backward, double* dst) {
    return guard([&] {
        auto& s = S(sys);
        need(dst && species >= 0 && species < (int)s.species.size(), "species");
        s.eng->get_propagator(s, species, sl, backward, dst);
    });
}

int pf_system_species_scalars(pf_system* sys, int species, double* Q, double* phi, double* mu) {
    return guard([&] {
        auto& s = S(sys);
        need(species >= 0 && species < (int)s.species.size(), "species");
        s.fetch_scalars();
        if (Q) *Q = s.Q[species];
        if (phi) *phi = s.phi[species];
        if (mu) *mu = s.mu[species];
    });
}

int pf_system_cell(pf_system* sys, double* params, int nparams, int set) {
    return guard([&] {
        auto& s = S(sys);
        need(params && nparams == (int)s.cell.params.size(), "cell parameter count");
        if (set) {
            s.cell.params.assign(params, params + nparams);
            s.cell.update();
        } else {
            for (int k = 0; k < nparams; ++k) params[k] = s.cell.params[k];
        }
    });
}

int pf_system_add_cell(pf_system* sys, int dim, const char* crystal, const double* params,
                       int nparams) {
    return guard([&] {
        auto& s = S(sys);
        need(crystal && params && nparams > 0, "cell arguments");
        if (dim != s.dim)
            throw pf::Error(-2, "Number of non-singleton dimensions in System and Cell do not match!");
        const int np = pf::crystal_nparams(dim, crystal);
        if (np < 0)
            throw pf::Error(-2, std::string("Unrecognized crystal system `") + crystal +
                                    "` for dim = " + std::to_string(dim) + ".");
        need(nparams == np, "cell parameter count for the crystal system");
        pf::Cell c;
        c.dim = dim;
        c.crystal = crystal;
        c.params.assign(params, params + nparams);
        c.version = s.cell.version;   // update() moves it past every table built so far
        c.update();
        s.sync();
        s.cell = c;
        s.dirty_species = 1;
    });
}

int pf_system_synchronize(pf_system* sys) {
    return guard([&] { S(sys).sync(); });
}

int pf_system_set_mde(pf_system* sys, int method) {
    return guard([&] {
        need(method == 0 || method == 1, "mde method (0 RK2, 1 RQM4)");
        auto& s = S(sys);
        if (s.mde != method) {
            s.mde = method;
            s.dirty_species = 1;
        }
    });
}

int pf_system_set_precision(pf_system* sys, int precision) {
    return guard([&] { S(sys).set_precision(precision); });
}

int pf_validate(pf_system* sys) {
    return guard([&] { S(sys).validate(); });
}

int pf_density(pf_system* sys) {
    return guard([&] { S(sys).density(); });
}

int pf_residuals(pf_system* sys) {
    return guard([&] { S(sys).residuals(); });
}

int pf_field_error(pf_system* sys, double* v) {
    return guard([&] {
        need(v, "output");
        *v = S(sys).field_error();
    });
}

int pf_free_energy(pf_system* sys, double* v) {
    return guard([&] {
        need(v, "output");
        *v = S(sys).free_energy();
    });
}

int pf_free_energy_bulk(pf_system* sys, double* v) {
    return guard([&] {
        need(v, "output");
        *v = S(sys).free_energy_bulk();
    });
}

int pf_scfstress(pf_system* sys, double* stress, int nparams) {
    return guard([&] {
        auto& s = S(sys);
        need(stress && nparams == (int)s.cell.params.size(), "stress output");
        auto v = s.scfstress();
        for (int k = 0; k < nparams; ++k) stress[k] = v[k];
    });
}


int pf_scfstress_noalloc(pf_system* sys, double* stress, int nparams) {
    return guard([&] {
        auto& s = S(sys);
        need(stress && nparams == (int)s.cell.params.size(), "stress output");
        PF_HIP(hipSetDevice(s.device));
        size_t cap = 0;
        char* scr = pf::noalloc_scratch(s.device, &cap);
        auto v = s.scfstress_scratch(scr, cap);
        for (int k = 0; k < nparams; ++k) stress[k] = v[k];
    });
}

int pf_fhelm_grad(pf_system* sys, int functional, double* fhelm, double* dw, double* dcell,
                  int nparams, double* dchi, double* du, double* dphi, double* dN, int ndN) {
    return guard([&] {
        auto& s = S(sys);
        need(functional == 0 || functional == 1, "functional");
        if (dcell) need(nparams == (int)s.cell.params.size(), "nparams");
        if (dN) need(ndN == chain_blocks(s), "ndN (total blocks of the chain species)");
        pf::GradResult g;
        pf::fhelm_grad(s, functional, nullptr, nullptr, g);
        if (fhelm) *fhelm = g.fhelm;
        copy_grad(s, g, dw, dcell, nparams, dN, ndN);
        if (dchi) std::memcpy(dchi, g.dchi.data(), sizeof(double) * g.dchi.size());
        if (du) std::memcpy(du, g.du.data(), sizeof(double) * g.du.size());
        if (dphi) std::memcpy(dphi, g.dphi.data(), sizeof(double) * g.dphi.size());
    });
}

int pf_density_vjp(pf_system* sys, const double* rho_bar, const double* Q_bar, double* dw,
                   double* dcell, int nparams, double* dN, int ndN) {
    return guard([&] {
        auto& s = S(sys);
        need(rho_bar && Q_bar, "cotangents");
        if (dcell) need(nparams == (int)s.cell.params.size(), "nparams");
        if (dN) need(ndN == chain_blocks(s), "ndN (total blocks of the chain species)");
        pf::GradResult g;
        pf::fhelm_grad(s, -1, rho_bar, Q_bar, g);
        copy_grad(s, g, dw, dcell, nparams, dN, ndN);
    });
}

int pf_system_set_checkpoint(pf_system* sys, int K) {
    return guard([&] {
        auto& s = S(sys);
        need(K >= -1, "segment length");
        s.sync();
        s.ckpt = K;
        s.dirty_species = 1;
    });
}

int pf_species_set_blocks(pf_system* sys, int isp, const double* N_block, int nblk) {
    return guard([&] {
        need(N_block && nblk > 0, "block lengths");
        S(sys).set_block_lengths(isp, std::vector<double>(N_block, N_block + nblk));
    });
}

int pf_updater_anderson(pf_updater** out, int nhist, int iter_euler, double lam_euler) {
    return guard([&] {
        need(out, "output");
        *out = new pf_updater{new pf::Anderson(nhist, iter_euler, lam_euler)};
    });
}

int pf_updater_euler(pf_updater** out, int method, double lam) {
    return guard([&] {
        need(out && (method == 0 || method == 1), "Euler method (0 PEC, 1 PECE)");
        *out = new pf_updater{new pf::Euler(method, lam)};
    });
}

int pf_updater_momentum(pf_updater** out, int method, double lam, double gamma, double beta1,
                        double beta2) {
    return guard([&] {
        need(out && method >= 0 && method <= 3, "Momentum method (0 MOM, 1 RMS, 2 ADAM, 3 NAG)");
        need(gamma > 0 && gamma < 1 && beta1 > 0 && beta1 < 1 && beta2 > 0 && beta2 < 1,
             "momentum coefficients in (0, 1)");
        *out = new pf_updater{new pf::Momentum(method, lam, gamma, beta1, beta2)};
    });
}

int pf_updater_setup(pf_updater* u, pf_system* sys) {
    return guard([&] {
        need(u && u->u, "updater");
        u->u->setup(&S(sys));
    });
}

int pf_updater_step(pf_updater* u) {
    return guard([&] {
        need(u && u->u && u->u->sys, "updater bound to a system");
        u->u->step();
    });
}

int pf_updater_destroy(pf_updater* u) {
    return guard([&] {
        if (!u) return;
        delete u->u;
        delete u;
    });
}

int pf_domain_create(pf_domain** out, int nper, int nskip, double lam, double pmax, double rlxn) {
    return guard([&] {
        need(out && nper >= 0 && nskip > 0, "domain parameters");
        *out = new pf_domain{new pf::Domain(nper, nskip, lam, pmax, rlxn)};
    });
}

int pf_domain_set_fd_step(pf_domain* d, double dx) {
    return guard([&] {
        need(d && d->d && dx > 0.0, "finite-difference step");
        d->d->fd_step = dx;
    });
}

int pf_domain_setup(pf_domain* d, pf_system* sys) {
    return guard([&] {
        need(d && d->d, "domain");
        d->d->setup(&S(sys));
    });
}

int pf_domain_step(pf_domain* d, int nsteps, double tol) {
    return guard([&] {
        need(d && d->d && d->d->sys, "domain bound to a system");
        d->d->step(nsteps, tol);
    });
}

int pf_domain_stress(pf_domain* d, int refresh, double* stress, int nparams, double* serr) {
    return guard([&] {
        need(d && d->d && d->d->sys, "domain bound to a system");
        if (refresh) d->d->scfstress_update();
        if (stress) {
            need(nparams == (int)d->d->stress.size(), "stress length");
            for (int k = 0; k < nparams; ++k) stress[k] = d->d->stress[k];
        }
        if (serr) *serr = d->d->stress_error();
    });
}

int pf_domain_destroy(pf_domain* d) {
    return guard([&] {
        if (!d) return;
        delete d->d;
        delete d;
    });
}

int pf_scft(pf_system* sys, pf_updater* u, pf_domain* d, const pf_scft_options* opts,
            pf_scft_results* res) {
    return pf_scft_save(sys, u, d, opts, res, nullptr, nullptr);
}

int pf_scft_save(pf_system* sys, pf_updater* u, pf_domain* d, const pf_scft_options* opts,
                 pf_scft_results* res, pf_save_fn save, void* user) {
    return guard([&] {
        need(u && u->u && opts && res, "scft arguments");
        pf::ScftOptions o{opts->nsteps, opts->nsout, opts->nsave, opts->ftol, opts->stol};
        std::function<void(int)> cb;
        if (save) cb = [save, user](int step) { save(user, step); };
        pf::ScftResults r = pf::scft(S(sys), *u->u, d ? d->d : nullptr, o, cb);
        res->time = r.time;
        res->nsteps = r.nsteps;
        res->converged = r.converged;
        res->ferr = r.ferr;
        res->serr = r.serr;
        res->fhelm = r.fhelm;
    });
}

int pf_batch_create(pf_batch** out, pf_system* const* systems, int nsys) {
    return guard([&] {
        need(out && systems && nsys > 0, "batch systems");
        std::vector<pf::System*> v;
        for (int i = 0; i < nsys; ++i) v.push_back(&S(systems[i]));
        auto* h = new pf_batch{nullptr, {}};
        try {
            h->b = new pf::Batch(v);
        } catch (...) {
            delete h;
            throw;
        }
        for (int i = 0; i < nsys; ++i) h->all.push_back(i);
        *out = h;
    });
}

int pf_batch_destroy(pf_batch* b) {
    return guard([&] {
        if (!b) return;
        delete b->b;
        delete b;
    });
}

static pf::Batch& B(pf_batch* b) {
    need(b && b->b, "batch handle");
    return *b->b;
}

int pf_batch_density(pf_batch* b) {
    return guard([&] {
        for (auto* s : B(b).sys) s->sync();
        B(b).density(b->all);
        PF_HIP(hipStreamSynchronize(B(b).stream));
    });
}

int pf_system_autotune(pf_system* sys, int reps, int* sched, double* best_ms) {
    return guard([&] {
        need(reps > 0, "repetitions");
        const pf::Schedule s = pf::autotune_system(S(sys), reps, best_ms);
        if (sched) {
            sched[0] = s.slab;
            sched[1] = s.pencil;
            sched[2] = s.planar;
            sched[3] = s.split;
        }
    });
}

int pf_schedule_set(int elem_bytes, int nx, int ny, int nz, const int* sched) {
    return guard([&] {
        need(sched && (elem_bytes == 4 || elem_bytes == 8), "schedule");
        pf::Schedule& s = pf::schedule(elem_bytes, nx, ny, nz);
        s.slab = sched[0];
        s.pencil = sched[1];
        s.planar = sched[2];
        s.split = sched[3];
    });
}

int pf_schedule_reset(void) {
    return guard([&] { pf::schedule_reset(); });
}

int pf_batch_autotune(pf_batch* b, int reps, int* sched, int cap, int* count) {
    return guard([&] {
        need(reps > 0, "repetitions");
        std::vector<std::pair<std::array<int, 4>, pf::Schedule>> out;
        pf::autotune_batch(B(b), reps, out);
        if (count) *count = (int)out.size();
        for (int k = 0; sched && k < (int)out.size() && k < cap; ++k) {
            int* r = sched + 8 * k;
            for (int j = 0; j < 4; ++j) r[j] = out[k].first[j];
            r[4] = out[k].second.slab;
            r[5] = out[k].second.pencil;
            r[6] = out[k].second.planar;
            r[7] = out[k].second.split;
        }
    });
}

int pf_batch_set_graph(pf_batch* b, int on) {
    return guard([&] {
        PF_HIP(hipStreamSynchronize(B(b).stream));
        B(b).use_graph = on != 0;
        B(b).drop_graphs();
    });
}

int pf_batch_set_precision(pf_batch* b, int precision) {
    return guard([&] { B(b).set_precision(precision); });
}

int pf_batch_sweeps(pf_batch* b) {
    return guard([&] {
        for (auto* s : B(b).sys) s->sync();
        B(b).sweeps(b->all);
        PF_HIP(hipStreamSynchronize(B(b).stream));
    });
}

int pf_batch_residuals(pf_batch* b) {
    return guard([&] {
        for (auto* s : B(b).sys) s->sync();
        B(b).residuals(b->all);
        PF_HIP(hipStreamSynchronize(B(b).stream));
    });
}

static void copy_batch_results(const std::vector<pf::BatchResult>& r, pf_scft_results* out) {
    for (size_t i = 0; i < r.size(); ++i) {
        out[i].time = r[i].time;
        out[i].nsteps = r[i].nsteps;
        out[i].converged = r[i].converged;
        out[i].ferr = r[i].ferr;
        out[i].serr = r[i].serr;
        out[i].fhelm = r[i].fhelm;
    }
}

int pf_batch_scft(pf_batch* b, int nhist, int iter_euler, double lam_euler, int nsteps,
                  double ftol, const int* mask, pf_scft_results* results) {
    return guard([&] {
        need(results, "results");
        need(nhist >= 1, "nhist must be at least 1");
        auto r = B(b).scft_anderson(nhist, iter_euler, lam_euler, nsteps, ftol, mask);
        copy_batch_results(r, results);
    });
}

int pf_batch_scft_momentum(pf_batch* b, int method, double lam, double gamma, double beta1,
                           double beta2, const pf_domain* domain, int nsteps, double ftol,
                           double stol, const int* mask, pf_scft_results* results) {
    return guard([&] {
        need(results, "results");
        need(!domain || domain->d, "domain");
        auto r = B(b).scft_momentum(method, lam, gamma, beta1, beta2, nsteps, ftol, mask,
                                    domain ? domain->d : nullptr, stol);
        copy_batch_results(r, results);
    });
}

int pf_batch_scft_euler(pf_batch* b, int method, double lam, const pf_domain* domain,
                        int nsteps, double ftol, double stol, const int* mask,
                        pf_scft_results* results) {
    return guard([&] {
        need(results, "results");
        need(!domain || domain->d, "domain");
        auto r = B(b).scft_euler(method, lam, nsteps, ftol, mask, domain ? domain->d : nullptr,
                                 stol);
        copy_batch_results(r, results);
    });
}

int pf_batch_scft_domain(pf_batch* b, int nhist, int iter_euler, double lam_euler,
                         const pf_domain* domain, int nsteps, double ftol, double stol,
                         const int* mask, pf_scft_results* results) {
    return guard([&] {
        need(results, "results");
        need(domain && domain->d, "domain");
        need(nhist >= 1, "nhist must be at least 1");
        auto r = B(b).scft_anderson(nhist, iter_euler, lam_euler, nsteps, ftol, mask,
                                    domain->d, stol);
        copy_batch_results(r, results);
    });
}

int pf_batch_timing(pf_batch* b, int on) {
    return guard([&] { B(b).eng->timing(on); });
}

int pf_batch_timing_read(pf_batch* b, int kind, double* total_ms, int* launches) {
    return guard([&] {
        need(total_ms, "output");
        *total_ms = B(b).eng->timing_ms(kind, launches);
    });
}

int pf_batch_timing_lanes(pf_batch* b, int kind, double* lanes) {
    return guard([&] {
        need(lanes, "output");
        *lanes = B(b).eng->timing_lanes(kind);
    });
}

int pf_system_timing(pf_system* sys, int on) {
    return guard([&] { S(sys).eng->timing(on); });
}

int pf_system_timing_read(pf_system* sys, int kind, double* total_ms, int* launches) {
    return guard([&] {
        need(total_ms, "output");
        *total_ms = S(sys).eng->timing_ms(kind, launches);
    });
}

}  // extern "C"
