// Batched SCFT: many independent FieldSystems advance in lockstep through the
// scft!(sys, AndersonUpdater) loop of src/simulate.jl:60-148 (fixed cell).  Each
// iteration is one set of launches for all active systems; the host solves each
// system's small Anderson least-squares problem (src/updaters/anderson.jl:184-189) and
// drops systems from the active set as they converge.
#include <chrono>
#include <cmath>
#include <cstring>
#include <algorithm>

#include "pf_batch.h"

namespace pf {

HostPool::HostPool(int nworkers) {
    for (int i = 0; i < nworkers; ++i) th.emplace_back([this, i] { worker(i); });
}

HostPool::~HostPool() {
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    go.notify_all();
    for (auto& t : th) t.join();
}

void HostPool::run(size_t n, const std::function<void(size_t, size_t)>& fn) {
    const size_t parts = th.size() + 1, per = (n + parts - 1) / parts;
    {
        std::lock_guard<std::mutex> lk(mu);
        job = &fn;
        count = n;
        pending = (int)th.size();
        ++gen;
    }
    go.notify_all();
    fn(0, std::min(n, per));
    std::unique_lock<std::mutex> lk(mu);
    done.wait(lk, [this] { return pending == 0; });
    job = nullptr;
}

void HostPool::worker(int id) {
    long seen = 0;
    while (true) {
        const std::function<void(size_t, size_t)>* fn;
        size_t n;
        {
            std::unique_lock<std::mutex> lk(mu);
            go.wait(lk, [&] { return stop || gen != seen; });
            if (stop) return;
            seen = gen;
            fn = job;
            n = count;
        }
        const size_t parts = th.size() + 1, per = (n + parts - 1) / parts;
        const size_t b = std::min(n, (size_t)(id + 1) * per), e = std::min(n, (size_t)(id + 2) * per);
        if (b < e) (*fn)(b, e);
        {
            std::lock_guard<std::mutex> lk(mu);
            --pending;
        }
        done.notify_one();
    }
}

Batch::Batch(const std::vector<System*>& s) : sys(s) {
    if (sys.empty()) throw Error(-1, "empty batch");
    precision = sys[0]->precision;
    for (System* x : sys) {
        if (x->precision != precision) throw Error(-2, "all systems of a batch need one precision");
        if (x->device != sys[0]->device) throw Error(-2, "all systems of a batch need one device");
        if (x->nmon() > PF_MAXMON) throw Error(-2, "too many monomers");
    }
    PF_HIP(hipSetDevice(sys[0]->device));
    PF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    PF_HIP(hipEventCreateWithFlags(&sums_ev, hipEventDisableTiming));
    const size_t ns = sys.size();
    sums = (double*)dev_alloc(sizeof(double) * 256 * ns);
    csums = (double*)dev_alloc(sizeof(double) * kCompact * ns);
    PF_HIP(hipHostMalloc((void**)&hcsums, sizeof(double) * kCompact * ns, hipHostMallocDefault));
    alloc_step_table(PF_MAXHIST);
    eng = make_batch_engine(*this);
    if (ns >= 64) {
        const unsigned hw = std::thread::hardware_concurrency();
        const int workers = (int)std::min(7u, hw > 1 ? hw - 1 : 0u);
        if (workers > 0) pool = std::make_unique<HostPool>(workers);
    }
}

Batch::~Batch() {
    (void)hipSetDevice(sys[0]->device);
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    eng.reset();
    free_momentum();
    dev_free(sums);
    dev_free(csums);
    dev_free(gpart);
    dev_free(d_step);
    dev_free(d_act);
    if (hcsums) (void)hipHostFree(hcsums);
    if (h_step) (void)hipHostFree(h_step);
    if (sums_ev) (void)hipEventDestroy(sums_ev);
    if (stream) (void)hipStreamDestroy(stream);
}

// Step table for histories of depth nhist (kept if the current one is large enough),
// and for depths above PF_MAXHIST the tiled-Gram scratch and a compact area of
// gram_tiles(nhist) * 64 values per system.  Called with the stream idle.
void Batch::alloc_step_table(int nhist) {
    const size_t ns = sys.size();
    if (!d_step || nhist > coef_stride) {
        dev_free(d_step);
        if (h_step) PF_HIP(hipHostFree(h_step));
        d_step = nullptr;
        h_step = nullptr;
        coef_stride = std::max(nhist, PF_MAXHIST);
        const size_t bytes = (sizeof(StepTask) + sizeof(double) * coef_stride) * ns;
        d_step = (StepTask*)dev_alloc(bytes);
        PF_HIP(hipHostMalloc((void**)&h_step, bytes, hipHostMallocDefault));
    }
    if (nhist > PF_MAXHIST) {
        const int nv = gram_tiles(nhist) * 64;
        if (nv > ncompact) {
            dev_free(csums);
            if (hcsums) PF_HIP(hipHostFree(hcsums));
            csums = nullptr;
            hcsums = nullptr;
            ncompact = nv;
            csums = (double*)dev_alloc(sizeof(double) * (size_t)ncompact * ns);
            PF_HIP(hipHostMalloc((void**)&hcsums, sizeof(double) * (size_t)ncompact * ns,
                                 hipHostMallocDefault));
        }
        size_t Nmax = 0;
        for (System* x : sys) Nmax = std::max(Nmax, x->n * x->nmon());
        const int parts = gram_tiled_parts(Nmax);
        if (parts != gparts || !gpart) {
            dev_free(gpart);
            gpart = nullptr;
            gparts = parts;
            gpart = (double*)dev_alloc(sizeof(double) * (size_t)nv * gparts * ns);
        }
    }
}

void Batch::drop_graphs() {
    for (auto& g : graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = Graph{};
    }
}

void Batch::set_precision(int p) {
    if (p == precision) return;
    PF_HIP(hipStreamSynchronize(stream));
    drop_graphs();
    eng.reset();
    for (System* s : sys) s->set_precision(p);
    precision = p;
    eng = make_batch_engine(*this);
}

void Batch::bind_active(const std::vector<int>& act) {
    if (act == act_cache && d_act) return;
    std::vector<SysTask> t;
    for (int i : act) {
        System& s = *sys[i];
        SysTask k{};
        k.ra = s.residual_args();
        k.pool.nmon = s.nmon();
        k.pool.dst = s.rho;
        for (size_t j = 0; j < s.species.size(); ++j) {
            const Species& sp = s.species[j];
            for (size_t q = 0; q < sp.slots.size(); ++q) {
                const int m = sp.slots[q];
                if (k.pool.nterm[m] >= 8) throw Error(-2, "too many species share one monomer type");
                k.pool.src[m][k.pool.nterm[m]] = sp.rho_sp + q * s.n;
                k.pool.isp[m][k.pool.nterm[m]] = (int)j;
                k.pool.nterm[m]++;
            }
        }
        k.n = s.n;
        k.nparts = reduce_parts(s.n);
        k.gparts = std::min(reduce_parts(s.n * s.nmon()), (int)(s.partial_len / (PF_MAXHIST * (PF_MAXHIST + 3) / 2)));
        k.nsp = (int)s.species.size();
        k.sc = s.sc;
        k.partial = s.partial;
        k.gpartial = s.partial;
        k.sums = sums + (size_t)i * 256;
        t.push_back(k);
    }
    // work queued earlier may still read the old descriptors
    PF_HIP(hipStreamSynchronize(stream));
    dev_free(d_act);
    d_act = (SysTask*)dev_alloc(sizeof(SysTask) * t.size());
    PF_HIP(hipMemcpyAsync(d_act, t.data(), sizeof(SysTask) * t.size(), hipMemcpyHostToDevice, stream));
    PF_HIP(hipStreamSynchronize(stream));
    tasks = t;
    act_cache = act;
}

static int max_parts(const std::vector<SysTask>& t, bool gram) {
    int m = 1;
    for (auto& k : t) m = std::max(m, gram ? k.gparts : k.nparts);
    return m;
}

// density! (kind 0) or residuals! (kind 1) of the active systems.  The launch sequence
// (exp tables, every lane group on its own stream, Q, Simpson, pooling, residuals) is
// captured once into a graph and replayed while the active set and the lane tables are
// unchanged; the graph holds kernel nodes only (the tables it reads are uploaded before
// capture).  Timing runs eagerly, so that events can bracket single launches.
void Batch::run(int kind, const std::vector<int>& act) {
    bind_active(act);
    eng->prepare(act);
    size_t nmax = 0;
    for (int i : act) nmax = std::max(nmax, sys[i]->n);
    auto seq = [&] {
        if (kind == 2) {
            eng->launch_sweeps(act);
            return;
        }
        eng->launch(act);
        launch_muphi_pool_b(d_act, (int)act.size(), nmax, stream);
        if (kind == 1)
            launch_residuals_b(d_act, (int)act.size(), max_parts(tasks, false), nmax, stream);
    };
    if (!use_graph || eng->timing_on() || !eng->capturable()) {
        seq();
        return;
    }
    Graph& g = graphs[kind];
    if (!g.exec || g.act != act || g.gen != eng->generation() || g.d_act != d_act) {
        if (g.exec) PF_HIP(hipGraphExecDestroy(g.exec));
        g = Graph{};
        hipGraph_t graph = nullptr;
        PF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        try {
            seq();
        } catch (...) {
            (void)hipStreamEndCapture(stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        PF_HIP(hipStreamEndCapture(stream, &graph));
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        PF_HIP(e);
        g.act = act;
        g.gen = eng->generation();
        g.d_act = d_act;
    }
    PF_HIP(hipGraphLaunch(g.exec, stream));
}

void Batch::density(const std::vector<int>& act) { run(0, act); }

void Batch::residuals(const std::vector<int>& act) { run(1, act); }

void Batch::sweeps(const std::vector<int>& act) { run(2, act); }

void Batch::sums_launch(const std::vector<int>& act) {
    bind_active(act);
    launch_sys_reduce_b(d_act, (int)act.size(), max_parts(tasks, false), csums, stream);
    PF_HIP(hipMemcpyAsync(hcsums, csums, sizeof(double) * 5 * act.size(), hipMemcpyDeviceToHost, stream));
    PF_HIP(hipEventRecord(sums_ev, stream));
}

void Batch::sums_read(const std::vector<int>& act, std::vector<double>& out) {
    PF_HIP(hipEventSynchronize(sums_ev));
    out.assign(sys.size() * 5, 0.0);
    for (size_t j = 0; j < act.size(); ++j)
        for (int k = 0; k < 5; ++k) out[(size_t)act[j] * 5 + k] = hcsums[j * 5 + k];
}

void Batch::sys_sums(const std::vector<int>& act, std::vector<double>& out) {
    sums_launch(act);
    sums_read(act, out);
}

static double ferr_of(const double* v) {
    return std::sqrt(v[0]) / std::sqrt(2.220446049250313e-16 + v[1]);
}

// free_energy (src/system/system.jl:402-433) from the batch's five system sums; in the
// grand ensemble the species fractions come from the device (muphi! of the last pass)
static double fhelm_of(System& s, const double* v) {
    if (!s.canonical) s.fetch_scalars();
    double ftrans = 0.0;
    for (size_t i = 0; i < s.species.size(); ++i)
        if (s.phi[i] > 1e-8) ftrans += s.phi[i] * (s.phi[i] - 1.0) / s.species[i].Nref;
    return ftrans + (-v[2] / (double)s.n) + (v[3] / (double)s.n + v[4] / (double)s.n);
}

void Batch::stress_all(const std::vector<int>& set, std::vector<std::vector<double>>& out) {
    PF_HIP(hipStreamSynchronize(stream));   // the batch launches wrote the stacks
    for (int i : set) sys[i]->stress_launch();
    for (int i : set) out[i] = sys[i]->stress_collect();
}

void Batch::domain_round(const std::vector<int>& due, double stol, std::vector<double>& serr,
                         std::vector<BatchResult>& res) {
    std::vector<std::vector<double>> st(sys.size());
    std::vector<int> need;
    {
        std::vector<int> live;
        for (int i : due)
            if (!res[i].diverged) live.push_back(i);
        stress_all(live, st);
        for (int i : live) {
            Domain& D = *dom_state[i];
            D.stress = st[i];
            serr[i] = D.stress_error();
            if (serr[i] > stol) need.push_back(i);
        }
    }
    if (need.empty()) return;
    // step! (src/updaters/domain.jl:50-92) of every system in `need`: its first
    // scfstress! reads the same stacks as the stress above, so that value is kept.  The
    // cells change at every evaluation, so the launches run eagerly (no graph capture).
    const bool graph_was = use_graph;
    use_graph = false;
    const int nper = dom_state[need[0]]->nper;
    int maxnp = 0;
    for (int i : need) maxnp = std::max(maxnp, (int)sys[i]->cell.params.size());
    for (int it = 0; it < nper && !need.empty(); ++it) {
        // dstress! (domain.jl:94-120): forward differences, one cell parameter at a time,
        // the perturbed cells of all systems evaluated together
        for (int i : need) dom_state[i]->oldparams = sys[i]->cell.params;
        for (int k = 0; k < maxnp; ++k) {
            std::vector<int> sub;
            for (int i : need)
                if ((int)sys[i]->cell.params.size() > k) sub.push_back(i);
            for (int i : sub) {
                Cell& c = sys[i]->cell;
                c.params = dom_state[i]->oldparams;
                c.params[k] += dom_state[i]->fd_step;
                c.update();
            }
            density(sub);
            stress_all(sub, st);
            for (int i : sub) {
                Domain& D = *dom_state[i];
                const int np = (int)sys[i]->cell.params.size();
                const double x = 1.0 / D.fd_step;
                for (int r = 0; r < np; ++r) D.dstress[r * np + k] = x * (st[i][r] - D.stress[r]);
            }
        }
        // the parameters are restored without update!, as the reference does
        for (int i : need) sys[i]->cell.params = dom_state[i]->oldparams;
        std::vector<int> ok;
        for (int i : need) {
            try {
                dom_state[i]->newton_update();
                ok.push_back(i);
            } catch (const Error&) {
                // step! raises on a singular Jacobian; in a batch the system is reported
                // as failed and dropped
                res[i].diverged = 1;
            }
        }
        need = ok;
        if (need.empty()) break;
        residuals(need);
        stress_all(need, st);
        for (int i : need) dom_state[i]->stress = st[i];
    }
    use_graph = graph_was;
    for (int i : need) serr[i] = dom_state[i]->stress_error();
}

namespace {

// AndersonUpdater over the batch (src/updaters/anderson.jl:98-225): the Gram matrices on
// the device, the least-squares solves on the host, the mix on the device
struct AndersonStepper : BatchStepper {
    int nhist, iter_euler;
    double lam_euler;
    bool deep = false;
    AndersonStepper(int nh, int ie, double le) : nhist(nh), iter_euler(ie), lam_euler(le) {}
    const char* noncanonical_error() const override {
        return "Anderson mixing can only be used for systems in the canonical ensemble.";
    }
    void setup(Batch& b, const std::vector<int>& run) override {
        b.alloc_step_table(nhist);
        deep = nhist > PF_MAXHIST;
        b.and_state.clear();
        for (size_t i = 0; i < b.sys.size(); ++i) {
            b.and_state.emplace_back(new Anderson(nhist, iter_euler, lam_euler));
            if (std::find(run.begin(), run.end(), (int)i) == run.end()) continue;
            Anderson& A = *b.and_state.back();
            A.sys = b.sys[i];
            const size_t N = b.sys[i]->n * b.sys[i]->nmon();
            A.hist_r = (double*)dev_alloc(sizeof(double) * N * (nhist + 1));
            A.hist_w = (double*)dev_alloc(sizeof(double) * N * (nhist + 1));
            PF_HIP(hipMemcpyAsync(A.wslot(0), b.sys[i]->w, sizeof(double) * N,
                                  hipMemcpyDeviceToDevice, b.stream));
            A.is_setup = true;
        }
    }
    void update(Batch& b, const std::vector<int>& act, std::vector<BatchResult>& res) override;
};

// MomentumUpdater (MOM, RMS, ADAM) over the batch (src/updaters/momentum.jl:47-139): one
// kernel updates the moments and the fields of every active system
struct MomentumStepper : BatchStepper {
    int method;
    double lam, gamma, b1, b2;
    MomentumStepper(int m, double l, double g, double bb1, double bb2)
        : method(m), lam(l), gamma(g), b1(bb1), b2(bb2) {}
    bool allows_grand() const override { return true; }
    // NAG evaluates residuals! at its look-ahead fields inside the step
    bool starts_with_residuals() const override { return method != 3; }
    void setup(Batch& b, const std::vector<int>& run) override {
        b.free_momentum();
        b.mom_state.assign(b.sys.size(), Batch::MomState{});
        for (int i : run) {
            System& s = *b.sys[i];
            const size_t N = s.n * s.nmon();
            Batch::MomState& m = b.mom_state[i];
            for (double** p : {&m.mt, &m.vt, &m.step})
                *p = (double*)dev_alloc(sizeof(double) * N);
            if (method == 3) m.tmp = (double*)dev_alloc(sizeof(double) * N);
            // setup! (momentum.jl:47-61): step = the first residual, zero moments
            PF_HIP(hipMemcpyAsync(m.step, s.res, sizeof(double) * N, hipMemcpyDeviceToDevice,
                                  b.stream));
            launch_fill<double>(m.mt, 0.0, N, b.stream);
            launch_fill<double>(m.vt, 0.0, N, b.stream);
        }
    }
    void update(Batch& b, const std::vector<int>& act, std::vector<BatchResult>&) override {
        size_t Nmax = 0;
        for (int i : act) Nmax = std::max(Nmax, b.sys[i]->n * b.sys[i]->nmon());
        b.bind_momentum(act);
        if (method == 3) {
            // NAG (momentum.jl:141-165): residuals! at w + lam m, the momentum update with
            // gamma, then the step from the saved fields
            launch_stage_b(b.d_mom, (int)act.size(), STAGE_NAG_LOOK, lam, 0.0, 0.0, Nmax,
                           b.stream);
            b.residuals(act);
            launch_stage_b(b.d_mom, (int)act.size(), STAGE_NAG_STEP, lam, gamma, b2, Nmax,
                           b.stream);
        } else {
            launch_momentum_b(b.d_mom, (int)act.size(), method, lam, b1, b2, Nmax, b.stream);
        }
    }
};

// EulerUpdater (src/updaters/euler.jl:36-66) over the batch: predictor, residuals! at the
// predicted fields, corrector, and for PECE residuals! at the corrected fields.  The step
// moves along the residual that the previous step left (its own last residuals! pass).
struct EulerStepper : BatchStepper {
    int method;   // 0 PEC, 1 PECE
    double lam;
    EulerStepper(int m, double l) : method(m), lam(l) {}
    bool allows_grand() const override { return true; }
    bool starts_with_residuals() const override { return false; }
    void setup(Batch& b, const std::vector<int>& run) override {
        b.free_momentum();
        b.mom_state.assign(b.sys.size(), Batch::MomState{});
        for (int i : run) {
            System& s = *b.sys[i];
            b.mom_state[i].tmp = (double*)dev_alloc(sizeof(double) * s.n * s.nmon());
        }
    }
    void update(Batch& b, const std::vector<int>& act, std::vector<BatchResult>&) override {
        size_t Nmax = 0;
        for (int i : act) Nmax = std::max(Nmax, b.sys[i]->n * b.sys[i]->nmon());
        const double half = 0.5 * lam;
        b.bind_momentum(act);
        launch_stage_b(b.d_mom, (int)act.size(), STAGE_EULER_PRED, lam, half, 0.0, Nmax,
                       b.stream);
        b.residuals(act);
        launch_stage_b(b.d_mom, (int)act.size(), STAGE_EULER_CORR, lam, half, 0.0, Nmax,
                       b.stream);
        if (method == 1) b.residuals(act);
    }
};

void AndersonStepper::update(Batch& b, const std::vector<int>& act, std::vector<BatchResult>& res) {
    // bookkeeping shared by the lockstep systems
    Anderson& A0 = *b.and_state[act[0]];
    const int iter = A0.iter + 1;
    int ih;
    double lam;
    if (iter < nhist + 1) {
        ih = iter - 1;
        lam = 1.0 - std::pow(0.9, (double)iter);
    } else {
        ih = nhist;
        lam = 1.0;
    }
    for (int i : act) {
        b.and_state[i]->iter = iter;
        b.and_state[i]->iter_hist = ih;
        b.and_state[i]->lam = lam;
    }
    const bool euler = iter <= b.and_state[act[0]]->iter_euler;
    for (size_t k = 0; k < act.size(); ++k) {
        Anderson& A = *b.and_state[act[k]];
        StepTask& s = b.h_step[k];
        std::memset(&s, 0, sizeof(StepTask));
        s.h = euler ? 0 : ih;
        s.euler = euler;
        s.lam = lam;
        s.lam_euler = A.lam_euler;
        s.r_out = A.rslot(ih);
        if (s.h > 0) s.ring = A.ring(ih);
        s.C = b.d_coef() + k * (size_t)b.coef_stride;
        s.f_iter = A.wslot(ih);
        s.f_out = (iter < nhist + 1) ? A.wslot(iter) : A.wslot(0);
    }
    // the StepTasks; the coefficients follow after the host solves below
    PF_HIP(hipMemcpyAsync(b.d_step, b.h_step, sizeof(StepTask) * act.size(), hipMemcpyHostToDevice,
                          b.stream));
    if (!euler) {
        const int nu = ih * (ih + 1) / 2;
        const int nv = deep ? gram_tiles(ih) * 64 : nu + ih;
        if (deep)
            launch_gram_tiled_b(ih, b.d_act, b.d_step, (int)act.size(), b.gparts, b.gpart,
                                b.csums, b.stream);
        else
            launch_gram_b(ih, b.d_act, b.d_step, (int)act.size(), max_parts(b.tasks, true),
                          b.csums, b.stream);
        PF_HIP(hipMemcpyAsync(b.hcsums, b.csums, sizeof(double) * nv * act.size(),
                              hipMemcpyDeviceToHost, b.stream));
        PF_HIP(hipStreamSynchronize(b.stream));
        // each system's solve is independent: large batches split them over host threads
        auto solve = [&](size_t k0, size_t k1) {
            std::vector<double> U(ih * ih), V(ih), inv(ih * ih), C(ih);
            for (size_t k = k0; k < k1; ++k) {
                const double* g = b.hcsums + k * (size_t)nv;
                double* hc = b.h_coef() + k * (size_t)b.coef_stride;
                if (deep) {
                    for (int m = 0; m < ih; ++m) {
                        for (int nn = m; nn < ih; ++nn)
                            U[m * ih + nn] = U[nn * ih + m] = g[gram_tile_index(m, nn)];
                        V[m] = g[gram_tile_index(m, ih)];
                    }
                } else {
                    int q = 0;
                    for (int m = 0; m < ih; ++m)
                        for (int nn = m; nn < ih; ++nn) U[m * ih + nn] = U[nn * ih + m] = g[q++];
                    for (int m = 0; m < ih; ++m) V[m] = g[nu + m];
                }
                if (!lu_inverse(ih, U.data(), inv.data())) {
                    // scft! raises on a singular Gram matrix (anderson.jl:203-208); in a
                    // batch the system is reported as failed and dropped
                    res[act[k]].diverged = 1;
                    for (int nn = 0; nn < ih; ++nn) hc[nn] = 0.0;
                    continue;
                }
                matvec(ih, inv.data(), V.data(), C.data());
                for (int nn = 0; nn < ih; ++nn) hc[nn] = C[nn];
            }
        };
        if (b.pool && act.size() >= 64)
            b.pool->run(act.size(), solve);
        else
            solve(0, act.size());
        PF_HIP(hipMemcpyAsync(const_cast<double*>(b.d_coef()), b.h_coef(),
                              sizeof(double) * b.coef_stride * act.size(), hipMemcpyHostToDevice,
                              b.stream));
    }
    size_t Nmax = 0;
    for (int i : act) Nmax = std::max(Nmax, b.sys[i]->n * b.sys[i]->nmon());
    launch_mix_b(b.d_act, b.d_step, (int)act.size(), Nmax, b.stream);
    if (iter >= nhist + 1)
        for (int i : act) b.and_state[i]->base = (b.and_state[i]->base + 1) % (nhist + 1);
}

}  // namespace

void Batch::free_momentum() {
    for (auto& m : mom_state) {
        dev_free(m.mt);
        dev_free(m.vt);
        dev_free(m.step);
        dev_free(m.tmp);
    }
    mom_state.clear();
    dev_free(d_mom);
    d_mom = nullptr;
    mom_act.clear();
}

std::vector<BatchResult> Batch::scft_anderson(int nhist, int iter_euler, double lam_euler,
                                              int nsteps, double ftol, const int* mask,
                                              const Domain* dom, double stol) {
    AndersonStepper u(nhist, iter_euler, lam_euler);
    return scft_run(u, nsteps, ftol, mask, dom, stol);
}

void Batch::bind_momentum(const std::vector<int>& act) {
    if (act == mom_act && d_mom) return;
    std::vector<MomTask> t;
    for (int i : act) {
        System& s = *sys[i];
        const MomState& m = mom_state[i];
        t.push_back({m.mt, m.vt, m.step, s.w, s.res, s.n * s.nmon(), m.tmp});
    }
    PF_HIP(hipStreamSynchronize(stream));   // queued work may read the old table
    dev_free(d_mom);
    d_mom = (MomTask*)dev_alloc(sizeof(MomTask) * t.size());
    PF_HIP(hipMemcpyAsync(d_mom, t.data(), sizeof(MomTask) * t.size(), hipMemcpyHostToDevice,
                          stream));
    PF_HIP(hipStreamSynchronize(stream));
    mom_act = act;
}

std::vector<BatchResult> Batch::scft_momentum(int method, double lam, double gamma, double b1,
                                              double b2, int nsteps, double ftol,
                                              const int* mask, const Domain* dom, double stol) {
    if (method < 0 || method > 3)
        throw Error(-1, "batched MomentumUpdater: method must be MOM, RMS, ADAM or NAG");
    MomentumStepper u(method, lam, gamma, b1, b2);
    return scft_run(u, nsteps, ftol, mask, dom, stol);
}

std::vector<BatchResult> Batch::scft_euler(int method, double lam, int nsteps, double ftol,
                                           const int* mask, const Domain* dom, double stol) {
    if (method < 0 || method > 1)
        throw Error(-1, "batched EulerUpdater: method must be PEC or PECE");
    EulerStepper u(method, lam);
    return scft_run(u, nsteps, ftol, mask, dom, stol);
}

std::vector<BatchResult> Batch::scft_run(BatchStepper& u, int nsteps, double ftol,
                                         const int* mask, const Domain* dom, double stol) {
    const auto t0 = std::chrono::steady_clock::now();
    const size_t ns = sys.size();
    std::vector<int> act;
    for (size_t i = 0; i < ns; ++i)
        if (!mask || mask[i]) act.push_back((int)i);
    std::vector<BatchResult> res(ns);
    for (auto& r : res) r = {0, 0, std::nan(""), std::nan(""), 0.0, 0, 0.0};
    const bool vcell = dom != nullptr;
    if (vcell && dom->nskip < 1) throw Error(-1, "DomainUpdater nskip must be at least 1");
    if (act.empty()) return res;
    bool grand = false;
    for (int i : act) {
        System* s = sys[i];
        if (!s->canonical && !u.allows_grand()) throw Error(-2, u.noncanonical_error());
        grand = grand || !s->canonical;
        s->validate();
        s->sync();
    }
    const std::vector<int> run = act;
    PF_HIP(hipStreamSynchronize(stream));
    residuals(act);
    u.setup(*this, run);
    std::vector<double> v;
    sys_sums(act, v);
    // variable cell: every system gets its own DomainUpdater with the given settings,
    // set up as scft! does after its first residuals! (src/simulate.jl:88-91); its stress
    // reads the stacks of that residuals! pass
    std::vector<double> serr(ns, 0.0);
    dom_state.clear();
    dom_state.resize(ns);
    if (vcell) {
        std::vector<std::vector<double>> st(ns);
        stress_all(act, st);
        for (int i : act) {
            Domain* D = new Domain(dom->nper, dom->nskip, dom->lam, dom->pmax, dom->rlxn);
            dom_state[i].reset(D);
            // setup! (src/updaters/domain.jl:28-37) with the stress computed above
            const int np = (int)sys[i]->cell.params.size();
            D->fd_step = dom->fd_step;
            D->sys = sys[i];
            D->stress = st[i];
            D->oldparams.assign(np, 0.0);
            D->stepv.assign(np, 0.0);
            D->dstress.assign((size_t)np * np, 0.0);
            serr[i] = D->stress_error();
        }
    }
    std::vector<int> next;
    for (int i : act) {
        const double fe = ferr_of(&v[(size_t)i * 5]);
        const bool conv = scft_converged(fe, ftol, serr[i], stol, vcell);
        res[i] = {conv, 0, fe, fhelm_of(*sys[i], &v[(size_t)i * 5]), 0.0, 0, serr[i]};
        if (!conv) next.push_back(i);
    }
    act = next;
    int step = 0;
    bool queued = false;   // the residuals of the current active set are already queued
    // fresh[i]: system i's last device work is a residuals pass queued after its last
    // update, which is then its final state (scft! ends with one residuals! after the
    // last step)
    std::vector<char> fresh(ns, 0);
    const bool lead = u.starts_with_residuals();
    while (!act.empty() && step < nsteps) {
        if (queued || !lead)
            bind_active(act);
        else
            residuals(act);
        queued = false;
        u.update(*this, act, res);
        for (int i : act) fresh[i] = 0;
        // the cell update of scft! (src/simulate.jl:113-121): the stress of the stacks of
        // this step's residuals!, and when serr > stol the Newton iterations, each ending
        // with a residuals! at the new cell (all due systems in lockstep)
        if (vcell && step % dom->nskip == 0) domain_round(act, stol, serr, res);
        sums_launch(act);
        // The next iteration's residuals (density! of the updated fields) are queued before
        // the convergence test reads its sums, so the device does not wait for the host
        // between iterations.  Each system's residuals depend on its own fields only: for
        // the systems that stay active they are what the next iteration computes, and for
        // a system that the test drops they are the final residuals! of scft!.  Not with a
        // grand-ensemble system, whose free energy reads species fractions that the queued
        // pass would overwrite.
        if (step + 1 < nsteps && !vcell && lead && !grand) {
            residuals(act);
            queued = true;
            for (int i : act) fresh[i] = 1;
        }
        sums_read(act, v);
        next.clear();
        for (int i : act) {
            const double fe = ferr_of(&v[(size_t)i * 5]);
            res[i].ferr = fe;
            res[i].fhelm = fhelm_of(*sys[i], &v[(size_t)i * 5]);
            res[i].nsteps = step;
            if (res[i].diverged) {
                res[i].ferr = std::nan("");
            } else if (std::isnan(fe)) {
                // scft! raises "Diverging field updater" (src/simulate.jl:109-111); in a
                // batch the system is reported (ferr = NaN, not converged) and dropped
                res[i].diverged = 1;
            } else if (scft_converged(fe, ftol, serr[i], stol, vcell)) {
                res[i].converged = 1;
            } else {
                next.push_back(i);
            }
        }
        // a smaller active set gets new descriptors and lane tables: the queued pass must
        // have read the old ones before they are freed
        if (queued && next != act) PF_HIP(hipStreamSynchronize(stream));
        act = next;
        step += 1;
    }
    for (int i : act) res[i].nsteps = step;
    // final state (src/simulate.jl:133-136): one residuals! after each system's last step
    // (already queued for the systems dropped with a pass in flight)
    const std::vector<int>& all = run;
    std::vector<int> need;
    for (int i : all)
        if (!fresh[i]) need.push_back(i);
    if (!need.empty()) residuals(need);
    sys_sums(all, v);
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i : all) {
        if (!res[i].diverged) {
            res[i].ferr = ferr_of(&v[(size_t)i * 5]);
            res[i].fhelm = fhelm_of(*sys[i], &v[(size_t)i * 5]);
        }
        res[i].serr = vcell ? dom_state[i]->stress_error() : 0.0;
        res[i].time = t;
    }
    PF_HIP(hipStreamSynchronize(stream));
    return res;
}

}  // namespace pf
