// Propagator engine: owns the q / q+ stacks, LW tables, FFT plans and twiddles of one
// system, and sequences the sweep kernels for density!(sys) and the stress pass.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pf_batch.h"

namespace pf {

static FftPlan make_plan(int n) {
    FftPlan p{};
    p.n = n;
    int m = n;
    auto push = [&](int r) {
        if (p.npass >= PF_MAXPASS) throw Error(-2, "transform length has too many factors");
        p.radix[p.npass++] = (int8_t)r;
        m /= r;
    };
    while (m % 4 == 0) push(4);
    while (m % 2 == 0) push(2);
    while (m % 3 == 0) push(3);
    if (m != 1)
        throw Error(-2, "grid size " + std::to_string(n) +
                            " has a prime factor other than 2 or 3 (unsupported)");
    return p;
}

// launches per event pair in timing mode
constexpr int kTimingReps = 4;

// Lanes [off, off + cnt) of a launch at step t: 1 + the last one still running (nsteps >
// t).  Lanes past it are finished and are not launched; a finished lane before it is
// launched and skips its work in the kernel.
template <typename T>
static int live_lanes(const std::vector<Lane<T>>& v, int off, int cnt, int t) {
    for (int i = cnt; i > 0; --i)
        if (v[(size_t)(off + i - 1)].nsteps > t) return i;
    return 0;
}

template <typename T>
static std::vector<cpx<T>> twiddles(int n) {
    std::vector<cpx<T>> t(std::max(n, 1));
    for (int m = 0; m < n; ++m) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * m / n;
        t[m] = {(T)(double)cosl(a), (T)(double)sinl(a)};
    }
    return t;
}

// 3-D RQM4 steps [t0, t1) of nl lanes (src/system/mde.jl:83-104): per step the pencils LD1
// on the full-step spectrum and LD2 on the half-step one, the RQ_MID slab, a pencil LD2 and
// the RQ_END slab (which starts the next step's spectra).  Compile-time forms where the
// size is specialised (slabs only where both planes fit in LDS: `two`), generic otherwise.
template <typename T>
void rqm4_sweep(const Lane<T>* dl, int nl, int nx, int ny, int nz, int t0, int t1, bool two,
                bool fixed, int ntiles, int tile, int lds_pencil, int rq_lds, hipStream_t st) {
    auto slab = [&](int mode, int t) {
        if (!two || !fixed || !fixed_slab_rqm4<T>(dl, nl, nx, ny, nz, mode, t, st))
            launch_slab_rqm4<T>(dl, nl, nz, mode, t, rq_lds, st);
    };
    auto pencil = [&](int t, int half) {
        if (!fixed || !fixed_pencil_rqm4<T>(dl, nl, nx, ny, nz, t, half, st))
            launch_pencil<T>(dl, nl, ny, ntiles, tile, t, lds_pencil, 256, st, half);
    };
    slab(RQ_START, t0);
    for (int t = t0; t < t1; ++t) {
        pencil(t, 0);
        pencil(t, 1);
        slab(RQ_MID, t);
        pencil(t, 1);
        slab(RQ_END, t);
    }
}

template <typename T>
class Engine : public EngineBase {
    struct SpBuf {
        T* q = nullptr;
        T* qc = nullptr;
        std::vector<T*> lw;
        std::vector<T*> lw2;       // RQM4 half-step tables
        int row0 = 0;              // first stress row of this species
        // checkpointed sweeps: q and q+ at the segment boundaries bnd[0..nseg], and the
        // Simpson accumulators; q / qc above are then one segment (K + 1 slices) each
        T* ckq = nullptr;
        T* ckqc = nullptr;
        double* acc = nullptr;     // [blk][n]
        std::vector<int> bnd;      // 1 = bnd[0] < ... < bnd[nseg] = Ns + 1
        int jq = -1, jc = -1;      // lanes of q and q+
    };
    int ckK = 0;                   // segment length in slices; 0: full stacks
    std::vector<Lane<T>> seglanes; // [lane][segment] copies with the stack shifted
    std::vector<int> segoff;       // first seglanes entry of each lane
    Lane<T>* d_seglanes = nullptr;
    System* S = nullptr;
    std::vector<SpBuf> sb;
    cpx<T>* tw = nullptr;
    const cpx<T>*twx = nullptr, *twy = nullptr, *twz = nullptr;
    FftPlan px{}, py{}, pz{};
    std::vector<cpx<T>*> kbufs;
    std::vector<Lane<T>> lanes;
    Lane<T>* d_lanes = nullptr;
    std::vector<SimpsonTask<T>> tasks;
    SimpsonTask<T>* d_tasks = nullptr;
    long lanes_cell_version = -1;
    int maxNs = 0;
    int planar = 1;
    int ntiles = 1, tile = 1;
    size_t nk = 0;
    int lds_plane = 0, lds_pencil = 0;
    bool rq_one = false;             // RQM4 with one xy plane of LDS (two do not fit)
    // stress
    std::vector<StressLane<T>> slanes;
    StressLane<T>* d_slanes = nullptr;
    long slanes_cell_version = -1;
    int nrows = 0;
    cpx<T>* kstress = nullptr;
    int schunk = 0;
    double* srows = nullptr;
    double* spart = nullptr;
    // timing of the propagator kernels: 0 slab, 1 pencil, 2 planar
    bool tim = false;
    bool fixed = true;               // use the compile-time specialised kernels when available
    std::vector<hipEvent_t> evpool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evlog;
    size_t evnext = 0;
    // reverse sweeps: one adjoint lane per forward lane, with its own scratch
    std::vector<int> lane_sp;        // species of each forward lane
    std::vector<Lane<T>> alanes;
    Lane<T>* d_alanes = nullptr;
    std::vector<void*> abuf;         // per adjoint lane: kbuf, kb2, wbar, gpart

    void free_adjoint() {
        for (void* p : abuf) dev_free(p);
        abuf.clear();
        alanes.clear();
        dev_free(d_alanes);
        d_alanes = nullptr;
    }

    void free_all() {
        free_adjoint();
        lane_sp.clear();
        dev_free(d_seglanes);
        d_seglanes = nullptr;
        seglanes.clear();
        segoff.clear();
        for (auto& b : sb) {
            dev_free(b.ckq);
            dev_free(b.ckqc);
            dev_free(b.acc);
            dev_free(b.q);
            dev_free(b.qc);
            for (T* p : b.lw) dev_free(p);
            for (T* p : b.lw2) dev_free(p);
        }
        sb.clear();
        for (auto* k : kbufs) dev_free(k);
        kbufs.clear();
        dev_free(d_lanes);
        dev_free(d_tasks);
        dev_free(d_slanes);
        dev_free(kstress);
        dev_free(srows);
        dev_free(spart);
        d_lanes = nullptr;
        d_tasks = nullptr;
        d_slanes = nullptr;
        kstress = nullptr;
        srows = spart = nullptr;
        slanes.clear();
    }

    void upload(void* dst, const void* src, size_t bytes) {
        PF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S->stream));
        PF_HIP(hipStreamSynchronize(S->stream));
    }

    void build() {
        System& s = *S;
        free_all();
        const int nx = s.dims[0], ny = s.dims[1], nz = s.dims[2];
        const size_t n = s.n;
        planar = (nz == 1);
        nk = (size_t)(nx / 2 + 1) * ny * nz;
        if (!tw) {
            px = make_plan(nx / 2);
            py = make_plan(ny);
            pz = make_plan(nz);
            auto tx = twiddles<T>(nx), ty = twiddles<T>(ny), tz = twiddles<T>(nz);
            tw = (cpx<T>*)dev_alloc(sizeof(cpx<T>) * (tx.size() + ty.size() + tz.size()));
            upload(tw, tx.data(), sizeof(cpx<T>) * tx.size());
            upload(tw + tx.size(), ty.data(), sizeof(cpx<T>) * ty.size());
            upload(tw + tx.size() + ty.size(), tz.data(), sizeof(cpx<T>) * tz.size());
            twx = tw;
            twy = tw + tx.size();
            twz = tw + tx.size() + ty.size();
        }
        const int nxh = nx / 2 + 1;
        ntiles = (nxh + 31) / 32;
        tile = (nxh + ntiles - 1) / ntiles;
        lds_plane = (int)(sizeof(cpx<T>) * ny * nxh);
        lds_pencil = (int)(sizeof(cpx<T>) * nz * tile);
        // RQM4 holds the full- and half-step planes in LDS together where two fit, else
        // one after the other (lanes marked rqm4 = 2)
        rq_one = s.mde == 1 && 2 * lds_plane > 160 * 1024;
        if (lds_plane > 160 * 1024)
            throw Error(-2, "xy plane of " + std::to_string(nx) + "x" + std::to_string(ny) +
                                " does not fit the 160 KiB LDS of one workgroup");
        maxNs = 0;
        lanes.clear();
        tasks.clear();
        sb.resize(s.species.size());
        ckK = s.ckpt;
        int row = 0;
        for (size_t i = 0; i < s.species.size(); ++i) {
            const Species& sp = s.species[i];
            if (sp.kind != 0) continue;
            SpBuf& b = sb[i];
            b.row0 = row;
            row += sp.Ns + 1;
            if (ckK) {
                // segments of K slices between checkpoints; every chain gets a real q+ sweep
                const int K = ckK > 0 ? ckK : (int)std::ceil(std::sqrt((double)sp.Ns + 1.0));
                b.bnd.clear();
                for (int s0 = 1; s0 < sp.Ns + 1; s0 += K) b.bnd.push_back(s0);
                b.bnd.push_back(sp.Ns + 1);
                const int nseg = (int)b.bnd.size() - 1;
                b.q = (T*)dev_alloc(sizeof(T) * n * (K + 1));
                b.qc = (T*)dev_alloc(sizeof(T) * n * (K + 1));
                b.ckq = (T*)dev_alloc(sizeof(T) * n * (nseg + 1));
                b.ckqc = (T*)dev_alloc(sizeof(T) * n * (nseg + 1));
                b.acc = (double*)dev_alloc(sizeof(double) * n * sp.nblk());
            } else {
                b.q = (T*)dev_alloc(sizeof(T) * n * (sp.Ns + 1));
                if (sp.nblk() > 1) b.qc = (T*)dev_alloc(sizeof(T) * n * (sp.Ns + 1));
            }
            for (int k = 0; k < sp.nblk(); ++k) b.lw.push_back((T*)dev_alloc(sizeof(T) * n));
            if (s.mde == 1)
                for (int k = 0; k < sp.nblk(); ++k) b.lw2.push_back((T*)dev_alloc(sizeof(T) * n));
            maxNs = std::max(maxNs, sp.Ns);
            for (int d = 0; d < (b.qc ? 2 : 1); ++d) {
                (d == 0 ? b.jq : b.jc) = (int)lanes.size();
                Lane<T> L{};
                L.nx = nx;
                L.ny = ny;
                L.nz = nz;
                L.dir = d == 0 ? 1 : -1;
                L.nsteps = sp.Ns;
                L.nblk = sp.nblk();
                for (int k = 0; k <= sp.nblk(); ++k) L.begin[k] = sp.begin[k];
                L.rqm4 = s.mde == 1 ? (rq_one ? 2 : 1) : 0;
                for (int k = 0; k < sp.nblk(); ++k) {
                    L.lw[k] = b.lw[k];
                    if (s.mde == 1) L.lw2[k] = b.lw2[k];
                    const double bb = sp.b_block[k];
                    L.ldc[k] = sp.ds_block[k] * (bb * bb) / 6.0;
                }
                L.stack = d == 0 ? b.q : b.qc;
                if (!planar) {
                    kbufs.push_back((cpx<T>*)dev_alloc(sizeof(cpx<T>) * nk));
                    L.kbuf = kbufs.back();
                    if (s.mde == 1) {
                        // RQM4: the half steps' spectrum and the full step's result
                        kbufs.push_back((cpx<T>*)dev_alloc(sizeof(cpx<T>) * nk));
                        L.kb2 = kbufs.back();
                        kbufs.push_back((cpx<T>*)dev_alloc(sizeof(T) * n));
                        L.rq = reinterpret_cast<T*>(kbufs.back());
                    }
                }
                L.twx = twx;
                L.twy = twy;
                L.twz = twz;
                L.px = px;
                L.py = py;
                L.pz = pz;
                lanes.push_back(L);
                lane_sp.push_back((int)i);
            }
            SimpsonTask<T> tk{};
            tk.q = b.q;
            tk.qc = b.qc;
            tk.Ns = sp.Ns;
            tk.nblk = sp.nblk();
            for (int k = 0; k <= sp.nblk(); ++k) tk.begin[k] = sp.begin[k];
            for (int k = 0; k < sp.nblk(); ++k) {
                tk.ds[k] = sp.ds_block[k];
                const int slot = sp.block_slot[k];
                tk.blk_mon[k] = (int)(std::find(sp.slots.begin(), sp.slots.end(), slot) -
                                      sp.slots.begin());
            }
            tk.Nref = sp.Nref;
            tk.isp = (int)i;
            tk.rho_sp = sp.rho_sp;
            tk.nmon = (int)sp.slots.size();
            tk.n = n;
            tk.sc = s.sc;
            tasks.push_back(tk);
        }
        nrows = row;
        if (!lanes.empty()) d_lanes = (Lane<T>*)dev_alloc(sizeof(Lane<T>) * lanes.size());
        if (!tasks.empty()) {
            d_tasks = (SimpsonTask<T>*)dev_alloc(sizeof(SimpsonTask<T>) * tasks.size());
            upload(d_tasks, tasks.data(), sizeof(SimpsonTask<T>) * tasks.size());
        }
        lanes_cell_version = -1;
        slanes_cell_version = -1;
        s.dirty_species = 0;
    }

    void refresh_lanes() {
        if (lanes.empty() || lanes_cell_version == S->cell.version) return;
        for (auto& L : lanes) std::memcpy(L.G, S->cell.G, sizeof(L.G));
        upload(d_lanes, lanes.data(), sizeof(Lane<T>) * lanes.size());
        if (ckK) build_seglanes();
        lanes_cell_version = S->cell.version;
    }

    // --- checkpointed sweeps ------------------------------------------------------------
    // Segment m of a chain covers slices bnd[m]..bnd[m+1]; a lane copy whose stack pointer
    // is shifted so that slice s lands in slot s - bnd[m] of the segment buffer.
    void build_seglanes() {
        seglanes.clear();
        segoff.assign(lanes.size(), 0);
        for (size_t j = 0; j < lanes.size(); ++j) {
            const SpBuf& b = sb[lane_sp[j]];
            segoff[j] = (int)seglanes.size();
            for (size_t m = 0; m + 1 < b.bnd.size(); ++m) {
                Lane<T> L = lanes[j];
                L.stack = L.stack - (ptrdiff_t)(b.bnd[m] - 1) * (ptrdiff_t)S->n;
                seglanes.push_back(L);
            }
        }
        dev_free(d_seglanes);
        d_seglanes = (Lane<T>*)dev_alloc(sizeof(Lane<T>) * seglanes.size());
        upload(d_seglanes, seglanes.data(), sizeof(Lane<T>) * seglanes.size());
    }

    // steps [t0, t1) of lane j that fill segment m
    void seg_steps(int j, int m, int* t0, int* t1) const {
        const SpBuf& b = sb[lane_sp[j]];
        const int lo = b.bnd[m], hi = b.bnd[m + 1], Ns = lanes[j].nsteps;
        if (lanes[j].dir > 0) {
            *t0 = lo - 1;
            *t1 = hi - 1;
        } else {
            *t0 = Ns + 1 - hi;
            *t1 = Ns + 1 - lo;
        }
    }

    // Recompute segment m of lane j from its checkpoint into the segment buffer; with
    // `store`, write the checkpoint at the segment's far end.
    void run_segment(int j, int m, bool store) {
        System& s = *S;
        const size_t n = s.n;
        hipStream_t st = s.stream;
        SpBuf& b = sb[lane_sp[j]];
        const Lane<T>& L = lanes[j];
        const int lo = b.bnd[m], hi = b.bnd[m + 1];
        T* seg = L.stack;
        T* ck = L.dir > 0 ? b.ckq : b.ckqc;
        const int from = L.dir > 0 ? m : m + 1, to = L.dir > 0 ? m + 1 : m;
        const int slot_from = L.dir > 0 ? 0 : hi - lo, slot_to = L.dir > 0 ? hi - lo : 0;
        PF_HIP(hipMemcpyAsync(seg + (size_t)slot_from * n, ck + (size_t)from * n, sizeof(T) * n,
                              hipMemcpyDeviceToDevice, st));
        int t0, t1;
        seg_steps(j, m, &t0, &t1);
        const Lane<T>* dl = d_seglanes + segoff[j] + m;
        if (planar && s.mde == 1) {
            if (!fixed || !fixed_planar_rqm4<T>(dl, 1, s.dims[0], s.dims[1], t0, t1, rq_one, st))
                launch_planar_steps<T>(dl, 1, t0, t1, rqm4_lds(), 256, st);
        } else if (planar) {
            if (!fixed || !fixed_planar<T>(dl, 1, s.dims[0], s.dims[1], t0, t1, st))
                launch_planar_steps<T>(dl, 1, t0, t1, lds_plane, 256, st);
        } else if (s.mde == 1) {
            // RQM4 steps of the segment (the one-system launch sequence from step t0)
            rqm4_sweep<T>(dl, 1, s.dims[0], s.dims[1], s.dims[2], t0, t1, !rq_one, fixed, ntiles,
                          tile, lds_pencil, rqm4_lds(), st);
        } else {
            const int nz = s.dims[2], ny = s.dims[1];
            if (!fixed || !fixed_slab<T>(dl, 1, s.dims[0], ny, nz, SLAB_FWD, t0, st))
                launch_slab<T>(dl, 1, nz, SLAB_FWD, t0, lds_plane, 256, st);
            for (int t = t0; t < t1; ++t) {
                if (!fixed || !fixed_pencil<T>(dl, 1, s.dims[0], ny, nz, true, t, st))
                    launch_pencil<T>(dl, 1, ny, ntiles, tile, t, lds_pencil, 256, st);
                if (!fixed || !fixed_slab<T>(dl, 1, s.dims[0], ny, nz, SLAB_INVFWD, t, st))
                    launch_slab<T>(dl, 1, nz, SLAB_INVFWD, t, lds_plane, 256, st);
            }
        }
        if (store)
            PF_HIP(hipMemcpyAsync(ck + (size_t)to * n, seg + (size_t)slot_to * n, sizeof(T) * n,
                                  hipMemcpyDeviceToDevice, st));
    }

    // density! with checkpointed sweeps: forward sweeps keep only the boundary slices;
    // Simpson runs segment by segment on recomputed slices.
    void density_ckpt(System& s) {
        const size_t n = s.n;
        hipStream_t st = s.stream;
        for (size_t i = 0; i < s.species.size(); ++i) {
            const Species& sp = s.species[i];
            if (sp.kind != 0) continue;
            SpBuf& b = sb[i];
            const int nseg = (int)b.bnd.size() - 1;
            launch_fill<T>(b.ckq, T(1), n, st);
            launch_fill<T>(b.ckqc + (size_t)nseg * n, T(1), n, st);
            for (int m = 0; m < nseg; ++m) run_segment(b.jq, m, true);
            for (int m = nseg - 1; m >= 0; --m) run_segment(b.jc, m, true);
        }
        // Q from the last checkpoint of q
        std::vector<SimpsonTask<T>> tq = tasks;
        for (auto& tk : tq) {
            const SpBuf& b = sb[tk.isp];
            tk.q = b.ckq + (ptrdiff_t)((int)b.bnd.size() - 1 - tk.Ns) * (ptrdiff_t)n;
        }
        auto* d = (SimpsonTask<T>*)dev_alloc(sizeof(SimpsonTask<T>) * tq.size());
        upload(d, tq.data(), sizeof(SimpsonTask<T>) * tq.size());
        launch_species_Q<T>(d, (int)tq.size(), n, s.partial, reduce_parts(n), st);
        PF_HIP(hipStreamSynchronize(st));
        dev_free(d);
    }

    void simpson_ckpt(System& s, const std::vector<double*>* blk) {
        const size_t n = s.n;
        hipStream_t st = s.stream;
        for (const SimpsonTask<T>& tk0 : tasks) {
            SpBuf& b = sb[tk0.isp];
            const Species& sp = s.species[tk0.isp];
            const int nseg = (int)b.bnd.size() - 1;
            launch_fill<double>(b.acc, 0.0, n * sp.nblk(), st);
            for (int m = 0; m < nseg; ++m) {
                run_segment(b.jq, m, false);
                run_segment(b.jc, m, false);
                SimpsonTask<T> tk = tk0;
                tk.q = seglanes[segoff[b.jq] + m].stack;
                tk.qc = seglanes[segoff[b.jc] + m].stack;
                const int hi = (m == nseg - 1) ? b.bnd[m + 1] : b.bnd[m + 1] - 1;
                launch_simpson_seg<T>(tk, b.bnd[m], hi, b.acc, st);
            }
            SimpsonTask<T> tk = tk0;
            tk.rho_blk = blk ? (*blk)[tk0.isp] : nullptr;
            launch_simpson_fin<T>(tk, b.acc, st);
        }
    }

    hipEvent_t ev() {
        if (evnext >= evpool.size()) {
            hipEvent_t e;
            PF_HIP(hipEventCreate(&e));
            evpool.push_back(e);
        }
        return evpool[evnext++];
    }
    // timing mode: one launch in 8 of each kind runs kTimingReps times back to back between
    // two events (the events' own overhead is spread over the repeats); the repeats redo
    // the launch on its own output, so the sweep's results are not meaningful then
    size_t ncalls[3] = {0, 0, 0};
    template <typename F>
    void timed(int cls, F&& f) {
        if (!tim || (ncalls[cls]++ % 8) != 0) {
            f();
            return;
        }
        hipEvent_t a = ev(), b = ev();
        PF_HIP(hipEventRecord(a, S->stream));
        for (int r = 0; r < kTimingReps; ++r) f();
        PF_HIP(hipEventRecord(b, S->stream));
        evlog.push_back({cls, {a, b}});
    }

   public:
    // --- access for the batch engine -------------------------------------------------
    void prepare() {
        if (S->ckpt) throw Error(-2, "systems in a batch keep full propagator stacks");
        if (S->dirty_species || sb.size() != S->species.size()) build();
        for (auto& L : lanes) std::memcpy(L.G, S->cell.G, sizeof(L.G));
    }
    const std::vector<Lane<T>>& host_lanes() const { return lanes; }
    const std::vector<SimpsonTask<T>>& host_tasks() const { return tasks; }
    T* q_stack(int isp) const { return sb[isp].q; }
    T* qc_stack(int isp) const { return sb[isp].qc; }
    T* lw_table(int isp, int blk) const { return sb[isp].lw[blk]; }
    T* lw_half_table(int isp, int blk) const {
        return sb[isp].lw2.empty() ? nullptr : sb[isp].lw2[blk];
    }
    int tiles() const { return ntiles; }
    int tile_w() const { return tile; }
    int plane_lds() const { return lds_plane; }
    // LDS of an RQM4 slab or planar workgroup: one or two xy planes
    int rqm4_lds() const { return rq_one ? lds_plane : 2 * lds_plane; }

    // The RK2 sweep of lanes [off, off + cnt) (device table dl) over `steps` contour steps
    // on stream ss: the first xy-slab forward pass, then pencil and slab per step.
    void rk2_sweep(System& s, const Lane<T>* dl, int off, int cnt, int steps, hipStream_t ss) {
        const int nz = s.dims[2], ny = s.dims[1];
        if (!fixed || !fixed_slab<T>(dl, cnt, s.dims[0], ny, nz, SLAB_FWD, 0, ss))
            launch_slab<T>(dl, cnt, nz, SLAB_FWD, 0, lds_plane, 256, ss);
        for (int t = 0; t < steps; ++t) {
            const int nla = live_lanes(lanes, off, cnt, t);
            if (!fixed || !fixed_pencil<T>(dl, nla, s.dims[0], ny, nz, true, t, ss))
                launch_pencil<T>(dl, nla, ny, ntiles, tile, t, lds_pencil, 256, ss);
            if (!fixed || !fixed_slab<T>(dl, nla, s.dims[0], ny, nz, SLAB_INVFWD, t, ss))
                launch_slab<T>(dl, nla, nz, SLAB_INVFWD, t, lds_plane, 256, ss);
        }
    }
    // one system's 3-D RK2 lanes as two stream groups: three or more lanes, unless the
    // grid's schedule asks for one stream (split = 1).  Two lanes stay on one stream
    // (measured: no gain at 128^3, a loss at 64^3), whatever a batch's tuned split is.
    bool split_lanes(int nl) const {
        const int sp = schedule(sizeof(T), S->dims[0], S->dims[1], S->dims[2]).split;
        return nl >= 3 && sp != 1;
    }
    hipStream_t sub[2] = {nullptr, nullptr};
    hipEvent_t sub_done[2] = {nullptr, nullptr}, sub_fork = nullptr;
    int pencil_lds() const { return lds_pencil; }

    explicit Engine(System& s) : S(&s) {}
    ~Engine() override {
        for (int p = 0; p < 2; ++p) {
            if (sub[p]) (void)hipStreamDestroy(sub[p]);
            if (sub_done[p]) (void)hipEventDestroy(sub_done[p]);
        }
        if (sub_fork) (void)hipEventDestroy(sub_fork);
        if (hrows) (void)hipHostFree(hrows);
        free_all();
        dev_free(tw);
        for (auto e : evpool) (void)hipEventDestroy(e);
    }

    void timing(int on) override {
        tim = on != 0;
        evlog.clear();
        evnext = 0;
        ncalls[0] = ncalls[1] = ncalls[2] = 0;
    }

    double timing_ms(int which, int* count) override {
        PF_HIP(hipStreamSynchronize(S->stream));
        double tot = 0.0;
        int c = 0;
        for (auto& e : evlog)
            if (e.first == which) {
                float ms = 0.f;
                PF_HIP(hipEventElapsedTime(&ms, e.second.first, e.second.second));
                tot += ms;
                c += kTimingReps;
            }
        if (count) *count = c;
        return tot;
    }

    void density(System& s) override {
        if (s.dirty_species || sb.size() != s.species.size() || ckK != s.ckpt) build();
        refresh_lanes();
        hipStream_t st = s.stream;
        const size_t n = s.n;
        for (size_t i = 0; i < s.species.size(); ++i) {
            const Species& sp = s.species[i];
            if (sp.kind != 0) continue;
            for (int k = 0; k < sp.nblk(); ++k) {
                const int slot = sp.block_slot[k];
                const double coef = (-(sp.ds_block[k] / 2.0)) * s.mon_vol[slot];
                launch_lw_table<T>(sb[i].lw[k], s.w + (size_t)slot * n, coef, n, st);
                if (s.mde == 1) launch_lw_half_table<T>(sb[i].lw2[k], s.w + (size_t)slot * n, coef, n, st);
            }
            if (ckK) continue;
            launch_fill<T>(sb[i].q, T(1), n, st);
            if (sb[i].qc) launch_fill<T>(sb[i].qc + (size_t)sp.Ns * n, T(1), n, st);
        }
        if (ckK) {
            density_ckpt(s);
            const int nparts = reduce_parts(n);
            for (size_t i = 0; i < s.species.size(); ++i) {
                const Species& sp = s.species[i];
                if (sp.kind != 1) continue;
                const int slot = sp.slots[0];
                launch_point_density(s.w + (size_t)slot * n, s.mon_vol[slot], n, sp.rho_sp,
                                     s.partial, nparts, s.sc, (int)i, st);
            }
            simpson_ckpt(s, nullptr);
            return;
        }
        const int nl = (int)lanes.size();
        if (nl) {
            if (planar) {
                timed(2, [&] {
                    if (s.mde == 1) {
                        if (!fixed || !fixed_planar_rqm4<T>(d_lanes, nl, s.dims[0], s.dims[1], 0,
                                                            maxNs, rq_one, st))
                            launch_planar_steps<T>(d_lanes, nl, 0, maxNs, rqm4_lds(), 256, st);
                    } else if (!fixed || !fixed_planar<T>(d_lanes, nl, s.dims[0], s.dims[1], 0, maxNs, st))
                        launch_planar_steps<T>(d_lanes, nl, 0, maxNs, lds_plane, 256, st);
                });
            } else if (s.mde == 1) {
                // RQM4: per contour step two spectra through the pencils and three slab
                // launches (the next step's forward transforms ride on RQ_END)
                rqm4_sweep<T>(d_lanes, nl, s.dims[0], s.dims[1], s.dims[2], 0, maxNs, !rq_one,
                              fixed, ntiles, tile, lds_pencil, rqm4_lds(), st);
            } else if (!tim && split_lanes(nl)) {
                // three or more lanes: two groups on streams of their own, so that one
                // group's slab launches fill the CUs the other's launches leave idle
                if (!sub[0]) {
                    for (int p = 0; p < 2; ++p) {
                        PF_HIP(hipStreamCreateWithFlags(&sub[p], hipStreamNonBlocking));
                        PF_HIP(hipEventCreateWithFlags(&sub_done[p], hipEventDisableTiming));
                    }
                    PF_HIP(hipEventCreateWithFlags(&sub_fork, hipEventDisableTiming));
                }
                PF_HIP(hipEventRecord(sub_fork, st));
                const int c0 = nl / 2;
                for (int p = 0; p < 2; ++p) {
                    const int off = p == 0 ? 0 : c0, cnt = p == 0 ? c0 : nl - c0;
                    int steps = 0;
                    for (int j = off; j < off + cnt; ++j) steps = std::max(steps, lanes[j].nsteps);
                    PF_HIP(hipStreamWaitEvent(sub[p], sub_fork, 0));
                    rk2_sweep(s, d_lanes + off, off, cnt, steps, sub[p]);
                    PF_HIP(hipEventRecord(sub_done[p], sub[p]));
                    PF_HIP(hipStreamWaitEvent(st, sub_done[p], 0));
                }
            } else {
                const int nz = s.dims[2], ny = s.dims[1];
                timed(0, [&] {
                    if (!fixed || !fixed_slab<T>(d_lanes, nl, s.dims[0], ny, nz, SLAB_FWD, 0, st))
                        launch_slab<T>(d_lanes, nl, nz, SLAB_FWD, 0, lds_plane, 256, st);
                });
                for (int t = 0; t < maxNs; ++t) {
                    const int nla = live_lanes(lanes, 0, nl, t);
                    timed(1, [&] {
                        if (!fixed || !fixed_pencil<T>(d_lanes, nla, s.dims[0], ny, nz, true, t, st))
                            launch_pencil<T>(d_lanes, nla, ny, ntiles, tile, t, lds_pencil, 256, st);
                    });
                    timed(0, [&] {
                        if (!fixed || !fixed_slab<T>(d_lanes, nla, s.dims[0], ny, nz, SLAB_INVFWD, t, st))
                            launch_slab<T>(d_lanes, nla, nz, SLAB_INVFWD, t, lds_plane, 256, st);
                    });
                }
            }
        }
        const int nparts = reduce_parts(n);
        if (!tasks.empty())
            launch_species_Q<T>(d_tasks, (int)tasks.size(), n, s.partial, nparts, st);
        for (size_t i = 0; i < s.species.size(); ++i) {
            const Species& sp = s.species[i];
            if (sp.kind != 1) continue;
            const int slot = sp.slots[0];
            launch_point_density(s.w + (size_t)slot * n, s.mon_vol[slot], n, sp.rho_sp,
                                 s.partial, nparts, s.sc, (int)i, st);
        }
        if (!tasks.empty()) launch_simpson<T>(d_tasks, (int)tasks.size(), n, st);
    }

    // Stress scratch (src/species/multiblock.jl:233-295): the row table, the per-row sums,
    // and for 3-D grids the k-space partial sums and the transforms of q_s and q+_s for
    // `chunk` rows at a time.
    struct StressBufs {
        StressLane<T>* d_slanes = nullptr;
        double* srows = nullptr;
        double* spart = nullptr;
        cpx<T>* kstress = nullptr;
        int chunk = 1;
    };
    bool stress_fused() const { return planar && 2 * lds_plane <= 160 * 1024; }
    size_t spart_len() const {
        return (size_t)std::max(nrows, 1) * S->dims[1] * ntiles * PF_MAXPAR;
    }

    // the row table for buffers `bf` (host copy in `out`)
    void fill_stress_lanes(std::vector<StressLane<T>>& out, const StressBufs& bf) {
        System& s = *S;
        const int nx = s.dims[0], ny = s.dims[1], nz = s.dims[2];
        const size_t n = s.n;
        const bool fused = stress_fused();
        out.assign(nrows, StressLane<T>{});
        const int np = (int)s.cell.params.size();
        for (size_t i = 0; i < s.species.size(); ++i) {
            const Species& sp = s.species[i];
            if (sp.kind != 0) continue;
            const SpBuf& b = sb[i];
            for (int sl = 1; sl <= sp.Ns + 1; ++sl) {
                const int r = b.row0 + sl - 1;
                StressLane<T>& L = out[r];
                L.nx = nx;
                L.ny = ny;
                L.nz = nz;
                if (ckK) {
                    // checkpointed: slice sl is read from the segment buffers while the
                    // segment that owns its row (stress_seg_rows) is recomputed
                    const size_t slot = (size_t)(sl - b.bnd[stress_seg(b, sl)]);
                    L.q = b.q + slot * n;
                    L.qc = b.qc + slot * n;
                } else {
                    L.q = b.q + (size_t)(sl - 1) * n;
                    L.qc = b.qc ? b.qc + (size_t)(sl - 1) * n : b.q + (size_t)(sp.Ns + 1 - sl) * n;
                }
                L.out = r;
                L.nparams = np;
                L.dim = s.cell.dim;
                std::memset(L.dGG, 0, sizeof(L.dGG));
                for (int k = 0; k < 9 * np; ++k) L.dGG[k] = s.cell.dGG[k];
                L.twx = twx;
                L.twy = twy;
                L.twz = twz;
                L.px = px;
                L.py = py;
                L.pz = pz;
                if (!fused) {
                    const int c = r % bf.chunk;
                    L.kq = bf.kstress + (size_t)(2 * c) * nk;
                    L.kqc = bf.kstress + (size_t)(2 * c + 1) * nk;
                }
            }
        }
    }

    void build_stress() {
        const bool fused = stress_fused();
        if (slanes.empty() || (int)slanes.size() != nrows) {
            dev_free(d_slanes);
            dev_free(kstress);
            dev_free(srows);
            dev_free(spart);
            d_slanes = nullptr;
            kstress = nullptr;
            spart = nullptr;
            if (nrows) d_slanes = (StressLane<T>*)dev_alloc(sizeof(StressLane<T>) * nrows);
            srows = (double*)dev_alloc(sizeof(double) * std::max(nrows, 1) * PF_MAXPAR);
            if (!fused) {
                schunk = (int)std::max<size_t>(1, std::min<size_t>(nrows, (size_t)(256u << 20) /
                                                                             (2 * nk * sizeof(cpx<T>))));
                kstress = (cpx<T>*)dev_alloc(sizeof(cpx<T>) * 2 * nk * schunk);
                spart = (double*)dev_alloc(sizeof(double) * spart_len());
            }
        }
        fill_stress_lanes(slanes, own_stress_bufs());
        if (nrows) upload(d_slanes, slanes.data(), sizeof(StressLane<T>) * nrows);
        slanes_cell_version = S->cell.version;
    }

    StressBufs own_stress_bufs() const {
        StressBufs bf;
        bf.d_slanes = d_slanes;
        bf.srows = srows;
        bf.spart = spart;
        bf.kstress = kstress;
        bf.chunk = std::max(schunk, 1);
        return bf;
    }

    // Segment of a checkpointed chain whose stress rows include slice sl: bnd[m] .. bnd[m+1]-1,
    // the last segment also its end slice.
    static int stress_seg(const SpBuf& b, int sl) {
        const int nseg = (int)b.bnd.size() - 1;
        int m = (int)(std::upper_bound(b.bnd.begin(), b.bnd.end(), sl) - b.bnd.begin()) - 1;
        return std::min(std::max(m, 0), nseg - 1);
    }

    // stress rows [r0, r0 + cnt) into srows (planar) or the k-space partial sums (3-D)
    void stress_launch(const StressBufs& bf, int r0, int cnt, hipStream_t st) {
        const System& s = *S;
        if (stress_fused()) {
            launch_stress_planar<T>(bf.d_slanes + r0, cnt, bf.srows, 2 * lds_plane, 256, st);
            return;
        }
        const int ny = s.dims[1], nz = s.dims[2];
        for (int a = r0; a < r0 + cnt; a += bf.chunk) {
            const int c = std::min(bf.chunk, r0 + cnt - a);
            launch_stress_slab<T>(bf.d_slanes + a, c, nz, lds_plane, 256, st);
            launch_stress_pencil<T>(bf.d_slanes + a, c, ny, ntiles, tile, bf.spart,
                                    2 * lds_pencil, 256, st);
        }
    }

    // every stress row with buffers `bf`, copied to the host
    void stress_run(System& s, const StressBufs& bf, std::vector<double>& rows) {
        rows.assign((size_t)nrows * PF_MAXPAR, 0.0);
        if (!nrows) return;
        stress_queue(s, bf, rows.data());
        PF_HIP(hipStreamSynchronize(s.stream));
    }

    // the stress launches for buffers `bf` and the copy of the rows to host memory `dst`
    // (nrows * PF_MAXPAR doubles), queued on the system's stream
    void stress_queue(System& s, const StressBufs& bf, double* dst) {
        hipStream_t st = s.stream;
        if (ckK) {
            // checkpointed sweeps: each segment of q and q+ is recomputed from its
            // checkpoints (as for Simpson) and its rows are taken before the next one
            for (size_t i = 0; i < s.species.size(); ++i) {
                if (s.species[i].kind != 0) continue;
                const SpBuf& b = sb[i];
                const int nseg = (int)b.bnd.size() - 1;
                for (int m = 0; m < nseg; ++m) {
                    run_segment(b.jq, m, false);
                    run_segment(b.jc, m, false);
                    const int lo = b.bnd[m], hi = (m == nseg - 1) ? b.bnd[m + 1] : b.bnd[m + 1] - 1;
                    stress_launch(bf, b.row0 + lo - 1, hi - lo + 1, st);
                }
            }
        } else {
            stress_launch(bf, 0, nrows, st);
        }
        if (!stress_fused())
            launch_segsum(bf.spart, nrows, s.dims[1] * ntiles, PF_MAXPAR, PF_MAXPAR, bf.srows, st);
        PF_HIP(hipMemcpyAsync(dst, bf.srows, sizeof(double) * nrows * PF_MAXPAR,
                              hipMemcpyDeviceToHost, st));
    }

    double* hrows = nullptr;         // pinned copy of the stress rows (queued passes)
    size_t hrows_len = 0;
    void stress_rows_launch(System& s) override {
        if (s.dirty_species || sb.size() != s.species.size()) build();
        if (slanes_cell_version != s.cell.version || (int)slanes.size() != nrows) build_stress();
        if (!nrows) return;
        const size_t len = (size_t)nrows * PF_MAXPAR;
        if (len > hrows_len) {
            PF_HIP(hipStreamSynchronize(s.stream));
            if (hrows) PF_HIP(hipHostFree(hrows));
            hrows = nullptr;
            PF_HIP(hipHostMalloc((void**)&hrows, sizeof(double) * len, hipHostMallocDefault));
            hrows_len = len;
        }
        stress_queue(s, own_stress_bufs(), hrows);
    }
    void stress_rows_collect(System& s, std::vector<double>& rows) override {
        PF_HIP(hipStreamSynchronize(s.stream));
        rows.assign((size_t)nrows * PF_MAXPAR, 0.0);
        if (nrows) std::copy(hrows, hrows + (size_t)nrows * PF_MAXPAR, rows.begin());
    }

    void stress_rows(System& s, std::vector<double>& rows) override {
        if (s.dirty_species || sb.size() != s.species.size()) build();
        if (slanes_cell_version != s.cell.version || (int)slanes.size() != nrows) build_stress();
        stress_run(s, own_stress_bufs(), rows);
    }

    // The same rows with every buffer carved from `scr` (cap bytes of device memory the
    // caller owns): nothing is allocated on the device.  The propagator tables must exist
    // already (a density! since the species last changed) and the stacks must be full.
    void stress_rows_scratch(System& s, std::vector<double>& rows, char* scr,
                             size_t cap) override {
        if (s.dirty_species || sb.size() != s.species.size())
            throw Error(-2, "the system has no propagator tables yet: run density! once "
                            "before the allocation-free stress");
        if (ckK)
            throw Error(-2, "the allocation-free stress needs full q stacks, not checkpoints");
        auto take = [&](size_t& off, size_t bytes) {
            char* p = scr + off;
            off += (bytes + 255) & ~(size_t)255;
            return p;
        };
        size_t off = 0;
        StressBufs bf;
        bf.d_slanes = (StressLane<T>*)take(off, sizeof(StressLane<T>) * std::max(nrows, 1));
        bf.srows = (double*)take(off, sizeof(double) * std::max(nrows, 1) * PF_MAXPAR);
        size_t need = off;
        if (!stress_fused()) {
            bf.spart = (double*)take(off, sizeof(double) * spart_len());
            const size_t per = 2 * nk * sizeof(cpx<T>);
            need = off + per;
            if (need <= cap)
                bf.chunk = (int)std::max<size_t>(1, std::min<size_t>(nrows, (cap - off) / per));
            bf.kstress = (cpx<T>*)(scr + off);
        }
        if (need > cap)
            throw Error(-1, "the allocation-free stress needs " + std::to_string(need) +
                                " bytes of scratch for this grid and contour; the library "
                                "holds " + std::to_string(cap));
        fill_stress_lanes(scratch_lanes, bf);
        if (nrows)
            PF_HIP(hipMemcpyAsync(bf.d_slanes, scratch_lanes.data(),
                                  sizeof(StressLane<T>) * nrows, hipMemcpyHostToDevice,
                                  s.stream));
        stress_run(s, bf, rows);
    }
    std::vector<StressLane<T>> scratch_lanes;

    // Reverse-sweep launches: the compile-time kernels when the size has them (and the
    // schedule does not ask for the generic ones), else the generic ones.  apper records
    // the pencil workgroups per lane, i.e. the rows of k-space partial sums.
    int apper = 0;
    void adj_slab(const Lane<T>* d, int nl, int mode, int t, int nextB = 1) {
        const System& s = *S;
        if (!fixed || !fixed_adj_slab<T>(d, nl, s.dims[0], s.dims[1], s.dims[2], mode, t,
                                         s.stream, nextB))
            launch_adj_slab<T>(d, nl, s.dims[2], mode, t, lds_plane, s.stream, nextB);
    }
    void adj_pencil(const Lane<T>* d, int nl, int t) {
        const System& s = *S;
        const int pp = fixed ? fixed_adj_pencil<T>(d, nl, s.dims[0], s.dims[1], s.dims[2], t,
                                                   s.stream)
                             : 0;
        if (!pp) launch_adj_pencil<T>(d, nl, s.dims[1], ntiles, tile, t, 2 * lds_pencil, s.stream);
        apper = pp ? pp : s.dims[1] * ntiles;
    }

    // Reverse sweeps on checkpointed stacks: each adjoint lane walks the segments in its
    // own order (q-bar from the top, q+-bar from the bottom); before a segment, q and q+
    // of that segment are recomputed from their checkpoints.  At a segment edge the
    // next step's B spectrum is made by a separate launch once its input slice exists.
    void adjoint_ckpt(System& s) {
        const size_t n = s.n;
        hipStream_t st = s.stream;
        std::vector<Lane<T>> aseg;
        std::vector<int> aoff(alanes.size(), 0);
        for (size_t j = 0; j < alanes.size(); ++j) {
            const SpBuf& b = sb[lane_sp[j]];
            aoff[j] = (int)aseg.size();
            for (size_t m = 0; m + 1 < b.bnd.size(); ++m) {
                Lane<T> A = alanes[j];
                const ptrdiff_t sh = (ptrdiff_t)(b.bnd[m] - 1) * (ptrdiff_t)n;
                A.fwd = (A.dir > 0 ? b.q : b.qc) - sh;
                A.other = (A.dir > 0 ? b.qc : b.q) - sh;
                aseg.push_back(A);
            }
        }
        auto* d_aseg = (Lane<T>*)dev_alloc(sizeof(Lane<T>) * aseg.size());
        upload(d_aseg, aseg.data(), sizeof(Lane<T>) * aseg.size());
        for (size_t i = 0; i < s.species.size(); ++i) {
            if (s.species[i].kind != 0) continue;
            const SpBuf& b = sb[i];
            const int nseg = (int)b.bnd.size() - 1;
            for (int which = 0; which < 2; ++which) {
                const int j = which == 0 ? b.jq : b.jc;
                const int NS = lanes[j].nsteps;
                for (int k = 0; k < nseg; ++k) {
                    const int m = which == 0 ? nseg - 1 - k : k;
                    run_segment(b.jq, m, false);
                    run_segment(b.jc, m, false);
                    const int lo = b.bnd[m], hi = b.bnd[m + 1];
                    const int t0 = which == 0 ? NS + 1 - hi : lo - 1;
                    const int t1 = which == 0 ? NS + 1 - lo : hi - 1;
                    const Lane<T>* dA = d_aseg + aoff[j] + m;
                    adj_slab(dA, 1, k == 0 ? 0 : 2, t0);
                    for (int t = t0; t < t1; ++t) {
                        adj_pencil(dA, 1, t);
                        adj_slab(dA, 1, 1, t, (t + 1 < t1 || k == nseg - 1) ? 1 : 0);
                    }
                }
            }
        }
        PF_HIP(hipStreamSynchronize(st));
        dev_free(d_aseg);
    }

    void block_densities(System& s, const std::vector<double*>& dst) override {
        if (tasks.empty()) return;
        if (ckK) {
            simpson_ckpt(s, &dst);
            PF_HIP(hipStreamSynchronize(s.stream));
            return;
        }
        std::vector<SimpsonTask<T>> tk = tasks;
        for (auto& t : tk) t.rho_blk = dst[t.isp];
        auto* d = (SimpsonTask<T>*)dev_alloc(sizeof(SimpsonTask<T>) * tk.size());
        upload(d, tk.data(), sizeof(SimpsonTask<T>) * tk.size());
        launch_simpson<T>(d, (int)tk.size(), s.n, s.stream);
        PF_HIP(hipStreamSynchronize(s.stream));
        dev_free(d);
    }

    // Reverse sweeps of every chain species (kernels in pf_adjoint.hip).  Requires the
    // stacks and LW tables of the density() call just made.
    void adjoint(System& s, AdjointIO& io) override {
        if (s.dirty_species || sb.size() != s.species.size())
            throw Error(-2, "density must run before the reverse sweeps");
        const int nl = (int)lanes.size();
        const size_t n = s.n;
        const int np = (int)s.cell.params.size();
        const int per = s.dims[1] * ntiles;
        constexpr int NG = PF_MAXPAR + 1;
        hipStream_t st = s.stream;
        io.dtheta.assign(np, 0.0);
        io.dds_ld.assign(s.species.size(), {});
        io.wdot.assign(s.species.size(), {});
        for (size_t i = 0; i < s.species.size(); ++i) {
            io.dds_ld[i].assign(s.species[i].nblk(), 0.0);
            io.wdot[i].assign(s.species[i].nblk(), 0.0);
        }
        if (!nl) return;
        if ((int)alanes.size() != nl) {
            free_adjoint();
            alanes.assign(nl, Lane<T>{});
            for (int j = 0; j < nl; ++j) {
                const int nb = lanes[j].nblk;
                abuf.push_back(dev_alloc(sizeof(cpx<T>) * nk));
                abuf.push_back(dev_alloc(sizeof(cpx<T>) * nk));
                abuf.push_back(dev_alloc(sizeof(double) * n * nb));
                abuf.push_back(dev_alloc(sizeof(double) * (size_t)per * nb * NG));
            }
            d_alanes = (Lane<T>*)dev_alloc(sizeof(Lane<T>) * nl);
        }
        for (int j = 0; j < nl; ++j) {
            Lane<T> A = lanes[j];
            const int i = lane_sp[j];
            const Species& sp = s.species[i];
            const bool homo = sb[i].qc == nullptr;
            A.kbuf = (cpx<T>*)abuf[4 * j];
            A.kb2 = (cpx<T>*)abuf[4 * j + 1];
            A.wbar = (double*)abuf[4 * j + 2];
            A.gpart = (double*)abuf[4 * j + 3];
            A.fwd = lanes[j].stack;
            A.other = (A.dir > 0 && !homo) ? sb[i].qc : sb[i].q;
            A.other_rev = homo ? 1 : 0;
            A.rbar = io.rbar;
            const double Q = s.Q[i], phi = s.phi[i];
            for (int k = 0; k < sp.nblk(); ++k) {
                const int slot = sp.block_slot[k];
                A.blk_slot[k] = slot;
                A.srcc[k] = phi * sp.ds_block[k] / (3.0 * Q * sp.Nref) * (homo ? 2.0 : 1.0);
                A.wco[k] = -(sp.ds_block[k] / 2.0) * s.mon_vol[slot];
            }
            A.qinit = A.dir > 0 ? io.qtot[i] / (double)n : 0.0;
            A.nparams = np;
            A.dim = s.cell.dim;
            std::memset(A.dGG, 0, sizeof(A.dGG));
            for (int k = 0; k < 9 * np; ++k) A.dGG[k] = s.cell.dGG[k];
            std::memcpy(A.G, s.cell.G, sizeof(A.G));
            alanes[j] = A;
            launch_fill<double>(A.wbar, 0.0, n * sp.nblk(), st);
            launch_fill<double>(A.gpart, 0.0, (size_t)per * sp.nblk() * NG, st);
        }
        upload(d_alanes, alanes.data(), sizeof(Lane<T>) * nl);
        apper = per;
        if (!ckK) {
            adj_slab(d_alanes, nl, 0, 0);
            for (int t = 0; t < maxNs; ++t) {
                adj_pencil(d_alanes, nl, t);
                adj_slab(d_alanes, nl, 1, t);
            }
        } else {
            adjoint_ckpt(s);
        }
        // pool the per-lane, per-block field derivatives; <w, wbar_b> for the ds terms
        DotArgs da{};
        std::vector<std::pair<int, int>> who;
        std::vector<double> dots;
        const int nparts = std::min(reduce_parts(n), 256);
        auto flush = [&] {
            if (!da.k) return;
            launch_dots(da, n, s.g_part, nparts, s.g_out, st);
            std::vector<double> h(da.k);
            PF_HIP(hipMemcpyAsync(h.data(), s.g_out, sizeof(double) * da.k,
                                  hipMemcpyDeviceToHost, st));
            PF_HIP(hipStreamSynchronize(st));
            dots.insert(dots.end(), h.begin(), h.end());
            da.k = 0;
        };
        for (int j = 0; j < nl; ++j) {
            const Species& sp = s.species[lane_sp[j]];
            for (int k = 0; k < sp.nblk(); ++k) {
                const int slot = sp.block_slot[k];
                const double* wb = alanes[j].wbar + (size_t)k * n;
                launch_axpy(io.dw + (size_t)slot * n, wb, 1.0, n, st);
                da.x[da.k] = s.w + (size_t)slot * n;
                da.y[da.k] = wb;
                ++da.k;
                who.push_back({lane_sp[j], k});
                if (da.k == 48) flush();
            }
        }
        flush();
        for (size_t q = 0; q < who.size(); ++q) io.wdot[who[q].first][who[q].second] += dots[q];
        // k-space sums: fixed-order host reduction over the pencil workgroups
        for (int j = 0; j < nl; ++j) {
            const int i = lane_sp[j];
            const Species& sp = s.species[i];
            const int nb = sp.nblk();
            std::vector<double> h((size_t)apper * nb * NG);
            PF_HIP(hipMemcpyAsync(h.data(), alanes[j].gpart, sizeof(double) * h.size(),
                                  hipMemcpyDeviceToHost, st));
            PF_HIP(hipStreamSynchronize(st));
            for (int k = 0; k < nb; ++k) {
                double G[NG] = {0};
                for (int r = 0; r < apper; ++r)
                    for (int g = 0; g < NG; ++g) G[g] += h[((size_t)r * nb + k) * NG + g];
                const double ldc = lanes[j].ldc[k];
                for (int p = 0; p < np; ++p) io.dtheta[p] += -(ldc / (double)n) * G[p];
                io.dds_ld[i][k] += -(ldc / (sp.ds_block[k] * (double)n)) * G[PF_MAXPAR];
            }
        }
    }

    void get_propagator(System& s, int isp, int slice, int backward, double* host) override {
        if (ckK) throw Error(-2, "propagator stacks are not kept with checkpointed sweeps");
        if (sb.size() != s.species.size() || s.species[isp].kind != 0)
            throw Error(-2, "no propagator for this species (run density first)");
        const Species& sp = s.species[isp];
        if (slice < 1 || slice > sp.Ns + 1) throw Error(-2, "contour index out of range");
        const T* src;
        if (!backward)
            src = sb[isp].q + (size_t)(slice - 1) * s.n;
        else if (sb[isp].qc)
            src = sb[isp].qc + (size_t)(slice - 1) * s.n;
        else
            src = sb[isp].q + (size_t)(sp.Ns + 1 - slice) * s.n;
        std::vector<T> tmp(s.n);
        PF_HIP(hipMemcpyAsync(tmp.data(), src, sizeof(T) * s.n, hipMemcpyDeviceToHost, s.stream));
        PF_HIP(hipStreamSynchronize(s.stream));
        for (size_t i = 0; i < s.n; ++i) host[i] = (double)tmp[i];
    }
};

// ---------------------------------------------------------------------------------
// Batch: lanes of every active system in one set of launches.  Lanes are grouped by grid
// size; planar groups of different sizes run concurrently on their own streams.
template <typename T>
class BatchEngine : public BatchEngineBase {
    Batch* B;
    struct Group {
        int nx, ny, nz, off, count, maxNs, ntiles, tile, lds_plane, lds_pencil;
        int mde;   // 0 RK2, 1 RQM4 (lanes are grouped by grid size and method)
        int rq_lds;   // LDS of an RQM4 slab or planar workgroup (one or two planes)
        hipStream_t st = nullptr;
        hipEvent_t done = nullptr;
    };
    std::vector<Group> groups;
    std::vector<Lane<T>> lanes;
    Lane<T>* d_lanes = nullptr;
    std::vector<SimpsonTask<T>> tasks;
    SimpsonTask<T>* d_tasks = nullptr;
    std::vector<LwTask<T>> lwt;
    LwTask<T>* d_lwt = nullptr;
    std::vector<FillTask<T>> fillt;
    FillTask<T>* d_fillt = nullptr;
    struct PointTask {
        const double* w;
        double vol;
        size_t n;
        double* rho_sp;
        double* partial;
        int nparts;
        DevScalars* sc;
        int isp;
    };
    std::vector<PointTask> points;
    double* partial = nullptr;
    size_t partial_len = 0;
    std::vector<int> cached;
    std::vector<long> cell_versions;
    hipEvent_t fork = nullptr;
    std::vector<std::pair<hipStream_t, hipEvent_t>> streams;   // one per (sub-)group
    // 3-D groups of at least 3 lanes run as 2 sub-groups on 2 streams by default: measured
    // on MI355X (64^3 FP32, 32 systems) 2 streams give +20% over 1, 3 no further gain
    static constexpr int split = 2;
    size_t nmax = 0;
    bool tim = false;
    std::vector<hipEvent_t> evpool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evlog;
    size_t evnext = 0;

    hipEvent_t ev() {
        if (evnext >= evpool.size()) {
            hipEvent_t e;
            PF_HIP(hipEventCreate(&e));
            evpool.push_back(e);
        }
        return evpool[evnext++];
    }
    size_t ncalls[3] = {0, 0, 0};
    // events bracket one launch in 8 of each kind (recording every launch would add a
    // gap between dependent kernels); the sample still spans the whole timed region.  A
    // sampled launch runs kTimingReps times back to back between its events, which spreads
    // the events' own overhead; the repeats redo the launch on its own output, so the
    // sweep's results are not meaningful in timing mode.
    std::vector<int> evlanes;        // lanes covered by each sampled launch
    template <typename F>
    void timed(int cls, hipStream_t s, int nl, F&& f) {
        if (!tim || (ncalls[cls]++ % 8) != 0) {
            f();
            return;
        }
        hipEvent_t a = ev(), b = ev();
        PF_HIP(hipEventRecord(a, s));
        for (int r = 0; r < kTimingReps; ++r) f();
        PF_HIP(hipEventRecord(b, s));
        evlog.push_back({cls, {a, b}});
        evlanes.push_back(nl);
    }

    Engine<T>& E(int i) { return *static_cast<Engine<T>*>(B->sys[i]->eng.get()); }

    template <typename V>
    V* upload_vec(V* old, const std::vector<V>& v) {
        dev_free(old);
        if (v.empty()) return nullptr;
        V* d = (V*)dev_alloc(sizeof(V) * v.size());
        PF_HIP(hipMemcpyAsync(d, v.data(), sizeof(V) * v.size(), hipMemcpyHostToDevice, B->stream));
        PF_HIP(hipStreamSynchronize(B->stream));
        return d;
    }

    void rebuild(const std::vector<int>& act) {
        std::vector<std::vector<Lane<T>>> by_size;
        std::vector<Group> g;
        tasks.clear();
        lwt.clear();
        fillt.clear();
        points.clear();
        nmax = 0;
        ck.clear();
        for (int i : act) {
            System& s = *B->sys[i];
            if (s.ckpt) {
                // checkpointed stacks: the system's own segment-wise density! (below)
                ck.push_back(i);
                continue;
            }
            Engine<T>& e = E(i);
            e.prepare();
            nmax = std::max(nmax, s.n);
            int gi = -1;
            for (size_t k = 0; k < g.size(); ++k)
                if (g[k].nx == s.dims[0] && g[k].ny == s.dims[1] && g[k].nz == s.dims[2] &&
                    g[k].mde == s.mde)
                    gi = (int)k;
            if (gi < 0) {
                Group ng{};
                ng.nx = s.dims[0];
                ng.ny = s.dims[1];
                ng.nz = s.dims[2];
                ng.mde = s.mde;
                ng.ntiles = e.tiles();
                ng.tile = e.tile_w();
                ng.lds_plane = e.plane_lds();
                ng.rq_lds = e.rqm4_lds();
                ng.lds_pencil = e.pencil_lds();
                g.push_back(ng);
                by_size.emplace_back();
                gi = (int)g.size() - 1;
            }
            for (const auto& L : e.host_lanes()) {
                by_size[gi].push_back(L);
                g[gi].maxNs = std::max(g[gi].maxNs, L.nsteps);
            }
            for (const auto& tk : e.host_tasks()) tasks.push_back(tk);
            for (size_t k = 0; k < s.species.size(); ++k) {
                const Species& sp = s.species[k];
                if (sp.kind != 0) {
                    // point species (src/species/point.jl:36-53): Q and density per system
                    const int slot = sp.slots[0];
                    points.push_back({s.w + (size_t)slot * s.n, s.mon_vol[slot], s.n, sp.rho_sp,
                                      s.partial, reduce_parts(s.n), s.sc, (int)k});
                    continue;
                }
                for (int b = 0; b < sp.nblk(); ++b) {
                    const int slot = sp.block_slot[b];
                    const double coef = (-(sp.ds_block[b] / 2.0)) * s.mon_vol[slot];
                    lwt.push_back({e.lw_table((int)k, b), s.w + (size_t)slot * s.n, coef, s.n});
                    // RQM4 half-step table exp(coef w / 2) (mde.jl:51): coef / 2 scales
                    // exactly, so the product rounds as the one-system table's does
                    if (s.mde == 1)
                        lwt.push_back({e.lw_half_table((int)k, b), s.w + (size_t)slot * s.n,
                                       coef / 2.0, s.n});
                }
                fillt.push_back({e.q_stack((int)k), s.n});
                if (e.qc_stack((int)k)) fillt.push_back({e.qc_stack((int)k) + (size_t)sp.Ns * s.n, s.n});
            }
        }
        // The groups are launched largest first (grid points x contour steps): they run
        // concurrently, and the longest one, which the batch waits for, then starts first.
        std::vector<size_t> order(g.size());
        for (size_t k = 0; k < g.size(); ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            return (double)g[a].nx * g[a].ny * g[a].nz * g[a].maxNs >
                   (double)g[b].nx * g[b].ny * g[b].nz * g[b].maxNs;
        });
        // 3-D groups are split into `split` sub-groups on streams of their own, so that
        // the slab launches of one sub-group overlap the pencil launches of another
        lanes.clear();
        std::vector<Group> fin;
        for (size_t k : order) {
            const int cnt = (int)by_size[k].size();
            const int sp = schedule(sizeof(T), g[k].nx, g[k].ny, g[k].nz).split;
            const int want = sp > 0 ? sp : split;
            // (a group of three lanes splits too: a 128^3 blend batched alone, FP64
            // 18.48 -> 16.43 ms per density!; two lanes split measured slower)
            const int parts = (g[k].nz > 1 && cnt >= 3)
                                  ? std::max(1, std::min(want, std::max(2, cnt / 2)))
                                  : 1;
            int done = 0;
            for (int p = 0; p < parts; ++p) {
                const int c = (cnt - done) / (parts - p);
                Group sg = g[k];
                sg.off = (int)lanes.size();
                sg.count = c;
                sg.maxNs = 0;
                for (int j = done; j < done + c; ++j) {
                    lanes.push_back(by_size[k][j]);
                    sg.maxNs = std::max(sg.maxNs, by_size[k][j].nsteps);
                }
                done += c;
                fin.push_back(sg);
            }
        }
        while (streams.size() < fin.size()) {
            hipStream_t st;
            hipEvent_t e;
            PF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            PF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            streams.push_back({st, e});
        }
        for (size_t k = 0; k < fin.size(); ++k) {
            fin[k].st = streams[k].first;
            fin[k].done = streams[k].second;
        }
        groups = fin;
        d_lanes = upload_vec(d_lanes, lanes);
        d_tasks = upload_vec(d_tasks, tasks);
        d_lwt = upload_vec(d_lwt, lwt);
        d_fillt = upload_vec(d_fillt, fillt);
        const size_t need = tasks.size() * (size_t)reduce_parts(nmax) + 1;
        if (need > partial_len) {
            dev_free(partial);
            partial = (double*)dev_alloc(sizeof(double) * need);
            partial_len = need;
        }
        cached = act;
        cell_versions.clear();
        for (int i : act) cell_versions.push_back(B->sys[i]->cell.version);
    }

   public:
    std::vector<int> ck;             // checkpointed systems of the prepared active set
    hipEvent_t ck_before = nullptr, ck_after = nullptr;
    explicit BatchEngine(Batch& b) : B(&b) {
        PF_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        PF_HIP(hipEventCreateWithFlags(&ck_before, hipEventDisableTiming));
        PF_HIP(hipEventCreateWithFlags(&ck_after, hipEventDisableTiming));
    }
    ~BatchEngine() override {
        dev_free(d_lanes);
        dev_free(d_tasks);
        dev_free(d_lwt);
        dev_free(d_fillt);
        dev_free(partial);
        for (auto& se : streams) {
            (void)hipStreamDestroy(se.first);
            (void)hipEventDestroy(se.second);
        }
        (void)hipEventDestroy(fork);
        (void)hipEventDestroy(ck_before);
        (void)hipEventDestroy(ck_after);
        for (auto e : evpool) (void)hipEventDestroy(e);
    }

    void timing(int on) override {
        tim = on != 0;
        evlog.clear();
        evlanes.clear();
        evnext = 0;
        ncalls[0] = ncalls[1] = ncalls[2] = 0;
    }

    double timing_lanes(int which) override {
        double tot = 0.0;
        int c = 0;
        for (size_t k = 0; k < evlog.size(); ++k)
            if (evlog[k].first == which) {
                tot += evlanes[k];
                ++c;
            }
        return c ? tot / c : 0.0;
    }

    double timing_ms(int which, int* count) override {
        PF_HIP(hipStreamSynchronize(B->stream));
        double tot = 0.0;
        int c = 0;
        for (auto& e : evlog)
            if (e.first == which) {
                float ms = 0.f;
                PF_HIP(hipEventElapsedTime(&ms, e.second.first, e.second.second));
                tot += ms;
                c += kTimingReps;
            }
        if (count) *count = c;
        return tot;
    }

    long gen = 0;

    void prepare(const std::vector<int>& act) override {
        bool stale = act != cached;
        for (size_t k = 0; !stale && k < act.size(); ++k)
            stale = B->sys[act[k]]->cell.version != cell_versions[k] || B->sys[act[k]]->dirty_species;
        if (stale) {
            rebuild(act);
            ++gen;
        }
    }

    long generation() const override { return gen; }
    void invalidate() override { cached.clear(); }
    bool timing_on() const override { return tim; }
    bool capturable() const override { return ck.empty(); }

    // density! of the checkpointed systems of the active set, each through its own engine
    // on its own stream (segments recomputed from their checkpoints, src/species/
    // multiblock.jl:125-231 in that order), ordered after the batch's earlier work and
    // before its later launches
    void launch_ckpt() {
        for (int i : ck) {
            System& s = *B->sys[i];
            PF_HIP(hipEventRecord(ck_before, B->stream));
            PF_HIP(hipStreamWaitEvent(s.stream, ck_before, 0));
            E(i).density(s);
            PF_HIP(hipEventRecord(ck_after, s.stream));
            PF_HIP(hipStreamWaitEvent(B->stream, ck_after, 0));
        }
    }

    void launch(const std::vector<int>& act) override {
        run_sweeps();
        hipStream_t st = B->stream;
        const int nparts = reduce_parts(nmax);
        if (!tasks.empty())
            launch_species_Q<T>(d_tasks, (int)tasks.size(), nmax, partial, nparts, st);
        for (const PointTask& p : points)
            launch_point_density(p.w, p.vol, p.n, p.rho_sp, p.partial, p.nparts, p.sc, p.isp, st);
        if (!tasks.empty()) launch_simpson<T>(d_tasks, (int)tasks.size(), nmax, st);
        launch_ckpt();
    }

    // a checkpointed system keeps no full stacks to sweep into: it runs its whole density!
    void launch_sweeps(const std::vector<int>&) override {
        run_sweeps();
        launch_ckpt();
    }

    void run_sweeps() {
        hipStream_t st = B->stream;
        if (lwt.empty() && fillt.empty()) return;   // checkpointed systems only
        launch_lw_fill_b<T>(d_lwt, (int)lwt.size(), d_fillt, (int)fillt.size(), nmax, st);
        PF_HIP(hipEventRecord(fork, st));
        hipEvent_t prev = fork;
        for (auto& gr : groups) {
            // with timing on, the groups run one after another, so that every timed
            // launch has the device to itself (the figure a kernel trace reports)
            PF_HIP(hipStreamWaitEvent(gr.st, tim ? prev : fork, 0));
            const Lane<T>* dl = d_lanes + gr.off;
            if (gr.nz == 1 && gr.mde == 1) {
                // RQM4 planes (mde.jl:83-104): the compile-time kernel where both planes fit
                // in LDS and the size is specialised, else the generic one
                timed(2, gr.st, gr.count, [&] {
                    if (!fixed_planar_rqm4<T>(dl, gr.count, gr.nx, gr.ny, 0, gr.maxNs,
                                              gr.rq_lds != 2 * gr.lds_plane, gr.st))
                        launch_planar_steps<T>(dl, gr.count, 0, gr.maxNs, gr.rq_lds, 256, gr.st);
                });
            } else if (gr.nz == 1) {
                timed(2, gr.st, gr.count, [&] {
                    if (!fixed_planar<T>(dl, gr.count, gr.nx, gr.ny, 0, gr.maxNs, gr.st))
                        launch_planar_steps<T>(dl, gr.count, 0, gr.maxNs, gr.lds_plane, 256, gr.st);
                });
            } else if (gr.mde == 1) {
                // RQM4 in 3-D: per contour step the full- and half-step spectra through the
                // pencils and three slab launches, as the one-system engine runs them
                rqm4_sweep<T>(dl, gr.count, gr.nx, gr.ny, gr.nz, 0, gr.maxNs,
                              gr.rq_lds == 2 * gr.lds_plane, true, gr.ntiles, gr.tile,
                              gr.lds_pencil, gr.rq_lds, gr.st);
            } else {
                timed(0, gr.st, gr.count, [&] {
                    if (!fixed_slab<T>(dl, gr.count, gr.nx, gr.ny, gr.nz, SLAB_FWD, 0, gr.st))
                        launch_slab<T>(dl, gr.count, gr.nz, SLAB_FWD, 0, gr.lds_plane, 256, gr.st);
                });
                for (int t = 0; t < gr.maxNs; ++t) {
                    const int nla = live_lanes(lanes, gr.off, gr.count, t);
                    timed(1, gr.st, nla, [&] {
                        if (!fixed_pencil<T>(dl, nla, gr.nx, gr.ny, gr.nz, false, t, gr.st))
                            launch_pencil<T>(dl, nla, gr.ny, gr.ntiles, gr.tile, t, gr.lds_pencil, 256, gr.st);
                    });
                    timed(0, gr.st, nla, [&] {
                        if (!fixed_slab<T>(dl, nla, gr.nx, gr.ny, gr.nz, SLAB_INVFWD, t, gr.st))
                            launch_slab<T>(dl, nla, gr.nz, SLAB_INVFWD, t, gr.lds_plane, 256, gr.st);
                    });
                }
            }
            PF_HIP(hipEventRecord(gr.done, gr.st));
            PF_HIP(hipStreamWaitEvent(st, gr.done, 0));
            prev = gr.done;
        }
    }
};

std::unique_ptr<BatchEngineBase> make_batch_engine(Batch& b) {
    if (b.precision == 0) return std::unique_ptr<BatchEngineBase>(new BatchEngine<double>(b));
    return std::unique_ptr<BatchEngineBase>(new BatchEngine<float>(b));
}

std::unique_ptr<EngineBase> make_engine(System& s) {
    if (s.precision == 0) return std::unique_ptr<EngineBase>(new Engine<double>(s));
    return std::unique_ptr<EngineBase>(new Engine<float>(s));
}

}  // namespace pf
