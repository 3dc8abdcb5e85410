// FieldSystem operations, field/cell updaters and the SCFT driver (host logic; every
// grid operation is a kernel on the system's stream).
//
// Follows src/system/system.jl, src/updaters/{anderson,domain,euler,momentum}.jl and
// src/simulate.jl.  Differences in work, never in results:
//   * Anderson history is a ring of nhist+1 slots instead of a copy-shift
//     (src/updaters/anderson.jl:228-244 moves every slot down by one).
//   * DomainUpdater skips two density! calls whose outputs the reference overwrites
//     before reading them: the one after restoring the parameters in dstress!
//     (src/updaters/domain.jl:117-118) and the one just before residuals!, which
//     recomputes the same densities (src/updaters/domain.jl:83-84).
#include <chrono>
#include <cmath>
#include <cstring>

#include "pf_host.h"

namespace pf {

void* dev_alloc(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) return nullptr;
    PF_HIP(hipMalloc(&p, bytes));
    return p;
}
void dev_free(void* p) {
    if (p) (void)hipFree(p);
}

static const double EPS = 2.220446049250313e-16;

static bool approx(double a, double b) {
    return a == b || std::fabs(a - b) <= std::sqrt(EPS) * std::max(std::fabs(a), std::fabs(b));
}

static bool radix23(int n) {
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

System::System(int nx, int ny, int nz, int dim_, const std::string& crystal,
               const std::vector<double>& params, int precision_, int canonical_,
               double inv_zeta_, int device_)
    : precision(precision_), canonical(canonical_), inv_zeta(inv_zeta_), device(device_) {
    dims[0] = nx;
    dims[1] = ny;
    dims[2] = nz;
    if (nx < ny || ny < nz || nz < 1)
        throw Error(-2, "dims must be given in descending order (the reference sorts them for "
                        "its grids but plans its FFT on the unsorted tuple)");
    dim = (nx > 1) + (ny > 1) + (nz > 1);
    if (dim != dim_)
        throw Error(-2, "Number of non-singleton dimensions in System and Cell do not match!");
    if (nx % 2 != 0) throw Error(-2, "the x grid size must be even");
    if (!radix23(nx) || !radix23(ny) || !radix23(nz))
        throw Error(-2, "grid sizes must be products of 2 and 3");
    if (nx / 2 > 2048 || ny > 2048 || nz > 2048) throw Error(-2, "grid size too large");
    if (precision != 0 && precision != 1) throw Error(-2, "precision must be 0 (fp64) or 1 (fp32)");
    compress_mode = (inv_zeta == 0.0) ? 0 : (inv_zeta > 0.0 ? 1 : 2);
    n = (size_t)nx * ny * nz;
    cell.dim = dim_;
    cell.crystal = crystal;
    cell.params = params;
    cell.update();
    PF_HIP(hipSetDevice(device));
    PF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    sc = (DevScalars*)dev_alloc(sizeof(DevScalars));
    PF_HIP(hipMemsetAsync(sc, 0, sizeof(DevScalars), stream));
    partial_len = 512 * 160;
    partial = (double*)dev_alloc(sizeof(double) * partial_len);
    sums = (double*)dev_alloc(sizeof(double) * 256);
    PF_HIP(hipHostMalloc((void**)&pinned, sizeof(double) * 4096, hipHostMallocDefault));
    eng = make_engine(*this);
}

System::~System() {
    eng.reset();
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& sp : species) dev_free(sp.rho_sp);
    dev_free(w);
    dev_free(pot);
    dev_free(res);
    dev_free(rho);
    dev_free(sc);
    dev_free(partial);
    dev_free(sums);
    dev_free(g_rbar);
    dev_free(g_dw);
    dev_free(g_blk);
    dev_free(g_part);
    dev_free(g_out);
    if (pinned) (void)hipHostFree(pinned);
    if (err_ev) (void)hipEventDestroy(err_ev);
    if (stream) (void)hipStreamDestroy(stream);
}

int System::slot_of(int id) const {
    for (int i = 0; i < nmon(); ++i)
        if (mon_id[i] == id) return i;
    return -1;
}

void System::sync() { PF_HIP(hipStreamSynchronize(stream)); }

void System::set_precision(int p) {
    if (p != 0 && p != 1) throw Error(-2, "precision must be 0 (fp64) or 1 (fp32)");
    if (p == precision) return;
    sync();
    eng.reset();
    precision = p;
    eng = make_engine(*this);
    dirty_species = 1;
}

void System::set_block_lengths(int isp, const std::vector<double>& N_block) {
    // Continuous block lengths on a fixed contour grid: every block keeps its step count
    // (the begin indices of src/species/base.jl:27-56) and ds_b = v_b N_b / steps_b.  This
    // is the argument the block-length derivative of the free energy is taken against.
    if (isp < 0 || isp >= (int)species.size()) throw Error(-2, "species index out of range");
    Species& sp = species[isp];
    if (sp.kind != 0) throw Error(-2, "block lengths are defined for chain species only");
    if ((int)N_block.size() != sp.nblk()) throw Error(-2, "wrong number of block lengths");
    double Nref = 0.0, N = 0.0;
    for (int b = 0; b < sp.nblk(); ++b) {
        if (!(N_block[b] > 0.0)) throw Error(-2, "block lengths must be positive");
        sp.Nref_block[b] = sp.vol_block[b] * N_block[b];
        sp.ds_block[b] = sp.Nref_block[b] / (double)(sp.begin[b + 1] - sp.begin[b]);
        Nref += sp.Nref_block[b];
        N += N_block[b];
    }
    sp.Nref = Nref;
    sp.N = N;
    for (int b = 0; b < sp.nblk(); ++b) sp.f_block[b] = sp.Nref_block[b] / Nref;
    dirty_species = 1;
}

void System::add_monomer(int id, double vol, double charge) {
    if (!species.empty()) throw Error(-2, "monomers must be added before species");
    if (vol <= 0.0) throw Error(-2, "monomer volume must be positive");
    int s = slot_of(id);
    if (s >= 0) {   // replacement (src/system/system.jl:114-116 warns and replaces)
        mon_vol[s] = vol;
        mon_charge[s] = charge;
        return;
    }
    if (nmon() >= PF_MAXMON) throw Error(-2, "too many monomer types");
    mon_order.push_back(id);
    size_t pos = 0;
    while (pos < mon_id.size() && mon_id[pos] < id) ++pos;
    mon_id.insert(mon_id.begin() + pos, id);
    mon_vol.insert(mon_vol.begin() + pos, vol);
    mon_charge.insert(mon_charge.begin() + pos, charge);
    dev_free(w);
    dev_free(pot);
    dev_free(res);
    dev_free(rho);
    const size_t bytes = sizeof(double) * n * nmon();
    w = (double*)dev_alloc(bytes);
    pot = (double*)dev_alloc(bytes);
    res = (double*)dev_alloc(bytes);
    rho = (double*)dev_alloc(bytes);
    for (double* p : {w, pot, res, rho}) PF_HIP(hipMemsetAsync(p, 0, bytes, stream));
    fields_set = false;
}

int System::add_chain(const std::vector<int>& mids, const std::vector<int>& N_block,
                      const std::vector<double>& b_block, int Ns, double val) {
    const int nb = (int)mids.size();
    if (nb < 1 || nb > PF_MAXBLK || (int)N_block.size() != nb || (int)b_block.size() != nb)
        throw Error(-2, "a chain needs 1..8 blocks with matching N and b");
    if ((int)species.size() >= 64) throw Error(-2, "too many species");
    Species sp;
    sp.kind = 0;
    for (int k = 0; k < nb; ++k) {
        const int s = slot_of(mids[k]);
        if (s < 0)
            throw Error(-2, "Species contains monomer with id = " + std::to_string(mids[k]) +
                                " not present in system.");
        if (N_block[k] < 1) throw Error(-2, "block lengths must be positive");
        sp.block_slot.push_back(s);
        sp.N_block.push_back(N_block[k]);
        sp.b_block.push_back(b_block[k]);
        sp.vol_block.push_back(mon_vol[s]);
        sp.Nref_block.push_back(mon_vol[s] * N_block[k]);
    }
    if (Ns < 1) throw Error(-2, "Ns must be positive");
    chain_grid(sp.Nref_block, Ns, &sp.Nref, sp.begin, sp.ds_block);
    sp.N = 0;
    for (int x : N_block) sp.N += x;
    for (int k = 0; k < nb; ++k) sp.f_block.push_back(sp.Nref_block[k] / sp.Nref);
    sp.Ns = sp.begin[nb] - 1;   // Ns adjusted as in src/species/multiblock.jl:55-59
    for (int s : sp.block_slot)
        if (std::find(sp.slots.begin(), sp.slots.end(), s) == sp.slots.end()) sp.slots.push_back(s);
    std::sort(sp.slots.begin(), sp.slots.end());
    if (canonical && val <= 0.0) throw Error(-2, "Bulk volume fraction must be 0 < phi <= 1.");
    sp.rho_sp = (double*)dev_alloc(sizeof(double) * n * sp.slots.size());
    PF_HIP(hipMemsetAsync(sp.rho_sp, 0, sizeof(double) * n * sp.slots.size(), stream));
    species.push_back(sp);
    phi.push_back(canonical ? val : 0.0);
    mu.push_back(canonical ? 0.0 : val);
    Q.push_back(0.0);
    dirty_species = 1;
    upload_scalars();
    return (int)species.size() - 1;
}

int System::add_point(int mid, double val) {
    const int s = slot_of(mid);
    if (s < 0)
        throw Error(-2, "Species contains monomer with id = " + std::to_string(mid) +
                            " not present in system.");
    if ((int)species.size() >= 64) throw Error(-2, "too many species");
    if (canonical && val <= 0.0) throw Error(-2, "Bulk volume fraction must be 0 < phi <= 1.");
    Species sp;
    sp.kind = 1;
    sp.slots = {s};
    sp.Nref = mon_vol[s];   // src/species/point.jl:19-21
    sp.rho_sp = (double*)dev_alloc(sizeof(double) * n);
    PF_HIP(hipMemsetAsync(sp.rho_sp, 0, sizeof(double) * n, stream));
    species.push_back(sp);
    phi.push_back(canonical ? val : 0.0);
    mu.push_back(canonical ? 0.0 : val);
    Q.push_back(0.0);
    dirty_species = 1;
    upload_scalars();
    return (int)species.size() - 1;
}

void System::set_flory(int a, int b, double c) {
    if (a == b) throw Error(-2, "a Flory pair needs two different monomers");
    const int sa = slot_of(a), sb = slot_of(b);
    if (sa < 0 || sb < 0)
        throw Error(-2, "Interaction contains monomer not present in system.");
    chi[sa * PF_MAXMON + sb] = chi[sb * PF_MAXMON + sa] = c;
    pair[sa * PF_MAXMON + sb] = pair[sb * PF_MAXMON + sa] = 1;
    has_flory = 1;
}

void System::set_edwards(int a, double uu) {
    const int s = slot_of(a);
    if (s < 0) throw Error(-2, "Interaction contains monomer not present in system.");
    u[s] = uu;
    u_set[s] = 1;
}

void System::upload_scalars() {
    DevScalars h{};
    for (size_t i = 0; i < species.size(); ++i) {
        h.phi[i] = phi[i];
        h.mu[i] = mu[i];
        h.Q[i] = Q[i];
    }
    std::memcpy(pinned, &h, sizeof(DevScalars));
    PF_HIP(hipMemcpyAsync(sc, pinned, sizeof(DevScalars), hipMemcpyHostToDevice, stream));
    sync();
}

void System::fetch_scalars() {
    PF_HIP(hipMemcpyAsync(pinned, sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
    sync();
    const DevScalars* h = reinterpret_cast<const DevScalars*>(pinned);
    for (size_t i = 0; i < species.size(); ++i) {
        Q[i] = h->Q[i];
        phi[i] = h->phi[i];
        mu[i] = h->mu[i];
    }
}

ResidualArgs System::residual_args() const {
    ResidualArgs a{};
    a.nmon = nmon();
    for (int m = 0; m < nmon(); ++m) {
        a.vol[m] = mon_vol[m];
        a.u[m] = u[m];
        a.u_set[m] = u_set[m];
    }
    a.has_flory = has_flory;
    std::memcpy(a.chi, chi, sizeof(chi));
    std::memcpy(a.pair, pair, sizeof(pair));
    a.compress = compress_mode;
    a.inv_zeta = inv_zeta;
    a.canonical = canonical;
    a.w = w;
    a.pot = pot;
    a.res = res;
    a.rho = rho;
    return a;
}

void System::monomer_fractions(std::vector<double>& fr) {
    fr.assign(nmon(), 0.0);
    for (size_t i = 0; i < species.size(); ++i) {
        const Species& sp = species[i];
        for (int s : sp.slots) {
            double f = 0.0;
            if (sp.kind == 1)
                f = 1.0;
            else
                for (int k = 0; k < sp.nblk(); ++k)
                    if (sp.block_slot[k] == s) f += sp.f_block[k];
            fr[s] += phi[i] * f;
        }
    }
}

void System::validate() {
    // src/system/system.jl:192-237
    if (species.empty() || nmon() == 0)
        throw Error(-2, "Fields must be initialized before SCFT simulation.");
    std::vector<double> hw(n * nmon());
    PF_HIP(hipMemcpyAsync(hw.data(), w, sizeof(double) * hw.size(), hipMemcpyDeviceToHost, stream));
    sync();
    bool init = false;
    for (int m = 0; m < nmon() && !init; ++m) {
        double s = 0.0;
        for (size_t i = 0; i < n; ++i) s += hw[m * n + i];
        const double mean = s / n;
        double v = 0.0;
        for (size_t i = 0; i < n; ++i) v += (hw[m * n + i] - mean) * (hw[m * n + i] - mean);
        if (mean != 0.0 || v != 0.0) init = true;
    }
    if (!init) throw Error(-2, "Fields must be initialized before SCFT simulation.");
    if (canonical) {
        double s = 0.0;
        for (double p : phi) s += p;
        if (!approx(s, 1.0)) throw Error(-2, "Species volume fractions do not sum to unity.");
    }
    std::vector<double> fr;
    monomer_fractions(fr);
    double charge = 0.0;
    for (int m = 0; m < nmon(); ++m) charge += fr[m] * mon_charge[m] / mon_vol[m];
    if (!approx(charge, 0.0)) throw Error(-2, "Bulk system not charge neutral.");
    for (int m = 0; m < nmon(); ++m) {
        bool found = false;
        for (auto& sp : species)
            if (std::find(sp.slots.begin(), sp.slots.end(), m) != sp.slots.end()) found = true;
        if (!found)
            throw Error(-2, "Monomer with ID = " + std::to_string(mon_id[m]) +
                                " not found in any species.");
    }
}

void System::density() {
    // src/system/system.jl:266-289
    PF_HIP(hipSetDevice(device));
    eng->density(*this);
    launch_muphi(sc, (int)species.size(), canonical, stream);
    PoolArgs a{};
    a.nmon = nmon();
    a.dst = rho;
    for (size_t i = 0; i < species.size(); ++i) {
        const Species& sp = species[i];
        for (size_t j = 0; j < sp.slots.size(); ++j) {
            const int m = sp.slots[j];
            if (a.nterm[m] >= 8) throw Error(-2, "too many species share one monomer type");
            a.src[m][a.nterm[m]] = sp.rho_sp + j * n;
            a.isp[m][a.nterm[m]] = (int)i;
            a.nterm[m]++;
        }
    }
    launch_pool(a, n, sc, stream);
}

void System::residuals() {
    // src/system/system.jl:317-340
    if (res_queued) {
        res_queued = false;
        return;
    }
    density();
    launch_residuals(residual_args(), n, partial, reduce_parts(n), sums, stream);
}

void System::uniform_fields() {
    // src/system/system.jl:348-357: w = potentials(rho); canonical: w -= mean(w).  The
    // residual kernel computes the potentials from the current densities (its eta/res
    // outputs are discarded and recomputed by the residuals! that follows a seeding).
    ResidualArgs a = residual_args();
    a.canonical = 0;
    launch_residuals(a, n, partial, reduce_parts(n), sums, stream);
    std::vector<double> hp(n * nmon());
    PF_HIP(hipMemcpyAsync(hp.data(), pot, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, stream));
    sync();
    if (canonical)
        for (int m = 0; m < nmon(); ++m) {
            double s = 0.0;
            for (size_t i = 0; i < n; ++i) s += hp[m * n + i];
            const double mean = s / (double)n;
            for (size_t i = 0; i < n; ++i) hp[m * n + i] -= mean;
        }
    PF_HIP(hipMemcpyAsync(w, hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice, stream));
    sync();
    fields_set = true;
}

// The five system sums at sums[32..36] and the muphi! domain flag at sums[37].
static void launch_sums(System& s) {
    launch_sys_reduce(s.residual_args(), s.n, s.partial, reduce_parts(s.n), s.sums + 32, s.stream);
    launch_domain_flag(s.sc, (int)s.species.size(), s.canonical, s.sums + 37, s.stream);
    PF_HIP(hipMemcpyAsync(s.pinned, s.sums + 32, sizeof(double) * 6, hipMemcpyDeviceToHost,
                          s.stream));
}

static void check_domain(const double* v) {
    if (v[5] != 0.0)
        throw Error(-5, "DomainError: log(phi / Q) of a negative partition-function ratio in "
                        "muphi! (the propagator's Q turned negative: the fields diverged)");
}

void System::sys_sums(double out[5]) {
    launch_sums(*this);
    sync();
    check_domain(pinned);
    for (int k = 0; k < 5; ++k) out[k] = pinned[k];
}

double System::field_error() {
    double v[5];
    sys_sums(v);
    return std::sqrt(v[0]) / std::sqrt(EPS + v[1]);
}

double System::free_energy() {
    double v[5];
    sys_sums(v);
    return free_energy_of(v);
}

void System::errors(double* ferr, double* fhelm) {
    errors_launch();
    errors_read(ferr, fhelm);
}

void System::errors_launch() {
    if (!err_ev) PF_HIP(hipEventCreateWithFlags(&err_ev, hipEventDisableTiming));
    launch_sums(*this);
    PF_HIP(hipEventRecord(err_ev, stream));
}

void System::errors_read(double* ferr, double* fhelm) {
    PF_HIP(hipEventSynchronize(err_ev));
    check_domain(pinned);
    double v[5];
    for (int k = 0; k < 5; ++k) v[k] = pinned[k];
    *ferr = std::sqrt(v[0]) / std::sqrt(EPS + v[1]);
    *fhelm = free_energy_of(v);
}

double System::free_energy_of(const double v[5]) {
    // src/system/system.jl:402-433 (mu = phi, as in the reference)
    if (!canonical) fetch_scalars();
    double ftrans = 0.0;
    for (size_t i = 0; i < species.size(); ++i) {
        const double p = phi[i], m = phi[i];
        if (p > 1e-8) ftrans += p * (m - 1.0) / species[i].Nref;
    }
    const double ffield = -v[2] / (double)n;
    const double fexc = v[3] / (double)n + v[4] / (double)n;
    return ftrans + ffield + fexc;
}

double System::free_energy_bulk() {
    // src/system/system.jl:440-458
    if (!canonical) fetch_scalars();
    double f = 0.0;
    for (size_t i = 0; i < species.size(); ++i)
        if (phi[i] > 1e-8) f += phi[i] * (std::log(phi[i]) - 1.0) / species[i].Nref;
    std::vector<double> fr;
    monomer_fractions(fr);
    for (int a = 0; a < nmon(); ++a)
        for (int b = a + 1; b < nmon(); ++b)
            if (pair[a * PF_MAXMON + b])
                f += chi[a * PF_MAXMON + b] * (fr[a] / mon_vol[a]) * (fr[b] / mon_vol[b]);
    for (int a = 0; a < nmon(); ++a)
        if (u_set[a]) f += (u[a] / 2) * ((fr[a] / mon_vol[a]) * (fr[a] / mon_vol[a]));
    return f;
}

// src/system/system.jl:384-395 with src/species/multiblock.jl:233-295: the per-slice rows
// of the k-space sums from the engine (its own buffers, or [scr, scr + cap) when scr is
// given), then the Simpson-weighted combination per species on the host
std::vector<double> System::stress_rows_combined(char* scr, size_t cap) {
    std::vector<double> rows;
    launch_muphi(sc, (int)species.size(), canonical, stream);
    fetch_scalars();
    if (scr)
        eng->stress_rows_scratch(*this, rows, scr, cap);
    else
        eng->stress_rows(*this, rows);
    return stress_combine(rows);
}

void System::stress_launch() {
    PF_HIP(hipSetDevice(device));
    launch_muphi(sc, (int)species.size(), canonical, stream);
    PF_HIP(hipMemcpyAsync(pinned, sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
    eng->stress_rows_launch(*this);
}

std::vector<double> System::stress_collect() {
    std::vector<double> rows;
    eng->stress_rows_collect(*this, rows);   // waits for the stream, so for the scalars too
    const DevScalars* h = reinterpret_cast<const DevScalars*>(pinned);
    for (size_t i = 0; i < species.size(); ++i) {
        Q[i] = h->Q[i];
        phi[i] = h->phi[i];
        mu[i] = h->mu[i];
    }
    return stress_combine(rows);
}

std::vector<double> System::scfstress(int exact) {
    if (exact) throw Error(-2, "only the reference stress formula is provided");
    return stress_rows_combined(nullptr, 0);
}

std::vector<double> System::scfstress_scratch(char* scr, size_t cap) {
    return stress_rows_combined(scr, cap);
}

std::vector<double> System::stress_combine(const std::vector<double>& rows) const {
    const int np = (int)cell.params.size();
    std::vector<double> stress(np, 0.0);
    const double n2 = (double)n * (double)n;
    int row0 = 0;
    for (size_t i = 0; i < species.size(); ++i) {
        const Species& sp = species[i];
        if (sp.kind != 0) continue;
        std::vector<double> dQ(np, 0.0);
        for (int b = 0; b < sp.nblk(); ++b) {
            const int bgn = sp.begin[b], end = sp.begin[b + 1];
            const double bb = sp.b_block[b], ds0 = sp.ds_block[b];
            for (int s = bgn; s <= end; ++s) {
                double ds = ds0;
                if (s != bgn && s != end) ds *= (s % 2 == 0) ? 4.0 : 2.0;
                const double* r = &rows[(size_t)(row0 + s - 1) * PF_MAXPAR];
                for (int k = 0; k < np; ++k) dQ[k] -= bb * bb * ds * r[k] / 6.0 / n2;
            }
        }
        for (int k = 0; k < np; ++k) {
            dQ[k] /= 3.0;
            dQ[k] /= Q[i] * sp.Nref;
            stress[k] -= phi[i] * dQ[k];
        }
        row0 += sp.Ns + 1;
    }
    return stress;
}

// ---------------------------------------------------------------------------------
// Anderson mixing (src/updaters/anderson.jl)

Anderson::Anderson(int nh, int ie, double le) : nhist(nh), iter_euler(ie <= 1 ? 1 : ie), lam_euler(le) {
    if (nhist < 1) throw Error(-2, "nhist must be at least 1");
}

static void free_deep(Anderson& a) {
    dev_free(a.gpart);
    dev_free(a.gout);
    dev_free(a.dcoef);
    if (a.hgram) (void)hipHostFree(a.hgram);
    if (a.hcoef) (void)hipHostFree(a.hcoef);
    a.gpart = a.gout = a.dcoef = a.hgram = a.hcoef = nullptr;
}

Anderson::~Anderson() {
    dev_free(hist_r);
    dev_free(hist_w);
    free_deep(*this);
}

HistRing Anderson::ring(int ih) const {
    HistRing g;
    g.r = hist_r;
    g.w = hist_w;
    g.N = sys->n * sys->nmon();
    g.nslots = nhist + 1;
    g.start = base + ih - 1;
    return g;
}

double* Anderson::rslot(int logical) const {
    const size_t N = sys->n * sys->nmon();
    return hist_r + (size_t)((base + logical) % (nhist + 1)) * N;
}
double* Anderson::wslot(int logical) const {
    const size_t N = sys->n * sys->nmon();
    return hist_w + (size_t)((base + logical) % (nhist + 1)) * N;
}

void Anderson::setup(System* s) {
    if (is_setup && s == sys) return;   // keeps its history (anderson.jl:57-59)
    if (!s->canonical)
        throw Error(-2, "Anderson mixing can only be used for systems in the canonical ensemble.");
    sys = s;
    dev_free(hist_r);
    dev_free(hist_w);
    const size_t N = s->n * s->nmon();
    hist_r = (double*)dev_alloc(sizeof(double) * N * (nhist + 1));
    hist_w = (double*)dev_alloc(sizeof(double) * N * (nhist + 1));
    PF_HIP(hipMemsetAsync(hist_r, 0, sizeof(double) * N * (nhist + 1), s->stream));
    PF_HIP(hipMemsetAsync(hist_w, 0, sizeof(double) * N * (nhist + 1), s->stream));
    base = 0;
    launch_copy(wslot(0), s->w, N, s->stream);
    free_deep(*this);
    if (nhist > PF_MAXHIST) {
        const int nt = gram_tiles(nhist);
        gparts = gram_tiled_parts(N);
        gpart = (double*)dev_alloc(sizeof(double) * (size_t)gparts * nt * 64);
        gout = (double*)dev_alloc(sizeof(double) * (size_t)nt * 64);
        dcoef = (double*)dev_alloc(sizeof(double) * nhist);
        PF_HIP(hipHostMalloc((void**)&hgram, sizeof(double) * (size_t)nt * 64, hipHostMallocDefault));
        PF_HIP(hipHostMalloc((void**)&hcoef, sizeof(double) * nhist, hipHostMallocDefault));
    }
    iter = 0;
    iter_hist = 0;
    lam = 0.0;
    is_setup = true;
}

void Anderson::step() {
    System& s = *sys;
    const size_t N = s.n * s.nmon();
    iter += 1;
    if (iter < nhist + 1) {
        iter_hist = iter - 1;
        lam = 1.0 - std::pow(0.9, (double)iter);
    } else {
        iter_hist = nhist;
        lam = 1.0;
    }
    const int ih = iter_hist;
    s.residuals();
    launch_copy(rslot(ih), s.res, N, s.stream);
    const bool euler = iter <= iter_euler;
    if (euler) {
        launch_axpy(s.w, s.res, lam_euler, N, s.stream);
    } else {
        const bool deep = nhist > PF_MAXHIST;
        lastU.assign(ih * ih, 0.0);
        lastV.assign(ih, 0.0);
        if (!deep) {
            HistPtrs hp{};
            for (int m = 0; m < ih; ++m) hp.p[m] = rslot(ih - (m + 1));
            const int nu = ih * (ih + 1) / 2;
            launch_gram(ih, s.res, hp, N, s.partial, reduce_parts(N), s.sums + 64, s.stream);
            PF_HIP(hipMemcpyAsync(s.pinned, s.sums + 64, sizeof(double) * (nu + ih),
                                  hipMemcpyDeviceToHost, s.stream));
            s.sync();
            int k = 0;
            for (int m = 0; m < ih; ++m)
                for (int nn = m; nn < ih; ++nn) {
                    lastU[m * ih + nn] = lastU[nn * ih + m] = s.pinned[k++];
                }
            for (int m = 0; m < ih; ++m) lastV[m] = s.pinned[nu + m];
        } else {
            const HistRing g = ring(ih);
            launch_gram_tiled(ih, s.res, g, gpart, gout, s.stream);
            PF_HIP(hipMemcpyAsync(hgram, gout, sizeof(double) * (size_t)gram_tiles(ih) * 64,
                                  hipMemcpyDeviceToHost, s.stream));
            s.sync();
            for (int m = 0; m < ih; ++m) {
                for (int nn = m; nn < ih; ++nn)
                    lastU[m * ih + nn] = lastU[nn * ih + m] = hgram[gram_tile_index(m, nn)];
                lastV[m] = hgram[gram_tile_index(m, ih)];
            }
        }
        std::vector<double> inv(ih * ih);
        if (!lu_inverse(ih, lastU.data(), inv.data()))
            throw Error(-4, "Error inverting Umn matrix on iteration " + std::to_string(iter) +
                                ": singular matrix");
        lastC.assign(ih, 0.0);
        matvec(ih, inv.data(), lastV.data(), lastC.data());
        double* f_out = (iter < nhist + 1) ? wslot(iter) : wslot(0);
        if (!deep) {
            HistPtrs fh{}, rh{};
            MixCoef c{};
            for (int nn = 0; nn < ih; ++nn) {
                fh.p[nn] = wslot(ih - (nn + 1));
                rh.p[nn] = rslot(ih - (nn + 1));
                c.C[nn] = lastC[nn];
            }
            launch_anderson_mix(ih, s.w, s.res, fh, rh, wslot(ih), c, lam, N, f_out, s.stream);
        } else {
            // the previous step's copy of hcoef finished before the Gram read above
            for (int nn = 0; nn < ih; ++nn) hcoef[nn] = lastC[nn];
            PF_HIP(hipMemcpyAsync(dcoef, hcoef, sizeof(double) * ih, hipMemcpyHostToDevice,
                                  s.stream));
            launch_anderson_mix_ring(ih, s.w, s.res, ring(ih), wslot(ih), dcoef, lam, f_out,
                                     s.stream);
        }
    }
    if (iter < nhist + 1) {
        if (euler) launch_copy(wslot(iter), s.w, N, s.stream);
    } else {
        base = (base + 1) % (nhist + 1);
        if (euler) launch_copy(wslot(nhist), s.w, N, s.stream);
    }
}

// Euler predictor-corrector (src/updaters/euler.jl:36-66)
Euler::~Euler() { dev_free(tmp); }
void Euler::setup(System* s) {
    sys = s;
    dev_free(tmp);
    tmp = (double*)dev_alloc(sizeof(double) * s->n * s->nmon());
    launch_copy(tmp, s->w, s->n * s->nmon(), s->stream);
}
void Euler::step() {
    System& s = *sys;
    const size_t N = s.n * s.nmon();
    launch_axpy(s.w, s.res, lam, N, s.stream);
    launch_copy(tmp, s.w, N, s.stream);
    launch_axpy(tmp, s.res, 0.5 * lam, N, s.stream);
    s.residuals();
    launch_axpy(tmp, s.res, 0.5 * lam, N, s.stream);
    launch_copy(s.w, tmp, N, s.stream);
    if (pece) s.residuals();
}

// Momentum family (src/updaters/momentum.jl:47-165)
Momentum::~Momentum() {
    dev_free(stepb);
    dev_free(mt);
    dev_free(vt);
    dev_free(tmp);
}
void Momentum::setup(System* s) {
    sys = s;
    const size_t N = s->n * s->nmon();
    for (double** p : {&stepb, &mt, &vt, &tmp}) {
        dev_free(*p);
        *p = (double*)dev_alloc(sizeof(double) * N);
    }
    launch_copy(stepb, s->res, N, s->stream);
    launch_copy(tmp, s->w, N, s->stream);
    PF_HIP(hipMemsetAsync(mt, 0, sizeof(double) * N, s->stream));
    PF_HIP(hipMemsetAsync(vt, 0, sizeof(double) * N, s->stream));
}
void Momentum::step() {
    System& s = *sys;
    const size_t N = s.n * s.nmon();
    if (method == 3) {   // NAG
        launch_copy(tmp, s.w, N, s.stream);
        launch_axpy(s.w, mt, lam, N, s.stream);
        s.residuals();
        launch_momentum(0, mt, vt, stepb, s.res, gamma, beta2, N, s.stream);
        launch_copy(s.w, tmp, N, s.stream);
    } else {
        s.residuals();
        launch_momentum(method, mt, vt, stepb, s.res, beta1, beta2, N, s.stream);
    }
    launch_axpy(s.w, stepb, lam, N, s.stream);
}

// ---------------------------------------------------------------------------------
// Variable cell (src/updaters/domain.jl)

static double norm2(const std::vector<double>& v) {
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

void Domain::setup(System* s) {
    sys = s;
    const int np = (int)s->cell.params.size();
    stress = s->scfstress();
    oldparams.assign(np, 0.0);
    stepv.assign(np, 0.0);
    dstress.assign(np * np, 0.0);
}

void Domain::scfstress_update() { stress = sys->scfstress(); }

double Domain::stress_error() const { return norm2(stress); }

void Domain::dstress_update() {
    System& s = *sys;
    Cell& cell = s.cell;
    const int np = (int)cell.params.size();
    oldparams = cell.params;
    const double dx = fd_step, x = 1.0 / dx;
    for (int k = 0; k < np; ++k) {
        cell.params = oldparams;
        cell.params[k] += dx;
        cell.update();
        s.density();
        std::vector<double> ps = s.scfstress();
        for (int i = 0; i < np; ++i) dstress[i * np + k] = x * (ps[i] - stress[i]);
    }
    cell.params = oldparams;
}

void Domain::newton_update() {
    Cell& cell = sys->cell;
    const int np = (int)cell.params.size();
    if (cell.dim == 1) {
        for (int k = 0; k < np; ++k) stepv[k] = -1.0 * stress[0] / dstress[0];
    } else {
        std::vector<double> inv(np * np);
        if (!lu_inverse(np, dstress.data(), inv.data()))
            throw Error(-4, "singular cell Jacobian");
        matvec(np, inv.data(), stress.data(), stepv.data());
        for (auto& v : stepv) v *= -1.0;
    }
    const double size = norm2(stepv);
    if (size > pmax)
        for (auto& v : stepv) v *= pmax / size;
    for (int k = 0; k < np; ++k) cell.params[k] += lam * stepv[k];
    for (int k = 0; k < np; ++k)
        cell.params[k] = rlxn * cell.params[k] + (1 - rlxn) * oldparams[k];
    cell.update();
}

void Domain::step(int nsteps, double /*tol: unused, as in the reference*/) {
    scfstress_update();
    for (int it = 0; it < nsteps; ++it) {
        dstress_update();
        newton_update();
        sys->residuals();
        scfstress_update();
    }
}

// ---------------------------------------------------------------------------------
bool scft_converged(double ferr, double ftol, double serr, double stol, bool vcell) {
    // src/simulate.jl:37-49
    bool f = false, sc = true;
    if (ferr < ftol) {
        f = true;
        if (vcell && serr > stol) sc = false;
    }
    return f && sc;
}

ScftResults scft(System& s, Updater& u, Domain* d, const ScftOptions& o,
                 const std::function<void(int)>& save) {
    // src/simulate.jl:60-148
    const auto t0 = std::chrono::steady_clock::now();
    s.validate();
    s.residuals();
    u.setup(&s);
    int step = 0;
    double ferr, fhelm;
    s.errors(&ferr, &fhelm);
    double serr = 0.0;
    const bool vcell = d != nullptr;
    if (vcell) {
        d->setup(&s);
        serr = d->stress_error();
    }
    bool conv = scft_converged(ferr, o.ftol, serr, o.stol, vcell);
    // The next step's residuals! are queued before the host reads this step's errors
    // when nothing can change the fields or the cell in between: an updater whose step
    // starts with residuals! (Anderson, and Momentum other than NAG), a fixed cell, and the canonical ensemble (the grand free
    // energy reads device scalars that the queued pass would overwrite).  A converged
    // system then takes the queued pass as its final residuals!.
    const bool queue = u.starts_with_residuals() && !vcell && s.canonical;
    struct Unqueue {
        System& s;
        ~Unqueue() { s.res_queued = false; }
    } unqueue{s};
    while (step < o.nsteps && !conv) {
        if (o.nsout > 0 && step % o.nsout == 0)
            std::printf("Step (%-5d / %5d): ferr = %8e, fhelm = %.5e\n", step, o.nsteps, ferr, fhelm);
        u.step();
        if (vcell && step % d->nskip == 0) {
            d->scfstress_update();
            serr = d->stress_error();
            if (serr > o.stol) {
                d->step(d->nper, o.stol);
                serr = d->stress_error();
                if (o.nsout > 0) std::printf("Updating cell dimensions: serr = %.5e\n", serr);
            }
        }
        // a save writes the fields as the update left them, so no residuals! (which
        // mean-subtracts the fields) is queued ahead of it
        const bool save_due = save && o.nsave > 0 && step % o.nsave == 0;
        s.errors_launch();
        if (queue && !save_due && step + 1 < o.nsteps) {
            s.residuals();
            s.res_queued = true;
        }
        s.errors_read(&ferr, &fhelm);
        if (std::isnan(ferr)) throw Error(-5, "Diverging field updater -- breaking.");
        conv = scft_converged(ferr, o.ftol, serr, o.stol, vcell);
        if (conv) break;
        if (save_due) save(step);
        step += 1;
    }
    s.residuals();
    s.errors(&ferr, &fhelm);
    if (vcell) serr = d->stress_error();
    if (save) save(-1);
    ScftResults r;
    r.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    r.nsteps = step;
    r.converged = conv;
    r.ferr = ferr;
    r.serr = serr;
    r.fhelm = fhelm;
    return r;
}

}  // namespace pf
