// Host-side state of a device-resident FieldSystem and its updaters.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "pf_internal.h"

namespace pf {

// --- geometry (pf_cell.cpp) ---------------------------------------------------------
struct Cell {
    int dim = 1;
    std::string crystal;
    std::vector<double> params;
    double R[9] = {0}, G[9] = {0}, volume = 0.0;
    std::vector<double> dR, dG, dRR, dGG;   // [(i*3+j)*np + k]
    long version = 0;                      // bumped by update()
    void basis();
    void dbasis();
    void ddbasis();
    void update();                         // src/system/cell.jl:96-109 (basis part)
};
int crystal_nparams(int dim, const std::string& crystal);
void chain_grid(const std::vector<double>& N_block, int Ns, double* N_out,
                std::vector<int>& begin, std::vector<double>& ds_block);
bool lu_inverse(int n, const double* U, double* inv);
void matvec(int n, const double* M, const double* v, double* out);

// --- species ------------------------------------------------------------------------
struct Species {
    int kind = 0;                    // 0 linear chain (Multiblock), 1 point particle
    std::vector<int> block_slot;     // monomer slot per block
    std::vector<int> N_block;
    std::vector<double> b_block, vol_block, Nref_block, ds_block, f_block;
    std::vector<int> begin;          // 1-based block starts, nblk+1 entries
    int Ns = 0;
    double N = 0, Nref = 0;
    std::vector<int> slots;          // monomer slots present (ascending id)
    int nblk() const { return (int)N_block.size(); }
    double* rho_sp = nullptr;        // device [slots.size()][n]
};

class EngineBase;

// A FieldSystem with all grids resident on one GPU (src/system/system.jl:22-49).
struct System {
    int dims[3];
    size_t n = 0;
    int dim = 0;
    int precision = 0;               // 0: FP64 everywhere, 1: FP32 propagator + FP64 sums
    int mde = 0;                     // 0 RK2, 1 RQM4 (src/system/mde.jl:18-30)
    int ckpt = 0;                    // checkpointed sweeps: segment length (-1: sqrt(Ns+1)), 0 off
    int canonical = 1;
    int compress_mode = 0;           // 0 strict, 1 compressible, 2 unconstrained
    double inv_zeta = 0.0;
    int device = 0;
    Cell cell;

    std::vector<int> mon_id;         // ascending monomer ids; slot = index
    std::vector<int> mon_order;      // monomer ids in the order they were added
    std::vector<double> mon_vol, mon_charge;
    int has_flory = 0;
    double chi[PF_MAXMON * PF_MAXMON] = {0};
    int pair[PF_MAXMON * PF_MAXMON] = {0};
    double u[PF_MAXMON] = {0};
    int u_set[PF_MAXMON] = {0};

    std::vector<Species> species;
    std::vector<double> phi, mu;     // host mirrors (canonical phi fixed, grand mu fixed)
    std::vector<double> Q;           // host mirror, refreshed by fetch_scalars()

    // device grids, FP64, [slot][n]
    double* w = nullptr;
    double* pot = nullptr;
    double* res = nullptr;
    double* rho = nullptr;
    DevScalars* sc = nullptr;
    double* partial = nullptr;       // reduction scratch
    size_t partial_len = 0;
    double* sums = nullptr;          // small result area (64 doubles)
    double* pinned = nullptr;        // pinned host staging (4096 doubles)
    hipStream_t stream = nullptr;
    hipEvent_t err_ev = nullptr;     // marks the copy of the five system sums to the host
    // residuals! of the current fields is already queued (scft! queues the next step's
    // before it reads the errors); the next residuals() call only clears the flag
    bool res_queued = false;

    std::unique_ptr<EngineBase> eng;
    bool fields_set = false;
    int dirty_species = 1;           // engine must rebuild lanes

    System(int nx, int ny, int nz, int dim_, const std::string& crystal,
           const std::vector<double>& params, int precision, int canonical, double inv_zeta,
           int device);
    ~System();

    int nmon() const { return (int)mon_id.size(); }
    int slot_of(int id) const;
    void add_monomer(int id, double vol, double charge);
    int add_chain(const std::vector<int>& mids, const std::vector<int>& N_block,
                  const std::vector<double>& b_block, int Ns, double val);
    int add_point(int mid, double val);
    void set_flory(int a, int b, double chi);
    void set_edwards(int a, double u);

    ResidualArgs residual_args() const;
    void upload_scalars();
    void fetch_scalars();            // Q, phi, mu -> host (synchronises)

    // reference operations
    void validate();
    void density();                  // density!(sys)
    void residuals();                // residuals!(sys)
    void uniform_fields();           // uniform_fields!(sys)
    void sys_sums(double out[5]);    // [res^2, w^2, rho*w, flory, edwards]
    double field_error();
    double free_energy();
    // both from one reduction (what scft! reads after every step); errors_launch queues
    // the reduction and its copy, errors_read waits for that copy only
    void errors(double* ferr, double* fhelm);
    void errors_launch();
    void errors_read(double* ferr, double* fhelm);
    double free_energy_of(const double v[5]);
    double free_energy_bulk();
    std::vector<double> scfstress(int exact = 0);
    // scfstress with its device scratch in [scr, scr + cap) (no device allocation)
    std::vector<double> scfstress_scratch(char* scr, size_t cap);
    std::vector<double> stress_rows_combined(char* scr, size_t cap);
    std::vector<double> stress_combine(const std::vector<double>& rows) const;
    // scfstress in two halves, so that many systems' stress passes run together: queue the
    // pass on this system's stream, then wait for it and combine
    void stress_launch();
    std::vector<double> stress_collect();
    void monomer_fractions(std::vector<double>& fr);
    void sync();
    void set_precision(int p);       // swap the propagator engine (fields stay FP64)
    void set_block_lengths(int isp, const std::vector<double>& N_block);

    // free-energy gradient scratch (allocated on first use)
    double* g_rbar = nullptr;        // [slot][n]
    double* g_dw = nullptr;          // [slot][n]
    double* g_blk = nullptr;         // per-block densities of all chain species
    size_t g_blk_len = 0;
    double* g_part = nullptr;        // dot partials
    double* g_out = nullptr;         // dot results
};

// Reverse-mode derivative of the free energy (pf_grad.cpp).  functional: 0 the reference's
// free_energy (mu = phi), 1 the same with mu = log(phi / Q); -1: density VJP only, with
// the cotangents rho_bar (host [slot][n]) and Q_bar (host, per species).
struct GradResult {
    double fhelm = 0.0;
    std::vector<double> dw;          // host [slot][n]
    std::vector<double> dtheta;      // [nparams]
    std::vector<double> dchi;        // [nmon * nmon], Flory pairs only
    std::vector<double> du;          // [nmon], Edwards parameters only
    std::vector<double> dphi;        // [nspecies]
    std::vector<double> dN;          // chain blocks in species order
};
void fhelm_grad(System& s, int functional, const double* rho_bar, const double* Q_bar,
                GradResult& g);

// Cotangents in and sums out of the reverse sweeps (Engine::adjoint).
struct AdjointIO {
    const double* rbar = nullptr;          // device [slot][n]: dF/d(system density)
    std::vector<double> qtot;              // per species: total dF/dQ (chains)
    double* dw = nullptr;                  // device [slot][n]: dF/dw, accumulated into
    std::vector<double> dtheta;            // out: dF/dtheta
    std::vector<std::vector<double>> dds_ld;  // out: LD part of dF/dds_b, per species/block
    std::vector<std::vector<double>> wdot;    // out: <w_mon(b), wbar_b>, per species/block
};

class EngineBase {
   public:
    // reverse sweeps of every chain species after density(); see pf_adjoint.hip
    virtual void adjoint(System& s, AdjointIO& io) = 0;
    // per-block Simpson densities of chain species into dst[isp] ([blk][n]; null: skip)
    virtual void block_densities(System& s, const std::vector<double*>& dst) = 0;
    virtual ~EngineBase() = default;
    virtual void density(System& s) = 0;
    virtual void stress_rows(System& s, std::vector<double>& rows) = 0;   // [row][PF_MAXPAR]
    // stress_rows in two halves: queue the pass and the copy of its rows to pinned host
    // memory on the system's stream; wait for it and read the rows
    virtual void stress_rows_launch(System& s) = 0;
    virtual void stress_rows_collect(System& s, std::vector<double>& rows) = 0;
    // the same rows with every device buffer carved from the caller's `cap` bytes at scr
    virtual void stress_rows_scratch(System& s, std::vector<double>& rows, char* scr,
                                     size_t cap) = 0;
    virtual void get_propagator(System& s, int isp, int slice, int backward, double* host) = 0;
    virtual void timing(int on) = 0;
    virtual double timing_ms(int which, int* count) = 0;
};
std::unique_ptr<EngineBase> make_engine(System& s);

// --- field updaters ------------------------------------------------------------------
struct Updater {
    virtual ~Updater() = default;
    virtual void setup(System* s) = 0;
    virtual void step() = 0;
    // true when step() starts with residuals! of the fields as they are (nothing changes
    // them before), so that scft! may queue that pass early
    virtual bool starts_with_residuals() const { return false; }
    System* sys = nullptr;
};

struct Anderson : Updater {
    int nhist, iter_euler;
    double lam_euler;
    int iter = 0, iter_hist = 0;
    double lam = 0.0;
    bool is_setup = false;
    double* hist_r = nullptr;        // (nhist+1) slots of nmon*n
    double* hist_w = nullptr;
    int base = 0;                    // physical slot of logical slot 0
    std::vector<double> lastU, lastV, lastC;
    // histories deeper than PF_MAXHIST: tiled Gram scratch and its pinned host copy,
    // mixing coefficients (pinned staging and device copy)
    double* gpart = nullptr;
    double* gout = nullptr;
    double* hgram = nullptr;
    double* hcoef = nullptr;
    double* dcoef = nullptr;
    int gparts = 0;
    HistRing ring(int ih) const;     // the history ring as seen by step ih
    Anderson(int nhist, int iter_euler, double lam_euler);
    ~Anderson();
    double* rslot(int logical) const;
    double* wslot(int logical) const;
    void setup(System* s) override;
    void step() override;
    bool starts_with_residuals() const override { return true; }
};

struct Euler : Updater {
    int pece;
    double lam;
    double* tmp = nullptr;
    Euler(int pece, double lam) : pece(pece), lam(lam) {}
    ~Euler();
    void setup(System* s) override;
    void step() override;
};

struct Momentum : Updater {
    int method;                      // 0 MOM, 1 RMS, 2 ADAM, 3 NAG
    double lam, gamma, beta1, beta2;
    double *stepb = nullptr, *mt = nullptr, *vt = nullptr, *tmp = nullptr;
    Momentum(int method, double lam, double gamma, double b1, double b2)
        : method(method), lam(lam), gamma(gamma), beta1(b1), beta2(b2) {}
    ~Momentum();
    void setup(System* s) override;
    void step() override;
    // NAG evaluates the residuals at a look-ahead point, the others at the fields as they are
    bool starts_with_residuals() const override { return method != 3; }
};

struct Domain {
    int nper, nskip;
    double lam, pmax, rlxn;
    double fd_step = 1e-8;           // forward-difference step of dstress! (domain.jl:101)
    System* sys = nullptr;
    std::vector<double> oldparams, stepv, stress, dstress;
    Domain(int nper, int nskip, double lam, double pmax, double rlxn)
        : nper(nper), nskip(nskip), lam(lam), pmax(pmax), rlxn(rlxn) {}
    void setup(System* s);
    void scfstress_update();
    double stress_error() const;
    void step(int nsteps, double tol);
    void dstress_update();
    // the Newton step on the cell from stress and dstress, its relaxation against
    // oldparams, and update! of the cell (src/updaters/domain.jl:64-84)
    void newton_update();
};

struct ScftOptions {
    int nsteps, nsout, nsave;
    double ftol, stol;
};
struct ScftResults {
    double time;
    int nsteps, converged;
    double ferr, serr, fhelm;
};
// save(step): called where the reference writes savefields(sys, opts.savepath) -- every
// nsave steps inside the loop (src/simulate.jl:124-128) -- and, with step = -1, once for
// the final state (the reference's own final write names a field SCFTOptions does not
// have, src/simulate.jl:140-142).  Empty: no saving.
ScftResults scft(System& s, Updater& u, Domain* d, const ScftOptions& o,
                 const std::function<void(int)>& save = {});
// isconverged (src/simulate.jl:37-49): ferr < ftol, and with a variable cell serr <= stol
bool scft_converged(double ferr, double ftol, double serr, double stol, bool vcell);

// the allocation-free entry point's device scratch on `device` (pf_noalloc.hip)
char* noalloc_scratch(int device, size_t* bytes);

void* dev_alloc(size_t bytes);
void dev_free(void* p);

}  // namespace pf
