// Batches of FieldSystems advanced together: every propagator launch covers the lanes
// of all systems, and every system-level kernel covers all systems (blockIdx.y).
#pragma once
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

#include <array>

#include "pf_host.h"

namespace pf {

template <typename T>
struct LwTask {
    T* out;
    const double* w;
    double coef;
    size_t n;
};

template <typename T>
struct FillTask {
    T* p;
    size_t n;
};

// Static per-system description (uploaded when the batch membership changes).
struct SysTask {
    ResidualArgs ra;
    PoolArgs pool;
    size_t n;
    int nparts;          // reduction workgroups for n points
    int gparts;          // reduction workgroups for nmon*n points (Gram)
    int nsp;
    DevScalars* sc;
    double* partial;     // reduction scratch of the system
    double* gpartial;
    double* sums;        // 256 doubles in the batch result table
};

// Per-iteration Anderson data of one system.
struct StepTask {
    HistRing ring;       // the history ring (logical slots ih-1 .. ih-h)
    const double* f_iter;
    double* f_out;       // history slot receiving the new field
    double* r_out;       // history slot receiving the residual
    const double* C;     // h mixing coefficients (device copy of the step table)
    double lam, lam_euler;
    int h, euler;
};

// Momentum family and Euler state of one system (device arrays of N = n * nmon doubles;
// tmp: NAG's saved fields, the Euler corrector's fields)
struct MomTask {
    double *mt, *vt, *step, *w;
    const double* r;
    size_t N;
    double* tmp;
};
void launch_momentum_b(const MomTask* d, int nsys, int method, double lam, double b1, double b2,
                       size_t Nmax, hipStream_t st);
// the pointwise stages of the updaters whose step evaluates residuals! inside it
enum { STAGE_NAG_LOOK = 0, STAGE_NAG_STEP = 1, STAGE_EULER_PRED = 2, STAGE_EULER_CORR = 3 };
void launch_stage_b(const MomTask* d, int nsys, int op, double lam, double b1, double b2,
                    size_t Nmax, hipStream_t st);

template <typename T>
void launch_lw_fill_b(const LwTask<T>* lw, int nlw, const FillTask<T>* fill, int nfill,
                      size_t nmax, hipStream_t st);
void launch_muphi_pool_b(const SysTask* d, int nsys, size_t nmax, hipStream_t st);
void launch_residuals_b(const SysTask* d, int nsys, int maxparts, size_t nmax, hipStream_t st);
void launch_sys_reduce_b(const SysTask* d, int nsys, int maxparts, double* compact, hipStream_t st);
void launch_gram_b(int h, const SysTask* d, const StepTask* s, int nsys, int maxparts,
                   double* compact, hipStream_t st);
// any depth h: tiled Gram, partial[(sys * parts + part) * tiles * 64] with `parts` the
// largest gram_tiled_parts of the systems, compact[sys][tiles * 64]
void launch_gram_tiled_b(int h, const SysTask* d, const StepTask* s, int nsys, int parts,
                         double* partial, double* compact, hipStream_t st);
void launch_mix_b(const SysTask* d, const StepTask* s, int nsys, size_t Nmax, hipStream_t st);

class BatchEngineBase {
   public:
    virtual ~BatchEngineBase() = default;
    // density! of the systems listed in `active` (indices into the batch)
    virtual void density(const std::vector<int>& active) {
        prepare(active);
        launch(active);
    }
    // rebuild the lane tables if the active set or a cell changed (host work, uploads)
    virtual void prepare(const std::vector<int>& active) = 0;
    // the launch sequence of density! for a prepared active set (stream-ordered only, so
    // it can be captured into a graph)
    virtual void launch(const std::vector<int>& active) = 0;
    // its first part: the exponential tables and the q / q+ sweeps of every lane group
    virtual void launch_sweeps(const std::vector<int>& active) = 0;
    virtual long generation() const = 0;   // bumped by every rebuild
    virtual void invalidate() = 0;          // force a rebuild (schedule change)
    virtual bool timing_on() const = 0;
    // false when the prepared active set holds a checkpointed system, whose density!
    // synchronises with the host between segments: the launch sequence then runs eagerly
    virtual bool capturable() const { return true; }
    // HIP-event timing of the propagator launches; kind: 0 slab, 1 pencil, 2 planar
    virtual void timing(int on) = 0;
    virtual double timing_ms(int kind, int* count) = 0;
    virtual double timing_lanes(int kind) = 0;   // mean lanes per sampled launch
};

struct BatchResult {
    int converged, nsteps;
    double ferr, fhelm, time;
    int diverged;
    double serr;
};

// A few host threads that split a loop over [0, n) with the calling thread (the batch's
// per-system Anderson solves).  Started with the batch, stopped with it.
class HostPool {
   public:
    explicit HostPool(int nworkers);
    ~HostPool();
    // fn(begin, end) over consecutive chunks of [0, n), one per thread; returns when all
    // are done
    void run(size_t n, const std::function<void(size_t, size_t)>& fn);
    int threads() const { return (int)th.size() + 1; }

   private:
    void worker(int id);
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable go, done;
    const std::function<void(size_t, size_t)>* job = nullptr;
    size_t count = 0;
    long gen = 0;
    int pending = 0;
    bool stop = false;
};

class Batch;
// A field updater run over a batch by Batch::scft_run.  For an updater whose step starts
// with residuals! (Anderson, MOM, RMS, ADAM) the loop runs that pass (possibly queued
// early), then update() moves the fields of every active system; NAG and the Euler
// predictor-corrector evaluate their residuals inside update().
struct BatchStepper {
    virtual ~BatchStepper() = default;
    // the grand ensemble is allowed (the reference refuses it for Anderson mixing only),
    // or the error a grand-ensemble system raises
    virtual bool allows_grand() const { return false; }
    virtual const char* noncanonical_error() const { return ""; }
    virtual bool starts_with_residuals() const { return true; }
    // setup! after the first residuals! of the systems in `run` (src/simulate.jl:86)
    virtual void setup(Batch& b, const std::vector<int>& run) = 0;
    // the rest of step! for the systems in `act`; may mark systems diverged in res
    virtual void update(Batch& b, const std::vector<int>& act, std::vector<BatchResult>& res) = 0;
};

class Batch {
   public:
    std::vector<System*> sys;
    std::unique_ptr<HostPool> pool;  // Anderson solves of large batches
    hipStream_t stream = nullptr;
    int precision = 0;
    std::unique_ptr<BatchEngineBase> eng;
    double* sums = nullptr;          // nsys * 256 device doubles
    // per active system, in active-set order, the values the host reads (the five system
    // sums, or the Gram entries): device copy and pinned mirror, nsys * ncompact doubles
    static constexpr int kCompact = PF_MAXHIST * (PF_MAXHIST + 3) / 2;
    int ncompact = kCompact;
    double* csums = nullptr;
    double* hcsums = nullptr;
    double* gpart = nullptr;         // tiled-Gram partial sums (histories above PF_MAXHIST)
    int gparts = 0;
    SysTask* d_sys = nullptr;        // all systems' descriptors
    // step table: nsys StepTasks, then nsys * coef_stride mixing coefficients; pinned
    // staging and device copy, uploaded as one block
    StepTask* d_step = nullptr;
    StepTask* h_step = nullptr;
    int coef_stride = PF_MAXHIST;
    double* h_coef() const { return reinterpret_cast<double*>(h_step + sys.size()); }
    const double* d_coef() const { return reinterpret_cast<const double*>(d_step + sys.size()); }
    void alloc_step_table(int nhist);
    std::vector<SysTask> tasks;
    std::vector<int> act_cache;      // active set the compacted descriptors were built for
    SysTask* d_act = nullptr;        // descriptors of the active systems, compacted
    std::vector<std::unique_ptr<Anderson>> and_state;
    // graphs of the density! / residuals! launch sequences, replayed while the active set
    // and the engine's lane tables stay the same
    bool use_graph = true;
    struct Graph {
        hipGraphExec_t exec = nullptr;
        std::vector<int> act;
        long gen = -1;
        const SysTask* d_act = nullptr;
    } graphs[3];
    void drop_graphs();
    void run(int kind, const std::vector<int>& act);   // 0 density, 1 residuals, 2 sweeps

    Batch(const std::vector<System*>& s);
    ~Batch();
    void bind_active(const std::vector<int>& act);
    void set_precision(int p);
    void density(const std::vector<int>& act);
    void residuals(const std::vector<int>& act);
    // the propagator sweeps alone (q and q+ of every chain), no Q, Simpson or pooling
    void sweeps(const std::vector<int>& act);
    void sys_sums(const std::vector<int>& act, std::vector<double>& out);  // 5 per system
    // sys_sums in two halves: queue the reduction and its copy to the host, then wait for
    // that copy only (work queued in between keeps running)
    void sums_launch(const std::vector<int>& act);
    void sums_read(const std::vector<int>& act, std::vector<double>& out);
    hipEvent_t sums_ev = nullptr;
    // dom: the DomainUpdater settings for a variable cell (each system gets its own
    // updater with these settings), or nullptr for a fixed cell
    std::vector<BatchResult> scft_anderson(int nhist, int iter_euler, double lam_euler,
                                           int nsteps, double ftol, const int* mask,
                                           const Domain* dom = nullptr, double stol = 0.0);
    // scft!(sys, MomentumUpdater(method, lam, gamma, beta1, beta2)) for method 0 MOM,
    // 1 RMS, 2 ADAM, 3 NAG
    std::vector<BatchResult> scft_momentum(int method, double lam, double gamma, double b1,
                                           double b2, int nsteps, double ftol, const int* mask,
                                           const Domain* dom = nullptr, double stol = 0.0);
    // scft!(sys, EulerUpdater(method, lam)) for method 0 PEC, 1 PECE
    std::vector<BatchResult> scft_euler(int method, double lam, int nsteps, double ftol,
                                        const int* mask, const Domain* dom = nullptr,
                                        double stol = 0.0);
    // the scft! loop of src/simulate.jl:60-148 over the batch with updater u
    std::vector<BatchResult> scft_run(BatchStepper& u, int nsteps, double ftol, const int* mask,
                                      const Domain* dom, double stol);
    // momentum state, per system (empty unless a momentum run set it up)
    struct MomState {
        double *mt = nullptr, *vt = nullptr, *step = nullptr, *tmp = nullptr;
    };
    // the MomTask table of the active set `act` on the device (rebuilt when the set changes)
    void bind_momentum(const std::vector<int>& act);
    std::vector<MomState> mom_state;
    MomTask* d_mom = nullptr;
    std::vector<int> mom_act;        // active set d_mom was built for
    void free_momentum();
    std::vector<std::unique_ptr<Domain>> dom_state;
    // the stress of every system in `set` from its current stacks: the passes are queued
    // on the systems' own streams, then collected (out[i] for i in set)
    void stress_all(const std::vector<int>& set, std::vector<std::vector<double>>& out);
    // one cell update of scft! (src/simulate.jl:113-121) for the systems in `due`, in
    // lockstep: the stresses of their last residuals!, and for those above stol the
    // DomainUpdater's Newton iterations, with every perturbed-cell density! and every
    // residuals! of the systems evaluated together as batch launches
    void domain_round(const std::vector<int>& due, double stol, std::vector<double>& serr,
                      std::vector<BatchResult>& res);
};

std::unique_ptr<BatchEngineBase> make_batch_engine(Batch& b);

// autotuner (pf_tune.cpp)
Schedule autotune_system(System& s, int reps, double* best_ms);
void autotune_batch(Batch& b, int reps,
                    std::vector<std::pair<std::array<int, 4>, Schedule>>& out);

}  // namespace pf
