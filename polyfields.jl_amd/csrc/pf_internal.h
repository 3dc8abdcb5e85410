// Internal declarations shared by the HIP kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#define PF_MAXBLK 8      // blocks per chain
#define PF_MAXPASS 16    // radix passes per transform
#define PF_MAXMON 8      // monomer types per system
#define PF_MAXPAR 6      // cell parameters
#define PF_MAXHIST 16    // Anderson history depth served by the one-pass Gram kernels;
                         // deeper histories take the tiled Gram (any depth)

namespace pf {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PF_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::pf::Error(-3, std::string("HIP error '") + hipGetErrorString(e_) +   \
                                      "' at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

template <typename T>
struct cpx {
    T x, y;
};

// Radix plan for one transform length (product of radix[0..npass)).
struct FftPlan {
    int n;
    int npass;
    int8_t radix[PF_MAXPASS];
};

// One propagator lane: a chain sweep in one direction over one grid.  All pointers are
// device pointers.  Contour indices s are 1-based as in the reference; slice s of the
// stack starts at stack + (s-1)*n.
template <typename T>
struct Lane {
    int nx, ny, nz;          // grid (x fastest)
    int dir;                 // +1: s increases (q), -1: s decreases (q+)
    int nsteps;              // Ns
    int nblk;
    int begin[PF_MAXBLK + 1];
    const T* lw[PF_MAXBLK];  // exp(-ds_b/2 v w) per block, n entries
    double ldc[PF_MAXBLK];   // ds_b b_b^2 / 6
    T* stack;
    cpx<T>* kbuf;            // nk complex scratch (3-D lanes)
    double G[9];             // reciprocal basis, row-major G[i][j]
    const cpx<T>* twx;       // exp(-2 pi i m / nx), m < nx
    const cpx<T>* twy;       // exp(-2 pi i m / ny)
    const cpx<T>* twz;       // exp(-2 pi i m / nz)
    FftPlan px, py, pz;      // plans for nx/2, ny, nz
    int rqm4;                // 1: RQM4 step (Richardson of one full and two half steps)
    const T* lw2[PF_MAXBLK]; // exp(-ds_b/4 v w) per block (RQM4 only)
    T* rq;                   // 3-D RQM4: the full step's result q1 (n entries); the half
                             // steps' spectrum is kb2
    // --- adjoint sweeps (free-energy gradient); unused by the density sweeps --------
    const T* fwd;            // forward stack of the chain this lane is the adjoint of
    const T* other;          // partner stack of the Simpson source (q+ for q, q for q+)
    int other_rev;           // 1: partner slice s is other[Ns+2-s] (homopolymer)
    const double* rbar;      // dF/d(system density) [slot][n]
    int blk_slot[PF_MAXBLK]; // system monomer slot of each block
    double srcc[PF_MAXBLK];  // phi ds_b / (3 Q Nref)  (doubled for a homopolymer)
    double wco[PF_MAXBLK];   // -(ds_b v_b / 2)
    double qinit;            // adjoint value at the starting slice from dF/dQ (q: Qbar/n)
    double* wbar;            // dF/dw accumulated by this lane, per block [blk][n]
    cpx<T>* kb2;             // nk scratch: spectrum of LW a_in
    double* gpart;           // per-workgroup k-space sums [wg][blk][PF_MAXPAR + 1]
    int nparams, dim;
    double dGG[3 * 3 * PF_MAXPAR];
};

// Stress task: transforms of q_s and q+_s reduced against dksq.
template <typename T>
struct StressLane {
    int nx, ny, nz;
    const T* q;              // slice q_s
    const T* qc;             // slice q+_s
    int out;                 // row of the partial-sum table
    int nparams, dim;
    double dGG[3 * 3 * PF_MAXPAR];  // dGG[(l*3+m)*np + k]
    const cpx<T>* twx;
    const cpx<T>* twy;
    const cpx<T>* twz;
    FftPlan px, py, pz;
    cpx<T>* kq;              // 3-D: nk scratch for q_s
    cpx<T>* kqc;             // 3-D: nk scratch for q+_s
};

struct DevScalars;

// Species entry for the Simpson density / Q kernels.
template <typename T>
struct SimpsonTask {
    const T* q;              // forward stack
    const T* qc;             // backward stack (nullptr: homopolymer, q+_s = q_{Ns+2-s})
    int Ns, nblk;
    int begin[PF_MAXBLK + 1];
    double ds[PF_MAXBLK];
    int blk_mon[PF_MAXBLK];  // monomer slot of each block
    double Nref;
    int isp;                 // species index (scalar slots)
    double* rho_sp;          // species density per monomer slot: rho_sp[slot*n + i]
    int nmon;
    double* rho_blk;         // optional: per-block densities [blk][n] (gradient pass)
    size_t n;                // grid points of the owning system
    DevScalars* sc;          // scalar slots of the owning system
};

// System-level scalar slots kept on the device.
struct DevScalars {
    double Q[64];
    double phi[64];
    double mu[64];
    double mean_w[PF_MAXMON];
    double mean_r[PF_MAXMON];
    double red[16];          // reduction results
};

// --- launchers (pf_fft.hip) -------------------------------------------------------
template <typename T>
void launch_planar_steps(const Lane<T>* d_lanes, int nlanes, int t0, int t1, int lds_bytes,
                         int threads, hipStream_t st);
template <typename T>
void launch_slab(const Lane<T>* d_lanes, int nlanes, int nz, int mode, int t, int lds_bytes,
                 int threads, hipStream_t st);
// half = 1: the RQM4 half-step spectrum kb2 with LD at half the contour step
template <typename T>
void launch_pencil(const Lane<T>* d_lanes, int nlanes, int ny, int ntiles, int tile, int t,
                   int lds_bytes, int threads, hipStream_t st, int half = 0);
// 3-D RQM4 xy-slab steps (two planes of LDS): RQ_START, RQ_MID, RQ_END
enum RqmMode { RQ_START = 0, RQ_MID = 1, RQ_END = 2 };
template <typename T>
void launch_slab_rqm4(const Lane<T>* d_lanes, int nlanes, int nz, int mode, int t,
                      int lds_bytes, hipStream_t st);
template <typename T>
void launch_stress_planar(const StressLane<T>* d_lanes, int nlanes, double* partial,
                          int lds_bytes, int threads, hipStream_t st);
template <typename T>
void launch_stress_slab(const StressLane<T>* d_lanes, int nlanes, int nz, int lds_bytes,
                        int threads, hipStream_t st);
template <typename T>
void launch_stress_pencil(const StressLane<T>* d_lanes, int nlanes, int ny, int ntiles,
                          int tile, double* partial, int lds_bytes, int threads,
                          hipStream_t st);

enum SlabMode { SLAB_FWD = 0, SLAB_INVFWD = 1 };

// adjoint sweeps (pf_adjoint.hip); mode 0 starts the sweep, 1 continues after step t
// mode 2: only the B spectrum of step t; nextB = 0: mode 1 stops after the A spectrum
template <typename T>
void launch_adj_slab(const Lane<T>* d_lanes, int nlanes, int nz, int mode, int t, int lds_bytes,
                     hipStream_t st, int nextB = 1);
template <typename T>
void launch_adj_pencil(const Lane<T>* d_lanes, int nlanes, int ny, int ntiles, int tile, int t,
                       int lds_bytes, hipStream_t st);
// compile-time forms (pf_adjoint.hip): false / 0 when the size has none or the schedule
// asks for the generic kernels; the pencil returns its workgroups per lane (partial-sum
// rows)
template <typename T>
bool fixed_adj_slab(const Lane<T>* d, int nl, int nx, int ny, int nz, int mode, int t,
                    hipStream_t st, int nextB = 1);
template <typename T>
int fixed_adj_pencil(const Lane<T>* d, int nl, int nx, int ny, int nz, int t, hipStream_t st);

// Kernel schedule of one grid size and precision, process-wide (pf_fft_fixed.hip).  -1
// keeps the built-in default; pf_autotune times the candidates and writes the fastest.
struct Schedule {
    int slab = -1;     // 0 generic, 1 compile-time kernel with 256 threads, 2 with 512,
                       // 3 shuffle-exchange (FP32 64 x 64 planes), 4 1024 threads (FP32 128^2)
    int pencil = -1;   // 0 generic, 1 compile-time LDS pencil, 2 register pencil, 3 pair pencil,
                       // 4 LDS pencil with one ky row per workgroup (64-point lines) or
                       // 512 threads (128-point lines), 5 one ky row and 512 threads
                       // (64-point lines) or 1024 threads (FP32 128), 6 half a ky row
                       // per workgroup (64- and 128-point lines)
    int planar = -1;   // 0 generic, compile-time planar kernel: 1 (2-D: 256 threads),
                       // 2 (2-D: 512), 3 (2-D: 1024)
    int split = -1;    // streams per 3-D lane group in a batch (1 or 2)
};
Schedule& schedule(int elem_bytes, int nx, int ny, int nz);
int device_cus();        // compute units of the current device (looked up once per device)
void schedule_reset();   // every grid back to the built-in defaults

// compile-time specialised variants (pf_fft_fixed.hip); false: size not specialised
template <typename T>
bool fixed_planar(const Lane<T>* d, int nl, int nx, int ny, int t0, int t1, hipStream_t st);
template <typename T>
bool fixed_planar_rqm4(const Lane<T>* d, int nl, int nx, int ny, int t0, int t1, bool one,
                       hipStream_t st);
template <typename T>
bool fixed_slab_rqm4(const Lane<T>* d, int nl, int nx, int ny, int nz, int mode, int t,
                     hipStream_t st);
template <typename T>
bool fixed_pencil_rqm4(const Lane<T>* d, int nl, int nx, int ny, int nz, int t, int half,
                       hipStream_t st);
template <typename T>
bool fixed_slab(const Lane<T>* d, int nl, int nx, int ny, int nz, int mode, int t,
                hipStream_t st);
// alone: no other lane group's launches run beside this one (a single system's sweep)
template <typename T>
bool fixed_pencil(const Lane<T>* d, int nl, int nx, int ny, int nz, bool alone, int t,
                  hipStream_t st);

// --- launchers (pf_kernels.hip) ---------------------------------------------------
struct ResidualArgs {
    int nmon;
    double vol[PF_MAXMON];
    int has_flory;
    double chi[PF_MAXMON * PF_MAXMON];   // symmetric
    int pair[PF_MAXMON * PF_MAXMON];     // 1 where a chi pair is set
    double u[PF_MAXMON];                 // Edwards excluded volume
    int u_set[PF_MAXMON];
    int compress;                        // 0 strict, 1 compressible, 2 off
    double inv_zeta;
    int canonical;
    double* w;                           // fields     [nmon][n]
    double* pot;                         // potentials [nmon][n]
    double* res;                         // residuals  [nmon][n]
    const double* rho;                   // densities  [nmon][n]
};

struct PoolArgs {
    int nmon;
    int nterm[PF_MAXMON];
    const double* src[PF_MAXMON][8];
    int isp[PF_MAXMON][8];
    double* dst;
};

struct HistPtrs {
    const double* p[PF_MAXHIST];
};

struct MixCoef {
    double C[PF_MAXHIST];
};

// Anderson history as a ring of `nslots` slots of N doubles each: logical slot m of a
// Gram or mix (m = 0 the newest earlier entry, m = h - 1 the oldest) is physical slot
// (start - m) % nslots, with start >= h - 1.
struct HistRing {
    const double* r;     // residual slots
    const double* w;     // field slots
    size_t N;
    int nslots, start;
};
__host__ __device__ inline size_t ring_off(const HistRing& g, int m) {
    return (size_t)((g.start - m) % g.nslots) * g.N;
}

// Tiled Gram of the vectors x_m = r - r_m (m < h) and x_h = r: 8 x 8 tiles (I <= J) of
// the (h+1) x (h+1) matrix, tile p = J (J + 1) / 2 + I, entry (a, b) at p * 64 + a * 8 + b.
inline int gram_tiles(int h) {
    const int nb = (h + 1 + 7) / 8;
    return nb * (nb + 1) / 2;
}
// workgroups per tile row of a system of N history points (one per 8192 points, at most
// 64): a function of N only, so that a batch sums in the same order as one system
__host__ __device__ inline int gram_tiled_parts(size_t N) {
    const size_t p = N / 8192;
    return (int)(p < 1 ? 1 : (p > 64 ? 64 : p));
}
inline int gram_tile_index(int i, int j) {   // entry (i, j), i <= j, of the Gram matrix
    const int I = i / 8, J = j / 8;
    return (J * (J + 1) / 2 + I) * 64 + (i % 8) * 8 + (j % 8);
}

int reduce_parts(size_t n);
template <typename T>
void launch_lw_table(T* out, const double* w, double coef, size_t n, hipStream_t st);
template <typename T>
void launch_lw_half_table(T* out, const double* w, double coef, size_t n, hipStream_t st);
template <typename T>
void launch_fill(T* p, T v, size_t n, hipStream_t st);
template <typename T>
void launch_species_Q(const SimpsonTask<T>* d_tasks, int ntasks, size_t nmax, double* partial,
                      int nparts, hipStream_t st);
template <typename T>
void launch_simpson(const SimpsonTask<T>* d_tasks, int ntasks, size_t nmax, hipStream_t st);
void launch_point_density(const double* w, double vol, size_t n, double* rho_sp,
                          double* partial, int nparts, DevScalars* sc, int isp,
                          hipStream_t st);
void launch_muphi(DevScalars* sc, int nsp, int canonical, hipStream_t st);
void launch_domain_flag(const DevScalars* sc, int nsp, int canonical, double* out,
                        hipStream_t st);
void launch_pool(const PoolArgs& a, size_t n, DevScalars* sc, hipStream_t st);
void launch_residuals(const ResidualArgs& a, size_t n, double* partial, int nparts,
                      double* sums, hipStream_t st);
void launch_sys_reduce(const ResidualArgs& a, size_t n, double* partial, int nparts,
                       double* out, hipStream_t st);
void launch_final_reduce(const double* partial, int nparts, int nvals, double* out,
                         hipStream_t st);
void launch_copy(double* dst, const double* src, size_t n, hipStream_t st);
void launch_axpy(double* w, const double* r, double lam, size_t n, hipStream_t st);
void launch_gram(int h, const double* r, const HistPtrs& hp, size_t N, double* partial,
                 int nparts, double* out, hipStream_t st);
// any depth: the tiled Gram (partial: gram_tiled_parts(N) * gram_tiles(h) * 64 doubles;
// out: gram_tiles(h) * 64 values) and the mix with coefficients C[h] in device memory
void launch_gram_tiled(int h, const double* r, const HistRing& g, double* partial, double* out,
                       hipStream_t st);
void launch_anderson_mix_ring(int h, double* w, const double* r, const HistRing& g,
                              const double* f_iter, const double* C, double lam, double* f_out,
                              hipStream_t st);
void launch_anderson_mix(int h, double* w, const double* r, const HistPtrs& fh,
                         const HistPtrs& rh, const double* f_iter, const MixCoef& c, double lam,
                         size_t N, double* f_out, hipStream_t st);
void launch_momentum(int method, double* mt, double* vt, double* step, const double* r,
                     double b1, double b2, size_t N, hipStream_t st);
void launch_segsum(const double* in, int nseg, int nparts, int stride, int nvals, double* out,
                   hipStream_t st);

// checkpointed Simpson (pf_adjoint.hip): tk.q / tk.qc index slices as in a full stack
template <typename T>
void launch_simpson_seg(const SimpsonTask<T>& tk, int s_lo, int s_hi, double* acc,
                        hipStream_t st);
template <typename T>
void launch_simpson_fin(const SimpsonTask<T>& tk, const double* acc, hipStream_t st);

// gradient helpers (pf_adjoint.hip)
struct DotArgs {
    int k;
    const double* x[48];
    const double* y[48];
};
void launch_dots(const DotArgs& a, size_t n, double* partial, int nparts, double* out,
                 hipStream_t st);
void launch_rbar(const ResidualArgs& a, size_t n, double* rbar, hipStream_t st);
void launch_point_grad(const double* w, double vol, size_t n, const double* rbar, double ca,
                       double cb, double* dw, hipStream_t st);

}  // namespace pf
