// Grid-point kernels and FP64 reductions of the SCFT system: exp tables, partition
// functions, Simpson densities, density pooling, potentials / pressure / residuals,
// field error and free-energy sums, Anderson Gram matrix and mixing, momentum steps.
//
// Reductions are two-stage and deterministic: each workgroup writes one partial row
// (fixed grid-stride order), then one workgroup sums the rows in index order.
#include "pf_device.h"

namespace pf {

static constexpr int RB = 256;   // threads per reduction workgroup

int reduce_parts(size_t n) {
    size_t p = (n + RB * 4 - 1) / (RB * 4);
    if (p < 1) p = 1;
    if (p > 512) p = 512;
    return (int)p;
}

__device__ __forceinline__ int reduce_parts_dev(size_t n) {
    size_t p = (n + RB * 4 - 1) / (RB * 4);
    return (int)(p < 1 ? 1 : (p > 512 ? 512 : p));
}

template <int NV>
__device__ void wg_reduce_store(double* v, double* out_row) {
    __shared__ double red[NV > 0 ? NV * 4 : 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) red[k * 4 + wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int k = threadIdx.x;
        out_row[k] = ((red[k * 4] + red[k * 4 + 1]) + red[k * 4 + 2]) + red[k * 4 + 3];
    }
}

// ---------------------------------------------------------------------------------
// LW = exp(-(ds/2) v w), computed as exp((-lw_coeff * vol) * w) (src/system/mde.jl:50);
// `coef` carries (-lw_coeff * vol) from the host.
template <typename T>
__global__ void k_lw_table(T* out, const double* __restrict__ w, double coef, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = (T)exp(coef * w[i]);
}

template <typename T>
void launch_lw_table(T* out, const double* w, double coef, size_t n, hipStream_t st) {
    int g = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_lw_table<T>, dim3(g), dim3(256), 0, st, out, w, coef, n);
    PF_HIP(hipGetLastError());
}

// LW2 = exp(-lw_coeff * vol * omega / 2) (src/system/mde.jl:51)
template <typename T>
__global__ void k_lw_half_table(T* out, const double* __restrict__ w, double coef, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = (T)exp(coef * w[i] / 2.0);
}

template <typename T>
void launch_lw_half_table(T* out, const double* w, double coef, size_t n, hipStream_t st) {
    int g = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_lw_half_table<T>, dim3(g), dim3(256), 0, st, out, w, coef, n);
    PF_HIP(hipGetLastError());
}

template <typename T>
__global__ void k_fill(T* p, T v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename T>
void launch_fill(T* p, T v, size_t n, hipStream_t st) {
    int g = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill<T>, dim3(g), dim3(256), 0, st, p, v, n);
    PF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------
// Q = sum(q_{Ns+1}) / n  (src/species/multiblock.jl:175); one row of partials per task.
template <typename T>
__global__ void __launch_bounds__(RB) k_sum_slice(const SimpsonTask<T>* __restrict__ tasks,
                                                  double* partial, int nparts) {
    const SimpsonTask<T>* tk = tasks + blockIdx.y;
    const size_t n = tk->n;
    // each task keeps the partition of its own size, so a task sums identically whether
    // it runs alone or inside a batch of larger systems
    const int own = reduce_parts_dev(n);
    if ((int)blockIdx.x >= own) return;
    const T* q = tk->q + (size_t)tk->Ns * n;
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)own * RB)
        acc += (double)q[i];
    double v[1] = {acc};
    wg_reduce_store<1>(v, partial + (size_t)blockIdx.y * nparts + blockIdx.x);
}

template <typename T>
__global__ void __launch_bounds__(64) k_finish_Q(const SimpsonTask<T>* __restrict__ tasks,
                                                 const double* partial, int nparts) {
    const SimpsonTask<T>* tk = tasks + blockIdx.x;
    // fixed-order two-level sum: lane l adds parts l, l+64, ...; then lane 0 adds lanes
    __shared__ double part[64];
    const int own = reduce_parts_dev(tk->n);
    double s = 0.0;
    for (int p = threadIdx.x; p < own; p += 64) s += partial[(size_t)blockIdx.x * nparts + p];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int l = 0; l < 64; ++l) t += part[l];
        tk->sc->Q[tk->isp] = t / (double)tk->n;
    }
}

template <typename T>
void launch_species_Q(const SimpsonTask<T>* d_tasks, int ntasks, size_t nmax, double* partial,
                      int nparts, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_slice<T>, dim3(nparts, ntasks), dim3(RB), 0, st, d_tasks, partial,
                       nparts);
    hipLaunchKernelGGL(k_finish_Q<T>, dim3(ntasks), dim3(64), 0, st, d_tasks, partial, nparts);
    PF_HIP(hipGetLastError());
}

// Simpson contour integration per block, normalisation by Q Nref, and block -> monomer
// sums (src/species/multiblock.jl:182-228), in the reference's order of operations.
template <typename T>
__global__ void k_simpson(const SimpsonTask<T>* __restrict__ tasks) {
    const SimpsonTask<T>* tk = tasks + blockIdx.y;
    const int Ns = tk->Ns;
    const size_t n = tk->n;
    const double QN = tk->sc->Q[tk->isp] * tk->Nref;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        double acc_m[PF_MAXMON];
        for (int m = 0; m < tk->nmon; ++m) acc_m[m] = 0.0;
        for (int b = 0; b < tk->nblk; ++b) {
            const int bgn = tk->begin[b], end = tk->begin[b + 1];
            const auto q = gps(tk->q);
            const auto qc = gps(tk->qc ? tk->qc : tk->q);
            const bool rev = tk->qc == nullptr;
            auto qv = [&](int s) { return (double)q[(size_t)(s - 1) * n + i]; };
            auto qcv = [&](int s) {
                return rev ? (double)qc[(size_t)(Ns + 1 - s) * n + i]
                           : (double)qc[(size_t)(s - 1) * n + i];
            };
            double rho = qv(bgn) * qcv(bgn);
            rho += qv(end) * qcv(end);
            for (int s = bgn + 1; s <= end - 1; s += 2) rho += 4.0 * qv(s) * qcv(s);
            for (int s = bgn + 2; s <= end - 2; s += 2) rho += 2.0 * qv(s) * qcv(s);
            rho *= tk->ds[b] / 3.0;
            rho /= QN;
            acc_m[tk->blk_mon[b]] += rho;
            if (tk->rho_blk) tk->rho_blk[(size_t)b * n + i] = rho;
        }
        for (int m = 0; m < tk->nmon; ++m) tk->rho_sp[(size_t)m * n + i] = acc_m[m];
    }
}

template <typename T>
void launch_simpson(const SimpsonTask<T>* d_tasks, int ntasks, size_t nmax, hipStream_t st) {
    int g = (int)std::min<size_t>((nmax + 255) / 256, 2048);
    hipLaunchKernelGGL(k_simpson<T>, dim3(g, ntasks), dim3(256), 0, st, d_tasks);
    PF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------
// Point species (src/species/point.jl:36-53).
__global__ void __launch_bounds__(RB) k_point_sum(const double* __restrict__ w, double mvol,
                                                  size_t n, double* partial, int nparts) {
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)nparts * RB)
        acc += exp(mvol * w[i]);
    double v[1] = {acc};
    wg_reduce_store<1>(v, partial + blockIdx.x);
}

__global__ void k_point_finish(const double* partial, int nparts, size_t n, DevScalars* sc,
                               int isp) {
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int p = 0; p < nparts; ++p) s += partial[p];
        sc->Q[isp] = s / (double)n;
    }
}

__global__ void k_point_rho(const double* __restrict__ w, double mvol, size_t n, double* rho,
                            const DevScalars* __restrict__ sc, int isp) {
    const double Q = sc->Q[isp];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        rho[i] = exp(mvol * w[i]) / Q;
}

void launch_point_density(const double* w, double vol, size_t n, double* rho_sp,
                          double* partial, int nparts, DevScalars* sc, int isp,
                          hipStream_t st) {
    const double mvol = -vol;
    hipLaunchKernelGGL(k_point_sum, dim3(nparts), dim3(RB), 0, st, w, mvol, n, partial, nparts);
    hipLaunchKernelGGL(k_point_finish, dim3(1), dim3(64), 0, st, partial, nparts, n, sc, isp);
    int g = (int)std::min<size_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_point_rho, dim3(g), dim3(256), 0, st, w, mvol, n, rho_sp, sc, isp);
    PF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------
// muphi! (src/system/system.jl:248-259)
__global__ void k_muphi(DevScalars* sc, int nsp, int canonical) {
    if (threadIdx.x == 0) {
        for (int i = 0; i < nsp; ++i) {
            if (canonical)
                sc->mu[i] = log(sc->phi[i] / sc->Q[i]);
            else
                sc->phi[i] = sc->Q[i] * exp(sc->mu[i]);
        }
    }
}

void launch_muphi(DevScalars* sc, int nsp, int canonical, hipStream_t st) {
    hipLaunchKernelGGL(k_muphi, dim3(1), dim3(64), 0, st, sc, nsp, canonical);
    PF_HIP(hipGetLastError());
}

// out[0] = 1 when a canonical species has phi / Q < 0: the reference's muphi! takes
// log(phi / Q) (src/system/system.jl:251), which raises a DomainError there; else 0.
__global__ void k_domain_flag(const DevScalars* __restrict__ sc, int nsp, int canonical,
                              double* out) {
    if (threadIdx.x == 0) {
        double f = 0.0;
        if (canonical)
            for (int i = 0; i < nsp; ++i)
                if (sc->phi[i] / sc->Q[i] < 0.0) f = 1.0;
        out[0] = f;
    }
}

void launch_domain_flag(const DevScalars* sc, int nsp, int canonical, double* out,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_domain_flag, dim3(1), dim3(64), 0, st, sc, nsp, canonical, out);
    PF_HIP(hipGetLastError());
}

// Pooling: rho_a = sum_p phi_p rho_{p,a} (src/system/system.jl:281-286).
__global__ void k_pool(PoolArgs a, size_t n, const DevScalars* __restrict__ sc) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        for (int m = 0; m < a.nmon; ++m) {
            double acc = 0.0;
            for (int t = 0; t < a.nterm[m]; ++t) acc += sc->phi[a.isp[m][t]] * a.src[m][t][i];
            a.dst[(size_t)m * n + i] = acc;
        }
    }
}

void launch_pool(const PoolArgs& a, size_t n, DevScalars* sc, hipStream_t st) {
    int g = (int)std::min<size_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_pool, dim3(g), dim3(256), 0, st, a, n, sc);
    PF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------
// potentials! + compressibility + residuals (src/system/system.jl:296-329,
// src/system/compress.jl:37-61, src/interactions/flory.jl:82-99,
// src/interactions/edwards.jl:72-87) with per-monomer sums of w and res for the
// canonical mean subtraction.
__global__ void __launch_bounds__(RB) k_residuals(ResidualArgs a, size_t n, double* partial,
                                                  int nparts) {
    double sw[PF_MAXMON], sr[PF_MAXMON];
    for (int m = 0; m < PF_MAXMON; ++m) sw[m] = sr[m] = 0.0;
    const int nm = a.nmon;
    for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)nparts * RB) {
        double rho[PF_MAXMON], pot[PF_MAXMON];
        double rhosum = 0.0;
        for (int m = 0; m < nm; ++m) {
            rho[m] = a.rho[(size_t)m * n + i];
            rhosum += rho[m];
        }
        for (int m = 0; m < nm; ++m) {
            double p = 0.0;
            if (a.has_flory)
                for (int b = 0; b < nm; ++b) {
                    const double chi = a.chi[m * PF_MAXMON + b];
                    if (a.pair[m * PF_MAXMON + b]) p += chi * (rho[b] / a.vol[b]);
                }
            if (a.u_set[m]) p += a.u[m] * (rho[m] / a.vol[m]);
            pot[m] = p;
        }
        double eta = 0.0;
        if (a.compress == 0) {
            eta = rhosum - 1.0;
            for (int m = 0; m < nm; ++m) eta += a.w[(size_t)m * n + i] - pot[m];
            eta /= (double)nm;
        } else if (a.compress == 1) {
            eta = (1.0 / a.inv_zeta) * (rhosum - 1.0);
        }
        for (int m = 0; m < nm; ++m) {
            const double w = a.w[(size_t)m * n + i];
            const double r = pot[m] + eta - w;
            a.pot[(size_t)m * n + i] = pot[m];
            a.res[(size_t)m * n + i] = r;
            sw[m] += w;
            sr[m] += r;
        }
    }
    double v[2 * PF_MAXMON];
    for (int m = 0; m < PF_MAXMON; ++m) {
        v[m] = sw[m];
        v[PF_MAXMON + m] = sr[m];
    }
    wg_reduce_store<2 * PF_MAXMON>(v, partial + (size_t)blockIdx.x * 2 * PF_MAXMON);
}

__global__ void k_mean_subtract(ResidualArgs a, size_t n, const double* __restrict__ sums) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        for (int m = 0; m < a.nmon; ++m) {
            const double mw = sums[m] / (double)n, mr = sums[PF_MAXMON + m] / (double)n;
            a.w[(size_t)m * n + i] -= mw;
            a.res[(size_t)m * n + i] -= mr;
        }
    }
}

// Final reduction: out[k] = sum_p partial[p*nvals + k], p in order.
__global__ void k_final(const double* __restrict__ partial, int nparts, int nvals, double* out) {
    for (int k = threadIdx.x; k < nvals; k += blockDim.x) {
        double s = 0.0;
        for (int p = 0; p < nparts; ++p) s += partial[(size_t)p * nvals + k];
        out[k] = s;
    }
}

void launch_final_reduce(const double* partial, int nparts, int nvals, double* out,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, st, partial, nparts, nvals, out);
    PF_HIP(hipGetLastError());
}

void launch_residuals(const ResidualArgs& a, size_t n, double* partial, int nparts,
                      double* sums, hipStream_t st) {
    hipLaunchKernelGGL(k_residuals, dim3(nparts), dim3(RB), 0, st, a, n, partial, nparts);
    launch_final_reduce(partial, nparts, 2 * PF_MAXMON, sums, st);
    if (a.canonical) {
        int g = (int)std::min<size_t>((n + 255) / 256, 2048);
        hipLaunchKernelGGL(k_mean_subtract, dim3(g), dim3(256), 0, st, a, n, sums);
    }
    PF_HIP(hipGetLastError());
}

// Sums for field_error (system.jl:470-486) and free_energy (system.jl:402-433,
// flory.jl:39-59, edwards.jl:34-52): [sum res^2, sum w^2, sum rho w, flory, edwards].
__global__ void __launch_bounds__(RB) k_sys_reduce(ResidualArgs a, size_t n, double* partial,
                                                   int nparts) {
    double v[5] = {0, 0, 0, 0, 0};
    const int nm = a.nmon;
    for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)nparts * RB) {
        for (int m = 0; m < nm; ++m) {
            const double r = a.res[(size_t)m * n + i], w = a.w[(size_t)m * n + i];
            const double rho = a.rho[(size_t)m * n + i];
            v[0] += r * r;
            v[1] += w * w;
            v[2] += rho * w;
            if (a.u_set[m]) v[4] += (a.u[m] / 2) * ((rho / a.vol[m]) * (rho / a.vol[m]));
            for (int b = m + 1; b < nm; ++b)
                if (a.pair[m * PF_MAXMON + b])
                    v[3] += a.chi[m * PF_MAXMON + b] * (rho / a.vol[m]) *
                            (a.rho[(size_t)b * n + i] / a.vol[b]);
        }
    }
    wg_reduce_store<5>(v, partial + (size_t)blockIdx.x * 5);
}

void launch_sys_reduce(const ResidualArgs& a, size_t n, double* partial, int nparts,
                       double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_sys_reduce, dim3(nparts), dim3(RB), 0, st, a, n, partial, nparts);
    launch_final_reduce(partial, nparts, 5, out, st);
}

void launch_copy(double* dst, const double* src, size_t n, hipStream_t st) {
    PF_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, st));
}

__global__ void k_axpy(double* w, const double* __restrict__ r, double lam, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        w[i] += lam * r[i];
}

void launch_axpy(double* w, const double* r, double lam, size_t n, hipStream_t st) {
    int g = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_axpy, dim3(g), dim3(256), 0, st, w, r, lam, n);
    PF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------
// Anderson Gram (src/updaters/anderson.jl:166-200): with d_m = r - r_m,
// U_mn = sum d_m d_n (m <= n), V_m = sum d_m r, over all monomers and grid points.
template <int H>
__global__ void __launch_bounds__(RB) k_gram(const double* __restrict__ r, HistPtrs hp, size_t N,
                                             double* partial, int nparts) {
    constexpr int NU = H * (H + 1) / 2;
    double acc[NU + H];
#pragma unroll
    for (int k = 0; k < NU + H; ++k) acc[k] = 0.0;
    for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < N; i += (size_t)nparts * RB) {
        const double ri = r[i];
        double d[H];
#pragma unroll
        for (int m = 0; m < H; ++m) d[m] = ri - hp.p[m][i];
        int k = 0;
#pragma unroll
        for (int m = 0; m < H; ++m)
#pragma unroll
            for (int nn = m; nn < H; ++nn) acc[k++] += d[m] * d[nn];
#pragma unroll
        for (int m = 0; m < H; ++m) acc[NU + m] += d[m] * ri;
    }
    wg_reduce_store<NU + H>(acc, partial + (size_t)blockIdx.x * (NU + H));
}

template <int H>
static void gram_h(const double* r, const HistPtrs& hp, size_t N, double* partial, int nparts,
                   double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_gram<H>, dim3(nparts), dim3(RB), 0, st, r, hp, N, partial, nparts);
    launch_final_reduce(partial, nparts, H * (H + 1) / 2 + H, out, st);
}

void launch_gram(int h, const double* r, const HistPtrs& hp, size_t N, double* partial,
                 int nparts, double* out, hipStream_t st) {
    switch (h) {
#define PF_G(H) case H: gram_h<H>(r, hp, N, partial, nparts, out, st); break;
        PF_G(1) PF_G(2) PF_G(3) PF_G(4) PF_G(5) PF_G(6) PF_G(7) PF_G(8)
        PF_G(9) PF_G(10) PF_G(11) PF_G(12) PF_G(13) PF_G(14) PF_G(15) PF_G(16)
#undef PF_G
        default: throw Error(-2, "Anderson history depth above 16 is not supported");
    }
    PF_HIP(hipGetLastError());
}

// Anderson Gram for any history depth h (src/updaters/anderson.jl:166-200): the
// (h+1) x (h+1) Gram matrix of x_m = r - r_m (m < h) and x_h = r, by 8 x 8 tiles.
// Workgroup (part, tile) sums its tile over a grid-stride share of the points; the
// entries past h of the last tile row are zero.  U_mn is entry (m, n), V_m entry (m, h).
__device__ __forceinline__ double gram_vec(const double* __restrict__ r, const HistRing& g,
                                           int h, int k, size_t i, double ri) {
    if (k < h) return ri - g.r[ring_off(g, k) + i];
    return k == h ? ri : 0.0;
}

__global__ void __launch_bounds__(RB) k_gram_tiled(const double* __restrict__ r, HistRing g,
                                                   int h, double* partial) {
    const int p = blockIdx.y;
    int J = 0;
    while ((J + 1) * (J + 2) / 2 <= p) ++J;
    const int I = p - J * (J + 1) / 2;
    double acc[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) acc[k] = 0.0;
    const size_t N = g.N;
    for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < N;
         i += (size_t)gridDim.x * RB) {
        const double ri = r[i];
        double a[8], b[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            a[t] = gram_vec(r, g, h, 8 * I + t, i, ri);
            b[t] = gram_vec(r, g, h, 8 * J + t, i, ri);
        }
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[t * 8 + u] += a[t] * b[u];
    }
    wg_reduce_store<64>(acc, partial + ((size_t)blockIdx.x * gridDim.y + p) * 64);
}

void launch_gram_tiled(int h, const double* r, const HistRing& g, double* partial, double* out,
                       hipStream_t st) {
    const int nt = gram_tiles(h), nparts = gram_tiled_parts(g.N);
    hipLaunchKernelGGL(k_gram_tiled, dim3(nparts, nt), dim3(RB), 0, st, r, g, h, partial);
    launch_final_reduce(partial, nparts, nt * 64, out, st);
    PF_HIP(hipGetLastError());
}

__global__ void k_anderson_mix_ring(int h, double* w, const double* __restrict__ r, HistRing g,
                                    const double* __restrict__ f_iter,
                                    const double* __restrict__ C, double lam, double* f_out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < g.N;
         i += (size_t)gridDim.x * blockDim.x) {
        double WW = w[i];
        double DD = r[i];
        const double fi = f_iter[i], ri = r[i];
        for (int nn = 0; nn < h; ++nn) {
            const size_t o = ring_off(g, nn) + i;
            WW += C[nn] * (g.w[o] - fi);
            DD += C[nn] * (g.r[o] - ri);
        }
        const double o = WW + lam * DD;
        w[i] = o;
        if (f_out) f_out[i] = o;
    }
}

void launch_anderson_mix_ring(int h, double* w, const double* r, const HistRing& g,
                              const double* f_iter, const double* C, double lam, double* f_out,
                              hipStream_t st) {
    int gr = (int)std::min<size_t>((g.N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_anderson_mix_ring, dim3(gr), dim3(256), 0, st, h, w, r, g, f_iter, C,
                       lam, f_out);
    PF_HIP(hipGetLastError());
}

// Anderson mix (src/updaters/anderson.jl:211-224) and history append.
__global__ void k_anderson_mix(int h, double* w, const double* __restrict__ r, HistPtrs fh,
                               HistPtrs rh, const double* __restrict__ f_iter, MixCoef c,
                               double lam, size_t N, double* f_out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x) {
        double WW = w[i];
        double DD = r[i];
        const double fi = f_iter[i], ri = r[i];
        for (int nn = 0; nn < h; ++nn) {
            WW += c.C[nn] * (fh.p[nn][i] - fi);
            DD += c.C[nn] * (rh.p[nn][i] - ri);
        }
        const double o = WW + lam * DD;
        w[i] = o;
        if (f_out) f_out[i] = o;
    }
}

void launch_anderson_mix(int h, double* w, const double* r, const HistPtrs& fh,
                         const HistPtrs& rh, const double* f_iter, const MixCoef& c, double lam,
                         size_t N, double* f_out, hipStream_t st) {
    int g = (int)std::min<size_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_anderson_mix, dim3(g), dim3(256), 0, st, h, w, r, fh, rh, f_iter, c, lam,
                       N, f_out);
    PF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------
// Momentum family updates (src/updaters/momentum.jl:86-139); `method`: 0 MOM, 1 RMS,
// 2 ADAM.  The NAG variant is composed on the host from fills/axpys.
__global__ void k_momentum(int method, double* mt, double* vt, double* step,
                           const double* __restrict__ r, double b1, double b2, size_t N) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x)
        momentum_point(method, mt, vt, step, r, b1, b2, i);
}

void launch_momentum(int method, double* mt, double* vt, double* step, const double* r,
                     double b1, double b2, size_t N, hipStream_t st) {
    int g = (int)std::min<size_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_momentum, dim3(g), dim3(256), 0, st, method, mt, vt, step, r, b1, b2, N);
    PF_HIP(hipGetLastError());
}

// Segmented sum: out[s*nvals + k] = sum_p in[(s*nparts + p)*stride + k].
__global__ void k_segsum(const double* __restrict__ in, int nparts, int stride, int nvals,
                         double* out) {
    const int s = blockIdx.x;
    for (int k = threadIdx.x; k < nvals; k += blockDim.x) {
        double acc = 0.0;
        for (int p = 0; p < nparts; ++p) acc += in[((size_t)s * nparts + p) * stride + k];
        out[(size_t)s * nvals + k] = acc;
    }
}

void launch_segsum(const double* in, int nseg, int nparts, int stride, int nvals, double* out,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_segsum, dim3(nseg), dim3(64), 0, st, in, nparts, stride, nvals, out);
    PF_HIP(hipGetLastError());
}

#define PF_INST2(T)                                                                         \
    template void launch_lw_table<T>(T*, const double*, double, size_t, hipStream_t);        \
    template void launch_lw_half_table<T>(T*, const double*, double, size_t, hipStream_t);   \
    template void launch_fill<T>(T*, T, size_t, hipStream_t);                                \
    template void launch_species_Q<T>(const SimpsonTask<T>*, int, size_t, double*, int,     \
                                      hipStream_t);                                         \
    template void launch_simpson<T>(const SimpsonTask<T>*, int, size_t, hipStream_t);
PF_INST2(float)
PF_INST2(double)

}  // namespace pf
