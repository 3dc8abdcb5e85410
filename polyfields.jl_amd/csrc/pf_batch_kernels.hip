// Batched system kernels: one launch serves every system of a batch (blockIdx.y is the
// system or task).  Per-point arithmetic is the same as in pf_kernels.hip.
#include "pf_batch.h"
#include "pf_device.h"

namespace pf {

static constexpr int RB = 256;

template <int NV>
__device__ void wg_store(double* v, double* out_row) {
    __shared__ double red[NV * 4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) red[k * 4 + wave] = x;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NV; k += RB)
        out_row[k] = ((red[k * 4] + red[k * 4 + 1]) + red[k * 4 + 2]) + red[k * 4 + 3];
}

template <typename T>
__global__ void k_lw_b(const LwTask<T>* __restrict__ tasks) {
    const LwTask<T>* t = tasks + blockIdx.y;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < t->n;
         i += (size_t)gridDim.x * blockDim.x)
        t->out[i] = (T)exp(t->coef * t->w[i]);
}

template <typename T>
__global__ void k_fill_b(const FillTask<T>* __restrict__ tasks) {
    const FillTask<T>* t = tasks + blockIdx.y;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < t->n;
         i += (size_t)gridDim.x * blockDim.x)
        t->p[i] = T(1);
}

template <typename T>
void launch_lw_fill_b(const LwTask<T>* lw, int nlw, const FillTask<T>* fill, int nfill,
                      size_t nmax, hipStream_t st) {
    const int g = (int)std::min<size_t>((nmax + 255) / 256, 1024);
    if (nlw) hipLaunchKernelGGL(k_lw_b<T>, dim3(g, nlw), dim3(256), 0, st, lw);
    if (nfill) hipLaunchKernelGGL(k_fill_b<T>, dim3(g, nfill), dim3(256), 0, st, fill);
    PF_HIP(hipGetLastError());
}

__global__ void k_muphi_b(const SysTask* __restrict__ t, int nsys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    DevScalars* sc = t[s].sc;
    for (int i = 0; i < t[s].nsp; ++i) {
        if (t[s].ra.canonical)
            sc->mu[i] = log(sc->phi[i] / sc->Q[i]);
        else
            sc->phi[i] = sc->Q[i] * exp(sc->mu[i]);
    }
}

__global__ void k_pool_b(const SysTask* __restrict__ tasks) {
    const SysTask* t = tasks + blockIdx.y;
    const PoolArgs& a = t->pool;
    const size_t n = t->n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        for (int m = 0; m < a.nmon; ++m) {
            double acc = 0.0;
            for (int k = 0; k < a.nterm[m]; ++k) acc += t->sc->phi[a.isp[m][k]] * a.src[m][k][i];
            a.dst[(size_t)m * n + i] = acc;
        }
    }
}

void launch_muphi_pool_b(const SysTask* d, int nsys, size_t nmax, hipStream_t st) {
    hipLaunchKernelGGL(k_muphi_b, dim3((nsys + 63) / 64), dim3(64), 0, st, d, nsys);
    const int g = (int)std::min<size_t>((nmax + 255) / 256, 1024);
    hipLaunchKernelGGL(k_pool_b, dim3(g, nsys), dim3(256), 0, st, d);
    PF_HIP(hipGetLastError());
}

// residuals (same per-point arithmetic as k_residuals)
__global__ void __launch_bounds__(RB) k_residuals_b(const SysTask* __restrict__ tasks) {
    const SysTask* t = tasks + blockIdx.y;
    if ((int)blockIdx.x >= t->nparts) return;
    const ResidualArgs& a = t->ra;
    const size_t n = t->n;
    const int nm = a.nmon;
    double sw[PF_MAXMON], sr[PF_MAXMON];
    for (int m = 0; m < PF_MAXMON; ++m) sw[m] = sr[m] = 0.0;
    for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)t->nparts * RB) {
        double rho[PF_MAXMON], pot[PF_MAXMON];
        double rhosum = 0.0;
        for (int m = 0; m < nm; ++m) {
            rho[m] = a.rho[(size_t)m * n + i];
            rhosum += rho[m];
        }
        for (int m = 0; m < nm; ++m) {
            double p = 0.0;
            if (a.has_flory)
                for (int b = 0; b < nm; ++b)
                    if (a.pair[m * PF_MAXMON + b]) p += a.chi[m * PF_MAXMON + b] * (rho[b] / a.vol[b]);
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
    wg_store<2 * PF_MAXMON>(v, t->partial + (size_t)blockIdx.x * 2 * PF_MAXMON);
}

// per system: sums[off + k] = sum_p partial[p*nv + k]  (p in order); with `compact`, the
// same values also go to compact[j*nv + k] for the j-th active system, so that the host
// copies nv values per active system rather than every system's sums row
__global__ void k_final_b(const SysTask* __restrict__ tasks, int nv, int off, int use_gram,
                          double* compact) {
    const SysTask* t = tasks + blockIdx.x;
    const double* part = use_gram ? t->gpartial : t->partial;
    const int np = use_gram ? t->gparts : t->nparts;
    // the system sums (off 32) carry the muphi! domain failure of one system as a NaN
    // field-error sum: the batch drops it as diverged (log(phi / Q) of a negative ratio
    // raises in the reference, src/system/system.jl:251)
    bool bad = false;
    if (off == 32 && t->ra.canonical)
        for (int i = 0; i < t->nsp; ++i)
            if (t->sc->phi[i] / t->sc->Q[i] < 0.0) bad = true;
    for (int k = threadIdx.x; k < nv; k += blockDim.x) {
        double s = 0.0;
        for (int p = 0; p < np; ++p) s += part[(size_t)p * nv + k];
        if (bad && k == 0) s = __builtin_nan("");
        t->sums[off + k] = s;
        if (compact) compact[(size_t)blockIdx.x * nv + k] = s;
    }
}

__global__ void k_mean_subtract_b(const SysTask* __restrict__ tasks) {
    const SysTask* t = tasks + blockIdx.y;
    const ResidualArgs& a = t->ra;
    if (!a.canonical) return;
    const size_t n = t->n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        for (int m = 0; m < a.nmon; ++m) {
            const double mw = t->sums[m] / (double)n, mr = t->sums[PF_MAXMON + m] / (double)n;
            a.w[(size_t)m * n + i] -= mw;
            a.res[(size_t)m * n + i] -= mr;
        }
    }
}

void launch_residuals_b(const SysTask* d, int nsys, int maxparts, size_t nmax, hipStream_t st) {
    hipLaunchKernelGGL(k_residuals_b, dim3(maxparts, nsys), dim3(RB), 0, st, d);
    hipLaunchKernelGGL(k_final_b, dim3(nsys), dim3(64), 0, st, d, 2 * PF_MAXMON, 0, 0, nullptr);
    const int g = (int)std::min<size_t>((nmax + 255) / 256, 1024);
    hipLaunchKernelGGL(k_mean_subtract_b, dim3(g, nsys), dim3(256), 0, st, d);
    PF_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(RB) k_sys_reduce_b(const SysTask* __restrict__ tasks) {
    const SysTask* t = tasks + blockIdx.y;
    if ((int)blockIdx.x >= t->nparts) return;
    const ResidualArgs& a = t->ra;
    const size_t n = t->n;
    double v[5] = {0, 0, 0, 0, 0};
    const int nm = a.nmon;
    for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)t->nparts * RB) {
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
    wg_store<5>(v, t->partial + (size_t)blockIdx.x * 5);
}

void launch_sys_reduce_b(const SysTask* d, int nsys, int maxparts, double* compact, hipStream_t st) {
    hipLaunchKernelGGL(k_sys_reduce_b, dim3(maxparts, nsys), dim3(RB), 0, st, d);
    hipLaunchKernelGGL(k_final_b, dim3(nsys), dim3(64), 0, st, d, 5, 32, 0, compact);
    PF_HIP(hipGetLastError());
}

template <int H>
__global__ void __launch_bounds__(RB) k_gram_b(const SysTask* __restrict__ tasks,
                                               const StepTask* __restrict__ steps) {
    const SysTask* t = tasks + blockIdx.y;
    const StepTask* st = steps + blockIdx.y;
    if ((int)blockIdx.x >= t->gparts) return;
    constexpr int NU = H * (H + 1) / 2;
    double acc[NU + H];
#pragma unroll
    for (int k = 0; k < NU + H; ++k) acc[k] = 0.0;
    const size_t N = t->n * t->ra.nmon;
    const double* r = t->ra.res;
    const HistRing g = st->ring;
    const double* hp[H];
#pragma unroll
    for (int m = 0; m < H; ++m) hp[m] = g.r + ring_off(g, m);
    for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < N; i += (size_t)t->gparts * RB) {
        const double ri = r[i];
        double d[H];
#pragma unroll
        for (int m = 0; m < H; ++m) d[m] = ri - hp[m][i];
        int k = 0;
#pragma unroll
        for (int m = 0; m < H; ++m)
#pragma unroll
            for (int nn = m; nn < H; ++nn) acc[k++] += d[m] * d[nn];
#pragma unroll
        for (int m = 0; m < H; ++m) acc[NU + m] += d[m] * ri;
    }
    wg_store<NU + H>(acc, t->gpartial + (size_t)blockIdx.x * (NU + H));
}

template <int H>
static void gram_b(const SysTask* d, const StepTask* s, int nsys, int maxparts, double* compact,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_gram_b<H>, dim3(maxparts, nsys), dim3(RB), 0, st, d, s);
    hipLaunchKernelGGL(k_final_b, dim3(nsys), dim3(64), 0, st, d, H * (H + 1) / 2 + H, 64, 1, compact);
}

void launch_gram_b(int h, const SysTask* d, const StepTask* s, int nsys, int maxparts,
                   double* compact, hipStream_t st) {
    switch (h) {
#define PF_G(H) case H: gram_b<H>(d, s, nsys, maxparts, compact, st); break;
        PF_G(1) PF_G(2) PF_G(3) PF_G(4) PF_G(5) PF_G(6) PF_G(7) PF_G(8)
        PF_G(9) PF_G(10) PF_G(11) PF_G(12) PF_G(13) PF_G(14) PF_G(15) PF_G(16)
#undef PF_G
        default: throw Error(-2, "launch_gram_b: history depth above PF_MAXHIST");
    }
    PF_HIP(hipGetLastError());
}

// Any depth: 8 x 8 tiles of the Gram matrix of x_m = r - r_m (m < h) and x_h = r, as
// k_gram_tiled does for one system; grid (parts, tiles, systems).
__global__ void __launch_bounds__(RB) k_gram_tiled_b(const SysTask* __restrict__ tasks,
                                                     const StepTask* __restrict__ steps,
                                                     double* partial) {
    const SysTask* t = tasks + blockIdx.z;
    const StepTask* st = steps + blockIdx.z;
    const HistRing g = st->ring;
    const int parts = gram_tiled_parts(g.N);
    if ((int)blockIdx.x >= parts) return;
    const int h = st->h;
    const int p = blockIdx.y;
    int J = 0;
    while ((J + 1) * (J + 2) / 2 <= p) ++J;
    const int I = p - J * (J + 1) / 2;
    const double* r = t->ra.res;
    const size_t N = g.N;
    double acc[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) acc[k] = 0.0;
    auto x = [&](int k, size_t i, double ri) {
        if (k < h) return ri - g.r[ring_off(g, k) + i];
        return k == h ? ri : 0.0;
    };
    for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < N; i += (size_t)parts * RB) {
        const double ri = r[i];
        double a[8], b[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            a[q] = x(8 * I + q, i, ri);
            b[q] = x(8 * J + q, i, ri);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[q * 8 + u] += a[q] * b[u];
    }
    const size_t row = ((size_t)blockIdx.z * gridDim.x + blockIdx.x) * gridDim.y + p;
    wg_store<64>(acc, partial + row * 64);
}

// compact[sys * nv + k] = sum over the system's parts (in order) of its partial rows
__global__ void k_gram_tiled_final_b(const StepTask* __restrict__ steps,
                                     const double* __restrict__ partial, int maxparts, int nv,
                                     double* compact) {
    const double* base = partial + (size_t)blockIdx.x * maxparts * nv;
    const int parts = gram_tiled_parts(steps[blockIdx.x].ring.N);
    for (int k = threadIdx.x; k < nv; k += blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < parts; ++q) s += base[(size_t)q * nv + k];
        compact[(size_t)blockIdx.x * nv + k] = s;
    }
}

void launch_gram_tiled_b(int h, const SysTask* d, const StepTask* s, int nsys, int parts,
                         double* partial, double* compact, hipStream_t st) {
    const int nt = gram_tiles(h);
    hipLaunchKernelGGL(k_gram_tiled_b, dim3(parts, nt, nsys), dim3(RB), 0, st, d, s, partial);
    hipLaunchKernelGGL(k_gram_tiled_final_b, dim3(nsys), dim3(256), 0, st, s, partial, parts,
                       nt * 64, compact);
    PF_HIP(hipGetLastError());
}

// Anderson mix or Euler step, residual and field history writes (src/updaters/anderson.jl:
// 138-147, 211-224).
__global__ void k_mix_b(const SysTask* __restrict__ tasks, const StepTask* __restrict__ steps) {
    const SysTask* t = tasks + blockIdx.y;
    const StepTask* s = steps + blockIdx.y;
    const size_t N = t->n * t->ra.nmon;
    double* w = t->ra.w;
    const double* r = t->ra.res;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x) {
        const double ri = r[i];
        if (s->r_out) s->r_out[i] = ri;
        double o;
        if (s->euler) {
            o = w[i] + s->lam_euler * ri;
        } else {
            double WW = w[i];
            double DD = ri;
            const double fi = s->f_iter[i];
            const HistRing& g = s->ring;
            for (int nn = 0; nn < s->h; ++nn) {
                const size_t off = ring_off(g, nn) + i;
                WW += s->C[nn] * (g.w[off] - fi);
                DD += s->C[nn] * (g.r[off] - ri);
            }
            o = WW + s->lam * DD;
        }
        w[i] = o;
        if (s->f_out) s->f_out[i] = o;
    }
}

// Momentum family step of every system (blockIdx.y): the update of src/updaters/
// momentum.jl:86-139 from this step's residual, then w += lam * step, per point in the
// one-system order (momentum update, then the axpy)
__global__ void k_momentum_b(const MomTask* __restrict__ t, int method, double lam, double b1,
                             double b2) {
    const MomTask m = t[blockIdx.y];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m.N;
         i += (size_t)gridDim.x * blockDim.x) {
        momentum_point(method, m.mt, m.vt, m.step, m.r, b1, b2, i);
        m.w[i] += lam * m.step[i];
    }
}

void launch_momentum_b(const MomTask* d, int nsys, int method, double lam, double b1, double b2,
                       size_t Nmax, hipStream_t st) {
    const int g = (int)std::min<size_t>((Nmax + 255) / 256, 1024);
    hipLaunchKernelGGL(k_momentum_b, dim3(g, nsys), dim3(256), 0, st, d, method, lam, b1, b2);
    PF_HIP(hipGetLastError());
}

// The pointwise parts of NAG (src/updaters/momentum.jl:141-165) and of the Euler
// predictor-corrector (src/updaters/euler.jl:36-66) for every system (blockIdx.y), each a
// fusion of the one-system copies and axpys with the same sums per point:
//   NAG_LOOK    tmp = w; w += lam mt                      (the look-ahead fields)
//   NAG_STEP    momentum update with b1 = gamma; w = tmp; w += lam step
//   EULER_PRED  w += lam r; tmp = w; tmp += b1 r          (b1 = lam / 2)
//   EULER_CORR  tmp += b1 r; w = tmp
__global__ void k_stage_b(const MomTask* __restrict__ t, int op, double lam, double b1,
                          double b2) {
    const MomTask m = t[blockIdx.y];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m.N;
         i += (size_t)gridDim.x * blockDim.x) {
        if (op == STAGE_NAG_LOOK) {
            const double wi = m.w[i];
            m.tmp[i] = wi;
            m.w[i] = wi + lam * m.mt[i];
        } else if (op == STAGE_NAG_STEP) {
            momentum_point(0, m.mt, m.vt, m.step, m.r, b1, b2, i);
            m.w[i] = m.tmp[i] + lam * m.step[i];
        } else if (op == STAGE_EULER_PRED) {
            const double ri = m.r[i];
            const double wp = m.w[i] + lam * ri;
            m.w[i] = wp;
            m.tmp[i] = wp + b1 * ri;
        } else {
            const double v = m.tmp[i] + b1 * m.r[i];
            m.tmp[i] = v;
            m.w[i] = v;
        }
    }
}

void launch_stage_b(const MomTask* d, int nsys, int op, double lam, double b1, double b2,
                    size_t Nmax, hipStream_t st) {
    const int g = (int)std::min<size_t>((Nmax + 255) / 256, 1024);
    hipLaunchKernelGGL(k_stage_b, dim3(g, nsys), dim3(256), 0, st, d, op, lam, b1, b2);
    PF_HIP(hipGetLastError());
}

void launch_mix_b(const SysTask* d, const StepTask* s, int nsys, size_t Nmax, hipStream_t st) {
    const int g = (int)std::min<size_t>((Nmax + 255) / 256, 1024);
    hipLaunchKernelGGL(k_mix_b, dim3(g, nsys), dim3(256), 0, st, d, s);
    PF_HIP(hipGetLastError());
}

template void launch_lw_fill_b<float>(const LwTask<float>*, int, const FillTask<float>*, int,
                                      size_t, hipStream_t);
template void launch_lw_fill_b<double>(const LwTask<double>*, int, const FillTask<double>*, int,
                                       size_t, hipStream_t);

}  // namespace pf
