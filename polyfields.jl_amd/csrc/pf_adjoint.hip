// Reverse sweeps of the pseudo-spectral propagator: the adjoint of density!(species)
// (src/species/multiblock.jl:125-231 with the RK2 step of src/system/mde.jl:65-81).
//
// The Strang operator P = LW F^-1 LD F LW is real-symmetric, so the adjoint of
// q_{s+1} = P q_s carries a-bar backwards through the same operator, with the Simpson
// sources srcc_b w_s rbar_{mon(b)} a'_s (a' the partner propagator) added at every slice.
// One adjoint step of forward step (s_in -> s_out), a-bar the adjoint at s_out:
//   wbar_b   += wco_b (a-bar a_out + a_in P a-bar)              wco_b = -(ds_b v_b / 2)
//   gpart_b  += sum_k m_k Re(conj(A) B) LD_k {dksq_theta(k), ksq(k)}
//   a-bar    <- P a-bar + source(s_in)
// with A = rfft(LW a-bar), B = rfft(LW a_in) and m_k the multiplicity of half-spectrum
// entry k in the full spectrum (1 on the kx = 0 and kx = nx/2 planes, 2 elsewhere).
//
// Every grid (1-D, 2-D, 3-D) runs as xy-slab + z-pencil launches; a 1-D or 2-D grid is
// one slab with nz = 1.  Slab launch, mode 0: form a-bar at the sweep's first slice and
// the transforms of the first step.  Pencil launch t: z transforms, the k-space sums and
// the LD multiply of step t.  Slab launch, mode 1, t: finish step t and start step t+1.
#include <algorithm>

#include "pf_fft_core.h"
#include "pf_fft_fixed.h"

namespace pf {

static constexpr int NG = PF_MAXPAR + 1;   // k-space sums per block: dksq_theta..., ksq

__device__ __forceinline__ double simpson_w(int s, int bgn, int end) {
    if (s == bgn || s == end) return 1.0;
    return ((s - bgn) & 1) ? 4.0 : 2.0;
}

// Source coefficient at slice s, grid point i: sum over the blocks that hold slice s of
// srcc_b w^b_s rbar_{mon(b)}(i) (the Simpson rule of src/species/multiblock.jl:182-216).
template <typename T>
__device__ __forceinline__ double src_coef(const Lane<T>* L, int s, size_t i, size_t n) {
    double c = 0.0;
    for (int b = 0; b < L->nblk; ++b) {
        const int bgn = L->begin[b], end = L->begin[b + 1];
        if (s >= bgn && s <= end)
            c += (L->srcc[b] * simpson_w(s, bgn, end)) * L->rbar[(size_t)L->blk_slot[b] * n + i];
    }
    return c;
}

template <typename T>
__device__ __forceinline__ const T* other_slice(const Lane<T>* L, int s, size_t n) {
    const int so = L->other_rev ? (L->nsteps + 2 - s) : s;
    return L->other + (size_t)(so - 1) * n;
}

// plane forward transform of the packed plane in S, written to the half spectrum kb
template <typename T>
__device__ void plane_to_spectrum(cpx<T>* S, cpx<T>* kb, const Lane<T>* L) {
    const int M = L->nx >> 1, ld = M + 1, nxh = M + 1, ny = L->ny, kplane = nxh * ny;
    __syncthreads();
    plane_forward(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
    for (int i = threadIdx.x; i < kplane; i += blockDim.x) {
        const int y = i / nxh;
        kb[i] = S[y * ld + (i - y * nxh)];
    }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_adj_slab(const Lane<T>* __restrict__ lanes, int nz,
                                                  int t, int nextB) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* S = reinterpret_cast<cpx<T>*>(smem_raw);
    const int lane = blockIdx.x / nz;
    const int z = blockIdx.x - lane * nz;
    const Lane<T>* L = lanes + lane;
    const int NS = L->nsteps;
    if (t >= NS) return;
    const int nx = L->nx, ny = L->ny, M = nx >> 1, ld = M + 1;
    const int plane = nx * ny, np = plane >> 1, nxh = M + 1, kplane = nxh * ny;
    const size_t n = (size_t)plane * nz, off = (size_t)z * plane;
    cpx<T>* kA = L->kbuf + (size_t)z * kplane;
    cpx<T>* kB = L->kb2 + (size_t)z * kplane;
    int s_in;
    int b = step_block(L, NS - 1 - t, &s_in);
    if (MODE == 2) {
        // B = rfft(LW a_in) of step t alone (its input slice was not in memory when the
        // previous step finished: checkpointed sweeps)
        const cpx<T>* ai = reinterpret_cast<const cpx<T>*>(L->fwd + (size_t)(s_in - 1) * n + off);
        const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b] + off);
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            const cpx<T> q = ai[i], w = lw[i];
            S[y * ld + (i - y * M)] = {q.x * w.x, q.y * w.y};
        }
        plane_to_spectrum(S, kB, L);
        return;
    }
    if (MODE == 0) {
        // a-bar at the first slice of the reverse sweep: dF/dQ part + Simpson source
        const int s_out = s_in + L->dir;
        const cpx<T>* ao = reinterpret_cast<const cpx<T>*>(L->fwd + (size_t)(s_out - 1) * n + off);
        const cpx<T>* ot = reinterpret_cast<const cpx<T>*>(other_slice(L, s_out, n) + off);
        const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b] + off);
        double* wb = L->wbar + (size_t)b * n + off;
        const double wc = L->wco[b];
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            const size_t r0 = 2 * (size_t)i;
            const cpx<T> o = ot[i], a = ao[i], w = lw[i];
            const double a0 = L->qinit + src_coef(L, s_out, off + r0, n) * (double)o.x;
            const double a1 = L->qinit + src_coef(L, s_out, off + r0 + 1, n) * (double)o.y;
            wb[r0] += wc * a0 * (double)a.x;
            wb[r0 + 1] += wc * a1 * (double)a.y;
            S[y * ld + (i - y * M)] = {(T)a0 * w.x, (T)a1 * w.y};
        }
    } else {
        // finish step t: P a-bar = LW * irfft(LD A)
        for (int i = threadIdx.x; i < kplane; i += blockDim.x) {
            const int y = i / nxh;
            S[y * ld + (i - y * nxh)] = kA[i];
        }
        __syncthreads();
        plane_inverse(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
        const T scale = T(2.0 / (double)n);
        const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b] + off);
        const cpx<T>* ai = reinterpret_cast<const cpx<T>*>(L->fwd + (size_t)(s_in - 1) * n + off);
        const cpx<T>* ot = reinterpret_cast<const cpx<T>*>(other_slice(L, s_in, n) + off);
        double* wb = L->wbar + (size_t)b * n + off;
        const double wc = L->wco[b];
        const bool more = t + 1 < NS;
        int s2 = s_in, b2 = b;
        if (more) b2 = step_block(L, NS - 2 - t, &s2);
        double* wb2 = L->wbar + (size_t)b2 * n + off;
        const double wc2 = L->wco[b2];
        const cpx<T>* lw2 = reinterpret_cast<const cpx<T>*>(L->lw[b2] + off);
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            const size_t r0 = 2 * (size_t)i;
            cpx<T>& v = S[y * ld + (i - y * M)];
            const cpx<T> w = lw[i], a = ai[i];
            const double p0 = (double)((v.x * scale) * w.x);
            const double p1 = (double)((v.y * scale) * w.y);
            double g0 = wc * (double)a.x * p0, g1 = wc * (double)a.y * p1;
            if (more) {
                // a-bar at s_in, which is the output slice of the next adjoint step
                const cpx<T> o = ot[i], w2 = lw2[i];
                const double a0 = p0 + src_coef(L, s_in, off + r0, n) * (double)o.x;
                const double a1 = p1 + src_coef(L, s_in, off + r0 + 1, n) * (double)o.y;
                if (b2 == b) {
                    g0 += wc2 * a0 * (double)a.x;
                    g1 += wc2 * a1 * (double)a.y;
                } else {
                    wb2[r0] += wc2 * a0 * (double)a.x;
                    wb2[r0 + 1] += wc2 * a1 * (double)a.y;
                }
                v = {(T)a0 * w2.x, (T)a1 * w2.y};
            }
            wb[r0] += g0;
            wb[r0 + 1] += g1;
        }
        if (!more) return;
        b = b2;
        s_in = s2;
    }
    // A = rfft(LW a-bar) of the step about to run, then B = rfft(LW a_in)
    plane_to_spectrum(S, kA, L);
    if (!nextB) return;
    __syncthreads();
    {
        const cpx<T>* ai = reinterpret_cast<const cpx<T>*>(L->fwd + (size_t)(s_in - 1) * n + off);
        const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b] + off);
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            const cpx<T> q = ai[i], w = lw[i];
            S[y * ld + (i - y * M)] = {q.x * w.x, q.y * w.y};
        }
    }
    plane_to_spectrum(S, kB, L);
}

template <int NV>
__device__ void wg_sum(double* v, double* out) {
    __shared__ double red[NV * 16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) red[k * 16 + wave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += red[k * 16 + w];
            out[k] = s;
        }
}

template <typename T>
__global__ void __launch_bounds__(256) k_adj_pencil(const Lane<T>* __restrict__ lanes, int ny,
                                                    int ntiles, int tile, int t) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int per = ny * ntiles;
    const int lane = blockIdx.x / per;
    const int r = blockIdx.x - lane * per;
    const int ky = r / ntiles, tl = r - ky * ntiles;
    const Lane<T>* L = lanes + lane;
    const int NS = L->nsteps;
    if (t >= NS) return;
    const int nz = L->nz, M = L->nx >> 1, nxh = M + 1;
    cpx<T>* PA = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* PB = PA + (size_t)nz * tile;
    const int kx0 = tl * tile;
    const int tw = min(tile, nxh - kx0);
    int s_in;
    const int b = step_block(L, NS - 1 - t, &s_in);
    cpx<T>* kA = L->kbuf + (size_t)ky * nxh + kx0;
    const cpx<T>* kB = L->kb2 + (size_t)ky * nxh + kx0;
    const size_t zstride = (size_t)ny * nxh;
    for (int i = threadIdx.x; i < nz * tw; i += blockDim.x) {
        const int z = i / tw, c = i - z * tw;
        PA[z * tile + c] = kA[z * zstride + c];
        PB[z * tile + c] = kB[z * zstride + c];
    }
    __syncthreads();
    lds_fft<T, false>(PA, L->pz, tw, tile, 1, L->twz, 1);
    lds_fft<T, false>(PB, L->pz, tw, tile, 1, L->twz, 1);
    const double mc = -L->ldc[b];
    const double jy = (double)fold(ky, ny);
    const int npar = L->nparams;
    double acc[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) acc[k] = 0.0;
    for (int i = threadIdx.x; i < nz * tw; i += blockDim.x) {
        const int z = i / tw, c = i - z * tw;
        const int kx = kx0 + c;
        const double jz = (double)fold(z, nz);
        const double k2 = ksq_at(L->G, (double)kx, jy, jz);
        const T f = exp_t<T>(mc * k2);
        cpx<T>& A = PA[z * tile + c];
        const cpx<T> B = PB[z * tile + c];
        const double m = (kx == 0 || kx == M) ? 1.0 : 2.0;
        const double g = m * ((double)A.x * (double)B.x + (double)A.y * (double)B.y) * (double)f;
        for (int k = 0; k < npar; ++k)
            acc[k] += g * dksq_at(L->dGG, npar, k, L->dim, (double)kx, jy, jz);
        acc[PF_MAXPAR] += g * k2;
        A.x *= f;
        A.y *= f;
    }
    __syncthreads();
    lds_fft<T, true>(PA, L->pz, tw, tile, 1, L->twz, 1);
    for (int i = threadIdx.x; i < nz * tw; i += blockDim.x) {
        const int z = i / tw, c = i - z * tw;
        kA[z * zstride + c] = PA[z * tile + c];
    }
    double out[NG];
    wg_sum<NG>(acc, out);
    if (threadIdx.x == 0) {
        double* gp = L->gpart + ((size_t)r * L->nblk + b) * NG;
        for (int k = 0; k < NG; ++k) gp[k] += out[k];
    }
}

// ---------------------------------------------------------------------------------
// Compile-time forms of the two reverse kernels for the production plane and line sizes:
// the same arithmetic, order of operations and Simpson sources as k_adj_slab /
// k_adj_pencil, with fixed radix plans, twiddles staged in LDS and unrolled loops.
//
// SPLIT (modes 0 and 1): grid (2 nz, lanes).  Workgroups nz.. make the B spectrum of the
// step whose A spectrum the first nz make (step t in mode 0, t + 1 in mode 1), as mode 2
// does, so that a launch of few lanes has twice the workgroups; the same arithmetic.
template <typename T, int MODE, int NX, int NY, int NT, bool SPLIT = false>
__global__ void __launch_bounds__(NT) k_adj_slab_fx(const Lane<T>* __restrict__ lanes, int nz,
                                                    int t, int nextB) {
    constexpr int M = NX / 2, LD = M + 1, PLANE = NX * NY, KPLANE = LD * NY;
    __shared__ cpx<T> S[KPLANE];
    __shared__ cpx<T> TWX[NX];
    __shared__ cpx<T> TWY[NY];
    // grid (nz, lanes): plane z of lane blockIdx.y (SPLIT: (2 nz, lanes), B part above nz)
    const bool bpart = SPLIT && (int)blockIdx.x >= nz;
    const int z = bpart ? (int)blockIdx.x - nz : (int)blockIdx.x;
    const Lane<T>* L = lanes + blockIdx.y;
    const int NS = L->nsteps;
    if (t >= NS) return;
    const size_t n = (size_t)PLANE * nz, off = (size_t)z * PLANE;
    cpx<T>* kA = L->kbuf + (size_t)z * KPLANE;
    cpx<T>* kB = L->kb2 + (size_t)z * KPLANE;
    // the twiddles are copied after the first plane's loads are out (before the first
    // barrier), so that those loads do not wait behind them
    auto twiddles_in = [&] {
        lds_copy_in<T, NX, NT>(TWX, L->twx);
        lds_copy_in<T, NY, NT>(TWY, L->twy);
    };
    int s_in;
    int b = step_block(L, NS - 1 - t, &s_in);
    auto to_spectrum = [&](cpx<T>* kb) {
        __syncthreads();
        cplane_forward<T, NX, NY, NT>(S, TWX, TWY);
#pragma unroll
        for (int i0 = 0; i0 < KPLANE; i0 += NT)
            if (i0 + (int)threadIdx.x < KPLANE) kb[i0 + threadIdx.x] = S[i0 + threadIdx.x];
    };
    auto lw_times_fwd = [&](int sl, int blk) {   // S <- LW_blk * a_sl
        const cpx<T>* ai = reinterpret_cast<const cpx<T>*>(L->fwd + (size_t)(sl - 1) * n + off);
        const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[blk] + off);
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> q = ai[i], w = lw[i];
            S[o] = {q.x * w.x, q.y * w.y};
        });
    };
    if constexpr (MODE == 2) {
        lw_times_fwd(s_in, b);
        twiddles_in();
        to_spectrum(kB);
        return;
    }
    if constexpr (SPLIT) {
        if (bpart) {
            const int tb = MODE == 0 ? t : t + 1;
            if (!nextB || tb >= NS) return;
            b = step_block(L, NS - 1 - tb, &s_in);
            lw_times_fwd(s_in, b);
            twiddles_in();
            to_spectrum(kB);
            return;
        }
        nextB = 0;
    }
    if constexpr (MODE == 0) {
        const int s_out = s_in + L->dir;
        const cpx<T>* ao = reinterpret_cast<const cpx<T>*>(L->fwd + (size_t)(s_out - 1) * n + off);
        const cpx<T>* ot = reinterpret_cast<const cpx<T>*>(other_slice(L, s_out, n) + off);
        const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b] + off);
        double* wb = L->wbar + (size_t)b * n + off;
        const double wc = L->wco[b];
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const size_t r0 = 2 * (size_t)i;
            const cpx<T> ov = ot[i], a = ao[i], w = lw[i];
            const double a0 = L->qinit + src_coef(L, s_out, off + r0, n) * (double)ov.x;
            const double a1 = L->qinit + src_coef(L, s_out, off + r0 + 1, n) * (double)ov.y;
            wb[r0] += wc * a0 * (double)a.x;
            wb[r0 + 1] += wc * a1 * (double)a.y;
            S[o] = {(T)a0 * w.x, (T)a1 * w.y};
        });
        twiddles_in();
    } else {
        lds_copy_in<T, KPLANE, NT>(S, kA);
        twiddles_in();
        __syncthreads();
        cplane_inverse<T, NX, NY, NT>(S, TWX, TWY);
        const T scale = T(2.0 / (double)n);
        const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b] + off);
        const cpx<T>* ai = reinterpret_cast<const cpx<T>*>(L->fwd + (size_t)(s_in - 1) * n + off);
        const cpx<T>* ot = reinterpret_cast<const cpx<T>*>(other_slice(L, s_in, n) + off);
        double* wb = L->wbar + (size_t)b * n + off;
        const double wc = L->wco[b];
        const bool more = t + 1 < NS;
        int s2 = s_in, b2 = b;
        if (more) b2 = step_block(L, NS - 2 - t, &s2);
        double* wb2 = L->wbar + (size_t)b2 * n + off;
        const double wc2 = L->wco[b2];
        const cpx<T>* lw2 = reinterpret_cast<const cpx<T>*>(L->lw[b2] + off);
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const size_t r0 = 2 * (size_t)i;
            cpx<T>& v = S[o];
            const cpx<T> w = lw[i], a = ai[i];
            const double p0 = (double)((v.x * scale) * w.x);
            const double p1 = (double)((v.y * scale) * w.y);
            double g0 = wc * (double)a.x * p0, g1 = wc * (double)a.y * p1;
            if (more) {
                const cpx<T> ov = ot[i], w2 = lw2[i];
                const double a0 = p0 + src_coef(L, s_in, off + r0, n) * (double)ov.x;
                const double a1 = p1 + src_coef(L, s_in, off + r0 + 1, n) * (double)ov.y;
                if (b2 == b) {
                    g0 += wc2 * a0 * (double)a.x;
                    g1 += wc2 * a1 * (double)a.y;
                } else {
                    wb2[r0] += wc2 * a0 * (double)a.x;
                    wb2[r0 + 1] += wc2 * a1 * (double)a.y;
                }
                v = {(T)a0 * w2.x, (T)a1 * w2.y};
            }
            wb[r0] += g0;
            wb[r0 + 1] += g1;
        });
        if (!more) return;
        b = b2;
        s_in = s2;
    }
    to_spectrum(kA);
    if (!nextB) return;
    __syncthreads();
    lw_times_fwd(s_in, b);
    to_spectrum(kB);
}

// Reverse z pass over KB ky rows: forward z of A and B, the k-space sums and LD on A,
// inverse z of A.  Partial-sum row r = the workgroup's index within its lane.
template <typename T, int NZ, int NXH, int KB, int NT>
__global__ void __launch_bounds__(NT) k_adj_pencil_fx(const Lane<T>* __restrict__ lanes, int ny,
                                                      int t) {
    constexpr int C = KB * NXH;
    __shared__ cpx<T> PA[NZ * C];
    __shared__ cpx<T> PB[NZ * C];
    __shared__ cpx<T> TWZ[NZ];
    // grid (ny / KB, lanes); partial-sum row r = blockIdx.x
    const int r = blockIdx.x;
    const int ky0 = r * KB;
    const Lane<T>* L = lanes + blockIdx.y;
    const int NS = L->nsteps;
    if (t >= NS) return;
    int s_in;
    const int b = step_block(L, NS - 1 - t, &s_in);
    constexpr int M = NXH - 1;
    cpx<T>* kA = L->kbuf + (size_t)ky0 * NXH;
    const cpx<T>* kB = L->kb2 + (size_t)ky0 * NXH;
    const size_t zstride = (size_t)ny * NXH;
#pragma unroll
    for (int i0 = 0; i0 < NZ * C; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < NZ * C) {
            const size_t g = (size_t)(i / C) * zstride + (i % C);
            PA[i] = kA[g];
            PB[i] = kB[g];
        }
    }
    lds_copy_in<T, NZ, NT>(TWZ, L->twz);
    __syncthreads();
    cfft<T, NZ, C, C, 1, NT, false, 1>(PA, TWZ);
    cfft<T, NZ, C, C, 1, NT, false, 1>(PB, TWZ);
    const double mc = -L->ldc[b];
    const int npar = L->nparams;
    double acc[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) acc[k] = 0.0;
#pragma unroll
    for (int i0 = 0; i0 < NZ * C; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < NZ * C) {
            const int z = i / C, c = i % C;
            const int ky = ky0 + c / NXH, kx = c % NXH;
            const double jy = (double)fold(ky, ny), jz = (double)fold(z, NZ);
            const double k2 = ksq_at(L->G, (double)kx, jy, jz);
            const T f = exp_t<T>(mc * k2);
            cpx<T>& A = PA[i];
            const cpx<T> B = PB[i];
            const double m = (kx == 0 || kx == M) ? 1.0 : 2.0;
            const double g = m * ((double)A.x * (double)B.x + (double)A.y * (double)B.y) * (double)f;
            for (int k = 0; k < npar; ++k)
                acc[k] += g * dksq_at(L->dGG, npar, k, L->dim, (double)kx, jy, jz);
            acc[PF_MAXPAR] += g * k2;
            A.x *= f;
            A.y *= f;
        }
    }
    __syncthreads();
    cfft<T, NZ, C, C, 1, NT, true, 1>(PA, TWZ);
#pragma unroll
    for (int i0 = 0; i0 < NZ * C; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < NZ * C) kA[(size_t)(i / C) * zstride + (i % C)] = PA[i];
    }
    double out[NG];
    wg_sum<NG>(acc, out);
    if (threadIdx.x == 0) {
        double* gp = L->gpart + ((size_t)r * L->nblk + b) * NG;
        for (int k = 0; k < NG; ++k) gp[k] += out[k];
    }
}

// Reverse z pass over half of one ky row (columns [h CW, min((h + 1) CW, NXH)), CW =
// ceil(NXH / 2)), for launches of few lanes: twice the workgroups of the one-row form,
// the same arithmetic per column.  Partial-sum row r = blockIdx.x = 2 ky + h.  The second
// half's missing column is loaded as a duplicate of a valid one, and neither summed nor
// stored.
template <typename T, int NZ, int NXH, int NT>
__global__ void __launch_bounds__(NT) k_adj_pencil_cs(const Lane<T>* __restrict__ lanes, int ny,
                                                      int t) {
    constexpr int CW = (NXH + 1) / 2, CNT = NZ * CW, M = NXH - 1;
    __shared__ cpx<T> PA[CNT];
    __shared__ cpx<T> PB[CNT];
    __shared__ cpx<T> TWZ[NZ];
    const int r = blockIdx.x, ky = r >> 1, h = r & 1;
    const int c0 = h * CW, cv = h ? NXH - CW : CW;
    const Lane<T>* L = lanes + blockIdx.y;
    const int NS = L->nsteps;
    if (t >= NS) return;
    int s_in;
    const int b = step_block(L, NS - 1 - t, &s_in);
    cpx<T>* kA = L->kbuf + (size_t)ky * NXH + c0;
    const cpx<T>* kB = L->kb2 + (size_t)ky * NXH + c0;
    const size_t zstride = (size_t)ny * NXH;
#pragma unroll
    for (int i0 = 0; i0 < CNT; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < CNT) {
            const int c = i % CW;
            const size_t g = (size_t)(i / CW) * zstride + (c < cv ? c : cv - 1);
            PA[i] = kA[g];
            PB[i] = kB[g];
        }
    }
    lds_copy_in<T, NZ, NT>(TWZ, L->twz);
    __syncthreads();
    cfft<T, NZ, CW, CW, 1, NT, false, 1>(PA, TWZ);
    cfft<T, NZ, CW, CW, 1, NT, false, 1>(PB, TWZ);
    const double mc = -L->ldc[b];
    const int npar = L->nparams;
    const double jy = (double)fold(ky, ny);
    double acc[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) acc[k] = 0.0;
#pragma unroll
    for (int i0 = 0; i0 < CNT; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < CNT && i % CW < cv) {
            const int z = i / CW, kx = c0 + i % CW;
            const double jz = (double)fold(z, NZ);
            const double k2 = ksq_at(L->G, (double)kx, jy, jz);
            const T f = exp_t<T>(mc * k2);
            cpx<T>& A = PA[i];
            const cpx<T> B = PB[i];
            const double m = (kx == 0 || kx == M) ? 1.0 : 2.0;
            const double g = m * ((double)A.x * (double)B.x + (double)A.y * (double)B.y) * (double)f;
            for (int k = 0; k < npar; ++k)
                acc[k] += g * dksq_at(L->dGG, npar, k, L->dim, (double)kx, jy, jz);
            acc[PF_MAXPAR] += g * k2;
            A.x *= f;
            A.y *= f;
        }
    }
    __syncthreads();
    cfft<T, NZ, CW, CW, 1, NT, true, 1>(PA, TWZ);
#pragma unroll
    for (int i0 = 0; i0 < CNT; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < CNT && i % CW < cv) kA[(size_t)(i / CW) * zstride + i % CW] = PA[i];
    }
    double out[NG];
    wg_sum<NG>(acc, out);
    if (threadIdx.x == 0) {
        double* gp = L->gpart + ((size_t)r * L->nblk + b) * NG;
        for (int k = 0; k < NG; ++k) gp[k] += out[k];
    }
}

template <typename T>
bool fixed_adj_slab(const Lane<T>* d, int nl, int nx, int ny, int nz, int mode, int t,
                    hipStream_t st, int nextB) {
    if (schedule(sizeof(T), nx, ny, nz).slab == 0) return false;
#define PF_AS(X, Y, NT)                                                                     \
    if (nx == X && ny == Y) {                                                               \
        if (mode == 0)                                                                      \
            hipLaunchKernelGGL((k_adj_slab_fx<T, 0, X, Y, NT>), dim3(nz, nl), dim3(NT), 0,   \
                               st, d, nz, t, nextB);                                        \
        else if (mode == 1)                                                                 \
            hipLaunchKernelGGL((k_adj_slab_fx<T, 1, X, Y, NT>), dim3(nz, nl), dim3(NT), 0,   \
                               st, d, nz, t, nextB);                                        \
        else                                                                                \
            hipLaunchKernelGGL((k_adj_slab_fx<T, 2, X, Y, NT>), dim3(nz, nl), dim3(NT), 0,   \
                               st, d, nz, t, 1);                                            \
        PF_HIP(hipGetLastError());                                                          \
        return true;                                                                        \
    }
    PF_AS(32, 32, 256)
    // 64x64 planes in launches with at most one plane per CU (a few lanes): the A and B
    // spectra from separate workgroups (SPLIT), twice the workgroups
    // (FP64: 1024 threads; measured on MI355X, 64^3 gradient: 18.5 -> 16.3 us per launch
    // against 512)
    if (nx == 64 && ny == 64 && mode != 2 && nl * nz <= device_cus()) {
        constexpr int NT = sizeof(T) == 8 ? 1024 : 256;
        if (mode == 0)
            hipLaunchKernelGGL((k_adj_slab_fx<T, 0, 64, 64, NT, true>), dim3(2 * nz, nl),
                               dim3(NT), 0, st, d, nz, t, nextB);
        else
            hipLaunchKernelGGL((k_adj_slab_fx<T, 1, 64, 64, NT, true>), dim3(2 * nz, nl),
                               dim3(NT), 0, st, d, nz, t, nextB);
        PF_HIP(hipGetLastError());
        return true;
    }
    // FP64 64x64: 512 threads per plane (and per ky row in the z pass) make a whole 64^3
    // gradient 15% faster on MI355X than 256; the LDS per workgroup is the same, the extra
    // waves hide load latency, and the results are bitwise the same.
    if constexpr (sizeof(T) == 8) { PF_AS(64, 64, 512) }
    PF_AS(64, 64, 256)
#undef PF_AS
    return false;
}

template <typename T>
int fixed_adj_pencil(const Lane<T>* d, int nl, int nx, int ny, int nz, int t, hipStream_t st) {
    if (schedule(sizeof(T), nx, ny, nz).pencil == 0) return 0;
#define PF_AP(X, Z, KB, NT)                                                                 \
    if (nx == X && nz == Z && ny % KB == 0) {                                               \
        hipLaunchKernelGGL((k_adj_pencil_fx<T, Z, X / 2 + 1, KB, NT>), dim3(ny / KB, nl),     \
                           dim3(NT), 0, st, d, ny, t);                                      \
        PF_HIP(hipGetLastError());                                                          \
        return ny / KB;                                                                     \
    }
    PF_AP(32, 32, 2, 256)
    // 64-point lines in launches with at most one ky row per CU: half rows (twice the
    // workgroups and partial-sum rows; the sums then run in another order)
    if (nx == 64 && nz == 64 && nl * ny <= device_cus()) {
        // 512 threads (measured on MI355X, 64^3 FP64 gradient: 17.2 -> 15.0 us per launch)
        hipLaunchKernelGGL((k_adj_pencil_cs<T, 64, 33, 512>), dim3(2 * ny, nl), dim3(512), 0,
                           st, d, ny, t);
        PF_HIP(hipGetLastError());
        return 2 * ny;
    }
    if constexpr (sizeof(T) == 8) { PF_AP(64, 64, 1, 512) }
    PF_AP(64, 64, 1, 256)
#undef PF_AP
    return 0;
}

template bool fixed_adj_slab<float>(const Lane<float>*, int, int, int, int, int, int, hipStream_t,
                                    int);
template bool fixed_adj_slab<double>(const Lane<double>*, int, int, int, int, int, int,
                                     hipStream_t, int);
template int fixed_adj_pencil<float>(const Lane<float>*, int, int, int, int, int, hipStream_t);
template int fixed_adj_pencil<double>(const Lane<double>*, int, int, int, int, int, hipStream_t);

template <typename Kern>
static void set_lds_adj(Kern k, int bytes) {
    if (bytes > 64 * 1024)
        PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
}

template <typename T>
void launch_adj_slab(const Lane<T>* d_lanes, int nlanes, int nz, int mode, int t, int lds_bytes,
                     hipStream_t st, int nextB) {
    if (mode == 0) {
        set_lds_adj(k_adj_slab<T, 0>, lds_bytes);
        hipLaunchKernelGGL((k_adj_slab<T, 0>), dim3(nlanes * nz), dim3(256), lds_bytes, st,
                           d_lanes, nz, t, nextB);
    } else if (mode == 1) {
        set_lds_adj(k_adj_slab<T, 1>, lds_bytes);
        hipLaunchKernelGGL((k_adj_slab<T, 1>), dim3(nlanes * nz), dim3(256), lds_bytes, st,
                           d_lanes, nz, t, nextB);
    } else {
        set_lds_adj(k_adj_slab<T, 2>, lds_bytes);
        hipLaunchKernelGGL((k_adj_slab<T, 2>), dim3(nlanes * nz), dim3(256), lds_bytes, st,
                           d_lanes, nz, t, 1);
    }
    PF_HIP(hipGetLastError());
}

template <typename T>
void launch_adj_pencil(const Lane<T>* d_lanes, int nlanes, int ny, int ntiles, int tile, int t,
                       int lds_bytes, hipStream_t st) {
    set_lds_adj(k_adj_pencil<T>, lds_bytes);
    hipLaunchKernelGGL(k_adj_pencil<T>, dim3(nlanes * ny * ntiles), dim3(256), lds_bytes, st,
                       d_lanes, ny, ntiles, tile, t);
    PF_HIP(hipGetLastError());
}

template void launch_adj_slab<float>(const Lane<float>*, int, int, int, int, int, hipStream_t,
                                     int);
template void launch_adj_slab<double>(const Lane<double>*, int, int, int, int, int, hipStream_t,
                                      int);
template void launch_adj_pencil<float>(const Lane<float>*, int, int, int, int, int, int,
                                       hipStream_t);
template void launch_adj_pencil<double>(const Lane<double>*, int, int, int, int, int, int,
                                        hipStream_t);

// ---------------------------------------------------------------------------------
// Grid-point helpers of the gradient.

// Dot products sum_i x_k[i] y_k[i] for k < a.k: one partial per (workgroup, k), then a
// fixed-order sum per k.
__global__ void __launch_bounds__(256) k_dots(DotArgs a, size_t n, double* partial, int nparts) {
    const int k = blockIdx.y;
    const double* x = a.x[k];
    const double* y = a.y[k];
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)nparts * 256)
        acc += x[i] * y[i];
    double v[1] = {acc}, o[1];
    wg_sum<1>(v, o);
    if (threadIdx.x == 0) partial[(size_t)k * nparts + blockIdx.x] = o[0];
}

__global__ void k_dots_finish(const double* partial, int nparts, double* out) {
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int p = 0; p < nparts; ++p) s += partial[(size_t)blockIdx.x * nparts + p];
        out[blockIdx.x] = s;
    }
}

void launch_dots(const DotArgs& a, size_t n, double* partial, int nparts, double* out,
                 hipStream_t st) {
    if (a.k <= 0) return;
    hipLaunchKernelGGL(k_dots, dim3(nparts, a.k), dim3(256), 0, st, a, n, partial, nparts);
    hipLaunchKernelGGL(k_dots_finish, dim3(a.k), dim3(64), 0, st, partial, nparts, out);
    PF_HIP(hipGetLastError());
}

// dF/d(rho_a) = (sum_b chi_ab rho_b / (v_a v_b) + u_a rho_a / v_a^2 - w_a) / n: the
// derivative of -<rho w> + the Flory (src/interactions/flory.jl:39-59) and Edwards
// (src/interactions/edwards.jl) energies.
__global__ void k_rbar(ResidualArgs a, size_t n, double* rbar) {
    const double inv_n = 1.0 / (double)n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        for (int m = 0; m < a.nmon; ++m) {
            double p = 0.0;
            for (int q = 0; q < a.nmon; ++q)
                if (a.pair[m * PF_MAXMON + q])
                    p += a.chi[m * PF_MAXMON + q] * (a.rho[(size_t)q * n + i] / a.vol[q]);
            if (a.u_set[m]) p += a.u[m] * (a.rho[(size_t)m * n + i] / a.vol[m]);
            rbar[(size_t)m * n + i] = (p / a.vol[m] - a.w[(size_t)m * n + i]) * inv_n;
        }
    }
}

void launch_rbar(const ResidualArgs& a, size_t n, double* rbar, hipStream_t st) {
    int g = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_rbar, dim3(g), dim3(256), 0, st, a, n, rbar);
    PF_HIP(hipGetLastError());
}

// Point species (src/species/point.jl:36-53): rho = e / Q, e = exp(-v w), Q = <e>;
// dw += -v e (ca rbar + cb) with ca = phi / Q and cb = (total dF/dQ) / n.
__global__ void k_point_grad(const double* w, double vol, size_t n, const double* rbar,
                             double ca, double cb, double* dw) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dw[i] += -vol * exp(-vol * w[i]) * (ca * rbar[i] + cb);
}

void launch_point_grad(const double* w, double vol, size_t n, const double* rbar, double ca,
                       double cb, double* dw, hipStream_t st) {
    int g = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_point_grad, dim3(g), dim3(256), 0, st, w, vol, n, rbar, ca, cb, dw);
    PF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------
// Simpson accumulation over a segment of slices [s_lo, s_hi] (checkpointed sweeps): the
// contour rule of src/species/multiblock.jl:182-216 applied slice by slice,
//   acc_b += w^b_s q_s q+_s   for every block b holding slice s,
// then normalised by ds_b / (3 Q Nref) and summed per monomer (multiblock.jl:219-228).
template <typename T>
__global__ void k_simpson_seg(SimpsonTask<T> tk, int s_lo, int s_hi, double* acc) {
    const size_t n = tk.n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        for (int s = s_lo; s <= s_hi; ++s) {
            const double p = (double)tk.q[(size_t)(s - 1) * n + i] *
                             (double)tk.qc[(size_t)(s - 1) * n + i];
            for (int b = 0; b < tk.nblk; ++b) {
                const int bgn = tk.begin[b], end = tk.begin[b + 1];
                if (s >= bgn && s <= end) acc[(size_t)b * n + i] += simpson_w(s, bgn, end) * p;
            }
        }
    }
}

template <typename T>
__global__ void k_simpson_fin(SimpsonTask<T> tk, const double* acc) {
    const size_t n = tk.n;
    const double QN = tk.sc->Q[tk.isp] * tk.Nref;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        double acc_m[PF_MAXMON];
        for (int m = 0; m < tk.nmon; ++m) acc_m[m] = 0.0;
        for (int b = 0; b < tk.nblk; ++b) {
            double rho = acc[(size_t)b * n + i] * (tk.ds[b] / 3.0);
            rho /= QN;
            acc_m[tk.blk_mon[b]] += rho;
            if (tk.rho_blk) tk.rho_blk[(size_t)b * n + i] = rho;
        }
        for (int m = 0; m < tk.nmon; ++m) tk.rho_sp[(size_t)m * n + i] = acc_m[m];
    }
}

template <typename T>
void launch_simpson_seg(const SimpsonTask<T>& tk, int s_lo, int s_hi, double* acc,
                        hipStream_t st) {
    const int g = (int)std::min<size_t>((tk.n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_simpson_seg<T>, dim3(g), dim3(256), 0, st, tk, s_lo, s_hi, acc);
    PF_HIP(hipGetLastError());
}

template <typename T>
void launch_simpson_fin(const SimpsonTask<T>& tk, const double* acc, hipStream_t st) {
    const int g = (int)std::min<size_t>((tk.n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_simpson_fin<T>, dim3(g), dim3(256), 0, st, tk, acc);
    PF_HIP(hipGetLastError());
}

template void launch_simpson_seg<float>(const SimpsonTask<float>&, int, int, double*, hipStream_t);
template void launch_simpson_seg<double>(const SimpsonTask<double>&, int, int, double*, hipStream_t);
template void launch_simpson_fin<float>(const SimpsonTask<float>&, const double*, hipStream_t);
template void launch_simpson_fin<double>(const SimpsonTask<double>&, const double*, hipStream_t);

}  // namespace pf
