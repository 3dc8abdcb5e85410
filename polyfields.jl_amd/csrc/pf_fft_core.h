// LDS-resident Stockham transforms shared by the propagator and adjoint kernels.
//
// Layout: real grids are x fastest, q[(z*ny + y)*nx + x]; the half spectrum keeps
// kx = 0..nx/2 fastest, k[(kz*ny + ky)*(nx/2+1) + kx] (the reference's FFTW r2c layout).
// The x transform of a real row of length nx = 2M runs as a length-M complex transform
// of the packed pairs (x[2m], x[2m+1]) followed by the split step that yields the M+1
// half-spectrum entries; the inverse runs the split step backwards.
#pragma once

#include "pf_device.h"

namespace pf {

// One Stockham pass of radix R over `nlines` independent lines held in LDS.  Element e
// of line l sits at buf[l*ls + e*es].  Each thread loads all butterflies it owns into
// registers, the workgroup synchronises, and the outputs are written back into the same
// buffer in autosorted order.  Lines are processed in rounds that fit the register
// budget; a round always covers whole lines, so later rounds read untouched data.
// Twiddle w_N^m is tw[m*twstride].
template <typename T, int R, bool INV>
__device__ void fft_pass(cpx<T>* buf, int N, int Ns, int nlines, int es, int ls,
                         const cpx<T>* __restrict__ tw, int twstride) {
    constexpr int MAXB = (R == 2) ? 4 : 2;
    const int nb = N / R;
    const int nthr = blockDim.x;
    const int lpr = max(1, (nthr * MAXB) / nb);
    const int twmul = (N / (Ns * R)) * twstride;
    // consecutive threads on consecutive lines (unit or odd line stride: minimum number
    // of hits per LDS bank)
    const bool lf = (ls == 1 || (ls & 1)) && nlines > 1;
    for (int l0 = 0; l0 < nlines; l0 += lpr) {
        const int nlr = min(lpr, nlines - l0);
        const int total = nlr * nb;
        cpx<T> v[MAXB][R];
#pragma unroll
        for (int i = 0; i < MAXB; ++i) {
            const int b = threadIdx.x + i * nthr;
            if (b < total) {
                const int line = lf ? l0 + b % nlr : l0 + b / nb;
                const int j = lf ? b / nlr : b - (b / nb) * nb;
                const cpx<T>* src = buf + line * ls;
#pragma unroll
                for (int r = 0; r < R; ++r) v[i][r] = src[(j + r * nb) * es];
                if (Ns > 1) {
                    const int k = j % Ns;
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        cpx<T> w = tw[k * r * twmul];
                        if (INV) w.y = -w.y;
                        v[i][r] = cmul(v[i][r], w);
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < MAXB; ++i) {
            const int b = threadIdx.x + i * nthr;
            if (b < total) {
                const int line = lf ? l0 + b % nlr : l0 + b / nb;
                const int j = lf ? b / nlr : b - (b / nb) * nb;
                const int k = j % Ns;
                dft<T, R, INV>(v[i]);
                const int d = (j / Ns) * Ns * R + k;
                cpx<T>* dst = buf + line * ls;
#pragma unroll
                for (int r = 0; r < R; ++r) dst[(d + r * Ns) * es] = v[i][r];
            }
        }
        __syncthreads();
    }
}

template <typename T, bool INV>
__device__ void lds_fft(cpx<T>* buf, const FftPlan& p, int nlines, int es, int ls,
                        const cpx<T>* __restrict__ tw, int twstride) {
    const int N = p.n;
    int Ns = 1;
    for (int ip = 0; ip < p.npass; ++ip) {
        const int R = p.radix[ip];
        if (R == 4)
            fft_pass<T, 4, INV>(buf, N, Ns, nlines, es, ls, tw, twstride);
        else if (R == 2)
            fft_pass<T, 2, INV>(buf, N, Ns, nlines, es, ls, tw, twstride);
        else
            fft_pass<T, 3, INV>(buf, N, Ns, nlines, es, ls, tw, twstride);
        Ns *= R;
    }
}

// Split step after the length-M transform of packed rows: Z -> X_k, k = 0..M.
// X_k = E + W^k O and X_{M-k} = conj(E - W^k O), E = (Z_k + Z*_{M-k})/2,
// O = -i (Z_k - Z*_{M-k})/2, W = exp(-2 pi i / 2M).
template <typename T>
__device__ void rows_r2c_split(cpx<T>* S, int M, int nrows, int ld,
                               const cpx<T>* __restrict__ twN) {
    const int half = M / 2 + 1;
    for (int i = threadIdx.x; i < nrows * half; i += blockDim.x) {
        const int k = i / nrows, y = i - k * nrows;
        cpx<T>* row = S + y * ld;
        const int km = (M - k) % M;
        const cpx<T> zk = row[k], zm = row[km];
        const cpx<T> e = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
        const cpx<T> d = {zk.x - zm.x, zk.y + zm.y};        // Z_k - conj(Z_{M-k})
        const cpx<T> o = {T(0.5) * d.y, T(-0.5) * d.x};      // -i d / 2
        const cpx<T> wo = cmul(twN[k], o);
        row[k] = cadd(e, wo);
        if (M - k != k) row[M - k] = cconj(csub(e, wo));
    }
    __syncthreads();
}

// Inverse split step: X_k (k = 0..M, Hermitian row) -> Z_k, k = 0..M-1, with the
// imaginary parts of X_0 and X_M dropped as a c2r transform does.
template <typename T>
__device__ void rows_c2r_split(cpx<T>* S, int M, int nrows, int ld,
                               const cpx<T>* __restrict__ twN) {
    const int half = M / 2 + 1;
    for (int i = threadIdx.x; i < nrows * half; i += blockDim.x) {
        const int k = i / nrows, y = i - k * nrows;
        cpx<T>* row = S + y * ld;
        cpx<T> xa = row[k], xb = row[M - k];
        if (k == 0) {
            xa.y = T(0);
            xb.y = T(0);
        }
        const cpx<T> e = {T(0.5) * (xa.x + xb.x), T(0.5) * (xa.y - xb.y)};
        const cpx<T> dd = {T(0.5) * (xa.x - xb.x), T(0.5) * (xa.y + xb.y)};
        const cpx<T> o = cmul(dd, cconj(twN[k]));
        row[k] = {e.x - o.y, e.y + o.x};                     // E + i O
        if (k != 0 && M - k != k) row[M - k] = {e.x + o.y, -e.y + o.x};   // conj(E) + i conj(O)
    }
    __syncthreads();
}

// x-forward then y-forward of a plane held as packed real rows in S.
template <typename T>
__device__ void plane_forward(cpx<T>* S, int M, int ny, int ld, const FftPlan& px,
                              const FftPlan& py, const cpx<T>* twx, const cpx<T>* twy) {
    lds_fft<T, false>(S, px, ny, 1, ld, twx, 2);
    rows_r2c_split(S, M, ny, ld, twx);
    if (ny > 1) lds_fft<T, false>(S, py, M + 1, ld, 1, twy, 1);
}

template <typename T>
__device__ void plane_inverse(cpx<T>* S, int M, int ny, int ld, const FftPlan& px,
                              const FftPlan& py, const cpx<T>* twx, const cpx<T>* twy) {
    if (ny > 1) lds_fft<T, true>(S, py, M + 1, ld, 1, twy, 1);
    rows_c2r_split(S, M, ny, ld, twx);
    lds_fft<T, true>(S, px, ny, 1, ld, twx, 2);
}

}  // namespace pf
