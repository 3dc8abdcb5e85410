// Compile-time transform building blocks shared by the specialised kernels: radix
// plans, Stockham passes over lines in LDS (in place or ping-pong), the split steps of the
// packed real-row transform, whole-plane forward / inverse transforms, and LDS staging
// helpers.  Every length, stride and trip count is a template constant.
#pragma once

#include "pf_device.h"

namespace pf {

struct CPlan {
    int npass;
    int r[12];
    int ns[12];
};

// Radix plans: the greedy 8/4/2/3 factorisation, except for the lengths where a radix of
// 6, 12 or 16 saves a pass (each pass is a round trip through LDS with two barriers):
// 48 = 8 x 6, 96 = 8 x 12, 128 = 8 x 16, 256 = 16 x 16.
constexpr CPlan cplan(int n) {
    CPlan p{};
    if (n == 48 || n == 96 || n == 128 || n == 256) {
        p.npass = 2;
        p.r[0] = n == 256 ? 16 : 8;
        p.ns[0] = 1;
        p.r[1] = n / p.r[0];
        p.ns[1] = p.r[0];
        return p;
    }
    int m = n, ns = 1;
    while (m % 8 == 0) { p.r[p.npass] = 8; p.ns[p.npass++] = ns; ns *= 8; m /= 8; }
    while (m % 4 == 0) { p.r[p.npass] = 4; p.ns[p.npass++] = ns; ns *= 4; m /= 4; }
    while (m % 2 == 0) { p.r[p.npass] = 2; p.ns[p.npass++] = ns; ns *= 2; m /= 2; }
    while (m % 3 == 0) { p.r[p.npass] = 3; p.ns[p.npass++] = ns; ns *= 3; m /= 3; }
    return p;
}

// One Stockham pass (compile-time).  NL lines of length N, element e of line l at
// buf[l*LS + e*ES]; NT threads; twiddle w_N^m = tw[m*TWS] (LDS).
// src == dst: in place (all reads of a round finish before its writes); otherwise the
// pass reads src and writes dst and needs one barrier instead of two.
template <typename T, int N, int R, int NS, int NL, int ES, int LS, int NT, bool INV, int TWS,
          bool PP = false>
__device__ __forceinline__ void cpass(const cpx<T>* src_buf, cpx<T>* dst_buf,
                                      const cpx<T>* __restrict__ tw) {
    constexpr int NB = N / R;
    // butterflies per thread and round: about 16 values in registers (32 in FP32 for the
    // radices 12 and 16)
    constexpr int BPR = R == 2 ? 8 : R == 8 || R == 6 ? 2 : R >= 12 ? (sizeof(T) == 4 ? 2 : 1) : 4;
    static_assert(NB <= BPR * NT, "transform too long for the workgroup");
    constexpr int LPR0 = (BPR * NT) / NB;
    constexpr int LPR = LPR0 < 1 ? 1 : (LPR0 > NL ? NL : LPR0);
    constexpr int TWM = (N / (NS * R)) * TWS;
    // Consecutive threads take consecutive lines.  With unit or odd line strides (every
    // layout here: rows of M+1 complex values with M even) a wavefront then hits each LDS
    // bank the minimum number of times; consecutive butterflies of one line would collide
    // across rows.
    constexpr bool LINE_FAST = (LS == 1 || LS % 2 == 1) && NL > 1;
#pragma unroll
    for (int l0 = 0; l0 < NL; l0 += LPR) {
        const int nlr = (NL - l0) < LPR ? (NL - l0) : LPR;
        const int total = nlr * NB;
        cpx<T> v[BPR][R];
#pragma unroll
        for (int i = 0; i < BPR; ++i) {
            const int b = threadIdx.x + i * NT;
            if (i * NT < total && b < total) {
                const int line = LINE_FAST ? l0 + b % nlr : l0 + b / NB;
                const int j = LINE_FAST ? b / nlr : b % NB;
                const cpx<T>* src = src_buf + line * LS + j * ES;
#pragma unroll
                for (int r = 0; r < R; ++r) v[i][r] = src[r * NB * ES];
                if constexpr (NS > 1) {
                    const int k = j % NS;
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        cpx<T> w = tw[k * r * TWM];
                        if (INV) w.y = -w.y;
                        v[i][r] = cmul(v[i][r], w);
                    }
                }
            }
        }
        if constexpr (!PP) __syncthreads();
#pragma unroll
        for (int i = 0; i < BPR; ++i) {
            const int b = threadIdx.x + i * NT;
            if (i * NT < total && b < total) {
                const int line = LINE_FAST ? l0 + b % nlr : l0 + b / NB;
                const int j = LINE_FAST ? b / nlr : b % NB;
                dft<T, R, INV>(v[i]);
                const int d = (j / NS) * NS * R + (j % NS);
                cpx<T>* dst = dst_buf + line * LS + d * ES;
#pragma unroll
                for (int r = 0; r < R; ++r) dst[r * NS * ES] = v[i][r];
            }
        }
        __syncthreads();
    }
}

template <typename T, int N, int NL, int ES, int LS, int NT, bool INV, int TWS, int IP = 0>
__device__ __forceinline__ void cfft(cpx<T>* buf, const cpx<T>* tw) {
    constexpr CPlan P = cplan(N);
    if constexpr (IP < P.npass) {
        cpass<T, N, P.r[IP], P.ns[IP], NL, ES, LS, NT, INV, TWS>(buf, buf, tw);
        cfft<T, N, NL, ES, LS, NT, INV, TWS, IP + 1>(buf, tw);
    }
}

// Ping-pong between buf and tmp (an even pass count leaves the result in buf).
template <typename T, int N, int NL, int ES, int LS, int NT, bool INV, int TWS, int IP = 0>
__device__ __forceinline__ void cfft_pp(cpx<T>* buf, cpx<T>* tmp, const cpx<T>* tw) {
    constexpr CPlan P = cplan(N);
    static_assert(P.npass % 2 == 0, "ping-pong transforms need an even number of passes");
    if constexpr (IP < P.npass) {
        if constexpr (IP % 2 == 0)
            cpass<T, N, P.r[IP], P.ns[IP], NL, ES, LS, NT, INV, TWS, true>(buf, tmp, tw);
        else
            cpass<T, N, P.r[IP], P.ns[IP], NL, ES, LS, NT, INV, TWS, true>(tmp, buf, tw);
        cfft_pp<T, N, NL, ES, LS, NT, INV, TWS, IP + 1>(buf, tmp, tw);
    }
}

// Split steps of the packed real-row transform (see pf_fft.hip for the formulas).
template <typename T, int M, int NROWS, int LD, int NT>
__device__ __forceinline__ void csplit_r2c(cpx<T>* S, const cpx<T>* twN) {
    constexpr int HALF = M / 2 + 1;
#pragma unroll
    for (int i0 = 0; i0 < NROWS * HALF; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < NROWS * HALF) {
            // consecutive threads on consecutive rows (odd row stride: no bank conflicts)
            const int y = i % NROWS, k = i / NROWS;
            cpx<T>* row = S + y * LD;
            const int km = (M - k) % M;
            const cpx<T> zk = row[k], zm = row[km];
            const cpx<T> e = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
            const cpx<T> d = {zk.x - zm.x, zk.y + zm.y};
            const cpx<T> o = {T(0.5) * d.y, T(-0.5) * d.x};
            const cpx<T> wo = cmul(twN[k], o);
            row[k] = cadd(e, wo);
            if (M - k != k) row[M - k] = cconj(csub(e, wo));
        }
    }
    __syncthreads();
}

template <typename T, int M, int NROWS, int LD, int NT>
__device__ __forceinline__ void csplit_c2r(cpx<T>* S, const cpx<T>* twN) {
    constexpr int HALF = M / 2 + 1;
#pragma unroll
    for (int i0 = 0; i0 < NROWS * HALF; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < NROWS * HALF) {
            const int y = i % NROWS, k = i / NROWS;
            cpx<T>* row = S + y * LD;
            cpx<T> xa = row[k], xb = row[M - k];
            if (k == 0) {
                xa.y = T(0);
                xb.y = T(0);
            }
            const cpx<T> e = {T(0.5) * (xa.x + xb.x), T(0.5) * (xa.y - xb.y)};
            const cpx<T> dd = {T(0.5) * (xa.x - xb.x), T(0.5) * (xa.y + xb.y)};
            const cpx<T> o = cmul(dd, cconj(twN[k]));
            row[k] = {e.x - o.y, e.y + o.x};
            if (k != 0 && M - k != k) row[M - k] = {e.x + o.y, -e.y + o.x};
        }
    }
    __syncthreads();
}

template <typename T, int NX, int NY, int NT>
__device__ __forceinline__ void cplane_forward(cpx<T>* S, const cpx<T>* twx, const cpx<T>* twy) {
    constexpr int M = NX / 2, LD = M + 1;
    cfft<T, M, NY, 1, LD, NT, false, 2>(S, twx);
    csplit_r2c<T, M, NY, LD, NT>(S, twx);
    if constexpr (NY > 1) cfft<T, NY, M + 1, LD, 1, NT, false, 1>(S, twy);
}

template <typename T, int NX, int NY, int NT>
__device__ __forceinline__ void cplane_forward_pp(cpx<T>* S, cpx<T>* S2, const cpx<T>* twx,
                                                  const cpx<T>* twy) {
    constexpr int M = NX / 2, LD = M + 1;
    cfft_pp<T, M, NY, 1, LD, NT, false, 2>(S, S2, twx);
    csplit_r2c<T, M, NY, LD, NT>(S, twx);
    cfft_pp<T, NY, M + 1, LD, 1, NT, false, 1>(S, S2, twy);
}

template <typename T, int NX, int NY, int NT>
__device__ __forceinline__ void cplane_inverse_pp(cpx<T>* S, cpx<T>* S2, const cpx<T>* twx,
                                                  const cpx<T>* twy) {
    constexpr int M = NX / 2, LD = M + 1;
    cfft_pp<T, NY, M + 1, LD, 1, NT, true, 1>(S, S2, twy);
    csplit_c2r<T, M, NY, LD, NT>(S, twx);
    cfft_pp<T, M, NY, 1, LD, NT, true, 2>(S, S2, twx);
}

template <typename T, int NX, int NY, int NT>
__device__ __forceinline__ void cplane_inverse(cpx<T>* S, const cpx<T>* twx, const cpx<T>* twy) {
    constexpr int M = NX / 2, LD = M + 1;
    if constexpr (NY > 1) cfft<T, NY, M + 1, LD, 1, NT, true, 1>(S, twy);
    csplit_c2r<T, M, NY, LD, NT>(S, twx);
    cfft<T, M, NY, 1, LD, NT, true, 2>(S, twx);
}

// COUNT values from global memory to LDS.  Loads go out in groups of up to 8 per thread
// before any of their stores.  A thread past the end copies element i % COUNT again (the
// same value to the same place, and distinct places across a wavefront, so the stores
// do not collide in one bank): no branch, so the compiler cannot move a load into a
// conditional block behind the waits for the others.
template <typename T, int COUNT, int NT>
__device__ __forceinline__ void lds_copy_in(cpx<T>* dst, const cpx<T>* __restrict__ src) {
    const auto g = gp(src);
    constexpr int IT = (COUNT + NT - 1) / NT, CH = 8;
#pragma unroll
    for (int c0 = 0; c0 < IT; c0 += CH) {
        cpx<T> v[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (c0 + c < IT) {
                const int i = (c0 + c) * NT + (int)threadIdx.x;
                v[c] = g[i < COUNT ? i : i % COUNT];
            }
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (c0 + c < IT) {
                const int i = (c0 + c) * NT + (int)threadIdx.x;
                dst[i < COUNT ? i : i % COUNT] = v[c];
            }
    }
}

// packed real plane (pairs, M per row, NY rows) <-> LDS rows of stride M+1
template <typename T, int NX, int NY, int NT, typename F>
__device__ __forceinline__ void for_pairs(F&& f) {
    constexpr int M = NX / 2, NP = M * NY;
#pragma unroll
    for (int i0 = 0; i0 < NP; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < NP) f(i, (i / M) * (M + 1) + (i % M));
    }
}

}  // namespace pf
