// Propagator kernels specialised at compile time for the grid sizes of the production
// workloads (64^3, 128^3, 32^3; 2-D planes 32^2..128^2; 1-D lines 32..256).  Same
// arithmetic as the generic kernels of pf_fft.hip (same Stockham formulation, same split
// steps, same order of the LW / LD / 1/n multiplies), with:
//   * every length, stride and trip count a constant, so index math is shifts and
//     multiplies and load / store loops are fully unrolled (all loads in flight at once);
//   * twiddle tables copied to LDS once per workgroup instead of read from global memory
//     in every pass;
//   * radix-8 passes (64 = 8*8, 128 = 8*8*2, 96 = 8*4*3) to cut the number of LDS
//     round trips and workgroup barriers.
#include <array>
#include <map>
#include <mutex>

#include "pf_fft_fixed.h"

namespace pf {

// ---------------------------------------------------------------------------------
template <typename T, int NX, int NY, int NT>
__global__ void __launch_bounds__(NT) k_planar_fx(const Lane<T>* __restrict__ lanes, int t0, int t1) {
    constexpr int M = NX / 2, LD = M + 1, N = NX * NY, NKP = NY * LD;
    // The LD factors exp(-c_b k^2) of the current block are computed when the block
    // changes, not in every step, where the plane, the twiddles and the factors fit in
    // 160 KiB of LDS.  In the 1024-thread forms (one workgroup fills a CU's wave slots,
    // so LDS does not limit occupancy), the block's LW pairs are copied there at the same
    // time when they fit too, instead of being read from global memory twice per step.
    // Entry i of either table is written and read only by thread i % NT, so they need no
    // barrier.
    constexpr int NP = M * NY;
    constexpr size_t LDS_BASE = (size_t)NKP * sizeof(cpx<T>) + (size_t)(NX + NY) * sizeof(cpx<T>);
    constexpr bool FTAB = LDS_BASE + (size_t)NKP * sizeof(T) <= 160u * 1024u;
    constexpr bool LTAB = FTAB && NT == 1024 &&
                          LDS_BASE + (size_t)NKP * sizeof(T) + (size_t)NP * sizeof(cpx<T>) <= 160u * 1024u;
    __shared__ cpx<T> S[NKP];
    __shared__ cpx<T> TWX[NX];
    __shared__ cpx<T> TWY[NY];
    __shared__ T FL[FTAB ? NKP : 1];
    __shared__ cpx<T> LWT[LTAB ? NP : 1];
    const Lane<T>* L = lanes + blockIdx.x;
    const int tend = min(t1, L->nsteps);
    if (t0 >= tend) return;
    lds_copy_in<T, NX, NT>(TWX, L->twx);
    lds_copy_in<T, NY, NT>(TWY, L->twy);
    const T scale = T(2.0 / (double)N);
    int s_in;
    step_block(L, t0, &s_in);
    {
        const auto qin = gp(reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * N));
        for_pairs<T, NX, NY, NT>([&](int i, int o) { S[o] = qin[i]; });
    }
    const double* G = L->G;
    int bf = -1;   // block whose tables FL (and LWT) hold
    for (int t = t0; t < tend; ++t) {
        const int b = step_block(L, t, &s_in);
        const auto lw = gp(reinterpret_cast<const cpx<T>*>(L->lw[b]));
        if constexpr (FTAB) {
            if (b != bf) {
                bf = b;
                if constexpr (LTAB) for_pairs<T, NX, NY, NT>([&](int i, int) { LWT[i] = lw[i]; });
                const double mc = -L->ldc[b];
                // not unrolled: the fill is rare, and one exp at a time keeps the
                // step loop's register budget
#pragma unroll 1
                for (int i0 = 0; i0 < NKP; i0 += NT) {
                    const int i = i0 + threadIdx.x;
                    if (i < NKP) {
                        const int y = i / LD, kx = i % LD;
                        FL[i] = exp_t<T>(mc * ksq_at(G, (double)kx, (double)fold(y, NY), 0.0));
                    }
                }
            }
        }
        __syncthreads();
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> w = LTAB ? LWT[i] : lw[i];
            S[o].x *= w.x;
            S[o].y *= w.y;
        });
        __syncthreads();
        cplane_forward<T, NX, NY, NT>(S, TWX, TWY);
        if constexpr (FTAB) {
#pragma unroll
            for (int i0 = 0; i0 < NKP; i0 += NT) {
                const int i = i0 + threadIdx.x;
                if (i < NKP) {
                    const T f = FL[i];
                    S[i].x *= f;
                    S[i].y *= f;
                }
            }
        } else {
            const double mc = -L->ldc[b];
#pragma unroll
            for (int i0 = 0; i0 < NKP; i0 += NT) {
                const int i = i0 + threadIdx.x;
                if (i < NKP) {
                    const int y = i / LD, kx = i % LD;
                    const T f = exp_t<T>(mc * ksq_at(G, (double)kx, (double)fold(y, NY), 0.0));
                    S[i].x *= f;
                    S[i].y *= f;
                }
            }
        }
        __syncthreads();
        cplane_inverse<T, NX, NY, NT>(S, TWX, TWY);
        const auto qout = gp(reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * N));
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> w = LTAB ? LWT[i] : lw[i];
            const cpx<T> v = {(S[o].x * scale) * w.x, (S[o].y * scale) * w.y};
            S[o] = v;
            qout[i] = v;
        });
    }
}

// RQM4 steps (src/system/mde.jl:83-104) of 1-D / 2-D lanes at compile-time sizes: the
// generic planar kernel's RQM4 branch (the same products in the same order: the full step
// q1 in S, the two half steps in S2 with LW1 between them, q_out = (4 q2 - q1) / 3), with
// fixed radix plans, LDS twiddles and one workgroup of NT threads per lane.  Where they
// fit beside the two planes, the block's full- and half-step LD factors are computed into
// LDS once per block (exp(a) and exp(a / 2), as the generic kernel forms them per pass).
template <typename T, int NX, int NY, int NT>
__global__ void __launch_bounds__(NT) k_planar_rq_fx(const Lane<T>* __restrict__ lanes, int t0,
                                                     int t1) {
    constexpr int M = NX / 2, LD = M + 1, N = NX * NY, NKP = NY * LD;
    constexpr size_t BASE = 2 * (size_t)NKP * sizeof(cpx<T>) + (size_t)(NX + NY) * sizeof(cpx<T>);
    static_assert(BASE <= 160u * 1024u, "both planes must fit in LDS");
    constexpr bool FTAB = BASE + 2 * (size_t)NKP * sizeof(T) <= 160u * 1024u;
    __shared__ cpx<T> S[NKP];
    __shared__ cpx<T> S2[NKP];
    __shared__ cpx<T> TWX[NX];
    __shared__ cpx<T> TWY[NY];
    __shared__ T FL[FTAB ? 2 * NKP : 1];
    const Lane<T>* L = lanes + blockIdx.x;
    const int tend = min(t1, L->nsteps);
    if (t0 >= tend) return;
    lds_copy_in<T, NX, NT>(TWX, L->twx);
    lds_copy_in<T, NY, NT>(TWY, L->twy);
    const T scale = T(2.0 / (double)N);
    int s_in;
    step_block(L, t0, &s_in);
    {
        const auto qin = gp(reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * N));
        for_pairs<T, NX, NY, NT>([&](int i, int o) { S[o] = qin[i]; });
    }
    const double* G = L->G;
    int bf = -1;   // block whose factors FL holds
    for (int t = t0; t < tend; ++t) {
        const int b = step_block(L, t, &s_in);
        const auto lw = gp(reinterpret_cast<const cpx<T>*>(L->lw[b]));
        const auto lw2 = gp(reinterpret_cast<const cpx<T>*>(L->lw2[b]));
        const double mc = -L->ldc[b];
        if constexpr (FTAB) {
            if (b != bf) {
                bf = b;
#pragma unroll 1
                for (int i0 = 0; i0 < NKP; i0 += NT) {
                    const int i = i0 + threadIdx.x;
                    if (i < NKP) {
                        const int y = i / LD, kx = i % LD;
                        const double a = mc * ksq_at(G, (double)kx, (double)fold(y, NY), 0.0);
                        FL[i] = exp_t<T>(a);
                        FL[NKP + i] = exp_t<T>(a / 2.0);
                    }
                }
            }
        }
        __syncthreads();
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> w = lw[i], w2 = lw2[i], v = S[o];
            S2[o] = {w2.x * v.x, w2.y * v.y};
            S[o] = {w.x * v.x, w.y * v.y};
        });
        __syncthreads();
#pragma unroll 1
        for (int pass = 0; pass < 3; ++pass) {
            cpx<T>* X = pass == 0 ? S : S2;
            const bool half = pass > 0;
            cplane_forward<T, NX, NY, NT>(X, TWX, TWY);
#pragma unroll
            for (int i0 = 0; i0 < NKP; i0 += NT) {
                const int i = i0 + threadIdx.x;
                if (i < NKP) {
                    T f;
                    if constexpr (FTAB) {
                        f = FL[(half ? NKP : 0) + i];
                    } else {
                        const int y = i / LD, kx = i % LD;
                        const double a = mc * ksq_at(G, (double)kx, (double)fold(y, NY), 0.0);
                        f = exp_t<T>(half ? a / 2.0 : a);
                    }
                    X[i].x *= f;
                    X[i].y *= f;
                }
            }
            __syncthreads();
            cplane_inverse<T, NX, NY, NT>(X, TWX, TWY);
            const auto wm = pass == 2 ? lw2 : lw;   // q1 *= LW1; mid *= LW1; q2 *= LW2
            for_pairs<T, NX, NY, NT>([&](int i, int o) {
                const cpx<T> w = wm[i];
                X[o] = {(X[o].x * scale) * w.x, (X[o].y * scale) * w.y};
            });
            __syncthreads();
        }
        const auto qout = gp(reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * N));
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> q1 = S[o], q2 = S2[o];
            const cpx<T> v = {(T(4) * q2.x - q1.x) / T(3), (T(4) * q2.y - q1.y) / T(3)};
            S[o] = v;
            qout[i] = v;
        });
    }
}

// RQM4 steps of 2-D lanes whose two planes do not fit in LDS together (FP64 128^2): the
// generic planar kernel's one-plane branch at a compile-time size.  The half steps re-read
// the step's input from the stack and the full step's result q1 waits in the output slice;
// each thread reads back only the entries it wrote itself.
template <typename T, int NX, int NY, int NT>
__global__ void __launch_bounds__(NT) k_planar_rq1_fx(const Lane<T>* __restrict__ lanes, int t0,
                                                      int t1) {
    constexpr int M = NX / 2, LD = M + 1, N = NX * NY, NKP = NY * LD;
    __shared__ cpx<T> S[NKP];
    __shared__ cpx<T> TWX[NX];
    __shared__ cpx<T> TWY[NY];
    const Lane<T>* L = lanes + blockIdx.x;
    const int tend = min(t1, L->nsteps);
    if (t0 >= tend) return;
    lds_copy_in<T, NX, NT>(TWX, L->twx);
    lds_copy_in<T, NY, NT>(TWY, L->twy);
    const T scale = T(2.0 / (double)N);
    int s_in;
    step_block(L, t0, &s_in);
    {
        const auto qin = gp(reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * N));
        for_pairs<T, NX, NY, NT>([&](int i, int o) { S[o] = qin[i]; });
    }
    const double* G = L->G;
    for (int t = t0; t < tend; ++t) {
        const int b = step_block(L, t, &s_in);
        const auto lw = gp(reinterpret_cast<const cpx<T>*>(L->lw[b]));
        const auto lw2 = gp(reinterpret_cast<const cpx<T>*>(L->lw2[b]));
        const auto qin = gp(reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * N));
        const auto qout = gp(reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * N));
        const double mc = -L->ldc[b];
#pragma unroll 1
        for (int pass = 0; pass < 3; ++pass) {
            const bool half = pass > 0;
            if (pass < 2) {
                __syncthreads();
                for_pairs<T, NX, NY, NT>([&](int i, int o) {
                    const cpx<T> w = pass == 0 ? lw[i] : lw2[i];
                    const cpx<T> v = pass == 0 ? S[o] : qin[i];
                    S[o] = {w.x * v.x, w.y * v.y};
                });
                __syncthreads();
            }
            cplane_forward<T, NX, NY, NT>(S, TWX, TWY);
#pragma unroll 1
            for (int i0 = 0; i0 < NKP; i0 += NT) {
                const int i = i0 + threadIdx.x;
                if (i < NKP) {
                    const int y = i / LD, kx = i % LD;
                    const double a = mc * ksq_at(G, (double)kx, (double)fold(y, NY), 0.0);
                    const T f = exp_t<T>(half ? a / 2.0 : a);
                    S[i].x *= f;
                    S[i].y *= f;
                }
            }
            __syncthreads();
            cplane_inverse<T, NX, NY, NT>(S, TWX, TWY);
            const auto wm = pass == 2 ? lw2 : lw;   // q1 *= LW1; mid *= LW1; q2 *= LW2
            for_pairs<T, NX, NY, NT>([&](int i, int o) {
                const cpx<T> w = wm[i];
                S[o] = {(S[o].x * scale) * w.x, (S[o].y * scale) * w.y};
                if (pass == 0) qout[i] = S[o];
            });
            __syncthreads();
        }
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> q1 = qout[i], q2 = S[o];
            const cpx<T> v = {(T(4) * q2.x - q1.x) / T(3), (T(4) * q2.y - q1.y) / T(3)};
            S[o] = v;
            qout[i] = v;
        });
    }
}

// 3-D RQM4 slab passes (src/system/mde.jl:83-104) at compile-time plane sizes, both planes
// in LDS: the modes, products and order of operations of the generic k_slab_rqm4, with
// fixed radix plans and LDS twiddles.
template <typename T, int NX, int NY, int MODE, int NT>
__global__ void __launch_bounds__(NT) k_slab_rq_fx(const Lane<T>* __restrict__ lanes, int nz, int t) {
    constexpr int M = NX / 2, PLANE = NX * NY, KPLANE = (M + 1) * NY;
    __shared__ cpx<T> S[KPLANE];
    __shared__ cpx<T> S2[KPLANE];
    __shared__ cpx<T> TWX[NX];
    __shared__ cpx<T> TWY[NY];
    // grid (nz, lanes): plane z of lane blockIdx.y
    const int z = blockIdx.x;
    const Lane<T>* L = lanes + blockIdx.y;
    if (t >= L->nsteps) return;
    const size_t n = (size_t)PLANE * nz;
    lds_copy_in<T, NX, NT>(TWX, L->twx);
    lds_copy_in<T, NY, NT>(TWY, L->twy);
    cpx<T>* const kbp = L->kbuf + (size_t)z * KPLANE;
    cpx<T>* const kb2p = L->kb2 + (size_t)z * KPLANE;
    const auto rq = gp(reinterpret_cast<cpx<T>*>(L->rq + (size_t)z * PLANE));
    const T scale = T(2.0 / (double)n);
    int s_in;
    const int b = step_block(L, t, &s_in);
    const auto lw1 = gp(reinterpret_cast<const cpx<T>*>(L->lw[b] + (size_t)z * PLANE));
    const auto lw2 = gp(reinterpret_cast<const cpx<T>*>(L->lw2[b] + (size_t)z * PLANE));
    auto spec_out = [&](cpx<T>* dst, const cpx<T>* src) {
        const auto g = gp(dst);
#pragma unroll
        for (int i0 = 0; i0 < KPLANE; i0 += NT)
            if (i0 + (int)threadIdx.x < KPLANE) g[i0 + threadIdx.x] = src[i0 + threadIdx.x];
    };
    auto both_forward = [&] {
        __syncthreads();
        cplane_forward<T, NX, NY, NT>(S, TWX, TWY);
        cplane_forward<T, NX, NY, NT>(S2, TWX, TWY);
        spec_out(kbp, S);
        spec_out(kb2p, S2);
    };
    if constexpr (MODE == RQ_START) {
        const auto q = gp(reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * n +
                                                          (size_t)z * PLANE));
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> v = q[i], w = lw1[i], w2 = lw2[i];
            S[o] = {v.x * w.x, v.y * w.y};
            S2[o] = {v.x * w2.x, v.y * w2.y};
        });
        both_forward();
    } else if constexpr (MODE == RQ_MID) {
        lds_copy_in<T, KPLANE, NT>(S, kbp);
        lds_copy_in<T, KPLANE, NT>(S2, kb2p);
        __syncthreads();
        cplane_inverse<T, NX, NY, NT>(S, TWX, TWY);
        cplane_inverse<T, NX, NY, NT>(S2, TWX, TWY);
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> w = lw1[i], v = S[o], m = S2[o];
            rq[i] = cpx<T>{(v.x * scale) * w.x, (v.y * scale) * w.y};
            S2[o] = {(m.x * scale) * w.x, (m.y * scale) * w.y};
        });
        __syncthreads();
        cplane_forward<T, NX, NY, NT>(S2, TWX, TWY);
        spec_out(kb2p, S2);
    } else {
        lds_copy_in<T, KPLANE, NT>(S2, kb2p);
        __syncthreads();
        cplane_inverse<T, NX, NY, NT>(S2, TWX, TWY);
        const auto qout = gp(reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * n +
                                                       (size_t)z * PLANE));
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> w = lw2[i], v = S2[o], q1 = rq[i];
            const cpx<T> q2 = {(v.x * scale) * w.x, (v.y * scale) * w.y};
            const cpx<T> r = {(T(4) * q2.x - q1.x) / T(3), (T(4) * q2.y - q1.y) / T(3)};
            qout[i] = r;
            S[o] = r;
        });
        if (t + 1 >= L->nsteps) return;
        int s2;
        const int b2 = step_block(L, t + 1, &s2);
        const auto lw1n = gp(reinterpret_cast<const cpx<T>*>(L->lw[b2] + (size_t)z * PLANE));
        const auto lw2n = gp(reinterpret_cast<const cpx<T>*>(L->lw2[b2] + (size_t)z * PLANE));
        __syncthreads();
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> v = S[o], w = lw1n[i], w2 = lw2n[i];
            S[o] = {v.x * w.x, v.y * w.y};
            S2[o] = {v.x * w2.x, v.y * w2.y};
        });
        both_forward();
    }
}

// forward xy transform of the plane in LDS and its half spectrum out to kb
template <typename T, int NX, int NY, int NT, bool PPS, typename K>
__device__ __forceinline__ void slab_forward_out(cpx<T>* S, cpx<T>* S2, const cpx<T>* TWX,
                                                 const cpx<T>* TWY, K kb) {
    constexpr int KPLANE = (NX / 2 + 1) * NY;
    if constexpr (PPS)
        cplane_forward_pp<T, NX, NY, NT>(S, S2, TWX, TWY);
    else
        cplane_forward<T, NX, NY, NT>(S, TWX, TWY);
#pragma unroll
    for (int i0 = 0; i0 < KPLANE; i0 += NT)
        if (i0 + (int)threadIdx.x < KPLANE) kb[i0 + threadIdx.x] = S[i0 + threadIdx.x];
}

// MINW: the waves per SIMD the register budget must allow (launch bound), 1 = no limit
template <typename T, int NX, int NY, int MODE, int NT, int MINW = 1>
__global__ void __launch_bounds__(NT, MINW) k_slab_fx(const Lane<T>* __restrict__ lanes, int nz, int t) {
    constexpr int M = NX / 2, LD = M + 1, PLANE = NX * NY, KPLANE = LD * NY;
    // ping-pong transforms (one barrier per pass) where a second plane of LDS is cheap
    // (FP32 128 x 128: two planes of 66.5 KB, so only in the one-plane-per-CU form)
    constexpr bool PPS = sizeof(T) == 4 && (NX <= 64 || (NX <= 128 && MINW == 1)) && NY > 1 &&
                         cplan(NX / 2).npass % 2 == 0 && cplan(NY).npass % 2 == 0;
    __shared__ cpx<T> S[KPLANE];
    __shared__ cpx<T> S2[PPS ? KPLANE : 1];
    __shared__ cpx<T> TWX[NX];
    __shared__ cpx<T> TWY[NY];
    // grid (nz, lanes): plane z of lane blockIdx.y
    const int z = blockIdx.x;
    const Lane<T>* L = lanes + blockIdx.y;
    const size_t n = (size_t)PLANE * nz;
    int s_in;
    if constexpr (MODE == SLAB_INVFWD) {
        // Every lane field the first loads need comes in one round of scalar loads (the
        // empty asm keeps them ahead of every branch).  The spectrum plane and the
        // twiddles are loaded first; the LW rows of this plane are issued once they have
        // arrived, and stream in during the inverse transforms (loads complete in order,
        // so LW rows issued first would hold up the spectrum's first LDS store).  A
        // finished lane (t >= Ns) loads its (valid) buffers too and returns before its
        // first barrier.
        unsigned long long kraw = __builtin_bit_cast(unsigned long long, L->kbuf),
                           txr = __builtin_bit_cast(unsigned long long, L->twx),
                           tyr = __builtin_bit_cast(unsigned long long, L->twy);
        int s2;
        const int b = step_block(L, t, &s_in);
        const int nsteps = L->nsteps;
        const bool more = t + 1 < nsteps;
        const int bn = step_block(L, t + 1, &s2), b2 = more ? bn : b;
        const T* const lwp = pick_blk(L->lw, b);
        unsigned long long lw2raw = __builtin_bit_cast(unsigned long long, pick_blk(L->lw, b2));
        asm volatile("" : "+s"(kraw), "+s"(txr), "+s"(tyr), "+s"(lw2raw));
        cpx<T>* const kbp = __builtin_bit_cast(cpx<T>*, kraw) + (size_t)z * KPLANE;
        const auto kb = gp(kbp);
        constexpr int NP = (NX / 2) * NY, PER = (NP + NT - 1) / NT;
        const auto lw = gp(reinterpret_cast<const cpx<T>*>(lwp + (size_t)z * PLANE));
        lds_copy_in<T, KPLANE, NT>(S, kbp);
        lds_copy_in<T, NX, NT>(TWX, __builtin_bit_cast(const cpx<T>*, txr));
        lds_copy_in<T, NY, NT>(TWY, __builtin_bit_cast(const cpx<T>*, tyr));
        // the LW loads stay behind the copies above (no global access moves across this)
        asm volatile("" ::: "memory");
        cpx<T> wr[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = k * NT + threadIdx.x;
            if (i < NP) wr[k] = lw[i];
        }
        if (t >= nsteps) return;
        __syncthreads();
        if constexpr (PPS)
            cplane_inverse_pp<T, NX, NY, NT>(S, S2, TWX, TWY);
        else
            cplane_inverse<T, NX, NY, NT>(S, TWX, TWY);
        const T scale = T(2.0 / (double)n);
        const auto qout = gp(reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * n +
                                                       (size_t)z * PLANE));
        const auto lw2 = gp(reinterpret_cast<const cpx<T>*>(__builtin_bit_cast(const T*, lw2raw) +
                                                            (size_t)z * PLANE));
        if (b2 == b) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = k * NT + threadIdx.x;
                if (i < NP) {
                    const int o = (i / (NX / 2)) * (NX / 2 + 1) + (i % (NX / 2));
                    const cpx<T> w = wr[k];
                    const cpx<T> v = {(S[o].x * scale) * w.x, (S[o].y * scale) * w.y};
                    qout[i] = v;
                    S[o] = {v.x * w.x, v.y * w.y};
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = k * NT + threadIdx.x;
                if (i < NP) {
                    const int o = (i / (NX / 2)) * (NX / 2 + 1) + (i % (NX / 2));
                    const cpx<T> w = wr[k];
                    const cpx<T> v = {(S[o].x * scale) * w.x, (S[o].y * scale) * w.y};
                    qout[i] = v;
                    const cpx<T> w2 = lw2[i];
                    S[o] = {v.x * w2.x, v.y * w2.y};
                }
            }
        }
        if (!more) return;
        __syncthreads();
        slab_forward_out<T, NX, NY, NT, PPS>(S, S2, TWX, TWY, kb);
    } else {
        // the first step: LW q_1 from the stack (one launch per sweep)
        if (t >= L->nsteps) return;
        const auto kb = gp(L->kbuf + (size_t)z * KPLANE);
        lds_copy_in<T, NX, NT>(TWX, L->twx);
        lds_copy_in<T, NY, NT>(TWY, L->twy);
        const int b = step_block(L, t, &s_in);
        const auto qin = gp(reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * n +
                                                            (size_t)z * PLANE));
        const auto lw = gp(reinterpret_cast<const cpx<T>*>(L->lw[b] + (size_t)z * PLANE));
        for_pairs<T, NX, NY, NT>([&](int i, int o) {
            const cpx<T> q = qin[i], w = lw[i];
            S[o] = {q.x * w.x, q.y * w.y};
        });
        __syncthreads();
        slab_forward_out<T, NX, NY, NT, PPS>(S, S2, TWX, TWY, kb);
    }
}

// xy-slab step for FP32 64 x 64 planes (steps after the first), 256 threads per plane,
// with the half-spectrum split carried into k-space so that every transform runs on
// 32 x 64 points and every thread holds 8 of them in every phase.
//
// The x split of the c2r / r2c transforms acts on the pair (kx, ky), (32 - kx, -ky): the
// conjugate of a y-transformed column flips ky.  Columns kx = 0 and 32 first take their
// Hermitian part in ky (the y-space rows' DC and Nyquist imaginary parts dropped, as the
// row-space split does).  Phases:
//   1  y inverse: 8 lanes per column, 8 points each; the split is applied as the column
//      is read.  64 = 8 (in lane) x 8 (across the 8 lanes, after a transpose).
//   2  x inverse: 4 lanes per row; 32 = 8 (in lane) x 4 (after a transpose).  Each lane
//      then holds four runs of 4 consecutive reals: scale, LW, store q, LW, and the x
//      forward transform backwards through the same steps.
//   3  y forward, as phase 1.
//   4  split r2c as the half spectrum is written out.
// Transposes inside a phase stay within one wavefront (a private LDS region, no
// workgroup barrier); phases exchange the plane through LDS with barriers.  LDS layouts
// are XOR-swizzled so that column and row accesses of a wavefront hit distinct banks.
namespace sx {
__device__ __forceinline__ int swz(int k, int y) {
    const int a = y & 7;
    return y * 32 + (k ^ (4 * (a & 3) + 2 * (a >> 2)));
}
// 8 x 8 transposes of the column phases: group g (8 lanes), lane a, value c (rows of 9:
// both the write and the transposed read of 16 lanes hit 16 distinct banks, and the
// offsets are linear, so they fold into the LDS instructions)
__device__ __forceinline__ int tcol(int g, int a, int c) { return g * 72 + c * 9 + a; }
// 4 x 8 transposes of the row phase: row rho (4 lanes), lane l, value m1
__device__ __forceinline__ int trow(int rho, int l, int m1) {
    return rho * 36 + m1 * 4 + (l ^ ((m1 >> 1) & 3));
}
// LDS exchange between the lanes of one wavefront: the writes above complete before the
// reads below (in-order LDS queue of a wavefront; this keeps the compiler from reordering)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
struct f4 {
    float x, y, z, w;
};
// 16-byte accesses of one float plane (a buffer resource over the plane; accesses past
// `bytes` read 0 and write nothing)
struct PlaneBuf {
    __amdgpu_buffer_rsrc_t r;
    using V4 = unsigned __attribute__((ext_vector_type(4)));
    __device__ __forceinline__ PlaneBuf(const float* base, unsigned bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0, (int)bytes, 0x00020000)) {}
    __device__ __forceinline__ f4 load(unsigned off) const {
        const V4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
        const unsigned a = v.x, b = v.y, c = v.z, d = v.w;
        return {__uint_as_float(a), __uint_as_float(b), __uint_as_float(c), __uint_as_float(d)};
    }
    __device__ __forceinline__ void store(const f4& q, unsigned off) const {
        const V4 v = {__float_as_uint(q.x), __float_as_uint(q.y), __float_as_uint(q.z),
                      __float_as_uint(q.w)};
        __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 0);
    }
};
}  // namespace sx

__global__ void __launch_bounds__(256, 4) k_slab_sx(const Lane<float>* __restrict__ lanes, int nz, int t) {
    using C = cpx<float>;
    constexpr int PLANE = 64 * 64, KPLANE = 33 * 64;
    __shared__ C P[2048];      // the plane between phases (32 columns, swizzled)
    __shared__ C NQ[64];       // column kx = 32 of the input spectrum
    __shared__ C W[4 * 576];   // per-wavefront transpose regions
    __shared__ C TW[64];       // e^{-2 pi i m / 64}
    // grid (nz, lanes): plane z of lane blockIdx.y.  Every lane field the first loads need
    // comes from one round of scalar loads; a finished lane (t >= Ns) loads its (valid)
    // buffers like the others and returns before its first barrier.
    const int z = blockIdx.x;
    const Lane<float>* L = lanes + blockIdx.y;
    const size_t n = (size_t)PLANE * nz;
    const int tid = threadIdx.x, w = tid >> 6, ln = tid & 63;
    unsigned long long kraw = __builtin_bit_cast(unsigned long long, L->kbuf),
                       twraw = __builtin_bit_cast(unsigned long long, L->twy);
    int s_in, s2;
    const int b = step_block(L, t, &s_in);
    const int nsteps = L->nsteps;
    const bool more = t + 1 < nsteps;
    const int bn = step_block(L, t + 1, &s2), b2 = more ? bn : b;   // a select, not a branch
    const float* const lwp = pick_blk(L->lw, b);
    unsigned long long lw2raw = __builtin_bit_cast(unsigned long long, pick_blk(L->lw, b2));
    // these pointers are used only from here on, so their loads join the round above
    asm volatile("" : "+s"(kraw), "+s"(twraw), "+s"(lw2raw));
    const auto kb = gp(__builtin_bit_cast(cpx<float>*, kraw) + (size_t)z * KPLANE);
    const cpx<float>* const twy = __builtin_bit_cast(const cpx<float>*, twraw);
    // row-phase position and its LW values, loaded first
    const int rho = ln >> 2, lp = ln & 3, yr = 16 * w + rho;
    const unsigned roff = (unsigned)(yr * 64 + 4 * lp) * 4u;
    const sx::PlaneBuf lwb(lwp + (size_t)z * PLANE, PLANE * 4);
    sx::f4 wr[4];
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) wr[m2] = lwb.load(roff + 64u * m2);
    // the spectrum plane, column kx = 32 and the twiddles: every load issued before the
    // first LDS store waits for them (loads complete in order)
    C kv[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int i = s * 256 + tid;
        kv[s] = kb[(i >> 5) * 33 + (i & 31)];
    }
    const C nq = kb[ln * 33 + 32], tw = gp(twy)[ln];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int i = s * 256 + tid;
        P[sx::swz(i & 31, i >> 5)] = kv[s];
    }
    // every wavefront stores the same 64 values (a condition here would move the two loads
    // behind the others' completion)
    NQ[ln] = nq;
    TW[ln] = tw;
    // the LW values count as used here, so their loads stay ahead of the exit test (they
    // were issued before the spectrum loads and have arrived by now)
    asm volatile("" : "+v"(wr[0].x), "+v"(wr[0].y), "+v"(wr[0].z), "+v"(wr[0].w),
                 "+v"(wr[1].x), "+v"(wr[1].y), "+v"(wr[1].z), "+v"(wr[1].w),
                 "+v"(wr[2].x), "+v"(wr[2].y), "+v"(wr[2].z), "+v"(wr[2].w),
                 "+v"(wr[3].x), "+v"(wr[3].y), "+v"(wr[3].z), "+v"(wr[3].w));
    if (t >= nsteps) return;
    __syncthreads();

    C* const Wr = W + 576 * w;
    const int g = ln >> 3, a = ln & 7, kc = 8 * w + g;   // column phases: column kc, lane a
    C v[8];
    // ---- phase 1: split and y inverse ----
    {
        const C om = {TW[kc].x, -TW[kc].y};   // e^{+2 pi i kc / 64}
        auto split = [&](C xa, C xb) {       // xb: the conjugated partner
            const C e = {0.5f * (xa.x + xb.x), 0.5f * (xa.y + xb.y)};
            const C dd = {0.5f * (xa.x - xb.x), 0.5f * (xa.y - xb.y)};
            const C o = cmul(dd, om);
            return C{e.x - o.y, e.y + o.x};
        };
        if (kc == 0) {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int ky = a + 8 * r, my = (64 - ky) & 63;
                const C x0 = P[sx::swz(0, ky)], x0m = P[sx::swz(0, my)];
                const C n0 = NQ[ky], n0m = NQ[my];
                const C xa = {0.5f * (x0.x + x0m.x), 0.5f * (x0.y - x0m.y)};
                const C xb = {0.5f * (n0.x + n0m.x), 0.5f * (n0.y - n0m.y)};
                v[r] = split(xa, xb);
            }
        } else {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int ky = a + 8 * r, my = (64 - ky) & 63;
                v[r] = split(P[sx::swz(kc, ky)], cconj(P[sx::swz(32 - kc, my)]));
            }
        }
    }
    dft<float, 8, true>(v);
#pragma unroll
    for (int c = 1; c < 8; ++c) v[c] = cmul(v[c], cconj(TW[a * c]));
#pragma unroll
    for (int c = 0; c < 8; ++c) Wr[sx::tcol(g, a, c)] = v[c];
    sx::wave_sync();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = Wr[sx::tcol(g, r, a)];   // lane a now holds output c = a
    dft<float, 8, true>(v);                                      // v[d] = Z(kc, a + 8d)
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 8; ++d) P[sx::swz(kc, a + 8 * d)] = v[d];
    __syncthreads();

    // ---- phase 2: x inverse, LW / q, x forward ----
    C u[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) u[r] = P[sx::swz(lp + 4 * r, yr)];
    dft<float, 8, true>(u);
#pragma unroll
    for (int m1 = 1; m1 < 8; ++m1) u[m1] = cmul(u[m1], cconj(TW[2 * lp * m1]));
    sx::wave_sync();
#pragma unroll
    for (int m1 = 0; m1 < 8; ++m1) Wr[sx::trow(rho, lp, m1)] = u[m1];
    sx::wave_sync();
    C z0[4], z1[4];   // m1 = 2 lp and 2 lp + 1
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        z0[l] = Wr[sx::trow(rho, l, 2 * lp)];
        z1[l] = Wr[sx::trow(rho, l, 2 * lp + 1)];
    }
    dft<float, 4, true>(z0);   // z_e[m2] = z(2 lp + e + 8 m2): reals x = 4 lp + 16 m2 + 0..3
    dft<float, 4, true>(z1);
    {
        const float scale = (float)(2.0 / (double)n);
        const sx::PlaneBuf qb(L->stack + (size_t)(s_in + L->dir - 1) * n + (size_t)z * PLANE,
                              PLANE * 4);
        const sx::PlaneBuf lw2b(__builtin_bit_cast(const float*, lw2raw) + (size_t)z * PLANE,
                                PLANE * 4);
#pragma unroll
        for (int m2 = 0; m2 < 4; ++m2) {
            const sx::f4 wv = wr[m2];
            const sx::f4 q = {(z0[m2].x * scale) * wv.x, (z0[m2].y * scale) * wv.y,
                              (z1[m2].x * scale) * wv.z, (z1[m2].y * scale) * wv.w};
            qb.store(q, roff + 64u * m2);
            const sx::f4 w2 = (b2 == b) ? wv : lw2b.load(roff + 64u * m2);
            z0[m2] = {q.x * w2.x, q.y * w2.y};
            z1[m2] = {q.z * w2.z, q.w * w2.w};
        }
    }
    if (!more) return;
    dft<float, 4, false>(z0);   // z_e[l] = U(2 lp + e, l)
    dft<float, 4, false>(z1);
#pragma unroll
    for (int l = 1; l < 4; ++l) {
        z0[l] = cmul(z0[l], TW[2 * l * (2 * lp)]);
        z1[l] = cmul(z1[l], TW[2 * l * (2 * lp + 1)]);
    }
    sx::wave_sync();
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        Wr[sx::trow(rho, l, 2 * lp)] = z0[l];
        Wr[sx::trow(rho, l, 2 * lp + 1)] = z1[l];
    }
    sx::wave_sync();
#pragma unroll
    for (int m1 = 0; m1 < 8; ++m1) u[m1] = Wr[sx::trow(rho, lp, m1)];
    dft<float, 8, false>(u);   // u[r] = Z(lp + 4 r, yr)
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) P[sx::swz(lp + 4 * r, yr)] = u[r];
    __syncthreads();

    // ---- phase 3: y forward ----
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = P[sx::swz(kc, a + 8 * r)];
    dft<float, 8, false>(v);
#pragma unroll
    for (int c = 1; c < 8; ++c) v[c] = cmul(v[c], TW[a * c]);
    sx::wave_sync();
#pragma unroll
    for (int c = 0; c < 8; ++c) Wr[sx::tcol(g, a, c)] = v[c];
    sx::wave_sync();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = Wr[sx::tcol(g, r, a)];
    dft<float, 8, false>(v);   // v[d] = Zhat(kc, a + 8d)
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 8; ++d) P[sx::swz(kc, a + 8 * d)] = v[d];
    __syncthreads();

    // ---- phase 4: split r2c into the half spectrum ----
    auto r2c = [&](C zk, C zm, C tw) {   // zm: the unconjugated partner
        const C e = {0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y)};
        const C d = {zk.x - zm.x, zk.y + zm.y};
        const C o = {0.5f * d.y, -0.5f * d.x};
        return cadd(e, cmul(tw, o));
    };
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int i = s * 256 + tid, ky = i >> 5, kx = i & 31, my = (64 - ky) & 63;
        kb[ky * 33 + kx] = r2c(P[sx::swz(kx, ky)], P[sx::swz((32 - kx) & 31, my)], TW[kx]);
    }
    if (tid < 64) {
        const int my = (64 - tid) & 63;
        kb[tid * 33 + 32] = r2c(P[sx::swz(0, tid)], P[sx::swz(0, my)], C{-1.f, 0.f});
    }
}

// z-pencil pass over KB consecutive ky rows and all NXH kx columns: in the half-spectrum
// layout [kz][ky][kx] those KB*NXH values are contiguous for every kz, so each z row is
// one coalesced run.
// HALF (RQM4): the half steps' spectrum kb2 with the LD exponent halved
template <typename T, int NZ, int NXH, int KB, int NT, bool HALF = false>
__global__ void __launch_bounds__(NT) k_pencil_fx(const Lane<T>* __restrict__ lanes, int ny, int t) {
    constexpr int C = KB * NXH;
    __shared__ cpx<T> P[NZ * C];
    __shared__ cpx<T> TWZ[NZ];
    // grid (ny / KB, lanes).  The lane fields come in one round of scalar loads (the empty
    // asm keeps them ahead of every branch); a finished lane (t >= Ns) loads its (valid)
    // rows too and returns before its first barrier.
    const Lane<T>* L = lanes + blockIdx.y;
    const int ky0 = blockIdx.x * KB;
    unsigned long long kraw = __builtin_bit_cast(unsigned long long, HALF ? L->kb2 : L->kbuf),
                       tzr = __builtin_bit_cast(unsigned long long, L->twz);
    int nsteps = L->nsteps, s_in;
    int b = __builtin_amdgcn_readfirstlane(step_block(L, t, &s_in));
    asm volatile("" : "+s"(kraw), "+s"(tzr), "+s"(nsteps), "+s"(b));
    const auto kb = gp(__builtin_bit_cast(cpx<T>*, kraw) + (size_t)ky0 * NXH);
    const size_t zstride = (size_t)ny * NXH;
    // the KB ky rows of every kz into LDS: loads in groups of 8 per thread ahead of their
    // stores, a thread past the end repeating element i % CNT (no branch)
    constexpr int CNT = NZ * C, IT = (CNT + NT - 1) / NT, CH = 8;
#pragma unroll
    for (int c0 = 0; c0 < IT; c0 += CH) {
        cpx<T> v[CH];
        int o[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (c0 + c < IT) {
                const int i0 = (c0 + c) * NT + (int)threadIdx.x, i = i0 < CNT ? i0 : i0 % CNT;
                o[c] = i;
                v[c] = kb[(i / C) * zstride + (i % C)];
            }
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (c0 + c < IT) P[o[c]] = v[c];
    }
    lds_copy_in<T, NZ, NT>(TWZ, __builtin_bit_cast(const cpx<T>*, tzr));
    if (t >= nsteps) return;
    __syncthreads();
    cfft<T, NZ, C, C, 1, NT, false, 1>(P, TWZ);
    const double mc = HALF ? -pick_blk(L->ldc, b) / 2.0 : -pick_blk(L->ldc, b);
    const double* G = L->G;
#pragma unroll
    for (int i0 = 0; i0 < NZ * C; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < NZ * C) {
            const int z = i / C, c = i % C;
            const int ky = ky0 + c / NXH, kx = c % NXH;
            const T f = exp_t<T>(mc * ksq_at(G, (double)kx, (double)fold(ky, ny), (double)fold(z, NZ)));
            P[i].x *= f;
            P[i].y *= f;
        }
    }
    __syncthreads();
    cfft<T, NZ, C, C, 1, NT, true, 1>(P, TWZ);
#pragma unroll
    for (int i0 = 0; i0 < NZ * C; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < NZ * C) kb[(i / C) * zstride + (i % C)] = P[i];
    }
}

#define PF_CS64_NT 256
#define PF_CS128_NT 512

// z-pencil pass over half of one ky row: workgroup 2 ky + h takes the kx columns
// [h CW, min((h + 1) CW, NXH)), CW = ceil(NXH / 2).  For launches of few lanes, where one
// ky row per workgroup leaves CUs idle (one 64^3 system: 2 lanes x 64 rows), this gives
// twice the workgroups at half the LDS each.  The second half has one column fewer: its
// threads load a valid duplicate for it (no branch) and do not store it.
template <typename T, int NZ, int NXH, int NT>
__global__ void __launch_bounds__(NT) k_pencil_cs(const Lane<T>* __restrict__ lanes, int ny, int t) {
    constexpr int CW = (NXH + 1) / 2;
    __shared__ cpx<T> P[NZ * CW];
    __shared__ cpx<T> TWZ[NZ];
    // grid (2 ny, lanes)
    const Lane<T>* L = lanes + blockIdx.y;
    const int ky = blockIdx.x >> 1, h = blockIdx.x & 1;
    const int c0 = h * CW, cv = h ? NXH - CW : CW;
    unsigned long long kraw = __builtin_bit_cast(unsigned long long, L->kbuf),
                       tzr = __builtin_bit_cast(unsigned long long, L->twz);
    int nsteps = L->nsteps, s_in;
    int b = __builtin_amdgcn_readfirstlane(step_block(L, t, &s_in));
    asm volatile("" : "+s"(kraw), "+s"(tzr), "+s"(nsteps), "+s"(b));
    const auto kb = gp(__builtin_bit_cast(cpx<T>*, kraw) + (size_t)ky * NXH + c0);
    const size_t zstride = (size_t)ny * NXH;
    constexpr int CNT = NZ * CW, IT = (CNT + NT - 1) / NT, CH = 8;
#pragma unroll
    for (int q0 = 0; q0 < IT; q0 += CH) {
        cpx<T> v[CH];
        int o[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q)
            if (q0 + q < IT) {
                const int i0 = (q0 + q) * NT + (int)threadIdx.x, i = i0 < CNT ? i0 : i0 % CNT;
                const int c = i % CW, cc = c < cv ? c : cv - 1;
                o[q] = i;
                v[q] = kb[(i / CW) * zstride + cc];
            }
#pragma unroll
        for (int q = 0; q < CH; ++q)
            if (q0 + q < IT) P[o[q]] = v[q];
    }
    lds_copy_in<T, NZ, NT>(TWZ, __builtin_bit_cast(const cpx<T>*, tzr));
    if (t >= nsteps) return;
    __syncthreads();
    cfft<T, NZ, CW, CW, 1, NT, false, 1>(P, TWZ);
    const double mc = -pick_blk(L->ldc, b);
    const double* G = L->G;
    const double fky = (double)fold(ky, ny);
#pragma unroll
    for (int i0 = 0; i0 < CNT; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < CNT) {
            const int z = i / CW, kx = c0 + i % CW;
            const T f = exp_t<T>(mc * ksq_at(G, (double)kx, fky, (double)fold(z, NZ)));
            P[i].x *= f;
            P[i].y *= f;
        }
    }
    __syncthreads();
    cfft<T, NZ, CW, CW, 1, NT, true, 1>(P, TWZ);
#pragma unroll
    for (int i0 = 0; i0 < CNT; i0 += NT) {
        const int i = i0 + threadIdx.x;
        if (i < CNT && i % CW < cv) kb[(i / CW) * zstride + i % CW] = P[i];
    }
}

// ---------------------------------------------------------------------------------
// Register-resident transforms: one thread holds a whole line of N <= 64 values and
// runs N = R1 * R2 as R2 transforms of length R1, twiddles w_N^(n2 k1), and R1 transforms
// of length R2 (n = R2 n1 + n2, k = k1 + R1 k2); every index is a compile-time constant,
// so the line stays in VGPRs.  cos(2 pi j / 64) for j = 0..16; the other octants follow
// by symmetry.
constexpr double kCos64[17] = {
    1.0, 0.9951847266721969, 0.9807852804032304, 0.9569403357322088,
    0.9238795325112867, 0.881921264348355, 0.8314696123025452, 0.773010453362737,
    0.7071067811865476, 0.6343932841636455, 0.5555702330196022, 0.47139673682599764,
    0.3826834323650898, 0.2902846772544624, 0.19509032201612828, 0.0980171403295606,
    0.0};

constexpr double cos64(int j) {
    j &= 63;
    if (j <= 16) return kCos64[j];
    if (j <= 32) return -kCos64[32 - j];
    if (j <= 48) return -kCos64[j - 32];
    return kCos64[64 - j];
}
constexpr double sin64(int j) { return cos64(j - 16 + 64); }

// w_N^m = exp(-2 pi i m / N) (forward) or its conjugate (INV)
template <typename T, int N, bool INV>
__device__ __forceinline__ cpx<T> twr(int m) {
    const int j = m * (64 / N);
    return {(T)cos64(j), INV ? (T)sin64(j) : (T)(-sin64(j))};
}

template <typename T, int N, int R1, int R2, bool INV>
__device__ __forceinline__ void fft_reg(cpx<T> (&v)[N]) {
    static_assert(R1 * R2 == N && 64 % N == 0, "register transform size");
    cpx<T> y[N];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
        cpx<T> a[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[R2 * n1 + n2];
        dft<T, R1, INV>(a);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1)
            y[n2 * R1 + k1] = (n2 * k1 == 0) ? a[k1] : cmul(a[k1], twr<T, N, INV>(n2 * k1));
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
        cpx<T> a[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) a[n2] = y[n2 * R1 + k1];
        dft<T, R2, INV>(a);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = a[k2];
    }
}

// FP32 LD factors exp(mc k^2) of one z column, k = a + g jz: k^2 = A + jz (B + C jz)
struct ColumnLd {
    double A, B, C;
    __device__ __forceinline__ ColumnLd(double mc, double a0, double a1, double a2, double g0,
                                        double g1, double g2) {
        const double m = mc * 1.4426950408889634074;   // mc log2(e)
        A = m * (a0 * a0 + a1 * a1 + a2 * a2);
        B = m * (2.0 * (a0 * g0 + a1 * g1 + a2 * g2));
        C = m * (g0 * g0 + g1 * g1 + g2 * g2);
    }
    __device__ __forceinline__ float operator()(double jz) const {
        return __builtin_amdgcn_exp2f((float)(A + jz * (B + C * jz)));
    }
};

template <int N> struct RegSplit;
template <> struct RegSplit<64> { static constexpr int R1 = 8, R2 = 8; };
template <> struct RegSplit<32> { static constexpr int R1 = 8, R2 = 4; };
template <> struct RegSplit<16> { static constexpr int R1 = 4, R2 = 4; };

// Buffer access to one lane's half spectrum.  The resource holds the uniform base and
// size (buffer accesses past the size read 0 and write nothing), the thread passes its
// byte offset in one register and each row's offset goes in a scalar register: a column
// of NZ rows then costs one address register instead of NZ 64-bit addresses held from
// the loads to the stores.  Word 3 of the resource: 0x00020000 for gfx9 targets (raw
// 32-bit elements, as composable_kernel sets it for gfx950).
template <typename T>
struct SpecBuf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ SpecBuf(cpx<T>* base, unsigned bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(base, 0, (int)bytes, 0x00020000)) {}
    using V2 = unsigned __attribute__((ext_vector_type(2)));
    using V4 = unsigned __attribute__((ext_vector_type(4)));
    __device__ __forceinline__ cpx<T> load(unsigned voff, unsigned soff) const {
        if constexpr (sizeof(T) == 4) {
            // elements copied out as values first (a bit cast of a vector element
            // lvalue would read element 0)
            const V2 w = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
            const unsigned w0 = w.x, w1 = w.y;
            return {__uint_as_float(w0), __uint_as_float(w1)};
        } else {
            const V4 w = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
            const unsigned w0 = w.x, w1 = w.y, w2 = w.z, w3 = w.w;
            return {__hiloint2double((int)w1, (int)w0), __hiloint2double((int)w3, (int)w2)};
        }
    }
    __device__ __forceinline__ void store(cpx<T> v, unsigned voff, unsigned soff) const {
        if constexpr (sizeof(T) == 4) {
            const V2 w = {__float_as_uint(v.x), __float_as_uint(v.y)};
            __builtin_amdgcn_raw_buffer_store_b64(w, r, voff, soff, 0);
        } else {
            const V4 w = {(unsigned)__double2loint(v.x), (unsigned)__double2hiint(v.x),
                          (unsigned)__double2loint(v.y), (unsigned)__double2hiint(v.y)};
            __builtin_amdgcn_raw_buffer_store_b128(w, r, voff, soff, 0);
        }
    }
};

// z-pencil pass with one thread per (ky, kx) column: the column is read straight from the
// half spectrum (consecutive threads read consecutive kx, ky -> one coalesced run per kz),
// transformed, multiplied by LD, transformed back and written, with no LDS and no barrier.
template <typename T, int NZ, int NT>
__global__ void __launch_bounds__(NT) k_pencil_reg(const Lane<T>* __restrict__ lanes, int ny,
                                                   int nxh, int t) {
    constexpr int R1 = RegSplit<NZ>::R1, R2 = RegSplit<NZ>::R2;
    // grid (column blocks, lanes).  The column loads wait for one round of scalar loads of
    // the lane record (the empty asm keeps those loads ahead of every branch).  A finished
    // lane (t >= Ns) runs through and its stores go to an empty buffer range, so no branch
    // lets the compiler move the column loads after the test.
    const Lane<T>* L = lanes + blockIdx.y;
    unsigned long long kraw = __builtin_bit_cast(unsigned long long, L->kbuf);
    int nsteps = L->nsteps, s_in;
    int b = __builtin_amdgcn_readfirstlane(step_block(L, t, &s_in));
    asm volatile("" : "+s"(kraw), "+s"(nsteps), "+s"(b));
    cpx<T>* const kbp = __builtin_bit_cast(cpx<T>*, kraw);
    const int c = blockIdx.x * NT + threadIdx.x, ncol = nxh * ny;
    if (c >= ncol) return;
    const unsigned row = (unsigned)ncol * (unsigned)sizeof(cpx<T>);
    const SpecBuf<T> kb(kbp, row * (unsigned)NZ);
    const SpecBuf<T> kout(kbp, t < nsteps ? row * (unsigned)NZ : 0u);
    const unsigned off = (unsigned)c * (unsigned)sizeof(cpx<T>);
    cpx<T> v[NZ];
#pragma unroll
    for (int z = 0; z < NZ; ++z) v[z] = kb.load(off, (unsigned)z * row);
    fft_reg<T, NZ, R1, R2, false>(v);
    const double mc = -pick_blk(L->ldc, b);
    const double* G = L->G;
    const int ky = c / nxh, kx = c - ky * nxh;
    const double jx = (double)kx, jy = (double)fold(ky, ny);
    const double a0 = G[0] * jx + G[1] * jy, a1 = G[3] * jx + G[4] * jy,
                 a2 = G[6] * jx + G[7] * jy;
    if constexpr (sizeof(T) == 4) {
        // along the column k^2 = |a + g jz|^2 = A + jz (B + C jz) exactly, so FP64 gives
        // each exponent in two FMAs (coefficients scaled by mc log2(e))
        const ColumnLd ld(mc, a0, a1, a2, G[2], G[5], G[8]);
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            const float f = ld((double)(2 * z <= NZ ? z : NZ - z));
            v[z].x *= f;
            v[z].y *= f;
        }
    } else {
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            const double jz = (double)(2 * z <= NZ ? z : NZ - z);
            const double k0 = a0 + G[2] * jz, k1 = a1 + G[5] * jz, k2 = a2 + G[8] * jz;
            const T f = exp_t<T>(mc * (k0 * k0 + k1 * k1 + k2 * k2));
            v[z].x *= f;
            v[z].y *= f;
        }
    }
    fft_reg<T, NZ, R1, R2, true>(v);
#pragma unroll
    for (int z = 0; z < NZ; ++z) kout.store(v[z], off, (unsigned)z * row);
}

// z-pencil pass with two threads per column (FP32, NZ = 64): lane pair (2c, 2c+1) holds
// the even and odd z samples of column c.  Each lane transforms its 32 samples in
// registers; one radix-2 stage across the pair (adjacent-lane swap) completes the 64-point
// transform, the LD multiply acts on the 32 wave numbers each lane then holds, and the
// inverse runs the same steps backwards.  Half the registers of the one-thread form, so
// two waves fit per SIMD.
__device__ __forceinline__ float swap_pair(float x) {
    // quad_perm [1, 0, 3, 2]: exchange with the adjacent lane
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
}

template <int NT>
__global__ void __launch_bounds__(NT) k_pencil_pair(const Lane<float>* __restrict__ lanes, int ny,
                                                    int nxh, int t) {
    constexpr int NZ = 64, H = 32;
    // grid (column blocks, lanes); the lane fields come in one round of scalar loads, and
    // a finished lane (t >= Ns) runs through with its stores going to an empty range
    const Lane<float>* L = lanes + blockIdx.y;
    unsigned long long kraw = __builtin_bit_cast(unsigned long long, L->kbuf);
    int nsteps = L->nsteps, s_in;
    int b = __builtin_amdgcn_readfirstlane(step_block(L, t, &s_in));
    asm volatile("" : "+s"(kraw), "+s"(nsteps), "+s"(b));
    cpx<float>* const kbp = __builtin_bit_cast(cpx<float>*, kraw);
    const int c = blockIdx.x * (NT / 2) + (threadIdx.x >> 1);
    const int half = threadIdx.x & 1;
    const int ncol = nxh * ny;
    const bool valid = c < ncol;
    // row 2m + half of column c: the half's row in the thread's offset, 2m rows in the
    // scalar one
    const unsigned row = (unsigned)ncol * 8u;
    const SpecBuf<float> kb(kbp, row * (unsigned)NZ);
    const SpecBuf<float> kout(kbp, t < nsteps ? row * (unsigned)NZ : 0u);
    const unsigned off = (unsigned)(valid ? c : 0) * 8u + (unsigned)half * row;
    cpx<float> v[H], u[H];
#pragma unroll
    for (int m = 0; m < H; ++m) {
        const cpx<float> x = kb.load(off, (unsigned)(2 * m) * row);
        v[m] = valid ? x : cpx<float>{0.f, 0.f};
    }
    // even (half 0) / odd (half 1) 32-point transforms, then X[k] = E + w^k O,
    // X[k+32] = E - w^k O
    fft_reg<float, H, RegSplit<H>::R1, RegSplit<H>::R2, false>(v);
#pragma unroll
    for (int k = 0; k < H; ++k) u[k] = {swap_pair(v[k].x), swap_pair(v[k].y)};
    const float sg = half ? -1.f : 1.f;
#pragma unroll
    for (int k = 0; k < H; ++k) {
        const cpx<float> e = half ? u[k] : v[k], o = half ? v[k] : u[k];
        const cpx<float> wo = cmul(twr<float, NZ, false>(k), o);
        v[k] = {e.x + sg * wo.x, e.y + sg * wo.y};
    }
    const double mc = -pick_blk(L->ldc, b);
    const double* G = L->G;
    const int ky = c / nxh, kx = c - ky * nxh;
    const double jx = (double)kx, jy = (double)fold(ky, ny);
    const ColumnLd ld(mc, G[0] * jx + G[1] * jy, G[3] * jx + G[4] * jy, G[6] * jx + G[7] * jy,
                      G[2], G[5], G[8]);
#pragma unroll
    for (int k = 0; k < H; ++k) {
        const int kz = k + H * half;
        const float f = ld((double)(2 * kz <= NZ ? kz : NZ - kz));
        v[k].x *= f;
        v[k].y *= f;
    }
    // inverse: even samples from X[k] + X[k+32], odd from (X[k] - X[k+32]) w^-k
#pragma unroll
    for (int k = 0; k < H; ++k) u[k] = {swap_pair(v[k].x), swap_pair(v[k].y)};
#pragma unroll
    for (int k = 0; k < H; ++k) {
        const cpx<float> lo = half ? u[k] : v[k], hi = half ? v[k] : u[k];
        const cpx<float> d = half ? csub(lo, hi) : cadd(lo, hi);
        v[k] = half ? cmul(d, twr<float, NZ, true>(k)) : d;
    }
    fft_reg<float, H, RegSplit<H>::R1, RegSplit<H>::R2, true>(v);
    if (valid) {
#pragma unroll
        for (int m = 0; m < H; ++m) kout.store(v[m], off, (unsigned)(2 * m) * row);
    }
}

// ---------------------------------------------------------------------------------
// Schedules and dispatch: the dispatchers return false when the schedule asks for the
// generic kernels or the size has no specialisation.

static std::mutex sched_mu;
static std::map<std::array<int, 4>, Schedule> sched_table;

Schedule& schedule(int elem_bytes, int nx, int ny, int nz) {
    std::lock_guard<std::mutex> lock(sched_mu);
    return sched_table[{elem_bytes, nx, ny, nz}];
}

void schedule_reset() {
    std::lock_guard<std::mutex> lock(sched_mu);
    sched_table.clear();
}

int device_cus() {
    static std::mutex mu;
    static std::map<int, int> cus;
    int dev = 0;
    PF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = cus.find(dev);
    if (it != cus.end()) return it->second;
    int n = 0;
    PF_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    return cus[dev] = n > 0 ? n : 1;
}

template <typename T>
bool fixed_planar(const Lane<T>* d, int nl, int nx, int ny, int t0, int t1, hipStream_t st) {
    const int pk = schedule(sizeof(T), nx, ny, 1).planar;
    if (pk == 0) return false;
#define PF_PL(X, Y, NT)                                                                    \
    if (nx == X && ny == Y) {                                                              \
        hipLaunchKernelGGL((k_planar_fx<T, X, Y, NT>), dim3(nl), dim3(NT), 0, st, d, t0, t1); \
        PF_HIP(hipGetLastError());                                                         \
        return true;                                                                       \
    }
    PF_PL(32, 1, 64)
    PF_PL(64, 1, 64)
    PF_PL(128, 1, 64)
    PF_PL(256, 1, 128)
    // 2-D planes: one workgroup per lane runs the whole sweep, so a launch holds at most a
    // few hundred workgroups and the threads per plane set the latency of every step.
    // Default 1024 threads (measured on MI355X, 256-instance hexagonal sweep: 9.0 -> 15.5
    // batched solves/s, plain FP64 0.35 -> 0.54); 256 for FP64 128 x 128, whose 1024-thread
    // form spills.
    if (pk == 3 || (pk < 0 && !(sizeof(T) == 8 && nx == 128))) {
        PF_PL(32, 32, 1024)
        PF_PL(48, 48, 1024)
        PF_PL(64, 64, 1024)
        PF_PL(96, 96, 1024)
        PF_PL(128, 128, 1024)
    } else if (pk == 2) {
        PF_PL(32, 32, 512)
        PF_PL(48, 48, 512)
        PF_PL(64, 64, 512)
        PF_PL(96, 96, 512)
        PF_PL(128, 128, 512)
    }
    PF_PL(32, 32, 256)
    PF_PL(48, 48, 256)
    PF_PL(64, 64, 256)
    PF_PL(96, 96, 256)
    PF_PL(128, 128, 256)
#undef PF_PL
    return false;
}

// RQM4 planes and lines; one: the two planes do not fit in LDS together (the caller's
// engine decides), else both planes in LDS
template <typename T>
bool fixed_planar_rqm4(const Lane<T>* d, int nl, int nx, int ny, int t0, int t1, bool one,
                       hipStream_t st) {
    if (schedule(sizeof(T), nx, ny, 1).planar == 0) return false;
    if (one) {
        if constexpr (sizeof(T) == 8) {
            if (nx == 128 && ny == 128) {
                hipLaunchKernelGGL((k_planar_rq1_fx<T, 128, 128, 256>), dim3(nl), dim3(256), 0, st,
                                   d, t0, t1);
                PF_HIP(hipGetLastError());
                return true;
            }
        }
        return false;
    }
#define PF_RQ(X, Y, NT)                                                                    \
    if (nx == X && ny == Y) {                                                              \
        hipLaunchKernelGGL((k_planar_rq_fx<T, X, Y, NT>), dim3(nl), dim3(NT), 0, st, d, t0, t1); \
        PF_HIP(hipGetLastError());                                                         \
        return true;                                                                       \
    }
    PF_RQ(32, 1, 64)
    PF_RQ(64, 1, 64)
    PF_RQ(128, 1, 64)
    PF_RQ(256, 1, 128)
    PF_RQ(32, 32, 1024)
    PF_RQ(48, 48, 1024)
    PF_RQ(64, 64, 1024)
    // FP64 96^2: 512 threads (the 1024-thread form spills 43 VGPRs); FP32 128^2 keeps
    // 1024 (a few spills; the 512-thread form spills more)
    if constexpr (sizeof(T) == 4) {
        PF_RQ(96, 96, 1024)
        PF_RQ(128, 128, 1024)
    } else {
        PF_RQ(96, 96, 512)
    }
#undef PF_RQ
    return false;
}

// 3-D RQM4 with both planes in LDS (the caller checks that two fit): the slab modes and
// the pencils (half: the half steps' spectrum) at compile-time sizes
template <typename T>
bool fixed_slab_rqm4(const Lane<T>* d, int nl, int nx, int ny, int nz, int mode, int t,
                     hipStream_t st) {
    if (schedule(sizeof(T), nx, ny, nz).slab == 0) return false;
#define PF_RS(X, Y, NT)                                                                    \
    if (nx == X && ny == Y) {                                                              \
        if (mode == RQ_START)                                                              \
            hipLaunchKernelGGL((k_slab_rq_fx<T, X, Y, RQ_START, NT>), dim3(nz, nl), dim3(NT), 0, \
                               st, d, nz, t);                                              \
        else if (mode == RQ_MID)                                                           \
            hipLaunchKernelGGL((k_slab_rq_fx<T, X, Y, RQ_MID, NT>), dim3(nz, nl), dim3(NT), 0, \
                               st, d, nz, t);                                              \
        else                                                                               \
            hipLaunchKernelGGL((k_slab_rq_fx<T, X, Y, RQ_END, NT>), dim3(nz, nl), dim3(NT), 0, \
                               st, d, nz, t);                                              \
        PF_HIP(hipGetLastError());                                                         \
        return true;                                                                       \
    }
    PF_RS(32, 32, 256)
    // launches of at most one plane per CU: 1024 threads (measured on MI355X, one 64^3
    // system's RQM4 density!: FP64 5.98 -> 5.85 ms, FP32 4.26 -> 4.16 ms)
    if (nl * nz <= device_cus()) { PF_RS(64, 64, 1024) }
    PF_RS(64, 64, 512)
    if constexpr (sizeof(T) == 4) { PF_RS(128, 128, 512) }
#undef PF_RS
    return false;
}

template <typename T>
bool fixed_pencil_rqm4(const Lane<T>* d, int nl, int nx, int ny, int nz, int t, int half,
                       hipStream_t st) {
    if (schedule(sizeof(T), nx, ny, nz).pencil == 0) return false;
#define PF_RP(X, Z, KB, NT)                                                                \
    if (nx == X && nz == Z && ny % KB == 0) {                                              \
        if (half)                                                                          \
            hipLaunchKernelGGL((k_pencil_fx<T, Z, X / 2 + 1, KB, NT, true>), dim3(ny / KB, nl), \
                               dim3(NT), 0, st, d, ny, t);                                 \
        else                                                                               \
            hipLaunchKernelGGL((k_pencil_fx<T, Z, X / 2 + 1, KB, NT, false>), dim3(ny / KB, nl), \
                               dim3(NT), 0, st, d, ny, t);                                 \
        PF_HIP(hipGetLastError());                                                         \
        return true;                                                                       \
    }
    PF_RP(32, 32, 4, 256)
    PF_RP(64, 64, 1, 512)
    PF_RP(128, 128, 1, 512)
#undef PF_RP
    return false;
}

template <typename T>
bool fixed_slab(const Lane<T>* d, int nl, int nx, int ny, int nz, int mode, int t,
                hipStream_t st) {
    const int sk = schedule(sizeof(T), nx, ny, nz).slab;
    if (sk == 0) return false;
#define PF_SL1(X, Y, NT)                                                                    \
    if (nx == X && ny == Y) {                                                               \
        if (mode == SLAB_FWD)                                                               \
            hipLaunchKernelGGL((k_slab_fx<T, X, Y, SLAB_FWD, NT>), dim3(nz, nl), dim3(NT), 0, \
                               st, d, nz, t);                                               \
        else                                                                                \
            hipLaunchKernelGGL((k_slab_fx<T, X, Y, SLAB_INVFWD, NT>), dim3(nz, nl), dim3(NT), \
                               0, st, d, nz, t);                                            \
        PF_HIP(hipGetLastError());                                                          \
        return true;                                                                        \
    }
#define PF_SL(X, Y) PF_SL1(X, Y, 256)
    PF_SL(32, 32)
    if (sk == 1) { PF_SL1(64, 64, 256) }
    if constexpr (sizeof(T) == 4) {
        // FP32 64 x 64 planes, steps after the first: the shuffle-exchange form by default
        // (measured on MI355X: equal at 32 systems per batch, +7% for one system)
        if ((sk == 3 || sk < 0) && nx == 64 && ny == 64 && mode == SLAB_INVFWD) {
            hipLaunchKernelGGL(k_slab_sx, dim3(nz, nl), dim3(256), 0, st, d, nz, t);
            PF_HIP(hipGetLastError());
            return true;
        }
    }
    PF_SL1(64, 64, 512)
    // 128 x 128 planes: 256 threads (form 1), 512 (form 2, the FP64 default: one plane of
    // 133 KiB per CU either way, with two waves per SIMD instead of one; measured on MI355X
    // on the 128^3 blend in FP64, 23.9 -> 19.4 ms per density! with the 512-thread pencil),
    // 1024 (form 4, FP32 only, whose default it is: the FP64 form spills; measured on the
    // FP32 blend, 13.4 -> 11.5 ms per density! with the 1024-thread pencil)
    if (nx == 128 && ny == 128) {
        if (sk == 1) { PF_SL1(128, 128, 256) }
        if constexpr (sizeof(T) == 4) {
            // FP32 default with more planes than CUs (three lanes of the 128^3 blend): 512
            // threads within 128 VGPRs (a few spills), so that two planes (2 x 67 KiB of LDS)
            // share a CU and the launch runs in one round instead of two; same arithmetic.
            // Measured on MI355X: 128^3 FP32 blend 11.08 -> 10.78 ms per density!; the
            // 1024-thread form within 64 VGPRs (two per CU, more spills) 12.56 ms.
            if (sk < 0 && mode == SLAB_INVFWD && nl * nz > device_cus()) {
                hipLaunchKernelGGL((k_slab_fx<T, 128, 128, SLAB_INVFWD, 512, 4>), dim3(nz, nl),
                                   dim3(512), 0, st, d, nz, t);
                PF_HIP(hipGetLastError());
                return true;
            }
            if (sk == 4 || sk < 0) { PF_SL1(128, 128, 1024) }
        }
        PF_SL1(128, 128, 512)
    }
#undef PF_SL
#undef PF_SL1
    return false;
}

// ky rows per workgroup: 4, 2, 1 for nx = 32, 64, 128 (NXH = 17, 33, 65)
template <typename T>
bool fixed_pencil(const Lane<T>* d, int nl, int nx, int ny, int nz, bool alone, int t,
                  hipStream_t st) {
    const int pk = schedule(sizeof(T), nx, ny, nz).pencil;
    if (pk == 0) return false;
#define PF_PE(X, Z, KB)                                                                     \
    if (nx == X && nz == Z && ny % KB == 0) {                                               \
        hipLaunchKernelGGL((k_pencil_fx<T, Z, X / 2 + 1, KB, 256>), dim3(ny / KB, nl),      \
                           dim3(256), 0, st, d, ny, t);                                     \
        PF_HIP(hipGetLastError());                                                          \
        return true;                                                                        \
    }
    if constexpr (sizeof(T) == 4) {
        if (pk == 3 && nz == 64) {
            const int ncol = (nx / 2 + 1) * ny, per = (ncol + 31) / 32;
            hipLaunchKernelGGL((k_pencil_pair<64>), dim3(per, nl), dim3(64), 0, st,
                               reinterpret_cast<const Lane<float>*>(d), ny, nx / 2 + 1, t);
            PF_HIP(hipGetLastError());
            return true;
        }
    }
    // half ky rows (form 6): 64- and 128-point lines.  The default for a single system's
    // launches of a few lanes, at most two ky rows per CU, of FP64 128-point lines, and of
    // FP32 128-point lines when one row per workgroup would take two rounds of the CUs.
    // Measured on MI355X, 128^3 blend per launch: FP64 26.4 -> 22.6 us (2 lanes), 46.5 ->
    // 31.6 us (3 lanes); FP32 26.4 -> 22.4 us (3 lanes), but 15.6 -> 17.8 us for 2 lanes,
    // which keep the 1024-thread row.  One 64^3 FP64 system: no change.  One 64^3 FP32
    // system: 1.64 -> 1.57 ms per density!, but not the default there: the variable-cell
    // SCFT of that system in FP32 (64^3 BCC-biased seed, momentum + DomainUpdater) then
    // ends above its stress tolerance of 3e-5 after 5000 steps, where the register pencil's
    // rounding lets it converge (the Domain step's finite-difference Jacobian, dx = 1e-8,
    // sits at the FP32 noise floor).  In a batch, where another lane group's launches run
    // beside these, the register pencil stays faster (FP32 64^3, 2 to 8 systems: 1% to
    // 13%).  Each column's transform is the same arithmetic as in the one-row LDS forms.
    const int rows = nl * ny, cus = device_cus();
    const bool cs_default = pk < 0 && alone && rows <= 2 * cus && nz == 128 &&
                            (sizeof(T) == 8 || rows > cus);
    // FP64 64-point lines of one system's launches (at most one ky row per CU): half rows
    // with 512 threads (measured on MI355X, one 64^3 density!: 8.8 -> 8.45 us per launch,
    // 2.12 -> 2.07 ms)
    if constexpr (sizeof(T) == 8) {
        if (pk < 0 && alone && rows <= cus && nx == 64 && nz == 64) {
            hipLaunchKernelGGL((k_pencil_cs<T, 64, 33, 512>), dim3(2 * ny, nl), dim3(512), 0,
                               st, d, ny, t);
            PF_HIP(hipGetLastError());
            return true;
        }
    }
    if (nx == nz && (nz == 64 || nz == 128) && (pk == 6 || cs_default)) {
        if (nz == 64)
            hipLaunchKernelGGL((k_pencil_cs<T, 64, 33, PF_CS64_NT>), dim3(2 * ny, nl),
                               dim3(PF_CS64_NT), 0, st, d, ny, t);
        else
            hipLaunchKernelGGL((k_pencil_cs<T, 128, 65, PF_CS128_NT>), dim3(2 * ny, nl),
                               dim3(PF_CS128_NT), 0, st, d, ny, t);
        PF_HIP(hipGetLastError());
        return true;
    }
    if (pk != 1 && pk != 4 && pk != 5 && pk != 6) {
        // one thread per column, the whole z line in registers
        const int ncol = (nx / 2 + 1) * ny;
#define PF_PR(Z)                                                                            \
        if (nz == Z) {                                                                      \
            const int per = (ncol + 63) / 64;                                               \
            hipLaunchKernelGGL((k_pencil_reg<T, Z, 64>), dim3(per, nl), dim3(64), 0, st, d,   \
                               ny, nx / 2 + 1, t);                                          \
            PF_HIP(hipGetLastError());                                                      \
            return true;                                                                    \
        }
        if constexpr (sizeof(T) == 4) {
            PF_PR(64)
            PF_PR(32)
        } else {
            PF_PR(32)
        }
#undef PF_PR
    }
    PF_PE(32, 32, 4)
    // 64-point lines: two ky rows per workgroup, or one (form 4: twice the workgroups;
    // measured on MI355X, one 64^3 FP64 system's density! 2.83 -> 2.42 ms), or one with 512
    // threads (form 5, the FP64 default: 2.28 -> 2.15 ms, bitwise the same densities)
    if (nx == 64 && nz == 64 && (pk == 5 || (pk < 0 && sizeof(T) == 8))) {
        hipLaunchKernelGGL((k_pencil_fx<T, 64, 33, 1, 512>), dim3(ny, nl), dim3(512), 0, st, d,
                           ny, t);
        PF_HIP(hipGetLastError());
        return true;
    }
    if (pk == 4) { PF_PE(64, 64, 1) }
    PF_PE(64, 64, 2)
    // 128-point lines: the LDS pencil with 512 threads (form 4, the FP64 default: 23.9 ->
    // 22.0 ms per FP64 blend density! on its own) or 1024 (form 5, FP32 only, its default)
#define PF_PE_NT(NT)                                                                        \
    if (nx == 128 && nz == 128) {                                                           \
        hipLaunchKernelGGL((k_pencil_fx<T, 128, 65, 1, NT>), dim3(ny, nl), dim3(NT), 0, st,    \
                           d, ny, t);                                                       \
        PF_HIP(hipGetLastError());                                                          \
        return true;                                                                        \
    }
    if constexpr (sizeof(T) == 4) {
        if (pk == 4) { PF_PE_NT(512) }
        if (pk == 5 || pk < 0) { PF_PE_NT(1024) }
    } else {
        if (pk == 4 || pk < 0) { PF_PE_NT(512) }
    }
#undef PF_PE_NT
    PF_PE(128, 128, 1)
#undef PF_PE
    return false;
}

template bool fixed_planar<float>(const Lane<float>*, int, int, int, int, int, hipStream_t);
template bool fixed_planar<double>(const Lane<double>*, int, int, int, int, int, hipStream_t);
template bool fixed_planar_rqm4<float>(const Lane<float>*, int, int, int, int, int, bool,
                                       hipStream_t);
template bool fixed_planar_rqm4<double>(const Lane<double>*, int, int, int, int, int, bool,
                                        hipStream_t);
template bool fixed_slab_rqm4<float>(const Lane<float>*, int, int, int, int, int, int, hipStream_t);
template bool fixed_slab_rqm4<double>(const Lane<double>*, int, int, int, int, int, int, hipStream_t);
template bool fixed_pencil_rqm4<float>(const Lane<float>*, int, int, int, int, int, int, hipStream_t);
template bool fixed_pencil_rqm4<double>(const Lane<double>*, int, int, int, int, int, int, hipStream_t);
template bool fixed_slab<float>(const Lane<float>*, int, int, int, int, int, int, hipStream_t);
template bool fixed_slab<double>(const Lane<double>*, int, int, int, int, int, int, hipStream_t);
template bool fixed_pencil<float>(const Lane<float>*, int, int, int, int, bool, int, hipStream_t);
template bool fixed_pencil<double>(const Lane<double>*, int, int, int, int, bool, int,
                                   hipStream_t);

}  // namespace pf
