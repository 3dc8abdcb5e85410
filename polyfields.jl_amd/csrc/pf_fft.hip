// Propagator kernels: LDS-resident Stockham transforms fused with the Strang-step
// multiplies of the pseudo-spectral MDE solver, and the k-space stress reductions.
//
// Reference semantics (PolyFields.jl): one RK2 step is
//   qr = LW .* q_s ; qk = rfft(qr) ; qk .*= LD ; qr = irfft(qk) (1/n) ; q_{s+1} = LW .* qr
// (src/system/mde.jl:76-81) with LW = exp(-ds/2 v w), LD = exp(-ds b^2/6 k^2)
// (src/system/mde.jl:47-54) and k^2 on folded wave indices (src/system/cell.jl:268-285).
//
// Layout: real grids are x fastest, q[(z*ny + y)*nx + x]; the half spectrum keeps
// kx = 0..nx/2 fastest, k[(kz*ny + ky)*(nx/2+1) + kx] (the reference's FFTW r2c layout).
// The x transform of a real row of length nx = 2M runs as a length-M complex transform
// of the packed pairs (x[2m], x[2m+1]) followed by the split step that yields the M+1
// half-spectrum entries; the inverse runs the split step backwards.
#include "pf_fft_core.h"

namespace pf {

// ---------------------------------------------------------------------------------
// 1-D / 2-D lanes: a workgroup owns one lane's whole plane and advances it through
// contour steps [t0, t1) without leaving LDS; every q_s is written to the stack.
template <typename T>
__global__ void __launch_bounds__(256) k_planar_steps(const Lane<T>* __restrict__ lanes, int t0,
                                                      int t1) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* S = reinterpret_cast<cpx<T>*>(smem_raw);
    const Lane<T>* L = lanes + blockIdx.x;
    const int nx = L->nx, ny = L->ny, M = nx >> 1, ld = M + 1, n = nx * ny;
    const int tend = min(t1, L->nsteps);
    if (t0 >= tend) return;
    const int np = n >> 1;
    const T scale = T(2.0 / (double)n);
    int s_in;
    step_block(L, t0, &s_in);
    {
        const cpx<T>* qin = reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * n);
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            S[y * ld + (i - y * M)] = qin[i];
        }
    }
    if (L->rqm4 == 2) {
        // RQM4 with one plane of LDS (planes too large for two): the half steps re-read the
        // step's input from the stack, and the full step's result q1 waits in the output
        // slice until the combination.  Each thread reads back only the entries it wrote
        // itself, and the products are those of the two-plane form below.
        const int nkp = ny * (M + 1);
        for (int t = t0; t < tend; ++t) {
            const int b = step_block(L, t, &s_in);
            const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b]);
            const cpx<T>* lw2 = reinterpret_cast<const cpx<T>*>(L->lw2[b]);
            const cpx<T>* qin = reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * n);
            cpx<T>* qout = reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * n);
            const double mc = -L->ldc[b];
            for (int pass = 0; pass < 3; ++pass) {
                const bool half = pass > 0;
                if (pass < 2) {
                    __syncthreads();
                    for (int i = threadIdx.x; i < np; i += blockDim.x) {
                        const int y = i / M, o = y * ld + (i - y * M);
                        const cpx<T> w = pass == 0 ? lw[i] : lw2[i];
                        const cpx<T> v = pass == 0 ? S[o] : qin[i];
                        S[o] = {w.x * v.x, w.y * v.y};
                    }
                    __syncthreads();
                }
                plane_forward(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
                for (int i = threadIdx.x; i < nkp; i += blockDim.x) {
                    const int y = i / (M + 1), kx = i - y * (M + 1);
                    const double a = mc * ksq_at(L->G, (double)kx, (double)fold(y, ny), 0.0);
                    const T f = exp_t<T>(half ? a / 2.0 : a);
                    S[y * ld + kx].x *= f;
                    S[y * ld + kx].y *= f;
                }
                __syncthreads();
                plane_inverse(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
                const cpx<T>* wm = pass == 2 ? lw2 : lw;   // q1 *= LW1; mid *= LW1; q2 *= LW2
                for (int i = threadIdx.x; i < np; i += blockDim.x) {
                    const int y = i / M, o = y * ld + (i - y * M);
                    const cpx<T> w = wm[i];
                    S[o] = {(S[o].x * scale) * w.x, (S[o].y * scale) * w.y};
                    if (pass == 0) qout[i] = S[o];
                }
                __syncthreads();
            }
            for (int i = threadIdx.x; i < np; i += blockDim.x) {
                const int y = i / M, o = y * ld + (i - y * M);
                const cpx<T> q1 = qout[i], q2 = S[o];
                const cpx<T> v = {(T(4) * q2.x - q1.x) / T(3), (T(4) * q2.y - q1.y) / T(3)};
                S[o] = v;
                qout[i] = v;
            }
        }
        return;
    }
    if (L->rqm4) {
        // RQM4 (src/system/mde.jl:83-104): q1 = full step, q2 = two half steps with LW1 in
        // the middle, q_out = (4 q2 - q1) / 3.  S2 is a second plane of LDS.
        cpx<T>* S2 = S + ny * ld;
        const int nkp = ny * (M + 1);
        for (int t = t0; t < tend; ++t) {
            const int b = step_block(L, t, &s_in);
            const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b]);
            const cpx<T>* lw2 = reinterpret_cast<const cpx<T>*>(L->lw2[b]);
            const double mc = -L->ldc[b];
            __syncthreads();
            for (int i = threadIdx.x; i < np; i += blockDim.x) {
                const int y = i / M, o = y * ld + (i - y * M);
                const cpx<T> w = lw[i], w2 = lw2[i], v = S[o];
                S2[o] = {w2.x * v.x, w2.y * v.y};
                S[o] = {w.x * v.x, w.y * v.y};
            }
            __syncthreads();
            for (int pass = 0; pass < 3; ++pass) {
                cpx<T>* X = pass == 0 ? S : S2;
                const bool half = pass > 0;
                plane_forward(X, M, ny, ld, L->px, L->py, L->twx, L->twy);
                for (int i = threadIdx.x; i < nkp; i += blockDim.x) {
                    const int y = i / (M + 1), kx = i - y * (M + 1);
                    const double a = mc * ksq_at(L->G, (double)kx, (double)fold(y, ny), 0.0);
                    const T f = exp_t<T>(half ? a / 2.0 : a);
                    X[y * ld + kx].x *= f;
                    X[y * ld + kx].y *= f;
                }
                __syncthreads();
                plane_inverse(X, M, ny, ld, L->px, L->py, L->twx, L->twy);
                const cpx<T>* wm = pass == 2 ? lw2 : lw;   // q1 *= LW1; mid *= LW1; q2 *= LW2
                for (int i = threadIdx.x; i < np; i += blockDim.x) {
                    const int y = i / M, o = y * ld + (i - y * M);
                    const cpx<T> w = wm[i];
                    X[o] = {(X[o].x * scale) * w.x, (X[o].y * scale) * w.y};
                }
                __syncthreads();
            }
            cpx<T>* qout = reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * n);
            for (int i = threadIdx.x; i < np; i += blockDim.x) {
                const int y = i / M, o = y * ld + (i - y * M);
                const cpx<T> q1 = S[o], q2 = S2[o];
                const cpx<T> v = {(T(4) * q2.x - q1.x) / T(3), (T(4) * q2.y - q1.y) / T(3)};
                S[o] = v;
                qout[i] = v;
            }
        }
        return;
    }
    for (int t = t0; t < tend; ++t) {
        const int b = step_block(L, t, &s_in);
        const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b]);
        __syncthreads();
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            cpx<T>& v = S[y * ld + (i - y * M)];
            const cpx<T> w = lw[i];
            v.x *= w.x;
            v.y *= w.y;
        }
        __syncthreads();
        plane_forward(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
        const double mc = -L->ldc[b];
        const int nkp = ny * (M + 1);
        for (int i = threadIdx.x; i < nkp; i += blockDim.x) {
            const int y = i / (M + 1), kx = i - y * (M + 1);
            const T f = exp_t<T>(mc * ksq_at(L->G, (double)kx, (double)fold(y, ny), 0.0));
            cpx<T>& v = S[y * ld + kx];
            v.x *= f;
            v.y *= f;
        }
        __syncthreads();
        plane_inverse(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
        cpx<T>* qout = reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * n);
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            cpx<T>& v = S[y * ld + (i - y * M)];
            const cpx<T> w = lw[i];
            const cpx<T> o = {(v.x * scale) * w.x, (v.y * scale) * w.y};
            v = o;
            qout[i] = o;
        }
    }
}

// ---------------------------------------------------------------------------------
// 3-D lanes, xy-slab pass.  SLAB_FWD: LW_b(t) * q_{s_in(t)} -> x,y forward -> kbuf.
// SLAB_INVFWD: kbuf -> y,x inverse -> scale, * LW_b(t) -> q_{s_out(t)}; then, when a
// step t+1 exists, * LW_b(t+1) -> x,y forward -> kbuf.
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_slab(const Lane<T>* __restrict__ lanes, int nz, int t) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* S = reinterpret_cast<cpx<T>*>(smem_raw);
    const int lane = blockIdx.x / nz;
    const int z = blockIdx.x - lane * nz;
    const Lane<T>* L = lanes + lane;
    if (t >= L->nsteps) return;
    const int nx = L->nx, ny = L->ny, M = nx >> 1, ld = M + 1;
    const int plane = nx * ny, n = plane * nz, np = plane >> 1;
    const int nxh = M + 1, kplane = nxh * ny;
    cpx<T>* kb = L->kbuf + (size_t)z * kplane;
    int s_in;
    int b = step_block(L, t, &s_in);
    bool fwd = (MODE == SLAB_FWD);
    if (MODE == SLAB_INVFWD) {
        for (int i = threadIdx.x; i < kplane; i += blockDim.x) {
            const int y = i / nxh;
            S[y * ld + (i - y * nxh)] = kb[i];
        }
        __syncthreads();
        plane_inverse(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
        const T scale = T(2.0 / (double)n);
        const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b] + (size_t)z * plane);
        cpx<T>* qout = reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * n +
                                                 (size_t)z * plane);
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            cpx<T>& v = S[y * ld + (i - y * M)];
            const cpx<T> w = lw[i];
            const cpx<T> o = {(v.x * scale) * w.x, (v.y * scale) * w.y};
            v = o;
            qout[i] = o;
        }
        if (t + 1 >= L->nsteps) return;
        b = step_block(L, t + 1, &s_in);
        fwd = true;
        __syncthreads();
        const cpx<T>* lw2 = reinterpret_cast<const cpx<T>*>(L->lw[b] + (size_t)z * plane);
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            cpx<T>& v = S[y * ld + (i - y * M)];
            const cpx<T> w = lw2[i];
            v.x *= w.x;
            v.y *= w.y;
        }
    } else {
        const cpx<T>* qin = reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * n +
                                                            (size_t)z * plane);
        const cpx<T>* lw = reinterpret_cast<const cpx<T>*>(L->lw[b] + (size_t)z * plane);
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            const cpx<T> q = qin[i], w = lw[i];
            S[y * ld + (i - y * M)] = {q.x * w.x, q.y * w.y};
        }
    }
    if (fwd) {
        __syncthreads();
        plane_forward(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
        for (int i = threadIdx.x; i < kplane; i += blockDim.x) {
            const int y = i / nxh;
            kb[i] = S[y * ld + (i - y * nxh)];
        }
    }
}

// 3-D lanes, RQM4 steps (src/system/mde.jl:83-104): the full step
// q1 = LW1 F^-1 LD1 F LW1 q and two half steps, mid = LW1 F^-1 LD2 F LW2 q and
// q2 = LW2 F^-1 LD2 F mid, give q_out = (4 q2 - q1) / 3.  The full step's spectrum is
// kbuf, the half steps' kb2, and q1 waits in rq between launches.  Per contour step:
//   RQ_START (first step only): kbuf <- F LW1 q_s, kb2 <- F LW2 q_s
//   pencils: LD1 on kbuf, LD2 on kb2
//   RQ_MID:  rq <- LW1 F^-1 kbuf;  kb2 <- F (LW1 F^-1 kb2)
//   pencil:  LD2 on kb2
//   RQ_END:  q_out <- (4 LW2 F^-1 kb2 - rq) / 3, then for the next step
//            kbuf <- F LW1' q_out, kb2 <- F LW2' q_out
// Two xy planes of LDS (S, S2), or one (lanes with rqm4 == 2, planes too large for two):
// the two transforms of a mode then run one after the other, re-reading the mode's input
// (q_s, or q_out just written by the same threads) from global memory.
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_slab_rqm4(const Lane<T>* __restrict__ lanes, int nz,
                                                   int t) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lane = blockIdx.x / nz;
    const int z = blockIdx.x - lane * nz;
    const Lane<T>* L = lanes + lane;
    if (t >= L->nsteps) return;
    const int nx = L->nx, ny = L->ny, M = nx >> 1, ld = M + 1;
    const int plane = nx * ny, n = plane * nz, np = plane >> 1;
    const int nxh = M + 1, kplane = nxh * ny;
    cpx<T>* S = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* S2 = S + ny * ld;
    cpx<T>* kb = L->kbuf + (size_t)z * kplane;
    cpx<T>* kb2 = L->kb2 + (size_t)z * kplane;
    cpx<T>* rq = reinterpret_cast<cpx<T>*>(L->rq + (size_t)z * plane);
    const T scale = T(2.0 / (double)n);
    int s_in;
    const int b = step_block(L, t, &s_in);
    const cpx<T>* lw1 = reinterpret_cast<const cpx<T>*>(L->lw[b] + (size_t)z * plane);
    const cpx<T>* lw2 = reinterpret_cast<const cpx<T>*>(L->lw2[b] + (size_t)z * plane);
    auto spec_in = [&](cpx<T>* dst, const cpx<T>* src) {
        for (int i = threadIdx.x; i < kplane; i += blockDim.x) {
            const int y = i / nxh;
            dst[y * ld + (i - y * nxh)] = src[i];
        }
    };
    auto spec_out = [&](cpx<T>* dst, const cpx<T>* src) {
        for (int i = threadIdx.x; i < kplane; i += blockDim.x) {
            const int y = i / nxh;
            dst[i] = src[y * ld + (i - y * nxh)];
        }
    };
    // packed real pairs i of the plane <-> LDS offset o
    auto pairs = [&](auto&& f) {
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const int y = i / M;
            f(i, y * ld + (i - y * M));
        }
    };
    auto both_forward = [&] {
        __syncthreads();
        plane_forward(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
        plane_forward(S2, M, ny, ld, L->px, L->py, L->twx, L->twy);
        spec_out(kb, S);
        spec_out(kb2, S2);
    };
    if (L->rqm4 == 2) {
        // one plane of LDS: the first transform's spectrum is out before the second starts
        // (src: the mode's input plane in global memory, multiplied by LW1 then LW2)
        auto fwd_pair = [&](const cpx<T>* src, const cpx<T>* w1, const cpx<T>* w2) {
            pairs([&](int i, int o) {
                const cpx<T> v = src[i], w = w1[i];
                S[o] = {v.x * w.x, v.y * w.y};
            });
            __syncthreads();
            plane_forward(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
            spec_out(kb, S);
            __syncthreads();
            pairs([&](int i, int o) {
                const cpx<T> v = src[i], w = w2[i];
                S[o] = {v.x * w.x, v.y * w.y};
            });
            __syncthreads();
            plane_forward(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
            spec_out(kb2, S);
        };
        if constexpr (MODE == RQ_START) {
            const cpx<T>* q = reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * n +
                                                              (size_t)z * plane);
            fwd_pair(q, lw1, lw2);
        } else if constexpr (MODE == RQ_MID) {
            spec_in(S, kb);
            __syncthreads();
            plane_inverse(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
            pairs([&](int i, int o) {
                const cpx<T> w = lw1[i], v = S[o];
                rq[i] = {(v.x * scale) * w.x, (v.y * scale) * w.y};
            });
            __syncthreads();
            spec_in(S, kb2);
            __syncthreads();
            plane_inverse(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
            pairs([&](int i, int o) {
                const cpx<T> w = lw1[i], m = S[o];
                S[o] = {(m.x * scale) * w.x, (m.y * scale) * w.y};
            });
            __syncthreads();
            plane_forward(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
            spec_out(kb2, S);
        } else {
            spec_in(S, kb2);
            __syncthreads();
            plane_inverse(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
            cpx<T>* qout = reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * n +
                                                     (size_t)z * plane);
            pairs([&](int i, int o) {
                const cpx<T> w = lw2[i], v = S[o], q1 = rq[i];
                const cpx<T> q2 = {(v.x * scale) * w.x, (v.y * scale) * w.y};
                qout[i] = {(T(4) * q2.x - q1.x) / T(3), (T(4) * q2.y - q1.y) / T(3)};
            });
            if (t + 1 >= L->nsteps) return;
            int s2;
            const int b2 = step_block(L, t + 1, &s2);
            const cpx<T>* lw1n = reinterpret_cast<const cpx<T>*>(L->lw[b2] + (size_t)z * plane);
            const cpx<T>* lw2n = reinterpret_cast<const cpx<T>*>(L->lw2[b2] + (size_t)z * plane);
            __syncthreads();
            fwd_pair(qout, lw1n, lw2n);
        }
        return;
    }
    if constexpr (MODE == RQ_START) {
        const cpx<T>* q = reinterpret_cast<const cpx<T>*>(L->stack + (size_t)(s_in - 1) * n +
                                                          (size_t)z * plane);
        pairs([&](int i, int o) {
            const cpx<T> v = q[i], w = lw1[i], w2 = lw2[i];
            S[o] = {v.x * w.x, v.y * w.y};
            S2[o] = {v.x * w2.x, v.y * w2.y};
        });
        both_forward();
    } else if constexpr (MODE == RQ_MID) {
        spec_in(S, kb);
        spec_in(S2, kb2);
        __syncthreads();
        plane_inverse(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
        plane_inverse(S2, M, ny, ld, L->px, L->py, L->twx, L->twy);
        pairs([&](int i, int o) {
            const cpx<T> w = lw1[i], v = S[o], m = S2[o];
            rq[i] = {(v.x * scale) * w.x, (v.y * scale) * w.y};
            S2[o] = {(m.x * scale) * w.x, (m.y * scale) * w.y};
        });
        __syncthreads();
        plane_forward(S2, M, ny, ld, L->px, L->py, L->twx, L->twy);
        spec_out(kb2, S2);
    } else {
        spec_in(S2, kb2);
        __syncthreads();
        plane_inverse(S2, M, ny, ld, L->px, L->py, L->twx, L->twy);
        cpx<T>* qout = reinterpret_cast<cpx<T>*>(L->stack + (size_t)(s_in + L->dir - 1) * n +
                                                 (size_t)z * plane);
        pairs([&](int i, int o) {
            const cpx<T> w = lw2[i], v = S2[o], q1 = rq[i];
            const cpx<T> q2 = {(v.x * scale) * w.x, (v.y * scale) * w.y};
            const cpx<T> r = {(T(4) * q2.x - q1.x) / T(3), (T(4) * q2.y - q1.y) / T(3)};
            qout[i] = r;
            S[o] = r;
        });
        if (t + 1 >= L->nsteps) return;
        int s2;
        const int b2 = step_block(L, t + 1, &s2);
        const cpx<T>* lw1n = reinterpret_cast<const cpx<T>*>(L->lw[b2] + (size_t)z * plane);
        const cpx<T>* lw2n = reinterpret_cast<const cpx<T>*>(L->lw2[b2] + (size_t)z * plane);
        __syncthreads();
        pairs([&](int i, int o) {
            const cpx<T> v = S[o], w = lw1n[i], w2 = lw2n[i];
            S[o] = {v.x * w.x, v.y * w.y};
            S2[o] = {v.x * w2.x, v.y * w2.y};
        });
        both_forward();
    }
}

// 3-D lanes, z-pencil pass: forward z, * LD_b(t), inverse z, for `tile` kx columns of
// one ky row.  half: the RQM4 half-step spectrum kb2, LD at half the contour step.
template <typename T>
__global__ void __launch_bounds__(256) k_pencil(const Lane<T>* __restrict__ lanes, int ny,
                                                int ntiles, int tile, int t, int half) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* P = reinterpret_cast<cpx<T>*>(smem_raw);
    const int per = ny * ntiles;
    const int lane = blockIdx.x / per;
    const int r = blockIdx.x - lane * per;
    const int ky = r / ntiles, tl = r - ky * ntiles;
    const Lane<T>* L = lanes + lane;
    if (t >= L->nsteps) return;
    const int nz = L->nz, nxh = (L->nx >> 1) + 1;
    const int kx0 = tl * tile;
    const int tw = min(tile, nxh - kx0);
    int s_in;
    const int b = step_block(L, t, &s_in);
    cpx<T>* kb = (half ? L->kb2 : L->kbuf) + (size_t)ky * nxh + kx0;
    const size_t zstride = (size_t)ny * nxh;
    for (int i = threadIdx.x; i < nz * tw; i += blockDim.x) {
        const int z = i / tw, c = i - z * tw;
        P[z * tile + c] = kb[z * zstride + c];
    }
    __syncthreads();
    lds_fft<T, false>(P, L->pz, tw, tile, 1, L->twz, 1);
    const double mc = half ? -L->ldc[b] / 2.0 : -L->ldc[b];
    const double jy = (double)fold(ky, ny);
    for (int i = threadIdx.x; i < nz * tw; i += blockDim.x) {
        const int z = i / tw, c = i - z * tw;
        const T f = exp_t<T>(mc * ksq_at(L->G, (double)(kx0 + c), jy, (double)fold(z, nz)));
        cpx<T>& v = P[z * tile + c];
        v.x *= f;
        v.y *= f;
    }
    __syncthreads();
    lds_fft<T, true>(P, L->pz, tw, tile, 1, L->twz, 1);
    for (int i = threadIdx.x; i < nz * tw; i += blockDim.x) {
        const int z = i / tw, c = i - z * tw;
        kb[z * zstride + c] = P[z * tile + c];
    }
}

// ---------------------------------------------------------------------------------
// Stress: sum_k c_k Re(qc^_k) Re(q^_k) dksq_theta(k), c_0 = 1 and every other half-
// spectrum entry doubled (src/species/multiblock.jl:276-282).

template <typename T>
__device__ void stress_accumulate(const cpx<T>* Sq, const cpx<T>* Sc, int ldq, int nkx,
                                  int nrows, int ny, int nz, int row0, int kz_fixed,
                                  int kx0, const StressLane<T>* L, double* acc) {
    // rows index ky (planar) or z (pencil, with kx0 offset)
    const int np = L->nparams;
    for (int i = threadIdx.x; i < nrows * nkx; i += blockDim.x) {
        const int rr = i / nkx, c = i - rr * nkx;
        int kx, ky, kz;
        if (kz_fixed >= 0) {
            kx = c; ky = rr; kz = kz_fixed;
        } else {
            kx = kx0 + c; ky = row0; kz = rr;
        }
        const size_t flat = ((size_t)kz * ny + ky) * (size_t)((L->nx >> 1) + 1) + kx;
        const double wgt = (flat == 0) ? 1.0 : 2.0;
        const double a = (double)Sc[rr * ldq + c].x;
        const double bq = (double)Sq[rr * ldq + c].x;
        const double jx = kx, jy = fold(ky, ny), jz = fold(kz, nz);
        for (int k = 0; k < np; ++k)
            acc[k] += (wgt * a) * (bq * dksq_at(L->dGG, np, k, L->dim, jx, jy, jz));
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_stress_planar(const StressLane<T>* __restrict__ lanes,
                                                       double* partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const StressLane<T>* L = lanes + blockIdx.x;
    const int nx = L->nx, ny = L->ny, M = nx >> 1, ld = M + 1, np2 = (nx * ny) >> 1;
    cpx<T>* Sq = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* Sc = Sq + ny * ld;
    const cpx<T>* q = reinterpret_cast<const cpx<T>*>(L->q);
    const cpx<T>* qc = reinterpret_cast<const cpx<T>*>(L->qc);
    for (int i = threadIdx.x; i < np2; i += blockDim.x) {
        const int y = i / M, o = y * ld + (i - y * M);
        Sq[o] = q[i];
        Sc[o] = qc[i];
    }
    __syncthreads();
    plane_forward(Sq, M, ny, ld, L->px, L->py, L->twx, L->twy);
    plane_forward(Sc, M, ny, ld, L->px, L->py, L->twx, L->twy);
    double acc[PF_MAXPAR] = {0, 0, 0, 0, 0, 0};
    stress_accumulate(Sq, Sc, ld, M + 1, ny, ny, 1, 0, 0, 0, L, acc);
    double out[PF_MAXPAR];
    block_sum<PF_MAXPAR>(acc, out);
    if (threadIdx.x == 0)
        for (int k = 0; k < L->nparams; ++k) partial[(size_t)L->out * PF_MAXPAR + k] = out[k];
}

// 3-D stress, slab part: plain r2c of q_s and q+_s planes (no LW) into kq / kqc.
template <typename T>
__global__ void __launch_bounds__(256) k_stress_slab(const StressLane<T>* __restrict__ lanes,
                                                     int nz) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* S = reinterpret_cast<cpx<T>*>(smem_raw);
    const int lane = blockIdx.x / (2 * nz);
    const int r = blockIdx.x - lane * 2 * nz;
    const int which = r / nz, z = r - which * nz;
    const StressLane<T>* L = lanes + lane;
    const int nx = L->nx, ny = L->ny, M = nx >> 1, ld = M + 1, plane = nx * ny, np2 = plane >> 1;
    const int nxh = M + 1, kplane = nxh * ny;
    const cpx<T>* src = reinterpret_cast<const cpx<T>*>((which ? L->qc : L->q) + (size_t)z * plane);
    cpx<T>* dst = (which ? L->kqc : L->kq) + (size_t)z * kplane;
    for (int i = threadIdx.x; i < np2; i += blockDim.x) {
        const int y = i / M;
        S[y * ld + (i - y * M)] = src[i];
    }
    __syncthreads();
    plane_forward(S, M, ny, ld, L->px, L->py, L->twx, L->twy);
    for (int i = threadIdx.x; i < kplane; i += blockDim.x) {
        const int y = i / nxh;
        dst[i] = S[y * ld + (i - y * nxh)];
    }
}

// 3-D stress, pencil part: z-forward of both spectra and the weighted reduction.
template <typename T>
__global__ void __launch_bounds__(256) k_stress_pencil(const StressLane<T>* __restrict__ lanes,
                                                       int ny, int ntiles, int tile,
                                                       double* partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int per = ny * ntiles;
    const int lane = blockIdx.x / per;
    const int r = blockIdx.x - lane * per;
    const int ky = r / ntiles, tl = r - ky * ntiles;
    const StressLane<T>* L = lanes + lane;
    const int nz = L->nz, nxh = (L->nx >> 1) + 1;
    cpx<T>* Pq = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* Pc = Pq + nz * tile;
    const int kx0 = tl * tile;
    const int tw = min(tile, nxh - kx0);
    const size_t zstride = (size_t)ny * nxh;
    const cpx<T>* a = L->kq + (size_t)ky * nxh + kx0;
    const cpx<T>* c = L->kqc + (size_t)ky * nxh + kx0;
    for (int i = threadIdx.x; i < nz * tw; i += blockDim.x) {
        const int z = i / tw, cc = i - z * tw;
        Pq[z * tile + cc] = a[z * zstride + cc];
        Pc[z * tile + cc] = c[z * zstride + cc];
    }
    __syncthreads();
    lds_fft<T, false>(Pq, L->pz, tw, tile, 1, L->twz, 1);
    lds_fft<T, false>(Pc, L->pz, tw, tile, 1, L->twz, 1);
    double acc[PF_MAXPAR] = {0, 0, 0, 0, 0, 0};
    stress_accumulate(Pq, Pc, tile, tw, nz, ny, nz, ky, -1, kx0, L, acc);
    double out[PF_MAXPAR];
    block_sum<PF_MAXPAR>(acc, out);
    if (threadIdx.x == 0)
        for (int k = 0; k < L->nparams; ++k)
            partial[((size_t)L->out * per + r) * PF_MAXPAR + k] = out[k];
}

// ---------------------------------------------------------------------------------
template <typename Kern>
static void set_lds(Kern k, int bytes) {
    if (bytes > 64 * 1024)
        PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
}

template <typename T>
void launch_planar_steps(const Lane<T>* d_lanes, int nlanes, int t0, int t1, int lds_bytes,
                         int threads, hipStream_t st) {
    set_lds(k_planar_steps<T>, lds_bytes);
    hipLaunchKernelGGL(k_planar_steps<T>, dim3(nlanes), dim3(threads), lds_bytes, st, d_lanes,
                       t0, t1);
    PF_HIP(hipGetLastError());
}

template <typename T>
void launch_slab(const Lane<T>* d_lanes, int nlanes, int nz, int mode, int t, int lds_bytes,
                 int threads, hipStream_t st) {
    if (mode == SLAB_FWD) {
        set_lds(k_slab<T, SLAB_FWD>, lds_bytes);
        hipLaunchKernelGGL((k_slab<T, SLAB_FWD>), dim3(nlanes * nz), dim3(threads), lds_bytes,
                           st, d_lanes, nz, t);
    } else {
        set_lds(k_slab<T, SLAB_INVFWD>, lds_bytes);
        hipLaunchKernelGGL((k_slab<T, SLAB_INVFWD>), dim3(nlanes * nz), dim3(threads),
                           lds_bytes, st, d_lanes, nz, t);
    }
    PF_HIP(hipGetLastError());
}

template <typename T>
void launch_pencil(const Lane<T>* d_lanes, int nlanes, int ny, int ntiles, int tile, int t,
                   int lds_bytes, int threads, hipStream_t st, int half) {
    set_lds(k_pencil<T>, lds_bytes);
    hipLaunchKernelGGL(k_pencil<T>, dim3(nlanes * ny * ntiles), dim3(threads), lds_bytes, st,
                       d_lanes, ny, ntiles, tile, t, half);
    PF_HIP(hipGetLastError());
}

template <typename T>
void launch_slab_rqm4(const Lane<T>* d_lanes, int nlanes, int nz, int mode, int t,
                      int lds_bytes, hipStream_t st) {
    auto go = [&](auto kern) {
        set_lds(kern, lds_bytes);
        hipLaunchKernelGGL(kern, dim3(nlanes * nz), dim3(256), lds_bytes, st, d_lanes, nz, t);
    };
    if (mode == RQ_START)
        go(k_slab_rqm4<T, RQ_START>);
    else if (mode == RQ_MID)
        go(k_slab_rqm4<T, RQ_MID>);
    else
        go(k_slab_rqm4<T, RQ_END>);
    PF_HIP(hipGetLastError());
}

template <typename T>
void launch_stress_planar(const StressLane<T>* d_lanes, int nlanes, double* partial,
                          int lds_bytes, int threads, hipStream_t st) {
    set_lds(k_stress_planar<T>, lds_bytes);
    hipLaunchKernelGGL(k_stress_planar<T>, dim3(nlanes), dim3(threads), lds_bytes, st, d_lanes,
                       partial);
    PF_HIP(hipGetLastError());
}

template <typename T>
void launch_stress_slab(const StressLane<T>* d_lanes, int nlanes, int nz, int lds_bytes,
                        int threads, hipStream_t st) {
    set_lds(k_stress_slab<T>, lds_bytes);
    hipLaunchKernelGGL(k_stress_slab<T>, dim3(nlanes * 2 * nz), dim3(threads), lds_bytes, st,
                       d_lanes, nz);
    PF_HIP(hipGetLastError());
}

template <typename T>
void launch_stress_pencil(const StressLane<T>* d_lanes, int nlanes, int ny, int ntiles,
                          int tile, double* partial, int lds_bytes, int threads,
                          hipStream_t st) {
    set_lds(k_stress_pencil<T>, lds_bytes);
    hipLaunchKernelGGL(k_stress_pencil<T>, dim3(nlanes * ny * ntiles), dim3(threads), lds_bytes,
                       st, d_lanes, ny, ntiles, tile, partial);
    PF_HIP(hipGetLastError());
}

#define PF_INST(T)                                                                          \
    template void launch_planar_steps<T>(const Lane<T>*, int, int, int, int, int, hipStream_t); \
    template void launch_slab<T>(const Lane<T>*, int, int, int, int, int, int, hipStream_t);  \
    template void launch_pencil<T>(const Lane<T>*, int, int, int, int, int, int, int,         \
                                   hipStream_t, int);                                       \
    template void launch_slab_rqm4<T>(const Lane<T>*, int, int, int, int, int, hipStream_t);   \
    template void launch_stress_planar<T>(const StressLane<T>*, int, double*, int, int,       \
                                          hipStream_t);                                     \
    template void launch_stress_slab<T>(const StressLane<T>*, int, int, int, int, hipStream_t); \
    template void launch_stress_pencil<T>(const StressLane<T>*, int, int, int, int, double*,  \
                                          int, int, hipStream_t);
PF_INST(float)
PF_INST(double)

}  // namespace pf
