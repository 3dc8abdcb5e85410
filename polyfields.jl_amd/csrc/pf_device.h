// Device helpers shared by the propagator kernels.
#pragma once

#include "pf_internal.h"

namespace pf {

// Pointers loaded from a Lane struct are generic to the compiler, which then emits flat
// memory instructions: they count against the LDS wait counter too, and a flat store must
// complete before any later LDS access.  Every buffer here is device memory, so the hot
// kernels access complex arrays through global-address-space views (element reads and
// writes go through the scalar members).
template <typename C>
struct GRef {
    __attribute__((address_space(1))) C* p;
    __device__ __forceinline__ operator cpx<decltype(+p->x)>() const { return {p->x, p->y}; }
    template <typename V>
    __device__ __forceinline__ const GRef& operator=(const V& v) const {
        p->x = v.x;
        p->y = v.y;
        return *this;
    }
};
template <typename C>
struct GArr {
    __attribute__((address_space(1))) C* p;
    __device__ __forceinline__ GRef<C> operator[](size_t i) const { return {p + i}; }
};
// real arrays: element reads of a global-address-space pointer
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) const T* gps(const T* p) {
    return (__attribute__((address_space(1))) const T*)p;
}
template <typename U>
__device__ __forceinline__ GArr<cpx<U>> gp(cpx<U>* p) {
    return {(__attribute__((address_space(1))) cpx<U>*)p};
}
template <typename U>
__device__ __forceinline__ GArr<const cpx<U>> gp(const cpx<U>* p) {
    return {(__attribute__((address_space(1))) const cpx<U>*)p};
}

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.x, -a.y}; }
// multiply by +i (INV) or -i (forward)
template <typename T, bool INV>
__device__ __forceinline__ cpx<T> mul_i(cpx<T> a) {
    return INV ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
}

template <typename T>
__device__ __forceinline__ T exp_t(double a);
template <>
__device__ __forceinline__ double exp_t<double>(double a) { return exp(a); }
// FP32 sweeps: the hardware 2^x (v_exp_f32, about 1 ulp) of a log2(e), rounded once to
// float, as expf would round its argument
template <>
__device__ __forceinline__ float exp_t<float>(double a) {
    return __builtin_amdgcn_exp2f((float)(a * 1.4426950408889634074));
}

// cos(2 pi m / 16) and cos(2 pi m / 12); sin(2 pi m / N) = cos(2 pi (m - N/4) / N)
constexpr double kCos16[16] = {
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
constexpr double kCos12[12] = {
    1.0, 0.86602540378443864676, 0.5, 0.0, -0.5, -0.86602540378443864676,
    -1.0, -0.86602540378443864676, -0.5, 0.0, 0.5, 0.86602540378443864676};

template <int N>
constexpr double cos_frac(int m) {
    static_assert(N == 16 || N == 12, "twiddle table");
    m = ((m % N) + N) % N;
    return N == 16 ? kCos16[m] : kCos12[m];
}

// v * w_N^m, w_N = exp(-/+ 2 pi i / N) (forward / inverse); m a constant after unrolling.
// The quarter turns are exact sign swaps.
template <typename T, int N, bool INV>
__device__ __forceinline__ cpx<T> twmul(cpx<T> v, int m) {
    if constexpr (N == 6) {
        return twmul<T, 12, INV>(v, 2 * m);   // w_6^m = w_12^(2m)
    } else {
        m %= N;
        if (m == 0) return v;
        if (2 * m == N) return {-v.x, -v.y};
        if (4 * m == N) return mul_i<T, INV>(v);
        if (4 * m == 3 * N) return mul_i<T, !INV>(v);
        const T c = T(cos_frac<N>(m)), sn = T(cos_frac<N>(m - N / 4));
        const cpx<T> w = {c, INV ? sn : -sn};
        return cmul(v, w);
    }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(cpx<T>* v);

// Length N1 * N2 from N2 length-N1 and N1 length-N2 DFTs (Cooley-Tukey): input n = N2 n1 +
// n2, output k = k1 + N1 k2, twiddles w_N^(n2 k1) between the stages.
template <typename T, int N1, int N2, bool INV>
__device__ __forceinline__ void dft_ct(cpx<T>* v) {
    constexpr int N = N1 * N2;
    cpx<T> y[N2][N1];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) y[n2][n1] = v[N2 * n1 + n2];
        dft<T, N1, INV>(y[n2]);
#pragma unroll
        for (int k1 = 1; k1 < N1; ++k1) y[n2][k1] = twmul<T, N, INV>(y[n2][k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        cpx<T> z[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) z[n2] = y[n2][k1];
        dft<T, N2, INV>(z);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) v[k1 + N1 * k2] = z[k2];
    }
}

// Small DFTs of length R in registers.  Forward uses exp(-2 pi i/R), inverse exp(+...).
template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(cpx<T>* v) {
    if constexpr (R == 2) {
        cpx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const T c = T(-0.5);
        const T s = T(0.86602540378443864676372317075294);
        cpx<T> t1 = cadd(v[1], v[2]);
        cpx<T> d = csub(v[1], v[2]);
        cpx<T> m = {v[0].x + c * t1.x, v[0].y + c * t1.y};
        cpx<T> sd = mul_i<T, INV>(cpx<T>{s * d.x, s * d.y});
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, sd);
        v[2] = csub(m, sd);
    } else if constexpr (R == 4) {
        cpx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        cpx<T> b0 = cadd(v[1], v[3]), b1 = mul_i<T, INV>(csub(v[1], v[3]));
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    } else if constexpr (R == 8) {
        // two length-4 DFTs on the even and odd samples, combined with w8^k
        cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
        dft<T, 4, INV>(e);
        dft<T, 4, INV>(o);
        const T h = T(0.70710678118654752440084436210485);
        // w8^1 = h(1 -/+ i), w8^2 = -/+ i, w8^3 = h(-1 -/+ i)   (forward / inverse)
        const cpx<T> o1 = INV ? cpx<T>{h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)}
                              : cpx<T>{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
        const cpx<T> o2 = mul_i<T, INV>(o[2]);
        const cpx<T> o3 = INV ? cpx<T>{-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)}
                              : cpx<T>{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    } else if constexpr (R == 6) {
        dft_ct<T, 3, 2, INV>(v);
    } else if constexpr (R == 12) {
        dft_ct<T, 4, 3, INV>(v);
    } else if constexpr (R == 16) {
        dft_ct<T, 4, 4, INV>(v);
    } else {
        static_assert(R == 2, "radix");
    }
}

__device__ __forceinline__ int fold(int i, int n) { return (2 * i <= n) ? i : n - i; }

__device__ __forceinline__ double ksq_at(const double* G, double jx, double jy, double jz) {
    const double kx = G[0] * jx + G[1] * jy + G[2] * jz;
    const double ky = G[3] * jx + G[4] * jy + G[5] * jz;
    const double kz = G[6] * jx + G[7] * jy + G[8] * jz;
    return kx * kx + ky * ky + kz * kz;
}

// Block b of step t and the slice s it reads.  q (dir > 0) reads slice s = 1 + t and is in
// the last block with begin[b] <= s; q+ reads s = Ns + 1 - t and is in the last block with
// begin[b] < s.  begin[] increases, so b counts the k >= 1 meeting that test: an unrolled
// count over PF_MAXBLK entries, whose loads all go out at once (a loop over begin[] would
// make each load wait for the previous compare, a chain of dependent scalar loads at the
// start of every launch).
// Momentum family update of point i (src/updaters/momentum.jl:86-139); `method`: 0 MOM,
// 1 RMS, 2 ADAM
__device__ __forceinline__ void momentum_point(int method, double* mt, double* vt, double* step,
                                               const double* __restrict__ r, double b1,
                                               double b2, size_t i) {
    const double ri = r[i];
    if (method == 0 || method == 2) mt[i] = b1 * mt[i] + (1 - b1) * ri;
    if (method == 1 || method == 2) vt[i] = b2 * vt[i] + (1 - b2) * (ri * ri);
    if (method == 0)
        step[i] = mt[i];
    else if (method == 1)
        step[i] = ri / (sqrt(vt[i]) + 1e-8);
    else
        step[i] = mt[i] / (sqrt(vt[i]) + 1e-8);
}

template <typename T>
__device__ __forceinline__ int step_block(const Lane<T>* L, int t, int* s_in) {
    // every field is read unconditionally (a short-circuit test would make the loads
    // conditional, hence sequential)
    const int dir = L->dir, ns = L->nsteps, nb = L->nblk;
    int bg[PF_MAXBLK];
#pragma unroll
    for (int k = 1; k < PF_MAXBLK; ++k) bg[k] = L->begin[k];
    const bool fwd = dir > 0;
    const int s = fwd ? 1 + t : ns + 1 - t;
    *s_in = s;
    const int lim = fwd ? s : s - 1;
    int b = 0;
#pragma unroll
    for (int k = 1; k < PF_MAXBLK; ++k) b += (int)(k < nb) & (int)(bg[k] <= lim);
    return b;
}

// a[b] of a per-block lane table with every entry loaded up front and b selecting among
// them (no load that waits for b)
template <typename P>
__device__ __forceinline__ P pick_blk(const P (&a)[PF_MAXBLK], int b) {
    static_assert(sizeof(P) == 8, "per-block entries of 8 bytes");
    static_assert(PF_MAXBLK == 8, "one operand per block below");
    unsigned long long e[PF_MAXBLK];
#pragma unroll
    for (int k = 0; k < PF_MAXBLK; ++k) e[k] = __builtin_bit_cast(unsigned long long, a[k]);
    // all eight loads go out together; the empty asm keeps the select below from becoming
    // a dependent load of a[b]
    asm volatile("" : "+s"(e[0]), "+s"(e[1]), "+s"(e[2]), "+s"(e[3]), "+s"(e[4]), "+s"(e[5]),
                 "+s"(e[6]), "+s"(e[7]));
    unsigned long long r = e[0];
#pragma unroll
    for (int k = 1; k < PF_MAXBLK; ++k) r = (b == k) ? e[k] : r;
    return __builtin_bit_cast(P, r);
}

__device__ __forceinline__ double dksq_at(const double* dGG, int np, int k, int dim, double j0,
                                          double j1, double j2) {
    const double j[3] = {j0, j1, j2};
    double acc = 0.0;
    for (int l = 0; l < dim; ++l)
        for (int m = 0; m < dim; ++m) acc += j[l] * j[m] * dGG[(l * 3 + m) * np + k];
    return acc;
}

template <int NV>
__device__ void block_sum(double* v, double* out) {
    // wavefront reduction then across waves through LDS; result in out[0..NV) of thread 0
    __shared__ double red[PF_MAXPAR * 16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) red[k * 16 + wave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += red[k * 16 + w];
            out[k] = s;
        }
    }
    __syncthreads();
}

}  // namespace pf
