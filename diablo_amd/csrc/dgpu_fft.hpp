// dgpu_fft.hpp -- LDS-staged Stockham FFTs for gfx950 (device side).
//
// A batch of transforms of one length n lives in LDS.  T = n / E threads work
// on each transform; every thread holds E complex values in registers during
// a radix-R stage (E/R butterflies), so a stage is: read E values from LDS,
// barrier, twiddle + butterfly in registers, write E values, barrier.  The
// stages are the Stockham autosort (no bit reversal): at a stage with
// span Ns and radix R, butterfly t reads x[t + r*n/R] and writes
// y[(t/Ns)*Ns*R + t%Ns + r*Ns].  Twiddles come from a per-length table
// tw[m] = exp(-2 pi i m / n) computed on the host in extended precision.
//
// Transforms are unnormalised; SIGN = -1 is the forward (FFTW_FORWARD)
// direction, SIGN = +1 the backward one.
#pragma once

#include <hip/hip_runtime.h>

namespace dgpu {

template <typename T>
struct cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> make_cx(T r, T i) { cx<T> c; c.re = r; c.im = i; return c; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return make_cx(a.re + b.re, a.im + b.im); }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return make_cx(a.re - b.re, a.im - b.im); }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b)
{
    return make_cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__device__ __forceinline__ cx<T> scale(cx<T> a, T s) { return make_cx(a.re * s, a.im * s); }
template <typename T>
__device__ __forceinline__ cx<T> conj_(cx<T> a) { return make_cx(a.re, -a.im); }
// multiply by i
template <typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) { return make_cx(-a.im, a.re); }

// Maximum number of radix stages of a supported length.
constexpr int kMaxStages = 16;

// Describes one transform length.
struct FftLen {
    int n;
    int nstages;
    int radix[kMaxStages];
    unsigned magic[kMaxStages];  // ceil(2^32 / ns) of each stage: t / ns = umulhi(t, magic)
    int e;      // values per thread (multiple of every radix)
    int t;      // threads per transform = n / e
};

// Quotient by a stage span with a precomputed reciprocal; exact for
// t, ns < 2^16 (transform lengths here are far below that).  magic == 0
// stands for a span of 1, whose reciprocal 2^32 does not fit.
__device__ __forceinline__ int div_magic(int t, unsigned magic)
{
    return magic ? (int)__umulhi((unsigned)t, magic) : t;
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void butterfly(cx<T> *v)
{
    if constexpr (R == 2) {
        cx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if constexpr (R == 3) {
        const T c1 = T(-0.5);
        const T s1 = T(SIGN) * T(0.86602540378443864676372317075294);
        cx<T> a = v[0], b = v[1], c = v[2];
        cx<T> s = b + c, d = b - c;
        cx<T> m = make_cx(a.re + c1 * s.re, a.im + c1 * s.im);
        cx<T> r = make_cx(-s1 * d.im, s1 * d.re);  // i * s1 * d
        v[0] = a + s;
        v[1] = m + r;
        v[2] = m - r;
    } else if constexpr (R == 4) {
        cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
        cx<T> b0 = v[1] + v[3], b1 = v[1] - v[3];
        // forward: -i*b1 ; backward: +i*b1
        cx<T> b1r = (SIGN < 0) ? make_cx(b1.im, -b1.re) : make_cx(-b1.im, b1.re);
        v[0] = a0 + b0;
        v[1] = a1 + b1r;
        v[2] = a0 - b0;
        v[3] = a1 - b1r;
    } else if constexpr (R == 5) {
        const T c1 = T(0.30901699437494742410229341718282);   // cos(2pi/5)
        const T c2 = T(-0.80901699437494742410229341718282);  // cos(4pi/5)
        const T s1 = T(SIGN) * T(0.95105651629515357211643933337938);  // sin(2pi/5)
        const T s2 = T(SIGN) * T(0.58778525229247312916870595463907);  // sin(4pi/5)
        cx<T> a = v[0];
        cx<T> p1 = v[1] + v[4], m1 = v[1] - v[4];
        cx<T> p2 = v[2] + v[3], m2 = v[2] - v[3];
        cx<T> t1 = make_cx(a.re + c1 * p1.re + c2 * p2.re, a.im + c1 * p1.im + c2 * p2.im);
        cx<T> t2 = make_cx(a.re + c2 * p1.re + c1 * p2.re, a.im + c2 * p1.im + c1 * p2.im);
        // i*(s1*m1 + s2*m2), i*(s2*m1 - s1*m2)
        cx<T> u1 = make_cx(-(s1 * m1.im + s2 * m2.im), s1 * m1.re + s2 * m2.re);
        cx<T> u2 = make_cx(-(s2 * m1.im - s1 * m2.im), s2 * m1.re - s1 * m2.re);
        v[0] = a + p1 + p2;
        v[1] = t1 + u1;
        v[4] = t1 - u1;
        v[2] = t2 + u2;
        v[3] = t2 - u2;
    }
}

// One Stockham stage of radix R for the transform whose element e sits at
// lds[e * stride].  th is this thread's index inside the transform.
template <typename T, int E, int R, int SIGN>
__device__ __forceinline__ void stage(cx<T> *lds, int stride, int n, int ns, unsigned magic, int th, int nthr,
                                      const cx<T> *__restrict__ tw, bool active)
{
    constexpr int NB = E / R;
    const int L = n / R;
    const int tws = n / (ns * R);
    cx<T> v[E];
    cx<T> w[E];
    if (active) {
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int t = th + nthr * u;
            const int q = div_magic(t, magic);
            const int k = t - q * ns;
#pragma unroll
            for (int r = 0; r < R; ++r) v[u * R + r] = lds[(t + r * L) * stride];
            // twiddle loads depend only on t: issue them before the barrier
#pragma unroll
            for (int r = 1; r < R; ++r) w[u * R + r] = tw[r * k * tws];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int t = th + nthr * u;
            const int q = div_magic(t, magic);
            const int k = t - q * ns;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cx<T> ww = w[u * R + r];
                if (SIGN > 0) ww.im = -ww.im;
                v[u * R + r] = cmul(v[u * R + r], ww);
            }
            butterfly<T, R, SIGN>(v + u * R);
            const int d = q * ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[(d + r * ns) * stride] = v[u * R + r];
        }
    }
    __syncthreads();
}

// Full transform.  Every thread of the block must call this (it contains
// barriers); threads with active == false only take part in the barriers.
template <typename T, int E, int SIGN>
__device__ void fft_lds(cx<T> *lds, int stride, const FftLen &f, int th, bool active,
                        const cx<T> *__restrict__ tw)
{
    int ns = 1;
    for (int s = 0; s < f.nstages; ++s) {
        const int R = f.radix[s];
        const unsigned mg = f.magic[s];
        if constexpr (E % 4 == 0) {
            if (R == 4) { stage<T, E, 4, SIGN>(lds, stride, f.n, ns, mg, th, f.t, tw, active); ns *= 4; continue; }
        }
        if constexpr (E % 2 == 0) {
            if (R == 2) { stage<T, E, 2, SIGN>(lds, stride, f.n, ns, mg, th, f.t, tw, active); ns *= 2; continue; }
        }
        if constexpr (E % 3 == 0) {
            if (R == 3) { stage<T, E, 3, SIGN>(lds, stride, f.n, ns, mg, th, f.t, tw, active); ns *= 3; continue; }
        }
        if constexpr (E % 5 == 0) {
            if (R == 5) { stage<T, E, 5, SIGN>(lds, stride, f.n, ns, mg, th, f.t, tw, active); ns *= 5; continue; }
        }
    }
}

}  // namespace dgpu
