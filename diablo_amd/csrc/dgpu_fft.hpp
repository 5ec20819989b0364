// dgpu_fft.hpp -- LDS-staged FFTs for gfx950 (device side).
//
// Two families share the butterflies below.
//
// In-place mixed-radix transforms (IpLen, ip_*) run the time-stepper's hot
// kernels.  A length n = R_0 R_1 ... R_{S-1} transform has S stages; stage i
// works on blocks of m_i = R_i R_{i+1} ... elements (m_0 = n) and its groups
// (block b, offset k < L_i = m_i / R_i) touch the R_i positions
// b m_i + k + q L_i.  Decimation in frequency (DIF: butterfly, then twiddle
// w_{m_i}^{kq}) takes natural-order input to output in mixed-radix
// digit-reversed positions; decimation in time (DIT: twiddle, then butterfly,
// stages in reverse order) is its transpose and takes that permuted order back
// to natural order.  A thread reads and writes the same positions, so a stage
// needs one barrier, not two, and the physical-space side of a transform pair
// stays in permuted order (the products there are pointwise).
//
// Stockham transforms (FftLen, fft_lds) serve the standalone
// FFT_XZ_TO_FOURIER / FFT_XZ_TO_PHYSICAL entry points, whose physical-space
// side is in natural order.
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

// Aligned to its size, so a value moves in one 16-byte (double) or 8-byte
// (float) access to LDS and global memory rather than two halves.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T re, im;
};

#define DGPU_HD __host__ __device__ __forceinline__

template <typename T>
DGPU_HD cx<T> make_cx(T r, T i) { cx<T> c; c.re = r; c.im = i; return c; }
template <typename T>
DGPU_HD cx<T> operator+(cx<T> a, cx<T> b) { return make_cx(a.re + b.re, a.im + b.im); }
template <typename T>
DGPU_HD cx<T> operator-(cx<T> a, cx<T> b) { return make_cx(a.re - b.re, a.im - b.im); }
template <typename T>
DGPU_HD cx<T> cmul(cx<T> a, cx<T> b)
{
    return make_cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
DGPU_HD cx<T> scale(cx<T> a, T s) { return make_cx(a.re * s, a.im * s); }
template <typename T>
DGPU_HD cx<T> conj_(cx<T> a) { return make_cx(a.re, -a.im); }
// multiply by i
template <typename T>
DGPU_HD cx<T> mul_i(cx<T> a) { return make_cx(-a.im, a.re); }

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
DGPU_HD int div_magic(int t, unsigned magic)
{
    return magic ? (int)(((unsigned long long)(unsigned)t * magic) >> 32) : t;
}

template <typename T, int R, int SIGN>
DGPU_HD void butterfly(cx<T> *v)
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
    } else if constexpr (R == 6) {
        // 6 = 2 x 3: DFT3 of the even and odd samples, then one radix-2 step
        // with the twiddles w6^k, k = 0..2 (w6 = exp(SIGN 2 pi i / 6)).
        cx<T> e[3] = {v[0], v[2], v[4]}, o[3] = {v[1], v[3], v[5]};
        butterfly<T, 3, SIGN>(e);
        butterfly<T, 3, SIGN>(o);
        const T h = T(0.5);
        const T s = T(SIGN) * T(0.86602540378443864676372317075294);
        const cx<T> o1 = make_cx(h * o[1].re - s * o[1].im, h * o[1].im + s * o[1].re);     // o1 * w6
        const cx<T> o2 = make_cx(-h * o[2].re - s * o[2].im, -h * o[2].im + s * o[2].re);   // o2 * w6^2
        v[0] = e[0] + o[0];
        v[3] = e[0] - o[0];
        v[1] = e[1] + o1;
        v[4] = e[1] - o1;
        v[2] = e[2] + o2;
        v[5] = e[2] - o2;
    } else if constexpr (R == 8) {
        // 8 = 2 x 4: DFT4 of the even and odd samples, then one radix-2 step
        // with w8^k, k = 0..3 (w8 = exp(SIGN 2 pi i / 8)).
        cx<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        butterfly<T, 4, SIGN>(e);
        butterfly<T, 4, SIGN>(o);
        const T r = T(0.70710678118654752440084436210485);
        const T sg = T(SIGN);
        // o1 * (r + i sg r), o2 * (i sg), o3 * (-r + i sg r)
        const cx<T> o1 = make_cx(r * (o[1].re - sg * o[1].im), r * (o[1].im + sg * o[1].re));
        const cx<T> o2 = make_cx(-sg * o[2].im, sg * o[2].re);
        const cx<T> o3 = make_cx(-r * (o[3].re + sg * o[3].im), r * (sg * o[3].re - o[3].im));
        v[0] = e[0] + o[0];
        v[4] = e[0] - o[0];
        v[1] = e[1] + o1;
        v[5] = e[1] - o1;
        v[2] = e[2] + o2;
        v[6] = e[2] - o2;
        v[3] = e[3] + o3;
        v[7] = e[3] - o3;
    }
}

// ---------------------------------------------------------------------------
// In-place mixed-radix transforms
// ---------------------------------------------------------------------------

struct IpLen {
    int n;
    int nstages;
    int radix[kMaxStages];      // R_i, stage 0 works on the whole transform
    int span[kMaxStages];       // L_i = m_i / R_i
    unsigned magic[kMaxStages]; // t / L_i = div_magic(t, magic[i])
    int twoff[kMaxStages];      // stage i's twiddles: tw[twoff[i] + k*(R_i-1) + q-1] = w_{m_i}^{kq}
    int twn;                    // length of the whole twiddle table
    int e;                      // values per thread (multiple of every R_i)
    int t;                      // threads per transform = n / e
    unsigned swz;               // row swizzle: the mask of pos >> swzs XORed into pos (0: none)
    int cswz;                   // column-batch swizzle: bit 0 of a position XORed with bit cswz (0: none)
    int swzs;                   // the row swizzle's shift
};

// Slot of position pos in a column batch (lds[slot * W + column]): swapping
// rows pairwise by a higher bit puts the two positions a 16-lane group reads
// in different halves of the banks.
DGPU_HD int ip_cswz(const IpLen &f, int pos)
{
    return f.cswz ? (pos ^ ((pos >> f.cswz) & 1)) : pos;
}

// LDS slot of position pos of a row: pos XORed with its own higher bits
// (pos >> swzs, masked), which spreads the strided accesses of the late
// stages over the banks.  A map x -> x ^ ((x >> s) & m) is invertible (the
// bits it changes are recovered from the higher ones, top down), and with
// m < 32 it stays inside each 32-aligned block of positions (the host
// checks that it is a bijection of 0..n-1).  Two instructions per access.
DGPU_HD int ip_swz(const IpLen &f, int pos)
{
    return pos ^ ((pos >> f.swzs) & (int)f.swz);
}

// Twiddle + butterfly of one group (values v[0..R-1] at positions
// base + q L_i); tw points at the stage's table, k is the group offset.
template <typename T, int R, int SIGN, bool DIF>
DGPU_HD void ip_group(cx<T> *v, const cx<T> *__restrict__ tw, int k)
{
    cx<T> w[R];
#pragma unroll
    for (int q = 1; q < R; ++q) {
        w[q] = tw[k * (R - 1) + q - 1];
        if (SIGN > 0) w[q].im = -w[q].im;
    }
    if (!DIF) {
#pragma unroll
        for (int q = 1; q < R; ++q) v[q] = cmul(v[q], w[q]);
    }
    butterfly<T, R, SIGN>(v);
    if (DIF) {
#pragma unroll
        for (int q = 1; q < R; ++q) v[q] = cmul(v[q], w[q]);
    }
}

// Groups of stage i handled by thread th: t = th + nthr*u, u < E/R.  fn(u, base, k)
template <int E, int R, class F>
DGPU_HD void ip_groups(const IpLen &f, int i, int th, F &&fn)
{
    constexpr int NB = E / R;
    const int L = f.span[i];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const int t = th + f.t * u;
        const int b = div_magic(t, f.magic[i]);
        const int k = t - b * L;
        fn(u, b * R * L + k, k);
    }
}

// One stage with arbitrary sources and sinks: ld(idx, pos) gives the value at
// position pos (idx = u*R + q is its slot in this thread's E values),
// st(idx, pos, v) takes the result.  All loads come before any store.
template <typename T, int E, int R, int SIGN, bool DIF, class Ld, class St>
DGPU_HD void ip_stage(const IpLen &f, int i, const cx<T> *__restrict__ tw, int th, Ld &&ld, St &&st)
{
    constexpr int NB = E / R;
    const int L = f.span[i];
    cx<T> v[E];
    int base[NB], kk[NB];
    ip_groups<E, R>(f, i, th, [&](int u, int b0, int k) {
        base[u] = b0;
        kk[u] = k;
#pragma unroll
        for (int q = 0; q < R; ++q) v[u * R + q] = ld(u * R + q, b0 + q * L);
    });
    if (L == 1) {
        // the last stage of a DIF (first of a DIT) transform: every group has
        // k = 0, whose twiddles are all 1
#pragma unroll
        for (int u = 0; u < NB; ++u) butterfly<T, R, SIGN>(v + u * R);
    } else {
#pragma unroll
        for (int u = 0; u < NB; ++u) ip_group<T, R, SIGN, DIF>(v + u * R, tw + f.twoff[i], kk[u]);
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
#pragma unroll
        for (int q = 0; q < R; ++q) st(u * R + q, base[u] + q * L, v[u * R + q]);
    }
}

// Runs fn(s) for s = s0..s1-1; fully unrolled when UNROLL (a plan known at
// compile time, whose stage count then is a constant).
template <bool UNROLL, class F>
DGPU_HD void stage_loop(int s0, int s1, F &&fn)
{
    if constexpr (UNROLL) {
#pragma unroll
        for (int s = s0; s < s1; ++s) fn(s);
    } else {
        for (int s = s0; s < s1; ++s) fn(s);
    }
}

// Plans known at compile time.  const_plan(id) builds the same IpLen as the
// host planner (ip_fill after a greedy factorisation, dgpu_plan.hpp) for the
// lengths the kernels are specialised for; with such a plan every span,
// radix and twiddle offset is a constant, divisions by power-of-two spans
// become shifts and the stage loops unroll.  Id 0 is "runtime plan".
struct ConstPlanSpec {
    int n, e, r0;  // r0 > 0: outermost radix fixed (x plans), 0: greedy (z plans)
    unsigned swz;
    int cswz;
    int swzs = 3;
};
constexpr ConstPlanSpec kConstPlanSpecs[] = {
    {0, 0, 0, 0u, 0},
    // x plans (rows of the fused nonlinear kernel); row swizzle mask, shift
    {384, 12, 6, 15u, 0, 2},  // 1: (6,4,4,4), FP64
    {192, 12, 6, 15u, 0, 2},  // 2: (6,4,4,2), FP32
    {96, 12, 6, 14u, 0, 1},   // 3: (6,4,4), FP32
    {128, 8, 4, 15u, 0, 3},   // 4: (4,8,4), FP64
    {64, 16, 8, 12u, 0, 1},   // 5: (8,8), FP32
    // z plans (column batches)
    {384, 12, 0, 0u, 2},           // 6: (6,4,4,4), FP64
    {384, 24, 0, 0u, 0},           // 7: (8,8,6), FP32 (DGPU_FFT_EZ=24)
    {64, 8, 0, 0u, 0},             // 8: (8,8)
    {96, 12, 0, 0u, 0},            // 9: (6,4,4)
    {128, 8, 0, 0u, 0},            // 10: (8,8,2)
    {192, 24, 0, 0u, 0},           // 11: (8,8,3), FP32 (DGPU_FFT_EZ=24)
    // x plans appended after the z plans (their ids stay)
    {256, 8, 4, 15u, 0, 3},   // 12: (4,8,8), FP64
    // FP32 x plans (8-byte slots)
    {256, 8, 4, 31u, 0, 3},   // 13: (4,8,8), FP32
    {128, 8, 4, 31u, 0, 2},   // 14: (4,8,4), FP32
    // z plans at up to 12 values per thread
    {192, 12, 0, 0u, 0},           // 15: (6,4,4,2)
    {256, 8, 0, 0u, 2},            // 16: (8,8,4)
    {384, 12, 6, 31u, 0, 2},       // 17: (6,4,4,4), FP32 x plan
};
constexpr int kNumConstPlans = (int)(sizeof(kConstPlanSpecs) / sizeof(kConstPlanSpecs[0]));

constexpr IpLen const_plan(int id)
{
    IpLen f{};
    if (id <= 0 || id >= kNumConstPlans) return f;
    const ConstPlanSpec c = kConstPlanSpecs[id];
    const int cand[6] = {8, 6, 5, 4, 3, 2};
    int radix[kMaxStages] = {};
    int ns = 0, m = c.n;
    if (c.r0 > 0) {
        radix[ns++] = c.r0;
        m /= c.r0;
    }
    while (m > 1 && ns < kMaxStages) {
        int r = 0;
        for (int q = 0; q < 6 && !r; ++q)
            if (c.e % cand[q] == 0 && m % cand[q] == 0) r = cand[q];
        if (!r) return IpLen{};
        radix[ns++] = r;
        m /= r;
    }
    f.n = c.n;
    f.nstages = ns;
    f.e = c.e;
    f.t = c.n / c.e;
    int mm = c.n, off = 0;
    for (int i = 0; i < ns; ++i) {
        const int R = radix[i];
        f.radix[i] = R;
        f.span[i] = mm / R;
        f.magic[i] = f.span[i] == 1 ? 0u : (unsigned)(((1ULL << 32) + f.span[i] - 1) / f.span[i]);
        f.twoff[i] = off;
        off += f.span[i] * (R - 1);
        mm /= R;
    }
    f.twn = off;
    f.swz = c.swz;
    f.cswz = c.cswz;
    f.swzs = c.swzs;
    return f;
}

// Calls fn(std::integral_constant<int, R>) for the runtime radix R (those
// that divide E; the plan only uses such radices).
template <int R>
struct RadixC {
    static constexpr int value = R;
};
template <int E, class F>
DGPU_HD void with_radix(int R, F &&fn)
{
    switch (R) {
    case 2: if constexpr (E % 2 == 0) fn(RadixC<2>{}); break;
    case 3: if constexpr (E % 3 == 0) fn(RadixC<3>{}); break;
    case 4: if constexpr (E % 4 == 0) fn(RadixC<4>{}); break;
    case 5: if constexpr (E % 5 == 0) fn(RadixC<5>{}); break;
    case 6: if constexpr (E % 6 == 0) fn(RadixC<6>{}); break;
    case 8: if constexpr (E % 8 == 0) fn(RadixC<8>{}); break;
    default: break;
    }
}

// In-place stage i on x[ip_cswz(pos) * stride] (LDS, the z kernels' column
// batches).
template <typename T, int E, int SIGN, bool DIF>
DGPU_HD void ip_stage_mem(cx<T> *x, int stride, const IpLen &f, int i, const cx<T> *__restrict__ tw, int th)
{
    with_radix<E>(f.radix[i], [&](auto rc) {
        constexpr int R = decltype(rc)::value;
        ip_stage<T, E, R, SIGN, DIF>(
            f, i, tw, th, [&](int, int pos) { return x[ip_cswz(f, pos) * stride]; },
            [&](int, int pos, cx<T> v) { x[ip_cswz(f, pos) * stride] = v; });
    });
}

// In-place stage i on one swizzled row x[ip_swz(pos)] (LDS).
template <typename T, int E, int SIGN, bool DIF>
DGPU_HD void ip_stage_row(cx<T> *x, const IpLen &f, int i, const cx<T> *__restrict__ tw, int th)
{
    with_radix<E>(f.radix[i], [&](auto rc) {
        constexpr int R = decltype(rc)::value;
        ip_stage<T, E, R, SIGN, DIF>(
            f, i, tw, th, [&](int, int pos) { return x[ip_swz(f, pos)]; },
            [&](int, int pos, cx<T> v) { x[ip_swz(f, pos)] = v; });
    });
}

// One Stockham stage of radix R for the transform whose element e sits at
// lds[e * stride].  th is this thread's index inside the transform.
// Orders a wavefront's own LDS accesses across this point: a wavefront's LDS
// operations execute in order, and the fences keep the compiler from moving
// them across.  Enough in place of a block barrier when the data exchanged
// stays inside the wavefront (a transform whose threads all lie in it).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// WS: every transform's threads lie inside one wavefront (T divides 64), so
// the stage synchronises with wave_lds_sync instead of block barriers.
template <typename T, int E, int R, int SIGN, bool WS = false>
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
    if constexpr (WS) wave_lds_sync();
    else __syncthreads();
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
    if constexpr (WS) wave_lds_sync();
    else __syncthreads();
}

// Stockham plan of length n: radix-4 stages, then 2s, then 3s.  E (values
// per thread) is the first of 4, 6, 8, 12, 16 that divides n, is a multiple
// of every radix and leaves at most 64 threads a transform, else the largest
// such.  nstages == 0: no plan (n < 4 or n has another prime factor).
// The host plans with it; the kernels also take it at compile time.
constexpr FftLen fft_len_of(int n)
{
    FftLen f{};
    f.n = n;
    if (n < 4) return f;
    int m = n, ns = 0;
    bool has2 = false, has3 = false, has4 = false;
    while (m % 4 == 0 && ns < kMaxStages) { f.radix[ns++] = 4; m /= 4; has4 = true; }
    while (m % 2 == 0 && ns < kMaxStages) { f.radix[ns++] = 2; m /= 2; has2 = true; }
    while (m % 3 == 0 && ns < kMaxStages) { f.radix[ns++] = 3; m /= 3; has3 = true; }
    if (m != 1) return FftLen{};
    int l = has4 ? 4 : (has2 ? 2 : 1);
    if (has3) l = (l == 4) ? 12 : (l == 2 ? 6 : 3);
    const int cand[5] = {4, 6, 8, 12, 16};
    int best = -1;
    for (int e : cand) {
        if (e % l || n % e) continue;
        best = e;
        if (n / e <= 64) break;
    }
    if (best < 0) return FftLen{};
    f.nstages = ns;
    f.e = best;
    f.t = n / best;
    unsigned long long p = 1;
    for (int s = 0; s < ns; ++s) {
        f.magic[s] = (p == 1) ? 0u : (unsigned)(((1ULL << 32) + p - 1) / p);
        p *= (unsigned long long)f.radix[s];
    }
    return f;
}

// Lengths the standalone transforms are also built for (FID > 0).
constexpr int kConstFftN[] = {0, 384, 256};
constexpr int kNumConstFft = (int)(sizeof(kConstFftN) / sizeof(kConstFftN[0]));

// Full transform.  Every thread of the block must call this (it contains
// barriers); threads with active == false only take part in the barriers.
// With WS the transform synchronises within its wavefront only, and the
// caller places a block barrier before other wavefronts read the result.
// With a plan known at compile time the stage loop unrolls and every
// radix, span and reciprocal is a constant.
template <typename T, int E, int SIGN, bool WS = false>
__device__ __forceinline__ void fft_lds(cx<T> *lds, int stride, const FftLen &f, int th, bool active,
                                        const cx<T> *__restrict__ tw)
{
    int ns = 1;
#pragma unroll
    for (int s = 0; s < f.nstages; ++s) {
        const int R = f.radix[s];
        const unsigned mg = f.magic[s];
        if constexpr (E % 4 == 0) {
            if (R == 4) { stage<T, E, 4, SIGN, WS>(lds, stride, f.n, ns, mg, th, f.t, tw, active); ns *= 4; continue; }
        }
        if constexpr (E % 2 == 0) {
            if (R == 2) { stage<T, E, 2, SIGN, WS>(lds, stride, f.n, ns, mg, th, f.t, tw, active); ns *= 2; continue; }
        }
        if constexpr (E % 3 == 0) {
            if (R == 3) { stage<T, E, 3, SIGN, WS>(lds, stride, f.n, ns, mg, th, f.t, tw, active); ns *= 3; continue; }
        }
        if constexpr (E % 5 == 0) {
            if (R == 5) { stage<T, E, 5, SIGN, WS>(lds, stride, f.n, ns, mg, th, f.t, tw, active); ns *= 5; continue; }
        }
    }
}

}  // namespace dgpu
