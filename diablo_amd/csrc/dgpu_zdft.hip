// dgpu_zdft.hip -- the z transforms of the mixed-precision substep as
// DFT-as-GEMM on the matrix cores, for NZ <= 256 (a multiple of 16): the
// same maps as zinv / zfwd_nl of dgpu_hot_z.hip (kept kz slots -> NZ points,
// unnormalised, sign +1; NZ points of G and H -> NL = [Z(G) + i kz Z(H)] / NZ
// at the kept slots, sign -1), in natural z order instead of the in-place
// transforms' digit-reversed positions (the x kernel is pointwise in z, so
// either order works as long as both directions agree).
//
// Complex products are real GEMMs on interleaved (re, im) rows and columns:
// out[2m + c'] = sum_k' A[2m + c'][k'] x[k'], with the 2x2 blocks
// [[Re w, -Im w], [Im w, Re w]] of each complex weight w.  Operands are split
// BF16 (v = hi + lo, hi = bf16(v), lo = bf16(v - hi)); three products
// hi*hi + hi*lo + lo*hi accumulate in FP32 (v_mfma_f32_32x32x16_bf16).  The
// weights A are fixed per context and stored fragment-major (one 16-byte
// load per lane per k-step); the data operand is split once per block.
//
// MFMA 32x32x16 operand layout (lane l of 64): A row l%32, k = (l/32)*8 + t;
// B column l%32, k = (l/32)*8 + t; D column l%32, row (r/4)*8 + (l/32)*4 + r%4
// for result register r.
#include <type_traits>

#include "dgpu_kernels.hpp"

namespace dgpu {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v)
{
    bf16x8 r;
    __builtin_memcpy(&r, &v, sizeof(r));
    return r;
}

// One block: all column tiles (NT of 32 columns) of one (field, plane) and
// the M tiles from mbase on; wave w owns M tiles mbase + 2w, mbase + 2w + 1.
// The data operand is split into BF16 halves once per block, a chunk of kKC
// k-steps at a time, into LDS in fragment order.  Each wave reads its two
// weight fragments per k-step once (the next k-step's are in flight while
// the current one's products run) and uses them for all NT column tiles.
// ld_b(q, i): complex input q of column i (zero outside); st(q, i, v):
// complex output q of column i.
constexpr int kKC = 8;

struct AFrag {
    bf16x8 h0, l0, h1, l1;
};

template <int NT, class LdB, class St>
__device__ __forceinline__ void zdft_block(const uint4 *__restrict__ ahi, const uint4 *__restrict__ alo, int nmt,
                                           int nks, int mk, int mbase, LdB &&ld_b, St &&st)
{
    __shared__ uint4 bsh[kKC * NT * 64], bsl[kKC * NT * 64];  // [ks][nt][lane], hi and lo halves
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nthr = blockDim.x;
    const int mt0 = mbase + 2 * w;
    const bool has0 = mt0 < nmt, has1 = mt0 + 1 < nmt;
    const int half = lane >> 5;
    f32x16 acc[2][NT];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[a][n] = f32x16{};
    auto load_a = [&](int k) {
        AFrag f;
        const long a0 = ((long)(has0 ? mt0 : 0) * nks + k) * 64 + lane;
        const long a1 = has1 ? a0 + (long)nks * 64 : a0;
        f.h0 = as_bf16x8(ahi[a0]);
        f.l0 = as_bf16x8(alo[a0]);
        f.h1 = as_bf16x8(ahi[a1]);
        f.l1 = as_bf16x8(alo[a1]);
        return f;
    };
    AFrag nxt = load_a(0);
    for (int k0 = 0; k0 < nks; k0 += kKC) {
        const int kc = min(kKC, nks - k0);
        // split this chunk of the data operand into LDS
        for (int fr = tid; fr < kc * NT * 64; fr += nthr) {
            const int l = fr & 63, nt = (fr >> 6) % NT, ks = (fr >> 6) / NT;
            const int i = nt * 32 + (l & 31);
            const int q0 = (k0 + ks) * 8 + (l >> 5) * 4;
            bf16x8 bh, bl;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float2 v = ld_b(q0 + u, i);
                const __bf16 hr = (__bf16)v.x, hi_ = (__bf16)v.y;
                bh[2 * u] = hr;
                bh[2 * u + 1] = hi_;
                bl[2 * u] = (__bf16)(v.x - (float)hr);
                bl[2 * u + 1] = (__bf16)(v.y - (float)hi_);
            }
            uint4 h4, l4;
            __builtin_memcpy(&h4, &bh, 16);
            __builtin_memcpy(&l4, &bl, 16);
            bsh[fr] = h4;
            bsl[fr] = l4;
        }
        __syncthreads();
        if (has0) {
            for (int ks = 0; ks < kc; ++ks) {
                const AFrag cur = nxt;
                if (k0 + ks + 1 < nks) nxt = load_a(k0 + ks + 1);
#pragma unroll
                for (int n = 0; n < NT; ++n) {
                    const int fr = (ks * NT + n) * 64 + lane;
                    const bf16x8 bh = as_bf16x8(bsh[fr]), bl = as_bf16x8(bsl[fr]);
                    acc[0][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.h0, bh, acc[0][n], 0, 0, 0);
                    acc[0][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.h0, bl, acc[0][n], 0, 0, 0);
                    acc[0][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.l0, bh, acc[0][n], 0, 0, 0);
                    if (has1) {
                        acc[1][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.h1, bh, acc[1][n], 0, 0, 0);
                        acc[1][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.h1, bl, acc[1][n], 0, 0, 0);
                        acc[1][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.l1, bh, acc[1][n], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (!has0) return;
    // rows m, m+1 (m even) of a register pair are (re, im) of output m/2
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int i = n * 32 + (lane & 31);
        if (i >= mk) continue;
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const int m = (r >> 2) * 8 + half * 4 + (r & 3);
            st((mt0 * 32 + m) >> 1, i, make_float2(acc[0][n][r], acc[0][n][r + 1]));
            if (has1) st(((mt0 + 1) * 32 + m) >> 1, i, make_float2(acc[1][n][r], acc[1][n][r + 1]));
        }
    }
}

template <typename TIN>
__device__ __forceinline__ float2 ld2(const cx<TIN> *p)
{
    const cx<TIN> v = *p;
    return make_float2((float)v.re, (float)v.im);
}

}  // namespace

// z backward: Zt[f][j][z][i] (natural z) from the kept slots of the state.
// grid (planes, fields, M blocks), block 64 * waves (zdft_waves).
template <typename TIN, int NT>
__global__ void __launch_bounds__(512) k_zdft_inv(const uint4 *__restrict__ ahi, const uint4 *__restrict__ alo,
                                                   int nmt, int nks, const cx<TIN> *__restrict__ in, long in_fs,
                                                   long in_ps, int mk, int mz, cx<float> *__restrict__ out,
                                                   long out_fs, long out_ps, int ld, int nz, int j0)
{
    const int j = j0 + blockIdx.x, f = blockIdx.y;
    const cx<TIN> *src = in + f * in_fs + (long)j * in_ps;
    cx<float> *dst = out + f * out_fs + (long)j * out_ps;
    zdft_block<NT>(
        ahi, alo, nmt, nks, mk, (int)blockIdx.z * 2 * (int)(blockDim.x >> 6),
        [&](int q, int i) {
            // unconditional load at a valid index, then select
            const float2 v = ld2(src + (long)min(q, mz - 1) * mk + min(i, mk - 1));
            return (i < mk && q < mz) ? v : make_float2(0.f, 0.f);
        },
        [&](int q, int i, float2 v) {
            if (q < nz) {
                cx<float> c;
                c.re = v.x;
                c.im = v.y;
                dst[(long)q * ld + i] = c;
            }
        });
}

// z forward: NL[o][j][s][i] at the kept slots s < mz from G = Wt[2o], H =
// Wt[2o+1] (natural z); the weights hold 1/NZ and i kz.  grid (planes
// jlo.., outputs, M blocks), block 64 * waves (zdft_waves).
template <int NT>
__global__ void __launch_bounds__(512) k_zdft_fwd(const uint4 *__restrict__ ahi, const uint4 *__restrict__ alo,
                                                   int nmt, int nks, const cx<float> *__restrict__ wt, long wt_fs,
                                                   long wt_ps, int ld, int nz, int mk, int mz,
                                                   cx<float> *const *__restrict__ nl, long ps, int jlo)
{
    const int j = jlo + blockIdx.x, o = blockIdx.y;
    const cx<float> *g = wt + (2 * o) * wt_fs + (long)j * wt_ps;
    cx<float> *dst = nl[o] + (long)j * ps;
    zdft_block<NT>(
        ahi, alo, nmt, nks, mk, (int)blockIdx.z * 2 * (int)(blockDim.x >> 6),
        [&](int q, int i) {
            // q < NZ: G at z = q; NZ <= q < 2 NZ: H at z = q - NZ
            const int h = q >= nz ? 1 : 0;
            const float2 v = ld2(g + h * wt_fs + (long)(q - h * nz) * ld + min(i, mk - 1));
            return i < mk ? v : make_float2(0.f, 0.f);
        },
        [&](int s, int i, float2 v) {
            if (s < mz) {
                cx<float> c;
                c.re = v.x;
                c.im = v.y;
                dst[(long)s * mk + i] = c;
            }
        });
}

// M tiles of a launch split over ceil(nmt / 8) blocks per plane: blocks of
// at most four waves keep three blocks resident per CU, so one block's
// staging overlaps another's products.
static void zdft_waves(int nmt, int &waves, int &mblocks)
{
    mblocks = (nmt + 7) / 8;
    waves = ((nmt + 1) / 2 + mblocks - 1) / mblocks;
}

template <class F>
static void with_nt(int mk, F &&f)
{
    const int nt = (mk + 31) / 32;
    if (nt <= 1) f(std::integral_constant<int, 1>{});
    else if (nt == 2) f(std::integral_constant<int, 2>{});
    else if (nt == 3) f(std::integral_constant<int, 3>{});
    else f(std::integral_constant<int, 4>{});
}

template <typename TIN>
void launch_zdft_inv(hipStream_t s, const Geo &g, const ZdftTab &t, const cx<TIN> *in, long in_fs, long in_ps,
                     cx<float> *out, long out_fs, long out_ps, int ld, int j0, int nplanes, int nf)
{
    with_nt(g.mk, [&](auto ntc) {
        constexpr int NT = decltype(ntc)::value;
        int waves, mb;
        zdft_waves(t.nmt, waves, mb);
        hipLaunchKernelGGL((k_zdft_inv<TIN, NT>), dim3(nplanes, nf, mb), dim3(64 * waves), 0, s, t.hi,
                           t.lo, t.nmt, t.nks, in, in_fs, in_ps, g.mk, g.mz, out, out_fs, out_ps, ld, g.nz, j0);
    });
}

void launch_zdft_fwd(hipStream_t s, const Geo &g, const ZdftTab &t, const cx<float> *wt, long wt_fs, int ld,
                     cx<float> *const *nl, int jlo, int jhi, int nout)
{
    with_nt(g.mk, [&](auto ntc) {
        constexpr int NT = decltype(ntc)::value;
        int waves, mb;
        zdft_waves(t.nmt, waves, mb);
        hipLaunchKernelGGL((k_zdft_fwd<NT>), dim3(jhi - jlo + 1, nout, mb), dim3(64 * waves), 0, s, t.hi,
                           t.lo, t.nmt, t.nks, wt, wt_fs, (long)g.nz * ld, ld, g.nz, g.mk, g.mz, nl, g.ps, jlo);
    });
}

template void launch_zdft_inv<float>(hipStream_t, const Geo &, const ZdftTab &, const cx<float> *, long, long,
                                     cx<float> *, long, long, int, int, int, int);
template void launch_zdft_inv<double>(hipStream_t, const Geo &, const ZdftTab &, const cx<double> *, long, long,
                                      cx<float> *, long, long, int, int, int, int);

}  // namespace dgpu
