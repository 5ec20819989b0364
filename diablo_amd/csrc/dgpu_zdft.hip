// dgpu_zdft.hip -- the z transforms of the mixed-precision substep as
// DFT-as-GEMM on the matrix cores, for NZ <= 256 (a multiple of 32): the
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
// load per lane per k-step); the data operand is split on the fly.
//
// MFMA 32x32x16 operand layout (lane l of 64): A row l%32, k = (l/32)*8 + t;
// B column l%32, k = (l/32)*8 + t; D column l%32, row (r/4)*8 + (l/32)*4 + r%4
// for result register r.
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

// One wave: M tiles mt0, mt0+1 of the output for column tile nt; ld_b(q, i)
// returns complex input q of column i (zero outside), st(q, i, v) stores
// complex output q of column i.
template <class LdB, class St>
__device__ __forceinline__ void zdft_wave(const uint4 *__restrict__ ahi, const uint4 *__restrict__ alo, int nmt,
                                          int nks, int mt0, int i, int lane, LdB &&ld_b, St &&st)
{
    f32x16 acc0 = {}, acc1 = {};
    const bool has1 = mt0 + 1 < nmt;
    const int half = lane >> 5;
    for (int ks = 0; ks < nks; ++ks) {
        const int q0 = ks * 8 + half * 4;
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
        {
            const long a = ((long)mt0 * nks + ks) * 64 + lane;
            const bf16x8 ah = as_bf16x8(ahi[a]), al = as_bf16x8(alo[a]);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc0, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc0, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc0, 0, 0, 0);
        }
        if (has1) {
            const long a = ((long)(mt0 + 1) * nks + ks) * 64 + lane;
            const bf16x8 ah = as_bf16x8(ahi[a]), al = as_bf16x8(alo[a]);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc1, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc1, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc1, 0, 0, 0);
        }
    }
    // rows m, m+1 (m even) of a register pair are (re, im) of output m/2
#pragma unroll
    for (int r = 0; r < 16; r += 2) {
        const int m = (r >> 2) * 8 + half * 4 + (r & 3);
        st((mt0 * 32 + m) >> 1, i, make_float2(acc0[r], acc0[r + 1]));
        if (has1) st(((mt0 + 1) * 32 + m) >> 1, i, make_float2(acc1[r], acc1[r + 1]));
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
// grid (column tiles, planes, fields), block 64 * ceil(nmt / 2).
template <typename TIN>
__global__ void __launch_bounds__(512) k_zdft_inv(const uint4 *__restrict__ ahi, const uint4 *__restrict__ alo,
                                                   int nmt, int nks, const cx<TIN> *__restrict__ in, long in_fs,
                                                   long in_ps, int mk, int mz, cx<float> *__restrict__ out,
                                                   long out_fs, long out_ps, int ld, int nz, int j0)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 32 + (lane & 31);
    const int j = j0 + blockIdx.y, f = blockIdx.z;
    const cx<TIN> *src = in + f * in_fs + (long)j * in_ps;
    cx<float> *dst = out + f * out_fs + (long)j * out_ps;
    const bool col = i < mk;
    zdft_wave(
        ahi, alo, nmt, nks, 2 * w, i, lane,
        [&](int q, int ii) {
            // unconditional load at a valid index, then select
            const float2 v = ld2(src + (long)min(q, mz - 1) * mk + (col ? ii : mk - 1));
            return (col && q < mz) ? v : make_float2(0.f, 0.f);
        },
        [&](int q, int ii, float2 v) {
            if (col && q < nz) {
                cx<float> c;
                c.re = v.x;
                c.im = v.y;
                dst[(long)q * ld + ii] = c;
            }
        });
}

// z forward: NL[o][j][s][i] at the kept slots s < mz from G = Wt[2o], H =
// Wt[2o+1] (natural z); the weights hold 1/NZ and i kz.  grid (column tiles,
// planes jlo.., outputs), block 64 * ceil(nmt / 2).
__global__ void __launch_bounds__(512) k_zdft_fwd(const uint4 *__restrict__ ahi, const uint4 *__restrict__ alo,
                                                   int nmt, int nks, const cx<float> *__restrict__ wt, long wt_fs,
                                                   long wt_ps, int ld, int nz, int mk, int mz,
                                                   cx<float> *const *__restrict__ nl, long ps, int jlo)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 32 + (lane & 31);
    const int j = jlo + blockIdx.y, o = blockIdx.z;
    const cx<float> *g = wt + (2 * o) * wt_fs + (long)j * wt_ps;
    cx<float> *dst = nl[o] + (long)j * ps;
    const bool col = i < mk;
    zdft_wave(
        ahi, alo, nmt, nks, 2 * w, i, lane,
        [&](int q, int ii) {
            // q < NZ: G at z = q; NZ <= q < 2 NZ: H at z = q - NZ
            const int h = q >= nz ? 1 : 0;
            const float2 v = ld2(g + h * wt_fs + (long)(q - h * nz) * ld + (col ? ii : mk - 1));
            return col ? v : make_float2(0.f, 0.f);
        },
        [&](int s, int ii, float2 v) {
            if (col && s < mz) {
                cx<float> c;
                c.re = v.x;
                c.im = v.y;
                dst[(long)s * mk + ii] = c;
            }
        });
}

template <typename TIN>
void launch_zdft_inv(hipStream_t s, const Geo &g, const ZdftTab &t, const cx<TIN> *in, long in_fs, long in_ps,
                     cx<float> *out, long out_fs, long out_ps, int ld, int j0, int nplanes, int nf)
{
    dim3 grid((unsigned)((g.mk + 31) / 32), nplanes, nf);
    hipLaunchKernelGGL(k_zdft_inv<TIN>, grid, dim3(64 * ((t.nmt + 1) / 2)), 0, s, t.hi, t.lo, t.nmt, t.nks, in,
                       in_fs, in_ps, g.mk, g.mz, out, out_fs, out_ps, ld, g.nz, j0);
}

void launch_zdft_fwd(hipStream_t s, const Geo &g, const ZdftTab &t, const cx<float> *wt, long wt_fs, int ld,
                     cx<float> *const *nl, int jlo, int jhi, int nout)
{
    dim3 grid((unsigned)((g.mk + 31) / 32), jhi - jlo + 1, nout);
    hipLaunchKernelGGL(k_zdft_fwd, grid, dim3(64 * ((t.nmt + 1) / 2)), 0, s, t.hi, t.lo, t.nmt, t.nks, wt, wt_fs,
                       (long)g.nz * ld, ld, g.nz, g.mk, g.mz, nl, g.ps, jlo);
}

template void launch_zdft_inv<float>(hipStream_t, const Geo &, const ZdftTab &, const cx<float> *, long, long,
                                     cx<float> *, long, long, int, int, int, int);
template void launch_zdft_inv<double>(hipStream_t, const Geo &, const ZdftTab &, const cx<double> *, long, long,
                                      cx<float> *, long, long, int, int, int, int);

}  // namespace dgpu
