// dgpu_hot_z.hip -- the z transform kernels of one RK substep (RK_CHAN_1,
// for/channel.f:62-1073), built on the in-place transforms of dgpu_fft.hpp;
// the x kernel is in dgpu_hot_x.hip.  The substep's transforms:
//
//   zinv     CU1,CU2,CU3,CTH (kept modes) --z backward, DIF--> Zt[f][j][zp][kx]
//   xfused   per z row zp, sweeping j with a window of three planes: x
//            backward (DIF) of the fields packed in pairs, the eleven products
//            of for/channel.f:410-571 and 670-720, x forward (DIT) of the
//            products packed in pairs, written as G = -(i kx X(a) + X(b)) and
//            H = -X(c) per momentum/scalar equation into Wt[o][j][zp][kx]
//   zfwd_nl  NL_f = [Z(G_f) + i kz Z(H_f)] / (NX NZ) at the kept slots (DIT)
//
// zp and the x index inside xfused are positions of the in-place transforms
// (digit-reversed), not z and x: every product is pointwise, and the forward
// transforms take the same permuted order back, so the order never matters.
//
// Intermediate rows (Zt, Wt) are ld >= NKX+1 complex values long, ld a
// multiple of 8, so a chunk of 8 columns is one 128-byte line.
#include "dgpu_hot.hpp"

namespace dgpu {

// -------------------------------------------------------------------------
// zinv: W columns per block, T = fz.t threads per column.  Stage 0 reads the
// kept slots straight from the state, the last stage writes Zt rows.
// grid (column chunks, planes, fields), block W*T, LDS W*NZ values.
// -------------------------------------------------------------------------
template <typename T, int E, typename TIN, bool CP>
__device__ __forceinline__ void zinv_body(const IpLen &fz, const cx<T> *__restrict__ tw,
                                          const cx<TIN> *__restrict__ in, long in_fs, long in_ps, int mk,
                                          cx<T> *__restrict__ out, long out_fs, long out_ps, int ld, int j0, int nkz,
                                          int tnkz, int lw, int bx, int by, int bz)
{
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const int W = 1 << lw;
    const int tid = threadIdx.x;
    const int b = tid & (W - 1), th = tid >> lw;
    const int i = bx * W + b;
    const bool col = i < mk;
    const int j = j0 + by;
    const cx<TIN> *src = in + bz * in_fs + (long)j * in_ps;
    cx<T> *dst = out + bz * out_fs + (long)j * out_ps;
    const int nz = fz.n, S = fz.nstages;
    cx<T> *x = lds + b;
    with_radix<E>(fz.radix[0], [&](auto rc) {
        constexpr int R = decltype(rc)::value;
        ip_stage<T, E, R, +1, true>(
            fz, 0, tw, th,
            [&](int, int f) {
                // unconditional load at a valid index, then select: keeps it
                // one 16-byte access
                const int s = slot_of_freq(f, nz, nkz, tnkz);
                const cx<TIN> v = src[(long)max(s, 0) * mk + (col ? i : mk - 1)];
                return (col && s >= 0) ? cvt_<T>(v) : make_cx<T>(0, 0);
            },
            [&](int, int pos, cx<T> v) { x[ip_cswz(fz, pos) * W] = v; });
    });
    __syncthreads();
    stage_loop<CP>(1, S - 1, [&](int s) {
        ip_stage_mem<T, E, +1, true>(x, W, fz, s, tw, th);
        __syncthreads();
    });
    with_radix<E>(fz.radix[S - 1], [&](auto rc) {
        constexpr int R = decltype(rc)::value;
        ip_stage<T, E, R, +1, true>(
            fz, S - 1, tw, th, [&](int, int pos) { return x[ip_cswz(fz, pos) * W]; },
            [&](int, int pos, cx<T> v) {
                if (col) dst[(long)pos * ld + i] = v;
            });
    });
}

// grid (8 * nchunks * ceil(nplanes / 8), 1, fields): workgroups go to the
// eight XCDs in turn, so block L runs on XCD L % 8; all column chunks of a
// plane are given to one XCD, one after another, because neighbouring chunks
// share the cache lines of every input row (a row of NKX+1 modes does not
// start on a line boundary) and then find them in that XCD's L2.
template <typename T, int E, typename TIN, int PID>
__global__ void __launch_bounds__(256) k_zinv_ip(IpLen fz, const cx<T> *__restrict__ tw,
                                                  const cx<TIN> *__restrict__ in, long in_fs, long in_ps, int mk,
                                                  cx<T> *__restrict__ out, long out_fs, long out_ps, int ld, int j0,
                                                  int nkz, int tnkz, int lw, int nchunks, int nplanes)
{
    const int L = blockIdx.x, r = L >> 3;
    const int chunk = r % nchunks, plane = 8 * (r / nchunks) + (L & 7);
    if (plane >= nplanes) return;
    zinv_body<T, E, TIN, PID != 0>(plan_ref<PID>(fz), tw, in, in_fs, in_ps, mk, out, out_fs, out_ps, ld, j0, nkz,
                                   tnkz, lw, chunk, plane, blockIdx.z);
}

// Members of one transform-length group: grid.x over (column chunks, planes
// of the tallest member) in plane_chunk's order, grid.z = member * nfields +
// field; planes beyond a member's own NY+2 exit at once.
template <typename T, int E, typename TIN, int PID>
__global__ void __launch_bounds__(256) k_zinv_ip_batch(const ZinvDesc<T, TIN> *__restrict__ d, IpLen fz,
                                                        const cx<T> *__restrict__ tw, long in_ps, int mk,
                                                        long out_ps, int ld, int nkz, int tnkz, int lw, int nfields)
{
    const int mem = blockIdx.z / nfields, f = blockIdx.z - mem * nfields;
    const ZinvDesc<T, TIN> D = d[mem];
    int chunk, plane;
    if (!plane_chunk(blockIdx.x, (mk + (1 << lw) - 1) >> lw, D.nplanes, chunk, plane)) return;
    zinv_body<T, E, TIN, PID != 0>(plan_ref<PID>(fz), tw, D.in, D.in_fs, in_ps, mk, D.out, D.out_fs, out_ps, ld, 0,
                                   nkz, tnkz, lw, chunk, plane, f);
}

// -------------------------------------------------------------------------
// zfwd_nl: W columns per block; both transforms of output o (G then H) in the
// same LDS; G's result stays in registers, H's last stage combines and writes
// the kept slots.  grid (column chunks, planes, outputs), block W*T.
// -------------------------------------------------------------------------
template <typename T, int E, typename TOUT, bool CP>
__device__ __forceinline__ void zfwd_body(const IpLen &fz, const cx<T> *__restrict__ tw,
                                          const cx<T> *__restrict__ wt, long wt_fs, long wt_ps, int ld, int mk,
                                          cx<TOUT> *const *__restrict__ nl, long ps, const double *__restrict__ kzv,
                                          int jlo, int nkz, int tnkz, int lw, int bx, int by, int bz)
{
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const int W = 1 << lw;
    const int tid = threadIdx.x;
    const int b = tid & (W - 1);
    int th = tid >> lw;  // re-laundered every stage (see the stage loops)
    const int i = bx * W + b;
    const bool col = i < mk;
    const int j = jlo + by;
    const int o = bz;
    const int nz = fz.n, S = fz.nstages;
    const cx<T> *sg = wt + (2 * o) * wt_fs + (long)j * wt_ps + (col ? i : mk - 1);
    const cx<T> *sh = sg + wt_fs;
    cx<T> *x = lds + b;
    const int rl = fz.radix[S - 1];
    cx<T> gin[E], hin[E], gv[E];
    // inputs of the first forward stage (stage S-1: groups of R consecutive positions)
    auto load_first = [&](const cx<T> *src, cx<T> *regs) {
        with_radix<E>(rl, [&](auto rc) {
            constexpr int R = decltype(rc)::value;
            ip_groups<E, R>(fz, S - 1, th, [&](int u, int base, int) {
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const cx<T> v = src[(long)(base + q * fz.span[S - 1]) * ld];
                    regs[u * R + q] = col ? v : make_cx<T>(0, 0);
                }
            });
        });
    };
    auto first = [&](cx<T> *regs) {
        with_radix<E>(rl, [&](auto rc) {
            constexpr int R = decltype(rc)::value;
            ip_stage<T, E, R, -1, false>(
                fz, S - 1, tw, th, [&](int idx, int) { return regs[idx]; },
                [&](int, int pos, cx<T> v) { x[ip_cswz(fz, pos) * W] = v; });
        });
    };
    load_first(sg, gin);
    load_first(sh, hin);
    first(gin);
    __syncthreads();
    stage_loop<CP>(1, S - 1, [&](int s) {
        // keeps the compiler from hoisting every stage's index arithmetic
        // out of the unrolled stages (both transforms' values are live)
        asm volatile("" : "+v"(th));
        ip_stage_mem<T, E, -1, false>(x, W, fz, S - 1 - s, tw, th);
        __syncthreads();
    });
    with_radix<E>(fz.radix[0], [&](auto rc) {
        constexpr int R = decltype(rc)::value;
        ip_stage<T, E, R, -1, false>(
            fz, 0, tw, th, [&](int, int pos) { return x[ip_cswz(fz, pos) * W]; }, [&](int idx, int, cx<T> v) { gv[idx] = v; });
    });
    __syncthreads();
    first(hin);
    __syncthreads();
    stage_loop<CP>(1, S - 1, [&](int s) {
        // keeps the compiler from hoisting every stage's index arithmetic
        // out of the unrolled stages (both transforms' values are live)
        asm volatile("" : "+v"(th));
        ip_stage_mem<T, E, -1, false>(x, W, fz, S - 1 - s, tw, th);
        __syncthreads();
    });
    const T inv_nz = T(1) / T(nz);
    cx<TOUT> *dst = nl[o] + (long)j * ps + i;
    with_radix<E>(fz.radix[0], [&](auto rc) {
        constexpr int R = decltype(rc)::value;
        ip_stage<T, E, R, -1, false>(
            fz, 0, tw, th, [&](int, int pos) { return x[ip_cswz(fz, pos) * W]; },
            [&](int idx, int f, cx<T> hv) {
                const int s = slot_of_freq(f, nz, nkz, tnkz);
                if (col && s >= 0) {
                    const cx<T> g = scale(gv[idx], inv_nz), hs = scale(hv, inv_nz);
                    const T kz = T(kzv[s]);
                    dst[(long)s * mk] = cvt_<TOUT>(make_cx<T>(g.re - kz * hs.im, g.im + kz * hs.re));
                }
            });
    });
}

// grid (8 * nchunks * ceil(nplanes / 8), 1, outputs), the XCD order of
// k_zinv_ip: the kept-slot rows it writes straddle cache lines shared by
// neighbouring column chunks.
template <typename T, int E, typename TOUT, int PID>
__global__ void __launch_bounds__(256) k_zfwd_nl_ip(IpLen fz, const cx<T> *__restrict__ tw,
                                                     const cx<T> *__restrict__ wt, long wt_fs, long wt_ps, int ld,
                                                     int mk, cx<TOUT> *const *__restrict__ nl, long ps,
                                                     const double *__restrict__ kzv, int jlo, int nkz, int tnkz, int lw,
                                                     int nchunks, int nplanes)
{
    const int L = blockIdx.x, r = L >> 3;
    const int chunk = r % nchunks, plane = 8 * (r / nchunks) + (L & 7);
    if (plane >= nplanes) return;
    zfwd_body<T, E, TOUT, PID != 0>(plan_ref<PID>(fz), tw, wt, wt_fs, wt_ps, ld, mk, nl, ps, kzv, jlo, nkz, tnkz, lw,
                                    chunk, plane, blockIdx.z);
}

// grid (column chunks x planes of the tallest member in plane_chunk's order,
// 1, members * nout)
template <typename T, int E, typename TOUT, int PID>
__global__ void __launch_bounds__(256) k_zfwd_nl_ip_batch(const ZfDesc<T, TOUT> *__restrict__ d, IpLen fz,
                                                           const cx<T> *__restrict__ tw, long wt_ps, int ld, int mk,
                                                           long ps, int nkz, int tnkz, int lw, int nout)
{
    const int mem = blockIdx.z / nout, o = blockIdx.z - mem * nout;
    const ZfDesc<T, TOUT> D = d[mem];
    int chunk, plane;
    if (!plane_chunk(blockIdx.x, (mk + (1 << lw) - 1) >> lw, D.nplanes, chunk, plane)) return;
    zfwd_body<T, E, TOUT, PID != 0>(plan_ref<PID>(fz), tw, D.wt, D.wt_fs, wt_ps, ld, mk, D.nl, ps, D.kz, D.jlo, nkz,
                                    tnkz, lw, chunk, plane, o);
}

// -------------------------------------------------------------------------
// launchers
// -------------------------------------------------------------------------

size_t zip_lds_bytes(const IpLen &fz, size_t elem) { return elem * ((size_t)1 << z_lw(fz.t)) * fz.n; }

// Id of the compile-time plan equal to f, or 0 (also when the environment
// variable DGPU_NO_CONST_PLANS is set, which keeps every kernel on its runtime
// plan).
int const_plan_id(const IpLen &f)
{
    static const bool off = std::getenv("DGPU_NO_CONST_PLANS") != nullptr;
    if (off) return 0;
    for (int id = 1; id < kNumConstPlans; ++id) {
        const IpLen c = const_plan(id);
        if (std::memcmp(&c, &f, sizeof(IpLen)) == 0) return id;
    }
    return 0;
}

template <typename T, int E, typename TIN, int PID>
static void zinv_ip_e(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw, const cx<TIN> *in, long in_fs,
                      long in_ps, cx<T> *out, long out_fs, long out_ps, int ld, int j0, int nplanes, int nfields)
{
    const int lw = z_lw(fz.t), W = 1 << lw;
    const int nchunks = (g.mk + W - 1) / W;
    dim3 grid((unsigned)(8 * nchunks * ((nplanes + 7) / 8)), 1, nfields);
    const size_t lds = sizeof(cx<T>) * W * fz.n;
    hipLaunchKernelGGL((k_zinv_ip<T, E, TIN, PID>), grid, dim3(W * fz.t), lds, s, fz, tw, in, in_fs, in_ps, g.mk, out,
                       out_fs, out_ps, ld, j0, g.nkz, g.tnkz, lw, nchunks, nplanes);
}

template <typename T, typename TIN>
void launch_zinv_ip(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw, const cx<TIN> *in, long in_fs,
                    long in_ps, cx<T> *out, long out_fs, long out_ps, int ld, int j0, int nplanes, int nfields)
{
    if (with_pid(const_plan_id(fz), ZPids<T>{}, [&](auto pc) {
            constexpr int P = decltype(pc)::value;
            zinv_ip_e<T, kConstPlanSpecs[P].e, TIN, P>(s, g, fz, tw, in, in_fs, in_ps, out, out_fs, out_ps, ld, j0,
                                                       nplanes, nfields);
        }))
        return;
    DGPU_DISPATCH_IPE(fz.e, (zinv_ip_e<T, EE, TIN, 0>(s, g, fz, tw, in, in_fs, in_ps, out, out_fs, out_ps, ld, j0,
                                                     nplanes, nfields)));
}

template <typename T, int E, typename TOUT, int PID>
static void zfwd_nl_ip_e(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw, const cx<T> *wt, long wt_fs,
                         int ld, cx<TOUT> *const *nl, const double *kz, int jlo, int jhi, int nout)
{
    const int lw = z_lw(fz.t), W = 1 << lw;
    const int nchunks = (g.mk + W - 1) / W, nplanes = jhi - jlo + 1;
    dim3 grid((unsigned)(8 * nchunks * ((nplanes + 7) / 8)), 1, nout);
    const size_t lds = sizeof(cx<T>) * W * fz.n;
    const long wt_ps = (long)g.nz * ld;
    hipLaunchKernelGGL((k_zfwd_nl_ip<T, E, TOUT, PID>), grid, dim3(W * fz.t), lds, s, fz, tw, wt, wt_fs, wt_ps, ld,
                       g.mk, nl, g.ps, kz, jlo, g.nkz, g.tnkz, lw, nchunks, nplanes);
}

template <typename T, typename TOUT>
void launch_zfwd_nl_ip(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw, const cx<T> *wt, long wt_fs,
                       int ld, cx<TOUT> *const *nl, const double *kz, int jlo, int jhi, int nout)
{
    if (with_pid(const_plan_id(fz), ZPids<T>{}, [&](auto pc) {
            constexpr int P = decltype(pc)::value;
            zfwd_nl_ip_e<T, kConstPlanSpecs[P].e, TOUT, P>(s, g, fz, tw, wt, wt_fs, ld, nl, kz, jlo, jhi, nout);
        }))
        return;
    DGPU_DISPATCH_IPE(fz.e, (zfwd_nl_ip_e<T, EE, TOUT, 0>(s, g, fz, tw, wt, wt_fs, ld, nl, kz, jlo, jhi, nout)));
}

// ---- batched (ensemble) launchers ----
template <typename T, int E, typename TIN, int PID>
static void zinv_ip_batch_e(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw,
                            const ZinvDesc<T, TIN> *d, int nmem, int max_planes, int ld, int nfields)
{
    const int lw = z_lw(fz.t), W = 1 << lw;
    // plane_chunk's order over (column chunks, planes of the tallest member)
    dim3 grid((unsigned)(8 * ((g.mk + W - 1) / W) * ((max_planes + 7) / 8)), 1, nfields * nmem);
    const size_t lds = sizeof(cx<T>) * W * fz.n;
    hipLaunchKernelGGL((k_zinv_ip_batch<T, E, TIN, PID>), grid, dim3(W * fz.t), lds, s, d, fz, tw, g.ps, g.mk,
                       (long)g.nz * ld, ld, g.nkz, g.tnkz, lw, nfields);
}

template <typename T, typename TIN>
void launch_zinv_ip_batch(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw, const ZinvDesc<T, TIN> *d,
                          int nmem, int max_planes, int ld, int nfields)
{
    if (with_pid(const_plan_id(fz), ZPids<T>{}, [&](auto pc) {
            constexpr int P = decltype(pc)::value;
            zinv_ip_batch_e<T, kConstPlanSpecs[P].e, TIN, P>(s, g, fz, tw, d, nmem, max_planes, ld, nfields);
        }))
        return;
    DGPU_DISPATCH_IPE(fz.e, (zinv_ip_batch_e<T, EE, TIN, 0>(s, g, fz, tw, d, nmem, max_planes, ld, nfields)));
}

template <typename T, int E, typename TOUT, int PID>
static void zfwd_nl_ip_batch_e(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw,
                               const ZfDesc<T, TOUT> *d, int nmem, int max_planes, int ld, int nout)
{
    const int lw = z_lw(fz.t), W = 1 << lw;
    // plane_chunk's order over (column chunks, planes of the tallest member)
    dim3 grid((unsigned)(8 * ((g.mk + W - 1) / W) * ((max_planes + 7) / 8)), 1, nout * nmem);
    const size_t lds = sizeof(cx<T>) * W * fz.n;
    hipLaunchKernelGGL((k_zfwd_nl_ip_batch<T, E, TOUT, PID>), grid, dim3(W * fz.t), lds, s, d, fz, tw,
                       (long)g.nz * ld, ld, g.mk, g.ps, g.nkz, g.tnkz, lw, nout);
}

template <typename T, typename TOUT>
void launch_zfwd_nl_ip_batch(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw, const ZfDesc<T, TOUT> *d,
                             int nmem, int max_planes, int ld, int nout)
{
    if (with_pid(const_plan_id(fz), ZPids<T>{}, [&](auto pc) {
            constexpr int P = decltype(pc)::value;
            zfwd_nl_ip_batch_e<T, kConstPlanSpecs[P].e, TOUT, P>(s, g, fz, tw, d, nmem, max_planes, ld, nout);
        }))
        return;
    DGPU_DISPATCH_IPE(fz.e, (zfwd_nl_ip_batch_e<T, EE, TOUT, 0>(s, g, fz, tw, d, nmem, max_planes, ld, nout)));
}

template void launch_zinv_ip_batch<double, double>(hipStream_t, const Geo &, const IpLen &, const cx<double> *,
                                                   const ZinvDesc<double, double> *, int, int, int, int);
template void launch_zinv_ip_batch<float, double>(hipStream_t, const Geo &, const IpLen &, const cx<float> *,
                                                  const ZinvDesc<float, double> *, int, int, int, int);
template void launch_zinv_ip_batch<float, float>(hipStream_t, const Geo &, const IpLen &, const cx<float> *,
                                                 const ZinvDesc<float, float> *, int, int, int, int);
template void launch_zfwd_nl_ip_batch<double, double>(hipStream_t, const Geo &, const IpLen &, const cx<double> *,
                                                      const ZfDesc<double, double> *, int, int, int, int);
template void launch_zfwd_nl_ip_batch<float, float>(hipStream_t, const Geo &, const IpLen &, const cx<float> *,
                                                    const ZfDesc<float, float> *, int, int, int, int);

template void launch_zinv_ip<double, double>(hipStream_t, const Geo &, const IpLen &, const cx<double> *,
                                             const cx<double> *, long, long, cx<double> *, long, long, int, int, int,
                                             int);
template void launch_zinv_ip<float, float>(hipStream_t, const Geo &, const IpLen &, const cx<float> *,
                                           const cx<float> *, long, long, cx<float> *, long, long, int, int, int,
                                           int);
template void launch_zinv_ip<float, double>(hipStream_t, const Geo &, const IpLen &, const cx<float> *,
                                            const cx<double> *, long, long, cx<float> *, long, long, int, int, int,
                                            int);
template void launch_zfwd_nl_ip<double, double>(hipStream_t, const Geo &, const IpLen &, const cx<double> *,
                                                const cx<double> *, long, int, cx<double> *const *, const double *,
                                                int, int, int);
template void launch_zfwd_nl_ip<float, float>(hipStream_t, const Geo &, const IpLen &, const cx<float> *,
                                              const cx<float> *, long, int, cx<float> *const *, const double *, int,
                                              int, int);

}  // namespace dgpu
