// dgpu_hot_x.hip -- the fused x kernel of one RK substep (the transforms and
// products of RK_CHAN_1, for/channel.f:62-1073) and COURANT's velocity maxima
// (for/courant.f:35-44), on the in-place transforms of dgpu_fft.hpp.
#include "dgpu_hot.hpp"

namespace dgpu {

// -------------------------------------------------------------------------
// xfused: zr z rows per block (one, unless the rows are short: then several
// rows fill the block, each with its own window and product rows), planes
// ja..jb (a chunk of the sweep).
//
// Rows of the transform passes (T = fx.t threads each): rows 0..npi-1 run the
// x backward of the next plane (pairs (u1,u2), (u3,th)), rows npi..npi+npo-1
// the x forward of the previous plane's products, in one pass of S stages.
// The backward's stage 0 takes its input from registers (fetched one plane
// ahead); the forward's stage 0 (the last) writes the output, which needs the
// products paired so that each G/H comes from one pair row:
//   row 0 (U1U1, d/dy U1U2) -> G1           row 3 (U3U2@GY, U3U3) -> H2, H3
//   row 1 (U1U2@GY, d/dy U2U2) -> G2        row 4 (TH U1, d/dy TH U2) -> G4
//   row 2 (U1U3, d/dy U3U2) -> G3, H1       row 5 (TH U3, 0) -> H4
// LDS: window of three physical planes [3][npi][NX], products [npo][NX],
// twiddles, the kept modes' KX.  Transform positions sit at swizzled slots
// (ip_swz) of their rows; the products read and write slot by slot, which is
// position by position in the same permutation for every row.
// -------------------------------------------------------------------------
template <typename T, int E, int NTH, bool CP, bool MZ>
__device__ __forceinline__ void xfused_body(const IpLen &fx, const cx<T> *__restrict__ twg, int mk, int ld,
                                            const cx<T> *__restrict__ zt, long zt_fs, long zt_ps,
                                            cx<T> *__restrict__ wt, long wt_fs, long wt_ps,
                                            const double *__restrict__ dy, const double *__restrict__ dyf,
                                            const double *__restrict__ kxv, int jlo, int jhi, int chunk, int bx,
                                            int by, int zr_rt, int nz)
{
    // MZ: several z rows per block (zr_rt); the single-run kernels of the
    // long-row plans keep one, known at compile time
    const int zr = MZ ? zr_rt : 1;
    using C = XCount<NTH>;
    constexpr int npi = C::npi, npo = C::npo;
    constexpr int nrow = npi + npo;  // transform rows per z row
    constexpr int R0 = E / 2;
    extern __shared__ unsigned char smem_raw[];
    const int nx = fx.n;
    const int zsz = (3 * npi + npo) * nx;  // a z row's window and product rows
    const int TT = fx.t, S = fx.nstages, L0 = fx.span[0];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int pg = tid / TT, zs = MZ ? pg / nrow : 0;
    const int p = pg - zs * nrow;
    int th = tid - pg * TT;  // re-laundered every plane (see the loop)
    cx<T> *win = reinterpret_cast<cx<T> *>(smem_raw) + zs * zsz;
    cx<T> *fb = win + 3 * npi * nx;
    cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw) + zr * zsz;
    T *kxs = reinterpret_cast<T *>(tw + fx.twn);  // KX of the kept modes, read by every plane's output
    for (int e = tid; e < fx.twn; e += nthr) tw[e] = twg[e];
    for (int e = tid; e < mk; e += nthr) kxs[e] = T(kxv[e]);
    // a block's last rows beyond NZ run on row NZ-1's data and store nothing
    const int zi = bx * zr + zs;
    const bool zvalid = zi < nz;
    const long zrow = (long)min(zi, nz - 1) * ld;
    const int ja = jlo + by * chunk;
    const int jb = min(jhi, ja + chunk - 1);

    // Spectral input of stage 0's two groups (k = th, th + T; frequencies
    // k + q L0) of pair `pair` of plane `plane`, kept in registers.
    cx<T> pa[E], pb[E];
    auto fetch = [&](int plane, int pair) {
        const cx<T> *fa = zt + (2 * pair) * zt_fs + (long)plane * zt_ps + zrow;
        const cx<T> *fbb = fa + zt_fs;
        const bool hasb = 2 * pair + 1 < C::nin;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
#pragma unroll
            for (int q = 0; q < R0; ++q) {
                const int f = th + TT * u + q * L0;
                const int kk = f < mk ? f : (f > nx - mk ? nx - f : -1);
                const cx<T> A = fa[max(kk, 0)];  // unconditional 16-byte loads, then select
                const cx<T> B = hasb ? fbb[max(kk, 0)] : make_cx<T>(0, 0);
                pa[u * R0 + q] = kk >= 0 ? A : make_cx<T>(0, 0);
                pb[u * R0 + q] = kk >= 0 ? B : make_cx<T>(0, 0);
            }
        }
    };
    // Backward stage 0 from the registers: spectrum of A + iB with the
    // Hermitian halves of both rows (C2R ignores Im of the DC term).
    auto dif0 = [&](cx<T> *row) {
        ip_stage<T, E, R0, +1, true>(
            fx, 0, tw, th,
            [&](int idx, int f) {
                cx<T> A = pa[idx], B = pb[idx];
                if (f == 0) {
                    A.im = T(0);
                    B.im = T(0);
                }
                if (f > nx - mk) {
                    A.im = -A.im;
                    B.im = -B.im;
                }
                return make_cx<T>(A.re - B.im, A.im + B.re);
            },
            [&](int, int pos, cx<T> v) { row[ip_swz(fx, pos)] = v; });
    };
    // Forward stage 0 of product pair row r (plane j): groups th and L0 - th
    // (0 and L0/2 for th == 0) hold Z(k') and Z(NX - k') of their frequencies.
    auto dit_out = [&](const cx<T> *row, int r, int j) {
        const int ga = th, gb = th == 0 ? TT : L0 - th;
        cx<T> va[R0], vb[R0];
#pragma unroll
        for (int q = 0; q < R0; ++q) {
            va[q] = row[ip_swz(fx, ga + q * L0)];
            vb[q] = row[ip_swz(fx, gb + q * L0)];
        }
        ip_group<T, R0, -1, false>(va, tw + fx.twoff[0], ga);
        ip_group<T, R0, -1, false>(vb, tw + fx.twoff[0], gb);
        const T h = T(0.5) / T(nx);
        cx<T> *o0 = wt + (long)j * wt_ps + zrow;
        // Outputs of pair row r (field slots of wt): G1 | G2 | G3, H1 | H2, H3
        // | G4 | H4.  A wavefront holds two rows.  FP32 takes the slots as
        // per-thread offsets (no branch per element; the second output
        // belongs to rows 2 and 3, which share a wavefront): 0.81 -> 0.79 ms.
        // FP64 keeps the switch: the offset form took it from 223 to 234
        // VGPRs and 1.77 -> 1.98 ms.
        const long off1 = (long)(r == 0 ? 0 : r == 1 ? 2 : r == 2 ? 4 : r == 3 ? 3 : r == 4 ? 6 : 7) * wt_fs;
        const long off2 = (long)(r == 2 ? 1 : 5) * wt_fs;
        const bool two = r == 2 || r == 3, xa1 = r == 3 || r == 5;
        auto emit = [&](int kp, cx<T> Zk, cx<T> Zn) {
            if (!zvalid) return;
            const cx<T> Xa = make_cx<T>((Zk.re + Zn.re) * h, (Zk.im - Zn.im) * h);
            const cx<T> Xb = make_cx<T>((Zk.im + Zn.im) * h, -(Zk.re - Zn.re) * h);
            const T kx = kxs[kp];
            // -(i kx Xa + Xb)
            const cx<T> G = make_cx<T>(kx * Xa.im - Xb.re, -kx * Xa.re - Xb.im);
            const cx<T> nXa = make_cx<T>(-Xa.re, -Xa.im), nXb = make_cx<T>(-Xb.re, -Xb.im);
            if constexpr (sizeof(T) == 4) {
                o0[off1 + kp] = xa1 ? nXa : G;
                if (two) o0[off2 + kp] = r == 2 ? nXa : nXb;
            } else {
                switch (r) {
                case 0: o0[0 * wt_fs + kp] = G; break;
                case 1: o0[2 * wt_fs + kp] = G; break;
                case 2: o0[4 * wt_fs + kp] = G; o0[1 * wt_fs + kp] = nXa; break;
                case 3: o0[3 * wt_fs + kp] = nXa; o0[5 * wt_fs + kp] = nXb; break;
                case 4: o0[6 * wt_fs + kp] = G; break;
                default: o0[7 * wt_fs + kp] = nXa; break;
                }
            }
        };
#pragma unroll
        for (int q = 0; q < R0; ++q) {
            const int ka = ga + q * L0;
            if (ka < mk) emit(ka, va[q], ga == 0 ? va[(R0 - q) % R0] : vb[R0 - 1 - q]);
            const int kb = gb + q * L0;
            if (kb < mk) emit(kb, vb[q], th == 0 ? vb[R0 - 1 - q] : va[R0 - 1 - q]);
        }
    };
    auto slot_row = [&](int plane, int pair) { return win + (plane % 3) * npi * nx + pair * nx; };
    // The eleven products of plane j at every position, from planes j-1, j, j+1.
    auto products = [&](int j) {
        cx<T> *const base = reinterpret_cast<cx<T> *>(smem_raw);
        const T dyfj = T(dyf[j]), dyfm = T(dyf[j - 1]);
        // the plane's divisors as reciprocals, once per plane (an IEEE
        // division is a dozen instructions; within an ulp of dividing)
        const T r2dy = T(1) / (T(2) * T(dy[j])), r2dyf = T(1) / (T(2) * dyfj), rdy = T(1) / T(dy[j]);
        // The positions beyond the last full pass go to the highest
        // wavefronts, not to wavefront 0, which also fetches the next plane
        // and runs its backward stage 0 alone after this phase.
        const int npos = zr * nx;  // positions of all the block's z rows
        const int extra = npos % nthr;
        const int e0 = extra ? (tid + nthr - extra) % nthr : tid;
        for (int e = e0; e < npos; e += nthr) {
            const int zq = e / nx, x = e - zq * nx;
            cx<T> *const zw = base + zq * zsz;
            const cx<T> *wm = zw + ((j - 1) % 3) * npi * nx, *w0 = zw + (j % 3) * npi * nx,
                        *wp = zw + ((j + 1) % 3) * npi * nx;
            cx<T> *const fq = zw + 3 * npi * nx;
            const cx<T> a0 = w0[x], am = wm[x], ap = wp[x];
            const cx<T> b0 = w0[nx + x], bm = wm[nx + x], bp = wp[nx + x];
            const T u1 = a0.re, u2 = a0.im, u3 = b0.re;
            const T u1m = am.re, u2m = am.im, u3m = bm.re;
            const T u1p = ap.re, u2p = ap.im, u3p = bp.re;
            const T p0 = u3 * u1;                                                      // U1*U3
            const T p1 = u1 * u1;                                                      // U1*U1
            const T p2 = u3 * u3;                                                      // U3*U3
            const T p3 = ((dyfj * u1 + dyfm * u1m) * r2dy) * u2;                 // U1*U2 at GY
            const T p4 = ((dyfj * u3 + dyfm * u3m) * r2dy) * u2;                 // U3*U2 at GY
            const T p5 = (u1p * u2p + u1 * u2p - u1 * u2 - u1m * u2) * r2dyf;    // d/dy(U1 U2)
            const T p6 = (u3p * u2p + u3 * u2p - u3 * u2 - u3m * u2) * r2dyf;    // d/dy(U3 U2)
            const T sp = u2 + u2p, sm = u2 + u2m;
            const T p7 = (T(0.25) * (sp * sp) - T(0.25) * (sm * sm)) * rdy;      // d/dy(U2 U2)
            fq[0 * nx + x] = make_cx<T>(p1, p5);
            fq[1 * nx + x] = make_cx<T>(p3, p7);
            fq[2 * nx + x] = make_cx<T>(p0, p6);
            fq[3 * nx + x] = make_cx<T>(p4, p2);
            if constexpr (NTH >= 1) {
                const T t0 = b0.im, tm = bm.im, tp = bp.im;
                const T p8 = t0 * u1;
                const T p9 = t0 * u3;
                const T p10 = (tp * u2p + t0 * u2p - t0 * u2 - tm * u2) * r2dyf;
                fq[4 * nx + x] = make_cx<T>(p8, p10);
                fq[5 * nx + x] = make_cx<T>(p9, T(0));
            }
        }
    };

    const bool dif_row = p < npi;
    const bool dit_row = p >= npi && p < npi + npo;
    // Prologue: planes ja-1, ja (rows 0..2 npi-1) through every backward
    // stage, plane ja+1 (rows 0..npi-1) through stage 0.
    const bool pro = p < 2 * npi;
    const int pro_plane = ja - 1 + (p >= npi ? 1 : 0), pro_pair = p < npi ? p : p - npi;
    if (pro) fetch(pro_plane, pro_pair);
    __syncthreads();  // twiddles, wavenumbers
    if (pro) dif0(slot_row(pro_plane, pro_pair));
    __syncthreads();
    stage_loop<CP>(1, S, [&](int s) {
        if (pro) ip_stage_row<T, E, +1, true>(slot_row(pro_plane, pro_pair), fx, s, tw, th);
        __syncthreads();
    });
    if (dif_row) {
        fetch(ja + 1, p);
        dif0(slot_row(ja + 1, p));
    }
    __syncthreads();
    // With T dividing 64 every transform row lies inside one wavefront, and a
    // row's stages exchange data only among its own T threads: between its
    // stages the wavefront's LDS operations need only stay in order (a
    // wavefront executes its LDS operations in order; the fences keep the
    // compiler from moving them across).  Cross-wavefront hand-overs (the
    // product rows, the window slots read by every thread) keep the block
    // barriers.
    const bool wave_rows = (64 % TT) == 0;
    auto row_sync = [&]() {
        if (wave_rows) wave_lds_sync();
        else __syncthreads();
    };
    // Iteration j: stages 1..S-1 of plane j+1's backward and the forward of
    // plane j-1's products (S phases), the products of plane j, then stage 0
    // of plane j+2's backward from registers fetched during the products.
    // Slot (j+2)%3 held plane j-1, free once the products of plane j are done
    // (the block barrier after them); plane j+2's backward then runs on the
    // wavefront that wrote its stage 0, and every other reader of the slot
    // (the products of plane j+1) comes after the next block barrier.
    for (int j = ja; j <= jb + 1; ++j) {
        // Keeps the compiler from hoisting the per-thread index arithmetic of
        // every stage out of the sweep, which would hold it in registers
        // across the whole loop: recomputing it is cheaper than the spills.
        asm volatile("" : "+v"(th));
        const bool inv = j <= jb && dif_row;  // backward of plane j+1
        const bool fwd = j > ja && dit_row;   // forward of the products of plane j-1
        cx<T> *drow = inv ? slot_row(j + 1, p) : win;
        cx<T> *frow = fb + (dit_row ? p - npi : 0) * nx;
        stage_loop<CP>(0, S - 1, [&](int s) {
            // FP32: once more per stage (its 128 VGPRs spill less: 0.645 ->
            // 0.622 ms at 384; FP64 measured no better)
            if constexpr (sizeof(T) == 4) asm volatile("" : "+v"(th));
            if (inv) ip_stage_row<T, E, +1, true>(drow, fx, s + 1, tw, th);
            if (fwd) ip_stage_row<T, E, -1, false>(frow, fx, S - 1 - s, tw, th);
            row_sync();
        });
        if (fwd) dit_out(frow, p - npi, j - 1);
        __syncthreads();
        if (j <= jb) {
            const bool nxt = dif_row && j + 2 <= jb + 1;
            if (nxt) fetch(j + 2, p);
            products(j);
            __syncthreads();
            if (nxt) dif0(slot_row(j + 2, p));
            row_sync();
        }
    }
}

// Which kernels take the z-row count at run time (MZ).  The long-row plans
// (T = 32: the 384 and 256 ones) always run one row per block; their FP32
// kernels are compiled for exactly one (the run-time form cost them 9% at
// 384), their FP64 ones keep the run-time form (6% faster at 384, 32 planes
// per block: the same work, scheduled differently by the compiler).
template <typename T>
constexpr bool xf_multi(int pid)
{
    return sizeof(T) == 8 || !(pid == 13 || pid == 17);
}

// FP32 with the 384 plan: four resident blocks per CU (LDS holds four FP32
// windows at NX = 384, 40 KB each) at the cost of a few spilled registers;
// FP64 stays at the two its LDS allows.
template <typename T, int E, int NTH, int PID>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && PID == 17 ? 4 : 2)))
k_xfused_ip(IpLen fx, const cx<T> *__restrict__ twg, int mk, int ld,
                                                    const cx<T> *__restrict__ zt, long zt_fs, long zt_ps,
                                                    cx<T> *__restrict__ wt, long wt_fs, long wt_ps,
                                                    const double *__restrict__ dy, const double *__restrict__ dyf,
                                                    const double *__restrict__ kxv, int jlo, int jhi, int chunk,
                                                    int zr, int nz)
{
    xfused_body<T, E, NTH, PID != 0, xf_multi<T>(PID)>(plan_ref<PID>(fx), twg, mk, ld, zt, zt_fs, zt_ps, wt, wt_fs, wt_ps,
                                                     dy, dyf, kxv, jlo, jhi, chunk, blockIdx.x, blockIdx.y, zr, nz);
}

// grid (z-row groups, chunks of the tallest member, members)
template <typename T, int E, int NTH, int PID>
__global__ void __launch_bounds__(512) k_xfused_ip_batch(const XfDesc<T> *__restrict__ d, IpLen fx,
                                                          const cx<T> *__restrict__ twg, int mk, int ld, long ps,
                                                          int chunk, int zr, int nz)
{
    const XfDesc<T> D = d[blockIdx.z];
    if ((int)blockIdx.y * chunk > D.jhi - D.jlo) return;
    xfused_body<T, E, NTH, PID != 0, xf_multi<T>(PID)>(plan_ref<PID>(fx), twg, mk, ld, D.zt, D.zt_fs, ps, D.wt, D.wt_fs,
                                                    ps, D.dy, D.dyf, D.kx, D.jlo, D.jhi, chunk, blockIdx.x, blockIdx.y, zr,
                                                    nz);
}

// -------------------------------------------------------------------------
// vmax: COURANT's velocity maxima (for/courant.f:35-44).  x backward (DIF) of
// the z-backward rows of CU1, CU2, CU3, packed as the pairs (u1, u2) and
// (u3, 0), for planes 1..NY; block (bx, j) takes rb z rows of plane j, thread
// row p = 2 * row + pair, and writes max |u1|, |u2|, |u3| over its points to
// part[((j - 1) * gridDim.x + bx) * 3 + q].  A maximum is exact and does not
// depend on the order of the comparisons.  The last stage never goes back to
// LDS: its outputs go straight into the maxima.
// -------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T absmax_(T m, T v)
{
    const T a = v < T(0) ? -v : v;
    return a > m ? a : m;
}

template <typename T, int E>
__global__ void __launch_bounds__(256) k_vmax_ip(IpLen fx, const cx<T> *__restrict__ twg, int mk, int ld,
                                                  const cx<T> *__restrict__ zt, long zt_fs, long zt_ps, int nz,
                                                  int rb, double *__restrict__ part)
{
    extern __shared__ unsigned char smem_raw[];
    __shared__ T red[3][256];
    const int nx = fx.n;
    cx<T> *rows = reinterpret_cast<cx<T> *>(smem_raw);  // [2 rb][nx]
    cx<T> *tw = rows + 2 * rb * nx;
    const int tid = threadIdx.x;
    for (int e = tid; e < fx.twn; e += 256) tw[e] = twg[e];
    const int TT = fx.t, S = fx.nstages;
    const int p = tid / TT, th = tid - p * TT;
    const int pair = p & 1, zp = blockIdx.x * rb + (p >> 1);
    const bool act = p < 2 * rb && zp < nz;
    const int j = 1 + blockIdx.y;
    cx<T> *row = rows + (act ? p : 0) * nx;
    __syncthreads();  // twiddles
    if (act) {
        const cx<T> *fa = zt + (2 * pair) * zt_fs + (long)j * zt_ps + (long)zp * ld;
        const cx<T> *fb = zt + zt_fs + (long)j * zt_ps + (long)zp * ld;
        with_radix<E>(fx.radix[0], [&](auto rc) {
            constexpr int R = decltype(rc)::value;
            ip_stage<T, E, R, +1, true>(
                fx, 0, tw, th,
                [&](int, int f) {
                    // spectrum of A + iB with the Hermitian halves of both
                    // rows; C2R ignores Im of the DC term
                    const int kk = f < mk ? f : (f > nx - mk ? nx - f : -1);
                    cx<T> A = fa[max(kk, 0)];
                    cx<T> B = fb[max(kk, 0)];
                    if (pair) B = make_cx<T>(0, 0);
                    if (kk < 0) {
                        A = make_cx<T>(0, 0);
                        B = make_cx<T>(0, 0);
                    }
                    if (f == 0) {
                        A.im = T(0);
                        B.im = T(0);
                    }
                    if (f > nx - mk) {
                        A.im = -A.im;
                        B.im = -B.im;
                    }
                    return make_cx<T>(A.re - B.im, A.im + B.re);
                },
                [&](int, int pos, cx<T> v) { row[ip_swz(fx, pos)] = v; });
        });
    }
    __syncthreads();
    for (int s = 1; s < S - 1; ++s) {
        if (act) ip_stage_row<T, E, +1, true>(row, fx, s, tw, th);
        __syncthreads();
    }
    T ma = T(0), mb = T(0);
    if (act) {
        with_radix<E>(fx.radix[S - 1], [&](auto rc) {
            constexpr int R = decltype(rc)::value;
            ip_stage<T, E, R, +1, true>(
                fx, S - 1, tw, th, [&](int, int pos) { return row[ip_swz(fx, pos)]; },
                [&](int, int, cx<T> v) {
                    ma = absmax_(ma, v.re);
                    mb = absmax_(mb, v.im);
                });
        });
    }
    red[0][tid] = (act && pair == 0) ? ma : T(0);
    red[1][tid] = (act && pair == 0) ? mb : T(0);
    red[2][tid] = (act && pair == 1) ? ma : T(0);
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
#pragma unroll
            for (int q = 0; q < 3; ++q) red[q][tid] = red[q][tid + h] > red[q][tid] ? red[q][tid + h] : red[q][tid];
        }
        __syncthreads();
    }
    if (tid < 3) part[((long)(j - 1) * gridDim.x + blockIdx.x) * 3 + tid] = (double)red[tid][0];
}

// Per plane j = 1..NY: out[q * NY + j - 1] = max over the plane's partials.
__global__ void __launch_bounds__(256) k_vmax_planes(const double *__restrict__ part, int nparts, int ny,
                                                      double *__restrict__ out)
{
    __shared__ double red[3][256];
    const int tid = threadIdx.x, j = blockIdx.x;
    double m[3] = {0.0, 0.0, 0.0};
    for (int b = tid; b < nparts; b += 256) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double v = part[((long)j * nparts + b) * 3 + q];
            m[q] = v > m[q] ? v : m[q];
        }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) red[q][tid] = m[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
#pragma unroll
            for (int q = 0; q < 3; ++q) red[q][tid] = red[q][tid + h] > red[q][tid] ? red[q][tid + h] : red[q][tid];
        }
        __syncthreads();
    }
    if (tid < 3) out[(long)tid * ny + j] = red[tid][0];
}

// -------------------------------------------------------------------------
// launchers
// -------------------------------------------------------------------------

// z rows per block: short rows (few threads per transform) are packed until
// a block has 256 threads, or until 64 KB of LDS
template <typename T, int NTH>
static int xfused_zrows(const IpLen &fx, int nz, int mk)
{
    using C = XCount<NTH>;
    const int per = fx.t * (C::npi + C::npo);
    const size_t row = sizeof(cx<T>) * (size_t)fx.n * (3 * C::npi + C::npo);
    const size_t fixed = sizeof(cx<T>) * fx.twn + sizeof(T) * mk;
    static const bool one = [] {
        const char *e = std::getenv("DGPU_XZR");
        return e && std::atoi(e) == 1;
    }();
    int zr = 1;
    if (one) return zr;
    while (2 * zr * per <= 256 && 2 * zr <= nz && fixed + 2 * zr * row <= 64 * 1024) zr *= 2;
    return zr;
}

template <typename T, int E, int NTH, int PID>
static void xfused_ip_e(hipStream_t s, const Geo &g, const IpLen &fx, const cx<T> *tw, int ld, const cx<T> *zt,
                        long zt_fs, cx<T> *wt, long wt_fs, const double *dy, const double *dyf, const double *kx,
                        int jlo, int jhi, int chunk)
{
    using C = XCount<NTH>;
    const int nchunks = (jhi - jlo + 1 + chunk - 1) / chunk;
    const int zr = xfused_zrows<T, NTH>(fx, g.nz, g.mk);
    const int threads = zr * fx.t * (C::npi + C::npo);
    const size_t lds =
        sizeof(cx<T>) * ((size_t)zr * fx.n * (3 * C::npi + C::npo) + fx.twn) + sizeof(T) * g.mk;
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_xfused_ip<T, E, NTH, PID>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    const long ps = (long)g.nz * ld;
    hipLaunchKernelGGL((k_xfused_ip<T, E, NTH, PID>), dim3((g.nz + zr - 1) / zr, nchunks), dim3(threads), lds, s, fx,
                       tw, g.mk, ld, zt, zt_fs, ps, wt, wt_fs, ps, dy, dyf, kx, jlo, jhi, chunk, zr, g.nz);
}


// The single-run x kernel is built for the 384 and 256 plans (the ensembles
// use the batched kernel) with one scalar; other shapes run on runtime plans.
template <typename T>
void launch_xfused_ip(hipStream_t s, const Geo &g, const IpLen &fx, const cx<T> *tw, int ld, const cx<T> *zt,
                      long zt_fs, cx<T> *wt, long wt_fs, const double *dy, const double *dyf, const double *kx,
                      int jlo, int jhi, int chunk)
{
    if (g.nth == 1 && with_pid(const_plan_id(fx), XPidsSingle<T>{}, [&](auto pc) {
            constexpr int P = decltype(pc)::value;
            xfused_ip_e<T, kConstPlanSpecs[P].e, 1, P>(s, g, fx, tw, ld, zt, zt_fs, wt, wt_fs, dy, dyf, kx, jlo, jhi,
                                                       chunk);
        }))
        return;
#define DGPU_XF(EE_)                                                                                             \
    do {                                                                                                         \
        if (g.nth == 0)                                                                                          \
            xfused_ip_e<T, EE_, 0, 0>(s, g, fx, tw, ld, zt, zt_fs, wt, wt_fs, dy, dyf, kx, jlo, jhi, chunk);     \
        else                                                                                                     \
            xfused_ip_e<T, EE_, 1, 0>(s, g, fx, tw, ld, zt, zt_fs, wt, wt_fs, dy, dyf, kx, jlo, jhi, chunk);     \
    } while (0)
    switch (fx.e) {
    case 4: DGPU_XF(4); break;
    case 6: DGPU_XF(6); break;
    case 8: DGPU_XF(8); break;
    case 12: DGPU_XF(12); break;
    case 16:
        if constexpr (sizeof(T) == 4) DGPU_XF(16);  // FP64 plans stop at 12 (16 spills)
        break;
    default: break;
    }
#undef DGPU_XF
}

template <typename T, int E, int NTH, int PID>
static void xfused_ip_batch_e(hipStream_t s, const Geo &g, const IpLen &fx, const cx<T> *tw, const XfDesc<T> *d,
                              int nmem, int max_span, int ld, int chunk)
{
    using C = XCount<NTH>;
    const int zr = xfused_zrows<T, NTH>(fx, g.nz, g.mk);
    const int threads = zr * fx.t * (C::npi + C::npo);
    const size_t lds =
        sizeof(cx<T>) * ((size_t)zr * fx.n * (3 * C::npi + C::npo) + fx.twn) + sizeof(T) * g.mk;
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_xfused_ip_batch<T, E, NTH, PID>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    dim3 grid((g.nz + zr - 1) / zr, (max_span + chunk - 1) / chunk, nmem);
    hipLaunchKernelGGL((k_xfused_ip_batch<T, E, NTH, PID>), grid, dim3(threads), lds, s, d, fx, tw, g.mk, ld,
                       (long)g.nz * ld, chunk, zr, g.nz);
}

template <typename T>
void launch_xfused_ip_batch(hipStream_t s, const Geo &g, const IpLen &fx, const cx<T> *tw, const XfDesc<T> *d,
                            int nmem, int max_span, int ld, int chunk)
{
    if (g.nth == 1 && with_pid(const_plan_id(fx), XPids<T>{}, [&](auto pc) {
            constexpr int P = decltype(pc)::value;
            xfused_ip_batch_e<T, kConstPlanSpecs[P].e, 1, P>(s, g, fx, tw, d, nmem, max_span, ld, chunk);
        }))
        return;
#define DGPU_XFB(EE_)                                                                               \
    do {                                                                                            \
        if (g.nth == 0)                                                                             \
            xfused_ip_batch_e<T, EE_, 0, 0>(s, g, fx, tw, d, nmem, max_span, ld, chunk);            \
        else                                                                                        \
            xfused_ip_batch_e<T, EE_, 1, 0>(s, g, fx, tw, d, nmem, max_span, ld, chunk);            \
    } while (0)
    switch (fx.e) {
    case 4: DGPU_XFB(4); break;
    case 6: DGPU_XFB(6); break;
    case 8: DGPU_XFB(8); break;
    case 12: DGPU_XFB(12); break;
    case 16:
        if constexpr (sizeof(T) == 4) DGPU_XFB(16);
        break;
    default: break;
    }
#undef DGPU_XFB
}

int vmax_rows_per_block(const IpLen &fx) { return 256 / (2 * fx.t); }

size_t vmax_part_count(const Geo &g, const IpLen &fx)
{
    const int rb = vmax_rows_per_block(fx);
    return rb > 0 ? (size_t)g.ny * ((g.nz + rb - 1) / rb) * 3 : 0;
}

template <typename T, int E>
static void vmax_ip_e(hipStream_t s, const Geo &g, const IpLen &fx, const cx<T> *tw, int ld, const cx<T> *zt,
                      long zt_fs, double *part, double *out)
{
    const int rb = vmax_rows_per_block(fx);
    const int nb = (g.nz + rb - 1) / rb;
    const size_t lds = sizeof(cx<T>) * ((size_t)2 * rb * fx.n + fx.twn);
    // above 64 KiB of dynamic LDS the kernel must be allowed it (its static
    // reduction array counts against the same 160 KiB); a refusal is left as
    // the launch error
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vmax_ip<T, E>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const long ps = (long)g.nz * ld;
    hipLaunchKernelGGL((k_vmax_ip<T, E>), dim3(nb, g.ny), dim3(256), lds, s, fx, tw, g.mk, ld, zt, zt_fs, ps, g.nz,
                       rb, part);
    hipLaunchKernelGGL(k_vmax_planes, dim3(g.ny), dim3(256), 0, s, part, nb, g.ny, out);
}

template <typename T>
void launch_vmax_ip(hipStream_t s, const Geo &g, const IpLen &fx, const cx<T> *tw, int ld, const cx<T> *zt,
                    long zt_fs, double *part, double *out)
{
    DGPU_DISPATCH_IPE(fx.e, (vmax_ip_e<T, EE>(s, g, fx, tw, ld, zt, zt_fs, part, out)));
}

template void launch_vmax_ip<double>(hipStream_t, const Geo &, const IpLen &, const cx<double> *, int,
                                     const cx<double> *, long, double *, double *);
template void launch_vmax_ip<float>(hipStream_t, const Geo &, const IpLen &, const cx<float> *, int,
                                    const cx<float> *, long, double *, double *);
template void launch_xfused_ip_batch<double>(hipStream_t, const Geo &, const IpLen &, const cx<double> *,
                                             const XfDesc<double> *, int, int, int, int);
template void launch_xfused_ip_batch<float>(hipStream_t, const Geo &, const IpLen &, const cx<float> *,
                                            const XfDesc<float> *, int, int, int, int);
template void launch_xfused_ip<double>(hipStream_t, const Geo &, const IpLen &, const cx<double> *, int,
                                       const cx<double> *, long, cx<double> *, long, const double *, const double *,
                                       const double *, int, int, int);
template void launch_xfused_ip<float>(hipStream_t, const Geo &, const IpLen &, const cx<float> *, int,
                                      const cx<float> *, long, cx<float> *, long, const double *, const double *,
                                      const double *, int, int, int);
}  // namespace dgpu
