// dgpu_colseg.hip -- the per-mode column stage of RK_CHAN_1 with each
// y-column split into segments (for/channel.f:88-290, 581-1029 for the
// Crank-Nicolson part, for/channel.f:1103-1214 and 1059-1065 for REM_DIV_CHAN
// and the CP update).
//
// A tridiagonal solve by elimination is two linear recurrences over the rows:
// g(j) = r(j) + m(j) g(j-1) forward, x(j) = (g(j) - U(j) x(j+1)) / B(j)
// backward.  A recurrence over a segment [a, b] of rows started from zero
// differs from the true one by its carry times a product of its
// coefficients:
//   g(j) = gl(j) + C(j) g(a-1),   C(j) = m(a) ... m(j),
//   x(j) = xl(j) + E(j) x(b+1),   E(j) = (-U(j)/B(j)) ... (-U(b)/B(b)).
// So every segment runs its rows locally (a lane holds its segment's rows in
// registers), the carries pass from segment to segment through LDS (one
// product and one addition per segment), and each lane corrects its rows.
// The solve's intermediate values never leave the chip, and the sequential
// chain of a lane is one segment (9 to 17 rows) long instead of a column.
//
// Block shape: NW waves (4, 8 or 16; a runtime choice, dgpu_set_column_shape)
// of 64 lanes; a lane is (mode, segment), MPW modes side by side so that a
// wave's loads of one row are MPW consecutive modes.  The segment count of
// k_cn_seg / k_remdiv_seg is fixed per precision (32 in FP64: rows per
// segment <= 9 up to NY = 286; 16 in FP32), so MPW = 64 NW / NSEG; the fused
// kernel has 30 segments at every shape.  The kernels ask for four waves per
// SIMD (<= 128 VGPRs): one 16-wave block per CU, or two / four smaller
// blocks whose phases (loads, carries, stores) overlap.
//
// k_cn_seg: one block per (MPW modes, system): the explicit right-hand
//   side, the new CF, the CN solve; the solution of every row goes to the
//   system's solve scratch, which holds the complete new field afterwards
//   (rows outside the solve's copy range keep the old value).  The state's
//   own fields are only read: the other systems' blocks read them as
//   explicit terms (Coriolis, buoyancy, background gradients).
// k_remdiv_seg: one block per mode group: the divergence of the new
//   velocity (from the scratch), the Poisson solve with the per-mode
//   factorisation, the projection into CU1..3, the CP update, and the copy
//   of the new scalar into CTH.
#include <algorithm>

#include "dgpu_kernels.hpp"

namespace dgpu {

// Lane -> (segment, mode) in a block of NW waves with MPW modes side by side
// (a segment's lanes read MPW consecutive modes of a row): NSEG = 64 NW / MPW
// segments.  When MPW does not divide 64 NW, the lanes past NSEG * MPW repeat
// lanes of the last segment: the same inputs, the same work, the same values
// stored to the same places.
constexpr int seg_count(int mpw, int nw) { return 64 * nw / mpw; }
template <int MPW, int NSEG>
__device__ __forceinline__ void seg_lane(int tid, int &seg, int &ml)
{
    const int s = tid / MPW;
    ml = tid - s * MPW;
    seg = min(s, NSEG - 1);
}

// The fused column kernel's segments per column, set by the column height
// and the same at every block shape (so 4, 8 and 16 waves give the same
// bits), at most 9 rows each: 16 up to NY = 142 (modes per block 16, 32, 64
// at 4, 8, 16 waves), 22 up to NY = 196 (11, 23, 46), else 30 (NY <= 268;
// 8, 17, 34).  The lanes past nseg * mpw repeat lanes of the last segment.
// The fewest segments a column fits in means the most modes per block:
// short columns on 30 segments would leave up to 40% of the lanes without
// rows, and fewer, fuller blocks take fewer rounds of resident blocks.
constexpr int kFusedSeg = 30, kFusedSegMid = 22, kFusedSegShort = 16;
constexpr int fused_nseg(int ny)
{
    return ny + 2 <= 9 * kFusedSegShort ? kFusedSegShort : ny + 2 <= 9 * kFusedSegMid ? kFusedSegMid : kFusedSeg;
}
constexpr int fused_mpw(int nw, int nseg) { return 64 * nw / nseg; }
// the segment count of a fused kernel built for (MPW, NW)
constexpr int fused_nseg_of(int mpw, int nw)
{
    return mpw == fused_mpw(nw, kFusedSeg) ? kFusedSeg : mpw == fused_mpw(nw, kFusedSegMid) ? kFusedSegMid
                                                                                            : kFusedSegShort;
}
static_assert(fused_mpw(4, kFusedSeg) != fused_mpw(4, kFusedSegMid) && fused_mpw(8, kFusedSeg) != fused_mpw(8, kFusedSegMid) &&
                  fused_mpw(16, kFusedSeg) != fused_mpw(16, kFusedSegMid) &&
                  fused_mpw(4, kFusedSegShort) != fused_mpw(4, kFusedSegMid) &&
                  fused_mpw(8, kFusedSegShort) != fused_mpw(8, kFusedSegMid) &&
                  fused_mpw(16, kFusedSegShort) != fused_mpw(16, kFusedSegMid),
              "each segment count has its own modes per block");

namespace {

template <typename T>
__device__ __forceinline__ cx<T> cfma(T a, cx<T> x, cx<T> y) { return make_cx<T>(y.re + a * x.re, y.im + a * x.im); }

// table entries kept in LDS, per row
enum { L_MULT = 0, L_INVPIV, L_UP, L_BC00, L_ISBC, L_WP, L_W0, L_WM, L_CPW, L_B0, L_BM, L_CFR, L_CPY, kLRow };
constexpr int kTabMap[kLRow] = {T_MULT, T_INVPIV, T_UP, T_BC00, T_ISBC, T_WP, T_W0, T_WM,
                                T_CPW,  T_B0,     T_BM, T_CFR,  T_CPY};

__device__ __forceinline__ int xcd_block(int L, int nb8) { return (L & 7) * (nb8 >> 3) + (L >> 3); }

}  // namespace

// ---------------------------------------------------------------------------
// k_cn_seg
// ---------------------------------------------------------------------------
// Constants of one mode's CN right-hand side (for/channel.f:88-290, 581-1029)
template <typename T>
struct CnMode {
    T kx, kz, pw, temp3, temp4, temp5, i_ro, rgx, rgz, drhodx, drhodz, force00, nu, nupr;
    bool is00, rk_gt1;
    int f_type;
};

template <typename T>
__device__ __forceinline__ CnMode<T> cn_mode(const CnArgs &A, bool live, int i, int s)
{
    CnMode<T> K;
    K.is00 = live && i == 0 && s == 0;
    K.kx = live ? T(A.kx[i]) : T(0);
    K.kz = live ? T(A.kz[s]) : T(0);
    const T k2 = live ? T(A.kx2[i]) + T(A.kz2[s]) : T(0);
    K.pw = (A.beta == 1.0) ? k2 : T(pow(double(k2), A.beta));
    K.temp3 = T(A.temp3);
    K.temp4 = T(A.temp4);
    K.temp5 = T(A.temp5);
    K.rk_gt1 = A.rk_gt1 != 0;
    K.f_type = A.f_type;
    K.i_ro = T(A.i_ro);
    K.rgx = T(A.rgx);
    K.rgz = T(A.rgz);
    K.drhodx = T(A.drhodx);
    K.drhodz = T(A.drhodz);
    K.force00 = T(A.force00);
    K.nu = T(A.nu);
    K.nupr = T(A.nupr);
    return K;
}

// One row of system q: the new explicit term CF (cfn) and the right-hand
// side of the CN solve (vec).  tb(k) reads the row's table entry L_k; f0, fm,
// fp: the system's field at j, j-1, j+1 (fp zero above NY); aux1, aux2: the
// partner fields at j (Coriolis / TH advection, buoyancy / TH advection);
// thm: CTH at j-1 (U2's buoyancy); cp0, cpm: CP at j, j-1; nlv, cfo: the
// nonlinear term and the old CF (zero on BC rows outside the CF range).
template <typename T, int NTH, typename TB>
__device__ __forceinline__ void cn_rhs(int q, const CnMode<T> &K, TB &&tb, cx<T> f0, cx<T> fm, cx<T> fp, cx<T> aux1,
                                       cx<T> aux2, cx<T> thm, cx<T> cp0, cx<T> cpm, cx<T> nlv, cx<T> cfo,
                                       cx<T> &cfn, cx<T> &vec)
{
    const cx<T> zero = make_cx<T>(0, 0);
    const T visc = (q == SYS_TH0) ? K.nupr * K.pw : K.nu * K.pw;
    cfn = make_cx<T>(-(visc * f0.re), -(visc * f0.im));
    if (q == SYS_U1) {
        cfn = cfma(K.i_ro, aux1, cfn);
        if (NTH) cfn = cfma(K.rgx, aux2, cfn);
    } else if (q == SYS_U3) {
        cfn = cfma(-K.i_ro, aux1, cfn);
        if (NTH) cfn = cfma(K.rgz, aux2, cfn);
    } else if (q == SYS_U2) {
        if (NTH && !K.is00) {
            const T b0 = tb(L_B0), bm = tb(L_BM);
            cfn = cfn + make_cx<T>(b0 * aux2.re + bm * thm.re, b0 * aux2.im + bm * thm.im);
        }
    } else {
        cfn = cfn - make_cx<T>(aux1.re * K.drhodx + aux2.re * K.drhodz, aux1.im * K.drhodx + aux2.im * K.drhodz);
    }
    cfn = cfn + nlv;
    if (tb(L_ISBC) != T(0)) {
        vec = K.is00 ? make_cx<T>(tb(L_BC00), 0) : zero;
    } else {
        cx<T> cr = f0;
        if (K.rk_gt1) cr = cfma(K.temp3, cfo, cr);
        if (q == SYS_U1) {
            cr = cr - make_cx<T>(-K.temp4 * (K.kx * cp0.im), K.temp4 * (K.kx * cp0.re));
            if (K.is00 && (K.f_type == 1 || K.f_type == 2)) cr.re = cr.re - K.temp4 * K.force00;
        } else if (q == SYS_U3) {
            cr = cr - make_cx<T>(-K.temp4 * (K.kz * cp0.im), K.temp4 * (K.kz * cp0.re));
        } else if (q == SYS_U2) {
            const T w = tb(L_CPW);
            cr = cr - make_cx<T>(w * (cp0.re - cpm.re), w * (cp0.im - cpm.im));
        }
        cr = cfma(K.temp5, cfn, cr);
        const T wp = tb(L_WP), w0 = tb(L_W0), wm = tb(L_WM);
        vec = make_cx<T>(cr.re + (wp * fp.re + w0 * f0.re + wm * fm.re),
                         cr.im + (wp * fp.im + w0 * f0.im + wm * fm.im));
    }
}

// The CN solve of system q for the block's modes: sol[r] = the new value of
// row a + r of the lane's segment (the rows outside the solve's copy range
// keep the old value).  The LDS holds the system's table and the carries
// (cn_seg_lds); a caller that solves several systems in a row puts a block
// barrier between them.
template <typename T, int NTH, typename NLT, int MPW, int NW, int RMAX, typename AT, typename PT,
          int NSEG = seg_count(MPW, NW)>
__device__ __forceinline__ void cn_seg_solve(AT A, PT P, int bx, int q, cx<T> (&sol)[RMAX])
{
    extern __shared__ unsigned char smem_raw[];
    const Geo &g = A.g;
    const int NY = g.ny, N = NY + 2;
    const long ps = g.ps;
    const int R = (N + NSEG - 1) / NSEG;  // rows per segment, <= RMAX (host check)
    const int nseg = (N + R - 1) / R;
    T *tl = reinterpret_cast<T *>(smem_raw);                               // [N][kLRow]
    cx<T> *GL = reinterpret_cast<cx<T> *>(tl + ((N * kLRow + 1) & ~1));   // [NSEG][MPW]
    cx<T> *XL = GL + NSEG * MPW;                                           // [NSEG][MPW]
    T *Cb = reinterpret_cast<T *>(XL + NSEG * MPW);                        // [NSEG]
    T *Ea = Cb + NSEG;                                                     // [NSEG]

    const int tid = threadIdx.x;
    int seg, ml;
    seg_lane<MPW, NSEG>(tid, seg, ml);
    const long m = (long)bx * MPW + ml;
    const bool live = m < ps;
    const int a = seg * R, b = min(a + R, N) - 1;
    const bool rows = live && a <= b;

    const double *tab = A.tab + (long)q * N * kRow;
    for (int e = tid; e < N * kLRow; e += blockDim.x) {
        const int j = e / kLRow, k = e - j * kLRow;
        tl[e] = T(tab[(long)j * kRow + kTabMap[k]]);
    }

    int s = 0, i = 0;
    const long gm = (live && A.map) ? (long)A.map[m] : m;
    if (live) { s = (int)(gm / g.mk); i = (int)(gm - (long)s * g.mk); }
    const long nps = A.nl_ps;
    const CnMode<T> K = cn_mode<T>(A, live, i, s);
    const bool rk_gt1 = K.rk_gt1;
    cx<T> *fld = (q == SYS_U1) ? P.cu1 : (q == SYS_U3) ? P.cu3 : (q == SYS_U2) ? P.cu2 : P.cth;
    cx<T> *cf = (q == SYS_U1) ? P.cf1 : (q == SYS_U3) ? P.cf3 : (q == SYS_U2) ? P.cf2 : P.cfth;
    const cx<NLT> *nl = P.nl[q];
    const cx<T> *a1p = (q == SYS_U1) ? P.cu3 : P.cu1;   // Coriolis partner / TH advection
    const cx<T> *a2p = (q == SYS_TH0) ? P.cu3 : P.cth;  // buoyancy / TH advection
    const bool has1 = (q != SYS_U2);
    const bool has2 = (q == SYS_TH0) || NTH;
    const bool hascp = (q != SYS_TH0);
    const cx<T> zero = make_cx<T>(0, 0);
    __syncthreads();  // table

    // ---- forward, local: right-hand side, CF, elimination from zero ----
    cx<T> yv[RMAX];
    if (rows) {
        struct In {
            cx<T> fp, nl, cf, aux1, aux2, cp;
        };
        auto load = [&](int j) {
            const long o = m + (long)j * ps;
            In in;
            in.fp = fld[o + (j <= NY ? ps : 0)];
            const cx<NLT> nv = nl[gm + (long)j * nps];
            in.nl = make_cx<T>(T(nv.re), T(nv.im));
            in.cf = rk_gt1 ? cf[o] : zero;
            in.aux1 = has1 ? a1p[o] : zero;
            in.aux2 = has2 ? a2p[o] : zero;
            in.cp = hascp ? P.cp[o] : zero;
            return in;
        };
        const long oa = m + (long)a * ps;
        cx<T> fm = a >= 1 ? fld[oa - ps] : zero, f0 = fld[oa];
        cx<T> thm = (a >= 1 && has2) ? a2p[oa - ps] : zero;
        cx<T> cpm = (a >= 1 && hascp) ? P.cp[oa - ps] : zero;
        cx<T> gl = zero;
        T cprod = T(1);
        In cur = load(a);
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            const int j = a + r;
            if (j <= b) {
                const In nxt = load(min(j + 1, b));
                const T *t = tl + j * kLRow;
                const T mult = t[L_MULT];
                const bool isbc = t[L_ISBC] != T(0);
                const bool cfr = t[L_CFR] != T(0);
                const cx<T> fp = (j <= NY) ? cur.fp : zero;
                const bool use = cfr || !isbc;
                const cx<T> nlv = use ? cur.nl : zero;
                const cx<T> cfo = use ? cur.cf : zero;
                const cx<T> aux2 = cur.aux2, cp0 = cur.cp;
                cx<T> cfn, vec;
                cn_rhs<T, NTH>(q, K, [&](int k) { return t[k]; }, f0, fm, fp, cur.aux1, aux2, thm, cp0, cpm, nlv, cfo,
                               cfn, vec);
                gl = (r == 0) ? vec : cfma(mult, gl, vec);
                cprod = (r == 0) ? mult : cprod * mult;
                yv[r] = gl;
                if (cfr) cf[m + (long)j * ps] = cfn;
                fm = f0;
                f0 = fp;
                thm = aux2;
                cpm = cp0;
                cur = nxt;
            }
        }
        GL[seg * MPW + ml] = gl;
        if (ml == 0) Cb[seg] = cprod;
    }
    __syncthreads();
    // ---- carries of the forward recurrence, back substitution, local ----
    if (rows) {
        cx<T> gin = zero;
        for (int t = 0; t < seg; ++t) gin = cfma(Cb[t], gin, GL[t * MPW + ml]);
        T c = T(1);
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            const int j = a + r;
            if (j <= b) {
                c *= tl[j * kLRow + L_MULT];
                yv[r] = cfma(c, gin, yv[r]);
            }
        }
        cx<T> xl = zero;
        T e = T(1);
        bool first = true;
#pragma unroll
        for (int r = RMAX - 1; r >= 0; --r) {
            const int j = a + r;
            if (j <= b) {
                const T ip = tl[j * kLRow + L_INVPIV], u = tl[j * kLRow + L_UP];
                if (first) {
                    xl = make_cx<T>(yv[r].re * ip, yv[r].im * ip);
                    e = -u * ip;
                    first = false;
                } else {
                    xl = make_cx<T>((yv[r].re - u * xl.re) * ip, (yv[r].im - u * xl.im) * ip);
                    e = (-u * ip) * e;
                }
                yv[r] = xl;
            }
        }
        XL[seg * MPW + ml] = xl;
        if (ml == 0) Ea[seg] = e;
    }
    __syncthreads();
    // ---- carries of the backward recurrence, the solution of every row ----
    if (rows) {
        cx<T> xin = zero;
        for (int t = nseg - 1; t > seg; --t) xin = cfma(Ea[t], xin, XL[t * MPW + ml]);
        T e = T(1);
        bool first = true;
#pragma unroll
        for (int r = RMAX - 1; r >= 0; --r) {
            const int j = a + r;
            if (j <= b) {
                const T *t = tl + j * kLRow;
                const T f = -t[L_UP] * t[L_INVPIV];
                e = first ? f : f * e;
                first = false;
                const long o = m + (long)j * ps;
                sol[r] = (t[L_CPY] != T(0)) ? cfma(e, xin, yv[r]) : fld[o];
            }
        }
    }
}

// k_cn_seg's block: system q's new field into its solve scratch P.gs[q]
template <typename T, int NTH, typename NLT, int MPW, int NW, int RMAX, typename AT, typename PT>
__device__ __forceinline__ void cn_seg_body(AT A, PT P, int bx, int q)
{
    constexpr int NSEG = seg_count(MPW, NW);
    const int N = A.g.ny + 2;
    const int R = (N + NSEG - 1) / NSEG;
    int seg, ml;
    seg_lane<MPW, NSEG>(threadIdx.x, seg, ml);
    const long m = (long)bx * MPW + ml;
    const int a = seg * R, b = min(a + R, N) - 1;
    cx<T> sol[RMAX];
    cn_seg_solve<T, NTH, NLT, MPW, NW, RMAX, AT, PT>(A, P, bx, q, sol);
    if (m < A.g.ps && a <= b) {
        cx<T> *out = P.gs[q];
#pragma unroll
        for (int r = 0; r < RMAX; ++r)
            if (a + r <= b) out[m + (long)(a + r) * A.g.ps] = sol[r];
    }
}

// One flat grid over (mode group, system), the systems of a mode group
// adjacent on one XCD: the fields every system reads as explicit terms
// (CU1, CU3, CTH, CP) come from HBM once per mode group and from that XCD's
// L2 for the other systems' blocks.
template <typename T, int NTH, typename NLT, int MPW, int NW, int RMAX>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4)))
k_cn_seg(CnArgs A, CnPtrs<T, NLT> P)
{
    constexpr int NSYS = 3 + NTH;
    const int L = xcd_block(blockIdx.x, gridDim.x);
    const int q = L % NSYS, bx = L / NSYS;
    if ((long)bx * MPW >= A.g.ps) return;  // padding blocks of the flat grid (whole block)
    cn_seg_body<T, NTH, NLT, MPW, NW, RMAX, const CnArgs &, const CnPtrs<T, NLT> &>(A, P, bx, q);
}

template <typename T, int NTH, typename NLT, int MPW, int NW, int RMAX>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4)))
k_cn_seg_batch(const ColDesc<T, NLT> *__restrict__ d)
{
    const ColDesc<T, NLT> &D = d[blockIdx.z];
    if ((long)blockIdx.x * MPW >= D.A.g.ps) return;
    cn_seg_body<T, NTH, NLT, MPW, NW, RMAX, const CnArgs &, const CnPtrs<T, NLT> &>(D.A, D.P, blockIdx.x,
                                                                                   blockIdx.y);
}

// ---------------------------------------------------------------------------
// k_remdiv_seg
// ---------------------------------------------------------------------------
// Rows of the Poisson solve: 1..NY (row 1 is the gauge / Neumann row with a
// zero right-hand side, row NY the top Neumann row).  The new velocity U*
// and scalar come from the solve scratch (P.gs, slot per system); with
// P.gs[q] equal to the state's own field the call works in place
// (REM_DIV_CHAN on the state).
template <typename T, typename NLT, int MPW, int NW, int RMAX, int NTH>
__device__ __forceinline__ void remdiv_seg_body(const RdArgs &A, const CnPtrs<T, NLT> &P, int bx)
{
    constexpr int NSEG = seg_count(MPW, NW);
    extern __shared__ unsigned char smem_raw[];
    const Geo &g = A.g;
    const int NY = g.ny;
    const long ps = g.ps, fps = A.fac_ps;
    const int R = (NY + NSEG - 1) / NSEG;
    const int nseg = (NY + R - 1) / R;
    cx<T> *GL = reinterpret_cast<cx<T> *>(smem_raw);  // [NSEG][MPW]
    cx<T> *XL = GL + NSEG * MPW;
    T *Cb = reinterpret_cast<T *>(XL + NSEG * MPW);   // [NSEG][MPW]
    T *Ea = Cb + NSEG * MPW;

    const int tid = threadIdx.x;
    int seg, ml;
    seg_lane<MPW, NSEG>(tid, seg, ml);
    const long m = (long)bx * MPW + ml;
    const bool live = m < ps;
    const int a = 1 + seg * R, b = min(a + R - 1, NY);
    const bool rows = live && a <= b;
    int s = 0, i = 0;
    const long gm = (live && A.map) ? (long)A.map[m] : m;
    if (live) { s = (int)(gm / g.mk); i = (int)(gm - (long)s * g.mk); }
    const bool is00 = live && i == 0 && s == 0;
    const T kx = live ? T(A.kx[i]) : T(0), kz = live ? T(A.kz[s]) : T(0);
    const cx<T> *u1 = P.gs[SYS_U1], *u2 = P.gs[SYS_U2], *u3 = P.gs[SYS_U3];
    const cx<T> zero = make_cx<T>(0, 0);
    auto pm = [&](int j) -> T {
        if constexpr (sizeof(T) == 4) return A.pmultf[gm + (long)j * fps];
        else return T(A.pmult[gm + (long)j * fps]);
    };
    auto pv = [&](int j) -> T {
        if constexpr (sizeof(T) == 4) return A.pinvf[gm + (long)j * fps];
        else return T(A.pinv[gm + (long)j * fps]);
    };
    // i k x
    auto ik = [](T k, cx<T> x) { return make_cx<T>(-k * x.im, k * x.re); };

    // ---- forward, local: divergence of U* and the Poisson elimination ----
    cx<T> yv[RMAX];
    T mr[RMAX], fr[RMAX];  // per row: the multiplier, -U / pivot
    if (rows) {
        cx<T> hl = zero;
        T cprod = T(1);
        cx<T> c2 = u2[m + (long)a * ps];
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            const int j = a + r;
            if (j <= b) {
                const long o = m + (long)j * ps;
                const cx<T> c2p = u2[o + ps];
                cx<T> vec = zero;
                if (j >= 2 && j < NY) {
                    // (i kx CU1 + (CU2(j+1) - CU2(j))/DYF(j)) + i kz CU3 (for/channel.f:1139-1141)
                    const T idf = T(A.rd[(long)j * 4 + 1]);
                    const cx<T> d1 = ik(kx, u1[o]), d3 = ik(kz, u3[o]);
                    vec = make_cx<T>((d1.re + (c2p.re - c2.re) * idf) + d3.re, (d1.im + (c2p.im - c2.im) * idf) + d3.im);
                }
                const T mu = j >= 2 ? pm(j) : T(0);
                hl = (r == 0) ? vec : cfma(mu, hl, vec);
                cprod = (r == 0) ? mu : cprod * mu;
                yv[r] = hl;
                mr[r] = mu;
                c2 = c2p;
            }
        }
        GL[seg * MPW + ml] = hl;
        Cb[seg * MPW + ml] = cprod;
    }
    __syncthreads();
    const T u1row = is00 ? T(0) : T(-1);
    if (rows) {
        cx<T> hin = zero;
        for (int t = 0; t < seg; ++t) hin = cfma(Cb[t * MPW + ml], hin, GL[t * MPW + ml]);
        T c = T(1);
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            const int j = a + r;
            if (j <= b) {
                c *= mr[r];
                yv[r] = cfma(c, hin, yv[r]);
            }
        }
        cx<T> xl = zero;
        T e = T(1);
        bool first = true;
#pragma unroll
        for (int r = RMAX - 1; r >= 0; --r) {
            const int j = a + r;
            if (j <= b) {
                const T ip = pv(j);
                const T u = (j == 1) ? u1row : T(A.rd[(long)j * 4 + 0]);
                fr[r] = -u * ip;
                if (first) {
                    xl = make_cx<T>((yv[r].re) * ip, (yv[r].im) * ip);
                    e = fr[r];
                    first = false;
                } else {
                    xl = make_cx<T>((yv[r].re - u * xl.re) * ip, (yv[r].im - u * xl.im) * ip);
                    e = fr[r] * e;
                }
                yv[r] = xl;
            }
        }
        XL[seg * MPW + ml] = xl;
        Ea[seg * MPW + ml] = e;
    }
    __syncthreads();
    // ---- phi of every row, projection, CP update, copies ----
    if (rows) {
        cx<T> xin = zero;
        for (int t = nseg - 1; t > seg; --t) xin = cfma(Ea[t * MPW + ml], xin, XL[t * MPW + ml]);
        const T itemp4 = T(A.inv_temp4);
        const bool upd = A.update_cp;
        const bool copy = P.gs[SYS_U1] != P.cu1;
        cx<T> phip = xin;  // phi(b+1)
        T e = T(1);
        bool first = true;
#pragma unroll
        for (int r = RMAX - 1; r >= 0; --r) {
            const int j = a + r;
            if (j <= b) {
                const long o = m + (long)j * ps;
                e = first ? fr[r] : fr[r] * e;
                first = false;
                const cx<T> phi = cfma(e, xin, yv[r]);
                // CU1 - i kx phi, CU3 - i kz phi (rows 1..NY), CU2(j+1) - (phi(j+1) - phi(j))/DY(j+1)
                P.cu1[o] = u1[o] - ik(kx, phi);
                P.cu3[o] = u3[o] - ik(kz, phi);
                if (j + 1 <= NY) {
                    const T idy = T(A.rd[(long)(j + 1) * 4 + 2]);
                    const cx<T> v = u2[o + ps];
                    P.cu2[o + ps] = make_cx<T>(v.re - (phip.re - phi.re) * idy, v.im - (phip.im - phi.im) * idy);
                }
                if (upd && j >= A.jstart && j <= A.jend) {
                    const cx<T> p0 = P.cp[o];
                    P.cp[o] = make_cx<T>(p0.re + phi.re * itemp4, p0.im + phi.im * itemp4);
                }
                if (NTH && copy) P.cth[o] = P.gs[SYS_TH0][o];
                phip = phi;
            }
        }
        if (copy) {
            // rows the projection does not touch: CU1, CU3 at 0 and NY+1, CU2 at
            // 0, 1 and NY+1, the scalar at 0 and NY+1
            if (seg == 0) {
                const long o0 = m, o1 = m + ps;
                P.cu1[o0] = u1[o0];
                P.cu3[o0] = u3[o0];
                P.cu2[o0] = u2[o0];
                P.cu2[o1] = u2[o1];
                if (NTH) P.cth[o0] = P.gs[SYS_TH0][o0];
            }
            if (b == NY) {
                const long o = m + (long)(NY + 1) * ps;
                P.cu1[o] = u1[o];
                P.cu3[o] = u3[o];
                P.cu2[o] = u2[o];
                if (NTH) P.cth[o] = P.gs[SYS_TH0][o];
            }
        }
    }
}

template <typename T, typename NLT, int MPW, int NW, int RMAX, int NTH>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4)))
k_remdiv_seg(RdArgs A, CnPtrs<T, NLT> P)
{
    remdiv_seg_body<T, NLT, MPW, NW, RMAX, NTH>(A, P, xcd_block(blockIdx.x, gridDim.x));
}

template <typename T, typename NLT, int MPW, int NW, int RMAX, int NTH>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4)))
k_remdiv_seg_batch(const ColDesc<T, NLT> *__restrict__ d)
{
    const ColDesc<T, NLT> &D = d[blockIdx.z];
    if ((long)blockIdx.x * MPW >= D.R.g.ps) return;
    remdiv_seg_body<T, NLT, MPW, NW, RMAX, NTH>(D.R, D.P, blockIdx.x);
}

// The CN solves of U1, U3 and U2 together, row by row (the fused column
// kernel's first pass): each row's CU1..3, CTH, CP, NL and CF are loaded
// once for the three systems (the partner fields of the explicit terms are
// the other systems' own rows), the three tables sit in LDS, and the three
// recurrences run side by side.  Carries through
// LDS as in cn_seg_solve, three systems per barrier.  u1s, u3s, u2s: the new
// values of the lane's rows (the rows outside a solve's copy range keep the
// old value).
template <typename T, int NTH, typename NLT, int MPW, int NW, int RMAX>
__device__ __forceinline__ void cn3_solve(const CnArgs &A, const CnPtrs<T, NLT> &P, int bx, cx<T> (&u1s)[RMAX],
                                          cx<T> (&u3s)[RMAX], cx<T> (&u2s)[RMAX])
{
    constexpr int NSEG = fused_nseg_of(MPW, NW);
    constexpr int NM = NSEG * MPW;
    extern __shared__ unsigned char smem_raw[];
    const Geo &g = A.g;
    const int NY = g.ny, N = NY + 2;
    const long ps = g.ps;
    const int R = (N + NSEG - 1) / NSEG;
    const int nseg = (N + R - 1) / R;
    cx<T> *GL3 = reinterpret_cast<cx<T> *>(smem_raw);  // [3][NSEG][MPW]
    cx<T> *XL3 = GL3 + 3 * NM;                          // [3][NSEG][MPW]
    T *Cb3 = reinterpret_cast<T *>(XL3 + 3 * NM);       // [3][NSEG]
    T *Ea3 = Cb3 + 3 * NSEG;                            // [3][NSEG]
    T *tl3 = Ea3 + 3 * NSEG;                            // [3][N][kLRow]
    for (int e = threadIdx.x; e < 3 * N * kLRow; e += blockDim.x) {
        const int qj = e / kLRow, k = e - qj * kLRow;  // qj = q * N + j
        tl3[e] = T(A.tab[(long)qj * kRow + kTabMap[k]]);
    }
    __syncthreads();

    const int tid = threadIdx.x;
    int seg, ml;
    seg_lane<MPW, NSEG>(tid, seg, ml);
    const long m = (long)bx * MPW + ml;
    const bool live = m < ps;
    const int a = seg * R, b = min(a + R, N) - 1;
    const bool rows = live && a <= b;
    int s = 0, i = 0;
    const long gm = (live && A.map) ? (long)A.map[m] : m;
    if (live) { s = (int)(gm / g.mk); i = (int)(gm - (long)s * g.mk); }
    const long nps = A.nl_ps;
    const CnMode<T> K = cn_mode<T>(A, live, i, s);
    const cx<T> zero = make_cx<T>(0, 0);
    // systems 0, 1, 2 = SYS_U1, SYS_U3, SYS_U2
    cx<T> *const fldp[3] = {P.cu1, P.cu3, P.cu2};
    cx<T> *const cfp[3] = {P.cf1, P.cf3, P.cf2};
    auto tab = [&](int q, int j, int k) -> T { return tl3[(q * N + j) * kLRow + k]; };
    auto put = [&](int q, int r, cx<T> v) {
        if (q == 0) u1s[r] = v;
        else if (q == 1) u3s[r] = v;
        else u2s[r] = v;
    };
    auto get = [&](int q, int r) -> cx<T> { return q == 0 ? u1s[r] : q == 1 ? u3s[r] : u2s[r]; };

    // ---- forward, local ----
    if (rows) {
        struct In3 {
            cx<T> fp[3], nl[3], cf[3], th, cp;
        };
        auto load = [&](int j) {
            const long o = m + (long)j * ps;
            In3 in;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                in.fp[q] = fldp[q][o + (j <= NY ? ps : 0)];
                const cx<NLT> nv = P.nl[q][gm + (long)j * nps];
                in.nl[q] = make_cx<T>(T(nv.re), T(nv.im));
                in.cf[q] = K.rk_gt1 ? cfp[q][o] : zero;
            }
            in.th = NTH ? P.cth[o] : zero;
            in.cp = P.cp[o];
            return in;
        };
        const long oa = m + (long)a * ps;
        cx<T> f0[3], fm[3], gl[3];
        T cprod[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            fm[q] = a >= 1 ? fldp[q][oa - ps] : zero;
            f0[q] = fldp[q][oa];
            gl[q] = zero;
            cprod[q] = T(1);
        }
        cx<T> thm = (a >= 1 && NTH) ? P.cth[oa - ps] : zero;
        cx<T> cpm = a >= 1 ? P.cp[oa - ps] : zero;
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            const int j = a + r;
            if (j <= b) {
                const In3 cur = load(j);
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    auto tb = [&](int k) { return tab(q, j, k); };
                    const T mult = tb(L_MULT);
                    const bool isbc = tb(L_ISBC) != T(0);
                    const bool cfr = tb(L_CFR) != T(0);
                    const cx<T> fp = (j <= NY) ? cur.fp[q] : zero;
                    const bool use = cfr || !isbc;
                    const cx<T> nlv = use ? cur.nl[q] : zero;
                    const cx<T> cfo = use ? cur.cf[q] : zero;
                    // U1's partner is CU3, U3's CU1 (their own rows here)
                    const cx<T> aux1 = q == 0 ? f0[1] : q == 1 ? f0[0] : zero;
                    cx<T> cfn, vec;
                    cn_rhs<T, NTH>(q, K, tb, f0[q], fm[q], fp, aux1, cur.th, thm, cur.cp, cpm, nlv, cfo, cfn, vec);
                    gl[q] = (r == 0) ? vec : cfma(mult, gl[q], vec);
                    cprod[q] = (r == 0) ? mult : cprod[q] * mult;
                    put(q, r, gl[q]);
                    if (cfr) cfp[q][m + (long)j * ps] = cfn;
                }
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    fm[q] = f0[q];
                    f0[q] = (j <= NY) ? cur.fp[q] : zero;
                }
                thm = cur.th;
                cpm = cur.cp;
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            GL3[q * NM + seg * MPW + ml] = gl[q];
            if (ml == 0) Cb3[q * NSEG + seg] = cprod[q];
        }
    }
    __syncthreads();
    // ---- carries of the forward recurrences, back substitution, local ----
    if (rows) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            cx<T> gin = zero;
            for (int t = 0; t < seg; ++t) gin = cfma(Cb3[q * NSEG + t], gin, GL3[q * NM + t * MPW + ml]);
            T c = T(1);
#pragma unroll
            for (int r = 0; r < RMAX; ++r) {
                const int j = a + r;
                if (j <= b) {
                    c *= tab(q, j, L_MULT);
                    put(q, r, cfma(c, gin, get(q, r)));
                }
            }
            cx<T> xl = zero;
            T e = T(1);
            bool first = true;
#pragma unroll
            for (int r = RMAX - 1; r >= 0; --r) {
                const int j = a + r;
                if (j <= b) {
                    const T ip = tab(q, j, L_INVPIV), u = tab(q, j, L_UP);
                    const cx<T> y = get(q, r);
                    if (first) {
                        xl = make_cx<T>(y.re * ip, y.im * ip);
                        e = -u * ip;
                        first = false;
                    } else {
                        xl = make_cx<T>((y.re - u * xl.re) * ip, (y.im - u * xl.im) * ip);
                        e = (-u * ip) * e;
                    }
                    put(q, r, xl);
                }
            }
            XL3[q * NM + seg * MPW + ml] = xl;
            if (ml == 0) Ea3[q * NSEG + seg] = e;
        }
    }
    __syncthreads();
    // ---- carries of the backward recurrences, the solution of every row ----
    if (rows) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            cx<T> xin = zero;
            for (int t = nseg - 1; t > seg; --t) xin = cfma(Ea3[q * NSEG + t], xin, XL3[q * NM + t * MPW + ml]);
            T e = T(1);
            bool first = true;
#pragma unroll
            for (int r = RMAX - 1; r >= 0; --r) {
                const int j = a + r;
                if (j <= b) {
                    const T f = -tab(q, j, L_UP) * tab(q, j, L_INVPIV);
                    e = first ? f : f * e;
                    first = false;
                    put(q, r, (tab(q, j, L_CPY) != T(0)) ? cfma(e, xin, get(q, r)) : fldp[q][m + (long)j * ps]);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_colfused: the whole column stage of a substep for a group of modes, with
// the FP32 state or in FP64.  The CN solves of U1, U3, U2 (cn3_solve: one pass over the
// rows for the three) stay in registers; TH's goes straight into CTH (every
// system of these modes has read the old CTH by then, and no other block
// reads these modes); then the divergence of U*,
// the Poisson solve, the projection into CU1..3 and the CP update.  The
// state is read once and written once, with no solve scratch.  The Poisson
// rows 1..NY run on the lanes that hold them in the CN part (the segments of
// rows 0..NY+1); a segment without Poisson rows passes the carries through.
// ---------------------------------------------------------------------------
template <typename T, typename NLT, int MPW, int NW, int RMAX, int NTH>
__device__ __forceinline__ void colfused_body(const CnArgs &A, const RdArgs &Rd, const CnPtrs<T, NLT> &P, int bx)
{
    constexpr int NSEG = fused_nseg_of(MPW, NW);
    extern __shared__ unsigned char smem_raw[];
    const Geo &g = A.g;
    const int NY = g.ny, N = NY + 2;
    const long ps = g.ps, fps = Rd.fac_ps;
    const int R = (N + NSEG - 1) / NSEG;
    const int nseg = (N + R - 1) / R;
    T *tl = reinterpret_cast<T *>(smem_raw);                               // [N][kLRow] (cn_seg_solve)
    cx<T> *GL = reinterpret_cast<cx<T> *>(tl + ((N * kLRow + 1) & ~1));   // [NSEG][MPW]
    cx<T> *XL = GL + NSEG * MPW;                                           // [NSEG][MPW]
    T *Cm = reinterpret_cast<T *>(XL + NSEG * MPW) + 2 * NSEG;             // [NSEG][MPW] (after CN's Cb, Ea)
    T *Em = Cm + NSEG * MPW;                                               // [NSEG][MPW]
    cx<T> *X2 = reinterpret_cast<cx<T> *>(Em + NSEG * MPW);                // [NSEG][MPW]: U2*(a)

    const int tid = threadIdx.x;
    int seg, ml;
    seg_lane<MPW, NSEG>(tid, seg, ml);
    const long m = (long)bx * MPW + ml;
    const bool live = m < ps;
    const int a = seg * R, b = min(a + R, N) - 1;
    const bool rows = live && a <= b;
    const cx<T> zero = make_cx<T>(0, 0);

    // ---- the CN solves: U1, U3, U2 together, then TH ----
    cx<T> u1s[RMAX], u3s[RMAX], u2s[RMAX];
    cn3_solve<T, NTH, NLT, MPW, NW, RMAX>(A, P, bx, u1s, u3s, u2s);
    if (NTH) {
        __syncthreads();  // the TH solve's table goes where the three systems' carries were
        cx<T> ts[RMAX];
        cn_seg_solve<T, NTH, NLT, MPW, NW, RMAX, const CnArgs &, const CnPtrs<T, NLT> &, NSEG>(A, P, bx, SYS_TH0,
                                                                                              ts);
        if (rows) {
#pragma unroll
            for (int r = 0; r < RMAX; ++r) {
                const int j = a + r;
                if (j <= b && tl[j * kLRow + L_CPY] != T(0)) P.cth[m + (long)j * ps] = ts[r];
            }
        }
    }
    __syncthreads();  // the carry arrays are reused below
    if (rows) X2[seg * MPW + ml] = u2s[0];
    __syncthreads();
    const cx<T> u2nx = (rows && seg + 1 < nseg) ? X2[(seg + 1) * MPW + ml] : zero;  // U2*(b+1)

    // ---- divergence of U* and the Poisson elimination, rows max(a,1)..min(b,NY) ----
    int s = 0, i = 0;
    const long gm = (live && Rd.map) ? (long)Rd.map[m] : m;
    if (live) { s = (int)(gm / g.mk); i = (int)(gm - (long)s * g.mk); }
    const bool is00 = live && i == 0 && s == 0;
    const T kx = live ? T(Rd.kx[i]) : T(0), kz = live ? T(Rd.kz[s]) : T(0);
    auto pm = [&](int j) -> T {
        if constexpr (sizeof(T) == 4) return Rd.pmultf[gm + (long)j * fps];
        else return T(Rd.pmult[gm + (long)j * fps]);
    };
    auto pv = [&](int j) -> T {
        if constexpr (sizeof(T) == 4) return Rd.pinvf[gm + (long)j * fps];
        else return T(Rd.pinv[gm + (long)j * fps]);
    };
    auto ik = [](T k, cx<T> x) { return make_cx<T>(-k * x.im, k * x.re); };
    // U2* of the row after a + r: the next register, or the next segment's first row
    auto u2up = [&](int r) -> cx<T> { return (a + r < b) ? u2s[r + 1 < RMAX ? r + 1 : r] : u2nx; };
    cx<T> yv[RMAX];
    T mr[RMAX], fr[RMAX];
    if (rows) {
        cx<T> hl = zero;
        T cprod = T(1);
        bool first = true;
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            const int j = a + r;
            if (j <= b && j >= 1 && j <= NY) {
                cx<T> vec = zero;
                if (j >= 2 && j < NY) {
                    // (i kx CU1 + (CU2(j+1) - CU2(j))/DYF(j)) + i kz CU3 (for/channel.f:1139-1141)
                    const T idf = T(Rd.rd[(long)j * 4 + 1]);
                    const cx<T> d1 = ik(kx, u1s[r]), d3 = ik(kz, u3s[r]), c2p = u2up(r), c2 = u2s[r];
                    vec = make_cx<T>((d1.re + (c2p.re - c2.re) * idf) + d3.re,
                                     (d1.im + (c2p.im - c2.im) * idf) + d3.im);
                }
                const T mu = j >= 2 ? pm(j) : T(0);
                hl = first ? vec : cfma(mu, hl, vec);
                cprod = first ? mu : cprod * mu;
                first = false;
                yv[r] = hl;
                mr[r] = mu;
            }
        }
        GL[seg * MPW + ml] = hl;
        Cm[seg * MPW + ml] = cprod;
    }
    __syncthreads();
    const T u1row = is00 ? T(0) : T(-1);
    if (rows) {
        cx<T> hin = zero;
        for (int t = 0; t < seg; ++t) hin = cfma(Cm[t * MPW + ml], hin, GL[t * MPW + ml]);
        T c = T(1);
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            const int j = a + r;
            if (j <= b && j >= 1 && j <= NY) {
                c *= mr[r];
                yv[r] = cfma(c, hin, yv[r]);
            }
        }
        cx<T> xl = zero;
        T e = T(1);
        bool first = true;
#pragma unroll
        for (int r = RMAX - 1; r >= 0; --r) {
            const int j = a + r;
            if (j <= b && j >= 1 && j <= NY) {
                const T ip = pv(j);
                const T u = (j == 1) ? u1row : T(Rd.rd[(long)j * 4 + 0]);
                fr[r] = -u * ip;
                if (first) {
                    xl = make_cx<T>(yv[r].re * ip, yv[r].im * ip);
                    e = fr[r];
                    first = false;
                } else {
                    xl = make_cx<T>((yv[r].re - u * xl.re) * ip, (yv[r].im - u * xl.im) * ip);
                    e = fr[r] * e;
                }
                yv[r] = xl;
            }
        }
        XL[seg * MPW + ml] = xl;
        Em[seg * MPW + ml] = e;
    }
    __syncthreads();
    // ---- phi, the projection into CU1..3, the CP update ----
    if (rows) {
        cx<T> xin = zero;
        for (int t = nseg - 1; t > seg; --t) xin = cfma(Em[t * MPW + ml], xin, XL[t * MPW + ml]);
        const T itemp4 = T(Rd.inv_temp4);
        const bool upd = Rd.update_cp;
        cx<T> phip = xin;  // phi of the row after the lane's last Poisson row
        T e = T(1);
        bool first = true;
#pragma unroll
        for (int r = RMAX - 1; r >= 0; --r) {
            const int j = a + r;
            if (j <= b) {
                const long o = m + (long)j * ps;
                if (j >= 1 && j <= NY) {
                    e = first ? fr[r] : fr[r] * e;
                    first = false;
                    const cx<T> phi = cfma(e, xin, yv[r]);
                    // CU1 - i kx phi, CU3 - i kz phi (rows 1..NY), CU2(j+1) - (phi(j+1) - phi(j))/DY(j+1)
                    P.cu1[o] = u1s[r] - ik(kx, phi);
                    P.cu3[o] = u3s[r] - ik(kz, phi);
                    if (j + 1 <= NY) {
                        const T idy = T(Rd.rd[(long)(j + 1) * 4 + 2]);
                        const cx<T> v = u2up(r);
                        P.cu2[o + ps] = make_cx<T>(v.re - (phip.re - phi.re) * idy, v.im - (phip.im - phi.im) * idy);
                    }
                    if (upd && j >= Rd.jstart && j <= Rd.jend) {
                        const cx<T> p0 = P.cp[o];
                        P.cp[o] = make_cx<T>(p0.re + phi.re * itemp4, p0.im + phi.im * itemp4);
                    }
                    if (j == 1) P.cu2[o] = u2s[r];  // CU2(1) is outside the projection
                    phip = phi;
                } else {
                    // rows 0 and NY+1: no projection
                    P.cu1[o] = u1s[r];
                    P.cu3[o] = u3s[r];
                    P.cu2[o] = u2s[r];
                }
            }
        }
    }
}

template <typename T, typename NLT, int MPW, int NW, int RMAX, int NTH>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2)))
k_colfused(CnArgs A, RdArgs Rd, CnPtrs<T, NLT> P)
{
    colfused_body<T, NLT, MPW, NW, RMAX, NTH>(A, Rd, P, xcd_block(blockIdx.x, gridDim.x));
}

template <typename T, typename NLT, int MPW, int NW, int RMAX, int NTH>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2)))
k_colfused_batch(const ColDesc<T, NLT> *__restrict__ d)
{
    const ColDesc<T, NLT> &D = d[blockIdx.z];
    const int bx = xcd_block(blockIdx.x, gridDim.x);  // a member's mode groups in 8 runs, one per XCD
    if ((long)bx * MPW >= D.A.g.ps) return;
    colfused_body<T, NLT, MPW, NW, RMAX, NTH>(D.A, D.R, D.P, bx);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {

// segments per column: fixed per precision
template <typename T>
constexpr int seg_nseg() { return sizeof(T) == 4 ? 16 : 32; }
// modes per segment group for NW waves per block
template <typename T, int NW>
constexpr int seg_mpw() { return 64 * NW / seg_nseg<T>(); }

template <typename T>
size_t cn_seg_lds(int ny, int nw)
{
    const int nseg = seg_nseg<T>(), mpw = 64 * nw / nseg;
    const int n = ny + 2;
    return sizeof(T) * (size_t)((n * kLRow + 1) & ~1) + 2 * sizeof(cx<T>) * nseg * mpw + 2 * sizeof(T) * nseg;
}

template <typename T>
size_t remdiv_seg_lds(int nw)
{
    const int nseg = seg_nseg<T>(), mpw = 64 * nw / nseg;
    return 2 * sizeof(cx<T>) * nseg * mpw + 2 * sizeof(T) * nseg * mpw;
}

// rows per segment of the tallest column -> the compile-time bound
template <typename F>
bool with_rmax(int rows, F &&f)
{
    if (rows <= 9) { f(std::integral_constant<int, 9>{}); return true; }
    if (rows <= 17) { f(std::integral_constant<int, 17>{}); return true; }
    return false;
}

// waves per block -> the compile-time shape
template <typename F>
bool with_nw(int nw, F &&f)
{
    switch (nw) {
    case 4: f(std::integral_constant<int, 4>{}); return true;
    case 8: f(std::integral_constant<int, 8>{}); return true;
    case 16: f(std::integral_constant<int, 16>{}); return true;
    default: return false;
    }
}

}  // namespace

template <typename T>
bool colseg_supported(int ny, int nw)
{
    const int rows = (ny + 2 + seg_nseg<T>() - 1) / seg_nseg<T>();
    return (nw == 4 || nw == 8 || nw == 16) && ny >= 4 && rows <= 17 && cn_seg_lds<T>(ny, nw) <= 64 * 1024;
}

template <typename T, typename NLT>
void launch_cn_seg(hipStream_t s, int nw, const CnArgs &a, const CnPtrs<T, NLT> &p)
{
    constexpr int NSEG = seg_nseg<T>();
    const size_t lds = cn_seg_lds<T>(a.g.ny, nw);
    with_nw(nw, [&](auto wc) {
        constexpr int NW = decltype(wc)::value, MPW = seg_mpw<T, NW>();
        const long groups = (a.g.ps + MPW - 1) / MPW;
        const long blocks = (groups * (3 + a.g.nth) + 7) / 8 * 8;  // a multiple of 8 (xcd_block)
        const dim3 grid(blocks);
        with_rmax((a.g.ny + 2 + NSEG - 1) / NSEG, [&](auto rc) {
            constexpr int RM = decltype(rc)::value;
            if (a.g.nth == 0)
                hipLaunchKernelGGL((k_cn_seg<T, 0, NLT, MPW, NW, RM>), grid, dim3(64 * NW), lds, s, a, p);
            else
                hipLaunchKernelGGL((k_cn_seg<T, 1, NLT, MPW, NW, RM>), grid, dim3(64 * NW), lds, s, a, p);
        });
    });
}

template <typename T, typename NLT>
void launch_remdiv_seg(hipStream_t s, int nw, const RdArgs &a, const CnPtrs<T, NLT> &p)
{
    constexpr int NSEG = seg_nseg<T>();
    const size_t lds = remdiv_seg_lds<T>(nw);
    with_nw(nw, [&](auto wc) {
        constexpr int NW = decltype(wc)::value, MPW = seg_mpw<T, NW>();
        const long blocks = ((a.g.ps + MPW - 1) / MPW + 7) / 8 * 8;
        with_rmax((a.g.ny + NSEG - 1) / NSEG, [&](auto rc) {
            constexpr int RM = decltype(rc)::value;
            if (a.g.nth == 0)
                hipLaunchKernelGGL((k_remdiv_seg<T, NLT, MPW, NW, RM, 0>), dim3(blocks), dim3(64 * NW), lds, s, a, p);
            else
                hipLaunchKernelGGL((k_remdiv_seg<T, NLT, MPW, NW, RM, 1>), dim3(blocks), dim3(64 * NW), lds, s, a, p);
        });
    });
}

template <typename T, typename NLT>
void launch_cn_seg_batch(hipStream_t s, int nw, const Geo &g, int max_ny, const ColDesc<T, NLT> *d, int nmem)
{
    constexpr int NSEG = seg_nseg<T>();
    const size_t lds = cn_seg_lds<T>(max_ny, nw);
    with_nw(nw, [&](auto wc) {
        constexpr int NW = decltype(wc)::value, MPW = seg_mpw<T, NW>();
        const dim3 grid((g.ps + MPW - 1) / MPW, 3 + g.nth, nmem);
        with_rmax((max_ny + 2 + NSEG - 1) / NSEG, [&](auto rc) {
            constexpr int RM = decltype(rc)::value;
            if (g.nth == 0)
                hipLaunchKernelGGL((k_cn_seg_batch<T, 0, NLT, MPW, NW, RM>), grid, dim3(64 * NW), lds, s, d);
            else
                hipLaunchKernelGGL((k_cn_seg_batch<T, 1, NLT, MPW, NW, RM>), grid, dim3(64 * NW), lds, s, d);
        });
    });
}

template <typename T, typename NLT>
void launch_remdiv_seg_batch(hipStream_t s, int nw, const Geo &g, int max_ny, const ColDesc<T, NLT> *d, int nmem)
{
    constexpr int NSEG = seg_nseg<T>();
    const size_t lds = remdiv_seg_lds<T>(nw);
    with_nw(nw, [&](auto wc) {
        constexpr int NW = decltype(wc)::value, MPW = seg_mpw<T, NW>();
        const dim3 grid((g.ps + MPW - 1) / MPW, 1, nmem);
        with_rmax((max_ny + NSEG - 1) / NSEG, [&](auto rc) {
            constexpr int RM = decltype(rc)::value;
            if (g.nth == 0)
                hipLaunchKernelGGL((k_remdiv_seg_batch<T, NLT, MPW, NW, RM, 0>), grid, dim3(64 * NW), lds, s, d);
            else
                hipLaunchKernelGGL((k_remdiv_seg_batch<T, NLT, MPW, NW, RM, 1>), grid, dim3(64 * NW), lds, s, d);
        });
    });
}

// The fused column kernel (fused_nseg(NY) segments of at most 9 rows), FP32
// state or FP64 (there at two waves per SIMD: 252 VGPRs, and the three
// systems' tables take 81 KB of LDS at NY = 257, so 4 or 8 waves per block).
// A round of resident blocks is latency-bound, so the mode count per block
// sets the number of rounds: at 384x257x384 (33153 modes, one FP64 block of
// 8 waves per CU) 16 modes per block took 8 full rounds and a ninth of 32
// blocks, 17 take 7.6.
template <typename T>
static size_t colfused_lds(int ny, int nw)
{
    const int nseg = fused_nseg(ny), mpw = fused_mpw(nw, nseg), n = ny + 2;
    const size_t one = sizeof(T) * (size_t)((n * kLRow + 1) & ~1) + 2 * sizeof(cx<T>) * nseg * mpw +
                       2 * sizeof(T) * nseg + 2 * sizeof(T) * nseg * mpw + sizeof(cx<T>) * nseg * mpw;
    const size_t three = 6 * sizeof(cx<T>) * nseg * mpw + 6 * sizeof(T) * nseg +
                         3 * sizeof(T) * (size_t)n * kLRow;  // cn3_solve: carries, tables
    return std::max(one, three);
}

template <typename T>
bool colfused_supported(int ny, int nw)
{
    // FP64 blocks of 16 waves would hold its 252 VGPRs in 128 (spills)
    const int nseg = fused_nseg(ny);
    return (nw == 4 || nw == 8 || (nw == 16 && sizeof(T) == 4)) && ny >= 4 && (ny + 2 + nseg - 1) / nseg <= 9 &&
           colfused_lds<T>(ny, nw) <= 150 * 1024;
}

namespace {

// Calls fn(std::integral_constant<int, MPW>) with the fused kernel's modes
// per block for NW waves and the segment count of columns of ny rows.
template <int NW, class F>
void with_fused_mpw(int ny, F &&fn)
{
    const int nseg = fused_nseg(ny);
    if (nseg == kFusedSegShort) fn(std::integral_constant<int, fused_mpw(NW, kFusedSegShort)>{});
    else if (nseg == kFusedSegMid) fn(std::integral_constant<int, fused_mpw(NW, kFusedSegMid)>{});
    else fn(std::integral_constant<int, fused_mpw(NW, kFusedSeg)>{});
}

template <typename T, typename NLT, int MPW, int NW>
void colfused_attr()
{
    static bool attr = false;
    if (attr) return;
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_colfused<T, NLT, MPW, NW, 9, 0>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_colfused<T, NLT, MPW, NW, 9, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_colfused_batch<T, NLT, MPW, NW, 9, 0>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_colfused_batch<T, NLT, MPW, NW, 9, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr = true;
}

}  // namespace

template <typename T, typename NLT>
void launch_colfused(hipStream_t s, int nw, const CnArgs &a, const RdArgs &r, const CnPtrs<T, NLT> &p)
{
    const size_t lds = colfused_lds<T>(a.g.ny, nw);
    with_nw(nw, [&](auto wc) {
        constexpr int NW = decltype(wc)::value;
        if constexpr (sizeof(T) == 8 && NW == 16) {
            return;  // not built: colfused_supported refuses the shape
        } else {
            with_fused_mpw<NW>(a.g.ny, [&](auto mc) {
                constexpr int MPW = decltype(mc)::value;
                const long blocks = ((a.g.ps + MPW - 1) / MPW + 7) / 8 * 8;  // a multiple of 8 (xcd_block)
                colfused_attr<T, NLT, MPW, NW>();
                if (a.g.nth == 0)
                    hipLaunchKernelGGL((k_colfused<T, NLT, MPW, NW, 9, 0>), dim3(blocks), dim3(64 * NW), lds, s, a, r,
                                       p);
                else
                    hipLaunchKernelGGL((k_colfused<T, NLT, MPW, NW, 9, 1>), dim3(blocks), dim3(64 * NW), lds, s, a, r,
                                       p);
            });
        }
    });
}

// Every member of the batch runs the kernel built for the tallest column.
template <typename T, typename NLT>
void launch_colfused_batch(hipStream_t s, int nw, const Geo &g, int max_ny, const ColDesc<T, NLT> *d, int nmem)
{
    const size_t lds = colfused_lds<T>(max_ny, nw);
    with_nw(nw, [&](auto wc) {
        constexpr int NW = decltype(wc)::value;
        if constexpr (sizeof(T) == 8 && NW == 16) {
            return;  // not built: colfused_supported refuses the shape
        } else {
            with_fused_mpw<NW>(max_ny, [&](auto mc) {
                constexpr int MPW = decltype(mc)::value;
                const dim3 grid((unsigned)(((g.ps + MPW - 1) / MPW + 7) / 8 * 8), 1, nmem);  // a multiple of 8 (xcd_block)
                colfused_attr<T, NLT, MPW, NW>();
                if (g.nth == 0)
                    hipLaunchKernelGGL((k_colfused_batch<T, NLT, MPW, NW, 9, 0>), grid, dim3(64 * NW), lds, s, d);
                else
                    hipLaunchKernelGGL((k_colfused_batch<T, NLT, MPW, NW, 9, 1>), grid, dim3(64 * NW), lds, s, d);
            });
        }
    });
}

template bool colfused_supported<float>(int, int);
template bool colfused_supported<double>(int, int);
template void launch_colfused<float, float>(hipStream_t, int, const CnArgs &, const RdArgs &,
                                            const CnPtrs<float, float> &);
template void launch_colfused<double, double>(hipStream_t, int, const CnArgs &, const RdArgs &,
                                              const CnPtrs<double, double> &);
template void launch_colfused_batch<float, float>(hipStream_t, int, const Geo &, int, const ColDesc<float, float> *,
                                                  int);
template void launch_colfused_batch<double, double>(hipStream_t, int, const Geo &, int,
                                                    const ColDesc<double, double> *, int);
template bool colseg_supported<double>(int, int);
template bool colseg_supported<float>(int, int);
template void launch_cn_seg<double, double>(hipStream_t, int, const CnArgs &, const CnPtrs<double, double> &);
template void launch_cn_seg<double, float>(hipStream_t, int, const CnArgs &, const CnPtrs<double, float> &);
template void launch_cn_seg<float, float>(hipStream_t, int, const CnArgs &, const CnPtrs<float, float> &);
template void launch_remdiv_seg<double, double>(hipStream_t, int, const RdArgs &, const CnPtrs<double, double> &);
template void launch_remdiv_seg<double, float>(hipStream_t, int, const RdArgs &, const CnPtrs<double, float> &);
template void launch_remdiv_seg<float, float>(hipStream_t, int, const RdArgs &, const CnPtrs<float, float> &);
template void launch_cn_seg_batch<double, double>(hipStream_t, int, const Geo &, int, const ColDesc<double, double> *, int);
template void launch_cn_seg_batch<double, float>(hipStream_t, int, const Geo &, int, const ColDesc<double, float> *, int);
template void launch_cn_seg_batch<float, float>(hipStream_t, int, const Geo &, int, const ColDesc<float, float> *, int);
template void launch_remdiv_seg_batch<double, double>(hipStream_t, int, const Geo &, int, const ColDesc<double, double> *,
                                                      int);
template void launch_remdiv_seg_batch<double, float>(hipStream_t, int, const Geo &, int, const ColDesc<double, float> *,
                                                     int);
template void launch_remdiv_seg_batch<float, float>(hipStream_t, int, const Geo &, int, const ColDesc<float, float> *,
                                                    int);

}  // namespace dgpu
