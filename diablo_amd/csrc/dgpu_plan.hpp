// dgpu_plan.hpp -- host-side plans for the in-place transforms (IpLen).
//
// A plan fixes the radices of a length, the values per thread E and the
// per-stage twiddle tables.  The x-direction plan of the fused nonlinear
// kernel has a constraint of its own: its outermost radix R_0 is E/2, so that
// in the last forward stage every thread holds two whole groups, k and
// L_0 - k, i.e. both Z(k') and Z(n - k') of every frequency k' it owns, and
// can split the transform of a pair of real rows without another pass
// through LDS.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "dgpu_fft.hpp"

namespace dgpu {

// Radices of m that divide e, fewest stages first (greedy on 8, 6, 5, 4, 3, 2).
inline bool ip_factor(int m, int e, std::vector<int> &out)
{
    static const int cand[] = {8, 6, 5, 4, 3, 2};
    while (m > 1) {
        bool ok = false;
        for (int r : cand) {
            if (e % r == 0 && m % r == 0) {
                out.push_back(r);
                m /= r;
                ok = true;
                break;
            }
        }
        if (!ok) return false;
    }
    return true;
}

inline bool ip_fill(int n, int e, const std::vector<int> &radix, IpLen &f)
{
    if ((int)radix.size() > kMaxStages || radix.empty() || n % e) return false;
    f = IpLen{};
    f.swzs = 3;
    f.n = n;
    f.nstages = (int)radix.size();
    f.e = e;
    f.t = n / e;
    int m = n, off = 0;
    for (int i = 0; i < f.nstages; ++i) {
        const int R = radix[i];
        if (e % R || m % R) return false;
        f.radix[i] = R;
        f.span[i] = m / R;
        f.magic[i] = f.span[i] == 1 ? 0u : (unsigned)(((1ULL << 32) + f.span[i] - 1) / f.span[i]);
        f.twoff[i] = off;
        off += f.span[i] * (R - 1);
        m /= R;
    }
    if (m != 1) return false;
    f.twn = off;
    return true;
}

// Twiddle tables of a plan: stage i, entry k*(R_i-1) + q-1 = exp(-2 pi i k q / m_i).
inline void ip_twiddles(const IpLen &f, std::vector<cx<double>> &tw)
{
    const long double twopi = 6.283185307179586476925286766559005768L;
    tw.assign(f.twn, cx<double>{0.0, 0.0});
    for (int i = 0; i < f.nstages; ++i) {
        const int R = f.radix[i], L = f.span[i], m = R * L;
        for (int k = 0; k < L; ++k)
            for (int q = 1; q < R; ++q) {
                const long double a = -twopi * (long double)(k * q) / (long double)m;
                cx<double> &w = tw[f.twoff[i] + k * (R - 1) + q - 1];
                w.re = (double)cosl(a);
                w.im = (double)sinl(a);
            }
    }
}

// LDS cost of a plan's stage accesses for one row layout (positions of a row
// at x[swz(pos)], rows of n values of wpe 4-byte words back to back: 4 for
// complex FP64, 2 for complex FP32; 64 lanes a wavefront, T lanes a row),
// relative to conflict-free, for `banks` four-byte banks serving as many
// lanes a cycle as fill them.  64 banks is what MI355X shows: the FP32 x
// kernel's LDS bank-conflict cycles (34% of its LDS cycles) are what this
// model gives its former swizzle (1.75; conflict-free with 32 banks).
inline double ip_bank_cost(const IpLen &f, unsigned swz, int swzs, int wpe, int banks)
{
    IpLen g = f;
    g.swz = swz;
    g.swzs = swzs;
    const int T = f.t;
    const int lpc = banks / wpe;
    double tot = 0, ideal = 0;
    for (int i = 0; i < f.nstages; ++i) {
        const int R = f.radix[i], L = f.span[i];
        for (int u = 0; u < f.e / R; ++u)
            for (int q = 0; q < R; ++q)
                for (int g0 = 0; g0 < 64; g0 += lpc) {
                    int cnt[64] = {0};
                    int mx = 0;
                    for (int lane = g0; lane < g0 + lpc; ++lane) {
                        const int row = lane / T, th = lane % T;
                        const int t = th + T * u, b = t / L, k = t - b * L;
                        const long a = ((long)row * f.n + ip_swz(g, b * R * L + k + q * L)) * wpe;
                        for (int w = 0; w < wpe; ++w) mx = std::max(mx, ++cnt[(a + w) % banks]);
                    }
                    tot += (double)mx;
                    ideal += 1.0;
                }
    }
    return ideal > 0 ? tot / ideal : 1.0;
}

// Picks the row swizzle x -> x ^ ((x >> s) & m) of an x-direction plan
// whose values are wpe 4-byte words: the fewest modelled conflicts with 64
// banks, then with 32, over s = 1..7 and m = 1..31 (the map must be a
// bijection of 0..n-1; none when nothing beats the plain layout).
inline void ip_pick_swizzle(IpLen &f, int wpe = 4)
{
    f.swz = 0;
    f.swzs = 3;
    double best64 = ip_bank_cost(f, 0, 3, wpe, 64), best32 = ip_bank_cost(f, 0, 3, wpe, 32);
    for (int sh = 1; sh <= 7; ++sh)
        for (unsigned m = 1; m < 32; ++m) {
            IpLen g = f;
            g.swz = m;
            g.swzs = sh;
            std::vector<char> seen(f.n, 0);
            bool ok = true;
            for (int p = 0; p < f.n && ok; ++p) {
                const int q = ip_swz(g, p);
                if (q < 0 || q >= f.n || seen[q]) ok = false;
                else seen[q] = 1;
            }
            if (!ok) continue;
            const double c64 = ip_bank_cost(f, m, sh, wpe, 64);
            if (c64 > best64 + 1e-9) continue;
            const double c32 = ip_bank_cost(f, m, sh, wpe, 32);
            if (c64 < best64 - 1e-9 || c32 < best32 - 1e-9) {
                best64 = c64;
                best32 = c32;
                f.swz = m;
                f.swzs = sh;
            }
        }
}

// Same model for a column batch (lds[slot * W + b], lanes b = lane % W
// fastest): picks the cheapest row swap ip_cswz (bit 0 ^= bit s), 0 = none.
inline double ip_bank_cost_cols(const IpLen &f, int W, int cswz)
{
    IpLen g = f;
    g.cswz = cswz;
    double tot = 0, ideal = 0;
    const int models[2][2] = {{32, 8}, {64, 16}};
    for (int i = 0; i < f.nstages; ++i) {
        const int R = f.radix[i], L = f.span[i];
        for (int u = 0; u < f.e / R; ++u)
            for (int q = 0; q < R; ++q)
                for (const auto &md : models) {
                    const int banks = md[0], lpc = md[1];
                    for (int g0 = 0; g0 < 64; g0 += lpc) {
                        int cnt[64] = {0};
                        int mx = 0;
                        for (int lane = g0; lane < g0 + lpc; ++lane) {
                            const int b = lane % W, th = lane / W;
                            if (th >= f.t) continue;
                            const int t = th + f.t * u, bb = t / L, k = t - bb * L;
                            const long a = ((long)ip_cswz(g, bb * R * L + k + q * L) * W + b) * 4;
                            for (int w = 0; w < 4; ++w) mx = std::max(mx, ++cnt[(a + w) % banks]);
                        }
                        const int per = lpc * 4 / banks;
                        tot += (double)std::max(1, (mx + per - 1) / per);
                        ideal += 1.0;
                    }
                }
    }
    return ideal > 0 ? tot / ideal : 1.0;
}

inline int ip_pick_cswz(const IpLen &f, int W)
{
    int best = 0;
    double best_c = ip_bank_cost_cols(f, W, 0);
    for (int sbit = 1; sbit <= 5; ++sbit) {
        if ((1 << sbit) >= f.n) break;
        const double c = ip_bank_cost_cols(f, W, sbit);
        if (c < best_c - 1e-9) {
            best_c = c;
            best = sbit;
        }
    }
    return best;
}

// Plan of a z-direction (column batch) transform: fewest stages with at most
// e_max values per thread (e_req > 0 asks for that E).  At least two threads
// per transform, hence at least two stages.
inline bool ip_plan_z(int n, int e_req, int e_max, IpLen &f)
{
    static const int cand[] = {12, 8, 16, 24, 6, 4};
    int best_s = 1 << 30;
    bool found = false;
    for (int e : cand) {
        if (e_req > 0 ? e != e_req : e > e_max) continue;
        if (n % e || n / e > 64 || n / e < 2) continue;
        std::vector<int> r;
        if (!ip_factor(n, e, r)) continue;
        // fewer stages first; on a tie the earlier candidate (moderate E) wins
        if ((int)r.size() < best_s) {
            IpLen g;
            if (!ip_fill(n, e, r, g)) continue;
            f = g;
            best_s = (int)r.size();
            found = true;
        }
    }
    return found;
}

// Plan of the x-direction transforms of the fused nonlinear kernel:
// R_0 = E/2 (see the file comment); at most max_t threads per transform and
// e_max values per thread unless e_req > 0 asks for an E.
inline bool ip_plan_x(int n, int e_req, int e_max, int max_t, IpLen &f)
{
    static const int r0s[] = {6, 4, 8, 3, 2};
    int best_s = 1 << 30, best_d = 1 << 30;
    bool found = false;
    for (int r0 : r0s) {
        const int e = 2 * r0;
        if (e_req > 0 ? e != e_req : e > e_max) continue;
        if (n % e || n / e > max_t || n / e < 2) continue;
        std::vector<int> r{r0};
        if (!ip_factor(n / r0, e, r)) continue;
        const int t = n / e;
        const int d = std::abs(t - 32);
        if ((int)r.size() < best_s || ((int)r.size() == best_s && d < best_d)) {
            IpLen g;
            if (!ip_fill(n, e, r, g)) continue;
            f = g;
            best_s = (int)r.size();
            best_d = d;
            found = true;
        }
    }
    return found;
}

}  // namespace dgpu
