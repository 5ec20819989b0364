// dgpu_hot.hpp -- pieces shared by the transform kernels of one RK substep
// (dgpu_hot_z.hip: zinv, zfwd_nl; dgpu_hot_x.hip: xfused, vmax).
#pragma once

#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "dgpu_kernels.hpp"

namespace dgpu {

namespace {

template <typename TO, typename FROM>
__device__ __forceinline__ cx<TO> cvt_(cx<FROM> a) { return make_cx<TO>(TO(a.re), TO(a.im)); }

// Kept slot of z frequency index f (0..NZ-1), or -1 in the dealiasing gap
// (slot K <= NKZ holds kz = K, slot K > NKZ holds kz = K - (TNKZ+1),
// for/mpi.f:1602-1606).
__device__ __forceinline__ int slot_of_freq(int f, int nz, int nkz, int tnkz)
{
    if (f <= nkz) return f;
    if (f >= nz - nkz) return f - nz + tnkz + 1;
    return -1;
}

template <int NTH>
struct XCount {
    static constexpr int nin = 3 + NTH;         // u1 u2 u3 th...
    static constexpr int npi = (nin + 1) / 2;   // backward pairs
    static constexpr int npo = (8 + 3 * NTH + 1) / 2;  // forward pairs of the products
};

}  // namespace

// The plans the kernels are also built for at compile time (PID > 0).
__device__ constexpr IpLen kDevConstPlans[] = {const_plan(0),  const_plan(1),  const_plan(2),  const_plan(3),
                                               const_plan(4),  const_plan(5),  const_plan(6),  const_plan(7),
                                               const_plan(8),  const_plan(9),  const_plan(10), const_plan(11),
                                               const_plan(12), const_plan(13), const_plan(14), const_plan(15),
                                               const_plan(16), const_plan(17)};
static_assert(sizeof(kDevConstPlans) / sizeof(IpLen) == (size_t)kNumConstPlans, "one entry per compile-time plan");

template <int PID>
__device__ __forceinline__ const IpLen &plan_ref(const IpLen &rt)
{
    if constexpr (PID == 0)
        return rt;
    else
        return kDevConstPlans[PID];
}

#define DGPU_DISPATCH_IPE(E_, CALL)                                \
    switch (E_) {                                                  \
    case 4: { constexpr int EE = 4; CALL; } break;                 \
    case 6: { constexpr int EE = 6; CALL; } break;                 \
    case 8: { constexpr int EE = 8; CALL; } break;                 \
    case 12: { constexpr int EE = 12; CALL; } break;               \
    case 16: { constexpr int EE = 16; CALL; } break;               \
    case 24: { constexpr int EE = 24; CALL; } break;               \
    default: break;                                                \
    }

static inline int z_lw(int t)
{
    int lw = 0;
    while ((2 << lw) * t <= 256) ++lw;
    return lw;
}

// Compile-time plans each kernel family is built for, per precision (ids of
// kConstPlanSpecs): x, the 384 and 256 cases in both precisions and the
// ensemble lengths of the FP32 path; z, every length with a plan, both
// precisions.
template <int... P>
struct PidList {};
template <typename T>
using XPids = std::conditional_t<sizeof(T) == 4, PidList<2, 3, 5, 13, 14, 17>, PidList<1, 12>>;
// the single-context x kernel: the 384 and 256 plans
template <typename T>
using XPidsSingle = std::conditional_t<sizeof(T) == 4, PidList<13, 17>, PidList<1, 12>>;
template <typename T>
using ZPids = PidList<6, 8, 9, 10, 15, 16>;

// Calls fn(std::integral_constant<int, P>) for the P of the list equal to
// pid; false when pid is not in the list.
template <class F, int... P>
static bool with_pid(int pid, PidList<P...>, F &&fn)
{
    bool done = false;
    ((!done && pid == P ? (fn(std::integral_constant<int, P>{}), done = true) : false), ...);
    return done;
}

// Id of the compile-time plan equal to f, or 0 (dgpu_hot_z.hip).
int const_plan_id(const IpLen &f);

}  // namespace dgpu
