// Overlapped trapezoid tiles — the up-triangle of the swept rule with its base
// widened by R*S points on each side so that after S stages its top is a full
// output tile of W_out = 64*PO points.
//
// One wave64 per tile.  Lane l holds PI = PO + SB consecutive points of the
// W_in = 64*PI point window in VGPRs (SB = 2*R*S/64 extra points per lane);
// every stage updates every point of the window, neighbours crossing lanes
// through DPP wave shifts.  A point's value after stage s is exact when its
// stage-s dependency cone lies inside the window, which holds for the
// W_in - 2*R*s central points, so after S stages the W_out central points are
// the classic scheme's values, bitwise (same formulas, same operands).  The
// points nearer the window edges are recomputed by the neighbouring tile: the
// redundant work is 2*R*S / W_in of the lanes (11% at PO = 8, SB = 1), against
// half the lanes idle in a triangle or diamond tile, and no per-point masks.
//
// Any nX works: tiles start at multiples of W_out, the last one stores only
// the points that exist, addresses wrap modulo nX.  The state crossing launches
// is the plain field at a full time step (heat T; KS u; Euler (rho, m, E)).
//
// Boundary conditions are applied by global index.  Tiles whose window
// touches point 0 or nX-1 of a non-periodic domain run a variant that carries
// the per-point checks; all other tiles run without them.

#include "sw1d_trap.hpp"

namespace sw1d {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;

// ---------------------------------------------------------------- kernel ---
// BCK: 0 periodic, 1 insulated (heat), 2 Dirichlet ghost cells (Euler).
template <typename T, int EQ, int PO, int SB, int BCK>
__global__ __launch_bounds__(kBlock) void k_trapezoid(const T* __restrict__ in, T* __restrict__ out,
                                                      int64_t n, int64_t ntiles, int64_t tile_first,
                                                      T fo, T fo2, KsCoef<T> ks, EulerCoef<T> eu)
{
    constexpr int PI = PO + SB, WIN = kWave * PI, WOUT = kWave * PO, RS = 32 * SB;
    constexpr int S = EQ == EQ_HEAT ? RS : RS / 2;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t tile = tile_first + (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (tile >= ntiles) return;
    const int64_t out0 = tile * WOUT;
    const int count = (int)(n - out0 < WOUT ? n - out0 : WOUT);
    const int64_t win = wrapn(out0 - RS, n);
    // does the window reach point 0 or point n-1 of a non-periodic domain?
    const bool bnd = BCK != 0 && (out0 < RS + 1 || out0 - RS + WIN > n - 1 || WIN >= n);
    if (EQ == EQ_HEAT) {
        if (bnd)
            trap_heat<T, PI, S, true>(in, out, n, win, out0, count, lane, fo, fo2);
        else
            trap_heat<T, PI, S, false>(in, out, n, win, out0, count, lane, fo, fo2);
    } else if (EQ == EQ_KS) {
        trap_ks<T, PI, S>(in, out, n, win, out0, count, lane, ks);
    } else {
        if (bnd)
            trap_euler<T, PI, S, true>(in, out, n, win, out0, count, lane, eu);
        else
            trap_euler<T, PI, S, false>(in, out, n, win, out0, count, lane, eu);
    }
}

template <typename T> struct CoefPick;
template <> struct CoefPick<double> {
    static double fo(const Coefs& c) { return c.fo_d; }
    static double fo2(const Coefs& c) { return c.fo2_d; }
    static const KsCoef<double>& ks(const Coefs& c) { return c.ks_d; }
    static const EulerCoef<double>& eu(const Coefs& c) { return c.eu_d; }
};
template <> struct CoefPick<float> {
    static float fo(const Coefs& c) { return c.fo_f; }
    static float fo2(const Coefs& c) { return c.fo2_f; }
    static const KsCoef<float>& ks(const Coefs& c) { return c.ks_f; }
    static const EulerCoef<float>& eu(const Coefs& c) { return c.eu_f; }
};

template <typename T, int EQ, int PO, int SB>
hipError_t trap_bc(const TrapLaunch& L, const Coefs& c, hipStream_t s)
{
    const int64_t count = L.ntiles - L.tile_first;
    if (count <= 0) return hipSuccess;
    const dim3 grid((unsigned)((count + kWavesPerBlock - 1) / kWavesPerBlock));
    const T* in = static_cast<const T*>(L.in);
    T* out = static_cast<T*>(L.out);
    using C = CoefPick<T>;
    if (L.bc == BC_PERIODIC)
        hipLaunchKernelGGL((k_trapezoid<T, EQ, PO, SB, 0>), grid, dim3(kBlock), 0, s, in, out, L.n,
                           L.ntiles, L.tile_first, C::fo(c), C::fo2(c), C::ks(c), C::eu(c));
    else if (EQ == EQ_HEAT)
        hipLaunchKernelGGL((k_trapezoid<T, EQ, PO, SB, 1>), grid, dim3(kBlock), 0, s, in, out, L.n,
                           L.ntiles, L.tile_first, C::fo(c), C::fo2(c), C::ks(c), C::eu(c));
    else
        hipLaunchKernelGGL((k_trapezoid<T, EQ, PO, SB, 2>), grid, dim3(kBlock), 0, s, in, out, L.n,
                           L.ntiles, L.tile_first, C::fo(c), C::fo2(c), C::ks(c), C::eu(c));
    return hipGetLastError();
}

template <typename T, int EQ>
hipError_t trap_eq(const TrapLaunch& L, const Coefs& c, hipStream_t s)
{
    // (PO, SB) pairs provided per equation
    if (EQ == EQ_EULER) {
        if (L.PO == 2 && L.SB == 1) return trap_bc<T, EQ, 2, 1>(L, c, s);
        if (L.PO == 4 && L.SB == 1) return trap_bc<T, EQ, 4, 1>(L, c, s);
        return hipErrorInvalidValue;
    }
    switch (L.PO * 4 + L.SB) {
    case 2 * 4 + 1: return trap_bc<T, EQ, 2, 1>(L, c, s);
    case 4 * 4 + 1: return trap_bc<T, EQ, 4, 1>(L, c, s);
    case 8 * 4 + 1: return trap_bc<T, EQ, 8, 1>(L, c, s);
    case 8 * 4 + 2: return trap_bc<T, EQ, 8, 2>(L, c, s);
    case 16 * 4 + 1: return trap_bc<T, EQ, 16, 1>(L, c, s);
    case 16 * 4 + 2: return trap_bc<T, EQ, 16, 2>(L, c, s);
    }
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t trap_prec(const TrapLaunch& L, const Coefs& c, hipStream_t s)
{
    switch (L.eq) {
    case EQ_HEAT: return trap_eq<T, EQ_HEAT>(L, c, s);
    case EQ_KS: return trap_eq<T, EQ_KS>(L, c, s);
    case EQ_EULER: return trap_eq<T, EQ_EULER>(L, c, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_trapezoid(const TrapLaunch& L, bool f64, const Coefs& c, hipStream_t s)
{
    return f64 ? trap_prec<double>(L, c, s) : trap_prec<float>(L, c, s);
}

bool trapezoid_supported(int eq, int PO, int SB)
{
    if (eq == EQ_EULER) return SB == 1 && (PO == 2 || PO == 4);
    return (SB == 1 && (PO == 2 || PO == 4 || PO == 8 || PO == 16)) ||
           (SB == 2 && (PO == 8 || PO == 16));
}

}  // namespace sw1d
