// Wave64 lane-shift primitives shared by the register-resident tile kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace sw1d {

constexpr int kWave = 64;

// ---------------------------------------------------------- lane shifts ---
// DPP wave_shr:1 (0x138): lane l receives lane l-1; lane 0 keeps `old`.
// DPP wave_shl:1 (0x130): lane l receives lane l+1; lane 63 keeps `old`.
__device__ __forceinline__ int shr1_i(int v, int old)
{
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int shl1_i(int v, int old)
{
    return __builtin_amdgcn_update_dpp(old, v, 0x130, 0xf, 0xf, false);
}

__device__ __forceinline__ double from_left(double v, double old)
{
    const int lo = shr1_i(__double2loint(v), __double2loint(old));
    const int hi = shr1_i(__double2hiint(v), __double2hiint(old));
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right(double v, double old)
{
    const int lo = shl1_i(__double2loint(v), __double2loint(old));
    const int hi = shl1_i(__double2hiint(v), __double2hiint(old));
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float from_left(float v, float old)
{
    return __int_as_float(shr1_i(__float_as_int(v), __float_as_int(old)));
}
__device__ __forceinline__ float from_right(float v, float old)
{
    return __int_as_float(shl1_i(__float_as_int(v), __float_as_int(old)));
}

// e[0 .. P+2R-1] = values at node-local x = l*P - R .. l*P + P + R - 1.
// The point at x = l*P - 1 - i is lane l-1's e[R - 1 - i + P]: its own point
// P-1-i when i < P, else (a radius wider than a lane, R > P) the value lane
// l-1 received i - P shifts earlier.  Likewise on the right.
template <typename T, int P, int R>
__device__ __forceinline__ void gather(const T (&v)[P], const T (&hl)[R], const T (&hr)[R],
                                       T (&e)[P + 2 * R])
{
#pragma unroll
    for (int p = 0; p < P; ++p) e[R + p] = v[p];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        e[R - 1 - i] = from_left(e[R - 1 - i + P], hl[i]);
        e[R + P + i] = from_right(e[R + i], hr[i]);
    }
}

// gather with the neighbour registers carried from one call to the next:
// nl / nr hold the halo in lane 0 / lane 63, which the shift never writes,
// and the last shifted values in the other lanes.  The shift's destination
// is then its own `old` operand, so no copy of the halo precedes it; the
// values in e[] are those of gather() with hl = nl, hr = nr.
template <typename T, int P, int R>
__device__ __forceinline__ void gather_carry(const T (&v)[P], T (&nl)[R], T (&nr)[R],
                                             T (&e)[P + 2 * R])
{
#pragma unroll
    for (int p = 0; p < P; ++p) e[R + p] = v[p];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        nl[i] = from_left(e[R - 1 - i + P], nl[i]);
        nr[i] = from_right(e[R + i], nr[i]);
        e[R - 1 - i] = nl[i];
        e[R + P + i] = nr[i];
    }
}

// dst = fma(a, b, c), one rounding, in the lanes where d <= thr; the other
// lanes keep dst.  v_cmpx narrows EXEC to those lanes for the one FMA, and
// EXEC is then set back to every lane: the tile kernels that call this run
// whole waves with every lane active (one node per wave).  Two issue slots
// instead of a compare and two 32-bit selects.
__device__ __forceinline__ void fma_where(double& dst, double a, double b, double c, int d, int thr)
{
    asm volatile("v_cmpx_ge_i32_e32 vcc, %4, %5\n\t"
                 "v_fma_f64 %0, %1, %2, %3\n\t"
                 "s_mov_b64 exec, -1"
                 : "+v"(dst)
                 : "v"(a), "v"(b), "v"(c), "s"(thr), "v"(d)
                 : "vcc");
}
__device__ __forceinline__ void fma_where(float& dst, float a, float b, float c, int d, int thr)
{
    asm volatile("v_cmpx_ge_i32_e32 vcc, %4, %5\n\t"
                 "v_fma_f32 %0, %1, %2, %3\n\t"
                 "s_mov_b64 exec, -1"
                 : "+v"(dst)
                 : "v"(a), "v"(b), "v"(c), "s"(thr), "v"(d)
                 : "vcc");
}

}  // namespace sw1d
