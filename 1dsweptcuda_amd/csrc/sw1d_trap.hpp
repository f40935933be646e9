// Device code of the overlapped trapezoid tile (one wave64 per tile), shared
// by the single-problem kernels (sw1d_trapezoid.hip) and the batched
// persistent kernel (sw1d_batch.hip).
//
// A tile's window is 64*PI consecutive points (mod n) starting at `win`;
// lane l holds points l*PI .. l*PI+PI-1 in VGPRs.  After S stages the window
// positions RS .. RS+count-1 (RS = R*S) hold the classic scheme's values and
// are stored to out0 .. out0+count-1 (mod n).
#pragma once

#include "sw1d_internal.hpp"
#include "sw1d_lanes.hpp"

namespace sw1d {

// i mod n in [0, n).  The tiles' indices lie within one period of the
// domain (window starts in [0, n), windows and output ranges no longer than
// n), so one add or subtract wraps them; the 64-bit remainder, a long
// software sequence on the GPU, is there for any other index and runs only
// in a wave that has one.
__device__ __forceinline__ int64_t wrapn(int64_t i, int64_t n)
{
    if (i >= -n && i < 2 * n) return i < 0 ? i + n : (i >= n ? i - n : i);
    i %= n;
    return i < 0 ? i + n : i;
}

// Lane l's PI window points win + l*PI + p (mod n).  Windows that do not wrap
// (all but the tiles next to the domain ends) index directly.
template <typename T, int PI>
__device__ __forceinline__ void load_window(const T* __restrict__ a, int64_t n, int64_t win,
                                            int lane, T (&v)[PI])
{
    const int64_t i0 = win + (int64_t)lane * PI;
    if (win + (int64_t)kWave * PI <= n) {
#pragma unroll
        for (int p = 0; p < PI; ++p) v[p] = a[i0 + p];
    } else {
#pragma unroll
        for (int p = 0; p < PI; ++p) v[p] = a[wrapn(i0 + p, n)];
    }
}

// Store the points at window positions RS .. RS+count-1 to out0 + (pos - RS).
template <typename T, int PI>
__device__ __forceinline__ void store_window(T* __restrict__ a, int64_t n, int64_t out0, int count,
                                             int lane, int RS, const T (&v)[PI])
{
    const bool direct = out0 + count <= n;
#pragma unroll
    for (int p = 0; p < PI; ++p) {
        const int pos = lane * PI + p - RS;
        if (pos >= 0 && pos < count) a[direct ? out0 + pos : wrapn(out0 + pos, n)] = v[p];
    }
}

// ------------------------------------------------------------------ heat ---
template <typename T, int PI, int S, bool BND>
__device__ __forceinline__ void trap_heat(const T* __restrict__ in, T* __restrict__ out, int64_t n,
                                          int64_t win, int64_t out0, int count, int lane, T fo,
                                          T fo2)
{
    constexpr int RS = S;  // R = 1
    T v[PI];
    int64_t g0 = 0;
    if (BND) g0 = wrapn(win + (int64_t)lane * PI, n);
    load_window<T, PI>(in, n, win, lane, v);
    // The window's outer points are discarded, so lane 0 / lane 63 may see
    // any finite value beyond it: the carried registers keep the first one.
    T hl[1] = {v[0]}, hr[1] = {v[PI - 1]};
    for (int s = 0; s < S; ++s) {
        T e[PI + 2];
        gather_carry<T, PI, 1>(v, hl, hr, e);
#pragma unroll
        for (int p = 0; p < PI; ++p) {
            T l = e[p], c = e[p + 1], r = e[p + 2];
            if (BND) {  // insulated ends (Heat:140-142)
                int64_t g = g0 + p;
                g = g >= n ? g - n : g;
                l = g == 0 ? r : l;
                r = g == n - 1 ? l : r;
            }
            v[p] = heat_ftcs(l, c, r, fo, fo2);
        }
    }
    store_window<T, PI>(out, n, out0, count, lane, RS, v);
}

// -------------------------------------------------------------------- KS ---
template <typename T, int PI, int S>
__device__ __forceinline__ void trap_ks(const T* __restrict__ in, T* __restrict__ out, int64_t n,
                                        int64_t win, int64_t out0, int count, int lane,
                                        const KsCoef<T>& k)
{
    constexpr int RS = 2 * S;
    T u[PI], uh[PI];
    load_window<T, PI>(in, n, win, lane, u);
    // beyond the window (discarded points): the carried registers' values
    T hl[2] = {u[0], u[0]}, hr[2] = {u[0], u[0]};
    T hlh[2] = {u[0], u[0]}, hrh[2] = {u[0], u[0]};
    for (int s = 0; s < S; s += 2) {
        {
            T e[PI + 4];
            gather_carry<T, PI, 2>(u, hl, hr, e);
#pragma unroll
            for (int p = 0; p < PI; ++p) uh[p] = ks_predict(e[p], e[p + 1], e[p + 2], e[p + 3], e[p + 4], k);
        }
        {
            T e[PI + 4];
            gather_carry<T, PI, 2>(uh, hlh, hrh, e);
#pragma unroll
            for (int p = 0; p < PI; ++p)
                u[p] = ks_correct(u[p], e[p], e[p + 1], e[p + 2], e[p + 3], e[p + 4], k);
        }
    }
    store_window<T, PI>(out, n, out0, count, lane, RS, u);
}

// ----------------------------------------------------------------- Euler ---
template <typename T, int PI, bool BND>
__device__ __forceinline__ void eu_trap_stage(const T (&src)[3][PI], T (&dst)[3][PI],
                                              const T (&base)[3][PI], T (&hl)[3][2], T (&hr)[3][2],
                                              int64_t g0, int64_t n, T c, const EulerCoef<T>& k)
{
    T ex[PI + 4], ey[PI + 4], ez[PI + 4];
    gather_carry<T, PI, 2>(src[0], hl[0], hr[0], ex);
    gather_carry<T, PI, 2>(src[1], hl[1], hr[1], ey);
    gather_carry<T, PI, 2>(src[2], hl[2], hr[2], ez);
    auto cell = [&](int j) { return Q3<T>{ex[j], ey[j], ez[j]}; };
    // global index of ext cell j (lane's point p is ext cell p + 2)
    auto gidx = [&](int j) {
        int64_t g = g0 + j - 2;
        return g < 0 ? g + n : (g >= n ? g - n : g);
    };
    T pd[PI + 3];
#pragma unroll
    for (int j = 0; j < PI + 3; ++j) pd[j] = eu_pdiff(cell(j), cell(j + 1));
    // interface j between ext cells j+1 and j+2 (Euler:311-341); the
    // interface right of cell 0 drops pdiff(-1,0), the one right of cell n-2
    // drops pdiff(n-1,n) (the flags of Euler:405, :424)
    auto iface = [&](int j, Q3<T>& F, Q3<T>& D) {
        T q0 = pd[j], q2 = pd[j + 2];
        if (BND) {
            const int64_t gl = gidx(j + 1);
            q0 = gl == 0 ? T(0) : q0;
            q2 = gl == n - 2 ? T(0) : q2;
        }
        eu_iface(cell(j), cell(j + 1), cell(j + 2), cell(j + 3), q0, pd[j + 1], q2, k, F, D);
    };
    Q3<T> FL, DL;
    iface(0, FL, DL);
#pragma unroll
    for (int p = 0; p < PI; ++p) {
        Q3<T> FR, DR;
        iface(p + 1, FR, DR);
        const Q3<T> f = eu_balance(FL, DL, FR, DR);
        Q3<T> nv = eu_apply(Q3<T>{base[0][p], base[1][p], base[2][p]}, f, c);
        if (BND) {  // ghost cells hold the Dirichlet states (Euler:407-416)
            const int64_t g = gidx(p + 2);
            nv = g == 0 ? k.bl : nv;
            nv = g == n - 1 ? k.br : nv;
        }
        dst[0][p] = nv.x;
        dst[1][p] = nv.y;
        dst[2][p] = nv.z;
        FL = FR;
        DL = DR;
    }
}

template <typename T, int PI, int S, bool BND>
__device__ __forceinline__ void trap_euler(const T* __restrict__ in, T* __restrict__ out, int64_t n,
                                           int64_t win, int64_t out0, int count, int lane,
                                           const EulerCoef<T>& k)
{
    constexpr int RS = 2 * S;
    T q[3][PI], qh[3][PI];
    const int64_t g0 = wrapn(win + (int64_t)lane * PI, n);
#pragma unroll
    for (int c = 0; c < 3; ++c) load_window<T, PI>(in + c * n, n, win, lane, q[c]);
    const T ca = T(0.25) * k.dtdx, cb = T(0.5) * k.dtdx;
    // beyond the window (discarded cells): the carried registers' values,
    // a physical state so that the discarded cells stay finite
    T hl[3][2], hr[3][2], hlh[3][2], hrh[3][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < 2; ++i) hl[c][i] = hr[c][i] = hlh[c][i] = hrh[c][i] = q[c][0];
    for (int s = 0; s < S; s += 2) {
        eu_trap_stage<T, PI, BND>(q, qh, q, hl, hr, g0, n, ca, k);     // predictor (Euler:307-350)
        eu_trap_stage<T, PI, BND>(qh, q, q, hlh, hrh, g0, n, cb, k);   // corrector (Euler:353-389)
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) store_window<T, PI>(out + c * n, n, out0, count, lane, RS, q[c]);
}


}  // namespace sw1d
