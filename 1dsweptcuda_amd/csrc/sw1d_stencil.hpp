// Point-update formulas of the three equations, written once for the GPU
// kernels and for the host-side hybrid node.  Every function takes its stencil
// values as arguments (no array indexing) so that the kernels can feed them
// from registers.  Operand order follows the reference expressions exactly and
// all code is compiled with -ffp-contract=off: each operator is one IEEE
// operation (the heat and KS stencils' explicit fused multiply-adds one each), on the
// GPU and on the host alike, so every path of this library produces the same
// bits.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

#define SW1D_HD __host__ __device__ __forceinline__

namespace sw1d {

// fused multiply-add, one rounding (FMA instructions on the GPU and the host)
SW1D_HD double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
SW1D_HD float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// ------------------------------------------------------------------ heat ---
// Heat1D_SweptShared.cu:121-130 (execFunc): Fo*(T[l]+T[r]) + (1-2Fo)*T[c],
// as fma(Fo, T[l]+T[r], (1-2Fo)*T[c]) (the CPU oracle uses the same form).
template <typename T>
SW1D_HD T heat_ftcs(T l, T c, T r, T fo, T fo2)
{
    return fmadd(fo, l + r, fo2 * c);
}

// -------------------------------------------------------------------- KS ---
// KS1D_SweptShared.cu:136-155 with the default (multiply by reciprocal)
// constants; c = {1/(4dx), 1/dx^2, 1/dx^4, dt, dt/2} (KS:917-926):
//   F = (r^2 - l^2) dx4x + (l + r - 2c) dx2 + (ll - 4l + 6c - 4r + rr) dx4,
// evaluated with fused multiply-adds at fixed places (one rounding each; the
// CPU oracle uses the same ones): twelve FP operations per stage.
template <typename T> struct KsCoef { T dx4x, dx2, dx4, dt, dth; };

// fma(-4, l, ll) into a new register.  On the GPU the compiler would form it
// as the two-address v_fmac and copy ll first (ll is still live as another
// point's neighbour); the three-address v_fma_f64 / v_fma_f32 needs no copy.
// The same single-rounding fused operation either way.
SW1D_HD double fma_m4(double l, double ll)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double t;
    asm("v_fma_f64 %0, -4.0, %1, %2" : "=v"(t) : "v"(l), "v"(ll));
    return t;
#else
    return fmadd(-4.0, l, ll);
#endif
}
SW1D_HD float fma_m4(float l, float ll)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float t;
    asm("v_fma_f32 %0, -4.0, %1, %2" : "=v"(t) : "v"(l), "v"(ll));
    return t;
#else
    return fmadd(-4.0f, l, ll);
#endif
}

template <typename T>
SW1D_HD T ks_rhs(T ll, T l, T c, T r, T rr, const KsCoef<T>& k)
{
    const T conv = fmadd(r, r, -(l * l));                                   // r^2 - l^2
    const T sec = fmadd(T(-2), c, l + r);                                   // l + r - 2c
    const T fourth = fmadd(T(-4), r, fmadd(T(6), c, fma_m4(l, ll))) + rr;
    return fmadd(conv, k.dx4x, fmadd(sec, k.dx2, fourth * k.dx4));
}
// KS:159-164 (stutterStep): predictor u_half = u - dt/2 * F(u).
template <typename T>
SW1D_HD T ks_predict(T ll, T l, T c, T r, T rr, const KsCoef<T>& k)
{
    return fmadd(-k.dth, ks_rhs(ll, l, c, r, rr, k), c);
}
// KS:166-171 (finalStep) and the "+=" of KS:202: u + (-dt * F(u_half)).
template <typename T>
SW1D_HD T ks_correct(T u, T ll, T l, T c, T r, T rr, const KsCoef<T>& k)
{
    return fmadd(-k.dt, ks_rhs(ll, l, c, r, rr, k), u);
}

// ----------------------------------------------------------------- Euler ---
// Conserved state (rho, rho*u, E) of one cell.  Own three-component type;
// the arithmetic is written out per component.
template <typename T> struct Q3 { T x, y, z; };

template <typename T> struct EulerCoef { T gam, mgam, dtdx; Q3<T> bl, br; };

// Euler1D_SweptShared.cu:183-193 (pressure).
template <typename T>
SW1D_HD T eu_pressure(const Q3<T>& c, T mgam)
{
    return mgam * (c.z - (T(0.5) * c.y * c.y / c.x));
}

// Euler:222-228 (limitor): cur + 1/2*min(ratio,1)*(other-cur).
template <typename T>
SW1D_HD Q3<T> eu_limit(const Q3<T>& cur, const Q3<T>& oth, T ratio)
{
    const T m = T(0.5) * (ratio < T(1) ? ratio : T(1));
    return Q3<T>{cur.x + m * (oth.x - cur.x), cur.y + m * (oth.y - cur.y),
                 cur.z + m * (oth.z - cur.z)};
}

// Euler:237-257 (eulerFlux): F(L)+F(R) without the factor 1/2, and
// Euler:266-292 (eulerSpectral): Roe-averaged (sqrt(gam*p/rho)+|u|)*(L-R).
template <typename T>
SW1D_HD void eu_flux_diss(const Q3<T>& L, const Q3<T>& R, const EulerCoef<T>& k,
                          Q3<T>& F, Q3<T>& D)
{
    const T uL = L.y / L.x;
    const T uR = R.y / R.x;
    const T pL = eu_pressure(L, k.mgam);
    const T pR = eu_pressure(R, k.mgam);
    F.x = (L.y + R.y);
    F.y = (L.y * uL + R.y * uR + pL + pR);
    F.z = (L.z * uL + R.z * uR + uL * pL + uR * pR);

    const T sl = sqrt(L.x);
    const T sr = sqrt(R.x);
    const T hx = sl * sr;
    const T den = T(1) / (hx * (sl + sr));
    const T hy = (sl * R.y + sr * L.y) * den;
    const T hz = (sl * R.z + sr * L.z) * den;
    const T ph = k.mgam * (hz - T(0.5) * hy * hy);  // Euler:202-212 (pressureHalf)
    const T s = (sqrt(ph * k.gam) + fabs(hy));
    D.x = s * (L.x - R.x);
    D.y = s * (L.y - R.y);
    D.z = s * (L.z - R.z);
}

// Division-free pressure difference of Euler:311-322.
template <typename T>
SW1D_HD T eu_pdiff(const Q3<T>& a, const Q3<T>& b)
{
    return (T(2) * b.x * a.x * (b.z - a.z) + (a.y * a.y * b.x - b.y * b.y * a.x));
}

// One interface between cells B and C (A and Dc are the next cells out), with
// q0 = pdiff(A,B), q1 = pdiff(B,C), q2 = pdiff(C,Dc).  The two limited states
// and the flux / dissipation of Euler:326-332 (left interface of a cell) and
// Euler:336-341 (right interface): the right interface of cell i and the left
// interface of cell i+1 are the same expressions on the same operands.
template <typename T>
SW1D_HD void eu_iface(const Q3<T>& A, const Q3<T>& B, const Q3<T>& C, const Q3<T>& Dc, T q0,
                      T q1, T q2, const EulerCoef<T>& k, Q3<T>& F, Q3<T>& D)
{
    const Q3<T> tL = (!q0 || !q1 || ((q0 < 0) != (q1 < 0))) ? B
                   : eu_limit(B, C, (A.x * q1 / (C.x * q0)));
    const Q3<T> tR = (!q1 || !q2 || ((q1 < 0) != (q2 < 0))) ? C
                   : eu_limit(C, B, (Dc.x * q1 / (B.x * q2)));
    eu_flux_diss(tL, tR, k, F, D);
}

// Cell balance in the reference's accumulation order (Euler:331-341):
// flux = F_left; flux += D_left; flux -= F_right; flux -= D_right.
template <typename T>
SW1D_HD Q3<T> eu_balance(const Q3<T>& FL, const Q3<T>& DL, const Q3<T>& FR, const Q3<T>& DR)
{
    Q3<T> a = FL;
    a.x = a.x + DL.x; a.y = a.y + DL.y; a.z = a.z + DL.z;
    a.x = a.x - FR.x; a.y = a.y - FR.y; a.z = a.z - FR.z;
    a.x = a.x - DR.x; a.y = a.y - DR.y; a.z = a.z - DR.z;
    return a;
}

// Net flux of cell 2 of the five cells s0..s4 (Euler:307-341 / 353-380).
// flagL / flagR drop the outermost pressure difference (cells 1 and n-2).
template <typename T>
SW1D_HD Q3<T> eu_net_flux(const Q3<T>& s0, const Q3<T>& s1, const Q3<T>& s2,
                          const Q3<T>& s3, const Q3<T>& s4, bool flagL, bool flagR,
                          const EulerCoef<T>& k)
{
    const T pLL = flagL ? T(0) : eu_pdiff(s0, s1);
    const T pL = eu_pdiff(s1, s2);
    const T pR = eu_pdiff(s2, s3);
    const T pRR = flagR ? T(0) : eu_pdiff(s3, s4);
    Q3<T> FL, DL, FR, DR;
    eu_iface(s0, s1, s2, s3, pLL, pL, pR, k, FL, DL);
    eu_iface(s1, s2, s3, s4, pL, pR, pRR, k, FR, DR);
    return eu_balance(FL, DL, FR, DR);
}

// Predictor / corrector update of one cell from its net flux.
template <typename T>
SW1D_HD Q3<T> eu_apply(const Q3<T>& base, const Q3<T>& f, T c)
{
    return Q3<T>{base.x + c * f.x, base.y + c * f.y, base.z + c * f.z};
}

// Euler:307-350 (eulerStutterStep): Q + (1/4 dt/dx) * flux.
template <typename T>
SW1D_HD Q3<T> eu_predict(const Q3<T>& s0, const Q3<T>& s1, const Q3<T>& s2,
                         const Q3<T>& s3, const Q3<T>& s4, bool fl, bool fr,
                         const EulerCoef<T>& k)
{
    const Q3<T> f = eu_net_flux(s0, s1, s2, s3, s4, fl, fr, k);
    const T c = T(0.25) * k.dtdx;
    return Q3<T>{s2.x + c * f.x, s2.y + c * f.y, s2.z + c * f.z};
}

// Euler:353-389 (eulerFinalStep) plus the "+=" of Euler:420: u + (1/2 dt/dx) flux.
template <typename T>
SW1D_HD Q3<T> eu_correct(const Q3<T>& u, const Q3<T>& s0, const Q3<T>& s1,
                         const Q3<T>& s2, const Q3<T>& s3, const Q3<T>& s4, bool fl,
                         bool fr, const EulerCoef<T>& k)
{
    const Q3<T> f = eu_net_flux(s0, s1, s2, s3, s4, fl, fr, k);
    const T c = T(0.5) * k.dtdx;
    return Q3<T>{u.x + c * f.x, u.y + c * f.y, u.z + c * f.z};
}

}  // namespace sw1d
