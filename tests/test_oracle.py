"""Pins the CPU oracle (oracle/swept1d_oracle.c) before it is trusted as the
checker: against the reference's golden fixture and published accuracy bands,
closed-form solutions, known-answer vectors and invariants — each chosen so
that a plausible mistake (a dropped term, a wrong sign or index, a wrong
boundary, a wrong precision) fails it.  CPU only."""
import os

import numpy as np
import pytest
from scipy import fft as sfft

from conftest import GOLDEN


# ---------------------------------------------------------------- helpers ---
def _fixture():
    rows = [l.split() for l in open(os.path.join(GOLDEN, "KSDouble_Official.txt"))]
    header = rows[0]
    data = {float(r[1]): np.array(r[2:], float) for r in rows[1:]}
    return header, data


def _rmse(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)))


# ---------------------------------------------------------- RNG and ICs ---
def test_splitmix64_known_answers(orc):
    # published reference outputs of splitmix64 for seed 1234567
    assert orc.splitmix64_stream(1234567, 5) == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]


def test_heat_reference_ic_closed_form(orc):
    n = 1024
    u = orc.heat_ic_ref(n)
    ds = 0.001
    lx = ds * (n - 1)
    x = np.arange(n) * ds
    assert np.allclose(u, 100 * x * (1 - x / lx), rtol=1e-13, atol=1e-13)
    assert u[0] == 0 and abs(u[-1]) < 1e-12


def test_ks_reference_ic_matches_fixture_row0(orc):
    header, data = _fixture()
    assert header == ["512", "1024", "0.5"]
    u0 = orc.ks_ic_ref(1024)
    assert np.max(np.abs(u0 - data[0.0])) < 1e-9  # 10 printed digits


def test_euler_reference_ic_and_constants(orc):
    n = 1024
    k, bd = orc.euler_consts(1e-4, n)
    assert k[0] == 1.4 and k[1] == 0.4
    assert k[2] == 1e-4 / (1.0 / (n - 2))  # dt/dx with dx = lx/(n-2), Euler:1144
    assert list(bd) == [1.0, 0.0, 1.0 / 0.4, 0.125, 0.0, 0.1 / 0.4]
    q = orc.euler_ic_ref(n).reshape(n, 3)
    assert (q[: n // 2] == bd[:3]).all() and (q[n // 2:] == bd[3:]).all()


def _uniforms(orc, seed, count):
    """The generator's uniform draws: the top 53 bits of each splitmix64 output."""
    return np.array([(v >> 11) * 2.0 ** -53 for v in orc.splitmix64_stream(seed, count)])


@pytest.mark.parametrize("seed,L", [(9, 1.0), (5, 1.024), (123456789, 3.5)])
def test_heat_fourier_ic(orc, seed, L):
    """The amplitudes are the first four uniform draws of the pinned generator,
    and a DFT of the field recovers exactly them as sine modes 1..4 (any wrong
    mode number, phase, wavelength or offset leaves other coefficients)."""
    n = 1024
    u, amp = orc.ic_heat_fourier(n, seed, L=L)
    assert np.array_equal(amp, _uniforms(orc, seed, 4))
    c = np.fft.rfft(u) / n
    assert abs(c[0] - 1.0) < 1e-14
    for k in range(1, 5):  # a sin(2 pi k j / n) has DFT coefficient -i a / 2
        assert abs(c[k].real) < 1e-14 and abs(-2 * c[k].imag - amp[k - 1]) < 1e-14
    assert np.max(np.abs(c[5:])) < 1e-14


def test_heat_fourier_amplitudes_are_uniform(orc):
    """Over many seeds the amplitudes fill [0, 1) uniformly (Kolmogorov-Smirnov)."""
    from scipy import stats

    amps = np.concatenate([orc.ic_heat_fourier(8, s)[1] for s in range(2000)])
    assert ((amps >= 0) & (amps < 1)).all()
    assert stats.kstest(amps, "uniform").pvalue > 1e-3


def test_ks_noise_is_gaussian_with_stated_sigma(orc):
    """u - 2cos(19 pi x/128) at 2^20 points is N(0, 1e-3) noise: sample mean
    within 5 standard errors of 0, standard deviation within 0.5% of 1e-3
    (7 standard errors), a Kolmogorov-Smirnov test against N(0, 1e-3) and no
    lag-1 correlation.  A wrong scale, a missing sqrt(-2 ln u1), a biased or
    reused uniform, or the sine in place of the cosine of the base field
    fails one of them."""
    from scipy import stats

    n, sigma = 1 << 20, 1e-3
    u = orc.ic_ks_noisy(n, 1)
    x = np.arange(n) * 0.5
    r = u - 2.0 * np.cos(19.0 * x * np.pi / 128.0)
    assert abs(r.mean()) < 5e-6
    assert abs(r.std() / sigma - 1.0) < 5e-3
    assert stats.kstest(r / sigma, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(r[:-1], r[1:])[0, 1]) < 5.0 / np.sqrt(n)
    # two uniform draws per point, in the Box-Muller cosine form
    u1, u2 = _uniforms(orc, 1, 2)
    assert abs(r[0] - sigma * np.sqrt(-2 * np.log(1 - u1)) * np.cos(2 * np.pi * u2)) < 1e-15
    base = 2.0 * np.cos(19.0 * np.arange(16) * 0.5 * np.pi / 128.0)
    assert np.max(np.abs(orc.ic_ks_noisy(16, 1, sigma=0.0) - base)) < 1e-15


# ------------------------------------------------------------------- heat ---
def test_heat_mirror_matches_discrete_modal_solution(orc):
    """Insulated ends: DCT-I modes cos(pi k j/(n-1)) are the exact eigenvectors
    of the mirrored FTCS operator, eigenvalue 1 - 4 Fo sin^2(pi k / (2(n-1)))."""
    n, steps = 512, 3000
    fo, fo2 = orc.heat_consts(0.002)
    u0 = orc.heat_ic_ref(n)
    got = orc.heat_run(u0, steps, fo, fo2, 0)
    k = np.arange(n)
    lam = 1 - 4 * fo * np.sin(np.pi * k / (2 * (n - 1))) ** 2
    want = sfft.idct(sfft.dct(u0, type=1) * lam ** steps, type=1)
    assert np.max(np.abs(got - want)) < 1e-11 * np.max(np.abs(u0))


def test_heat_periodic_matches_discrete_fourier_solution(orc):
    n, steps = 1024, 2000
    dx = 1.0 / n
    dt = 0.25 * dx * dx / 8.418e-5  # Fo = 0.25
    fo, fo2 = orc.heat_consts(dt, ds=dx)
    assert abs(fo - 0.25) < 1e-12
    u0, _ = orc.ic_heat_fourier(n, 4)
    got = orc.heat_run(u0, steps, fo, fo2, 1)
    k = np.fft.fftfreq(n, 1.0 / n)
    lam = 1 - 4 * fo * np.sin(np.pi * k / n) ** 2
    want = np.fft.ifft(np.fft.fft(u0) * lam ** steps).real
    assert np.max(np.abs(got - want)) < 1e-12


def test_heat_reference_accuracy_band(orc):
    """testProcedure.py:63-82 series solution vs classic heat, nX=1024:
    published RMSE 1.84e-2 .. 1.865e-2 at t = 15..50 (the L = n*dx of the
    series against the program's lx = dx*(n-1) dominates it)."""
    n, dt, alpha = 1024, 1e-3, 8.418e-5
    fo, fo2 = orc.heat_consts(dt)
    u = orc.heat_ic_ref(n)
    L = n * 0.001
    xm = np.linspace(0, L, n)
    c = np.arange(2, 2 * 5002 + 1, 2)[:, None]
    done = 0
    for t in (15.0, 30.0, 45.0, 50.0):
        steps = int(round(t / dt))
        u = orc.heat_run(u, steps - done, fo, fo2, 0)
        done = steps
        h = np.sum(1.0 / c ** 2 * np.exp(-alpha * t * (c * np.pi / L) ** 2)
                   * np.cos(c * xm[None, :] * np.pi / L), axis=0)
        exact = 50.0 * L / 3.0 - 400.0 * L / np.pi ** 2 * h
        assert 1.80e-2 < _rmse(u, exact) < 1.90e-2, t


# --------------------------------------------------------------------- KS ---
@pytest.mark.parametrize("dt,bounds", [
    (5e-3, {6: 7.7e-7, 12: 3.2e-7, 18: 9.6e-8, 20: 6.3e-8}),
    (1e-3, {6: 3.1e-8, 12: 1.3e-8, 18: 3.9e-9, 20: 2.5e-9}),
])
def test_ks_against_reference_fixture(orc, dt, bounds):
    """KSDouble_Official.txt was written by the reference's swept binary at
    dt=1e-7 (testProcedure.py:51-57); the oracle must converge to it as
    O(dt^2)."""
    _, data = _fixture()
    c = orc.ks_consts(dt)
    u = orc.ks_ic_ref(1024)
    t = 0
    for T in (6, 12, 18, 20):
        u = orc.ks_run(u, int(round((T - t) / dt)), c)
        t = T
        r = _rmse(u, data[float(T)])
        assert 0.7 * bounds[T] < r < 1.3 * bounds[T], (T, r)


def test_ks_linear_mode_amplification(orc):
    """A tiny single Fourier mode obeys the linearised scheme: per step the
    RK2-midpoint factor 1 + z + z^2/2 with z = dt*(4s/dx^2 - 16 s^2/dx^4),
    s = sin^2(theta/2) (central 2nd and 4th differences)."""
    n, dt, dx, steps, eps = 256, 0.005, 0.5, 400, 1e-9
    for kk in (3, 17, 40):
        th = 2 * np.pi * kk / n
        u0 = eps * np.cos(th * np.arange(n))
        u = orc.ks_run(u0, steps, orc.ks_consts(dt))
        s = np.sin(th / 2) ** 2
        z = dt * (4 * s / dx ** 2 - 16 * s ** 2 / dx ** 4)
        g = 1 + z + z * z / 2
        assert np.allclose(u, u0 * g ** steps, rtol=1e-6, atol=1e-6 * eps * abs(g) ** steps)


def test_ks_conserves_sum_and_is_shift_equivariant(orc):
    n = 1024
    u0 = orc.ic_ks_noisy(n, 3)
    c = orc.ks_consts(0.005)
    u = orc.ks_run(u0, 500, c)
    assert abs(u.sum() - u0.sum()) < 1e-10 * np.abs(u0).sum()
    us = orc.ks_run(np.roll(u0, 37), 500, c)
    assert np.array_equal(us, np.roll(u, 37))


def test_ks_single_precision_tracks_double(orc):
    n = 1024
    u0 = orc.ks_ic_ref(n)
    d = orc.ks_run(u0, 400, orc.ks_consts(0.005))
    f = orc.ks_run(u0.astype(np.float32), 400, orc.ks_consts(0.005, np.float32))
    assert f.dtype == np.float32
    assert np.max(np.abs(f - d)) < 1e-4


# ------------------------------------------------------------------ Euler ---
GAM = 1.4


def sod_exact(x, t):
    """Toro's exact Sod solution, from the analysis module (pinned below by the
    published star state)."""
    import importlib

    return importlib.import_module("1dsweptcuda_amd.harness").sod_exact(x, t)


def test_sod_exact_star_state():
    # Toro, table 4.2 (test 1)
    _, (p, u, rl, rr) = sod_exact(np.array([0.5]), 0.1)
    assert abs(p - 0.30313) < 1e-5 and abs(u - 0.92745) < 1e-5
    assert abs(rl - 0.42632) < 1e-5 and abs(rr - 0.26557) < 1e-5


def test_euler_against_exact_sod_bands(orc):
    """testProcedure.py:84-96 / published RMSE at nX=1024, t = 0.066..0.22:
    rho <= 8.8e-3, p <= 4.5e-3, u <= 1.8e-2, e <= 5.7e-2."""
    n, dt = 1024, 1e-4
    k, bd = orc.euler_consts(dt, n)
    q = orc.euler_ic_ref(n)
    steps = 2200
    q = orc.euler_run(q, steps, k, bd)
    rho, vel, en, pr = orc.euler_outputs(q, k[1])
    dx = 1.0 / (n - 2)
    r = np.arange(0.5 * dx, 1.0 - 0.5 * dx, dx)[: n - 2]
    ex, _ = sod_exact(r, steps * dt)
    errs = [_rmse(v[1:-1], e) for v, e in zip((rho, vel, en, pr), ex)]
    bands = (8.8e-3, 1.8e-2, 5.7e-2, 4.5e-3)
    for e, b in zip(errs, bands):
        assert 0.3 * b < e < 1.15 * b, (errs, bands)


def test_euler_interior_uniform_state_and_mass(orc):
    n, dt = 2048, 1e-4
    k, bd = orc.euler_consts(dt, n)
    q0 = orc.euler_ic_ref(n)
    q = orc.euler_run(q0, 20, k, bd).reshape(n, 3)
    # far from the discontinuity and the ends a uniform state stays bit-exact
    assert (q[10:900] == q0.reshape(n, 3)[10:900]).all()
    assert (q[1150:2030] == q0.reshape(n, 3)[1150:2030]).all()
    # mass is conserved to rounding while no wave reaches the ends (u=0 there)
    q2 = orc.euler_run(q0, 1500, k, bd).reshape(n, 3)
    m0, m1 = q0.reshape(n, 3)[1:-1, 0].sum(), q2[1:-1, 0].sum()
    assert abs(m1 - m0) < 1e-12 * m0


def test_euler_single_precision_evaluation(orc):
    """The float build evaluates every operation in float (sqrtf, fabsf and
    float additions, Euler1D_SweptShared.cu:47-57), not in promoted double:
    one corrector stage at a state with small momenta, against an independent
    float32 evaluation of the same formulas."""
    f = np.float32
    n = 64
    k, bd = orc.euler_consts(1e-3, n, f)
    rng = np.random.default_rng(5)
    rho = (1 + 0.1 * rng.standard_normal(n)).astype(f)
    m = (1e-3 * rng.standard_normal(n)).astype(f)
    E = (2.5 + 0.1 * rng.standard_normal(n)).astype(f)
    q = np.stack([rho, m, E], 1).astype(f).ravel()
    out = q.copy()
    orc.euler_stage(q, out, k, bd, 1)  # u = q + 1/2 dt/dx F(q)
    gam, mgam, dtdx = k

    def pdiff(a, b):
        return f(f(f(f(2) * b[0]) * a[0]) * f(b[2] - a[2]) + f(f(f(a[1] * a[1]) * b[0]) - f(f(b[1] * b[1]) * a[0])))

    def limit(c, o, r):
        mm = f(f(0.5) * (r if r < f(1) else f(1)))
        return tuple(f(c[i] + f(mm * f(o[i] - c[i]))) for i in range(3))

    def pres(c):
        return f(mgam * f(c[2] - f(f(f(f(0.5) * c[1]) * c[1]) / c[0])))

    def fd(L, R):
        uL, uR = f(L[1] / L[0]), f(R[1] / R[0])
        pL, pR = pres(L), pres(R)
        F = (f(L[1] + R[1]), f(f(f(f(L[1] * uL) + f(R[1] * uR)) + pL) + pR),
             f(f(f(f(L[2] * uL) + f(R[2] * uR)) + f(uL * pL)) + f(uR * pR)))
        sl, sr = f(np.sqrt(L[0])), f(np.sqrt(R[0]))
        den = f(f(1) / f(f(sl * sr) * f(sl + sr)))
        hy = f(f(f(sl * R[1]) + f(sr * L[1])) * den)
        hz = f(f(f(sl * R[2]) + f(sr * L[2])) * den)
        s = f(f(np.sqrt(f(f(mgam * f(hz - f(f(f(0.5) * hy) * hy))) * gam))) + f(abs(hy)))
        return F, tuple(f(s * f(L[i] - R[i])) for i in range(3))

    qq = q.reshape(n, 3)
    with np.errstate(divide="ignore", invalid="ignore"):
        for c in range(3, n - 3):
            s0, s1, s2, s3, s4 = (tuple(qq[c - 2 + j]) for j in range(5))
            pLL, pL, pR, pRR = pdiff(s0, s1), pdiff(s1, s2), pdiff(s2, s3), pdiff(s3, s4)
            mid = (not pL) or (not pR) or ((pL < 0) != (pR < 0))
            tL = s1 if ((not pLL) or (not pL) or ((pLL < 0) != (pL < 0))) else \
                limit(s1, s2, f(f(s0[0] * pL) / f(s2[0] * pLL)))
            tR = s2 if mid else limit(s2, s1, f(f(s3[0] * pL) / f(s1[0] * pR)))
            F, D = fd(tL, tR)
            acc = [f(F[i] + D[i]) for i in range(3)]
            tL = s2 if mid else limit(s2, s3, f(f(s1[0] * pR) / f(s3[0] * pL)))
            tR = s3 if ((not pRR) or (not pR) or ((pRR < 0) != (pR < 0))) else \
                limit(s3, s2, f(f(s4[0] * pR) / f(s2[0] * pRR)))
            F, D = fd(tL, tR)
            acc = [f(f(acc[i] - F[i]) - D[i]) for i in range(3)]
            cc = f(f(0.5) * dtdx)
            want = [f(qq[c][i] + f(cc * acc[i])) for i in range(3)]
            assert list(out.reshape(n, 3)[c]) == want, c
