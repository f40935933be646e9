/*
 * swept1d_oracle — CPU oracle for the 1D swept-rule solver.
 *
 * TEST INFRASTRUCTURE.  Only the tests, the smoke check and the CPU-baseline
 * leg of bench.py load this library, and only as the checker.  The HIP library
 * (libswept1d) neither links nor calls it.
 *
 * It restates the reference's classic algorithm (one pass per RK stage, the
 * formulas of Heat1D_SweptShared.cu, KS1D_SweptShared.cu and
 * Euler1D_SweptShared.cu) serially in plain C, in both precisions, plus the
 * reference's initial conditions and constant setup and the seeded synthetic
 * initial conditions of the benchmark workloads.
 *
 * Parity status of each part (what pins it, see tests/test_oracle.py):
 *   heat stage      discrete-exact modal solution (mirror: DCT-I modes,
 *                   periodic: DFT modes), published series band
 *   KS stages       reference fixture KSDouble_Official.txt (O(dt^2) RMSE),
 *                   linearised single-mode amplification, sum conservation,
 *                   shift equivariance
 *   Euler stages    exact Sod Riemann solution (published RMSE bands),
 *                   uniform state stays uniform, mass conservation
 *   constants, ICs  fixture row t=0 (KS), closed forms (heat, Sod);
 *                   splitmix64 known answers; the seeded heat amplitudes are
 *                   the generator's first uniform draws and are recovered from
 *                   the field by a DFT; the KS noise is N(0, sigma) by its
 *                   sample mean, variance, Kolmogorov-Smirnov distance and
 *                   lag-1 correlation at 2^20 points
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* The heat and KS stages are built twice, for CPUs with the FMA instructions
 * and for those without (glibc's correctly rounded fma(), the same results,
 * slower); the loader picks the clone once. */
#if defined(__x86_64__)
#define ORC_FMA_CLONES __attribute__((target_clones("fma", "default")))
#else
#define ORC_FMA_CLONES
#endif

#define REAL double
#define SFX _d
#define RSQRT sqrt
#define RFMA fma
#define RFABS fabs
#include "oracle_body.inc"
#undef REAL
#undef SFX
#undef RSQRT
#undef RFMA
#undef RFABS

#define REAL float
#define SFX _f
#define RSQRT sqrtf
#define RFMA fmaf
#define RFABS fabsf
#include "oracle_body.inc"
#undef REAL
#undef SFX
#undef RSQRT
#undef RFMA
#undef RFABS

/* ------------------------------------------------ seeded synthetic ICs --- */
/* splitmix64 (Steele, Lea, Flood 2014), a portable generator so that the same
 * seed gives the same initial condition on every host. */
static uint64_t orc_splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double orc_uniform(uint64_t *s) /* [0,1) */
{
    return (double)(orc_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}
static double orc_normal(uint64_t *s) /* Box-Muller, cosine branch */
{
    const double u1 = 1.0 - orc_uniform(s);
    const double u2 = orc_uniform(s);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* Periodic heat workload: T0(x) = 1 + sum_{k=1..4} a_k sin(2 pi k x / L),
 * a_k ~ U[0,1), x = j*L/n.  Writes the 4 amplitudes to amp (may be NULL). */
void orc_ic_heat_fourier(double *u, int64_t n, double L, uint64_t seed, double *amp)
{
    uint64_t s = seed;
    double a[4];
    for (int k = 0; k < 4; k++) a[k] = orc_uniform(&s);
    if (amp) for (int k = 0; k < 4; k++) amp[k] = a[k];
    for (int64_t j = 0; j < n; j++) {
        const double x = (double)j * (L / (double)n);
        double v = 1.0;
        for (int k = 0; k < 4; k++) v += a[k] * sin(2.0 * M_PI * (double)(k + 1) * x / L);
        u[j] = v;
    }
}

/* KS workload: reference IC 2cos(19 pi x/128) plus N(0, sigma) noise per point. */
void orc_ic_ks_noisy(double *u, int64_t n, double dx, double sigma, uint64_t seed)
{
    uint64_t s = seed;
    for (int64_t k = 0; k < n; k++) {
        const double x = (double)k * dx;
        u[k] = 2.0 * cos(19.0 * x * M_PI / 128.0) + sigma * orc_normal(&s);
    }
}

uint64_t orc_splitmix64_next(uint64_t *state) { return orc_splitmix64(state); }
