// Batched persistent solve: thousands of independent periodic heat / KS
// problems of arbitrary sizes advanced by nsteps in one launch.
//
// Each workgroup (NW = 2, 4 or 8 waves: the tiles it works on at once, a
// batch setting) takes whole problems from a device work queue (an atomic
// counter over a table sorted largest first), and advances its problem through
// all the steps itself: trapezoid tiles (sw1d_trap.hpp) over the problem, one
// wave per tile, ping-ponging between the two packed state arrays with a
// workgroup barrier between time blocks.  The tiles' arithmetic does not
// depend on NW, so neither do the results.  Steps that do not fill
// a time block are taken one stage at a time.  No workgroup ever waits for
// another, so the launch cannot deadlock, and every wave leaves when the
// queue is empty.
//
// The counter is reset by a one-thread kernel launched just before (no
// memset), so the pair can be captured in a hipGraph and replayed.

#include "sw1d_trap.hpp"

namespace sw1d {
namespace {

__device__ __forceinline__ int64_t wrap1(int64_t i, int64_t n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

template <typename T, int EQ, int PO, int NW>
__device__ void batch_problem(const BatchJob<T>& J, T* a, T* b, int64_t nsteps, int wave, int lane)
{
    constexpr int NT = NW * kWave;
    constexpr int PI = PO + 1, WOUT = kWave * PO, RS = 32;
    constexpr int S = EQ == EQ_HEAT ? 32 : 16;          // stages per time block
    constexpr int BLOCK_STEPS = EQ == EQ_HEAT ? 32 : 8;
    const int64_t n = J.n;
    const int64_t ntiles = (n + WOUT - 1) / WOUT;
    const int64_t blocks = nsteps / BLOCK_STEPS;
    KsCoef<T> k;
    k.dx4x = J.c[0]; k.dx2 = J.c[1]; k.dx4 = J.c[2]; k.dt = J.c[3]; k.dth = J.c[4];
    for (int64_t tb = 0; tb < blocks; ++tb) {
        for (int64_t t = wave; t < ntiles; t += NW) {
            const int64_t out0 = t * WOUT;
            const int count = (int)(n - out0 < WOUT ? n - out0 : WOUT);
            const int64_t win = wrapn(out0 - RS, n);
            if (EQ == EQ_HEAT)
                trap_heat<T, PI, S, false>(a, b, n, win, out0, count, lane, J.c[0], J.c[1]);
            else
                trap_ks<T, PI, S>(a, b, n, win, out0, count, lane, k);
        }
        __syncthreads();
        T* t = a; a = b; b = t;
    }
    // remaining steps, one stage per pass (Heat:132-145, KS:189-207)
    const int64_t rem = nsteps - blocks * BLOCK_STEPS;
    for (int64_t r = 0; r < rem; ++r) {
        if (EQ == EQ_HEAT) {
            for (int64_t i = threadIdx.x; i < n; i += NT)
                b[i] = heat_ftcs(a[wrap1(i - 1, n)], a[i], a[wrap1(i + 1, n)], J.c[0], J.c[1]);
            __syncthreads();
            T* t = a; a = b; b = t;
        } else {
            for (int64_t i = threadIdx.x; i < n; i += NT)
                b[i] = ks_predict(a[wrap1(i - 2, n)], a[wrap1(i - 1, n)], a[i], a[wrap1(i + 1, n)],
                                  a[wrap1(i + 2, n)], k);
            __syncthreads();
            for (int64_t i = threadIdx.x; i < n; i += NT)
                a[i] = ks_correct(a[i], b[wrap1(i - 2, n)], b[wrap1(i - 1, n)], b[i],
                                  b[wrap1(i + 1, n)], b[wrap1(i + 2, n)], k);
            __syncthreads();
        }
    }
}

template <typename T, int NW>
__global__ __launch_bounds__(NW * kWave) void k_batch(const BatchJob<T>* __restrict__ jobs,
                                                         int njobs, int* __restrict__ next,
                                                         T* __restrict__ A, T* __restrict__ B,
                                                         int64_t nsteps)
{
    __shared__ int s_job;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
    for (;;) {
        __syncthreads();  // every wave has read the previous s_job
        if (threadIdx.x == 0) s_job = atomicAdd(next, 1);
        __syncthreads();
        const int j = s_job;
        if (j >= njobs) break;
        const BatchJob<T> J = jobs[j];
        T* a = A + J.offset;
        T* b = B + J.offset;
        const int blk = J.eq == EQ_HEAT ? 32 : 8;
        int64_t swaps = nsteps / blk + (J.eq == EQ_HEAT ? nsteps % blk : 0);
        if (J.eq == EQ_HEAT) {
            if (J.po == 2) batch_problem<T, EQ_HEAT, 2, NW>(J, a, b, nsteps, wave, lane);
            else if (J.po == 4) batch_problem<T, EQ_HEAT, 4, NW>(J, a, b, nsteps, wave, lane);
            else if (J.po == 8) batch_problem<T, EQ_HEAT, 8, NW>(J, a, b, nsteps, wave, lane);
            else batch_problem<T, EQ_HEAT, 16, NW>(J, a, b, nsteps, wave, lane);
        } else {
            if (J.po == 2) batch_problem<T, EQ_KS, 2, NW>(J, a, b, nsteps, wave, lane);
            else if (J.po == 4) batch_problem<T, EQ_KS, 4, NW>(J, a, b, nsteps, wave, lane);
            else if (J.po == 8) batch_problem<T, EQ_KS, 8, NW>(J, a, b, nsteps, wave, lane);
            else batch_problem<T, EQ_KS, 16, NW>(J, a, b, nsteps, wave, lane);
        }
        // the result ends in b after an odd number of ping-pongs: copy it home
        if (swaps & 1) {
            for (int64_t i = threadIdx.x; i < J.n; i += NW * kWave) a[i] = b[i];
        }
    }
}

__global__ void k_reset_counter(int* c)
{
    if (threadIdx.x == 0) *c = 0;
}

template <typename T, int NW>
hipError_t batch_launch(const BatchLaunch& L, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, L.next);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_batch<T, NW>), dim3(L.grid), dim3(NW * kWave), 0, s,
                       static_cast<const BatchJob<T>*>(L.jobs), L.njobs, L.next,
                       static_cast<T*>(L.a), static_cast<T*>(L.b), L.nsteps);
    return hipGetLastError();
}

template <typename T>
hipError_t batch_prec(const BatchLaunch& L, hipStream_t s)
{
    switch (L.waves) {
    case 2: return batch_launch<T, 2>(L, s);
    case 8: return batch_launch<T, 8>(L, s);
    default: return batch_launch<T, 4>(L, s);
    }
}

template <typename T, int NW>
int occupancy()
{
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<T, NW>, NW * kWave, 0);
    return e == hipSuccess && nb > 0 ? nb : 1;
}

}  // namespace

hipError_t launch_batch(const BatchLaunch& L, bool f64, hipStream_t s)
{
    return f64 ? batch_prec<double>(L, s) : batch_prec<float>(L, s);
}

int batch_blocks_per_cu(bool f64, int waves)
{
    switch (waves) {
    case 2: return f64 ? occupancy<double, 2>() : occupancy<float, 2>();
    case 8: return f64 ? occupancy<double, 8>() : occupancy<float, 8>();
    default: return f64 ? occupancy<double, 4>() : occupancy<float, 4>();
    }
}

}  // namespace sw1d
