// Internal interface between the host plan code (sw1d_plan.cpp) and the
// kernels (sw1d_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "sw1d_stencil.hpp"

namespace sw1d {

enum Eq : int { EQ_HEAT = 0, EQ_KS = 1, EQ_EULER = 2 };
enum Bc : int { BC_MIRROR = 1, BC_PERIODIC = 2, BC_DIRICHLET = 3 };

// Swept tile kinds.  UP: up-triangle only (Heat:147-196, KS:209-263,
// Euler:437-484).  DOWN: down-triangle only (Heat:198-244, KS:265-308,
// Euler:495-536).  DIAMOND: down then up (wholeDiamond / splitDiamond).
enum Tile : int { TILE_UP = 0, TILE_DOWN = 1, TILE_DIAMOND = 2 };

// Number of values each point carries in the staggered-level state:
// heat (even row, odd row); KS (u, u_half); Euler (rho, m, E, rho_half,
// m_half, E_half).
inline int eq_components(int eq) { return eq == EQ_EULER ? 6 : 2; }
// Stencil radius of one stage (heat: one FTCS step; KS/Euler: one RK2 stage).
inline int eq_radius(int eq) { return eq == EQ_HEAT ? 1 : 2; }
// RK stages per time step.
inline int eq_stages(int eq) { return eq == EQ_HEAT ? 1 : 2; }

// One swept launch over `nnodes` nodes of W = 64*P points.  Node k covers the
// points (k*W + offset + j) mod n, j = 0..W-1.  `in` and `out` point at
// component-major arrays of nc*n values.
struct SweptLaunch {
    int eq, bc, P, tile;
    int64_t n, nnodes, node_first, offset;
    const void* in;
    void* out;
    int simds = 0;  // SIMDs of the device (0: unknown), which picks the stage form
};

struct Coefs {
    // heat
    double fo_d, fo2_d; float fo_f, fo2_f;
    KsCoef<double> ks_d; KsCoef<float> ks_f;
    EulerCoef<double> eu_d; EulerCoef<float> eu_f;
};

// One trapezoid launch: tile t stores the points [t*W_out, t*W_out + W_out)
// (the last tile clipped at nX) after S stages, W_out = 64*PO, from a window
// of 64*(PO+SB) points; S = 32*SB stages for heat and 16*SB for KS / Euler.
struct TrapLaunch {
    int eq, bc, PO, SB;
    int64_t n, ntiles, tile_first;
    const void* in;
    void* out;
};

// One problem of a batch.  Periodic heat (c = {Fo, 1-2Fo}) or KS
// (c = {1/(4dx), 1/dx^2, 1/dx^4, dt, dt/2}), stored at `offset` in the packed
// state arrays, advanced with trapezoid tiles of 64*po output points.
template <typename T> struct BatchJob {
    int64_t offset;
    int32_t n, eq, po, pad;
    T c[5];
};

struct BatchLaunch {
    const void* jobs;  // device BatchJob<T>[njobs], in queue order
    int32_t njobs;
    int* next;         // device work-queue counter (set to 0 by the launch)
    void* a;           // packed state; receives the result
    void* b;           // packed scratch of the same size
    int64_t nsteps;
    int32_t grid;      // persistent workgroups
    int32_t waves;     // waves (tiles worked on at once) per workgroup: 2, 4 or 8
};

}  // namespace sw1d

struct sw1d_batch;
namespace sw1d {
// Device view of a batch for the surrogate integrator (sw1d_nn.hip).
struct BatchView {
    void* state;            // packed state (problem i at offset i*n when uniform)
    int64_t count, n;       // problems; their common size (0 if sizes differ)
    bool f64, all_ks;
    hipStream_t stream;
};
bool batch_view(sw1d_batch* b, BatchView& v);

// Launchers (sw1d_kernels.hip, sw1d_trapezoid.hip, sw1d_batch.hip).  All
// return a hipError_t.
hipError_t launch_batch(const BatchLaunch& L, bool f64, hipStream_t s);
int batch_blocks_per_cu(bool f64, int waves);
inline bool batch_waves_supported(int w) { return w == 2 || w == 4 || w == 8; }
inline bool batch_tile_supported(int po) { return po == 2 || po == 4 || po == 8 || po == 16; }
hipError_t launch_trapezoid(const TrapLaunch& L, bool f64, const Coefs& c, hipStream_t s);
bool trapezoid_supported(int eq, int PO, int SB);
inline int64_t trapezoid_steps(int eq, int SB) { return eq == EQ_HEAT ? 32 * SB : 8 * SB; }
hipError_t launch_swept(const SweptLaunch& L, bool f64, const Coefs& c, hipStream_t s);
// One classic stage: heat (stage ignored), KS / Euler stage 0 = predictor,
// 1 = corrector.  State layout as above; heat ping-pongs between in and out,
// KS / Euler keep u and u_half in the same buffer.
hipError_t launch_classic(int eq, int bc, bool f64, int stage, int64_t n, const void* in,
                          void* out, const Coefs& c, hipStream_t s);
// Checks the lane-shift primitive the swept kernels rely on; returns 1 if
// every lane received its neighbour's value, 0 otherwise, <0 on HIP error.
int check_lane_shifts();
// The same check on caller-provided device memory (one int; no allocation).
int check_lane_shifts_at(int* d);

// Points per lane of a swept node (node width = 64 x that).  One is enough
// for every stencil (a radius-2 neighbour two lanes away arrives by a second
// lane shift); an Euler node's LDS image (six components) limits it to 8 in
// either precision, the others to 16.
inline int max_points_per_lane(int eq, bool /*f64*/) { return eq == EQ_EULER ? 8 : 16; }
inline int min_points_per_lane(int) { return 1; }

}  // namespace sw1d
