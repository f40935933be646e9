// gfx950 kernels of the 1D swept-rule solver.
//
// Swept tiles (up-triangle, down-triangle, diamond) — one wave64 per node.
//   A node is W = 64*P consecutive points; lane l owns points l*P .. l*P+P-1
//   in VGPRs for the whole launch.  The stencil neighbours of a lane's first
//   and last points come from the adjacent lanes through DPP wave shifts
//   (v_mov_b32_dpp wave_shr:1 / wave_shl:1), so a whole triangle or diamond
//   runs with no LDS traffic and no barrier.  Lanes 0 and 63 receive the
//   values beyond the node (its halo) through the DPP "old" operand.
//
//   Each point carries two time levels of its state — heat: the even and odd
//   level rows; KS: (u, u_half); Euler: the conserved triple and its
//   predictor — at its own position in the tile, so the data
//   exchanged between launches (the reference's right/left edge arrays,
//   Heat:69-119, KS:75-109, Euler:95-174) is simply one staggered-level state
//   array per launch, read once and written once, coalesced.
//
//   Stage s of a tile updates exactly the points inside the triangle or
//   diamond at that level: down phase |2x-(W-1)| < 2Rs, up phase
//   |2x-(W-1)| < W - 2Rm (R = stencil radius of one stage).  A point outside
//   keeps its value, so a tile computes every point-update of the classic
//   scheme once, with the same operands, and the result is bitwise the
//   classic one (the reference's own consistency invariant,
//   ResultPlots/ExactTesting/consistency.out).
//
//   Node placement: node k covers (k*W + offset + j) mod n.  offset 0 gives
//   the node-aligned tiles (upTriangle, wholeDiamond, downTriangle); offset
//   n - W/2 gives the tiles centred on node boundaries, whose node 0 straddles
//   the domain ends (splitDiamond, Heat:335-408, Euler:666-746).  Boundary
//   conditions are applied by global index wherever the boundary falls.
//
// Classic kernels — one thread per point, one launch per stage
//   (classicHeat Heat:132-145, classicKS KS:189-207, classicEuler Euler:400-426).

#include <type_traits>
#include <utility>

#include "sw1d_internal.hpp"
#include "sw1d_lanes.hpp"

namespace sw1d {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr int kClassicBlock = 256;


template <typename T, int P>
__device__ __forceinline__ void load_pts(const T* __restrict__ a, int64_t i0, T (&v)[P])
{
#pragma unroll
    for (int p = 0; p < P; ++p) v[p] = a[i0 + p];
}
template <typename T, int P>
__device__ __forceinline__ void store_pts(T* __restrict__ a, int64_t i0, const T (&v)[P])
{
#pragma unroll
    for (int p = 0; p < P; ++p) a[i0 + p] = v[p];
}

struct NodeGeom {
    int64_t base;  // global index of node-local x = 0
    int64_t lb;    // global index of this lane's first point
    int lane;
};

template <int P>
__device__ __forceinline__ bool node_geom(int64_t n, int64_t nnodes, int64_t node_first,
                                          int64_t offset, NodeGeom& g)
{
    constexpr int W = kWave * P;
    g.lane = threadIdx.x & (kWave - 1);
    const int64_t node = node_first + (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (node >= nnodes) return false;
    int64_t b = node * W + offset;
    b = b >= n ? b - n : b;
    g.base = b;
    int64_t lb = b + (int64_t)g.lane * P;
    g.lb = lb >= n ? lb - n : lb;
    return true;
}

__device__ __forceinline__ int64_t wrap(int64_t i, int64_t n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// Calls f(integral_constant<int, i>) for i = 0, 1, .. in order.
template <typename F, int... I>
__device__ __forceinline__ void for_each_index(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}

// Calls f(integral_constant<int, j>) for j = A, A + 1, .. B in order (none if A > B).
template <int A, int B, typename F>
__device__ __forceinline__ void for_range(F&& f)
{
    if constexpr (A <= B) {
        f(std::integral_constant<int, A>{});
        for_range<A + 1, B>(f);
    }
}

// EXEC-masked stage fmas with a compile-time threshold.  Lane l of a centred
// region of PC points per lane holds d = |2(l PC + p) - 64 PC + 1| at point
// p, so each point's set of active lanes is known with the threshold: a
// point active in all 64 lanes takes its fma with no mask, a point active in
// none is not evaluated, and the points of one lane set run their fmas under
// one branch on the first such point's d, so EXEC is narrowed and restored
// once per set instead of once per point.  The fmas are inline asm, which
// the compiler keeps inside the branch.
template <int PC>
struct LaneSets {
    static constexpr uint64_t of(int p, int thr)
    {
        uint64_t m = 0;
        for (int l = 0; l < 64; ++l) {
            const int v = 2 * (l * PC + p) - 64 * PC + 1;
            if ((v < 0 ? -v : v) <= thr) m |= uint64_t(1) << l;
        }
        return m;
    }
    static constexpr int first(int p, int thr)
    {
        for (int q = 0; q < p; ++q)
            if (of(q, thr) == of(p, thr)) return q;
        return p;
    }
};
// (FP32 only: in FP64 the branches measured slower than one mask per point)
__device__ __forceinline__ void fma_lanes(float& dst, float a, float b, float c)
{
    asm volatile("v_fma_f32 %0, %1, %2, %3" : "+v"(dst) : "v"(a), "v"(b), "v"(c));
}
// dst[p] = fma(a, rhs(p), base(p)) in the lanes where d[p] <= THR
template <typename T, int PC, int THR, typename FR, typename FB>
__device__ __forceinline__ void fma_lane_sets(T (&dst)[PC], const int (&d)[PC], T a, FR&& rhs, FB&& base)
{
    using S = LaneSets<PC>;
    // the stage's own copy of the threshold: the compiler would otherwise
    // merge equal compares of distant stages (the two halves of a diamond
    // repeat thresholds) and keep their lane masks live in between
    int thr = THR;
    asm volatile("" : "+s"(thr));
    T f[PC];
    for_range<0, PC - 1>([&](auto pi) {
        constexpr int p = decltype(pi)::value;
        if constexpr (S::of(p, THR) != 0) f[p] = rhs(p);
    });
    for_range<0, PC - 1>([&](auto pi) {
        constexpr int p = decltype(pi)::value;
        constexpr uint64_t m = S::of(p, THR);
        if constexpr (m == ~uint64_t(0)) {
            dst[p] = fmadd(a, f[p], base(p));
        } else if constexpr (m != 0 && S::first(p, THR) == p) {
            if (d[p] <= thr) {
                for_range<p, PC - 1>([&](auto qi) {
                    constexpr int q = decltype(qi)::value;
                    if constexpr (S::of(q, THR) == m) fma_lanes(dst[q], a, f[q], base(q));
                });
            }
        }
    });
}

// ----------------------------------------------------------- heat tiles ---
// Two values per point, the even-level and the odd-level row (the two rows of
// the reference's working array, Heat:151-164): a point just inside a
// triangle edge is one level ahead of its outer neighbour, and the neighbour
// needs the level before.  Odd stages write row 1 from row 0, even stages
// row 0 from row 1.  HeatRegion holds PC points per lane of the middle
// 64*PC points of a node (all of it when PC = P).
template <typename T, int PC, bool MIRROR, bool XM>
struct HeatRegion {
    T r0[PC], r1[PC];
    T hl0[1], hr0[1], hl1[1], hr1[1];
    int d[PC];
    bool g0, gN;  // this lane's first / last point is global point 0 / n-1
    bool bnd;     // some lane of the wave has g0 or gN (the same in every lane)

    // MB: this wave holds point 0 or n-1 and applies the insulated ends.
    template <bool MB>
    __device__ __forceinline__ void stage_bc(int thr, const T (&src)[PC], T (&hl)[1], T (&hr)[1],
                                             T (&dst)[PC], T fo, T fo2)
    {
        T e[PC + 2];
        gather_carry<T, PC, 1>(src, hl, hr, e);
        // Insulated ends (Heat:140-142): the left neighbour of point 0 is
        // point 1, the right neighbour of point n-1 is point n-2.
        if constexpr (MB) {
            e[0] = g0 ? e[2] : e[0];
            e[PC + 1] = gN ? e[PC - 1] : e[PC + 1];
        }
#pragma unroll
        for (int p = 0; p < PC; ++p) {
            if constexpr (XM) {
                // heat_ftcs = fma(Fo, l + r, (1-2Fo) c), its fma in the active lanes
                fma_where(dst[p], fo, e[p] + e[p + 2], fo2 * e[p + 1], d[p], thr);
            } else {
                const T nv = heat_ftcs(e[p], e[p + 1], e[p + 2], fo, fo2);
                dst[p] = (d[p] <= thr) ? nv : dst[p];
            }
        }
    }
    // One stage; the level loops below are unswitched on bnd, so the nodes
    // without a domain end (all but one or two) run no boundary select.
    __device__ __forceinline__ void stage(int thr, const T (&src)[PC], T (&hl)[1], T (&hr)[1],
                                          T (&dst)[PC], T fo, T fo2)
    {
        if (MIRROR && bnd)
            stage_bc<MIRROR>(thr, src, hl, hr, dst, fo, fo2);
        else
            stage_bc<false>(thr, src, hl, hr, dst, fo, fo2);
    }
    // down pairs s = s0, s0+2, .. s1 (thresholds 2s-1, 2s+1); up pairs
    // m = m0, .. m1 (thresholds W-1-2m, W-3-2m); U pairs per loop trip
    template <bool MB, int U>
    __device__ __forceinline__ void down_bc(int s0, int s1, T fo, T fo2)
    {
#pragma unroll U
        for (int s = s0; s <= s1; s += 2) {
            stage_bc<MB>(2 * s - 1, r0, hl0, hr0, r1, fo, fo2);
            stage_bc<MB>(2 * s + 1, r1, hl1, hr1, r0, fo, fo2);
        }
    }
    template <bool MB, int U>
    __device__ __forceinline__ void up_bc(int W, int m0, int m1, T fo, T fo2)
    {
#pragma unroll U
        for (int m = m0; m <= m1; m += 2) {
            stage_bc<MB>(W - 1 - 2 * m, r0, hl0, hr0, r1, fo, fo2);
            stage_bc<MB>(W - 3 - 2 * m, r1, hl1, hr1, r0, fo, fo2);
        }
    }
    template <int U>
    __device__ __forceinline__ void down(int s0, int s1, T fo, T fo2)
    {
        if (MIRROR && bnd)
            down_bc<MIRROR, U>(s0, s1, fo, fo2);
        else
            down_bc<false, U>(s0, s1, fo, fo2);
    }
    template <int U>
    __device__ __forceinline__ void up(int W, int m0, int m1, T fo, T fo2)
    {
        if (MIRROR && bnd)
            up_bc<MIRROR, U>(W, m0, m1, fo, fo2);
        else
            up_bc<false, U>(W, m0, m1, fo, fo2);
    }
    // flags and distance codes for a lane whose first point is node-local x0
    __device__ __forceinline__ void place(int W, int x0, int64_t base, int64_t n)
    {
#pragma unroll
        for (int p = 0; p < PC; ++p) d[p] = abs(2 * (x0 + p) - (W - 1));
        int64_t gi = base + x0;
        gi = gi >= n ? gi - n : gi;
        int64_t gl = gi + PC - 1;
        gl = gl >= n ? gl - n : gl;
        g0 = MIRROR && gi == 0;
        gN = MIRROR && gl == n - 1;
        bnd = MIRROR && __builtin_amdgcn_ballot_w64(g0 || gN) != 0;
    }
    template <int W>
    __device__ __forceinline__ void load(const T* img, int lane, int64_t base, int64_t n)
    {
        constexpr int xs = W / 2 - 32 * PC;
        static_assert(xs >= 1 && xs + 64 * PC + 1 <= W, "region and halo inside the node");
        const int x0 = xs + lane * PC;
#pragma unroll
        for (int p = 0; p < PC; ++p) {
            r0[p] = img[x0 + p];
            r1[p] = img[W + x0 + p];
        }
        hl0[0] = img[xs - 1];
        hr0[0] = img[xs + 64 * PC];
        hl1[0] = img[W + xs - 1];
        hr1[0] = img[W + xs + 64 * PC];
        place(W, x0, base, n);
    }
    template <int W>
    __device__ __forceinline__ void save(T* img, int lane) const
    {
        const int x0 = W / 2 - 32 * PC + lane * PC;
#pragma unroll
        for (int p = 0; p < PC; ++p) {
            img[x0 + p] = r0[p];
            img[W + x0 + p] = r1[p];
        }
    }
};

// Ring PC of a node (1 <= PC <= P) is the 64 points that the centred region
// of PC points per lane holds and the region of PC - 1 does not: 32 on each
// side of the centre.  Lane l takes ring point l, so a ring moves between HBM
// and the LDS image as two coalesced 256-byte runs per state array.
template <int W, int PC>
__device__ __forceinline__ int ring_x(int lane)
{
    return lane < 32 ? W / 2 - 32 * PC + lane : W / 2 + 32 * (PC - 1) + (lane - 32);
}

// A node's HBM traffic ring by ring, for tiles whose levels run on centred
// regions: NC state arrays of n values, array c at a + c n in HBM and at
// img + c W in the node's LDS image.  A tile with a down phase issues the
// loads of every ring from the centre out and puts ring PC + 1 (region PC's
// halo) into the image just before region PC, so the outer rings arrive
// while the inner regions compute.  A tile with an up phase stores ring PC as
// soon as region PC has finished it, so the stores drain while the inner
// regions compute.
template <typename T, int P, int NC>
struct NodeRings {
    static constexpr int W = kWave * P;
    T v[P + 1][NC];  // ring PC's point of this lane, PC = 1 .. P
    int64_t hbase;
    int lane;

    // Node bases are multiples of W/2 and n of W, so each half of a node is
    // contiguous in a state array: ring point x (in this lane's half) sits at
    // hbase + x.
    __device__ __forceinline__ NodeRings(const NodeGeom& g, int64_t n)
        : hbase(g.lane < 32 ? g.base : wrap(g.base + W / 2, n) - W / 2), lane(g.lane)
    {
    }
    template <int PC>
    __device__ __forceinline__ void load(const T* a, int64_t n)
    {
        const int64_t gi = hbase + ring_x<W, PC>(lane);
#pragma unroll
        for (int c = 0; c < NC; ++c) v[PC][c] = a[c * n + gi];
    }
    __device__ __forceinline__ void load_all(const T* a, int64_t n)
    {
        for_each_index([&](auto i) { load<decltype(i)::value + 1>(a, n); },
                       std::make_integer_sequence<int, P>{});
    }
    template <int PC>
    __device__ __forceinline__ void put(T* img) const
    {
        const int x = ring_x<W, PC>(lane);
#pragma unroll
        for (int c = 0; c < NC; ++c) img[c * W + x] = v[PC][c];
    }
    template <int PC>
    __device__ __forceinline__ void store(const T* img, T* a, int64_t n) const
    {
        const int x = ring_x<W, PC>(lane);
        const int64_t gi = hbase + x;
#pragma unroll
        // Streaming stores: the next launch reads these values, from whichever
        // XCD its wave lands on; kept out of this XCD's L2 they leave less to
        // write back at the end of the launch (0.9% more point-updates/s, KS
        // diamond at 2^20).  Loads marked the same way gained nothing.
        for (int c = 0; c < NC; ++c) __builtin_nontemporal_store(img[c * W + x], &a[c * n + gi]);
    }
};

// Orders a wave's own LDS writes before its own reads of other lanes' data:
// LDS executes a wave's instructions in order, and the wavefront-scope
// fences keep the compiler from moving one lane's loads above other lanes'
// stores.  Each wave of the tile kernels owns its slice of the LDS image.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Swept heat tiles: one wave per node of W = 64*P points.  For P >= 2 the
// node's image sits in LDS and the diamond's levels run on centred regions of
// 1, 2, .. P points per lane, each level on the narrowest region holding its
// active set and the one-point stencil reach:
// the same operations, so the same bits, on fewer points.  The node moves
// between HBM and the image ring by ring (NodeRings).
template <typename T, int P, int TILE, bool MIRROR, bool XM>
__global__ __launch_bounds__(kBlock) void k_heat_swept(const T* __restrict__ in,
                                                       T* __restrict__ out, int64_t n,
                                                       int64_t nnodes, int64_t node_first,
                                                       int64_t offset, T fo, T fo2)
{
    constexpr int W = kWave * P, H = W / 2;
    constexpr bool PHASED = P >= 2;
    constexpr bool RING_IN = PHASED && TILE != TILE_UP;
    constexpr bool RING_OUT = PHASED && TILE != TILE_DOWN;
    // level pairs per loop trip: a heat level is three FP64 operations per
    // point, so the loop's counter and branch are a visible share of it at up
    // to 8 points per lane (+8-13% at W = 128 .. 512), none at 16
    constexpr int U = P <= 8 ? 8 : 1;
    __shared__ T img_all[PHASED ? kWavesPerBlock * 2 * W : 1];
    NodeGeom g;
    if (!node_geom<P>(n, nnodes, node_first, offset, g)) return;

    const T* __restrict__ in1 = in + n;
    HeatRegion<T, P, MIRROR, XM> full;
    NodeRings<T, P, 2> rings(g, n);
    if constexpr (RING_IN) {
        rings.load_all(in, n);
    } else {
        load_pts<T, P>(in, g.lb, full.r0);
        load_pts<T, P>(in1, g.lb, full.r1);
    }
    full.hl0[0] = in[wrap(g.base - 1, n)];
    full.hr0[0] = in[wrap(g.base + W, n)];
    full.hl1[0] = in1[wrap(g.base - 1, n)];
    full.hr1[0] = in1[wrap(g.base + W, n)];
    full.place(W, g.lane * P, g.base, n);

    if constexpr (!PHASED) {
        if (TILE != TILE_UP) full.template down<U>(1, H - 1, fo, fo2);
        if (TILE != TILE_DOWN) {
            full.template up<U>(W, 1, H - 3, fo, fo2);
            full.stage(W - 1 - 2 * (H - 1), full.r0, full.hl0, full.hr0, full.r1, fo, fo2);
        }
    } else {
        // Regions of PC = 1, 2, .. P - 1 points per lane (widths 64 PC):
        // region PC takes down pairs 32(PC-1)+1 <= s <= 32 PC - 1 and up pairs
        // 32(P-PC)+1 <= m <= 32(P-PC+1)-1 (to H-3 for PC = 1), where every
        // threshold is <= 64 PC - 1.
        T* img = img_all + (threadIdx.x >> 6) * 2 * W;
        const int lane = g.lane;
        auto reload_full = [&] {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                full.r0[p] = img[lane * P + p];
                full.r1[p] = img[W + lane * P + p];
            }
        };
        auto down_region = [&](auto pc_tag) {
            constexpr int PC = decltype(pc_tag)::value;
            if constexpr (RING_IN) {
                if constexpr (PC == 1) rings.template put<1>(img);
                rings.template put<PC + 1>(img);  // region PC's halo
                wave_sync();
            }
            HeatRegion<T, PC, MIRROR, XM> r;
            r.template load<W>(img, lane, g.base, n);
            r.template down<U>(32 * (PC - 1) + 1, 32 * PC - 1, fo, fo2);
            r.template save<W>(img, lane);
            wave_sync();
        };
        auto up_region = [&](auto pc_tag) {
            constexpr int PC = decltype(pc_tag)::value;
            HeatRegion<T, PC, MIRROR, XM> r;
            r.template load<W>(img, lane, g.base, n);
            if constexpr (PC == 1) {
                r.template up<U>(W, 32 * (P - PC) + 1, H - 3, fo, fo2);
                r.stage(W - 1 - 2 * (H - 1), r.r0, r.hl0, r.hr0, r.r1, fo, fo2);
            } else {
                r.template up<U>(W, 32 * (P - PC) + 1, 32 * (P - PC + 1) - 1, fo, fo2);
            }
            r.template save<W>(img, lane);
            wave_sync();
            // later regions lie inside region PC - 1: ring PC is final
            rings.template store<PC>(img, out, n);
        };
        using Regions = std::make_integer_sequence<int, P - 1>;  // i -> PC = i + 1
        if (TILE != TILE_UP) {
            if constexpr (!RING_IN) {
                full.template save<W>(img, lane);
                wave_sync();
            }
            for_each_index([&](auto i) { down_region(std::integral_constant<int, decltype(i)::value + 1>{}); },
                           Regions{});
            reload_full();
            full.template down<U>(32 * (P - 1) + 1, H - 1, fo, fo2);
        }
        if (TILE != TILE_DOWN) {
            full.template up<U>(W, 1, 31, fo, fo2);
            full.template save<W>(img, lane);
            wave_sync();
            rings.template store<P>(img, out, n);
            for_each_index([&](auto i) { up_region(std::integral_constant<int, P - 1 - decltype(i)::value>{}); },
                           Regions{});
        }
    }
    if constexpr (!RING_OUT) {
        store_pts<T, P>(out, g.lb, full.r0);
        store_pts<T, P>(out + n, g.lb, full.r1);
    }
}

// ------------------------------------------------------------- KS tiles ---
// One RK2 stage pair on PC points per lane.  e[] holds the lane's points with
// two neighbours on each side (DPP from the adjacent lanes; the region's
// halos in lane 0 / lane 63 of the carried neighbour registers); d[p] is the point's distance code |2x - (W-1)| from
// the node centre, and a stage with threshold thr updates the points with
// d <= thr (the diamond's active set) and keeps the others.
template <typename T, int PC, bool XM>
struct KsRegion {
    T u[PC], uh[PC];
    T hlu[2], hru[2], hlh[2], hrh[2];
    int d[PC];

    // Predictor stage (KS:159-164): u_half = u - dt/2 F(u).
    __device__ __forceinline__ void stage_a(int thr, const KsCoef<T>& k)
    {
        T e[PC + 4];
        gather_carry<T, PC, 2>(u, hlu, hru, e);
#pragma unroll
        for (int p = 0; p < PC; ++p) {
            if constexpr (XM) {
                // ks_predict = fma(-dt/2, F, c), its fma in the active lanes
                fma_where(uh[p], -k.dth, ks_rhs(e[p], e[p + 1], e[p + 2], e[p + 3], e[p + 4], k), u[p], d[p],
                          thr);
            } else {
                const T nv = ks_predict(e[p], e[p + 1], e[p + 2], e[p + 3], e[p + 4], k);
                uh[p] = (d[p] <= thr) ? nv : uh[p];
            }
        }
    }
    // Corrector stage (KS:166-171, :202): u += -dt F(u_half).
    __device__ __forceinline__ void stage_b(int thr, const KsCoef<T>& k)
    {
        T e[PC + 4];
        gather_carry<T, PC, 2>(uh, hlh, hrh, e);
#pragma unroll
        for (int p = 0; p < PC; ++p) {
            if constexpr (XM) {
                // ks_correct = fma(-dt, F, u), its fma in the active lanes
                fma_where(u[p], -k.dt, ks_rhs(e[p], e[p + 1], e[p + 2], e[p + 3], e[p + 4], k), u[p], d[p],
                          thr);
            } else {
                const T nv = ks_correct(u[p], e[p], e[p + 1], e[p + 2], e[p + 3], e[p + 4], k);
                u[p] = (d[p] <= thr) ? nv : u[p];
            }
        }
    }
    // Down-triangle pairs s = s0, s0+2, .. s1 (thresholds 4s-1, 4s+3) and
    // up-triangle pairs m = m0, m0+2, .. m1 (thresholds W-1-4m, W-5-4m), U
    // pairs per loop trip.
    template <int U>
    __device__ __forceinline__ void down(int s0, int s1, const KsCoef<T>& k)
    {
#pragma unroll U
        for (int s = s0; s <= s1; s += 2) { stage_a(4 * s - 1, k); stage_b(4 * s + 3, k); }
    }
    template <int U>
    __device__ __forceinline__ void up(int W, int m0, int m1, const KsCoef<T>& k)
    {
#pragma unroll U
        for (int m = m0; m <= m1; m += 2) { stage_a(W - 1 - 4 * m, k); stage_b(W - 5 - 4 * m, k); }
    }

    // The masked stages with compile-time thresholds (fma_lane_sets): the
    // final fma of ks_predict / ks_correct, fma(-dt/2, F(u), u) into uh and
    // fma(-dt, F(uh), u) into u, on the same values.
    template <int THR>
    __device__ __forceinline__ void stage_a_c(const KsCoef<T>& k)
    {
        T e[PC + 4];
        gather_carry<T, PC, 2>(u, hlu, hru, e);
        fma_lane_sets<T, PC, THR>(uh, d, -k.dth,
                                  [&](int p) { return ks_rhs(e[p], e[p + 1], e[p + 2], e[p + 3], e[p + 4], k); },
                                  [&](int p) { return u[p]; });
    }
    template <int THR>
    __device__ __forceinline__ void stage_b_c(const KsCoef<T>& k)
    {
        T e[PC + 4];
        gather_carry<T, PC, 2>(uh, hlh, hrh, e);
        fma_lane_sets<T, PC, THR>(u, d, -k.dt,
                                  [&](int p) { return ks_rhs(e[p], e[p + 1], e[p + 2], e[p + 3], e[p + 4], k); },
                                  [&](int p) { return u[p]; });
    }
    template <int S0, int S1>
    __device__ __forceinline__ void down_c(const KsCoef<T>& k)
    {
        if constexpr (S0 <= S1) {
            stage_a_c<4 * S0 - 1>(k);
            stage_b_c<4 * S0 + 3>(k);
            down_c<S0 + 2, S1>(k);
        }
    }
    template <int W, int M0, int M1>
    __device__ __forceinline__ void up_c(const KsCoef<T>& k)
    {
        if constexpr (M0 <= M1) {
            stage_a_c<W - 1 - 4 * M0>(k);
            stage_b_c<W - 5 - 4 * M0>(k);
            up_c<W, M0 + 2, M1>(k);
        }
    }
    // Load the middle 64*PC points of the node image (u at img[0..W), u_half
    // at img[W..2W)) with their two-point halos, which stay constant while
    // every stage's active set lies inside the region.
    template <int W>
    __device__ __forceinline__ void load(const T* img, int lane)
    {
        constexpr int xs = W / 2 - 32 * PC;
        static_assert(xs >= 2 && xs + 64 * PC + 2 <= W, "region and halo inside the node");
        const int x0 = xs + lane * PC;
#pragma unroll
        for (int p = 0; p < PC; ++p) {
            u[p] = img[x0 + p];
            uh[p] = img[W + x0 + p];
            d[p] = abs(2 * (x0 + p) - (W - 1));
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            hlu[i] = img[xs - 1 - i];
            hru[i] = img[xs + 64 * PC + i];
            hlh[i] = img[W + xs - 1 - i];
            hrh[i] = img[W + xs + 64 * PC + i];
        }
    }
    template <int W>
    __device__ __forceinline__ void save(T* img, int lane) const
    {
        const int x0 = W / 2 - 32 * PC + lane * PC;
#pragma unroll
        for (int p = 0; p < PC; ++p) {
            img[x0 + p] = u[p];
            img[W + x0 + p] = uh[p];
        }
    }
};

// Region sizes of the phased KS tiles (points per lane): R_i = R_0 + i
// kKsRegionStep, i = 0 .. ks_regions<P>() - 1, all below P, from R_0 =
// kKsRegionFirst (P - 1 if that is not below P).
constexpr int kKsRegionFirst = 1, kKsRegionStep = 1;
template <int P>
__host__ __device__ constexpr int ks_region(int i)
{
    return (kKsRegionFirst < P - 1 ? kKsRegionFirst : P - 1) + i * kKsRegionStep;
}
template <int P>
__host__ __device__ constexpr int ks_regions()
{
    return (P - 1 - ks_region<P>(0)) / kKsRegionStep + 1;
}

// Swept KS tiles: one wave per node of W = 64*P points, lane l owning points
// l*P .. l*P+P-1 in VGPRs.  A whole diamond runs W/2 stage levels but its
// active set is W wide only in the middle; for P >= 4 the node's image is
// kept in LDS and the levels run on centred regions of 1, 2, .. P points per
// lane, each level on the narrowest region that holds its active set and the
// two-point stencil reach.  Same operations on the same values as the
// one-width form, so the same bits, with 47% fewer point evaluations at
// P = 16 (44% at P = 8, 37% at P = 4).  The node moves between HBM and the
// image ring by ring (NodeRings).
template <typename T, int P, int TILE, bool XM>
__global__ __launch_bounds__(kBlock) void k_ks_swept(const T* __restrict__ in,
                                                     T* __restrict__ out, int64_t n,
                                                     int64_t nnodes, int64_t node_first,
                                                     int64_t offset, KsCoef<T> k)
{
    constexpr int W = kWave * P, H = W / 4;
    constexpr bool PHASED = P >= 4;
    constexpr bool RING_IN = PHASED && TILE != TILE_UP;
    constexpr bool RING_OUT = PHASED && TILE != TILE_DOWN;
    // level pairs per loop trip: the loop's counter and branch are a visible
    // share of a pair at up to 8 points per lane (+6-20% at W = 64 .. 512),
    // none at 16
    constexpr int U = P <= 8 ? 4 : 1;
    // FP32 diamonds of the masked form at 8 points per lane run their stages
    // unrolled with compile-time thresholds (KsRegion::down_c / up_c): +3.6%
    // at 2^20, +8% at 2^21; the same form measured -5% in FP64
    constexpr bool CT = XM && PHASED && P == 8 && TILE == TILE_DIAMOND && std::is_same_v<T, float>;
    __shared__ T img_all[PHASED ? kWavesPerBlock * 2 * W : 1];
    NodeGeom g;
    if (!node_geom<P>(n, nnodes, node_first, offset, g)) return;

    const T* __restrict__ in_h = in + n;
    KsRegion<T, P, XM> full;
    NodeRings<T, P, 2> rings(g, n);
    if constexpr (RING_IN) {
        rings.load_all(in, n);
    } else {
        load_pts<T, P>(in, g.lb, full.u);
        load_pts<T, P>(in_h, g.lb, full.uh);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int64_t il = wrap(g.base - 1 - i, n), ir = wrap(g.base + W + i, n);
        full.hlu[i] = in[il]; full.hru[i] = in[ir];
        full.hlh[i] = in_h[il]; full.hrh[i] = in_h[ir];
    }
#pragma unroll
    for (int p = 0; p < P; ++p) full.d[p] = abs(2 * (g.lane * P + p) - (W - 1));

    if constexpr (!PHASED) {
        if (TILE != TILE_UP) full.template down<U>(1, H - 1, k);
        if (TILE != TILE_DOWN) {
            full.template up<U>(W, 1, H - 3, k);
            full.stage_a(W - 1 - 4 * (H - 1), k);
        }
    } else {
        // Regions R_1 < R_2 < .. < R_K < P points per lane (widths 64 R_i,
        // ks_region): region R_i takes down pairs 16 R_(i-1) + 1 <= s <=
        // 16 R_i - 1 (from 1 for i = 1) and up pairs 16(P - R_i) + 1 <= m <=
        // 16(P - R_(i-1)) - 1 (to H - 3 for i = 1), where every threshold is
        // <= 64 R_i - 1; the full node takes the pairs above 16 R_K.
        constexpr int K = ks_regions<P>();
        static_assert(K >= 1, "phased KS tiles have at least one region");
        constexpr int RK = ks_region<P>(K - 1);
        T* img = img_all + (threadIdx.x >> 6) * 2 * W;
        const int lane = g.lane;
        auto reload_full = [&] {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                full.u[p] = img[lane * P + p];
                full.uh[p] = img[W + lane * P + p];
            }
        };
        auto down_region = [&](auto i_tag) {
            constexpr int I = decltype(i_tag)::value;
            constexpr int PC = ks_region<P>(I), PREV = I == 0 ? 0 : ks_region<P>(I - 1);
            if constexpr (RING_IN) {
                // rings 1 .. PC + 1 (PC + 1: the region's halo) are in the image
                for_range<PREV + 2 - (I == 0 ? 1 : 0), PC + 1>(
                    [&](auto j) { rings.template put<decltype(j)::value>(img); });
                wave_sync();
            }
            KsRegion<T, PC, XM> r;
            r.template load<W>(img, lane);
            if constexpr (CT)
                r.template down_c<I == 0 ? 1 : 16 * PREV + 1, 16 * PC - 1>(k);
            else
                r.template down<U>(I == 0 ? 1 : 16 * PREV + 1, 16 * PC - 1, k);
            r.template save<W>(img, lane);
            wave_sync();
        };
        auto up_region = [&](auto j_tag) {
            constexpr int I = K - 1 - decltype(j_tag)::value;
            constexpr int PC = ks_region<P>(I), PREV = I == 0 ? 0 : ks_region<P>(I - 1);
            KsRegion<T, PC, XM> r;
            r.template load<W>(img, lane);
            if constexpr (I == 0 && CT) {
                r.template up_c<W, 16 * (P - PC) + 1, H - 3>(k);
                r.template stage_a_c<W - 1 - 4 * (H - 1)>(k);
            } else if constexpr (I == 0) {
                r.template up<U>(W, 16 * (P - PC) + 1, H - 3, k);
                r.stage_a(W - 1 - 4 * (H - 1), k);
            } else if constexpr (CT) {
                r.template up_c<W, 16 * (P - PC) + 1, 16 * (P - PREV) - 1>(k);
            } else {
                r.template up<U>(W, 16 * (P - PC) + 1, 16 * (P - PREV) - 1, k);
            }
            r.template save<W>(img, lane);
            wave_sync();
            // later regions lie inside region PREV: rings PREV + 1 .. PC are final
            for_range<PREV + 1, PC>([&](auto j) { rings.template store<decltype(j)::value>(img, out, n); });
        };
        using Regions = std::make_integer_sequence<int, K>;
        if (TILE != TILE_UP) {
            if constexpr (!RING_IN) {
                full.template save<W>(img, lane);
                wave_sync();
            }
            for_each_index(down_region, Regions{});
            if constexpr (RING_IN && RK + 2 <= P) {
                for_range<RK + 2, P>([&](auto j) { rings.template put<decltype(j)::value>(img); });
                wave_sync();
            }
            reload_full();
            if constexpr (CT)
                full.template down_c<16 * RK + 1, H - 1>(k);
            else
                full.template down<U>(16 * RK + 1, H - 1, k);
        }
        if (TILE != TILE_DOWN) {
            if constexpr (CT)
                full.template up_c<W, 1, 16 * (P - RK) - 1>(k);
            else
                full.template up<U>(W, 1, 16 * (P - RK) - 1, k);
            full.template save<W>(img, lane);
            wave_sync();
            for_range<RK + 1, P>([&](auto j) { rings.template store<decltype(j)::value>(img, out, n); });
            for_each_index(up_region, Regions{});
        }
    }
    if constexpr (!RING_OUT) {
        store_pts<T, P>(out, g.lb, full.u);
        store_pts<T, P>(out + n, g.lb, full.uh);
    }
}

// ---------------------------------------------------------- Euler tiles ---
// Components 0..2 are the conserved values (rho, rho*u, E), 3..5 their
// predictor.  EulerRegion holds PC cells per lane of the middle 64*PC cells
// of a node (all of it when PC = P).
template <typename T, int PC>
struct EulerRegion {
    T q[6][PC];
    T hl[6][2], hr[6][2];
    int d[PC];
    bool g0, gN;  // this lane's first / last cell is global cell 0 / n-1
    // one cell per lane: the lane's cell is global cell 1 / n-2 (with two or
    // more per lane, cells 0 and 1, and n-2 and n-1, share a lane)
    bool g1 = false, gNm1 = false;

    // src = 0: predictor from the conserved values (Euler:307-350);
    // src = 3: corrector from the predictor values (Euler:353-389, :420).
    // Each interface flux is evaluated once and used by the two cells beside
    // it (PC+1 interfaces per lane instead of 2 PC).
    template <int src>
    __device__ __forceinline__ void stage(int thr, const EulerCoef<T>& k)
    {
        constexpr int dst = 3 - src;
        T ex[PC + 4], ey[PC + 4], ez[PC + 4];
        gather_carry<T, PC, 2>(q[src + 0], hl[src + 0], hr[src + 0], ex);
        gather_carry<T, PC, 2>(q[src + 1], hl[src + 1], hr[src + 1], ey);
        gather_carry<T, PC, 2>(q[src + 2], hl[src + 2], hr[src + 2], ez);
        auto cell = [&](int j) { return Q3<T>{ex[j], ey[j], ez[j]}; };
        // pressure differences between ext cells j and j+1
        T pd[PC + 3];
#pragma unroll
        for (int j = 0; j < PC + 3; ++j) pd[j] = eu_pdiff(cell(j), cell(j + 1));
        // interface j lies between ext cells j+1 and j+2; the interface left
        // of global cell 1 drops pdiff(-1,0), the one right of cell n-2 drops
        // pdiff(n-1,n) (the flags of Euler:405, :424).
        auto iface = [&](int j, Q3<T>& F, Q3<T>& D) {
            bool drop0 = j == 1 && g0, drop2 = j == PC - 1 && gN;
            if constexpr (PC == 1) {
                drop0 = drop0 || (j == 0 && g1);
                drop2 = drop2 || (j == 1 && gNm1);
            }
            const T q0 = drop0 ? T(0) : pd[j];
            const T q2 = drop2 ? T(0) : pd[j + 2];
            eu_iface(cell(j), cell(j + 1), cell(j + 2), cell(j + 3), q0, pd[j + 1], q2, k, F, D);
        };
        const T c = src == 0 ? T(0.25) * k.dtdx : T(0.5) * k.dtdx;
        Q3<T> FL, DL;
        iface(0, FL, DL);
#pragma unroll
        for (int p = 0; p < PC; ++p) {
            Q3<T> FR, DR;
            iface(p + 1, FR, DR);
            const Q3<T> f = eu_balance(FL, DL, FR, DR);
            const Q3<T> base = src == 0 ? cell(p + 2) : Q3<T>{q[0][p], q[1][p], q[2][p]};
            Q3<T> nv = eu_apply(base, f, c);
            // ghost cells 0 and n-1 hold the Dirichlet states (Euler:405-425)
            if (p == 0 && g0) nv = k.bl;
            if (p == PC - 1 && gN) nv = k.br;
            const bool act = d[p] <= thr;
            q[dst + 0][p] = act ? nv.x : q[dst + 0][p];
            q[dst + 1][p] = act ? nv.y : q[dst + 1][p];
            q[dst + 2][p] = act ? nv.z : q[dst + 2][p];
            FL = FR;
            DL = DR;
        }
    }
    __device__ __forceinline__ void down(int s0, int s1, const EulerCoef<T>& k)
    {
        for (int s = s0; s <= s1; s += 2) { stage<0>(4 * s - 1, k); stage<3>(4 * s + 3, k); }
    }
    __device__ __forceinline__ void up(int W, int m0, int m1, const EulerCoef<T>& k)
    {
        for (int m = m0; m <= m1; m += 2) { stage<0>(W - 1 - 4 * m, k); stage<3>(W - 5 - 4 * m, k); }
    }
    __device__ __forceinline__ void place(int W, int x0, int64_t base, int64_t n)
    {
#pragma unroll
        for (int p = 0; p < PC; ++p) d[p] = abs(2 * (x0 + p) - (W - 1));
        int64_t gi = base + x0;
        gi = gi >= n ? gi - n : gi;
        int64_t gl = gi + PC - 1;
        gl = gl >= n ? gl - n : gl;
        g0 = gi == 0;
        gN = gl == n - 1;
        if (PC == 1) {
            g1 = gi == 1;
            gNm1 = gl == n - 2;
        }
    }
    // The region is the middle 64*PC cells of a node of W cells (all of it
    // when PC = W/64): lane l holds node-local x0 = W/2 - 32 PC + l PC ..
    // x0 + PC - 1, and the halo is the two cells beyond each end.
    template <int W>
    static constexpr int xs() { return W / 2 - 32 * PC; }
    // every component and the halo from the state arrays (node base `base`)
    template <int W>
    __device__ __forceinline__ void load_global(const T* in, int64_t n, int64_t base, int lane)
    {
        const int x0 = xs<W>() + lane * PC;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
#pragma unroll
            for (int p = 0; p < PC; ++p) q[c][p] = in[c * n + wrap(base + x0 + p, n)];
        }
        halo_global<W>(in, n, base);
    }
    template <int W>
    __device__ __forceinline__ void halo_global(const T* in, int64_t n, int64_t base)
    {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                hl[c][i] = in[c * n + wrap(base + xs<W>() - 1 - i, n)];
                hr[c][i] = in[c * n + wrap(base + xs<W>() + 64 * PC + i, n)];
            }
        }
    }
    // Components C0 .. C0+2 to / from a three-component node image (component
    // c at img + (c - C0) W, node-local cell order).
    template <int W, int C0>
    __device__ __forceinline__ void save_half(T* img, int lane) const
    {
        const int x0 = xs<W>() + lane * PC;
#pragma unroll
        for (int c = C0; c < C0 + 3; ++c)
#pragma unroll
            for (int p = 0; p < PC; ++p) img[(c - C0) * W + x0 + p] = q[c][p];
    }
    template <int W, int C0>
    __device__ __forceinline__ void load_half(const T* img, int lane)
    {
        const int x0 = xs<W>() + lane * PC;
#pragma unroll
        for (int c = C0; c < C0 + 3; ++c)
#pragma unroll
            for (int p = 0; p < PC; ++p) q[c][p] = img[(c - C0) * W + x0 + p];
    }
    template <int W, int C0>
    __device__ __forceinline__ void halo_half(const T* img)
    {
        static_assert(xs<W>() >= 2 && xs<W>() + 64 * PC + 2 <= W, "halo inside the node");
#pragma unroll
        for (int c = C0; c < C0 + 3; ++c) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                hl[c][i] = img[(c - C0) * W + xs<W>() - 1 - i];
                hr[c][i] = img[(c - C0) * W + xs<W>() + 64 * PC + i];
            }
        }
    }
    // the region's cells to the state arrays
    template <int W>
    __device__ __forceinline__ void store_global(T* out, int64_t n, int64_t base, int lane) const
    {
        const int x0 = xs<W>() + lane * PC;
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int p = 0; p < PC; ++p) out[c * n + wrap(base + x0 + p, n)] = q[c][p];
    }
};

// Ring PC's cells of components C0 .. C0+2: from the state arrays into a
// three-component node image, and from the image to the state arrays.
template <typename T, int W, int PC, int C0>
__device__ __forceinline__ void eu_ring_in(T* img, const T* in, int64_t n, int64_t base, int lane)
{
    const int x = ring_x<W, PC>(lane);
#pragma unroll
    for (int c = C0; c < C0 + 3; ++c) img[(c - C0) * W + x] = in[c * n + wrap(base + x, n)];
}
template <typename T, int W, int PC, int C0>
__device__ __forceinline__ void eu_ring_out(const T* img, T* out, int64_t n, int64_t base, int lane)
{
    const int x = ring_x<W, PC>(lane);
#pragma unroll
    for (int c = C0; c < C0 + 3; ++c) out[c * n + wrap(base + x, n)] = img[(c - C0) * W + x];
}

// Smallest region of the phased Euler tiles, in cells per lane.
constexpr int kEuRegionFirst = 1;

// A phased Euler tile's node: its base in the state arrays and the wave's
// three-component LDS image (3 W values).
template <typename T>
struct EuNode {
    T* img;
    const T* in;
    T* out;
    int64_t n, base;
    int lane;
};

// Region A -> region B = A + 1 cells per lane (B = W/64: the whole node), one
// half of the state at a time through the image.  The cells of ring B, and
// the halo of region B, lie outside every region so far, so they still hold
// the tile's input.
template <typename T, int W, int A, int B>
__device__ __forceinline__ void eu_grow(const EulerRegion<T, A>& a, EulerRegion<T, B>& b, const EuNode<T>& d)
{
    for_each_index(
        [&](auto h) {
            constexpr int C0 = 3 * decltype(h)::value;
            a.template save_half<W, C0>(d.img, d.lane);
            eu_ring_in<T, W, B, C0>(d.img, d.in, d.n, d.base, d.lane);
            wave_sync();
            b.template load_half<W, C0>(d.img, d.lane);
            wave_sync();
        },
        std::make_integer_sequence<int, 2>{});
    b.template halo_global<W>(d.in, d.n, d.base);
    b.place(W, EulerRegion<T, B>::template xs<W>() + d.lane * B, d.base, d.n);
}

// Region A -> region B = A - 1 cells per lane: ring A is final and goes to
// the output; region B's halo lies in ring A and comes from the image.
template <typename T, int W, int A, int B>
__device__ __forceinline__ void eu_shrink(const EulerRegion<T, A>& a, EulerRegion<T, B>& b, const EuNode<T>& d)
{
    for_each_index(
        [&](auto h) {
            constexpr int C0 = 3 * decltype(h)::value;
            a.template save_half<W, C0>(d.img, d.lane);
            wave_sync();
            eu_ring_out<T, W, A, C0>(d.img, d.out, d.n, d.base, d.lane);
            b.template load_half<W, C0>(d.img, d.lane);
            b.template halo_half<W, C0>(d.img);
            wave_sync();
        },
        std::make_integer_sequence<int, 2>{});
    b.place(W, EulerRegion<T, B>::template xs<W>() + d.lane * B, d.base, d.n);
}

// The down phase from region PC (loaded) on: region PC takes down pairs
// 16 PC - 15 <= s <= 16 PC - 1 (from 1 for the first region), where every
// threshold is <= 64 PC - 1; the whole node ends it in `full`.
template <typename T, int P, int PC>
__device__ __forceinline__ void eu_down(EulerRegion<T, PC>& r, EulerRegion<T, P>& full, const EuNode<T>& d,
                                        const EulerCoef<T>& k)
{
    constexpr int W = kWave * P, H = W / 4;
    r.down(PC == kEuRegionFirst ? 1 : 16 * PC - 15, 16 * PC - 1, k);
    if constexpr (PC + 1 < P) {
        EulerRegion<T, PC + 1> r2;
        eu_grow<T, W, PC, PC + 1>(r, r2, d);
        eu_down<T, P, PC + 1>(r2, full, d, k);
    } else {
        eu_grow<T, W, PC, P>(r, full, d);
        full.down(16 * (P - 1) + 1, H - 1, k);
    }
}

// The up phase from region PC + 1 (or the whole node, PC + 1 = P) down:
// region PC takes up pairs 16(P-PC)+1 <= m <= 16(P-PC)+15 (to H-3 for the
// first region); each ring goes to the output once final, the first region
// last.
template <typename T, int P, int PC>
__device__ __forceinline__ void eu_up(const EulerRegion<T, PC + 1>& a, const EuNode<T>& d, const EulerCoef<T>& k)
{
    constexpr int W = kWave * P, H = W / 4;
    EulerRegion<T, PC> b;
    eu_shrink<T, W, PC + 1, PC>(a, b, d);
    if constexpr (PC == kEuRegionFirst) {
        b.up(W, 16 * (P - PC) + 1, H - 3, k);
        b.template stage<0>(W - 1 - 4 * (H - 1), k);
        b.template store_global<W>(d.out, d.n, d.base, d.lane);
    } else {
        b.up(W, 16 * (P - PC) + 1, 16 * (P - PC) + 15, k);
        eu_up<T, P, PC - 1>(b, d, k);
    }
}

// Swept Euler tiles: one wave per node of W = 64*P cells.  For P >= 4 the
// diamond's levels run on the narrowest centred region of 1, 2, .. P cells
// per lane that holds their active set and the two-cell stencil reach (same
// operations, same bits), and a change of region moves the lane's cells
// through an LDS image of three of the six components at a time: 12 KB per
// wave at W = 512, so that two waves fit on a SIMD (the kernel is held to
// 256 registers for the same reason).  The node's cells come from the input
// ring by ring as the regions grow and go to the output ring by ring as they
// shrink.
template <typename T, int P, int TILE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
void k_euler_swept(const T* __restrict__ in, T* __restrict__ out, int64_t n, int64_t nnodes,
                   int64_t node_first, int64_t offset, EulerCoef<T> k)
{
    constexpr int W = kWave * P, H = W / 4;
    constexpr bool PHASED = P >= 4;
    __shared__ T img_all[PHASED ? kWavesPerBlock * 3 * W : 1];
    NodeGeom g;
    if (!node_geom<P>(n, nnodes, node_first, offset, g)) return;

    EulerRegion<T, P> full;
    if constexpr (!PHASED) {
        full.template load_global<W>(in, n, g.base, g.lane);
        full.place(W, g.lane * P, g.base, n);
        if (TILE != TILE_UP) full.down(1, H - 1, k);
        if (TILE != TILE_DOWN) {
            full.up(W, 1, H - 3, k);
            full.template stage<0>(W - 1 - 4 * (H - 1), k);
        }
        full.template store_global<W>(out, n, g.base, g.lane);
    } else {
        constexpr int F = kEuRegionFirst;
        const EuNode<T> d{img_all + (threadIdx.x >> 6) * 3 * W, in, out, n, g.base, g.lane};
        if (TILE != TILE_UP) {
            EulerRegion<T, F> r;
            r.template load_global<W>(in, n, g.base, g.lane);
            r.place(W, EulerRegion<T, F>::template xs<W>() + g.lane * F, g.base, n);
            eu_down<T, P, F>(r, full, d, k);
        } else {
            full.template load_global<W>(in, n, g.base, g.lane);
            full.place(W, g.lane * P, g.base, n);
        }
        if (TILE != TILE_DOWN) {
            full.up(W, 1, 15, k);
            eu_up<T, P, P - 1>(full, d, k);
        } else {
            full.template store_global<W>(out, n, g.base, g.lane);
        }
    }
}

// ------------------------------------------------------ classic kernels ---
template <typename T, bool PERIODIC>
__global__ __launch_bounds__(kClassicBlock) void k_heat_classic(const T* __restrict__ in,
                                                                T* __restrict__ out, int64_t n,
                                                                T fo, T fo2)
{
    const int64_t i = (int64_t)blockIdx.x * kClassicBlock + threadIdx.x;
    if (i >= n) return;
    int64_t l, r;
    if (PERIODIC) {
        l = i == 0 ? n - 1 : i - 1;
        r = i == n - 1 ? 0 : i + 1;
    } else {
        l = i == 0 ? 1 : i - 1;
        r = i == n - 1 ? n - 2 : i + 1;
    }
    out[i] = heat_ftcs(in[l], in[i], in[r], fo, fo2);
}

// One KS stage, one point per thread (any n).
template <typename T, int STAGE>
__global__ __launch_bounds__(kClassicBlock) void k_ks_classic1(T* __restrict__ buf, int64_t n,
                                                               KsCoef<T> k)
{
    const int64_t i = (int64_t)blockIdx.x * kClassicBlock + threadIdx.x;
    if (i >= n) return;
    const T* __restrict__ src = STAGE == 0 ? buf : buf + n;
    const T ll = src[wrap(i - 2, n)], l = src[wrap(i - 1, n)], c = src[i];
    const T r = src[wrap(i + 1, n)], rr = src[wrap(i + 2, n)];
    if (STAGE == 0)
        buf[n + i] = ks_predict(ll, l, c, r, rr, k);
    else
        buf[i] = ks_correct(buf[i], ll, l, c, r, rr, k);
}

// One KS stage, two points per thread (n even, so both halves of buf are
// 16-byte aligned): the six inputs i-2 .. i+3 of points i, i+1 arrive as three
// aligned pairs, and the two results leave as one pair.  The threads at the
// ends of the array wrap.
template <typename T, int STAGE>
__global__ __launch_bounds__(kClassicBlock) void k_ks_classic(T* __restrict__ buf, int64_t n,
                                                              KsCoef<T> k)
{
    typedef T pair_t __attribute__((ext_vector_type(2)));
    const int64_t i = 2 * ((int64_t)blockIdx.x * kClassicBlock + threadIdx.x);
    if (i >= n) return;
    const T* __restrict__ src = STAGE == 0 ? buf : buf + n;
    T e[6];
    if (i >= 2 && i + 4 <= n) {
        const pair_t* p = reinterpret_cast<const pair_t*>(src + i - 2);
        const pair_t x0 = p[0], x1 = p[1], x2 = p[2];
        e[0] = x0.x; e[1] = x0.y; e[2] = x1.x; e[3] = x1.y; e[4] = x2.x; e[5] = x2.y;
    } else {
#pragma unroll
        for (int j = 0; j < 6; ++j) e[j] = src[wrap(i - 2 + j, n)];
    }
    pair_t o;
    if (STAGE == 0) {
        o.x = ks_predict(e[0], e[1], e[2], e[3], e[4], k);
        o.y = ks_predict(e[1], e[2], e[3], e[4], e[5], k);
        *reinterpret_cast<pair_t*>(buf + n + i) = o;
    } else {
        const pair_t u = *reinterpret_cast<const pair_t*>(buf + i);
        o.x = ks_correct(u.x, e[0], e[1], e[2], e[3], e[4], k);
        o.y = ks_correct(u.y, e[1], e[2], e[3], e[4], e[5], k);
        *reinterpret_cast<pair_t*>(buf + i) = o;
    }
}

// Euler classic stage (Euler:391-426) with each interface flux evaluated once:
// lane l of a wave computes the interface between cells x = 63 w + l - 1 and
// x + 1, and (l < 63) updates cell x + 1 from it and from lane l + 1's,
// which a DPP shift brings: 63 cells per wave, and the values the per-cell
// form computes twice, so the same bits.  Cell indices are clamped at the
// domain ends as in the per-cell form (the values so read are unused there).
template <typename T, int STAGE>
__global__ __launch_bounds__(kClassicBlock) void k_euler_classic(T* __restrict__ buf, int64_t n,
                                                                 EulerCoef<T> k)
{
    constexpr int src = STAGE == 0 ? 0 : 3, dst = 3 - src;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = (int64_t)blockIdx.x * (kClassicBlock / kWave) + (threadIdx.x >> 6);
    const int64_t x = wave * (kWave - 1) + lane - 1;
    auto cell = [&](int64_t j) {
        j = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
        return Q3<T>{buf[(src + 0) * n + j], buf[(src + 1) * n + j], buf[(src + 2) * n + j]};
    };
    const Q3<T> A = cell(x - 1), B = cell(x), C = cell(x + 1), Dc = cell(x + 2);
    // the interface right of cell 0 drops pdiff(-1,0), the one left of cell
    // n-1 drops pdiff(n-1,n) (the flags of Euler:405, :424)
    const T q0 = x == 0 ? T(0) : eu_pdiff(A, B);
    const T q1 = eu_pdiff(B, C);
    const T q2 = x == n - 2 ? T(0) : eu_pdiff(C, Dc);
    Q3<T> FL, DL;
    eu_iface(A, B, C, Dc, q0, q1, q2, k, FL, DL);
    // every lane takes part in the shifts before any leaves
    const Q3<T> FR{from_right(FL.x, FL.x), from_right(FL.y, FL.y), from_right(FL.z, FL.z)};
    const Q3<T> DR{from_right(DL.x, DL.x), from_right(DL.y, DL.y), from_right(DL.z, DL.z)};
    const int64_t i = x + 1;
    if (lane == kWave - 1 || i >= n) return;
    Q3<T> nv;
    if (i == 0) {
        nv = k.bl;
    } else if (i == n - 1) {
        nv = k.br;
    } else {
        const Q3<T> f = eu_balance(FL, DL, FR, DR);
        if (STAGE == 0)
            nv = eu_apply(C, f, T(0.25) * k.dtdx);
        else
            nv = eu_apply(Q3<T>{buf[i], buf[n + i], buf[2 * n + i]}, f, T(0.5) * k.dtdx);
    }
    buf[(dst + 0) * n + i] = nv.x;
    buf[(dst + 1) * n + i] = nv.y;
    buf[(dst + 2) * n + i] = nv.z;
}

// ------------------------------------------------------------ lane check ---
__global__ void k_lane_check(int* ok)
{
    const int lane = threadIdx.x;
    const double v = 1000.0 + lane;
    const double l = from_left(v, -1.0);
    const double r = from_right(v, -2.0);
    const float lf = from_left((float)v, -1.0f);
    const float rf = from_right((float)v, -2.0f);
    const double el = lane == 0 ? -1.0 : v - 1.0;
    const double er = lane == 63 ? -2.0 : v + 1.0;
    const bool good = (l == el) && (r == er) && (lf == (float)el) && (rf == (float)er);
    if (!good) atomicAnd(ok, 0);
}

// ------------------------------------------------------------- dispatch ---
template <typename T, int P>
hipError_t swept_eq(const SweptLaunch& L, const Coefs& c, hipStream_t s);

template <typename T> T fo_of(const Coefs& c);
template <> double fo_of<double>(const Coefs& c) { return c.fo_d; }
template <> float fo_of<float>(const Coefs& c) { return c.fo_f; }
template <typename T> T fo2_of(const Coefs& c);
template <> double fo2_of<double>(const Coefs& c) { return c.fo2_d; }
template <> float fo2_of<float>(const Coefs& c) { return c.fo2_f; }
template <typename T> const KsCoef<T>& ks_of(const Coefs& c);
template <> const KsCoef<double>& ks_of<double>(const Coefs& c) { return c.ks_d; }
template <> const KsCoef<float>& ks_of<float>(const Coefs& c) { return c.ks_f; }
template <typename T> const EulerCoef<T>& eu_of(const Coefs& c);
template <> const EulerCoef<double>& eu_of<double>(const Coefs& c) { return c.eu_d; }
template <> const EulerCoef<float>& eu_of<float>(const Coefs& c) { return c.eu_f; }

template <typename T, int P, int TILE>
hipError_t swept_tile(const SweptLaunch& L, const Coefs& c, hipStream_t s)
{
    const int64_t count = L.nnodes - L.node_first;
    if (count <= 0) return hipSuccess;
    const dim3 grid((unsigned)((count + kWavesPerBlock - 1) / kWavesPerBlock));
    const T* in = static_cast<const T*>(L.in);
    T* out = static_cast<T*>(L.out);
    // About two or more nodes (waves) per SIMD: the stages' final fma runs
    // under an EXEC mask (v_cmpx), and another wave issues while one runs its
    // scalar EXEC restore.  About one wave per SIMD: the compare-and-select
    // form, which needs no scalar instruction between the FP64 work.  The
    // threshold, 1.5 nodes per SIMD, puts the hybrid's nodes 1..K-1 with the
    // full launch.  Same bits.
    const bool xm = L.simds > 0 && 2 * count >= 3 * (int64_t)L.simds;
    if (L.eq == EQ_HEAT) {
        if (L.bc == BC_PERIODIC) {
            if (xm)
                hipLaunchKernelGGL((k_heat_swept<T, P, TILE, false, true>), grid, dim3(kBlock), 0, s, in, out,
                                   L.n, L.nnodes, L.node_first, L.offset, fo_of<T>(c), fo2_of<T>(c));
            else
                hipLaunchKernelGGL((k_heat_swept<T, P, TILE, false, false>), grid, dim3(kBlock), 0, s, in, out,
                                   L.n, L.nnodes, L.node_first, L.offset, fo_of<T>(c), fo2_of<T>(c));
        } else {
            if (xm)
                hipLaunchKernelGGL((k_heat_swept<T, P, TILE, true, true>), grid, dim3(kBlock), 0, s, in, out,
                                   L.n, L.nnodes, L.node_first, L.offset, fo_of<T>(c), fo2_of<T>(c));
            else
                hipLaunchKernelGGL((k_heat_swept<T, P, TILE, true, false>), grid, dim3(kBlock), 0, s, in, out,
                                   L.n, L.nnodes, L.node_first, L.offset, fo_of<T>(c), fo2_of<T>(c));
        }
    } else if (L.eq == EQ_KS) {
        if (xm)
            hipLaunchKernelGGL((k_ks_swept<T, P, TILE, true>), grid, dim3(kBlock), 0, s, in, out, L.n,
                               L.nnodes, L.node_first, L.offset, ks_of<T>(c));
        else
            hipLaunchKernelGGL((k_ks_swept<T, P, TILE, false>), grid, dim3(kBlock), 0, s, in, out, L.n,
                               L.nnodes, L.node_first, L.offset, ks_of<T>(c));
    } else {
        if constexpr (P <= 8)
            hipLaunchKernelGGL((k_euler_swept<T, P, TILE>), grid, dim3(kBlock), 0, s, in, out,
                               L.n, L.nnodes, L.node_first, L.offset, eu_of<T>(c));
        else
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, int P>
hipError_t swept_eq(const SweptLaunch& L, const Coefs& c, hipStream_t s)
{
    switch (L.tile) {
    case TILE_UP: return swept_tile<T, P, TILE_UP>(L, c, s);
    case TILE_DOWN: return swept_tile<T, P, TILE_DOWN>(L, c, s);
    case TILE_DIAMOND: return swept_tile<T, P, TILE_DIAMOND>(L, c, s);
    }
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t swept_prec(const SweptLaunch& L, const Coefs& c, hipStream_t s)
{
    switch (L.P) {
    case 1: return swept_eq<T, 1>(L, c, s);
    case 2: return swept_eq<T, 2>(L, c, s);
    case 4: return swept_eq<T, 4>(L, c, s);
    case 8: return swept_eq<T, 8>(L, c, s);
    case 16: return swept_eq<T, 16>(L, c, s);
    }
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t classic_prec(int eq, int bc, int stage, int64_t n, const void* in, void* out,
                        const Coefs& c, hipStream_t s)
{
    const dim3 grid((unsigned)((n + kClassicBlock - 1) / kClassicBlock));
    if (eq == EQ_HEAT) {
        const T* a = static_cast<const T*>(in);
        T* b = static_cast<T*>(out);
        if (bc == BC_PERIODIC)
            hipLaunchKernelGGL((k_heat_classic<T, true>), grid, dim3(kClassicBlock), 0, s, a, b, n,
                               fo_of<T>(c), fo2_of<T>(c));
        else
            hipLaunchKernelGGL((k_heat_classic<T, false>), grid, dim3(kClassicBlock), 0, s, a, b, n,
                               fo_of<T>(c), fo2_of<T>(c));
    } else if (eq == EQ_KS) {
        T* b = static_cast<T*>(out);
        const dim3 g2((unsigned)((n / 2 + kClassicBlock - 1) / kClassicBlock));
        if (n % 2 != 0) {
            if (stage == 0)
                hipLaunchKernelGGL((k_ks_classic1<T, 0>), grid, dim3(kClassicBlock), 0, s, b, n, ks_of<T>(c));
            else
                hipLaunchKernelGGL((k_ks_classic1<T, 1>), grid, dim3(kClassicBlock), 0, s, b, n, ks_of<T>(c));
        } else if (stage == 0) {
            hipLaunchKernelGGL((k_ks_classic<T, 0>), g2, dim3(kClassicBlock), 0, s, b, n, ks_of<T>(c));
        } else {
            hipLaunchKernelGGL((k_ks_classic<T, 1>), g2, dim3(kClassicBlock), 0, s, b, n, ks_of<T>(c));
        }
    } else {
        T* b = static_cast<T*>(out);
        const int64_t waves = (n + kWave - 2) / (kWave - 1);  // 63 cells per wave
        const dim3 ge((unsigned)((waves + kClassicBlock / kWave - 1) / (kClassicBlock / kWave)));
        if (stage == 0)
            hipLaunchKernelGGL((k_euler_classic<T, 0>), ge, dim3(kClassicBlock), 0, s, b, n, eu_of<T>(c));
        else
            hipLaunchKernelGGL((k_euler_classic<T, 1>), ge, dim3(kClassicBlock), 0, s, b, n, eu_of<T>(c));
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_swept(const SweptLaunch& L, bool f64, const Coefs& c, hipStream_t s)
{
    return f64 ? swept_prec<double>(L, c, s) : swept_prec<float>(L, c, s);
}

hipError_t launch_classic(int eq, int bc, bool f64, int stage, int64_t n, const void* in,
                          void* out, const Coefs& c, hipStream_t s)
{
    return f64 ? classic_prec<double>(eq, bc, stage, n, in, out, c, s)
               : classic_prec<float>(eq, bc, stage, n, in, out, c, s);
}

int check_lane_shifts_at(int* d)
{
    const int one = 1;
    int ok = 0;
    hipError_t e = hipMemcpy(d, &one, sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_lane_check, dim3(1), dim3(64), 0, 0, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&ok, d, sizeof(int), hipMemcpyDeviceToHost);
    return e == hipSuccess ? ok : -1;
}

int check_lane_shifts()
{
    int* d = nullptr;
    if (hipMalloc(&d, sizeof(int)) != hipSuccess) return -1;
    const int r = check_lane_shifts_at(d);
    (void)hipFree(d);
    return r;
}

}  // namespace sw1d
