// Host side of libswept1d: plans, the C ABI, the hybrid CPU node and the
// schedule entry point.  See include/swept1d.h for the contract.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/swept1d.h"
#include "sw1d_internal.hpp"
#include "sw1d_schedule.hpp"

using namespace sw1d;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_OR_FAIL(expr)                                                               \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(SW1D_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));  \
    } while (0)

enum KernelKind { KK_CLASSIC = 0, KK_UP = 1, KK_DOWN = 2, KK_DIAMOND = 3, KK_SPLIT = 4, KK_COUNT = 5 };

}  // namespace

struct sw1d_plan {
    sw1d_problem prob{};
    int eq = 0, bc = 0, alg = 0;
    bool f64 = true;
    int64_t n = 0;
    size_t S = 8;      // sizeof(REAL)
    int nc = 1;        // values per point
    int W = 0, P = 0;  // swept node width, points per lane (0: no swept path)
    int SB = 0;        // trapezoid: extra points per lane (stages per launch / (32/R))
    int R = 1;         // stage radius
    int64_t tile_steps = 0;  // steps advanced by one swept tile launch
    int simds = 0;           // SIMDs of the plan's device (4 per compute unit)
    int64_t cycle_steps = 0; // reference t_fullstep / dt
    Coefs coef{};
    double consts[9]{};
    int nconsts = 0;

    void* buf[2]{nullptr, nullptr};  // each nc*n REALs, component-major
    int cur = 0;                     // buffer holding the current (flat) state
    hipStream_t stream = nullptr;
    // classic algorithm: a hipGraph of kClassicGraphSteps steps per starting buffer
    hipGraphExec_t classic_exec[2]{nullptr, nullptr};
    bool classic_graph_failed = false;
    // trapezoid launches: a hipGraph of kTileGraphLaunches launches per
    // starting buffer
    hipGraphExec_t tile_exec[2]{nullptr, nullptr};
    bool tile_graph_failed = false;

    // hybrid resources
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    void* host_win = nullptr;  // pinned (W + 2R) * nc REALs
    std::vector<unsigned char> host_flux;  // the host node's interface fluxes: 2 (W + 1) Q3<REAL>

    // profiling
    bool profiling = false;
    struct Pending { int kind; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    int64_t k_launches[KK_COUNT]{};
    double k_ms[KK_COUNT]{};
};

namespace {

// ------------------------------------------------------------ constants ---
// Heat1D_SweptShared.cu:56-58, 719-728: fou = th_diff*dt/(ds*ds) in double
// with REAL th_diff and a REAL ds*ds; Fo and 1-2Fo stored as REAL.
template <typename T>
void heat_consts(double dt, double alpha, double ds, T& fo, T& fo2)
{
    const T a = (T)alpha, d = (T)ds;
    const T dd = d * d;
    const double fou = (double)a * dt / (double)dd;
    fo = (T)fou;
    fo2 = (T)(1.0 - 2.0 * fou);
}

// KS1D_SweptShared.cu:45, 917-926.
template <typename T>
KsCoef<T> ks_consts(double dt, double dx_)
{
    const T dx = (T)dx_;
    KsCoef<T> k;
    k.dx4x = T(1) / (T(4) * dx);
    k.dx2 = T(1) / (dx * dx);
    k.dx4 = T(1) / (dx * dx * dx * dx);
    k.dt = (T)dt;
    k.dth = (T)(dt * 0.5);
    return k;
}

// Euler1D_SweptShared.cu:69, 1127-1135, 1144, 1149.
template <typename T>
EulerCoef<T> euler_consts(double dt, int64_t n, double dx_override, double* dx_out)
{
    EulerCoef<T> k;
    k.gam = (T)1.4;
    k.mgam = (T)0.4;
    const T lx = (T)1.0;
    const double dx = dx_override > 0 ? dx_override : (double)(lx / ((T)n - T(2)));
    *dx_out = dx;
    k.dtdx = (T)(dt / dx);
    k.bl = Q3<T>{T(1), T(0), T(1) / k.mgam};
    k.br = Q3<T>{(T)0.125, T(0), (T)(0.1 / (double)k.mgam)};
    return k;
}

void fill_coefs(sw1d_plan* p, double* euler_dx)
{
    const sw1d_problem& q = p->prob;
    Coefs& c = p->coef;
    const double alpha = q.alpha > 0 ? q.alpha : 8.418e-5;
    if (p->eq == EQ_HEAT) {
        const double ds = q.dx > 0 ? q.dx : 0.001;
        heat_consts<double>(q.dt, alpha, ds, c.fo_d, c.fo2_d);
        heat_consts<float>(q.dt, alpha, ds, c.fo_f, c.fo2_f);
        p->nconsts = 2;
        p->consts[0] = p->f64 ? c.fo_d : c.fo_f;
        p->consts[1] = p->f64 ? c.fo2_d : c.fo2_f;
    } else if (p->eq == EQ_KS) {
        const double dx = q.dx > 0 ? q.dx : 0.5;
        c.ks_d = ks_consts<double>(q.dt, dx);
        c.ks_f = ks_consts<float>(q.dt, dx);
        p->nconsts = 5;
        if (p->f64) {
            const KsCoef<double>& k = c.ks_d;
            const double v[5] = {k.dx4x, k.dx2, k.dx4, k.dt, k.dth};
            std::copy(v, v + 5, p->consts);
        } else {
            const KsCoef<float>& k = c.ks_f;
            const double v[5] = {k.dx4x, k.dx2, k.dx4, k.dt, k.dth};
            std::copy(v, v + 5, p->consts);
        }
    } else {
        double dxd, dxf;
        c.eu_d = euler_consts<double>(q.dt, p->n, q.dx, &dxd);
        c.eu_f = euler_consts<float>(q.dt, p->n, q.dx, &dxf);
        *euler_dx = p->f64 ? dxd : dxf;
        p->nconsts = 9;
        if (p->f64) {
            const EulerCoef<double>& k = c.eu_d;
            const double v[9] = {k.gam, k.mgam, k.dtdx, k.bl.x, k.bl.y, k.bl.z, k.br.x, k.br.y, k.br.z};
            std::copy(v, v + 9, p->consts);
        } else {
            const EulerCoef<float>& k = c.eu_f;
            const double v[9] = {k.gam, k.mgam, k.dtdx, k.bl.x, k.bl.y, k.bl.z, k.br.x, k.br.y, k.br.z};
            std::copy(v, v + 9, p->consts);
        }
    }
}

bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

// SW1D_LAUNCH_GRAPHS=0 in the environment: plans and batches launch their
// kernels one by one instead of replaying hipGraphs (the same kernels in the
// same order, so the same bits).  For tools that do not follow graph launches,
// such as hardware-counter collection.
bool launch_graphs_enabled()
{
    static const bool on = [] {
        const char* v = std::getenv("SW1D_LAUNCH_GRAPHS");
        return !(v && v[0] == '0');
    }();
    return on;
}

// splitmix64 (Steele, Lea, Flood 2014): a portable generator, so the same seed
// gives the same synthetic initial condition on every host.
uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double splitmix_uniform(uint64_t& s)  // [0, 1), 53 random bits
{
    return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}
double splitmix_normal(uint64_t& s)  // Box-Muller, cosine branch
{
    const double u1 = 1.0 - splitmix_uniform(s);
    const double u2 = splitmix_uniform(s);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

// SIMDs of the current device: 4 per compute unit on CDNA (0 if the query
// fails).  Cached per device; a plain attribute query, no allocation.
int device_simds()
{
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 0;
    cache[dev].store(4 * cus, std::memory_order_relaxed);
    return 4 * cus;
}

// ------------------------------------------------------------ profiling ---
hipEvent_t take_event(sw1d_plan* p)
{
    if (!p->event_pool.empty()) {
        hipEvent_t e = p->event_pool.back();
        p->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void drain_profile(sw1d_plan* p)
{
    for (auto& q : p->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, q.a, q.b) == hipSuccess) {
            p->k_launches[q.kind] += 1;
            p->k_ms[q.kind] += ms;
        }
        p->event_pool.push_back(q.a);
        p->event_pool.push_back(q.b);
    }
    p->pending.clear();
}

struct ProfScope {
    sw1d_plan* p;
    int kind;
    hipEvent_t a = nullptr;
    ProfScope(sw1d_plan* p_, int k) : p(p_), kind(k)
    {
        if (p->profiling) {
            a = take_event(p);
            if (a) (void)hipEventRecord(a, p->stream);
        }
    }
    ~ProfScope()
    {
        if (a) {
            hipEvent_t b = take_event(p);
            if (b) {
                (void)hipEventRecord(b, p->stream);
                p->pending.push_back({kind, a, b});
            } else {
                p->event_pool.push_back(a);
            }
        }
    }
};

// --------------------------------------------------------- GPU launches ---
int classic_step(sw1d_plan* p)
{
    if (p->eq == EQ_HEAT) {
        ProfScope ps(p, KK_CLASSIC);
        HIP_OR_FAIL(launch_classic(p->eq, p->bc, p->f64, 0, p->n, p->buf[p->cur],
                                   p->buf[p->cur ^ 1], p->coef, p->stream));
        p->cur ^= 1;
    } else {
        for (int st = 0; st < 2; ++st) {
            ProfScope ps(p, KK_CLASSIC);
            HIP_OR_FAIL(launch_classic(p->eq, p->bc, p->f64, st, p->n, p->buf[p->cur],
                                       p->buf[p->cur], p->coef, p->stream));
        }
    }
    return SW1D_OK;
}

// Classic steps in blocks of kClassicGraphSteps replayed from a hipGraph (one
// graph launch instead of 64-128 kernel launches; the kernels and their order
// are the same, so the bits are the same).  Heat alternates its two buffers,
// so a graph is kept per starting buffer; an even block returns to it.  With
// per-kernel profiling on, or if capture fails, the steps are launched one by one.
constexpr int64_t kClassicGraphSteps = 64;

int classic_advance(sw1d_plan* p, int64_t nsteps)
{
    if (!p->profiling && !p->classic_graph_failed && nsteps >= kClassicGraphSteps) {
        const int slot = p->cur;
        if (!p->classic_exec[slot]) {
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal);
            int rc = SW1D_OK;
            for (int64_t s = 0; s < kClassicGraphSteps && e == hipSuccess && rc == SW1D_OK; ++s)
                rc = classic_step(p);
            const hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(p->stream, &g) : e;
            p->cur = slot;  // capture launched nothing
            if (rc == SW1D_OK && e2 == hipSuccess && g)
                e = hipGraphInstantiate(&p->classic_exec[slot], g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (rc != SW1D_OK || e2 != hipSuccess || e != hipSuccess) {
                p->classic_exec[slot] = nullptr;
                p->classic_graph_failed = true;
                (void)hipGetLastError();
            }
        }
        if (p->classic_exec[slot]) {
            for (; nsteps >= kClassicGraphSteps; nsteps -= kClassicGraphSteps)
                HIP_OR_FAIL(hipGraphLaunch(p->classic_exec[slot], p->stream));
        }
    }
    int rc = SW1D_OK;
    for (int64_t s = 0; s < nsteps && rc == SW1D_OK; ++s) rc = classic_step(p);
    return rc;
}

int swept_launch(sw1d_plan* p, int tile, bool shifted, int64_t node_first)
{
    SweptLaunch L;
    L.eq = p->eq;
    L.bc = p->bc;
    L.P = p->P;
    L.tile = tile;
    L.n = p->n;
    L.nnodes = p->n / p->W;
    L.node_first = node_first;
    L.offset = shifted ? p->n - p->W / 2 : 0;
    L.in = p->buf[p->cur];
    L.out = p->buf[p->cur ^ 1];
    L.simds = p->simds;
    const int kind = tile == TILE_UP ? KK_UP : tile == TILE_DOWN ? KK_DOWN
                   : (shifted ? KK_SPLIT : KK_DIAMOND);
    {
        ProfScope ps(p, kind);
        HIP_OR_FAIL(launch_swept(L, p->f64, p->coef, p->stream));
    }
    return SW1D_OK;
}

// ------------------------------------------------ hybrid CPU node (alg 2) ---
// Node 0 of a boundary-straddling diamond computed on the host while the GPU
// computes nodes 1..K-1 (the CPU_diamond of Heat:410-468 / Euler:761-835 and
// its orchestration, Heat:565-611, Euler:953-1003).  The window holds the
// node's W points plus R halo points on each side, component-major; the same
// stage formulas and active sets as the GPU tile give the same bits.
template <typename T>
void cpu_node_diamond(sw1d_plan* p, T* win)
{
    const int W = p->W, R = p->R, WW = W + 2 * R, H = W / (2 * R);
    // node-local x = 0 is global n - W/2: global n-1 at x = W/2-1, global 0 at x = W/2.
    const int xg0 = W / 2, xgN = W / 2 - 1;
    // component c of the window, indexed by node-local x (halo x = -R .. -1, W .. W+R-1)
    auto row = [&](int c) { return win + (size_t)c * WW + R; };

    // A stage updates its active set |2x - (W-1)| <= thr, the points x = lo .. W-1-lo
    // (thr is odd, W even), and keeps the others.  Sources and destinations are
    // different rows, or the same point only (the corrector's base), so each
    // stage writes its destination row directly.
    auto stage = [&](int thr, int st) {
        const int lo = std::max(0, (W - 1 - thr) / 2), hi = W - 1 - lo;
        if (p->eq == EQ_HEAT) {
            const T fo = (T)(p->f64 ? p->coef.fo_d : p->coef.fo_f);
            const T fo2 = (T)(p->f64 ? p->coef.fo2_d : p->coef.fo2_f);
            const T* __restrict__ s = row(st == 0 ? 0 : 1);
            T* __restrict__ d = row(st == 0 ? 1 : 0);
            for (int x = lo; x <= hi; ++x) d[x] = heat_ftcs(s[x - 1], s[x], s[x + 1], fo, fo2);
            if (p->bc == BC_MIRROR) {  // insulated ends (Heat:140-142)
                if (lo <= xg0 && xg0 <= hi) d[xg0] = heat_ftcs(s[xg0 + 1], s[xg0], s[xg0 + 1], fo, fo2);
                if (lo <= xgN && xgN <= hi) d[xgN] = heat_ftcs(s[xgN - 1], s[xgN], s[xgN - 1], fo, fo2);
            }
        } else {
            const EulerCoef<T>& k = *reinterpret_cast<const EulerCoef<T>*>(
                p->f64 ? (const void*)&p->coef.eu_d : (const void*)&p->coef.eu_f);
            const int src = st == 0 ? 0 : 3, dst = 3 - src;
            const T *s0 = row(src), *s1 = row(src + 1), *s2 = row(src + 2);
            T *d0 = row(dst), *d1 = row(dst + 1), *d2 = row(dst + 2);
            auto cell = [&](int x) { return Q3<T>{s0[x], s1[x], s2[x]}; };
            // Each interface x+1/2 (x = lo-1 .. hi) once, for the two cells beside it:
            // the values eu_net_flux gives each cell (Euler:307-389), the interface
            // right of the ghost cell 0 dropping pdiff(-1,0) and the one left of
            // the ghost cell n-1 dropping pdiff(n-1,n) (the flags of Euler:405, :424).
            Q3<T>* F = reinterpret_cast<Q3<T>*>(p->host_flux.data());
            Q3<T>* D = F + (W + 1);
            for (int x = lo - 1; x <= hi; ++x) {
                const T q0 = x == xg0 ? T(0) : eu_pdiff(cell(x - 1), cell(x));
                const T q1 = eu_pdiff(cell(x), cell(x + 1));
                const T q2 = x == xgN - 1 ? T(0) : eu_pdiff(cell(x + 1), cell(x + 2));
                eu_iface(cell(x - 1), cell(x), cell(x + 1), cell(x + 2), q0, q1, q2, k, F[x - lo + 1],
                         D[x - lo + 1]);
            }
            const T c = st == 0 ? T(0.25) * k.dtdx : T(0.5) * k.dtdx;
            for (int x = lo; x <= hi; ++x) {
                Q3<T> q;
                if (x == xg0) {
                    q = k.bl;
                } else if (x == xgN) {
                    q = k.br;
                } else {
                    const Q3<T> f = eu_balance(F[x - lo], D[x - lo], F[x - lo + 1], D[x - lo + 1]);
                    q = eu_apply(st == 0 ? cell(x) : Q3<T>{d0[x], d1[x], d2[x]}, f, c);
                }
                d0[x] = q.x;
                d1[x] = q.y;
                d2[x] = q.z;
            }
        }
    };
    for (int s = 1; s <= H; s += 2) { stage(2 * R * s - 1, 0); stage(2 * R * (s + 1) - 1, 1); }
    for (int m = 1; m < H - 1; m += 2) { stage(W - 1 - 2 * R * m, 0); stage(W - 1 - 2 * R * (m + 1), 1); }
    stage(W - 1 - 2 * R * (H - 1), 0);
}

int hybrid_split_diamond(sw1d_plan* p)
{
    const int W = p->W, R = p->R, WW = W + 2 * R;
    const int64_t n = p->n;
    const size_t S = p->S;
    char* in = static_cast<char*>(p->buf[p->cur]);
    char* out = static_cast<char*>(p->buf[p->cur ^ 1]);
    char* win = static_cast<char*>(p->host_win);

    HIP_OR_FAIL(hipEventRecord(p->ev_in, p->stream));
    // GPU part: nodes 1..K-1 of the boundary-straddling diamond.
    int rc = swept_launch(p, TILE_DIAMOND, true, 1);
    if (rc) return rc;
    // Node 0 window: global [n - W/2 - R, n) then [0, W/2 + R).
    HIP_OR_FAIL(hipStreamWaitEvent(p->copy_stream, p->ev_in, 0));
    const int64_t left = W / 2 + R, right = W / 2 + R;
    for (int c = 0; c < p->nc; ++c) {
        char* wc = win + (size_t)c * WW * S;
        const char* ic = in + (size_t)c * n * S;
        HIP_OR_FAIL(hipMemcpyAsync(wc, ic + (size_t)(n - left) * S, (size_t)left * S,
                                   hipMemcpyDeviceToHost, p->copy_stream));
        HIP_OR_FAIL(hipMemcpyAsync(wc + (size_t)left * S, ic, (size_t)right * S,
                                   hipMemcpyDeviceToHost, p->copy_stream));
    }
    HIP_OR_FAIL(hipStreamSynchronize(p->copy_stream));
    if (p->f64)
        cpu_node_diamond<double>(p, reinterpret_cast<double*>(win));
    else
        cpu_node_diamond<float>(p, reinterpret_cast<float*>(win));
    // Node 0 result: window [R, R + W/2) -> [n - W/2, n), [R + W/2, R + W) -> [0, W/2).
    for (int c = 0; c < p->nc; ++c) {
        const char* wc = win + (size_t)c * WW * S;
        char* oc = out + (size_t)c * n * S;
        HIP_OR_FAIL(hipMemcpyAsync(oc + (size_t)(n - W / 2) * S, wc + (size_t)R * S,
                                   (size_t)(W / 2) * S, hipMemcpyHostToDevice, p->copy_stream));
        HIP_OR_FAIL(hipMemcpyAsync(oc, wc + (size_t)(R + W / 2) * S, (size_t)(W / 2) * S,
                                   hipMemcpyHostToDevice, p->copy_stream));
    }
    HIP_OR_FAIL(hipEventRecord(p->ev_out, p->copy_stream));
    HIP_OR_FAIL(hipStreamWaitEvent(p->stream, p->ev_out, 0));
    // The pinned window is reused by the next CPU node; its H2D copies must
    // be complete before the host writes it again.
    HIP_OR_FAIL(hipEventSynchronize(p->ev_out));
    return SW1D_OK;
}

// Trapezoid launches in blocks of kTileGraphLaunches replayed from a hipGraph:
// one graph launch instead of 16 kernel launches.  Small plans (launches of a
// few microseconds, many plans at once on their own streams) are otherwise
// bound by the host's launch rate: a loop of 512 such plans runs 1.4x faster.
// A block is an even number of launches, so it returns to its starting
// buffer; a graph is kept per starting buffer and both are captured together,
// so a later call never captures.  The kernels and their order are those of
// the launches one by one, so the bits are the same.  With per-kernel
// profiling on, or if capture fails, the launches go one by one.
// issue_block launches one block (flipping p->cur).
constexpr int64_t kTileGraphLaunches = 16;

template <typename F>
int tile_graph_blocks(sw1d_plan* p, int64_t& launches, F&& issue_block)
{
    if (p->profiling || p->tile_graph_failed || launches < kTileGraphLaunches) return SW1D_OK;
    for (int slot = 0; slot < 2 && !p->tile_graph_failed; ++slot) {
        if (p->tile_exec[slot]) continue;
        const int keep = p->cur;
        p->cur = slot;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal);
        const int rc = e == hipSuccess ? issue_block() : SW1D_OK;
        const hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(p->stream, &g) : e;
        if (rc == SW1D_OK && e2 == hipSuccess && g)
            e = hipGraphInstantiate(&p->tile_exec[slot], g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        p->cur = keep;  // capture launched nothing
        if (rc != SW1D_OK || e2 != hipSuccess || e != hipSuccess) {
            p->tile_exec[slot] = nullptr;
            p->tile_graph_failed = true;
            (void)hipGetLastError();
        }
    }
    if (p->tile_graph_failed) return SW1D_OK;
    for (; launches >= kTileGraphLaunches; launches -= kTileGraphLaunches)
        HIP_OR_FAIL(hipGraphLaunch(p->tile_exec[p->cur], p->stream));
    return SW1D_OK;
}

int swept_advance(sw1d_plan* p, int64_t nsteps)
{
    const int64_t H = p->tile_steps;
    int64_t rem = nsteps;
    if (nsteps >= H) {
        const int64_t tiles = nsteps / H;  // up + (tiles-1) diamonds + down
        rem = nsteps - tiles * H;
        bool shifted = false;
        int rc = swept_launch(p, TILE_UP, shifted, 0);
        if (rc) return rc;
        p->cur ^= 1;
        // Diamonds go one launch at a time.  A diamond lasts longer than the
        // host takes to launch the next, so a hipGraph saves nothing here,
        // and replayed from one they ran 0.3% (2^20) to 2% (2^14) slower.
        for (int64_t i = 0; i + 1 < tiles; ++i) {
            shifted = !shifted;
            if (shifted && p->alg == SW1D_ALT && p->eq != EQ_KS)
                rc = hybrid_split_diamond(p);
            else
                rc = swept_launch(p, TILE_DIAMOND, shifted, 0);
            if (rc) return rc;
            p->cur ^= 1;
        }
        shifted = !shifted;
        rc = swept_launch(p, TILE_DOWN, shifted, 0);
        if (rc) return rc;
        p->cur ^= 1;
    }
    for (int64_t s = 0; s < rem; ++s) {
        int rc = classic_step(p);
        if (rc) return rc;
    }
    return SW1D_OK;
}

int trapezoid_advance(sw1d_plan* p, int64_t nsteps)
{
    const int64_t H = p->tile_steps;
    const int64_t launches = nsteps / H;
    TrapLaunch L;
    L.eq = p->eq;
    L.bc = p->bc;
    L.PO = p->P;
    L.SB = p->SB;
    L.n = p->n;
    L.ntiles = (p->n + p->W - 1) / p->W;
    L.tile_first = 0;
    auto one = [p, &L]() -> int {
        L.in = p->buf[p->cur];
        L.out = p->buf[p->cur ^ 1];
        {
            ProfScope ps(p, KK_DIAMOND);
            HIP_OR_FAIL(launch_trapezoid(L, p->f64, p->coef, p->stream));
        }
        p->cur ^= 1;
        return SW1D_OK;
    };
    int64_t left = launches;
    int rc = tile_graph_blocks(p, left, [&one] {
        for (int64_t j = 0; j < kTileGraphLaunches; ++j) {
            const int r = one();
            if (r) return r;
        }
        return (int)SW1D_OK;
    });
    if (rc) return rc;
    for (int64_t i = 0; i < left; ++i) {
        rc = one();
        if (rc) return rc;
    }
    for (int64_t s = launches * H; s < nsteps; ++s) {
        int rc = classic_step(p);
        if (rc) return rc;
    }
    return SW1D_OK;
}

void release(sw1d_plan* p)
{
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (auto& q : p->pending) {
        (void)hipEventDestroy(q.a);
        (void)hipEventDestroy(q.b);
    }
    for (auto e : p->event_pool) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; ++i)
        if (p->buf[i]) (void)hipFree(p->buf[i]);
    if (p->host_win) (void)hipHostFree(p->host_win);
    if (p->ev_in) (void)hipEventDestroy(p->ev_in);
    if (p->ev_out) (void)hipEventDestroy(p->ev_out);
    if (p->copy_stream) (void)hipStreamDestroy(p->copy_stream);
    for (auto x : p->classic_exec)
        if (x) (void)hipGraphExecDestroy(x);
    for (auto x : p->tile_exec)
        if (x) (void)hipGraphExecDestroy(x);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

// Result of the DPP lane-shift self check: 0 not run yet, 1 passed, -1
// failed.  Plans may be created from several host threads at once.
std::atomic<int> g_lane_check{0};

// Euler states: callers hold interleaved (rho, m, E) cells (the reference's
// REALthree arrays), the plan three component arrays plus their predictor
// copies.  Bit copies of 4- or 8-byte words, one thread per cell.
template <typename U>
__global__ __launch_bounds__(256) void k_split3(const U* __restrict__ in, U* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const U v = in[3 * i + c];
        out[c * n + i] = v;
        out[(c + 3) * n + i] = v;  // the predictor half starts as a copy
    }
}
template <typename U>
__global__ __launch_bounds__(256) void k_join3(const U* __restrict__ in, U* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[3 * i + c] = in[c * n + i];
}

hipError_t euler_layout(sw1d_plan* p, const void* in, void* out, bool split)
{
    const int64_t n = p->n;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (p->S == 8) {
        if (split)
            hipLaunchKernelGGL(k_split3<uint64_t>, grid, block, 0, p->stream, static_cast<const uint64_t*>(in),
                               static_cast<uint64_t*>(out), n);
        else
            hipLaunchKernelGGL(k_join3<uint64_t>, grid, block, 0, p->stream, static_cast<const uint64_t*>(in),
                               static_cast<uint64_t*>(out), n);
    } else {
        if (split)
            hipLaunchKernelGGL(k_split3<uint32_t>, grid, block, 0, p->stream, static_cast<const uint32_t*>(in),
                               static_cast<uint32_t*>(out), n);
        else
            hipLaunchKernelGGL(k_join3<uint32_t>, grid, block, 0, p->stream, static_cast<const uint32_t*>(in),
                               static_cast<uint32_t*>(out), n);
    }
    return hipGetLastError();
}

// Validation, constants and tile geometry of a problem: everything a plan
// holds except its stream and memory.  Allocates nothing and makes no device
// call.
int plan_setup(sw1d_plan* p, const sw1d_problem* q)
{
    if (q->equation < SW1D_HEAT || q->equation > SW1D_EULER)
        return fail(SW1D_EINVAL, "unknown equation %d", q->equation);
    if (q->precision != SW1D_F64 && q->precision != SW1D_F32)
        return fail(SW1D_EINVAL, "unknown precision %d", q->precision);
    if (q->algorithm < SW1D_CLASSIC || q->algorithm > SW1D_TRAPEZOID)
        return fail(SW1D_EINVAL, "unknown algorithm %d", q->algorithm);
    if (q->nx < 8) return fail(SW1D_EINVAL, "nx must be >= 8 (got %lld)", (long long)q->nx);
    if (!(q->dt > 0) || !std::isfinite(q->dt)) return fail(SW1D_EINVAL, "dt must be positive");

    p->prob = *q;
    p->eq = q->equation;
    p->alg = q->algorithm;
    p->f64 = q->precision == SW1D_F64;
    p->n = q->nx;
    p->S = p->f64 ? 8 : 4;
    p->nc = eq_components(p->eq);
    p->R = eq_radius(p->eq);
    int bc = q->bc;
    if (bc == SW1D_BC_DEFAULT) bc = p->eq == EQ_HEAT ? BC_MIRROR : (p->eq == EQ_KS ? BC_PERIODIC : BC_DIRICHLET);
    if ((p->eq == EQ_HEAT && bc != BC_MIRROR && bc != BC_PERIODIC) ||
        (p->eq == EQ_KS && bc != BC_PERIODIC) || (p->eq == EQ_EULER && bc != BC_DIRICHLET)) {
        return fail(SW1D_EINVAL, "boundary condition %d is not defined for this equation", q->bc);
    }
    p->bc = bc;

    double euler_dx = 0;
    fill_coefs(p, &euler_dx);
    if (p->eq == EQ_EULER && p->consts[2] > 0.21) {
        return fail(SW1D_EUNSTABLE,
                    "The value of dt/dx (%g) is too high.  In general it must be <=.21 for stability.",
                    p->consts[2]);
    }

    // Reference cycle: t_fullstep / dt (Heat:483/541, KS:616/676/755, Euler:863/924).
    const int tpb = q->tpb > 0 ? q->tpb : 64;
    if (p->alg == SW1D_CLASSIC) p->cycle_steps = p->eq == EQ_HEAT ? 2 : 1;
    else if (p->alg == SW1D_TRAPEZOID) p->cycle_steps = 0;
    else if (p->eq == EQ_HEAT) p->cycle_steps = tpb;
    else if (p->eq == EQ_KS && p->alg == SW1D_ALT) p->cycle_steps = 8;
    else p->cycle_steps = tpb / 4;

    if (p->alg == SW1D_TRAPEZOID) {
        const int SB = q->tile_depth > 0 ? q->tile_depth : 1;
        int W = q->node_width > 0 ? q->node_width
              : (p->eq == EQ_HEAT ? 1024 : (p->eq == EQ_KS ? 512 : 256));
        if (!is_pow2(W) || W < 128 || !trapezoid_supported(p->eq, W / 64, SB)) {
            return fail(SW1D_ENOTSUP, "no trapezoid tile of %d points and depth %d for this equation",
                        W, SB);
        }
        if (p->n < 2 * (int64_t)W) {
            return fail(SW1D_ENOTSUP, "nx=%lld is below two trapezoid tiles of %d points",
                        (long long)p->n, W);
        }
        p->W = W;
        p->P = W / 64;
        p->SB = SB;
        p->tile_steps = trapezoid_steps(p->eq, SB);
    } else if (p->alg != SW1D_CLASSIC) {
        int W = q->node_width > 0 ? q->node_width : tpb;
        if (p->eq == EQ_KS && p->alg == SW1D_ALT && q->node_width <= 0) W = 128;
        const int pmin = min_points_per_lane(p->eq), pmax = max_points_per_lane(p->eq, p->f64);
        if (!is_pow2(W)) {
            return fail(SW1D_EINVAL, "swept node width must be a power of two (got %d)", W);
        }
        W = std::max(W, 64 * pmin);
        W = std::min(W, 64 * pmax);
        while (W > 64 * pmin && (p->n % W != 0 || p->n < 2 * (int64_t)W)) W /= 2;
        if (p->n % W != 0 || p->n < 2 * (int64_t)W) {
            return fail(SW1D_ENOTSUP, "nx=%lld cannot be split into swept nodes of >= %d points",
                        (long long)p->n, 64 * pmin);
        }
        p->W = W;
        p->P = W / 64;
        p->tile_steps = p->eq == EQ_HEAT ? W / 2 : W / 8;
    }
    return SW1D_OK;
}

}  // namespace

// ================================================ allocation-free advance ===
// Workspace of sw1d_advance_noalloc: the two component-major state buffers of
// the plan it stands in for, as file-scope device memory (one copy per GPU,
// held until the process ends).  On the host `scratch` is only a stand-in
// symbol, never a device address: host code uses the address that
// hipGetSymbolAddress returns on the current device (g_scratch_ptr).
static __device__ __align__(256) unsigned char scratch[SW1D_NOALLOC_MAX_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Once per device, with the device current: resolve the workspace's device
// address and run the lane-shift check in its first word (synchronous, on
// the null stream).
int noalloc_device_init(int dev, void** ws)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* a = nullptr;
        const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !a)
            return fail(SW1D_EDEVICE, "hipGetSymbolAddress(workspace) on device %d: %s", dev,
                        hipGetErrorString(e));
        if (g_lane_check == 0) {
            const int r = check_lane_shifts_at(static_cast<int*>(a));
            if (r < 0) return fail(SW1D_EDEVICE, "lane-shift check could not run on device %d", dev);
            g_lane_check = r == 1 ? 1 : -1;
        }
        g_scratch_ptr[dev] = a;
    }
    *ws = g_scratch_ptr[dev];
    return SW1D_OK;
}

}  // namespace

// ================================================================= C ABI ===
extern "C" {

int sw1d_abi_version(void) { return SW1D_ABI_VERSION; }

const char* sw1d_last_error(void) { return g_err.c_str(); }

int sw1d_device_check(void)
{
    const int r = check_lane_shifts();
    if (r < 0) return fail(SW1D_EDEVICE, "lane-shift check could not run on the device");
    g_lane_check = r == 1 ? 1 : -1;
    return r;
}

int sw1d_plan_create(sw1d_plan** out, const sw1d_problem* q)
{
    if (!out || !q) return fail(SW1D_EINVAL, "null argument");
    *out = nullptr;
    sw1d_plan* p = new sw1d_plan();
    int rc = plan_setup(p, q);
    if (rc) {
        delete p;
        return rc;
    }
    if (!launch_graphs_enabled()) p->classic_graph_failed = p->tile_graph_failed = true;
    p->simds = device_simds();
    if (p->alg != SW1D_CLASSIC) {
        if (g_lane_check == 0) sw1d_device_check();
        if (g_lane_check != 1) {
            delete p;
            return fail(SW1D_EDEVICE, "DPP lane-shift self check failed on this device");
        }
    }

    hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        release(p);
        return fail(SW1D_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    const size_t bytes = (size_t)p->nc * (size_t)p->n * p->S;
    for (int i = 0; i < 2; ++i) {
        e = hipMalloc(&p->buf[i], bytes);
        if (e != hipSuccess) {
            release(p);
            return fail(SW1D_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        }
        e = hipMemsetAsync(p->buf[i], 0, bytes, p->stream);
        if (e != hipSuccess) {
            release(p);
            return fail(SW1D_EDEVICE, "hipMemset: %s", hipGetErrorString(e));
        }
    }
    if (p->alg == SW1D_ALT && p->eq != EQ_KS) {
        const size_t wb = (size_t)(p->W + 2 * p->R) * p->nc * p->S;
        p->host_flux.assign((size_t)2 * (p->W + 1) * 3 * p->S, 0);
        if (hipHostMalloc(&p->host_win, wb, hipHostMallocDefault) != hipSuccess ||
            hipStreamCreateWithFlags(&p->copy_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming) != hipSuccess) {
            release(p);
            return fail(SW1D_ENOMEM, "hybrid host resources could not be allocated");
        }
    }
    e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) {
        release(p);
        return fail(SW1D_EDEVICE, "hipStreamSynchronize: %s", hipGetErrorString(e));
    }
    *out = p;
    return SW1D_OK;
}

int sw1d_plan_destroy(sw1d_plan* p)
{
    release(p);
    return SW1D_OK;
}

// Copies the state in or out of the plan's current buffer on its stream; with
// sync, waits for the copy.
static int copy_state(sw1d_plan* p, const void* src, void* dst, bool to_device, hipMemcpyKind kind,
                      bool sync = true)
{
    const size_t S = p->S;
    const int64_t n = p->n;
    char* dev = static_cast<char*>(p->buf[p->cur]);
    if (p->eq != EQ_EULER) {
        if (to_device) {
            HIP_OR_FAIL(hipMemcpyAsync(dev, src, n * S, kind, p->stream));
            // second value (odd row / u_half) starts as a copy; it is written
            // before it is read
            HIP_OR_FAIL(hipMemcpyAsync(dev + n * S, dev, n * S, hipMemcpyDeviceToDevice, p->stream));
        } else {
            HIP_OR_FAIL(hipMemcpyAsync(dst, dev, n * S, kind, p->stream));
        }
    } else {
        // Host buffers go through the plan's other state buffer (idle between
        // advances, 6n values) in one contiguous copy; device buffers are
        // converted in place of the copy.
        const bool host = kind != hipMemcpyDeviceToDevice;
        void* scratch = p->buf[p->cur ^ 1];
        if (to_device) {
            const void* in = src;
            if (host) {
                HIP_OR_FAIL(hipMemcpyAsync(scratch, src, 3 * n * S, kind, p->stream));
                in = scratch;
            }
            HIP_OR_FAIL(euler_layout(p, in, dev, true));
        } else {
            HIP_OR_FAIL(euler_layout(p, dev, host ? scratch : dst, false));
            if (host) HIP_OR_FAIL(hipMemcpyAsync(dst, scratch, 3 * n * S, kind, p->stream));
        }
    }
    if (sync) HIP_OR_FAIL(hipStreamSynchronize(p->stream));
    return SW1D_OK;
}

int sw1d_set_state(sw1d_plan* p, const void* host)
{
    if (!p || !host) return fail(SW1D_EINVAL, "null argument");
    return copy_state(p, host, nullptr, true, hipMemcpyHostToDevice);
}

int sw1d_get_state(sw1d_plan* p, void* host)
{
    if (!p || !host) return fail(SW1D_EINVAL, "null argument");
    return copy_state(p, nullptr, host, false, hipMemcpyDeviceToHost);
}

int sw1d_set_state_device(sw1d_plan* p, const void* dptr)
{
    if (!p || !dptr) return fail(SW1D_EINVAL, "null argument");
    return copy_state(p, dptr, nullptr, true, hipMemcpyDeviceToDevice);
}

int sw1d_get_state_device(sw1d_plan* p, void* dptr)
{
    if (!p || !dptr) return fail(SW1D_EINVAL, "null argument");
    return copy_state(p, nullptr, dptr, false, hipMemcpyDeviceToDevice);
}

int sw1d_advance(sw1d_plan* p, int64_t nsteps)
{
    if (!p) return fail(SW1D_EINVAL, "null plan");
    if (nsteps < 0) return fail(SW1D_EINVAL, "nsteps must be >= 0");
    int rc;
    if (p->alg == SW1D_CLASSIC) {
        rc = classic_advance(p, nsteps);
    } else if (p->alg == SW1D_TRAPEZOID) {
        rc = trapezoid_advance(p, nsteps);
    } else {
        rc = swept_advance(p, nsteps);
    }
    return rc;
}

int sw1d_advance_noalloc(const sw1d_problem* q, void* d_state, int64_t nsteps, void* stream)
{
    if (!q || !d_state) return fail(SW1D_EINVAL, "null argument");
    if (nsteps < 0) return fail(SW1D_EINVAL, "nsteps must be >= 0");
    sw1d_plan p;  // on the stack, no buffers of its own
    int rc = plan_setup(&p, q);
    if (rc) return rc;
    if (p.alg == SW1D_ALT && p.eq != EQ_KS)
        return fail(SW1D_ENOTSUP, "the hybrid CPU+GPU algorithm needs pinned host memory; "
                                  "the allocation-free advance runs the GPU-only algorithms");
    const size_t half = (size_t)p.nc * (size_t)p.n * p.S;
    if (2 * half > (size_t)SW1D_NOALLOC_MAX_BYTES)
        return fail(SW1D_EINVAL, "nx=%lld needs %zu bytes of workspace; the allocation-free advance "
                    "holds %zu per GPU (use a plan)", (long long)p.n, 2 * half,
                    (size_t)SW1D_NOALLOC_MAX_BYTES);
    int dev = -1;
    HIP_OR_FAIL(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(SW1D_EDEVICE, "device %d out of range", dev);
    hipStream_t s = static_cast<hipStream_t>(stream);
    void* ws = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        ws = g_scratch_ptr[dev];
    }
    if (!ws) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIP_OR_FAIL(hipStreamIsCapturing(s, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(SW1D_EINVAL, "the first allocation-free advance on device %d sets up the "
                        "workspace synchronously and cannot run under stream capture", dev);
        rc = noalloc_device_init(dev, &ws);
        if (rc) return rc;
    }
    if (p.alg != SW1D_CLASSIC && g_lane_check != 1)
        return fail(SW1D_EDEVICE, "DPP lane-shift self check failed on this device");
    p.buf[0] = ws;
    p.buf[1] = static_cast<unsigned char*>(ws) + half;
    p.stream = s;
    p.simds = device_simds();
    // launches one by one: no launch graph is captured or instantiated
    p.classic_graph_failed = true;
    p.tile_graph_failed = true;
    rc = copy_state(&p, d_state, nullptr, true, hipMemcpyDeviceToDevice, false);
    if (rc == SW1D_OK) rc = sw1d_advance(&p, nsteps);
    if (rc == SW1D_OK) rc = copy_state(&p, nullptr, d_state, false, hipMemcpyDeviceToDevice, false);
    if (rc == SW1D_OK) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(SW1D_EDEVICE, "launch: %s", hipGetErrorString(e));
    }
    return rc;
}

int sw1d_synchronize(sw1d_plan* p)
{
    if (!p) return fail(SW1D_EINVAL, "null plan");
    HIP_OR_FAIL(hipStreamSynchronize(p->stream));
    HIP_OR_FAIL(hipGetLastError());
    if (p->profiling) drain_profile(p);
    return SW1D_OK;
}

int sw1d_advance_timed(sw1d_plan* p, int64_t nsteps, double* ms)
{
    if (!p || !ms) return fail(SW1D_EINVAL, "null argument");
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, p->stream);
    int rc = e == hipSuccess ? sw1d_advance(p, nsteps) : SW1D_EDEVICE;
    if (rc == SW1D_OK) {
        float f = 0.f;
        e = hipEventRecord(b, p->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&f, a, b);
        if (e == hipSuccess) *ms = f;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (e != hipSuccess) return fail(SW1D_EDEVICE, "timing events: %s", hipGetErrorString(e));
    return rc == SW1D_OK ? sw1d_synchronize(p) : rc;
}

int sw1d_plan_info(const sw1d_plan* p, int64_t info[5])
{
    if (!p || !info) return fail(SW1D_EINVAL, "null argument");
    info[0] = p->W;
    info[1] = p->alg == SW1D_CLASSIC ? 1 : p->tile_steps;
    info[2] = p->nc;
    info[3] = (int64_t)(2 * (size_t)p->nc * (size_t)p->n * p->S);
    info[4] = p->cycle_steps;
    return SW1D_OK;
}

int sw1d_plan_constants(const sw1d_plan* p, double* out, int32_t cap)
{
    if (!p || !out) return fail(SW1D_EINVAL, "null argument");
    if (cap < p->nconsts) return fail(SW1D_EINVAL, "need %d slots", p->nconsts);
    std::copy(p->consts, p->consts + p->nconsts, out);
    return p->nconsts;
}

int sw1d_plan_stream(const sw1d_plan* p, void** s)
{
    if (!p || !s) return fail(SW1D_EINVAL, "null argument");
    *s = (void*)p->stream;
    return SW1D_OK;
}

int sw1d_set_profiling(sw1d_plan* p, int32_t on)
{
    if (!p) return fail(SW1D_EINVAL, "null plan");
    HIP_OR_FAIL(hipStreamSynchronize(p->stream));
    drain_profile(p);
    p->profiling = on != 0;
    for (int k = 0; k < KK_COUNT; ++k) { p->k_launches[k] = 0; p->k_ms[k] = 0; }
    return SW1D_OK;
}

int sw1d_kernel_stats(sw1d_plan* p, int32_t kind, int64_t* launches, double* ms)
{
    if (!p || !launches || !ms) return fail(SW1D_EINVAL, "null argument");
    if (kind < 0 || kind >= KK_COUNT) return fail(SW1D_EINVAL, "unknown kernel kind %d", kind);
    HIP_OR_FAIL(hipStreamSynchronize(p->stream));
    drain_profile(p);
    *launches = p->k_launches[kind];
    *ms = p->k_ms[kind];
    return SW1D_OK;
}

int sw1d_schedule(int32_t kind, double dt, double tf_arg, double freq, int32_t tpb,
                  double dt_sched, int64_t* steps, double* times, int64_t cap, int64_t* count)
{
    if (!count) return fail(SW1D_EINVAL, "null count");
    std::vector<SchedRow> rows;
    if (!run_schedule(kind, dt, tf_arg, freq, tpb, dt_sched, rows))
        return fail(SW1D_EINVAL, "unknown schedule kind %d or tpb %d", kind, tpb);
    *count = (int64_t)rows.size();
    if ((int64_t)rows.size() > cap || (!steps && cap > 0) || (!times && cap > 0))
        return fail(SW1D_EINVAL, "capacity %lld < %zu rows", (long long)cap, rows.size());
    for (size_t i = 0; i < rows.size(); ++i) {
        steps[i] = rows[i].steps;
        times[i] = rows[i].time;
    }
    return SW1D_OK;
}

int sw1d_initial_condition(int32_t kind, int32_t precision, int64_t n, double length, double param,
                           uint64_t seed, void* host)
{
    if (!host) return fail(SW1D_EINVAL, "null output buffer");
    if (n < 1) return fail(SW1D_EINVAL, "n must be positive, got %lld", (long long)n);
    if (precision != SW1D_F64 && precision != SW1D_F32)
        return fail(SW1D_EINVAL, "unknown precision %d", precision);
    if (!(length > 0.0)) return fail(SW1D_EINVAL, "length/spacing must be positive");
    std::vector<double> u((size_t)n);
    uint64_t st = seed;
    if (kind == SW1D_IC_KS_NOISY) {
        // 2cos(19 pi x/128) at x = k dx (KS1D_SweptShared.cu:68-72, 945), plus
        // param * N(0,1) per point
        for (int64_t k = 0; k < n; ++k) {
            const double x = (double)k * length;
            u[(size_t)k] = 2.0 * cos(19.0 * x * M_PI / 128.0) + param * splitmix_normal(st);
        }
    } else if (kind == SW1D_IC_HEAT_FOURIER) {
        // 1 + sum_{k=1..4} a_k sin(2 pi k x / L), a_k ~ U[0,1), x = j L / n
        double a[4];
        for (double& v : a) v = splitmix_uniform(st);
        for (int64_t j = 0; j < n; ++j) {
            const double x = (double)j * (length / (double)n);
            double v = 1.0;
            for (int k = 0; k < 4; ++k) v += a[k] * sin(2.0 * M_PI * (double)(k + 1) * x / length);
            u[(size_t)j] = v;
        }
    } else {
        return fail(SW1D_EINVAL, "unknown initial condition kind %d", kind);
    }
    if (precision == SW1D_F64) {
        std::memcpy(host, u.data(), sizeof(double) * (size_t)n);
    } else {
        float* f = static_cast<float*>(host);
        for (int64_t k = 0; k < n; ++k) f[k] = (float)u[(size_t)k];
    }
    return SW1D_OK;
}

}  // extern "C"

// ============================================================ batches ===
struct sw1d_batch {
    bool f64 = true;
    bool all_ks = true;
    size_t S = 8;
    int64_t count = 0, total = 0;
    std::vector<int64_t> offset, n;
    std::vector<char> jobs_host;  // BatchJob<T>[count], queue order
    void* jobs = nullptr;
    int* next = nullptr;
    void* a = nullptr;
    void* b = nullptr;
    int32_t grid = 0;
    int32_t waves = 4;  // waves per workgroup (tiles worked on at once)
    int32_t cus = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t graph_steps = -1;
};

namespace {

template <typename T>
void build_jobs(sw1d_batch* B, const int32_t* eq, const double* dt, const double* dx,
                const int32_t* tile_points)
{
    std::vector<int64_t> order((size_t)B->count);
    for (int64_t i = 0; i < B->count; ++i) order[i] = i;
    // largest problems first: a greedy longest-processing-time queue
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return B->n[x] > B->n[y]; });
    B->jobs_host.assign(sizeof(BatchJob<T>) * (size_t)B->count, 0);
    BatchJob<T>* J = reinterpret_cast<BatchJob<T>*>(B->jobs_host.data());
    for (int64_t k = 0; k < B->count; ++k) {
        const int64_t i = order[k];
        BatchJob<T>& j = J[k];
        j.offset = B->offset[i];
        j.n = (int32_t)B->n[i];
        j.eq = eq[i];
        int tp = tile_points ? tile_points[i] : 0;
        if (tp <= 0) tp = 512;  // widest tile that won every size class of the tuned workload
        j.po = tp / 64;
        if (eq[i] == EQ_HEAT) {
            heat_consts<T>(dt[i], 8.418e-5, dx[i], j.c[0], j.c[1]);
        } else {
            const KsCoef<T> k2 = ks_consts<T>(dt[i], dx[i]);
            j.c[0] = k2.dx4x; j.c[1] = k2.dx2; j.c[2] = k2.dx4; j.c[3] = k2.dt; j.c[4] = k2.dth;
        }
    }
}

void batch_release(sw1d_batch* B)
{
    if (!B) return;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    if (B->jobs) (void)hipFree(B->jobs);
    if (B->next) (void)hipFree(B->next);
    if (B->a) (void)hipFree(B->a);
    if (B->b) (void)hipFree(B->b);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
}


}  // namespace

extern "C" {

int sw1d_batch_create(sw1d_batch** out, int32_t precision, int64_t count, const int32_t* eq,
                      const int64_t* nx, const double* dt, const double* dx,
                      const int32_t* tile_points)
{
    if (!out || !eq || !nx || !dt || !dx) return fail(SW1D_EINVAL, "null argument");
    *out = nullptr;
    if (count <= 0 || count > (1 << 30)) return fail(SW1D_EINVAL, "bad problem count %lld", (long long)count);
    if (precision != SW1D_F64 && precision != SW1D_F32) return fail(SW1D_EINVAL, "unknown precision");
    sw1d_batch* B = new sw1d_batch();
    B->f64 = precision == SW1D_F64;
    B->S = B->f64 ? 8 : 4;
    B->count = count;
    B->offset.resize((size_t)count);
    B->n.resize((size_t)count);
    for (int64_t i = 0; i < count; ++i) {
        if (eq[i] != SW1D_HEAT && eq[i] != SW1D_KS) {
            delete B;
            return fail(SW1D_EINVAL, "problem %lld: batches hold periodic heat or KS", (long long)i);
        }
        if (nx[i] < 128 || nx[i] > (int64_t)1 << 30 || !(dt[i] > 0) || !(dx[i] > 0)) {
            delete B;
            return fail(SW1D_EINVAL, "problem %lld: nx >= 128, dt > 0, dx > 0 required", (long long)i);
        }
        const int tp = tile_points ? tile_points[i] : 0;
        if (tp != 0 && !batch_tile_supported(tp / 64)) {
            delete B;
            return fail(SW1D_EINVAL, "problem %lld: tile_points must be 128, 256, 512 or 1024", (long long)i);
        }
        B->offset[i] = B->total;
        B->n[i] = nx[i];
        B->all_ks = B->all_ks && eq[i] == SW1D_KS;
        B->total += nx[i];
    }
    if (B->f64) build_jobs<double>(B, eq, dt, dx, tile_points);
    else build_jobs<float>(B, eq, dt, dx, tile_points);

    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        delete B;
        return fail(SW1D_EDEVICE, "no HIP device");
    }
    B->cus = cus;
    const int64_t g = (int64_t)cus * batch_blocks_per_cu(B->f64, B->waves);
    B->grid = (int32_t)std::min<int64_t>(g, count);
    const size_t bytes = (size_t)B->total * B->S;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&B->a, bytes) != hipSuccess || hipMalloc(&B->b, bytes) != hipSuccess ||
        hipMalloc(&B->jobs, B->jobs_host.size()) != hipSuccess ||
        hipMalloc((void**)&B->next, sizeof(int)) != hipSuccess) {
        batch_release(B);
        return fail(SW1D_ENOMEM, "batch device allocation of %zu bytes failed", 2 * bytes);
    }
    hipError_t e = hipMemcpy(B->jobs, B->jobs_host.data(), B->jobs_host.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        batch_release(B);
        return fail(SW1D_EDEVICE, "job table upload: %s", hipGetErrorString(e));
    }
    *out = B;
    return SW1D_OK;
}

int sw1d_batch_destroy(sw1d_batch* B)
{
    batch_release(B);
    return SW1D_OK;
}

int sw1d_batch_set_state(sw1d_batch* B, const void* host)
{
    if (!B || !host) return fail(SW1D_EINVAL, "null argument");
    HIP_OR_FAIL(hipMemcpyAsync(B->a, host, (size_t)B->total * B->S, hipMemcpyHostToDevice, B->stream));
    HIP_OR_FAIL(hipStreamSynchronize(B->stream));
    return SW1D_OK;
}

int sw1d_batch_get_state(sw1d_batch* B, void* host)
{
    if (!B || !host) return fail(SW1D_EINVAL, "null argument");
    HIP_OR_FAIL(hipMemcpyAsync(host, B->a, (size_t)B->total * B->S, hipMemcpyDeviceToHost, B->stream));
    HIP_OR_FAIL(hipStreamSynchronize(B->stream));
    return SW1D_OK;
}

int sw1d_batch_advance(sw1d_batch* B, int64_t nsteps)
{
    if (!B) return fail(SW1D_EINVAL, "null batch");
    if (nsteps < 0) return fail(SW1D_EINVAL, "nsteps must be >= 0");
    if (nsteps == 0) return SW1D_OK;
    BatchLaunch L;
    L.jobs = B->jobs;
    L.njobs = (int32_t)B->count;
    L.next = B->next;
    L.a = B->a;
    L.b = B->b;
    L.nsteps = nsteps;
    L.grid = B->grid;
    L.waves = B->waves;
    if (!launch_graphs_enabled()) {
        HIP_OR_FAIL(launch_batch(L, B->f64, B->stream));
        return SW1D_OK;
    }
    if (B->graph_steps != nsteps) {
        // the previous graph's launch may still run on the stream: let it
        // finish before its executable is destroyed
        HIP_OR_FAIL(hipStreamSynchronize(B->stream));
        if (B->exec) { (void)hipGraphExecDestroy(B->exec); B->exec = nullptr; }
        if (B->graph) { (void)hipGraphDestroy(B->graph); B->graph = nullptr; }
        B->graph_steps = -1;
        HIP_OR_FAIL(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = launch_batch(L, B->f64, B->stream);
        const hipError_t ce = hipStreamEndCapture(B->stream, &B->graph);
        if (le != hipSuccess) return fail(SW1D_EDEVICE, "batch launch capture: %s", hipGetErrorString(le));
        if (ce != hipSuccess) return fail(SW1D_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ce));
        HIP_OR_FAIL(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
        B->graph_steps = nsteps;
    }
    HIP_OR_FAIL(hipGraphLaunch(B->exec, B->stream));
    return SW1D_OK;
}

int sw1d_batch_set_workgroup(sw1d_batch* B, int32_t waves)
{
    if (!B) return fail(SW1D_EINVAL, "null batch");
    if (!batch_waves_supported(waves)) return fail(SW1D_EINVAL, "waves per workgroup must be 2, 4 or 8");
    if (waves == B->waves) return SW1D_OK;
    HIP_OR_FAIL(hipStreamSynchronize(B->stream));
    if (B->exec) { (void)hipGraphExecDestroy(B->exec); B->exec = nullptr; }
    if (B->graph) { (void)hipGraphDestroy(B->graph); B->graph = nullptr; }
    B->graph_steps = -1;
    B->waves = waves;
    const int64_t g = (int64_t)B->cus * batch_blocks_per_cu(B->f64, waves);
    B->grid = (int32_t)std::min<int64_t>(g, B->count);
    return SW1D_OK;
}

int sw1d_batch_advance_timed(sw1d_batch* B, int64_t nsteps, double* ms)
{
    if (!B || !ms) return fail(SW1D_EINVAL, "null argument");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, B->stream);
    int rc = e == hipSuccess ? sw1d_batch_advance(B, nsteps) : SW1D_EDEVICE;
    if (rc == SW1D_OK) {
        float f = 0;
        e = hipEventRecord(e1, B->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
        if (e == hipSuccess) *ms = f;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(SW1D_EDEVICE, "batch timing: %s", hipGetErrorString(e));
    return rc;
}

int sw1d_batch_info(const sw1d_batch* B, int64_t info[4])
{
    if (!B || !info) return fail(SW1D_EINVAL, "null argument");
    info[0] = B->total;
    info[1] = B->grid;
    info[2] = B->count;
    info[3] = (int64_t)(2 * (size_t)B->total * B->S + B->jobs_host.size());
    return SW1D_OK;
}

}  // extern "C"

namespace sw1d {
bool batch_view(sw1d_batch* B, BatchView& v)
{
    if (!B) return false;
    v.state = B->a;
    v.count = B->count;
    v.n = B->n.empty() ? 0 : B->n[0];
    for (int64_t x : B->n)
        if (x != v.n) v.n = 0;
    v.f64 = B->f64;
    v.all_ks = B->all_ks;
    v.stream = B->stream;
    return true;
}
}  // namespace sw1d
