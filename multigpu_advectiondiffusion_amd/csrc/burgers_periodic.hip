// burgers_periodic.hip -- fused SSP-RK3 stage for periodic inviscid Burgers
// with WENO5 (Z or JS) global Lax-Friedrichs flux splitting.
//
// One launch per RK stage computes, for every node,
//   L(u) = -(F_{i+1/2}-F_{i-1/2})/dx - (G_{j+1/2}-G_{j-1/2})/dy - (H_{k+1/2}-H_{k-1/2})/dz
// (the accumulation order of Call_Adv_x -> Call_Adv_y -> Call_Adv_z,
// SingleGPU/Burgers3d_WENO5/main.cpp:144-146), the SSP-RK3 combine of
// Compute_RK (kernels.cu:520-527) and max|u_out| for the next stage's LF
// alpha, so u is read once and written once per stage (16-24 B per node).
//
// Work decomposition: a workgroup owns a TX x TY column tile over a band of z
// planes; workgroups are numbered band by band.  Bands are kchunk planes deep
// (by default from default_kchunk), or, with kchunk = MGAD_BANDS_SHRINK, they
// shrink towards the top of the grid: the dispatcher hands out workgroups in
// order as resident ones finish, so the launch then ends with short ones (a
// CU's first-dispatched workgroup issues ahead of its second, so workgroups
// end at uneven times).  Each thread owns one (x,y) column and marches up the
// band with a 7-plane register queue of split fluxes (and of the 13/12
// smoothness terms), so every z face is computed once and reused by the two
// nodes that share it.  For x and y faces the current plane's split fluxes (tile + 3-node
// halo) are staged in LDS; each thread computes the face on its +1/2 side and
// one wave, rotating with the plane index, computes the TY + TX faces on the
// tile's -1/2 edges, so every face of the plane is computed once.
//
// Storage type T is double, or float with all arithmetic in double (split
// fluxes, smoothness indicators, weights, RK combine): loads widen, the store
// rounds once to float.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <tuple>
#include <type_traits>
#include <vector>

#include "burgers_periodic.hpp"

namespace mgad {

// XW: nx or ny is not a multiple of the tile; the last tile of each row or
// column reaches past the grid and stores only its real nodes.
// Workgroup w covers band w / ntiles; within a band, the tiles (x-fastest,
// then y) are remapped so that each XCD's workgroups (w mod 8 under
// round-robin dispatch) take a contiguous run of them and share their halo
// rows in that XCD's L2.
template <class T, int WENO, int P, int TX, int TY, bool XW>
__global__ __launch_bounds__(TX * TY, 4) void k_burgers_stage(const T *__restrict__ uin, const T *un,
                                                           T *uout, GridArgs g, ZBands bands, double eps,
                                                           int stage, Ctrl *ctrl) {
    __shared__ TileSmem<TX, TY> sm;
    double dt, alpha;
    const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
    if (!stage_control(ctrl, stage, g.dx, lead, dt, alpha)) return;
    const unsigned ntx = XW ? (g.nx + TX - 1) / TX : g.nx / TX;
    const unsigned ntiles = ntx * (XW ? (g.ny + TY - 1) / TY : g.ny / TY);
    const unsigned b = blockIdx.x / ntiles;
    const unsigned t = xcd_tile(blockIdx.x - b * ntiles, ntiles);
    const int x0 = static_cast<int>(t % ntx) * TX, y0 = static_cast<int>(t / ntx) * TY;
    const int z0 = bands.z[b], z1 = bands.z[b + 1];
    if (fast_division_ok<WENO>(alpha, eps))
        stage_cols<T, WENO, P, TX, TY, true, false, XW, XW>(uin, un, uout, g, stage, x0, y0, z0, z1, dt, alpha, eps, sm,
                                                        &ctrl->amax[stage % 3], nullptr, nullptr);
    else
        stage_cols<T, WENO, P, TX, TY, false, false, XW, XW>(uin, un, uout, g, stage, x0, y0, z0, z1, dt, alpha, eps,
                                                         sm, &ctrl->amax[stage % 3], nullptr, nullptr);
}

// max|u| over n values into *slot (atomicMax of the bit pattern).
template <class T>
__global__ __launch_bounds__(256) void k_maxabs(const T *__restrict__ u, size_t n,
                                                unsigned long long *slot) {
    __shared__ unsigned long long sRed[4];
    unsigned long long m = 0ull;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
        const unsigned long long b = abs_bits(static_cast<double>(u[i]));
        m = b > m ? b : m;
    }
    m = wave_max_u64(m);
    if ((threadIdx.x & 63) == 0) sRed[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = sRed[0];
        for (int w = 1; w < 4; w++) m = sRed[w] > m ? sRed[w] : m;
        atomicMax(slot, m);
    }
}

__global__ void k_ctrl_init(Ctrl *c, double t0, double tend, double cfl) {
    c->t = t0;
    c->dt = 0.0;
    c->tend = tend;
    c->cfl = cfl;
    c->amax[0] = 0ull;
    c->amax[1] = 0ull;
    c->amax[2] = 0ull;
    c->steps = 0ull;
    c->rec = nullptr;
    c->rec_step0 = 0ull;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

template <class T>
using StageKernel = void (*)(const T *, const T *, T *, GridArgs, ZBands, double, int, Ctrl *);

template <class T>
struct Shape {
    int tx, ty;
    StageKernel<T> fn;
};

// Workgroups of `threads` threads the device keeps resident: the stage
// kernels' launch bounds hold them to 4 waves per SIMD, 16 waves per CU.
int resident_groups(int threads) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    return cus * 16 / (threads / 64);
}

// 32x32 tiles when asked for (kSquare; they reach past the grid unless 32
// divides nx and ny); else tiles that divide the grid when they can: 32x16
// (unless 16 is asked for), else 16x16; otherwise tiles that reach past the
// grid, 32x16 for rows of 64 nodes or more (or when 32 is asked for), else
// 16x16.
constexpr int kSquare = 1024;
template <class T, int WENO, int P>
bool pick(int nx, int ny, int tile_pref, Shape<T> &s) {
    if (tile_pref == kSquare) {
        if (nx % 32 == 0 && ny % 32 == 0)
            s = {32, 32, k_burgers_stage<T, WENO, P, 32, 32, false>};
        else
            s = {32, 32, k_burgers_stage<T, WENO, P, 32, 32, true>};
        return true;
    }
    if (tile_pref != 16 && nx % 32 == 0 && ny % 16 == 0) {
        s = {32, 16, k_burgers_stage<T, WENO, P, 32, 16, false>};
        return true;
    }
    if (nx % 16 == 0 && ny % 16 == 0) {
        s = {16, 16, k_burgers_stage<T, WENO, P, 16, 16, false>};
        return true;
    }
    if (tile_pref == 32 || (tile_pref != 16 && nx >= 64))
        s = {32, 16, k_burgers_stage<T, WENO, P, 32, 16, true>};
    else
        s = {16, 16, k_burgers_stage<T, WENO, P, 16, 16, true>};
    return true;
}

template <class T>
bool pick_kernel(const mgad_scheme &sc, int nx, int ny, int tile_pref, Shape<T> &s) {
    if (sc.weno == MGAD_WENO_Z && sc.p == 1) return pick<T, kWenoZ, 1>(nx, ny, tile_pref, s);
    if (sc.weno == MGAD_WENO_Z && sc.p == 2) return pick<T, kWenoZ, 2>(nx, ny, tile_pref, s);
    if (sc.weno == MGAD_WENO_JS) return pick<T, kWenoJS, 2>(nx, ny, tile_pref, s);
    return false;
}

}  // namespace

ZBands make_bands(int nz, long ntiles, long groups, int kchunk) {
    ZBands b{};
    if (kchunk > 0) {
        const int len = std::max(kchunk, (nz + kMaxBands - 1) / kMaxBands);
        for (int z = 0; z < nz; z += len) b.z[b.n++] = z;
        b.z[b.n] = nz;
        return b;
    }
    // Shrinking bands: with R planes left, the next band has
    // ceil(R * ntiles / (2 * max(groups, ntiles))) planes, at least lmin, so
    // the first round of resident workgroups covers about half the work and
    // the sizes then fall geometrically.  lmin is 8 planes (the 6-plane queue
    // warm-up of a workgroup stays small), or 4 when the grid gives the
    // resident workgroups fewer than 32 tile planes each.
    const long den = 2 * std::max(groups, ntiles);
    const long lmin = ntiles * nz >= 32 * groups ? 8 : 4;
    for (long pos = 0; pos < nz;) {
        b.z[b.n++] = static_cast<int>(pos);
        const long r = nz - pos;
        long len = std::max(lmin, (r * ntiles + den - 1) / den);
        if (b.n == kMaxBands) len = r;
        pos += std::min(len, r);
    }
    b.z[b.n] = nz;
    return b;
}

int square_bands(long tiles, unsigned nz) {
    const long cus = resident_groups(1024);
    if (tiles < 1) return 0;
    const long b = std::min<long>({cus / tiles, static_cast<long>(nz) / 4, kMaxBands});
    return b >= 1 && 10 * tiles * b >= 9 * cus ? static_cast<int>(b) : 0;
}

// square_bands of a grid in 32x32 column tiles; where 32 does not divide nx
// or ny the last tiles reach past the grid, and the tiles may cover at most
// 7 % more columns than the grid has (else 0).
int square_bands(unsigned nx, unsigned ny, unsigned nz) {
    const long tx = (nx + 31) / 32, ty = (ny + 31) / 32;
    if (100 * (tx * 32) * (ty * 32) > 107l * nx * ny) return 0;
    return square_bands(tx * ty, nz);
}

int default_kchunk(int nx, int ny, int nz, int tx, int ty) {
    const long tiles = static_cast<long>((nx + tx - 1) / tx) * ((ny + ty - 1) / ty);
    // Enough blocks to give every CU several tiles; chunks of >= 8 planes so
    // the 6-plane queue warm-up stays small.
    int kc = nz;
    while (kc > 8 && tiles * ((nz + kc - 1) / kc) < 2048) kc = (kc + 1) / 2;
    return kc;
}

// The z queue wraps plane indices by at most one period (z0-3 and k+4), so
// nz >= 4; x and y halo indices wrap from -3, so nx, ny >= 4 as well; indices
// stay below 2^31.
bool valid_grid(unsigned nx, unsigned ny, unsigned nz) {
    if (nx < 4 || ny < 4 || nz < 4) return false;
    const unsigned long long n = 1ull * nx * ny * nz;
    return n < (1ull << 31);
}

GridArgs make_grid(unsigned nx, unsigned ny, unsigned nz, double dx, double dy, double dz, int kchunk) {
    GridArgs g;
    g.nx = static_cast<int>(nx);
    g.ny = static_cast<int>(ny);
    g.nz = static_cast<int>(nz);
    g.kchunk = kchunk;
    g.dx = dx;
    g.dy = dy;
    g.dz = dz;
    g.rdx = 1.0 / dx;
    g.rdy = 1.0 / dy;
    g.rdz = 1.0 / dz;
    return g;
}

template <class T>
int burgers_stage_launch_t(hipStream_t stream, const mgad_scheme &sc, unsigned nx, unsigned ny,
                           unsigned nz, double dx, double dy, double dz, int stage, const T *u_in,
                           const T *u_n, T *u_out, void *ctrl, int kchunk, int tile_pref) {
    if (!valid_grid(nx, ny, nz) || stage < 1 || stage > 3 || !u_in || !u_n || !u_out || !ctrl ||
        u_out == u_in || !(dx > 0) || !(dy > 0) || !(dz > 0)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    if (tile_pref == 0 && kchunk == 0) {
        const int b = square_bands(nx, ny, nz);
        if (b > 0) {
            tile_pref = kSquare;
            kchunk = static_cast<int>((nz + b - 1) / b);
        }
    }
    Shape<T> s;
    if (!pick_kernel<T>(sc, static_cast<int>(nx), static_cast<int>(ny), tile_pref, s)) {
        set_error(MGAD_ERR_UNSUPPORTED);
        return MGAD_ERR_UNSUPPORTED;
    }
    const long ntiles = static_cast<long>((nx + s.tx - 1) / s.tx) * ((ny + s.ty - 1) / s.ty);
    if (kchunk == 0) kchunk = default_kchunk(nx, ny, nz, s.tx, s.ty);
    const ZBands bands = make_bands(static_cast<int>(nz), ntiles, resident_groups(s.tx * s.ty), kchunk);
    const long groups = ntiles * bands.n;
    if (groups >= (1l << 31)) {
        set_error(MGAD_ERR_UNSUPPORTED);
        return MGAD_ERR_UNSUPPORTED;
    }
    const GridArgs g = make_grid(nx, ny, nz, dx, dy, dz, kchunk);
    hipLaunchKernelGGL(s.fn, dim3(static_cast<unsigned>(groups)), dim3(s.tx * s.ty), 0, stream, u_in, u_n, u_out,
                       g, bands, sc.eps, stage, static_cast<Ctrl *>(ctrl));
    return check_launch();
}

template int burgers_stage_launch_t<double>(hipStream_t, const mgad_scheme &, unsigned, unsigned,
                                            unsigned, double, double, double, int, const double *,
                                            const double *, double *, void *, int, int);
template int burgers_stage_launch_t<float>(hipStream_t, const mgad_scheme &, unsigned, unsigned,
                                           unsigned, double, double, double, int, const float *,
                                           const float *, float *, void *, int, int);

template <class T>
int burgers_init_t(hipStream_t stream, unsigned nx, unsigned ny, unsigned nz, const T *u, void *ctrl,
                   double t0, double tend, double cfl) {
    if (!valid_grid(nx, ny, nz) || !u || !ctrl) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    Ctrl *c = static_cast<Ctrl *>(ctrl);
    hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(1), 0, stream, c, t0, tend, cfl);
    int st = check_launch();
    if (st) return st;
    const size_t n = 1ull * nx * ny * nz;
    const unsigned nb = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 2048));
    hipLaunchKernelGGL(k_maxabs<T>, dim3(nb), dim3(256), 0, stream, u, n, &c->amax[0]);
    return check_launch();
}

template int burgers_init_t<double>(hipStream_t, unsigned, unsigned, unsigned, const double *, void *,
                                    double, double, double);
template int burgers_init_t<float>(hipStream_t, unsigned, unsigned, unsigned, const float *, void *,
                                   double, double, double);

// ---------------------------------------------------------------------------
// Launch configuration: environment override (MGAD_TILE, MGAD_KCHUNK), then
// the per-shape table that mgad_burgers_autotune / mgad_burgers_set_launch
// fill, then the default (32-wide tiles when they divide nx, kchunk from
// default_kchunk).  Every configuration computes bit-identical results.
// ---------------------------------------------------------------------------
namespace {

int env_int(const char *name) {
    const char *v = std::getenv(name);
    return v ? std::atoi(v) : 0;
}

using CfgKey = std::tuple<int, int, int, double, unsigned, unsigned, unsigned>;

std::map<CfgKey, LaunchCfg> &cfg_table() {
    static std::map<CfgKey, LaunchCfg> t;
    return t;
}

CfgKey cfg_key(int precision, const mgad_scheme &sc, unsigned nx, unsigned ny, unsigned nz) {
    return CfgKey{precision, sc.weno, sc.p, sc.eps, nx, ny, nz};
}

}  // namespace

LaunchCfg launch_cfg(int precision, const mgad_scheme &sc, unsigned nx, unsigned ny, unsigned nz) {
    std::lock_guard<std::recursive_mutex> lock(host_mutex());
    static const int env_tile = env_int("MGAD_TILE"), env_kc = env_int("MGAD_KCHUNK");
    LaunchCfg c{0, 0};
    auto it = cfg_table().find(cfg_key(precision, sc, nx, ny, nz));
    if (it != cfg_table().end()) c = it->second;
    if (env_tile) c.tile = env_tile;
    if (env_kc) c.kchunk = env_kc;
    return c;
}

void set_launch_cfg(int precision, const mgad_scheme &sc, unsigned nx, unsigned ny, unsigned nz, LaunchCfg c) {
    std::lock_guard<std::recursive_mutex> lock(host_mutex());
    cfg_table()[cfg_key(precision, sc, nx, ny, nz)] = c;
}

// Times stage-1 launches (u -> w1) for each candidate tile width and z band
// rule (shrinking bands, bands of 4..128 planes and of nz) with HIP events on
// the caller's stream; keeps the fastest.
template <class T>
int autotune_t(hipStream_t stream, const mgad_scheme &sc, unsigned nx, unsigned ny, unsigned nz, double dx,
               double dy, double dz, const T *u, T *w1, void *ctrl, LaunchCfg *best_out) {
    int st = burgers_init_t<T>(stream, nx, ny, nz, u, ctrl, 0.0, 1e300, 0.4);
    if (st) return st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if ((st = check_hip(hipEventCreate(&e0))) || (st = check_hip(hipEventCreate(&e1)))) return st;
    std::vector<int> tiles;
    if (nx % 32 == 0 || (nx % 16 != 0 && nx >= 32) || ny % 16 != 0) tiles.push_back(32);
    tiles.push_back(16);
    std::vector<int> chunks{0, MGAD_BANDS_SHRINK};  // 0: default_kchunk
    for (int kc : {4, 8, 16, 32, 64, 128})
        if (kc <= static_cast<int>(nz)) chunks.push_back(kc);
    chunks.push_back(static_cast<int>(nz));
    LaunchCfg best{0, 0};
    float best_ms = 1e30f;
    const int reps = 3;
    // the default launch (one round of 32x32 workgroups where it fits) first
    std::vector<LaunchCfg> cands{LaunchCfg{0, 0}};
    for (int t : tiles)
        for (int kc : chunks) cands.push_back(LaunchCfg{t, kc});
    for (const LaunchCfg &cand : cands) {
        const int t = cand.tile, kc = cand.kchunk;
        st = burgers_stage_launch_t<T>(stream, sc, nx, ny, nz, dx, dy, dz, 1, u, u, w1, ctrl, kc, t);
        if (!st) st = check_hip(hipEventRecord(e0, stream));
        for (int r = 0; r < reps && !st; r++)
            st = burgers_stage_launch_t<T>(stream, sc, nx, ny, nz, dx, dy, dz, 1, u, u, w1, ctrl, kc, t);
        if (!st) st = check_hip(hipEventRecord(e1, stream));
        if (!st) st = check_hip(hipEventSynchronize(e1));
        float ms = 0.f;
        if (!st) st = check_hip(hipEventElapsedTime(&ms, e0, e1));
        if (st) break;
        if (ms < best_ms) {
            best_ms = ms;
            best = cand;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    set_launch_cfg(std::is_same<T, float>::value ? MGAD_F32 : MGAD_F64, sc, nx, ny, nz, best);
    if (best_out) *best_out = best;
    return MGAD_OK;
}

}  // namespace mgad

using namespace mgad;

namespace {

int precision_ok(int precision) { return precision == MGAD_F64 || precision == MGAD_F32; }

template <class T>
int enqueue_step(hipStream_t stream, const mgad_scheme &sc, unsigned nx, unsigned ny, unsigned nz,
                 double dx, double dy, double dz, T *u, T *w1, T *w2, void *ctrl) {
    const LaunchCfg c = launch_cfg(std::is_same<T, float>::value ? MGAD_F32 : MGAD_F64, sc, nx, ny, nz);
    int st = burgers_stage_launch_t<T>(stream, sc, nx, ny, nz, dx, dy, dz, 1, u, u, w1, ctrl, c.kchunk, c.tile);
    if (!st)
        st = burgers_stage_launch_t<T>(stream, sc, nx, ny, nz, dx, dy, dz, 2, w1, u, w2, ctrl, c.kchunk, c.tile);
    if (!st)
        st = burgers_stage_launch_t<T>(stream, sc, nx, ny, nz, dx, dy, dz, 3, w2, u, u, ctrl, c.kchunk, c.tile);
    return st;
}

int enqueue_step_p(int precision, hipStream_t stream, const mgad_scheme &sc, unsigned nx, unsigned ny,
                   unsigned nz, double dx, double dy, double dz, void *u, void *w1, void *w2, void *ctrl) {
    if (precision == MGAD_F32)
        return enqueue_step<float>(stream, sc, nx, ny, nz, dx, dy, dz, static_cast<float *>(u),
                                   static_cast<float *>(w1), static_cast<float *>(w2), ctrl);
    return enqueue_step<double>(stream, sc, nx, ny, nz, dx, dy, dz, static_cast<double *>(u),
                                static_cast<double *>(w1), static_cast<double *>(w2), ctrl);
}

// A captured step is replayed only while its launch shape is current: the
// tile width and z chunk (mgad_burgers_set_launch, autotune) are part of the key.
struct GraphKey {
    hipStream_t stream;
    int precision, weno, p;
    double eps;
    unsigned nx, ny, nz;
    double dx, dy, dz;
    void *u, *w1, *w2, *ctrl;
    int tile, kchunk;
    bool operator<(const GraphKey &o) const {
        return std::tie(stream, precision, weno, p, eps, nx, ny, nz, dx, dy, dz, u, w1, w2, ctrl, tile, kchunk) <
               std::tie(o.stream, o.precision, o.weno, o.p, o.eps, o.nx, o.ny, o.nz, o.dx, o.dy, o.dz,
                        o.u, o.w1, o.w2, o.ctrl, o.tile, o.kchunk);
    }
};

std::map<GraphKey, hipGraphExec_t> &graph_cache() {
    static std::map<GraphKey, hipGraphExec_t> m;
    return m;
}

int run_steps(int precision, hipStream_t stream, const mgad_scheme *scheme, unsigned nx, unsigned ny,
              unsigned nz, double dx, double dy, double dz, void *u, void *w1, void *w2, void *ctrl,
              int nsteps, int use_graph) {
    if (!scheme || !u || !w1 || !w2 || !ctrl || nsteps < 0 || u == w1 || u == w2 || w1 == w2 ||
        !precision_ok(precision)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    if (!use_graph || stream == nullptr) {
        for (int s = 0; s < nsteps; s++) {
            const int st = enqueue_step_p(precision, stream, *scheme, nx, ny, nz, dx, dy, dz, u, w1, w2, ctrl);
            if (st) return st;
        }
        return MGAD_OK;
    }
    std::lock_guard<std::recursive_mutex> lock(host_mutex());
    const LaunchCfg cfg = launch_cfg(precision, *scheme, nx, ny, nz);
    const GraphKey key{stream, precision, scheme->weno, scheme->p, scheme->eps, nx, ny, nz, dx, dy, dz,
                       u,      w1,        w2,           ctrl,       cfg.tile,  cfg.kchunk};
    auto &cache = graph_cache();
    auto it = cache.find(key);
    hipGraphExec_t exec = nullptr;
    if (it == cache.end()) {
        if (cache.size() > 32) {
            clear_graph_cache(cache);
        }
        hipGraph_t g = nullptr;
        int st = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        if (st) return st;
        const int est = enqueue_step_p(precision, stream, *scheme, nx, ny, nz, dx, dy, dz, u, w1, w2, ctrl);
        st = check_hip(hipStreamEndCapture(stream, &g));
        if (est) {
            if (g) (void)hipGraphDestroy(g);
            return est;
        }
        if (st) return st;
        st = check_hip(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        if (st) return st;
        cache[key] = exec;
    } else {
        exec = it->second;
    }
    for (int s = 0; s < nsteps; s++) {
        const int st = check_hip(hipGraphLaunch(exec, stream));
        if (st) return st;
    }
    return MGAD_OK;
}

}  // namespace

extern "C" int mgad_burgers_init(hipStream_t stream, unsigned int nx, unsigned int ny, unsigned int nz,
                                 const double *u, void *ctrl, double t0, double tend, double cfl) {
    return burgers_init_t<double>(stream, nx, ny, nz, u, ctrl, t0, tend, cfl);
}

extern "C" int mgad_burgers_stage(hipStream_t stream, const mgad_scheme *scheme, unsigned int nx,
                                  unsigned int ny, unsigned int nz, double dx, double dy, double dz,
                                  int stage, const double *u_in, const double *u_n, double *u_out,
                                  void *ctrl) {
    if (!scheme) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const LaunchCfg c = launch_cfg(MGAD_F64, *scheme, nx, ny, nz);
    return burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, stage, u_in, u_n,
                                          u_out, ctrl, c.kchunk, c.tile);
}

extern "C" int mgad_burgers_steps(hipStream_t stream, const mgad_scheme *scheme, unsigned int nx,
                                  unsigned int ny, unsigned int nz, double dx, double dy, double dz,
                                  double *u, double *w1, double *w2, void *ctrl, int nsteps,
                                  int use_graph) {
    return run_steps(MGAD_F64, stream, scheme, nx, ny, nz, dx, dy, dz, u, w1, w2, ctrl, nsteps, use_graph);
}

extern "C" int mgad_burgers_init_p(int precision, hipStream_t stream, unsigned int nx, unsigned int ny,
                                   unsigned int nz, const void *u, void *ctrl, double t0, double tend,
                                   double cfl) {
    if (!precision_ok(precision)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    if (precision == MGAD_F32)
        return burgers_init_t<float>(stream, nx, ny, nz, static_cast<const float *>(u), ctrl, t0, tend, cfl);
    return burgers_init_t<double>(stream, nx, ny, nz, static_cast<const double *>(u), ctrl, t0, tend, cfl);
}

extern "C" int mgad_burgers_steps_p(int precision, hipStream_t stream, const mgad_scheme *scheme,
                                    unsigned int nx, unsigned int ny, unsigned int nz, double dx,
                                    double dy, double dz, void *u, void *w1, void *w2, void *ctrl,
                                    int nsteps, int use_graph) {
    return run_steps(precision, stream, scheme, nx, ny, nz, dx, dy, dz, u, w1, w2, ctrl, nsteps, use_graph);
}

extern "C" int mgad_burgers_read_ctrl(hipStream_t stream, const void *ctrl, mgad_burgers_state *out) {
    if (!ctrl || !out) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    Ctrl c;
    int st = check_hip(hipMemcpyAsync(&c, ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
    if (!st) st = check_hip(hipStreamSynchronize(stream));
    if (st) return st;
    out->t = c.t;
    out->dt = c.dt;
    out->tend = c.tend;
    out->cfl = c.cfl;
    double a[3];
    static_assert(sizeof(a) == sizeof(c.amax), "double size");
    __builtin_memcpy(a, c.amax, sizeof(a));
    out->amax = a[0];
    out->steps = c.steps;
    // a stage that found its input's max|u| NaN or infinite left it in its
    // slot and stopped the run (dt = 0)
    if (!(a[0] <= 1.7976931348623157e308) || !(a[1] <= 1.7976931348623157e308) ||
        !(a[2] <= 1.7976931348623157e308)) {
        set_error(MGAD_ERR_NONFINITE);
        return MGAD_ERR_NONFINITE;
    }
    return MGAD_OK;
}

extern "C" int mgad_burgers_autotune(hipStream_t stream, int precision, const mgad_scheme *scheme,
                                     unsigned int nx, unsigned int ny, unsigned int nz, double dx, double dy,
                                     double dz, const void *u, void *w1, void *ctrl, int *tile, int *kchunk) {
    if (!scheme || !u || !w1 || !ctrl || u == w1 || !precision_ok(precision) || !valid_grid(nx, ny, nz)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    LaunchCfg best{0, 0};
    const int st = precision == MGAD_F32
                       ? autotune_t<float>(stream, *scheme, nx, ny, nz, dx, dy, dz, static_cast<const float *>(u),
                                           static_cast<float *>(w1), ctrl, &best)
                       : autotune_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz,
                                            static_cast<const double *>(u), static_cast<double *>(w1), ctrl, &best);
    if (st) return st;
    if (tile) *tile = best.tile;
    if (kchunk) *kchunk = best.kchunk;
    return MGAD_OK;
}

extern "C" int mgad_burgers_set_launch(int precision, const mgad_scheme *scheme, unsigned int nx,
                                       unsigned int ny, unsigned int nz, int tile, int kchunk) {
    if (!scheme || !precision_ok(precision) || (tile != 0 && tile != 16 && tile != 32) ||
        (kchunk < 0 && kchunk != MGAD_BANDS_SHRINK)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    set_launch_cfg(precision, *scheme, nx, ny, nz, LaunchCfg{tile, kchunk});
    return MGAD_OK;
}
