// slab.hip -- the periodic WENO5 Burgers step on z slabs with ghost planes:
// the data layout and halo exchange of a z-slab decomposition
// (MultiGPU/Burgers3d_Baseline/main.c:64-66, RADIUS 3 BurgersMPI.h:50), with
// the fused stage kernel of burgers_periodic.hpp on each slab.
//
// A slab set is nslab slabs of nzs planes of an nx x ny x (nslab*nzs)
// periodic grid, stored one after another; slab s keeps 3 ghost planes below
// and 3 above its own planes (nzs + 6 planes, own plane k at storage plane
// k + 3), so one GPU's share is one slab and its ghosts are the 3 boundary
// planes of each neighbour.  The x-fastest layout makes every boundary region
// 3 contiguous planes: a neighbour exchange is a plain copy of
// 3*nx*ny values per side, with no packing kernel.  Here all slabs live on
// one device, and each stage writes the boundary planes of its output into
// the neighbours' ghost planes as it stores them (the periodic wrap
// included), so the next stage finds its input's ghosts filled;
// mgad_slab_exchange does the same copy on its own (the initial field).
// Across devices the same stores go to the neighbour's memory (peer stores)
// or become the send/recv of each stage.  The exchange is of u after every
// stage (the reference exchanges the RHS).
// All slabs of a stage run in one launch and share one control block, so
// max|u| and dt are global over the slabs.
#include <hip/hip_runtime.h>

#include <map>
#include <tuple>

#include "burgers_periodic.hpp"

namespace mgad {

template <class T, int WENO, int P, int TX, int TY>
__global__ __launch_bounds__(TX * TY, 4) void k_burgers_slab_stage(const T *__restrict__ uin, const T *un,
                                                                T *uout, GridArgs g, unsigned nslab,
                                                                double eps, int stage, Ctrl *ctrl) {
    __shared__ TileSmem<TX, TY> sm;
    double dt, alpha;
    const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
    if (!stage_control(ctrl, stage, g.dx, lead, dt, alpha)) return;
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x), per = gridDim.x / nslab;
    const unsigned slab = t / per, tile = t % per;
    const size_t stride = static_cast<size_t>(g.nz + 6) * static_cast<size_t>(g.nx) * g.ny;
    const size_t off = slab * stride;
    // the neighbours whose ghost planes this slab's boundary planes fill
    T *lo = uout + ((slab + nslab - 1) % nslab) * stride, *hi = uout + ((slab + 1) % nslab) * stride;
    if (fast_division_ok<WENO>(alpha, eps))
        stage_tile<T, WENO, P, TX, TY, true, true>(uin + off, un + off, uout + off, g, stage, tile, dt, alpha,
                                                   eps, sm, &ctrl->amax[stage % 3], lo, hi);
    else
        stage_tile<T, WENO, P, TX, TY, false, true>(uin + off, un + off, uout + off, g, stage, tile, dt, alpha,
                                                    eps, sm, &ctrl->amax[stage % 3], lo, hi);
}

// Ghost planes of every slab from its neighbours: storage planes 0..2 of
// slab s <- storage planes nzs..nzs+2 of slab s-1, planes nzs+3..nzs+5 <-
// planes 3..5 of slab s+1 (slab indices periodic).  Grid: (x, 6 ghost
// planes, slabs).
__global__ __launch_bounds__(256) void k_slab_exchange(double *u, unsigned plane, unsigned nzs, unsigned nslab) {
    const unsigned g = blockIdx.y, s = blockIdx.z;
    const size_t stride = static_cast<size_t>(nzs + 6) * plane;
    const unsigned src_slab = g < 3 ? (s + nslab - 1) % nslab : (s + 1) % nslab;
    const unsigned src_plane = g < 3 ? nzs + g : g;  // g >= 3: planes 3..5
    const unsigned dst_plane = g < 3 ? g : nzs + g;
    const double *src = u + src_slab * stride + static_cast<size_t>(src_plane) * plane;
    double *dst = u + s * stride + static_cast<size_t>(dst_plane) * plane;
    const unsigned i = blockIdx.x * 256 + threadIdx.x;
    if (i < plane) dst[i] = src[i];
}

__global__ void k_slab_ctrl_init(Ctrl *c, double t0, double tend, double cfl) {
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

// max|u| over the own planes of every slab into *slot.
__global__ __launch_bounds__(256) void k_slab_maxabs(const double *__restrict__ u, unsigned plane, unsigned nzs,
                                                     unsigned nslab, unsigned long long *slot) {
    __shared__ unsigned long long red[4];
    const size_t own = static_cast<size_t>(nzs) * plane, n = own * nslab;
    unsigned long long m = 0ull;
    for (size_t q = blockIdx.x * 256ull + threadIdx.x; q < n; q += 256ull * gridDim.x) {
        const size_t s = q / own, r = q % own;
        const unsigned long long b = abs_bits(u[s * (own + 6ull * plane) + 3ull * plane + r]);
        m = b > m ? b : m;
    }
    m = wave_max_u64(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = red[0];
        for (int w = 1; w < 4; w++) m = red[w] > m ? red[w] : m;
        atomicMax(slot, m);
    }
}

namespace {

using SlabKernel = void (*)(const double *, const double *, double *, GridArgs, unsigned, double, int, Ctrl *);

// tile_pref kSlabSquare: 32x32 tiles (one 16-wave workgroup per CU).
constexpr int kSlabSquare = 1024;

template <int WENO, int P>
bool pick_slab(unsigned nx, unsigned ny, int tile_pref, SlabKernel &fn, int &tx, int &ty) {
    if (tile_pref == kSlabSquare && nx % 32 == 0 && ny % 32 == 0) {
        fn = k_burgers_slab_stage<double, WENO, P, 32, 32>;
        tx = 32;
        ty = 32;
        return true;
    }
    if (tile_pref != 16 && nx % 32 == 0 && ny % 16 == 0) {
        fn = k_burgers_slab_stage<double, WENO, P, 32, 16>;
        tx = 32;
        ty = 16;
        return true;
    }
    if (nx % 16 == 0 && ny % 16 == 0) {
        fn = k_burgers_slab_stage<double, WENO, P, 16, 16>;
        tx = 16;
        ty = 16;
        return true;
    }
    return false;
}

bool pick_slab_kernel(const mgad_scheme &sc, unsigned nx, unsigned ny, int tile_pref, SlabKernel &fn, int &tx,
                      int &ty) {
    if (sc.weno == MGAD_WENO_Z && sc.p == 1) return pick_slab<kWenoZ, 1>(nx, ny, tile_pref, fn, tx, ty);
    if (sc.weno == MGAD_WENO_Z && sc.p == 2) return pick_slab<kWenoZ, 2>(nx, ny, tile_pref, fn, tx, ty);
    if (sc.weno == MGAD_WENO_JS) return pick_slab<kWenoJS, 2>(nx, ny, tile_pref, fn, tx, ty);
    return false;
}

// nzs >= 3 so a ghost region is one neighbour's boundary planes; storage of
// the whole set below 2^31 values per slab offset and 32-bit plane indexing.
bool slab_ok(unsigned nx, unsigned ny, unsigned nzs, unsigned nslab) {
    if (nx < 4 || ny < 4 || nzs < 3 || nslab < 1) return false;
    const unsigned long long per = 1ull * nx * ny * (nzs + 6);
    return per < (1ull << 31) && per * nslab < (1ull << 40);
}

int exchange(hipStream_t s, unsigned nx, unsigned ny, unsigned nzs, unsigned nslab, double *u) {
    const unsigned plane = nx * ny;
    hipLaunchKernelGGL(k_slab_exchange, dim3((plane + 255) / 256, 6, nslab), dim3(256), 0, s, u, plane, nzs, nslab);
    return check_launch();
}

int enqueue_slab_step(hipStream_t s, const mgad_scheme &sc, unsigned nx, unsigned ny, unsigned nzs, unsigned nslab,
                      double dx, double dy, double dz, double *u, double *w1, double *w2, Ctrl *ctrl) {
    // the launch table of the unsplit grid of the same size (environment
    // overrides, mgad_burgers_set_launch / autotune entries), chunks capped at
    // one slab
    LaunchCfg cfg = launch_cfg(MGAD_F64, sc, nx, ny, nzs * nslab);
    // as on the unsplit grid, the default launch is one round of 32x32
    // workgroups, one per CU, where the slabs' tiles fill the CUs
    if (cfg.tile == 0 && cfg.kchunk == 0 && nx % 32 == 0 && ny % 32 == 0) {
        const int b = square_bands(static_cast<long>(nx / 32) * (ny / 32) * nslab, nzs);
        if (b > 0) cfg = LaunchCfg{kSlabSquare, static_cast<int>((nzs + b - 1) / b)};
    }
    SlabKernel fn = nullptr;
    int tx = 0, ty = 0;
    if (!pick_slab_kernel(sc, nx, ny, cfg.tile, fn, tx, ty)) {
        set_error(MGAD_ERR_UNSUPPORTED);
        return MGAD_ERR_UNSUPPORTED;
    }
    const int kc = cfg.kchunk > 0 ? cfg.kchunk
                                  : default_kchunk(static_cast<int>(nx), static_cast<int>(ny),
                                                   static_cast<int>(nzs * nslab), tx, ty);
    const int kchunk = kc < static_cast<int>(nzs) ? kc : static_cast<int>(nzs);
    const GridArgs g = make_grid(nx, ny, nzs, dx, dy, dz, kchunk);
    const unsigned per = (nx / tx) * (ny / ty) * ((nzs + kchunk - 1) / kchunk);
    const dim3 grid(per * nslab), block(tx * ty);
    const double *in[3] = {u, w1, w2};
    double *out[3] = {w1, w2, u};
    for (int stage = 1; stage <= 3; stage++) {
        hipLaunchKernelGGL(fn, grid, block, 0, s, in[stage - 1], u, out[stage - 1], g, nslab, sc.eps, stage, ctrl);
        const int st = check_launch();
        if (st) return st;
    }
    return MGAD_OK;
}

}  // namespace
}  // namespace mgad

using namespace mgad;

extern "C" int mgad_slab_exchange(hipStream_t stream, unsigned int nx, unsigned int ny, unsigned int nzs,
                                  unsigned int nslab, double *u) {
    if (!u || !slab_ok(nx, ny, nzs, nslab)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    return exchange(stream, nx, ny, nzs, nslab, u);
}

extern "C" int mgad_burgers_slab_init(hipStream_t stream, unsigned int nx, unsigned int ny, unsigned int nzs,
                                      unsigned int nslab, double *u, void *ctrl, double t0, double tend,
                                      double cfl) {
    if (!u || !ctrl || !slab_ok(nx, ny, nzs, nslab)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    int st = exchange(stream, nx, ny, nzs, nslab, u);
    if (st) return st;
    Ctrl *c = static_cast<Ctrl *>(ctrl);
    hipLaunchKernelGGL(k_slab_ctrl_init, dim3(1), dim3(1), 0, stream, c, t0, tend, cfl);
    if ((st = check_launch())) return st;
    hipLaunchKernelGGL(k_slab_maxabs, dim3(1024), dim3(256), 0, stream, u, nx * ny, nzs, nslab, &c->amax[0]);
    return check_launch();
}

extern "C" int mgad_burgers_slab_steps(hipStream_t stream, const mgad_scheme *scheme, unsigned int nx,
                                       unsigned int ny, unsigned int nzs, unsigned int nslab, double dx, double dy,
                                       double dz, double *u, double *w1, double *w2, void *ctrl, int nsteps,
                                       int use_graph) {
    if (!scheme || !u || !w1 || !w2 || !ctrl || u == w1 || u == w2 || w1 == w2 || nsteps < 0 ||
        !slab_ok(nx, ny, nzs, nslab) || !(dx > 0) || !(dy > 0) || !(dz > 0)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    Ctrl *c = static_cast<Ctrl *>(ctrl);
    if (!use_graph || stream == nullptr) {
        for (int s = 0; s < nsteps; s++) {
            const int st = enqueue_slab_step(stream, *scheme, nx, ny, nzs, nslab, dx, dy, dz, u, w1, w2, c);
            if (st) return st;
        }
        return MGAD_OK;
    }
    std::lock_guard<std::recursive_mutex> lock(host_mutex());
    // the launch shape taken from the unsplit grid's table is part of the key,
    // so a changed table entry is captured afresh
    const LaunchCfg cfg = launch_cfg(MGAD_F64, *scheme, nx, ny, nzs * nslab);
    using Key = std::tuple<hipStream_t, int, int, double, unsigned, unsigned, unsigned, unsigned, double, double,
                           double, double *, double *, double *, void *, int, int>;
    static std::map<Key, hipGraphExec_t> cache;
    const Key key{stream, scheme->weno, scheme->p, scheme->eps, nx, ny, nzs, nslab, dx, dy, dz, u, w1, w2, ctrl,
                  cfg.tile, cfg.kchunk};
    hipGraphExec_t exec = nullptr;
    auto it = cache.find(key);
    if (it == cache.end()) {
        if (cache.size() > 16) {
            clear_graph_cache(cache);
        }
        hipGraph_t g = nullptr;
        int st = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        if (st) return st;
        const int est = enqueue_slab_step(stream, *scheme, nx, ny, nzs, nslab, dx, dy, dz, u, w1, w2, c);
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
