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
// Work decomposition: a workgroup owns a TX x TY column tile and a chunk of
// z planes.  Each thread owns one (x,y) column and marches up the chunk with
// a 7-plane register queue of split fluxes, so every z face is computed once
// and reused by the two nodes that share it.  For x and y faces the current
// plane's split fluxes (tile + 3-node halo) are staged in LDS; each thread
// computes the face on its +1/2 side and wave 0 computes the TY + TX faces on
// the tile's -1/2 edges, so every face of the plane is computed once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <tuple>

#include "mgad_common.hpp"

namespace mgad {

// Layout of the caller's MGAD_CTRL_BYTES device buffer.
struct Ctrl {
    double t;                    // time at the start of the next step
    double dt;                   // dt of the current step, 0 once done
    double tend;
    double cfl;
    unsigned long long amax[3];  // max|u| of: [0] u^n, [1] u^(1), [2] u^(2)
    unsigned long long steps;
};
static_assert(sizeof(Ctrl) <= MGAD_CTRL_BYTES, "ctrl layout");

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Maps the launch's linear block id to a tile so that the blocks one XCD
// runs (ids congruent mod 8) are contiguous tiles sharing halos in L2.
__device__ __forceinline__ unsigned xcd_tile(unsigned bid, unsigned nblocks) {
    if ((nblocks & 7u) != 0u) return bid;
    return (bid & 7u) * (nblocks >> 3) + (bid >> 3);
}

template <int WENO, int P, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_burgers_stage(
    const double *__restrict__ uin, const double *un, double *uout, int nx, int ny, int nz,
    int kchunk, double dx, double dy, double dz, double eps, int stage, Ctrl *ctrl) {
    constexpr int NT = TX * TY;
    constexpr int NW = NT / 64;
    static_assert(NT % 64 == 0, "whole waves");
    static_assert(6 * TX + 6 * TY <= NT, "one halo node per thread at most");
    static_assert(TX + TY <= 64, "edge faces fit in wave 0");
    __shared__ double2 sP[TY + 6][TX + 6];  // (f+, f-) of the current plane, tile + halo
    __shared__ double sFX[TY][TX + 1];      // x faces x0-1/2 .. x0+TX-1/2
    __shared__ double sFY[TY + 1][TX];      // y faces y0-1/2 .. y0+TY-1/2
    __shared__ unsigned long long sRed[NW];

    const int tid = threadIdx.x;
    const unsigned nblocks = gridDim.x;
    const bool lead = (blockIdx.x == 0 && tid == 0);

    // ---- step control (every block derives the same dt) ----
    double dt, alpha;
    if (stage == 1) {
        const double t = ctrl->t, tend = ctrl->tend;
        if (!(t < tend)) {
            if (lead) ctrl->dt = 0.0;
            return;
        }
        const double a = bits_double(ctrl->amax[0]);
        dt = ctrl->cfl * dx / a;
        if (t + dt > tend) dt = tend - t;
        alpha = a;
        if (lead) {
            ctrl->dt = dt;
            ctrl->amax[2] = 0ull;
        }
    } else {
        dt = ctrl->dt;
        if (dt == 0.0) return;
        alpha = bits_double(ctrl->amax[stage - 1]);
        if (lead) {
            if (stage == 2) {
                ctrl->amax[0] = 0ull;
                ctrl->t = ctrl->t + dt;
                ctrl->steps = ctrl->steps + 1ull;
            } else {
                ctrl->amax[1] = 0ull;
            }
        }
    }

    // ---- tile of this block ----
    const int ntx = nx / TX, nty = ny / TY;
    const unsigned tile = xcd_tile(blockIdx.x, nblocks);
    const int bx = tile % ntx;
    const int by = (tile / ntx) % nty;
    const int bz = tile / (ntx * nty);
    const int x0 = bx * TX, y0 = by * TY;
    const int z0 = bz * kchunk;
    const int z1 = min(z0 + kchunk, nz);
    const int tx = tid % TX, ty = tid / TX;
    const unsigned plane = static_cast<unsigned>(nx) * static_cast<unsigned>(ny);
    const unsigned col = static_cast<unsigned>(x0 + tx) + static_cast<unsigned>(nx) * (y0 + ty);

    // Halo node handled by this thread: 6*TY x-halo nodes, then 6*TX y-halo nodes.
    int hl_r = 0, hl_c = 0;
    unsigned hcol = 0;
    const bool has_halo = tid < 6 * TY + 6 * TX;
    if (tid < 6 * TY) {
        const int r = tid / 6, c = tid % 6;
        hl_r = r + 3;
        hl_c = c < 3 ? c : TX + c;
        hcol = static_cast<unsigned>(wrapi(x0 + hl_c - 3, nx)) + static_cast<unsigned>(nx) * (y0 + r);
    } else if (has_halo) {
        const int h = tid - 6 * TY;
        const int c = h % TX, q = h / TX;
        hl_r = q < 3 ? q : TY + q;
        hl_c = c + 3;
        hcol = static_cast<unsigned>(x0 + c) + static_cast<unsigned>(nx) * wrapi(y0 + hl_r - 3, ny);
    }

    // ---- z register queue: index q holds plane k-3+q ----
    double qp[7], qm[7], qu[4];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const double v = uin[col + plane * static_cast<unsigned>(wrapi(z0 - 3 + q, nz))];
        lf_split(v, alpha, qp[q], qm[q]);
        if (q >= 3) qu[q - 3] = v;
    }
    double fz_prev = weno_face<WENO, P>(qp[0], qp[1], qp[2], qp[3], qp[4],
                                        qm[1], qm[2], qm[3], qm[4], qm[5], eps);

    double own_next = uin[col + plane * static_cast<unsigned>(wrapi(z0 + 3, nz))];
    double halo_next = has_halo ? uin[hcol + plane * static_cast<unsigned>(z0)] : 0.0;
    double un_next = (stage == 1) ? 0.0 : un[col + plane * static_cast<unsigned>(z0)];
    unsigned long long vmax = 0ull;

    for (int k = z0; k < z1; k++) {
        const double own = own_next, halo = halo_next, unv = un_next;
        if (k + 1 < z1) {
            own_next = uin[col + plane * static_cast<unsigned>(wrapi(k + 4, nz))];
            if (has_halo) halo_next = uin[hcol + plane * static_cast<unsigned>(k + 1)];
            if (stage != 1) un_next = un[col + plane * static_cast<unsigned>(k + 1)];
        }
        lf_split(own, alpha, qp[6], qm[6]);
        qu[3] = own;
        sP[ty + 3][tx + 3] = make_double2(qp[3], qm[3]);
        if (has_halo) {
            double hp, hm;
            lf_split(halo, alpha, hp, hm);
            sP[hl_r][hl_c] = make_double2(hp, hm);
        }
        __syncthreads();

        // x face at x+1/2 and y face at y+1/2
        {
            const double2 a0 = sP[ty + 3][tx + 1], a1 = sP[ty + 3][tx + 2], a2 = sP[ty + 3][tx + 3],
                          a3 = sP[ty + 3][tx + 4], a4 = sP[ty + 3][tx + 5], a5 = sP[ty + 3][tx + 6];
            sFX[ty][tx + 1] = weno_face<WENO, P>(a0.x, a1.x, a2.x, a3.x, a4.x,
                                                 a1.y, a2.y, a3.y, a4.y, a5.y, eps);
        }
        {
            const double2 a0 = sP[ty + 1][tx + 3], a1 = sP[ty + 2][tx + 3], a2 = sP[ty + 3][tx + 3],
                          a3 = sP[ty + 4][tx + 3], a4 = sP[ty + 5][tx + 3], a5 = sP[ty + 6][tx + 3];
            sFY[ty + 1][tx] = weno_face<WENO, P>(a0.x, a1.x, a2.x, a3.x, a4.x,
                                                 a1.y, a2.y, a3.y, a4.y, a5.y, eps);
        }
        // faces on the tile's low edges (wave 0 only when TX + TY <= 64)
        if (tid < TY + TX) {
            double2 a0, a1, a2, a3, a4, a5;
            if (tid < TY) {
                const int r = tid + 3;
                a0 = sP[r][0]; a1 = sP[r][1]; a2 = sP[r][2]; a3 = sP[r][3]; a4 = sP[r][4]; a5 = sP[r][5];
            } else {
                const int c = tid - TY + 3;
                a0 = sP[0][c]; a1 = sP[1][c]; a2 = sP[2][c]; a3 = sP[3][c]; a4 = sP[4][c]; a5 = sP[5][c];
            }
            const double f = weno_face<WENO, P>(a0.x, a1.x, a2.x, a3.x, a4.x,
                                                a1.y, a2.y, a3.y, a4.y, a5.y, eps);
            if (tid < TY)
                sFX[tid][0] = f;
            else
                sFY[0][tid - TY] = f;
        }
        // z face at k+1/2 from the register queue
        const double fz = weno_face<WENO, P>(qp[1], qp[2], qp[3], qp[4], qp[5],
                                             qm[2], qm[3], qm[4], qm[5], qm[6], eps);
        __syncthreads();

        double Lu = -(sFX[ty][tx + 1] - sFX[ty][tx]) / dx;
        Lu -= (sFY[ty + 1][tx] - sFY[ty][tx]) / dy;
        Lu -= (fz - fz_prev) / dz;
        const double uo = (stage == 1) ? qu[0] : unv;
        const double unew = rk_stage(stage, dt, qu[0], uo, Lu);
        uout[col + plane * static_cast<unsigned>(k)] = unew;
        const unsigned long long b = abs_bits(unew);
        vmax = b > vmax ? b : vmax;

        fz_prev = fz;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            qp[q] = qp[q + 1];
            qm[q] = qm[q + 1];
        }
        qu[0] = qu[1];
        qu[1] = qu[2];
        qu[2] = qu[3];
    }

    // ---- block max -> one atomic per block ----
    vmax = wave_max_u64(vmax);
    if ((tid & 63) == 0) sRed[tid >> 6] = vmax;
    __syncthreads();
    if (tid == 0) {
        unsigned long long m = sRed[0];
#pragma unroll
        for (int w = 1; w < NW; w++) m = sRed[w] > m ? sRed[w] : m;
        atomicMax(&ctrl->amax[stage % 3], m);
    }
}

// max|u| over n values into *slot (atomicMax of the bit pattern).
__global__ __launch_bounds__(256) void k_maxabs(const double *__restrict__ u, size_t n,
                                                unsigned long long *slot) {
    __shared__ unsigned long long sRed[4];
    unsigned long long m = 0ull;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
        const unsigned long long b = abs_bits(u[i]);
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
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

using StageKernel = void (*)(const double *, const double *, double *, int, int, int, int, double,
                             double, double, double, int, Ctrl *);

struct Shape {
    int tx, ty;
    StageKernel fn;
};

template <int WENO, int P>
bool pick(int nx, int ny, Shape &s) {
    if (nx % 32 == 0 && ny % 16 == 0) {
        s = {32, 16, k_burgers_stage<WENO, P, 32, 16>};
        return true;
    }
    if (nx % 16 == 0 && ny % 16 == 0) {
        s = {16, 16, k_burgers_stage<WENO, P, 16, 16>};
        return true;
    }
    return false;
}

bool pick_kernel(const mgad_scheme &sc, int nx, int ny, Shape &s) {
    if (sc.weno == MGAD_WENO_Z && sc.p == 1) return pick<kWenoZ, 1>(nx, ny, s);
    if (sc.weno == MGAD_WENO_Z && sc.p == 2) return pick<kWenoZ, 2>(nx, ny, s);
    if (sc.weno == MGAD_WENO_JS) return pick<kWenoJS, 2>(nx, ny, s);
    return false;
}

int default_kchunk(int nx, int ny, int nz, int tx, int ty) {
    // Enough blocks to give every CU several tiles; chunks of >= 8 planes so
    // the 6-plane queue warm-up stays small.
    const long tiles = static_cast<long>(nx / tx) * (ny / ty);
    int kc = nz;
    while (kc > 8 && tiles * ((nz + kc - 1) / kc) < 2048) kc = (kc + 1) / 2;
    return kc;
}

// The z queue wraps plane indices by at most one period (z0-3 and k+4), so
// nz >= 4; x and y tiles divide the extents, and indices stay below 2^31.
bool valid_grid(unsigned nx, unsigned ny, unsigned nz) {
    if (nx < 4 || ny < 4 || nz < 4) return false;
    const unsigned long long n = 1ull * nx * ny * nz;
    return n < (1ull << 31);
}

}  // namespace

int burgers_stage_launch(hipStream_t stream, const mgad_scheme &sc, unsigned nx, unsigned ny,
                         unsigned nz, double dx, double dy, double dz, int stage,
                         const double *u_in, const double *u_n, double *u_out, void *ctrl,
                         int kchunk) {
    if (!valid_grid(nx, ny, nz) || stage < 1 || stage > 3 || !u_in || !u_n || !u_out || !ctrl ||
        u_out == u_in) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    Shape s;
    if (!pick_kernel(sc, static_cast<int>(nx), static_cast<int>(ny), s)) {
        set_error(MGAD_ERR_UNSUPPORTED);
        return MGAD_ERR_UNSUPPORTED;
    }
    if (kchunk <= 0) kchunk = default_kchunk(nx, ny, nz, s.tx, s.ty);
    const unsigned nblocks = (nx / s.tx) * (ny / s.ty) * ((nz + kchunk - 1) / kchunk);
    hipLaunchKernelGGL(s.fn, dim3(nblocks), dim3(s.tx * s.ty), 0, stream, u_in, u_n, u_out,
                       static_cast<int>(nx), static_cast<int>(ny), static_cast<int>(nz), kchunk,
                       dx, dy, dz, sc.eps, stage, static_cast<Ctrl *>(ctrl));
    return check_launch();
}

int burgers_init(hipStream_t stream, unsigned nx, unsigned ny, unsigned nz, const double *u,
                 void *ctrl, double t0, double tend, double cfl) {
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
    hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(256), 0, stream, u, n, &c->amax[0]);
    return check_launch();
}

}  // namespace mgad

using namespace mgad;

extern "C" int mgad_burgers_init(hipStream_t stream, unsigned int nx, unsigned int ny,
                                 unsigned int nz, const double *u, void *ctrl, double t0,
                                 double tend, double cfl) {
    return burgers_init(stream, nx, ny, nz, u, ctrl, t0, tend, cfl);
}

extern "C" int mgad_burgers_stage(hipStream_t stream, const mgad_scheme *scheme, unsigned int nx,
                                  unsigned int ny, unsigned int nz, double dx, double dy,
                                  double dz, int stage, const double *u_in, const double *u_n,
                                  double *u_out, void *ctrl) {
    if (!scheme) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    return burgers_stage_launch(stream, *scheme, nx, ny, nz, dx, dy, dz, stage, u_in, u_n, u_out,
                                ctrl, 0);
}

namespace {

int enqueue_step(hipStream_t stream, const mgad_scheme &sc, unsigned nx, unsigned ny, unsigned nz,
                 double dx, double dy, double dz, double *u, double *w1, double *w2, void *ctrl) {
    int st = burgers_stage_launch(stream, sc, nx, ny, nz, dx, dy, dz, 1, u, u, w1, ctrl, 0);
    if (!st) st = burgers_stage_launch(stream, sc, nx, ny, nz, dx, dy, dz, 2, w1, u, w2, ctrl, 0);
    if (!st) st = burgers_stage_launch(stream, sc, nx, ny, nz, dx, dy, dz, 3, w2, u, u, ctrl, 0);
    return st;
}

struct GraphKey {
    hipStream_t stream;
    int weno, p;
    double eps;
    unsigned nx, ny, nz;
    double dx, dy, dz;
    double *u, *w1, *w2;
    void *ctrl;
    bool operator<(const GraphKey &o) const {
        return std::tie(stream, weno, p, eps, nx, ny, nz, dx, dy, dz, u, w1, w2, ctrl) <
               std::tie(o.stream, o.weno, o.p, o.eps, o.nx, o.ny, o.nz, o.dx, o.dy, o.dz, o.u,
                        o.w1, o.w2, o.ctrl);
    }
};

std::map<GraphKey, hipGraphExec_t> &graph_cache() {
    static std::map<GraphKey, hipGraphExec_t> m;
    return m;
}

}  // namespace

extern "C" int mgad_burgers_steps(hipStream_t stream, const mgad_scheme *scheme, unsigned int nx,
                                  unsigned int ny, unsigned int nz, double dx, double dy,
                                  double dz, double *u, double *w1, double *w2, void *ctrl,
                                  int nsteps, int use_graph) {
    if (!scheme || !u || !w1 || !w2 || !ctrl || nsteps < 0 || u == w1 || u == w2 || w1 == w2) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    if (!use_graph || stream == nullptr) {
        for (int s = 0; s < nsteps; s++) {
            const int st = enqueue_step(stream, *scheme, nx, ny, nz, dx, dy, dz, u, w1, w2, ctrl);
            if (st) return st;
        }
        return MGAD_OK;
    }
    const GraphKey key{stream, scheme->weno, scheme->p, scheme->eps, nx, ny, nz, dx, dy, dz,
                       u,      w1,          w2,        ctrl};
    auto &cache = graph_cache();
    auto it = cache.find(key);
    hipGraphExec_t exec = nullptr;
    if (it == cache.end()) {
        if (cache.size() > 32) {
            for (auto &kv : cache) (void)hipGraphExecDestroy(kv.second);
            cache.clear();
        }
        hipGraph_t g = nullptr;
        int st = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        if (st) return st;
        const int est = enqueue_step(stream, *scheme, nx, ny, nz, dx, dy, dz, u, w1, w2, ctrl);
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
    double a;
    const unsigned long long b = c.amax[0];
    static_assert(sizeof(a) == sizeof(b), "double size");
    __builtin_memcpy(&a, &b, sizeof(a));
    out->amax = a;
    out->steps = c.steps;
    if (!(a <= 1.7976931348623157e308)) {
        set_error(MGAD_ERR_NONFINITE);
        return MGAD_ERR_NONFINITE;
    }
    return MGAD_OK;
}
