// burgers_ref.hip -- the reference's own Burgers step, fused: one launch per
// SSP-RK3 stage that gives the same u as the drop-in sequence
//   Call_Adv_x, Call_Adv_y, Call_Adv_z, [Call_Diff_], Call_sspRK
// (SingleGPU/Burgers3d_WENO5/main.cpp:139-149): local Lax-Friedrichs split
// (alpha = |u_i|), WENO5 faces, the optional 13-point viscous term with the
// CopyToConstantMemory coefficients, and the SSP-RK3 combine on the interior
// nodes 3..n-4 of every axis; the three ghost layers are never written.
//
// Per interior node the right-hand side is accumulated in the drop-ins' order
//   Lu = -(F_{i+1/2}-F_{i-1/2})/dx;  Lu -= (G..)/dy;  Lu -= (H..)/dz;
//   Lu += kx*(-u[i-2] + 16u[i-1] - 30u[i] + 16u[i+1] - u[i+2]) + ky*(..) + kz*(..)
// (kernels.cu:197-394, 474-476) and the stage is Compute_RK (kernels.cu:520-527).
// The tile / z-march structure is that of the periodic stage kernel, over the
// interior instead of the whole periodic grid: tiles start at node 3, the last
// ones may be partial (their extra columns load clamped, valid addresses and
// store nothing), and halos read the ghost layers instead of wrapping.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <tuple>

#include "burgers_periodic.hpp"

namespace mgad {

struct RefArgs {
    int nx, ny, nz, kchunk, ntx, nty;
    double dx, dy, dz, rdx, rdy, rdz;
    // max|u| slots: [0] u^n, [1] u^(1), [2] u^(2) (interior values written by
    // the stages), [3] all of u at the call's start (covers the ghost layers,
    // which no stage writes)
    unsigned long long *amax;
};

template <int TX, int TY>
struct RefSmem {
    double2 P[TY + 6][TX + 6];  // (f+, f-) of the current plane, tile + 3-node halo
    double U[TY + 6][TX + 6];   // u of the same nodes (viscous term)
    double FX[TY][TX + 1];
    double FY[TY + 1][TX];
    unsigned long long red[TX * TY / 64];
};

// One tile x z-chunk of one stage; returns the largest |u_out| bit pattern
// this thread stored.  FD: fast_division_ok<WENO>(max|u|, eps) holds for the
// stage's input (the weight divisions then take the shortened sequence).
template <int WENO, int P, bool VISC, int TX, int TY, bool FD>
__device__ __forceinline__ unsigned long long ref_stage_tile(const double *__restrict__ uin, const double *un,
                                                             double *uout, const RefArgs &g, int stage, double dt,
                                                             double eps, const DiffCoef &kc, RefSmem<TX, TY> &sm) {
    constexpr int NT = TX * TY;
    constexpr int NW = NT / 64;
    static_assert(NT % 64 == 0 && 6 * TX + 6 * TY <= NT && TX + TY <= 64, "tile shape");
    unsigned long long vmax = 0ull;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = tile % g.ntx, by = (tile / g.ntx) % g.nty, bz = tile / (g.ntx * g.nty);
    const int x0 = 3 + bx * TX, y0 = 3 + by * TY;
    const int z0 = 3 + bz * g.kchunk, z1 = min(z0 + g.kchunk, nz - 3);
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int wave = tid >> 6, lane = tid & 63;
    const bool active = x0 + tx < nx - 3 && y0 + ty < ny - 3;
    const unsigned plane = static_cast<unsigned>(nx) * static_cast<unsigned>(ny);
    auto cx = [&](int x) { return static_cast<unsigned>(min(x, nx - 1)); };
    auto cy = [&](int y) { return static_cast<unsigned>(min(y, ny - 1)); };
    const unsigned col = cx(x0 + tx) + static_cast<unsigned>(nx) * cy(y0 + ty);

    // halo node of this thread: 6*TY x-halo nodes, then 6*TX y-halo nodes
    int hl_r = 0, hl_c = 0;
    unsigned hcol = 0;
    const bool has_halo = tid < 6 * TY + 6 * TX;
    if (tid < 6 * TY) {
        const int r = tid / 6, c = tid % 6;
        hl_r = r + 3;
        hl_c = c < 3 ? c : TX + c;
        hcol = cx(x0 + hl_c - 3) + static_cast<unsigned>(nx) * cy(y0 + r);
    } else if (has_halo) {
        const int h = tid - 6 * TY, c = h % TX, q = h / TX;
        hl_r = q < 3 ? q : TY + q;
        hl_c = c + 3;
        hcol = cx(x0 + c) + static_cast<unsigned>(nx) * cy(y0 + hl_r - 3);
    }

    // z queues: qp/qm[q] split fluxes of plane k-3+q, uq[q] u of plane k-2+q,
    // sp/sn[q] the 13/12 terms centred on plane k-1+q
    double qp[7], qm[7], uq[6], sp[4], sn[4];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const double v = uin[col + plane * static_cast<unsigned>(z0 - 3 + q)];
        lf_split_local(v, qp[q], qm[q]);
        if (q >= 1) uq[q - 1] = v;
    }
    const double sp0 = sterm(qp[0], qp[1], qp[2]);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        sp[q] = sterm(qp[q + 1], qp[q + 2], qp[q + 3]);
        sn[q] = sterm(qm[q + 1], qm[q + 2], qm[q + 3]);
    }
    double fz_prev = weno_face_s<WENO, P, FD>(qp[0], qp[1], qp[2], qp[3], qp[4], qm[1], qm[2], qm[3], qm[4], qm[5],
                                              sp0, sp[0], sp[1], sn[0], sn[1], sn[2], eps);
    double own_next = uin[col + plane * static_cast<unsigned>(z0 + 3)];
    double halo_next = has_halo ? uin[hcol + plane * static_cast<unsigned>(z0)] : 0.0;
    double un_next = stage == 1 ? 0.0 : un[col + plane * static_cast<unsigned>(z0)];

    for (int k = z0; k < z1; k++) {
        const double own = own_next, halo = halo_next, unv = un_next;
        if (k + 1 < z1) {
            own_next = uin[col + plane * static_cast<unsigned>(k + 4)];
            if (has_halo) halo_next = uin[hcol + plane * static_cast<unsigned>(k + 1)];
            if (stage != 1) un_next = un[col + plane * static_cast<unsigned>(k + 1)];
        }
        lf_split_local(own, qp[6], qm[6]);
        uq[5] = own;
        sp[3] = sterm(qp[4], qp[5], qp[6]);
        sn[3] = sterm(qm[4], qm[5], qm[6]);
        sm.P[ty + 3][tx + 3] = make_double2(qp[3], qm[3]);
        if (VISC) sm.U[ty + 3][tx + 3] = uq[2];
        if (has_halo) {
            double hp, hm;
            lf_split_local(halo, hp, hm);
            sm.P[hl_r][hl_c] = make_double2(hp, hm);
            if (VISC) sm.U[hl_r][hl_c] = halo;
        }
        __syncthreads();

        // x and y second differences of the viscous term: U is read only
        // between the two barriers (the next plane overwrites it after the second)
        double X = 0.0, Y = 0.0;
        if (VISC) {
            const double c = uq[2];
            X = d4(sm.U[ty + 3][tx + 1], sm.U[ty + 3][tx + 2], c, sm.U[ty + 3][tx + 4], sm.U[ty + 3][tx + 5]);
            Y = d4(sm.U[ty + 1][tx + 3], sm.U[ty + 2][tx + 3], c, sm.U[ty + 4][tx + 3], sm.U[ty + 5][tx + 3]);
        }
        const double fz = weno_face_s<WENO, P, FD>(qp[1], qp[2], qp[3], qp[4], qp[5], qm[2], qm[3], qm[4], qm[5],
                                                   qm[6], sp[0], sp[1], sp[2], sn[1], sn[2], sn[3], eps);
        {
            const double2 a0 = sm.P[ty + 3][tx + 1], a1 = sm.P[ty + 3][tx + 2], a2 = sm.P[ty + 3][tx + 3],
                          a3 = sm.P[ty + 3][tx + 4], a4 = sm.P[ty + 3][tx + 5], a5 = sm.P[ty + 3][tx + 6];
            sm.FX[ty][tx + 1] =
                weno_face2<WENO, P, FD>(a0.x, a1.x, a2.x, a3.x, a4.x, a1.y, a2.y, a3.y, a4.y, a5.y, eps);
        }
        {
            const double2 a0 = sm.P[ty + 1][tx + 3], a1 = sm.P[ty + 2][tx + 3], a2 = sm.P[ty + 3][tx + 3],
                          a3 = sm.P[ty + 4][tx + 3], a4 = sm.P[ty + 5][tx + 3], a5 = sm.P[ty + 6][tx + 3];
            sm.FY[ty + 1][tx] =
                weno_face2<WENO, P, FD>(a0.x, a1.x, a2.x, a3.x, a4.x, a1.y, a2.y, a3.y, a4.y, a5.y, eps);
        }
        if (wave == k % NW && lane < TY + TX) {
            double2 a0, a1, a2, a3, a4, a5;
            if (lane < TY) {
                const int r = lane + 3;
                a0 = sm.P[r][0]; a1 = sm.P[r][1]; a2 = sm.P[r][2];
                a3 = sm.P[r][3]; a4 = sm.P[r][4]; a5 = sm.P[r][5];
            } else {
                const int c = lane - TY + 3;
                a0 = sm.P[0][c]; a1 = sm.P[1][c]; a2 = sm.P[2][c];
                a3 = sm.P[3][c]; a4 = sm.P[4][c]; a5 = sm.P[5][c];
            }
            const double f =
                weno_face2<WENO, P, FD>(a0.x, a1.x, a2.x, a3.x, a4.x, a1.y, a2.y, a3.y, a4.y, a5.y, eps);
            if (lane < TY)
                sm.FX[lane][0] = f;
            else
                sm.FY[0][lane - TY] = f;
        }
        __syncthreads();

        double Lu = -div_by(sm.FX[ty][tx + 1] - sm.FX[ty][tx], g.dx, g.rdx);
        Lu -= div_by(sm.FY[ty + 1][tx] - sm.FY[ty][tx], g.dy, g.rdy);
        Lu -= div_by(fz - fz_prev, g.dz, g.rdz);
        if (VISC) {
            const double Z = d4(uq[0], uq[1], uq[2], uq[3], uq[4]);
            Lu += (kc.kx * X + kc.ky * Y) + kc.kz * Z;
        }
        const double uo = stage == 1 ? uq[2] : unv;
        double unew;
        if (stage == 1)
            unew = uo + dt * Lu;
        else if (stage == 2)
            unew = 0.75 * uo + 0.25 * (uq[2] + dt * Lu);
        else
            unew = div_by(p2fma(2.0, uq[2] + dt * Lu, uo), 3.0, 1.0 / 3.0);  // (uo + 2(u + dt Lu)) / 3
        if (active) {
            uout[col + plane * static_cast<unsigned>(k)] = unew;
            const unsigned long long bb = abs_bits(unew);
            vmax = bb > vmax ? bb : vmax;
        }

        fz_prev = fz;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            qp[q] = qp[q + 1];
            qm[q] = qm[q + 1];
        }
#pragma unroll
        for (int q = 0; q < 5; q++) uq[q] = uq[q + 1];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            sp[q] = sp[q + 1];
            sn[q] = sn[q + 1];
        }
    }
    return vmax;
}

// One stage over all tiles: the stage's max|u| bound (its input's slot and
// the ghost floor) picks the division sequence for every block alike; each
// block folds its stored |u_out| into the output slot, and the lead clears
// the slot the next-but-one stage fills.
template <int WENO, int P, bool VISC, int TX, int TY>
__global__ __launch_bounds__(TX * TY, 4) void k_burgers_ref_stage(const double *__restrict__ uin,
                                                                  const double *un, double *uout, RefArgs g,
                                                                  int stage, double dt, double eps,
                                                                  DiffCoef kc) {
    constexpr int NW = TX * TY / 64;
    __shared__ RefSmem<TX, TY> sm;
    const unsigned long long a_in = g.amax[stage - 1], a_gh = g.amax[3];
    const double amax = bits_double(a_in > a_gh ? a_in : a_gh);
    if (blockIdx.x == 0 && threadIdx.x == 0) g.amax[(stage + 1) % 3] = 0ull;
    unsigned long long vmax;
    if (fast_division_ok<WENO>(amax, eps))
        vmax = ref_stage_tile<WENO, P, VISC, TX, TY, true>(uin, un, uout, g, stage, dt, eps, kc, sm);
    else
        vmax = ref_stage_tile<WENO, P, VISC, TX, TY, false>(uin, un, uout, g, stage, dt, eps, kc, sm);
    vmax = wave_max_u64(vmax);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sm.red[wave] = vmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = sm.red[0];
#pragma unroll
        for (int w = 1; w < NW; w++) m = sm.red[w] > m ? sm.red[w] : m;
        atomicMax(&g.amax[stage % 3], m);
    }
}

// Slots for a call: [0] = [3] = max|u| over all nodes, [1] = [2] = 0.
__global__ void k_ref_amax_clear(unsigned long long *amax) {
    amax[0] = amax[1] = amax[2] = amax[3] = 0ull;
}

__global__ __launch_bounds__(256) void k_ref_amax_init(const double *__restrict__ u, size_t n,
                                                       unsigned long long *amax) {
    __shared__ unsigned long long red[4];
    unsigned long long m = 0ull;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
        const unsigned long long b = abs_bits(u[i]);
        m = b > m ? b : m;
    }
    m = wave_max_u64(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = red[0];
        for (int w = 1; w < 4; w++) m = red[w] > m ? red[w] : m;
        atomicMax(&amax[0], m);
        atomicMax(&amax[3], m);
    }
}


namespace {

using RefKernel = void (*)(const double *, const double *, double *, RefArgs, int, double, double, DiffCoef);

template <int WENO, int P, int TX, int TY>
RefKernel ref_kernel(bool visc) {
    return visc ? k_burgers_ref_stage<WENO, P, true, TX, TY> : k_burgers_ref_stage<WENO, P, false, TX, TY>;
}

template <int TX, int TY>
RefKernel ref_kernel_for(const mgad_scheme &sc, bool visc) {
    if (sc.weno == MGAD_WENO_JS) return ref_kernel<kWenoJS, 2, TX, TY>(visc);
    if (sc.p == 2) return ref_kernel<kWenoZ, 2, TX, TY>(visc);
    return ref_kernel<kWenoZ, 1, TX, TY>(visc);
}

struct RefLaunch {
    RefKernel fn;
    RefArgs g;
    unsigned blocks, threads;
};

RefLaunch ref_launch(const mgad_scheme &sc, bool visc, unsigned nx, unsigned ny, unsigned nz, double dx, double dy,
                     double dz) {
    RefLaunch L{};
    const int ix = static_cast<int>(nx) - 6, iy = static_cast<int>(ny) - 6, iz = static_cast<int>(nz) - 6;
    // one round of 32x32 workgroups, one per CU, when the interior's tiles
    // fill the CUs (as for the periodic stage); else 32x16 or 16x16 tiles
    const int sq = ix > 0 && iy > 0 && iz > 0
                       ? square_bands(static_cast<long>((ix + 31) / 32) * ((iy + 31) / 32), static_cast<unsigned>(iz))
                       : 0;
    const bool wide = ix > 16;
    const int TX = sq > 0 || wide ? 32 : 16, TY = sq > 0 ? 32 : 16;
    L.fn = sq > 0 ? ref_kernel_for<32, 32>(sc, visc)
                  : (wide ? ref_kernel_for<32, 16>(sc, visc) : ref_kernel_for<16, 16>(sc, visc));
    L.threads = TX * TY;
    RefArgs &g = L.g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.ntx = (ix + TX - 1) / TX;
    g.nty = (iy + TY - 1) / TY;
    // z chunks: with 32x32 tiles the one round's bands; else enough blocks
    // for several per CU, at least 8 planes each
    int kc = iz;
    if (sq > 0)
        kc = (iz + sq - 1) / sq;
    else
        while (kc > 8 && static_cast<long>(g.ntx) * g.nty * ((iz + kc - 1) / kc) < 2048) kc = (kc + 1) / 2;
    g.kchunk = kc;
    g.dx = dx;
    g.dy = dy;
    g.dz = dz;
    g.rdx = 1.0 / dx;
    g.rdy = 1.0 / dy;
    g.rdz = 1.0 / dz;
    L.blocks = static_cast<unsigned>(g.ntx) * g.nty * ((iz + kc - 1) / kc);
    return L;
}

int enqueue_ref_step(hipStream_t s, const RefLaunch &L, double dt, double eps, DiffCoef kc, double *u, double *w1,
                     double *w2) {
    hipLaunchKernelGGL(L.fn, dim3(L.blocks), dim3(L.threads), 0, s, u, u, w1, L.g, 1, dt, eps, kc);
    hipLaunchKernelGGL(L.fn, dim3(L.blocks), dim3(L.threads), 0, s, w1, u, w2, L.g, 2, dt, eps, kc);
    hipLaunchKernelGGL(L.fn, dim3(L.blocks), dim3(L.threads), 0, s, w2, u, u, L.g, 3, dt, eps, kc);
    return check_launch();
}

// The max|u| slots of reference-step runs on a stream (4 per stream, kept
// for the process: graphs captured on the stream point at them).
unsigned long long *ref_amax_slots(hipStream_t s) {
    static std::map<hipStream_t, unsigned long long *> slots;
    auto it = slots.find(s);
    if (it != slots.end()) return it->second;
    unsigned long long *p = nullptr;
    if (check_hip(hipMalloc(&p, 4 * sizeof(unsigned long long)))) return nullptr;
    slots[s] = p;
    return p;
}

using RefGraphKey = std::tuple<hipStream_t, int, int, double, int, unsigned, unsigned, unsigned, double, double,
                               double, double, double, double, double, double *, double *, double *>;

std::map<RefGraphKey, hipGraphExec_t> &ref_graphs() {
    static std::map<RefGraphKey, hipGraphExec_t> m;
    return m;
}

}  // namespace

}  // namespace mgad

using namespace mgad;

extern "C" int mgad_burgers_ref_steps(hipStream_t stream, const mgad_scheme *scheme, int viscous, unsigned int nx,
                                      unsigned int ny, unsigned int nz, double dx, double dy, double dz, double dt,
                                      double *u, double *w1, double *w2, int nsteps, int use_graph) {
    if (!scheme || !u || !w1 || !w2 || u == w1 || u == w2 || w1 == w2 || nsteps < 0 || nx < 7 || ny < 7 ||
        nz < 7 || 1ull * nx * ny * nz >= (1ull << 31) || !(dx > 0) || !(dy > 0) || !(dz > 0) ||
        (scheme->weno != MGAD_WENO_Z && scheme->weno != MGAD_WENO_JS) || (scheme->p != 1 && scheme->p != 2)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    // the stage buffers carry u's ghost layers (never written by any stage)
    const size_t bytes = sizeof(double) * nx * ny * nz;
    int st = check_hip(hipMemcpyAsync(w1, u, bytes, hipMemcpyDeviceToDevice, stream));
    if (!st) st = check_hip(hipMemcpyAsync(w2, u, bytes, hipMemcpyDeviceToDevice, stream));
    if (st || nsteps == 0) return st;
    std::lock_guard<std::recursive_mutex> lock(host_mutex());
    RefLaunch L = ref_launch(*scheme, viscous != 0, nx, ny, nz, dx, dy, dz);
    L.g.amax = ref_amax_slots(stream);
    if (!L.g.amax) return MGAD_ERR_HIP;
    hipLaunchKernelGGL(k_ref_amax_clear, dim3(1), dim3(1), 0, stream, L.g.amax);
    const size_t n = 1ull * nx * ny * nz;
    hipLaunchKernelGGL(k_ref_amax_init, dim3(static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 2048))),
                       dim3(256), 0, stream, u, n, L.g.amax);
    if ((st = check_launch())) return st;
    const DiffCoef kc = viscous ? diff_coef() : DiffCoef{0.0, 0.0, 0.0};
    if (!use_graph || stream == nullptr) {
        for (int s = 0; s < nsteps && !st; s++) st = enqueue_ref_step(stream, L, dt, scheme->eps, kc, u, w1, w2);
        return st;
    }
    const RefGraphKey key{stream, scheme->weno, scheme->p, scheme->eps, viscous != 0, nx, ny, nz, dx, dy, dz, dt,
                          kc.kx, kc.ky, kc.kz, u, w1, w2};
    auto &cache = ref_graphs();
    hipGraphExec_t exec = nullptr;
    auto it = cache.find(key);
    if (it == cache.end()) {
        if (cache.size() > 16) {
            clear_graph_cache(cache);
        }
        hipGraph_t gr = nullptr;
        if ((st = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal)))) return st;
        const int est = enqueue_ref_step(stream, L, dt, scheme->eps, kc, u, w1, w2);
        st = check_hip(hipStreamEndCapture(stream, &gr));
        if (est) {
            if (gr) (void)hipGraphDestroy(gr);
            return est;
        }
        if (st) return st;
        st = check_hip(hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
        if (st) return st;
        cache[key] = exec;
    } else {
        exec = it->second;
    }
    for (int s = 0; s < nsteps && !st; s++) st = check_hip(hipGraphLaunch(exec, stream));
    return st;
}
