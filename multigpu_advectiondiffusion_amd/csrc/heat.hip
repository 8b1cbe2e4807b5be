// heat.hip -- 3-D heat equation, 7-point Laplacian + forward Euler fused in
// one pass: u' = u + dt*L7(u) on the Dirichlet interior, boundary copied.
// L7 is Compute_Laplace3d_O2 (SingleGPU/Diffusion3d_baselineCode/
// kernels.cu:59-62) and the update is stage 1 of Compute_RK (:130-132), so a
// step reads u once and writes u' once (16 B per node) instead of the
// reference's Lu round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <type_traits>

#include "mgad_common.hpp"

namespace mgad {

// Both heat kernels march a 64 x TY tile of (x, y) columns through a chunk of
// z planes.  Wave w of the 256-thread block owns the tile rows w, w + 4, ...
// and keeps its columns' z neighbours in a register ring; each plane of the
// tile, with its x/y halo, passes through LDS once (double buffered, one
// barrier per plane), so every node comes from memory once per tile and the
// x/y neighbours are LDS reads.  Loads for the next plane are issued before
// the current one is computed, and every thread issues the same loads each
// plane (indices clamped into the grid), so the compiler's wait counts stay
// exact across the loop.  Tiles are ordered so that the blocks of one XCD
// take neighbouring tiles (their halos meet in that XCD's L2).
constexpr int kHeatTX = 64;

template <int N>
using Ic = std::integral_constant<int, N>;

typedef unsigned int v2u32 __attribute__((ext_vector_type(2)));

// One z plane of a field as a buffer: the plane's base and byte size in
// scalar registers, a 32-bit byte offset per lane (a plane holds fewer than
// 2^29 nodes).  An offset at the plane's size drops the store (the buffer
// range check), which masks nodes outside the grid or the updated region.
struct PlaneBuf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ PlaneBuf(const double *base, unsigned bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), 0, static_cast<int>(bytes), 0x00020000)) {}
    __device__ __forceinline__ double load(unsigned off) const {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
    __device__ __forceinline__ void store(unsigned off, double x) const {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u32, x), r, off, 0, 0);
    }
};

// A halo entry of one thread: LDS (row, column) and the byte offset within a
// plane it is loaded from (clamped into the grid).
struct HaloSlot {
    int row = -1, col = 0;
    unsigned off = 0;
};

// The halo a thread loads, R nodes wide, for a tile of TY rows whose first
// node is (i0, j0): waves 0 and 1 the R rows below and above (R entries each),
// waves 2 and 3 the R columns left and right of every row.
template <int R, int TY>
__device__ __forceinline__ void heat_halo(HaloSlot (&h)[R], int w, int lane, int i0, int j0, int nx, int ny,
                                          int px) {
    constexpr int W = kHeatTX + 2 * R;
    static_assert(2 * R * TY <= 128, "x halo exceeds waves 2 and 3");
    if (w < 2) {
#pragma unroll
        for (int e = 0; e < R; e++) {
            const int jj = w == 0 ? max(j0 - R + e, 0) : min(j0 + TY + e, ny - 1);
            h[e].row = w == 0 ? e : TY + R + e;
            h[e].col = lane + R;
            h[e].off = 8u * static_cast<unsigned>(min(i0 + lane, nx - 1) + px * jj);
        }
    } else {
        const int t = (w - 2) * 64 + lane;
        if (t < 2 * R * TY) {
            const int c = t / TY, rr = t % TY;  // c: column 0..2R-1 left to right
            const int ii = c < R ? max(i0 - R + c, 0) : min(i0 + kHeatTX + c - R, nx - 1);
            h[0].row = rr + R;
            h[0].col = c < R ? c : W - 2 * R + c;
            h[0].off = 8u * static_cast<unsigned>(ii + px * min(j0 + rr, ny - 1));
        }
    }
}

// One forward-Euler step, u' = u + dt*L7(u) on the Dirichlet interior and
// u' = u on the boundary nodes.  L7 is Compute_Laplace3d_O2
// (Diffusion3d_baselineCode/kernels.cu:59-62) with 12*k as the source
// evaluates it, the update stage 1 of Compute_RK (:130-132); a step reads u
// once and writes u' once (16 B per node) instead of the reference's Lu
// round trip.
template <int TY>
__global__ __launch_bounds__(256) void k_heat7_fe(const double *__restrict__ u, double *__restrict__ v, int nx,
                                                  int ny, int nz, int ntx, int nty, int kchunk, double kx, double ky,
                                                  double kz, double dt) {
    constexpr int RPT = TY / 4, W = kHeatTX + 2;
    __shared__ double sp[2][TY + 2][W];
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tix = tile % ntx, tiy = (tile / ntx) % nty, tiz = tile / (ntx * nty);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = tix * kHeatTX, j0 = tiy * TY, z0 = tiz * kchunk, z1 = min(z0 + kchunk, nz);
    const int i = i0 + lane;
    const size_t xy = static_cast<size_t>(nx) * ny;
    const unsigned pb = static_cast<unsigned>(xy * 8);
    unsigned lo[RPT], so[RPT];  // load and store offsets; so = pb drops the store
    bool keep[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int j = j0 + w + 4 * r;
        lo[r] = 8u * static_cast<unsigned>(min(i, nx - 1) + nx * min(j, ny - 1));
        so[r] = i < nx && j < ny ? lo[r] : pb;
        keep[r] = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
    }
    HaloSlot hs[1];
    heat_halo<1, TY>(hs, w, lane, i0, j0, nx, ny, nx);
    const bool hal = hs[0].row >= 0;
    const unsigned ho = hal ? hs[0].off : lo[0];  // lanes without a halo node repeat their own
    const double ckx = 12 * kx, cky = 12 * ky, ckz = 12 * kz;
    const int zl = min(z1, nz - 1);  // last plane this chunk reads
    // plane p of the own columns in slot (p - z0 + 1) & 3, of the halo in (p - z0) & 1
    double q[RPT][4] = {}, h[2] = {};
    {
        const PlaneBuf pa(u + xy * max(z0 - 1, 0), pb), pc(u + xy * z0, pb), pd(u + xy * min(z0 + 1, zl), pb);
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            q[r][0] = pa.load(lo[r]);
            q[r][1] = pc.load(lo[r]);
            q[r][2] = pd.load(lo[r]);
        }
        h[0] = pc.load(ho);
    }
    int kb = z0;
    auto plane = [&](auto S) {
        constexpr int s = decltype(S)::value;
        const int k = kb + s;
        const PlaneBuf pn(u + xy * min(k + 2, zl), pb), ph(u + xy * min(k + 1, z1 - 1), pb);
#pragma unroll
        for (int r = 0; r < RPT; r++) q[r][(s + 3) & 3] = pn.load(lo[r]);
        h[(s + 1) & 1] = ph.load(ho);
        double(*P)[W] = sp[s & 1];
#pragma unroll
        for (int r = 0; r < RPT; r++) P[w + 4 * r + 1][lane + 1] = q[r][(s + 1) & 3];
        if (hal) P[hs[0].row][hs[0].col] = h[s & 1];
        __syncthreads();
        const bool zb = k == 0 || k == nz - 1;
        const PlaneBuf pv(v + xy * k, pb);
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            const int row = w + 4 * r + 1;
            const double cur = q[r][(s + 1) & 3], below = q[r][s & 3], above = q[r][(s + 2) & 3];
            // (u_e - 2 u_o) + u_w etc., the 2*u_o product exact
            const double Lu = ckx * (p2fma(-2.0, cur, P[row][lane + 2]) + P[row][lane]) +
                              cky * (p2fma(-2.0, cur, P[row + 1][lane + 1]) + P[row - 1][lane + 1]) +
                              ckz * (p2fma(-2.0, cur, above) + below);
            pv.store(so[r], (keep[r] || zb) ? cur : cur + dt * Lu);
        }
    };
    for (; kb < z1; kb += 4) {
        plane(Ic<0>{});
        if (kb + 1 >= z1) break;
        plane(Ic<1>{});
        if (kb + 2 >= z1) break;
        plane(Ic<2>{});
        if (kb + 3 >= z1) break;
        plane(Ic<3>{});
    }
}

// What k_heat13_rk writes: an SSP-RK3 stage of the heat step, the 13-point
// operator itself (Call_Lu3d), the operator added to Lu over the Burgers
// viscous region (Call_Diff_), or stored there (the blocked heat code's
// Call_Lu3d).
enum : int { kHeatRk = 0, kHeatLu = 1, kHeatDiffB5 = 2, kHeatLuB5 = 3 };

// One SSP-RK3 stage of the reference heat step fused: the 13-point Laplacian
// of Compute_Laplace3d (Diffusion3d_baselineCode/kernels.cu:94-97, with the
// CopyToConstantMemory coefficients) and Compute_RK (kernels.cu:127-140) on
// the interior 2..n-3 of every axis, Lu kept in a register.  Loads run D
// planes ahead of their first use: planes of the own columns sit in an
// 8-slot register ring (k-2..k+2 in use), the halo and un in 4-slot rings.
// MODE kHeatLu stores Lu instead.  MODE kHeatDiffB5 is Compute_Laplace_Async
// (SingleGPU/Burgers3d_WENO5/kernels.cu:444-494): nodes 3..n-4 in x and y,
// planes 3..nz-3, each axis's 4th-order difference summed from the low side
// (-u_{-2} + 16 u_{-1} ... instead of from the high side), added to Lu,
// which the un ring then reads (un = uout = Lu).  MODE kHeatLuB5 is
// Compute_Laplace3d_Async of SingleGPU/Diffusion3d_Blocking (kernels.cu:
// 37-88, Call_Lu3d :187-191): the same region and sums, stored into Lu.
template <int TY, int D, int MODE = kHeatRk>
__global__ __launch_bounds__(256) void k_heat13_rk(const double *__restrict__ uin, const double *un, double *uout,
                                                   int nx, int ny, int nz, int px, int ntx, int nty, int kchunk,
                                                   double kx, double ky, double kz, double dt, int stage) {
    static_assert(D >= 1 && D <= 3, "prefetch distance");
    constexpr bool LU = MODE == kHeatLu || MODE == kHeatLuB5, B5 = MODE == kHeatDiffB5 || MODE == kHeatLuB5;
    constexpr int LO = B5 ? 3 : 2;  // first interior node of x and y (and z)
    constexpr int RPT = TY / 4, W = kHeatTX + 4;
    __shared__ double sp[2][TY + 4][W];
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tix = tile % ntx, tiy = (tile / ntx) % nty, tiz = tile / (ntx * nty);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = LO + tix * kHeatTX, j0 = LO + tiy * TY, z0 = LO + tiz * kchunk, z1 = min(z0 + kchunk, nz - 2);
    const int i = i0 + lane;
    const size_t xy = static_cast<size_t>(px) * ny;
    const unsigned pb = static_cast<unsigned>(xy * 8);
    unsigned lo[RPT], so[RPT];  // load and store offsets; so = pb drops the store
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int j = j0 + w + 4 * r;
        lo[r] = 8u * static_cast<unsigned>(min(i, nx - 1) + px * min(j, ny - 1));
        so[r] = i < nx - LO && j < ny - LO ? lo[r] : pb;
    }
    HaloSlot hs[2];
    heat_halo<2, TY>(hs, w, lane, i0, j0, nx, ny, px);
    const unsigned ho0 = hs[0].row >= 0 ? hs[0].off : lo[0], ho1 = hs[1].row >= 0 ? hs[1].off : lo[0];
    // un is read only by stages 2 and 3 (and as Lu by kHeatDiffB5); stage 1
    // repeats uin loads instead
    const double *ur = stage == 1 && !B5 ? uin : un;
    // own plane p in slot (p - z0 + 2) & 7, halo and un plane p in (p - z0) & 3;
    // planes past the chunk's last read are clamped to it
    const int zq = z1 + 1, zh = z1 - 1;
    double q[RPT][8] = {}, h0[4] = {}, h1[4] = {}, o[RPT][4] = {};
#pragma unroll
    for (int t = 0; t < 4 + D; t++) {
        const PlaneBuf pq(uin + xy * min(z0 - 2 + t, zq), pb);
#pragma unroll
        for (int r = 0; r < RPT; r++) q[r][t] = pq.load(lo[r]);
    }
#pragma unroll
    for (int t = 0; t < D; t++) {
        const PlaneBuf ph(uin + xy * min(z0 + t, zh), pb), po(ur + xy * min(z0 + t, zh), pb);
        h0[t] = ph.load(ho0);
        h1[t] = ph.load(ho1);
        if (!LU) {
#pragma unroll
            for (int r = 0; r < RPT; r++) o[r][t] = po.load(lo[r]);
        }
    }
    int kb = z0;
    auto plane = [&](auto S) {
        constexpr int s = decltype(S)::value;
        const int k = kb + s;
        const PlaneBuf pn(uin + xy * min(k + 2 + D, zq), pb), ph(uin + xy * min(k + D, zh), pb),
            po(ur + xy * min(k + D, zh), pb);
#pragma unroll
        for (int r = 0; r < RPT; r++) q[r][(s + 4 + D) & 7] = pn.load(lo[r]);
        h0[(s + D) & 3] = ph.load(ho0);
        h1[(s + D) & 3] = ph.load(ho1);
        if (!LU) {
#pragma unroll
            for (int r = 0; r < RPT; r++) o[r][(s + D) & 3] = po.load(lo[r]);
        }
        double(*P)[W] = sp[s & 1];
#pragma unroll
        for (int r = 0; r < RPT; r++) P[w + 4 * r + 2][lane + 2] = q[r][(s + 2) & 7];
        if (hs[0].row >= 0) P[hs[0].row][hs[0].col] = h0[s & 3];
        if (hs[1].row >= 0) P[hs[1].row][hs[1].col] = h1[s & 3];
        __syncthreads();
        const PlaneBuf pv(uout + xy * k, pb);
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            const int row = w + 4 * r + 2, x = lane + 2;
            const double c = q[r][(s + 2) & 7];
            double X, Y, Z;
            if (B5) {
                X = d4(P[row][x - 2], P[row][x - 1], c, P[row][x + 1], P[row][x + 2]);
                Y = d4(P[row - 2][x], P[row - 1][x], c, P[row + 1][x], P[row + 2][x]);
                Z = d4(q[r][s & 7], q[r][(s + 1) & 7], c, q[r][(s + 3) & 7], q[r][(s + 4) & 7]);
            } else {
                X = d4(P[row][x + 2], P[row][x + 1], c, P[row][x - 1], P[row][x - 2]);
                Y = d4(P[row + 2][x], P[row + 1][x], c, P[row - 1][x], P[row - 2][x]);
                Z = d4(q[r][(s + 4) & 7], q[r][(s + 3) & 7], c, q[r][(s + 1) & 7], q[r][s & 7]);
            }
            const double Lu = (kx * X + ky * Y) + kz * Z;
            if (LU) {
                pv.store(so[r], Lu);
            } else if (B5) {  // kHeatDiffB5
                pv.store(so[r], o[r][s & 3] + Lu);
            } else {
                const double uo = stage == 1 ? c : o[r][s & 3];
                pv.store(so[r], rk_stage(stage, dt, c, uo, Lu));
            }
        }
    };
    for (; kb < z1; kb += 8) {
        plane(Ic<0>{});
        if (kb + 1 >= z1) break;
        plane(Ic<1>{});
        if (kb + 2 >= z1) break;
        plane(Ic<2>{});
        if (kb + 3 >= z1) break;
        plane(Ic<3>{});
        if (kb + 4 >= z1) break;
        plane(Ic<4>{});
        if (kb + 5 >= z1) break;
        plane(Ic<5>{});
        if (kb + 6 >= z1) break;
        plane(Ic<6>{});
        if (kb + 7 >= z1) break;
        plane(Ic<7>{});
    }
}

namespace {

// Tile heights and the 13-point kernel's prefetch distance, from sweeps at
// 512^3 (7-point: 16 rows 5.6 TB/s vs 32 rows 5.5; 13-point: 32 rows
// 4.9 TB/s vs 16 rows 4.0, distance 2 no faster than 1).
constexpr int kHeat7TY = 16, kHeat13TY = 32, kHeat13D = 1;

struct HeatGrid {
    int ntx, nty, kc;
    unsigned blocks;
};

HeatGrid heat_tiles(unsigned ex, unsigned ey, int ty) {
    HeatGrid g{};
    g.ntx = static_cast<int>((ex + kHeatTX - 1) / kHeatTX);
    g.nty = static_cast<int>((ey + ty - 1) / ty);
    return g;
}

// 7-point: z chunks as long as possible while the grid keeps >= 512 blocks
// (two per CU), down to one plane.
HeatGrid heat7_grid(unsigned nx, unsigned ny, unsigned nz) {
    HeatGrid g = heat_tiles(nx, ny, kHeat7TY);
    const long bxy = static_cast<long>(g.ntx) * g.nty;
    int kc = static_cast<int>(nz);
    while (kc > 1 && bxy * ((nz + kc - 1) / kc) < 512) kc = (kc + 1) / 2;
    g.kc = kc;
    g.blocks = static_cast<unsigned>(bxy * ((nz + kc - 1) / kc));
    return g;
}

// 13-point over the (ex, ey, ez) interior: the z chunk among 8..64 planes
// that minimises (rounds of resident blocks) x (planes each block reads,
// chunk + 4), so a small grid runs in one round of long chunks and a large
// one in a few rounds.
HeatGrid heat13_grid(unsigned ex, unsigned ey, unsigned ez) {
    HeatGrid g = heat_tiles(ex, ey, kHeat13TY);
    const long bxy = static_cast<long>(g.ntx) * g.nty;
    int dev = 0, cus = 0, per = 0;
    long cap = 256;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, reinterpret_cast<const void *>(&k_heat13_rk<kHeat13TY, kHeat13D>), 256, 0) == hipSuccess &&
        per > 0)
        cap = static_cast<long>(cus) * per;
    long best = -1;
    for (int kc : {8, 12, 16, 24, 32, 48, 64}) {
        const int k = std::min(kc, static_cast<int>(ez));
        const long blocks = bxy * ((ez + k - 1) / k);
        const long cost = ((blocks + cap - 1) / cap) * (k + 4);
        if (best < 0 || cost < best) {
            best = cost;
            g.kc = k;
            g.blocks = static_cast<unsigned>(blocks);
        }
    }
    return g;
}

int enqueue_heat13(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, DiffCoef c, double dt, double *u,
                   double *w1, double *w2) {
    const HeatGrid g = heat13_grid(nx - 4, ny - 4, nz - 4);
    const auto kern = k_heat13_rk<kHeat13TY, kHeat13D>;
    hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(256), 0, s, u, u, w1, (int)nx, (int)ny, (int)nz, (int)nx,
                       g.ntx, g.nty, g.kc, c.kx, c.ky, c.kz, dt, 1);
    hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(256), 0, s, w1, u, w2, (int)nx, (int)ny, (int)nz, (int)nx,
                       g.ntx, g.nty, g.kc, c.kx, c.ky, c.kz, dt, 2);
    hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(256), 0, s, w2, u, u, (int)nx, (int)ny, (int)nz, (int)nx,
                       g.ntx, g.nty, g.kc, c.kx, c.ky, c.kz, dt, 3);
    return check_launch();
}

int enqueue_heat(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                 double kz, double dt, double *u, double *w, int nsteps) {
    const HeatGrid g = heat7_grid(nx, ny, nz);
    for (int it = 0; it < nsteps; it++) {
        const double *src = (it & 1) ? w : u;
        double *dst = (it & 1) ? u : w;
        hipLaunchKernelGGL(k_heat7_fe<kHeat7TY>, dim3(g.blocks), dim3(256), 0, s, src, dst, (int)nx, (int)ny,
                           (int)nz, g.ntx, g.nty, g.kc, kx, ky, kz, dt);
        const int st = check_launch();
        if (st) return st;
    }
    if (nsteps & 1)
        return check_hip(hipMemcpyAsync(u, w, sizeof(double) * nx * ny * nz,
                                        hipMemcpyDeviceToDevice, s));
    return MGAD_OK;
}

}  // namespace

// Compute_Laplace3d (Diffusion3d_baselineCode/kernels.cu:94-97) into Lu on
// the interior 2..n-3 with the heat stage kernel's LDS z-march (Call_Lu3d):
// row pitch px, planes of px*ny < 2^29 values.
int heat13_lu_launch(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, unsigned px, const double *u,
                     double *Lu) {
    const HeatGrid g = heat13_grid(nx - 4, ny - 4, nz - 4);
    const DiffCoef c = diff_coef();
    const auto kern = k_heat13_rk<kHeat13TY, kHeat13D, kHeatLu>;
    hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(256), 0, s, u, u, Lu, (int)nx, (int)ny, (int)nz, (int)px, g.ntx,
                       g.nty, g.kc, c.kx, c.ky, c.kz, 0.0, 1);
    return check_launch();
}

// Compute_Laplace_Async (SingleGPU/Burgers3d_WENO5/kernels.cu:444-494,
// Call_Diff_): Lu += the 13-point operator on nodes 3..n-4 in x and y and
// planes 3..nz-3, with the same z-march; row pitch px, planes of
// px*ny < 2^29 values, nx, ny >= 7, nz >= 6.
int diff13_b5_launch(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, unsigned px, const double *u, double *Lu,
                     double kx, double ky, double kz) {
    const HeatGrid g = heat13_grid(nx - 6, ny - 6, nz - 5);
    const auto kern = k_heat13_rk<kHeat13TY, kHeat13D, kHeatDiffB5>;
    hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(256), 0, s, u, Lu, Lu, (int)nx, (int)ny, (int)nz, (int)px, g.ntx,
                       g.nty, g.kc, kx, ky, kz, 0.0, 1);
    return check_launch();
}

// Compute_Laplace3d_Async of SingleGPU/Diffusion3d_Blocking (kernels.cu:
// 37-88, launched by Call_Lu3d :187-191 with kstart 3, kstop nz-2): Lu = the
// 13-point operator on nodes 3..n-4 in x and y and planes 3..nz-3, summed
// as Call_Diff_'s operator; row pitch px, planes of px*ny < 2^29 values,
// nx, ny >= 7, nz >= 6.
int lu13_b5_launch(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, unsigned px, const double *u, double *Lu) {
    const HeatGrid g = heat13_grid(nx - 6, ny - 6, nz - 5);
    const DiffCoef c = diff_coef();
    const auto kern = k_heat13_rk<kHeat13TY, kHeat13D, kHeatLuB5>;
    hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(256), 0, s, u, u, Lu, (int)nx, (int)ny, (int)nz, (int)px, g.ntx,
                       g.nty, g.kc, c.kx, c.ky, c.kz, 0.0, 1);
    return check_launch();
}

namespace {

struct HeatKey {
    hipStream_t s;
    unsigned nx, ny, nz;
    double kx, ky, kz, dt;
    double *u, *w;
    int nsteps;
    bool operator<(const HeatKey &o) const {
        return std::tie(s, nx, ny, nz, kx, ky, kz, dt, u, w, nsteps) <
               std::tie(o.s, o.nx, o.ny, o.nz, o.kx, o.ky, o.kz, o.dt, o.u, o.w, o.nsteps);
    }
};

}  // namespace
}  // namespace mgad

using namespace mgad;

extern "C" int mgad_heat7_fe_steps(hipStream_t stream, unsigned int nx, unsigned int ny,
                                   unsigned int nz, double kx, double ky, double kz, double dt,
                                   double *u, double *w, int nsteps, int use_graph) {
    if (nx < 3 || ny < 3 || nz < 3 || 1ull * nx * ny * nz >= (1ull << 31) || 1ull * nx * ny >= (1ull << 29) || !u ||
        !w || u == w || nsteps < 0) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    if (!use_graph || stream == nullptr || nsteps == 0)
        return enqueue_heat(stream, nx, ny, nz, kx, ky, kz, dt, u, w, nsteps);
    std::lock_guard<std::recursive_mutex> lock(host_mutex());
    static std::map<HeatKey, hipGraphExec_t> cache;
    const HeatKey key{stream, nx, ny, nz, kx, ky, kz, dt, u, w, nsteps};
    auto it = cache.find(key);
    hipGraphExec_t exec = nullptr;
    if (it == cache.end()) {
        if (cache.size() > 16) {
            clear_graph_cache(cache);
        }
        hipGraph_t g = nullptr;
        int st = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        if (st) return st;
        const int est = enqueue_heat(stream, nx, ny, nz, kx, ky, kz, dt, u, w, nsteps);
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
    return check_hip(hipGraphLaunch(exec, stream));
}

// The reference's heat step fused (SingleGPU/Diffusion3d_baselineCode/
// main.cpp:108-121): per step uo <- u, then for rk = 1..3 Call_Lu3d and
// Call_RK3d(rk), one launch per stage; the two ghost layers are not written.
extern "C" int mgad_heat13_rk3_steps(hipStream_t stream, unsigned int nx, unsigned int ny, unsigned int nz, double dt,
                                     double *u, double *w1, double *w2, int nsteps, int use_graph) {
    if (nx < 5 || ny < 5 || nz < 5 || 1ull * nx * ny * nz >= (1ull << 31) || 1ull * nx * ny >= (1ull << 29) || !u ||
        !w1 || !w2 || u == w1 || u == w2 || w1 == w2 || nsteps < 0) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const size_t bytes = sizeof(double) * nx * ny * nz;
    int st = check_hip(hipMemcpyAsync(w1, u, bytes, hipMemcpyDeviceToDevice, stream));
    if (!st) st = check_hip(hipMemcpyAsync(w2, u, bytes, hipMemcpyDeviceToDevice, stream));
    if (st || nsteps == 0) return st;
    const DiffCoef c = diff_coef();
    if (!use_graph || stream == nullptr) {
        for (int s = 0; s < nsteps && !st; s++) st = enqueue_heat13(stream, nx, ny, nz, c, dt, u, w1, w2);
        return st;
    }
    std::lock_guard<std::recursive_mutex> lock(host_mutex());
    using Key = std::tuple<hipStream_t, unsigned, unsigned, unsigned, double, double, double, double, double *,
                           double *, double *>;
    static std::map<Key, hipGraphExec_t> cache;
    const Key key{stream, nx, ny, nz, c.kx, c.ky, c.kz, dt, u, w1, w2};
    hipGraphExec_t exec = nullptr;
    auto it = cache.find(key);
    if (it == cache.end()) {
        if (cache.size() > 16) {
            clear_graph_cache(cache);
        }
        hipGraph_t g = nullptr;
        if ((st = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal)))) return st;
        const int est = enqueue_heat13(stream, nx, ny, nz, c, dt, u, w1, w2);
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
    for (int s = 0; s < nsteps && !st; s++) st = check_hip(hipGraphLaunch(exec, stream));
    return st;
}
