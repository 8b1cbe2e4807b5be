// burgers_periodic.hpp -- shared declarations of the fused periodic Burgers
// stage (single-problem launches and the batched persistent kernel).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mgad_common.hpp"

namespace mgad {

// Per-step record of a run (kept by the adjoint): dt, the LF alpha of each
// stage and the unclipped CFL step cfl*dx/a1 (a clipped last step has
// dt != dt_cfl).
struct StepRec {
    double dt, a1, a2, a3, dt_cfl;
};

// Layout of the caller's MGAD_CTRL_BYTES device buffer.
struct Ctrl {
    double t;                    // time at the start of the next step
    double dt;                   // dt of the current step, 0 once done
    double tend;
    double cfl;
    unsigned long long amax[3];  // max|u| of: [0] u^n, [1] u^(1), [2] u^(2)
    unsigned long long steps;
    StepRec *rec;                // if set, step s records into rec[s - rec_step0]
    unsigned long long rec_step0;
};
static_assert(sizeof(Ctrl) <= MGAD_CTRL_BYTES, "ctrl layout");

// Grid of one problem: extents, spacings and their correctly rounded
// reciprocals (for div_by), planes per z chunk.
struct GridArgs {
    int nx, ny, nz, kchunk;
    double dx, dy, dz, rdx, rdy, rdz;
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// A value every lane of the wave holds alike (loaded from one address),
// moved to scalar registers: it then takes no vector registers for the rest
// of the kernel.
__device__ __forceinline__ double wave_uniform(double x) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(b));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(b >> 32));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Step control shared by every block of a stage.  Stage 1 derives dt from
// max|u^n| (all blocks compute the same value) and records it; stage 2 moves
// t forward.  Each stage's lead thread clears the max|u| slot the next-but-one
// stage accumulates into, and fills the step record when one is kept.
// Returns false once the run has reached tend, or when a stage input's
// max|u| is NaN or infinite: then dt is set to 0, so the rest of the step is
// skipped too (u^n stays as it was; only stage 3 writes it), t and the step
// count stay, and the non-finite max|u| is left in its slot for the host to
// report.  Stage 1 also checks the slots of stages 2 and 3, so a run stopped
// inside a step stays stopped.
__device__ __forceinline__ bool stage_control(Ctrl *ctrl, int stage, double dx, bool lead, double &dt,
                                              double &alpha) {
    if (stage == 1) {
        const double t = ctrl->t, tend = ctrl->tend;
        const double a = bits_double(ctrl->amax[0]);
        const unsigned long long inf_bits = 0x7FF0000000000000ull;
        if (!(t < tend) || !(a <= 1.7976931348623157e308) || ctrl->amax[1] >= inf_bits ||
            ctrl->amax[2] >= inf_bits) {
            if (lead) ctrl->dt = 0.0;
            return false;
        }
        dt = ctrl->cfl * dx / a;
        if (t + dt > tend) dt = tend - t;
        alpha = a;
        if (lead) {
            ctrl->dt = dt;
            ctrl->amax[2] = 0ull;
            if (ctrl->rec) {
                StepRec &r = ctrl->rec[ctrl->steps - ctrl->rec_step0];
                r.dt = dt;
                r.a1 = a;
                r.dt_cfl = ctrl->cfl * dx / a;
            }
        }
        dt = wave_uniform(dt);
        alpha = wave_uniform(alpha);
        return true;
    }
    dt = ctrl->dt;
    if (dt == 0.0) return false;
    alpha = bits_double(ctrl->amax[stage - 1]);
    if (!(alpha <= 1.7976931348623157e308)) {
        if (lead) ctrl->dt = 0.0;
        return false;
    }
    if (lead) {
        if (stage == 2) {
            if (ctrl->rec) ctrl->rec[ctrl->steps - ctrl->rec_step0].a2 = alpha;
            ctrl->amax[0] = 0ull;
            ctrl->t = ctrl->t + dt;
            ctrl->steps = ctrl->steps + 1ull;
        } else {
            if (ctrl->rec) ctrl->rec[ctrl->steps - 1ull - ctrl->rec_step0].a3 = alpha;
            ctrl->amax[1] = 0ull;
        }
    }
    dt = wave_uniform(dt);
    alpha = wave_uniform(alpha);
    return true;
}

template <class T>
__device__ __forceinline__ double ld(const T *p, unsigned o) {
    return static_cast<double>(p[o]);
}

// LDS of one tile.
template <int TX, int TY>
struct TileSmem {
    static constexpr int RS = TX + 6;  // row stride of P
    // (f+, f-) of the current plane over the tile and its halo, node (r, c)
    // at P[r*RS + c] for column x0 + c - 3, row y0 + r - 3
    double2 P[(TY + 6) * RS];
    double FX[TY][TX + 1];  // x faces x0-1/2 .. x0+TX-1/2
    double FY[TY + 1][TX];  // y faces y0-1/2 .. y0+TY-1/2
    unsigned long long red[TX * TY / 64];
};

// One stage over the tile whose first column is (x0, y0), planes z0 .. z1-1
// (0 <= z0 < z1 <= nz); ends with one atomicMax per block into the stage's
// max|u| slot.  FD: fast_division_ok<WENO>(alpha, eps) holds.
// ZG: the field is one z slab of nz planes stored between 3 ghost planes
// below and 3 above (plane k at storage plane k + 3, read for k in -3..nz+2,
// written for k in 0..nz-1); otherwise z is periodic over nz planes.  With
// ZG, planes 0..2 are also stored into the top ghost planes of the slab set
// at gh_lo and planes nz-3..nz-1 into the bottom ghost planes at gh_hi (the
// neighbours' copies of them, so no separate exchange is needed).
// XW: the tile may reach past nx or ny (x0 + TX > nx, y0 + TY > ny); its
// columns x >= nx stand for x mod nx and its rows y >= ny for y mod ny.  They
// are computed like the others, so every face of the real columns is there,
// but are neither stored nor counted in max|u| (in stage 3 their u^n may
// already hold the owner's new value).
template <class T, int WENO, int P, int TX, int TY, bool FD, bool ZG = false, bool XW = false,
          bool LEAN = false>
__device__ __forceinline__ void stage_cols(const T *__restrict__ uin, const T *un, T *uout,
                                           const GridArgs &g, int stage, int x0, int y0, int z0, int z1,
                                           double dt, double alpha, double eps, TileSmem<TX, TY> &sm,
                                           unsigned long long *amax_slot, T *gh_lo, T *gh_hi) {
    using Sm = TileSmem<TX, TY>;
    constexpr int RS = Sm::RS;
    constexpr int NT = TX * TY;
    constexpr int NW = NT / 64;
    static_assert(NT % 64 == 0, "whole waves");
    static_assert(6 * TX + 6 * TY <= NT, "one halo node per thread at most");
    static_assert(TX + TY <= 64, "edge faces fit in one wave");
    static_assert(!(XW && ZG), "x-wrapped tiles are for periodic fields");
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int wave = tid >> 6, lane = tid & 63;
    double2 *const L = sm.P;
    const int o = ty * RS + tx;  // this thread's base index in P
    const unsigned plane = static_cast<unsigned>(nx) * static_cast<unsigned>(ny);
    // x (y) index of a column (row) or halo node at x0 + i (i >= -3): wrapped
    // once by periodicity, or reduced mod nx (ny) when the tile reaches past it
    auto xmod = [&](int i) { return XW ? (i + nx) % nx : wrapi(i, nx); };
    auto ymod = [&](int i) { return XW ? (i + ny) % ny : wrapi(i, ny); };
    const bool own_col = !XW || (x0 + tx < nx && y0 + ty < ny);
    const unsigned col = static_cast<unsigned>(XW ? xmod(x0 + tx) : x0 + tx) +
                         static_cast<unsigned>(nx) * static_cast<unsigned>(XW ? ymod(y0 + ty) : y0 + ty);
    // storage plane of plane k: zw for k in -3..nz+3 (wrapped when periodic), zi for 0 <= k < nz
    auto zw = [&](int k) { return static_cast<unsigned>(ZG ? k + 3 : wrapi(k, nz)); };
    auto zi = [&](int k) { return static_cast<unsigned>(ZG ? k + 3 : k); };

    // Halo node handled by this thread: 6*TY x-halo nodes, then 6*TX y-halo
    // nodes; hl is its index in P.
    int hl = 0;
    unsigned hcol = 0;
    const bool has_halo = tid < 6 * TY + 6 * TX;
    if (tid < 6 * TY) {
        const int r = tid / 6, c = tid % 6;
        const int hl_c = c < 3 ? c : TX + c;
        hl = (r + 3) * RS + hl_c;
        hcol = static_cast<unsigned>(xmod(x0 + hl_c - 3)) +
               static_cast<unsigned>(nx) * static_cast<unsigned>(XW ? ymod(y0 + r) : y0 + r);
    } else if (has_halo) {
        const int h = tid - 6 * TY;
        const int c = h % TX, q = h / TX;
        const int hl_r = q < 3 ? q : TY + q;
        hl = hl_r * RS + c + 3;
        hcol = static_cast<unsigned>(XW ? xmod(x0 + c) : x0 + c) +
               static_cast<unsigned>(nx) * static_cast<unsigned>(ymod(y0 + hl_r - 3));
    }

    // z register rings (see the march below): qp/qm hold the split fluxes of
    // planes k-3..k+3, plane p in slot (p - z0 + 3) & 7; sp/sn the 13/12 terms
    // centred on planes k-1..k+2, centre c in slot (c - z0 + 1) & 3; qu u_in of
    // planes k..k+3, plane p in slot (p - z0) & 3.  At k = z0 these are the
    // slots q of plane k-3+q, centre k-1+q and plane k+q.
    double qp[8], qm[8], sp[4], sn[4], qu[4];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const double v = ld(uin, col + plane * zw(z0 - 3 + q));
        lf_split(v, alpha, qp[q], qm[q]);
        if (q >= 3) qu[q - 3] = v;
    }
    // 13/12 terms at centres z0-2 .. z0+1 (index 0..3 here, shifted below)
    const double sp0 = sterm(qp[0], qp[1], qp[2]);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        sp[q] = sterm(qp[q + 1], qp[q + 2], qp[q + 3]);
        sn[q] = sterm(qm[q + 1], qm[q + 2], qm[q + 3]);
    }
    // face z0-1/2: f+ at z0-3..z0+1 (centres z0-2..z0), f- at z0-2..z0+2 (centres z0-1..z0+1)
    double fz_prev = weno_face_s<WENO, P, FD>(qp[0], qp[1], qp[2], qp[3], qp[4], qm[1], qm[2], qm[3],
                                          qm[4], qm[5], sp0, sp[0], sp[1], sn[0], sn[1], sn[2], eps);

    const T *own_ptr = uin + col;
    const T *un_ptr = un + col;
    T *out_ptr = uout + col;
    const T *halo_ptr = uin + hcol;
    // storage plane of the next own-column load (plane k + 3)
    unsigned kn = zw(z0 + 3);
    double own_next = ld(own_ptr, plane * kn);
    double halo_next = has_halo ? ld(halo_ptr, plane * zi(z0)) : 0.0;
    double un_next = (stage == 1) ? 0.0 : ld(un_ptr, plane * zi(z0));
    unsigned long long vmax = 0ull;

    // The queues are rings over the z march, indexed by the plane's offset
    // from z0 modulo 8 (split fluxes, planes k-3..k+3) or 4 (13/12 terms,
    // centres k-1..k+2; u_in, planes k..k+3), and the march is unrolled by 8,
    // so every index is a constant and no register moves shift the queues.
    int kb = z0;
    auto zstep = [&](auto S) {
        constexpr int s = decltype(S)::value;
        auto R8 = [](int q) { return (s + q) & 7; };
        auto R4 = [](int q) { return (s + q) & 3; };
        const int k = kb + s;
        const double own = own_next, halo = halo_next, unv = un_next;
        if (k + 1 < z1) {
            kn = (!ZG && kn + 1 == static_cast<unsigned>(nz)) ? 0u : kn + 1;
            own_next = ld(own_ptr, plane * kn);
            if (has_halo) halo_next = ld(halo_ptr, plane * zi(k + 1));
            if (stage != 1) un_next = ld(un_ptr, plane * zi(k + 1));
        }
        lf_split(own, alpha, qp[R8(6)], qm[R8(6)]);
        qu[R4(3)] = own;
        sp[R4(3)] = sterm(qp[R8(4)], qp[R8(5)], qp[R8(6)]);  // centre k+2
        sn[R4(3)] = sterm(qm[R8(4)], qm[R8(5)], qm[R8(6)]);
        L[o + 3 * RS + 3] = make_double2(qp[R8(3)], qm[R8(3)]);
        if (has_halo) {
            double hp, hm;
            lf_split(halo, alpha, hp, hm);
            L[hl] = make_double2(hp, hm);
        }
        __syncthreads();

        // z face k+1/2 from the register queue; x face x+1/2 and y face y+1/2 from LDS
        const double fz = weno_face_s<WENO, P, FD>(qp[R8(1)], qp[R8(2)], qp[R8(3)], qp[R8(4)], qp[R8(5)],
                                                   qm[R8(2)], qm[R8(3)], qm[R8(4)], qm[R8(5)], qm[R8(6)],
                                                   sp[R4(0)], sp[R4(1)], sp[R4(2)], sn[R4(1)], sn[R4(2)],
                                                   sn[R4(3)], eps);
        {
            const double2 *a = L + o + 3 * RS + 1;  // columns x-2 .. x+3
            sm.FX[ty][tx + 1] = weno_face2<WENO, P, FD>(a[0].x, a[1].x, a[2].x, a[3].x, a[4].x, a[1].y,
                                                         a[2].y, a[3].y, a[4].y, a[5].y, eps);
        }
        {
            const double2 *a = L + o + RS + 3;  // rows y-2 .. y+3
            sm.FY[ty + 1][tx] =
                weno_face2<WENO, P, FD>(a[0].x, a[RS].x, a[2 * RS].x, a[3 * RS].x, a[4 * RS].x, a[RS].y,
                                        a[2 * RS].y, a[3 * RS].y, a[4 * RS].y, a[5 * RS].y, eps);
        }
        // faces on the tile's low edges, by one wave that rotates with k:
        // lanes < TY the x faces x0-1/2 of rows 0..TY-1 (nodes (r, 0..5)),
        // the next TX lanes the y faces y0-1/2 of columns 0..TX-1 (nodes
        // (0..5, c))
        if (wave == k % NW && lane < TY + TX) {
            int e0 = lane < TY ? (lane + 3) * RS : lane - TY + 3;
            int es = lane < TY ? 1 : RS;
            // LEAN (callers near the 128-register bound): computed here, not
            // kept in registers over the march
            if (LEAN) asm volatile("" : "+v"(e0), "+v"(es));
            const double2 *a = L + e0;
            const double2 a0 = a[0], a1 = a[es], a2 = a[2 * es], a3 = a[3 * es], a4 = a[4 * es],
                          a5 = a[5 * es];
            const double f = weno_face2<WENO, P, FD>(a0.x, a1.x, a2.x, a3.x, a4.x, a1.y, a2.y, a3.y, a4.y,
                                                 a5.y, eps);
            if (lane < TY)
                sm.FX[lane][0] = f;
            else
                sm.FY[0][lane - TY] = f;
        }
        __syncthreads();

        double Lu = -div_by(sm.FX[ty][tx + 1] - sm.FX[ty][tx], g.dx, g.rdx);
        Lu -= div_by(sm.FY[ty + 1][tx] - sm.FY[ty][tx], g.dy, g.rdy);
        Lu -= div_by(fz - fz_prev, g.dz, g.rdz);
        const double uo = (stage == 1) ? qu[R4(0)] : unv;
        double unew;
        if (stage == 1)
            unew = uo + dt * Lu;
        else if (stage == 2)
            unew = 0.75 * uo + 0.25 * (qu[R4(0)] + dt * Lu);
        else
            unew = div_by(p2fma(2.0, qu[R4(0)] + dt * Lu, uo), 3.0, 1.0 / 3.0);  // (uo + 2(..)) / 3
        const T stored = static_cast<T>(unew);
        if (own_col) out_ptr[plane * zi(k)] = stored;
        if (ZG) {
            if (k < 3) gh_lo[col + plane * static_cast<unsigned>(nz + 3 + k)] = stored;
            if (k >= nz - 3) gh_hi[col + plane * static_cast<unsigned>(k - nz + 3)] = stored;
        }
        const unsigned long long b = own_col ? abs_bits(static_cast<double>(stored)) : 0ull;
        vmax = b > vmax ? b : vmax;

        fz_prev = fz;
    };
    for (; kb < z1; kb += 8) {
        zstep(std::integral_constant<int, 0>{});
        if (kb + 1 >= z1) break;
        zstep(std::integral_constant<int, 1>{});
        if (kb + 2 >= z1) break;
        zstep(std::integral_constant<int, 2>{});
        if (kb + 3 >= z1) break;
        zstep(std::integral_constant<int, 3>{});
        if (kb + 4 >= z1) break;
        zstep(std::integral_constant<int, 4>{});
        if (kb + 5 >= z1) break;
        zstep(std::integral_constant<int, 5>{});
        if (kb + 6 >= z1) break;
        zstep(std::integral_constant<int, 6>{});
        if (kb + 7 >= z1) break;
        zstep(std::integral_constant<int, 7>{});
    }

    // ---- block max -> one atomic per block ----
    vmax = wave_max_u64(vmax);
    if (lane == 0) sm.red[wave] = vmax;
    __syncthreads();
    if (tid == 0) {
        unsigned long long m = sm.red[0];
#pragma unroll
        for (int w = 1; w < NW; w++) m = sm.red[w] > m ? sm.red[w] : m;
        atomicMax(amax_slot, m);
    }
}

// stage_cols for block `tile` of the chunked decomposition: tiles x-fastest,
// then y, then z chunks of g.kchunk planes.  XW: the last tile of a row or
// column may reach past nx or ny (see stage_cols).
template <class T, int WENO, int P, int TX, int TY, bool FD, bool ZG = false, bool XW = false>
__device__ __forceinline__ void stage_tile(const T *__restrict__ uin, const T *un, T *uout,
                                           const GridArgs &g, int stage, unsigned tile, double dt,
                                           double alpha, double eps, TileSmem<TX, TY> &sm,
                                           unsigned long long *amax_slot, T *gh_lo = nullptr,
                                           T *gh_hi = nullptr) {
    const int ntx = XW ? (g.nx + TX - 1) / TX : g.nx / TX, nty = XW ? (g.ny + TY - 1) / TY : g.ny / TY;
    const int bx = tile % ntx;
    const int by = (tile / ntx) % nty;
    const int bz = tile / (ntx * nty);
    const int z0 = bz * g.kchunk;
    stage_cols<T, WENO, P, TX, TY, FD, ZG, XW, ZG>(uin, un, uout, g, stage, bx * TX, by * TY, z0,
                                               min(z0 + g.kchunk, g.nz), dt, alpha, eps, sm, amax_slot, gh_lo,
                                               gh_hi);
}

int default_kchunk(int nx, int ny, int nz, int tx, int ty);

// Bands of the default launch's 32x32 tiles.  A 16-wave workgroup of 32x32
// columns fills a CU, so one round of one workgroup per CU over the whole grid
// has no CU holding two workgroups (the first-dispatched of two issues ahead
// of the second, so pairs end at uneven times) and no CU waiting for a next
// one.  For `tiles` 32x32 column tiles over nz planes: the band count b (bands
// of at least 4 planes) with which tiles x b workgroups fill at least 90 % of
// the CUs in one round, or 0 when no such b exists.
int square_bands(long tiles, unsigned nz);

// z bands of a stage launch: band b covers planes z[b] .. z[b+1]-1.
constexpr int kMaxBands = 128;
struct ZBands {
    int n;
    int z[kMaxBands + 1];
};
// Bands of kchunk > 0 planes (more when nz / kchunk exceeds kMaxBands), or for
// kchunk = MGAD_BANDS_SHRINK bands that shrink towards the top, sized from the
// tiles per plane and the workgroups the device keeps resident.
ZBands make_bands(int nz, long ntiles, long groups, int kchunk);

// Tile width (16 or 32; 0 = widest that divides nx) and planes per block
// (0 = default_kchunk) of the single-problem stage launch.
struct LaunchCfg {
    int tile, kchunk;
};
LaunchCfg launch_cfg(int precision, const mgad_scheme &sc, unsigned nx, unsigned ny, unsigned nz);
void set_launch_cfg(int precision, const mgad_scheme &sc, unsigned nx, unsigned ny, unsigned nz, LaunchCfg c);
bool valid_grid(unsigned nx, unsigned ny, unsigned nz);
GridArgs make_grid(unsigned nx, unsigned ny, unsigned nz, double dx, double dy, double dz, int kchunk);

template <class T>
int burgers_stage_launch_t(hipStream_t stream, const mgad_scheme &sc, unsigned nx, unsigned ny,
                           unsigned nz, double dx, double dy, double dz, int stage, const T *u_in,
                           const T *u_n, T *u_out, void *ctrl, int kchunk, int tile_pref);
template <class T>
int burgers_init_t(hipStream_t stream, unsigned nx, unsigned ny, unsigned nz, const T *u, void *ctrl,
                   double t0, double tend, double cfl);

}  // namespace mgad
