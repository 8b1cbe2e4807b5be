// batch.hip -- many independent periodic Burgers problems of uneven sizes
// advanced together: one persistent launch per SSP-RK3 stage.
//
// The batch's work list holds (problem, tile) items; every item is a 16x16
// column tile by kchunk planes of one problem, so items cost about the same
// whatever the problem size.  A grid of as many 256-thread blocks as the
// device keeps resident walks the list (block b takes items b, b+G, ...,
// with b remapped so that the blocks of one XCD take adjacent items).  Each
// item runs the same stage_tile body as the single-problem kernel, with the
// problem's own dt, LF alpha and max|u| slot (its own ctrl), so results are
// bit-identical to advancing each problem alone.
//
// Workspace (device memory given by the caller):
//   BatchProblem[nprob] | item_prob[nitems] | item_tile[nitems]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

#include "burgers_periodic.hpp"

namespace mgad {

constexpr int kBT = 16;  // batch tile: 16 x 16 columns (divides 32, 48, 64, 96, 128)

struct BatchProblem {
    void *u, *w1, *w2;
    Ctrl *ctrl;
    GridArgs g;
};

template <class T, int WENO, int P>
__global__ __launch_bounds__(kBT *kBT, 4) void k_burgers_batch(const BatchProblem *__restrict__ probs,
                                                               const unsigned *__restrict__ item_prob,
                                                               const unsigned *__restrict__ item_tile,
                                                               int nitems, double eps, int stage) {
    __shared__ TileSmem<kBT, kBT> sm;
    const unsigned G = gridDim.x;
    for (unsigned i = xcd_tile(blockIdx.x, G); i < static_cast<unsigned>(nitems); i += G) {
        const BatchProblem &bp = probs[item_prob[i]];
        const unsigned tile = item_tile[i];
        double dt, alpha;
        if (!stage_control(bp.ctrl, stage, bp.g.dx, tile == 0 && threadIdx.x == 0, dt, alpha)) continue;
        const T *u = static_cast<const T *>(bp.u);
        const T *uin = stage == 1 ? u : (stage == 2 ? static_cast<const T *>(bp.w1) : static_cast<const T *>(bp.w2));
        T *uout = stage == 1 ? static_cast<T *>(bp.w1) : (stage == 2 ? static_cast<T *>(bp.w2) : static_cast<T *>(bp.u));
        if (fast_division_ok<WENO>(alpha, eps))
            stage_tile<T, WENO, P, kBT, kBT, true>(uin, u, uout, bp.g, stage, tile, dt, alpha, eps, sm,
                                                   &bp.ctrl->amax[stage % 3]);
        else
            stage_tile<T, WENO, P, kBT, kBT, false>(uin, u, uout, bp.g, stage, tile, dt, alpha, eps, sm,
                                                    &bp.ctrl->amax[stage % 3]);
    }
}

namespace {

template <class T>
using BatchKernel = void (*)(const BatchProblem *, const unsigned *, const unsigned *, int, double, int);

template <class T>
BatchKernel<T> pick_batch(const mgad_scheme &sc) {
    if (sc.weno == MGAD_WENO_Z && sc.p == 1) return k_burgers_batch<T, kWenoZ, 1>;
    if (sc.weno == MGAD_WENO_Z && sc.p == 2) return k_burgers_batch<T, kWenoZ, 2>;
    if (sc.weno == MGAD_WENO_JS) return k_burgers_batch<T, kWenoJS, 2>;
    return nullptr;
}

size_t ws_bytes(int nprob, long nitems) {
    return sizeof(BatchProblem) * nprob + 2 * sizeof(unsigned) * static_cast<size_t>(nitems);
}

long count_items(int nprob, const mgad_problem *p, int kchunk) {
    long n = 0;
    for (int i = 0; i < nprob; i++)
        n += static_cast<long>(p[i].nx / kBT) * (p[i].ny / kBT) * ((p[i].nz + kchunk - 1) / kchunk);
    return n;
}

int resident_blocks() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 1024;
    return cus * 4;  // four 4-wave blocks per CU at the kernel's 4 waves/SIMD occupancy
}

bool problems_ok(int nprob, const mgad_problem *p) {
    if (nprob <= 0 || !p) return false;
    for (int i = 0; i < nprob; i++) {
        if (!p[i].u || !p[i].w1 || !p[i].w2 || !p[i].ctrl || p[i].u == p[i].w1 || p[i].u == p[i].w2 ||
            p[i].w1 == p[i].w2)
            return false;
        if (!valid_grid(p[i].nx, p[i].ny, p[i].nz) || p[i].nx % kBT || p[i].ny % kBT) return false;
        if (!(p[i].dx > 0) || !(p[i].dy > 0) || !(p[i].dz > 0)) return false;
    }
    return true;
}

// Writes descriptors and the work list (problems in the given order, tiles
// x-fastest, then y, then z chunk) into the workspace.
int write_workspace(hipStream_t stream, int nprob, const mgad_problem *p, int kchunk, void *ws,
                    size_t ws_size, int *nitems_out) {
    const long nitems = count_items(nprob, p, kchunk);
    if (ws_bytes(nprob, nitems) > ws_size || nitems >= (1l << 31)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    std::vector<unsigned char> host(ws_bytes(nprob, nitems));
    BatchProblem *bp = reinterpret_cast<BatchProblem *>(host.data());
    unsigned *ip = reinterpret_cast<unsigned *>(bp + nprob);
    unsigned *it = ip + nitems;
    long k = 0;
    for (int i = 0; i < nprob; i++) {
        bp[i].u = p[i].u;
        bp[i].w1 = p[i].w1;
        bp[i].w2 = p[i].w2;
        bp[i].ctrl = static_cast<Ctrl *>(p[i].ctrl);
        bp[i].g = make_grid(p[i].nx, p[i].ny, p[i].nz, p[i].dx, p[i].dy, p[i].dz, kchunk);
        const long nt = static_cast<long>(p[i].nx / kBT) * (p[i].ny / kBT) * ((p[i].nz + kchunk - 1) / kchunk);
        for (long t = 0; t < nt; t++, k++) {
            ip[k] = static_cast<unsigned>(i);
            it[k] = static_cast<unsigned>(t);
        }
    }
    int st = check_hip(hipMemcpyAsync(ws, host.data(), host.size(), hipMemcpyHostToDevice, stream));
    if (!st) st = check_hip(hipStreamSynchronize(stream));
    *nitems_out = static_cast<int>(nitems);
    return st;
}

template <class T>
int launch_batch_stage(hipStream_t stream, const mgad_batch &b, int stage) {
    BatchKernel<T> fn = pick_batch<T>(b.scheme);
    if (!fn) {
        set_error(MGAD_ERR_UNSUPPORTED);
        return MGAD_ERR_UNSUPPORTED;
    }
    const BatchProblem *probs = static_cast<const BatchProblem *>(b.workspace);
    const unsigned *ip = reinterpret_cast<const unsigned *>(probs + b.nprob);
    const unsigned *it = ip + b.nitems;
    hipLaunchKernelGGL(fn, dim3(b.grid), dim3(kBT * kBT), 0, stream, probs, ip, it, b.nitems, b.scheme.eps,
                       stage);
    return check_launch();
}

int batch_stage(hipStream_t stream, const mgad_batch &b, int stage) {
    return b.precision == MGAD_F32 ? launch_batch_stage<float>(stream, b, stage)
                                   : launch_batch_stage<double>(stream, b, stage);
}

int enqueue_batch_step(hipStream_t stream, const mgad_batch &b) {
    int st = batch_stage(stream, b, 1);
    if (!st) st = batch_stage(stream, b, 2);
    if (!st) st = batch_stage(stream, b, 3);
    return st;
}

// One step of every problem with the single-problem kernel, problem after
// problem (the one-by-one baseline the batch is measured against).
int enqueue_serial_step(hipStream_t stream, const mgad_batch &b, const mgad_problem *p) {
    for (int i = 0; i < b.nprob; i++) {
        const LaunchCfg c = launch_cfg(b.precision, b.scheme, p[i].nx, p[i].ny, p[i].nz);
        for (int stage = 1; stage <= 3 ; stage++) {
            int st;
            if (b.precision == MGAD_F32) {
                float *u = static_cast<float *>(p[i].u), *w1 = static_cast<float *>(p[i].w1),
                      *w2 = static_cast<float *>(p[i].w2);
                const float *in = stage == 1 ? u : (stage == 2 ? w1 : w2);
                float *out = stage == 1 ? w1 : (stage == 2 ? w2 : u);
                st = burgers_stage_launch_t<float>(stream, b.scheme, p[i].nx, p[i].ny, p[i].nz, p[i].dx, p[i].dy,
                                                   p[i].dz, stage, in, u, out, p[i].ctrl, c.kchunk, c.tile);
            } else {
                double *u = static_cast<double *>(p[i].u), *w1 = static_cast<double *>(p[i].w1),
                       *w2 = static_cast<double *>(p[i].w2);
                const double *in = stage == 1 ? u : (stage == 2 ? w1 : w2);
                double *out = stage == 1 ? w1 : (stage == 2 ? w2 : u);
                st = burgers_stage_launch_t<double>(stream, b.scheme, p[i].nx, p[i].ny, p[i].nz, p[i].dx, p[i].dy,
                                                    p[i].dz, stage, in, u, out, p[i].ctrl, c.kchunk, c.tile);
            }
            if (st) return st;
        }
    }
    return MGAD_OK;
}

template <class F>
int run_graph(hipStream_t stream, int nsteps, int use_graph, F &&enqueue) {
    if (!use_graph || stream == nullptr) {
        for (int s = 0; s < nsteps; s++) {
            const int st = enqueue();
            if (st) return st;
        }
        return MGAD_OK;
    }
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    int st = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    if (st) return st;
    const int est = enqueue();
    st = check_hip(hipStreamEndCapture(stream, &g));
    if (est) {
        if (g) (void)hipGraphDestroy(g);
        return est;
    }
    if (st) return st;
    st = check_hip(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    if (st) return st;
    for (int s = 0; s < nsteps && !st; s++) st = check_hip(hipGraphLaunch(exec, stream));
    const int sst = check_hip(hipStreamSynchronize(stream));
    (void)hipGraphExecDestroy(exec);
    return st ? st : sst;
}

int init_ctrls(hipStream_t stream, int precision, int nprob, const mgad_problem *p, double t0, double tend,
               double cfl) {
    for (int i = 0; i < nprob; i++) {
        const int st = precision == MGAD_F32
                           ? burgers_init_t<float>(stream, p[i].nx, p[i].ny, p[i].nz,
                                                   static_cast<const float *>(p[i].u), p[i].ctrl, t0, tend, cfl)
                           : burgers_init_t<double>(stream, p[i].nx, p[i].ny, p[i].nz,
                                                    static_cast<const double *>(p[i].u), p[i].ctrl, t0, tend, cfl);
        if (st) return st;
    }
    return MGAD_OK;
}

}  // namespace
}  // namespace mgad

using namespace mgad;

extern "C" size_t mgad_batch_workspace_bytes(int nprob, const mgad_problem *probs) {
    if (nprob <= 0 || !probs) return 0;
    return ws_bytes(nprob, count_items(nprob, probs, 4));  // the smallest chunk the tuner tries
}

extern "C" int mgad_batch_init(hipStream_t stream, int precision, const mgad_scheme *scheme, int nprob,
                               const mgad_problem *probs, void *workspace, size_t workspace_bytes, int kchunk,
                               double t0, double tend, double cfl, mgad_batch *out) {
    if (!scheme || !out || !workspace || (precision != MGAD_F64 && precision != MGAD_F32) ||
        !problems_ok(nprob, probs) || kchunk < 0) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    if (!pick_batch<double>(*scheme)) {
        set_error(MGAD_ERR_UNSUPPORTED);
        return MGAD_ERR_UNSUPPORTED;
    }
    mgad_batch b{};
    b.workspace = workspace;
    b.workspace_bytes = workspace_bytes;
    b.nprob = nprob;
    b.precision = precision;
    b.scheme = *scheme;
    int st = init_ctrls(stream, precision, nprob, probs, t0, 1e300, cfl);
    if (st) return st;
    std::vector<int> cands;
    if (kchunk > 0)
        cands.push_back(kchunk);
    else
        cands = {8, 16, 32};
    float best_ms = 1e30f;
    int best_kc = cands[0];
    if (cands.size() > 1) {
        // autotune: time the batched stage 1 (u -> w1) for each z chunk depth
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if ((st = check_hip(hipEventCreate(&e0))) || (st = check_hip(hipEventCreate(&e1)))) return st;
        for (int kc : cands) {
            if (ws_bytes(nprob, count_items(nprob, probs, kc)) > workspace_bytes) continue;
            if ((st = write_workspace(stream, nprob, probs, kc, workspace, workspace_bytes, &b.nitems))) break;
            b.kchunk = kc;
            b.grid = std::max(1, std::min(b.nitems, resident_blocks()));
            st = batch_stage(stream, b, 1);
            if (!st) st = check_hip(hipEventRecord(e0, stream));
            for (int r = 0; r < 3 && !st; r++) st = batch_stage(stream, b, 1);
            if (!st) st = check_hip(hipEventRecord(e1, stream));
            if (!st) st = check_hip(hipEventSynchronize(e1));
            float ms = 0.f;
            if (!st) st = check_hip(hipEventElapsedTime(&ms, e0, e1));
            if (st) break;
            if (ms < best_ms) {
                best_ms = ms;
                best_kc = kc;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (st) return st;
    }
    if ((st = write_workspace(stream, nprob, probs, best_kc, workspace, workspace_bytes, &b.nitems))) return st;
    b.kchunk = best_kc;
    b.grid = std::max(1, std::min(b.nitems, resident_blocks()));
    if ((st = init_ctrls(stream, precision, nprob, probs, t0, tend, cfl))) return st;
    *out = b;
    return MGAD_OK;
}

extern "C" int mgad_batch_steps(hipStream_t stream, const mgad_batch *batch, int nsteps, int use_graph) {
    if (!batch || !batch->workspace || nsteps < 0) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    return run_graph(stream, nsteps, use_graph, [&] { return enqueue_batch_step(stream, *batch); });
}

extern "C" int mgad_batch_steps_serial(hipStream_t stream, const mgad_batch *batch, const mgad_problem *probs,
                                       int nsteps, int use_graph) {
    if (!batch || nsteps < 0 || !problems_ok(batch->nprob, probs)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    return run_graph(stream, nsteps, use_graph, [&] { return enqueue_serial_step(stream, *batch, probs); });
}
