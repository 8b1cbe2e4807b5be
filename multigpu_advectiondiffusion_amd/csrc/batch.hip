// batch.hip -- many independent periodic Burgers problems of uneven sizes
// advanced together: one persistent launch per SSP-RK3 stage.
//
// Two work lists, chosen by timing (mgad_batch_init):
//  - chunked: every item is a 16x16 column tile by kchunk planes of one
//    problem, so items cost about the same whatever the problem size; a grid
//    of as many 256-thread blocks as the device keeps resident walks the list
//    (block b takes items b, b+G, ...);
//  - split: every item is a whole column tile of one problem, and the items'
//    planes, in list order, are cut into shares that shrink towards the end
//    of the list; the resident blocks draw shares from a counter until none
//    is left, so a block that runs faster takes more of them and the blocks
//    end together.
// Chunked lists remap block ids so that the blocks of one XCD take adjacent
// items.
// Each piece of an item runs the same stage body as the single-problem
// kernel, with the problem's own dt, LF alpha and max|u| slot (its own ctrl),
// so results are bit-identical to advancing each problem alone.
//
// Workspace (device memory given by the caller):
//   BatchProblem[nprob] | item_prob[nitems] | item_tile[nitems]
//   | ShareQueue | shares[nshares + 1]   (split lists; 8-byte aligned)
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

// Hand-out state of a split list's shares.  A launch of G blocks adds exactly
// period = nshares + G to `next`: each block draws shares until it draws one
// past the last, and then stops.  Launches on one stream run one after the
// other, so `next % period` numbers the current launch's draws from 0, and
// no reset is needed between launches or graph replays.  Written as zero with
// the work list.
struct ShareQueue {
    unsigned long long next;
    unsigned nshares, period;
};

// One piece of a work item: planes z0 .. z1-1 of the tile's columns.  The
// piece holding the problem's first column and plane 0 leads its stage
// control.
// Tiles are TX x 16 columns; a 32-wide tile row of a problem with nx % 32 == 16
// ends in a tile that reaches past nx (its last 16 columns are not stored).
template <class T, int WENO, int P, int TX>
__device__ __forceinline__ void batch_piece(const BatchProblem &bp, unsigned tile, int z0, int z1, double eps,
                                            int stage, TileSmem<TX, kBT> &sm) {
    constexpr bool XW = TX != kBT;
    double dt, alpha;
    if (!stage_control(bp.ctrl, stage, bp.g.dx, tile == 0 && z0 == 0 && threadIdx.x == 0, dt, alpha)) return;
    const T *u = static_cast<const T *>(bp.u);
    const T *uin = stage == 1 ? u : (stage == 2 ? static_cast<const T *>(bp.w1) : static_cast<const T *>(bp.w2));
    T *uout = stage == 1 ? static_cast<T *>(bp.w1) : (stage == 2 ? static_cast<T *>(bp.w2) : static_cast<T *>(bp.u));
    const int ntx = (bp.g.nx + TX - 1) / TX;
    const int x0 = static_cast<int>(tile % ntx) * TX, y0 = static_cast<int>(tile / ntx) * kBT;
    if (fast_division_ok<WENO>(alpha, eps))
        stage_cols<T, WENO, P, TX, kBT, true, false, XW, true>(uin, u, uout, bp.g, stage, x0, y0, z0, z1, dt, alpha, eps,
                                                         sm, &bp.ctrl->amax[stage % 3], nullptr, nullptr);
    else
        stage_cols<T, WENO, P, TX, kBT, false, false, XW, true>(uin, u, uout, bp.g, stage, x0, y0, z0, z1, dt, alpha, eps,
                                                          sm, &bp.ctrl->amax[stage % 3], nullptr, nullptr);
}

// Chunked list: item i is tile item_tile[i] of problem item_prob[i], tiles
// x-fastest, then y, then z chunks of kchunk planes; block b takes items
// b, b+G, ... (b remapped so that the blocks of one XCD take adjacent items).
template <class T, int WENO, int P>
__global__ __launch_bounds__(kBT *kBT, 4) void k_burgers_batch(const BatchProblem *__restrict__ probs,
                                                               const unsigned *__restrict__ item_prob,
                                                               const unsigned *__restrict__ item_tile,
                                                               int nitems, double eps, int stage) {
    __shared__ TileSmem<kBT, kBT> sm;
    const unsigned G = gridDim.x;
    for (unsigned i = xcd_tile(blockIdx.x, G); i < static_cast<unsigned>(nitems); i += G) {
        const BatchProblem &bp = probs[item_prob[i]];
        const unsigned ntile = static_cast<unsigned>(bp.g.nx / kBT) * static_cast<unsigned>(bp.g.ny / kBT);
        const unsigned tile = item_tile[i] % ntile;
        const int z0 = static_cast<int>(item_tile[i] / ntile) * bp.g.kchunk;
        batch_piece<T, WENO, P, kBT>(bp, tile, z0, min(z0 + bp.g.kchunk, bp.g.nz), eps, stage, sm);
    }
}

// Split list: item i is a whole column tile (all nz planes) and the items'
// planes, taken in list order, are cut into q->nshares shares; share s runs
// from shares[s] to shares[s + 1] (item, plane) pairs, shares[nshares] =
// (nitems, 0).  Blocks draw share numbers from q (see ShareQueue) until the
// list is used up.  The first-dispatched block on a CU issues ahead of the
// second, so blocks march at different speeds; drawn shares let the faster
// ones take more.  Tiles are TX x 16 (TX = 16 or 32) columns.
template <class T, int WENO, int P, int TX>
__global__ __launch_bounds__(TX *kBT, 4) void k_burgers_batch_split(const BatchProblem *__restrict__ probs,
                                                                    const unsigned *__restrict__ item_prob,
                                                                    const unsigned *__restrict__ item_tile,
                                                                    ShareQueue *q, const uint2 *__restrict__ shares,
                                                                    int nitems, double eps, int stage) {
    __shared__ TileSmem<TX, kBT> sm;
    __shared__ unsigned long long drawn;
    const unsigned nshares = q->nshares;
    const unsigned long long period = q->period;
    for (;;) {
        const unsigned s = static_cast<unsigned>(block_draw(&q->next, drawn) % period);
        if (s >= nshares) break;
        const uint2 a = shares[s], e = shares[s + 1];
        for (unsigned i = a.x; i <= e.x && i < static_cast<unsigned>(nitems); i++) {
            const BatchProblem &bp = probs[item_prob[i]];
            const int z0 = i == a.x ? static_cast<int>(a.y) : 0;
            const int z1 = i == e.x ? static_cast<int>(e.y) : bp.g.nz;
            if (z0 < z1) batch_piece<T, WENO, P, TX>(bp, item_tile[i], z0, z1, eps, stage, sm);
        }
    }
}

namespace {

template <class T>
using BatchKernel = void (*)(const BatchProblem *, const unsigned *, const unsigned *, int, double, int);
template <class T>
using SplitKernel = void (*)(const BatchProblem *, const unsigned *, const unsigned *, ShareQueue *, const uint2 *,
                             int, double, int);

template <class T>
BatchKernel<T> pick_batch(const mgad_scheme &sc) {
    if (sc.weno == MGAD_WENO_Z && sc.p == 1) return k_burgers_batch<T, kWenoZ, 1>;
    if (sc.weno == MGAD_WENO_Z && sc.p == 2) return k_burgers_batch<T, kWenoZ, 2>;
    if (sc.weno == MGAD_WENO_JS) return k_burgers_batch<T, kWenoJS, 2>;
    return nullptr;
}

template <class T, int TX>
SplitKernel<T> pick_split(const mgad_scheme &sc) {
    if (sc.weno == MGAD_WENO_Z && sc.p == 1) return k_burgers_batch_split<T, kWenoZ, 1, TX>;
    if (sc.weno == MGAD_WENO_Z && sc.p == 2) return k_burgers_batch_split<T, kWenoZ, 2, TX>;
    if (sc.weno == MGAD_WENO_JS) return k_burgers_batch_split<T, kWenoJS, 2, TX>;
    return nullptr;
}

bool is_split(int kchunk) { return kchunk == MGAD_BATCH_SPLIT || kchunk == MGAD_BATCH_SPLIT_WIDE; }
// tile width of a work list: 32 for the wide split list, else 16
int tile_w(int kchunk) { return kchunk == MGAD_BATCH_SPLIT_WIDE ? 2 * kBT : kBT; }
// column tiles of one problem in a list of tile width tw
long col_tiles(const mgad_problem &p, int tw) {
    return static_cast<long>((p.nx + tw - 1) / tw) * (p.ny / kBT);
}

// Byte offset of the ShareQueue of a split list (after the items, 8-aligned).
size_t queue_offset(int nprob, long nitems) {
    const size_t o = sizeof(BatchProblem) * nprob + 2 * sizeof(unsigned) * static_cast<size_t>(nitems);
    return (o + 7) & ~static_cast<size_t>(7);
}

// descriptors | item_prob[nitems] | item_tile[nitems] | ShareQueue | shares[nshares + 1]
// (the queue and shares for a split list only)
size_t ws_bytes(int nprob, long nitems, long nshares) {
    if (nshares <= 0) return sizeof(BatchProblem) * nprob + 2 * sizeof(unsigned) * static_cast<size_t>(nitems);
    return queue_offset(nprob, nitems) + sizeof(ShareQueue) + sizeof(uint2) * static_cast<size_t>(nshares + 1);
}

// Items of the chunked list with kchunk planes per item, or of a split list
// (one per column tile).
long count_items(int nprob, const mgad_problem *p, int kchunk) {
    long n = 0;
    for (int i = 0; i < nprob; i++)
        n += col_tiles(p[i], tile_w(kchunk)) * (is_split(kchunk) ? 1 : (p[i].nz + kchunk - 1) / kchunk);
    return n;
}

// Blocks of tw x 16 threads the device keeps resident at the kernels' 4
// waves per SIMD: 16 waves per CU.
int resident_blocks(int tw) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    return cus * 16 / (tw * kBT / 64);
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

// Share sizes of a split list: with R tile planes not yet cut, the next share
// takes max(kShareMin, ceil(R / (kShareDiv * grid))) of them (the last share
// what is left).  The first round of draws covers 1/kShareDiv of the work in
// big shares, which keeps the per-share set-up (stage control, the z queue's
// warm-up, the block max) rare; the shares then shrink geometrically, and the
// last ones are kShareMin planes, so the blocks end within a few planes of
// each other.
constexpr long kShareDiv = 2;
constexpr long kShareMin = 8;

// Global first plane of every share of a split list of `total` tile planes.
std::vector<long> share_starts(long total, long grid) {
    std::vector<long> st;
    for (long pos = 0; pos < total;) {
        st.push_back(pos);
        const long r = total - pos;
        pos += std::min(r, std::max(kShareMin, (r + kShareDiv * grid - 1) / (kShareDiv * grid)));
    }
    return st;
}

// Writes the shares (starts as global planes of the items' planes in list
// order; item k has planes[k] planes) as the (item, plane) pairs they fall in,
// and the end marker (nitems, 0) after them.
void cut_shares(const std::vector<unsigned> &planes, const std::vector<long> &starts, uint2 *shares) {
    size_t k = 0;
    long before = 0;  // planes of the items ahead of item k
    for (size_t s = 0; s < starts.size(); s++) {
        while (k < planes.size() && before + static_cast<long>(planes[k]) <= starts[s]) before += planes[k++];
        shares[s] = make_uint2(static_cast<unsigned>(k), static_cast<unsigned>(starts[s] - before));
    }
    shares[starts.size()] = make_uint2(static_cast<unsigned>(planes.size()), 0u);
}

// Tile planes of a split list (every item a whole column of nz planes).
long split_planes(int nprob, const mgad_problem *p, int kchunk) {
    long work = 0;
    for (int i = 0; i < nprob; i++) work += col_tiles(p[i], tile_w(kchunk)) * p[i].nz;
    return work;
}

// Persistent blocks of a list's stage launch: the resident blocks, or fewer
// when the list has less work (tile planes of a split list, items of a
// chunked one).
long list_grid(int nprob, const mgad_problem *p, int kchunk) {
    const long work = is_split(kchunk) ? split_planes(nprob, p, kchunk) : count_items(nprob, p, kchunk);
    return std::max(1l, std::min(work, static_cast<long>(resident_blocks(tile_w(kchunk)))));
}

// Shares of a split list (0 for a chunked one).
long list_shares(int nprob, const mgad_problem *p, int kchunk) {
    if (!is_split(kchunk)) return 0;
    return static_cast<long>(share_starts(split_planes(nprob, p, kchunk), list_grid(nprob, p, kchunk)).size());
}

// Workspace bytes of one list.
size_t list_bytes(int nprob, const mgad_problem *p, int kchunk) {
    return ws_bytes(nprob, count_items(nprob, p, kchunk), list_shares(nprob, p, kchunk));
}

// Writes descriptors and the work list (problems in the given order, tiles
// x-fastest, then y, then z chunk; one item per column tile and the shares for
// the split list) into the workspace, and sets nitems, kchunk and grid.
int write_workspace(hipStream_t stream, int nprob, const mgad_problem *p, int kchunk, void *ws, size_t ws_size,
                    mgad_batch &b) {
    const long nitems = count_items(nprob, p, kchunk);
    const bool split = is_split(kchunk);
    const int tw = tile_w(kchunk);
    const long grid = list_grid(nprob, p, kchunk);
    const std::vector<long> starts = split ? share_starts(split_planes(nprob, p, kchunk), grid) : std::vector<long>();
    const long nshares = static_cast<long>(starts.size());
    if (ws_bytes(nprob, nitems, nshares) > ws_size || nitems >= (1l << 31) || nshares >= (1l << 31) - grid) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    std::vector<unsigned char> host(ws_bytes(nprob, nitems, nshares));
    BatchProblem *bp = reinterpret_cast<BatchProblem *>(host.data());
    unsigned *ip = reinterpret_cast<unsigned *>(bp + nprob);
    unsigned *it = ip + nitems;
    std::vector<unsigned> planes;
    long k = 0;
    for (int i = 0; i < nprob; i++) {
        bp[i].u = p[i].u;
        bp[i].w1 = p[i].w1;
        bp[i].w2 = p[i].w2;
        bp[i].ctrl = static_cast<Ctrl *>(p[i].ctrl);
        bp[i].g = make_grid(p[i].nx, p[i].ny, p[i].nz, p[i].dx, p[i].dy, p[i].dz,
                            split ? static_cast<int>(p[i].nz) : kchunk);
        const long nt = col_tiles(p[i], tw) * (split ? 1 : (p[i].nz + kchunk - 1) / kchunk);
        for (long t = 0; t < nt; t++, k++) {
            ip[k] = static_cast<unsigned>(i);
            it[k] = static_cast<unsigned>(t);
            if (split) planes.push_back(p[i].nz);
        }
    }
    if (split) {
        ShareQueue *q = reinterpret_cast<ShareQueue *>(host.data() + queue_offset(nprob, nitems));
        q->next = 0ull;
        q->nshares = static_cast<unsigned>(nshares);
        q->period = static_cast<unsigned>(nshares + grid);
        cut_shares(planes, starts, reinterpret_cast<uint2 *>(q + 1));
    }
    int st = check_hip(hipMemcpyAsync(ws, host.data(), host.size(), hipMemcpyHostToDevice, stream));
    if (!st) st = check_hip(hipStreamSynchronize(stream));
    b.nitems = static_cast<int>(nitems);
    b.kchunk = kchunk;
    b.grid = static_cast<int>(grid);
    return st;
}

template <class T>
int launch_batch_stage(hipStream_t stream, const mgad_batch &b, int stage) {
    const BatchProblem *probs = static_cast<const BatchProblem *>(b.workspace);
    const unsigned *ip = reinterpret_cast<const unsigned *>(probs + b.nprob);
    const unsigned *it = ip + b.nitems;
    if (is_split(b.kchunk)) {
        const int tw = tile_w(b.kchunk);
        SplitKernel<T> fn = tw == kBT ? pick_split<T, kBT>(b.scheme) : pick_split<T, 2 * kBT>(b.scheme);
        if (!fn) {
            set_error(MGAD_ERR_UNSUPPORTED);
            return MGAD_ERR_UNSUPPORTED;
        }
        ShareQueue *q = reinterpret_cast<ShareQueue *>(static_cast<unsigned char *>(b.workspace) +
                                                       queue_offset(b.nprob, b.nitems));
        hipLaunchKernelGGL(fn, dim3(b.grid), dim3(tw * kBT), 0, stream, probs, ip, it, q,
                           reinterpret_cast<const uint2 *>(q + 1), b.nitems, b.scheme.eps, stage);
    } else {
        BatchKernel<T> fn = pick_batch<T>(b.scheme);
        if (!fn) {
            set_error(MGAD_ERR_UNSUPPORTED);
            return MGAD_ERR_UNSUPPORTED;
        }
        hipLaunchKernelGGL(fn, dim3(b.grid), dim3(kBT * kBT), 0, stream, probs, ip, it, b.nitems, b.scheme.eps,
                           stage);
    }
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

// The work list of mgad_batch_init_static: one array per GPU, held from
// library load until the process ends.  Host code never passes its name to a
// kernel or a copy (on the host it is only a stand-in for the symbol): it uses
// the device address scratch_ptr() looked up for the current device.
static __device__ __align__(256) unsigned char scratch[MGAD_BATCH_SCRATCH_BYTES];

namespace mgad {
namespace {

constexpr size_t kBatchScratchBytes = MGAD_BATCH_SCRATCH_BYTES;

// Device address of `scratch` on the current device, looked up once per
// device with that device current (each GPU holds its own copy of the array).
void *scratch_ptr() {
    constexpr int kMaxDevices = 64;
    static void *ptrs[kMaxDevices] = {};
    std::lock_guard<std::recursive_mutex> lock(host_mutex());
    int dev = -1;
    if (check_hip(hipGetDevice(&dev))) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error(MGAD_ERR_UNSUPPORTED);
        return nullptr;
    }
    if (!ptrs[dev]) {
        void *p = nullptr;
        if (check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)))) return nullptr;
        ptrs[dev] = p;
    }
    return ptrs[dev];
}

}  // namespace
}  // namespace mgad

using namespace mgad;

extern "C" size_t mgad_batch_workspace_bytes(int nprob, const mgad_problem *probs) {
    if (nprob <= 0 || !probs) return 0;
    // the chunked list of the smallest chunk the tuner tries holds at least
    // as many items as the other chunked lists; the split lists add shares
    size_t n = list_bytes(nprob, probs, 4);
    for (int kc : {MGAD_BATCH_SPLIT, MGAD_BATCH_SPLIT_WIDE}) n = std::max(n, list_bytes(nprob, probs, kc));
    return n;
}

namespace {

// mgad_batch_init on a given workspace: kchunk != 0 writes that list, 0
// autotunes among the five lists.
int batch_init_ws(hipStream_t stream, int precision, const mgad_scheme *scheme, int nprob,
                  const mgad_problem *probs, void *workspace, size_t workspace_bytes, int kchunk, double t0,
                  double tend, double cfl, mgad_batch *out) {
    mgad_batch b{};
    b.workspace = workspace;
    b.workspace_bytes = workspace_bytes;
    b.nprob = nprob;
    b.precision = precision;
    b.scheme = *scheme;
    int st = MGAD_OK;
    std::vector<int> cands;
    if (kchunk != 0)
        cands.push_back(kchunk);
    else
        cands = {MGAD_BATCH_SPLIT_WIDE, MGAD_BATCH_SPLIT, 8, 16, 32};
    int best_kc = cands[0];
    if (cands.size() > 1) {
        // autotune: time stages 1 and 2 of a step (u -> w1 -> w2; they leave
        // every u as given, and each pair starts from freshly initialised
        // ctrls) for each work list, in two passes over the candidates, and
        // keep the fastest by its better pass
        std::vector<float> ms_best(cands.size(), 1e30f);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if ((st = check_hip(hipEventCreate(&e0))) || (st = check_hip(hipEventCreate(&e1)))) return st;
        for (int pass = 0; pass < 2 && !st; pass++) {
            for (size_t c = 0; c < cands.size(); c++) {
                const int kc = cands[c];
                if (list_bytes(nprob, probs, kc) > workspace_bytes) continue;
                if ((st = write_workspace(stream, nprob, probs, kc, workspace, workspace_bytes, b))) break;
                st = init_ctrls(stream, precision, nprob, probs, t0, 1e300, cfl);
                if (!st) st = batch_stage(stream, b, 1);  // warm-up
                if (!st) st = check_hip(hipEventRecord(e0, stream));
                if (!st) st = batch_stage(stream, b, 1);
                if (!st) st = batch_stage(stream, b, 2);
                if (!st) st = check_hip(hipEventRecord(e1, stream));
                if (!st) st = check_hip(hipEventSynchronize(e1));
                float ms = 0.f;
                if (!st) st = check_hip(hipEventElapsedTime(&ms, e0, e1));
                if (st) break;
                ms_best[c] = std::min(ms_best[c], ms);
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (st) return st;
        const size_t c = std::min_element(ms_best.begin(), ms_best.end()) - ms_best.begin();
        if (!(ms_best[c] < 1e30f)) {  // no list fits the workspace
            set_error(MGAD_ERR_INVALID_ARG);
            return MGAD_ERR_INVALID_ARG;
        }
        best_kc = cands[c];
    }
    if ((st = write_workspace(stream, nprob, probs, best_kc, workspace, workspace_bytes, b))) return st;
    if ((st = init_ctrls(stream, precision, nprob, probs, t0, tend, cfl))) return st;
    *out = b;
    return MGAD_OK;
}

bool batch_args_ok(int precision, const mgad_scheme *scheme, int nprob, const mgad_problem *probs, int kchunk,
                   const mgad_batch *out) {
    return scheme && out && (precision == MGAD_F64 || precision == MGAD_F32) && problems_ok(nprob, probs) &&
           (kchunk >= 0 || is_split(kchunk));
}

}  // namespace

extern "C" int mgad_batch_init(hipStream_t stream, int precision, const mgad_scheme *scheme, int nprob,
                               const mgad_problem *probs, void *workspace, size_t workspace_bytes, int kchunk,
                               double t0, double tend, double cfl, mgad_batch *out) {
    if (!workspace || !batch_args_ok(precision, scheme, nprob, probs, kchunk, out)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    if (!pick_batch<double>(*scheme)) {
        set_error(MGAD_ERR_UNSUPPORTED);
        return MGAD_ERR_UNSUPPORTED;
    }
    return batch_init_ws(stream, precision, scheme, nprob, probs, workspace, workspace_bytes, kchunk, t0, tend, cfl,
                         out);
}

extern "C" int mgad_batch_init_static(hipStream_t stream, int precision, const mgad_scheme *scheme, int nprob,
                                      const mgad_problem *probs, int kchunk, double t0, double tend, double cfl,
                                      mgad_batch *out) {
    if (kchunk == 0 || !batch_args_ok(precision, scheme, nprob, probs, kchunk, out)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    if (!pick_batch<double>(*scheme)) {
        set_error(MGAD_ERR_UNSUPPORTED);
        return MGAD_ERR_UNSUPPORTED;
    }
    if (list_bytes(nprob, probs, kchunk) > kBatchScratchBytes) {
        set_error(MGAD_ERR_SCRATCH_FULL);
        return MGAD_ERR_SCRATCH_FULL;
    }
    void *ws = scratch_ptr();
    if (!ws) return MGAD_ERR_HIP;
    return batch_init_ws(stream, precision, scheme, nprob, probs, ws, kBatchScratchBytes, kchunk, t0, tend, cfl,
                         out);
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
