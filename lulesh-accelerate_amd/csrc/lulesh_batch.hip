// lulesh_batch.hip -- ensembles of independent domains (the lulesh_batch_*
// entry points of include/lulesh_mi355x.h).
//
// All members live in one device arena.  A cycle is five launches whatever
// the number of members: TimeIncrement of every member (one thread each),
// the force sweep over the tile list of all members, kinematics and
// Q / EOS / dt over the element-chunk lists of all members (one block per
// tile or chunk, or a persistent grid of fewer blocks looping over the
// list: a launch-shape choice), and one closing kernel that applies the last
// minima and counts the members with cycles left.  Members run their own cycle loops:
// a member whose time reached its stoptime (or that failed) takes no part;
// the host drops its tiles and chunks from the lists when it notices.
//
// The cycles are replayed from a captured HIP graph (cycles_per_graph
// cycles per graph), so the host issues one launch per graph.
#include "lulesh_mi355x.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "lulesh_kernels.hpp"
#include "lulesh_mesh.hpp"
#include "lulesh_setup.hpp"

using namespace lulesh;

namespace {

// the autotuner's pick on the 64-member ensemble of nx 16..64 (16 x 8
// columns, 16 planes, one block per tile and per chunk): 2.84 s to t = 1e-2
// against 2.97 s for 32 x 8, measured on MI355X
constexpr SweepShape kDefaultBatchSweep = {16, 8, 16, 2};
constexpr int kDefaultCyclesPerGraph = 16;
constexpr int kTimeBlock = 256;

size_t round_up(size_t n, size_t m) { return (n + m - 1) / m * m; }

// splitmix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): the project's seeded stream for ensembles
uint64_t splitmix64(uint64_t &state)
{
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

struct lulesh_batch {
    int n = 0;
    int ic_mode = 2;
    std::vector<lulesh_config> cfg;
    std::vector<DevFields> f;  // host copies of the members' field tables
    PhysParams prm{};

    void *arena = nullptr;
    size_t arena_bytes = 0;
    DevScalars *d_scalars = nullptr;  // n contiguous members' scalars
    BatchMember *d_members = nullptr;
    BatchWork *d_work = nullptr;
    int4 *d_tiles = nullptr;
    int2 *d_chunks = nullptr;
    int2 *d_eos_chunks = nullptr;
    int *h_pending = nullptr;  // pinned
    std::vector<int4> tiles;
    std::vector<int2> chunks;
    std::vector<int2> eos_chunks;
    std::vector<std::vector<int64_t>> eos_passes;  // EOS passes of each member's chunks
    std::vector<char> dropped;  // members no longer in the work lists
    int full_tiles = 0, full_chunks = 0;

    SweepShape sweep = kDefaultBatchSweep;
    int sweep_blocks = 0, elem_blocks = 0;
    bool fused = false;  // the phases of lulesh::fused (lulesh_batch_set_fused_multiply_add)
    int cycles_per_graph = kDefaultCyclesPerGraph;
    int cu_count = 256;

    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_stale = true;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    std::string err;
};

namespace {

int bfail(lulesh_batch *b, int code, const std::string &msg)
{
    if (b) {
        b->err = msg;
    }
    return code;
}

#define BHIP_TRY(b, expr)                                                    \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            return bfail((b), LULESH_HIP_ERROR,                              \
                         std::string(#expr ": ") + hipGetErrorString(e_));   \
        }                                                                    \
    } while (0)

// The Q/EOS chunks of the members still in the lists, in dispatch order.  A
// chunk is one of its member's Q/EOS blocks (kChunk = kEosBlockElems), and
// the blocks' costs are as uneven over the ensemble as over one domain, so
// the whole list is ordered as the single domain orders its blocks
// (spread_costly_blocks): the costly chunks of all members spread over each
// XCD's first slots, the rest member by member in mesh order.
// The list has as many entries as the chunk list, whose size the device
// copy of both is allocated for.
static_assert(kChunk == kEosBlockElems, "a Q/EOS chunk is one Q/EOS block");
void order_eos_chunks(lulesh_batch *b, bool all)
{
    std::vector<int2> base;
    std::vector<int64_t> passes;
    for (int m = 0; m < b->n; ++m) {
        if (!all && b->dropped[m]) {
            continue;
        }
        const std::vector<int64_t> &p = b->eos_passes[size_t(m)];
        for (size_t c = 0; c < p.size(); ++c) {
            base.push_back(make_int2(m, int(c) * kChunk));
            passes.push_back(p[c]);
        }
    }
    const std::vector<int32_t> order = spread_costly_blocks(passes);
    b->eos_chunks.resize(base.size());
    for (size_t t = 0; t < order.size(); ++t) {
        b->eos_chunks[t] = base[size_t(order[t])];
    }
}

// tiles of the sweep (for the current shape) and element chunks of the
// members still in the lists
void cut_work(lulesh_batch *b, bool all)
{
    b->tiles.clear();
    b->chunks.clear();
    const int tx = b->sweep.tx, ty = b->sweep.ty, nz = b->sweep.nz;
    for (int m = 0; m < b->n; ++m) {
        if (!all && b->dropped[m]) {
            continue;
        }
        const int en = b->f[m].nx + 1;
        const int gx = (en + tx - 2) / (tx - 1), gy = (en + ty - 2) / (ty - 1);
        const int gz = (en + nz - 1) / nz;
        for (int kz = 0; kz < gz; ++kz) {
            for (int jy = 0; jy < gy; ++jy) {
                for (int ix = 0; ix < gx; ++ix) {
                    b->tiles.push_back(make_int4(m, ix, jy, kz));
                }
            }
        }
        // member by member, last chunk first
        const int nchunk = (b->f[m].num_elem + kChunk - 1) / kChunk;
        for (int c = nchunk - 1; c >= 0; --c) {
            b->chunks.push_back(make_int2(m, c * kChunk));
        }
    }
    order_eos_chunks(b, all);
}

// uploads the work lists (and their counts) for the members not dropped;
// stream-ordered after every launch queued so far
int upload_work(lulesh_batch *b)
{
    cut_work(b, false);
    BatchWork w{};
    w.num_members = b->n;
    w.num_tiles = int(b->tiles.size());
    w.num_chunks = int(b->chunks.size());
    w.num_pending = 0;
    w.tiles = b->d_tiles;
    w.chunks = b->d_chunks;
    w.eos_chunks = b->d_eos_chunks;
    if (!b->tiles.empty()) {
        BHIP_TRY(b, hipMemcpyAsync(b->d_tiles, b->tiles.data(),
                                   b->tiles.size() * sizeof(int4),
                                   hipMemcpyHostToDevice, b->stream));
    }
    if (!b->chunks.empty()) {
        BHIP_TRY(b, hipMemcpyAsync(b->d_chunks, b->chunks.data(),
                                   b->chunks.size() * sizeof(int2),
                                   hipMemcpyHostToDevice, b->stream));
        BHIP_TRY(b, hipMemcpyAsync(b->d_eos_chunks, b->eos_chunks.data(),
                                   b->eos_chunks.size() * sizeof(int2),
                                   hipMemcpyHostToDevice, b->stream));
    }
    BHIP_TRY(b, hipMemcpyAsync(b->d_work, &w, sizeof(w), hipMemcpyHostToDevice,
                               b->stream));
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    return LULESH_OK;
}

// the full list sizes fix the grids of the captured launches
void size_grids(lulesh_batch *b)
{
    cut_work(b, true);
    b->full_tiles = int(b->tiles.size());
    b->full_chunks = int(b->chunks.size());
}

int sweep_grid(const lulesh_batch *b)
{
    return b->sweep_blocks > 0 ? std::min(b->sweep_blocks, b->full_tiles) : b->full_tiles;
}

int elem_grid(const lulesh_batch *b)
{
    return b->elem_blocks > 0 ? std::min(b->elem_blocks, b->full_chunks) : b->full_chunks;
}

// queues ncycles cycles and the closing kernel on the batch stream; the
// TimeIncrement and dt kernels are the same in both builds, as in the
// single domain's cycle
int queue_cycles(lulesh_batch *b, int ncycles)
{
    hipStream_t st = b->stream;
    const int gs = sweep_grid(b), ge = elem_grid(b);
    const bool fz = b->fused;
    for (int c = 0; c < ncycles; ++c) {
        hipLaunchKernelGGL(k_time_increment_batch, dim3(1), dim3(kTimeBlock), 0, st,
                           b->d_members, b->d_work, b->ic_mode);
        BHIP_TRY(b, hipGetLastError());
        BHIP_TRY(b, fz ? fused::launch_force_sweep_batch(b->sweep, gs, b->d_members,
                                                         b->d_work, b->prm, st)
                       : launch_force_sweep_batch(b->sweep, gs, b->d_members, b->d_work,
                                                  b->prm, st));
        if (b->elem_blocks == 0) {
            // one chunk per block over the full chunk list
            hipLaunchKernelGGL(fz ? fused::k_kinematics_q_chunks : k_kinematics_q_chunks,
                               dim3(ge), dim3(kChunk), 0, st, b->d_members, b->d_work,
                               b->prm);
            BHIP_TRY(b, hipGetLastError());
            hipLaunchKernelGGL(fz ? fused::k_q_eos_dt_chunks : k_q_eos_dt_chunks, dim3(ge),
                               dim3(kChunk), 0, st, b->d_members, b->d_work, b->prm);
        } else {
            hipLaunchKernelGGL(fz ? fused::k_kinematics_q_batch : k_kinematics_q_batch,
                               dim3(ge), dim3(kChunk), 0, st, b->d_members, b->d_work,
                               b->prm);
            BHIP_TRY(b, hipGetLastError());
            hipLaunchKernelGGL(fz ? fused::k_q_eos_dt_batch : k_q_eos_dt_batch, dim3(ge),
                               dim3(kChunk), 0, st, b->d_members, b->d_work, b->prm);
        }
        BHIP_TRY(b, hipGetLastError());
    }
    hipLaunchKernelGGL(k_dt_finish_batch, dim3(1), dim3(kTimeBlock), 0, st, b->d_members,
                       b->d_work, b->ic_mode);
    BHIP_TRY(b, hipGetLastError());
    return LULESH_OK;
}

void drop_graph(lulesh_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_stale = true;
}

int ensure_graph(lulesh_batch *b)
{
    if (!b->graph_stale || b->cycles_per_graph <= 0) {
        return LULESH_OK;
    }
    drop_graph(b);
    BHIP_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = queue_cycles(b, b->cycles_per_graph);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (rc != LULESH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) {
        return bfail(b, LULESH_HIP_ERROR,
                     std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    }
    b->graph = g;
    BHIP_TRY(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_stale = false;
    return LULESH_OK;
}

int read_all_scalars(lulesh_batch *b, std::vector<DevScalars> &out)
{
    out.resize(size_t(b->n));
    BHIP_TRY(b, hipMemcpyAsync(out.data(), b->d_scalars, out.size() * sizeof(DevScalars),
                               hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    return LULESH_OK;
}

bool has_cycles_left(const DevScalars &s)
{
    return s.error == 0 && s.time < s.stoptime && s.cycle < s.max_cycles;
}

// drops the members that have stopped from the work lists
int prune(lulesh_batch *b)
{
    std::vector<DevScalars> s;
    int rc = read_all_scalars(b, s);
    if (rc) return rc;
    bool changed = false;
    for (int m = 0; m < b->n; ++m) {
        if (!b->dropped[m] && !has_cycles_left(s[m])) {
            b->dropped[m] = 1;
            changed = true;
        }
    }
    return changed ? upload_work(b) : LULESH_OK;
}

// runs ncycles cycles (graph replays plus plain launches for the rest) and
// leaves the count of members with cycles left in *b->h_pending
int advance(lulesh_batch *b, int ncycles)
{
    int rc = ensure_graph(b);
    if (rc) return rc;
    BHIP_TRY(b, hipEventRecord(b->ev0, b->stream));
    int left = ncycles;
    if (b->cycles_per_graph > 0) {
        for (; left >= b->cycles_per_graph; left -= b->cycles_per_graph) {
            BHIP_TRY(b, hipGraphLaunch(b->exec, b->stream));
        }
    }
    if (left > 0 || ncycles == 0) {
        if ((rc = queue_cycles(b, left))) return rc;
    }
    BHIP_TRY(b, hipEventRecord(b->ev1, b->stream));
    BHIP_TRY(b, hipMemcpyAsync(b->h_pending, &b->d_work->num_pending, sizeof(int),
                               hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    float ms = 0.f;
    BHIP_TRY(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    return LULESH_OK;
}

int first_error(lulesh_batch *b, int *code)
{
    std::vector<DevScalars> s;
    int rc = read_all_scalars(b, s);
    if (rc) return rc;
    *code = LULESH_OK;
    for (const DevScalars &x : s) {
        if (x.error != 0) {
            *code = x.error;
            break;
        }
    }
    return LULESH_OK;
}

int batch_build(lulesh_batch *b, const lulesh_config *cfg, const double *efactor)
{
    std::vector<HostMesh> meshes(size_t(b->n));
    std::vector<size_t> offset(size_t(b->n));
    size_t total = 0;
    for (int m = 0; m < b->n; ++m) {
        MeshConfig mc;
        mc.nx = cfg[m].nx;
        mc.ic_mode = cfg[m].ic_mode;
        mc.num_reg = cfg[m].num_reg;
        mc.balance = cfg[m].balance;
        mc.cost = cfg[m].cost;
        mc.efactor = efactor ? efactor[m] : 1.0;
        try {
            meshes[m] = build_mesh(mc);
        } catch (const std::exception &ex) {
            return bfail(b, LULESH_BAD_ARG,
                         "member " + std::to_string(m) + ": " + ex.what());
        }
        offset[m] = total;
        total += round_up(domain_arena_bytes(meshes[m]), 256);
    }
    b->arena_bytes = total;
    BHIP_TRY(b, hipMalloc(&b->arena, total));
    BHIP_TRY(b, hipMemsetAsync(b->arena, 0, total, b->stream));
    BHIP_TRY(b, hipMalloc(&b->d_scalars, size_t(b->n) * sizeof(DevScalars)));
    BHIP_TRY(b, hipMalloc(&b->d_members, size_t(b->n) * sizeof(BatchMember)));
    BHIP_TRY(b, hipMalloc(&b->d_work, sizeof(BatchWork)));
    BHIP_TRY(b, hipHostMalloc(&b->h_pending, sizeof(int), hipHostMallocDefault));

    std::vector<BatchMember> members(size_t(b->n));
    std::vector<DevScalars> scalars(size_t(b->n));
    b->f.resize(size_t(b->n));
    b->eos_passes.resize(size_t(b->n));
    for (int m = 0; m < b->n; ++m) {
        DevScalars *unused = nullptr;  // the members' scalars live contiguously
        carve_domain(static_cast<char *>(b->arena) + offset[m], meshes[m], b->f[m],
                     &unused);
        BHIP_TRY(b, upload_mesh(meshes[m], b->f[m], b->stream));
        members[m].f = b->f[m];
        b->eos_passes[size_t(m)] = eos_block_passes(meshes[m].elem_rep);
        members[m].s = b->d_scalars + m;
        scalars[m] = initial_scalars(meshes[m], cfg[m].stoptime, cfg[m].max_cycles);
    }
    BHIP_TRY(b, hipMemcpyAsync(b->d_members, members.data(),
                               members.size() * sizeof(BatchMember),
                               hipMemcpyHostToDevice, b->stream));
    BHIP_TRY(b, hipMemcpyAsync(b->d_scalars, scalars.data(),
                               scalars.size() * sizeof(DevScalars),
                               hipMemcpyHostToDevice, b->stream));
    BHIP_TRY(b, hipStreamSynchronize(b->stream));

    // work lists: room for the largest cut of any supported shape is not
    // known in advance, so size for the finest one (16 x 4 tiles, nz = 1)
    // lazily in set_shape; here for the default shape
    b->dropped.assign(size_t(b->n), 0);
    size_grids(b);
    BHIP_TRY(b, hipMalloc(&b->d_tiles, size_t(std::max(b->full_tiles, 1)) * sizeof(int4)));
    BHIP_TRY(b, hipMalloc(&b->d_chunks, size_t(std::max(b->full_chunks, 1)) * sizeof(int2)));
    BHIP_TRY(b, hipMalloc(&b->d_eos_chunks, size_t(std::max(b->full_chunks, 1)) * sizeof(int2)));
    return upload_work(b);
}

}  // namespace

extern "C" {

void lulesh_ensemble_default(lulesh_ensemble *e)
{
    if (!e) {
        return;
    }
    e->num_members = 64;
    e->nx_min = 16;
    e->nx_max = 64;
    e->efactor_min = 0.8;
    e->efactor_max = 1.2;
    e->seed = 0;
}

int lulesh_ensemble_draw(const lulesh_ensemble *e, int *nx, double *efactor)
{
    if (!e || !nx || !efactor || e->num_members < 1 || e->nx_min < 1 ||
        e->nx_max < e->nx_min || !(e->efactor_min > 0.0) ||
        e->efactor_max < e->efactor_min) {
        return LULESH_BAD_ARG;
    }
    uint64_t state = e->seed;
    const uint64_t span = uint64_t(e->nx_max - e->nx_min + 1);
    for (int m = 0; m < e->num_members; ++m) {
        // the modulo bias is below span / 2^64
        nx[m] = e->nx_min + int(splitmix64(state) % span);
        const double u = double(splitmix64(state) >> 11) * 0x1.0p-53;  // [0, 1)
        efactor[m] = e->efactor_min + (e->efactor_max - e->efactor_min) * u;
    }
    return LULESH_OK;
}

int lulesh_batch_create(const lulesh_config *cfg, const double *efactor, int n,
                        lulesh_batch **out)
{
    if (!cfg || !out || n < 1) {
        return LULESH_BAD_ARG;
    }
    *out = nullptr;
    for (int m = 0; m < n; ++m) {
        if (cfg[m].ic_mode != cfg[0].ic_mode || !(cfg[m].stoptime > 0.0) ||
            cfg[m].max_cycles < 0) {
            return LULESH_BAD_ARG;
        }
    }
    lulesh_batch *b = new (std::nothrow) lulesh_batch();
    if (!b) {
        return LULESH_BAD_ARG;
    }
    b->n = n;
    b->ic_mode = cfg[0].ic_mode;
    b->cfg.assign(cfg, cfg + n);
    b->prm = make_params(b->ic_mode);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        lulesh_batch_destroy(b);
        return LULESH_HIP_ERROR;
    }
    const int rc = batch_build(b, cfg, efactor);
    if (rc != LULESH_OK) {
        lulesh_batch_destroy(b);
        return rc;
    }
    *out = b;
    return LULESH_OK;
}

void lulesh_batch_destroy(lulesh_batch *b)
{
    if (!b) {
        return;
    }
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
    }
    drop_graph(b);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->arena) (void)hipFree(b->arena);
    if (b->d_scalars) (void)hipFree(b->d_scalars);
    if (b->d_members) (void)hipFree(b->d_members);
    if (b->d_work) (void)hipFree(b->d_work);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_chunks) (void)hipFree(b->d_chunks);
    if (b->d_eos_chunks) (void)hipFree(b->d_eos_chunks);
    if (b->h_pending) (void)hipHostFree(b->h_pending);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int lulesh_batch_step(lulesh_batch *b, int ncycles, int *pending)
{
    if (!b || ncycles < 0) {
        return LULESH_BAD_ARG;
    }
    int rc = advance(b, ncycles);
    if (rc) return rc;
    if (pending) {
        *pending = *b->h_pending;
    }
    return prune(b);
}

int lulesh_batch_run(lulesh_batch *b)
{
    if (!b) {
        return LULESH_BAD_ARG;
    }
    const int chunk = b->cycles_per_graph > 0 ? b->cycles_per_graph : 32;
    int last = -1;
    double total_ms = 0.0;
    for (;;) {
        int rc = advance(b, chunk);
        if (rc) return rc;
        total_ms += b->last_ms;
        const int pending = *b->h_pending;
        if (pending == 0) {
            break;
        }
        if (pending != last) {
            if ((rc = prune(b))) return rc;
            last = pending;
        }
    }
    b->last_ms = total_ms;
    int code = LULESH_OK;
    int rc = first_error(b, &code);
    if (rc) return rc;
    if ((rc = prune(b))) return rc;
    return code;
}

int lulesh_batch_get_stats(lulesh_batch *b, int member, lulesh_stats *out)
{
    if (!b || !out || member < 0 || member >= b->n) {
        return LULESH_BAD_ARG;
    }
    DevScalars s{};
    BHIP_TRY(b, hipMemcpyAsync(&s, b->d_scalars + member, sizeof(s),
                               hipMemcpyDeviceToHost, b->stream));
    double e0 = 0.0;
    BHIP_TRY(b, hipMemcpyAsync(&e0, b->f[member].e, sizeof(double),
                               hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    out->time = s.time;
    out->deltatime = s.deltatime;
    out->dtcourant = s.dtcourant;
    out->dthydro = s.dthydro;
    out->e_origin = e0;
    out->cycle = s.cycle;
    out->done = (!(s.time < s.stoptime) || !(s.cycle < s.max_cycles)) ? 1 : 0;
    out->num_elem = b->f[member].num_elem;
    out->num_node = b->f[member].num_node;
    return s.error;
}

int lulesh_batch_get_field(lulesh_batch *b, int member, int field, double *host_out)
{
    if (!b || !host_out || member < 0 || member >= b->n) {
        return LULESH_BAD_ARG;
    }
    DevScalars s{};
    BHIP_TRY(b, hipMemcpyAsync(&s, b->d_scalars + member, sizeof(s),
                               hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    int64_t count = 0;
    const double *p = field_pointer(b->f[member], field, s.cur, &count);
    if (!p) {
        return bfail(b, LULESH_BAD_ARG, "unknown field id");
    }
    BHIP_TRY(b, hipMemcpyAsync(host_out, p, size_t(count) * sizeof(double),
                               hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    return LULESH_OK;
}

int lulesh_batch_set_fused_multiply_add(lulesh_batch *b, int enable)
{
    if (!b) {
        return LULESH_BAD_ARG;
    }
    if (b->fused != (enable != 0)) {
        b->fused = enable != 0;
        drop_graph(b);
    }
    return LULESH_OK;
}

int lulesh_batch_set_shape(lulesh_batch *b, int tx, int ty, int nz, int waves,
                           int sweep_blocks, int elem_blocks, int cycles_per_graph)
{
    if (!b || sweep_blocks < 0 || elem_blocks < 0 || cycles_per_graph < 0 ||
        cycles_per_graph > 1024) {
        return LULESH_BAD_ARG;
    }
    const SweepShape sh = {tx, ty, nz, waves};
    if (!force_sweep_shape_supported(sh)) {
        return bfail(b, LULESH_BAD_ARG, "sweep shape not built into the library");
    }
    const bool recut = sh.tx != b->sweep.tx || sh.ty != b->sweep.ty || sh.nz != b->sweep.nz;
    b->sweep = sh;
    b->sweep_blocks = sweep_blocks;
    b->elem_blocks = elem_blocks;
    b->cycles_per_graph = cycles_per_graph;
    drop_graph(b);
    if (recut) {
        const int old_tiles = b->full_tiles;
        size_grids(b);
        if (b->full_tiles > old_tiles) {
            BHIP_TRY(b, hipStreamSynchronize(b->stream));
            BHIP_TRY(b, hipFree(b->d_tiles));
            b->d_tiles = nullptr;
            BHIP_TRY(b, hipMalloc(&b->d_tiles, size_t(b->full_tiles) * sizeof(int4)));
        }
        return upload_work(b);
    }
    return LULESH_OK;
}

int lulesh_batch_get_shape(lulesh_batch *b, int *tx, int *ty, int *nz, int *waves,
                           int *sweep_blocks, int *elem_blocks, int *cycles_per_graph)
{
    if (!b) {
        return LULESH_BAD_ARG;
    }
    if (tx) *tx = b->sweep.tx;
    if (ty) *ty = b->sweep.ty;
    if (nz) *nz = b->sweep.nz;
    if (waves) *waves = b->sweep.waves;
    if (sweep_blocks) *sweep_blocks = b->sweep_blocks;
    if (elem_blocks) *elem_blocks = b->elem_blocks;
    if (cycles_per_graph) *cycles_per_graph = b->cycles_per_graph;
    return LULESH_OK;
}

int lulesh_batch_autotune(lulesh_batch *b, int trial_cycles, double *best_ms)
{
    if (!b || trial_cycles < 1) {
        return LULESH_BAD_ARG;
    }
    // save the state: the arena and the scalars
    void *save = nullptr;
    DevScalars *save_s = nullptr;
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    BHIP_TRY(b, hipMalloc(&save, b->arena_bytes));
    if (hipMalloc(&save_s, size_t(b->n) * sizeof(DevScalars)) != hipSuccess) {
        (void)hipFree(save);
        return bfail(b, LULESH_HIP_ERROR, "autotune: out of device memory");
    }
    const std::vector<char> dropped = b->dropped;
    auto restore = [&]() -> int {
        BHIP_TRY(b, hipMemcpyAsync(b->arena, save, b->arena_bytes,
                                   hipMemcpyDeviceToDevice, b->stream));
        BHIP_TRY(b, hipMemcpyAsync(b->d_scalars, save_s,
                                   size_t(b->n) * sizeof(DevScalars),
                                   hipMemcpyDeviceToDevice, b->stream));
        BHIP_TRY(b, hipStreamSynchronize(b->stream));
        b->dropped = dropped;
        return upload_work(b);
    };
    int rc = LULESH_OK;
    struct Cand { SweepShape sh; int sb, eb; };
    const Cand start = {b->sweep, b->sweep_blocks, b->elem_blocks};
    Cand best = start;
    double best_t = 1e300;
    // a trial times exactly trial_cycles cycles as plain launches, so no
    // graph is built per candidate; the chosen shape gets the caller's
    // cycles per graph back
    const int graph_cycles = b->cycles_per_graph;
    auto trial = [&](const Cand &c, double *t) -> int {
        int r = lulesh_batch_set_shape(b, c.sh.tx, c.sh.ty, c.sh.nz, c.sh.waves, c.sb,
                                       c.eb, 0);
        if (r) return r;
        if ((r = restore())) return r;
        if ((r = advance(b, 1))) return r;  // warm-up
        if ((r = advance(b, trial_cycles))) return r;
        *t = b->last_ms / trial_cycles;
        return LULESH_OK;
    };
    if (hipMemcpyAsync(save, b->arena, b->arena_bytes, hipMemcpyDeviceToDevice,
                       b->stream) != hipSuccess ||
        hipMemcpyAsync(save_s, b->d_scalars, size_t(b->n) * sizeof(DevScalars),
                       hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        (void)hipFree(save);
        (void)hipFree(save_s);
        return bfail(b, LULESH_HIP_ERROR, "autotune: saving the state failed");
    }

    // 1) tile shape and planes per block, one block per tile
    const SweepShape shapes[] = {{32, 8, 0, 2}, {16, 16, 0, 2}, {16, 8, 0, 2},
                                 {32, 4, 0, 2}};
    const int depths[] = {8, 16, 32};
    for (const SweepShape &s0 : shapes) {
        for (int nz : depths) {
            Cand c = {s0, 0, start.eb};
            c.sh.nz = nz;
            double t = 0.0;
            if ((rc = trial(c, &t))) break;
            if (t < best_t) {
                best_t = t;
                best = c;
            }
        }
        if (rc) break;
    }
    // 2) persistent grid of the sweep, 3) of the element kernels
    if (!rc) {
        const int mult[] = {1, 2, 4, 8};
        for (int k : mult) {
            Cand c = best;
            c.sb = k * b->cu_count;
            double t = 0.0;
            if ((rc = trial(c, &t))) break;
            if (t < best_t) {
                best_t = t;
                best = c;
            }
        }
    }
    if (!rc) {
        const int mult[] = {0, 2, 4, 8, 16};
        const Cand base = best;
        for (int k : mult) {
            Cand c = base;
            c.eb = k * b->cu_count;
            if (c.eb == base.eb) continue;
            double t = 0.0;
            if ((rc = trial(c, &t))) break;
            if (t < best_t) {
                best_t = t;
                best = c;
            }
        }
    }
    int rc2 = lulesh_batch_set_shape(b, best.sh.tx, best.sh.ty, best.sh.nz,
                                     best.sh.waves, best.sb, best.eb, graph_cycles);
    if (!rc2) rc2 = restore();
    (void)hipFree(save);
    (void)hipFree(save_s);
    if (rc) return rc;
    if (rc2) return rc2;
    if (best_ms) *best_ms = best_t;
    return LULESH_OK;
}

double lulesh_batch_last_ms(lulesh_batch *b) { return b ? b->last_ms : 0.0; }

const char *lulesh_batch_last_error(const lulesh_batch *b)
{
    return b ? b->err.c_str() : "null batch";
}

}  // extern "C"
