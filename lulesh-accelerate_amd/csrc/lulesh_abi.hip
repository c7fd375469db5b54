// lulesh_abi.hip -- the C ABI of include/lulesh_mi355x.h.
//
// Owns the device state of one domain, builds it from the host mesh, queues
// cycles of kernels on the domain's stream and copies fields in and out.
#include "lulesh_mi355x.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "lulesh_kernels.hpp"
#include "lulesh_mesh.hpp"
#include "lulesh_mlp.hpp"
#include "lulesh_setup.hpp"

using namespace lulesh;

namespace {

enum KernelId {
    KID_TIME_INCREMENT = 0,
    KID_FORCE_SWEEP,
    KID_KINEMATICS_Q,
    KID_Q_EOS_DT,
    KID_DT_FINISH,
    KID_FORCES_U,         // unstructured meshes: element blocks and their interior nodes
    KID_NODES_U,          //   and the boundary nodes
    KID_COUNT
};

const char *const kKernelNames[KID_COUNT] = {
    "k_time_increment", "k_force_sweep", "k_kinematics_q", "k_q_eos_dt",
    "k_dt_finish", "k_forces_u_blocks", "k_nodes_u",
};

// default launch shape of the force sweep (32 x 8 element columns, at most
// 17 node planes per block, register budget for 2 waves per SIMD; the planes
// per block shrink on small domains, see default_sweep_planes);
// lulesh_set_sweep_shape overrides it
constexpr SweepShape kDefaultSweep = {32, 8, 17, 2};
// the mixed-precision sweep holds float32 corner forces: three blocks fit
// on a CU
constexpr SweepShape kDefaultSweepMixed = {32, 8, 12, 3};

constexpr int kBlock = 256;
static_assert(kBlock == kEosBlockElems, "the Q/EOS block order is cut for this block size");
constexpr int kBatch = 32;  // cycles queued per host round trip in run()
// cycles per captured graph: lulesh_step replays a graph of this many cycles
// as often as it fits and queues the rest one launch at a time
constexpr int kGraphCycles = 32;

// Node planes per block of the default sweep on an nx^3 domain: the largest
// nz <= 17 whose grid still fills 90% of the resident block slots (two
// 256-thread blocks per CU).  A block sweeps its planes one after another,
// so on small domains parallelism comes from cutting the columns into
// short runs: at 30^3 nz = 1 takes the cycle from 138 to 45 us, at 64^3
// nz = 4 from 162 to 95 us; from 120^3 on nz = 17 is kept (at 150^3 it
// runs 0.4% faster than 16: 990 blocks instead of 1100; measured on
// MI355X).
int default_sweep_planes(int nx)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess ||
        cus < 1) {
        cus = 256;
    }
    const long long want = (long long)(0.9 * 2 * cus);
    const int en = nx + 1;
    const long long cols = (long long)((en + kDefaultSweep.tx - 2) / (kDefaultSweep.tx - 1)) *
                           ((en + kDefaultSweep.ty - 2) / (kDefaultSweep.ty - 1));
    for (int nz = kDefaultSweep.nz; nz > 1; --nz) {
        if (cols * ((en + nz - 1) / nz) >= want) {
            return nz;
        }
    }
    return 1;
}

struct TimedLaunch {
    int kernel;
    hipEvent_t start, stop;
};

}  // namespace

struct lulesh_domain {
    lulesh_config cfg{};
    int nx = 0;
    int64_t num_elem = 0, num_node = 0;
    int num_reg = 1;
    std::vector<int32_t> reg_num_list;

    bool mixed = false;   // float32 element-local arithmetic (fm), else f
    bool fused = false;   // float64 phases with multiply-adds fused (lulesh::fused)
    bool unstructured = false;  // an explicit-connectivity mesh (mesh), f
    DevMeshU mesh{};
    DevForceBlocksU fblocks{};  // its element blocks
    int64_t boundary_corners = 0;
    // An unstructured mesh is held renumbered (renumber_mesh_u): the new
    // number of each element and node of the mesh as given, which the field
    // copies map, and the nodelist as given ([8][num_elem]).
    std::vector<int32_t> elem_new, node_new, nodelist_given;
    int origin = 0;             // the element holding the initial energy
    int origin_dev = 0;         //   and its number on the device
    std::vector<void *> bufs;   // device allocations of an unstructured domain
    int eos_bands = 0;   // 0: mesh-order Q/EOS; else the interleave of eos_chunk
    const MlpFrags *learned = nullptr;  // learned-Q network, if set
    float learned_z_max = 0.f;
    int learned_form = 0;               // LULESH_LEARNED_FACTORS / _STANDIN
    DevFields f{};
    DevFieldsMixed fm{};
    DevScalars *s = nullptr;
    PhysParams prm{};
    SweepShape sweep = kDefaultSweep;
    void *arena = nullptr;
    hipStream_t stream = nullptr;

    bool timing = false;
    std::vector<hipEvent_t> event_pool;
    size_t events_used = 0;
    std::vector<TimedLaunch> pending;
    double kernel_ms[KID_COUNT] = {};
    long long kernel_launches[KID_COUNT] = {};
    hipEvent_t step_begin = nullptr, step_end = nullptr;
    double last_step_ms = 0.0;

    // kGraphCycles cycles captured as one graph, and the launch state it was
    // captured with (launch_signature): any change to what the cycle's
    // launches depend on (the parameters, the mode, the network, the sweep
    // shape) makes the next lulesh_step capture it again
    hipGraphExec_t graph = nullptr;
    std::vector<unsigned char> graph_sig;
    bool graphs = true;  // lulesh_set_cycle_graphs

    std::string err;
};

namespace {

int fail(lulesh_domain *d, int code, const std::string &msg)
{
    if (d) {
        d->err = msg;
    }
    return code;
}

#define HIP_TRY(d, expr)                                                     \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            return fail((d), LULESH_HIP_ERROR,                               \
                        std::string(#expr ": ") + hipGetErrorString(e_));    \
        }                                                                    \
    } while (0)

hipEvent_t next_event(lulesh_domain *d)
{
    if (d->events_used == d->event_pool.size()) {
        hipEvent_t ev = nullptr;
        if (hipEventCreate(&ev) != hipSuccess) {
            return nullptr;
        }
        d->event_pool.push_back(ev);
    }
    return d->event_pool[d->events_used++];
}

// launches one kernel, bracketed by events when kernel timing is on.  A body
// that returns a hipError_t (a launcher that checks its own launch, which
// also clears the runtime's last error) has that value checked; any other
// body is checked through hipGetLastError.
template <typename F>
int launch(lulesh_domain *d, int kid, F &&body)
{
    hipEvent_t a = nullptr, b = nullptr;
    if (d->timing) {
        a = next_event(d);
        b = next_event(d);
        if (!a || !b) {
            return fail(d, LULESH_HIP_ERROR, "hipEventCreate failed");
        }
        HIP_TRY(d, hipEventRecord(a, d->stream));
    }
    hipError_t e = hipSuccess;
    if constexpr (std::is_same_v<decltype(body()), hipError_t>) {
        e = body();
    } else {
        body();
    }
    if (e == hipSuccess) {
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        return fail(d, LULESH_HIP_ERROR,
                    std::string(kKernelNames[kid]) + " launch: " + hipGetErrorString(e));
    }
    if (d->timing) {
        HIP_TRY(d, hipEventRecord(b, d->stream));
        d->pending.push_back({kid, a, b});
    }
    return LULESH_OK;
}

int queue_cycle_u(lulesh_domain *d)
{
    const unsigned ge = unsigned((d->num_elem + kBlock - 1) / kBlock);
    const DevForceBlocksU &b = d->fblocks;  // a renumbered mesh
    // k_nodes_u: eight lanes per boundary node (num_boundary <= 2^27)
    const unsigned gb = unsigned((8 * int64_t(b.num_boundary) + kBlock - 1) / kBlock);
    hipStream_t st = d->stream;
    DevScalars *s = d->s;
    const DevFields &f = d->f;
    const DevMeshU &m = d->mesh;
    const PhysParams &p = d->prm;
    const bool fused = d->fused;
    int rc;
    if ((rc = launch(d, KID_TIME_INCREMENT, [&] {
             hipLaunchKernelGGL(k_time_increment, dim3(1), dim3(64), 0, st, s, p.ic_mode);
         })))
        return rc;
    if (b.num_blocks > 0 && (rc = launch(d, KID_FORCES_U, [&] {
             hipLaunchKernelGGL(fused ? fused::k_forces_u_blocks : k_forces_u_blocks,
                                dim3(unsigned(b.num_blocks)), dim3(kForceBlockElems), 0, st,
                                f, m, b, s, p);
         })))
        return rc;
    if (gb > 0 && (rc = launch(d, KID_NODES_U, [&] {
                       hipLaunchKernelGGL(fused ? fused::k_nodes_u : k_nodes_u, dim3(gb),
                                          dim3(kBlock), 0, st, f, m, s, p, b.first_boundary,
                                          b.num_boundary);
                   })))
        return rc;
    if ((rc = launch(d, KID_KINEMATICS_Q, [&] {
             hipLaunchKernelGGL(fused ? fused::k_kinematics_q_u : k_kinematics_q_u, dim3(ge),
                                dim3(kBlock), 0, st, f, m, s, p);
         })))
        return rc;
    if ((rc = launch(d, KID_Q_EOS_DT, [&] {
             hipLaunchKernelGGL(fused ? fused::k_q_eos_dt_u : k_q_eos_dt_u, dim3(ge),
                                dim3(kBlock), 0, st, f, m, s, p);
         })))
        return rc;
    return LULESH_OK;
}

int queue_cycle(lulesh_domain *d)
{
    if (d->unstructured) {
        return queue_cycle_u(d);
    }
    const unsigned ge = unsigned((d->num_elem + kBlock - 1) / kBlock);
    hipStream_t st = d->stream;
    DevScalars *s = d->s;
    const DevFields &f = d->f;
    const DevFieldsMixed &fm = d->fm;
    const PhysParams &p = d->prm;
    const bool mixed = d->mixed;
    int rc;
    if ((rc = launch(d, KID_TIME_INCREMENT, [&] {
             hipLaunchKernelGGL(k_time_increment, dim3(1), dim3(64), 0, st, s, p.ic_mode);
         })))
        return rc;
    const bool fused = d->fused;
    if ((rc = launch(d, KID_FORCE_SWEEP, [&]() -> hipError_t {
             if (mixed) {
                 return launch_force_sweep_mixed(d->sweep, fm, s, p, st);
             }
             return fused ? fused::launch_force_sweep(d->sweep, f, s, p, st)
                          : launch_force_sweep(d->sweep, f, s, p, st);
         })))
        return rc;
    if ((rc = launch(d, KID_KINEMATICS_Q, [&] {
             if (mixed) {
                 hipLaunchKernelGGL(k_kinematics_q_mixed, dim3(ge), dim3(kBlock), 0, st, fm,
                                    s, p);
             } else if (fused) {
                 hipLaunchKernelGGL(fused::k_kinematics_q, dim3(ge), dim3(kBlock), 0, st, f, s,
                                    p);
             } else {
                 hipLaunchKernelGGL(k_kinematics_q, dim3(ge), dim3(kBlock), 0, st, f, s, p);
             }
         })))
        return rc;
    if ((rc = launch(d, KID_Q_EOS_DT, [&] {
             if (mixed && d->eos_bands == 0) {
                 hipLaunchKernelGGL(k_q_eos_dt_mixed_nat, dim3(ge), dim3(kBlock), 0, st, fm,
                                    s, p);
             } else if (mixed) {
                 hipLaunchKernelGGL(k_q_eos_dt_mixed, dim3(ge), dim3(kBlock), 0, st, fm, s,
                                    p, d->eos_bands);
             } else if (d->learned) {
                 hipLaunchKernelGGL(fused ? fused::k_q_eos_dt_learned : k_q_eos_dt_learned,
                                    dim3(ge), dim3(kBlock), 0, st, f, s, p, d->learned,
                                    d->learned_z_max, d->learned_form);
             } else if (d->eos_bands == 0) {
                 hipLaunchKernelGGL(fused ? fused::k_q_eos_dt_nat : k_q_eos_dt_nat, dim3(ge),
                                    dim3(kBlock), 0, st, f, s, p);
             } else {
                 hipLaunchKernelGGL(fused ? fused::k_q_eos_dt : k_q_eos_dt, dim3(ge),
                                    dim3(kBlock), 0, st, f, s, p, d->eos_bands);
             }
         })))
        return rc;
    return LULESH_OK;
}

// The bytes of everything queue_cycle reads from the domain: two equal
// signatures queue the same launches with the same arguments.
std::vector<unsigned char> launch_signature(const lulesh_domain *d)
{
    std::vector<unsigned char> sig;
    auto put = [&sig](const void *p, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        sig.insert(sig.end(), b, b + n);
    };
    const unsigned char flags[3] = {d->mixed, d->fused, d->unstructured};
    put(flags, sizeof(flags));
    put(&d->num_elem, sizeof(d->num_elem));
    put(&d->eos_bands, sizeof(d->eos_bands));
    put(&d->learned, sizeof(d->learned));
    put(&d->learned_z_max, sizeof(d->learned_z_max));
    put(&d->learned_form, sizeof(d->learned_form));
    put(&d->f, sizeof(d->f));
    put(&d->fm, sizeof(d->fm));
    put(&d->s, sizeof(d->s));
    put(&d->prm, sizeof(d->prm));
    put(&d->sweep, sizeof(d->sweep));
    put(&d->mesh, sizeof(d->mesh));
    put(&d->fblocks, sizeof(d->fblocks));
    return sig;
}

void drop_graph(lulesh_domain *d)
{
    if (d->graph) {
        (void)hipGraphExecDestroy(d->graph);
        d->graph = nullptr;
    }
    d->graph_sig.clear();
}

// the graph of kGraphCycles cycles for the current launch state, captured
// from the domain's stream unless the one held matches it
int ensure_graph(lulesh_domain *d)
{
    std::vector<unsigned char> sig = launch_signature(d);
    if (d->graph && sig == d->graph_sig) {
        return LULESH_OK;
    }
    drop_graph(d);
    HIP_TRY(d, hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
    int rc = LULESH_OK;
    for (int c = 0; c < kGraphCycles && rc == LULESH_OK; ++c) {
        rc = queue_cycle(d);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(d->stream, &g);
    if (rc == LULESH_OK && e != hipSuccess) {
        rc = fail(d, LULESH_HIP_ERROR,
                  std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    if (rc == LULESH_OK) {
        const hipError_t ei = hipGraphInstantiate(&d->graph, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            d->graph = nullptr;
            rc = fail(d, LULESH_HIP_ERROR,
                      std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
    }
    if (g) {
        (void)hipGraphDestroy(g);
    }
    if (rc == LULESH_OK) {
        d->graph_sig = std::move(sig);
    }
    return rc;
}

int read_scalars(lulesh_domain *d, DevScalars *out)
{
    HIP_TRY(d, hipMemcpyAsync(out, d->s, sizeof(DevScalars), hipMemcpyDeviceToHost,
                              d->stream));
    HIP_TRY(d, hipStreamSynchronize(d->stream));
    return LULESH_OK;
}

int write_scalars(lulesh_domain *d, const DevScalars &in)
{
    HIP_TRY(d, hipMemcpyAsync(d->s, &in, sizeof(DevScalars), hipMemcpyHostToDevice,
                              d->stream));
    HIP_TRY(d, hipStreamSynchronize(d->stream));
    return LULESH_OK;
}

// folds the event pairs of finished launches into the per-kernel totals
int collect_timings(lulesh_domain *d)
{
    for (const TimedLaunch &t : d->pending) {
        float ms = 0.f;
        HIP_TRY(d, hipEventElapsedTime(&ms, t.start, t.stop));
        d->kernel_ms[t.kernel] += ms;
        d->kernel_launches[t.kernel] += 1;
    }
    d->pending.clear();
    d->events_used = 0;
    return LULESH_OK;
}

struct FieldRef {
    void *ptr;
    int64_t n;
    bool f32;                  // stored as float32 (mixed precision)
    const int32_t *renumber;   // device place of each value, or null (in order)
};

// par: the copy of the node arrays that holds the current state
FieldRef field_ref(lulesh_domain *d, int id, int par)
{
    FieldRef r{nullptr, 0, false, nullptr};
    if (d->mixed) {
        r.ptr = field_pointer_mixed(d->fm, id, par, &r.n, &r.f32);
    } else {
        r.ptr = field_pointer(d->f, id, par, &r.n);
    }
    if (d->unstructured && r.ptr) {
        r.renumber = (id < LULESH_F_E) ? d->node_new.data() : d->elem_new.data();
    }
    return r;
}

// host float64 <-> device field (converted when the field is float32, put
// in the device's order when the mesh is renumbered); synchronous
int copy_in(lulesh_domain *d, const FieldRef &r, const double *src)
{
    if (r.renumber) {
        std::vector<double> tmp(static_cast<size_t>(r.n));
        for (int64_t i = 0; i < r.n; ++i) {
            tmp[size_t(r.renumber[i])] = src[i];
        }
        HIP_TRY(d, hipMemcpyAsync(r.ptr, tmp.data(), tmp.size() * sizeof(double),
                                  hipMemcpyHostToDevice, d->stream));
        HIP_TRY(d, hipStreamSynchronize(d->stream));
        return LULESH_OK;
    }
    if (!r.f32) {
        HIP_TRY(d, hipMemcpyAsync(r.ptr, src, size_t(r.n) * sizeof(double),
                                  hipMemcpyHostToDevice, d->stream));
        HIP_TRY(d, hipStreamSynchronize(d->stream));
        return LULESH_OK;
    }
    std::vector<float> tmp(static_cast<size_t>(r.n));
    for (int64_t i = 0; i < r.n; ++i) {
        tmp[size_t(i)] = static_cast<float>(src[i]);
    }
    HIP_TRY(d, hipMemcpyAsync(r.ptr, tmp.data(), tmp.size() * sizeof(float),
                              hipMemcpyHostToDevice, d->stream));
    HIP_TRY(d, hipStreamSynchronize(d->stream));
    return LULESH_OK;
}

int copy_out(lulesh_domain *d, double *dst, const FieldRef &r)
{
    if (r.renumber) {
        std::vector<double> tmp(static_cast<size_t>(r.n));
        HIP_TRY(d, hipMemcpyAsync(tmp.data(), r.ptr, tmp.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(d, hipStreamSynchronize(d->stream));
        for (int64_t i = 0; i < r.n; ++i) {
            dst[i] = tmp[size_t(r.renumber[i])];
        }
        return LULESH_OK;
    }
    if (!r.f32) {
        HIP_TRY(d, hipMemcpyAsync(dst, r.ptr, size_t(r.n) * sizeof(double),
                                  hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(d, hipStreamSynchronize(d->stream));
        return LULESH_OK;
    }
    std::vector<float> tmp(static_cast<size_t>(r.n));
    HIP_TRY(d, hipMemcpyAsync(tmp.data(), r.ptr, tmp.size() * sizeof(float),
                              hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(d, hipStreamSynchronize(d->stream));
    for (int64_t i = 0; i < r.n; ++i) {
        dst[i] = static_cast<double>(tmp[size_t(i)]);
    }
    return LULESH_OK;
}

// the domain tuple's fields in the order of upload/download_state
constexpr int kStateFields[11] = {LULESH_F_X, LULESH_F_Y, LULESH_F_Z, LULESH_F_XD,
                                  LULESH_F_YD, LULESH_F_ZD, LULESH_F_E, LULESH_F_P,
                                  LULESH_F_Q, LULESH_F_V, LULESH_F_SS};

double *energy_ptr(lulesh_domain *d) { return d->mixed ? d->fm.e : d->f.e; }

int build_device_domain(lulesh_domain *d, const HostMesh &m)
{
    const size_t bytes = d->mixed ? domain_arena_bytes_mixed(m) : domain_arena_bytes(m);
    HIP_TRY(d, hipMalloc(&d->arena, bytes));
    HIP_TRY(d, hipMemsetAsync(d->arena, 0, bytes, d->stream));
    if (d->mixed) {
        carve_domain_mixed(static_cast<char *>(d->arena), m, d->fm, &d->s);
        HIP_TRY(d, upload_mesh_mixed(m, d->fm, d->stream));
    } else {
        carve_domain(static_cast<char *>(d->arena), m, d->f, &d->s);
        HIP_TRY(d, upload_mesh(m, d->f, d->stream));
    }
    return write_scalars(d, initial_scalars(m, d->cfg.stoptime, d->cfg.max_cycles));
}

}  // namespace

extern "C" {

void lulesh_config_default(lulesh_config *cfg)
{
    if (!cfg) {
        return;
    }
    cfg->nx = 30;
    cfg->ic_mode = LULESH_IC_2_0;
    cfg->num_reg = 11;
    cfg->balance = 1;
    cfg->cost = 1;
    cfg->max_cycles = 9999999;
    cfg->stoptime = 1.0e-2;
}

}  // extern "C"

namespace {

int create_domain(const lulesh_config *cfg, bool mixed, double efactor,
                  lulesh_domain **out)
{
    if (!cfg || !out) {
        return LULESH_BAD_ARG;
    }
    *out = nullptr;
    if (!(cfg->stoptime > 0.0) || cfg->max_cycles < 0) {
        return LULESH_BAD_ARG;
    }
    lulesh_domain *d = new (std::nothrow) lulesh_domain();
    if (!d) {
        return LULESH_BAD_ARG;
    }
    d->cfg = *cfg;
    d->mixed = mixed;
    d->sweep = mixed ? kDefaultSweepMixed : kDefaultSweep;
    if (!mixed && cfg->nx >= 1 && cfg->nx <= 600) {
        d->sweep.nz = default_sweep_planes(cfg->nx);
    }
    HostMesh m;
    try {
        MeshConfig mc;
        mc.nx = cfg->nx;
        mc.ic_mode = cfg->ic_mode;
        mc.num_reg = cfg->num_reg;
        mc.balance = cfg->balance;
        mc.cost = cfg->cost;
        mc.efactor = efactor;
        m = build_mesh(mc);
    } catch (const std::exception &ex) {
        delete d;
        return LULESH_BAD_ARG;
    }
    d->nx = m.nx;
    d->num_elem = m.num_elem;
    d->num_node = m.num_node;
    d->num_reg = m.num_reg;
    d->reg_num_list = m.reg_num_list;
    d->prm = make_params(cfg->ic_mode);
    int rc = LULESH_OK;
    hipError_t he = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete d;
        return LULESH_HIP_ERROR;
    }
    if (hipEventCreate(&d->step_begin) != hipSuccess ||
        hipEventCreate(&d->step_end) != hipSuccess) {
        lulesh_destroy(d);
        return LULESH_HIP_ERROR;
    }
    rc = build_device_domain(d, m);
    if (rc != LULESH_OK) {
        lulesh_destroy(d);
        return rc;
    }
    *out = d;
    return LULESH_OK;
}

}  // namespace

extern "C" {

int lulesh_create(const lulesh_config *cfg, lulesh_domain **out)
{
    return create_domain(cfg, false, 1.0, out);
}

int lulesh_create_mixed(const lulesh_config *cfg, lulesh_domain **out)
{
    return create_domain(cfg, true, 1.0, out);
}

int lulesh_create_scaled(const lulesh_config *cfg, double efactor, int mixed,
                         lulesh_domain **out)
{
    return create_domain(cfg, mixed != 0, efactor, out);
}

int lulesh_is_mixed(const lulesh_domain *d) { return (d && d->mixed) ? 1 : 0; }

}  // extern "C"

namespace {

// device copy of a host vector, owned by the domain
template <typename T>
int upload_vec(lulesh_domain *d, const std::vector<T> &v, T **out, size_t min_count = 0)
{
    const size_t n = std::max(v.size(), min_count);
    void *p = nullptr;
    HIP_TRY(d, hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    d->bufs.push_back(p);
    HIP_TRY(d, hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), d->stream));
    if (!v.empty()) {
        HIP_TRY(d, hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice,
                                  d->stream));
    }
    *out = static_cast<T *>(p);
    return LULESH_OK;
}

// m and fb are the sources of asynchronous copies: they must outlive the
// domain's stream work (the caller destroys the domain on failure first)
int build_device_domain_u(lulesh_domain *d, const HostMeshU &m, const ForceBlocksU &fb)
{
    const size_t ne = size_t(m.num_elem), nn = size_t(m.num_node);
    DevFields &f = d->f;
    f.nx = 0;
    f.num_elem = m.num_elem;
    f.num_node = m.num_node;
    f.div_nx = make_fastdiv(1u);
    int rc;
    const std::vector<double> zn(nn, 0.0), ze(ne, 0.0), ones(ne, 1.0);
    // node arrays: one copy, updated in place by k_nodes_u
    if ((rc = upload_vec(d, m.x, &f.x[0])) || (rc = upload_vec(d, m.y, &f.y[0])) ||
        (rc = upload_vec(d, m.z, &f.z[0])) || (rc = upload_vec(d, zn, &f.xd[0])) ||
        (rc = upload_vec(d, zn, &f.yd[0])) || (rc = upload_vec(d, zn, &f.zd[0])) ||
        (rc = upload_vec(d, m.nodal_mass, &f.nodal_mass)))
        return rc;
    f.x[1] = f.x[0]; f.y[1] = f.y[0]; f.z[1] = f.z[0];
    f.xd[1] = f.xd[0]; f.yd[1] = f.yd[0]; f.zd[1] = f.zd[0];
    double **elem[] = {&f.p, &f.q, &f.ql, &f.qq, &f.ss, &f.vnew, &f.delv,
                       &f.arealg, &f.vdov, &f.delv_xi, &f.delv_eta, &f.delv_zeta,
                       &f.delx_xi, &f.delx_eta, &f.delx_zeta};
    for (double **a : elem) {
        if ((rc = upload_vec(d, ze, a))) return rc;
    }
    if ((rc = upload_vec(d, m.e, &f.e)) || (rc = upload_vec(d, ones, &f.v)) ||
        (rc = upload_vec(d, m.volo, &f.volo)) || (rc = upload_vec(d, m.volo, &f.elem_mass)))
        return rc;
    // one region, one EOS evaluation, element order
    std::vector<int32_t> order(ne);
    for (size_t i = 0; i < ne; ++i) order[i] = int32_t(i);
    const std::vector<uint16_t> rep(ne, 1);
    int32_t *ord = nullptr;
    uint16_t *rp = nullptr;
    if ((rc = upload_vec(d, order, &ord)) || (rc = upload_vec(d, rep, &rp))) return rc;
    f.eos_order = ord;
    f.eos_rep = rp;
    f.rep = rp;  // the identity order: the same array in mesh order
    DevMeshU &g = d->mesh;
    int32_t *a = nullptr;
    if ((rc = upload_vec(d, m.nodelist, &a))) return rc;
    g.nodelist = a;
    if ((rc = upload_vec(d, m.nbr, &a))) return rc;
    g.nbr = a;
    if ((rc = upload_vec(d, m.elem_bc, &a))) return rc;
    g.elem_bc = a;
    if ((rc = upload_vec(d, m.corner_start, &a))) return rc;
    g.corner_start = a;
    uint8_t *sy = nullptr;
    if ((rc = upload_vec(d, m.symm, &sy))) return rc;
    g.symm = sy;
    // the element blocks of the force kernel, and the slots of the boundary
    // nodes' corner forces
    DevForceBlocksU &b = d->fblocks;
    b.num_blocks = fb.num_blocks;
    b.first_boundary = fb.node_start.back();
    b.num_boundary = int(nn) - b.first_boundary;
    g.corner_base = m.corner_start[size_t(b.first_boundary)];
    d->boundary_corners = int64_t(m.corner_start[nn]) - g.corner_base;
    {
        const std::vector<int32_t> slots = boundary_corner_slots(m, b.first_boundary);
        if ((rc = upload_vec(d, slots, &a))) return rc;
        g.corner_slot = a;
        const std::vector<double> none;
        // at least one slot: a node with no corner reads slot 0 (and adds 0.0)
        const size_t n4 = 4 * size_t(std::max<int64_t>(d->boundary_corners, 1));
        if ((rc = upload_vec(d, none, &g.bs, n4)) || (rc = upload_vec(d, none, &g.bh, n4)))
            return rc;
    }
    {
        int32_t *p = nullptr;
        if ((rc = upload_vec(d, fb.elem_start, &p))) return rc;
        b.elem_start = p;
        if ((rc = upload_vec(d, fb.node_start, &p))) return rc;
        b.node_start = p;
        if ((rc = upload_vec(d, fb.ncorner_start, &p))) return rc;
        b.ncorner_start = p;
        uint16_t *u16 = nullptr;
        if ((rc = upload_vec(d, fb.ncorners, &u16))) return rc;
        b.ncorners = u16;
    }
    DevScalars *sc = nullptr;
    const std::vector<DevScalars> one(1);
    if ((rc = upload_vec(d, one, &sc))) return rc;
    d->s = sc;
    HIP_TRY(d, hipStreamSynchronize(d->stream));
    HostMesh hm;  // the loop scalars of the 1.0 ICs
    hm.deltatime0 = 1.0e-7;
    hm.dtfixed = -1.0e-7;
    return write_scalars(d, initial_scalars(hm, d->cfg.stoptime, d->cfg.max_cycles));
}

int create_domain_u(const MeshLists &lists, const lulesh_config *cfg, lulesh_domain **out)
{
    if (!out) {
        return LULESH_BAD_ARG;
    }
    *out = nullptr;
    lulesh_config c;
    lulesh_config_default(&c);
    if (cfg) {
        c = *cfg;
    }
    if (!(c.stoptime > 0.0) || c.max_cycles < 0 || c.ic_mode != LULESH_IC_1_0) {
        return LULESH_BAD_ARG;
    }
    HostMeshU given, m;
    ForceBlocksU fb;
    try {
        given = build_mesh_u(lists);
        fb = build_force_blocks(given);
        m = renumber_mesh_u(given, fb);
    } catch (const std::exception &) {
        return LULESH_BAD_ARG;
    }
    lulesh_domain *d = new (std::nothrow) lulesh_domain();
    if (!d) {
        return LULESH_BAD_ARG;
    }
    d->cfg = c;
    d->unstructured = true;
    d->nx = 0;
    d->num_elem = m.num_elem;
    d->num_node = m.num_node;
    d->num_reg = 1;
    d->reg_num_list.assign(size_t(m.num_elem), 1);
    d->origin = given.origin;
    d->origin_dev = m.origin;
    d->elem_new = fb.elem_new;
    d->node_new = fb.node_new;
    d->nodelist_given = std::move(given.nodelist);
    d->prm = make_params(LULESH_IC_1_0);
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
        delete d;
        return LULESH_HIP_ERROR;
    }
    if (hipEventCreate(&d->step_begin) != hipSuccess ||
        hipEventCreate(&d->step_end) != hipSuccess) {
        lulesh_destroy(d);
        return LULESH_HIP_ERROR;
    }
    const int rc = build_device_domain_u(d, m, fb);
    if (rc != LULESH_OK) {
        lulesh_destroy(d);
        return rc;
    }
    *out = d;
    return LULESH_OK;
}

}  // namespace

extern "C" {

int lulesh_create_lmesh(const char *path, const lulesh_config *cfg, lulesh_domain **out)
{
    if (!path || !out) {
        return LULESH_BAD_ARG;
    }
    MeshLists l;
    try {
        l = read_lmesh(path);
    } catch (const std::exception &) {
        return LULESH_BAD_ARG;
    }
    return create_domain_u(l, cfg, out);
}

int lulesh_create_mesh(int num_elem, int num_node, const double *x, const double *y,
                       const double *z, const int *nodelist, const int *neighbors,
                       int nsx, const int *symm_x, int nsy, const int *symm_y, int nsz,
                       const int *symm_z, int nfree, const int *free_surf,
                       const lulesh_config *cfg, lulesh_domain **out)
{
    // the bounds of the .lmesh reader: corner-major indices c * num_elem + el
    // stay within int32 in the kernels
    if (num_elem < 1 || num_node < 8 || num_elem > (1 << 26) || num_node > (1 << 27) ||
        !x || !y || !z || !nodelist || !neighbors ||
        nsx < 0 || nsy < 0 || nsz < 0 || nfree < 0 || (nsx && !symm_x) ||
        (nsy && !symm_y) || (nsz && !symm_z) || (nfree && !free_surf)) {
        return LULESH_BAD_ARG;
    }
    MeshLists l;
    l.num_elem = num_elem;
    l.num_node = num_node;
    l.x.assign(x, x + num_node);
    l.y.assign(y, y + num_node);
    l.z.assign(z, z + num_node);
    l.nodelist.assign(nodelist, nodelist + size_t(num_elem) * 8);
    l.neighbors.assign(neighbors, neighbors + size_t(num_elem) * 6);
    if (nsx) l.symm_x.assign(symm_x, symm_x + nsx);
    if (nsy) l.symm_y.assign(symm_y, symm_y + nsy);
    if (nsz) l.symm_z.assign(symm_z, symm_z + nsz);
    if (nfree) l.free_surf.assign(free_surf, free_surf + nfree);
    return create_domain_u(l, cfg, out);
}

int lulesh_origin_element(const lulesh_domain *d) { return d ? d->origin : -1; }

int lulesh_get_nodelist(lulesh_domain *d, int *nodelist)
{
    if (!d || !nodelist) {
        return LULESH_BAD_ARG;
    }
    const int64_t ne = d->num_elem;
    if (!d->unstructured) {
        // the cube's corners: bottom face counter-clockwise, then the top face
        const int n = d->nx, en = n + 1, pl = en * en;
        int64_t el = 0;
        for (int k = 0; k < n; ++k) {
            for (int j = 0; j < n; ++j) {
                for (int i = 0; i < n; ++i, ++el) {
                    const int n0 = (k * en + j) * en + i;
                    const int c[8] = {n0, n0 + 1, n0 + en + 1, n0 + en,
                                      n0 + pl, n0 + pl + 1, n0 + pl + en + 1, n0 + pl + en};
                    for (int q = 0; q < 8; ++q) nodelist[el * 8 + q] = c[q];
                }
            }
        }
        return LULESH_OK;
    }
    const std::vector<int32_t> &cm = d->nodelist_given;  // corner-major
    for (int64_t el = 0; el < ne; ++el) {
        for (int q = 0; q < 8; ++q) nodelist[el * 8 + q] = cm[size_t(q * ne + el)];
    }
    return LULESH_OK;
}

int lulesh_force_blocks(const lulesh_domain *d, long long out[4])
{
    if (!d || !out || !d->unstructured) {
        return LULESH_BAD_ARG;
    }
    out[0] = d->fblocks.num_blocks;
    out[1] = d->fblocks.first_boundary;
    out[2] = d->fblocks.num_boundary;
    out[3] = d->boundary_corners;
    return LULESH_OK;
}

void lulesh_destroy(lulesh_domain *d)
{
    if (!d) {
        return;
    }
    if (d->stream) {
        (void)hipStreamSynchronize(d->stream);
    }
    for (hipEvent_t ev : d->event_pool) {
        (void)hipEventDestroy(ev);
    }
    if (d->step_begin) (void)hipEventDestroy(d->step_begin);
    if (d->step_end) (void)hipEventDestroy(d->step_end);
    drop_graph(d);
    if (d->arena) (void)hipFree(d->arena);
    for (void *b : d->bufs) (void)hipFree(b);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

int lulesh_step(lulesh_domain *d, int ncycles, int *cycles_done)
{
    if (cycles_done) {
        *cycles_done = 0;
    }
    if (!d || ncycles < 0) {
        return LULESH_BAD_ARG;
    }
    DevScalars before{}, after{};
    int rc = read_scalars(d, &before);
    if (rc) return rc;
    if (before.error != 0) {
        return fail(d, before.error, "domain is in an error state");
    }
    // whole graphs of cycles first, then the rest (none while kernels are
    // timed one by one: the timing brackets every launch with events).  The
    // graph is brought up to date on every call, also on one too short to
    // replay it, so that a caller's first calls capture it and later ones
    // only replay it.
    const bool graphs = d->graphs && !d->timing;
    if (graphs && ncycles > 0 && (rc = ensure_graph(d))) return rc;
    HIP_TRY(d, hipEventRecord(d->step_begin, d->stream));
    int c = 0;
    if (graphs) {
        for (; c + kGraphCycles <= ncycles; c += kGraphCycles) {
            HIP_TRY(d, hipGraphLaunch(d->graph, d->stream));
        }
    }
    for (; c < ncycles; ++c) {
        if ((rc = queue_cycle(d))) return rc;
    }
    // the last cycle's time constraints (the others were applied by the
    // following cycle's TimeIncrement)
    if ((rc = launch(d, KID_DT_FINISH, [&] {
             hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(64), 0, d->stream, d->s,
                                d->prm.ic_mode);
         })))
        return rc;
    HIP_TRY(d, hipEventRecord(d->step_end, d->stream));
    rc = read_scalars(d, &after);
    if (rc) return rc;
    {
        float ms = 0.f;
        HIP_TRY(d, hipEventElapsedTime(&ms, d->step_begin, d->step_end));
        d->last_step_ms = ms;
    }
    if (d->timing && (rc = collect_timings(d))) return rc;
    if (cycles_done) {
        *cycles_done = after.cycle - before.cycle;
    }
    if (after.error != 0) {
        return fail(d, after.error,
                    after.error == LULESH_VOLUME_ERROR ? "VolumeError" : "QStopError");
    }
    return LULESH_OK;
}

int lulesh_run(lulesh_domain *d, int *cycles_done)
{
    if (cycles_done) {
        *cycles_done = 0;
    }
    if (!d) {
        return LULESH_BAD_ARG;
    }
    for (;;) {
        int n = 0;
        const int rc = lulesh_step(d, kBatch, &n);
        if (cycles_done) {
            *cycles_done += n;
        }
        if (rc != LULESH_OK) {
            return rc;
        }
        if (n < kBatch) {
            return LULESH_OK;
        }
    }
}

int lulesh_upload_state(lulesh_domain *d, const double *x, const double *y,
                        const double *z, const double *xd, const double *yd,
                        const double *zd, const double *e, const double *p,
                        const double *q, const double *v, const double *ss,
                        double time, double deltatime)
{
    if (!d || !x || !y || !z || !xd || !yd || !zd || !e || !p || !q || !v || !ss) {
        return LULESH_BAD_ARG;
    }
    DevScalars s{};
    int rc = read_scalars(d, &s);
    if (rc) return rc;
    const double *src[11] = {x, y, z, xd, yd, zd, e, p, q, v, ss};
    for (int i = 0; i < 11; ++i) {
        if ((rc = copy_in(d, field_ref(d, kStateFields[i], s.cur), src[i]))) return rc;
    }
    return lulesh_set_scalars(d, &time, &deltatime, nullptr, nullptr, nullptr);
}

int lulesh_download_state(lulesh_domain *d, double *x, double *y, double *z,
                          double *xd, double *yd, double *zd, double *e,
                          double *p, double *q, double *v, double *ss,
                          double *time, double *deltatime)
{
    if (!d || !x || !y || !z || !xd || !yd || !zd || !e || !p || !q || !v || !ss) {
        return LULESH_BAD_ARG;
    }
    DevScalars s{};
    int rc = read_scalars(d, &s);
    if (rc) return rc;
    double *dst[11] = {x, y, z, xd, yd, zd, e, p, q, v, ss};
    for (int i = 0; i < 11; ++i) {
        if ((rc = copy_out(d, dst[i], field_ref(d, kStateFields[i], s.cur)))) return rc;
    }
    if (time) *time = s.time;
    if (deltatime) *deltatime = s.deltatime;
    return LULESH_OK;
}

int lulesh_get_field(lulesh_domain *d, int field, double *host_out)
{
    if (!d || !host_out) {
        return LULESH_BAD_ARG;
    }
    DevScalars s{};
    int rc = read_scalars(d, &s);
    if (rc) return rc;
    const FieldRef r = field_ref(d, field, s.cur);
    if (!r.ptr) {
        return fail(d, LULESH_BAD_ARG, "unknown field id");
    }
    return copy_out(d, host_out, r);
}

int lulesh_set_field(lulesh_domain *d, int field, const double *host_in)
{
    if (!d || !host_in) {
        return LULESH_BAD_ARG;
    }
    DevScalars s{};
    int rc = read_scalars(d, &s);
    if (rc) return rc;
    const FieldRef r = field_ref(d, field, s.cur);
    if (!r.ptr) {
        return fail(d, LULESH_BAD_ARG, "unknown field id");
    }
    return copy_in(d, r, host_in);
}

int lulesh_get_stats(lulesh_domain *d, lulesh_stats *out)
{
    if (!d || !out) {
        return LULESH_BAD_ARG;
    }
    DevScalars s{};
    int rc = read_scalars(d, &s);
    if (rc) return rc;
    double e0 = 0.0;
    const int origin = d->unstructured ? d->origin_dev : d->origin;
    HIP_TRY(d, hipMemcpy(&e0, energy_ptr(d) + origin, sizeof(double), hipMemcpyDeviceToHost));
    out->time = s.time;
    out->deltatime = s.deltatime;
    out->dtcourant = s.dtcourant;
    out->dthydro = s.dthydro;
    out->e_origin = e0;
    out->cycle = s.cycle;
    out->done = (!(s.time < s.stoptime) || !(s.cycle < s.max_cycles)) ? 1 : 0;
    out->num_elem = int(d->num_elem);
    out->num_node = int(d->num_node);
    return LULESH_OK;
}

int lulesh_set_scalars(lulesh_domain *d, const double *time,
                       const double *deltatime, const double *dtcourant,
                       const double *dthydro, const int *cycle)
{
    if (!d) {
        return LULESH_BAD_ARG;
    }
    DevScalars s{};
    int rc = read_scalars(d, &s);
    if (rc) return rc;
    if (time) s.time = *time;
    if (deltatime) s.deltatime = *deltatime;
    if (dtcourant) s.dtcourant = *dtcourant;
    if (dthydro) s.dthydro = *dthydro;
    if (cycle) s.cycle = *cycle;
    s.error = 0;
    return write_scalars(d, s);
}

int lulesh_set_parameter(lulesh_domain *d, const char *name, double value)
{
    if (!d || !name) {
        return LULESH_BAD_ARG;
    }
    struct Entry { const char *name; double *field; };
    PhysParams &p = d->prm;
    const Entry table[] = {
        {"hgcoef", &p.hgcoef}, {"qstop", &p.qstop},
        {"monoq_max_slope", &p.monoq_max_slope},
        {"monoq_limiter_mult", &p.monoq_limiter_mult},
        {"qlc_monoq", &p.qlc_monoq}, {"qqc_monoq", &p.qqc_monoq},
        {"qqc", &p.qqc}, {"eosvmax", &p.eosvmax}, {"eosvmin", &p.eosvmin},
        {"pmin", &p.pmin}, {"emin", &p.emin}, {"dvovmax", &p.dvovmax},
        {"refdens", &p.refdens}, {"e_cut", &p.e_cut}, {"p_cut", &p.p_cut},
        {"q_cut", &p.q_cut}, {"u_cut", &p.u_cut}, {"v_cut", &p.v_cut},
    };
    for (const Entry &e : table) {
        if (std::strcmp(e.name, name) == 0) {
            *e.field = value;
            return LULESH_OK;
        }
    }
    // the remaining ones live with the device-resident loop scalars
    DevScalars s{};
    int rc = read_scalars(d, &s);
    if (rc) return rc;
    if (std::strcmp(name, "stoptime") == 0) {
        s.stoptime = value;
        d->cfg.stoptime = value;
    } else if (std::strcmp(name, "dtmax") == 0) {
        s.dtmax = value;
    } else if (std::strcmp(name, "dtfixed") == 0) {
        s.dtfixed = value;
    } else {
        return fail(d, LULESH_BAD_ARG, std::string("unknown parameter ") + name);
    }
    return write_scalars(d, s);
}

int lulesh_set_eos_bands(lulesh_domain *d, int bands)
{
    if (!d || bands < 0 || bands > 4096) {
        return LULESH_BAD_ARG;
    }
    d->eos_bands = bands;
    return LULESH_OK;
}

int lulesh_set_sweep_shape(lulesh_domain *d, int tx, int ty, int nz, int waves)
{
    if (!d) {
        return LULESH_BAD_ARG;
    }
    if (d->unstructured) {
        return fail(d, LULESH_BAD_ARG, "an unstructured mesh has no sweep");
    }
    const SweepShape sh = {tx, ty, nz, waves};
    if (!(d->mixed ? force_sweep_mixed_shape_supported(sh)
                   : force_sweep_shape_supported(sh))) {
        return fail(d, LULESH_BAD_ARG, "sweep shape not built into the library");
    }
    d->sweep = sh;
    return LULESH_OK;
}

int lulesh_get_sweep_shape(lulesh_domain *d, int *tx, int *ty, int *nz, int *waves)
{
    if (!d) {
        return LULESH_BAD_ARG;
    }
    if (tx) *tx = d->sweep.tx;
    if (ty) *ty = d->sweep.ty;
    if (nz) *nz = d->sweep.nz;
    if (waves) *waves = d->sweep.waves;
    return LULESH_OK;
}

int lulesh_get_regions(lulesh_domain *d, int *reg_num_list)
{
    if (!d || !reg_num_list) {
        return LULESH_BAD_ARG;
    }
    std::memcpy(reg_num_list, d->reg_num_list.data(),
                d->reg_num_list.size() * sizeof(int32_t));
    return LULESH_OK;
}

int lulesh_symmetry(lulesh_domain *d, double *max_abs_diff,
                    double *total_abs_diff, double *max_rel_diff)
{
    if (!d) {
        return LULESH_BAD_ARG;
    }
    if (d->unstructured) {
        return fail(d, LULESH_BAD_ARG, "the plane-0 symmetry check needs the cube");
    }
    const int nx = d->nx;
    std::vector<double> e(size_t(nx) * nx);
    HIP_TRY(d, hipMemcpy(e.data(), energy_ptr(d), e.size() * sizeof(double),
                         hipMemcpyDeviceToHost));
    double mabs = 0.0, tabs = 0.0, mrel = 0.0;
    for (int j = 0; j < nx; ++j) {
        for (int k = j + 1; k < nx; ++k) {
            const double a = std::fabs(e[size_t(j) * nx + k] - e[size_t(k) * nx + j]);
            tabs += a;
            if (mabs < a) mabs = a;
            const double r = a / e[size_t(k) * nx + j];
            if (mrel < r) mrel = r;
        }
    }
    if (max_abs_diff) *max_abs_diff = mabs;
    if (total_abs_diff) *total_abs_diff = tabs;
    if (max_rel_diff) *max_rel_diff = mrel;
    return LULESH_OK;
}

int lulesh_kernel_count(void) { return KID_COUNT; }

const char *lulesh_kernel_name(int k)
{
    return (k >= 0 && k < KID_COUNT) ? kKernelNames[k] : "";
}

int lulesh_set_cycle_graphs(lulesh_domain *d, int enable)
{
    if (!d) {
        return LULESH_BAD_ARG;
    }
    d->graphs = enable != 0;
    if (!d->graphs) {
        drop_graph(d);
    }
    return LULESH_OK;
}

int lulesh_set_kernel_timing(lulesh_domain *d, int enable)
{
    if (!d) {
        return LULESH_BAD_ARG;
    }
    d->timing = enable != 0;
    for (int k = 0; k < KID_COUNT; ++k) {
        d->kernel_ms[k] = 0.0;
        d->kernel_launches[k] = 0;
    }
    return LULESH_OK;
}

int lulesh_kernel_times(lulesh_domain *d, double *ms, long long *launches)
{
    if (!d) {
        return LULESH_BAD_ARG;
    }
    for (int k = 0; k < KID_COUNT; ++k) {
        if (ms) ms[k] = d->kernel_ms[k];
        if (launches) launches[k] = d->kernel_launches[k];
    }
    return LULESH_OK;
}

double lulesh_last_step_ms(lulesh_domain *d) { return d ? d->last_step_ms : 0.0; }

int lulesh_set_learned_q(lulesh_domain *d, const lulesh_mlp *net, double z_max)
{
    if (!d || (net && !(z_max > 0.0))) {
        return LULESH_BAD_ARG;
    }
    if (net && (d->mixed || d->unstructured)) {
        return fail(d, LULESH_BAD_ARG,
                    "the learned closure runs on float64 structured domains");
    }
    d->learned = net ? lulesh_mlp_device_frags(net) : nullptr;
    d->learned_z_max = float(z_max);
    return LULESH_OK;
}

int lulesh_set_fused_multiply_add(lulesh_domain *d, int enable)
{
    if (!d) {
        return LULESH_BAD_ARG;
    }
    if (d->mixed) {
        return fail(d, LULESH_BAD_ARG,
                    "fused multiply-adds are a mode of the float64 step");
    }
    d->fused = enable != 0;
    return LULESH_OK;
}

int lulesh_set_learned_q_form(lulesh_domain *d, int form)
{
    if (!d || (form != LULESH_LEARNED_FACTORS && form != LULESH_LEARNED_STANDIN)) {
        return LULESH_BAD_ARG;
    }
    d->learned_form = form;
    return LULESH_OK;
}

int lulesh_q_features(lulesh_domain *d, float *features, double *q)
{
    if (!d || !features || !q) {
        return LULESH_BAD_ARG;
    }
    if (d->mixed || d->unstructured) {
        return fail(d, LULESH_BAD_ARG, "features are formed on float64 structured domains");
    }
    // kinematics and gradients of the current state (dt = 0: the half-step
    // geometry is the current one) under a temporarily active cycle
    DevScalars saved{};
    int rc = read_scalars(d, &saved);
    if (rc) return rc;
    DevScalars tmp = saved;
    tmp.active = 1;
    tmp.error = 0;
    tmp.deltatime = 0.0;
    if ((rc = write_scalars(d, tmp))) return rc;
    const unsigned ge = unsigned((d->num_elem + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_kinematics_q, dim3(ge), dim3(kBlock), 0, d->stream, d->f, d->s,
                       d->prm);
    HIP_TRY(d, hipGetLastError());
    float *dfeat = nullptr;
    double *dq = nullptr;
    const size_t ne = size_t(d->num_elem);
    HIP_TRY(d, hipMalloc(&dfeat, ne * kMlpIn * sizeof(float)));
    if (hipMalloc(&dq, ne * 4 * sizeof(double)) != hipSuccess) {
        (void)hipFree(dfeat);
        return fail(d, LULESH_HIP_ERROR, "hipMalloc failed");
    }
    hipLaunchKernelGGL(k_q_features, dim3(ge), dim3(kBlock), 0, d->stream, d->f, d->prm,
                       d->learned_form, dfeat, dq);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(features, dfeat, ne * kMlpIn * sizeof(float),
                           hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(q, dq, ne * 4 * sizeof(double), hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    (void)hipFree(dfeat);
    (void)hipFree(dq);
    if (e != hipSuccess) {
        return fail(d, LULESH_HIP_ERROR, std::string("q features: ") + hipGetErrorString(e));
    }
    DevScalars after{};
    if ((rc = read_scalars(d, &after))) return rc;
    saved.error = after.error;  // a non-positive volume is reported
    if ((rc = write_scalars(d, saved))) return rc;
    return after.error;
}

const char *lulesh_last_error(const lulesh_domain *d)
{
    return d ? d->err.c_str() : "null domain";
}

int lulesh_device_math(int op, int n, const double *a, const double *b,
                       double *out)
{
    const bool two = op == 2 || op == 4 || op == 5;
    if (n < 0 || !a || !out || (two && !b) || op < 0 || op > 5) {
        return LULESH_BAD_ARG;
    }
    if (n == 0) {
        return LULESH_OK;
    }
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    const size_t bytes = size_t(n) * sizeof(double);
    int rc = LULESH_OK;
    if (hipMalloc(&da, bytes) != hipSuccess || hipMalloc(&dout, bytes) != hipSuccess ||
        (two && hipMalloc(&db, bytes) != hipSuccess)) {
        rc = LULESH_HIP_ERROR;
    }
    if (rc == LULESH_OK && hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = LULESH_HIP_ERROR;
    if (rc == LULESH_OK && two &&
        hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = LULESH_HIP_ERROR;
    if (rc == LULESH_OK) {
        if (op >= 4) {
            // the fused build's quotient (4) and a / cbrt(b) (5)
            hipLaunchKernelGGL(fused::k_device_math, dim3((n + 255) / 256), dim3(256), 0, 0,
                               op == 4 ? 2 : 4, n, da, db, dout);
        } else {
            hipLaunchKernelGGL(k_device_math, dim3((n + 255) / 256), dim3(256), 0, 0,
                               op, n, da, db, dout);
        }
        if (hipGetLastError() != hipSuccess ||
            hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = LULESH_HIP_ERROR;
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dout) (void)hipFree(dout);
    return rc;
}

}  // extern "C"
