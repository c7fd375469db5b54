// lulesh_abi.hip -- the C ABI of include/lulesh_mi355x.h.
//
// Owns the device state of one domain, builds it from the host mesh, queues
// cycles of kernels on the domain's stream and copies fields in and out.
#include "lulesh_mi355x.h"

#include <hip/hip_runtime.h>

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
    KID_COUNT
};

const char *const kKernelNames[KID_COUNT] = {
    "k_time_increment", "k_force_sweep", "k_kinematics_q", "k_q_eos_dt",
    "k_dt_finish",
};

// default launch shape of the force sweep (32 x 8 element columns, 16 node
// planes per block, register budget for 2 waves per SIMD);
// lulesh_set_sweep_shape overrides it
constexpr SweepShape kDefaultSweep = {32, 8, 16, 2};
// the mixed-precision sweep holds float32 corner forces: three blocks fit
// on a CU
constexpr SweepShape kDefaultSweepMixed = {32, 8, 12, 3};

constexpr int kBlock = 256;
constexpr int kBatch = 32;  // cycles queued per host round trip in run()

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
    const MlpFrags *learned = nullptr;  // learned-Q network, if set
    float learned_z_max = 0.f;
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

// launches one kernel, bracketed by events when kernel timing is on
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
    body();
    HIP_TRY(d, hipGetLastError());
    if (d->timing) {
        HIP_TRY(d, hipEventRecord(b, d->stream));
        d->pending.push_back({kid, a, b});
    }
    return LULESH_OK;
}

int queue_cycle(lulesh_domain *d)
{
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
    if ((rc = launch(d, KID_FORCE_SWEEP, [&] {
             if (mixed) {
                 (void)launch_force_sweep_mixed(d->sweep, fm, s, p, st);
             } else {
                 (void)launch_force_sweep(d->sweep, f, s, p, st);
             }
         })))
        return rc;
    if ((rc = launch(d, KID_KINEMATICS_Q, [&] {
             if (mixed) {
                 hipLaunchKernelGGL(k_kinematics_q_mixed, dim3(ge), dim3(kBlock), 0, st, fm,
                                    s, p);
             } else {
                 hipLaunchKernelGGL(k_kinematics_q, dim3(ge), dim3(kBlock), 0, st, f, s, p);
             }
         })))
        return rc;
    if ((rc = launch(d, KID_Q_EOS_DT, [&] {
             if (mixed) {
                 hipLaunchKernelGGL(k_q_eos_dt_mixed, dim3(ge), dim3(kBlock), 0, st, fm, s,
                                    p);
             } else if (d->learned) {
                 hipLaunchKernelGGL(k_q_eos_dt_learned, dim3(ge), dim3(kBlock), 0, st, f, s,
                                    p, d->learned, d->learned_z_max);
             } else {
                 hipLaunchKernelGGL(k_q_eos_dt, dim3(ge), dim3(kBlock), 0, st, f, s, p);
             }
         })))
        return rc;
    return LULESH_OK;
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
    bool f32;  // stored as float32 (mixed precision)
};

// par: the copy of the node arrays that holds the current state
FieldRef field_ref(lulesh_domain *d, int id, int par)
{
    FieldRef r{nullptr, 0, false};
    if (d->mixed) {
        r.ptr = field_pointer_mixed(d->fm, id, par, &r.n, &r.f32);
    } else {
        r.ptr = field_pointer(d->f, id, par, &r.n);
    }
    return r;
}

// host float64 <-> device field (converted when the field is float32);
// synchronous
int copy_in(lulesh_domain *d, const FieldRef &r, const double *src)
{
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
    if (d->arena) (void)hipFree(d->arena);
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
    HIP_TRY(d, hipEventRecord(d->step_begin, d->stream));
    for (int c = 0; c < ncycles; ++c) {
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
    HIP_TRY(d, hipMemcpy(&e0, energy_ptr(d), sizeof(double), hipMemcpyDeviceToHost));
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

int lulesh_set_sweep_shape(lulesh_domain *d, int tx, int ty, int nz, int waves)
{
    if (!d) {
        return LULESH_BAD_ARG;
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
    if (net && d->mixed) {
        return fail(d, LULESH_BAD_ARG, "the learned closure runs on float64 domains");
    }
    d->learned = net ? lulesh_mlp_device_frags(net) : nullptr;
    d->learned_z_max = float(z_max);
    return LULESH_OK;
}

int lulesh_q_features(lulesh_domain *d, float *features, double *q)
{
    if (!d || !features || !q) {
        return LULESH_BAD_ARG;
    }
    if (d->mixed) {
        return fail(d, LULESH_BAD_ARG, "features are formed on float64 domains");
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
                       dfeat, dq);
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
    if (n < 0 || !a || !out || (op == 2 && !b) || op < 0 || op > 2) {
        return LULESH_BAD_ARG;
    }
    if (n == 0) {
        return LULESH_OK;
    }
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    const size_t bytes = size_t(n) * sizeof(double);
    int rc = LULESH_OK;
    if (hipMalloc(&da, bytes) != hipSuccess || hipMalloc(&dout, bytes) != hipSuccess ||
        (op == 2 && hipMalloc(&db, bytes) != hipSuccess)) {
        rc = LULESH_HIP_ERROR;
    }
    if (rc == LULESH_OK && hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = LULESH_HIP_ERROR;
    if (rc == LULESH_OK && op == 2 &&
        hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = LULESH_HIP_ERROR;
    if (rc == LULESH_OK) {
        hipLaunchKernelGGL(k_device_math, dim3((n + 255) / 256), dim3(256), 0, 0,
                           op, n, da, db, dout);
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
