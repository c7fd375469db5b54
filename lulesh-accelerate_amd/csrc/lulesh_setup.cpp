// lulesh_setup.cpp -- placing a host-built domain in device memory.
#include "lulesh_setup.hpp"

#include <cstring>
#include <vector>

#include "lulesh_mi355x.h"

namespace lulesh {

namespace {

size_t round_up(size_t n, size_t m) { return (n + m - 1) / m * m; }

constexpr size_t kNodeArrays = 13, kElemArrays = 19;
// mixed precision: 5 float64 and 14 float32 element arrays
constexpr size_t kElemArrays64 = 5, kElemArrays32 = 14;

}  // namespace

PhysParams make_params(int ic_mode)
{
    PhysParams p{};
    p.u_cut = 1.0e-7;
    p.hgcoef = 3.0;
    p.qstop = 1.0e+12;
    p.monoq_max_slope = 1.0;
    p.monoq_limiter_mult = 2.0;
    p.e_cut = 1.0e-7;
    p.p_cut = 1.0e-7;
    p.q_cut = 1.0e-7;
    p.v_cut = 1.0e-10;
    p.qlc_monoq = 0.5;
    p.qqc_monoq = 2.0 / 3.0;
    p.qqc = 2.0;
    p.eosvmax = 1.0e+9;
    p.eosvmin = 1.0e-9;
    p.pmin = 0.;
    p.emin = -1.0e+15;
    p.dvovmax = 0.1;
    p.refdens = 1.0;
    // sound-speed floor: reference/C/LULESH-OMP.cc:2453-2454 (1.0),
    // lulesh-2.0.3/lulesh.cc:2113-2114 (2.0)
    p.ss_tiny = (ic_mode == 1) ? .111111e-36 : .1111111e-36;
    p.ss_floor = (ic_mode == 1) ? .333333e-18 : .3333333e-18;
    p.ic_mode = ic_mode;
    return p;
}

size_t domain_arena_bytes(const HostMesh &m)
{
    const size_t pe = round_up(size_t(m.num_elem), 64);
    const size_t pn = round_up(size_t(m.num_node), 64);
    size_t bytes = (kNodeArrays * pn + kElemArrays * pe) * sizeof(double);
    bytes += round_up(pe * sizeof(int32_t), 256) + 2 * round_up(pe * sizeof(uint16_t), 256);
    bytes += round_up(m.eos_blocks.size() * sizeof(int32_t), 256);
    bytes += round_up(sizeof(DevScalars), 256);
    return bytes;
}

void carve_domain(char *base, const HostMesh &m, DevFields &f, DevScalars **s)
{
    const size_t pe = round_up(size_t(m.num_elem), 64);
    const size_t pn = round_up(size_t(m.num_node), 64);
    double *cur = reinterpret_cast<double *>(base);
    auto take_node = [&]() { double *p = cur; cur += pn; return p; };
    auto take_elem = [&]() { double *p = cur; cur += pe; return p; };
    f.nx = m.nx;
    f.num_elem = int(m.num_elem);
    f.num_node = int(m.num_node);
    f.div_nx = make_fastdiv(unsigned(m.nx));
    for (int c = 0; c < 2; ++c) {
        f.x[c] = take_node(); f.y[c] = take_node(); f.z[c] = take_node();
        f.xd[c] = take_node(); f.yd[c] = take_node(); f.zd[c] = take_node();
    }
    f.nodal_mass = take_node();
    f.e = take_elem(); f.p = take_elem(); f.q = take_elem();
    f.ql = take_elem(); f.qq = take_elem(); f.v = take_elem();
    f.volo = take_elem(); f.elem_mass = take_elem(); f.ss = take_elem();
    f.vnew = take_elem(); f.delv = take_elem(); f.arealg = take_elem();
    f.vdov = take_elem();
    f.delv_xi = take_elem(); f.delv_eta = take_elem(); f.delv_zeta = take_elem();
    f.delx_xi = take_elem(); f.delx_eta = take_elem(); f.delx_zeta = take_elem();
    char *sched = reinterpret_cast<char *>(cur);
    f.eos_order = reinterpret_cast<int32_t *>(sched);
    sched += round_up(pe * sizeof(int32_t), 256);
    f.eos_rep = reinterpret_cast<uint16_t *>(sched);
    sched += round_up(pe * sizeof(uint16_t), 256);
    f.rep = reinterpret_cast<uint16_t *>(sched);
    sched += round_up(pe * sizeof(uint16_t), 256);
    f.eos_blocks = reinterpret_cast<int32_t *>(sched);
    sched += round_up(m.eos_blocks.size() * sizeof(int32_t), 256);
    *s = reinterpret_cast<DevScalars *>(sched);
}

hipError_t upload_mesh(const HostMesh &m, const DevFields &f, hipStream_t stream)
{
    const size_t ne = size_t(m.num_elem), nn = size_t(m.num_node);
    struct Copy { void *dst; const void *src; size_t bytes; };
    std::vector<double> ones(ne, 1.0);
    const Copy copies[] = {
        {f.x[0], m.x.data(), nn * sizeof(double)},
        {f.y[0], m.y.data(), nn * sizeof(double)},
        {f.z[0], m.z.data(), nn * sizeof(double)},
        {f.nodal_mass, m.nodal_mass.data(), nn * sizeof(double)},
        {f.e, m.e.data(), ne * sizeof(double)},
        {f.volo, m.volo.data(), ne * sizeof(double)},
        {f.elem_mass, m.volo.data(), ne * sizeof(double)},  // rho = 1
        {f.v, ones.data(), ne * sizeof(double)},
        {const_cast<int32_t *>(f.eos_order), m.eos_order.data(), ne * sizeof(int32_t)},
        {const_cast<uint16_t *>(f.eos_rep), m.eos_rep.data(), ne * sizeof(uint16_t)},
        {const_cast<uint16_t *>(f.rep), m.elem_rep.data(), ne * sizeof(uint16_t)},
        {const_cast<int32_t *>(f.eos_blocks), m.eos_blocks.data(),
         m.eos_blocks.size() * sizeof(int32_t)},
    };
    for (const Copy &c : copies) {
        const hipError_t e =
            hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) {
            return e;
        }
    }
    // the host buffers must outlive the copies
    return hipStreamSynchronize(stream);
}

DevScalars initial_scalars(const HostMesh &m, double stoptime, int max_cycles)
{
    DevScalars s{};
    s.time = 0.0;
    s.deltatime = m.deltatime0;
    s.dtcourant = 1.0e+20;
    s.dthydro = 1.0e+20;
    s.stoptime = stoptime;
    s.dtfixed = m.dtfixed;
    s.dtmax = 1.0e-2;
    s.deltatimemultlb = 1.1;
    s.deltatimemultub = 1.2;
    const double big = 1.0e+20;
    std::memcpy(&s.acc_courant, &big, sizeof(big));
    std::memcpy(&s.acc_hydro, &big, sizeof(big));
    s.cycle = 0;
    s.max_cycles = max_cycles;
    s.active = 0;
    s.error = 0;
    return s;
}

size_t domain_arena_bytes_mixed(const HostMesh &m)
{
    const size_t pe = round_up(size_t(m.num_elem), 64);
    const size_t pn = round_up(size_t(m.num_node), 64);
    size_t bytes = (kNodeArrays * pn + kElemArrays64 * pe) * sizeof(double) +
                   kElemArrays32 * pe * sizeof(float);
    bytes += round_up(pe * sizeof(int32_t), 256) + 2 * round_up(pe * sizeof(uint16_t), 256);
    bytes += round_up(m.eos_blocks.size() * sizeof(int32_t), 256);
    bytes += round_up(sizeof(DevScalars), 256);
    return bytes;
}

void carve_domain_mixed(char *base, const HostMesh &m, DevFieldsMixed &f, DevScalars **s)
{
    const size_t pe = round_up(size_t(m.num_elem), 64);
    const size_t pn = round_up(size_t(m.num_node), 64);
    double *cur = reinterpret_cast<double *>(base);
    auto take_node = [&]() { double *p = cur; cur += pn; return p; };
    auto take_elem = [&]() { double *p = cur; cur += pe; return p; };
    f.nx = m.nx;
    f.num_elem = int(m.num_elem);
    f.num_node = int(m.num_node);
    f.div_nx = make_fastdiv(unsigned(m.nx));
    for (int c = 0; c < 2; ++c) {
        f.x[c] = take_node(); f.y[c] = take_node(); f.z[c] = take_node();
        f.xd[c] = take_node(); f.yd[c] = take_node(); f.zd[c] = take_node();
    }
    f.nodal_mass = take_node();
    f.e = take_elem(); f.v = take_elem(); f.volo = take_elem();
    f.elem_mass = take_elem(); f.vnew = take_elem();
    float *cf = reinterpret_cast<float *>(cur);
    auto take_f32 = [&]() { float *p = cf; cf += pe; return p; };
    f.p = take_f32(); f.q = take_f32(); f.ql = take_f32(); f.qq = take_f32();
    f.ss = take_f32(); f.delv = take_f32(); f.arealg = take_f32(); f.vdov = take_f32();
    f.delv_xi = take_f32(); f.delv_eta = take_f32(); f.delv_zeta = take_f32();
    f.delx_xi = take_f32(); f.delx_eta = take_f32(); f.delx_zeta = take_f32();
    char *sched = reinterpret_cast<char *>(cf);
    f.eos_order = reinterpret_cast<int32_t *>(sched);
    sched += round_up(pe * sizeof(int32_t), 256);
    f.eos_rep = reinterpret_cast<uint16_t *>(sched);
    sched += round_up(pe * sizeof(uint16_t), 256);
    f.rep = reinterpret_cast<uint16_t *>(sched);
    sched += round_up(pe * sizeof(uint16_t), 256);
    f.eos_blocks = reinterpret_cast<int32_t *>(sched);
    sched += round_up(m.eos_blocks.size() * sizeof(int32_t), 256);
    *s = reinterpret_cast<DevScalars *>(sched);
}

hipError_t upload_mesh_mixed(const HostMesh &m, const DevFieldsMixed &f,
                             hipStream_t stream)
{
    // the float64 arrays the mesh initialises have the same places in a
    // DevFields view
    DevFields v{};
    for (int c = 0; c < 2; ++c) {
        v.x[c] = f.x[c]; v.y[c] = f.y[c]; v.z[c] = f.z[c];
    }
    v.nodal_mass = f.nodal_mass;
    v.e = f.e;
    v.volo = f.volo;
    v.elem_mass = f.elem_mass;
    v.v = f.v;
    v.eos_order = f.eos_order;
    v.eos_rep = f.eos_rep;
    v.rep = f.rep;
    v.eos_blocks = f.eos_blocks;
    return upload_mesh(m, v, stream);
}

void *field_pointer_mixed(const DevFieldsMixed &f, int id, int par, int64_t *count,
                          bool *f32)
{
    const int64_t ne = f.num_elem, nn = f.num_node;
    void *p = nullptr;
    int64_t n = ne;
    bool single = true;
    switch (id) {
    case LULESH_F_X: p = f.x[par]; n = nn; single = false; break;
    case LULESH_F_Y: p = f.y[par]; n = nn; single = false; break;
    case LULESH_F_Z: p = f.z[par]; n = nn; single = false; break;
    case LULESH_F_XD: p = f.xd[par]; n = nn; single = false; break;
    case LULESH_F_YD: p = f.yd[par]; n = nn; single = false; break;
    case LULESH_F_ZD: p = f.zd[par]; n = nn; single = false; break;
    case LULESH_F_NODAL_MASS: p = f.nodal_mass; n = nn; single = false; break;
    case LULESH_F_E: p = f.e; single = false; break;
    case LULESH_F_V: p = f.v; single = false; break;
    case LULESH_F_VOLO: p = f.volo; single = false; break;
    case LULESH_F_ELEM_MASS: p = f.elem_mass; single = false; break;
    case LULESH_F_VNEW: p = f.vnew; single = false; break;
    case LULESH_F_P: p = f.p; break;
    case LULESH_F_Q: p = f.q; break;
    case LULESH_F_QL: p = f.ql; break;
    case LULESH_F_QQ: p = f.qq; break;
    case LULESH_F_DELV: p = f.delv; break;
    case LULESH_F_VDOV: p = f.vdov; break;
    case LULESH_F_AREALG: p = f.arealg; break;
    case LULESH_F_SS: p = f.ss; break;
    case LULESH_F_DELV_XI: p = f.delv_xi; break;
    case LULESH_F_DELV_ETA: p = f.delv_eta; break;
    case LULESH_F_DELV_ZETA: p = f.delv_zeta; break;
    case LULESH_F_DELX_XI: p = f.delx_xi; break;
    case LULESH_F_DELX_ETA: p = f.delx_eta; break;
    case LULESH_F_DELX_ZETA: p = f.delx_zeta; break;
    default: n = 0; break;
    }
    *count = n;
    *f32 = single;
    return p;
}

double *field_pointer(const DevFields &f, int id, int par, int64_t *count)
{
    const int64_t ne = f.num_elem, nn = f.num_node;
    double *p = nullptr;
    int64_t n = ne;
    switch (id) {
    case LULESH_F_X: p = f.x[par]; n = nn; break;
    case LULESH_F_Y: p = f.y[par]; n = nn; break;
    case LULESH_F_Z: p = f.z[par]; n = nn; break;
    case LULESH_F_XD: p = f.xd[par]; n = nn; break;
    case LULESH_F_YD: p = f.yd[par]; n = nn; break;
    case LULESH_F_ZD: p = f.zd[par]; n = nn; break;
    case LULESH_F_NODAL_MASS: p = f.nodal_mass; n = nn; break;
    case LULESH_F_E: p = f.e; break;
    case LULESH_F_P: p = f.p; break;
    case LULESH_F_Q: p = f.q; break;
    case LULESH_F_QL: p = f.ql; break;
    case LULESH_F_QQ: p = f.qq; break;
    case LULESH_F_V: p = f.v; break;
    case LULESH_F_VOLO: p = f.volo; break;
    case LULESH_F_DELV: p = f.delv; break;
    case LULESH_F_VDOV: p = f.vdov; break;
    case LULESH_F_AREALG: p = f.arealg; break;
    case LULESH_F_SS: p = f.ss; break;
    case LULESH_F_ELEM_MASS: p = f.elem_mass; break;
    case LULESH_F_VNEW: p = f.vnew; break;
    case LULESH_F_DELV_XI: p = f.delv_xi; break;
    case LULESH_F_DELV_ETA: p = f.delv_eta; break;
    case LULESH_F_DELV_ZETA: p = f.delv_zeta; break;
    case LULESH_F_DELX_XI: p = f.delx_xi; break;
    case LULESH_F_DELX_ETA: p = f.delx_eta; break;
    case LULESH_F_DELX_ZETA: p = f.delx_zeta; break;
    default: n = 0; break;
    }
    *count = n;
    return p;
}

}  // namespace lulesh
