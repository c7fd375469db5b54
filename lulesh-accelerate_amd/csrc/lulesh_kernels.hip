// lulesh_kernels.hip -- the element phases and the time step of a cycle.
//
// Cycle = TimeIncrement + LagrangeLeapFrog (reference/C/LULESH-OMP.cc:
// 3179-3188; lulesh-2.0.3/lulesh.cc:2771-2781).  Kernels, in launch order:
//
//   k_time_increment   TimeIncrement, on the device (:557-604)
//   k_force_sweep      LagrangeNodal: forces, acceleration, boundary
//                      conditions, velocity, position (lulesh_force.hip)
//   k_kinematics_q     per element: kinematics and the monotonic-Q gradients
//                      (:1927-2180)
//   k_q_eos_dt         per element in region order: monotonic-Q limiter
//                      (:2182-2320), qstop check (:2368-2381), EOS with the
//                      region repetitions (:2384-2761; lulesh-2.0.3/lulesh.cc:
//                      2357-2430), volume update (:2763-2782) and the
//                      Courant / hydro minima (:2799-2942), which the last
//                      block to finish applies to dtcourant / dthydro
//
// Every kernel returns at once when the cycle is not active (the run is
// over or an earlier phase raised an error), so whole batches of cycles can
// be queued without a host round trip per cycle.
#include <type_traits>

#include "lulesh_kernels.hpp"

#include "lulesh_elem.hpp"
#include "lulesh_mlp.hpp"

// FMA contraction off: every value rounded as the reference rounds it;
// the fused build (LULESH_FUSED) lets multiply-adds fuse
#ifdef LULESH_FUSED
#pragma clang fp contract(fast)
#else
#pragma clang fp contract(off)
#endif

namespace lulesh {
LULESH_VARIANT_BEGIN

namespace {

constexpr double kBig = 1.0e+20;

__device__ __forceinline__ unsigned long long as_bits(double v)
{
    return static_cast<unsigned long long>(__double_as_longlong(v));
}

__device__ __forceinline__ double from_bits(unsigned long long b)
{
    return __longlong_as_double(static_cast<long long>(b));
}

__device__ __forceinline__ void raise_error(DevScalars *s, int code)
{
    // the first error of a cycle wins; later phases see it and return
    atomicCAS(&s->error, 0, code);
}

// The Courant and hydro minima of a cycle are gathered in the accumulators
// by k_q_eos_dt and applied here, at the next kernel boundary: by the next
// cycle's TimeIncrement or by the k_dt_finish that closes a batch.
__device__ __forceinline__ void apply_constraints(DevScalars *s, int ic_mode)
{
    if (!s->dt_pending) {
        return;
    }
    s->dt_pending = 0;
    if (s->error != 0) {
        return;
    }
    const double c = from_bits(s->acc_courant);
    const double h = from_bits(s->acc_hydro);
    if (ic_mode == 2) {
        s->dtcourant = kBig;
        s->dthydro = kBig;
    }
    if (c < kBig) s->dtcourant = c;
    if (h < kBig) s->dthydro = h;
}

// element (i, j, k) of a flat element index
template <class F>
__device__ __forceinline__ void elem_ijk(const F &f, int el, int &i, int &j,
                                         int &k)
{
    const unsigned q1 = fast_div(unsigned(el), f.div_nx);
    i = el - int(q1) * f.nx;
    const unsigned q2 = fast_div(q1, f.div_nx);
    j = int(q1) - int(q2) * f.nx;
    k = int(q2);
}

}  // namespace

// ---------------------------------------------------------------------------
// TimeIncrement (reference/C/LULESH-OMP.cc:557-604; lulesh-2.0.3/lulesh.cc:
// 194-251).  Decides whether this cycle runs at all: the loop condition of
// the drivers (time < stoptime, and cycle < its for 2.0) is evaluated here.
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ void time_increment(DevScalars *s, int ic_mode)
{
    apply_constraints(s, ic_mode);
    if (s->error != 0 || !(s->time < s->stoptime) || !(s->cycle < s->max_cycles)) {
        s->active = 0;
        return;
    }
    s->active = 1;
    double targetdt = s->stoptime - s->time;
    if ((s->dtfixed <= 0.0) && (s->cycle != 0)) {
        double ratio;
        const double olddt = s->deltatime;
        double newdt = kBig;
        if (s->dtcourant < newdt) {
            newdt = s->dtcourant / 2.0;
        }
        if (s->dthydro < newdt) {
            newdt = s->dthydro * 2.0 / 3.0;
        }
        ratio = newdt / olddt;
        if (ratio >= 1.0) {
            if (ratio < s->deltatimemultlb) {
                newdt = olddt;
            } else if (ratio > s->deltatimemultub) {
                newdt = olddt * s->deltatimemultub;
            }
        }
        if (newdt > s->dtmax) {
            newdt = s->dtmax;
        }
        s->deltatime = newdt;
    }
    if ((targetdt > s->deltatime) && (targetdt < (4.0 * s->deltatime / 3.0))) {
        targetdt = 2.0 * s->deltatime / 3.0;
    }
    if (targetdt < s->deltatime) {
        s->deltatime = targetdt;
    }
    s->time += s->deltatime;
    ++s->cycle;
    s->cur ^= 1;  // this cycle's nodes go to the other copy
    s->acc_courant = as_bits(kBig);
    s->acc_hydro = as_bits(kBig);
    s->dt_pending = 1;
}

}  // namespace

__global__ void k_time_increment(DevScalars *s, int ic_mode)
{
    if (threadIdx.x == 0) {
        // on a copy in registers: one batch of loads and one of stores, not a
        // wait for memory at every field the pointer form might alias
        DevScalars t = *s;
        time_increment(&t, ic_mode);
        *s = t;
    }
}

// one thread per member
__global__ void k_time_increment_batch(const BatchMember *__restrict__ members,
                                       const BatchWork *__restrict__ work, int ic_mode)
{
    const int n = work->num_members;
    for (int m = int(threadIdx.x); m < n; m += int(blockDim.x)) {
        DevScalars *const s = members[m].s;
        DevScalars t = *s;  // on a register copy, as in k_time_increment
        time_increment(&t, ic_mode);
        *s = t;
    }
}

// ---------------------------------------------------------------------------
// Kinematics (CalcKinematicsForElems + CalcLagrangeElements,
// reference/C/LULESH-OMP.cc:1927-2032) and the monotonic-Q gradients
// (CalcMonotonicQGradientsForElems, :2034-2180) of each element, sharing one
// read of its eight advanced nodes.
// ---------------------------------------------------------------------------
namespace {

// the element's eight nodes are given (structured or from a node list);
// OneTrip: all node loads before any arithmetic (see below)
template <bool OneTrip, class F>
__device__ __forceinline__ void kinematics_q_nodes(const F &f, DevScalars *s,
                                                   int el, const int (&nodes)[8])
{
    // the loop scalars are read through a global pointer: scalar loads
    const LULESH_AS_GLOBAL DevScalars *gs = (const LULESH_AS_GLOBAL DevScalars *)s;
    const int par = wave_uniform(gs->cur);
    const double dt = wave_uniform(gs->deltatime);
    typename F::CPtr __restrict__ px = parity_copy(f.x, par);
    typename F::CPtr __restrict__ py = parity_copy(f.y, par);
    typename F::CPtr __restrict__ pz = parity_copy(f.z, par);
    typename F::CPtr __restrict__ pxd = parity_copy(f.xd, par);
    typename F::CPtr __restrict__ pyd = parity_copy(f.yd, par);
    typename F::CPtr __restrict__ pzd = parity_copy(f.zd, par);
    double x[8], y[8], z[8], xv[8], yv[8], zv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        x[c] = px[nodes[c]];
        y[c] = py[nodes[c]];
        z[c] = pz[nodes[c]];
        xv[c] = pxd[nodes[c]];
        yv[c] = pyd[nodes[c]];
        zv[c] = pzd[nodes[c]];
    }
    const double volo = f.volo[el];
    const double v_old = f.v[el];  // read before the stores below may alias it
    // The empty asm takes every node value as an operand, so all the loads are
    // issued before any arithmetic and the element waits for memory once.
    // Without it the compiler issues the velocity loads after the volume
    // arithmetic, and each element waits for two round trips.  The list form
    // (whose node indices are loads themselves) runs faster without it.
    if constexpr (OneTrip) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            asm volatile("" : "+v"(x[c]), "+v"(y[c]), "+v"(z[c]), "+v"(xv[c]), "+v"(yv[c]),
                         "+v"(zv[c]));
        }
        __builtin_amdgcn_s_setprio(1);  // a wave with its data finishes first
    }

    // kinematics
    const double volume = elem_volume(x, y, z);
    const double vnew = elem_quot(volume, volo);
    f.vnew[el] = vnew;
    f.delv[el] = vnew - v_old;
    f.arealg[el] = characteristic_length(x, y, z, volume);
    {
        const double dt2 = 0.5 * dt;
        double xh[8], yh[8], zh[8], b[3][4];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            xh[c] = x[c] - dt2 * xv[c];
            yh[c] = y[c] - dt2 * yv[c];
            zh[c] = z[c] - dt2 * zv[c];
        }
        const double det_j = shape_derivatives(xh, yh, zh, b);
        const double inv_det_j = elem_recip(det_j);
        const double dxx = inv_det_j * (b[0][0] * (xv[0] - xv[6]) + b[0][1] * (xv[1] - xv[7]) +
                                        b[0][2] * (xv[2] - xv[4]) + b[0][3] * (xv[3] - xv[5]));
        const double dyy = inv_det_j * (b[1][0] * (yv[0] - yv[6]) + b[1][1] * (yv[1] - yv[7]) +
                                        b[1][2] * (yv[2] - yv[4]) + b[1][3] * (yv[3] - yv[5]));
        const double dzz = inv_det_j * (b[2][0] * (zv[0] - zv[6]) + b[2][1] * (zv[1] - zv[7]) +
                                        b[2][2] * (zv[2] - zv[4]) + b[2][3] * (zv[3] - zv[5]));
        f.vdov[el] = dxx + dyy + dzz;
    }
    if (vnew <= 0.0) {
        raise_error(s, -1);
    }

    // monotonic-Q gradients; each root below is of a sum of squares plus
    // ptiny, so its argument is >= 1e-36 and sqrt_above_tiny is sqrt there
    const double ptiny = 1.e-36;
    const double vol = volo * vnew;
    const double norm = elem_recip(vol + ptiny);
    const double dxj = -0.25 * ((x[0] + x[1] + x[5] + x[4]) - (x[3] + x[2] + x[6] + x[7]));
    const double dyj = -0.25 * ((y[0] + y[1] + y[5] + y[4]) - (y[3] + y[2] + y[6] + y[7]));
    const double dzj = -0.25 * ((z[0] + z[1] + z[5] + z[4]) - (z[3] + z[2] + z[6] + z[7]));
    const double dxi = 0.25 * ((x[1] + x[2] + x[6] + x[5]) - (x[0] + x[3] + x[7] + x[4]));
    const double dyi = 0.25 * ((y[1] + y[2] + y[6] + y[5]) - (y[0] + y[3] + y[7] + y[4]));
    const double dzi = 0.25 * ((z[1] + z[2] + z[6] + z[5]) - (z[0] + z[3] + z[7] + z[4]));
    const double dxk = 0.25 * ((x[4] + x[5] + x[6] + x[7]) - (x[0] + x[1] + x[2] + x[3]));
    const double dyk = 0.25 * ((y[4] + y[5] + y[6] + y[7]) - (y[0] + y[1] + y[2] + y[3]));
    const double dzk = 0.25 * ((z[4] + z[5] + z[6] + z[7]) - (z[0] + z[1] + z[2] + z[3]));
    double ax, ay, az, dxv, dyv, dzv;

    ax = dyi * dzj - dzi * dyj;
    ay = dzi * dxj - dxi * dzj;
    az = dxi * dyj - dyi * dxj;
    f.delx_zeta[el] = elem_quot(vol, sqrt_above_tiny(ax * ax + ay * ay + az * az + ptiny));
    ax *= norm;
    ay *= norm;
    az *= norm;
    dxv = 0.25 * ((xv[4] + xv[5] + xv[6] + xv[7]) - (xv[0] + xv[1] + xv[2] + xv[3]));
    dyv = 0.25 * ((yv[4] + yv[5] + yv[6] + yv[7]) - (yv[0] + yv[1] + yv[2] + yv[3]));
    dzv = 0.25 * ((zv[4] + zv[5] + zv[6] + zv[7]) - (zv[0] + zv[1] + zv[2] + zv[3]));
    f.delv_zeta[el] = ax * dxv + ay * dyv + az * dzv;

    ax = dyj * dzk - dzj * dyk;
    ay = dzj * dxk - dxj * dzk;
    az = dxj * dyk - dyj * dxk;
    f.delx_xi[el] = elem_quot(vol, sqrt_above_tiny(ax * ax + ay * ay + az * az + ptiny));
    ax *= norm;
    ay *= norm;
    az *= norm;
    dxv = 0.25 * ((xv[1] + xv[2] + xv[6] + xv[5]) - (xv[0] + xv[3] + xv[7] + xv[4]));
    dyv = 0.25 * ((yv[1] + yv[2] + yv[6] + yv[5]) - (yv[0] + yv[3] + yv[7] + yv[4]));
    dzv = 0.25 * ((zv[1] + zv[2] + zv[6] + zv[5]) - (zv[0] + zv[3] + zv[7] + zv[4]));
    f.delv_xi[el] = ax * dxv + ay * dyv + az * dzv;

    ax = dyk * dzi - dzk * dyi;
    ay = dzk * dxi - dxk * dzi;
    az = dxk * dyi - dyk * dxi;
    f.delx_eta[el] = elem_quot(vol, sqrt_above_tiny(ax * ax + ay * ay + az * az + ptiny));
    ax *= norm;
    ay *= norm;
    az *= norm;
    dxv = -0.25 * ((xv[0] + xv[1] + xv[5] + xv[4]) - (xv[3] + xv[2] + xv[6] + xv[7]));
    dyv = -0.25 * ((yv[0] + yv[1] + yv[5] + yv[4]) - (yv[3] + yv[2] + yv[6] + yv[7]));
    dzv = -0.25 * ((zv[0] + zv[1] + zv[5] + zv[4]) - (zv[3] + zv[2] + zv[6] + zv[7]));
    f.delv_eta[el] = ax * dxv + ay * dyv + az * dzv;
}

template <class F>
__device__ __forceinline__ void kinematics_q(const F &f, DevScalars *s, int el)
{
    int i, j, k;
    elem_ijk(f, el, i, j, k);
    const int en = f.nx + 1;
    const int n0 = (k * en + j) * en + i;
    const int pl = en * en;
    const int nodes[8] = {n0,      n0 + 1,      n0 + en + 1,      n0 + en,
                          n0 + pl, n0 + pl + 1, n0 + pl + en + 1, n0 + pl + en};
    kinematics_q_nodes<true>(f, s, el, nodes);
}

}  // namespace

__global__ __launch_bounds__(256) void k_kinematics_q(DevFields f, DevScalars *s,
                                                      PhysParams prm)
{
    if (s->active == 0 || s->error != 0) {
        return;
    }
    const int el = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    if (el < f.num_elem) {
        kinematics_q(f, s, el);
    }
}

// persistent blocks over the element chunks of all members
__global__ __launch_bounds__(256) void k_kinematics_q_batch(const BatchMember *__restrict__ members,
                                                            const BatchWork *__restrict__ work,
                                                            PhysParams prm)
{
    const int n = work->num_chunks;
    const LULESH_AS_GLOBAL int32_t *chunks = (const LULESH_AS_GLOBAL int32_t *)work->chunks;
    for (int c = int(blockIdx.x); c < n; c += int(gridDim.x)) {
        const BatchMember &m = members[chunks[2 * c]];
        if (!member_running(m.s)) {
            continue;
        }
        const int el = chunks[2 * c + 1] + int(threadIdx.x);
        if (el < m.f.num_elem) {
            kinematics_q(global_view(m.f), m.s, el);
        }
    }
}

// The ensemble's element phases with one chunk per block (the default
// launch: a grid of the full chunk list; blocks beyond the current list
// return), without the persistent loop.
__global__ __launch_bounds__(256, 3) void k_kinematics_q_chunks(const BatchMember *__restrict__ members,
                                                             const BatchWork *__restrict__ work,
                                                             PhysParams prm)
{
    const int c = int(blockIdx.x);
    if (c >= work->num_chunks) {
        return;
    }
    const LULESH_AS_GLOBAL int32_t *chunks = (const LULESH_AS_GLOBAL int32_t *)work->chunks;
    const BatchMember &m = members[chunks[2 * c]];
    if (!member_running(m.s)) {
        return;
    }
    const int el = chunks[2 * c + 1] + int(threadIdx.x);
    if (el < m.f.num_elem) {
        kinematics_q(global_view(m.f), m.s, el);
    }
}

namespace {

// limiter of one direction (reference/C/LULESH-OMP.cc:2201-2228)
__device__ __forceinline__ double mono_phi(double self, double delvm,
                                           double delvp, double limiter_mult,
                                           double max_slope)
{
    const double ptiny = 1.e-36;
    const double norm = elem_recip(self + ptiny);
    double phi;
    delvm = delvm * norm;
    delvp = delvp * norm;
    phi = .5 * (delvm + delvp);
    delvm *= limiter_mult;
    delvp *= limiter_mult;
    if (delvm < phi) phi = delvm;
    if (delvp < phi) phi = delvp;
    if (phi < 0.) phi = 0.;
    if (phi > max_slope) phi = max_slope;
    return phi;
}

}  // namespace

// ---------------------------------------------------------------------------
// Monotonic Q, EOS, volume update and time constraints of each element.
// k_q_eos_dt_nat (the default) walks the elements in mesh order and shares
// a block's EOS repetitions over its threads; k_q_eos_dt walks them region
// by region (eos_order), so a wavefront mostly shares one repetition count.
// The neighbour across a face is the adjacent element of the structured
// cube; x = 0, y = 0 and z = 0 faces are symmetry planes (neighbour value =
// own value), the far faces free surfaces (neighbour value = 0)
// (reference/C/LULESH-OMP.cc:3153-3171, :2204-2287).  The Courant and hydro
// minima are exact and order independent: each block folds its minimum into
// the cycle's accumulator with an unsigned 64-bit atomic min on the bit
// pattern, and the next kernel boundary applies them (apply_constraints).
// ---------------------------------------------------------------------------
namespace {

// What the monotonic-Q limiter of an element works from
// (reference/C/LULESH-OMP.cc:2182-2320): the limiter values of the three
// directions and, for a compressing element (vdov <= 0), the velocity
// jumps across it (delv * delx, clamped to <= 0) and its density.
struct QInputs {
    double phi[3];
    // the face values the limiter of each direction compares with the
    // element's own (low side, high side), boundary rules applied
    double face[3][2];
    double du[3];
    double rho;
    double vdov, vnew, volo;
    // the element's fields that the EOS and the time constraints read; they
    // are loaded with the limiter's inputs, before the kernel stores anything
    double q_old, v_old, e_old, delv, p_old, arealg;
};

// the rest of the inputs once the limiter values are known
template <class F>
__device__ __forceinline__ void q_inputs_rest(const F &f, int el, double dvxi,
                                              double dveta, double dvzeta, QInputs &in);

template <class F>
__device__ __forceinline__ void q_inputs(const F &f, const PhysParams &prm,
                                         int el, int i, int j, int k, QInputs &in)
{
    const int nx = f.nx;
    const int sy = nx, sz = nx * nx;
    const double dvxi = f.delv_xi[el], dveta = f.delv_eta[el], dvzeta = f.delv_zeta[el];
    // every neighbour is loaded, from the element itself across a boundary
    // face (the value is then replaced), so that no load waits on a branch
    const double xm = f.delv_xi[(i == 0) ? el : el - 1];
    const double xp = f.delv_xi[(i == nx - 1) ? el : el + 1];
    const double ym = f.delv_eta[(j == 0) ? el : el - sy];
    const double yp = f.delv_eta[(j == nx - 1) ? el : el + sy];
    const double zm = f.delv_zeta[(k == 0) ? el : el - sz];
    const double zp = f.delv_zeta[(k == nx - 1) ? el : el + sz];
    q_inputs_rest(f, el, dvxi, dveta, dvzeta, in);
    const double lm = prm.monoq_limiter_mult, ms = prm.monoq_max_slope;
    in.face[0][0] = (i == 0) ? dvxi : xm;
    in.face[0][1] = (i == nx - 1) ? 0.0 : xp;
    in.face[1][0] = (j == 0) ? dveta : ym;
    in.face[1][1] = (j == nx - 1) ? 0.0 : yp;
    in.face[2][0] = (k == 0) ? dvzeta : zm;
    in.face[2][1] = (k == nx - 1) ? 0.0 : zp;
    in.phi[0] = mono_phi(dvxi, in.face[0][0], in.face[0][1], lm, ms);
    in.phi[1] = mono_phi(dveta, in.face[1][0], in.face[1][1], lm, ms);
    in.phi[2] = mono_phi(dvzeta, in.face[2][0], in.face[2][1], lm, ms);
}

// the neighbour value across one face of an unstructured element: own value
// on a symmetry face, 0 on a free face, else the neighbour's
// (reference/C/LULESH-OMP.cc:2201-2287).  The neighbour's value is loaded in
// every case (the mesh builder checks that every neighbour index is an
// element), so that no load waits on the boundary bits.
__device__ __forceinline__ double face_value(double own, double nbr_value, int bc, int symm,
                                             int freeb)
{
    return (bc & symm) ? own : ((bc & freeb) ? 0.0 : nbr_value);
}

__device__ __forceinline__ void q_inputs_u(const DevFields &f, const DevMeshU &m,
                                           const PhysParams &prm, int el, QInputs &in)
{
    const int ne = f.num_elem;
    const int bc = m.elem_bc[el];
    const double dvxi = f.delv_xi[el], dveta = f.delv_eta[el], dvzeta = f.delv_zeta[el];
    const double xm = f.delv_xi[m.nbr[0 * ne + el]], xp = f.delv_xi[m.nbr[1 * ne + el]];
    const double ym = f.delv_eta[m.nbr[2 * ne + el]], yp = f.delv_eta[m.nbr[3 * ne + el]];
    const double zm = f.delv_zeta[m.nbr[4 * ne + el]], zp = f.delv_zeta[m.nbr[5 * ne + el]];
    q_inputs_rest(f, el, dvxi, dveta, dvzeta, in);
    const double lm = prm.monoq_limiter_mult, ms = prm.monoq_max_slope;
    in.phi[0] = mono_phi(dvxi, face_value(dvxi, xm, bc, kBcXiMSymm, kBcXiMFree),
                         face_value(dvxi, xp, bc, kBcXiPSymm, kBcXiPFree), lm, ms);
    in.phi[1] = mono_phi(dveta, face_value(dveta, ym, bc, kBcEtaMSymm, kBcEtaMFree),
                         face_value(dveta, yp, bc, kBcEtaPSymm, kBcEtaPFree), lm, ms);
    in.phi[2] = mono_phi(dvzeta, face_value(dvzeta, zm, bc, kBcZetaMSymm, kBcZetaMFree),
                         face_value(dvzeta, zp, bc, kBcZetaPSymm, kBcZetaPFree), lm, ms);
}

template <class F>
__device__ __forceinline__ void q_inputs_rest(const F &f, int el, double dvxi,
                                              double dveta, double dvzeta, QInputs &in)
{
    in.vdov = f.vdov[el];
    in.vnew = f.vnew[el];
    in.volo = f.volo[el];
    in.q_old = f.q[el];
    in.v_old = f.v[el];
    in.e_old = f.e[el];
    in.delv = f.delv[el];
    in.p_old = f.p[el];
    // arealg is read only by a moving element, delx and the mass only by a
    // compressing or resting one (a quiescent element reads neither arealg)
    in.arealg = (in.vdov != 0.) ? f.arealg[el] : 0.0;
    in.du[0] = in.du[1] = in.du[2] = 0.0;
    in.rho = 0.0;
    if (!(in.vdov > 0.)) {
        double dxi = dvxi * f.delx_xi[el];
        double deta = dveta * f.delx_eta[el];
        double dzeta = dvzeta * f.delx_zeta[el];
        if (dxi > 0.) dxi = 0.;
        if (deta > 0.) deta = 0.;
        if (dzeta > 0.) dzeta = 0.;
        in.du[0] = dxi;
        in.du[1] = deta;
        in.du[2] = dzeta;
        in.rho = elem_quot(f.elem_mass[el], in.volo * in.vnew);
    }
}

// linear and quadratic Q of the monotonic limiter (:2289-2316)
__device__ __forceinline__ void monotonic_q(const PhysParams &prm, const QInputs &in,
                                            double &qlin, double &qquad)
{
    if (in.vdov > 0.) {
        qlin = 0.;
        qquad = 0.;
    } else {
        const double dxi = in.du[0], deta = in.du[1], dzeta = in.du[2];
        const double phixi = in.phi[0], phieta = in.phi[1], phizeta = in.phi[2];
        qlin = -prm.qlc_monoq * in.rho *
               (dxi * (1. - phixi) + deta * (1. - phieta) + dzeta * (1. - phizeta));
        qquad = prm.qqc_monoq * in.rho *
                (dxi * dxi * (1. - phixi * phixi) + deta * deta * (1. - phieta * phieta) +
                 dzeta * dzeta * (1. - phizeta * phizeta));
    }
}

// What one EOS evaluation of an element reads (the clamped new volume and
// the values of the last cycle), formed once per element.
struct EosIn {
    double vnewc, e_old, delvc, p_old, q_old, qquad, qlin;
};

__device__ __forceinline__ EosConst eos_const(const PhysParams &prm)
{
    EosConst ec;
    ec.e_cut = prm.e_cut;
    ec.p_cut = prm.p_cut;
    ec.q_cut = prm.q_cut;
    ec.eosvmax = prm.eosvmax;
    ec.eosvmin = prm.eosvmin;
    ec.pmin = prm.pmin;
    ec.emin = prm.emin;
    ec.rho0 = prm.refdens;
    ec.unit_rho0 = prm.refdens == 1.0;
    ec.ss_tiny = prm.ss_tiny;
    ec.ss_floor = prm.ss_floor;
    return ec;
}

// The element's side of the EOS before its evaluations: ql / qq stored,
// qstop tested on the q of the previous cycle (:2368-2381), the new and old
// relative volumes clamped (:2704-2753; 2.0 clamps vnew in place) and the
// clamped old volume tested.
template <class F>
__device__ __forceinline__ EosIn eos_prepare(const F &f, DevScalars *s,
                                             const PhysParams &prm, int el,
                                             const QInputs &in, double qlin, double qquad)
{
    f.ql[el] = qlin;
    f.qq[el] = qquad;
    const double q_old = in.q_old;
    if (q_old > prm.qstop) {
        raise_error(s, -2);
    }
    double vnewc = in.vnew;
    if (prm.eosvmin != 0.) {
        if (vnewc < prm.eosvmin) vnewc = prm.eosvmin;
    }
    if (prm.eosvmax != 0.) {
        if (vnewc > prm.eosvmax) vnewc = prm.eosvmax;
    }
    double vc = in.v_old;
    if (prm.eosvmin != 0.) {
        if (vc < prm.eosvmin) vc = prm.eosvmin;
    }
    if (prm.eosvmax != 0.) {
        if (vc > prm.eosvmax) vc = prm.eosvmax;
    }
    if (vc <= 0.) {
        raise_error(s, -1);
    }
    return EosIn{vnewc, in.e_old, in.delv, in.p_old, q_old, qquad, qlin};
}

// `rep` evaluations of the EOS on the same inputs, as the reference repeats
// them (lulesh-2.0.3/lulesh.cc:2267, :2413-2424).  The inputs pass through an
// empty asm statement at the top of every repetition, so the compiler cannot
// know they are unchanged and computes each repetition in full; they are
// carried from one repetition to the next in the same registers (no copies),
// and every repetition's result is marked as used.
template <bool Unit>
__device__ __forceinline__ EosOut eos_repeat(const EosIn &a, int rep, const EosConst &ec)
{
    double a0 = a.vnewc, a1 = a.e_old, a2 = a.delvc, a3 = a.p_old, a4 = a.q_old;
    double a5 = a.qquad, a6 = a.qlin;
    EosOut o;
    o.p = o.e = o.q = o.ss = 0.0;
    for (int r = 0; r < rep; ++r) {
        asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5),
                     "+v"(a6));
        o = eos_element<Unit>(a0, a1, a2, a3, a4, a5, a6, ec);
        asm volatile("" ::"v"(o.p), "v"(o.e), "v"(o.q), "v"(o.ss));
    }
    return o;
}

// The element's side after the EOS: p, e, q, ss stored, the volume update
// (UpdateVolumesForElems: 1.0 uses the unclamped vnew, 2.0 the clamped one,
// lulesh-2.0.3/lulesh.cc:2439-2452) and the Courant and hydro candidates
// (:2799-2933).
template <class F>
__device__ __forceinline__ void eos_finish(const F &f, const PhysParams &prm, int el,
                                           const QInputs &in, double vnewc, const EosOut &o,
                                           unsigned long long &cand_c,
                                           unsigned long long &cand_h)
{
    const double vdov = in.vdov, vnew = in.vnew;
    f.p[el] = o.p;
    f.e[el] = o.e;
    f.q[el] = o.q;
    f.ss[el] = o.ss;
    double tmpv = (prm.ic_mode == 1) ? vnew : vnewc;
    if (fabs(tmpv - 1.0) < prm.v_cut) tmpv = 1.0;
    f.v[el] = tmpv;
    if (prm.ic_mode == 2) {
        f.vnew[el] = vnewc;
    }
    if (vdov != 0.) {
        const double arealg = in.arealg;
        const double qqc2 = 64.0 * prm.qqc * prm.qqc;
        double dtf = o.ss * o.ss;
        if (vdov < 0.) {
            dtf = dtf + qqc2 * arealg * arealg * vdov * vdov;
        }
        dtf = sqrt(dtf);
        dtf = elem_quot(arealg, dtf);
        cand_c = as_bits(dtf);
        const double dtdvov = elem_quot(prm.dvovmax, fabs(vdov) + 1.e-20);
        cand_h = as_bits(dtdvov);
    }
}

// Everything after ql / qq for one element: qstop, the volume clamps, the
// EOS with the region repetitions, the volume update and the element's
// Courant / hydro candidates.
template <class F>
__device__ __forceinline__ void eos_dt_elem(const F &f, DevScalars *s,
                                            const PhysParams &prm, int el, int rep,
                                            const QInputs &in, double qlin, double qquad,
                                            unsigned long long &cand_c,
                                            unsigned long long &cand_h)
{
    const EosIn a = eos_prepare(f, s, prm, el, in, qlin, qquad);
    const EosConst ec = eos_const(prm);
    EosOut o;
    __builtin_amdgcn_s_setprio(0);  // the repetitions yield to waves still loading
    if (ec.unit_rho0) {
        o = eos_repeat<true>(a, rep, ec);
    } else {
        o = eos_repeat<false>(a, rep, ec);
    }
    __builtin_amdgcn_s_setprio(1);  // stores and the block's reduction
    eos_finish(f, prm, el, in, a.vnewc, o, cand_c, cand_h);
}

// fold the block minima into the cycle's accumulators; they are applied to
// dtcourant / dthydro after the kernel (apply_constraints).  No fence is
// needed: the next kernel on the stream sees the atomics' results.
__device__ __forceinline__ void fold_minima(DevScalars *s, unsigned long long cand_c,
                                            unsigned long long cand_h,
                                            unsigned long long *lds_c,
                                            unsigned long long *lds_h)
{
    const unsigned long long mc = block_min(cand_c, lds_c);
    const unsigned long long mh = block_min(cand_h, lds_h);
    if (threadIdx.x == 0) {
        if (mc < as_bits(kBig)) atomicMin(&s->acc_courant, mc);
        if (mh < as_bits(kBig)) atomicMin(&s->acc_hydro, mh);
    }
}

// one block of 256 elements in schedule order; every thread of the block
// calls it (the minima are reduced over the block)
template <class F>
__device__ __forceinline__ void q_eos_dt_block(const F &f, DevScalars *s,
                                               const PhysParams &prm, int slot,
                                               unsigned long long *lds_c,
                                               unsigned long long *lds_h)
{
    unsigned long long cand_c = as_bits(kBig), cand_h = as_bits(kBig);
    if (slot < f.num_elem) {
        const int el = f.eos_order[slot];
        const int rep = f.eos_rep[slot];
        int i, j, k;
        elem_ijk(f, el, i, j, k);
        QInputs in;
        q_inputs(f, prm, el, i, j, k, in);
        double qlin, qquad;
        monotonic_q(prm, in, qlin, qquad);
        eos_dt_elem(f, s, prm, el, rep, in, qlin, qquad, cand_c, cand_h);
    }
    fold_minima(s, cand_c, cand_h, lds_c, lds_h);
}

}  // namespace

__global__ __launch_bounds__(256) void k_q_eos_dt(DevFields f, DevScalars *s,
                                                  PhysParams prm, int bands)
{
    __shared__ unsigned long long lds_c[4], lds_h[4];
    __shared__ int skip;
    // other blocks may raise an error while this one runs: read the flags
    // once per block so that every thread takes the same path to the
    // barriers of the reduction
    if (threadIdx.x == 0) {
        skip = (s->active == 0 || s->error != 0) ? 1 : 0;
    }
    __syncthreads();
    if (skip) {
        return;
    }
    const int chunk = eos_chunk(int(blockIdx.x), int(gridDim.x), bands);
    __builtin_amdgcn_s_setprio(1);  // the element's loads first
    q_eos_dt_block(f, s, prm, chunk * int(blockDim.x) + int(threadIdx.x), lds_c, lds_h);
}

namespace {

// Per block: the EOS inputs of each thread's element, the inclusive prefix of
// the evaluations beyond each element's first, and each wave's total.
struct RepLds {
    double in[7][256];
    int end[256];
    int wave_total[4];
};

// The EOS evaluations of a block's 256 elements beyond each one's first.  The
// thread that owns an element evaluates it once (its result is the
// element's); the other rep - 1 evaluations of every element are dealt out
// over all 256 threads, so a block of elements with 1, 2 and 20 repetitions
// takes ceil(sum(rep - 1) / 256) extra passes instead of running every
// thread 20 times.  An element's evaluations all read the same inputs and
// give the same results: each is computed in full (its inputs pass through
// an empty asm statement) and its result marked as used, as the reference
// repeats the evaluation (lulesh-2.0.3/lulesh.cc:2413-2424).  Every thread
// of the block (256, four waves) calls it.
template <bool Unit>
__device__ __forceinline__ void eos_extra_evaluations(const EosIn &a, int extra,
                                                      const EosConst &ec, RepLds &L)
{
    const int t = int(threadIdx.x), lane = t & 63, wave = t >> 6;
    int incl = extra;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) L.wave_total[wave] = incl;
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const int v = L.wave_total[w];
        if (w < wave) base += v;
        total += v;
    }
    if (total == 0) {  // the same for every thread of the block
        return;
    }
    L.end[t] = incl + base;
    L.in[0][t] = a.vnewc;
    L.in[1][t] = a.e_old;
    L.in[2][t] = a.delvc;
    L.in[3][t] = a.p_old;
    L.in[4][t] = a.q_old;
    L.in[5][t] = a.qquad;
    L.in[6][t] = a.qlin;
    __syncthreads();
    for (int w = t; w < total; w += 256) {
        // the thread whose evaluations [end - extra, end) hold number w: the
        // count of threads whose range ends at or before w (end is
        // non-decreasing and end[255] = total > w, so src <= 255)
        int src = 0;
#pragma unroll
        for (int step = 128; step >= 1; step >>= 1) {
            if (L.end[src + step - 1] <= w) src += step;
        }
        double b0 = L.in[0][src], b1 = L.in[1][src], b2 = L.in[2][src], b3 = L.in[3][src];
        double b4 = L.in[4][src], b5 = L.in[5][src], b6 = L.in[6][src];
        asm volatile("" : "+v"(b0), "+v"(b1), "+v"(b2), "+v"(b3), "+v"(b4), "+v"(b5),
                     "+v"(b6));
        const EosOut o = eos_element<Unit>(b0, b1, b2, b3, b4, b5, b6, ec);
        asm volatile("" ::"v"(o.p), "v"(o.e), "v"(o.q), "v"(o.ss));
    }
}

}  // namespace

namespace {

// One block of 256 elements in mesh order, el the thread's element; every
// thread of the block calls it (the minima are reduced over the block).
template <class F>
__device__ __forceinline__ void q_eos_dt_nat_block(const F &f, DevScalars *s,
                                                   const PhysParams &prm, int el,
                                                   unsigned long long *lds_c,
                                                   unsigned long long *lds_h, RepLds &L)
{
    const bool valid = el < f.num_elem;
    unsigned long long cand_c = as_bits(kBig), cand_h = as_bits(kBig);
    const EosConst ec = eos_const(prm);
    QInputs in;
    EosIn a{1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    EosOut o;
    int extra = 0;
    if (valid) {
        const int rep = f.rep[el];
        extra = rep - 1;
        int i, j, k;
        elem_ijk(f, el, i, j, k);
        q_inputs(f, prm, el, i, j, k, in);
        double qlin, qquad;
        monotonic_q(prm, in, qlin, qquad);
        a = eos_prepare(f, s, prm, el, in, qlin, qquad);
    }
    __builtin_amdgcn_s_setprio(0);  // the evaluations yield to waves still loading
    if (ec.unit_rho0) {
        o = eos_repeat<true>(a, 1, ec);
        eos_extra_evaluations<true>(a, extra, ec, L);
    } else {
        o = eos_repeat<false>(a, 1, ec);
        eos_extra_evaluations<false>(a, extra, ec, L);
    }
    __builtin_amdgcn_s_setprio(1);  // stores and the block's reduction
    if (valid) {
        eos_finish(f, prm, el, in, a.vnewc, o, cand_c, cand_h);
    }
    fold_minima(s, cand_c, cand_h, lds_c, lds_h);
}

}  // namespace

// Monotonic Q, EOS, volume update and time constraints with the elements in
// mesh order, one per thread: each block takes 256 consecutive elements, so
// the limiter's neighbour values and the element's own fields are read in
// the order they lie in memory.  The regions' EOS repetitions are shared
// over each block's threads (eos_extra_evaluations), and the blocks run in
// the order f.eos_blocks gives (costliest first, then the rest in mesh
// order, each XCD on one contiguous part of the mesh; eos_block_order).
__global__ __launch_bounds__(256) void k_q_eos_dt_nat(DevFields f, DevScalars *s,
                                                      PhysParams prm)
{
    __shared__ unsigned long long lds_c[4], lds_h[4];
    __shared__ int skip;
    __shared__ RepLds rep_lds;
    if (threadIdx.x == 0) {
        skip = (s->active == 0 || s->error != 0) ? 1 : 0;
    }
    __syncthreads();
    if (skip) {
        return;
    }
    __builtin_amdgcn_s_setprio(1);  // the element's loads first
    const int b = f.eos_blocks[blockIdx.x];
    q_eos_dt_nat_block(f, s, prm, b * int(blockDim.x) + int(threadIdx.x), lds_c, lds_h,
                       rep_lds);
}

// persistent blocks over the element chunks of all members; a chunk lies
// inside one member, so its minima go to that member's accumulators
__global__ __launch_bounds__(256) void k_q_eos_dt_batch(const BatchMember *__restrict__ members,
                                                        const BatchWork *__restrict__ work,
                                                        PhysParams prm)
{
    __shared__ unsigned long long lds_c[4], lds_h[4];
    __shared__ int skip;
    __shared__ RepLds rep_lds;
    const int n = work->num_chunks;
    const LULESH_AS_GLOBAL int32_t *chunks = (const LULESH_AS_GLOBAL int32_t *)work->eos_chunks;
    for (int c = int(blockIdx.x); c < n; c += int(gridDim.x)) {
        const BatchMember &m = members[chunks[2 * c]];
        __syncthreads();  // the previous chunk's reads of the LDS are done
        if (threadIdx.x == 0) {
            skip = member_running(m.s) ? 0 : 1;
        }
        __syncthreads();
        if (skip) {
            continue;
        }
        __builtin_amdgcn_s_setprio(1);  // the element's loads first
        q_eos_dt_nat_block(global_view(m.f), m.s, prm, chunks[2 * c + 1] + int(threadIdx.x),
                           lds_c, lds_h, rep_lds);
    }
}

__global__ __launch_bounds__(256) void k_q_eos_dt_chunks(const BatchMember *__restrict__ members,
                                                         const BatchWork *__restrict__ work,
                                                         PhysParams prm)
{
    __shared__ unsigned long long lds_c[4], lds_h[4];
    __shared__ int skip;
    __shared__ RepLds rep_lds;
    const int c = int(blockIdx.x);
    if (c >= work->num_chunks) {
        return;
    }
    const LULESH_AS_GLOBAL int32_t *chunks = (const LULESH_AS_GLOBAL int32_t *)work->eos_chunks;
    const BatchMember &m = members[chunks[2 * c]];
    if (threadIdx.x == 0) {
        skip = member_running(m.s) ? 0 : 1;
    }
    __syncthreads();
    if (skip) {
        return;
    }
    __builtin_amdgcn_s_setprio(1);  // the element's loads first
    q_eos_dt_nat_block(global_view(m.f), m.s, prm, chunks[2 * c + 1] + int(threadIdx.x),
                       lds_c, lds_h, rep_lds);
}

// ---------------------------------------------------------------------------
// The learned closure, in one of two forms (LearnedForm).
//
// Factors (form 0): the network sees twelve dimensionless inputs formed
// from what the monotonic limiter works from and its result:
//   0-2   the limiter values phi of the xi, eta, zeta directions
//   3-5   the velocity jumps du (delv * delx, <= 0) over the sound speed c,
//         as du / (c + |du|)
//   6-8   delx of each direction over their mean hbar
//   9     vdov * hbar / (c + |vdov| hbar)
//   10    log of the relative volume vnew (the density is m / (V0 vnew))
//   11    log of the sound speed c of the last EOS evaluation
// and returns two factors on the monotonic limiter's linear and quadratic
// Q of the element, ql = m_l ql_limiter, qq = m_q qq_limiter.
//
// Stand-in (form 1): the network takes the limiter's place.  It sees the
// limiter's inputs rather than its result:
//   0-5   the velocity gradients across the element's faces relative to its
//         own, r = delv_face / delv (low and high face of xi, eta, zeta; the
//         boundary rules of the limiter applied), as r / (1 + |r|)
//   6-8   the velocity jumps du / (c + |du|)
//   9     vdov * hbar / (c + |vdov| hbar)
//   10    log of the relative volume vnew (the density is m / (V0 vnew))
//   11    log of the sound speed c
// and its two outputs are the linear and quadratic limiting factors
// themselves: ql = -qlc rho m_l sum du, qq = qqc rho m_q sum du^2 (the
// plain limiter's factors are sum du (1 - phi) / sum du and
// sum du^2 (1 - phi^2) / sum du^2); the limiter is evaluated only for the
// elements that fall back to it.
//
// In both forms each output is clamped to [0, 3], and an element keeps the
// plain limiter when an input lies more than z_max standard deviations from
// the training mean, when its volume or energy is not positive, or when it
// has no compression to limit.
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ void learned_inputs(const DevFields &f, const QInputs &in,
                                               int el, int form, float feat[kMlpIn],
                                               double &ql_unlim, double &qq_unlim,
                                               bool &usable)
{
    const double c = f.ss[el];
    const double hbar = (f.delx_xi[el] + f.delx_eta[el] + f.delx_zeta[el]) * (1.0 / 3.0);
    const double delx[3] = {f.delx_xi[el], f.delx_eta[el], f.delx_zeta[el]};
    const double self[3] = {f.delv_xi[el], f.delv_eta[el], f.delv_zeta[el]};
    if (form == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            feat[d] = float(in.phi[d]);
            feat[3 + d] = float(in.du[d] / (c + fabs(in.du[d]) + 1e-300));
            feat[6 + d] = float(delx[d] / hbar);
        }
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double norm = 1.0 / (self[d] + 1.e-36);
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const double r = in.face[d][side] * norm;
                feat[2 * d + side] = float(r / (1.0 + fabs(r)));
            }
            feat[6 + d] = float(in.du[d] / (c + fabs(in.du[d]) + 1e-300));
        }
    }
    feat[9] = float(in.vdov * hbar / (c + fabs(in.vdov) * hbar + 1e-300));
    feat[10] = float(log(in.vnew > 0.0 ? in.vnew : 1.0));
    feat[11] = float(log(c > 0.0 ? c : 1e-300));
    const double sdu = in.du[0] + in.du[1] + in.du[2];
    const double sdu2 = in.du[0] * in.du[0] + in.du[1] * in.du[1] + in.du[2] * in.du[2];
    ql_unlim = -in.rho * sdu;
    qq_unlim = in.rho * sdu2;
    usable = !(in.vdov > 0.) && sdu < 0.0 && in.vnew > 0.0 && f.e[el] >= 0.0 &&
             hbar > 0.0 && c > 0.0;
}

}  // namespace

// features of every element for training: feat[el][12] and
// q[el][4] = (unlimited ql / qlc, unlimited qq / qqc, plain ql, plain qq);
// the unlimited values are 0 where the element has no compression
__global__ __launch_bounds__(256) void k_q_features(DevFields f, PhysParams prm, int form,
                                                    float *feat, double *q)
{
    const int el = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    if (el >= f.num_elem) {
        return;
    }
    int i, j, k;
    elem_ijk(f, el, i, j, k);
    QInputs in;
    q_inputs(f, prm, el, i, j, k, in);
    double qlin, qquad;
    monotonic_q(prm, in, qlin, qquad);
    float x[kMlpIn];
    double ql_u, qq_u;
    bool usable;
    learned_inputs(f, in, el, form, x, ql_u, qq_u, usable);
    for (int t = 0; t < kMlpIn; ++t) {
        feat[size_t(el) * kMlpIn + t] = x[t];
    }
    q[size_t(el) * 4 + 0] = usable ? ql_u : 0.0;
    q[size_t(el) * 4 + 1] = usable ? qq_u : 0.0;
    q[size_t(el) * 4 + 2] = qlin;
    q[size_t(el) * 4 + 3] = qquad;
}

// k_q_eos_dt with the learned closure in place of the limiter: every wave
// evaluates the network on its 64 elements on the matrix cores
__global__ __launch_bounds__(256) void k_q_eos_dt_learned(DevFields f, DevScalars *s,
                                                          PhysParams prm,
                                                          const MlpFrags *net,
                                                          float z_max, int form)
{
    __shared__ unsigned long long lds_c[4], lds_h[4];
    __shared__ int skip;
    __shared__ MlpWaveLds lds[4];
    if (threadIdx.x == 0) {
        skip = (s->active == 0 || s->error != 0) ? 1 : 0;
    }
    __syncthreads();
    if (skip) {
        return;
    }
    const int lane = int(threadIdx.x) & 63;
    MlpWaveLds &L = lds[threadIdx.x >> 6];
    const MlpFrags &F = *net;
    const int chunk = int(gridDim.x) - 1 - int(blockIdx.x);
    const int slot = chunk * int(blockDim.x) + int(threadIdx.x);
    const bool valid = slot < f.num_elem;
    int el = 0, rep = 0;
    QInputs in;
    double qlin = 0.0, qquad = 0.0, ql_u = 0.0, qq_u = 0.0;
    bool use_net = false;
    float x[kMlpIn];
#pragma unroll
    for (int t = 0; t < kMlpIn; ++t) {
        x[t] = 0.f;
    }
    if (valid) {
        el = f.eos_order[slot];
        rep = f.eos_rep[slot];
        int i, j, k;
        elem_ijk(f, el, i, j, k);
        q_inputs(f, prm, el, i, j, k, in);
        if (form == 0) {
            monotonic_q(prm, in, qlin, qquad);
        }
        learned_inputs(f, in, el, form, x, ql_u, qq_u, use_net);
    }
    // standardise; out-of-range inputs keep the plain limiter
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        float z = 0.f;
        if (t < kMlpIn) {
            z = (x[t] - F.in_mean[t]) * F.in_inv_std[t];
            if (!(fabsf(z) <= z_max)) {
                use_net = false;
                z = 0.f;
            }
        }
        L.x[lane][t] = to_bf16(z);
    }
    __builtin_amdgcn_wave_barrier();
    mlp_forward_wave(F, L, lane);
    if (use_net) {
        const float ml = fminf(fmaxf(L.y[lane][0], 0.f), 3.f);
        const float mq = fminf(fmaxf(L.y[lane][1], 0.f), 3.f);
        if (form == 0) {
            qlin = double(ml) * qlin;
            qquad = double(mq) * qquad;
        } else {
            // the network's limiting factors on the unlimited sums
            qlin = prm.qlc_monoq * (double(ml) * ql_u);
            qquad = prm.qqc_monoq * (double(mq) * qq_u);
        }
    } else if (valid && form != 0) {
        monotonic_q(prm, in, qlin, qquad);  // the fallback: the plain limiter
    }
    unsigned long long cand_c = as_bits(kBig), cand_h = as_bits(kBig);
    if (valid) {
        eos_dt_elem(f, s, prm, el, rep, in, qlin, qquad, cand_c, cand_h);
    }
    fold_minima(s, cand_c, cand_h, lds_c, lds_h);
}

// ---------------------------------------------------------------------------
// Unstructured meshes: the same element phases, nodes and face neighbours
// from the mesh's lists
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_kinematics_q_u(DevFields f, DevMeshU m,
                                                        DevScalars *s, PhysParams prm)
{
    if (s->active == 0 || s->error != 0) {
        return;
    }
    const int el = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    const int ne = f.num_elem;
    if (el < ne) {
        int nodes[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            nodes[c] = m.nodelist[c * ne + el];
        }
        kinematics_q_nodes<false>(f, s, el, nodes);
    }
}

__global__ __launch_bounds__(256) void k_q_eos_dt_u(DevFields f, DevMeshU m, DevScalars *s,
                                                    PhysParams prm)
{
    __shared__ unsigned long long lds_c[4], lds_h[4];
    __shared__ int skip;
    if (threadIdx.x == 0) {
        skip = (s->active == 0 || s->error != 0) ? 1 : 0;
    }
    __syncthreads();
    if (skip) {
        return;
    }
    const int slot = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    unsigned long long cand_c = as_bits(kBig), cand_h = as_bits(kBig);
    if (slot < f.num_elem) {
        const int el = f.eos_order[slot];
        const int rep = f.eos_rep[slot];
        QInputs in;
        q_inputs_u(f, m, prm, el, in);
        double qlin, qquad;
        monotonic_q(prm, in, qlin, qquad);
        eos_dt_elem(f, s, prm, el, rep, in, qlin, qquad, cand_c, cand_h);
    }
    fold_minima(s, cand_c, cand_h, lds_c, lds_h);
}

// applies the minima of the last cycle: 2.0 starts every cycle from 1e20
// (lulesh-2.0.3/lulesh.cc:2616-2620); 1.0 keeps the previous constraint when
// no element gave one (reference/C/LULESH-OMP.cc:2867-2871)
__global__ void k_dt_finish(DevScalars *s, int ic_mode)
{
    if (threadIdx.x == 0) {
        apply_constraints(s, ic_mode);
    }
}

// the same for every member, and the count of members that have cycles
// left (not failed, time < stoptime, cycle < max_cycles); one block
__global__ void k_dt_finish_batch(const BatchMember *members, BatchWork *work,
                                  int ic_mode)
{
    __shared__ int pending;
    if (threadIdx.x == 0) {
        pending = 0;
    }
    __syncthreads();
    const int n = work->num_members;
    int mine = 0;
    for (int m = int(threadIdx.x); m < n; m += int(blockDim.x)) {
        DevScalars *s = members[m].s;
        apply_constraints(s, ic_mode);
        if (s->error == 0 && s->time < s->stoptime && s->cycle < s->max_cycles) {
            ++mine;
        }
    }
    if (mine) {
        atomicAdd(&pending, mine);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        work->num_pending = pending;
    }
}

// device math probe: op 0 = lulesh_cbrt(a), 1 = sqrt(a), 2 = the element
// arithmetic's quotient a / b (elem_quot: a divide here, the refined
// reciprocal in the fused build), 3 = sqrt_above_tiny(a) (the EOS's square
// root), 4 = the hourglass coefficient's a / cbrt(b) (elem_over_cbrt)
__global__ void k_device_math(int op, int n, const double *a, const double *b,
                              double *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) {
        return;
    }
    if (op == 0) {
        out[t] = lulesh_cbrt(a[t]);
    } else if (op == 1) {
        out[t] = sqrt(a[t]);
    } else if (op == 3) {
        out[t] = sqrt_above_tiny(a[t]);
    } else if (op == 4) {
        out[t] = elem_over_cbrt(a[t], b[t]);
    } else {
        out[t] = elem_quot(a[t], b[t]);
    }
}

LULESH_VARIANT_END
}  // namespace lulesh
