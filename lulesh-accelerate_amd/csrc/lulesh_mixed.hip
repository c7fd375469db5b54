// lulesh_mixed.hip -- the mixed-precision form of the cycle.
//
// Same phases, launch structure and data flow as the float64 step
// (lulesh_force.hip, lulesh_kernels.hip), with the element-local arithmetic
// of the stress integration, the hourglass control, the monotonic Q and the
// equation of state in float32.  What stays float64:
//   - node coordinates, velocities, masses and the whole nodal update;
//   - element volumes: the Jacobian determinant of the VolumeError check,
//     the element volume behind vnew, and the half-step Jacobian of the
//     velocity gradient (cancellation in these differences of coordinates
//     is where float32 would lose the most);
//   - the energy update of the EOS (e and its three increments);
//   - the Courant / hydro reductions (each element's candidate is widened
//     to float64 before the minimum).
// Float32 geometry is formed on coordinates relative to the element's node
// 0 (differences taken in float64, then rounded): every quantity it feeds
// is translation invariant, and the local coordinates keep the relative
// precision of the element size rather than of the domain size.
// Multiply-adds may be fused here; reciprocals, square roots and the cube
// root use the hardware float32 approximations (1 ulp).
#define LULESH_MIXED_TU 1
#include "lulesh_kernels.hpp"

#include "lulesh_elem.hpp"

#pragma clang fp contract(fast)

namespace lulesh {

namespace {

constexpr double kBig = 1.0e+20;

constexpr int kStencil[8][6] = {{1, 2, 3, 4, 5, 7}, {2, 3, 0, 5, 6, 4},
                                {3, 0, 1, 6, 7, 5}, {0, 1, 2, 7, 4, 6},
                                {7, 6, 5, 0, 3, 1}, {4, 7, 6, 1, 0, 2},
                                {5, 4, 7, 2, 1, 3}, {6, 5, 4, 3, 2, 0}};

__device__ __forceinline__ float frcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ float frsq(float v) { return __builtin_amdgcn_rsqf(v); }
__device__ __forceinline__ float fsqrt(float v) { return __builtin_amdgcn_sqrtf(v); }
// cube root of a positive normal float: 2^(log2(v) / 3)
__device__ __forceinline__ float fcbrt(float v)
{
    return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(v) * (1.0f / 3.0f));
}

__device__ __forceinline__ unsigned long long as_bits(double v)
{
    return static_cast<unsigned long long>(__double_as_longlong(v));
}

__device__ __forceinline__ void raise_error(DevScalars *s, int code)
{
    atomicCAS(&s->error, 0, code);
}

__device__ __forceinline__ void elem_ijk(const DevFieldsMixed &f, int el, int &i, int &j,
                                         int &k)
{
    const unsigned q1 = fast_div(unsigned(el), f.div_nx);
    i = el - int(q1) * f.nx;
    const unsigned q2 = fast_div(q1, f.div_nx);
    j = int(q1) - int(q2) * f.nx;
    k = int(q2);
}

// ---------------------------------------------------------------------------
// Force sweep: the tiling, staging and node order of sweep_tile
// (lulesh_force.hip), with float32 corner forces.
// ---------------------------------------------------------------------------
template <int TX, int TY>
__device__ __forceinline__ void sweep_tile_mixed(const DevFieldsMixed &f, DevScalars *s,
                                                 const PhysParams &prm, int nz, int bx,
                                                 int by, int bz)
{
    constexpr int NT = TX * TY;
    constexpr int PX = TX + 1;
    constexpr int PN = (TX + 1) * (TY + 1);
    __shared__ double plane[2][6][PN];
    __shared__ float corner[3][8][NT];
    __shared__ int s_go, s_par;
    __shared__ double s_dt;

    const int tid = threadIdx.x;
    if (tid == 0) {
        s_go = (s->active != 0 && s->error == 0) ? 1 : 0;
        s_par = s->cur;
        s_dt = s->deltatime;
    }
    __syncthreads();
    if (!s_go) {
        return;
    }
    const int par = s_par;
    const double dt = s_dt;
    const int nx = f.nx;
    const int en = nx + 1;
    const double *__restrict__ in[6] = {f.x[par ^ 1], f.y[par ^ 1], f.z[par ^ 1],
                                        f.xd[par ^ 1], f.yd[par ^ 1], f.zd[par ^ 1]};
    const int tx = tid % TX;
    const int ty = tid / TX;
    const int i0 = bx * (TX - 1);
    const int j0 = by * (TY - 1);
    const int kc0 = bz * nz;
    const int kc1 = min(kc0 + nz, en);
    const int ie = i0 - 1 + tx;
    const int je = j0 - 1 + ty;
    const bool col_valid = ie >= 0 && ie < nx && je >= 0 && je < nx;
    const bool owner = tx >= 1 && ty >= 1 && ie <= nx && je <= nx;
    const bool hourglass = prm.hgcoef > 0.0;
    // corner slots of element columns outside the domain hold +0.0, so the
    // fold adds four corners per layer without testing which elements exist
    // (adding +0.0 to a sum started from +0.0 changes no bits)
    if (!col_valid) {
#pragma unroll
        for (int t = 0; t < 3; ++t) {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                corner[t][c][tid] = 0.0f;
            }
        }
    }
    const float hg = float(-prm.hgcoef * 0.01);

    constexpr int kSlots = (PN + NT - 1) / NT;
    int gnode[kSlots];
#pragma unroll
    for (int r = 0; r < kSlots; ++r) {
        const int idx = tid + r * NT;
        const int lx = idx % PX;
        const int ly = idx / PX;
        const int gi = i0 - 1 + lx;
        const int gj = j0 - 1 + ly;
        gnode[r] = (idx < PN && gi >= 0 && gi <= nx && gj >= 0 && gj <= nx)
                       ? gj * en + gi
                       : -1;
    }
    const int plane_stride = en * en;
    double pre[kSlots][6];
    auto fetch_plane = [&](int kp) {
#pragma unroll
        for (int r = 0; r < kSlots; ++r) {
            if (gnode[r] >= 0) {
                const int n = kp * plane_stride + gnode[r];
#pragma unroll
                for (int fld = 0; fld < 6; ++fld) {
                    pre[r][fld] = in[fld][n];
                }
            } else {
#pragma unroll
                for (int fld = 0; fld < 6; ++fld) {
                    pre[r][fld] = 0.0;
                }
            }
        }
    };
    auto store_plane = [&](int slot) {
#pragma unroll
        for (int r = 0; r < kSlots; ++r) {
            const int idx = tid + r * NT;
            if (idx < PN) {
#pragma unroll
                for (int fld = 0; fld < 6; ++fld) {
                    plane[slot][fld][idx] = pre[r][fld];
                }
            }
        }
    };
    // wave priority rises as a barrier approaches, as in the float64 sweep:
    // 2 in the folds, the nodal update and the plane store, 1 in the stress
    // arithmetic and the hourglass projection, 0 in the volume derivatives
    auto fold = [&](bool finalize, bool carry, bool layer, float *F, float *P) {
        __syncthreads();
        __builtin_amdgcn_s_setprio(2);
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            if (finalize) {
                float acc = P[t];
                if (layer) {
                    acc += corner[t][2][tid - TX - 1];
                    acc += corner[t][3][tid - TX];
                    acc += corner[t][1][tid - 1];
                    acc += corner[t][0][tid];
                }
                F[t] = acc;
            }
            if (carry) {
                float acc = 0.0f;
                acc += corner[t][6][tid - TX - 1];
                acc += corner[t][7][tid - TX];
                acc += corner[t][5][tid - 1];
                acc += corner[t][4][tid];
                P[t] = acc;
            }
        }
        __syncthreads();
    };

    const int base = ty * PX + tx;
    const int nodeoff[4] = {base, base + 1, base + PX + 1, base + PX};

    float P[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    // element fields of the next layer and the owned node's mass of the next
    // plane, loaded one iteration ahead (see sweep_tile)
    float pq_pre[3] = {0.f, 0.f, 0.f};        // p, q, ss
    double vm_pre[3] = {0.0, 0.0, 0.0};       // v, volo, elemMass
    double mass_pre = 1.0;
    auto fetch_elem = [&](int kk) {
        if (kk < nx && col_valid) {
            const int e = (kk * nx + je) * nx + ie;
            pq_pre[0] = f.p[e];
            pq_pre[1] = f.q[e];
            pq_pre[2] = f.ss[e];
            vm_pre[0] = f.v[e];
            vm_pre[1] = f.volo[e];
            vm_pre[2] = f.elem_mass[e];
        }
    };
    auto fetch_mass = [&](int kk) {
        if (owner && kk >= kc0 && kk < kc1) {
            mass_pre = f.nodal_mass[(kk * en + je) * en + ie];
        }
    };
    const int kbeg = (kc0 > 0) ? kc0 - 1 : 0;
    fetch_elem(kbeg);
    fetch_mass(kbeg);
    fetch_plane(kbeg);
    store_plane(kbeg & 1);
    if (kbeg < nx) {
        fetch_plane(kbeg + 1);
    }
    for (int k = kbeg; k < kc1; ++k) {
        const bool layer = k < nx;
        __syncthreads();
        if (layer) {
            store_plane((k + 1) & 1);
        }
        __syncthreads();
        if (layer && k + 1 < nx && k + 1 < kc1) {
            fetch_plane(k + 2);
        }
        __builtin_amdgcn_s_setprio(1);
        const int lo = k & 1;
        const int hi = lo ^ 1;
        const bool finalize = owner && k >= kc0;
        const bool carry = owner && layer && k + 1 < kc1;
        const bool compute = layer && col_valid;
        const float ep = pq_pre[0], eq = pq_pre[1], ess = pq_pre[2];
        const double ev = vm_pre[0], evolo = vm_pre[1], emass = vm_pre[2];
        const double nmass = mass_pre;
        fetch_elem(k + 1);
        fetch_mass(k + 1);

        float F[6];
        // float64 corner coordinates of component t from the LDS planes
        auto corner8 = [&](int t, double *X) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                X[c] = plane[lo][t][nodeoff[c]];
                X[c + 4] = plane[hi][t][nodeoff[c]];
            }
        };
        // float32 copy relative to corner 0
        auto local8 = [&](const double *X, float *x) {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                x[c] = float(X[c] - X[0]);
            }
        };

        // ---- stress: f = -sigma * node normal ----
        if (compute) {
            float x[8], y[8], z[8], pf[3][8];
            {
                double X[8], Y[8], Z[8];
                corner8(0, X);
                corner8(1, Y);
                corner8(2, Z);
                // VolumeError on the Jacobian volume, then on v, which the
                // reference tests for every element before it looks at
                // hgcoef (reference/C/LULESH-OMP.cc:1579-1584, :1540-1543)
                if (jacobian_volume(X, Y, Z) <= 0.0 || ev <= 0.0) {
                    raise_error(s, -1);
                }
                local8(X, x);
                local8(Y, y);
                local8(Z, z);
            }
            const float sig = -ep - eq;
            node_normals(pf[0], pf[1], pf[2], x, y, z);
#pragma unroll
            for (int t = 0; t < 3; ++t) {
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    corner[t][c][tid] = -(sig * pf[t][c]);
                }
            }
        }
        fold(finalize, carry, layer, F, P);

        // ---- hourglass control ----
        if (hourglass) {
            __builtin_amdgcn_s_setprio(0);
            if (compute) {
                const double vv = ev;
                const float determ = float(evolo * vv);
                const float volinv = frcp(determ);
                float hourgam[8][4];
                {
                    float x[8], y[8], z[8];
                    {
                        double X[8];
                        corner8(0, X);
                        local8(X, x);
                        corner8(1, X);
                        local8(X, y);
                        corner8(2, X);
                        local8(X, z);
                    }
                    float hm[4][3];
#pragma unroll
                    for (int i1 = 0; i1 < 4; ++i1) {
                        float hmx = x[0] * fb_gamma<float>(i1, 0);
                        float hmy = y[0] * fb_gamma<float>(i1, 0);
                        float hmz = z[0] * fb_gamma<float>(i1, 0);
#pragma unroll
                        for (int n = 1; n < 8; ++n) {
                            hmx = hmx + x[n] * fb_gamma<float>(i1, n);
                            hmy = hmy + y[n] * fb_gamma<float>(i1, n);
                            hmz = hmz + z[n] * fb_gamma<float>(i1, n);
                        }
                        hm[i1][0] = hmx;
                        hm[i1][1] = hmy;
                        hm[i1][2] = hmz;
                    }
#pragma unroll
                    for (int n = 0; n < 8; ++n) {
                        const int *st = kStencil[n];
                        float dvx, dvy, dvz;
                        volu_der(x, y, z, st[0], st[1], st[2], st[3], st[4], st[5], dvx,
                                 dvy, dvz);
#pragma unroll
                        for (int i1 = 0; i1 < 4; ++i1) {
                            hourgam[n][i1] =
                                fb_gamma<float>(i1, n) -
                                volinv * (dvx * hm[i1][0] + dvy * hm[i1][1] + dvz * hm[i1][2]);
                        }
                    }
                }
                __builtin_amdgcn_s_setprio(1);
                const float coefficient =
                    hg * ess * float(emass) * frcp(fcbrt(determ));
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    float u[8];
                    {
                        // the hourglass modes annihilate uniform velocities:
                        // project the velocities relative to corner 0
                        double U[8];
                        corner8(3 + t, U);
                        local8(U, u);
                    }
                    float h[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        h[i] = hourgam[0][i] * u[0] + hourgam[1][i] * u[1] +
                               hourgam[2][i] * u[2] + hourgam[3][i] * u[3] +
                               hourgam[4][i] * u[4] + hourgam[5][i] * u[5] +
                               hourgam[6][i] * u[6] + hourgam[7][i] * u[7];
                    }
#pragma unroll
                    for (int n = 0; n < 8; ++n) {
                        corner[t][n][tid] =
                            coefficient * (hourgam[n][0] * h[0] + hourgam[n][1] * h[1] +
                                           hourgam[n][2] * h[2] + hourgam[n][3] * h[3]);
                    }
                }
            }
            fold(finalize, carry, layer, F + 3, P + 3);
        }

        // ---- acceleration, symmetry planes, velocity, position: float64 ----
        __builtin_amdgcn_s_setprio(2);
        if (finalize) {
            double fx = double(F[0]), fy = double(F[1]), fz = double(F[2]);
            if (hourglass) {
                fx += double(F[3]);
                fy += double(F[4]);
                fz += double(F[5]);
            }
            const int n = (k * en + je) * en + ie;
            const double minv = 1.0 / nmass;
            double xdd = fx * minv;
            double ydd = fy * minv;
            double zdd = fz * minv;
            if (ie == 0) xdd = 0.0;
            if (je == 0) ydd = 0.0;
            if (k == 0) zdd = 0.0;
            double xd = plane[lo][3][base] + xdd * dt;
            if (fabs(xd) < prm.u_cut) xd = 0.0;
            double yd = plane[lo][4][base] + ydd * dt;
            if (fabs(yd) < prm.u_cut) yd = 0.0;
            double zd = plane[lo][5][base] + zdd * dt;
            if (fabs(zd) < prm.u_cut) zd = 0.0;
            f.xd[par][n] = xd;
            f.yd[par][n] = yd;
            f.zd[par][n] = zd;
            f.x[par][n] = plane[lo][0][base] + xd * dt;
            f.y[par][n] = plane[lo][1][base] + yd * dt;
            f.z[par][n] = plane[lo][2][base] + zd * dt;
        }
    }
}

template <int TX, int TY, int W>
__global__ __launch_bounds__(TX *TY, W) void k_force_sweep_mixed(DevFieldsMixed f,
                                                              DevScalars *s,
                                                              PhysParams prm, int nz)
{
    sweep_tile_mixed<TX, TY>(f, s, prm, nz, int(blockIdx.x), int(blockIdx.y),
                             int(blockIdx.z));
}

template <int TX, int TY, int W>
hipError_t launch_mixed_shape(int nz, const DevFieldsMixed &f, DevScalars *s,
                              const PhysParams &prm, hipStream_t stream)
{
    const int en = f.nx + 1;
    const dim3 grid((en + TX - 2) / (TX - 1), (en + TY - 2) / (TY - 1),
                    (en + nz - 1) / nz);
    hipLaunchKernelGGL((k_force_sweep_mixed<TX, TY, W>), grid, dim3(TX * TY), 0, stream,
                       f, s, prm, nz);
    return hipGetLastError();
}

// limiter of one direction (reference/C/LULESH-OMP.cc:2201-2228), float32
__device__ __forceinline__ float mono_phi(float self, float delvm, float delvp,
                                          float limiter_mult, float max_slope)
{
    const float norm = frcp(self + 1.e-36f);
    delvm = delvm * norm;
    delvp = delvp * norm;
    float phi = .5f * (delvm + delvp);
    delvm *= limiter_mult;
    delvp *= limiter_mult;
    if (delvm < phi) phi = delvm;
    if (delvp < phi) phi = delvp;
    if (phi < 0.f) phi = 0.f;
    if (phi > max_slope) phi = max_slope;
    return phi;
}

// EOS of one element (reference/C/LULESH-OMP.cc:2384-2561; lulesh-2.0.3/
// lulesh.cc:2043-2231): pressures, q and sound speed in float32, the energy
// and its increments in float64
struct EosConstMixed {
    double e_cut, emin, eosvmax, eosvmin;
    float p_cut, q_cut, pmin, rho0_inv, ss_tiny, ss_floor;
    bool unit_rho0;
};

struct EosOutMixed {
    double e;
    float p, q, ss;
};

__device__ __forceinline__ float eos_sound_m(float pbvc, float e, float vv, float bvc,
                                             float p, const EosConstMixed &c)
{
    float s = pbvc * e + vv * bvc * p;
    if (!c.unit_rho0) {
        s = s * c.rho0_inv;
    }
    return (s <= c.ss_tiny) ? c.ss_floor : fsqrt(s);
}

__device__ __forceinline__ float eos_pressure_m(float bvc, double e, double vnewc,
                                                const EosConstMixed &c)
{
    float p = bvc * float(e);
    if (fabsf(p) < c.p_cut) p = 0.f;
    if (vnewc >= c.eosvmax) p = 0.f;
    if (p < c.pmin) p = c.pmin;
    return p;
}

__device__ __forceinline__ EosOutMixed eos_element_mixed(double vnewc, double e_old,
                                                         float delvc, float p_old,
                                                         float q_old, float qq_old,
                                                         float ql_old,
                                                         const EosConstMixed &c)
{
    const float c1s = 2.0f / 3.0f;
    const float vn = float(vnewc);
    float compression = frcp(vn) - 1.f;
    const float vchalf = vn - delvc * .5f;
    float comp_half_step = frcp(vchalf) - 1.f;
    if (c.eosvmin != 0.) {
        if (vnewc <= c.eosvmin) comp_half_step = compression;
    }
    if (c.eosvmax != 0.) {
        if (vnewc >= c.eosvmax) {
            p_old = 0.f;
            compression = 0.f;
            comp_half_step = 0.f;
        }
    }
    const double dv = double(delvc);
    double e_new = e_old - 0.5 * dv * double(p_old + q_old);
    if (e_new < c.emin) e_new = c.emin;

    float bvc = c1s * (comp_half_step + 1.f);
    const float pbvc = c1s;
    const float p_half_step = eos_pressure_m(bvc, e_new, vnewc, c);
    const float vhalf = frcp(1.f + comp_half_step);
    float q_new;
    if (delvc > 0.f) {
        q_new = 0.f;
    } else {
        const float ssc = eos_sound_m(pbvc, float(e_new), vhalf * vhalf, bvc, p_half_step, c);
        q_new = ssc * ql_old + qq_old;
    }
    e_new = e_new + 0.5 * dv * double(3.f * (p_old + q_old) - 4.f * (p_half_step + q_new));
    if (fabs(e_new) < c.e_cut) e_new = 0.;
    if (e_new < c.emin) e_new = c.emin;

    bvc = c1s * (compression + 1.f);
    float p_new = eos_pressure_m(bvc, e_new, vnewc, c);
    {
        float q_tilde;
        if (delvc > 0.f) {
            q_tilde = 0.f;
        } else {
            const float ssc = eos_sound_m(pbvc, float(e_new), vn * vn, bvc, p_new, c);
            q_tilde = ssc * ql_old + qq_old;
        }
        e_new = e_new - double(7.f * (p_old + q_old) - 8.f * (p_half_step + q_new) +
                               (p_new + q_tilde)) *
                            dv * (1.0 / 6.0);
        if (fabs(e_new) < c.e_cut) e_new = 0.;
        if (e_new < c.emin) e_new = c.emin;
    }
    p_new = eos_pressure_m(bvc, e_new, vnewc, c);
    if (delvc <= 0.f) {
        const float ssc = eos_sound_m(pbvc, float(e_new), vn * vn, bvc, p_new, c);
        q_new = ssc * ql_old + qq_old;
        if (fabsf(q_new) < c.q_cut) q_new = 0.f;
    }
    EosOutMixed o;
    o.ss = eos_sound_m(pbvc, float(e_new), vn * vn, bvc, p_new, c);
    o.p = p_new;
    o.e = e_new;
    o.q = q_new;
    return o;
}

}  // namespace

// ---------------------------------------------------------------------------
// Kinematics and monotonic-Q gradients (reference/C/LULESH-OMP.cc:1927-2180)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_kinematics_q_mixed(DevFieldsMixed f,
                                                            DevScalars *s,
                                                            PhysParams prm)
{
    if (s->active == 0 || s->error != 0) {
        return;
    }
    const int el = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    if (el >= f.num_elem) {
        return;
    }
    const int par = s->cur;
    const double dt = s->deltatime;
    int i, j, k;
    elem_ijk(f, el, i, j, k);
    const int en = f.nx + 1;
    const int n0 = (k * en + j) * en + i;
    const int pl = en * en;
    const int nodes[8] = {n0,      n0 + 1,      n0 + en + 1,      n0 + en,
                          n0 + pl, n0 + pl + 1, n0 + pl + en + 1, n0 + pl + en};
    double X[8], Y[8], Z[8], XV[8], YV[8], ZV[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        X[c] = f.x[par][nodes[c]];
        Y[c] = f.y[par][nodes[c]];
        Z[c] = f.z[par][nodes[c]];
        XV[c] = f.xd[par][nodes[c]];
        YV[c] = f.yd[par][nodes[c]];
        ZV[c] = f.zd[par][nodes[c]];
    }
    const double volo = f.volo[el];
    const double v_old = f.v[el];  // read before the stores below may alias it
    // all loads issued before any arithmetic: one wait for memory (as in
    // k_kinematics_q)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        asm volatile("" : "+v"(X[c]), "+v"(Y[c]), "+v"(Z[c]), "+v"(XV[c]), "+v"(YV[c]),
                     "+v"(ZV[c]));
    }
    __builtin_amdgcn_s_setprio(1);  // a wave with its data finishes first
    // float64: element volume, relative volume and its change
    const double volume = elem_volume(X, Y, Z);
    const double vnew = volume / volo;
    f.vnew[el] = vnew;
    f.delv[el] = float(vnew - v_old);
    if (vnew <= 0.0) {
        raise_error(s, -1);
    }
    float x[8], y[8], z[8], xv[8], yv[8], zv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        x[c] = float(X[c] - X[0]);
        y[c] = float(Y[c] - Y[0]);
        z[c] = float(Z[c] - Z[0]);
        xv[c] = float(XV[c] - XV[0]);
        yv[c] = float(YV[c] - YV[0]);
        zv[c] = float(ZV[c] - ZV[0]);
    }
    f.arealg[el] = characteristic_length(x, y, z, float(volume));
    {
        // velocity gradient on the half-step geometry; its Jacobian in
        // float64 on coordinates relative to corner 0
        const double dt2 = 0.5 * dt;
        double xh[8], yh[8], zh[8], b[3][4];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            xh[c] = (X[c] - X[0]) - dt2 * (XV[c] - XV[0]);
            yh[c] = (Y[c] - Y[0]) - dt2 * (YV[c] - YV[0]);
            zh[c] = (Z[c] - Z[0]) - dt2 * (ZV[c] - ZV[0]);
        }
        const double inv_det_j = 1.0 / shape_derivatives(xh, yh, zh, b);
        const float dxx = float(b[0][0]) * (xv[0] - xv[6]) + float(b[0][1]) * (xv[1] - xv[7]) +
                          float(b[0][2]) * (xv[2] - xv[4]) + float(b[0][3]) * (xv[3] - xv[5]);
        const float dyy = float(b[1][0]) * (yv[0] - yv[6]) + float(b[1][1]) * (yv[1] - yv[7]) +
                          float(b[1][2]) * (yv[2] - yv[4]) + float(b[1][3]) * (yv[3] - yv[5]);
        const float dzz = float(b[2][0]) * (zv[0] - zv[6]) + float(b[2][1]) * (zv[1] - zv[7]) +
                          float(b[2][2]) * (zv[2] - zv[4]) + float(b[2][3]) * (zv[3] - zv[5]);
        f.vdov[el] = float(inv_det_j * double(dxx + dyy + dzz));
    }

    // monotonic-Q gradients, float32
    const float ptiny = 1.e-36f;
    const float vol = float(volo * vnew);
    const float norm = frcp(vol + ptiny);
    const float dxj = -0.25f * ((x[0] + x[1] + x[5] + x[4]) - (x[3] + x[2] + x[6] + x[7]));
    const float dyj = -0.25f * ((y[0] + y[1] + y[5] + y[4]) - (y[3] + y[2] + y[6] + y[7]));
    const float dzj = -0.25f * ((z[0] + z[1] + z[5] + z[4]) - (z[3] + z[2] + z[6] + z[7]));
    const float dxi = 0.25f * ((x[1] + x[2] + x[6] + x[5]) - (x[0] + x[3] + x[7] + x[4]));
    const float dyi = 0.25f * ((y[1] + y[2] + y[6] + y[5]) - (y[0] + y[3] + y[7] + y[4]));
    const float dzi = 0.25f * ((z[1] + z[2] + z[6] + z[5]) - (z[0] + z[3] + z[7] + z[4]));
    const float dxk = 0.25f * ((x[4] + x[5] + x[6] + x[7]) - (x[0] + x[1] + x[2] + x[3]));
    const float dyk = 0.25f * ((y[4] + y[5] + y[6] + y[7]) - (y[0] + y[1] + y[2] + y[3]));
    const float dzk = 0.25f * ((z[4] + z[5] + z[6] + z[7]) - (z[0] + z[1] + z[2] + z[3]));
    float ax, ay, az, dxv, dyv, dzv;

    ax = dyi * dzj - dzi * dyj;
    ay = dzi * dxj - dxi * dzj;
    az = dxi * dyj - dyi * dxj;
    f.delx_zeta[el] = vol * frsq(ax * ax + ay * ay + az * az + ptiny);
    ax *= norm;
    ay *= norm;
    az *= norm;
    dxv = 0.25f * ((xv[4] + xv[5] + xv[6] + xv[7]) - (xv[0] + xv[1] + xv[2] + xv[3]));
    dyv = 0.25f * ((yv[4] + yv[5] + yv[6] + yv[7]) - (yv[0] + yv[1] + yv[2] + yv[3]));
    dzv = 0.25f * ((zv[4] + zv[5] + zv[6] + zv[7]) - (zv[0] + zv[1] + zv[2] + zv[3]));
    f.delv_zeta[el] = ax * dxv + ay * dyv + az * dzv;

    ax = dyj * dzk - dzj * dyk;
    ay = dzj * dxk - dxj * dzk;
    az = dxj * dyk - dyj * dxk;
    f.delx_xi[el] = vol * frsq(ax * ax + ay * ay + az * az + ptiny);
    ax *= norm;
    ay *= norm;
    az *= norm;
    dxv = 0.25f * ((xv[1] + xv[2] + xv[6] + xv[5]) - (xv[0] + xv[3] + xv[7] + xv[4]));
    dyv = 0.25f * ((yv[1] + yv[2] + yv[6] + yv[5]) - (yv[0] + yv[3] + yv[7] + yv[4]));
    dzv = 0.25f * ((zv[1] + zv[2] + zv[6] + zv[5]) - (zv[0] + zv[3] + zv[7] + zv[4]));
    f.delv_xi[el] = ax * dxv + ay * dyv + az * dzv;

    ax = dyk * dzi - dzk * dyi;
    ay = dzk * dxi - dxk * dzi;
    az = dxk * dyi - dyk * dxi;
    f.delx_eta[el] = vol * frsq(ax * ax + ay * ay + az * az + ptiny);
    ax *= norm;
    ay *= norm;
    az *= norm;
    dxv = -0.25f * ((xv[0] + xv[1] + xv[5] + xv[4]) - (xv[3] + xv[2] + xv[6] + xv[7]));
    dyv = -0.25f * ((yv[0] + yv[1] + yv[5] + yv[4]) - (yv[3] + yv[2] + yv[6] + yv[7]));
    dzv = -0.25f * ((zv[0] + zv[1] + zv[5] + zv[4]) - (zv[3] + zv[2] + zv[6] + zv[7]));
    f.delv_eta[el] = ax * dxv + ay * dyv + az * dzv;
}

// ---------------------------------------------------------------------------
// Monotonic Q, EOS, volume update and time constraints; the structure of
// k_q_eos_dt and k_q_eos_dt_nat (lulesh_kernels.hip)
// ---------------------------------------------------------------------------
namespace {

struct EosInMixed {
    double vnewc, e_old;
    float delvc, p_old, q_old, qquad, qlin;
};

// what an element keeps for its stores after the EOS
struct QKeepMixed {
    double vnew;
    float vdov, arealg;
};

__device__ __forceinline__ EosConstMixed eos_const_mixed(const PhysParams &prm)
{
    EosConstMixed ec;
    ec.e_cut = prm.e_cut;
    ec.emin = prm.emin;
    ec.eosvmax = prm.eosvmax;
    ec.eosvmin = prm.eosvmin;
    ec.p_cut = float(prm.p_cut);
    ec.q_cut = float(prm.q_cut);
    ec.pmin = float(prm.pmin);
    ec.rho0_inv = float(1.0 / prm.refdens);
    ec.unit_rho0 = prm.refdens == 1.0;
    ec.ss_tiny = float(prm.ss_tiny);
    ec.ss_floor = float(prm.ss_floor);
    return ec;
}

// The element's monotonic Q (ql / qq stored) and its side of the EOS before
// the evaluations: qstop tested on the previous cycle's q, the new and old
// relative volumes clamped and the clamped old volume tested
// (reference/C/LULESH-OMP.cc:2183-2320, :2368-2381; lulesh-2.0.3/
// lulesh.cc:2704-2753).  The element's inputs are all loaded before its
// first store.
__device__ __forceinline__ EosInMixed q_eos_prepare_mixed(const DevFieldsMixed &f,
                                                          DevScalars *s, const PhysParams &prm,
                                                          int el, QKeepMixed &kp)
{
    const int nx = f.nx;
    int i, j, k;
    elem_ijk(f, el, i, j, k);
    const int sy = nx, sz = nx * nx;

    const float dvxi = f.delv_xi[el], dveta = f.delv_eta[el], dvzeta = f.delv_zeta[el];
    // neighbours across a boundary face come from the element itself and
    // are then replaced
    const float xm = f.delv_xi[(i == 0) ? el : el - 1];
    const float xp = f.delv_xi[(i == nx - 1) ? el : el + 1];
    const float ym = f.delv_eta[(j == 0) ? el : el - sy];
    const float yp = f.delv_eta[(j == nx - 1) ? el : el + sy];
    const float zm = f.delv_zeta[(k == 0) ? el : el - sz];
    const float zp = f.delv_zeta[(k == nx - 1) ? el : el + sz];
    const float vdov = f.vdov[el];
    const double vnew = f.vnew[el];
    const double volo = f.volo[el];
    const float q_old = f.q[el];
    const double v_old = f.v[el];
    const double e_old = f.e[el];
    const float delvc = f.delv[el], p_old = f.p[el];
    const float arealg = (vdov != 0.f) ? f.arealg[el] : 0.f;
    const float lm = float(prm.monoq_limiter_mult), ms = float(prm.monoq_max_slope);
    const float phixi = mono_phi(dvxi, (i == 0) ? dvxi : xm, (i == nx - 1) ? 0.f : xp, lm, ms);
    const float phieta = mono_phi(dveta, (j == 0) ? dveta : ym, (j == nx - 1) ? 0.f : yp, lm, ms);
    const float phizeta =
        mono_phi(dvzeta, (k == 0) ? dvzeta : zm, (k == nx - 1) ? 0.f : zp, lm, ms);
    float qlin, qquad;
    if (vdov > 0.f) {
        qlin = 0.f;
        qquad = 0.f;
    } else {
        float dxi = dvxi * f.delx_xi[el];
        float deta = dveta * f.delx_eta[el];
        float dzeta = dvzeta * f.delx_zeta[el];
        if (dxi > 0.f) dxi = 0.f;
        if (deta > 0.f) deta = 0.f;
        if (dzeta > 0.f) dzeta = 0.f;
        const float rho = float(f.elem_mass[el]) * frcp(float(volo * vnew));
        qlin = -float(prm.qlc_monoq) * rho *
               (dxi * (1.f - phixi) + deta * (1.f - phieta) + dzeta * (1.f - phizeta));
        qquad = float(prm.qqc_monoq) * rho *
                (dxi * dxi * (1.f - phixi * phixi) + deta * deta * (1.f - phieta * phieta) +
                 dzeta * dzeta * (1.f - phizeta * phizeta));
    }
    f.ql[el] = qlin;
    f.qq[el] = qquad;

    if (double(q_old) > prm.qstop) {
        raise_error(s, -2);
    }
    double vnewc = vnew;
    if (prm.eosvmin != 0.) {
        if (vnewc < prm.eosvmin) vnewc = prm.eosvmin;
    }
    if (prm.eosvmax != 0.) {
        if (vnewc > prm.eosvmax) vnewc = prm.eosvmax;
    }
    double vc = v_old;
    if (prm.eosvmin != 0.) {
        if (vc < prm.eosvmin) vc = prm.eosvmin;
    }
    if (prm.eosvmax != 0.) {
        if (vc > prm.eosvmax) vc = prm.eosvmax;
    }
    if (vc <= 0.) {
        raise_error(s, -1);
    }
    kp.vnew = vnew;
    kp.vdov = vdov;
    kp.arealg = arealg;
    return EosInMixed{vnewc, e_old, delvc, p_old, q_old, qquad, qlin};
}

// `rep` evaluations of the EOS on the same inputs, as the reference repeats
// them (lulesh-2.0.3/lulesh.cc:2413-2424): the inputs pass through an empty
// asm statement at the top of every repetition, carried in the same
// registers, so each repetition is computed in full, and every result is
// marked as used.
__device__ __forceinline__ EosOutMixed eos_repeat_mixed(const EosInMixed &a, int rep,
                                                        const EosConstMixed &ec)
{
    double a0 = a.vnewc, a1 = a.e_old;
    float a2 = a.delvc, a3 = a.p_old, a4 = a.q_old, a5 = a.qquad, a6 = a.qlin;
    EosOutMixed o;
    o.e = 0.0;
    o.p = o.q = o.ss = 0.f;
    for (int r = 0; r < rep; ++r) {
        asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5),
                     "+v"(a6));
        o = eos_element_mixed(a0, a1, a2, a3, a4, a5, a6, ec);
        asm volatile("" ::"v"(o.p), "v"(o.e), "v"(o.q), "v"(o.ss));
    }
    return o;
}

// The element's side after the EOS: p, e, q, ss stored, the volume update
// (1.0 keeps the unclamped vnew, 2.0 the clamped one, lulesh-2.0.3/
// lulesh.cc:2439-2452) and the Courant and hydro candidates (:2799-2933).
__device__ __forceinline__ void eos_finish_mixed(const DevFieldsMixed &f,
                                                 const PhysParams &prm, int el,
                                                 const QKeepMixed &kp, double vnewc,
                                                 const EosOutMixed &o,
                                                 unsigned long long &cand_c,
                                                 unsigned long long &cand_h)
{
    f.p[el] = o.p;
    f.e[el] = o.e;
    f.q[el] = o.q;
    f.ss[el] = o.ss;
    double tmpv = (prm.ic_mode == 1) ? kp.vnew : vnewc;
    if (fabs(tmpv - 1.0) < prm.v_cut) tmpv = 1.0;
    f.v[el] = tmpv;
    if (prm.ic_mode == 2) {
        f.vnew[el] = vnewc;
    }
    const float vdov = kp.vdov;
    if (vdov != 0.f) {
        const float qqc2 = float(64.0 * prm.qqc * prm.qqc);
        float dtf = o.ss * o.ss;
        if (vdov < 0.f) {
            dtf = dtf + qqc2 * kp.arealg * kp.arealg * vdov * vdov;
        }
        cand_c = as_bits(double(kp.arealg * frsq(dtf)));
        cand_h = as_bits(prm.dvovmax / (double(fabsf(vdov)) + 1.e-20));
    }
}

__device__ __forceinline__ void fold_minima_mixed(DevScalars *s, unsigned long long cand_c,
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

// Per block: the EOS inputs of each thread's element, the inclusive prefix of
// the evaluations beyond each element's first, and each wave's total.
struct RepLdsMixed {
    double in_d[2][256];
    float in_f[5][256];
    int end[256];
    int wave_total[4];
};

// The EOS evaluations of a block's 256 elements beyond each one's first,
// dealt out over all 256 threads (the scheme of eos_extra_evaluations in
// lulesh_kernels.hip): the thread that owns an element evaluates it once and
// keeps that result; the other rep - 1 evaluations of every element take
// ceil(sum(rep - 1) / 256) passes of the block.  Each is computed in full
// and its result marked as used.  Every thread of the block calls it.
__device__ __forceinline__ void eos_extra_evaluations_mixed(const EosInMixed &a, int extra,
                                                            const EosConstMixed &ec,
                                                            RepLdsMixed &L)
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
    L.in_d[0][t] = a.vnewc;
    L.in_d[1][t] = a.e_old;
    L.in_f[0][t] = a.delvc;
    L.in_f[1][t] = a.p_old;
    L.in_f[2][t] = a.q_old;
    L.in_f[3][t] = a.qquad;
    L.in_f[4][t] = a.qlin;
    __syncthreads();
    for (int w = t; w < total; w += 256) {
        // the thread whose evaluations [end - extra, end) hold number w
        // (end is non-decreasing and end[255] = total > w, so src <= 255)
        int src = 0;
#pragma unroll
        for (int step = 128; step >= 1; step >>= 1) {
            if (L.end[src + step - 1] <= w) src += step;
        }
        double b0 = L.in_d[0][src], b1 = L.in_d[1][src];
        float b2 = L.in_f[0][src], b3 = L.in_f[1][src], b4 = L.in_f[2][src];
        float b5 = L.in_f[3][src], b6 = L.in_f[4][src];
        asm volatile("" : "+v"(b0), "+v"(b1), "+v"(b2), "+v"(b3), "+v"(b4), "+v"(b5),
                     "+v"(b6));
        const EosOutMixed o = eos_element_mixed(b0, b1, b2, b3, b4, b5, b6, ec);
        asm volatile("" ::"v"(o.p), "v"(o.e), "v"(o.q), "v"(o.ss));
    }
}

}  // namespace

// The region-ordered schedule: block b takes chunk eos_chunk(b) of the
// elements sorted by region, each thread running its element's repetitions.
__global__ __launch_bounds__(256) void k_q_eos_dt_mixed(DevFieldsMixed f, DevScalars *s,
                                                        PhysParams prm, int bands)
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
    const int slot = eos_chunk(int(blockIdx.x), int(gridDim.x), bands) * int(blockDim.x) +
                     int(threadIdx.x);
    __builtin_amdgcn_s_setprio(1);  // the element's loads first
    unsigned long long cand_c = as_bits(kBig), cand_h = as_bits(kBig);
    if (slot < f.num_elem) {
        const int el = f.eos_order[slot];
        const int rep = f.eos_rep[slot];
        QKeepMixed kp;
        const EosInMixed a = q_eos_prepare_mixed(f, s, prm, el, kp);
        const EosConstMixed ec = eos_const_mixed(prm);
        __builtin_amdgcn_s_setprio(0);  // the repetitions yield to waves still loading
        const EosOutMixed o = eos_repeat_mixed(a, rep, ec);
        __builtin_amdgcn_s_setprio(1);  // stores and the block's reduction
        eos_finish_mixed(f, prm, el, kp, a.vnewc, o, cand_c, cand_h);
    }
    fold_minima_mixed(s, cand_c, cand_h, lds_c, lds_h);
}

// Mesh order, one element per thread, blocks of 256 consecutive elements in
// the order f.eos_blocks gives, the regions' EOS repetitions shared over
// each block's threads (as k_q_eos_dt_nat).  Each element's arithmetic is
// that of k_q_eos_dt_mixed, so both give the same bits.
__global__ __launch_bounds__(256) void k_q_eos_dt_mixed_nat(DevFieldsMixed f, DevScalars *s,
                                                            PhysParams prm)
{
    __shared__ unsigned long long lds_c[4], lds_h[4];
    __shared__ int skip;
    __shared__ RepLdsMixed rep_lds;
    if (threadIdx.x == 0) {
        skip = (s->active == 0 || s->error != 0) ? 1 : 0;
    }
    __syncthreads();
    if (skip) {
        return;
    }
    __builtin_amdgcn_s_setprio(1);  // the element's loads first
    const int el = f.eos_blocks[blockIdx.x] * int(blockDim.x) + int(threadIdx.x);
    const bool valid = el < f.num_elem;
    const EosConstMixed ec = eos_const_mixed(prm);
    unsigned long long cand_c = as_bits(kBig), cand_h = as_bits(kBig);
    QKeepMixed kp{1.0, 0.f, 0.f};
    EosInMixed a{1.0, 0.0, 0.f, 0.f, 0.f, 0.f, 0.f};
    int extra = 0;
    if (valid) {
        extra = int(f.rep[el]) - 1;
        a = q_eos_prepare_mixed(f, s, prm, el, kp);
    }
    __builtin_amdgcn_s_setprio(0);  // the evaluations yield to waves still loading
    const EosOutMixed o = eos_repeat_mixed(a, 1, ec);
    eos_extra_evaluations_mixed(a, extra, ec, rep_lds);
    __builtin_amdgcn_s_setprio(1);  // stores and the block's reduction
    if (valid) {
        eos_finish_mixed(f, prm, el, kp, a.vnewc, o, cand_c, cand_h);
    }
    fold_minima_mixed(s, cand_c, cand_h, lds_c, lds_h);
}

#define LULESH_MIXED_SHAPES(X) \
    X(32, 8, 2)                \
    X(32, 8, 3)                \
    X(16, 16, 3)               \
    X(16, 8, 3)                \
    X(32, 16, 2)

bool force_sweep_mixed_shape_supported(const SweepShape &sh)
{
    if (sh.nz < 1 || sh.nz > 4096) {
        return false;
    }
#define LULESH_MATCH(a, b, w) \
    if (sh.tx == a && sh.ty == b && sh.waves == w) return true;
    LULESH_MIXED_SHAPES(LULESH_MATCH)
#undef LULESH_MATCH
    return false;
}

hipError_t launch_force_sweep_mixed(const SweepShape &sh, const DevFieldsMixed &f,
                                    DevScalars *s, const PhysParams &prm,
                                    hipStream_t stream)
{
    if (sh.nz < 1) {
        return hipErrorInvalidValue;
    }
#define LULESH_LAUNCH(a, b, w)                     \
    if (sh.tx == a && sh.ty == b && sh.waves == w) \
        return launch_mixed_shape<a, b, w>(sh.nz, f, s, prm, stream);
    LULESH_MIXED_SHAPES(LULESH_LAUNCH)
#undef LULESH_LAUNCH
    return hipErrorInvalidValue;
}

}  // namespace lulesh
