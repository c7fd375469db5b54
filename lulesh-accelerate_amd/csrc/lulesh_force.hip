// lulesh_force.hip -- the nodal phase of a cycle as one fused sweep.
//
// LagrangeNodal (reference/C/LULESH-OMP.cc:1685-1704): the stress forces
// (InitStressTermsForElems + IntegrateStressForElems, :606-931), the
// Flanagan-Belytschko hourglass forces (CalcHourglassControlForElems +
// CalcFBHourglassForceForElems, :1233-1558), then acceleration, symmetry
// boundary conditions, velocity and position of every node (:1613-1683).
//
// The reference writes 8 corner forces per element to memory and gathers
// them per node in ascending element order (:905-926, :1475-1496).  Here a
// block owns a (TX-1) x (TY-1) column of nodes and nz node planes.  Its TX x
// TY threads each compute one element column (the owned node columns plus a
// one-element halo on the low side), sweeping upward one element layer at a
// time.  The corner forces of a layer are staged in LDS, the three stress
// components together and then the three hourglass components, and each
// owned node folds them in the reference's order: the eight elements around
// node (I,J,K) in ascending index are (I-1,J-1,K-1) corner 6, (I,J-1,K-1) 7,
// (I-1,J,K-1) 5, (I,J,K-1) 4, (I-1,J-1,K) 2, (I,J-1,K) 3, (I-1,J,K) 1,
// (I,J,K) 0.  The first four come from the layer below and are carried in
// registers between layers, so each sum is formed exactly as the reference
// forms it: stress first (f = sum), then hourglass (f += sum), every sum
// started from 0.0.  The corner forces never reach HBM.
//
// Node planes are read from the previous copy of the node arrays through
// LDS (one global read per node per block) and the advanced nodes are
// written to the other copy.
#include "lulesh_kernels.hpp"
#include "lulesh_mesh.hpp"

#include "lulesh_elem.hpp"

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

// node-by-node stencils of the volume derivative (reference/C/LULESH-OMP.cc:
// 1008-1048)
constexpr int kStencil[8][6] = {{1, 2, 3, 4, 5, 7}, {2, 3, 0, 5, 6, 4},
                                {3, 0, 1, 6, 7, 5}, {0, 1, 2, 7, 4, 6},
                                {7, 6, 5, 0, 3, 1}, {4, 7, 6, 1, 0, 2},
                                {5, 4, 7, 2, 1, 3}, {6, 5, 4, 3, 2, 0}};

// One block's share of the sweep: node columns [bx*(TX-1), (bx+1)*(TX-1)) x
// [by*(TY-1), (by+1)*(TY-1)) and node planes [bz*nz, (bz+1)*nz) of one
// domain.
//
// The element fields of the next layer and the owned node's mass of the
// next plane are loaded one layer ahead (the barriers keep the compiler from
// moving loads up, so a load issued in the layer that uses it waits for
// HBM), and the owned node is copied to registers so that the next node
// plane is stored inside the layer's last fold: four barriers per layer.
template <int TX, int TY, class Fields>
__device__ __forceinline__ void sweep_tile(const Fields &f, DevScalars *s,
                                           const PhysParams &prm, int nz, int bx,
                                           int by, int bz)
{
    constexpr int NT = TX * TY;
    constexpr int PX = TX + 1;
    constexpr int PN = (TX + 1) * (TY + 1);
    __shared__ double plane[2][6][PN];   // node planes k and k+1: x y z xd yd zd
    __shared__ double corner[3][8][NT];  // three force components of a layer
    __shared__ int s_go, s_par;
    __shared__ double s_dt;

    const int tid = threadIdx.x;
    if (tid == 0) {
        // the flags are only changed by earlier kernels: read them once
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
    typename Fields::CPtr __restrict__ in[6] = {
        parity_copy(f.x, par ^ 1),  parity_copy(f.y, par ^ 1),  parity_copy(f.z, par ^ 1),
        parity_copy(f.xd, par ^ 1), parity_copy(f.yd, par ^ 1), parity_copy(f.zd, par ^ 1)};

    const int tx = tid % TX;
    const int ty = tid / TX;
    const int i0 = bx * (TX - 1);
    const int j0 = by * (TY - 1);
    const int kc0 = bz * nz;
    const int kc1 = min(kc0 + nz, en);  // owned node planes [kc0, kc1)
    // this thread's element column; the node column (ie, je) is owned when
    // tx, ty >= 1
    const int ie = i0 - 1 + tx;
    const int je = j0 - 1 + ty;
    const bool col_valid = ie >= 0 && ie < nx && je >= 0 && je < nx;
    const bool owner = tx >= 1 && ty >= 1 && ie <= nx && je <= nx;
    const bool hourglass = prm.hgcoef > 0.0;
    // The corner slots of a thread whose element column lies outside the
    // domain hold +0.0, so an owned node adds the four corners of a layer
    // without testing which elements exist.  Adding +0.0 leaves the sum's
    // bits as skipping the element does: the sum starts from +0.0, and a sum
    // started from +0.0 is never -0.0.  Only this thread writes these slots;
    // the first fold's barrier orders the stores before any read.
    if (!col_valid) {
#pragma unroll
        for (int t = 0; t < 3; ++t) {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                corner[t][c][tid] = 0.0;
            }
        }
    }

    // A node plane of the tile is (TX+1) x (TY+1) nodes; thread tid holds the
    // slots tid + r * NT (r < kSlots) of it.  Planes are fetched into
    // registers one layer ahead and written to LDS when they are needed, so
    // the global loads overlap the element arithmetic of the layer before.
    constexpr int kSlots = (PN + NT - 1) / NT;
    int gnode[kSlots];  // global node index of each slot in plane 0, or -1
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

    // fold the three staged components into the owned node:
    // F = (sum carried from the layer below) + corners 2,3,1,0; the carried
    // sum for the plane above restarts from corners 6,7,5,4.  The last fold
    // of a layer also moves the node planes on (`advance`): the owned node of
    // plane k is copied to registers before its barrier, since plane k's
    // slot takes plane k+2 after it.
    auto fold = [&](bool finalize, bool carry, bool layer, double *F, double *P, bool last,
                    auto &&advance) {
        if (last) {
            advance(0);
        }
        __syncthreads();
        // Wave priority rises as a barrier approaches: 2 in a fold and the
        // nodal update, 1 in the stress arithmetic and the hourglass
        // projection, 0 in the hourglass's volume derivatives.  The two
        // blocks on a CU share every SIMD and each waits at its barriers for
        // its slowest wave, so a wave near its barrier issues ahead of the
        // other block's wave far from one (150^3: 301 -> 291 us).
        __builtin_amdgcn_s_setprio(2);
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            if (finalize) {
                double acc = P[t];
                if (layer) {
                    acc += corner[t][2][tid - TX - 1];
                    acc += corner[t][3][tid - TX];
                    acc += corner[t][1][tid - 1];
                    acc += corner[t][0][tid];
                }
                F[t] = acc;
            }
            if (carry) {
                double acc = 0.0;
                acc += corner[t][6][tid - TX - 1];
                acc += corner[t][7][tid - TX];
                acc += corner[t][5][tid - 1];
                acc += corner[t][4][tid];
                P[t] = acc;
            }
        }
        if (last) {
            advance(1);
        }
        __syncthreads();
        if (last) {
            __builtin_amdgcn_s_setprio(2);  // nodal update
        } else {
            __builtin_amdgcn_s_setprio(0);  // the hourglass's volume derivatives
        }
    };

    const int base = ty * PX + tx;
    const int nodeoff[4] = {base, base + 1, base + PX + 1, base + PX};

    double P[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // the element fields of a layer (p, q, v, volo, ss, elemMass) and the
    // owned node's mass of a plane are loaded one iteration ahead too: the
    // barriers keep the compiler from moving loads up, so a load issued in
    // the iteration that uses it would wait for HBM
    double epre[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double mpre = 1.0;
    auto fetch_elem = [&](int kk) {
        if (kk < nx && col_valid) {
            const int e = (kk * nx + je) * nx + ie;
            epre[0] = f.p[e];
            epre[1] = f.q[e];
            epre[2] = f.v[e];
            epre[3] = f.volo[e];
            epre[4] = f.ss[e];
            epre[5] = f.elem_mass[e];
        }
    };
    auto fetch_mass = [&](int kk) {
        if (owner && kk >= kc0 && kk < kc1) {
            mpre = f.nodal_mass[(kk * en + je) * en + ie];
        }
    };
    // iteration k uses node planes k (slot k & 1) and, for an element layer
    // (k < nx), k+1 (the other slot); both are in LDS when it starts, and
    // plane k+2 waits in registers
    const int kbeg = (kc0 > 0) ? kc0 - 1 : 0;
    fetch_elem(kbeg);
    fetch_mass(kbeg);
    fetch_plane(kbeg);
    store_plane(kbeg & 1);
    if (kbeg < nx) {
        fetch_plane(kbeg + 1);
        store_plane((kbeg + 1) & 1);
        if (kbeg + 1 < nx && kbeg + 1 < kc1) {
            fetch_plane(kbeg + 2);
        }
    }
    __syncthreads();
    for (int k = kbeg; k < kc1; ++k) {
        const bool layer = k < nx;
        const int lo = k & 1;
        const int hi = lo ^ 1;
        const bool finalize = owner && k >= kc0;
        const bool carry = owner && layer && k + 1 < kc1;
        const bool compute = layer && col_valid;
        // this layer's element fields (p q v volo ss elemMass) and the owned
        // node's mass, prefetched one layer ago
        double ev[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            ev[i] = epre[i];
        }
        const double mass = mpre;
        fetch_elem(k + 1);
        fetch_mass(k + 1);
        auto field = [&](int i) -> double { return ev[i]; };
        // the owned node of plane k (x y z xd yd zd), kept for the update
        double own[6];
        // before the last barrier of the layer: copy the owned node out of
        // plane k's slot; after it (no thread reads plane k any more): store
        // plane k+2 there when iteration k+1 is an element layer, and fetch
        // plane k+3 for the iteration after
        auto advance = [&](int phase) {
            if (phase == 0) {
                if (finalize) {
#pragma unroll
                    for (int fld = 0; fld < 6; ++fld) {
                        own[fld] = plane[lo][fld][base];
                    }
                }
            } else if (k + 1 < kc1 && k + 1 < nx) {
                store_plane(lo);
                if (k + 2 < kc1 && k + 2 < nx) {
                    fetch_plane(k + 3);
                }
            }
        };

        double F[6];
        // element corner coordinates from the LDS planes, read where needed
        // rather than kept live across the staging
        auto coords = [&](double *x, double *y, double *z) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                x[c] = plane[lo][0][nodeoff[c]];
                y[c] = plane[lo][1][nodeoff[c]];
                z[c] = plane[lo][2][nodeoff[c]];
                x[c + 4] = plane[hi][0][nodeoff[c]];
                y[c + 4] = plane[hi][1][nodeoff[c]];
                z[c + 4] = plane[hi][2][nodeoff[c]];
            }
        };

        __builtin_amdgcn_s_setprio(1);
        // ---- stress: f = -sigma * node normal (:864-903) ----
        if (compute) {
            double x[8], y[8], z[8], pf[3][8];
            coords(x, y, z);
            const double sig = -field(0) - field(1);
            const double det = jacobian_volume(x, y, z);
            // VolumeError on the Jacobian volume, then on v, which the
            // reference tests for every element before it looks at hgcoef
            // (:1579-1584, :1540-1543)
            if (det <= 0.0 || field(2) <= 0.0) {
                atomicCAS(&s->error, 0, -1);
            }
            node_normals(pf[0], pf[1], pf[2], x, y, z);
#pragma unroll
            for (int t = 0; t < 3; ++t) {
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    corner[t][c][tid] = -(sig * pf[t][c]);
                }
            }
        }
        fold(finalize, carry, layer, F, P, !hourglass, advance);

        // ---- hourglass control (:1291-1473, :1516-1547) ----
        if (hourglass) {
            if (compute) {
                const double vv = field(2);
                const double determ = field(3) * vv;
                const double volinv = elem_recip(determ);
                // the coefficient before hourgam, which holds 32 values live
                const double root = hg_root(determ);
                const double coefficient =
                    hg_over_root(-prm.hgcoef * 0.01 * field(4) * field(5), root);
                double hourgam[8][4];
                {
                    double x[8], y[8], z[8];
                    coords(x, y, z);
                    // hourglass modes of the coordinates (:1308-1324)
                    double hm[4][3];
#pragma unroll
                    for (int i1 = 0; i1 < 4; ++i1) {
                        double hmx = x[0] * fb_gamma(i1, 0);
                        double hmy = y[0] * fb_gamma(i1, 0);
                        double hmz = z[0] * fb_gamma(i1, 0);
#pragma unroll
                        for (int n = 1; n < 8; ++n) {
                            hmx = hmx + x[n] * fb_gamma(i1, n);
                            hmy = hmy + y[n] * fb_gamma(i1, n);
                            hmz = hmz + z[n] * fb_gamma(i1, n);
                        }
                        hm[i1][0] = hmx;
                        hm[i1][1] = hmy;
                        hm[i1][2] = hmz;
                    }
                    // node by node: its volume derivative, then its row of
                    // hourgam (:1326-1356)
#pragma unroll
                    for (int n = 0; n < 8; ++n) {
                        const int *st = kStencil[n];
                        double dvx, dvy, dvz;
                        volu_der(x, y, z, st[0], st[1], st[2], st[3], st[4], st[5], dvx,
                                 dvy, dvz);
#pragma unroll
                        for (int i1 = 0; i1 < 4; ++i1) {
                            hourgam[n][i1] =
                                fb_gamma(i1, n) -
                                volinv * (dvx * hm[i1][0] + dvy * hm[i1][1] + dvz * hm[i1][2]);
                        }
                    }
                }
                __builtin_amdgcn_s_setprio(1);  // the projection, before the fold
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    double u[8];
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        u[c] = plane[lo][3 + t][nodeoff[c]];
                        u[c + 4] = plane[hi][3 + t][nodeoff[c]];
                    }
                    double h[4];
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
            fold(finalize, carry, layer, F + 3, P + 3, true, advance);
        }

        // ---- acceleration, symmetry planes, velocity, position (:1613-1683)
        if (finalize) {
            double fx = F[0], fy = F[1], fz = F[2];
            if (hourglass) {
                fx += F[3];
                fy += F[4];
                fz += F[5];
            }
            const int n = (k * en + je) * en + ie;
            double xdd = elem_quot(fx, mass);
            double ydd = elem_quot(fy, mass);
            double zdd = elem_quot(fz, mass);
            if (ie == 0) xdd = 0.0;
            if (je == 0) ydd = 0.0;
            if (k == 0) zdd = 0.0;
            double xd = own[3] + xdd * dt;
            if (fabs(xd) < prm.u_cut) xd = 0.0;
            double yd = own[4] + ydd * dt;
            if (fabs(yd) < prm.u_cut) yd = 0.0;
            double zd = own[5] + zdd * dt;
            if (fabs(zd) < prm.u_cut) zd = 0.0;
            parity_copy(f.xd, par)[n] = xd;
            parity_copy(f.yd, par)[n] = yd;
            parity_copy(f.zd, par)[n] = zd;
            parity_copy(f.x, par)[n] = own[0] + xd * dt;
            parity_copy(f.y, par)[n] = own[1] + yd * dt;
            parity_copy(f.z, par)[n] = own[2] + zd * dt;
        }
    }
}

template <int TX, int TY, int W>
__global__ __launch_bounds__(TX *TY, W) void k_force_sweep(DevFields f, DevScalars *s,
                                                        PhysParams prm, int nz)
{
    sweep_tile<TX, TY>(f, s, prm, nz, int(blockIdx.x), int(blockIdx.y),
                             int(blockIdx.z));
}

// The ensemble form: a grid walks the tile list of all members (one block
// per tile, or fewer blocks looping over it); a tile names (member, bx, by,
// bz).  Members whose run is over or failed
// leave their tiles at once (sweep_tile reads their flags).
template <int TX, int TY, int W>
__global__ __launch_bounds__(TX *TY, W) void k_force_sweep_batch(
    const BatchMember *__restrict__ members, const BatchWork *__restrict__ work, PhysParams prm, int nz)
{
    const int n = work->num_tiles;
    const LULESH_AS_GLOBAL int32_t *tiles = (const LULESH_AS_GLOBAL int32_t *)work->tiles;
    for (int t = int(blockIdx.x); t < n; t += int(gridDim.x)) {
        __syncthreads();  // the LDS of the previous tile is no longer read
        const BatchMember &m = members[tiles[4 * t]];
        sweep_tile<TX, TY>(global_view(m.f), m.s, prm, nz, tiles[4 * t + 1],
                                  tiles[4 * t + 2], tiles[4 * t + 3]);
    }
}

template <int TX, int TY, int W>
hipError_t launch_shape(int nz, const DevFields &f, DevScalars *s,
                        const PhysParams &prm, hipStream_t stream)
{
    const int en = f.nx + 1;
    const dim3 grid((en + TX - 2) / (TX - 1), (en + TY - 2) / (TY - 1),
                    (en + nz - 1) / nz);
    hipLaunchKernelGGL((k_force_sweep<TX, TY, W>), grid, dim3(TX * TY), 0, stream, f,
                       s, prm, nz);
    return hipGetLastError();
}

template <int TX, int TY, int W>
hipError_t launch_shape_batch(int nz, int blocks, const BatchMember *members,
                              const BatchWork *work, const PhysParams &prm,
                              hipStream_t stream)
{
    hipLaunchKernelGGL((k_force_sweep_batch<TX, TY, W>), dim3(blocks), dim3(TX * TY), 0,
                       stream, members, work, prm, nz);
    return hipGetLastError();
}

}  // namespace

// ---------------------------------------------------------------------------
// Unstructured meshes.  With an arbitrary connectivity a node's elements are
// not a fixed stencil.  The reference's scheme (reference/C/
// LULESH-OMP.cc:853-931, :1475-1496, :209-268) writes every element's 8
// corner forces (stress and hourglass apart) and gathers each node's corners
// in ascending element order from a corner list.  Here the elements are cut
// into blocks (ForceBlocksU, lulesh_mesh.hpp; the mesh is renumbered so that
// each block's elements and interior nodes are consecutive): a block stages
// its elements' corner forces in LDS and sums and moves the nodes whose
// elements all lie in it; only the corners of the other nodes go to device
// memory, in those nodes' corner-list order, and k_nodes_u sums and moves
// them.  Every node's sums keep the corner list's order, so the bits are the
// reference scheme's.  The element arithmetic is the sweep's, expression for
// expression.
// ---------------------------------------------------------------------------
namespace {

// one element's inputs: corner positions and velocities, p, q, v, volo, ss
// and mass
struct ElemU {
    double x[8], y[8], z[8];
    double vel[3][8];
    double p, q, v, volo, ss, emass;
};

// everything the element reads is loaded before any store, which the
// compiler must assume may alias it: one wait for memory, not three
__device__ __forceinline__ void load_elem_u(const DevFields &f, const DevMeshU &m, int par,
                                            int el, ElemU &E)
{
    const int ne = f.num_elem;
    int nodes[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        nodes[c] = m.nodelist[c * ne + el];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        E.x[c] = f.x[par][nodes[c]];
        E.y[c] = f.y[par][nodes[c]];
        E.z[c] = f.z[par][nodes[c]];
    }
    E.p = f.p[el];
    E.q = f.q[el];
    E.v = f.v[el];
    E.volo = f.volo[el];
    E.ss = f.ss[el];
    E.emass = f.elem_mass[el];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        E.vel[0][c] = f.xd[par][nodes[c]];
        E.vel[1][c] = f.yd[par][nodes[c]];
        E.vel[2][c] = f.zd[par][nodes[c]];
    }
}

// stress corner forces (IntegrateStressForElems); returns whether the
// element raises VolumeError: on the Jacobian volume, then on v, which the
// reference tests for every element before it looks at hgcoef (:1579-1584,
// :1540-1543)
__device__ __forceinline__ bool stress_corners_u(const ElemU &E, double fs[3][8])
{
    const double sig = -E.p - E.q;
    const double det = jacobian_volume(E.x, E.y, E.z);
    double pf[3][8];
    node_normals(pf[0], pf[1], pf[2], E.x, E.y, E.z);
#pragma unroll
    for (int t = 0; t < 3; ++t) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            fs[t][c] = -(sig * pf[t][c]);
        }
    }
    return det <= 0.0 || E.v <= 0.0;
}

// hourglass corner forces (CalcFBHourglassForceForElems)
__device__ __forceinline__ void hourglass_corners_u(const ElemU &E, const PhysParams &prm,
                                                    double fh[3][8])
{
    const double determ = E.volo * E.v;
    const double volinv = elem_recip(determ);
    double hourgam[8][4];
    {
        double hm[4][3];
#pragma unroll
        for (int i1 = 0; i1 < 4; ++i1) {
            double hmx = E.x[0] * fb_gamma(i1, 0);
            double hmy = E.y[0] * fb_gamma(i1, 0);
            double hmz = E.z[0] * fb_gamma(i1, 0);
#pragma unroll
            for (int n = 1; n < 8; ++n) {
                hmx = hmx + E.x[n] * fb_gamma(i1, n);
                hmy = hmy + E.y[n] * fb_gamma(i1, n);
                hmz = hmz + E.z[n] * fb_gamma(i1, n);
            }
            hm[i1][0] = hmx;
            hm[i1][1] = hmy;
            hm[i1][2] = hmz;
        }
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const int *st = kStencil[n];
            double dvx, dvy, dvz;
            volu_der(E.x, E.y, E.z, st[0], st[1], st[2], st[3], st[4], st[5], dvx, dvy, dvz);
#pragma unroll
            for (int i1 = 0; i1 < 4; ++i1) {
                hourgam[n][i1] = fb_gamma(i1, n) -
                                 volinv * (dvx * hm[i1][0] + dvy * hm[i1][1] + dvz * hm[i1][2]);
            }
        }
    }
    const double root = hg_root(determ);
    const double coefficient = hg_over_root(-prm.hgcoef * 0.01 * E.ss * E.emass, root);
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const double *u = E.vel[t];
        double h[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            h[i] = hourgam[0][i] * u[0] + hourgam[1][i] * u[1] + hourgam[2][i] * u[2] +
                   hourgam[3][i] * u[3] + hourgam[4][i] * u[4] + hourgam[5][i] * u[5] +
                   hourgam[6][i] * u[6] + hourgam[7][i] * u[7];
        }
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            fh[t][n] = coefficient * (hourgam[n][0] * h[0] + hourgam[n][1] * h[1] +
                                      hourgam[n][2] * h[2] + hourgam[n][3] * h[3]);
        }
    }
}

// a node's state before the move: mass, symmetry bits, position, velocity
struct NodeU {
    double mass, x, y, z, xd, yd, zd;
    int sym;
};

__device__ __forceinline__ NodeU load_node_u(const DevFields &f, const DevMeshU &m, int par,
                                             int n)
{
    return NodeU{f.nodal_mass[n], f.x[par][n], f.y[par][n], f.z[par][n],
                 f.xd[par][n],    f.yd[par][n], f.zd[par][n], int(m.symm[n])};
}

// acceleration, symmetry planes, velocity and position of node n from its
// force (:1613-1683), in place
__device__ __forceinline__ void move_node_u(const DevFields &f, int par, double dt,
                                            const PhysParams &prm, int n, const NodeU &nd,
                                            double fx, double fy, double fz)
{
    double xdd = elem_quot(fx, nd.mass);
    double ydd = elem_quot(fy, nd.mass);
    double zdd = elem_quot(fz, nd.mass);
    if (nd.sym & 1) xdd = 0.0;
    if (nd.sym & 2) ydd = 0.0;
    if (nd.sym & 4) zdd = 0.0;
    double xd = nd.xd + xdd * dt;
    if (fabs(xd) < prm.u_cut) xd = 0.0;
    double yd = nd.yd + ydd * dt;
    if (fabs(yd) < prm.u_cut) yd = 0.0;
    double zd = nd.zd + zdd * dt;
    if (fabs(zd) < prm.u_cut) zd = 0.0;
    f.xd[par][n] = xd;
    f.yd[par][n] = yd;
    f.zd[par][n] = zd;
    f.x[par][n] = nd.x + xd * dt;
    f.y[par][n] = nd.y + yd * dt;
    f.z[par][n] = nd.z + zd * dt;
}

// one boundary corner's force into its slot: x, y, z and a zero, the whole
// 32-byte sector in two 16-byte stores (as a 64-byte line holding the stress
// and the hourglass force of the corner, 3% slower here and 10% slower to
// gather)
__device__ __forceinline__ void store_corner(double *a, int slot, double x, double y, double z)
{
    double2 *p = reinterpret_cast<double2 *>(a + 4 * size_t(slot));
    p[0] = make_double2(x, y);
    p[1] = make_double2(z, 0.0);
}

}  // namespace

// The bit-identical build writes a boundary corner's stress and hourglass
// forces to two slot arrays, and k_nodes_u sums each kind in the node's
// corner order, then adds the two sums, as the reference does.  The fused
// build (results within the reference's tolerances, not its bits) writes
// their sum to one slot array: half the boundary bytes, out and back.
#ifdef LULESH_FUSED
constexpr bool kOneSlot = true;
#else
constexpr bool kOneSlot = false;
#endif

// One block of elements (b.elem_start[blockIdx.x] ..., consecutive), one
// element per thread, then the block's interior nodes (consecutive), one per
// thread.  The stress corners and then the hourglass corners are staged in
// LDS (96 KB, one kind at a time); a corner whose node is a boundary node
// (m.corner_slot >= 0) is also written to its slot, one whole 32-byte sector
// (a partial sector written per corner cost the kernel 70%).  Each interior
// node sums its stress corners from 0.0 in corner-list order, then its
// hourglass corners from 0.0, adds the two and moves.  The block's interior
// nodes are read by no other block's elements, so they can move while other
// blocks still compute.
__global__ __launch_bounds__(kForceBlockElems) void k_forces_u_blocks(DevFields f, DevMeshU m,
                                                                      DevForceBlocksU b,
                                                                      DevScalars *s,
                                                                      PhysParams prm)
{
    constexpr int kN = kForceBlockElems;
    __shared__ double corner[3][8][kN];
    __shared__ int skip;
    // other blocks may raise an error meanwhile: read the flags once per
    // block, so that every thread takes the same path to the barriers
    if (threadIdx.x == 0) {
        skip = (s->active == 0 || s->error != 0) ? 1 : 0;
    }
    __syncthreads();
    if (skip) {
        return;
    }
    const int blk = int(blockIdx.x), l = int(threadIdx.x);
    const int ne = f.num_elem;
    const int par = s->cur;
    const double dt = s->deltatime;
    const int e0 = b.elem_start[blk];
    const int el = e0 + l;
    const bool has = el < b.elem_start[blk + 1];
    const int node = b.node_start[blk] + l;
    const bool own = node < b.node_start[blk + 1];
    int slot[8];
    ElemU E;
    if (has) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            slot[c] = m.corner_slot[c * ne + el];
        }
        load_elem_u(f, m, par, el, E);
    }
    // the interior node's state is loaded with the element's (no other block
    // reads or moves it): its loads wait with the element's, not after the
    // sums (2.6% of the kernel)
    int cb = 0, ce = 0;
    NodeU nd{1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
    if (own) {
        cb = b.ncorner_start[node];
        ce = b.ncorner_start[node + 1];
        nd = load_node_u(f, m, par, node);
    }
    const bool hourglass = prm.hgcoef > 0.0;
    if (has) {
        double fs[3][8];
        if (stress_corners_u(E, fs)) {
            atomicCAS(&s->error, 0, -1);
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                corner[t][c][l] = fs[t][c];
            }
            if (slot[c] >= 0 && (!kOneSlot || !hourglass)) {
                store_corner(m.bs, slot[c], fs[0][c], fs[1][c], fs[2][c]);
            }
        }
    }
    __syncthreads();
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int i = cb; i < ce; ++i) {
        const int k = b.ncorners[i];
        fx += corner[0][k & 7][k >> 3];
        fy += corner[1][k & 7][k >> 3];
        fz += corner[2][k & 7][k >> 3];
    }
    if (hourglass) {
        __syncthreads();  // the stress corners are summed
        if (has) {
            double fh[3][8];
            hourglass_corners_u(E, prm, fh);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                if (kOneSlot && slot[c] >= 0) {
                    // the element's own stress corner, still in LDS
                    store_corner(m.bs, slot[c], corner[0][c][l] + fh[0][c],
                                 corner[1][c][l] + fh[1][c], corner[2][c][l] + fh[2][c]);
                }
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    corner[t][c][l] = fh[t][c];
                }
                if (!kOneSlot && slot[c] >= 0) {
                    store_corner(m.bh, slot[c], fh[0][c], fh[1][c], fh[2][c]);
                }
            }
        }
        __syncthreads();
        double hx = 0.0, hy = 0.0, hz = 0.0;
        for (int i = cb; i < ce; ++i) {
            const int k = b.ncorners[i];
            hx += corner[0][k & 7][k >> 3];
            hy += corner[1][k & 7][k >> 3];
            hz += corner[2][k & 7][k >> 3];
        }
        fx += hx;
        fy += hy;
        fz += hz;
    }
    if (own) {
        move_node_u(f, par, dt, prm, node, nd, fx, fy, fz);
    }
}

// The boundary nodes first .. first + count - 1, eight lanes per node:
// gather of their corners from their slots, acceleration, symmetry planes,
// velocity and position (:905-926, :1475-1496, :1613-1683), in place (the
// corner forces are complete before this kernel runs).  Lane q of a node's
// group loads corners q, q + 8, ... of the node's run of slots, stress and
// hourglass, so a wave's loads cover the consecutive runs of its eight nodes
// sector after sector; the group's first lane adds them up from LDS in the
// run's order and moves the node.
__global__ __launch_bounds__(256) void k_nodes_u(DevFields f, DevMeshU m, DevScalars *s,
                                                 PhysParams prm, int first, int count)
{
    constexpr int kLanes = 8;
    __shared__ double stage[6][256];  // one corner per lane: stress x y z, hourglass x y z
    if (s->active == 0 || s->error != 0) {
        return;
    }
    const int t = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    const int i = t / kLanes, q = t % kLanes, l = int(threadIdx.x);
    // a group lies in one wave and leaves whole: its lanes share i
    if (i >= count) {
        return;
    }
    const int n = first + i;
    const int par = s->cur;
    const double dt = s->deltatime;
    // the fused build's slots hold stress plus hourglass
    const bool hourglass = !kOneSlot && prm.hgcoef > 0.0;
    // the node's corners are slots b .. e - 1, consecutive
    const int b = m.corner_start[n] - m.corner_base, e = m.corner_start[n + 1] - m.corner_base;
    const int cnt = e - b;
    // every lane of the group loads the node (one line for the group), so
    // that these loads are in flight with the corners'
    const NodeU nd = load_node_u(f, m, par, n);
    // Each sum starts from 0.0 and adds the corners in slot order, as the
    // reference adds a node's corners in element order.  A slot past the
    // run adds +0.0, which leaves the sum's bits as skipping it does (a sum
    // started from +0.0 is never -0.0).
    double F[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int c0 = 0; c0 < cnt; c0 += kLanes) {
        double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (c0 + q < cnt) {
            const size_t j = 4 * size_t(b + c0 + q);
            const double2 *ps = reinterpret_cast<const double2 *>(m.bs + j);
            const double2 sxy = ps[0], szw = ps[1];
            v[0] = sxy.x;
            v[1] = sxy.y;
            v[2] = szw.x;
            if (hourglass) {
                const double2 *ph = reinterpret_cast<const double2 *>(m.bh + j);
                const double2 hxy = ph[0], hzw = ph[1];
                v[3] = hxy.x;
                v[4] = hxy.y;
                v[5] = hzw.x;
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            stage[k][l] = v[k];
        }
        // the group's lanes are in this wave, whose LDS accesses run in
        // order: only the compiler must keep the reads after the writes
        __builtin_amdgcn_wave_barrier();
        if (q == 0) {
            const int last = min(kLanes, cnt - c0);
#pragma unroll
            for (int r = 0; r < kLanes; ++r) {
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const double w = stage[k][l + r];
                    F[k] += (r < last) ? w : 0.0;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();  // read before the next round writes
    }
    if (q == 0) {
        double fx = F[0], fy = F[1], fz = F[2];
        if (hourglass) {
            fx += F[3];
            fy += F[4];
            fz += F[5];
        }
        move_node_u(f, par, dt, prm, n, nd, fx, fy, fz);
    }
}

// (TX, TY, minimum waves per SIMD the register budget must allow); the
// number of node planes per block is a launch argument
#define LULESH_SWEEP_SHAPES(X) \
    X(32, 8, 1)                \
    X(32, 8, 2)                \
    X(32, 16, 2)               \
    X(16, 16, 2)               \
    X(16, 8, 2)                \
    X(32, 4, 2)                \
    X(64, 4, 1)

bool force_sweep_shape_supported(const SweepShape &sh)
{
    if (sh.nz < 1 || sh.nz > 4096) {
        return false;
    }
#define LULESH_MATCH(a, b, w) \
    if (sh.tx == a && sh.ty == b && sh.waves == w) return true;
    LULESH_SWEEP_SHAPES(LULESH_MATCH)
#undef LULESH_MATCH
    return false;
}

hipError_t launch_force_sweep(const SweepShape &sh, const DevFields &f,
                              DevScalars *s, const PhysParams &prm,
                              hipStream_t stream)
{
    if (sh.nz < 1) {
        return hipErrorInvalidValue;
    }
#define LULESH_LAUNCH(a, b, w)                            \
    if (sh.tx == a && sh.ty == b && sh.waves == w)        \
        return launch_shape<a, b, w>(sh.nz, f, s, prm, stream);
    LULESH_SWEEP_SHAPES(LULESH_LAUNCH)
#undef LULESH_LAUNCH
    return hipErrorInvalidValue;
}

hipError_t launch_force_sweep_batch(const SweepShape &sh, int blocks,
                                    const BatchMember *members, const BatchWork *work,
                                    const PhysParams &prm, hipStream_t stream)
{
    if (sh.nz < 1 || blocks < 1) {
        return hipErrorInvalidValue;
    }
#define LULESH_LAUNCH(a, b, w)                                                  \
    if (sh.tx == a && sh.ty == b && sh.waves == w)                              \
        return launch_shape_batch<a, b, w>(sh.nz, blocks, members, work, prm, stream);
    LULESH_SWEEP_SHAPES(LULESH_LAUNCH)
#undef LULESH_LAUNCH
    return hipErrorInvalidValue;
}

LULESH_VARIANT_END
}  // namespace lulesh
