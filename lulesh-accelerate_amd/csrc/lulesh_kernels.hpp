// lulesh_kernels.hpp -- device data layout and kernel declarations.
//
// Layout in HBM: structure of arrays, one array per field, each allocation
// rounded up to a multiple of 64 values and aligned to 256 bytes so that a
// wavefront's 64 consecutive float64 loads are two full 256-byte lines.
// Nodes are indexed (k * (nx+1) + j) * (nx+1) + i and elements
// (k * nx + j) * nx + i, x fastest, as in the reference
// (reference/C/LULESH-OMP.cc:2996-3042), so host buffers copy 1:1.
//
// Node positions and velocities are double-buffered: the force sweep reads
// one copy and writes the other, so no block can see a neighbour's already
// advanced nodes.  DevScalars::cur names the copy holding the current state;
// TimeIncrement flips it when a cycle starts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

// The element phases (lulesh_force.hip, lulesh_kernels.hip) are compiled
// twice: with FMA contraction off, every value rounded as the reference
// rounds it (namespace lulesh), and with LULESH_FUSED defined, multiply-adds
// fused into one rounding (namespace lulesh::fused; selected per domain by
// lulesh_set_fused_multiply_add).
#ifdef LULESH_FUSED
#define LULESH_VARIANT_BEGIN namespace fused {
#define LULESH_VARIANT_END }
#else
#define LULESH_VARIANT_BEGIN
#define LULESH_VARIANT_END
#endif

namespace lulesh {

// Scalars of the cycle loop, resident in device memory so that TimeIncrement
// and the stop test run on the GPU.
struct DevScalars {
    double time;
    double deltatime;
    double dtcourant;
    double dthydro;
    double stoptime;
    double dtfixed;
    double dtmax;
    double deltatimemultlb;
    double deltatimemultub;
    unsigned long long acc_courant;  // bit pattern of this cycle's minimum
    unsigned long long acc_hydro;
    int cycle;
    int max_cycles;
    int active;        // 1 while the current cycle runs
    int error;         // 0, -1 VolumeError, -2 QStopError
    int dt_pending;  // the accumulators hold minima not yet applied
    int cur;         // copy of the node arrays holding the state
};

// Constants of the method (reference/C/LULESH-OMP.cc:3063-3087;
// lulesh-2.0.3/lulesh-init.cc:19-41).
struct PhysParams {
    double u_cut, hgcoef, qstop, monoq_max_slope, monoq_limiter_mult;
    double e_cut, p_cut, q_cut, v_cut, qlc_monoq, qqc_monoq, qqc;
    double eosvmax, eosvmin, pmin, emin, dvovmax, refdens;
    double ss_tiny, ss_floor;  // sound-speed floor (differs 1.0 / 2.0)
    int ic_mode;
};

// n / d and n % d for 0 <= n < 2^31 by multiply-high (Granlund-Montgomery):
// with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (umulhi(n, m) + n) >> l.
struct FastDiv {
    unsigned int d;
    unsigned int m;
    unsigned int l;
};

inline FastDiv make_fastdiv(unsigned int d)
{
    unsigned int l = 0;
    while ((1ull << l) < d) {
        ++l;
    }
    const unsigned long long m =
        ((1ull << 32) * ((1ull << l) - d)) / d + 1;
    return FastDiv{d, static_cast<unsigned int>(m), l};
}

__device__ __forceinline__ unsigned int fast_div(unsigned int n, const FastDiv &f)
{
    return (__umulhi(n, f.m) + n) >> f.l;
}

// The chunk of the EOS schedule that block b of g takes.  The schedule holds
// the elements sorted by region, the costly regions (more EOS repetitions)
// last; blocks take it from the end, so long blocks start first.  With
// bands > 1 the reversed schedule is cut into `bands` near-equal bands
// (the first g % bands one chunk longer) and consecutive blocks take the
// next chunk of each band in turn, so compute-bound chunks of the costly
// regions run beside memory-bound chunks of the cheap ones.  A bijection of
// [0, g) for every bands >= 1.
__device__ __forceinline__ int eos_chunk(int b, int g, int bands)
{
    const int q = g / bands, r = g - q * bands;
    int band, pos;
    if (b < q * bands) {
        band = b % bands;
        pos = b / bands;
    } else {
        band = b - q * bands;
        pos = q;
    }
    return g - 1 - (band * q + min(band, r) + pos);
}

struct DevFields {
    using CPtr = const double *;  // a read-only field (see DevFieldsG)
    int nx;
    int num_elem, num_node;
    FastDiv div_nx;
    // node-centred: positions and velocities in two copies (see above)
    double *x[2], *y[2], *z[2], *xd[2], *yd[2], *zd[2];
    double *nodal_mass;
    // element-centred
    double *e, *p, *q, *ql, *qq, *v, *volo, *elem_mass, *ss;
    double *vnew, *delv, *arealg, *vdov;
    double *delv_xi, *delv_eta, *delv_zeta, *delx_xi, *delx_eta, *delx_zeta;
    // region-ordered EOS schedule
    const int32_t *eos_order;
    const uint16_t *eos_rep;
    // EOS repetitions of each element, in mesh order
    const uint16_t *rep;
    // dispatch order of the mesh-order Q/EOS blocks (eos_block_order)
    const int32_t *eos_blocks;
};

// The same fields as pointers into the global address space, for device code
// that finds a domain's field table in memory (the ensemble's members).  A
// pointer loaded from memory is generic to the compiler, which then reaches
// every field through flat instructions, each of which counts against both
// memory counters, so that every wait for a load waits for all of them; a
// global pointer gives global loads and stores, as the single domain's kernel
// arguments do.  The element functions take either table (template class F).
#define LULESH_AS_GLOBAL __attribute__((address_space(1)))
struct DevFieldsG {
    using CPtr = const LULESH_AS_GLOBAL double *;
    int nx;
    int num_elem, num_node;
    FastDiv div_nx;
    LULESH_AS_GLOBAL double *x[2], *y[2], *z[2], *xd[2], *yd[2], *zd[2];
    LULESH_AS_GLOBAL double *nodal_mass;
    LULESH_AS_GLOBAL double *e, *p, *q, *ql, *qq, *v, *volo, *elem_mass, *ss;
    LULESH_AS_GLOBAL double *vnew, *delv, *arealg, *vdov;
    LULESH_AS_GLOBAL double *delv_xi, *delv_eta, *delv_zeta, *delx_xi, *delx_eta, *delx_zeta;
    const LULESH_AS_GLOBAL int32_t *eos_order;
    const LULESH_AS_GLOBAL uint16_t *eos_rep;
    const LULESH_AS_GLOBAL uint16_t *rep;
    const LULESH_AS_GLOBAL int32_t *eos_blocks;
};

// the global view of a field table (every pointer of it addresses device
// memory allocated by the library)
__device__ __forceinline__ DevFieldsG global_view(const DevFields &f)
{
    DevFieldsG g;
    g.nx = f.nx;
    g.num_elem = f.num_elem;
    g.num_node = f.num_node;
    g.div_nx = f.div_nx;
    for (int c = 0; c < 2; ++c) {
        g.x[c] = (LULESH_AS_GLOBAL double *)f.x[c];
        g.y[c] = (LULESH_AS_GLOBAL double *)f.y[c];
        g.z[c] = (LULESH_AS_GLOBAL double *)f.z[c];
        g.xd[c] = (LULESH_AS_GLOBAL double *)f.xd[c];
        g.yd[c] = (LULESH_AS_GLOBAL double *)f.yd[c];
        g.zd[c] = (LULESH_AS_GLOBAL double *)f.zd[c];
    }
    g.nodal_mass = (LULESH_AS_GLOBAL double *)f.nodal_mass;
    g.e = (LULESH_AS_GLOBAL double *)f.e;
    g.p = (LULESH_AS_GLOBAL double *)f.p;
    g.q = (LULESH_AS_GLOBAL double *)f.q;
    g.ql = (LULESH_AS_GLOBAL double *)f.ql;
    g.qq = (LULESH_AS_GLOBAL double *)f.qq;
    g.v = (LULESH_AS_GLOBAL double *)f.v;
    g.volo = (LULESH_AS_GLOBAL double *)f.volo;
    g.elem_mass = (LULESH_AS_GLOBAL double *)f.elem_mass;
    g.ss = (LULESH_AS_GLOBAL double *)f.ss;
    g.vnew = (LULESH_AS_GLOBAL double *)f.vnew;
    g.delv = (LULESH_AS_GLOBAL double *)f.delv;
    g.arealg = (LULESH_AS_GLOBAL double *)f.arealg;
    g.vdov = (LULESH_AS_GLOBAL double *)f.vdov;
    g.delv_xi = (LULESH_AS_GLOBAL double *)f.delv_xi;
    g.delv_eta = (LULESH_AS_GLOBAL double *)f.delv_eta;
    g.delv_zeta = (LULESH_AS_GLOBAL double *)f.delv_zeta;
    g.delx_xi = (LULESH_AS_GLOBAL double *)f.delx_xi;
    g.delx_eta = (LULESH_AS_GLOBAL double *)f.delx_eta;
    g.delx_zeta = (LULESH_AS_GLOBAL double *)f.delx_zeta;
    g.eos_order = (const LULESH_AS_GLOBAL int32_t *)f.eos_order;
    g.eos_rep = (const LULESH_AS_GLOBAL uint16_t *)f.eos_rep;
    g.rep = (const LULESH_AS_GLOBAL uint16_t *)f.rep;
    g.eos_blocks = (const LULESH_AS_GLOBAL int32_t *)f.eos_blocks;
    return g;
}

// A value every lane of the wave holds, moved to a scalar register.  A
// field table found in memory reaches the loop scalars through a generic
// pointer, whose loads land in vector registers; the parity then selects
// every field pointer per lane.  Where the value is already scalar (the
// single domain's kernel arguments) this compiles to nothing.
__device__ __forceinline__ int wave_uniform(int v)
{
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ double wave_uniform(double v)
{
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(b));
    const unsigned hi = __builtin_amdgcn_readfirstlane(unsigned(b >> 32));
    return __longlong_as_double((long long)((unsigned long long)hi << 32 | lo));
}

// One DPP step of a wavefront minimum: each lane keeps the smaller of its
// value and that of the lane the DPP control picks in its row of 16.
template <int Ctrl>
__device__ __forceinline__ unsigned long long dpp_min_step(unsigned long long v)
{
    const unsigned lo =
        unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(v)), Ctrl, 0xF, 0xF, false));
    const unsigned hi =
        unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(v >> 32)), Ctrl, 0xF, 0xF, false));
    const unsigned long long o = (unsigned long long)hi << 32 | lo;
    return o < v ? o : v;
}

// Minimum over the wavefront of 64-bit values (every lane active), exact
// and order independent: DPP moves within each row of 16 lanes (swap
// neighbours, swap pairs, mirror each half row, mirror the row), after
// which every lane holds its row's minimum, then the four rows' minima read
// from lanes 0, 16, 32 and 48.
__device__ __forceinline__ unsigned long long wave_min(unsigned long long v)
{
    v = dpp_min_step<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v = dpp_min_step<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v = dpp_min_step<0x141>(v);  // row_half_mirror
    v = dpp_min_step<0x140>(v);  // row_mirror
    unsigned long long m = ~0ull;
#pragma unroll
    for (int r = 0; r < 64; r += 16) {
        const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(unsigned(v)), r));
        const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(unsigned(v >> 32)), r));
        const unsigned long long w = (unsigned long long)hi << 32 | lo;
        m = w < m ? w : m;
    }
    return m;
}

// Minimum over the block of a value given as the bit pattern of a
// non-negative double (ordered like the values), returned to every thread;
// every thread of the block calls it.
__device__ __forceinline__ unsigned long long block_min(unsigned long long v,
                                                        unsigned long long *lds)
{
    v = wave_min(v);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[wave] = v;
    }
    __syncthreads();
    unsigned long long m = lds[0];
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 1; w < nw; ++w) {
        m = (lds[w] < m) ? lds[w] : m;
    }
    return m;
}

// One of a node field's two parity copies.  The single domain's table is a
// kernel argument and is indexed where it lies; the global view is a copy
// in registers, which a run-time index would move to scratch memory, so its
// pointer is selected instead.
__device__ __forceinline__ double *parity_copy(double *const (&t)[2], int par)
{
    return t[par];
}
__device__ __forceinline__ LULESH_AS_GLOBAL double *
parity_copy(LULESH_AS_GLOBAL double *const (&t)[2], int par)
{
    return par != 0 ? t[1] : t[0];
}

// Whether an ensemble member takes part in the cycle: active and without
// error.  Both flags are read before either is tested (a short-circuit test
// waits for the first load before it issues the second).
__device__ __forceinline__ bool member_running(const DevScalars *s)
{
    const LULESH_AS_GLOBAL DevScalars *g = (const LULESH_AS_GLOBAL DevScalars *)s;
    const int active = g->active;
    const int error = g->error;
    return (active != 0) & (error == 0);
}

// The mixed-precision domain: the node arrays, the energy, the volumes and
// the masses stay float64; the element-local quantities of the stress,
// hourglass, Q and EOS arithmetic (pressure, q and its linear / quadratic
// parts, sound speed, the kinematic rates and the Q gradients) are float32.
struct DevFieldsMixed {
    int nx;
    int num_elem, num_node;
    FastDiv div_nx;
    double *x[2], *y[2], *z[2], *xd[2], *yd[2], *zd[2];
    double *nodal_mass;
    double *e, *v, *volo, *elem_mass, *vnew;
    float *p, *q, *ql, *qq, *ss, *delv, *arealg, *vdov;
    float *delv_xi, *delv_eta, *delv_zeta, *delx_xi, *delx_eta, *delx_zeta;
    const int32_t *eos_order;
    const uint16_t *eos_rep;
    const uint16_t *rep;
    const int32_t *eos_blocks;
};

// Launch shape of the force sweep: TX x TY element columns per block (one
// thread each), NZ node planes per block, and the number of waves per SIMD
// the kernel's register budget must leave room for.
struct SweepShape {
    int tx, ty, nz, waves;
};

// One member of an ensemble: its fields and loop scalars.  Members share the
// method constants (PhysParams) and IC mode; each keeps its own time, dt and
// stop test.
struct BatchMember {
    DevFields f;
    DevScalars *s;
};

// Work lists of the persistent ensemble kernels, in device memory.  The
// host rewrites them (between cycles) when members finish, so the kernels
// read the counts from here rather than from their launch arguments.
constexpr int kChunk = 256;  // elements per chunk = threads per element block
struct BatchWork {
    int num_members;
    int num_tiles;    // force-sweep tiles (member, bx, by, bz)
    int num_chunks;   // element chunks (member, first element)
    int num_pending;  // members with cycles left, set by k_dt_finish_batch
    int4 *tiles;
    int2 *chunks;      // in mesh order, for the kinematics
    int2 *eos_chunks;  // the same chunks in the Q/EOS dispatch order
};

__global__ void k_time_increment(DevScalars *s, int ic_mode);
__global__ void k_kinematics_q(DevFields f, DevScalars *s, PhysParams prm);
__global__ void k_q_eos_dt(DevFields f, DevScalars *s, PhysParams prm, int bands);
__global__ void k_q_eos_dt_nat(DevFields f, DevScalars *s, PhysParams prm);
__global__ void k_dt_finish(DevScalars *s, int ic_mode);
__global__ void k_device_math(int op, int n, const double *a, const double *b,
                              double *out);

// ensemble forms of the same phases
__global__ void k_time_increment_batch(const BatchMember *members,
                                       const BatchWork *work, int ic_mode);
__global__ void k_kinematics_q_batch(const BatchMember *members,
                                     const BatchWork *work, PhysParams prm);
__global__ void k_q_eos_dt_batch(const BatchMember *members, const BatchWork *work,
                                 PhysParams prm);
__global__ void k_kinematics_q_chunks(const BatchMember *members, const BatchWork *work,
                                      PhysParams prm);
__global__ void k_q_eos_dt_chunks(const BatchMember *members, const BatchWork *work,
                                  PhysParams prm);
__global__ void k_dt_finish_batch(const BatchMember *members, BatchWork *work,
                                  int ic_mode);

// Launches the fused force / nodal-update sweep with the given shape.
hipError_t launch_force_sweep(const SweepShape &shape, const DevFields &f,
                              DevScalars *s, const PhysParams &prm,
                              hipStream_t stream);
// The ensemble sweep: `blocks` persistent blocks over work->tiles, which
// must have been cut for the same shape (tx, ty, nz).
hipError_t launch_force_sweep_batch(const SweepShape &shape, int blocks,
                                    const BatchMember *members, const BatchWork *work,
                                    const PhysParams &prm, hipStream_t stream);
bool force_sweep_shape_supported(const SweepShape &shape);

// Unstructured hexahedral meshes (the .lmesh input of the reference's CUDA
// port): explicit connectivity and boundary masks.  Corner-major layouts
// ([c][num_elem]) keep a wavefront's accesses contiguous.
struct DevMeshU {
    const int32_t *nodelist;      // [8][num_elem]
    const int32_t *nbr;           // [6][num_elem]: xi-, xi+, eta-, eta+, zeta-, zeta+
    const int32_t *elem_bc;       // [num_elem] face masks (kBc*)
    const int32_t *corner_start;  // [num_node + 1]: each node's corners in corner-list order
    const uint8_t *symm;          // [num_node] bit 0 / 1 / 2: x / y / z symmetry plane
    // The boundary nodes' corner forces, in their corner lists' order: slot
    // j - corner_base for corner-list entry j, 4 doubles (x, y, z, 0) each,
    // so that every corner is one whole 32-byte sector.
    const int32_t *corner_slot;   // [8][num_elem] slot of each corner, -1: interior node
    double *bs;                   // stress [slot][4]
    double *bh;                   // hourglass [slot][4]
    int corner_base;
};

// element face boundary masks (reference/C/LULESH-OMP.cc:532-554)
constexpr int kBcXiMSymm = 0x001, kBcXiMFree = 0x002, kBcXiPSymm = 0x004,
              kBcXiPFree = 0x008, kBcEtaMSymm = 0x010, kBcEtaMFree = 0x020,
              kBcEtaPSymm = 0x040, kBcEtaPFree = 0x080, kBcZetaMSymm = 0x100,
              kBcZetaMFree = 0x200, kBcZetaPSymm = 0x400, kBcZetaPFree = 0x800;

// The element blocks of an unstructured mesh in device memory (ForceBlocksU,
// lulesh_mesh.hpp), in the renumbered mesh: block k is elements
// [elem_start[k], elem_start[k + 1]) with interior nodes [node_start[k],
// node_start[k + 1]); the boundary nodes are [first_boundary, num_node).
struct DevForceBlocksU {
    const int32_t *elem_start;     // [num_blocks + 1]
    const int32_t *node_start;     // [num_blocks + 1]
    const int32_t *ncorner_start;  // [first_boundary + 1], into ncorners
    const uint16_t *ncorners;      // 8 * (element - elem_start[k]) + corner
    int num_blocks, first_boundary, num_boundary;
};

__global__ void k_forces_u_blocks(DevFields f, DevMeshU m, DevForceBlocksU b, DevScalars *s,
                                  PhysParams prm);
__global__ void k_nodes_u(DevFields f, DevMeshU m, DevScalars *s, PhysParams prm, int first,
                          int count);
__global__ void k_kinematics_q_u(DevFields f, DevMeshU m, DevScalars *s, PhysParams prm);
__global__ void k_q_eos_dt_u(DevFields f, DevMeshU m, DevScalars *s, PhysParams prm);

// The learned closure (network in lulesh_mlp.hpp).
struct MlpFrags;
__global__ void k_q_features(DevFields f, PhysParams prm, int form, float *feat, double *q);
__global__ void k_q_eos_dt_learned(DevFields f, DevScalars *s, PhysParams prm,
                                   const MlpFrags *net, float z_max, int form);

// The mixed-precision phases (the time step kernels are shared).
__global__ void k_kinematics_q_mixed(DevFieldsMixed f, DevScalars *s, PhysParams prm);
__global__ void k_q_eos_dt_mixed(DevFieldsMixed f, DevScalars *s, PhysParams prm, int bands);
__global__ void k_q_eos_dt_mixed_nat(DevFieldsMixed f, DevScalars *s, PhysParams prm);
hipError_t launch_force_sweep_mixed(const SweepShape &shape, const DevFieldsMixed &f,
                                    DevScalars *s, const PhysParams &prm,
                                    hipStream_t stream);
bool force_sweep_mixed_shape_supported(const SweepShape &shape);

// The float64 phases with multiply-adds fused (the same sources built with
// LULESH_FUSED).
namespace fused {
__global__ void k_kinematics_q(DevFields f, DevScalars *s, PhysParams prm);
__global__ void k_q_eos_dt(DevFields f, DevScalars *s, PhysParams prm, int bands);
__global__ void k_q_eos_dt_nat(DevFields f, DevScalars *s, PhysParams prm);
__global__ void k_q_eos_dt_learned(DevFields f, DevScalars *s, PhysParams prm,
                                   const MlpFrags *net, float z_max, int form);
__global__ void k_device_math(int op, int n, const double *a, const double *b,
                              double *out);
hipError_t launch_force_sweep(const SweepShape &shape, const DevFields &f,
                              DevScalars *s, const PhysParams &prm,
                              hipStream_t stream);
__global__ void k_kinematics_q_batch(const BatchMember *members,
                                     const BatchWork *work, PhysParams prm);
__global__ void k_q_eos_dt_batch(const BatchMember *members, const BatchWork *work,
                                 PhysParams prm);
__global__ void k_kinematics_q_chunks(const BatchMember *members, const BatchWork *work,
                                      PhysParams prm);
__global__ void k_q_eos_dt_chunks(const BatchMember *members, const BatchWork *work,
                                  PhysParams prm);
hipError_t launch_force_sweep_batch(const SweepShape &shape, int blocks,
                                    const BatchMember *members, const BatchWork *work,
                                    const PhysParams &prm, hipStream_t stream);
__global__ void k_forces_u_blocks(DevFields f, DevMeshU m, DevForceBlocksU b, DevScalars *s,
                                  PhysParams prm);
__global__ void k_nodes_u(DevFields f, DevMeshU m, DevScalars *s, PhysParams prm, int first,
                          int count);
__global__ void k_kinematics_q_u(DevFields f, DevMeshU m, DevScalars *s, PhysParams prm);
__global__ void k_q_eos_dt_u(DevFields f, DevMeshU m, DevScalars *s, PhysParams prm);
}  // namespace fused

}  // namespace lulesh
