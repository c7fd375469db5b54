// lulesh_kernels.hpp -- device data layout and kernel declarations.
//
// Layout in HBM: structure of arrays, one array per field, each allocation
// rounded up to a multiple of 64 values and aligned to 256 bytes so that a
// wavefront's 64 consecutive float64 loads are two full 256-byte lines.
// Nodes are indexed (k * (nx+1) + j) * (nx+1) + i and elements
// (k * nx + j) * nx + i, x fastest, as in the reference
// (reference/C/LULESH-OMP.cc:2996-3042), so host buffers copy 1:1.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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
    int active;  // 1 while the current cycle runs
    int error;   // 0, -1 VolumeError, -2 QStopError
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

struct DevFields {
    int nx;
    int64_t num_elem, num_node;
    // node-centred
    double *x, *y, *z, *xd, *yd, *zd, *nodal_mass;
    // element-centred
    double *e, *p, *q, *ql, *qq, *v, *volo, *elem_mass, *ss;
    double *vnew, *delv, *arealg, *vdov;
    double *delv_xi, *delv_eta, *delv_zeta, *delx_xi, *delx_eta, *delx_zeta;
    // corner forces: 6 components x 8 corners x num_elem
    double *corner;
    // region-ordered EOS schedule
    const int32_t *eos_order;
    const uint16_t *eos_rep;
};

__global__ void k_time_increment(DevScalars *s);
__global__ void k_corner_forces(DevFields f, DevScalars *s, PhysParams prm);
__global__ void k_nodal_update(DevFields f, DevScalars *s, PhysParams prm);
__global__ void k_kinematics_q(DevFields f, DevScalars *s, PhysParams prm);
__global__ void k_q_eos_dt(DevFields f, DevScalars *s, PhysParams prm);
__global__ void k_dt_finish(DevScalars *s, int ic_mode);
__global__ void k_device_math(int op, int n, const double *a, const double *b,
                              double *out);

}  // namespace lulesh
