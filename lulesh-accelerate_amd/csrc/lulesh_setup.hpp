// lulesh_setup.hpp -- placing a host-built domain in device memory.
//
// Shared by the single-domain and the ensemble entry points.  A domain's
// arrays are carved out of one arena: 13 node arrays (two copies of x..zd,
// the nodal mass), 19 element arrays, the region-ordered EOS schedule and
// the loop scalars, each array padded to a multiple of 64 values so every
// array starts on a 512-byte boundary.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "lulesh_kernels.hpp"
#include "lulesh_mesh.hpp"

namespace lulesh {

// Constants of the method for the given IC mode (reference/C/
// LULESH-OMP.cc:3063-3087; lulesh-2.0.3/lulesh-init.cc:19-41).
PhysParams make_params(int ic_mode);

// Bytes of the arena holding the domain of m.
size_t domain_arena_bytes(const HostMesh &m);

// Points f and *s into the arena at base (domain_arena_bytes(m) bytes).
void carve_domain(char *base, const HostMesh &m, DevFields &f, DevScalars **s);

// Queues the copies of the mesh, ICs and schedule into f (which must be
// zeroed) and waits for them.
hipError_t upload_mesh(const HostMesh &m, const DevFields &f, hipStream_t stream);

// Device address and length of field `id` (LULESH_F_*; par: the copy of
// the node arrays holding the state); nullptr for an unknown id.
double *field_pointer(const DevFields &f, int id, int par, int64_t *count);

// The same for the mixed-precision domain (float32 element-local arrays);
// field_pointer_mixed sets *f32 when the field is stored as float32.
size_t domain_arena_bytes_mixed(const HostMesh &m);
void carve_domain_mixed(char *base, const HostMesh &m, DevFieldsMixed &f,
                        DevScalars **s);
hipError_t upload_mesh_mixed(const HostMesh &m, const DevFieldsMixed &f,
                             hipStream_t stream);
void *field_pointer_mixed(const DevFieldsMixed &f, int id, int par, int64_t *count,
                          bool *f32);

// The loop scalars before the first cycle.
DevScalars initial_scalars(const HostMesh &m, double stoptime, int max_cycles);

}  // namespace lulesh
