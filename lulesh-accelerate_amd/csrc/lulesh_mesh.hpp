// lulesh_mesh.hpp -- host-side construction of the Sedov domain.
//
// Builds, on the host, everything Domain::Domain builds before the first
// cycle (reference/C/lulesh-2.0.3/lulesh-init.cc:16-184, 189-237, 377-484,
// 486-648; reference/C/LULESH-OMP.cc:2960-3171): nodal coordinates, element
// volumes and masses, nodal masses, the initial energy and time step, and
// the region layout.  Connectivity, symmetry planes and face boundary
// conditions of the structured cube are implied by the element and node
// coordinates, so the GPU computes them instead of storing them.
#pragma once

#include <cstdint>
#include <vector>

namespace lulesh {

struct MeshConfig {
    int nx = 30;
    int ic_mode = 2;  // 1: LULESH 1.0 ICs, 2: LULESH 2.0.3 ICs
    int num_reg = 11;
    int balance = 1;
    int cost = 1;
    double efactor = 1.0;  // multiplies the initial origin energy
};

struct HostMesh {
    int nx = 0, edge_nodes = 0;
    int64_t num_elem = 0, num_node = 0;
    std::vector<double> x, y, z;        // node coordinates
    std::vector<double> nodal_mass;     // sum of V/8 of the adjacent elements
    std::vector<double> volo;           // reference volume (= element mass)
    std::vector<double> e;              // initial energy (only e(0) != 0)
    double deltatime0 = 0.0;            // initial time step
    double dtfixed = 0.0;               // negative: Courant-controlled dt
    int num_reg = 1;
    std::vector<int32_t> reg_num_list;  // region of each element, 1-based
    std::vector<int32_t> reg_elem_size; // elements per region
    std::vector<int32_t> eos_order;     // elements sorted by region, ascending
    std::vector<uint16_t> eos_rep;      // EOS repetitions at each eos_order slot
    std::vector<int32_t> region_offset; // start of each region in eos_order
};

// Builds the mesh, masses, ICs and regions; throws std::invalid_argument.
HostMesh build_mesh(const MeshConfig &cfg);

// Number of EOS evaluations for region r (lulesh-2.0.3/lulesh.cc:2413-2424).
int region_rep(int ic_mode, int num_reg, int cost, int r);

// Volume of the hexahedron with the given corner coordinates
// (reference/C/LULESH-OMP.cc:1706-1789).
double hex_volume(const double x[8], const double y[8], const double z[8]);

// The C library rand() sequence of glibc (TYPE_3 additive feedback
// generator), kept private to the library so it neither depends on nor
// disturbs the caller's rand() state.
class GlibcRand {
public:
    explicit GlibcRand(unsigned seed);
    int next();

private:
    uint32_t ring_[31];
    int pos_ = 0;
};

}  // namespace lulesh
