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
    std::vector<uint16_t> elem_rep;     // EOS repetitions of each element
    std::vector<int32_t> eos_blocks;    // dispatch order of the mesh-order Q/EOS blocks
};

// Elements per block of the mesh-order Q/EOS kernels (one per thread).
constexpr int kEosBlockElems = 256;

// The order in which the blocks of a mesh-order Q/EOS launch take the mesh's
// blocks of kEosBlockElems consecutive elements (eos_block_order in
// lulesh_mesh.cpp): slot t of the launch processes block order[t].
std::vector<int32_t> eos_block_order(const std::vector<uint16_t> &rep);
// Its two parts: the EOS passes of each block (one, plus the evaluations
// beyond each element's first shared over the block), and the launch order
// of blocks of the given costs, listed in the order whose locality the
// launch should keep.
std::vector<int64_t> eos_block_passes(const std::vector<uint16_t> &rep);
std::vector<int32_t> spread_costly_blocks(const std::vector<int64_t> &passes);

// Builds the mesh, masses, ICs and regions; throws std::invalid_argument.
HostMesh build_mesh(const MeshConfig &cfg);

// An unstructured hexahedral mesh in LULESH 1.0 semantics, set up as the
// reference's .lmesh reader sets it up (reference/CUDA/
// lulesh-kepler-singlegpu/lulesh.cu:728-925): element face masks derived
// from the node sets, the energy in the octant corner, masses from the
// element volumes.  Arrays the GPU reads are corner-major ([c][num_elem]).
struct HostMeshU {
    int num_elem = 0, num_node = 0;
    std::vector<double> x, y, z, nodal_mass, volo, e;
    std::vector<int32_t> nodelist;      // [8][num_elem]
    std::vector<int32_t> nbr;           // [6][num_elem]
    std::vector<int32_t> elem_bc;       // face masks
    std::vector<int32_t> corner_start;  // [num_node + 1]
    std::vector<int32_t> corner_list;   // c * num_elem + el, per node in ascending order of
                                        // the elements' numbers in the mesh as given
    std::vector<uint8_t> symm;          // bit 0 / 1 / 2: on the x / y / z symmetry set
    int origin = 0;                     // element holding the initial energy
};

// The raw contents of a mesh: nodelist [num_elem][8] and neighbors
// [num_elem][6] element-major, as the .lmesh file lists them.
struct MeshLists {
    int num_elem = 0, num_node = 0;
    std::vector<double> x, y, z;
    std::vector<int32_t> nodelist, neighbors;
    std::vector<int32_t> symm_x, symm_y, symm_z, free_surf;
};

// Reads a .lmesh file (the format of lulesh.cu:728-848: counts, node
// coordinates, 8 nodes per element, 6 face neighbours per element, the x, y,
// z symmetry node sets and the free-surface node set, each set preceded by
// its size); throws std::invalid_argument on a malformed file.
MeshLists read_lmesh(const char *path);

// Validates the lists and builds the mesh; throws std::invalid_argument.
HostMeshU build_mesh_u(const MeshLists &l);

// Elements per block of the unstructured force kernel (one per thread).
constexpr int kForceBlockElems = 512;

// The element blocks of the unstructured force kernel (k_forces_u_blocks):
// the elements cut into blocks of at most kForceBlockElems by recursive
// coordinate bisection of their centroids.  A node whose elements all lie
// in one block is an interior node of that block (at most kForceBlockElems
// per block): the block sums its corner forces from LDS and moves it.  Every
// other node is a boundary node: its elements' corner forces for it go
// through device memory, and k_nodes_u gathers and moves it.
// The device holds the mesh renumbered (renumber_mesh_u) so that each
// block's elements, each block's interior nodes and the boundary nodes are
// contiguous: block k is elements [elem_start[k], elem_start[k + 1]) with
// interior nodes [node_start[k], node_start[k + 1]), and the boundary nodes
// are [node_start[num_blocks], num_node).
struct ForceBlocksU {
    int num_blocks = 0;
    std::vector<int32_t> elem_start;     // [num_blocks + 1], new element numbers
    std::vector<int32_t> node_start;     // [num_blocks + 1], new node numbers
    std::vector<int32_t> elem_new;       // new number of each element of the mesh
    std::vector<int32_t> node_new;       // new number of each node of the mesh
    std::vector<int32_t> ncorner_start;  // [node_start[num_blocks] + 1] into ncorners
    std::vector<uint16_t> ncorners;      // 8 * (element - elem_start[k]) + corner, in the
                                         // node's corner-list order
};

// Builds the blocks of an unstructured mesh (block_elems <= kForceBlockElems).
ForceBlocksU build_force_blocks(const HostMeshU &m, int block_elems = kForceBlockElems);

// The mesh in the blocks' numbering: every array permuted and every index
// mapped.  A node's corner list keeps its order, ascending in the mesh's own
// element numbers, so its sums are formed in the reference's order.
HostMeshU renumber_mesh_u(const HostMeshU &m, const ForceBlocksU &b);

// The boundary nodes' corner forces lie in the renumbered mesh's corner-list
// order: slot j - corner_start[first_boundary] for corner-list entry j, so
// that k_nodes_u reads each node's corners in one run (slots in element
// order, one contiguous range per block, made the block kernel 4% faster and
// the gather 35% slower).  The slot of each element corner c * num_elem + el
// (-1: its node is interior).
std::vector<int32_t> boundary_corner_slots(const HostMeshU &r, int32_t first_boundary);

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
