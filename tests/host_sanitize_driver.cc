// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
// structured mesh builder over many configurations, the .lmesh reader and
// the unstructured mesh builder on the files named on the command line, and
// short oracle runs on the cube and on the first mesh.  Built and run by
// tests/test_host_sanitizers.py; any sanitizer report aborts the program.
#include <cstdio>
#include <exception>
#include <vector>

#include "lulesh_mesh.hpp"

extern "C" {
#include "lulesh_oracle.h"
}

namespace {

int check_cube(const lulesh::MeshConfig &c)
{
    const lulesh::HostMesh m = lulesh::build_mesh(c);
    const long long ne = (long long)c.nx * c.nx * c.nx;
    if (m.num_elem != ne || (long long)m.eos_order.size() != ne ||
        (long long)m.volo.size() != ne || (long long)m.x.size() != m.num_node) {
        std::printf("cube nx=%d: sizes\n", c.nx);
        return 1;
    }
    std::vector<char> seen(size_t(ne), 0);
    for (int32_t el : m.eos_order) {
        if (el < 0 || el >= ne || seen[size_t(el)]) {
            std::printf("cube nx=%d: eos_order is not a permutation\n", c.nx);
            return 1;
        }
        seen[size_t(el)] = 1;
    }
    const size_t g = (size_t(ne) + lulesh::kEosBlockElems - 1) / lulesh::kEosBlockElems;
    std::vector<char> seen_block(g, 0);
    for (int32_t b : m.eos_blocks) {
        if (b < 0 || size_t(b) >= g || seen_block[size_t(b)]) {
            std::printf("cube nx=%d: eos_blocks is not a permutation\n", c.nx);
            return 1;
        }
        seen_block[size_t(b)] = 1;
    }
    return m.eos_blocks.size() == g ? 0 : 1;
}

// The Q/EOS block order on synthetic repetition counts: g blocks, the blocks
// in `costly` with 20 repetitions per element (20 passes), the rest with one.
// Checks: a permutation; every costly block in the first 3/4 of its XCD's
// slots (slot t goes to XCD t % 8); per XCD, the costly blocks' slot numbers
// spread over the 32 residues mod 32 (the XCD's CUs) within 1.5x the mean
// + 2; the other blocks in mesh order, each XCD one contiguous range.
int check_block_order(size_t g, const std::vector<size_t> &costly, const char *name)
{
    std::vector<uint16_t> rep(g * lulesh::kEosBlockElems, 1);
    std::vector<char> is_costly(g, 0);
    for (size_t b : costly) {
        is_costly[b] = 1;
        for (int i = 0; i < lulesh::kEosBlockElems; ++i) rep[b * lulesh::kEosBlockElems + i] = 20;
    }
    const std::vector<int32_t> order = lulesh::eos_block_order(rep);
    std::vector<int64_t> rank(g, -1);  // place among the cheap blocks, in mesh order
    for (size_t b = 0, k = 0; b < g; ++b) {
        if (!is_costly[b]) rank[b] = int64_t(k++);
    }
    std::vector<char> seen(g, 0);
    int bad = order.size() != g;
    for (int32_t b : order) {
        bad |= (b < 0 || size_t(b) >= g || seen[size_t(b)]);
        if (!bad) seen[size_t(b)] = 1;
    }
    for (size_t x = 0; x < 8 && !bad; ++x) {
        const size_t n_x = (g - x + 7) / 8;
        size_t bins[32] = {}, h_x = 0;
        int64_t last = -1;
        for (size_t r = 0; r < n_x; ++r) {
            const int32_t b = order[8 * r + x];
            if (is_costly[size_t(b)]) {
                bad |= r >= (3 * n_x + 3) / 4;
                bins[r % 32]++;
                h_x++;
            } else {
                bad |= (last >= 0 && rank[size_t(b)] != last + 1);  // contiguous, ascending
                last = rank[size_t(b)];
            }
        }
        for (size_t c = 0; c < 32; ++c) {
            bad |= double(bins[c]) > 1.5 * double(h_x) / 32.0 + 2.0;
        }
    }
    if (bad) std::printf("block order %s: failed\n", name);
    return bad;
}

// The force blocks of an unstructured mesh and the renumbered mesh: the new
// element and node numbers are permutations; every block holds at most
// kForceBlockElems elements, ascending in the mesh's numbers; each interior
// node's elements all lie in its block, and its local corner list names
// them and their corners in the corner list's order; the renumbered mesh
// holds the same connectivity, coordinates, masses and corner lists (in the
// old order) under the new numbers; every corner of a boundary node has the
// slot of its corner-list entry, and no other corner has one.
int check_force_blocks(const lulesh::HostMeshU &m, const char *name)
{
    const lulesh::ForceBlocksU b = lulesh::build_force_blocks(m);
    const lulesh::HostMeshU r = lulesh::renumber_mesh_u(m, b);
    const size_t ne = size_t(m.num_elem), nn = size_t(m.num_node);
    int bad = 0;
    std::vector<int32_t> elem_old(ne, -1), node_old(nn, -1);
    bad |= b.elem_new.size() != ne || b.node_new.size() != nn;
    for (size_t el = 0; el < ne && !bad; ++el) {
        const int32_t k = b.elem_new[el];
        bad |= k < 0 || size_t(k) >= ne || elem_old[size_t(k)] != -1;
        if (!bad) elem_old[size_t(k)] = int32_t(el);
    }
    for (size_t n = 0; n < nn && !bad; ++n) {
        const int32_t k = b.node_new[n];
        bad |= k < 0 || size_t(k) >= nn || node_old[size_t(k)] != -1;
        if (!bad) node_old[size_t(k)] = int32_t(n);
    }
    bad |= b.elem_start.size() != size_t(b.num_blocks) + 1 || b.elem_start[0] != 0 ||
           size_t(b.elem_start.back()) != ne || b.node_start.size() != b.elem_start.size() ||
           b.node_start[0] != 0;
    const int32_t first_boundary = bad ? 0 : b.node_start.back();
    const std::vector<int32_t> slots = lulesh::boundary_corner_slots(r, first_boundary);
    for (int k = 0; k < b.num_blocks && !bad; ++k) {
        const int32_t e0 = b.elem_start[size_t(k)], e1 = b.elem_start[size_t(k) + 1];
        const int32_t n0 = b.node_start[size_t(k)], n1 = b.node_start[size_t(k) + 1];
        bad |= e1 <= e0 || e1 - e0 > lulesh::kForceBlockElems ||
               n1 < n0 || n1 - n0 > lulesh::kForceBlockElems;
        for (int32_t e = e0 + 1; e < e1 && !bad; ++e) {
            bad |= elem_old[size_t(e)] <= elem_old[size_t(e) - 1];
        }
        for (int32_t n = n0; n < n1 && !bad; ++n) {
            const int32_t on = node_old[size_t(n)];
            const int32_t c0 = m.corner_start[size_t(on)], c1 = m.corner_start[size_t(on) + 1];
            bad |= c1 == c0 ||
                   b.ncorner_start[size_t(n) + 1] - b.ncorner_start[size_t(n)] != c1 - c0;
            for (int32_t j = 0; j < c1 - c0 && !bad; ++j) {
                const size_t cl = size_t(m.corner_list[size_t(c0 + j)]);
                const uint16_t lc = b.ncorners[size_t(b.ncorner_start[size_t(n)] + j)];
                bad |= e0 + lc / 8 >= e1 || elem_old[size_t(e0 + lc / 8)] != int32_t(cl % ne) ||
                       lc % 8 != cl / ne;
            }
        }
    }
    // the renumbered mesh
    bad |= r.num_elem != m.num_elem || r.num_node != m.num_node ||
           r.origin != (bad ? 0 : b.elem_new[size_t(m.origin)]);
    for (size_t n = 0; n < nn && !bad; ++n) {
        const size_t k = size_t(b.node_new[n]);
        bad |= r.x[k] != m.x[n] || r.y[k] != m.y[n] || r.z[k] != m.z[n] ||
               r.nodal_mass[k] != m.nodal_mass[n] || r.symm[k] != m.symm[n] ||
               r.corner_start[k + 1] - r.corner_start[k] !=
                   m.corner_start[n + 1] - m.corner_start[n];
        for (int32_t i = 0; !bad && i < m.corner_start[n + 1] - m.corner_start[n]; ++i) {
            const size_t cl = size_t(m.corner_list[size_t(m.corner_start[n] + i)]);
            bad |= size_t(r.corner_list[size_t(r.corner_start[k] + i)]) !=
                   (cl / ne) * ne + size_t(b.elem_new[cl % ne]);
        }
    }
    for (size_t el = 0; el < ne && !bad; ++el) {
        const size_t k = size_t(b.elem_new[el]);
        bad |= r.volo[k] != m.volo[el] || r.e[k] != m.e[el] || r.elem_bc[k] != m.elem_bc[el];
        for (size_t c = 0; c < 8; ++c) {
            const int32_t node = m.nodelist[c * ne + el];
            bad |= r.nodelist[c * ne + k] != b.node_new[size_t(node)];
            // a boundary node's corner has the slot of its corner-list entry
            const int32_t sl = slots[c * ne + k];
            const int32_t nn_new = b.node_new[size_t(node)];
            if (nn_new < first_boundary) {
                bad |= sl != -1;
            } else {
                const int32_t j = sl + r.corner_start[size_t(first_boundary)];
                bad |= sl < 0 || j < r.corner_start[size_t(nn_new)] ||
                       j >= r.corner_start[size_t(nn_new) + 1] ||
                       size_t(r.corner_list[size_t(j)]) != c * ne + k;
            }
        }
        for (size_t f = 0; f < 6; ++f) {
            bad |= r.nbr[f * ne + k] != b.elem_new[size_t(m.nbr[f * ne + el])];
        }
    }
    std::printf("force blocks %s: %s %d blocks, %d interior and %zu boundary nodes\n", name,
                bad ? "FAILED" : "ok", b.num_blocks, first_boundary,
                nn - size_t(first_boundary));
    return bad;
}

}  // namespace

int main(int argc, char **argv)
{
    int bad = 0;
    const int sizes[] = {1, 2, 3, 5, 8};
    for (int nx : sizes) {
        for (int mode = 1; mode <= 2; ++mode) {
            for (int reg : {1, 11, 20}) {
                lulesh::MeshConfig c;
                c.nx = nx;
                c.ic_mode = mode;
                c.num_reg = reg;
                c.balance = reg % 3;
                c.cost = reg % 4;
                bad |= check_cube(c);
            }
        }
    }
    {
        // 4096 blocks, 768 costly: 96 of each XCD's 512 slots, whose first
        // 3/4 (384) at a regular spacing would be exactly every 4th slot
        std::vector<size_t> costly;
        for (size_t b = 0; b < 768; ++b) costly.push_back((b * 4099) % 4096);
        bad |= check_block_order(4096, costly, "period 4");
        costly.clear();
        for (size_t b = 0; b < 1000; ++b) costly.push_back((b * 7919) % 3371);
        bad |= check_block_order(3371, costly, "ragged");
        bad |= check_block_order(1000, {}, "even");
        bad |= check_block_order(5, {1, 2, 3, 4, 0}, "all costly");
        bad |= check_block_order(3, {2}, "tiny");
    }
    for (int nx : {0, 601, -3}) {
        lulesh::MeshConfig c;
        c.nx = nx;
        try {
            (void)lulesh::build_mesh(c);
            std::printf("cube nx=%d accepted\n", nx);
            bad = 1;
        } catch (const std::exception &) {
        }
    }

    lulesh::MeshLists first;
    bool have_first = false;
    for (int i = 1; i < argc; ++i) {
        try {
            lulesh::MeshLists l = lulesh::read_lmesh(argv[i]);
            const lulesh::HostMeshU m = lulesh::build_mesh_u(l);
            std::printf("mesh %s: ok %d %d %d\n", argv[i], m.num_elem, m.num_node, m.origin);
            bad |= check_force_blocks(m, argv[i]);
            if (!have_first) {
                first = l;
                have_first = true;
            }
        } catch (const std::exception &ex) {
            std::printf("mesh %s: rejected\n", argv[i]);
        }
    }

    lo_set_threads(2);
    lo_domain *d = lo_create(5, 2, 11, 1, 1);
    if (!d || lo_run(d, 30) != 0) bad = 1;
    lo_destroy(d);
    d = lo_create(4, 1, 1, 1, 1);
    if (!d || lo_run(d, 20) != 0) bad = 1;
    lo_destroy(d);
    if (have_first) {
        d = lo_create_mesh(first.num_elem, first.num_node, first.x.data(), first.y.data(),
                           first.z.data(), first.nodelist.data(), first.neighbors.data(),
                           int(first.symm_x.size()), first.symm_x.data(),
                           int(first.symm_y.size()), first.symm_y.data(),
                           int(first.symm_z.size()), first.symm_z.data(),
                           int(first.free_surf.size()), first.free_surf.data());
        if (!d || lo_run(d, 20) != 0) bad = 1;
        lo_destroy(d);
    }
    std::printf(bad ? "FAILED\n" : "host code clean\n");
    return bad;
}
