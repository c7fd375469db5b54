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
    return 0;
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
