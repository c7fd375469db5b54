// lulesh_output.cpp -- the domain as a file for visual verification.
//
// The reference writes each dumped cycle as a Silo database through
// writeDomain (cbits/visit.cc:147-175, called from src/VisIt.hs:71-86;
// lulesh-2.0.3/lulesh-viz.cc).  Silo is not available here, so the same
// entry point writes a legacy VTK file instead, which VisIt and ParaView
// read: an unstructured grid of hexahedra with the node velocities as a
// point vector field and e, p, v, q as cell fields, plus the time and cycle.
// LULESH's corner order (bottom face counter-clockwise, then the top face)
// is VTK's hexahedron order, so the connectivity (the cube's, or an
// unstructured mesh's own lists) is written as it is.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lulesh_mi355x.h"

namespace {

// VTK legacy binary data is big-endian
void put_be32(std::vector<unsigned char> &out, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) out.push_back(static_cast<unsigned char>(v >> s));
}

void put_be64(std::vector<unsigned char> &out, double d)
{
    uint64_t v;
    std::memcpy(&v, &d, 8);
    for (int s = 56; s >= 0; s -= 8) out.push_back(static_cast<unsigned char>(v >> s));
}

int edge_of(int num_elem)
{
    const int n = static_cast<int>(std::lround(std::cbrt(static_cast<double>(num_elem))));
    for (int c = n - 1; c <= n + 1; ++c) {
        if (c > 0 && int64_t(c) * c * c == num_elem) return c;
    }
    return -1;
}

// Legacy binary VTK of num_elem hexahedra over num_node nodes; nodelist
// holds 8 node indices per element in LULESH's corner order.
int write_vtk(const char *path, int num_elem, int64_t num_node, const int *nodelist,
              int step, double time, const double *x, const double *y, const double *z,
              const double *xd, const double *yd, const double *zd, const double *e,
              const double *p, const double *v, const double *q)
{
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) {
        return LULESH_BAD_ARG;
    }
    const int64_t nn = num_node;
    std::vector<unsigned char> buf;
    auto text = [&](const std::string &s) { buf.insert(buf.end(), s.begin(), s.end()); };
    char line[256];
    text("# vtk DataFile Version 3.0\n");
    std::snprintf(line, sizeof line, "LULESH Sedov cycle %d time %.17g\n", step, time);
    text(line);
    text("BINARY\nDATASET UNSTRUCTURED_GRID\n");
    text("FIELD FieldData 2\nTIME 1 1 double\n");
    put_be64(buf, time);
    text("\nCYCLE 1 1 int\n");
    put_be32(buf, uint32_t(step));
    std::snprintf(line, sizeof line, "\nPOINTS %lld double\n", static_cast<long long>(nn));
    text(line);
    for (int64_t i = 0; i < nn; ++i) {
        put_be64(buf, x[i]);
        put_be64(buf, y[i]);
        put_be64(buf, z[i]);
    }
    std::snprintf(line, sizeof line, "\nCELLS %d %lld\n", num_elem,
                  static_cast<long long>(num_elem) * 9);
    text(line);
    for (int64_t c = 0; c < num_elem; ++c) {
        put_be32(buf, 8);
        for (int k = 0; k < 8; ++k) put_be32(buf, uint32_t(nodelist[c * 8 + k]));
    }
    std::snprintf(line, sizeof line, "\nCELL_TYPES %d\n", num_elem);
    text(line);
    for (int c = 0; c < num_elem; ++c) put_be32(buf, 12);  // VTK_HEXAHEDRON
    std::snprintf(line, sizeof line, "\nCELL_DATA %d\n", num_elem);
    text(line);
    const struct { const char *name; const double *a; } cells[] = {
        {"e", e}, {"p", p}, {"v", v}, {"q", q}};
    for (const auto &cf : cells) {
        std::snprintf(line, sizeof line, "SCALARS %s double 1\nLOOKUP_TABLE default\n",
                      cf.name);
        text(line);
        for (int c = 0; c < num_elem; ++c) put_be64(buf, cf.a[c]);
        text("\n");
    }
    std::snprintf(line, sizeof line, "POINT_DATA %lld\nVECTORS velocity double\n",
                  static_cast<long long>(nn));
    text(line);
    for (int64_t i = 0; i < nn; ++i) {
        put_be64(buf, xd[i]);
        put_be64(buf, yd[i]);
        put_be64(buf, zd[i]);
    }
    text("\n");
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
    return (std::fclose(fp) == 0 && ok) ? LULESH_OK : LULESH_BAD_ARG;
}

}  // namespace

extern "C" {

int lulesh_write_domain_arrays(const char *path, int num_elem, int step, double time,
                               const double *x, const double *y, const double *z,
                               const double *xd, const double *yd, const double *zd,
                               const double *e, const double *p, const double *v,
                               const double *q)
{
    if (!path || !x || !y || !z || !xd || !yd || !zd || !e || !p || !v || !q) {
        return LULESH_BAD_ARG;
    }
    const int n = edge_of(num_elem);
    if (n < 1) {
        return LULESH_BAD_ARG;  // the structured cube: num_elem = n^3
    }
    const int en = n + 1, pl = en * en;
    std::vector<int> nodelist(size_t(num_elem) * 8);
    size_t w = 0;
    for (int k = 0; k < n; ++k) {
        for (int j = 0; j < n; ++j) {
            for (int i = 0; i < n; ++i) {
                const int n0 = (k * en + j) * en + i;
                const int c[8] = {n0, n0 + 1, n0 + en + 1, n0 + en,
                                  n0 + pl, n0 + pl + 1, n0 + pl + en + 1, n0 + pl + en};
                for (int cc : c) nodelist[w++] = cc;
            }
        }
    }
    return write_vtk(path, num_elem, int64_t(en) * en * en, nodelist.data(), step, time, x,
                     y, z, xd, yd, zd, e, p, v, q);
}

int lulesh_write_domain(lulesh_domain *d, const char *path)
{
    if (!d || !path) {
        return LULESH_BAD_ARG;
    }
    lulesh_stats st;
    int rc = lulesh_get_stats(d, &st);
    if (rc) return rc;
    const size_t nn = size_t(st.num_node), ne = size_t(st.num_elem);
    std::vector<double> node(6 * nn), elem(5 * ne);
    std::vector<int> nodelist(8 * ne);
    double t = 0.0, dt = 0.0;
    rc = lulesh_download_state(d, &node[0], &node[nn], &node[2 * nn], &node[3 * nn],
                               &node[4 * nn], &node[5 * nn], &elem[0], &elem[ne],
                               &elem[2 * ne], &elem[3 * ne], &elem[4 * ne], &t, &dt);
    if (rc) return rc;
    if ((rc = lulesh_get_nodelist(d, nodelist.data()))) return rc;
    return write_vtk(path, st.num_elem, int64_t(nn), nodelist.data(), st.cycle, t, &node[0],
                     &node[nn], &node[2 * nn], &node[3 * nn], &node[4 * nn], &node[5 * nn],
                     &elem[0], &elem[ne], &elem[3 * ne], &elem[2 * ne]);
}

}  // extern "C"
