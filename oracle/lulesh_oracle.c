/*
 * lulesh_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement of the LULESH Lagrange step used to check the HIP path.
 * Every floating-point expression keeps the operand order and association of
 * the reference C code it cites, so that (with no FMA contraction) the
 * results are bit-identical to binaries built from the reference sources.
 * Node forces are assembled by a gather over a corner list built in
 * ascending element order, as the threaded reference does
 * (reference/C/LULESH-OMP.cc:209-268, :905-926; lulesh-2.0.3/lulesh-init.cc:
 * 226-283, lulesh-2.0.3/lulesh.cc:588-612); that makes the result
 * independent of the number of threads.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this code.
 */
#include "lulesh_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LULESH_FN static inline
#include "../lulesh-accelerate_amd/csrc/lulesh_cbrt.h"

/* element face boundary-condition bits (reference/C/LULESH-OMP.cc:532-554;
 * the 2.0 layout in lulesh-2.0.3/lulesh.h:60-88 differs only in bit
 * positions and the COMM flags, which a single domain never sets) */
enum {
    BC_XI_M_SYMM = 0x001, BC_XI_M_FREE = 0x002, BC_XI_M = 0x003,
    BC_XI_P_SYMM = 0x004, BC_XI_P_FREE = 0x008, BC_XI_P = 0x00c,
    BC_ETA_M_SYMM = 0x010, BC_ETA_M_FREE = 0x020, BC_ETA_M = 0x030,
    BC_ETA_P_SYMM = 0x040, BC_ETA_P_FREE = 0x080, BC_ETA_P = 0x0c0,
    BC_ZETA_M_SYMM = 0x100, BC_ZETA_M_FREE = 0x200, BC_ZETA_M = 0x300,
    BC_ZETA_P_SYMM = 0x400, BC_ZETA_P_FREE = 0x800, BC_ZETA_P = 0xc00
};

enum { LO_VOLUME_ERROR = -1, LO_QSTOP_ERROR = -2 };

struct lo_domain {
    int nx, mode, num_elem, num_node, edge_nodes;
    int num_symm_x, num_symm_y, num_symm_z; /* sizes of the symmetry node sets */
    int origin;                             /* the element holding e(0) */
    int num_reg, balance, cost, cbrt_mode;
    double efactor; /* multiplies the initial origin energy (ensembles) */

    /* node-centred */
    double *x, *y, *z, *xd, *yd, *zd, *xdd, *ydd, *zdd, *fx, *fy, *fz;
    double *nodal_mass;
    int *symm_x, *symm_y, *symm_z;
    int *node_elem_start, *node_elem_corner;

    /* element-centred */
    int *nodelist, *lxim, *lxip, *letam, *letap, *lzetam, *lzetap, *elem_bc;
    double *e, *p, *q, *ql, *qq, *v, *volo, *vnew, *delv, *vdov, *arealg;
    double *ss, *elem_mass;
    double *dxx, *dyy, *dzz;
    double *delv_xi, *delv_eta, *delv_zeta, *delx_xi, *delx_eta, *delx_zeta;
    int *reg_num_list, *reg_elem_size, **reg_elem_list;

    /* scratch for the corner forces and the hourglass pass */
    double *fx_elem, *fy_elem, *fz_elem, *determ;
    double *vnewc;

    /* parameters (reference/C/LULESH-OMP.cc:3052-3087;
     * lulesh-2.0.3/lulesh-init.cc:19-41, :140-182) */
    double dtfixed, time, deltatime, deltatimemultlb, deltatimemultub;
    double stoptime, dtcourant, dthydro, dtmax;
    double u_cut, hgcoef, qstop, monoq_max_slope, monoq_limiter_mult;
    double e_cut, p_cut, ss4o3, q_cut, v_cut, qlc_monoq, qqc_monoq, qqc;
    double eosvmax, eosvmin, pmin, emin, dvovmax, refdens;
    int cycle;
};

static int g_threads = 0;

void lo_set_threads(int nthreads) { g_threads = nthreads; }

static void apply_threads(void)
{
#ifdef _OPENMP
    if (g_threads > 0) {
        omp_set_num_threads(g_threads);
    }
#endif
}

static double *dalloc(size_t n)
{
    double *p = (double *)calloc(n ? n : 1, sizeof(double));
    if (!p) {
        fprintf(stderr, "lulesh_oracle: out of memory\n");
        abort();
    }
    return p;
}

static int *ialloc(size_t n)
{
    int *p = (int *)calloc(n ? n : 1, sizeof(int));
    if (!p) {
        fprintf(stderr, "lulesh_oracle: out of memory\n");
        abort();
    }
    return p;
}

/* ------------------------------------------------------------------------ */
/* element-local geometry                                                    */
/* ------------------------------------------------------------------------ */

/* reference/C/LULESH-OMP.cc:620-706 */
static void shape_function_derivatives(const double *x, const double *y,
                                       const double *z, double b[3][8],
                                       double *volume)
{
    const double fjxxi = .125 * ((x[6] - x[0]) + (x[5] - x[3]) - (x[7] - x[1]) - (x[4] - x[2]));
    const double fjxet = .125 * ((x[6] - x[0]) - (x[5] - x[3]) + (x[7] - x[1]) - (x[4] - x[2]));
    const double fjxze = .125 * ((x[6] - x[0]) + (x[5] - x[3]) + (x[7] - x[1]) + (x[4] - x[2]));

    const double fjyxi = .125 * ((y[6] - y[0]) + (y[5] - y[3]) - (y[7] - y[1]) - (y[4] - y[2]));
    const double fjyet = .125 * ((y[6] - y[0]) - (y[5] - y[3]) + (y[7] - y[1]) - (y[4] - y[2]));
    const double fjyze = .125 * ((y[6] - y[0]) + (y[5] - y[3]) + (y[7] - y[1]) + (y[4] - y[2]));

    const double fjzxi = .125 * ((z[6] - z[0]) + (z[5] - z[3]) - (z[7] - z[1]) - (z[4] - z[2]));
    const double fjzet = .125 * ((z[6] - z[0]) - (z[5] - z[3]) + (z[7] - z[1]) - (z[4] - z[2]));
    const double fjzze = .125 * ((z[6] - z[0]) + (z[5] - z[3]) + (z[7] - z[1]) + (z[4] - z[2]));

    const double cjxxi = (fjyet * fjzze) - (fjzet * fjyze);
    const double cjxet = -(fjyxi * fjzze) + (fjzxi * fjyze);
    const double cjxze = (fjyxi * fjzet) - (fjzxi * fjyet);

    const double cjyxi = -(fjxet * fjzze) + (fjzet * fjxze);
    const double cjyet = (fjxxi * fjzze) - (fjzxi * fjxze);
    const double cjyze = -(fjxxi * fjzet) + (fjzxi * fjxet);

    const double cjzxi = (fjxet * fjyze) - (fjyet * fjxze);
    const double cjzet = -(fjxxi * fjyze) + (fjyxi * fjxze);
    const double cjzze = (fjxxi * fjyet) - (fjyxi * fjxet);

    /* nodes 4..7 mirror nodes 2,3,0,1 with opposite sign (:681-702) */
    b[0][0] = -cjxxi - cjxet - cjxze;
    b[0][1] = cjxxi - cjxet - cjxze;
    b[0][2] = cjxxi + cjxet - cjxze;
    b[0][3] = -cjxxi + cjxet - cjxze;
    b[0][4] = -b[0][2];
    b[0][5] = -b[0][3];
    b[0][6] = -b[0][0];
    b[0][7] = -b[0][1];

    b[1][0] = -cjyxi - cjyet - cjyze;
    b[1][1] = cjyxi - cjyet - cjyze;
    b[1][2] = cjyxi + cjyet - cjyze;
    b[1][3] = -cjyxi + cjyet - cjyze;
    b[1][4] = -b[1][2];
    b[1][5] = -b[1][3];
    b[1][6] = -b[1][0];
    b[1][7] = -b[1][1];

    b[2][0] = -cjzxi - cjzet - cjzze;
    b[2][1] = cjzxi - cjzet - cjzze;
    b[2][2] = cjzxi + cjzet - cjzze;
    b[2][3] = -cjzxi + cjzet - cjzze;
    b[2][4] = -b[2][2];
    b[2][5] = -b[2][3];
    b[2][6] = -b[2][0];
    b[2][7] = -b[2][1];

    *volume = 8. * (fjxet * cjxet + fjyet * cjyet + fjzet * cjzet);
}

/* reference/C/LULESH-OMP.cc:709-742: adds the area vector of face
 * (n0,n1,n2,n3) to the normals of its four nodes */
static void sum_face_normal(double *pfx, double *pfy, double *pfz,
                            const double *x, const double *y, const double *z,
                            int n0, int n1, int n2, int n3)
{
    const double bx0 = 0.5 * (x[n3] + x[n2] - x[n1] - x[n0]);
    const double by0 = 0.5 * (y[n3] + y[n2] - y[n1] - y[n0]);
    const double bz0 = 0.5 * (z[n3] + z[n2] - z[n1] - z[n0]);
    const double bx1 = 0.5 * (x[n2] + x[n1] - x[n3] - x[n0]);
    const double by1 = 0.5 * (y[n2] + y[n1] - y[n3] - y[n0]);
    const double bz1 = 0.5 * (z[n2] + z[n1] - z[n3] - z[n0]);
    const double ax = 0.25 * (by0 * bz1 - bz0 * by1);
    const double ay = 0.25 * (bz0 * bx1 - bx0 * bz1);
    const double az = 0.25 * (bx0 * by1 - by0 * bx1);
    pfx[n0] += ax; pfx[n1] += ax; pfx[n2] += ax; pfx[n3] += ax;
    pfy[n0] += ay; pfy[n1] += ay; pfy[n2] += ay; pfy[n3] += ay;
    pfz[n0] += az; pfz[n1] += az; pfz[n2] += az; pfz[n3] += az;
}

/* reference/C/LULESH-OMP.cc:745-799 (face order and node order kept) */
static void elem_node_normals(double *pfx, double *pfy, double *pfz,
                              const double *x, const double *y, const double *z)
{
    for (int i = 0; i < 8; ++i) {
        pfx[i] = 0.0;
        pfy[i] = 0.0;
        pfz[i] = 0.0;
    }
    sum_face_normal(pfx, pfy, pfz, x, y, z, 0, 1, 2, 3);
    sum_face_normal(pfx, pfy, pfz, x, y, z, 0, 4, 5, 1);
    sum_face_normal(pfx, pfy, pfz, x, y, z, 1, 5, 6, 2);
    sum_face_normal(pfx, pfy, pfz, x, y, z, 2, 6, 7, 3);
    sum_face_normal(pfx, pfy, pfz, x, y, z, 3, 7, 4, 0);
    sum_face_normal(pfx, pfy, pfz, x, y, z, 4, 7, 6, 5);
}

/* reference/C/LULESH-OMP.cc:979-1006, with the six nodes passed by index */
static void volu_der(const double *x, const double *y, const double *z,
                     int a0, int a1, int a2, int a3, int a4, int a5,
                     double *dvdx, double *dvdy, double *dvdz)
{
    const double twelfth = 1.0 / 12.0;
    const double x0 = x[a0], x1 = x[a1], x2 = x[a2], x3 = x[a3], x4 = x[a4], x5 = x[a5];
    const double y0 = y[a0], y1 = y[a1], y2 = y[a2], y3 = y[a3], y4 = y[a4], y5 = y[a5];
    const double z0 = z[a0], z1 = z[a1], z2 = z[a2], z3 = z[a3], z4 = z[a4], z5 = z[a5];

    *dvdx = (y1 + y2) * (z0 + z1) - (y0 + y1) * (z1 + z2) +
            (y0 + y4) * (z3 + z4) - (y3 + y4) * (z0 + z4) -
            (y2 + y5) * (z3 + z5) + (y3 + y5) * (z2 + z5);
    *dvdy = -(x1 + x2) * (z0 + z1) + (x0 + x1) * (z1 + z2) -
            (x0 + x4) * (z3 + z4) + (x3 + x4) * (z0 + z4) +
            (x2 + x5) * (z3 + z5) - (x3 + x5) * (z2 + z5);
    *dvdz = -(y1 + y2) * (x0 + x1) + (y0 + y1) * (x1 + x2) -
            (y0 + y4) * (x3 + x4) + (y3 + y4) * (x0 + x4) +
            (y2 + y5) * (x3 + x5) - (y3 + y5) * (x2 + x5);

    *dvdx *= twelfth;
    *dvdy *= twelfth;
    *dvdz *= twelfth;
}

/* reference/C/LULESH-OMP.cc:1008-1048: the six-node stencil of each node */
static const int k_volu_der_nodes[8][6] = {
    {1, 2, 3, 4, 5, 7}, /* node 0 */
    {2, 3, 0, 5, 6, 4}, /* node 1 */
    {3, 0, 1, 6, 7, 5}, /* node 2 */
    {0, 1, 2, 7, 4, 6}, /* node 3 */
    {7, 6, 5, 0, 3, 1}, /* node 4 */
    {4, 7, 6, 1, 0, 2}, /* node 5 */
    {5, 4, 7, 2, 1, 3}, /* node 6 */
    {6, 5, 4, 3, 2, 0}, /* node 7 */
};

static void elem_volume_derivative(double *dvdx, double *dvdy, double *dvdz,
                                   const double *x, const double *y,
                                   const double *z)
{
    for (int n = 0; n < 8; ++n) {
        const int *s = k_volu_der_nodes[n];
        volu_der(x, y, z, s[0], s[1], s[2], s[3], s[4], s[5],
                 &dvdx[n], &dvdy[n], &dvdz[n]);
    }
}

/* reference/C/LULESH-OMP.cc:1706-1789 */
static double triple_product(double x1, double y1, double z1, double x2,
                             double y2, double z2, double x3, double y3,
                             double z3)
{
    return x1 * (y2 * z3 - z2 * y3) + x2 * (z1 * y3 - y1 * z3) +
           x3 * (y1 * z2 - z1 * y2);
}

static double elem_volume(const double *x, const double *y, const double *z)
{
    const double twelveth = 1.0 / 12.0;
    const double dx61 = x[6] - x[1], dy61 = y[6] - y[1], dz61 = z[6] - z[1];
    const double dx70 = x[7] - x[0], dy70 = y[7] - y[0], dz70 = z[7] - z[0];
    const double dx63 = x[6] - x[3], dy63 = y[6] - y[3], dz63 = z[6] - z[3];
    const double dx20 = x[2] - x[0], dy20 = y[2] - y[0], dz20 = z[2] - z[0];
    const double dx50 = x[5] - x[0], dy50 = y[5] - y[0], dz50 = z[5] - z[0];
    const double dx64 = x[6] - x[4], dy64 = y[6] - y[4], dz64 = z[6] - z[4];
    const double dx31 = x[3] - x[1], dy31 = y[3] - y[1], dz31 = z[3] - z[1];
    const double dx72 = x[7] - x[2], dy72 = y[7] - y[2], dz72 = z[7] - z[2];
    const double dx43 = x[4] - x[3], dy43 = y[4] - y[3], dz43 = z[4] - z[3];
    const double dx57 = x[5] - x[7], dy57 = y[5] - y[7], dz57 = z[5] - z[7];
    const double dx14 = x[1] - x[4], dy14 = y[1] - y[4], dz14 = z[1] - z[4];
    const double dx25 = x[2] - x[5], dy25 = y[2] - y[5], dz25 = z[2] - z[5];

    double volume =
        triple_product(dx31 + dx72, dx63, dx20, dy31 + dy72, dy63, dy20,
                       dz31 + dz72, dz63, dz20) +
        triple_product(dx43 + dx57, dx64, dx70, dy43 + dy57, dy64, dy70,
                       dz43 + dz57, dz64, dz70) +
        triple_product(dx14 + dx25, dx61, dx50, dy14 + dy25, dy61, dy50,
                       dz14 + dz25, dz61, dz50);
    volume *= twelveth;
    return volume;
}

/* reference/C/LULESH-OMP.cc:1799-1819 */
static double area_face(const double *x, const double *y, const double *z,
                        int n0, int n1, int n2, int n3)
{
    const double fx = (x[n2] - x[n0]) - (x[n3] - x[n1]);
    const double fy = (y[n2] - y[n0]) - (y[n3] - y[n1]);
    const double fz = (z[n2] - z[n0]) - (z[n3] - z[n1]);
    const double gx = (x[n2] - x[n0]) + (x[n3] - x[n1]);
    const double gy = (y[n2] - y[n0]) + (y[n3] - y[n1]);
    const double gz = (z[n2] - z[n0]) + (z[n3] - z[n1]);
    return (fx * fx + fy * fy + fz * fz) * (gx * gx + gy * gy + gz * gz) -
           (fx * gx + fy * gy + fz * gz) * (fx * gx + fy * gy + fz * gz);
}

/* std::max(a, b) of the reference: b if a < b, else a */
static double ref_max(double a, double b) { return (a < b) ? b : a; }

/* reference/C/LULESH-OMP.cc:1821-1862 */
static double elem_characteristic_length(const double *x, const double *y,
                                         const double *z, double volume)
{
    double cl = 0.0;
    cl = ref_max(area_face(x, y, z, 0, 1, 2, 3), cl);
    cl = ref_max(area_face(x, y, z, 4, 5, 6, 7), cl);
    cl = ref_max(area_face(x, y, z, 0, 1, 5, 4), cl);
    cl = ref_max(area_face(x, y, z, 1, 2, 6, 5), cl);
    cl = ref_max(area_face(x, y, z, 2, 3, 7, 6), cl);
    cl = ref_max(area_face(x, y, z, 3, 0, 4, 7), cl);
    return 4.0 * volume / sqrt(cl);
}

/* reference/C/LULESH-OMP.cc:1864-1891: only the diagonal of the velocity
 * gradient is consumed (vdov); the off-diagonal terms are never read */
static void elem_velocity_gradient_diag(const double *xv, const double *yv,
                                        const double *zv, double b[3][8],
                                        double det_j, double d[3])
{
    const double inv_det_j = 1.0 / det_j;
    const double *pfx = b[0], *pfy = b[1], *pfz = b[2];
    d[0] = inv_det_j * (pfx[0] * (xv[0] - xv[6]) + pfx[1] * (xv[1] - xv[7]) +
                        pfx[2] * (xv[2] - xv[4]) + pfx[3] * (xv[3] - xv[5]));
    d[1] = inv_det_j * (pfy[0] * (yv[0] - yv[6]) + pfy[1] * (yv[1] - yv[7]) +
                        pfy[2] * (yv[2] - yv[4]) + pfy[3] * (yv[3] - yv[5]));
    d[2] = inv_det_j * (pfz[0] * (zv[0] - zv[6]) + pfz[1] * (zv[1] - zv[7]) +
                        pfz[2] * (zv[2] - zv[4]) + pfz[3] * (zv[3] - zv[5]));
}

/* Flanagan-Belytschko hourglass mode vectors (reference/C/LULESH-OMP.cc:
 * 1252-1285) */
static const double k_gamma[4][8] = {
    {1., 1., -1., -1., -1., -1., 1., 1.},
    {1., -1., -1., 1., -1., 1., 1., -1.},
    {1., -1., 1., -1., 1., -1., 1., -1.},
    {-1., 1., -1., 1., 1., -1., 1., -1.},
};

/* hourgam = Gamma - (1/determ) (dV/dx . hourmod)  (reference/C/LULESH-OMP.cc:
 * 1304-1358), then the force of one element (:1050-1231) */
static void elem_fb_hourglass(const double *x8, const double *y8,
                              const double *z8, const double *dvdx,
                              const double *dvdy, const double *dvdz,
                              const double *xd, const double *yd,
                              const double *zd, double determ,
                              double coefficient, double *hgfx, double *hgfy,
                              double *hgfz)
{
    double hourgam[8][4];
    const double volinv = 1.0 / determ;
    for (int i1 = 0; i1 < 4; ++i1) {
        const double *g = k_gamma[i1];
        const double hmx = x8[0] * g[0] + x8[1] * g[1] + x8[2] * g[2] + x8[3] * g[3] +
                           x8[4] * g[4] + x8[5] * g[5] + x8[6] * g[6] + x8[7] * g[7];
        const double hmy = y8[0] * g[0] + y8[1] * g[1] + y8[2] * g[2] + y8[3] * g[3] +
                           y8[4] * g[4] + y8[5] * g[5] + y8[6] * g[6] + y8[7] * g[7];
        const double hmz = z8[0] * g[0] + z8[1] * g[1] + z8[2] * g[2] + z8[3] * g[3] +
                           z8[4] * g[4] + z8[5] * g[5] + z8[6] * g[6] + z8[7] * g[7];
        for (int n = 0; n < 8; ++n) {
            hourgam[n][i1] =
                g[n] - volinv * (dvdx[n] * hmx + dvdy[n] * hmy + dvdz[n] * hmz);
        }
    }

    const double *vel[3] = {xd, yd, zd};
    double *out[3] = {hgfx, hgfy, hgfz};
    for (int c = 0; c < 3; ++c) {
        const double *u = vel[c];
        double h[4];
        for (int i = 0; i < 4; ++i) {
            h[i] = hourgam[0][i] * u[0] + hourgam[1][i] * u[1] +
                   hourgam[2][i] * u[2] + hourgam[3][i] * u[3] +
                   hourgam[4][i] * u[4] + hourgam[5][i] * u[5] +
                   hourgam[6][i] * u[6] + hourgam[7][i] * u[7];
        }
        for (int n = 0; n < 8; ++n) {
            out[c][n] = coefficient * (hourgam[n][0] * h[0] + hourgam[n][1] * h[1] +
                                       hourgam[n][2] * h[2] + hourgam[n][3] * h[3]);
        }
    }
}

static void gather8(const lo_domain *d, const double *src, const int *nl,
                    double *dst)
{
    for (int j = 0; j < 8; ++j) {
        dst[j] = src[nl[j]];
    }
}

/* ------------------------------------------------------------------------ */
/* nodal phase                                                               */
/* ------------------------------------------------------------------------ */

/* gather of per-corner values into nodes: out = (init) + sum over the
 * node's corners in ascending element order, the sum started from 0.0
 * (reference/C/LULESH-OMP.cc:905-926 assigns, :1475-1496 adds) */
static void gather_corners(lo_domain *d, int accumulate)
{
    const int num_node = d->num_node;
#pragma omp parallel for
    for (int n = 0; n < num_node; ++n) {
        const int start = d->node_elem_start[n];
        const int end = d->node_elem_start[n + 1];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int i = start; i < end; ++i) {
            const int corner = d->node_elem_corner[i];
            sx += d->fx_elem[corner];
            sy += d->fy_elem[corner];
            sz += d->fz_elem[corner];
        }
        if (accumulate) {
            d->fx[n] += sx;
            d->fy[n] += sy;
            d->fz[n] += sz;
        } else {
            d->fx[n] = sx;
            d->fy[n] = sy;
            d->fz[n] = sz;
        }
    }
}

/* InitStressTermsForElems + IntegrateStressForElems
 * (reference/C/LULESH-OMP.cc:606-617, :853-931) */
static void integrate_stress(lo_domain *d)
{
    const int num_elem = d->num_elem;
#pragma omp parallel for
    for (int k = 0; k < num_elem; ++k) {
        const int *nl = &d->nodelist[8 * k];
        double xl[8], yl[8], zl[8], b[3][8];
        const double sig = -d->p[k] - d->q[k];
        gather8(d, d->x, nl, xl);
        gather8(d, d->y, nl, yl);
        gather8(d, d->z, nl, zl);
        shape_function_derivatives(xl, yl, zl, b, &d->determ[k]);
        elem_node_normals(b[0], b[1], b[2], xl, yl, zl);
        for (int j = 0; j < 8; ++j) {
            d->fx_elem[8 * k + j] = -(sig * b[0][j]);
            d->fy_elem[8 * k + j] = -(sig * b[1][j]);
            d->fz_elem[8 * k + j] = -(sig * b[2][j]);
        }
    }
    gather_corners(d, 0);
}

static double hourglass_cbrt(const lo_domain *d, double v)
{
    return d->cbrt_mode ? lulesh_cbrt(v) : cbrt(v);
}

/* CalcHourglassControlForElems + CalcFBHourglassForceForElems
 * (reference/C/LULESH-OMP.cc:1233-1558) */
static int hourglass_control(lo_domain *d)
{
    const int num_elem = d->num_elem;
    int bad = 0;
#pragma omp parallel for reduction(| : bad)
    for (int i = 0; i < num_elem; ++i) {
        d->determ[i] = d->volo[i] * d->v[i];
        if (d->v[i] <= 0.0) {
            bad = 1;
        }
    }
    if (bad) {
        return LO_VOLUME_ERROR;
    }
    if (!(d->hgcoef > 0.0)) {
        return 0;
    }
    const double hourg = d->hgcoef;
#pragma omp parallel for
    for (int i2 = 0; i2 < num_elem; ++i2) {
        const int *nl = &d->nodelist[8 * i2];
        double x8[8], y8[8], z8[8], dvdx[8], dvdy[8], dvdz[8];
        double xd1[8], yd1[8], zd1[8];
        gather8(d, d->x, nl, x8);
        gather8(d, d->y, nl, y8);
        gather8(d, d->z, nl, z8);
        elem_volume_derivative(dvdx, dvdy, dvdz, x8, y8, z8);
        gather8(d, d->xd, nl, xd1);
        gather8(d, d->yd, nl, yd1);
        gather8(d, d->zd, nl, zd1);
        const double ss1 = d->ss[i2];
        const double mass1 = d->elem_mass[i2];
        const double volume13 = hourglass_cbrt(d, d->determ[i2]);
        const double coefficient = -hourg * 0.01 * ss1 * mass1 / volume13;
        elem_fb_hourglass(x8, y8, z8, dvdx, dvdy, dvdz, xd1, yd1, zd1,
                          d->determ[i2], coefficient, &d->fx_elem[8 * i2],
                          &d->fy_elem[8 * i2], &d->fz_elem[8 * i2]);
    }
    gather_corners(d, 1);
    return 0;
}

/* CalcForceForNodes / CalcVolumeForceForElems (reference/C/LULESH-OMP.cc:
 * 1560-1611) */
static int calc_force_for_nodes(lo_domain *d)
{
    const int num_elem = d->num_elem;
    int bad = 0;
    integrate_stress(d);
#pragma omp parallel for reduction(| : bad)
    for (int k = 0; k < num_elem; ++k) {
        if (d->determ[k] <= 0.0) {
            bad = 1;
        }
    }
    if (bad) {
        return LO_VOLUME_ERROR;
    }
    return hourglass_control(d);
}

/* LagrangeNodal (reference/C/LULESH-OMP.cc:1613-1704) */
int lo_lagrange_nodal(lo_domain *d)
{
    apply_threads();
    const double delt = d->deltatime;
    const double u_cut = d->u_cut;
    const int num_node = d->num_node;
    int rc = calc_force_for_nodes(d);
    if (rc) {
        return rc;
    }
#pragma omp parallel for
    for (int i = 0; i < num_node; ++i) {
        d->xdd[i] = d->fx[i] / d->nodal_mass[i];
        d->ydd[i] = d->fy[i] / d->nodal_mass[i];
        d->zdd[i] = d->fz[i] / d->nodal_mass[i];
    }
    for (int i = 0; i < d->num_symm_x; ++i) {
        d->xdd[d->symm_x[i]] = 0.0;
    }
    for (int i = 0; i < d->num_symm_y; ++i) {
        d->ydd[d->symm_y[i]] = 0.0;
    }
    for (int i = 0; i < d->num_symm_z; ++i) {
        d->zdd[d->symm_z[i]] = 0.0;
    }
#pragma omp parallel for
    for (int i = 0; i < num_node; ++i) {
        double t;
        t = d->xd[i] + d->xdd[i] * delt;
        if (fabs(t) < u_cut) t = 0.0;
        d->xd[i] = t;
        t = d->yd[i] + d->ydd[i] * delt;
        if (fabs(t) < u_cut) t = 0.0;
        d->yd[i] = t;
        t = d->zd[i] + d->zdd[i] * delt;
        if (fabs(t) < u_cut) t = 0.0;
        d->zd[i] = t;
    }
#pragma omp parallel for
    for (int i = 0; i < num_node; ++i) {
        d->x[i] += d->xd[i] * delt;
        d->y[i] += d->yd[i] * delt;
        d->z[i] += d->zd[i] * delt;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* element phase                                                             */
/* ------------------------------------------------------------------------ */

/* CalcKinematicsForElems + CalcLagrangeElements (reference/C/LULESH-OMP.cc:
 * 1927-2032; lulesh-2.0.3/lulesh.cc:1533-1639) */
static int lagrange_elements_kinematics(lo_domain *d)
{
    const int num_elem = d->num_elem;
    const double dt = d->deltatime;
    int bad = 0;
#pragma omp parallel for reduction(| : bad)
    for (int k = 0; k < num_elem; ++k) {
        const int *nl = &d->nodelist[8 * k];
        double xl[8], yl[8], zl[8], xdl[8], ydl[8], zdl[8], b[3][8], dd[3];
        double det_j = 0.0;
        gather8(d, d->x, nl, xl);
        gather8(d, d->y, nl, yl);
        gather8(d, d->z, nl, zl);
        const double volume = elem_volume(xl, yl, zl);
        const double rel = volume / d->volo[k];
        d->vnew[k] = rel;
        d->delv[k] = rel - d->v[k];
        d->arealg[k] = elem_characteristic_length(xl, yl, zl, volume);
        gather8(d, d->xd, nl, xdl);
        gather8(d, d->yd, nl, ydl);
        gather8(d, d->zd, nl, zdl);
        const double dt2 = 0.5 * dt;
        for (int j = 0; j < 8; ++j) {
            xl[j] -= dt2 * xdl[j];
            yl[j] -= dt2 * ydl[j];
            zl[j] -= dt2 * zdl[j];
        }
        shape_function_derivatives(xl, yl, zl, b, &det_j);
        elem_velocity_gradient_diag(xdl, ydl, zdl, b, det_j, dd);
        const double vdov = dd[0] + dd[1] + dd[2];
        const double vdovthird = vdov / 3.0;
        d->vdov[k] = vdov;
        d->dxx[k] = dd[0] - vdovthird;
        d->dyy[k] = dd[1] - vdovthird;
        d->dzz[k] = dd[2] - vdovthird;
        if (rel <= 0.0) {
            bad = 1;
        }
    }
    return bad ? LO_VOLUME_ERROR : 0;
}

/* CalcMonotonicQGradientsForElems (reference/C/LULESH-OMP.cc:2034-2180) */
static void monotonic_q_gradients(lo_domain *d)
{
    const int num_elem = d->num_elem;
#pragma omp parallel for
    for (int i = 0; i < num_elem; ++i) {
        const double ptiny = 1.e-36;
        const int *nl = &d->nodelist[8 * i];
        double x[8], y[8], z[8], xv[8], yv[8], zv[8];
        gather8(d, d->x, nl, x);
        gather8(d, d->y, nl, y);
        gather8(d, d->z, nl, z);
        gather8(d, d->xd, nl, xv);
        gather8(d, d->yd, nl, yv);
        gather8(d, d->zd, nl, zv);

        const double vol = d->volo[i] * d->vnew[i];
        const double norm = 1.0 / (vol + ptiny);

        const double dxj = -0.25 * ((x[0] + x[1] + x[5] + x[4]) - (x[3] + x[2] + x[6] + x[7]));
        const double dyj = -0.25 * ((y[0] + y[1] + y[5] + y[4]) - (y[3] + y[2] + y[6] + y[7]));
        const double dzj = -0.25 * ((z[0] + z[1] + z[5] + z[4]) - (z[3] + z[2] + z[6] + z[7]));

        const double dxi = 0.25 * ((x[1] + x[2] + x[6] + x[5]) - (x[0] + x[3] + x[7] + x[4]));
        const double dyi = 0.25 * ((y[1] + y[2] + y[6] + y[5]) - (y[0] + y[3] + y[7] + y[4]));
        const double dzi = 0.25 * ((z[1] + z[2] + z[6] + z[5]) - (z[0] + z[3] + z[7] + z[4]));

        const double dxk = 0.25 * ((x[4] + x[5] + x[6] + x[7]) - (x[0] + x[1] + x[2] + x[3]));
        const double dyk = 0.25 * ((y[4] + y[5] + y[6] + y[7]) - (y[0] + y[1] + y[2] + y[3]));
        const double dzk = 0.25 * ((z[4] + z[5] + z[6] + z[7]) - (z[0] + z[1] + z[2] + z[3]));

        double ax, ay, az, dxv, dyv, dzv;

        /* zeta: i cross j */
        ax = dyi * dzj - dzi * dyj;
        ay = dzi * dxj - dxi * dzj;
        az = dxi * dyj - dyi * dxj;
        d->delx_zeta[i] = vol / sqrt(ax * ax + ay * ay + az * az + ptiny);
        ax *= norm;
        ay *= norm;
        az *= norm;
        dxv = 0.25 * ((xv[4] + xv[5] + xv[6] + xv[7]) - (xv[0] + xv[1] + xv[2] + xv[3]));
        dyv = 0.25 * ((yv[4] + yv[5] + yv[6] + yv[7]) - (yv[0] + yv[1] + yv[2] + yv[3]));
        dzv = 0.25 * ((zv[4] + zv[5] + zv[6] + zv[7]) - (zv[0] + zv[1] + zv[2] + zv[3]));
        d->delv_zeta[i] = ax * dxv + ay * dyv + az * dzv;

        /* xi: j cross k */
        ax = dyj * dzk - dzj * dyk;
        ay = dzj * dxk - dxj * dzk;
        az = dxj * dyk - dyj * dxk;
        d->delx_xi[i] = vol / sqrt(ax * ax + ay * ay + az * az + ptiny);
        ax *= norm;
        ay *= norm;
        az *= norm;
        dxv = 0.25 * ((xv[1] + xv[2] + xv[6] + xv[5]) - (xv[0] + xv[3] + xv[7] + xv[4]));
        dyv = 0.25 * ((yv[1] + yv[2] + yv[6] + yv[5]) - (yv[0] + yv[3] + yv[7] + yv[4]));
        dzv = 0.25 * ((zv[1] + zv[2] + zv[6] + zv[5]) - (zv[0] + zv[3] + zv[7] + zv[4]));
        d->delv_xi[i] = ax * dxv + ay * dyv + az * dzv;

        /* eta: k cross i */
        ax = dyk * dzi - dzk * dyi;
        ay = dzk * dxi - dxk * dzi;
        az = dxk * dyi - dyk * dxi;
        d->delx_eta[i] = vol / sqrt(ax * ax + ay * ay + az * az + ptiny);
        ax *= norm;
        ay *= norm;
        az *= norm;
        dxv = -0.25 * ((xv[0] + xv[1] + xv[5] + xv[4]) - (xv[3] + xv[2] + xv[6] + xv[7]));
        dyv = -0.25 * ((yv[0] + yv[1] + yv[5] + yv[4]) - (yv[3] + yv[2] + yv[6] + yv[7]));
        dzv = -0.25 * ((zv[0] + zv[1] + zv[5] + zv[4]) - (zv[3] + zv[2] + zv[6] + zv[7]));
        d->delv_eta[i] = ax * dxv + ay * dyv + az * dzv;
    }
}

/* limiter of one direction (reference/C/LULESH-OMP.cc:2201-2228) */
static double mono_phi(double delv_self, double delvm, double delvp,
                       double limiter_mult, double max_slope, double ptiny)
{
    const double norm = 1. / (delv_self + ptiny);
    double phi;
    delvm = delvm * norm;
    delvp = delvp * norm;
    phi = .5 * (delvm + delvp);
    delvm *= limiter_mult;
    delvp *= limiter_mult;
    if (delvm < phi) phi = delvm;
    if (delvp < phi) phi = delvp;
    if (phi < 0.) phi = 0.;
    if (phi > max_slope) phi = max_slope;
    return phi;
}

/* neighbour value across one face: interior -> neighbour, symmetry plane ->
 * self, free surface -> 0 (reference/C/LULESH-OMP.cc:2204-2215) */
static double face_neighbour(const double *f, int i, int nbr, int bc,
                             int symm, int freeb)
{
    if (bc & symm) return f[i];
    if (bc & freeb) return 0.0;
    return f[nbr];
}

/* CalcMonotonicQRegionForElems for one element (reference/C/LULESH-OMP.cc:
 * 2182-2320; lulesh-2.0.3/lulesh.cc:1789-1951) */
static void monotonic_q_region_elem(lo_domain *d, int i)
{
    const double ptiny = 1.e-36;
    const double lm = d->monoq_limiter_mult, ms = d->monoq_max_slope;
    const int bc = d->elem_bc[i];
    double qlin, qquad;

    const double phixi = mono_phi(
        d->delv_xi[i],
        face_neighbour(d->delv_xi, i, d->lxim[i], bc, BC_XI_M_SYMM, BC_XI_M_FREE),
        face_neighbour(d->delv_xi, i, d->lxip[i], bc, BC_XI_P_SYMM, BC_XI_P_FREE),
        lm, ms, ptiny);
    const double phieta = mono_phi(
        d->delv_eta[i],
        face_neighbour(d->delv_eta, i, d->letam[i], bc, BC_ETA_M_SYMM, BC_ETA_M_FREE),
        face_neighbour(d->delv_eta, i, d->letap[i], bc, BC_ETA_P_SYMM, BC_ETA_P_FREE),
        lm, ms, ptiny);
    const double phizeta = mono_phi(
        d->delv_zeta[i],
        face_neighbour(d->delv_zeta, i, d->lzetam[i], bc, BC_ZETA_M_SYMM, BC_ZETA_M_FREE),
        face_neighbour(d->delv_zeta, i, d->lzetap[i], bc, BC_ZETA_P_SYMM, BC_ZETA_P_FREE),
        lm, ms, ptiny);

    if (d->vdov[i] > 0.) {
        qlin = 0.;
        qquad = 0.;
    } else {
        double dxi = d->delv_xi[i] * d->delx_xi[i];
        double deta = d->delv_eta[i] * d->delx_eta[i];
        double dzeta = d->delv_zeta[i] * d->delx_zeta[i];
        if (dxi > 0.) dxi = 0.;
        if (deta > 0.) deta = 0.;
        if (dzeta > 0.) dzeta = 0.;
        const double rho = d->elem_mass[i] / (d->volo[i] * d->vnew[i]);
        qlin = -d->qlc_monoq * rho *
               (dxi * (1. - phixi) + deta * (1. - phieta) + dzeta * (1. - phizeta));
        qquad = d->qqc_monoq * rho *
                (dxi * dxi * (1. - phixi * phixi) + deta * deta * (1. - phieta * phieta) +
                 dzeta * dzeta * (1. - phizeta * phizeta));
    }
    d->qq[i] = qquad;
    d->ql[i] = qlin;
}

/* CalcQForElems (reference/C/LULESH-OMP.cc:2351-2382) */
static int calc_q(lo_domain *d)
{
    const int num_elem = d->num_elem;
    monotonic_q_gradients(d);
#pragma omp parallel for
    for (int i = 0; i < num_elem; ++i) {
        monotonic_q_region_elem(d, i);
    }
    for (int i = 0; i < num_elem; ++i) {
        if (d->q[i] > d->qstop) {
            return LO_QSTOP_ERROR;
        }
    }
    return 0;
}

/* the equation of state of one element: CalcEnergyForElems with its three
 * CalcPressureForElems calls, then CalcSoundSpeedForElems
 * (reference/C/LULESH-OMP.cc:2384-2561; lulesh-2.0.3/lulesh.cc:2043-2231).
 * vnewc is the clamped relative volume. */
static void eos_element(const lo_domain *d, double vnewc, double e_old,
                        double delvc, double p_old_in, double q_old,
                        double qq_old, double ql_old, double *p_out,
                        double *e_out, double *q_out, double *ss_out)
{
    const double e_cut = d->e_cut, p_cut = d->p_cut, q_cut = d->q_cut;
    const double eosvmax = d->eosvmax, eosvmin = d->eosvmin;
    const double pmin = d->pmin, emin = d->emin, rho0 = d->refdens;
    /* sound-speed floor: 1.0 uses .111111e-36/.333333e-18 (LULESH-OMP.cc:
     * 2453-2454), 2.0 uses .1111111e-36/.3333333e-18 (lulesh.cc:2113-2114) */
    const double ss_tiny = (d->mode == 1) ? .111111e-36 : .1111111e-36;
    const double ss_floor = (d->mode == 1) ? .333333e-18 : .3333333e-18;
    const double c1s = 2.0 / 3.0;
    const double work = 0.;
    double p_old = p_old_in;
    double compression, comp_half_step, vchalf;
    double bvc, pbvc, p_half_step, e_new, q_new, p_new, ssc, vhalf;

    compression = 1. / vnewc - 1.;
    vchalf = vnewc - delvc * .5;
    comp_half_step = 1. / vchalf - 1.;
    if (eosvmin != 0.) {
        if (vnewc <= eosvmin) {
            comp_half_step = compression;
        }
    }
    if (eosvmax != 0.) {
        if (vnewc >= eosvmax) {
            p_old = 0.;
            compression = 0.;
            comp_half_step = 0.;
        }
    }

    e_new = e_old - 0.5 * delvc * (p_old + q_old) + 0.5 * work;
    if (e_new < emin) e_new = emin;

    /* CalcPressureForElems(pHalfStep, ..., e_new, compHalfStep) */
    bvc = c1s * (comp_half_step + 1.);
    pbvc = c1s;
    p_half_step = bvc * e_new;
    if (fabs(p_half_step) < p_cut) p_half_step = 0.0;
    if (vnewc >= eosvmax) p_half_step = 0.0;
    if (p_half_step < pmin) p_half_step = pmin;

    vhalf = 1. / (1. + comp_half_step);
    if (delvc > 0.) {
        q_new = 0.;
    } else {
        ssc = (pbvc * e_new + vhalf * vhalf * bvc * p_half_step) / rho0;
        if (ssc <= ss_tiny) {
            ssc = ss_floor;
        } else {
            ssc = sqrt(ssc);
        }
        q_new = (ssc * ql_old + qq_old);
    }
    e_new = e_new + 0.5 * delvc * (3.0 * (p_old + q_old) - 4.0 * (p_half_step + q_new));

    e_new += 0.5 * work;
    if (fabs(e_new) < e_cut) e_new = 0.;
    if (e_new < emin) e_new = emin;

    /* CalcPressureForElems(p_new, ..., e_new, compression) */
    bvc = c1s * (compression + 1.);
    pbvc = c1s;
    p_new = bvc * e_new;
    if (fabs(p_new) < p_cut) p_new = 0.0;
    if (vnewc >= eosvmax) p_new = 0.0;
    if (p_new < pmin) p_new = pmin;

    {
        const double sixth = 1.0 / 6.0;
        double q_tilde;
        if (delvc > 0.) {
            q_tilde = 0.;
        } else {
            ssc = (pbvc * e_new + vnewc * vnewc * bvc * p_new) / rho0;
            if (ssc <= ss_tiny) {
                ssc = ss_floor;
            } else {
                ssc = sqrt(ssc);
            }
            q_tilde = (ssc * ql_old + qq_old);
        }
        e_new = e_new - (7.0 * (p_old + q_old) - 8.0 * (p_half_step + q_new) +
                         (p_new + q_tilde)) * delvc * sixth;
        if (fabs(e_new) < e_cut) e_new = 0.;
        if (e_new < emin) e_new = emin;
    }

    bvc = c1s * (compression + 1.);
    pbvc = c1s;
    p_new = bvc * e_new;
    if (fabs(p_new) < p_cut) p_new = 0.0;
    if (vnewc >= eosvmax) p_new = 0.0;
    if (p_new < pmin) p_new = pmin;

    if (delvc <= 0.) {
        ssc = (pbvc * e_new + vnewc * vnewc * bvc * p_new) / rho0;
        if (ssc <= ss_tiny) {
            ssc = ss_floor;
        } else {
            ssc = sqrt(ssc);
        }
        q_new = (ssc * ql_old + qq_old);
        if (fabs(q_new) < q_cut) q_new = 0.;
    }

    /* CalcSoundSpeedForElems (reference/C/LULESH-OMP.cc:2543-2561) */
    {
        double ss_tmp = (pbvc * e_new + vnewc * vnewc * bvc * p_new) / rho0;
        if (ss_tmp <= ss_tiny) {
            ss_tmp = ss_floor;
        } else {
            ss_tmp = sqrt(ss_tmp);
        }
        *ss_out = ss_tmp;
    }
    *p_out = p_new;
    *e_out = e_new;
    *q_out = q_new;
}

/* number of EOS repetitions of region r (lulesh-2.0.3/lulesh.cc:2413-2424) */
static int region_rep(const lo_domain *d, int r)
{
    if (d->mode == 1) return 1;
    if (r < d->num_reg / 2) return 1;
    if (r < (d->num_reg - (d->num_reg + 15) / 20)) return 1 + d->cost;
    return 10 * (1 + d->cost);
}

/* ApplyMaterialPropertiesForElems + EvalEOSForElems
 * (reference/C/LULESH-OMP.cc:2563-2761; lulesh-2.0.3/lulesh.cc:2233-2430) */
static int apply_material_properties(lo_domain *d)
{
    const int num_elem = d->num_elem;
    const double eosvmin = d->eosvmin, eosvmax = d->eosvmax;
    int bad = 0;
    /* 1.0 clamps a copy, 2.0 clamps vnew in place (lulesh.cc:2371-2385) */
    double *vnewc = (d->mode == 1) ? d->vnewc : d->vnew;
#pragma omp parallel for reduction(| : bad)
    for (int i = 0; i < num_elem; ++i) {
        double vc = d->vnew[i];
        if (eosvmin != 0.) {
            if (vc < eosvmin) vc = eosvmin;
        }
        if (eosvmax != 0.) {
            if (vc > eosvmax) vc = eosvmax;
        }
        vnewc[i] = vc;
        double vo = d->v[i];
        if (eosvmin != 0.) {
            if (vo < eosvmin) vo = eosvmin;
        }
        if (eosvmax != 0.) {
            if (vo > eosvmax) vo = eosvmax;
        }
        if (vo <= 0.) {
            bad = 1;
        }
    }
    if (bad) {
        return LO_VOLUME_ERROR;
    }
    for (int r = 0; r < d->num_reg; ++r) {
        const int n = d->reg_elem_size[r];
        const int *list = d->reg_elem_list[r];
        const int rep = region_rep(d, r);
        for (int j = 0; j < rep; ++j) {
            const int last = (j == rep - 1);
#pragma omp parallel for
            for (int ii = 0; ii < n; ++ii) {
                const int i = list[ii];
                double pn, en, qn, sn;
                eos_element(d, vnewc[i], d->e[i], d->delv[i], d->p[i], d->q[i],
                            d->qq[i], d->ql[i], &pn, &en, &qn, &sn);
                if (last) {
                    /* the repetitions recompute identical values from the
                     * same inputs; only the last one is stored, as in the
                     * reference (lulesh.cc:2267-2340) */
                    d->fx_elem[i] = pn;
                    d->fy_elem[i] = en;
                    d->fz_elem[i] = qn;
                    d->determ[i] = sn;
                }
            }
        }
#pragma omp parallel for
        for (int ii = 0; ii < n; ++ii) {
            const int i = list[ii];
            d->p[i] = d->fx_elem[i];
            d->e[i] = d->fy_elem[i];
            d->q[i] = d->fz_elem[i];
            d->ss[i] = d->determ[i];
        }
    }
    return 0;
}

/* UpdateVolumesForElems (reference/C/LULESH-OMP.cc:2763-2782) */
static void update_volumes(lo_domain *d)
{
    const int num_elem = d->num_elem;
    const double v_cut = d->v_cut;
#pragma omp parallel for
    for (int i = 0; i < num_elem; ++i) {
        double t = d->vnew[i];
        if (fabs(t - 1.0) < v_cut) t = 1.0;
        d->v[i] = t;
    }
}

/* LagrangeElements (reference/C/LULESH-OMP.cc:2784-2797) */
int lo_lagrange_elements(lo_domain *d)
{
    int rc;
    apply_threads();
    rc = lagrange_elements_kinematics(d);
    if (rc) return rc;
    rc = calc_q(d);
    if (rc) return rc;
    rc = apply_material_properties(d);
    if (rc) return rc;
    update_volumes(d);
    return 0;
}

/* CalcTimeConstraintsForElems (reference/C/LULESH-OMP.cc:2799-2942;
 * lulesh-2.0.3/lulesh.cc:2473-2637).  Both minima are order independent; a
 * constraint is only replaced when some active element (vdov != 0) gave a
 * value below 1e20.  2.0 resets both to 1e20 at the start of every cycle. */
void lo_time_constraints(lo_domain *d)
{
    const int num_elem = d->num_elem;
    const double qqc2 = 64.0 * d->qqc * d->qqc;
    const double dvovmax = d->dvovmax;
    double dtc = 1.0e+20, dth = 1.0e+20;
    apply_threads();
    if (d->mode == 2) {
        d->dtcourant = 1.0e+20;
        d->dthydro = 1.0e+20;
    }
#pragma omp parallel
    {
        double c = 1.0e+20, h = 1.0e+20;
#pragma omp for
        for (int i = 0; i < num_elem; ++i) {
            double dtf = d->ss[i] * d->ss[i];
            if (d->vdov[i] < 0.) {
                dtf = dtf + qqc2 * d->arealg[i] * d->arealg[i] * d->vdov[i] * d->vdov[i];
            }
            dtf = sqrt(dtf);
            dtf = d->arealg[i] / dtf;
            if (d->vdov[i] != 0.) {
                if (dtf < c) c = dtf;
                {
                    const double dtdvov = dvovmax / (fabs(d->vdov[i]) + 1.e-20);
                    if (h > dtdvov) h = dtdvov;
                }
            }
        }
#pragma omp critical
        {
            if (c < dtc) dtc = c;
            if (h < dth) dth = h;
        }
    }
    if (dtc < 1.0e+20) d->dtcourant = dtc;
    if (dth < 1.0e+20) d->dthydro = dth;
}

/* TimeIncrement (reference/C/LULESH-OMP.cc:557-604; lulesh-2.0.3/lulesh.cc:
 * 194-251) */
void lo_time_increment(lo_domain *d)
{
    double targetdt = d->stoptime - d->time;
    if ((d->dtfixed <= 0.0) && (d->cycle != 0)) {
        double ratio;
        const double olddt = d->deltatime;
        double newdt = 1.0e+20;
        if (d->dtcourant < newdt) {
            newdt = d->dtcourant / 2.0;
        }
        if (d->dthydro < newdt) {
            newdt = d->dthydro * 2.0 / 3.0;
        }
        ratio = newdt / olddt;
        if (ratio >= 1.0) {
            if (ratio < d->deltatimemultlb) {
                newdt = olddt;
            } else if (ratio > d->deltatimemultub) {
                newdt = olddt * d->deltatimemultub;
            }
        }
        if (newdt > d->dtmax) {
            newdt = d->dtmax;
        }
        d->deltatime = newdt;
    }
    if ((targetdt > d->deltatime) && (targetdt < (4.0 * d->deltatime / 3.0))) {
        targetdt = 2.0 * d->deltatime / 3.0;
    }
    if (targetdt < d->deltatime) {
        d->deltatime = targetdt;
    }
    d->time += d->deltatime;
    ++d->cycle;
}

int lo_run(lo_domain *d, int max_cycles)
{
    while ((d->time < d->stoptime) && (d->cycle < max_cycles)) {
        int rc;
        lo_time_increment(d);
        rc = lo_lagrange_nodal(d);
        if (rc) return rc;
        rc = lo_lagrange_elements(d);
        if (rc) return rc;
        lo_time_constraints(d);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* set-up                                                                    */
/* ------------------------------------------------------------------------ */

/* glibc's rand() is the additive feedback generator of random_r.c (TYPE_3,
 * degree 31, separation 3): the state is seeded by r[i] = 16807 r[i-1]
 * mod (2^31 - 1), the first 310 outputs are discarded, and each output is
 * (r[i-3] + r[i-31]) mod 2^32, shifted right by one.  srand(0) seeds as
 * srand(1).  The oracle uses its own generator so it neither depends on nor
 * disturbs the host program's rand() state. */
typedef struct {
    int32_t r[34 + 310 + 4];
    uint32_t ring[31];
    int pos;
} lo_rand;

static void lo_srand(lo_rand *g, unsigned seed)
{
    int32_t r[34];
    if (seed == 0) seed = 1;
    r[0] = (int32_t)seed;
    for (int i = 1; i < 31; ++i) {
        const int64_t hi = r[i - 1] / 127773;
        const int64_t lo = r[i - 1] % 127773;
        int64_t word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        r[i] = (int32_t)word;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    /* ring holds the last 31 values; next = ring[pos-31] + ring[pos-3] */
    for (int i = 0; i < 31; ++i) g->ring[i] = (uint32_t)r[i + 3];
    g->pos = 0;
    for (int i = 0; i < 310; ++i) {
        const uint32_t v = g->ring[g->pos] + g->ring[(g->pos + 28) % 31];
        g->ring[g->pos] = v;
        g->pos = (g->pos + 1) % 31;
    }
}

static int lo_rand_next(lo_rand *g)
{
    const uint32_t v = g->ring[g->pos] + g->ring[(g->pos + 28) % 31];
    g->ring[g->pos] = v;
    g->pos = (g->pos + 1) % 31;
    return (int)(v >> 1);
}

/* CreateRegionIndexSets (lulesh-2.0.3/lulesh-init.cc:377-484), one domain */
static void create_regions(lo_domain *d)
{
    const int num_elem = d->num_elem;
    int num_reg = d->num_reg;
    d->reg_num_list = ialloc((size_t)num_elem);
    d->reg_elem_size = ialloc((size_t)num_reg);
    d->reg_elem_list = (int **)calloc((size_t)num_reg, sizeof(int *));
    if (num_reg == 1) {
        for (int i = 0; i < num_elem; ++i) d->reg_num_list[i] = 1;
    } else {
        lo_rand g;
        int *bin_end = ialloc((size_t)num_reg);
        int cost_den = 0, last_reg = -1, next = 0;
        lo_srand(&g, 0);
        for (int i = 0; i < num_reg; ++i) {
            cost_den += (int)pow((double)(i + 1), (double)d->balance);
            bin_end[i] = cost_den;
        }
        while (next < num_elem) {
            int var = lo_rand_next(&g) % cost_den;
            int i = 0, reg, bin, elements, runto;
            while (var >= bin_end[i]) i++;
            reg = (i % num_reg) + 1;
            while (reg == last_reg) {
                var = lo_rand_next(&g) % cost_den;
                i = 0;
                while (var >= bin_end[i]) i++;
                reg = (i % num_reg) + 1;
            }
            bin = lo_rand_next(&g) % 1000;
            if (bin < 773) elements = lo_rand_next(&g) % 15 + 1;
            else if (bin < 937) elements = lo_rand_next(&g) % 16 + 16;
            else if (bin < 970) elements = lo_rand_next(&g) % 32 + 32;
            else if (bin < 974) elements = lo_rand_next(&g) % 64 + 64;
            else if (bin < 978) elements = lo_rand_next(&g) % 128 + 128;
            else if (bin < 981) elements = lo_rand_next(&g) % 256 + 256;
            else elements = lo_rand_next(&g) % 1537 + 512;
            runto = elements + next;
            while (next < runto && next < num_elem) {
                d->reg_num_list[next] = reg;
                next++;
            }
            last_reg = reg;
        }
        free(bin_end);
    }
    for (int i = 0; i < num_elem; ++i) d->reg_elem_size[d->reg_num_list[i] - 1]++;
    for (int r = 0; r < num_reg; ++r) {
        d->reg_elem_list[r] = ialloc((size_t)d->reg_elem_size[r]);
        d->reg_elem_size[r] = 0;
    }
    for (int i = 0; i < num_elem; ++i) {
        const int r = d->reg_num_list[i] - 1;
        d->reg_elem_list[r][d->reg_elem_size[r]++] = i;
    }
}

lo_domain *lo_create(int nx, int mode, int num_reg, int balance, int cost)
{
    return lo_create_scaled(nx, mode, num_reg, balance, cost, 1.0);
}

lo_domain *lo_create_scaled(int nx, int mode, int num_reg, int balance, int cost,
                            double efactor)
{
    if (nx < 1 || (mode != 1 && mode != 2) || !(efactor > 0.0)) return NULL;
    lo_domain *d = (lo_domain *)calloc(1, sizeof(lo_domain));
    const int en = nx + 1;
    d->efactor = efactor;
    d->nx = nx;
    d->mode = mode;
    d->edge_nodes = en;
    d->num_elem = nx * nx * nx;
    d->num_node = en * en * en;
    d->num_reg = (mode == 1) ? 1 : (num_reg < 1 ? 1 : num_reg);
    d->balance = balance;
    d->cost = cost;
    d->cbrt_mode = 0;

    const size_t ne = (size_t)d->num_elem, nn = (size_t)d->num_node;
    d->x = dalloc(nn); d->y = dalloc(nn); d->z = dalloc(nn);
    d->xd = dalloc(nn); d->yd = dalloc(nn); d->zd = dalloc(nn);
    d->xdd = dalloc(nn); d->ydd = dalloc(nn); d->zdd = dalloc(nn);
    d->fx = dalloc(nn); d->fy = dalloc(nn); d->fz = dalloc(nn);
    d->nodal_mass = dalloc(nn);
    d->symm_x = ialloc((size_t)en * en);
    d->symm_y = ialloc((size_t)en * en);
    d->symm_z = ialloc((size_t)en * en);
    d->num_symm_x = d->num_symm_y = d->num_symm_z = en * en;
    d->origin = 0;
    d->nodelist = ialloc(8 * ne);
    d->lxim = ialloc(ne); d->lxip = ialloc(ne);
    d->letam = ialloc(ne); d->letap = ialloc(ne);
    d->lzetam = ialloc(ne); d->lzetap = ialloc(ne);
    d->elem_bc = ialloc(ne);
    d->e = dalloc(ne); d->p = dalloc(ne); d->q = dalloc(ne);
    d->ql = dalloc(ne); d->qq = dalloc(ne); d->v = dalloc(ne);
    d->volo = dalloc(ne); d->vnew = dalloc(ne); d->delv = dalloc(ne);
    d->vdov = dalloc(ne); d->arealg = dalloc(ne); d->ss = dalloc(ne);
    d->elem_mass = dalloc(ne);
    d->dxx = dalloc(ne); d->dyy = dalloc(ne); d->dzz = dalloc(ne);
    d->delv_xi = dalloc(ne); d->delv_eta = dalloc(ne); d->delv_zeta = dalloc(ne);
    d->delx_xi = dalloc(ne); d->delx_eta = dalloc(ne); d->delx_zeta = dalloc(ne);
    d->fx_elem = dalloc(8 * ne); d->fy_elem = dalloc(8 * ne); d->fz_elem = dalloc(8 * ne);
    d->determ = dalloc(ne);
    d->vnewc = dalloc(ne);

    /* nodal coordinates (reference/C/LULESH-OMP.cc:2996-3017) */
    {
        int nidx = 0;
        double tz = 0.;
        for (int plane = 0; plane < en; ++plane) {
            double ty = 0.;
            for (int row = 0; row < en; ++row) {
                double tx = 0.;
                for (int col = 0; col < en; ++col) {
                    d->x[nidx] = tx;
                    d->y[nidx] = ty;
                    d->z[nidx] = tz;
                    ++nidx;
                    tx = 1.125 * (double)(col + 1) / (double)nx;
                }
                ty = 1.125 * (double)(row + 1) / (double)nx;
            }
            tz = 1.125 * (double)(plane + 1) / (double)nx;
        }
    }
    /* element -> node connectivity (:3020-3042) */
    {
        int nidx = 0, zidx = 0;
        for (int plane = 0; plane < nx; ++plane) {
            for (int row = 0; row < nx; ++row) {
                for (int col = 0; col < nx; ++col) {
                    int *ln = &d->nodelist[8 * zidx];
                    ln[0] = nidx;
                    ln[1] = nidx + 1;
                    ln[2] = nidx + en + 1;
                    ln[3] = nidx + en;
                    ln[4] = nidx + en * en;
                    ln[5] = nidx + en * en + 1;
                    ln[6] = nidx + en * en + en + 1;
                    ln[7] = nidx + en * en + en;
                    ++zidx;
                    ++nidx;
                }
                ++nidx;
            }
            nidx += en;
        }
    }
    /* node -> corner list in ascending element order (:209-268) */
    {
        int *count = ialloc(nn);
        d->node_elem_start = ialloc(nn + 1);
        for (size_t i = 0; i < ne; ++i)
            for (int j = 0; j < 8; ++j) ++count[d->nodelist[8 * i + j]];
        d->node_elem_start[0] = 0;
        for (size_t i = 1; i <= nn; ++i)
            d->node_elem_start[i] = d->node_elem_start[i - 1] + count[i - 1];
        d->node_elem_corner = ialloc((size_t)d->node_elem_start[nn]);
        memset(count, 0, nn * sizeof(int));
        for (size_t i = 0; i < ne; ++i) {
            for (int j = 0; j < 8; ++j) {
                const int m = d->nodelist[8 * i + j];
                d->node_elem_corner[d->node_elem_start[m] + count[m]] = (int)(i * 8 + j);
                ++count[m];
            }
        }
        free(count);
    }

    /* parameters */
    d->e_cut = 1.0e-7;
    d->p_cut = 1.0e-7;
    d->q_cut = 1.0e-7;
    d->u_cut = 1.0e-7;
    d->v_cut = 1.0e-10;
    d->hgcoef = 3.0;
    d->ss4o3 = 4.0 / 3.0;
    d->qstop = 1.0e+12;
    d->monoq_max_slope = 1.0;
    d->monoq_limiter_mult = 2.0;
    d->qlc_monoq = 0.5;
    d->qqc_monoq = 2.0 / 3.0;
    d->qqc = 2.0;
    d->pmin = 0.;
    d->emin = -1.0e+15;
    d->dvovmax = 0.1;
    d->eosvmax = 1.0e+9;
    d->eosvmin = 1.0e-9;
    d->refdens = 1.0;
    d->dtfixed = (mode == 1) ? -1.0e-7 : -1.0e-6;
    d->deltatimemultlb = 1.1;
    d->deltatimemultub = 1.2;
    d->stoptime = 1.0e-2;
    d->dtcourant = 1.0e+20;
    d->dthydro = 1.0e+20;
    d->dtmax = 1.0e-2;
    d->time = 0.;
    d->cycle = 0;

    for (size_t i = 0; i < ne; ++i) d->v[i] = 1.0;

    /* volumes and masses (:3089-3109) */
    for (size_t i = 0; i < ne; ++i) {
        double xl[8], yl[8], zl[8];
        const int *nl = &d->nodelist[8 * i];
        gather8(d, d->x, nl, xl);
        gather8(d, d->y, nl, yl);
        gather8(d, d->z, nl, zl);
        const double volume = elem_volume(xl, yl, zl);
        d->volo[i] = volume;
        d->elem_mass[i] = volume;
        for (int j = 0; j < 8; ++j) d->nodal_mass[nl[j]] += volume / 8.0;
    }

    /* initial energy and time step: 1.0 (:3053, :3112); 2.0 scales the
     * energy by (nx/45)^3 and derives dt from it (lulesh-init.cc:167-182) */
    if (mode == 1) {
        d->e[0] = 3.948746e+7 * d->efactor;
        d->deltatime = 1.0e-7;
    } else {
        const double ebase = 3.948746e+7;
        const double scale = (double)(nx * 1) / 45.0;
        const double einit = ebase * scale * scale * scale * d->efactor;
        d->e[0] = einit;
        d->deltatime = (.5 * cbrt(d->volo[0])) / sqrt(2.0 * einit);
    }

    /* symmetry node sets (:3114-3125) */
    {
        int nidx = 0;
        for (int i = 0; i < en; ++i) {
            const int plane_inc = i * en * en, row_inc = i * en;
            for (int j = 0; j < en; ++j) {
                d->symm_x[nidx] = plane_inc + j * en;
                d->symm_y[nidx] = plane_inc + j;
                d->symm_z[nidx] = row_inc + j;
                ++nidx;
            }
        }
    }
    /* face neighbours (:3127-3151) */
    {
        const int n = d->num_elem;
        d->lxim[0] = 0;
        for (int i = 1; i < n; ++i) {
            d->lxim[i] = i - 1;
            d->lxip[i - 1] = i;
        }
        d->lxip[n - 1] = n - 1;
        for (int i = 0; i < nx; ++i) {
            d->letam[i] = i;
            d->letap[n - nx + i] = n - nx + i;
        }
        for (int i = nx; i < n; ++i) {
            d->letam[i] = i - nx;
            d->letap[i - nx] = i;
        }
        for (int i = 0; i < nx * nx; ++i) {
            d->lzetam[i] = i;
            d->lzetap[n - nx * nx + i] = n - nx * nx + i;
        }
        for (int i = nx * nx; i < n; ++i) {
            d->lzetam[i] = i - nx * nx;
            d->lzetap[i - nx * nx] = i;
        }
    }
    /* boundary conditions (:3153-3171) */
    for (int i = 0; i < nx; ++i) {
        const int plane_inc = i * nx * nx, row_inc = i * nx;
        for (int j = 0; j < nx; ++j) {
            d->elem_bc[plane_inc + j * nx] |= BC_XI_M_SYMM;
            d->elem_bc[plane_inc + j * nx + nx - 1] |= BC_XI_P_FREE;
            d->elem_bc[plane_inc + j] |= BC_ETA_M_SYMM;
            d->elem_bc[plane_inc + j + nx * nx - nx] |= BC_ETA_P_FREE;
            d->elem_bc[row_inc + j] |= BC_ZETA_M_SYMM;
            d->elem_bc[row_inc + j + d->num_elem - nx * nx] |= BC_ZETA_P_FREE;
        }
    }
    create_regions(d);
    return d;
}

/* An unstructured hexahedral mesh, LULESH 1.0 semantics, set up as the
 * reference's mesh reader sets up a .lmesh input
 * (reference/CUDA/lulesh-kepler-singlegpu/lulesh.cu:728-925): coordinates,
 * node lists, face neighbours and node sets from the caller; the element
 * boundary masks derived from the node sets -- a face whose four nodes all
 * lie in the union of the symmetry sets is a symmetry face, and so is (as
 * the reader does it) a face whose four nodes all lie in the free-surface
 * set; the energy deposited in the first element whose mask is exactly the
 * three low symmetry faces (the octant corner). */
lo_domain *lo_create_mesh(int num_elem, int num_node, const double *x, const double *y,
                          const double *z, const int *nodelist, const int *neighbors,
                          int nsx, const int *symm_x, int nsy, const int *symm_y,
                          int nsz, const int *symm_z, int nfree, const int *free_surf)
{
    if (num_elem < 1 || num_node < 8 || !x || !y || !z || !nodelist || !neighbors ||
        nsx < 0 || nsy < 0 || nsz < 0 || nfree < 0) return NULL;
    for (size_t i = 0; i < 8 * (size_t)num_elem; ++i)
        if (nodelist[i] < 0 || nodelist[i] >= num_node) return NULL;
    for (size_t i = 0; i < 6 * (size_t)num_elem; ++i)
        if (neighbors[i] < 0 || neighbors[i] >= num_elem) return NULL;
    lo_domain *d = (lo_domain *)calloc(1, sizeof(lo_domain));
    const size_t ne = (size_t)num_elem, nn = (size_t)num_node;
    d->nx = 0;
    d->mode = 1;
    d->edge_nodes = 0;
    d->num_elem = num_elem;
    d->num_node = num_node;
    d->num_reg = 1;
    d->balance = 1;
    d->cost = 1;
    d->cbrt_mode = 0;
    d->efactor = 1.0;
    d->x = dalloc(nn); d->y = dalloc(nn); d->z = dalloc(nn);
    d->xd = dalloc(nn); d->yd = dalloc(nn); d->zd = dalloc(nn);
    d->xdd = dalloc(nn); d->ydd = dalloc(nn); d->zdd = dalloc(nn);
    d->fx = dalloc(nn); d->fy = dalloc(nn); d->fz = dalloc(nn);
    d->nodal_mass = dalloc(nn);
    d->num_symm_x = nsx; d->num_symm_y = nsy; d->num_symm_z = nsz;
    d->symm_x = ialloc((size_t)nsx + 1);
    d->symm_y = ialloc((size_t)nsy + 1);
    d->symm_z = ialloc((size_t)nsz + 1);
    d->nodelist = ialloc(8 * ne);
    d->lxim = ialloc(ne); d->lxip = ialloc(ne);
    d->letam = ialloc(ne); d->letap = ialloc(ne);
    d->lzetam = ialloc(ne); d->lzetap = ialloc(ne);
    d->elem_bc = ialloc(ne);
    d->e = dalloc(ne); d->p = dalloc(ne); d->q = dalloc(ne);
    d->ql = dalloc(ne); d->qq = dalloc(ne); d->v = dalloc(ne);
    d->volo = dalloc(ne); d->vnew = dalloc(ne); d->delv = dalloc(ne);
    d->vdov = dalloc(ne); d->arealg = dalloc(ne); d->ss = dalloc(ne);
    d->elem_mass = dalloc(ne);
    d->dxx = dalloc(ne); d->dyy = dalloc(ne); d->dzz = dalloc(ne);
    d->delv_xi = dalloc(ne); d->delv_eta = dalloc(ne); d->delv_zeta = dalloc(ne);
    d->delx_xi = dalloc(ne); d->delx_eta = dalloc(ne); d->delx_zeta = dalloc(ne);
    d->fx_elem = dalloc(8 * ne); d->fy_elem = dalloc(8 * ne); d->fz_elem = dalloc(8 * ne);
    d->determ = dalloc(ne);
    d->vnewc = dalloc(ne);

    memcpy(d->x, x, nn * sizeof(double));
    memcpy(d->y, y, nn * sizeof(double));
    memcpy(d->z, z, nn * sizeof(double));
    memcpy(d->nodelist, nodelist, 8 * ne * sizeof(int));
    for (size_t i = 0; i < ne; ++i) {
        d->lxim[i] = neighbors[6 * i + 0];
        d->lxip[i] = neighbors[6 * i + 1];
        d->letam[i] = neighbors[6 * i + 2];
        d->letap[i] = neighbors[6 * i + 3];
        d->lzetam[i] = neighbors[6 * i + 4];
        d->lzetap[i] = neighbors[6 * i + 5];
    }
    if (nsx) memcpy(d->symm_x, symm_x, (size_t)nsx * sizeof(int));
    if (nsy) memcpy(d->symm_y, symm_y, (size_t)nsy * sizeof(int));
    if (nsz) memcpy(d->symm_z, symm_z, (size_t)nsz * sizeof(int));

    /* node -> corner list in ascending element order */
    {
        int *count = ialloc(nn);
        d->node_elem_start = ialloc(nn + 1);
        for (size_t i = 0; i < ne; ++i)
            for (int j = 0; j < 8; ++j) ++count[d->nodelist[8 * i + j]];
        d->node_elem_start[0] = 0;
        for (size_t i = 1; i <= nn; ++i)
            d->node_elem_start[i] = d->node_elem_start[i - 1] + count[i - 1];
        d->node_elem_corner = ialloc((size_t)d->node_elem_start[nn]);
        memset(count, 0, nn * sizeof(int));
        for (size_t i = 0; i < ne; ++i) {
            for (int j = 0; j < 8; ++j) {
                const int m = d->nodelist[8 * i + j];
                d->node_elem_corner[d->node_elem_start[m] + count[m]] = (int)(i * 8 + j);
                ++count[m];
            }
        }
        free(count);
    }
    /* boundary masks from the node sets */
    {
        int *surf = ialloc(nn);
        const int *sets[2][3] = {{symm_x, symm_y, symm_z}, {free_surf, NULL, NULL}};
        const int sizes[2][3] = {{nsx, nsy, nsz}, {nfree, 0, 0}};
        for (int pass = 0; pass < 2; ++pass) {
            memset(surf, 0, nn * sizeof(int));
            for (int s = 0; s < 3; ++s)
                for (int i = 0; i < sizes[pass][s]; ++i) surf[sets[pass][s][i]] = 1;
            for (size_t i = 0; i < ne; ++i) {
                int mask = 0;
                for (int j = 0; j < 8; ++j) mask |= surf[d->nodelist[8 * i + j]] << j;
                if ((mask & 0x0f) == 0x0f) d->elem_bc[i] |= BC_ZETA_M_SYMM;
                if ((mask & 0xf0) == 0xf0) d->elem_bc[i] |= BC_ZETA_P_SYMM;
                if ((mask & 0x33) == 0x33) d->elem_bc[i] |= BC_ETA_M_SYMM;
                if ((mask & 0xcc) == 0xcc) d->elem_bc[i] |= BC_ETA_P_SYMM;
                if ((mask & 0x99) == 0x99) d->elem_bc[i] |= BC_XI_M_SYMM;
                if ((mask & 0x66) == 0x66) d->elem_bc[i] |= BC_XI_P_SYMM;
            }
            if (pass == 0) {
                d->origin = 0;
                for (size_t i = 0; i < ne; ++i) {
                    if (d->elem_bc[i] == (BC_XI_M_SYMM | BC_ETA_M_SYMM | BC_ZETA_M_SYMM)) {
                        d->origin = (int)i;
                        break;
                    }
                }
            }
        }
        free(surf);
    }

    d->e_cut = 1.0e-7;
    d->p_cut = 1.0e-7;
    d->q_cut = 1.0e-7;
    d->u_cut = 1.0e-7;
    d->v_cut = 1.0e-10;
    d->hgcoef = 3.0;
    d->ss4o3 = 4.0 / 3.0;
    d->qstop = 1.0e+12;
    d->monoq_max_slope = 1.0;
    d->monoq_limiter_mult = 2.0;
    d->qlc_monoq = 0.5;
    d->qqc_monoq = 2.0 / 3.0;
    d->qqc = 2.0;
    d->pmin = 0.;
    d->emin = -1.0e+15;
    d->dvovmax = 0.1;
    d->eosvmax = 1.0e+9;
    d->eosvmin = 1.0e-9;
    d->refdens = 1.0;
    d->dtfixed = -1.0e-7;
    d->deltatimemultlb = 1.1;
    d->deltatimemultub = 1.2;
    d->stoptime = 1.0e-2;
    d->dtcourant = 1.0e+20;
    d->dthydro = 1.0e+20;
    d->dtmax = 1.0e-2;
    d->time = 0.;
    d->cycle = 0;
    for (size_t i = 0; i < ne; ++i) d->v[i] = 1.0;
    for (size_t i = 0; i < ne; ++i) {
        double xl[8], yl[8], zl[8];
        const int *nl = &d->nodelist[8 * i];
        gather8(d, d->x, nl, xl);
        gather8(d, d->y, nl, yl);
        gather8(d, d->z, nl, zl);
        const double volume = elem_volume(xl, yl, zl);
        d->volo[i] = volume;
        d->elem_mass[i] = volume;
        for (int j = 0; j < 8; ++j) d->nodal_mass[nl[j]] += volume / 8.0;
    }
    d->e[d->origin] = 3.948746e+7;
    d->deltatime = 1.0e-7;
    create_regions(d);
    return d;
}

void lo_destroy(lo_domain *d)
{
    if (!d) return;
    double *dp[] = {d->x, d->y, d->z, d->xd, d->yd, d->zd, d->xdd, d->ydd,
                    d->zdd, d->fx, d->fy, d->fz, d->nodal_mass, d->e, d->p,
                    d->q, d->ql, d->qq, d->v, d->volo, d->vnew, d->delv,
                    d->vdov, d->arealg, d->ss, d->elem_mass, d->dxx, d->dyy,
                    d->dzz, d->delv_xi, d->delv_eta, d->delv_zeta, d->delx_xi,
                    d->delx_eta, d->delx_zeta, d->fx_elem, d->fy_elem,
                    d->fz_elem, d->determ, d->vnewc};
    int *ip[] = {d->symm_x, d->symm_y, d->symm_z, d->node_elem_start,
                 d->node_elem_corner, d->nodelist, d->lxim, d->lxip,
                 d->letam, d->letap, d->lzetam, d->lzetap, d->elem_bc,
                 d->reg_num_list, d->reg_elem_size};
    for (size_t i = 0; i < sizeof(dp) / sizeof(dp[0]); ++i) free(dp[i]);
    for (int r = 0; r < d->num_reg; ++r) free(d->reg_elem_list[r]);
    free(d->reg_elem_list);
    for (size_t i = 0; i < sizeof(ip) / sizeof(ip[0]); ++i) free(ip[i]);
    free(d);
}

void lo_set_cbrt_mode(lo_domain *d, int mode) { d->cbrt_mode = mode; }
void lo_set_stoptime(lo_domain *d, double stoptime) { d->stoptime = stoptime; }
int lo_num_elem(const lo_domain *d) { return d->num_elem; }
int lo_num_node(const lo_domain *d) { return d->num_node; }

static double *field_ptr(const lo_domain *d, const char *name, int *n)
{
    struct { const char *name; double *p; int node; } t[] = {
        {"x", d->x, 1}, {"y", d->y, 1}, {"z", d->z, 1},
        {"xd", d->xd, 1}, {"yd", d->yd, 1}, {"zd", d->zd, 1},
        {"xdd", d->xdd, 1}, {"ydd", d->ydd, 1}, {"zdd", d->zdd, 1},
        {"fx", d->fx, 1}, {"fy", d->fy, 1}, {"fz", d->fz, 1},
        {"nodalMass", d->nodal_mass, 1},
        {"e", d->e, 0}, {"p", d->p, 0}, {"q", d->q, 0}, {"ql", d->ql, 0},
        {"qq", d->qq, 0}, {"v", d->v, 0}, {"volo", d->volo, 0},
        {"vnew", d->vnew, 0}, {"delv", d->delv, 0}, {"vdov", d->vdov, 0},
        {"arealg", d->arealg, 0}, {"ss", d->ss, 0},
        {"elemMass", d->elem_mass, 0},
        {"delv_xi", d->delv_xi, 0}, {"delv_eta", d->delv_eta, 0},
        {"delv_zeta", d->delv_zeta, 0}, {"delx_xi", d->delx_xi, 0},
        {"delx_eta", d->delx_eta, 0}, {"delx_zeta", d->delx_zeta, 0},
    };
    for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); ++i) {
        if (strcmp(t[i].name, name) == 0) {
            *n = t[i].node ? d->num_node : d->num_elem;
            return t[i].p;
        }
    }
    return NULL;
}

int lo_get(const lo_domain *d, const char *name, double *out)
{
    int n = 0;
    double *p = field_ptr(d, name, &n);
    if (!p) return -1;
    memcpy(out, p, (size_t)n * sizeof(double));
    return n;
}

int lo_set(lo_domain *d, const char *name, const double *in)
{
    int n = 0;
    double *p = field_ptr(d, name, &n);
    if (!p) return -1;
    memcpy(p, in, (size_t)n * sizeof(double));
    return n;
}

int lo_get_int(const lo_domain *d, const char *name, int *out)
{
    if (strcmp(name, "regNumList") == 0) {
        memcpy(out, d->reg_num_list, (size_t)d->num_elem * sizeof(int));
        return d->num_elem;
    }
    if (strcmp(name, "elemBC") == 0) {
        memcpy(out, d->elem_bc, (size_t)d->num_elem * sizeof(int));
        return d->num_elem;
    }
    if (strcmp(name, "nodelist") == 0) {
        memcpy(out, d->nodelist, (size_t)d->num_elem * 8 * sizeof(int));
        return d->num_elem * 8;
    }
    if (strcmp(name, "regElemSize") == 0) {
        memcpy(out, d->reg_elem_size, (size_t)d->num_reg * sizeof(int));
        return d->num_reg;
    }
    return -1;
}

double lo_scalar(const lo_domain *d, const char *name)
{
    if (strcmp(name, "time") == 0) return d->time;
    if (strcmp(name, "deltatime") == 0) return d->deltatime;
    if (strcmp(name, "dtcourant") == 0) return d->dtcourant;
    if (strcmp(name, "dthydro") == 0) return d->dthydro;
    if (strcmp(name, "stoptime") == 0) return d->stoptime;
    if (strcmp(name, "cycle") == 0) return (double)d->cycle;
    if (strcmp(name, "origin") == 0) return (double)d->origin;
    if (strcmp(name, "e_origin") == 0) return d->e[d->origin];
    return NAN;
}

void lo_set_scalar(lo_domain *d, const char *name, double value)
{
    struct { const char *name; double *p; } t[] = {
        {"time", &d->time}, {"deltatime", &d->deltatime},
        {"dtcourant", &d->dtcourant}, {"dthydro", &d->dthydro},
        {"stoptime", &d->stoptime}, {"dtmax", &d->dtmax},
        {"dtfixed", &d->dtfixed}, {"hgcoef", &d->hgcoef},
        {"qstop", &d->qstop}, {"monoq_max_slope", &d->monoq_max_slope},
        {"monoq_limiter_mult", &d->monoq_limiter_mult},
        {"qlc_monoq", &d->qlc_monoq}, {"qqc_monoq", &d->qqc_monoq},
        {"qqc", &d->qqc}, {"eosvmax", &d->eosvmax}, {"eosvmin", &d->eosvmin},
        {"pmin", &d->pmin}, {"emin", &d->emin}, {"dvovmax", &d->dvovmax},
        {"refdens", &d->refdens}, {"e_cut", &d->e_cut}, {"p_cut", &d->p_cut},
        {"q_cut", &d->q_cut}, {"u_cut", &d->u_cut}, {"v_cut", &d->v_cut},
    };
    if (strcmp(name, "cycle") == 0) {
        d->cycle = (int)value;
        return;
    }
    for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); ++i) {
        if (strcmp(t[i].name, name) == 0) {
            *t[i].p = value;
            return;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* element-local kernels for closed-form tests                               */
/* ------------------------------------------------------------------------ */

double lo_elem_volume(const double x[8], const double y[8], const double z[8])
{
    return elem_volume(x, y, z);
}

void lo_elem_volume_derivative(double dvdx[8], double dvdy[8], double dvdz[8],
                               const double x[8], const double y[8],
                               const double z[8])
{
    elem_volume_derivative(dvdx, dvdy, dvdz, x, y, z);
}

void lo_elem_node_normals(double pfx[8], double pfy[8], double pfz[8],
                          const double x[8], const double y[8],
                          const double z[8])
{
    elem_node_normals(pfx, pfy, pfz, x, y, z);
}

double lo_elem_shape_derivatives(const double x[8], const double y[8],
                                 const double z[8], double b[24])
{
    double bb[3][8], vol;
    shape_function_derivatives(x, y, z, bb, &vol);
    memcpy(b, bb, sizeof(bb));
    return vol;
}

void lo_elem_hourglass_force(const double x[8], const double y[8],
                             const double z[8], const double xd[8],
                             const double yd[8], const double zd[8],
                             double determ, double coefficient,
                             double hgfx[8], double hgfy[8], double hgfz[8])
{
    double dvdx[8], dvdy[8], dvdz[8];
    elem_volume_derivative(dvdx, dvdy, dvdz, x, y, z);
    elem_fb_hourglass(x, y, z, dvdx, dvdy, dvdz, xd, yd, zd, determ,
                      coefficient, hgfx, hgfy, hgfz);
}

double lo_cbrt_project(double x) { return lulesh_cbrt(x); }

/* first n outputs of the rand() restatement after seeding, for the test that
 * compares it with the C library's rand() */
void lo_rand_sequence(unsigned seed, int n, int *out)
{
    lo_rand g;
    lo_srand(&g, seed);
    for (int i = 0; i < n; ++i) out[i] = lo_rand_next(&g);
}
