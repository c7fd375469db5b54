/*
 * lulesh_oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C restatement of the LULESH Lagrange time step (TimeIncrement +
 * LagrangeLeapFrog) of the reference C programs:
 *   mode 1 -> LULESH 1.0 semantics, reference/C/LULESH-OMP.cc
 *   mode 2 -> LULESH 2.0.3 semantics, reference/C/lulesh-2.0.3/lulesh.cc,
 *             lulesh-init.cc, lulesh.h
 * It is the checker for the HIP path.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - bitwise against binaries built from the reference sources
 *     (oracle/Makefile -> oracle/_ref/), full final state, both modes;
 *   - the published known-answer tables (cycles, origin energy);
 *   - closed forms (unit-cube volume, translation/Euler identities of the
 *     volume derivative, zero hourglass force for affine velocity fields).
 */
#ifndef LULESH_ORACLE_H
#define LULESH_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lo_domain lo_domain;

/* mode: 1 = LULESH 1.0 (LULESH-OMP.cc), 2 = LULESH 2.0.3.
 * num_reg/balance/cost are the 2.0 -r/-b/-c options (ignored in mode 1). */
lo_domain *lo_create(int nx, int mode, int num_reg, int balance, int cost);
/* An ensemble member: the initial origin energy (and, in mode 2, the
 * initial dt derived from it) is that of lo_create times efactor, applied
 * as the last factor of the energy product. */
lo_domain *lo_create_scaled(int nx, int mode, int num_reg, int balance, int cost,
                            double efactor);
/* An unstructured hexahedral mesh in mode 1 (LULESH 1.0), as the
 * reference's .lmesh reader builds it (reference/CUDA/
 * lulesh-kepler-singlegpu/lulesh.cu:728-925): nodelist [num_elem][8],
 * neighbors [num_elem][6] (xi-, xi+, eta-, eta+, zeta-, zeta+), the
 * symmetry node sets and the free-surface node set.  lo_scalar(d, "origin")
 * is the element holding the initial energy. */
lo_domain *lo_create_mesh(int num_elem, int num_node, const double *x, const double *y,
                          const double *z, const int *nodelist, const int *neighbors,
                          int nsx, const int *symm_x, int nsy, const int *symm_y,
                          int nsz, const int *symm_z, int nfree, const int *free_surf);
void lo_destroy(lo_domain *d);

/* cube root used by the hourglass coefficient: 0 = C library cbrt (the
 * reference), 1 = the project cube root (lulesh_cbrt.h, as on the GPU). */
void lo_set_cbrt_mode(lo_domain *d, int mode);
void lo_set_stoptime(lo_domain *d, double stoptime);
void lo_set_threads(int nthreads);

/* Runs cycles while time < stoptime and cycle < max_cycles (the 2.0 -i
 * semantics); returns 0, -1 (VolumeError) or -2 (QStopError). */
int lo_run(lo_domain *d, int max_cycles);

/* Phase entry points, in LagrangeLeapFrog order. */
void lo_time_increment(lo_domain *d);
int lo_lagrange_nodal(lo_domain *d);
int lo_lagrange_elements(lo_domain *d);
void lo_time_constraints(lo_domain *d);

int lo_num_elem(const lo_domain *d);
int lo_num_node(const lo_domain *d);

/* Field access by name; returns the number of values copied, or -1. */
int lo_get(const lo_domain *d, const char *name, double *out);
int lo_set(lo_domain *d, const char *name, const double *in);
int lo_get_int(const lo_domain *d, const char *name, int *out);
double lo_scalar(const lo_domain *d, const char *name);
void lo_set_scalar(lo_domain *d, const char *name, double value);

/* Element-local kernels exposed for closed-form tests. */
double lo_elem_volume(const double x[8], const double y[8], const double z[8]);
void lo_elem_volume_derivative(double dvdx[8], double dvdy[8], double dvdz[8],
                               const double x[8], const double y[8],
                               const double z[8]);
void lo_elem_node_normals(double pfx[8], double pfy[8], double pfz[8],
                          const double x[8], const double y[8],
                          const double z[8]);
double lo_elem_shape_derivatives(const double x[8], const double y[8],
                                 const double z[8], double b[24]);
void lo_elem_hourglass_force(const double x[8], const double y[8],
                             const double z[8], const double xd[8],
                             const double yd[8], const double zd[8],
                             double determ, double coefficient,
                             double hgfx[8], double hgfy[8], double hgfz[8]);
double lo_cbrt_project(double x);
void lo_rand_sequence(unsigned seed, int n, int *out);

#ifdef __cplusplus
}
#endif

#endif /* LULESH_ORACLE_H */
