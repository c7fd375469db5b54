/*
 * ref_dump.h -- TEST INFRASTRUCTURE.  Writes the final state of a reference
 * LULESH run (built from the reference sources by oracle/Makefile) to the
 * file named by $LULESH_REF_DUMP, so the tests can compare the oracle with
 * the reference bit for bit.
 *
 * Layout (little endian): "LULDUMP1", int32 nx, num_elem, num_node, cycle,
 * float64 time, deltatime, dtcourant, dthydro, then the node arrays
 * x y z xd yd zd nodalMass (num_node each), then the element arrays
 * e p q ql qq v volo delv vdov arealg ss elemMass (num_elem each), then
 * int32 regNumList (num_elem; all ones for LULESH 1.0).
 */
#ifndef LULESH_REF_DUMP_H
#define LULESH_REF_DUMP_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

template <typename Dom>
static void lulesh_ref_dump(Dom &d, int nx, const int *reg_num_list)
{
    const char *path = getenv("LULESH_REF_DUMP");
    if (!path) {
        return;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        perror("LULESH_REF_DUMP");
        return;
    }
    const int32_t ne = d.numElem(), nn = d.numNode();
    const int32_t hdr[4] = {nx, ne, nn, (int32_t)d.cycle()};
    const double sc[4] = {d.time(), d.deltatime(), d.dtcourant(), d.dthydro()};
    fwrite("LULDUMP1", 1, 8, f);
    fwrite(hdr, sizeof(int32_t), 4, f);
    fwrite(sc, sizeof(double), 4, f);
    for (int32_t i = 0; i < nn; ++i) { double v = d.x(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < nn; ++i) { double v = d.y(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < nn; ++i) { double v = d.z(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < nn; ++i) { double v = d.xd(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < nn; ++i) { double v = d.yd(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < nn; ++i) { double v = d.zd(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < nn; ++i) { double v = d.nodalMass(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.e(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.p(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.q(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.ql(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.qq(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.v(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.volo(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.delv(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.vdov(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.arealg(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.ss(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) { double v = d.elemMass(i); fwrite(&v, 8, 1, f); }
    for (int32_t i = 0; i < ne; ++i) {
        int32_t r = reg_num_list ? reg_num_list[i] : 1;
        fwrite(&r, 4, 1, f);
    }
    fclose(f);
}

#endif /* LULESH_REF_DUMP_H */
