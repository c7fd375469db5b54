/*
 * ref_dump_omp.cc -- TEST INFRASTRUCTURE.  Builds the reference LULESH 1.0
 * OpenMP program (reference/C/LULESH-OMP.cc, included from where it lies;
 * the path comes from oracle/Makefile as LULESH_OMP_SOURCE) with its main()
 * renamed, runs it unchanged, and dumps its global `domain` after the run.
 *
 * $LULESH_REF_STOPTIME, when set, replaces domain.stoptime() just before the
 * cycle loop starts (the reference reads the wall clock exactly there,
 * reference/C/LULESH-OMP.cc:3175-3176), to stop a run early at a given
 * simulated time without editing the reference.
 */
#include <stdlib.h>
#include <sys/time.h>

static int lulesh_ref_loop_clock(struct timeval *tv, void *tz);

#define gettimeofday(tv, tz) lulesh_ref_loop_clock(tv, tz)
#define main lulesh_omp_reference_main
#include LULESH_OMP_SOURCE
#undef main
#undef gettimeofday

#include "ref_dump.h"

static int lulesh_ref_loop_clock(struct timeval *tv, void *tz)
{
    static int calls = 0;
    if (calls++ == 0) {
        const char *s = getenv("LULESH_REF_STOPTIME");
        if (s) {
            domain.stoptime() = atof(s);
        }
    }
    return gettimeofday(tv, (struct timezone *)tz);
}

int main(int argc, char *argv[])
{
    const int rc = lulesh_omp_reference_main(argc, argv);
    lulesh_ref_dump(domain, domain.sizeX(), (const int *)0);
    return rc;
}
