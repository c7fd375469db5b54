/*
 * ref_dump2.cc -- TEST INFRASTRUCTURE.  Linked with the reference LULESH
 * 2.0.3 sources (compiled where they lie by oracle/Makefile, with the
 * reference's VerifyAndWriteFinalOutput renamed to
 * lulesh_ref_verify_output).  The reference main() calls
 * VerifyAndWriteFinalOutput once after its cycle loop
 * (reference/C/lulesh-2.0.3/lulesh.cc:2804-2806); this definition prints the
 * reference's own summary and then dumps the domain.
 */
#include "lulesh.h"
#include "ref_dump.h"

void lulesh_ref_verify_output(Real_t elapsed_time, Domain &locDom, Int_t nx,
                              Int_t numRanks);

void VerifyAndWriteFinalOutput(Real_t elapsed_time, Domain &locDom, Int_t nx,
                               Int_t numRanks)
{
    lulesh_ref_verify_output(elapsed_time, locDom, nx, numRanks);
    lulesh_ref_dump(locDom, nx, locDom.regNumList());
}
