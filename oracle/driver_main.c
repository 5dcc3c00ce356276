/*
 * driver_main.c -- the reference's own driver flow (test.c:3-42) with fixed
 * arguments and a fixed seed, stopped after its first measurement block.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Linked (oracle/Makefile) against the reference's objects compiled from its
 * sources in place, twice:
 *   _ref/ref_driver     the reference alone (-DREF_ONLY);
 *   _ref/dropin_driver  the same objects with the drop-in's functions made
 *                       local (objcopy --localize-symbol), so every call the
 *                       driver makes into them -- the LFMN2VF trajectory,
 *                       the condensate, the checkpoint -- binds to
 *                       libgnx_ref.so instead.
 * Both run UpdateLattice / TakeMeasures (observables.c:178-310) unchanged
 * and write the observables, acceptance and terma files through the
 * reference's own code (inputoutput.c, observables.c:259-304,
 * moleculardynamics.c:1426, simulation.c:129-185).
 *
 * Two calls are wrapped at link time (-Wl,--wrap):
 *   StartSeed (random.c:72-80, time-based) -> Initrand(DRIVER_SEED), as a
 *     reproducible run needs;
 *   WriteBinaryLattice -> the real one, then exit(0): the first checkpoint
 *     is written right after the first observables line
 *     (observables.c:290-303), so this ends the run after IMEAS = 100
 *     updates instead of MEASURE * IMEAS; exit() flushes the observables
 *     line, which is still buffered at that point.
 *
 * usage: driver WORKDIR   (output files under WORKDIR/dati_simulation3d/)
 */
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "observables.h"

#define DRIVER_SEED 20241

void myrand_init(unsigned long int initstate, unsigned long int initseq); /* new_random.h:24 */

#ifndef REF_ONLY
int GnxRefInit(int lspace, int ltime, int device, double (*gauss1)(void), double (*casuale)(void));
void GnxRefSetTolerances(double high, double medium);
#endif

void __real_WriteBinaryLattice(const struct Simulation *simulation);

void __wrap_WriteBinaryLattice(const struct Simulation *simulation)
{
    __real_WriteBinaryLattice(simulation);
    exit(0);
}

void __wrap_StartSeed(int seed)
{
    (void)seed;
    myrand_init(24912, 2413);
    Initrand(DRIVER_SEED);
}

int main(int argc, char **argv)
{
    if (argc < 2 || chdir(argv[1]) != 0) {
        fprintf(stderr, "usage: %s WORKDIR\n", argv[0]);
        return 1;
    }
#ifndef REF_ONLY
    if (GnxRefInit(LSPACE, LTIME, 0, Gauss1, Casuale) != 0) return 3;
    GnxRefSetTolerances(HIGH_PRECISION_INVERSION, MEDIUM_PRECISION_INVERSION);
#endif
    CreateAllRequiredDirectories();
    /* zeroed first: with PI_INTERACTION off StartInfo leaves the pi fields of
     * Info unset, and the checkpoint writes them (simulation.c:176-179), so
     * test.c's stack-allocated Simulation puts whatever the stack held into
     * the file */
    struct Simulation simulation;
    memset(&simulation, 0, sizeof(simulation));
    AllocateSimulation(&simulation);
    struct Observables obs;
    AllocateObservables(&obs);
    /* beta, beta_scalar, mass, epsilon link/sigma/pi, steps link/sigma/pi,
     * sosia (test.c:11-20 reads them from argv) */
    StartSimulation(&simulation, 0.0, 0.85, 0.0, 0.1, 0.1, 0.1, 2, 2, 2, 10);
    ReadBinaryLattice(&simulation, simulation.terma_file, true);
    TakeMeasures(&simulation, &obs);
    return 2; /* not reached: the first checkpoint ends the run */
}
