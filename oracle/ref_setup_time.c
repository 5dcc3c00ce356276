/*
 * ref_setup_time.c -- times the reference's own lattice set-up
 * (StartLattice, lattice.c:225-252: per site StartSite, lattice.c:15-63,
 * whose InitializeFermionPPArray / InitializeFermionMMArray,
 * coordinate.c:322-374, each search all VOLUME sites, coordinate.c:253-267)
 * at the size the reference's macro.h ships (8^3).  Linked against the
 * reference's sources compiled in place by oracle/Makefile; prints
 * "<VOLUME> <seconds per StartLattice>" (the median of 21 calls).  The set-up
 * is O(V^2); the benchmark's CPU baseline extrapolates this figure to larger
 * lattices.
 *
 * TEST INFRASTRUCTURE ONLY.
 */
#include "lattice.h"

#include <stdio.h>
#include <stdlib.h>
#include <time.h>

static int cmp(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int main(void)
{
    struct Lattice lat;
    AllocateLattice(&lat);
    double t[21];
    for (int k = 0; k < 21; k++) {
        struct timespec a, b;
        clock_gettime(CLOCK_MONOTONIC, &a);
        StartLattice(&lat);
        clock_gettime(CLOCK_MONOTONIC, &b);
        t[k] = (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
    }
    qsort(t, 21, sizeof(double), cmp);
    printf("%d %.9e\n", VOLUME, t[10]);
    return 0;
}
