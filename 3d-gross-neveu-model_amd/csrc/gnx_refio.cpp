// gnx_refio.cpp -- the driver's terma checkpoint (simulation.c:104-401) over
// the reference's struct layouts, so that a run driven through the drop-in
// writes, and resumes from, the files the reference's own driver writes and
// reads.  Byte layout of the shipped configuration (SIGMA_INTERACTION on,
// PI_INTERACTION off, macro.h:49-50), in write order:
//   U_d(x)   complex, x = 0..V-1 in lex order, d = 0..2 (COOR(x,d), macro.h:39)
//   sigma[x] double, x = 0..V-1 in f-order
//   test     double: the sum over x, then d, of carg U_d(x) + sm[x]/3
//   epsilon_link_HMC double, story_link (8 bytes), hybrid_accepted_link int,
//   hybrid_done_link int, story_sigma (8 bytes), hybrid_accepted_sigma int,
//   hybrid_done_sigma int, epsilon_pi_HMC double, story_pi (8 bytes),
//   hybrid_accepted_pi int, hybrid_done_pi int, measure int
// The story fields are ints that the reference writes and reads with
// sizeof(double) (simulation.c:168,172,177 and :278,308,348), so each slot
// carries the field and the 4 bytes after it in struct Info -- the next
// story field, or the struct's tail padding -- and a read overwrites those
// 4 bytes too.  That is kept, so the files are interchangeable byte for
// byte.  Errors follow the reference: a file that
// cannot be opened for writing, or a damaged file read with try_copy false,
// prints the reference's message and exits with status 1.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/stat.h>

#include "../../include/gnx_refapi.h"

// lattice volume bound by GnxRefInit (gnx_refshim.cpp)
long gnx_ref_bound_volume();

namespace {

constexpr int kDim = 3;  // DIM, macro.h:23

}  // namespace

extern "C" {

// simulation.c:104-121 (SIGMA_INTERACTION on, PI_INTERACTION off)
double TestWriteReadLattice(const struct Simulation *simulation) {
  const struct Lattice *lat = simulation->lattice;
  const long V = gnx_ref_bound_volume();
  double total = 0.0;
  for (long x = 0; x < V; x++)
    for (int d = 0; d < kDim; d++) {
      const gnx_ref_complex u = lat->dlink[x * kDim + d].link;
      total += atan2(u.imag(), u.real());  // carg
      total += lat->sigmamass[x] / kDim;
    }
  return total;
}

// simulation.c:129-185
void WriteBinaryLattice(const struct Simulation *simulation) {
  // the previous file becomes the copy, an older copy is removed first
  struct stat buffer;
  if (stat(simulation->terma_file_copy, &buffer) == 0) remove(simulation->terma_file_copy);
  rename(simulation->terma_file, simulation->terma_file_copy);

  FILE *opf = fopen(simulation->terma_file, "w");
  if (opf == NULL) {
    fprintf(stderr, "It is impossible to open the file file in writing mode\n");
    exit(1);
  }
  const long V = gnx_ref_bound_volume();
  const struct Lattice *lat = simulation->lattice;
  const struct Info *info = &(simulation->info);
  fwrite(lat->dlink, sizeof(struct Link), (size_t)(V * kDim), opf);
  fwrite(lat->sigma, sizeof(double), (size_t)V, opf);
  const double test = TestWriteReadLattice(simulation);
  fwrite(&test, sizeof(double), 1, opf);
  fwrite(&(info->epsilon_link_HMC), sizeof(double), 1, opf);
  fwrite(&(info->story_link), sizeof(double), 1, opf);
  fwrite(&(info->hybrid_accepted_link), sizeof(int), 1, opf);
  fwrite(&(info->hybrid_done_link), sizeof(int), 1, opf);
  fwrite(&(info->story_sigma), sizeof(double), 1, opf);
  fwrite(&(info->hybrid_accepted_sigma), sizeof(int), 1, opf);
  fwrite(&(info->hybrid_done_sigma), sizeof(int), 1, opf);
  fwrite(&(info->epsilon_pi_HMC), sizeof(double), 1, opf);
  fwrite(&(info->story_pi), sizeof(double), 1, opf);
  fwrite(&(info->hybrid_accepted_pi), sizeof(int), 1, opf);
  fwrite(&(info->hybrid_done_pi), sizeof(int), 1, opf);
  fwrite(&(info->measure), sizeof(int), 1, opf);
  fclose(opf);
}

// simulation.c:191-401.  The sigma mass of the read sigma comes from the
// drop-in's ComputeFullSigmaMass (the GPU kernel, bit-identical to
// sigma.c:75-94), so the check value is the one the writer stored.
void ReadBinaryLattice(struct Simulation *simulation, const char *terma_file, bool try_copy) {
  FILE *opf = fopen(terma_file, "r");
  // no file: the first run with these parameters; else try the copy
  if (opf == NULL) {
    opf = fopen(simulation->terma_file_copy, "r");
    if (opf == NULL) return;
  }
  int reading_errors = 0;
  auto failed = [&]() {
    reading_errors++;
    if (!try_copy) {
      fprintf(stderr, "Error when reading the file!\n");
      exit(1);
    }
  };
  const long V = gnx_ref_bound_volume();
  struct Lattice *lat = simulation->lattice;
  struct Info *info = &(simulation->info);
  if (fread(lat->dlink, sizeof(struct Link), (size_t)(V * kDim), opf) != (size_t)(V * kDim)) failed();
  if (fread(lat->sigma, sizeof(double), (size_t)V, opf) != (size_t)V) failed();
  ComputeFullSigmaMass(lat, simulation->fcor, lat->sigmamass);
  double test_old = 0.0;
  const double test_new = TestWriteReadLattice(simulation);
  if (fread(&test_old, sizeof(double), 1, opf) != 1 || test_old != test_new) failed();
  if (fread(&(info->epsilon_link_HMC), sizeof(double), 1, opf) != 1) failed();
  if (fread(&(info->story_link), sizeof(double), 1, opf) != 1) failed();
  if (fread(&(info->hybrid_accepted_link), sizeof(int), 1, opf) != 1) failed();
  if (fread(&(info->hybrid_done_link), sizeof(int), 1, opf) != 1) failed();
  if (fread(&(info->story_sigma), sizeof(double), 1, opf) != 1) failed();
  if (fread(&(info->hybrid_accepted_sigma), sizeof(int), 1, opf) != 1) failed();
  if (fread(&(info->hybrid_done_sigma), sizeof(int), 1, opf) != 1) failed();
  if (fread(&(info->epsilon_pi_HMC), sizeof(double), 1, opf) != 1) failed();
  if (fread(&(info->story_pi), sizeof(double), 1, opf) != 1) failed();
  if (fread(&(info->hybrid_accepted_pi), sizeof(int), 1, opf) != 1) failed();
  if (fread(&(info->hybrid_done_pi), sizeof(int), 1, opf) != 1) failed();
  // simulation.c:380-382
  info->hybrid_done_acceptance_link = ((double)info->hybrid_accepted_link) / ((double)info->hybrid_done_link);
  info->hybrid_done_acceptance_sigma = ((double)info->hybrid_accepted_sigma) / ((double)info->hybrid_done_sigma);
  info->hybrid_done_acceptance_pi = ((double)info->hybrid_accepted_pi) / ((double)info->hybrid_done_pi);
  if (fread(&(info->measure), sizeof(int), 1, opf) != 1) failed();
  fclose(opf);
  // any error: read the copy from the beginning, now without a fallback
  if (reading_errors > 0) ReadBinaryLattice(simulation, simulation->terma_file_copy, false);
}

}  // extern "C"
