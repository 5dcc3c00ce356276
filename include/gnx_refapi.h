/*
 * gnx_refapi.h -- drop-in replacement of the reference's hot-path functions
 * (alefrajoe/3D-Gross-Neveu-model), implemented on the MI355X kernels of
 * gnx.h.  Every function keeps the reference's name, C signature, argument
 * meaning, in-place updates of the caller's structs and return convention,
 * so the reference's own driver (test.c, observables.c, simulation.c) links
 * against libgnx_ref.so instead of compiling the replaced function bodies
 * (the integration steps are in INTEGRATION.md).
 *
 * The struct declarations below are layout mirrors of the reference's types
 * (same member order and C types) for the build the reference ships:
 * DIM 3, NFLAV 4, SIGMA_INTERACTION and HYBRID_MC on, PI_INTERACTION and
 * METROPOLIS_LINK off (macro.h:23-25,47-50,113-114).  A driver compiled
 * against the reference's headers passes pointers to its own structs; only
 * the layout has to agree.  Vectors are host memory, as in the reference;
 * the library stages them to the GPU per call, except inside
 * LeapFrog2MinimumNormSigmaVelocityFirst, which keeps the whole trajectory
 * device-resident.
 *
 * Host random numbers stay with the caller: GnxRefInit binds the driver's
 * Gauss1 and Casuale (random.c:28-35, 88-102) so that the replaced
 * trajectory and condensate draw in the reference's order.
 */
#ifndef GNX_REFAPI_H
#define GNX_REFAPI_H

#include <stdbool.h>

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> gnx_ref_complex;
extern "C" {
#else
#include <complex.h>
typedef double _Complex gnx_ref_complex;
#endif

/* ---- layout mirrors of the reference's types --------------------------- */
/* link.h:7-10 */
struct Link { gnx_ref_complex link; };

/* lattice.h:12-56 */
struct Lattice {
    struct Link *dlink;
    double *sigma;
    double *sigmamass;
    double *pi;
    double *epsilon_pi;
    double *pimass;
    int *site;
    int *lattice_site;
    int *pp;
    int *mm;
    int *fermionsite;
    int *fermionpp;
    int *fermionmm;
    double *eta_pp;
    double *eta_mm;
    double *eta_pbc;
    double distance_link;
    double distance_sigma;
    double distance_pi;
    int done_link;
    int done_sigma;
    int done_pi;
};

/* fermion.h:12-33 */
struct Fermioncoor {
    int *site;
    int *pp;
    int *mm;
    int *fermionsite;
    int *fermionpp;
    int *fermionmm;
    double *eta_pp;
    double *eta_mm;
    double *eta_pbc;
};

/* fermion.h:49-52, 77-80, 93-112 */
struct Fermion { gnx_ref_complex *fermion; };
struct ManyFermions { struct Fermion *flav; };
struct Fermionkit {
    struct ManyFermions r_k;
    struct ManyFermions p_k;
    struct ManyFermions chi;
    struct ManyFermions phi;
    struct ManyFermions inverted_phi;
    struct ManyFermions inverted_phi_obs;
    struct Fermion Ddag_inverted_phi_obs;
    struct Fermion temp1;
    struct Fermion temp2;
    struct Fermion tempbminusx0;
    struct Fermion bminusax0;
    double iterations;
    int CG_call;
};

/* info.h:7-66 with HYBRID_MC on and METROPOLIS_LINK off */
struct Info {
    double beta;
    double beta_scalar;
    double mass;
    int sosia;
    int measure;
    double epsilon_link_HMC;
    double distance_link_HMC;
    int steps_link_HMC;
    double epsilon_sigma_HMC;
    double distance_sigma_HMC;
    int steps_sigma_HMC;
    double epsilon_pi_HMC;
    double distance_pi_HMC;
    int steps_pi_HMC;
    int hybrid_accepted_link;
    int hybrid_done_link;
    double hybrid_done_acceptance_link;
    int hybrid_accepted_sigma;
    int hybrid_done_sigma;
    double hybrid_done_acceptance_sigma;
    int local_accepted_sigma;
    int local_done_sigma;
    double local_acceptance_sigma;
    int hybrid_accepted_pi;
    int hybrid_done_pi;
    double hybrid_done_acceptance_pi;
    double delta_H_link;
    double delta_H_sigma;
    double delta_H_pi;
    int story_link;
    int story_sigma;
    int story_pi;
};

/* moleculardynamics.h:17-30, 62-88, condensate.h:11-15 */
struct Momenta { double *momenta; };
struct Force { double *force; };
struct Momentasigma { double *momenta; };
struct Forcesigma { double *force; };
struct RandomVec { struct Fermion *vec; };
struct Momentapi;   /* pi.h: not used by the drop-in (PI_INTERACTION off) */
struct Forcepi;

/* simulation.h:14-31 */
struct Simulation {
    struct Lattice *lattice;
    struct Info info;
    struct Fermioncoor *fcor;
    struct Fermionkit *kit;
    struct Force *force;
    struct Momenta *momenta;
    struct Momentasigma *momentasigma;
    struct Forcesigma *forcesigma;
    struct Momentapi *momentapi;
    struct Forcepi *forcepi;
    char observables_file[200];
    char acceptance_file[200];
    char terma_file[200];
    char terma_file_copy[200];
};

/* ---- set-up (no counterpart in the reference: it fixes L at compile time) -- */
/* Bind the lattice size (macro.h:21-24 LSPACE/LTIME), the HIP device and the
 * driver's host RNG.  Returns 0 on success. */
int GnxRefInit(int lspace, int ltime, int device, double (*gauss1)(void), double (*casuale)(void));
void GnxRefFinalize(void);
/* HIGH/MEDIUM_PRECISION_INVERSION of the driver's macro.h (macro.h:135-136);
 * the trajectory uses them as the reference's integrator does
 * (defaults 1e-9 / 1e-3, the shipped values). */
void GnxRefSetTolerances(double high, double medium);
/* The driver's Gauss2MeanVariance (random.c:129-143), used by
 * LeapFrog2Link's momentum draw (moleculardynamics.c:38-46). */
void GnxRefBindGauss2(void (*gauss2mv)(double *ris1, double *ris2, double mean, double variance));
/* PURE_GAUGE of the driver's macro.h:47 (off in the shipped build): adds the
 * pure-gauge force and action to ComputeForce and LeapFrog2Link. */
void GnxRefSetPureGauge(int on);

/* ---- replaced functions, reference signatures ------------------------- */
/* dirac.h:10-15 (dirac.c:46-82, 88-123, 213-247, 253-287) */
void DiracFermion(const struct Lattice *lat, const struct Fermion *fermion, struct Fermion *target,
                  const struct Info *info, const struct Fermioncoor *fermioncoor);
void AdjDiracFermion(const struct Lattice *lat, const struct Fermion *fermion, struct Fermion *target,
                     const struct Info *info, const struct Fermioncoor *fermioncoor);
void SubtractDiracFermion(const struct Lattice *lat, const struct Fermion *fermion, struct Fermion *target,
                          const struct Info *info, const struct Fermioncoor *fermioncoor);
void SubtractAdjDiracFermion(const struct Lattice *lat, const struct Fermion *fermion, struct Fermion *target,
                             const struct Info *info, const struct Fermioncoor *fermioncoor);

/* conjugategradient.h:9-10 (conjugategradient.c:13-107, 115-182) */
bool ConjugateGradientFastDiracAdjFastDiracToTheMinusOne(const struct Lattice *lat, const struct Fermion *fermion,
                                                         struct Fermion *target_inverted, const struct Info *info,
                                                         const struct Fermioncoor *fermioncoor,
                                                         double inversion_precision, struct Fermionkit *kit, int f);
bool StartX0ImposedConjugateGradientFastDiracAdjFastDiracToTheMinusOne(
    const struct Lattice *lat, const struct Fermion *fermion, struct Fermion *target_inverted,
    const struct Info *info, const struct Fermioncoor *fermioncoor, double inversion_precision,
    struct Fermionkit *kit, int f);

/* sigma.h:11 (sigma.c:75-94) */
void ComputeFullSigmaMass(struct Lattice *lat, const struct Fermioncoor *fcor, double *sigmamass);

/* moleculardynamics.h:104-105 (moleculardynamics.c:590-622, 631-647) */
bool AddForceFromFermionSectorRespectToSigma(const struct Lattice *lat, const struct Info *info,
                                             const struct Fermioncoor *fcor, struct Forcesigma *force,
                                             struct Fermionkit *kit, double inversion_precision);
bool ComputeSigmaForce(const struct Lattice *lat, const struct Info *info, const struct Fermioncoor *fcor,
                       struct Forcesigma *force, struct Fermionkit *kit, double inversion_precision);

/* moleculardynamics.h:48-49, 79 (moleculardynamics.c:164-222, 419-542).  As
 * in the reference, the solves use the live operator (dirac.c ignores the
 * links) and the force differentiates the link-coupled one. */
bool AddForceFromFermionSectorRespectToLink(const struct Lattice *lat, const struct Info *info,
                                            const struct Fermioncoor *fcor, struct Force *force,
                                            struct Fermionkit *kit);
bool ComputeForce(const struct Lattice *lattice, const struct Info *info, struct Force *force,
                  const struct Fermioncoor *fermioncoor, struct Fermionkit *kit);
bool LeapFrog2Link(struct Lattice *lattice, struct Info *info, struct Force *force, struct Momenta *momenta,
                   const struct Fermioncoor *fermioncoor, struct Fermionkit *kit, double epsilon, const int step,
                   const char *acceptance_file);

/* moleculardynamics.h:130 (moleculardynamics.c:1234-1434) */
bool LeapFrog2MinimumNormSigmaVelocityFirst(struct Lattice *lattice, struct Info *info, struct Forcesigma *force,
                                            struct Momentasigma *momenta, const struct Fermioncoor *fermioncoor,
                                            struct Fermionkit *kit, double epsilon, const int step,
                                            const char *acceptance_file);

/* condensate.h:27 (condensate.c:114-244) */
void ComputeCondensateDirac(const struct Lattice *lat, double *condensate_tr1, double *condensate_tr1_abs,
                            double *condensate_tr2, double *condensate_tr4, double *condensate_mu2,
                            double *condensate_mu4, double *condensate_chi, double *condensate_chi_abs,
                            double *condensate_chip, double *condensate12_nosigma, double *condensate12_withsigma,
                            struct RandomVec *vec, struct Fermion *target_inverted, const struct Info *info,
                            const struct Fermioncoor *fermioncoor, double inversion_precision,
                            struct Fermionkit *kit, int f);

/* simulation.h:39-41 (simulation.c:104-401): the terma checkpoint, byte for
 * byte the reference's file (SIGMA_INTERACTION on, PI_INTERACTION off, as
 * shipped).  WriteBinaryLattice renames an existing terma_file to
 * terma_file_copy first; ReadBinaryLattice falls back to the copy on a
 * missing or damaged file and, like the reference, prints an error and exits
 * with status 1 when the file it must read (try_copy false) is damaged, or
 * when the file cannot be opened for writing.  ReadBinaryLattice recomputes
 * the sigma mass with ComputeFullSigmaMass above. */
double TestWriteReadLattice(const struct Simulation *simulation);
void WriteBinaryLattice(const struct Simulation *simulation);
void ReadBinaryLattice(struct Simulation *simulation, const char *terma_file, bool try_copy);

#ifdef __cplusplus
}
#endif
#endif /* GNX_REFAPI_H */
