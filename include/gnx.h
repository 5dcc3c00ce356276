/*
 * gnx.h -- C ABI of the MI355X-native fermion hot path of the 3D Z2
 * Gross-Neveu model (gfx950 HIP kernels).
 *
 * Conventions
 *   - Every vector/field argument is a DEVICE pointer (hipMalloc'd or any
 *     device allocation of the caller), never a host pointer.
 *   - Complex vectors are double[2*V] with (re, im) interleaved per site, in
 *     the reference's checkerboard "fermion lexicographic" order
 *     f = n/2 + parity*V/2 (coordinate.c:301-313), i.e. exactly the layout of
 *     the reference's `Fermion` (fermion.h:49-52).  Real fields (sigma, sigma
 *     mass, momenta, forces) are double[V] in the same f-order
 *     (lattice.h:16-20).
 *   - V = L*L*LT.  L and LT must be even and >= 4 (the checkerboard index
 *     is a bijection only for even LT, coordinate.c:301-313).
 *   - Work is enqueued on the context's stream (gnx_set_stream); calls that
 *     return host scalars (CG results, energies, dots) synchronise it.
 *   - The caller owns every buffer it passes; the context owns its scratch.
 *   - No function calls exit(); errors come back as gnx_status, with a
 *     message in gnx_last_error().
 */
#ifndef GNX_H
#define GNX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 4: gnx_cg_shape reports the one-reduction kernels in bit 1 of xcd_mapped
 * 5: launch shapes by a fixed rule unless autotuning is on (gnx_set_autotune,
 *    gnx_tune_runs); geometries outside the exact site decode are refused;
 *    bit 2 of gnx_cg_shape's xcd_mapped marks the deferred-update kernels
 * 6: bit 3 of xcd_mapped marks the fused kernels (one per iteration);
 *    gnx_multi_link_cg, gnx_multi_cg_prec; link-coupled lockstep chains
 * 7: bit 4 of xcd_mapped marks the one-launch solve of small lattices */
#define GNX_ABI_VERSION 9

typedef struct gnx_ctx gnx_ctx;
typedef struct gnx_multi gnx_multi;

typedef enum gnx_status {
    GNX_OK = 0,
    GNX_ERR_ARG = 1,      /* bad argument (null pointer, bad size, bad enum) */
    GNX_ERR_HIP = 2,      /* a HIP runtime call failed                        */
    GNX_ERR_NOCONV = 3,   /* CG exhausted max_iter (reference returns false)  */
    GNX_ERR_SIZE = 4,     /* request exceeds a compiled or allocated limit    */
    GNX_ERR_STATE = 5     /* call sequence error                              */
} gnx_status;

/* operator selector for gnx_dirac */
typedef enum gnx_op {
    GNX_OP_M = 0,         /* out = M in        DiracFermion         dirac.c:46-82   */
    GNX_OP_MDAG = 1,      /* out = M^dag in    AdjDiracFermion      dirac.c:88-123  */
    GNX_OP_MMDAG = 2,     /* out = M M^dag in  (the CG normal operator, conjugategradient.c:68,83) */
    GNX_OP_MDAGM = 3,     /* out = M^dag M in  (offered beside it; not used by the CG) */
    GNX_OP_SUB_M = 4,     /* out -= M in       SubtractDiracFermion dirac.c:213-247 */
    GNX_OP_SUB_MDAG = 5   /* out -= M^dag in   SubtractAdjDiracFermion dirac.c:253-287 */
} gnx_op;

/* CG start modes */
typedef enum gnx_cg_start {
    GNX_START_NULL = 0,     /* START_X0_NULL prologue: r=b-MM^dag x_in, early exit
                               keeps x_in; else restart from x=0
                               (conjugategradient.c:13-107) */
    GNX_START_IMPOSED = 1   /* x=0, r=b (conjugategradient.c:115-182) */
} gnx_cg_start;

/* CG arithmetic */
typedef enum gnx_prec {
    GNX_PREC_F64 = 0,       /* plain FP64 drop-in                                */
    GNX_PREC_F32 = 1,       /* FP32 inner iterations, FP64 reliable updates      */
    GNX_PREC_BF16 = 2,      /* BF16 storage / FP32 arithmetic inner iterations,
                               FP64 reliable updates                             */
    GNX_PREC_F16 = 3        /* FP16 storage (scaled per system at every reliable
                               update) / FP32 arithmetic inner iterations, FP64
                               reliable updates; the first slot of a solve is an
                               update, counted in reliable_updates           */
} gnx_prec;

typedef struct gnx_cg_result {
    int converged;        /* 1: ||r|| < tol (absolute, conjugategradient.c:87-89)  */
    int early_exit;       /* 1: returned before the loop (prologue or ||b|| < tol) */
    int iterations;       /* loop iterations (k+1 at exit); 0 on early exit        */
    int reliable_updates; /* mixed precision: FP64 residual recomputations         */
    double residual;      /* ||r|| of the recursively updated residual at exit     */
} gnx_cg_result;

int         gnx_abi_version(void);
const char *gnx_last_error(void);

/* Context: one lattice of L x L x LT sites on one HIP device. */
gnx_status gnx_create(int lspace, int ltime, int device, gnx_ctx **out);
gnx_status gnx_destroy(gnx_ctx *ctx);
gnx_status gnx_set_stream(gnx_ctx *ctx, void *hip_stream);   /* NULL: default stream */
gnx_status gnx_synchronize(gnx_ctx *ctx);
long       gnx_volume(const gnx_ctx *ctx);

/* A1. Geometry tables in the reference's f-order layout (fermion.h:12-33),
 * written to device buffers of V*3 ints/doubles and V ints.  The kernels do
 * not read tables; this export exists so the tables can be checked bit-exactly. */
gnx_status gnx_geometry_tables(gnx_ctx *ctx, int *fermionpp, int *fermionmm, double *eta_pp,
                               double *eta_mm, int *f2lex);

/* A2. ComputeFullSigmaMass (sigma.c:75-94): sm[x] = (1/8) sum of sigma over
 * the 8 dual sites below x, reference summation order. */
gnx_status gnx_sigma_mass(gnx_ctx *ctx, const double *sigma, double *sm);

/* A3/A4/A5 and the normal operators.  `in` and `out` must not alias.  M M^dag
 * and M^dag M are the compositions of the bit-exact single operators,
 * applied in one launch with the intermediate in LDS where the lattice is
 * made of 4 x 8 x 16 tiles (L divisible by 8, LT by 16) and V <= 300 000
 * (bit-identical to the two launches through a staged vector, used
 * otherwise). */
gnx_status gnx_dirac(gnx_ctx *ctx, gnx_op op, const double *sm, const double *in, double *out);
/* The same operator on nvec vectors (in, out: [nvec][V] complex, not
 * overlapping) with one sigma mass, one launch; op is GNX_OP_M, GNX_OP_MDAG,
 * GNX_OP_SUB_M or GNX_OP_SUB_MDAG, or GNX_OP_MMDAG / GNX_OP_MDAGM on a
 * lattice of 4 x 8 x 16 tiles.  Each vector's result is bit-identical to
 * its own gnx_dirac.  No counterpart in the reference (one vector per call). */
gnx_status gnx_dirac_batch(gnx_ctx *ctx, gnx_op op, const double *sm, const double *in, double *out, int nvec);

/* A6/A7. Solve (M M^dag) x = b to ||r||_2 < tol (absolute).
 * max_iter <= 0 selects the reference cap V+1 (macro.h:134).
 * Returns GNX_ERR_NOCONV when the cap is reached (x holds the last iterate). */
gnx_status gnx_cg(gnx_ctx *ctx, const double *sm, const double *b, double *x, double tol,
                  long max_iter, gnx_cg_start start, gnx_prec prec, gnx_cg_result *res);

/* Batched CG: nrhs right-hand sides b[k] = b + k*2V, solutions x + k*2V,
 * one fused persistent launch; res[nrhs].  Each system keeps its own alpha,
 * beta and exit iteration. */
gnx_status gnx_cg_batch(gnx_ctx *ctx, const double *sm, const double *b, double *x, int nrhs,
                        double tol, long max_iter, gnx_cg_start start, gnx_prec prec,
                        gnx_cg_result *res);

/* gnx_cg with prec GNX_PREC_F64 and one right-hand side, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The call allocates nothing: the solver's buffers live in one
 * static device array of 80 MiB per device that loads the library (held
 * until the process ends), enough for V <= about 950 000 sites (98^3);
 * larger lattices return GNX_ERR_SIZE.  The launch shape is the default
 * rule's, so results equal gnx_cg's without GNX_CG_SHAPE or autotuning (on
 * small lattices the one-launch solve, which needs no buffer of its own: its
 * result block sits in the context's reduction scalars, allocated by
 * gnx_create).  All calls on one device share the array: calls that overlap on one device
 * (other threads, other streams) race. */
gnx_status gnx_cg_noalloc(gnx_ctx *ctx, const double *sm, const double *b, double *x, double tol, long max_iter,
                          gnx_cg_start start, gnx_cg_result *res);

/* Uneven batches (extension 2: independent chains on lattices of uneven
 * size).  n systems, system k on an lspace[k]^2 x ltime[k] lattice with its
 * own sigma mass, solved by the FP64 CG of gnx_cg (both starts, the same exit
 * rules, the cap V_k + 1 when max_iter <= 0) in one launch per kernel: one
 * graph-replayed kernel pair per iteration over all systems, each block
 * reading its system's geometry from a descriptor table; a system that has
 * left the loop costs one state read per block per kernel.  Each system's
 * solution and iteration count are bit-identical to its gnx_cg solve with
 * one site per thread in the XCD slab mapping (GNX_CG_SHAPE=1) when every
 * system has at most 512 blocks of 256 sites, or every one more.  sm[k], b[k], x[k]
 * are device pointers (V_k doubles, 2 V_k doubles); res[n].  Returns
 * GNX_ERR_NOCONV when a system reaches its cap. */
gnx_status gnx_multi_create(int n, const int *lspace, const int *ltime, int device, gnx_multi **out);
gnx_status gnx_multi_destroy(gnx_multi *m);
gnx_status gnx_multi_set_stream(gnx_multi *m, void *stream);
gnx_status gnx_multi_cg(gnx_multi *m, const double *const *sm, const double *const *b, double *const *x,
                        double tol, long max_iter, gnx_cg_start start, gnx_cg_result *res);
/* The same on the link-coupled operator M_U M_U^dag (as gnx_link_cg_batch,
 * copy_dirac.c:46-123), system k with its own links U[k] (device, 3 V_k
 * complex, f-order, three per site); bit-identical to its gnx_link_cg_batch
 * solve under the same condition. */
gnx_status gnx_multi_link_cg(gnx_multi *m, const double *const *sm, const double *const *U, const double *const *b,
                             double *const *x, double tol, long max_iter, gnx_cg_start start, gnx_cg_result *res);
/* gnx_multi_cg in another arithmetic: GNX_PREC_F64 is gnx_multi_cg; F32,
 * BF16 and F16 keep r, p, t in that storage with FP64 reliable updates, as
 * gnx_cg's mixed precision (the default reliable-update factors), in one
 * launch per kernel over all systems.  Each system's solution and counters
 * are bit-identical to its gnx_cg solve with GNX_MIX_SHAPE=11 (loads-first
 * kernels, one site per thread, XCD slab mapping). */
gnx_status gnx_multi_cg_prec(gnx_multi *m, const double *const *sm, const double *const *b, double *const *x,
                             double tol, long max_iter, gnx_cg_start start, gnx_prec prec, gnx_cg_result *res);

/* A12, batched.  Stochastic trace term of the condensate for nsrc noise
 * vectors at once (condensate.c:127-145 is this for one vector at a time):
 * noise[k] = noise + k*2V, one batched solve M M^dag X_k = v_k (imposed
 * start, precision prec, absolute tol; X receives the solutions), then
 *   per_source[k] = sum_x Re(conj(v_k[x]) (M^dag X_k)[x]) / V = Re(v_k^dag M^-1 v_k) / V
 * on the device.  per_source is HOST memory [nsrc]; res[nsrc].  nsrc <= 64. */
gnx_status gnx_condensate_batch(gnx_ctx *ctx, const double *sm, const double *noise, int nsrc, double tol,
                                long max_iter, gnx_prec prec, double *X, double *per_source,
                                gnx_cg_result *res);

/* Mixed precision: residual reduction factor since the last FP64 reliable
 * update that triggers the next one.  0 (the default) selects it per
 * precision: 0.05 for FP32 and FP16, 0.2 for BF16. */
gnx_status gnx_set_reliable_delta(gnx_ctx *ctx, double delta);

/* Launch shapes.  A solve's iterates depend on its launch shape (the
 * one-reduction and two-reduction forms follow different recurrences, and
 * sites per thread set the partition of every global sum), so by default the
 * shape is a fixed function of the lattice size and batch size: the same
 * solve gives the same bits in every process, run and restart.  With
 * autotuning on (gnx_set_autotune(ctx, 1), or GNX_AUTOTUNE=1 in the
 * environment when the context's first solve runs) each new (batch size,
 * operator or storage format) is timed over every shape once and the fastest
 * kept for the context's lifetime: faster where the rule is not the
 * optimum, but the choice may differ between processes. */
gnx_status gnx_set_autotune(gnx_ctx *ctx, int on);
/* how many keys the timing tuner has chosen so far (FP64 and mixed solvers);
 * each key is timed once per context unless it was evicted from the
 * 16-entry cache of choices (least recently used first) */
gnx_status gnx_tune_runs(gnx_ctx *ctx, long *fp64, long *mixed);
/* Iteration slots the context's graph-replayed solves (FP64 and mixed) have
 * launched and the times they polled their state, since it was created.
 * A replay runs a graph of 32 iterations or a tail of 8, 16 or 24, and a
 * slot after a system stopped costs its kernels' launch; each solve first
 * launches about as many slots as the last solve of the same kind (graph,
 * batch size, tolerance, start) needed, as whole graphs and one tail, and
 * polls once. */
gnx_status gnx_replay_stats(gnx_ctx *ctx, long *slots, long *polls);
/* Launch shape of the FP64 CG on the live operator (M M^dag) for batches of
 * nrhs systems on this context (sites per thread; xcd_mapped bit 0:
 * XCD-slab block mapping, bit 1: the one-reduction (Chronopoulos-Gear)
 * kernels, bit 2: their form with deferred solution updates, the same
 * iterates bit for bit, bit 3: the fused form, one kernel per iteration with
 * each block's one-site halo recomputed into LDS); GNX_ERR_STATE if no solve
 * of that batch size has run yet.  GNX_CG_SHAPE=<n> in the environment pins
 * the shape: n = 0..5 the two-reduction kernels with 1, 1, 2, 2, 4, 4 sites
 * per thread, n = 6..11 the one-reduction kernels likewise, n = 12..17 the
 * deferred form likewise; odd n use the XCD slab mapping; n = 18..21 the
 * fused form on 4 x 8 x 16 and 8 x 8 x 8 tiles, without and with deferral
 * (where the tiles divide the lattice; elsewhere the fixed rule's shape).
 * The fixed rule chooses the fused form (n = 18) on the live operator where
 * the batch's 4 x 8 x 16 tiles number 160 to 256 (one per CU: 48^3 single,
 * 32^3 x 3 and x 4), where it saves a kernel boundary per iteration; with
 * more tiles its halo work costs more than the boundary.  The link-coupled
 * operator's shape is chosen the same way (without the fused form) but not
 * reported.  On lattices of at most 1728 sites (12^3), with an iteration cap
 * of at most 100 000, the fixed rule runs each solve as ONE launch instead:
 * one 512-thread workgroup per system iterates to the exit with the
 * direction and M^dag of it in LDS (the two-reduction recurrences, every
 * global sum a workgroup sum); it reports bit 4 (16) and its sites per
 * thread (1, 2 or 4).  GNX_CG_SHAPE or autotuning turns it off (they choose
 * among the graph forms). */
gnx_status gnx_cg_shape(gnx_ctx *ctx, int nrhs, int *sites_per_thread, int *xcd_mapped);
/* The same for the mixed-precision solver (prec GNX_PREC_F32 or GNX_PREC_BF16);
 * pair = 1: two systems share each block's site arithmetic; loads_first = 1:
 * kernels that issue every load before their first wait (more registers,
 * less latency).  GNX_MIX_SHAPE=<n> pins one: n = 0..5 one system per block
 * with 1, 1, 2, 2, 4, 4 sites per thread, n = 6..9 two systems per block with
 * 1, 1, 2, 2, n = 10..15 as 0..5 with loads-first kernels; odd n use the XCD
 * slab mapping; n = 16 loads-first kernels on groups of four consecutive
 * sites (16-byte loads; BF16 and FP16 storage with LT a multiple of 8 only, a
 * pin that does not fit is ignored), reported as xcd_mapped = 3 (bit 1: the
 * groups). */
gnx_status gnx_cg_mixed_shape(gnx_ctx *ctx, int nrhs, gnx_prec prec, int *sites_per_thread, int *xcd_mapped,
                              int *pair, int *loads_first);

/* A9. Sigma force (moleculardynamics.c:590-647):
 *   X_f = CG(phi_f) with START_X0_NULL from the incoming X_f, Y_f = M^dag X_f,
 *   F = sum_f sum_corners 0.125 Re(conj X_f Y_f) - (nflav*beta_s/2) sigma.
 * phi, X, Y are arrays of two device vectors.  X is in/out. */
gnx_status gnx_sigma_force(gnx_ctx *ctx, const double *sigma, const double *sm, double beta_s,
                           int nflav, double tol, long max_iter, const double *const phi[2],
                           double *const X[2], double *const Y[2], double *force,
                           gnx_cg_result res[2]);

/* Force from given solutions only (the kernel after the solves). */
gnx_status gnx_sigma_force_from_solutions(gnx_ctx *ctx, const double *sigma, double beta_s, int nflav,
                                          const double *X0, const double *Y0, const double *X1,
                                          const double *Y1, double *force);

/* The two force terms separately: accumulate != 0 adds the fermion part onto
 * the incoming force (AddForceFromFermionSectorRespectToSigma,
 * moleculardynamics.c:606-620); sigma_term != 0 then subtracts
 * (nflav*beta_s/2) sigma (AddForceFromSigmaSector, :576-581). */
gnx_status gnx_sigma_force_terms(gnx_ctx *ctx, const double *sigma, double beta_s, int nflav,
                                 const double *X0, const double *Y0, const double *X1, const double *Y1,
                                 double *force, int accumulate, int sigma_term);

/* A10. HMC pieces. */
gnx_status gnx_evolve_momenta(gnx_ctx *ctx, double *mom, const double *force, double eps);
gnx_status gnx_evolve_sigma(gnx_ctx *ctx, double *sigma, double *sm, const double *mom, double eps);
/* H = sum p^2/2 + sum_f sum_x 0.5 Re(conj phi_f X_f) + (nflav beta_s/4) sum sigma^2
 * (moleculardynamics.c:299-319), deterministic FP64 tree reductions. */
gnx_status gnx_energy(gnx_ctx *ctx, const double *mom, const double *sigma, double beta_s, int nflav,
                      const double *const phi[2], const double *const X[2], double *H);

/* A12 helpers: deterministic reductions (host result). */
gnx_status gnx_dot(gnx_ctx *ctx, const double *a, const double *b, double *re, double *im);
gnx_status gnx_norm2(gnx_ctx *ctx, const double *a, double *out);
/* y = x * field (complex vector times real field, condensate.c:221,225) */
gnx_status gnx_scale_by_field(gnx_ctx *ctx, const double *field, const double *x, double *y);


/* ------------------------------------------------------------------------
 * A10/A12/A14. Device-resident sigma-HMC state of one lattice.
 * The host keeps the random-number generator (random.c) and passes each
 * trajectory's draws in the reference's order (momenta V x Gauss1, then chi_0,
 * then chi_1, each V x Gauss1; moleculardynamics.c:1241-1250); the
 * Metropolis uniform is drawn by the caller after the trajectory and the
 * decision handed back with gnx_hmc_accept (moleculardynamics.c:1385-1409).
 * ------------------------------------------------------------------------ */
typedef struct gnx_hmc gnx_hmc;

typedef struct gnx_traj_result {
    double dH, H0, H1;      /* final - starting energy, moleculardynamics.c:1379-1385 */
    int ok;                 /* 0: a CG exhausted max_iter; sigma restored (:1262-1283) */
    int cg_solves;          /* solves that converged inside the loop */
    long cg_iterations;     /* sum of their iteration counts (kit->iterations) */
} gnx_traj_result;

gnx_status gnx_hmc_create(gnx_ctx *ctx, double beta_s, int nflav, gnx_hmc **out);
gnx_status gnx_hmc_destroy(gnx_hmc *h);
/* host <-> device copies of sigma (V doubles, f-order); set recomputes sm */
gnx_status gnx_hmc_set_sigma(gnx_hmc *h, const double *sigma_host);
gnx_status gnx_hmc_get_sigma(gnx_hmc *h, double *sigma_host, double *sm_host);
/* device pointers of the state: sigma, sm, momenta, force [V]; X, Y, phi [2][2V] */
gnx_status gnx_hmc_device_fields(gnx_hmc *h, double **sigma, double **sm, double **mom, double **force,
                                 double **X, double **Y, double **phi);
/* One LeapFrog2MinimumNormSigmaVelocityFirst trajectory (moleculardynamics.c:1234-1379).
 * mom_host, chi0_host, chi1_host: V real draws each.  On return sigma holds
 * the proposal; res->dH = H1 - H0. */
gnx_status gnx_hmc_trajectory(gnx_hmc *h, const double *mom_host, const double *chi0_host,
                              const double *chi1_host, double eps, int steps, double tol_high,
                              double tol_medium, long max_iter, gnx_traj_result *res);
/* accepted = 0 restores sigma and its mass (moleculardynamics.c:1407-1408) */
gnx_status gnx_hmc_accept(gnx_hmc *h, int accepted);
/* Arithmetic of the trajectory's MEDIUM-tolerance force solves
 * (moleculardynamics.c:1279,1291,1308,1321,1335,1348): GNX_PREC_F64 (the
 * default, the reference's) or mixed precision (FP32, BF16 or FP16 storage
 * with FP64 reliable updates, converged on the FP64 true residual).  The
 * HIGH-tolerance solves (:1263, :1362) -- the energies and the last force --
 * stay FP64.  No counterpart in the reference, which is FP64 throughout. */
gnx_status gnx_hmc_set_force_precision(gnx_hmc *h, gnx_prec prec);
/* ComputeCondensateDirac (condensate.c:114-244) with the caller's real Z2
 * draws z2_host[6*V] (4 trace vectors, then the two susceptibility vectors);
 * the first solve starts from Y_0 of the last force, as the driver passes
 * inverted_phi_obs[0] (observables.c:247).  out[11] = tr1, tr1_abs, tr2, tr4,
 * mu2, mu4, chi, chi_abs, chip, 12_nosigma, 12_withsigma. */
gnx_status gnx_hmc_condensate(gnx_hmc *h, const double *z2_host, double tol, long max_iter, double *out);
/* CondensateSigmaPi (observables.c:323-369): out[8] = sigma, |sigma|, mu2,
 * mu4, chi, chip, chip_pbc1, chip_pbc2 */
gnx_status gnx_hmc_condensate_sigma(gnx_hmc *h, double *out);
/* accumulated CG statistics (kit->iterations, kit->CG_call) */
gnx_status gnx_hmc_cg_stats(gnx_hmc *h, double *iterations, int *calls);
/* n independent chains (configuration 5: lattices of uneven size) in
 * lockstep: gnx_hmc_trajectory for every chain, with each solve point of the
 * integrator one launch per CG kernel over all live chains.  m: an uneven
 * solver (gnx_multi_create) holding 2n systems, systems 2k and 2k+1 on chain
 * k's lattice (its two flavours).  Per-chain arguments and results as
 * gnx_hmc_trajectory; a chain whose solve fails gets res[k].ok = 0 and its
 * sigma restored while the others go on (status GNX_ERR_NOCONV).  No
 * counterpart in the reference, which runs one chain per process. */
gnx_status gnx_hmc_ensemble_trajectory(gnx_multi *m, int n, gnx_hmc *const *hs, const double *const *mom_host,
                                       const double *const *chi0_host, const double *const *chi1_host,
                                       double eps, int steps, double tol_high, double tol_medium,
                                       long max_iter, gnx_traj_result *res);

/* ------------------------------------------------------------------------
 * A11: the U(1) link sector.  The reference's live operator (dirac.c)
 * ignores the links; the link-coupled operator is the commented-out
 * copy_dirac.c:46-123:
 *   M_U v(x) = sm v(x) + sum_d [0.5 eta_pp(x,d) U_d(x) v(x+d) - 0.5 eta_mm(x,d) conj(U_d(x-d)) v(x-d)]
 * and M_U^dag the same with the hop terms negated.  On the device, links are
 * f-order with three per site, uf[2*(3 f + d)] = (Re, Im) U_d(f); the
 * reference's dlink is lex order [V][3] (lattice.h:15), converted by
 * gnx_links_from_lex / gnx_links_to_lex.  Momenta and forces are f-order
 * [V][3] (moleculardynamics.c:389).  No reference path reaches this code,
 * so parity is pinned by invariants (gauge covariance, U = 1 reduction,
 * finite differences), not by reference outputs.
 * ------------------------------------------------------------------------ */
gnx_status gnx_links_from_lex(gnx_ctx *ctx, const double *ulex, double *uf);
gnx_status gnx_links_to_lex(gnx_ctx *ctx, const double *uf, double *ulex);
/* op: GNX_OP_M (copy_dirac.c:46-82), GNX_OP_MDAG (:88-123), GNX_OP_SUB_M (out -= M_U in) */
gnx_status gnx_link_dirac(gnx_ctx *ctx, gnx_op op, const double *sm, const double *uf, const double *in,
                          double *out);
/* FP64 CG on M_U M_U^dag, conjugategradient.c semantics (null or imposed start) */
gnx_status gnx_link_cg_batch(gnx_ctx *ctx, const double *sm, const double *uf, const double *b, double *x,
                             int nrhs, double tol, long max_iter, gnx_cg_start start, gnx_cg_result *res);
/* AddForceFromPureGaugeSector (moleculardynamics.c:127-138, linklattice.c:146-172): overwrites force */
gnx_status gnx_gauge_force(gnx_ctx *ctx, const double *uf, double beta, double *force);
/* AddForceFromFermionSectorRespectToLink after its solves (moleculardynamics.c:181-190):
 * force (+)= sum_f Re(...) with X_f, Y_f = M^dag X_f; accumulate = 0 starts from zero */
gnx_status gnx_link_force_from_solutions(gnx_ctx *ctx, const double *uf, const double *X0, const double *Y0,
                                         const double *X1, const double *Y1, double *force, int accumulate);
/* sum_x sum_{mu > nu} Re Pi_{mu nu}(x) (TotalRealTracePlaquette, linklattice.c:62-80) */
gnx_status gnx_plaquette_sum(gnx_ctx *ctx, const double *uf, double *out);

/* Device-resident LeapFrog2Link (moleculardynamics.c:419-542) at fixed sigma
 * mass.  pure_gauge adds -beta sum Re Pi to the action (PURE_GAUGE,
 * macro.h:47); linked = 0 solves with the live operator as the reference
 * does, linked = 1 with M_U.  The caller draws (3V/2 Gauss2MeanVariance
 * momentum pairs, then chi_0, chi_1: V x Gauss1 each) and decides the
 * Metropolis test with gnx_link_hmc_accept. */
typedef struct gnx_link_hmc gnx_link_hmc;
gnx_status gnx_link_hmc_create(gnx_ctx *ctx, double beta, int pure_gauge, int linked, gnx_link_hmc **out);
gnx_status gnx_link_hmc_destroy(gnx_link_hmc *h);
/* host links in lex order [V][3] complex (NULL: keep) and sigma mass [V] (NULL: keep) */
gnx_status gnx_link_hmc_set_fields(gnx_link_hmc *h, const double *ulex_host, const double *sm_host);
/* starting guesses X_0, X_1 (kit->inverted_phi) for the next null-start solves */
gnx_status gnx_link_hmc_set_solutions(gnx_link_hmc *h, const double *X0_host, const double *X1_host);
gnx_status gnx_link_hmc_get_links(gnx_link_hmc *h, double *ulex_host);
/* device pointers: links (f-order), sm, momenta [3V], force [3V], X, Y, phi [2][2V] */
gnx_status gnx_link_hmc_device_fields(gnx_link_hmc *h, double **uf, double **sm, double **mom, double **force,
                                      double **X, double **Y, double **phi);
/* ComputeForce (moleculardynamics.c:203-222) on the current fields; *ok = 0 if a solve failed */
gnx_status gnx_link_hmc_compute_force(gnx_link_hmc *h, double tol, long max_iter, int *ok);
gnx_status gnx_link_hmc_trajectory(gnx_link_hmc *h, const double *mom_host, const double *chi0_host,
                                   const double *chi1_host, double eps, int steps, double tol, long max_iter,
                                   gnx_traj_result *res);
/* accepted = 0 restores the links of the trajectory's start (:518-520) */
gnx_status gnx_link_hmc_accept(gnx_link_hmc *h, int accepted);
/* ComputeFullEnergyWithPhiWithoutSigma (moleculardynamics.c:268-294) on the current fields */
gnx_status gnx_link_hmc_energy(gnx_link_hmc *h, double *H);
gnx_status gnx_link_hmc_cg_stats(gnx_link_hmc *h, double *iterations, int *calls);
/* n chains' LeapFrog2Link trajectories in lockstep, each solve point one
 * launch per CG kernel over all live chains (as gnx_hmc_ensemble_trajectory):
 * the live operator for chains created with linked = 0, the link-coupled one
 * (gnx_multi_link_cg, each chain's links) for linked = 1; all chains of a
 * call alike.  No counterpart in the reference. */
gnx_status gnx_link_hmc_ensemble_trajectory(gnx_multi *m, int n, gnx_link_hmc *const *hs,
                                            const double *const *mom_host, const double *const *chi0_host,
                                            const double *const *chi1_host, double eps, int steps, double tol,
                                            long max_iter, gnx_traj_result *res);

#ifdef __cplusplus
}
#endif
#endif /* GNX_H */
