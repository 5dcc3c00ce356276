/*
 * gnx_oracle.h -- CPU restatement of the fermion hot path of the 3D Z2
 * Gross-Neveu model (staggered Dirac operator, CG on M M^dagger, sigma force,
 * sigma HMC trajectory, Z2-noise condensate estimators, dSFMT host RNG).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * serial CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path (libgnx) never
 * links or calls it.
 *
 * Every routine keeps the reference's loop order and floating-point
 * expression shapes (C99 double _Complex arithmetic, pow(cabs(z),2) norms,
 * the same left-to-right accumulation) so that results are bit-identical to
 * the reference compiled with gcc -O3.  Lattice sizes are runtime values
 * (the reference fixes them at compile time, macro.h:21-24).
 *
 * Pinning (tests/test_oracle_golden.py against vectors the reference's own
 * sources produced at 8^3; tests/test_oracle_invariants.py at other shapes):
 *   geometry  : golden tables at 8^3; closed forms of coordinates, f-index,
 *               neighbours, eta and the time boundary on 4^2x6, 6^2x10, ...;
 *               pp(mm(x)) = x, parity flip.
 *   sigma mass, M, M^dag, r-=Mv : golden vectors (bit-exact) at 8^3; dense
 *               M^dag == transpose(M) bit for bit, M + M^dag = 2 sm.
 *   CG (both starts): golden solutions, iteration counts, early exits and
 *               counters at 8^3 (three tolerances); free-field plane-wave
 *               eigenvalues c^2 + sum sin^2 p and CG solution b / lambda.
 *   sigma force: golden force, X and Y at 8^3; central finite difference of
 *               the action (dense linear algebra) at 4^3.
 *   HMC trajectory, energy: golden Delta-H, accept, sigma after one and after
 *               all trajectories, CG counters at 8^3 (45 trajectories,
 *               three step sizes).
 *   condensate, CondensateSigmaPi: golden outputs after every trajectory.
 *   RNG       : golden dSFMT-132049 stream after seeding, Gauss1/Gauss2 values.
 *   link sector, live operator (the reference as written): golden
 *               AddForceFromFermionSectorRespectToLink and three LeapFrog2Link
 *               trajectories of the reference, called directly (no driver path
 *               reaches them), reproduced bit for bit.
 *   link sector, link-coupled operator (copy_dirac.c, commented out in the
 *               reference): tests/test_oracle_links.py invariants (U = 1
 *               reduction, M^dag = M^H, gauge covariance, dense CG, forces by
 *               finite differences, second-order integrator).
 */
#ifndef GNX_ORACLE_H
#define GNX_ORACLE_H

#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DIM 3
#define ORC_NFLAV 4
#define ORC_HALFNFLAV 2
#define ORC_MAGIC_LAMBDA 0.1931833275037836

/* ---------------- geometry (coordinate.c, lattice.c, fermion.c) ---------- */
typedef struct orc_geom {
    int L, LT, V, half;
    int *lex2f;      /* [V]    lex n -> f        (lattice.h:35 fermionsite)   */
    int *f2lex;      /* [V]    f -> lex n        (fermion.h:15 site)          */
    int *coords;     /* [V*3]  lex n -> (x0,x1,x2) (lattice.h:29 lattice_site) */
    int *fpp, *fmm;  /* [V*3]  f-order f-index neighbours (fermion.h:23-25)   */
    double *eta_pp, *eta_mm, *eta_pbc; /* [V*3] f-order (fermion.h:27-31)     */
} orc_geom;

int  orc_geom_init(orc_geom *g, int L, int LT);
void orc_geom_free(orc_geom *g);
/* The reference's own table set-up, by linear search (coordinate.c:253-267
 * called twice per site through InitializeFermionPP/MMArray, coordinate.c:
 * 322-374, from StartSite, lattice.c:32-33): O(V^2).  Fills fpp and fmm
 * [V][3] in f-order; timed by the benchmark's CPU baseline as the set-up the
 * reference pays before its first solve. */
void orc_setup_tables_by_search(const orc_geom *g, int *fpp, int *fmm);

/* ---------------- host RNG (random.c, dSFMT-132049) ---------------------- */
#define ORC_DSFMT_N 1269
typedef struct orc_rng {
    uint64_t st[(ORC_DSFMT_N + 1) * 2];
    int idx;
} orc_rng;

void   orc_rng_init(orc_rng *r, uint32_t seed);        /* Initrand, random.c:52-66 */
double orc_casuale(orc_rng *r);                         /* random.c:28-35 */
double orc_gauss1(orc_rng *r);                          /* random.c:88-102 */
void   orc_gauss2mv(orc_rng *r, double *a, double *b, double mean, double var); /* random.c:129-143 */
void   orc_fill_casuale(orc_rng *r, double *out, long n);
/* sigma field draw used by every workload: Gauss2MeanVariance pairs (x, x+V/2)
 * with mean 0 (sigma.c:29-30 pattern). */
void   orc_draw_sigma(orc_rng *r, int V, double var, double *sigma);
/* complex Z2 source: re then im, each Casuale()<=0.5 ? +1 : -1, per f-index */
void   orc_draw_z2_complex(orc_rng *r, int V, double *v /* [2V] */);
/* real Z2 source (condensate.c:52-60) */
void   orc_draw_z2_real(orc_rng *r, int V, double *v /* [2V] */);

/* ---------------- operators (sigma.c, dirac.c) --------------------------- */
/* complex vectors are double[2*V], (re,im) interleaved, f-order */
void orc_sigma_mass(const orc_geom *g, const double *sigma, double *sm);
void orc_dirac(const orc_geom *g, const double *sm, const double *in, double *out);
void orc_adj_dirac(const orc_geom *g, const double *sm, const double *in, double *out);
void orc_sub_dirac(const orc_geom *g, const double *sm, const double *in, double *out);
void orc_sub_adj_dirac(const orc_geom *g, const double *sm, const double *in, double *out);
double orc_norm(int V, const double *v);               /* fermion.h:67 */

/* ---------------- CG (conjugategradient.c) ------------------------------- */
typedef struct orc_cgstat {
    double iterations;  /* kit->iterations */
    int cg_call;        /* kit->CG_call    */
    int last_iters;     /* iterations of the last call (-1: early exit, no loop) */
} orc_cgstat;

/* work: 4 complex vectors (r, p, temp1, temp2) = 8V doubles */
bool orc_cg_null(const orc_geom *g, const double *sm, const double *b, double *x,
                 double tol, long max_iter, double *work, orc_cgstat *st);
bool orc_cg_imposed(const orc_geom *g, const double *sm, const double *b, double *x,
                    double tol, long max_iter, double *work, orc_cgstat *st);

/* ---------------- sigma HMC (moleculardynamics.c) ------------------------ */
typedef struct orc_kit {
    int V;
    double *work;           /* 8V: CG scratch r,p,temp1,temp2 */
    double *chi[2], *phi[2], *inv_phi[2], *inv_phi_obs[2]; /* 2V each */
    double *ddag;           /* 2V: Ddag_inverted_phi_obs */
    double *vec;            /* 2V: RandomVec */
    orc_cgstat st;
} orc_kit;

int  orc_kit_alloc(orc_kit *k, int V);
void orc_kit_free(orc_kit *k);

/* ComputeSigmaForce (moleculardynamics.c:631-647): force = fermion + sigma parts */
bool orc_sigma_force(const orc_geom *g, const double *sigma, const double *sm,
                     double beta_s, double tol, long max_iter, orc_kit *k, double *force);
/* force from given solutions X_f and Y_f = M^dag X_f (moleculardynamics.c:606-620, 576-581) */
void orc_sigma_force_from_solutions(const orc_geom *g, const double *sigma, double beta_s,
                                    double *const X[2], double *const Y[2], double *force);
double orc_energy(const orc_geom *g, const double *mom, const double *sigma, double beta_s,
                  double *const phi[2], double *const X[2]);   /* :299-319 */

typedef struct orc_traj {
    double dH, H0, H1;
    int accepted;
    int ok;             /* false if a CG failed */
} orc_traj;

/* One LeapFrog2MinimumNormSigmaVelocityFirst trajectory (:1234-1434),
 * with RNG draws in the reference order. sigma/sm are updated in place. */
void orc_lfmn2vf(const orc_geom *g, orc_rng *rng, double *sigma, double *sm, double *mom,
                 double *force, double beta_s, double eps, int steps, double tol_high,
                 double tol_medium, long max_iter, orc_kit *k, orc_traj *out);

/* ComputeCondensateDirac (condensate.c:114-244); x_inout aliases the
 * caller's inverted_phi_obs[0]; out[11] = tr1, tr1_abs, tr2, tr4, mu2, mu4,
 * chi, chi_abs, chip, 12_nosigma, 12_withsigma */
void orc_condensate(const orc_geom *g, orc_rng *rng, const double *sm, double tol, long max_iter,
                    orc_kit *k, double *x_inout, double *out);
/* CondensateSigmaPi (observables.c:323-369); out[8] = sigma, |sigma|, mu2, mu4,
 * chi, chip, chip_pbc1, chip_pbc2 */
void orc_condensate_sigma(const orc_geom *g, const double *sigma, double *out);

/* ---------------- link sector (config with links on) --------------------- */
/* The link-coupled operator of the commented-out copy_dirac.c:46-123; the
 * live operator (dirac.c) ignores the links.  U: lex order [V][3] complex
 * (lattice.h:15).  mom/force: f-order [V][3] (moleculardynamics.c:389).
 * With linked = 0 these follow the reference as written (golden-pinned);
 * the link-coupled operator itself is pinned by invariants. */
void orc_link_dirac(const orc_geom *g, const double *sm, const double *U, const double *in, double *out);
void orc_link_adj_dirac(const orc_geom *g, const double *sm, const double *U, const double *in, double *out);
void orc_link_sub_dirac(const orc_geom *g, const double *sm, const double *U, const double *in, double *out);
bool orc_link_cg_null(const orc_geom *g, const double *sm, const double *U, const double *b, double *x,
                      double tol, long max_iter, double *work, orc_cgstat *st);
bool orc_link_cg_imposed(const orc_geom *g, const double *sm, const double *U, const double *b, double *x,
                         double tol, long max_iter, double *work, orc_cgstat *st);
double orc_plaquette_sum(const orc_geom *g, const double *U);                         /* linklattice.c:62-80 */
void orc_gauge_force(const orc_geom *g, const double *U, double beta, double *force);  /* md.c:127-138 */
void orc_link_force_from_solutions(const orc_geom *g, const double *U, double *const X[2], double *const Y[2],
                                   double *force);                                     /* md.c:181-190 */
bool orc_link_force(const orc_geom *g, const double *sm, const double *U, double beta, int pure_gauge, int linked,
                    double tol, long max_iter, orc_kit *k, double *force);             /* md.c:164-222 */
double orc_link_energy(const orc_geom *g, const double *U, const double *mom, double beta, int pure_gauge,
                       double *const phi[2], double *const X[2]);                      /* md.c:268-294 */
void orc_leapfrog2link(const orc_geom *g, orc_rng *rng, const double *sm, double *U, double *mom, double *force,
                       double beta, double eps, int steps, double tol_high, long max_iter, int pure_gauge,
                       int linked, orc_kit *k, orc_traj *out);                         /* md.c:419-542 */

#ifdef __cplusplus
}
#endif
#endif
