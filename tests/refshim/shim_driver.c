/*
 * shim_driver.c -- drives the drop-in library (include/gnx_refapi.h) the way
 * the reference's own driver drives the functions it replaces, through the
 * reference's struct layouts, and dumps the same named records as
 * oracle/ref_harness.c so tests/test_refshim.py can compare them with
 * the golden vectors of the reference itself.  The host RNG is the oracle's
 * dSFMT restatement (pinned to the reference's stream).
 *
 * TEST INFRASTRUCTURE ONLY.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/gnx_refapi.h"
#include "../../oracle/gnx_oracle.h"

#define L 8
#define V (L * L * L)
#define HALF (V / 2)

static FILE *g_out;
static orc_rng R;
static double drv_gauss1(void) { return orc_gauss1(&R); }
static double drv_casuale(void) { return orc_casuale(&R); }
static void drv_gauss2mv(double *a, double *b, double mean, double var) { orc_gauss2mv(&R, a, b, mean, var); }

static void rec(const char *name, char kind, const void *data, int64_t count)
{
    char hdr[40];
    memset(hdr, 0, sizeof(hdr));
    strncpy(hdr, name, 31);
    hdr[32] = kind;
    fwrite(hdr, 1, 40, g_out);
    fwrite(&count, sizeof(int64_t), 1, g_out);
    fwrite(data, kind == 'd' ? sizeof(double) : kind == 'i' ? sizeof(int) : 1, (size_t)count, g_out);
}

/* a whole file as a byte record ('b') */
static void rec_file(const char *name, const char *path)
{
    static unsigned char buf[1 << 20];
    FILE *f = fopen(path, "rb");
    size_t n = f ? fread(buf, 1, sizeof(buf), f) : 0;
    if (f) fclose(f);
    rec(name, 'b', buf, (int64_t)n);
}

/* the checkpoint fields ReadBinaryLattice restores, as doubles */
static void rec_terma_state(const char *prefix, const struct Simulation *s)
{
    char name[64];
    const struct Info *in = &(s->info);
    double st[15] = {in->epsilon_link_HMC, in->story_link, in->hybrid_accepted_link, in->hybrid_done_link,
                     in->hybrid_done_acceptance_link, in->story_sigma, in->hybrid_accepted_sigma,
                     in->hybrid_done_sigma, in->hybrid_done_acceptance_sigma, in->epsilon_pi_HMC, in->story_pi,
                     in->hybrid_accepted_pi, in->hybrid_done_pi, in->hybrid_done_acceptance_pi, in->measure};
    snprintf(name, sizeof(name), "%s_info", prefix); rec(name, 'd', st, 15);
    snprintf(name, sizeof(name), "%s_U", prefix); rec(name, 'd', s->lattice->dlink, 6 * V);
    snprintf(name, sizeof(name), "%s_sigma", prefix); rec(name, 'd', s->lattice->sigma, V);
    snprintf(name, sizeof(name), "%s_sigmamass", prefix); rec(name, 'd', s->lattice->sigmamass, V);
}

/* a Simulation with fresh, zeroed lattice fields (for reads) */
static void fresh_sim(struct Simulation *s, const struct Simulation *from)
{
    memset(s, 0, sizeof(*s));
    s->lattice = (struct Lattice *)calloc(1, sizeof(struct Lattice));
    s->lattice->dlink = (struct Link *)calloc(3 * V, sizeof(struct Link));
    s->lattice->sigma = (double *)calloc(V, sizeof(double));
    s->lattice->sigmamass = (double *)calloc(V, sizeof(double));
    strcpy(s->terma_file, from->terma_file);
    strcpy(s->terma_file_copy, from->terma_file_copy);
}

static void rec_f(const char *name, const struct Fermion *f) { rec(name, 'd', f->fermion, 2 * V); }

static gnx_ref_complex *cvec(void) { return (gnx_ref_complex *)calloc(V, sizeof(gnx_ref_complex)); }

static void many(struct ManyFermions *m)
{
    m->flav = (struct Fermion *)calloc(2, sizeof(struct Fermion));
    for (int f = 0; f < 2; f++) m->flav[f].fermion = cvec();
}

static void z2_complex(struct Fermion *f)
{
    for (int x = 0; x < V; x++) {
        double re = (drv_casuale() <= 0.5) ? 1.0 : -1.0;
        double im = (drv_casuale() <= 0.5) ? 1.0 : -1.0;
        f->fermion[x] = re + I * im;
    }
}

static void draw_sigma(struct Lattice *lat)
{
    double var = 2.0 / (4 * 0.85);
    for (int x = 0; x < HALF; x++) orc_gauss2mv(&R, &lat->sigma[x], &lat->sigma[x + HALF], 0.0, var);
    ComputeFullSigmaMass(lat, NULL, lat->sigmamass);
}

static double last_dh(const char *path)
{
    FILE *f = fopen(path, "r");
    char line[4096], last[4096];
    last[0] = 0;
    while (fgets(line, sizeof(line), f)) strcpy(last, line);
    fclose(f);
    double v[13];
    int n = 0;
    char *tok = strtok(last, "\t\n");
    while (tok && n < 13) { v[n++] = atof(tok); tok = strtok(NULL, "\t\n"); }
    return n >= 12 ? v[11] : NAN;
}

int main(int argc, char **argv)
{
    if (argc < 3) { fprintf(stderr, "usage: %s out.bin accfile\n", argv[0]); return 1; }
    g_out = fopen(argv[1], "wb");
    const char *accfile = argv[2];
    if (GnxRefInit(L, L, 0, drv_gauss1, drv_casuale) != 0) return 3;
    GnxRefSetTolerances(1e-9, 1e-3);

    struct Lattice lat;
    memset(&lat, 0, sizeof(lat));
    lat.sigma = (double *)calloc(V, sizeof(double));
    lat.sigmamass = (double *)calloc(V, sizeof(double));
    struct Info info;
    memset(&info, 0, sizeof(info));
    info.beta_scalar = 0.85;
    struct Fermionkit kit;
    memset(&kit, 0, sizeof(kit));
    many(&kit.r_k); many(&kit.p_k); many(&kit.chi); many(&kit.phi);
    many(&kit.inverted_phi); many(&kit.inverted_phi_obs);
    kit.Ddag_inverted_phi_obs.fermion = cvec();
    struct Forcesigma force = {(double *)calloc(V, sizeof(double))};
    struct Momentasigma mom = {(double *)calloc(V, sizeof(double))};
    struct Fermion rvf = {cvec()};
    struct RandomVec rv = {&rvf};

    /* ---------------- operators ---------------- */
    orc_rng_init(&R, 1231);
    draw_sigma(&lat);
    rec("op_sigmamass", 'd', lat.sigmamass, V);
    struct Fermion b = {cvec()}, c = {cvec()}, y = {cvec()}, x = {cvec()};
    z2_complex(&b);
    z2_complex(&c);
    DiracFermion(&lat, &b, &y, &info, NULL);         rec_f("op_Mb", &y);
    AdjDiracFermion(&lat, &b, &y, &info, NULL);      rec_f("op_Mdagb", &y);
    memcpy(y.fermion, c.fermion, sizeof(gnx_ref_complex) * V);
    SubtractDiracFermion(&lat, &b, &y, &info, NULL); rec_f("op_c_minus_Mb", &y);
    memcpy(y.fermion, c.fermion, sizeof(gnx_ref_complex) * V);
    SubtractAdjDiracFermion(&lat, &b, &y, &info, NULL); rec_f("op_c_minus_Mdagb", &y);

    /* ---------------- CG ---------------- */
    double tols[3] = {1e-10, 1e-6, 1e-3};
    const char *tn[3] = {"1e-10", "1e-6", "1e-3"};
    for (int i = 0; i < 3; i++) {
        char name[64];
        memcpy(x.fermion, c.fermion, sizeof(gnx_ref_complex) * V);
        kit.iterations = 0.0; kit.CG_call = 0;
        bool ok = ConjugateGradientFastDiracAdjFastDiracToTheMinusOne(&lat, &b, &x, &info, NULL, tols[i], &kit, 0);
        snprintf(name, sizeof(name), "cg_null_x_%s", tn[i]); rec_f(name, &x);
        double st[3] = {(double)ok, kit.iterations, (double)kit.CG_call};
        snprintf(name, sizeof(name), "cg_null_stat_%s", tn[i]); rec(name, 'd', st, 3);
        kit.iterations = 0.0; kit.CG_call = 0;
        ok = ConjugateGradientFastDiracAdjFastDiracToTheMinusOne(&lat, &b, &x, &info, NULL, tols[i], &kit, 0);
        double st2[3] = {(double)ok, kit.iterations, (double)kit.CG_call};
        snprintf(name, sizeof(name), "cg_null_again_stat_%s", tn[i]); rec(name, 'd', st2, 3);
        snprintf(name, sizeof(name), "cg_null_again_x_%s", tn[i]); rec_f(name, &x);
        memcpy(x.fermion, c.fermion, sizeof(gnx_ref_complex) * V);
        kit.iterations = 0.0; kit.CG_call = 0;
        ok = StartX0ImposedConjugateGradientFastDiracAdjFastDiracToTheMinusOne(&lat, &b, &x, &info, NULL, tols[i], &kit, 0);
        snprintf(name, sizeof(name), "cg_imp_x_%s", tn[i]); rec_f(name, &x);
        double st3[3] = {(double)ok, kit.iterations, (double)kit.CG_call};
        snprintf(name, sizeof(name), "cg_imp_stat_%s", tn[i]); rec(name, 'd', st3, 3);
    }

    /* ---------------- sigma force ---------------- */
    orc_rng_init(&R, 4242);
    for (int f = 0; f < 2; f++) {
        for (int r = 0; r < V; r++) kit.chi.flav[f].fermion[r] = drv_gauss1();
        DiracFermion(&lat, &kit.chi.flav[f], &kit.phi.flav[f], &info, NULL);
        memset(kit.inverted_phi.flav[f].fermion, 0, sizeof(gnx_ref_complex) * V);
    }
    kit.iterations = 0.0; kit.CG_call = 0;
    ComputeSigmaForce(&lat, &info, NULL, &force, &kit, 1e-10);
    rec("force_F", 'd', force.force, V);
    rec_f("force_X0", &kit.inverted_phi.flav[0]);
    rec_f("force_X1", &kit.inverted_phi.flav[1]);
    rec_f("force_Y0", &kit.inverted_phi_obs.flav[0]);
    rec_f("force_Y1", &kit.inverted_phi_obs.flav[1]);
    double fst[2] = {kit.iterations, (double)kit.CG_call};
    rec("force_stat", 'd', fst, 2);
    /* the accumulate-only entry point on a zeroed force gives the fermion part */
    double *fsave = (double *)malloc(sizeof(double) * V);
    memcpy(fsave, force.force, sizeof(double) * V);
    memset(force.force, 0, sizeof(double) * V);
    AddForceFromFermionSectorRespectToSigma(&lat, &info, NULL, &force, &kit, 1e-10);
    rec("force_fermion_only", 'd', force.force, V);

    /* ---------------- HMC + measurements ---------------- */
    const char *tags[3] = {"hmc_e0.1", "hmc_e0.2", "hmc_e0.05s4"};
    double eps[3] = {0.1, 0.2, 0.05};
    int steps[3] = {2, 2, 4}, ntraj[3] = {20, 20, 5};
    for (int t = 0; t < 3; t++) {
        memset(&info, 0, sizeof(info));
        info.beta_scalar = 0.85;
        info.epsilon_sigma_HMC = eps[t];
        info.steps_sigma_HMC = steps[t];
        orc_rng_init(&R, 1231);
        draw_sigma(&lat);
        for (int f = 0; f < 2; f++) {
            memset(kit.inverted_phi.flav[f].fermion, 0, sizeof(gnx_ref_complex) * V);
            memset(kit.inverted_phi_obs.flav[f].fermion, 0, sizeof(gnx_ref_complex) * V);
        }
        kit.iterations = 0.0; kit.CG_call = 0;
        remove(accfile);
        double dh[64], acc[64], cond[64 * 11], iters[64], calls[64];
        char name[64];
        for (int k = 0; k < ntraj[t]; k++) {
            int before = info.hybrid_accepted_sigma;
            if (!LeapFrog2MinimumNormSigmaVelocityFirst(&lat, &info, &force, &mom, NULL, &kit, eps[t], steps[t], accfile)) {
                fprintf(stderr, "inversion failure\n");
                return 2;
            }
            acc[k] = (double)(info.hybrid_accepted_sigma - before);
            dh[k] = last_dh(accfile);
            double *o = &cond[11 * k];
            ComputeCondensateDirac(&lat, &o[0], &o[1], &o[2], &o[3], &o[4], &o[5], &o[6], &o[7], &o[8], &o[9], &o[10],
                                   &rv, &kit.inverted_phi_obs.flav[0], &info, NULL, 1e-9, &kit, 0);
            iters[k] = kit.iterations;
            calls[k] = (double)kit.CG_call;
            if (k == 0) {
                snprintf(name, sizeof(name), "%s_sigma1", tags[t]);
                rec(name, 'd', lat.sigma, V);
            }
        }
        snprintf(name, sizeof(name), "%s_dh", tags[t]);     rec(name, 'd', dh, ntraj[t]);
        snprintf(name, sizeof(name), "%s_acc", tags[t]);    rec(name, 'd', acc, ntraj[t]);
        snprintf(name, sizeof(name), "%s_cond", tags[t]);   rec(name, 'd', cond, 11 * ntraj[t]);
        snprintf(name, sizeof(name), "%s_iters", tags[t]);  rec(name, 'd', iters, ntraj[t]);
        snprintf(name, sizeof(name), "%s_calls", tags[t]);  rec(name, 'd', calls, ntraj[t]);
        snprintf(name, sizeof(name), "%s_sigmaN", tags[t]); rec(name, 'd', lat.sigma, V);
    }
    /* ---------------- link sector (reference link functions through the drop-in) ---------------- */
    GnxRefBindGauss2(drv_gauss2mv);
    lat.dlink = (struct Link *)calloc(3 * V, sizeof(struct Link));
    orc_rng_init(&R, 555);
    for (int i = 0; i < 3 * V; i++) lat.dlink[i].link = cexp(I * 2.0 * M_PI * drv_casuale());
    orc_rng_init(&R, 4343);
    draw_sigma(&lat);
    memset(&info, 0, sizeof(info));
    info.beta_scalar = 0.85;
    info.beta = 1.0;
    for (int f = 0; f < 2; f++) {
        for (int r = 0; r < V; r++) kit.chi.flav[f].fermion[r] = drv_gauss1();
        DiracFermion(&lat, &kit.chi.flav[f], &kit.phi.flav[f], &info, NULL);
        memset(kit.inverted_phi.flav[f].fermion, 0, sizeof(gnx_ref_complex) * V);
    }
    struct Force lforce = {(double *)calloc(3 * V, sizeof(double))};
    kit.iterations = 0.0; kit.CG_call = 0;
    double lfst[3];
    lfst[0] = (double)AddForceFromFermionSectorRespectToLink(&lat, &info, NULL, &lforce, &kit);
    lfst[1] = kit.iterations; lfst[2] = (double)kit.CG_call;
    rec("link_force_F", 'd', lforce.force, 3 * V);
    rec_f("link_force_X0", &kit.inverted_phi.flav[0]);
    rec_f("link_force_X1", &kit.inverted_phi.flav[1]);
    rec_f("link_force_Y0", &kit.inverted_phi_obs.flav[0]);
    rec_f("link_force_Y1", &kit.inverted_phi_obs.flav[1]);
    rec("link_force_stat", 'd', lfst, 3);
    struct Momenta lmom = {(double *)calloc(3 * V, sizeof(double))};
    info.epsilon_link_HMC = 0.02;
    info.steps_link_HMC = 10;
    orc_rng_init(&R, 4444);
    for (int f = 0; f < 2; f++) memset(kit.inverted_phi.flav[f].fermion, 0, sizeof(gnx_ref_complex) * V);
    kit.iterations = 0.0; kit.CG_call = 0;
    char laccfile[1024];
    snprintf(laccfile, sizeof(laccfile), "%s.link", accfile);
    remove(laccfile);
    double ldh[3], lacc[3];
    for (int t = 0; t < 3; t++) {
        int before = info.hybrid_accepted_link;
        if (!LeapFrog2Link(&lat, &info, &lforce, &lmom, NULL, &kit, 0.02, 10, laccfile)) {
            fprintf(stderr, "link inversion failure\n");
            return 2;
        }
        ldh[t] = info.delta_H_link;
        lacc[t] = (double)(info.hybrid_accepted_link - before);
        if (t == 0) {
            rec("link_traj_mom1", 'd', lmom.momenta, 3 * V);
            rec("link_traj_force1", 'd', lforce.force, 3 * V);
        }
    }
    double lst[2] = {kit.iterations, (double)kit.CG_call};
    rec("link_traj_dh", 'd', ldh, 3);
    rec("link_traj_acc", 'd', lacc, 3);
    rec("link_traj_U", 'd', lat.dlink, 6 * V);
    rec("link_traj_stat", 'd', lst, 2);

    /* ---------------- terma checkpoint through the drop-in (as the harness) ---------------- */
    orc_rng_init(&R, 555);
    for (int i = 0; i < 3 * V; i++) lat.dlink[i].link = cexp(I * 2.0 * M_PI * drv_casuale());
    orc_rng_init(&R, 4343);
    draw_sigma(&lat);
    struct Simulation sim;
    memset(&sim, 0, sizeof(sim));
    sim.lattice = &lat;
    sim.info.epsilon_link_HMC = 0.02; sim.info.story_link = 3;
    sim.info.hybrid_accepted_link = 5; sim.info.hybrid_done_link = 7;
    sim.info.story_sigma = 11; sim.info.hybrid_accepted_sigma = 13; sim.info.hybrid_done_sigma = 17;
    sim.info.epsilon_pi_HMC = 0.05; sim.info.story_pi = 19;
    sim.info.hybrid_accepted_pi = 23; sim.info.hybrid_done_pi = 29;
    sim.info.measure = 31;
    snprintf(sim.terma_file, sizeof(sim.terma_file), "%s.terma", accfile);
    snprintf(sim.terma_file_copy, sizeof(sim.terma_file_copy), "%s.termacopy", accfile);
    remove(sim.terma_file);
    remove(sim.terma_file_copy);
    double tv = TestWriteReadLattice(&sim);
    rec("terma_test", 'd', &tv, 1);
    WriteBinaryLattice(&sim);
    rec_file("terma_first", sim.terma_file);
    sim.info.measure = 32;
    sim.info.hybrid_done_sigma = 18;
    WriteBinaryLattice(&sim);
    rec_file("terma_main", sim.terma_file);
    rec_file("terma_copy", sim.terma_file_copy);
    struct Simulation rd;
    fresh_sim(&rd, &sim);
    ReadBinaryLattice(&rd, rd.terma_file, true);
    rec_terma_state("terma_read", &rd);
    {
        static unsigned char buf[1 << 20];
        FILE *f = fopen(sim.terma_file, "rb");
        size_t n = fread(buf, 1, sizeof(buf), f);
        fclose(f);
        f = fopen(sim.terma_file, "wb");
        fwrite(buf, 1, n / 2, f);
        fclose(f);
    }
    fresh_sim(&rd, &sim);
    ReadBinaryLattice(&rd, rd.terma_file, true);
    rec_terma_state("terma_fallback", &rd);
    remove(sim.terma_file);
    fresh_sim(&rd, &sim);
    ReadBinaryLattice(&rd, rd.terma_file, true);
    rec_terma_state("terma_missing", &rd);
    remove(sim.terma_file_copy);

    GnxRefFinalize();
    fclose(g_out);
    return 0;
}
