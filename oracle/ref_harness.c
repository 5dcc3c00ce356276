/*
 * ref_harness.c -- golden-vector generator.  Linked against the reference's
 * own sources (all *.c except test.c, compiled in place from the reference
 * checkout by oracle/Makefile), at the lattice size its macro.h ships
 * (LSPACE = LTIME = 8, macro.h:21-22).  It replaces the reference's main
 * (test.c:3-42) with a fixed-seed driver: Initrand(seed) instead of the
 * time-based StartSeed (random.c:72-80), sigma drawn explicitly, and the
 * hot-path functions called directly.
 *
 * Output: a flat binary file of named records
 *   char name[32] | char kind ('d' double, 'i' int, 'b' byte) | pad[7] | int64 count | data
 * read by tests/golden/make_golden.py.
 *
 * TEST INFRASTRUCTURE ONLY.
 */
#include "simulation.h"
#include "observables.h"

#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "../tests/refshim/layout_fields.h"

static FILE *g_out;

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
    snprintf(name, sizeof(name), "%s_U", prefix); rec(name, 'd', s->lattice->dlink, 2 * DIM * VOLUME);
    snprintf(name, sizeof(name), "%s_sigma", prefix); rec(name, 'd', s->lattice->sigma, VOLUME);
    snprintf(name, sizeof(name), "%s_sigmamass", prefix); rec(name, 'd', s->lattice->sigmamass, VOLUME);
}

/* a Simulation with fresh, zeroed lattice fields sharing fcor (for reads) */
static void fresh_sim(struct Simulation *s, const struct Simulation *from)
{
    memset(s, 0, sizeof(*s));
    s->lattice = malloc(sizeof(Lattice));
    AllocateLattice(s->lattice);
    memset(s->lattice->dlink, 0, sizeof(struct Link) * DIM * VOLUME);
    memset(s->lattice->sigma, 0, sizeof(double) * VOLUME);
    memset(s->lattice->sigmamass, 0, sizeof(double) * VOLUME);
    s->fcor = from->fcor;
    strcpy(s->terma_file, from->terma_file);
    strcpy(s->terma_file_copy, from->terma_file_copy);
}

static void rec_f(const char *name, const struct Fermion *f)
{
    rec(name, 'd', f->fermion, 2 * VOLUME);
}

static void rec_scalar(const char *name, double v) { rec(name, 'd', &v, 1); }

static void z2_complex(struct Fermion *f)
{
    for (int x = 0; x < VOLUME; x++) {
        double re = (Casuale() <= 0.5) ? 1.0 : -1.0;
        double im = (Casuale() <= 0.5) ? 1.0 : -1.0;
        f->fermion[x] = re + I * im;
    }
}

static void draw_sigma(struct Lattice *lat, const struct Fermioncoor *fcor, double beta_s)
{
    double var = 2.0 / (NFLAV * beta_s);
    for (int x = 0; x < HALFVOLUME; x++)
        Gauss2MeanVariance(&(lat->sigma[x]), &(lat->sigma[x + HALFVOLUME]), 0.0, var);
    ComputeFullSigmaMass(lat, fcor, lat->sigmamass);
}

static void setup_info(struct Info *info, double beta_s, double eps, int steps)
{
    memset(info, 0, sizeof(*info));
    info->beta = 0.0;
    info->beta_scalar = beta_s;
    info->mass = 0.0;
    info->epsilon_sigma_HMC = eps;
    info->steps_sigma_HMC = steps;
    info->distance_sigma_HMC = eps * steps;
}

/* parse the Delta-H column (12th) of the last line of the acceptance file */
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

static void run_hmc(struct Lattice *lat, struct Fermioncoor *fcor, struct Fermionkit *kit,
                    struct Momentasigma *mom, struct Forcesigma *force, struct RandomVec *rv,
                    double eps, int steps, int ntraj, const char *tag, const char *accfile)
{
    struct Info info;
    setup_info(&info, 0.85, eps, steps);
    Initrand(1231);
    draw_sigma(lat, fcor, 0.85);
    for (int x = 0; x < VOLUME; x++) for (int f = 0; f < HALFNFLAV; f++) {
        kit->inverted_phi.flav[f].fermion[x] = 0.0;
        kit->inverted_phi_obs.flav[f].fermion[x] = 0.0;
    }
    kit->iterations = 0.0;
    kit->CG_call = 0;
    char name[64];
    snprintf(name, sizeof(name), "%s_sigma0", tag);
    rec(name, 'd', lat->sigma, VOLUME);
    remove(accfile);
    double dh[64], acc[64], cond[64 * 11], csig[64 * 8], iters[64], calls[64];
    for (int t = 0; t < ntraj; t++) {
        int before = info.hybrid_accepted_sigma;
        if (!LeapFrog2MinimumNormSigmaVelocityFirst(lat, &info, force, mom, fcor, kit, eps, steps, accfile)) {
            fprintf(stderr, "inversion failure\n");
            exit(2);
        }
        acc[t] = (double)(info.hybrid_accepted_sigma - before);
        dh[t] = last_dh(accfile);
        double *o = &cond[11 * t];
        ComputeCondensateDirac(lat, &o[0], &o[1], &o[2], &o[3], &o[4], &o[5], &o[6], &o[7], &o[8], &o[9], &o[10],
                               rv, &(kit->inverted_phi_obs.flav[0]), &info, fcor, HIGH_PRECISION_INVERSION, kit, 0);
        double *s = &csig[8 * t];
        CondensateSigmaPi(lat, &info, &s[0], &s[1], &s[2], &s[3], &s[4], &s[5], &s[6], &s[7]);
        iters[t] = kit->iterations;
        calls[t] = (double)kit->CG_call;
        if (t == 0) {
            snprintf(name, sizeof(name), "%s_sigma1", tag);
            rec(name, 'd', lat->sigma, VOLUME);
            snprintf(name, sizeof(name), "%s_mom1", tag);
            rec(name, 'd', mom->momenta, VOLUME);
        }
    }
    snprintf(name, sizeof(name), "%s_dh", tag);      rec(name, 'd', dh, ntraj);
    snprintf(name, sizeof(name), "%s_acc", tag);     rec(name, 'd', acc, ntraj);
    snprintf(name, sizeof(name), "%s_cond", tag);    rec(name, 'd', cond, 11 * ntraj);
    snprintf(name, sizeof(name), "%s_csig", tag);    rec(name, 'd', csig, 8 * ntraj);
    snprintf(name, sizeof(name), "%s_iters", tag);   rec(name, 'd', iters, ntraj);
    snprintf(name, sizeof(name), "%s_calls", tag);   rec(name, 'd', calls, ntraj);
    snprintf(name, sizeof(name), "%s_sigmaN", tag);  rec(name, 'd', lat->sigma, VOLUME);
}

int main(int argc, char **argv)
{
    if (argc < 3) { fprintf(stderr, "usage: %s out.bin accfile\n", argv[0]); return 1; }
    g_out = fopen(argv[1], "wb");
    const char *accfile = argv[2];

    int meta[4] = {LSPACE, LTIME, VOLUME, NFLAV};
    rec("meta", 'i', meta, 4);

    /* struct layouts the drop-in library mirrors */
    int layout[64], nl = 0;
#define OFF(T, F) layout[nl++] = (int)offsetof(struct T, F);
#define SZ(T) layout[nl++] = (int)sizeof(struct T);
    GNX_LAYOUT_FIELDS(OFF)
    GNX_LAYOUT_TYPES(SZ)
#undef OFF
#undef SZ
    rec("struct_layout", 'i', layout, nl);

    /* ---------------- RNG stream ---------------- */
    Initrand(1231);
    static double u[6000];
    for (int i = 0; i < 6000; i++) u[i] = Casuale();
    rec("rng_casuale_1231", 'd', u, 6000);
    Initrand(77);
    double gs[400];
    for (int i = 0; i < 200; i++) gs[i] = Gauss1();
    for (int i = 0; i < 100; i++) Gauss2MeanVariance(&gs[200 + 2 * i], &gs[201 + 2 * i], 0.0, 0.5);
    rec("rng_gauss_77", 'd', gs, 400);

    /* ---------------- geometry ---------------- */
    struct Simulation sim;
    sim.lattice = malloc(sizeof(Lattice));
    sim.fcor = malloc(sizeof(Fermioncoor));
    sim.kit = malloc(sizeof(Fermionkit));
    sim.momentasigma = malloc(sizeof(Momentasigma));
    sim.forcesigma = malloc(sizeof(Forcesigma));
    AllocateLattice(sim.lattice);
    AllocateFermionkit(sim.kit);
    AllocateFermioncoor(sim.fcor);
    AllocateMomentaSigma(sim.momentasigma);
    AllocateForceSigma(sim.forcesigma);
    StartLattice(sim.lattice);
    StartFermioncoorFromLattice(sim.fcor, sim.lattice);
    struct Lattice *lat = sim.lattice;
    struct Fermioncoor *fcor = sim.fcor;
    struct Fermionkit *kit = sim.kit;

    rec("fcor_fermionpp", 'i', fcor->fermionpp, VOLUME * DIM);
    rec("fcor_fermionmm", 'i', fcor->fermionmm, VOLUME * DIM);
    rec("fcor_eta_pp", 'd', fcor->eta_pp, VOLUME * DIM);
    rec("fcor_eta_mm", 'd', fcor->eta_mm, VOLUME * DIM);
    rec("fcor_eta_pbc", 'd', fcor->eta_pbc, VOLUME * DIM);
    rec("fcor_site", 'i', fcor->site, VOLUME);
    rec("lat_fermionsite", 'i', lat->fermionsite, VOLUME);
    rec("lat_lattice_site", 'i', lat->lattice_site, VOLUME * DIM);

    /* ---------------- operators on a drawn sigma and complex Z2 source ---------------- */
    struct Info info;
    setup_info(&info, 0.85, 0.1, 2);
    Initrand(1231);
    draw_sigma(lat, fcor, 0.85);
    rec("op_sigma", 'd', lat->sigma, VOLUME);
    rec("op_sigmamass", 'd', lat->sigmamass, VOLUME);

    struct Fermion b, c, y;
    AllocateFermion(&b); AllocateFermion(&c); AllocateFermion(&y);
    z2_complex(&b);
    z2_complex(&c);
    rec_f("op_b", &b);
    rec_f("op_c", &c);
    DiracFermion(lat, &b, &y, &info, fcor);            rec_f("op_Mb", &y);
    AdjDiracFermion(lat, &b, &y, &info, fcor);         rec_f("op_Mdagb", &y);
    CopyFermion(&c, &y);
    SubtractDiracFermion(lat, &b, &y, &info, fcor);    rec_f("op_c_minus_Mb", &y);
    CopyFermion(&c, &y);
    SubtractAdjDiracFermion(lat, &b, &y, &info, fcor); rec_f("op_c_minus_Mdagb", &y);

    /* ---------------- CG ---------------- */
    struct Fermion x;
    AllocateFermion(&x);
    double tols[3] = {1e-10, 1e-6, 1e-3};
    const char *tn[3] = {"1e-10", "1e-6", "1e-3"};
    for (int i = 0; i < 3; i++) {
        char name[64];
        /* START_X0_NULL with a nonzero x_in = c (prologue does not exit) */
        CopyFermion(&c, &x);
        kit->iterations = 0.0; kit->CG_call = 0;
        bool ok = ConjugateGradientFastDiracAdjFastDiracToTheMinusOne(lat, &b, &x, &info, fcor, tols[i], kit, 0);
        snprintf(name, sizeof(name), "cg_null_x_%s", tn[i]);  rec_f(name, &x);
        double st[3] = {(double)ok, kit->iterations, (double)kit->CG_call};
        snprintf(name, sizeof(name), "cg_null_stat_%s", tn[i]); rec(name, 'd', st, 3);
        /* a second call with the converged x_in exits in the prologue */
        kit->iterations = 0.0; kit->CG_call = 0;
        ok = ConjugateGradientFastDiracAdjFastDiracToTheMinusOne(lat, &b, &x, &info, fcor, tols[i], kit, 0);
        double st2[3] = {(double)ok, kit->iterations, (double)kit->CG_call};
        snprintf(name, sizeof(name), "cg_null_again_stat_%s", tn[i]); rec(name, 'd', st2, 3);
        snprintf(name, sizeof(name), "cg_null_again_x_%s", tn[i]);  rec_f(name, &x);
        /* StartX0Imposed */
        CopyFermion(&c, &x);
        kit->iterations = 0.0; kit->CG_call = 0;
        ok = StartX0ImposedConjugateGradientFastDiracAdjFastDiracToTheMinusOne(lat, &b, &x, &info, fcor, tols[i], kit, 0);
        snprintf(name, sizeof(name), "cg_imp_x_%s", tn[i]);  rec_f(name, &x);
        double st3[3] = {(double)ok, kit->iterations, (double)kit->CG_call};
        snprintf(name, sizeof(name), "cg_imp_stat_%s", tn[i]); rec(name, 'd', st3, 3);
    }

    /* ---------------- sigma force from real pseudo-fermions ---------------- */
    Initrand(4242);
    for (int f = 0; f < HALFNFLAV; f++) {
        DrawFullFermionFromGaussianDistribution(&(kit->chi.flav[f]));
        DiracFermion(lat, &(kit->chi.flav[f]), &(kit->phi.flav[f]), &info, fcor);
        SetFermionToZero(&(kit->inverted_phi.flav[f]));
    }
    rec_f("force_phi0", &(kit->phi.flav[0]));
    rec_f("force_phi1", &(kit->phi.flav[1]));
    kit->iterations = 0.0; kit->CG_call = 0;
    ComputeSigmaForce(lat, &info, fcor, sim.forcesigma, kit, 1e-10);
    rec("force_F", 'd', sim.forcesigma->force, VOLUME);
    rec_f("force_X0", &(kit->inverted_phi.flav[0]));
    rec_f("force_X1", &(kit->inverted_phi.flav[1]));
    rec_f("force_Y0", &(kit->inverted_phi_obs.flav[0]));
    rec_f("force_Y1", &(kit->inverted_phi_obs.flav[1]));
    double fst[2] = {kit->iterations, (double)kit->CG_call};
    rec("force_stat", 'd', fst, 2);

    /* ---------------- HMC + measurements (20 trajectories, 2 steps) ---------------- */
    struct RandomVec rv;
    AllocateRandomVec(&rv);
    run_hmc(lat, fcor, kit, sim.momentasigma, sim.forcesigma, &rv, 0.1, 2, 20, "hmc_e0.1", accfile);
    run_hmc(lat, fcor, kit, sim.momentasigma, sim.forcesigma, &rv, 0.2, 2, 20, "hmc_e0.2", accfile);
    run_hmc(lat, fcor, kit, sim.momentasigma, sim.forcesigma, &rv, 0.05, 4, 5, "hmc_e0.05s4", accfile);

    /* ---------------- link sector: the reference's own link functions ---------------- */
    /* (reachable from no driver path, observables.c:127,183; called directly here) */
    Initrand(555);
    for (int i = 0; i < DIM * VOLUME; i++) lat->dlink[i].link = cexp(I_UNIT * 2.0 * PI * Casuale());
    rec("link_U", 'd', lat->dlink, 2 * DIM * VOLUME);
    Initrand(4343);
    draw_sigma(lat, fcor, 0.85);
    rec("link_sigmamass", 'd', lat->sigmamass, VOLUME);
    for (int f = 0; f < HALFNFLAV; f++) {
        DrawFullFermionFromGaussianDistribution(&(kit->chi.flav[f]));
        DiracFermion(lat, &(kit->chi.flav[f]), &(kit->phi.flav[f]), &info, fcor);
        SetFermionToZero(&(kit->inverted_phi.flav[f]));
    }
    rec_f("link_phi0", &(kit->phi.flav[0]));
    rec_f("link_phi1", &(kit->phi.flav[1]));
    struct Force lforce;
    AllocateForce(&lforce);
    SetForceToZero(&lforce);
    kit->iterations = 0.0; kit->CG_call = 0;
    double lfst[3];
    lfst[0] = (double)AddForceFromFermionSectorRespectToLink(lat, &info, fcor, &lforce, kit);
    lfst[1] = kit->iterations; lfst[2] = (double)kit->CG_call;
    rec("link_force_F", 'd', lforce.force, DIM * VOLUME);
    rec_f("link_force_X0", &(kit->inverted_phi.flav[0]));
    rec_f("link_force_X1", &(kit->inverted_phi.flav[1]));
    rec_f("link_force_Y0", &(kit->inverted_phi_obs.flav[0]));
    rec_f("link_force_Y1", &(kit->inverted_phi_obs.flav[1]));
    rec("link_force_stat", 'd', lfst, 3);

    /* LeapFrog2Link as shipped (PURE_GAUGE off): 3 trajectories, eps 0.02, 10 steps */
    struct Momenta lmom;
    AllocateMomenta(&lmom);
    struct Info linfo;
    setup_info(&linfo, 0.85, 0.1, 2);
    linfo.beta = 1.0;
    linfo.epsilon_link_HMC = 0.02;
    linfo.steps_link_HMC = 10;
    Initrand(4444);
    for (int f = 0; f < HALFNFLAV; f++) SetFermionToZero(&(kit->inverted_phi.flav[f]));
    kit->iterations = 0.0; kit->CG_call = 0;
    remove(accfile);
    double ldh[3], lacc[3];
    for (int t = 0; t < 3; t++) {
        int before = linfo.hybrid_accepted_link;
        if (!LeapFrog2Link(lat, &linfo, &lforce, &lmom, fcor, kit, 0.02, 10, accfile)) {
            fprintf(stderr, "link inversion failure\n");
            exit(2);
        }
        ldh[t] = linfo.delta_H_link;
        lacc[t] = (double)(linfo.hybrid_accepted_link - before);
        if (t == 0) {
            rec("link_traj_mom1", 'd', lmom.momenta, DIM * VOLUME);
            rec("link_traj_force1", 'd', lforce.force, DIM * VOLUME);
        }
    }
    double lst[2] = {kit->iterations, (double)kit->CG_call};
    rec("link_traj_dh", 'd', ldh, 3);
    rec("link_traj_acc", 'd', lacc, 3);
    rec("link_traj_U", 'd', lat->dlink, 2 * DIM * VOLUME);
    rec("link_traj_stat", 'd', lst, 2);

    /* ---------------- terma checkpoint (simulation.c:110-401) ---------------- */
    /* Links of seed 555 and sigma of seed 4343 (as drawn above), distinct
     * counters.  Info is zeroed first: the story fields are ints written and
     * read as 8-byte doubles (simulation.c:169,173,178), so each slot also
     * carries the next field or the struct's tail padding. */
    Initrand(555);
    for (int i = 0; i < DIM * VOLUME; i++) lat->dlink[i].link = cexp(I_UNIT * 2.0 * PI * Casuale());
    Initrand(4343);
    draw_sigma(lat, fcor, 0.85);
    memset(&(sim.info), 0, sizeof(sim.info));
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
    rec_scalar("terma_test", TestWriteReadLattice(&sim));
    WriteBinaryLattice(&sim);                       /* no earlier file: main only */
    rec_file("terma_first", sim.terma_file);
    sim.info.measure = 32;
    sim.info.hybrid_done_sigma = 18;
    WriteBinaryLattice(&sim);                       /* the first file becomes the copy */
    rec_file("terma_main", sim.terma_file);
    rec_file("terma_copy", sim.terma_file_copy);
    struct Simulation rd;
    fresh_sim(&rd, &sim);
    ReadBinaryLattice(&rd, rd.terma_file, true);    /* intact main file */
    rec_terma_state("terma_read", &rd);
    {   /* main file cut short: the reader falls back to the copy */
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
    remove(sim.terma_file);                         /* main file missing: the copy is read */
    fresh_sim(&rd, &sim);
    ReadBinaryLattice(&rd, rd.terma_file, true);
    rec_terma_state("terma_missing", &rd);
    remove(sim.terma_file_copy);

    fclose(g_out);
    return 0;
}
