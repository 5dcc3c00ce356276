/* struct fields whose offsets the drop-in's layout mirrors must share with
 * the reference's types (included by oracle/ref_harness.c against the
 * reference headers and by tests/refshim/layout.c against gnx_refapi.h) */
#define GNX_LAYOUT_FIELDS(X)                                  \
    X(Lattice, dlink) X(Lattice, sigma) X(Lattice, sigmamass) \
    X(Lattice, eta_pbc) X(Lattice, distance_sigma)            \
    X(Lattice, done_sigma) X(Lattice, done_pi)                \
    X(Fermioncoor, fermionmm) X(Fermioncoor, eta_pbc)         \
    X(Fermionkit, chi) X(Fermionkit, phi)                     \
    X(Fermionkit, inverted_phi) X(Fermionkit, inverted_phi_obs) \
    X(Fermionkit, Ddag_inverted_phi_obs) X(Fermionkit, bminusax0) \
    X(Fermionkit, iterations) X(Fermionkit, CG_call)          \
    X(Info, beta_scalar) X(Info, measure) X(Info, epsilon_sigma_HMC) \
    X(Info, steps_sigma_HMC) X(Info, hybrid_accepted_sigma)   \
    X(Info, hybrid_done_acceptance_sigma) X(Info, local_acceptance_sigma) \
    X(Info, delta_H_sigma) X(Info, story_sigma) X(Info, story_pi) \
    X(Simulation, info) X(Simulation, fcor) X(Simulation, observables_file) \
    X(Simulation, terma_file) X(Simulation, terma_file_copy)
#define GNX_LAYOUT_TYPES(X) X(Lattice) X(Fermioncoor) X(Fermion) X(Fermionkit) X(Info) \
    X(Momentasigma) X(Forcesigma) X(RandomVec) X(Link) X(Simulation)
