/*
 * mbpol_hip.h -- C ABI of the MI355X (gfx950) MB-pol platform.
 *
 * This is the drop-in boundary for one MB-pol energy+force evaluation.  It
 * replaces, for the HIP platform, what the reference's OpenMM plugin exposes
 * through its Force / KernelImpl layers:
 *
 *   system description  <- MBPolOneBodyForce   (openmmapi/include/openmm/MBPolOneBodyForce.h:51-147)
 *                          MBPolTwoBodyForce   (openmmapi/include/openmm/MBPolTwoBodyForce.h:57-178)
 *                          MBPolThreeBodyForce (openmmapi/include/openmm/MBPolThreeBodyForce.h:57-178)
 *                          MBPolElectrostaticsForce (openmmapi/include/openmm/MBPolElectrostaticsForce.h:56-323)
 *                          the dispersion CustomNonbondedForce built by python/mbpol.xml:37-108
 *                          ThreeParticleAverageSite M sites (python/mbpol.xml:14)
 *   context create      <- Calc*ForceKernel::initialize (openmmapi/include/openmm/mbpolKernels.h:49-207,
 *                          platforms/reference/src/MBPolReferenceKernels.cpp:86-99,141-207,341-360,430-449)
 *   evaluate            <- Calc*ForceKernel::execute, summed over the five Forces
 *                          (MBPolReferenceKernels.cpp:101-114,246-259,362-399,451-483)
 *   update parameters   <- copyParametersToContext (MBPolReferenceKernels.cpp:116-127,305-325,401-414,485-498)
 *   electrostatic potential <- getElectrostaticPotential (MBPolReferenceKernels.cpp:260-283)
 *   system moments      <- getSystemElectrostaticsMoments (MBPolReferenceKernels.cpp:285-303)
 *
 * Units follow the reference API: positions and lengths in nm, energies in
 * kJ/mol, forces in kJ/mol/nm, charges in e, polarisabilities in nm^3.
 * Plain pointers and sizes only; errors are returned as status codes with a
 * per-thread message (mbpol_last_error) -- no exceptions cross this ABI.
 * A context binds one GPU and two HIP streams: the electrostatics run on the
 * main stream and the 1-, 2- and 3-body terms and the dispersion on the second
 * one, forked from and joined to the main stream by events.  A host thread owned
 * by the context enqueues the second stream's work while the calling thread
 * enqueues the electrostatics ("fork_early" 2, the default; the context's launch
 * tuning may instead have the calling thread enqueue both); that thread's job is
 * waited for before the join, and every call returns after both streams' work
 * has finished.  Calls on one context must not overlap (the reference kernels are
 * not re-entrant either).
 */
#ifndef MBPOL_HIP_H
#define MBPOL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MBPOL_OK 0
#define MBPOL_ERR_INVALID_ARGUMENT 1    /* bad sizes / indices / options (OpenMMException in the reference) */
#define MBPOL_ERR_HIP 2                 /* a HIP runtime or library call failed */
#define MBPOL_ERR_NOT_CONVERGED 3       /* induced dipoles did not converge
                                           (MBPolReferenceElectrostaticsForce.cpp:888-894) */
#define MBPOL_ERR_BOX_TOO_SMALL 4       /* box < 2 * cutoff (MBPolReferenceKernels.cpp:224-227,387-390,471-474) */
#define MBPOL_ERR_CUTOFF_TOO_LARGE 5    /* cutoff > box/2 at initialisation (MBPol*ForceImpl.cpp:60-66) */
#define MBPOL_ERR_COUNT_CHANGED 6       /* parameter-count change on update (MBPolReferenceKernels.cpp:117-118) */
#define MBPOL_ERR_NO_DEVICE 7           /* no usable GPU */

/* nonbonded methods, numbered as in the reference enums */
#define MBPOL_ONEBODY_NONPERIODIC 0     /* MBPolOneBodyForce::NonPeriodic */
#define MBPOL_ONEBODY_PERIODIC 1        /* MBPolOneBodyForce::Periodic */
#define MBPOL_NOCUTOFF 0                /* MBPolTwoBodyForce / MBPolThreeBodyForce::NoCutoff */
#define MBPOL_CUTOFF_PERIODIC 1         /* ::CutoffPeriodic */
#define MBPOL_CUTOFF_NONPERIODIC 2      /* ::CutoffNonPeriodic */
#define MBPOL_ELEC_NOCUTOFF 0           /* MBPolElectrostaticsForce::NoCutoff */
#define MBPOL_ELEC_PME 1                /* MBPolElectrostaticsForce::PME */

/* component bits for evaluate(); also the order of energies_out[] */
#define MBPOL_ONEBODY 1u
#define MBPOL_TWOBODY 2u
#define MBPOL_THREEBODY 4u
#define MBPOL_DISPERSION 8u
#define MBPOL_ELECTROSTATICS 16u
#define MBPOL_ALL 31u

/* arithmetic of the PIP and pair kernels */
#define MBPOL_PRECISION_MIXED 0         /* FP64 polynomials and electrostatics, FP32 dispersion pair math,
                                           FP64 accumulation */
#define MBPOL_PRECISION_DOUBLE 1        /* FP64 everywhere */
#define MBPOL_PRECISION_SINGLE_PIP 2    /* MIXED with FP32 2-/3-body polynomials: faster, but the 2-/3-body
                                           energies carry relative errors of order 1e-3 */

/* initial guess of the induced-dipole iteration */
#define MBPOL_GUESS_ALPHA_E 0           /* mu0 = alpha * E_fixed, as the reference (:628-639) */
#define MBPOL_GUESS_ZERO 1
#define MBPOL_GUESS_PREVIOUS 2          /* converged dipoles of the previous evaluate() */
#define MBPOL_GUESS_USER 3              /* dipoles passed with mbpol_context_set_dipole_guess() */
#define MBPOL_GUESS_LEARNED 4           /* the network attached with mbpol_guess_attach() */

/* induced-dipole solver (mbpol_context_set_solver); both stop when the reference's convergence measure
 * (:504-511, :588) of the current dipoles falls below the target epsilon */
#define MBPOL_SOLVER_SOR 0              /* the reference's SOR iteration, omega 0.55 (:557-616); default of DOUBLE */
#define MBPOL_SOLVER_PCG 1              /* conjugate gradients with a block-Jacobi preconditioner (the inverse of
                                           each water's 9x9 polarisation block) on the symmetric form of the same
                                           equations; default of MIXED and SINGLE_PIP */

typedef struct mbpol_system mbpol_system;
typedef struct mbpol_context mbpol_context;

typedef struct {
    int32_t iterations;        /* induced-field evaluations of the last evaluate()'s dipole solve (SOR: its
                                  iterations; PCG: its steps plus the initial residual) */
    double epsilon;            /* final convergence measure (reference definition, :504-511, :588) */
    int32_t converged;
    int64_t num_dimers;        /* entries of the 2-body list */
    int64_t num_trimers;       /* entries of the 3-body work list */
    int64_t num_elec_pairs;    /* real-space site pairs within the electrostatics cutoff (PME) */
    int64_t num_disp_pairs;    /* dispersion pairs within the cutoff */
    int64_t num_pol_pairs;     /* directed (polarisable, polarisable) site pairs the dipole iteration visits */
} mbpol_eval_stats;

/* thread-local message of the last failed call on this thread */
const char* mbpol_last_error(void);
/* number of usable GPUs */
int mbpol_device_count(int32_t* count);

/* ---- system description (host side; mirrors the Force classes) ---- */
int mbpol_system_create(int32_t num_particles, mbpol_system** out);
void mbpol_system_destroy(mbpol_system* sys);
/* rectangular periodic box (System::setDefaultPeriodicBoxVectors); zeros = none */
int mbpol_system_set_box(mbpol_system* sys, const double box_nm[3]);
/* ThreeParticleAverageSite: r_site = w1 r_p1 + w2 r_p2 + w3 r_p3; its force is folded back with the same weights */
int mbpol_system_add_virtual_site(mbpol_system* sys, int32_t site, int32_t p1, int32_t p2, int32_t p3, double w1,
                                  double w2, double w3);
/* MBPolOneBodyForce::addOneBody / setNonbondedMethod; idx = n x {O, H1, H2} */
int mbpol_system_set_onebody(mbpol_system* sys, int32_t n, const int32_t* idx, int32_t method);
/* MBPolTwoBodyForce::addParticle / setNonbondedMethod / setCutoff */
int mbpol_system_set_twobody(mbpol_system* sys, int32_t n, const int32_t* idx, int32_t method, double cutoff_nm);
/* MBPolThreeBodyForce::addParticle / setNonbondedMethod / setCutoff */
int mbpol_system_set_threebody(mbpol_system* sys, int32_t n, const int32_t* idx, int32_t method, double cutoff_nm);
/* MBPolElectrostaticsForce::addElectrostatics (one entry per particle, O,H1,H2,M order per water when
 * charge redistribution is on) and its settings; ewald_alpha = 0 or grid[0] = 0 selects the automatic
 * choice from ewald_tol, as NonbondedForceImpl::calcPMEParameters (MBPolReferenceKernels.cpp:186-201) */
int mbpol_system_set_electrostatics(mbpol_system* sys, int32_t n, const double* charge, const int32_t* molecule,
                                    const int32_t* atom_type, const double* damping, const double* polarity,
                                    int32_t method, double cutoff_nm, double ewald_alpha, const int32_t grid[3],
                                    double ewald_tol, int32_t max_iter, double target_eps,
                                    int32_t charge_redistribution, const double thole[5]);
/* damped-C6 dispersion (python/mbpol.xml:37-108): per-particle type (num_types <= 16) and molecule (0 <= m < 2^26);
 * pairs within one molecule and particles whose type has an all-zero C6 row are excluded; c6/d6 are
 * num_types x num_types
 * (kJ nm^6/mol, nm^-1); method as the two-body enum */
int mbpol_system_set_dispersion(mbpol_system* sys, int32_t n, const int32_t* type, const int32_t* molecule,
                                int32_t num_types, const double* c6, const double* d6, int32_t method,
                                double cutoff_nm);

/* independent groups of waters evaluated together (no counterpart in the reference, which evaluates one system per
 * Context): waters [water_start[g], water_start[g+1]) form group g (water w = particles 4w..4w+3, O H1 H2 M; the
 * 1-, 2- and 3-body molecule lists are the waters in order; non-periodic methods only).  No interaction crosses a
 * group; the dipole solve (conjugate gradients) runs until every group meets the target; energies are also
 * reported per group (mbpol_context_get_group_energies).  ngroups = 0 removes the groups. */
int mbpol_system_set_groups(mbpol_system* sys, int32_t ngroups, const int32_t* water_start);

/* ---- context (device side) ---- */
int mbpol_context_create(const mbpol_system* sys, int32_t precision, int32_t device, mbpol_context** out);
void mbpol_context_destroy(mbpol_context* ctx);
/* positions: host, N x 3 nm (virtual-site rows are recomputed);
 * forces_out: host N x 3 (overwritten, virtual-site forces folded back and zeroed) or NULL;
 * energies_out: host [6] = {1B, 2B, 3B, dispersion, electrostatics, total} kJ/mol, or NULL */
int mbpol_context_evaluate(mbpol_context* ctx, const double* positions, double* forces_out, double* energies_out,
                           uint32_t components);
/* same with device pointers (positions and forces stay in HBM; energies_out is host) */
int mbpol_context_evaluate_device(mbpol_context* ctx, const double* d_positions, double* d_forces,
                                  double* energies_out, uint32_t components);
/* device-resident inputs: copy positions into the context once, then evaluate without host transfers;
 * forces stay in HBM until downloaded */
int mbpol_context_upload_positions(mbpol_context* ctx, const double* positions);
int mbpol_context_evaluate_resident(mbpol_context* ctx, double* energies_out, uint32_t components);
int mbpol_context_download_forces(mbpol_context* ctx, double* forces_out);
/* per-kernel timing with HIP events on the context's stream (off by default);
 * name one of the phases "evaluate", "onebody", "lists", "twobody", "threebody", "dispersion", "elec_setup",
 * "elec_iteration", "elec_energy", or one of the single kernels of the dipole iteration "pme_spread",
 * "fft_z_r2c", "fft_y", "fft_x_conv", "fft_z_c2r", "fft_plane_fwd", "fft_plane_inv", "elec_iterate",
 * "spread_fft_z" (the dipole spreading fused with the z pass, used instead of "pme_spread" + "fft_z_r2c" for
 * the dipole grids unless MBPOL_HIP_SPREAD_FUSED=0); returns
 * accumulated milliseconds and the number of timed spans since timing was last enabled */
int mbpol_context_set_timing(mbpol_context* ctx, int32_t enabled);
int mbpol_context_get_timing(mbpol_context* ctx, const char* name, double* total_ms, int64_t* count);
int mbpol_context_set_box(mbpol_context* ctx, const double box_nm[3]);
int mbpol_context_get_stats(const mbpol_context* ctx, mbpol_eval_stats* stats);
/* energies of every group of the last evaluate: ngroups x {1B, 2B, 3B, dispersion, electrostatics, total} */
int mbpol_context_get_group_energies(const mbpol_context* ctx, double* energies);
/* converged induced dipoles of the last evaluate (electrostatics particles x 3, e nm) */
int mbpol_context_get_dipoles(const mbpol_context* ctx, double* dipoles);
/* charges used by the last evaluate (after redistribution) */
int mbpol_context_get_charges(const mbpol_context* ctx, double* charges);
int mbpol_context_set_dipole_guess(mbpol_context* ctx, int32_t mode, const double* dipoles);
/* setMutualInducedTargetEpsilon / setMutualInducedMaxIterations on a live context (the reference sets them on the
 * Force before the context exists, MBPolElectrostaticsForce.h:233-262) */
int mbpol_context_set_dipole_convergence(mbpol_context* ctx, double target_eps, int32_t max_iter);
/* MBPOL_SOLVER_SOR or MBPOL_SOLVER_PCG for the following evaluations (no counterpart in the reference, which
 * always iterates by SOR; the default depends on the precision mode given at creation) */
int mbpol_context_set_solver(mbpol_context* ctx, int32_t solver);
/* launch shapes, capacities and counters of the context (no counterpart in the reference): "site_lanes" =
 * lanes per site of the electrostatics row build, picked by timing 8, 16 and 32 on the first evaluation (0
 * before it); "row_capacity_pol" / "row_capacity_other" = entries per electrostatics neighbour row;
 * "dimer_capacity", "lower_capacity", "trimer_capacity" = entries of the 2- and 3-body work lists (rows and
 * lists grow when a configuration needs more: an overflow is detected at the end of the evaluation, which
 * is then run again); "pip_blocks", "pip3_blocks", "disp_blocks" = grid caps of the 2-body, 3-body and
 * dispersion kernels (0 = one block per 64 units); "pip2_kernel" / "pip3_kernel" = 2- / 3-body kernel
 * (0 relabelling-group loop, 1 term by term; the 2-body kernel is term by term unless
 * MBPOL_HIP_2B_KERNEL=orbit, the 3-body one the loop unless MBPOL_HIP_3B_KERNEL=terms); "fork_early" = where the short-range stream forks (1 after the virtual sites, 0 after
 * the electrostatics set-up, 2 after the virtual sites with its kernels enqueued by a host thread of the
 * context's own); "pip_waves", "pip3_waves", "disp_waves" = the caps in waves per SIMD;
 * "tune_evaluations" = trial evaluations the tuning of the second evaluation ran; "graph_steps",
 * "graph_captures", "graph_replays", "graph_fallbacks" = field evaluations of the captured dipole solve and
 * the graph mode's counters (mbpol_context_set_graph) */
int mbpol_context_get_tuning(const mbpol_context* ctx, const char* name, int32_t* value);
/* copyParametersToContext of MBPolOneBodyForce / MBPolTwoBodyForce / MBPolThreeBodyForce
 * (MBPolReferenceKernels.cpp:116-127, 401-414, 485-498): new {O,H1,H2} index lists for the same number of
 * molecules (MBPOL_ERR_COUNT_CHANGED otherwise) */
int mbpol_context_update_onebody(mbpol_context* ctx, int32_t n, const int32_t* idx);
int mbpol_context_update_twobody(mbpol_context* ctx, int32_t n, const int32_t* idx);
int mbpol_context_update_threebody(mbpol_context* ctx, int32_t n, const int32_t* idx);
/* copyParametersToContext for the electrostatics (same count required) */
int mbpol_context_update_electrostatics(mbpol_context* ctx, int32_t n, const double* charge, const int32_t* molecule,
                                        const int32_t* atom_type, const double* damping, const double* polarity);
/* neighbour lists of the last evaluate, for parity checks: pairs (i<j) and chain triplets (i>j>k), molecule
 * indices of the two/three-body Force.  The trimer list is the full reference chain list when
 * full_chain_list was requested at the evaluate (see mbpol_context_set_list_mode), else the work list. */
int mbpol_context_get_dimer_list(const mbpol_context* ctx, int32_t* pairs, int64_t capacity, int64_t* count);
int mbpol_context_get_trimer_list(const mbpol_context* ctx, int32_t* triplets, int64_t capacity, int64_t* count);
/* 1 = keep every chain triplet (reference list); 0 (default) = drop triplets whose switch product is exactly 0 */
int mbpol_context_set_list_mode(mbpol_context* ctx, int32_t full_chain_list);
/* Graph mode (no counterpart in the reference; OpenMM's CUDA platform launches kernel by kernel): 1 =
 * evaluations are captured in a HIP graph once and replayed while the host state they were captured from
 * holds -- the same arguments, parameters, list and row capacities and guess start; every setter above
 * marks a captured graph stale.  The dipole solve is captured as a fixed number of conjugate-gradient steps
 * (the last solve's + 1; steps after the stopping rule return at once); after each replay the host checks
 * the stopping rule, the lists and the rows, and runs the evaluation again eagerly when one of them did not
 * fit (the next call captures anew).  Evaluations a graph cannot hold run eagerly: timing on, the SOR
 * solver, a learned guess or its history, hipFFT transforms, and the first evaluation of a context, which
 * sizes every buffer.  Results equal the eager path's bit for bit.  0 (default) = eager. */
int mbpol_context_set_graph(mbpol_context* ctx, int32_t enabled);
/* MBPolElectrostaticsForce::getElectrostaticPotential at npoints points (nm), kJ/mol/e, from the charges and
 * induced dipoles of the last evaluate (MBPolReferenceElectrostaticsForce.cpp:1023-1086) */
int mbpol_context_get_electrostatic_potential(mbpol_context* ctx, int32_t npoints, const double* points,
                                              double* potential);
/* No counterpart in the reference: mbpol_context_get_electrostatic_potential without any allocation, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops).  The points and
 * results are staged through a fixed 2 MiB device array of the library (one per GPU, held until the process
 * ends): at most 65536 points per call (MBPOL_ERR_INVALID_ARGUMENT beyond).  Calls that overlap on one GPU
 * share that array and race. */
int mbpol_context_get_electrostatic_potential_noalloc(mbpol_context* ctx, int32_t npoints, const double* points,
                                                      double* potential);
/* MBPolElectrostaticsForce::getSystemElectrostaticsMoments (mbpolKernels.h:120; Reference glue
 * MBPolReferenceKernels.cpp:285-303, sums MBPolReferenceElectrostaticsForce.cpp:923-1021): moments[13] = net
 * charge (e), dipole (3, Debye), traceless quadrupole (3x3 row-major, Buckingham) about the centre of mass of
 * masses[n] (amu, one per particle), from the positions, charges and induced dipoles of the last evaluate */
int mbpol_context_get_system_electrostatics_moments(mbpol_context* ctx, const double* masses, double* moments);

/* ---- learned initial guess of the induced dipoles (no counterpart in the reference) ----
 * A per-atom network (48 radial features of the neighbourhood within 0.6 nm -> 64 -> 64 -> 3, SiLU, BF16 MFMA with
 * FP32 accumulation) predicts (mu - alpha E_fixed) / |alpha E_fixed| in a frame of the atom anchored on its
 * alpha E_fixed.  It is trained on the GPU from snapshots of
 * converged dipoles taken from a context's evaluations.  Water systems with the O H H M site layout only. */
typedef struct mbpol_guess mbpol_guess;
#define MBPOL_GUESS_PARAMS 7491         /* W1[48][64] b1[64] W2[64][64] b2[64] W3[64][3] b3[3], matrices [in][out] */
#define MBPOL_GUESS_FEATURES 48
/* what the network corrects: alpha E_fixed, anchored on it (the default), or the previous evaluation's
 * converged dipoles, anchored on the change of alpha E_fixed since that evaluation (the context then keeps
 * the last two evaluations' dipoles and fields; a snapshot needs two earlier evaluations) */
#define MBPOL_GUESS_KIND_ALPHA_E 0
#define MBPOL_GUESS_KIND_PREVIOUS 1
int mbpol_guess_create(mbpol_context* ctx, int32_t max_snapshots, uint64_t seed, mbpol_guess** out);
int mbpol_guess_create_kind(mbpol_context* ctx, int32_t max_snapshots, uint64_t seed, int32_t kind,
                            mbpol_guess** out);
void mbpol_guess_destroy(mbpol_guess* g);
/* features and targets (converged dipoles) of the context's last evaluate */
int mbpol_guess_add_snapshot(mbpol_guess* g);
int mbpol_guess_num_snapshots(const mbpol_guess* g, int32_t* count);
/* the network's guess (dipoles, particles x 3) for the context's last evaluated configuration, from alpha
 * E_fixed or (previous-step kind) from the evaluation before it -- as a snapshot pairs them */
int mbpol_guess_predict_last(mbpol_guess* g, double* dipoles);
/* stored snapshot k: features (3 per water x 48) and targets (3 per water x 3, frame components of
 * (mu - alpha E_fixed) / |alpha E_fixed|) */
int mbpol_guess_get_snapshot(const mbpol_guess* g, int32_t k, float* features, float* targets);
/* Adam on the stored snapshots (input normalisation and output scale set from them on the first call);
 * loss_per_epoch[epochs]: mean squared error in units of the output scale */
int mbpol_guess_train(mbpol_guess* g, int32_t epochs, int32_t batch, double learning_rate, float* loss_per_epoch);
/* the same on the first ntrain stored snapshots only (the rest held out) */
int mbpol_guess_train_range(mbpol_guess* g, int32_t ntrain, int32_t epochs, int32_t batch, double learning_rate,
                            float* loss_per_epoch);
/* mean squared error (units of the output scale) of the trained network on stored snapshots
 * [first, first + count), in batches of `batch` rows */
int mbpol_guess_validation_loss(mbpol_guess* g, int32_t first, int32_t count, int32_t batch, double* loss);
/* attach (1) or detach (0) the network as the context's MBPOL_GUESS_LEARNED guess (the mode is set too) */
int mbpol_guess_attach(mbpol_guess* g, int32_t on);
/* parameters (MBPOL_GUESS_PARAMS floats) and normalisation (48 means, 48 inverse deviations, output scale) */
int mbpol_guess_get_params(const mbpol_guess* g, float* params, float* norm);
int mbpol_guess_set_params(mbpol_guess* g, const float* params, const float* norm);
/* gradient of (1/n) sum |net(x_i) - t_i / scale|^2 over n raw feature rows x (n x 48) and targets t (n x 3), as
 * the training step computes it; grad[MBPOL_GUESS_PARAMS] */
int mbpol_guess_gradient(mbpol_guess* g, int32_t n, const float* x, const float* t, float* grad, float* loss);
/* network outputs (n x 3, unscaled) for n raw feature rows */
int mbpol_guess_forward(mbpol_guess* g, int32_t n, const float* x, float* y);
/* features (3 per water x 48) and frames (3 x 3 per atom, rows e1 e2 e3: e1 along the atom's alpha E_fixed, e2
 * from the HOH bisector (O) or O -> H (H) orthogonalised, e3 = e1 x e2) of the context's last evaluate */
int mbpol_guess_features(mbpol_guess* g, float* features, float* frames);

/* ---- Langevin dynamics on the resident positions (BAOAB; no counterpart in the reference, whose examples drive
 * OpenMM integrators, python/water14.py:36-126) ---- */
typedef struct mbpol_md mbpol_md;
/* masses in amu per particle (0: virtual site); velocities drawn at temperature_K; forces of the components
 * (MBPOL_ALL or a subset) evaluated once */
int mbpol_md_create(mbpol_context* ctx, const double* masses, double temperature_K, double friction_per_ps,
                    double dt_ps, uint64_t seed, uint32_t components, mbpol_md** out);
void mbpol_md_destroy(mbpol_md* md);
/* nsteps steps; energies_out[6] of the last force evaluation and the kinetic energy (kJ/mol) after it, or NULL */
int mbpol_md_step(mbpol_md* md, int32_t nsteps, double* energies_out, double* kinetic_out);
int mbpol_md_get_positions(const mbpol_md* md, double* positions);

#ifdef __cplusplus
}
#endif
#endif /* MBPOL_HIP_H */
