/*
 * origami MI355X engine — C ABI.
 *
 * A drop-in for origami's energy-and-gradient entry point and its BAOAB /
 * SHAKE / REMD step loop.  All inputs are resolved per-atom and per-term
 * arrays (the content of origami's ForceScratch plus the resolved bonded
 * lists), so the caller's structure/topology/force-field code (origami's
 * crates/geom, chem) stays as it is.  Plain pointers and sizes only; every
 * function returns 0 on success or a negative og_status, with a message in
 * og_last_error().  Units: Å, kJ/mol, kJ/mol/Å, fs, K, Da, e.
 *
 * One og_engine holds a batch of 1..N independent systems (chains or
 * replicas) resident on one GPU; single-system use is a batch of one.
 * Handles are not thread-safe; distinct handles may be used concurrently.
 */
#ifndef ORIGAMI_GPU_H
#define ORIGAMI_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_engine og_engine;

enum og_status {
    OG_OK = 0,
    OG_ERR_ARG = -1,        /* invalid argument / size */
    OG_ERR_HIP = -2,        /* HIP runtime failure */
    OG_ERR_STATE = -3,      /* call sequence (e.g. run before setup) */
    OG_ERR_NONFINITE = -4,  /* a system diverged or was stopped (see og_get_status) */
};

/* Per-system status bits (og_get_status's diverged[], og_run's stop reasons). */
#define OG_STATUS_NONFINITE 1u  /* non-finite velocities after a step (langevin.rs:316-319) */
#define OG_STATUS_RANGE 2u      /* a force/energy contribution NaN or >= 2^30 kJ/mol(/Å), or a summed force >= 2^30: outside
                                   the int64 fixed-point accumulators' range */
#define OG_STATUS_SASA 4u       /* a SASA atom exceeded the arrangement's capacity
                                   (128 caps, 256 arcs, 256 vertices) */

/* Force flags. */
#define OG_MODE_A 1u        /* origami's integrator cutoffs: LJ + reaction-field Coulomb at
                               cutoff_a, GB pair force at 10 Å, descreening at 20 Å.
                               Without it: all pairs, no cutoff (bare Coulomb). */
#define OG_SASA 2u          /* add analytical (PowerSasa) hydrophobic SASA forces */
#define OG_ENERGY 4u        /* also accumulate per-system energy terms */
#define OG_SURROGATE 8u     /* Born radii from the learned surrogate (og_surrogate_*) */
#define OG_SASA_SR 16u      /* og_energies only: the SASA energy term by Shrake-Rupley dot
                               density (960 dots, sasa.rs:113-178), the term origami's REMD
                               adds to its swap energies (remd.rs:529-539); not with OG_SASA */

/* Per-system energy terms written by og_forces / og_energies
 * (n_systems x OG_N_ENERGY doubles, kJ/mol). */
enum og_energy_term {
    OG_EN_BOND = 0, OG_EN_ANGLE, OG_EN_DIHEDRAL, OG_EN_IMPROPER, OG_EN_LJ, OG_EN_COULOMB,
    OG_EN_GB_SELF, OG_EN_GB_PAIR, OG_EN_SASA, OG_EN_KINETIC, OG_N_ENERGY
};

typedef struct og_system_desc {
    int32_t n_atoms;
    const double* pos;              /* [n_atoms*3] Å, atom-major x,y,z */
    const double* charge;           /* [n_atoms] e */
    const double* rmin_half;        /* [n_atoms] CHARMM Rmin/2, Å */
    const double* epsilon;          /* [n_atoms] LJ well depth, kcal/mol (positive) */
    const double* rmin_half_14;     /* [n_atoms] 1-4 Rmin/2 (= rmin_half if absent) */
    const double* epsilon_14;       /* [n_atoms] 1-4 epsilon (= epsilon if absent) */
    const double* gb_rho;           /* [n_atoms] intrinsic Born radius rho, Å */
    const double* gb_scale;         /* [n_atoms] HCT descreening scale */
    const double* sasa_radius;      /* [n_atoms] vdW + 1.4 Å probe */
    const double* sasa_gamma;       /* [n_atoms] surface tension, kJ/mol/Å^2 (unscaled) */
    const double* mass;             /* [n_atoms] Da */
    int32_t n_excl;  const int32_t* excl;    /* [n_excl*2] local i<j 1-2 / 1-3 pairs */
    int32_t n_14;    const int32_t* pair14;  /* [n_14*2]  local i<j 1-4 pairs */
    int32_t n_bonds; const int32_t* bonds; const double* bond_params;           /* (K kcal/mol/Å^2, r0 Å) */
    int32_t n_angles; const int32_t* angles; const double* angle_params;        /* (K kcal/mol/rad^2, theta0 rad) */
    int32_t n_dihedrals; const int32_t* dihedrals; const double* dihedral_params; /* one row per periodic term:
                                                                                  (k kcal/mol, n, delta rad) */
    int32_t n_impropers; const int32_t* impropers; const double* improper_params; /* (K, psi0 rad); atom 0 central */
    int32_t n_constraints; const int32_t* constraints; const double* constraint_len; /* X-H bonds, (i,j), r0 Å */
} og_system_desc;

/* Replaces ForceScratch::new (energy/src/scratch.rs:211-262): uploads
 * parameters, builds the tile-pair work list, exclusion masks and SHAKE
 * clusters once.  All index arrays are local to their system. */
int og_engine_create(const og_system_desc* systems, int32_t n_systems, og_engine** out);
void og_engine_destroy(og_engine* e);
const char* og_last_error(void);

/* info[0]=total atoms, [1]=systems, [2]=tiles, [3]=tile pairs, [4]=masked
 * tile pairs, [5]=clusters, [6]=sum over systems of N(N-1)/2 pairs,
 * [7]=SASA work atoms (gamma != 0). */
int og_engine_info(const og_engine* e, int64_t* info);

/* Host <-> device state, all systems concatenated in atom order. */
int og_set_positions(og_engine* e, const double* xyz);
int og_get_positions(og_engine* e, double* xyz);
int og_set_velocities(og_engine* e, const double* v);
int og_get_velocities(og_engine* e, double* v);
int og_get_forces(og_engine* e, double* f);

/* Replaces total_force_with_scratch (energy/src/forces.rs:78-128):
 * overwrites f_out (if not NULL) with bonded + LJ + Coulomb + GB (+ SASA)
 * forces at xyz (if NULL, at the device-resident positions).  energies
 * (if not NULL, needs OG_ENERGY) receives n_systems x OG_N_ENERGY terms.
 * cutoff_a is origami's cutoff_a argument (forces.rs:82): the LJ and
 * reaction-field Coulomb cutoff under OG_MODE_A (finite, > 0; origami's
 * callers pass DEFAULT_CUTOFF_A = 10); without OG_MODE_A every pair counts
 * and cutoff_a is not used.  sasa_gamma_scale replaces the
 * ORIGAMI_SASA_GAMMA_SCALE environment variable
 * (energy/src/powersasa/analytical.rs:477-480).  Returns OG_ERR_NONFINITE
 * (energies of the system set to NaN) when a system's contributions leave the
 * fixed-point range (OG_STATUS_RANGE), OG_ERR_ARG when a SASA atom exceeds
 * the arrangement's capacity. */
int og_forces(og_engine* e, const double* xyz, double* f_out, uint32_t flags, double cutoff_a,
              double sasa_gamma_scale, double* energies);

/* Replaces total_energy_with_options (dynamics/src/energy_eval.rs:20-34) and
 * remd::potential_energy (dynamics/src/remd.rs:529-539): per-system energy
 * terms; GB pair energy is uncut in both modes, as in gb_energy.  Computes
 * no forces: the engine's force array (the forces og_run's next step starts
 * from) is left as the last og_forces / og_run step made it. */
int og_energies(og_engine* e, const double* xyz, uint32_t flags, double cutoff_a, double* energies);

/* Effective Born radii and descreening integrals after the last force call
 * (gb.rs:203-342); either pointer may be NULL. */
int og_born_radii(og_engine* e, double* radii, double* integral);

typedef struct og_dynamics_params {
    double dt_fs;                   /* timestep, fs */
    double friction_ps_inv;         /* gamma, 1/ps */
    uint64_t seed;                  /* Philox key for the O-step noise */
    int32_t constrain_h_bonds;      /* SHAKE X-H (tol 1e-6 Å^2, 64 iterations) */
    uint32_t flags;                 /* force flags (OG_MODE_A, OG_SASA, OG_SURROGATE) */
    double sasa_gamma_scale;
    double cutoff_a;                /* LJ / Coulomb cutoff under OG_MODE_A (langevin.rs passes
                                       DEFAULT_CUTOFF_A = 10 Å); not used without it */
} og_dynamics_params;

/* Replaces run_langevin's set-up (dynamics/src/langevin.rs:138-199):
 * per-system bath temperatures, DOF correction for SHAKE, and the initial
 * force evaluation.  Velocities are whatever og_set_velocities left (zero
 * after create). */
int og_dynamics_setup(og_engine* e, const og_dynamics_params* p, const double* temperatures_k);

/* n_steps BAOAB steps for every system (langevin.rs:223-342 /
 * remd.rs:286-405).  ke_per_step (optional, n_steps x n_systems) receives
 * each step's kinetic energy in kJ/mol.  A system stops at the step where
 * its velocities turn non-finite (remd.rs:382-389), a force leaves the
 * fixed-point range or a SASA atom exceeds capacity; the others run all
 * n_steps.  Returns OG_ERR_NONFINITE if any system has stopped;
 * og_get_status gives each system's OG_STATUS_* bits. */
int og_run(og_engine* e, int32_t n_steps, double* ke_per_step);

/* og_run without device allocation.  No counterpart in origami: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same steps as
 * og_run, with the per-step kinetic-energy accumulators in a file-scope
 * device array of OG_NOALLOC_KE_MAX int64 entries (one per GPU, held until
 * the process ends) instead of a per-call allocation, and direct launches
 * (no captured graph is instantiated).  With ke_per_step, n_steps x
 * n_systems must not exceed OG_NOALLOC_KE_MAX (else OG_ERR_ARG).  Every
 * call on one GPU shares that array: calls that overlap on one GPU race. */
#define OG_NOALLOC_KE_MAX 65536
int og_run_noalloc(og_engine* e, int32_t n_steps, double* ke_per_step);

/* Batched L-BFGS energy minimisation on the device, every system of the
 * engine independently (lbfgs, dynamics/src/lbfgs.rs:61-254, with
 * backtracking, line_search.rs:52-101, and minimize.rs:27-38's defaults):
 * history ring, two-loop recursion, Armijo line search on FP64 energies,
 * steepest-descent retry after a failed search and stop after a second,
 * origami's convergence tests.  Forces are the engine's (FP32 pairs, int64
 * accumulation), energies its FP64 totals; flags OG_MODE_A (origami's
 * minimiser uses DEFAULT_CUTOFF_A = 10 Å, energy_eval.rs:20-50) and OG_SASA.
 * Positions are updated in place; results (optional) get one entry per
 * system.  The control state stays on the device; between phases the host
 * reads four counters. */
typedef struct og_minimize_params {
    int32_t max_steps;              /* 500 */
    int32_t history;                /* 10 (1..32) */
    double gradient_tol;            /* 1.0 kJ/mol/Å, |g|inf */
    double energy_tol;              /* 0.01 kJ/mol */
    double max_step_a;              /* 0.1 Å: caps alpha0 by max |p| */
    double alpha0;                  /* 1.0 */
    double contraction;             /* 0.5 */
    double c1;                      /* 1e-4 */
    double min_alpha;               /* 1e-12 */
    uint32_t flags;                 /* OG_MODE_A, OG_SASA */
    double cutoff_a;                /* LJ / Coulomb cutoff under OG_MODE_A (10 Å) */
    double sasa_gamma_scale;        /* SASA force scale under OG_SASA */
} og_minimize_params;

typedef struct og_minimize_result {
    int32_t steps;                  /* LbfgsResult.steps */
    int32_t converged;
    int32_t line_search_failures;
    int32_t stopped;                /* 1: both searches of a step failed */
    double initial_energy, final_energy, max_force;
} og_minimize_result;

int og_minimize(og_engine* e, const og_minimize_params* p, og_minimize_result* results);

/* The force pass's tile shape (no counterpart in origami): on systems above
 * four tiles a wave evaluates two i tiles (128 atoms) against a strip of up
 * to `set` j tiles of 64 atoms (1..8), keeping the i-side sums in registers
 * across the strip.  og_autotune times 2, 4 and 8 with the launch widths and
 * keeps the fastest; set > 0 sets it directly (e.g. to restore an autotuned
 * choice), get (optional) receives the current value.  Forces change only
 * by FP32 rounding (the per-strip partial sums), not in their pairs. */
int og_strip_tiles(og_engine* e, int32_t set, int32_t* get);

/* External cylindrical exit tunnel for dynamics (CylindricalTunnel of
 * dynamics/src/cotranslate.rs:105-155, the co-translational driver's
 * ExternalPotential): atoms with 0 <= (x - origin).direction <= length_a
 * whose distance from the axis exceeds radius_a feel k_confine (d - radius_a)
 * towards the axis, added to every dynamics force evaluation (og_run's steps
 * and og_dynamics_setup's initial forces; not og_forces / og_energies).
 * direction must be a unit vector; NULL removes the tunnel.  Set it before
 * og_dynamics_setup so that the first step's forces carry it. */
typedef struct og_tunnel {
    double origin[3];
    double direction[3];
    double radius_a;                /* 12 Å */
    double length_a;                /* 80 Å */
    double k_confine;               /* 50 kJ/mol/Å^2 */
} og_tunnel;
int og_set_tunnel(og_engine* e, const og_tunnel* t);

/* Trajectory frames (langevin.rs save_every and its frame callback,
 * :205-220, :332-341): with capacity > 0 and every > 0, og_run and
 * og_run_noalloc copy every system's positions into a device ring of
 * `capacity` frames after each step whose number is a multiple of `every`
 * (a device-to-device copy in order with the steps; no host round trip).
 * Set-up records the current positions as the first frame, as origami emits
 * step 0.  A run that would add more frames than the ring has free slots is
 * refused with OG_ERR_ARG before any step; og_frames_read frees them.
 * capacity = every = 0 switches frames off. */
int og_frames_setup(og_engine* e, int32_t capacity, int32_t every);

/* The oldest unread frames, up to max_frames: xyz (max_frames x N x 3, all
 * systems in engine atom order) and their step numbers; *n_frames receives
 * the count.  Either output may be NULL (the frames are still consumed). */
int og_frames_read(og_engine* e, double* xyz, uint32_t* steps, int32_t max_frames, int32_t* n_frames);

/* How og_run issues a step (no counterpart in origami; its step loop is
 * langevin.rs:195-367, run per replica by remd.rs:286):
 *   0 direct launches (five per step on small batches);
 *   1 one captured hipGraph per BAOAB step (all systems, every launch of the
 *     step), replayed, when profiling is off and no per-step kinetic energy
 *     is requested;
 *   2 the engine's choice (the default): direct launches, measured fastest
 *     on the REMD batch and the 64-chain batch;
 *   3 persistent replica steps: one launch per run of steps, one block of
 *     512 threads per system, for batches of systems of up to 512 atoms
 *     without SASA or the surrogate (else direct launches).
 * Every mode gives bit-identical trajectories. */
int og_use_graphs(og_engine* e, int32_t on);

/* Autotuner (no counterpart in origami): at the current positions, times
 * the force pass's strip length (og_strip_tiles: 2, 4, 8 j tiles; the pass
 * runs one wave per work item), the descreening pass's launch width (2..32
 * blocks of 256 threads per CU) and the SASA arrangement kernel's (2..8
 * blocks of 2 waves per CU), each candidate twice, and keeps the fastest.
 * With OG_SASA in flags the arrangement width, and then the descreening
 * width again, are timed on the whole force evaluation (SASA on its own
 * stream beside the pair kernels), with OG_SURROGATE (surrogate enabled) on
 * the evaluation with learned radii.  choice[2] (optional) receives the pair
 * and SASA block counts, best_ms[2] the force and descreening passes' time
 * alone and the best evaluation's.  Other than the strip length (FP32
 * rounding of the per-strip partial sums), results do not change:
 * accumulation is order-independent fixed point. */
int og_autotune(og_engine* e, uint32_t flags, int32_t* choice, double* best_ms);

/* Launch widths without timing (no counterpart in origami), e.g. to restore
 * an earlier og_autotune choice: pair_blocks and sasa_blocks as in choice[]
 * (<= 0 keeps the current value; clamped to the staging the engine holds).
 * sasa_passes: 0 = the arrangement runs in one pass holding 128 caps per atom
 * when that grid fits the LDS, else a pass holding 80 and a second pass over
 * the atoms with more; 1 or 2 forces the one- or two-pass form.  Results do
 * not depend on any of these. */
int og_set_launch(og_engine* e, int32_t pair_blocks, int32_t sasa_blocks, int32_t sasa_passes);

/* Set the bath temperature of one system (REMD ladder assignment). */
int og_set_temperature(og_engine* e, int32_t system, double temperature_k);

/* REMD exchange (remd.rs:447-466): swap positions, velocities and forces of
 * systems a and b and rescale velocities by sqrt(T_a/T_b) and sqrt(T_b/T_a).
 * OG_ERR_ARG unless the two systems have identical parameters and term
 * lists (checked against a hash taken at og_engine_create). */
int og_swap_systems(og_engine* e, int32_t a, int32_t b);

/* Per-system status (OG_STATUS_* bits; 0 = running), SHAKE failure counts,
 * Born-radius clamp counts. */
int og_get_status(og_engine* e, int32_t* diverged, int32_t* shake_failures, int32_t* clamped);

/* Kernel timing: when enabled, every launch is bracketed by HIP events on
 * the engine's stream.  og_profile_read returns, per kernel class
 * (OG_K_*), total milliseconds and launch counts since the last reset. */
enum og_kernel_class {
    OG_K_CLUSTER = 0, OG_K_BONDED, OG_K_DESCREEN, OG_K_RADII, OG_K_FORCE, OG_K_SASA, OG_K_BACK,
    OG_K_SURROGATE, OG_K_OTHER, OG_N_KCLASS
};
int og_profile_enable(og_engine* e, int32_t on);
int og_profile_read(og_engine* e, double* ms, int64_t* count);

/* ---- Learned Born-radius surrogate --------------------------------------
 * No counterpart in origami: it stands in for the descreening pass of
 * compute_born_radii_into_scratch (crates/energy/src/gb.rs:203-342) and is
 * trained on the radii that pass produces along the engine's trajectories.
 * Descriptor per atom: 5 element classes (H, C, N, O, S/P by mass) x 32
 * radial bins of neighbour counts over 0-12 Å (0.375 Å bins, /8) plus
 * (rho - 1.5)*4; 161 features.  MLP 161 -> 128 -> 128 -> 1 (SiLU), BF16
 * operands on MFMA with FP32 accumulation; R = 1.5 + 0.25 y.
 * Parameters (41345 floats): W1 [192][128] input-major (rows 161..191
 * zero), b1 [128], W2 [128][128] input-major, b2 [128], w3 [128], b3. */

/* Allocate the surrogate (training set of capacity_rows rows, weights drawn
 * from a seeded N(0, 1/fan_in)). */
int og_surrogate_setup(og_engine* e, int64_t capacity_rows, uint64_t seed);
/* Append one row per atom: descriptor and plain Born radius at the current
 * positions (OG_MODE_A in flags selects the 20 Å descreening cutoff). */
int og_surrogate_collect(og_engine* e, uint32_t flags, int64_t* rows_total);
/* n_steps Adam steps (beta 0.9 / 0.999) on batches sampled from the
 * collected rows; loss_per_step (optional) receives the mean squared error
 * in units of (0.25 Å)^2.  batch: multiple of 32, at most 8192. */
int og_surrogate_train(og_engine* e, int32_t n_steps, int32_t batch, double lr, uint64_t seed,
                       float* loss_per_step);
/* Use surrogate radii for every system under OG_SURROGATE.  Every
 * check_every force evaluations (0 = never) the plain radii are computed as
 * well and a system whose max |R_sur - R_plain| / R_plain exceeds tolerance
 * returns to plain GB from the next evaluation on. */
int og_surrogate_enable(og_engine* e, int32_t check_every, double tolerance);
/* Surrogate radii of every atom at xyz (or the current positions if NULL). */
int og_surrogate_radii(og_engine* e, const double* xyz, double* radii);
/* Descriptors of every atom at xyz (or current), n_atoms x 161 floats. */
int og_surrogate_descriptors(og_engine* e, const double* xyz, float* x);
/* Read (get) and/or replace (set) the FP32 parameters. */
int og_surrogate_params(og_engine* e, float* get, const float* set);
/* Gradient of the mean squared loss over an explicit batch (x: n x 161,
 * target radii in Å; n a multiple of 32, at most 8192) with the GPU's own
 * forward and backward kernels; no update. */
int og_surrogate_gradient(og_engine* e, const float* x, const double* r_target, int32_t n, float* grad,
                          double* loss);
/* Per system: 1 while it uses surrogate radii, and the last check's max
 * relative deviation from the plain radii. */
int og_surrogate_status(og_engine* e, int32_t* surrogate_mode, float* deviation);

#ifdef __cplusplus
}
#endif
#endif
