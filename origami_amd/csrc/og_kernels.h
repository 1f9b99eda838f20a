// Host-visible kernel launchers and argument blocks shared by the engine
// (engine.cpp) and the kernel translation units (*.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct og_work;
struct og_strip;

struct og_pair_args {
    const double* pos;               // [N*3] FP64 master positions (atom index space)
    const float4* par;               // [P] (se, rm, se14, rm14) per padded slot
    float4* gb;                      // [P] (rho~, S = scale*rho~, R, sqrt(log2e/4)/R)
    const float* qs;                 // [P] charge * sqrt(k_C * 4.184)
    const int* tile_atom0;           // [T] first atom index of each tile
    const int* tile_cnt;             // [T] atoms in each tile (1..64)
    const og_work* work;             // [n_work] tile pairs
    int n_work;
    const unsigned long long* masks; // [n_masked*128]: lanes' 1-2/1-3 bits then 1-4 bits
    long long* facc;                 // [N*3] force accumulator (fixed point)
    long long* iacc;                 // [N]   descreening integral accumulator
    long long* eacc;                 // [S*OG_NE] energy accumulators
    float rc2_nb, rc2_gb, rc2_ds;    // squared cutoffs (Mode A)
    float inv_rc3, kappa, crf;       // reaction-field constants (Mode A)
    const int* skip_sys;             // [S] descreening skipped where nonzero (may be null)
    int grid;                        // blocks of 256 threads (0: default)
    int* sflag;                      // [S] status bits (OG_ST_RANGE on a contribution out of range)
    // force pass only: the strips of large systems (taken first), then the
    // tile pairs that no strip covers (fwork: diagonal, adjacent and masked
    // tile pairs, and every tile pair of small systems)
    const og_strip* strips;
    int n_strip;
    const og_work* fwork;
    int n_fwork;
    // descreening: tile pairs whose atoms are all at least ds_far apart take
    // the far-field series of the HCT term (ds_far >= 10 S and >= 1.25
    // (rho~ + S) for every atom; 0 switches the series off)
    float ds_far;
};

// energy slots per system
enum {
    OG_E_BOND = 0, OG_E_ANGLE, OG_E_DIHEDRAL, OG_E_IMPROPER, OG_E_LJ, OG_E_COUL,
    OG_E_GB_SELF, OG_E_GB_PAIR, OG_E_SASA, OG_E_KIN, OG_NE
};

// The O step's three normals per atom for the step *step_dev + ahead, written
// into half (step & 1) of noise [2][3N]; they depend only on the key, the
// system, the atom and the step, so the next step's are formed while this
// step's forces are computed.
struct og_noise_args {
    uint32_t key0, key1;
    const int* atom_sys;
    const int* atom_local;
    int n_atoms;
    const uint32_t* step_dev;
    uint32_t ahead;
    double* noise;
};
void og_launch_noise(const og_noise_args& a, hipStream_t s);

struct og_bonded_args {
    const double* pos;
    long long* facc;
    long long* eacc;
    const int* atom_sys;
    int nb, na, nd, ni;
    const int* bonds; const double* bond_p;       // (K, r0)
    const int* angles; const double* angle_p;     // (K, θ0)
    const int* dih; const double* dih_p;          // (k, n, δ)
    const int* imp; const double* imp_p;          // (K, ψ0)
    // per-(term, atom) force slots, gathered per atom in slot order:
    // bonds 2 per term, then angles 3, dihedrals 4, impropers 4
    double* tf;                                   // [slots][3]
    const int* csr_start;                         // [N+1]
    const int* csr_slot;                          // [slots]
    int n_atoms;
    og_noise_args noise;                          // if noise.noise is set, formed beside the gather
    int* sflag;                                   // [S] status bits
};

struct og_radii_args {
    const long long* iacc;
    const int* atom_sys;
    const int* atom_slot;            // atom → padded slot
    const double* rho;               // intrinsic radius ρ
    const double* charge;
    float4* gb;
    double* radii;                   // [N] FP64 effective radii
    int* clamped;                    // [S]
    long long* eacc;
    int n_atoms;
    const int* sur_mode;             // [S] nonzero: take R from r_sur (may be null)
    const double* r_sur;             // [N] surrogate radii
    unsigned* dev;                   // [S] max |R_sur - R_plain| / R_plain (float bits), when check
    int check;                       // plain radii were computed for every system this call
    int* sflag;                      // [S] status bits
};

#define OG_MAX_CL 8                  // atoms per SHAKE cluster (X + up to 7 H)
#define OG_SMALL_CL 4                // the narrow cluster kernel's limit (CH3, NH3+)

struct og_cluster_args {
    double* pos; double* vel; double* ref; const double* force;
    long long* facc; long long* iacc;
    const double* mass;
    const int* atom_sys;
    const int* atom_local;           // index within its system (RNG counter)
    const int* cl_start;             // [C+1] atoms of cluster c: cl_atoms[cl_start[c]..],
    const int* cl_atoms;             //   heavy atom first, then the hydrogen of each constraint
    const int* cn_start;             // [C+1] constraints of cluster c (constraint m: atoms 0, m + 1)
    const double* cn_len;            // [K]
    int n_clusters;
    int n_small;                     // clusters [0, n_small) hold at most OG_SMALL_CL atoms
    const double* sys_temp;          // [S]
    const double* sys_dofcorr;       // [S]
    int* shake_fail;                 // [S]
    const int* diverged;             // [S]
    double dt, alpha;
    int use_shake;
    uint32_t key0, key1;
    uint32_t step;
    const uint32_t* step_dev;        // if set, the step counter is read from device memory
    const double* noise;             // if set, the O step's normals: [2][3N], step s in half s & 1
    int n_atoms;
};

// External cylindrical exit tunnel on every atom (the cotranslate driver's
// ExternalPotential, cotranslate.rs:105-155): radial pushback -k (d - R)
// beyond radius R for atoms within the axial extent.
struct og_tunnel_dev {
    double ox, oy, oz, dx, dy, dz;  // axis origin and unit direction
    double radius, length, k;
    int on;
};

struct og_back_args {
    const double* pos;               // read for the external tunnel only
    double* vel; double* force;
    const long long* facc;
    const double* mass;
    const int* atom_sys;
    int n_atoms;
    double dt;
    int do_kick;
    long long* ke_acc;               // [S] or null
    int* diverged;                   // [S]
    uint32_t* step_dev;              // if set (and do_kick), advanced by one per step
    int* sflag;                      // [S] status bits of this evaluation: with do_kick, a flagged
                                     //   system gets them in diverged and is not kicked
    og_tunnel_dev tun;               // added to every atom's force when tun.on
};

void og_launch_pair_descreen(const og_pair_args& a, bool mode_a, hipStream_t s);
void og_launch_pair_force(const og_pair_args& a, bool mode_a, hipStream_t s);

struct og_energy64_args {            // FP64 per-atom data for the energy pass
    const double* charge;
    const double* eps; const double* rmin;      // kcal/mol, Å
    const double* eps14; const double* rmin14;
    const double* radii;                        // effective Born radii
    double rc2_nb, kappa, crf;                  // Mode A cutoff and reaction field
};
void og_launch_pair_energy64(const og_pair_args& a, const og_energy64_args& e, bool mode_a,
                             hipStream_t s);
void og_launch_bonded(const og_bonded_args& a, bool energy, hipStream_t s);
#define OG_GATHER_LDS_MAX_ATOMS 20000   // batches below this size gather through LDS
// the descreening pass and the bonded terms (forces) in one launch; batches
// below OG_GATHER_LDS_MAX_ATOMS
void og_launch_descreen_bonded(const og_pair_args& a, const og_bonded_args& b, bool mode_a, hipStream_t s);
// the bonded terms alone (forces), the gather left to og_launch_radii_gather
void og_launch_bonded_terms(const og_bonded_args& a, hipStream_t s);
// Born radii (forces only, no surrogate) and the bonded gather's LDS form
// in one launch; batches below 20 000 atoms
void og_launch_radii_gather(const og_radii_args& r, const og_bonded_args& b, hipStream_t s);
void og_launch_radii(const og_radii_args& a, bool energy, hipStream_t s);
void og_launch_cluster_step(const og_cluster_args& a, hipStream_t s);
void og_launch_zero_acc(long long* facc, long long* iacc, int n_atoms, hipStream_t s);
void og_launch_back(const og_back_args& a, hipStream_t s);
void og_launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s);
// p[0..n) = 0 with a kernel (no memset node when a caller captures the stream)
void og_launch_zero_i64(long long* p, long long n, hipStream_t s);
void og_launch_swap(double* pos, double* vel, double* force, int off_a, int off_b, int n,
                    double scale_a, double scale_b, hipStream_t s);

// ---------------------------------------------------------------------------
// Device-resident batched L-BFGS (lbfgs.hip); one block per system.
// ---------------------------------------------------------------------------
#define OG_LB_MAX_HISTORY 32
enum { OG_LB_ACTIVE = 0, OG_LB_SEARCHING, OG_LB_ACCEPTED, OG_LB_FAILED, OG_LB_CONVERGED, OG_LB_STOPPED };
enum { OG_LB_K_INIT = 0, OG_LB_K_BEGIN, OG_LB_K_TRIAL, OG_LB_K_ACCEPT, OG_LB_K_COUNT, OG_LB_K_FALLBACK,
       OG_LB_K_UPDATE };

struct og_lb_sys {                 // one system's minimiser state
    int state, phase;              // OG_LB_*; 0 = L-BFGS direction, 1 = steepest-descent retry
    int count, head;               // history entries and the ring slot of the oldest
    int steps, failures;           // origami's LbfgsResult fields
    int pad0, pad1;
    double alpha, gdotp;           // current step length and p.g
    double e_now, e_prev, e_new;   // energies at x0, at the last accepted point, at the trial
    double e_init, max_force;
};

struct og_lbfgs_args {
    int n_sys, n_atoms, history;
    const int* atom_off;           // [S+1]
    double* pos;                   // engine positions (trial points are written here)
    const double* force;           // engine forces (FP64, after the last force evaluation)
    const long long* eacc;         // [S*OG_NE] energies of the last evaluation
    double *x0, *g, *p;            // [3N] each
    double *hs, *hy;               // [history][3N] each
    double* rho;                   // [S][OG_LB_MAX_HISTORY]
    og_lb_sys* sys;                // [S]
    int* skip;                     // [S] 1 = leave out of the next energy evaluation
    int* counters;                 // [4] searching, failed, active, accepted
    double gradient_tol, energy_tol, max_step, alpha0, contraction, c1, min_alpha;
};
void og_launch_lbfgs(int phase, const og_lbfgs_args& a, int step, hipStream_t s);

// A system's total energy from its fixed-point terms, summed as the host sums
// them: bonded four, LJ, Coulomb, GB (self + pair), SASA.
__host__ __device__ inline double og_lb_total_energy(const long long* acc) {
#pragma clang fp contract(off)
    const double k = 1.0 / 4294967296.0;
    const double gb = (double)acc[OG_E_GB_SELF] * k + (double)acc[OG_E_GB_PAIR] * k;
    return (double)acc[OG_E_BOND] * k + (double)acc[OG_E_ANGLE] * k + (double)acc[OG_E_DIHEDRAL] * k +
           (double)acc[OG_E_IMPROPER] * k + (double)acc[OG_E_LJ] * k + (double)acc[OG_E_COUL] * k + gb +
           (double)acc[OG_E_SASA] * k;
}

// ---------------------------------------------------------------------------
// Persistent replica steps (pair_kernels.hip): one block per system runs
// n_steps whole BAOAB steps (normals, cluster B-A-SHAKE-O-A,
// bonded terms, descreening, radii + bonded gather, pair forces, closing
// kick) with block barriers between the phases; systems share nothing, so
// no grid-wide synchronisation is needed.  Per-system index ranges below.
// ---------------------------------------------------------------------------
struct og_persist_args {
    int n_steps;
    uint32_t step0;                  // Philox step counter of the first step
    int n_sys;
    const int* atom_off;             // [S+1]
    const int* cl_small_off;         // [S+1] into cl_small (clusters of <= OG_SMALL_CL atoms; the
    const int* cl_small;             //   kernel runs batches that have no larger ones)
    const int* cl_big_off;           // [S+1] into cl_big
    const int* cl_big;
    const int* term_off;             // [4][S+1]: bonds, angles, dihedrals, impropers
    const int* work_off;             // [S+1] descreening items (og_pair_args::work)
    const int* fwork_off;            // [S+1] force items (og_pair_args::fwork)
    const int* strip_off;            // [S+1] strips (og_pair_args::strips)
    long long* ke_acc;               // [n_steps][S] or null
};
void og_launch_replica_steps(const og_pair_args& pa, const og_bonded_args& ba, const og_radii_args& ra,
                             const og_cluster_args& ca, const og_back_args& bk, const og_noise_args& nz,
                             const og_persist_args& pp, bool mode_a, hipStream_t s);

