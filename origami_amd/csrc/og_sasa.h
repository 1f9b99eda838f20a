// Analytical SASA force kernels (PowerSasa) — host interface.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

// Pool capacity: arcs (and, separately, vertices) per SASA atom on average
// over a batch.  Protein atoms carry ~5-20; exceeding it is reported as an
// error, never as silently missing forces.
#define OG_SASA_POOL_ARCS 64
// probe points of the accessible-component count (arrangement.rs:327-389)
#define OG_SASA_PROBES 500

struct og_sasa_data {
    int n_work = 0;               // atoms with gamma != 0
    int n_cu = 0;
    int blocks = 0;               // launch width of the arrangement kernel
    bool one_pass = false;        // arrangement in one pass of MAXC caps (no second pass)
    int passes = 0;               // 0: one pass when the blocks fit its LDS; 1 or 2: forced
    long long lds_cu = 0;         // LDS bytes per CU
    int stride = 0;               // waves launched (= per-wave staging slots)
    int n_tiles = 0;
    double max_radius = 0.0;      // largest SASA radius (neighbour culling)
    int* work = nullptr;          // [n_work] global atom index
    int* tile_begin = nullptr;    // [n_work] first tile of the work atom's system
    int* tile_end = nullptr;      // [n_work] one past its last tile
    const int* tile_atom0 = nullptr;  // [n_tiles] (engine-owned)
    const int* tile_cnt = nullptr;    // [n_tiles] (engine-owned)
    double4* tile_bound = nullptr;    // [n_tiles] bounding sphere, refreshed per launch
    double* radius = nullptr;     // [N] vdW + probe
    double* gamma = nullptr;      // [N] kJ/mol/Å^2 (unscaled)
    double* scratch = nullptr;    // per-wave FP64 staging slots
    int* iscratch = nullptr;      // per-wave int staging slots
    int* overflow = nullptr;      // [1] count of atoms whose arrangement exceeded capacity
    int* counters = nullptr;      // pool fill: arcs, vertices, cap owners [0..2]; deferred
                                  //     atoms [3]; the arrangement's work queues [64 (g + 1)]
    int* deferred = nullptr;      // [n_work] atoms with more caps than the main pass holds
    int4* rec = nullptr;          // [n_work] (arc_off, vtx_off, owner_off, na | nv<<9 | nc<<18)
    // pool
    int arc_cap = 0, vtx_cap = 0, own_cap = 0;
    int4* arc_atoms = nullptr;    // (i, k, l, m) atoms; l = m = -1 for arcs without vertices
    double4* arc_par = nullptr;   // (theta, start sign, end sign, -)
    unsigned* arc_code = nullptr; // local roles k, l, m as bytes (1 + cap; 0xFF absent)
    double* arc_terms = nullptr;  // [12 per arc] role x axis
    int4* vtx_atoms = nullptr;    // (i, in, out, -)
    double* vtx_sign = nullptr;
    unsigned* vtx_code = nullptr;
    double* vtx_terms = nullptr;  // [9 per vertex]
    int* owners = nullptr;        // cap owners per atom
    // accessible area (energy evaluations only)
    double4* probes = nullptr;    // [OG_SASA_PROBES] golden-spiral points
    int* padj_start = nullptr;    // [OG_SASA_PROBES + 1] probes within 2.5 spacings
    int* padj = nullptr;
    long long* eacc = nullptr;    // engine energy accumulators (set per launch)
    const int* atom_sys = nullptr;    // [N] system of each atom (engine-owned, set per launch)
    int* sflag = nullptr;         // [S] per-system status bits (engine-owned, set per launch):
                                  //     OG_ST_SASA on an atom over capacity, OG_ST_RANGE on a
                                  //     force or energy out of the fixed-point range
};

// sys_tile_off[s] .. sys_tile_off[s+1] are system s's tiles (64-atom, never
// straddling systems); tile_atom0/tile_cnt are device arrays owned by the caller.
int og_sasa_build(og_sasa_data* d, int n_atoms, const std::vector<int>& atom_sys,
                  const std::vector<int>& atom_off, const std::vector<double>& radius,
                  const std::vector<double>& gamma, const std::vector<int>& sys_tile_off,
                  const int* tile_atom0_dev, const int* tile_cnt_dev, int n_tiles);
void og_sasa_free(og_sasa_data* d);
// eacc != null: also accumulate gamma_i * A_i (unscaled gamma) into each
// system's SASA energy slot (powersasa_energy, mod.rs:55-111).
// atom_sys and sflag (engine-owned) are required.
int og_sasa_launch(og_sasa_data* d, const double* pos, long long* facc, double gamma_scale,
                   hipStream_t s, long long* eacc, const int* atom_sys, int* sflag);
// Shrake-Rupley energy (sasa.rs:113-178): gamma_i * A_i of every SASA atom into
// each system's SASA energy slot, from n_dots unit-sphere dots (device, 3 per dot).
int og_sasa_sr_launch(og_sasa_data* d, const double* pos, const double* dots, int n_dots, hipStream_t s,
                      long long* eacc, const int* atom_sys, int* sflag);
// Arrangement-kernel width (blocks of 2 waves); at most the staging slots allocated.
int og_sasa_max_blocks(const og_sasa_data* d);
void og_sasa_set_blocks(og_sasa_data* d, int blocks);
// 0: one pass when the grid fits the LDS of MAXC caps, else two; 1 or 2: forced
void og_sasa_set_passes(og_sasa_data* d, int passes);
// The JVP-terms kernel of a force evaluation (sasa_terms.hip).
void og_launch_sasa_terms(const og_sasa_data& d, const double* pos, hipStream_t s);
// Reads and clears the overflow count.
int og_sasa_overflow(og_sasa_data* d, int* count);
const char* og_sasa_error();
