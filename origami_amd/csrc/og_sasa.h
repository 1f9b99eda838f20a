// Analytical SASA force kernel (PowerSasa) — host interface.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

struct og_sasa_data {
    int n_work = 0;               // atoms with gamma != 0
    int blocks = 0;               // launch width of the wave-per-atom kernel
    int stride = 0;               // waves launched (= per-wave scratch slots)
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
    double* scratch = nullptr;    // per-wave FP64 scratch slots
    int* iscratch = nullptr;      // per-wave int scratch slots
    int* overflow = nullptr;      // [1] count of atoms whose arrangement exceeded capacity
};

// sys_tile_off[s] .. sys_tile_off[s+1] are system s's tiles (64-atom, never
// straddling systems); tile_atom0/tile_cnt are device arrays owned by the caller.
int og_sasa_build(og_sasa_data* d, int n_atoms, const std::vector<int>& atom_sys,
                  const std::vector<int>& atom_off, const std::vector<double>& radius,
                  const std::vector<double>& gamma, const std::vector<int>& sys_tile_off,
                  const int* tile_atom0_dev, const int* tile_cnt_dev, int n_tiles);
void og_sasa_free(og_sasa_data* d);
int og_sasa_launch(og_sasa_data* d, const double* pos, long long* facc, double gamma_scale,
                   hipStream_t s);
int og_sasa_overflow(og_sasa_data* d, int* count);
const char* og_sasa_error();
