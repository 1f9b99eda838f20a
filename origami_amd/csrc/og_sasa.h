// Analytical SASA force kernel (PowerSasa) — host interface.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

struct og_sasa_data {
    int n_work = 0;            // atoms with gamma != 0
    int stride = 0;            // threads launched (scratch interleave stride)
    int* work = nullptr;       // [n_work] global atom index
    int* sys_begin = nullptr;  // [n_work] first atom of the work atom's system
    int* sys_end = nullptr;    // [n_work] one past its last atom
    int* atom_sys = nullptr;   // [N]
    double* radius = nullptr;  // [N] vdW + probe
    double* gamma = nullptr;   // [N] kJ/mol/Å^2 (unscaled)
    double* scratch = nullptr; // interleaved per-thread FP64 scratch
    int* iscratch = nullptr;   // interleaved per-thread int scratch
    int* overflow = nullptr;   // [1] count of atoms whose arrangement exceeded capacity
};

int og_sasa_build(og_sasa_data* d, int n_atoms, const std::vector<int>& atom_sys,
                  const std::vector<int>& atom_off, const std::vector<double>& radius,
                  const std::vector<double>& gamma);
void og_sasa_free(og_sasa_data* d);
int og_sasa_launch(og_sasa_data* d, const double* pos, long long* facc, long long* eacc,
                   double gamma_scale, bool energy, hipStream_t s);
int og_sasa_overflow(og_sasa_data* d, int* count);
const char* og_sasa_error();
