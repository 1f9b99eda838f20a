// Analytical hydrophobic-SASA forces (PowerSasa, Klenin et al. 2011 spherical
// Gauss-Bonnet), one thread per atom with nonzero surface tension, FP64.
//
// Per atom i: overlapping neighbours -> buried caps on i's unit sphere ->
// accessible-region boundary (arcs on cap circles joined at vertices) ->
// forward-mode derivatives of A_i / R_i^2 = 2*pi*chi - sum_arcs cos(a) theta
// - sum_vertices eps, scattered as F_x -= gamma_i * dA_i/dr_x.
//
// Restates origami:
//   caps / boundary      energy/src/powersasa/arrangement.rs:69-289
//   circle intersection  energy/src/powersasa/geometry.rs:37-100
//   vertex sign / JVPs   energy/src/powersasa/derivatives.rs:92-545
//   cache + gradient     energy/src/powersasa/analytical.rs:101-214, 285-421, 583-649
// Topology decisions use origami's FP64 tolerances unchanged (1e-12, 1e-9,
// 1e-10), which is why this kernel stays in FP64.
//
// Gradient evaluation order: origami evaluates, for every affected atom x and
// axis, a sum over all arcs and vertices with only x perturbed.  Terms whose
// arc/vertex does not involve x are exactly zero (every JVP is linear in the
// perturbation), so this kernel visits each arc/vertex once and evaluates it
// for the (at most four) distinct atoms it involves, perturbing every role
// that atom plays together; per-(x, axis) arc and vertex sums are kept apart
// and combined as -sum_arc - sum_vert, so each sum receives the same nonzero
// terms in the same order as origami's.
//
// Neighbour candidates: all atoms of the same system within r_i + r_j, in
// ascending index order (analytical.rs:616-625).  Per-thread working arrays
// live in global scratch interleaved across threads (element k of thread t at
// k*stride + t) so that lanes walking the same k touch consecutive addresses.
#include <string>

#include "og_device.h"
#include "og_sasa.h"
#include "sasa_core.h"

static std::string g_sasa_err;
const char* og_sasa_error() { return g_sasa_err.c_str(); }

__global__ void __launch_bounds__(64) og_k_sasa_forces(og_sasa_data d, const double* pos,
                                                       long long* facc, double gamma_scale) {
    using namespace og_sasa;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.n_work) return;
    const Scr S{d.scratch, d.iscratch, d.stride, t};
    const int ia = d.work[t];
    int ovf = 0;
    atom_forces(S, ia, pos, d.radius, d.sys_begin[t], d.sys_end[t], d.gamma[ia] * gamma_scale, &ovf,
                [&](int atom, int axis, double f) {
                    og_atomic_add_fix(facc + 3 * (size_t)atom + axis, og_to_fix(f));
                });
    if (ovf) atomicAdd(d.overflow, 1);
}

int og_sasa_build(og_sasa_data* d, int n_atoms, const std::vector<int>& atom_sys,
                  const std::vector<int>& atom_off, const std::vector<double>& radius,
                  const std::vector<double>& gamma) {
    std::vector<int> work, b, e;
    for (int i = 0; i < n_atoms; ++i) {
        if (gamma[i] == 0.0) continue;
        work.push_back(i);
        b.push_back(atom_off[atom_sys[i]]);
        e.push_back(atom_off[atom_sys[i] + 1]);
    }
    d->n_work = (int)work.size();
    d->stride = ((d->n_work + 63) / 64) * 64;
    auto up = [](auto** dst, const auto& v) -> bool {
        using T = typename std::remove_reference<decltype(v)>::type::value_type;
        size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
        if (hipMalloc((void**)dst, bytes) != hipSuccess) return false;
        if (!v.empty() && hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return false;
        return true;
    };
    if (!up(&d->work, work) || !up(&d->sys_begin, b) || !up(&d->sys_end, e) ||
        !up(&d->atom_sys, atom_sys) || !up(&d->radius, radius) || !up(&d->gamma, gamma)) {
        g_sasa_err = "device allocation failed";
        return -2;
    }
    if (d->n_work > 0) {
        if (hipMalloc((void**)&d->scratch, (size_t)S_TOTAL * d->stride * sizeof(double)) != hipSuccess ||
            hipMalloc((void**)&d->iscratch, (size_t)I_TOTAL * d->stride * sizeof(int)) != hipSuccess) {
            g_sasa_err = "SASA scratch allocation failed";
            return -2;
        }
    }
    if (hipMalloc((void**)&d->overflow, sizeof(int)) != hipSuccess ||
        hipMemset(d->overflow, 0, sizeof(int)) != hipSuccess) {
        g_sasa_err = "device allocation failed";
        return -2;
    }
    return 0;
}

void og_sasa_free(og_sasa_data* d) {
    void* p[] = {d->work, d->sys_begin, d->sys_end, d->atom_sys, d->radius, d->gamma,
                 d->scratch, d->iscratch, d->overflow};
    for (void* q : p)
        if (q) hipFree(q);
    *d = og_sasa_data{};
}

int og_sasa_launch(og_sasa_data* d, const double* pos, long long* facc, long long* eacc,
                   double gamma_scale, bool energy, hipStream_t s) {
    (void)eacc;
    (void)energy;
    if (d->n_work == 0) return 0;
    hipLaunchKernelGGL(og_k_sasa_forces, dim3((d->n_work + 63) / 64), dim3(64), 0, s, *d, pos, facc,
                       gamma_scale);
    if (hipGetLastError() != hipSuccess) {
        g_sasa_err = "kernel launch failed";
        return -2;
    }
    return 0;
}

int og_sasa_overflow(og_sasa_data* d, int* count) {
    if (hipMemcpy(count, d->overflow, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return -2;
    return 0;
}
