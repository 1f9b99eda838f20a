// O(N) kernels in FP64: bonded forces/energies, the Born-radius transform,
// the fused BAOAB front half with SHAKE, and the force finalisation / second
// half-kick.  These are HBM/latency-bound; they keep origami's FP64 state.
//
//   bonded terms     energy/src/forces_bonded.rs:47-236, bonded.rs:59-191
//   Born radii       energy/src/gb.rs:327-341 (tanh transform + clamp)
//   BAOAB            dynamics/src/langevin.rs:223-313
//   SHAKE            dynamics/src/shake.rs:86-125
//   kinetic energy   dynamics/src/langevin.rs:469-475
#include "og_device.h"
#include "og_kernels.h"

#define OG_PI 3.14159265358979323846

#include "bonded_core.h"
#include "md_core.h"


__global__ void __launch_bounds__(256) og_k_bonded(og_bonded_args a) {
    bonded_body<false>(a, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) og_k_bonded_e(og_bonded_args a) {
    bonded_body<true>(a, blockIdx.x, gridDim.x);
}

// Per atom: sum its (term, atom) force slots in slot order (bonds, angles,
// dihedrals, impropers, each in list order, as origami's term loops add
// them) in FP64, then one fixed-point add per component.  A block takes 64
// consecutive atoms, whose slots are one contiguous CSR range: all 256
// threads first copy the slot forces of that range into LDS (independent
// loads, all in flight together, where a thread walking its own atom's
// slots would wait on each one), then wave 0 sums each atom's slots from
// LDS in slot order.  Ranges longer than the LDS copy go in chunks.
// NOISE: wave 1 meanwhile forms the block's atoms' O-step normals for the
// next step.
#define OG_GATHER_ATOMS 64
#define OG_GATHER_CHUNK 2048                // 8 entries per thread of a 256-thread block
template <bool NOISE>
__device__ __forceinline__ void gather_lds_body(const og_bonded_args& a, int blk) {
    __shared__ double s_t[3 * OG_GATHER_CHUNK];
    const int a0 = blk * OG_GATHER_ATOMS;
    const int a1 = min(a0 + OG_GATHER_ATOMS, a.n_atoms);
    const int t = threadIdx.x;
    const int i = a0 + t;
    const bool own = t < OG_GATHER_ATOMS && i < a1;
    const int q_begin = a.csr_start[a0], q_end = a.csr_start[a1];
    int q = own ? a.csr_start[i] : 0;
    const int qe = own ? a.csr_start[i + 1] : 0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int c0 = q_begin; c0 < q_end; c0 += OG_GATHER_CHUNK) {
        const int c1 = min(c0 + OG_GATHER_CHUNK, q_end);
        __syncthreads();                                  // the previous chunk is summed
        // OG_GATHER_CHUNK / 256 = 8 entries per thread, every load issued
        // before the first LDS store
        size_t sl[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = c0 + t + 256 * u;
            sl[u] = k < c1 ? 3 * (size_t)a.csr_slot[k] : 0;
        }
        double v[24];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (c0 + t + 256 * u < c1) {
                v[3 * u] = a.tf[sl[u]];
                v[3 * u + 1] = a.tf[sl[u] + 1];
                v[3 * u + 2] = a.tf[sl[u] + 2];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = t + 256 * u;
            if (c0 + e < c1) {
                s_t[3 * e] = v[3 * u];
                s_t[3 * e + 1] = v[3 * u + 1];
                s_t[3 * e + 2] = v[3 * u + 2];
            }
        }
        __syncthreads();
        for (; q < qe && q < c1; ++q) {
            const double* d = s_t + 3 * (q - c0);
            fx += d[0];
            fy += d[1];
            fz += d[2];
        }
    }
    if (NOISE && t >= OG_GATHER_ATOMS && t < 2 * OG_GATHER_ATOMS && a0 + t - OG_GATHER_ATOMS < a1)
        noise_atom(a.noise, a0 + t - OG_GATHER_ATOMS);
    if (own) {
        int* fl = a.sflag ? a.sflag + a.atom_sys[i] : nullptr;
        og_atomic_add_fix(a.facc + 3 * (size_t)i + 0, og_to_fix_chk(fx, fl));
        og_atomic_add_fix(a.facc + 3 * (size_t)i + 1, og_to_fix_chk(fy, fl));
        og_atomic_add_fix(a.facc + 3 * (size_t)i + 2, og_to_fix_chk(fz, fl));
    }
}
template <bool NOISE>
__global__ void __launch_bounds__(256) og_k_bonded_gather_lds(og_bonded_args a) {
    gather_lds_body<NOISE>(a, blockIdx.x);
}

// The same sums, a thread per atom for large batches (where the LDS copy
// would take LDS from the SASA kernels running beside it): the slots are
// fetched eight at a time (indices, then their forces) so that the loads of
// a batch are in flight together; the adds keep slot order.  NOISE: the
// thread also forms its atom's O-step normals for the next step (ALU work
// beside the gather's load latency).
template <bool NOISE>
__global__ void __launch_bounds__(256) og_k_bonded_gather(og_bonded_args a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_atoms) return;
    if (NOISE) noise_atom(a.noise, i);
    gather_atom(a, i);
}

static int bonded_grid(int total) {
    const int g = (total + 255) / 256;
    return g > 4096 ? 4096 : g;
}

void og_launch_bonded_terms(const og_bonded_args& a, hipStream_t s) {
    const int total = a.nb + a.na + a.nd + a.ni;
    if (total > 0) hipLaunchKernelGGL(og_k_bonded, dim3(bonded_grid(total)), dim3(256), 0, s, a);
}

void og_launch_bonded(const og_bonded_args& a, bool energy, hipStream_t s) {
    const int total = a.nb + a.na + a.nd + a.ni;
    if (total <= 0) {
        if (a.noise.noise) og_launch_noise(a.noise, s);
        return;
    }
    if (energy) hipLaunchKernelGGL(og_k_bonded_e, dim3(bonded_grid(total)), dim3(256), 0, s, a);
    else og_launch_bonded_terms(a, s);
    if (a.n_atoms <= 0) return;
    if (a.n_atoms < OG_GATHER_LDS_MAX_ATOMS) {
        const dim3 gg((a.n_atoms + OG_GATHER_ATOMS - 1) / OG_GATHER_ATOMS);
        if (a.noise.noise) hipLaunchKernelGGL(og_k_bonded_gather_lds<true>, gg, dim3(256), 0, s, a);
        else hipLaunchKernelGGL(og_k_bonded_gather_lds<false>, gg, dim3(256), 0, s, a);
    } else {
        const dim3 gg((a.n_atoms + 255) / 256);
        if (a.noise.noise) hipLaunchKernelGGL(og_k_bonded_gather<true>, gg, dim3(256), 0, s, a);
        else hipLaunchKernelGGL(og_k_bonded_gather<false>, gg, dim3(256), 0, s, a);
    }
}

template <bool ENERGY>
__global__ void __launch_bounds__(256) og_k_radii(og_radii_args a) {
    radii_body<ENERGY>(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// Born radii and the bonded gather in one launch, for small batches whose
// force evaluation is a chain of latency-bound launches: the two are
// independent (the gather needs only the bonded terms, the radii only the
// descreening), so blocks [0, nr) take the radii and the rest the gather's
// LDS form, side by side.
template <bool NOISE>
__global__ void __launch_bounds__(256) og_k_radii_gather(og_radii_args r, og_bonded_args b, int nr) {
    if ((int)blockIdx.x < nr) radii_body<false>(r, blockIdx.x * blockDim.x + threadIdx.x);
    else gather_lds_body<NOISE>(b, blockIdx.x - nr);
}

void og_launch_radii_gather(const og_radii_args& r, const og_bonded_args& b, hipStream_t s) {
    const int nr = (r.n_atoms + 255) / 256;
    const int ng = b.n_atoms > 0 ? (b.n_atoms + OG_GATHER_ATOMS - 1) / OG_GATHER_ATOMS : 0;
    if (nr + ng <= 0) return;
    if (b.noise.noise) hipLaunchKernelGGL(og_k_radii_gather<true>, dim3(nr + ng), dim3(256), 0, s, r, b, nr);
    else hipLaunchKernelGGL(og_k_radii_gather<false>, dim3(nr + ng), dim3(256), 0, s, r, b, nr);
}

void og_launch_radii(const og_radii_args& a, bool energy, hipStream_t s) {
    if (a.n_atoms <= 0) return;
    const int g = (a.n_atoms + 255) / 256;
    if (energy) hipLaunchKernelGGL(og_k_radii<true>, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(og_k_radii<false>, dim3(g), dim3(256), 0, s, a);
}

// One thread per cluster of at most MC atoms: clusters [c_begin, c_end).
// PRE: the O step reads its normals from a.noise (formed by og_k_noise)
// rather than forming them; a separate instance keeps the other one's
// registers as they are.
template <int MC, bool PRE>
__global__ void __launch_bounds__(256) og_k_cluster_step(og_cluster_args a, int c_begin, int c_end) {
    const int c = c_begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c_end) return;
    cluster_body<MC, PRE>(a, c, a.step_dev ? *a.step_dev : a.step);
}

__global__ void __launch_bounds__(256) og_k_noise(og_noise_args a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n_atoms) noise_atom(a, i);
}

void og_launch_noise(const og_noise_args& a, hipStream_t s) {
    if (a.n_atoms <= 0) return;
    hipLaunchKernelGGL(og_k_noise, dim3((a.n_atoms + 255) / 256), dim3(256), 0, s, a);
}

void og_launch_cluster_step(const og_cluster_args& a, hipStream_t s) {
    if (a.n_clusters <= 0) return;
    const dim3 g_small((a.n_small + 255) / 256), g_large((a.n_clusters - a.n_small + 255) / 256);
    if (a.noise) {
        if (a.n_small > 0)
            hipLaunchKernelGGL((og_k_cluster_step<OG_SMALL_CL, true>), g_small, dim3(256), 0, s, a, 0, a.n_small);
        if (a.n_clusters > a.n_small)
            hipLaunchKernelGGL((og_k_cluster_step<OG_MAX_CL, true>), g_large, dim3(256), 0, s, a, a.n_small,
                               a.n_clusters);
    } else {
        if (a.n_small > 0)
            hipLaunchKernelGGL((og_k_cluster_step<OG_SMALL_CL, false>), g_small, dim3(256), 0, s, a, 0, a.n_small);
        if (a.n_clusters > a.n_small)
            hipLaunchKernelGGL((og_k_cluster_step<OG_MAX_CL, false>), g_large, dim3(256), 0, s, a, a.n_small,
                               a.n_clusters);
    }
}

__global__ void __launch_bounds__(256) og_k_zero_acc(long long* facc, long long* iacc, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    facc[3 * (size_t)i] = 0; facc[3 * (size_t)i + 1] = 0; facc[3 * (size_t)i + 2] = 0;
    iacc[i] = 0;
}

void og_launch_zero_acc(long long* facc, long long* iacc, int n_atoms, hipStream_t s) {
    if (n_atoms <= 0) return;
    hipLaunchKernelGGL(og_k_zero_acc, dim3((n_atoms + 255) / 256), dim3(256), 0, s, facc, iacc, n_atoms);
}

// ---------------------------------------------------------------------------
// Finalise forces (fixed point -> FP64), second half-kick, kinetic energy,
// divergence flag (langevin.rs:309-319).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) og_k_back(og_back_args a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_atoms) return;
    if (i == 0 && a.do_kick && a.step_dev) *a.step_dev += 1u;
    back_atom(a, i);
}

__global__ void __launch_bounds__(256) og_k_zero_i64(long long* p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void og_launch_zero_i64(long long* p, long long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(og_k_zero_i64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}

__global__ void og_k_set_u32(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0) *p = v;
}

void og_launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(og_k_set_u32, dim3(1), dim3(64), 0, s, p, v);
}

void og_launch_back(const og_back_args& a, hipStream_t s) {
    if (a.n_atoms <= 0) return;
    hipLaunchKernelGGL(og_k_back, dim3((a.n_atoms + 255) / 256), dim3(256), 0, s, a);
}

__global__ void og_k_swap(double* pos, double* vel, double* force, int off_a, int off_b, int n,
                          double sa, double sb) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 3 * n) return;
    const size_t ia = 3 * (size_t)off_a + k, ib = 3 * (size_t)off_b + k;
    double t = pos[ia]; pos[ia] = pos[ib]; pos[ib] = t;
    t = force[ia]; force[ia] = force[ib]; force[ib] = t;
    const double va = vel[ib] * sa, vb = vel[ia] * sb;
    vel[ia] = va;
    vel[ib] = vb;
}

void og_launch_swap(double* pos, double* vel, double* force, int off_a, int off_b, int n,
                    double scale_a, double scale_b, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(og_k_swap, dim3((3 * n + 255) / 256), dim3(256), 0, s, pos, vel, force,
                       off_a, off_b, n, scale_a, scale_b);
}
