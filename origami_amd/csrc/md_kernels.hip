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


__global__ void __launch_bounds__(256) og_k_bonded(og_bonded_args a) {
    bonded_body<false>(a, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) og_k_bonded_e(og_bonded_args a) {
    bonded_body<true>(a, blockIdx.x, gridDim.x);
}

__device__ void noise3(uint32_t k0, uint32_t k1, uint32_t sys, uint32_t atom, uint32_t step,
                       double& g0, double& g1, double& g2);

// The O-step normals of atom i for step *step_dev + ahead.
__device__ __forceinline__ void noise_atom(const og_noise_args& a, int i) {
    const uint32_t step = *a.step_dev + a.ahead;
    double g0, g1, g2;
    noise3(a.key0, a.key1, (uint32_t)a.atom_sys[i], (uint32_t)a.atom_local[i], step, g0, g1, g2);
    double* o = a.noise + (size_t)(step & 1u) * 3 * a.n_atoms + 3 * (size_t)i;
    o[0] = g0;
    o[1] = g1;
    o[2] = g2;
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
    constexpr int B = 8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_atoms) return;
    if (NOISE) noise_atom(a.noise, i);
    const int q1 = a.csr_start[i + 1];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int q = a.csr_start[i]; q < q1; q += B) {
        long long s[B];
#pragma unroll
        for (int u = 0; u < B; ++u) s[u] = q + u < q1 ? a.csr_slot[q + u] : -1;
        double t[3 * B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            if (s[u] >= 0) {
                t[3 * u] = a.tf[3 * s[u]];
                t[3 * u + 1] = a.tf[3 * s[u] + 1];
                t[3 * u + 2] = a.tf[3 * s[u] + 2];
            }
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            if (s[u] >= 0) {
                fx += t[3 * u];
                fy += t[3 * u + 1];
                fz += t[3 * u + 2];
            }
        }
    }
    int* fl = a.sflag ? a.sflag + a.atom_sys[i] : nullptr;
    og_atomic_add_fix(a.facc + 3 * (size_t)i + 0, og_to_fix_chk(fx, fl));
    og_atomic_add_fix(a.facc + 3 * (size_t)i + 1, og_to_fix_chk(fy, fl));
    og_atomic_add_fix(a.facc + 3 * (size_t)i + 2, og_to_fix_chk(fz, fl));
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

// ---------------------------------------------------------------------------
// Born radii from the descreening integral (gb.rs:327-341), GB self energy.
// ---------------------------------------------------------------------------
template <bool ENERGY>
__device__ __forceinline__ void radii_body(const og_radii_args& a, int i) {
    int sys = 0;
    long long e_fix = 0;
    bool has_e = false;
    if (i < a.n_atoms) {
        sys = a.atom_sys[i];
        const bool sur = a.sur_mode && a.sur_mode[sys];
        double R = 0.0;
        if (!sur || a.check) {
            const double rho = a.rho[i];
            const double rt = rho - OG_OBC_OFFSET;
            const double integral = (double)a.iacc[i] * OG_FIX_INV;
            const double psi = integral * rt;
            const double arg = 1.0 * psi - 0.8 * psi * psi + 4.85 * psi * psi * psi;
            const double inv = 1.0 / rt - tanh(arg) / rho;
            const double floor_r = rt > 0.5 ? rt : 0.5;
            if (inv <= 0.0 || !isfinite(inv)) {
                if (!sur) atomicAdd(a.clamped + sys, 1);
                R = floor_r;
            } else {
                R = 1.0 / inv;
                R = R > floor_r ? R : floor_r;
            }
        }
        if (sur) {
            const double Rs = a.r_sur[i];
            if (a.check) {
                const float dev = (float)(fabs(Rs - R) / R);
                atomicMax(a.dev + sys, isfinite(dev) ? __float_as_uint(dev) : 0x7F800000u);
            }
            R = Rs;
        }
        a.radii[i] = R;
        float4 g = a.gb[a.atom_slot[i]];
        g.z = (float)R;
        g.w = (float)(0.60056120439322491 / R);  // sqrt(log2(e)/4) / R
        a.gb[a.atom_slot[i]] = g;
        if (ENERGY) {
            const double q = a.charge[i];
            if (q != 0.0) {
                const double pref = -0.5 * (1.0 - 1.0 / OG_EPS_WATER) * OG_COULOMB_KCAL * OG_KCAL_TO_KJ;
                e_fix = og_to_fix_chk(pref * q * q / R, a.sflag ? a.sflag + sys : nullptr);
                has_e = true;
            }
        }
    }
    // GB self energy: one atomic per system in the wave
    if (ENERGY) og_wave_add_fix_keyed(a.eacc, sys, OG_NE, OG_E_GB_SELF, e_fix, has_e);
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

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and Box-Muller normals.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t k0, uint32_t k1, uint32_t c[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c[0];
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2];
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
    const unsigned long long v = (((unsigned long long)hi << 32) | lo) >> 11;
    return ((double)v + 0.5) * (1.0 / 9007199254740992.0);
}

// Three standard normals for (system, atom, step).  Counter layout:
// (atom_local, step, system, sub) with sub in {0, 1}; key = seed.
__device__ void noise3(uint32_t k0, uint32_t k1, uint32_t sys, uint32_t atom, uint32_t step,
                       double& g0, double& g1, double& g2) {
    uint32_t c[4] = {atom, step, sys, 0u};
    philox4x32_10(k0, k1, c);
    double u1 = u53(c[0], c[1]), u2 = u53(c[2], c[3]);
    double r = sqrt(-2.0 * log(u1)), th = 2.0 * OG_PI * u2;
    g0 = r * cos(th);
    g1 = r * sin(th);
    uint32_t d[4] = {atom, step, sys, 1u};
    philox4x32_10(k0, k1, d);
    u1 = u53(d[0], d[1]);
    u2 = u53(d[2], d[3]);
    r = sqrt(-2.0 * log(u1));
    th = 2.0 * OG_PI * u2;
    g2 = r * cos(th);
}

// ---------------------------------------------------------------------------
// Fused BAOAB front half, one thread per constraint cluster (a heavy atom
// with its X-H partners, or a lone atom).  X-H constraints couple only
// through their heavy atom, so per-cluster Gauss-Seidel SHAKE reproduces
// shake_iterate's global sweep (converged constraints are skipped there).
// B, A, SHAKE, O, A, SHAKE; also zeroes this cluster's force and integral
// accumulators for the force evaluation that follows.
// ---------------------------------------------------------------------------

// SHAKE on one X-H cluster (shake.rs:35-125): local atom 0 is the heavy
// atom X and local atom m + 1 the hydrogen of the cluster's constraint m,
// the constraints in origami's sorted bond order (Gauss-Seidel).  Every
// constraint is taken as (X, H); for a bond given as (H, X) that flips the
// sign of both the current and the reference pair difference, which leaves
// err, denom and lambda unchanged and flips the correction, so each update is
// bit-identical to the given orientation.  All indices are compile-time, so
// the cluster lives in registers.
template <int MC>
__device__ __forceinline__ bool shake_cluster(double (&xs)[3 * MC], const double (&rs)[3 * MC],
                                              const double (&im)[MC], const double (&dl)[MC], int nc) {
    if (nc == 0) return true;
    for (int iter = 1; iter <= 64; ++iter) {
        double max_err = 0.0;
#pragma unroll
        for (int m = 0; m < MC - 1; ++m) {
            if (m < nc) {
                const int j = m + 1;
                const double d = dl[m];
                const double rnx = xs[0] - xs[3 * j], rny = xs[1] - xs[3 * j + 1], rnz = xs[2] - xs[3 * j + 2];
                const double err = (rnx * rnx + rny * rny + rnz * rnz) - d * d;
                if (fabs(err) > max_err) max_err = fabs(err);
                if (!(fabs(err) < 1e-6)) {
                    const double rox = rs[0] - rs[3 * j], roy = rs[1] - rs[3 * j + 1], roz = rs[2] - rs[3 * j + 2];
                    const double imi = im[0], imj = im[j];
                    const double denom = 2.0 * (imi + imj) * (rnx * rox + rny * roy + rnz * roz);
                    if (!(fabs(denom) < 1e-18)) {
                        const double lam = err / denom;
                        const double dx = rox * lam, dy = roy * lam, dz = roz * lam;
                        xs[0] -= dx * imi; xs[1] -= dy * imi; xs[2] -= dz * imi;
                        xs[3 * j] += dx * imj; xs[3 * j + 1] += dy * imj; xs[3 * j + 2] += dz * imj;
                    }
                }
            }
        }
        if (max_err < 1e-6) return true;
    }
    return false;
}

// One thread per cluster of at most MC atoms: clusters [c_begin, c_end).
// PRE: the O step reads its normals from a.noise (formed by og_k_noise)
// rather than forming them; a separate instance keeps the other one's
// registers as they are.
template <int MC, bool PRE>
__global__ void __launch_bounds__(256) og_k_cluster_step(og_cluster_args a, int c_begin, int c_end) {
    const int c = c_begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c_end) return;
    const int s0 = a.cl_start[c], na = a.cl_start[c + 1] - s0;
    const int k0 = a.cn_start[c], nc = a.cn_start[c + 1] - k0;
    int at[MC];
    double xs[3 * MC], vs[3 * MC], rs[3 * MC], im[MC], dl[MC];
    const int sys = a.atom_sys[a.cl_atoms[s0]];
    const bool frozen = a.diverged[sys] != 0;
    const double hdt = 0.5 * a.dt;
    const double kbt = OG_KB * a.sys_temp[sys];
    const double dofc = a.sys_dofcorr[sys];
    const double om = 1.0 - a.alpha * a.alpha;
#pragma unroll
    for (int t = 0; t < MC; ++t) {
        at[t] = 0;
        im[t] = 0.0;
        dl[t] = t < nc ? a.cn_len[k0 + t] : 0.0;
        for (int d = 0; d < 3; ++d) xs[3 * t + d] = vs[3 * t + d] = rs[3 * t + d] = 0.0;
        if (t < na) {
            const int g = a.cl_atoms[s0 + t];
            at[t] = g;
            im[t] = 1.0 / a.mass[g];
            for (int d = 0; d < 3; ++d) {
                xs[3 * t + d] = a.pos[3 * (size_t)g + d];
                vs[3 * t + d] = a.vel[3 * (size_t)g + d];
            }
            a.facc[3 * (size_t)g] = 0; a.facc[3 * (size_t)g + 1] = 0; a.facc[3 * (size_t)g + 2] = 0;
            a.iacc[g] = 0;
        }
    }
    if (frozen) return;
    int fails = 0;
    const uint32_t step = a.step_dev ? *a.step_dev : a.step;
    // B: v += F/m * ACCEL * dt/2
#pragma unroll
    for (int t = 0; t < MC; ++t) {
        if (t < na) {
            const int g = at[t];
            const double k = OG_ACCEL / a.mass[g] * hdt;
            for (int d = 0; d < 3; ++d) vs[3 * t + d] += a.force[3 * (size_t)g + d] * k;
        }
    }
    for (int half = 0; half < 2; ++half) {
        if (half == 1) {  // O
#pragma unroll
            for (int t = 0; t < MC; ++t) {
                if (t < na) {
                    const int g = at[t];
                    const double sigma = sqrt(om * kbt * dofc * OG_ACCEL / a.mass[g]);
                    double g0, g1, g2;
                    if (PRE) {
                        const double* nz = a.noise + (size_t)(step & 1u) * 3 * a.n_atoms + 3 * (size_t)g;
                        g0 = nz[0]; g1 = nz[1]; g2 = nz[2];
                    } else {
                        noise3(a.key0, a.key1, (uint32_t)sys, (uint32_t)a.atom_local[g], step, g0, g1, g2);
                    }
                    vs[3 * t] = a.alpha * vs[3 * t] + sigma * g0;
                    vs[3 * t + 1] = a.alpha * vs[3 * t + 1] + sigma * g1;
                    vs[3 * t + 2] = a.alpha * vs[3 * t + 2] + sigma * g2;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 3 * MC; ++k) {  // A
            rs[k] = xs[k];
            xs[k] += vs[k] * hdt;
        }
        if (a.use_shake) {
            if (!shake_cluster<MC>(xs, rs, im, dl, nc)) fails++;
#pragma unroll
            for (int k = 0; k < 3 * MC; ++k) vs[k] = (xs[k] - rs[k]) / hdt;
        }
    }
#pragma unroll
    for (int t = 0; t < MC; ++t) {
        if (t < na) {
            const int g = at[t];
            for (int d = 0; d < 3; ++d) {
                a.pos[3 * (size_t)g + d] = xs[3 * t + d];
                a.vel[3 * (size_t)g + d] = vs[3 * t + d];
            }
        }
    }
    if (fails) atomicAdd(a.shake_fail + sys, fails);
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
    double f[3];
    bool big = false;
    for (int d = 0; d < 3; ++d) {
        f[d] = (double)a.facc[3 * (size_t)i + d] * OG_FIX_INV;
        a.force[3 * (size_t)i + d] = f[d];
        big = big || !(fabs(f[d]) < OG_FIX_LIMIT);
    }
    const int sys = a.atom_sys[i];
    // a summed force at or beyond the contribution limit: the int64 sum may
    // have wrapped, so the system is reported as out of range
    if (big && a.sflag) atomicOr(a.sflag + sys, OG_ST_RANGE);
    if (!a.do_kick) return;
    if (a.diverged[sys]) return;
    // a contribution out of range (or a SASA atom over capacity) in this
    // step's forces: the system stops here, as on non-finite velocities
    const int st = a.sflag ? a.sflag[sys] : 0;
    if (st) {
        atomicOr(a.diverged + sys, st);
        return;
    }
    const double m = a.mass[i];
    const double k = OG_ACCEL / m * (0.5 * a.dt);
    double v2 = 0.0;
    bool finite = true;
    for (int d = 0; d < 3; ++d) {
        const double v = a.vel[3 * (size_t)i + d] + f[d] * k;
        a.vel[3 * (size_t)i + d] = v;
        v2 += v * v;
        finite = finite && isfinite(v);
    }
    if (!finite) atomicOr(a.diverged + sys, OG_ST_NONFINITE);
    if (a.ke_acc) og_atomic_add_fix(a.ke_acc + sys, og_to_fix_chk(0.5 * m * v2 / OG_ACCEL, nullptr));
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
