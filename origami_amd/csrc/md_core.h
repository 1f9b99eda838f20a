// Device bodies of the O(N) FP64 kernels (md_kernels.hip), shared with the
// persistent replica-step kernel (pair_kernels.hip): Philox normals, Born
// radii, the per-cluster BAOAB front half with SHAKE, the bonded gather and
// the force finalisation.
//
//   Born radii       energy/src/gb.rs:327-341 (tanh transform + clamp)
//   BAOAB            dynamics/src/langevin.rs:223-313
//   SHAKE            dynamics/src/shake.rs:86-125
//   kinetic energy   dynamics/src/langevin.rs:469-475
#pragma once
#include "og_device.h"
#include "og_kernels.h"

#ifndef OG_PI
#define OG_PI 3.14159265358979323846
#endif

namespace {

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
__device__ __forceinline__ void noise3(uint32_t k0, uint32_t k1, uint32_t sys, uint32_t atom, uint32_t step,
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

// The O-step normals of atom i for an explicit step.
__device__ __forceinline__ void noise_atom_at(const og_noise_args& a, int i, uint32_t step) {
    double g0, g1, g2;
    noise3(a.key0, a.key1, (uint32_t)a.atom_sys[i], (uint32_t)a.atom_local[i], step, g0, g1, g2);
    double* o = a.noise + (size_t)(step & 1u) * 3 * a.n_atoms + 3 * (size_t)i;
    o[0] = g0;
    o[1] = g1;
    o[2] = g2;
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
            const double integral =
                (double)__hip_atomic_load(a.iacc + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) * OG_FIX_INV;
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

// Cluster c (at most MC atoms) for step `step` (the Philox counter).  PRE:
// the O step reads its normals from a.noise rather than forming them.
template <int MC, bool PRE>
__device__ __forceinline__ void cluster_body(const og_cluster_args& a, int c, uint32_t step) {
    const int s0 = a.cl_start[c], na = a.cl_start[c + 1] - s0;
    const int k0 = a.cn_start[c], nc = a.cn_start[c + 1] - k0;
    int at[MC];
    double xs[3 * MC], vs[3 * MC], rs[3 * MC], im[MC], dl[MC];
    const int sys = a.atom_sys[a.cl_atoms[s0]];
    const bool frozen = __hip_atomic_load(a.diverged + sys, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
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

// Atom i's bonded slots, fetched eight at a time (indices, then their
// forces, so that a batch of loads is in flight together), added in slot
// order, one fixed-point add per component.
__device__ __forceinline__ void gather_atom(const og_bonded_args& a, int i) {
    constexpr int B = 8;
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

// Atom i: fixed point -> FP64 force; with do_kick the closing half-kick,
// kinetic energy and divergence flag.  facc is read at device scope (it is
// summed by atomics in L2).
__device__ __forceinline__ void back_atom(const og_back_args& a, int i) {
    double f[3];
    bool big = false;
    for (int d = 0; d < 3; ++d) {
        f[d] = (double)__hip_atomic_load(a.facc + 3 * (size_t)i + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) *
               OG_FIX_INV;
        big = big || !(fabs(f[d]) < OG_FIX_LIMIT);
    }
    if (a.tun.on) {          // CylindricalTunnel::add_force (cotranslate.rs:131-155)
        const og_tunnel_dev& t = a.tun;
        const double rx = a.pos[3 * (size_t)i] - t.ox, ry = a.pos[3 * (size_t)i + 1] - t.oy,
                     rz = a.pos[3 * (size_t)i + 2] - t.oz;
        const double along = rx * t.dx + ry * t.dy + rz * t.dz;
        if (!(along < 0.0 || along > t.length)) {
            const double px = rx - t.dx * along, py = ry - t.dy * along, pz = rz - t.dz * along;
            const double dist = sqrt(px * px + py * py + pz * pz);
            if (dist > t.radius) {
                const double push = t.k * (dist - t.radius);
                f[0] -= px / dist * push;
                f[1] -= py / dist * push;
                f[2] -= pz / dist * push;
            }
        }
    }
    for (int d = 0; d < 3; ++d) a.force[3 * (size_t)i + d] = f[d];
    const int sys = a.atom_sys[i];
    // a summed force at or beyond the contribution limit: the int64 sum may
    // have wrapped, so the system is reported as out of range
    if (big && a.sflag) atomicOr(a.sflag + sys, OG_ST_RANGE);
    if (!a.do_kick) return;
    if (__hip_atomic_load(a.diverged + sys, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    // a contribution out of range (or a SASA atom over capacity) in this
    // step's forces: the system stops here, as on non-finite velocities
    const int st = a.sflag ? __hip_atomic_load(a.sflag + sys, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
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

}  // namespace
