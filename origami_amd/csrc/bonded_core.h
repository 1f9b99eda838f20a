// Bonded terms in FP64 (energy/src/forces_bonded.rs:47-236,
// bonded.rs:59-191): one thread per term, each writing its per-atom forces
// into slots of its own.  Included by the kernels that launch them: the
// bonded kernels, and the descreening kernel that shares a launch with them
// on small batches.
#pragma once
#include "og_device.h"
#include "og_kernels.h"

#ifndef OG_PI
#define OG_PI 3.14159265358979323846
#endif

namespace {

struct d3 { double x, y, z; };
__device__ __forceinline__ d3 ld3(const double* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 add(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 mul(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ d3 dv(d3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ d3 neg(d3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double nrm(d3 a) { return sqrt(dot(a, a)); }

__device__ __forceinline__ void put(double* tf, long long slot, d3 f) {
    tf[3 * slot] = f.x;
    tf[3 * slot + 1] = f.y;
    tf[3 * slot + 2] = f.z;
}

// forces_bonded.rs:124-167
__device__ bool dihedral_gradient(d3 pa, d3 pb, d3 pc, d3 pd, d3& da, d3& db, d3& dc, d3& dd,
                                  double& phi) {
    d3 b1 = sub(pb, pa), b2 = sub(pc, pb), b3 = sub(pd, pc);
    double b2n = nrm(b2);
    if (b2n < 1e-9) return false;
    d3 s = cross(b1, b2), t = cross(b2, b3);
    double s2 = dot(s, s), t2 = dot(t, t);
    if (s2 < 1e-18 || t2 < 1e-18) return false;
    da = mul(s, b2n / s2);
    dd = mul(neg(t), b2n / t2);
    double ib2 = 1.0 / (b2n * b2n);
    double f1 = dot(b1, b2) * ib2, f2 = dot(b3, b2) * ib2;
    db = add(mul(da, -(f1 + 1.0)), mul(dd, f2));
    dc = neg(add(add(da, db), dd));
    d3 m1 = cross(s, dv(b2, b2n));
    phi = atan2(dot(m1, t), dot(s, t));
    return true;
}

// geom/src/measure.rs:17-28
__device__ double measure_dihedral(d3 a, d3 b, d3 c, d3 d) {
    d3 b0 = sub(a, b);
    d3 b1 = sub(c, b);
    b1 = dv(b1, nrm(b1));
    d3 b2 = sub(d, c);
    d3 v = sub(b0, mul(b1, dot(b0, b1)));
    d3 w = sub(b2, mul(b1, dot(b2, b1)));
    return atan2(dot(cross(b1, v), w), dot(v, w));
}

__device__ __forceinline__ double wrap_pi(double x) {
    while (x > OG_PI) x -= 2.0 * OG_PI;
    while (x < -OG_PI) x += 2.0 * OG_PI;
    return x;
}

// One bonded term k: its per-atom forces into its own slots; with ENERGY
// its energy (fixed point), energy slot and system are returned.
template <bool ENERGY>
__device__ __forceinline__ void bonded_term(const og_bonded_args& a, int k, int& e_slot, int& e_sys, long long& e_fix) {
    if (k < a.nb) {
        const int ia = a.bonds[2 * k], ib = a.bonds[2 * k + 1];
        const double K = a.bond_p[2 * k], r0 = a.bond_p[2 * k + 1];
        d3 d = sub(ld3(a.pos + 3 * ia), ld3(a.pos + 3 * ib));
        d = neg(d);  // d = x_b - x_a
        const double r = nrm(d);
        d3 f = {0.0, 0.0, 0.0};
        if (r >= 1e-9) {
            const double mag = 2.0 * K * (r - r0) * OG_KCAL_TO_KJ;
            f = mul(d, mag / r);
        }
        put(a.tf, 2ll * k, f);
        put(a.tf, 2ll * k + 1, neg(f));
        if (ENERGY) {
            const double dr = r - r0;
            e_slot = OG_E_BOND;
            e_sys = a.atom_sys[ia];
            e_fix = og_to_fix_chk(K * dr * dr * OG_KCAL_TO_KJ, a.sflag ? a.sflag + e_sys : nullptr);
        }
        return;
    }
    int kk = k - a.nb;
    if (kk < a.na) {
        const int ia = a.angles[3 * kk], ib = a.angles[3 * kk + 1], ic = a.angles[3 * kk + 2];
        const double K = a.angle_p[2 * kk], th0 = a.angle_p[2 * kk + 1];
        const d3 pb = ld3(a.pos + 3 * ib);
        const d3 u = sub(ld3(a.pos + 3 * ia), pb), v = sub(ld3(a.pos + 3 * ic), pb);
        const double un = nrm(u), vn = nrm(v);
        const long long s0 = 2ll * a.nb + 3ll * kk;
        d3 fa = {0.0, 0.0, 0.0}, fc = {0.0, 0.0, 0.0};
        if (un >= 1e-9 && vn >= 1e-9) {
            const d3 uh = dv(u, un), vh = dv(v, vn);
            double ct = fmin(fmax(dot(uh, vh), -1.0), 1.0);
            const double st = sqrt(1.0 - ct * ct);
            if (st >= 1e-9) {
                const double theta = acos(ct);
                const double coeff = 2.0 * K * (theta - th0) * OG_KCAL_TO_KJ / st;
                fa = mul(sub(vh, mul(uh, ct)), coeff / un);
                fc = mul(sub(uh, mul(vh, ct)), coeff / vn);
            }
            if (ENERGY) {
                const double dt = acos(fmin(fmax(dot(uh, vh), -1.0), 1.0)) - th0;
                e_slot = OG_E_ANGLE;
                e_sys = a.atom_sys[ia];
                e_fix = og_to_fix_chk(K * dt * dt * OG_KCAL_TO_KJ, a.sflag ? a.sflag + e_sys : nullptr);
            }
        }
        put(a.tf, s0, fa);
        put(a.tf, s0 + 1, neg(add(fa, fc)));
        put(a.tf, s0 + 2, fc);
        return;
    }
    kk -= a.na;
    const bool is_dih = kk < a.nd;
    if (!is_dih) kk -= a.nd;
    const int* q = is_dih ? a.dih + 4 * kk : a.imp + 4 * kk;
    const d3 pa = ld3(a.pos + 3 * q[0]), pb = ld3(a.pos + 3 * q[1]), pc = ld3(a.pos + 3 * q[2]),
             pd = ld3(a.pos + 3 * q[3]);
    d3 da, db, dc, dd;
    double phi;
    const long long s0 = 2ll * a.nb + 3ll * a.na + (is_dih ? 4ll * kk : 4ll * a.nd + 4ll * kk);
    d3 f0 = {0.0, 0.0, 0.0}, f1 = f0, f2 = f0, f3 = f0;
    if (dihedral_gradient(pa, pb, pc, pd, da, db, dc, dd, phi)) {
        double dvdphi;
        if (is_dih) {
            const double kd = a.dih_p[3 * kk], nn = a.dih_p[3 * kk + 1], delta = a.dih_p[3 * kk + 2];
            dvdphi = -kd * nn * sin(nn * phi - delta) * OG_KCAL_TO_KJ;
        } else {
            dvdphi = 2.0 * a.imp_p[2 * kk] * wrap_pi(phi - a.imp_p[2 * kk + 1]) * OG_KCAL_TO_KJ;
        }
        f0 = mul(da, -dvdphi);
        f1 = mul(db, -dvdphi);
        f2 = mul(dc, -dvdphi);
        f3 = mul(dd, -dvdphi);
    }
    put(a.tf, s0, f0);
    put(a.tf, s0 + 1, f1);
    put(a.tf, s0 + 2, f2);
    put(a.tf, s0 + 3, f3);
    if (ENERGY) {
        const double chi = measure_dihedral(pa, pb, pc, pd);
        double e;
        int slot;
        if (is_dih) {
            e = a.dih_p[3 * kk] * (1.0 + cos(a.dih_p[3 * kk + 1] * chi - a.dih_p[3 * kk + 2]));
            slot = OG_E_DIHEDRAL;
        } else {
            const double dom = wrap_pi(chi - a.imp_p[2 * kk + 1]);
            e = a.imp_p[2 * kk] * dom * dom;
            slot = OG_E_IMPROPER;
        }
        e_slot = slot;
        e_sys = a.atom_sys[q[0]];
        e_fix = og_to_fix_chk(e * OG_KCAL_TO_KJ, a.sflag ? a.sflag + e_sys : nullptr);
    }
}

// Grid-stride over the terms, a wave at a time so that the energy sums run
// in uniform control flow: one atomic per (system, term kind) in the wave.
// blk / nblk: this block's index among the nblk blocks that share the terms
// (a launch of its own, or the bonded part of a shared launch).
template <bool ENERGY>
__device__ void bonded_body(const og_bonded_args& a, int blk, int nblk) {
    const int total = a.nb + a.na + a.nd + a.ni;
    const int lane = threadIdx.x & 63;
    const int stride = nblk * blockDim.x;
    for (int kb = blk * blockDim.x + (threadIdx.x & ~63); kb < total; kb += stride) {
        const int k = kb + lane;
        int e_slot = -1, e_sys = 0;
        long long e_fix = 0;
        if (k < total) bonded_term<ENERGY>(a, k, e_slot, e_sys, e_fix);
        if (ENERGY) {
            og_wave_add_fix_keyed(a.eacc, e_sys, OG_NE, OG_E_BOND, e_fix, e_slot == OG_E_BOND);
            og_wave_add_fix_keyed(a.eacc, e_sys, OG_NE, OG_E_ANGLE, e_fix, e_slot == OG_E_ANGLE);
            og_wave_add_fix_keyed(a.eacc, e_sys, OG_NE, OG_E_DIHEDRAL, e_fix, e_slot == OG_E_DIHEDRAL);
            og_wave_add_fix_keyed(a.eacc, e_sys, OG_NE, OG_E_IMPROPER, e_fix, e_slot == OG_E_IMPROPER);
        }
    }
}

}  // namespace
