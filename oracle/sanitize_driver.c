/* CPU driver for `make sanitize` (test infrastructure): runs every oracle
 * entry point and the host build of the SASA core on a small deterministic
 * system under AddressSanitizer + UndefinedBehaviorSanitizer.  A chain of 48
 * atoms with bonds, angles, dihedrals, impropers, exclusions, 1-4 pairs and
 * X-H constraints along it; positions from a fixed linear congruential
 * stream.  Exits non-zero on a non-finite result (the sanitizers abort on
 * their own findings). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

void host_intersect_counts(int n, const double *caps, int *counts, int *fast);
void host_intersect_points(int n, const double *caps, double *pts);
void host_midpoints(int n, const double *in, double *out);
void host_cover_arcs(int n, const double *caps, float *arcs, double *basis);
void host_jvp_split(int n, const double *rows, double *arc_full, double *arc_split, double *vtx_full,
                    double *vtx_split);

static uint64_t lcg = 88172645463325252ULL;
static double uni(void) {
    lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(lcg >> 11) * (1.0 / 9007199254740992.0);
}

static int finite_all(const double *x, int n, const char *what) {
    for (int i = 0; i < n; ++i)
        if (!isfinite(x[i])) {
            fprintf(stderr, "sanitize driver: non-finite %s[%d]\n", what, i);
            return 0;
        }
    return 1;
}

#define N 48

int main(void) {
    double pos[3 * N], q[N], rmin[N], eps[N], rmin14[N], eps14[N], rho[N], scale[N], mass[N], sr[N], sg[N];
    for (int i = 0; i < N; ++i) {
        /* a helix-like chain, 1.5 Å between neighbours, with a small jitter */
        pos[3 * i] = 2.3 * cos(1.7 * i) + 0.1 * uni();
        pos[3 * i + 1] = 2.3 * sin(1.7 * i) + 0.1 * uni();
        pos[3 * i + 2] = 1.1 * i + 0.1 * uni();
        q[i] = (i % 3 == 0) ? -0.3 : 0.15;
        rmin[i] = 1.6 + 0.4 * uni();
        eps[i] = 0.05 + 0.1 * uni();
        rmin14[i] = rmin[i] - 0.1;
        eps14[i] = 0.5 * eps[i];
        rho[i] = 1.5 + 0.3 * uni();
        scale[i] = 0.72 + 0.1 * uni();
        mass[i] = (i % 4 == 3) ? 1.008 : 12.011;
        sr[i] = rho[i] + 1.4;
        sg[i] = (i % 2) ? 0.005 * 4.184 : 0.0;
    }
    int bonds[2 * (N - 1)], excl[2 * 2 * N], p14[2 * N], angles[3 * (N - 2)], dih[4 * (N - 3)], imp[4 * 4],
        cons[2 * N];
    double bond_p[2 * (N - 1)], angle_p[2 * (N - 2)], dih_p[3 * (N - 3)], imp_p[2 * 4], cons_d[N];
    int nb = 0, ne = 0, n14 = 0, na = 0, nd = 0, ni = 0, nc = 0;
    for (int i = 0; i + 1 < N; ++i) {
        bonds[2 * nb] = i; bonds[2 * nb + 1] = i + 1;
        bond_p[2 * nb] = 300.0; bond_p[2 * nb + 1] = 1.5; ++nb;
        excl[2 * ne] = i; excl[2 * ne + 1] = i + 1; ++ne;
        if (mass[i + 1] < 2.0) { cons[2 * nc] = i; cons[2 * nc + 1] = i + 1; cons_d[nc] = 1.09; ++nc; }
    }
    for (int i = 0; i + 2 < N; ++i) {
        angles[3 * na] = i; angles[3 * na + 1] = i + 1; angles[3 * na + 2] = i + 2;
        angle_p[2 * na] = 50.0; angle_p[2 * na + 1] = 1.9; ++na;
        excl[2 * ne] = i; excl[2 * ne + 1] = i + 2; ++ne;
    }
    for (int i = 0; i + 3 < N; ++i) {
        for (int c = 0; c < 4; ++c) dih[4 * nd + c] = i + c;
        dih_p[3 * nd] = 0.2; dih_p[3 * nd + 1] = 3.0; dih_p[3 * nd + 2] = 0.0; ++nd;
        p14[2 * n14] = i; p14[2 * n14 + 1] = i + 3; ++n14;
    }
    for (int k = 0; k < 4; ++k) {
        for (int c = 0; c < 4; ++c) imp[4 * ni + c] = 10 * k + c;
        imp_p[2 * ni] = 120.0; imp_p[2 * ni + 1] = 0.0; ++ni;
    }
    orc_sys *s = orc_create(N, q, rmin, eps, rmin14, eps14, rho, scale, mass, sr, sg, ne, excl, n14, p14, nb, bonds,
                            bond_p, na, angles, angle_p, nd, dih, dih_p, ni, imp, imp_p, nc, cons, cons_d);
    if (!s || orc_n_atoms(s) != N) return 1;
    double f[3 * N], radii[N], integral[N], terms[9 * N], e[16], vel[3 * N], xi[3 * N], ref[3 * N];
    int ok = 1;
    for (int mode = 0; mode < 2; ++mode) {
        const double cnb = mode ? 10.0 : INFINITY, cgb = mode ? 10.0 : INFINITY, cds = mode ? 20.0 : INFINITY;
        orc_total_force(s, pos, cnb, cgb, cds, f, radii, terms);
        ok &= finite_all(f, 3 * N, "total_force");
        orc_total_energy(s, pos, cnb, cds, e);
        ok &= finite_all(e, 10, "total_energy");
        orc_nonbonded_forces(s, pos, cnb, f);
        orc_nonbonded_energy(s, pos, cnb, e);
        orc_born_radii(s, pos, cds, integral, radii);
        orc_gb_forces(s, pos, radii, cgb, f);
        orc_gb_energy(s, pos, radii, e);
        ok &= finite_all(f, 3 * N, "gb_forces");
    }
    orc_bonded_forces(s, pos, f);
    orc_bonded_energy(s, pos, e);
    ok &= finite_all(e, 4, "bonded_energy");
    for (int i = 0; i < 3 * N; ++i) { vel[i] = 0.01 * (uni() - 0.5); xi[i] = uni() - 0.5; ref[i] = pos[i]; }
    orc_total_force(s, pos, INFINITY, INFINITY, INFINITY, f, NULL, NULL);
    orc_baoab_step(s, pos, vel, f, xi, 1.0, 300.0, 2.0, 1, INFINITY, INFINITY, INFINITY);
    orc_shake(s, pos, ref, 1e-6, 64);
    ok &= finite_all(pos, 3 * N, "positions after a step");
    (void)orc_kinetic_energy(s, vel);
    double stats[8];
    orc_psa_forces(N, pos, sr, sg, f, stats);
    ok &= finite_all(f, 3 * N, "psa_forces");
    double per_atom[N];
    (void)orc_psa_area(N, pos, sr, per_atom);
    int nbl[N];
    for (int i = 0; i < N; ++i) nbl[i] = i;
    double grad[3 * N];
    (void)orc_psa_atom_gradient(N, pos, sr, 5, N, nbl, grad);
    const double axes[6] = {0, 0, 1, 0.6, 0.0, 0.8};
    const double cosa[2] = {0.3, -0.2};
    (void)orc_psa_area_from_caps(2, axes, cosa, 3.1);
    uint64_t u[16];
    orc_xoshiro_u64(1234, 16, u);
    orc_splitmix64(1234567, 4, u);
    uint64_t st[4] = {1, 2, 3, 4};
    orc_xoshiro_from_state(st, 8, u);
    uint32_t ph[4];
    orc_philox4x32(1, 2, 3, 4, 5, 6, ph);
    orc_philox_normals(7, 8, 0, N, 3, xi);
    ok &= finite_all(xi, 3 * N, "philox normals");
    orc_destroy(s);
    /* the SASA core's host build on random caps */
    enum { M = 64 };
    double caps[8 * M], pts[6 * M], mid_in[8 * M], mid[3 * M], basis[6 * M], rows[18 * M];
    double af[12 * M], as[12 * M], vf[9 * M], vs[9 * M];
    float arcs[2 * M];
    int cnt[M], fast[M];
    for (int i = 0; i < M; ++i) {
        for (int c = 0; c < 2; ++c) {
            double x = uni() - 0.5, y = uni() - 0.5, z = uni() - 0.5, n = sqrt(x * x + y * y + z * z) + 1e-12;
            caps[8 * i + 4 * c] = x / n; caps[8 * i + 4 * c + 1] = y / n; caps[8 * i + 4 * c + 2] = z / n;
            caps[8 * i + 4 * c + 3] = 1.6 * uni() - 0.8;
            mid_in[8 * i + 3 * c] = x / n; mid_in[8 * i + 3 * c + 1] = y / n; mid_in[8 * i + 3 * c + 2] = z / n;
        }
        mid_in[8 * i + 6] = caps[8 * i + 3];
        mid_in[8 * i + 7] = 6.0 * uni();
        for (int k = 0; k < 12; ++k) rows[18 * i + k] = 3.0 * uni();
        for (int k = 12; k < 16; ++k) rows[18 * i + k] = 2.5 + uni();
        rows[18 * i + 16] = 1.0; rows[18 * i + 17] = -1.0;
    }
    host_intersect_counts(M, caps, cnt, fast);
    host_intersect_points(M, caps, pts);
    host_midpoints(M, mid_in, mid);
    host_cover_arcs(M, caps, arcs, basis);
    host_jvp_split(M, rows, af, as, vf, vs);
    if (!ok) return 1;
    printf("sanitize driver: oracle and SASA core ran clean under ASan/UBSan\n");
    return 0;
}
