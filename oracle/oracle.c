/*
 * FP64 CPU oracle for the origami force-and-integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it restates, in plain
 * C and in double precision, the arithmetic of origami's energy and dynamics
 * crates so the HIP engine can be compared against it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product library never links, calls or falls back to it.
 *
 * Every function cites the origami source it follows (paths relative to the
 * reference repository).  Pair loops visit i < j in index order; origami
 * visits the same pair set in cell-list / Verlet order, which only changes
 * floating-point summation order.
 *
 * Two modes share one code path through the cutoff arguments:
 *   Mode A (origami's integrator): LJ/Coulomb cut at 10 Å with reaction-field
 *     Coulomb, GB pair force cut at 10 Å, descreening at 20 Å.
 *   Mode B (all pairs): every cutoff = +inf, for which the reaction-field
 *     constants 1/(2Rc^3), 3/(2Rc), 1/Rc^3 are exactly 0 (bare Coulomb).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define KCAL_TO_KJ 4.184                 /* energy/src/units.rs:5 */
#define COULOMB_KCAL 332.0637            /* energy/src/nonbonded.rs:34 */
#define EPS_WATER 78.5                   /* energy/src/gb.rs:23 */
#define OBC_ALPHA 1.0                    /* energy/src/gb.rs:28-30 */
#define OBC_BETA 0.8
#define OBC_GAMMA 4.85
#define OBC_OFFSET 0.09                  /* energy/src/gb.rs:32 */
#define ACCEL_FACTOR 1.0e-4              /* dynamics/src/langevin.rs:49 */
#define BOLTZMANN_KJ 8.314462618e-3      /* dynamics/src/langevin.rs:46 */

#define EXCLUDED_BIT 1u                  /* energy/src/scratch.rs:31 */
#define ONE_FOUR_BIT 2u                  /* energy/src/scratch.rs:32 */

struct orc_sys {
    int n;
    double *q, *rmin, *eps, *rmin14, *eps14, *rho, *scale, *mass;
    double *sasa_r, *sasa_g;
    unsigned char *mask; /* n*n, scratch.rs:233-251 */
    int nb, na, nd, ni, nc;
    int *bonds, *angles, *dihedrals, *impropers, *cons;
    double *bond_p, *angle_p, *dih_p, *imp_p, *cons_d;
};

static double *dup_d(const double *s, size_t n) {
    double *d = (double *)malloc((n ? n : 1) * sizeof(double));
    if (n) memcpy(d, s, n * sizeof(double));
    return d;
}
static int *dup_i(const int *s, size_t n) {
    int *d = (int *)malloc((n ? n : 1) * sizeof(int));
    if (n) memcpy(d, s, n * sizeof(int));
    return d;
}

orc_sys *orc_create(int n, const double *q, const double *rmin, const double *eps,
                    const double *rmin14, const double *eps14, const double *rho,
                    const double *scale, const double *mass, const double *sasa_r,
                    const double *sasa_g, int n_excl, const int *excl, int n14, const int *p14,
                    int nb, const int *bonds, const double *bond_p, int na, const int *angles,
                    const double *angle_p, int nd, const int *dih, const double *dih_p, int ni,
                    const int *imp, const double *imp_p, int nc, const int *cons,
                    const double *cons_d) {
    orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
    s->n = n;
    s->q = dup_d(q, n);
    s->rmin = dup_d(rmin, n);
    s->eps = dup_d(eps, n);
    s->rmin14 = dup_d(rmin14, n);
    s->eps14 = dup_d(eps14, n);
    s->rho = dup_d(rho, n);
    s->scale = dup_d(scale, n);
    s->mass = dup_d(mass, n);
    s->sasa_r = dup_d(sasa_r, n);
    s->sasa_g = dup_d(sasa_g, n);
    s->mask = (unsigned char *)calloc((size_t)n * n + 1, 1);
    for (int k = 0; k < n_excl; ++k) {
        int a = excl[2 * k], b = excl[2 * k + 1];
        s->mask[(size_t)a * n + b] = EXCLUDED_BIT;
        s->mask[(size_t)b * n + a] = EXCLUDED_BIT;
    }
    for (int k = 0; k < n14; ++k) {
        int a = p14[2 * k], b = p14[2 * k + 1];
        s->mask[(size_t)a * n + b] = ONE_FOUR_BIT;
        s->mask[(size_t)b * n + a] = ONE_FOUR_BIT;
    }
    s->nb = nb; s->na = na; s->nd = nd; s->ni = ni; s->nc = nc;
    s->bonds = dup_i(bonds, 2 * (size_t)nb);
    s->bond_p = dup_d(bond_p, 2 * (size_t)nb);
    s->angles = dup_i(angles, 3 * (size_t)na);
    s->angle_p = dup_d(angle_p, 2 * (size_t)na);
    s->dihedrals = dup_i(dih, 4 * (size_t)nd);
    s->dih_p = dup_d(dih_p, 3 * (size_t)nd);
    s->impropers = dup_i(imp, 4 * (size_t)ni);
    s->imp_p = dup_d(imp_p, 2 * (size_t)ni);
    s->cons = dup_i(cons, 2 * (size_t)nc);
    s->cons_d = dup_d(cons_d, (size_t)nc);
    return s;
}

void orc_destroy(orc_sys *s) {
    if (!s) return;
    free(s->q); free(s->rmin); free(s->eps); free(s->rmin14); free(s->eps14);
    free(s->rho); free(s->scale); free(s->mass); free(s->sasa_r); free(s->sasa_g);
    free(s->mask); free(s->bonds); free(s->bond_p); free(s->angles); free(s->angle_p);
    free(s->dihedrals); free(s->dih_p); free(s->impropers); free(s->imp_p);
    free(s->cons); free(s->cons_d);
    free(s);
}

int orc_n_atoms(const orc_sys *s) { return s->n; }

/* ------------------------------------------------------------------------ */
/* Small vector helpers (nalgebra Vector3<f64> semantics).                   */
/* ------------------------------------------------------------------------ */
typedef struct { double x, y, z; } v3;
static inline v3 V(const double *p) { v3 r = {p[0], p[1], p[2]}; return r; }
static inline v3 vsub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline v3 vadd(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline v3 vmul(v3 a, double s) { v3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static inline v3 vdiv(v3 a, double s) { v3 r = {a.x / s, a.y / s, a.z / s}; return r; }
static inline v3 vneg(v3 a) { v3 r = {-a.x, -a.y, -a.z}; return r; }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static inline double vnorm2(v3 a) { return vdot(a, a); }
static inline double vnorm(v3 a) { return sqrt(vdot(a, a)); }
static inline void vacc(double *f, v3 a) { f[0] += a.x; f[1] += a.y; f[2] += a.z; }
static inline void vdec(double *f, v3 a) { f[0] -= a.x; f[1] -= a.y; f[2] -= a.z; }

/* ------------------------------------------------------------------------ */
/* LJ + reaction-field Coulomb forces                                        */
/* energy/src/forces_nonbonded.rs:125-303 (SoA kernel, compute_pair :196-232)*/
/* energy/src/nonbonded.rs:41-84 (CoulombRf)                                 */
/* ------------------------------------------------------------------------ */
void orc_nonbonded_forces(const orc_sys *s, const double *pos, double cutoff, double *f) {
    const int n = s->n;
    const double cutoff_sq = cutoff * cutoff;
    const double kj_per_kcal = KCAL_TO_KJ;
    const double rc3 = cutoff * cutoff * cutoff;
    const double inv_rc3 = 1.0 / rc3;
    const double coulomb_kj = COULOMB_KCAL * KCAL_TO_KJ;
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            unsigned char m = s->mask[(size_t)i * n + j];
            if (m & EXCLUDED_BIT) continue;
            double dx = pos[3 * j] - pos[3 * i];
            double dy = pos[3 * j + 1] - pos[3 * i + 1];
            double dz = pos[3 * j + 2] - pos[3 * i + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > cutoff_sq || r2 < 1e-18) continue;
            double r = sqrt(r2);
            int one_four = (m & ONE_FOUR_BIT) != 0;
            double ri = one_four ? s->rmin14[i] : s->rmin[i];
            double ei = one_four ? s->eps14[i] : s->eps[i];
            double rj = one_four ? s->rmin14[j] : s->rmin[j];
            double ej = one_four ? s->eps14[j] : s->eps[j];
            double rmin_ij = ri + rj;
            double eps_ij = sqrt(ei * ej);
            double inv_r2 = 1.0 / r2;
            double ratio2 = rmin_ij * rmin_ij * inv_r2;
            double ratio6 = ratio2 * ratio2 * ratio2;
            double ratio12 = ratio6 * ratio6;
            double lj_coeff = 12.0 * eps_ij * inv_r2 * (ratio6 - ratio12) * kj_per_kcal;
            double qq = s->q[i] * s->q[j];
            double coul_coeff = 0.0;
            if (qq != 0.0) coul_coeff = -coulomb_kj * qq * (inv_r2 / r - inv_rc3);
            double coeff = lj_coeff + coul_coeff;
            f[3 * i] += dx * coeff; f[3 * i + 1] += dy * coeff; f[3 * i + 2] += dz * coeff;
            f[3 * j] -= dx * coeff; f[3 * j + 1] -= dy * coeff; f[3 * j + 2] -= dz * coeff;
        }
    }
}

/* nonbonded_energy (energy/src/nonbonded.rs:97-174). out[0]=LJ kJ/mol,
 * out[1]=Coulomb kJ/mol, out[2]=pair count, out[3]=1-4 count. */
void orc_nonbonded_energy(const orc_sys *s, const double *pos, double cutoff, double *out) {
    const int n = s->n;
    const double cutoff_sq = cutoff * cutoff;
    const double rc3 = cutoff * cutoff * cutoff;
    const double kappa = 1.0 / (2.0 * rc3);
    const double c = 3.0 / (2.0 * cutoff);
    double lj = 0.0, coul = 0.0, pairs = 0.0, n14 = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            double dx = pos[3 * i] - pos[3 * j];
            double dy = pos[3 * i + 1] - pos[3 * j + 1];
            double dz = pos[3 * i + 2] - pos[3 * j + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 <= cutoff_sq)) continue;
            unsigned char m = s->mask[(size_t)i * n + j];
            if (m & EXCLUDED_BIT) continue;
            double r = sqrt(r2);
            int one_four = (m & ONE_FOUR_BIT) != 0;
            double ri = one_four ? s->rmin14[i] : s->rmin[i];
            double ei = one_four ? s->eps14[i] : s->eps[i];
            double rj = one_four ? s->rmin14[j] : s->rmin[j];
            double ej = one_four ? s->eps14[j] : s->eps[j];
            double rmin_ij = ri + rj;
            double eps_ij = sqrt(ei * ej);
            double t = rmin_ij / r;
            double t2 = t * t;
            double ratio6 = t2 * t2 * t2; /* powi(6) */
            double ratio12 = ratio6 * ratio6;
            lj += eps_ij * (ratio12 - 2.0 * ratio6);
            double qq = s->q[i] * s->q[j];
            if (qq != 0.0) coul += COULOMB_KCAL * qq * (1.0 / r + kappa * r * r - c);
            pairs += 1.0;
            if (one_four) n14 += 1.0;
        }
    }
    out[0] = lj * KCAL_TO_KJ;
    out[1] = coul * KCAL_TO_KJ;
    out[2] = pairs;
    out[3] = n14;
}

/* ------------------------------------------------------------------------ */
/* Generalized Born (OBC II form as written in origami)                      */
/* ------------------------------------------------------------------------ */
/* pairwise_descreening, energy/src/gb.rs:347-369 */
double orc_pairwise_descreening(double r, double rho_i_tilde, double s_rho_j_tilde) {
    if (r + s_rho_j_tilde <= rho_i_tilde) return 0.0;
    double l = (r - s_rho_j_tilde < rho_i_tilde) ? rho_i_tilde : r - s_rho_j_tilde;
    double u = r + s_rho_j_tilde;
    if (u <= 0.0 || l <= 0.0) return 0.0;
    double inv_l = 1.0 / l;
    double inv_u = 1.0 / u;
    double term1 = 0.5 * (inv_l - inv_u);
    double term2 = (r / 4.0) * (inv_u * inv_u - inv_l * inv_l);
    double term3 = (1.0 / (2.0 * r)) * log(l / u);
    double term4 = (s_rho_j_tilde * s_rho_j_tilde - r * r) / (4.0 * r) * (inv_u * inv_u - inv_l * inv_l);
    return term1 + term2 + term3 + term4;
}

/* compute_born_radii_into_scratch, energy/src/gb.rs:203-342.  Writes the
 * descreening integral and the effective radii; returns the clamp count. */
int orc_born_radii(const orc_sys *s, const double *pos, double cutoff, double *integral,
                   double *radii) {
    const int n = s->n;
    const double cutoff_sq = cutoff * cutoff;
    for (int i = 0; i < n; ++i) integral[i] = 0.0;
    for (int i = 0; i < n; ++i) {
        double rti = s->rho[i] - OBC_OFFSET;
        for (int j = i + 1; j < n; ++j) {
            double dx = pos[3 * j] - pos[3 * i];
            double dy = pos[3 * j + 1] - pos[3 * i + 1];
            double dz = pos[3 * j + 2] - pos[3 * i + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > cutoff_sq || r2 < 1e-18) continue;
            double r = sqrt(r2);
            double rtj = s->rho[j] - OBC_OFFSET;
            integral[i] += orc_pairwise_descreening(r, rti, s->scale[j] * rtj);
            integral[j] += orc_pairwise_descreening(r, rtj, s->scale[i] * rti);
        }
    }
    int clamped = 0;
    for (int i = 0; i < n; ++i) {
        double rt = s->rho[i] - OBC_OFFSET;
        double rho = s->rho[i];
        double psi = integral[i] * rt;
        double arg = OBC_ALPHA * psi - OBC_BETA * psi * psi + OBC_GAMMA * psi * psi * psi;
        double inv = 1.0 / rt - tanh(arg) / rho;
        double floor_r = rt > 0.5 ? rt : 0.5;
        if (inv <= 0.0 || !isfinite(inv)) {
            clamped++;
            radii[i] = floor_r;
        } else {
            double r_eff = 1.0 / inv;
            radii[i] = r_eff > floor_r ? r_eff : floor_r;
        }
    }
    return clamped;
}

/* add_gb_forces_soa pair loop, energy/src/forces_gb.rs:64-149 */
void orc_gb_forces(const orc_sys *s, const double *pos, const double *radii, double cutoff,
                   double *f) {
    const int n = s->n;
    const double prefactor_kj = (1.0 - 1.0 / EPS_WATER) * COULOMB_KCAL * KCAL_TO_KJ;
    const double cutoff_sq = cutoff * cutoff;
    for (int i = 0; i < n; ++i) {
        double qi = s->q[i];
        if (qi == 0.0) continue;
        for (int j = i + 1; j < n; ++j) {
            double qj = s->q[j];
            if (qj == 0.0) continue;
            double dx = pos[3 * j] - pos[3 * i];
            double dy = pos[3 * j + 1] - pos[3 * i + 1];
            double dz = pos[3 * j + 2] - pos[3 * i + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > cutoff_sq || r2 < 1e-18) continue;
            double r = sqrt(r2);
            double rij = radii[i] * radii[j];
            double e = exp(-r2 / (4.0 * rij));
            double fgb2 = r2 + rij * e;
            double fgb = sqrt(fgb2);
            double dfgb2 = 2.0 * r - 0.5 * r * e;
            double dfgb = dfgb2 / (2.0 * fgb);
            double coeff = prefactor_kj * (qi * qj) / fgb2 * dfgb / r;
            f[3 * i] += dx * coeff; f[3 * i + 1] += dy * coeff; f[3 * i + 2] += dz * coeff;
            f[3 * j] -= dx * coeff; f[3 * j + 1] -= dy * coeff; f[3 * j + 2] -= dz * coeff;
        }
    }
}

/* gb_energy, energy/src/gb.rs:146-182 (uncut pair sum).  out[0]=total,
 * out[1]=self, out[2]=pair (kJ/mol). */
void orc_gb_energy(const orc_sys *s, const double *pos, const double *radii, double *out) {
    const int n = s->n;
    const double prefactor = -0.5 * (1.0 - 1.0 / EPS_WATER) * COULOMB_KCAL;
    double self_e = 0.0, pair_e = 0.0;
    for (int i = 0; i < n; ++i) {
        if (s->q[i] != 0.0) self_e += prefactor * s->q[i] * s->q[i] / radii[i];
    }
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            double qq = s->q[i] * s->q[j];
            if (qq == 0.0) continue;
            double dx = pos[3 * i] - pos[3 * j];
            double dy = pos[3 * i + 1] - pos[3 * j + 1];
            double dz = pos[3 * i + 2] - pos[3 * j + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            double rp = radii[i] * radii[j];
            double fgb = sqrt(r2 + rp * exp(-r2 / (4.0 * rp)));
            pair_e += 2.0 * prefactor * qq / fgb;
        }
    }
    out[1] = self_e * KCAL_TO_KJ;
    out[2] = pair_e * KCAL_TO_KJ;
    out[0] = (self_e + pair_e) * KCAL_TO_KJ;
}

/* ------------------------------------------------------------------------ */
/* Bonded terms, energy/src/forces_bonded.rs:47-236, bonded.rs:59-191       */
/* ------------------------------------------------------------------------ */
/* dihedral_gradient, forces_bonded.rs:124-167.  Returns 0 if degenerate. */
static int dihedral_gradient(v3 pa, v3 pb, v3 pc, v3 pd, v3 *da, v3 *db, v3 *dc, v3 *dd,
                             double *phi) {
    v3 b1 = vsub(pb, pa), b2 = vsub(pc, pb), b3 = vsub(pd, pc);
    double b2n = vnorm(b2);
    if (b2n < 1e-9) return 0;
    v3 s = vcross(b1, b2), t = vcross(b2, b3);
    double s2 = vnorm2(s), t2 = vnorm2(t);
    if (s2 < 1e-18 || t2 < 1e-18) return 0;
    *da = vmul(s, b2n / s2);
    *dd = vmul(vneg(t), b2n / t2);
    double inv_b2 = 1.0 / (b2n * b2n);
    double f1 = vdot(b1, b2) * inv_b2;
    double f2 = vdot(b3, b2) * inv_b2;
    *db = vadd(vmul(*da, -(f1 + 1.0)), vmul(*dd, f2));
    *dc = vneg(vadd(vadd(*da, *db), *dd));
    v3 m1 = vcross(s, vdiv(b2, b2n));
    *phi = atan2(vdot(m1, t), vdot(s, t));
    return 1;
}

void orc_bonded_forces(const orc_sys *s, const double *pos, double *f) {
    for (int k = 0; k < s->nb; ++k) { /* add_bond_forces :47-70 */
        int a = s->bonds[2 * k], b = s->bonds[2 * k + 1];
        v3 d = vsub(V(pos + 3 * b), V(pos + 3 * a));
        double r = vnorm(d);
        if (r < 1e-9) continue;
        double mag = 2.0 * s->bond_p[2 * k] * (r - s->bond_p[2 * k + 1]) * KCAL_TO_KJ;
        v3 fv = vmul(d, mag / r);
        vacc(f + 3 * a, fv);
        vdec(f + 3 * b, fv);
    }
    for (int k = 0; k < s->na; ++k) { /* add_angle_forces :77-116 */
        int a = s->angles[3 * k], b = s->angles[3 * k + 1], c = s->angles[3 * k + 2];
        v3 u = vsub(V(pos + 3 * a), V(pos + 3 * b));
        v3 v = vsub(V(pos + 3 * c), V(pos + 3 * b));
        double un = vnorm(u), vn = vnorm(v);
        if (un < 1e-9 || vn < 1e-9) continue;
        v3 uh = vdiv(u, un), vh = vdiv(v, vn);
        double ct = vdot(uh, vh);
        if (ct < -1.0) ct = -1.0;
        if (ct > 1.0) ct = 1.0;
        double st = sqrt(1.0 - ct * ct);
        if (st < 1e-9) continue;
        double theta = acos(ct);
        double dv = 2.0 * s->angle_p[2 * k] * (theta - s->angle_p[2 * k + 1]);
        double coeff = dv * KCAL_TO_KJ / st;
        v3 fa = vmul(vsub(vh, vmul(uh, ct)), coeff / un);
        v3 fc = vmul(vsub(uh, vmul(vh, ct)), coeff / vn);
        v3 fb = vneg(vadd(fa, fc));
        vacc(f + 3 * a, fa);
        vacc(f + 3 * b, fb);
        vacc(f + 3 * c, fc);
    }
    /* add_dihedral_forces :172-202; one row per (dihedral, periodic term):
     * the sum of per-term forces equals origami's summed dV/dφ times ∂φ/∂r. */
    for (int k = 0; k < s->nd; ++k) {
        const int *q = s->dihedrals + 4 * k;
        v3 da, db, dc, dd;
        double phi;
        if (!dihedral_gradient(V(pos + 3 * q[0]), V(pos + 3 * q[1]), V(pos + 3 * q[2]),
                               V(pos + 3 * q[3]), &da, &db, &dc, &dd, &phi))
            continue;
        double kk = s->dih_p[3 * k], nn = s->dih_p[3 * k + 1], delta = s->dih_p[3 * k + 2];
        double dvdphi = -kk * nn * sin(nn * phi - delta) * KCAL_TO_KJ;
        vdec(f + 3 * q[0], vmul(da, dvdphi));
        vdec(f + 3 * q[1], vmul(db, dvdphi));
        vdec(f + 3 * q[2], vmul(dc, dvdphi));
        vdec(f + 3 * q[3], vmul(dd, dvdphi));
    }
    for (int k = 0; k < s->ni; ++k) { /* add_improper_forces :207-236 */
        const int *q = s->impropers + 4 * k;
        v3 da, db, dc, dd;
        double om;
        if (!dihedral_gradient(V(pos + 3 * q[0]), V(pos + 3 * q[1]), V(pos + 3 * q[2]),
                               V(pos + 3 * q[3]), &da, &db, &dc, &dd, &om))
            continue;
        double dom = om - s->imp_p[2 * k + 1];
        while (dom > M_PI) dom -= 2.0 * M_PI;
        while (dom < -M_PI) dom += 2.0 * M_PI;
        double dv = 2.0 * s->imp_p[2 * k] * dom * KCAL_TO_KJ;
        vdec(f + 3 * q[0], vmul(da, dv));
        vdec(f + 3 * q[1], vmul(db, dv));
        vdec(f + 3 * q[2], vmul(dc, dv));
        vdec(f + 3 * q[3], vmul(dd, dv));
    }
}

/* measure::dihedral, geom/src/measure.rs:17-28 */
static double measure_dihedral(v3 a, v3 b, v3 c, v3 d) {
    v3 b0 = vsub(a, b);
    v3 b1 = vsub(c, b);
    b1 = vdiv(b1, vnorm(b1));
    v3 b2 = vsub(d, c);
    v3 v = vsub(b0, vmul(b1, vdot(b0, b1)));
    v3 w = vsub(b2, vmul(b1, vdot(b2, b1)));
    return atan2(vdot(vcross(b1, v), w), vdot(v, w));
}

/* bonded_energy, energy/src/bonded.rs:34-191.  out[0..3] = bond, angle,
 * dihedral, improper (kJ/mol). */
void orc_bonded_energy(const orc_sys *s, const double *pos, double *out) {
    double eb = 0, ea = 0, ed = 0, ei = 0;
    for (int k = 0; k < s->nb; ++k) {
        int a = s->bonds[2 * k], b = s->bonds[2 * k + 1];
        double r = vnorm(vsub(V(pos + 3 * b), V(pos + 3 * a)));
        double dr = r - s->bond_p[2 * k + 1];
        eb += s->bond_p[2 * k] * dr * dr;
    }
    for (int k = 0; k < s->na; ++k) {
        int a = s->angles[3 * k], b = s->angles[3 * k + 1], c = s->angles[3 * k + 2];
        v3 ba = vsub(V(pos + 3 * a), V(pos + 3 * b));
        v3 bc = vsub(V(pos + 3 * c), V(pos + 3 * b));
        ba = vdiv(ba, vnorm(ba));
        bc = vdiv(bc, vnorm(bc));
        double ct = vdot(ba, bc);
        if (ct < -1.0) ct = -1.0;
        if (ct > 1.0) ct = 1.0;
        double dt = acos(ct) - s->angle_p[2 * k + 1];
        ea += s->angle_p[2 * k] * dt * dt;
    }
    for (int k = 0; k < s->nd; ++k) {
        const int *q = s->dihedrals + 4 * k;
        double chi = measure_dihedral(V(pos + 3 * q[0]), V(pos + 3 * q[1]), V(pos + 3 * q[2]),
                                      V(pos + 3 * q[3]));
        ed += s->dih_p[3 * k] * (1.0 + cos(s->dih_p[3 * k + 1] * chi - s->dih_p[3 * k + 2]));
    }
    for (int k = 0; k < s->ni; ++k) {
        const int *q = s->impropers + 4 * k;
        double om = measure_dihedral(V(pos + 3 * q[0]), V(pos + 3 * q[1]), V(pos + 3 * q[2]),
                                     V(pos + 3 * q[3]));
        double dom = om - s->imp_p[2 * k + 1];
        while (dom > M_PI) dom -= 2.0 * M_PI;
        while (dom < -M_PI) dom += 2.0 * M_PI;
        ei += s->imp_p[2 * k] * dom * dom;
    }
    out[0] = eb * KCAL_TO_KJ;
    out[1] = ea * KCAL_TO_KJ;
    out[2] = ed * KCAL_TO_KJ;
    out[3] = ei * KCAL_TO_KJ;
}

/* ------------------------------------------------------------------------ */
/* Aggregates                                                                */
/* ------------------------------------------------------------------------ */
/* total_force_with_scratch, energy/src/forces.rs:78-128 (without SASA; the
 * SASA term is added by orc_sasa_forces in oracle_sasa.c).  terms (may be
 * NULL) receives 3 N-vectors: bonded, LJ+Coulomb, GB. */
void orc_total_force(const orc_sys *s, const double *pos, double cut_nb, double cut_gb,
                     double cut_descreen, double *f, double *radii_out, double *terms) {
    const int n = s->n;
    double *fb = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *fn = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *fg = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *integral = (double *)malloc((size_t)n * sizeof(double));
    double *radii = radii_out ? radii_out : (double *)malloc((size_t)n * sizeof(double));
    orc_bonded_forces(s, pos, fb);
    orc_nonbonded_forces(s, pos, cut_nb, fn);
    orc_born_radii(s, pos, cut_descreen, integral, radii);
    orc_gb_forces(s, pos, radii, cut_gb, fg);
    for (int k = 0; k < 3 * n; ++k) f[k] = fb[k] + (fn[k] + fg[k]);
    if (terms) {
        memcpy(terms, fb, 3 * (size_t)n * sizeof(double));
        memcpy(terms + 3 * (size_t)n, fn, 3 * (size_t)n * sizeof(double));
        memcpy(terms + 6 * (size_t)n, fg, 3 * (size_t)n * sizeof(double));
    }
    free(fb); free(fn); free(fg); free(integral);
    if (!radii_out) free(radii);
}

/* total_energy_with_options, dynamics/src/energy_eval.rs:20-34 (no SASA).
 * out: [0]bond [1]angle [2]dihedral [3]improper [4]LJ [5]Coulomb [6]GB total
 * [7]GB self [8]GB pair [9]total. */
void orc_total_energy(const orc_sys *s, const double *pos, double cut_nb, double cut_descreen,
                      double *out) {
    const int n = s->n;
    double b[4], nb[4], gb[3];
    double *integral = (double *)malloc((size_t)n * sizeof(double));
    double *radii = (double *)malloc((size_t)n * sizeof(double));
    orc_bonded_energy(s, pos, b);
    orc_nonbonded_energy(s, pos, cut_nb, nb);
    orc_born_radii(s, pos, cut_descreen, integral, radii);
    orc_gb_energy(s, pos, radii, gb);
    out[0] = b[0]; out[1] = b[1]; out[2] = b[2]; out[3] = b[3];
    out[4] = nb[0]; out[5] = nb[1];
    out[6] = gb[0]; out[7] = gb[1]; out[8] = gb[2];
    out[9] = (b[0] + b[1] + b[2] + b[3]) + nb[0] + nb[1] + gb[0];
    free(integral); free(radii);
}

/* ------------------------------------------------------------------------ */
/* SHAKE and BAOAB                                                           */
/* ------------------------------------------------------------------------ */
/* shake_iterate, dynamics/src/shake.rs:86-125.  Returns iterations or -1. */
int orc_shake(const orc_sys *s, double *pos, const double *ref, double tol_sq, int max_iters) {
    for (int iter = 1; iter <= max_iters; ++iter) {
        double max_err = 0.0;
        for (int k = 0; k < s->nc; ++k) {
            int i = s->cons[2 * k], j = s->cons[2 * k + 1];
            double d = s->cons_d[k];
            v3 rn = vsub(V(pos + 3 * i), V(pos + 3 * j));
            double err = vnorm2(rn) - d * d;
            if (fabs(err) > max_err) max_err = fabs(err);
            if (fabs(err) < tol_sq) continue;
            v3 ro = vsub(V(ref + 3 * i), V(ref + 3 * j));
            double denom = 2.0 * (1.0 / s->mass[i] + 1.0 / s->mass[j]) * vdot(rn, ro);
            if (fabs(denom) < 1e-18) continue;
            v3 delta = vmul(ro, err / denom);
            vdec(pos + 3 * i, vmul(delta, 1.0 / s->mass[i]));
            vacc(pos + 3 * j, vmul(delta, 1.0 / s->mass[j]));
        }
        if (max_err < tol_sq) return iter;
    }
    return -1;
}

/* One BAOAB step, dynamics/src/langevin.rs:223-313, with the O-step noise
 * supplied by the caller (xi: 3N standard normals, atom-major x,y,z).
 * forces: in = F(x_n), out = F(x_{n+1}).  Returns SHAKE failures (0..2). */
int orc_baoab_step(const orc_sys *s, double *pos, double *vel, double *forces, const double *xi,
                   double dt, double temperature, double gamma_ps, int use_shake,
                   double cut_nb, double cut_gb, double cut_descreen) {
    const int n = s->n;
    const double half_dt = 0.5 * dt;
    const double alpha = exp(-gamma_ps * dt * 1.0e-3);
    const double kbt = BOLTZMANN_KJ * temperature;
    const double total_dof = 3.0 * n;
    double dof_correction = 1.0;
    if (use_shake && s->nc > 0) {
        double rem = total_dof - s->nc;
        dof_correction = total_dof / (rem > 1.0 ? rem : 1.0);
    }
    int shake = use_shake && s->nc > 0;
    double *refb = (double *)malloc(3 * (size_t)n * sizeof(double));
    int failures = 0;
    for (int i = 0; i < n; ++i) {
        double k = ACCEL_FACTOR / s->mass[i] * half_dt;
        for (int c = 0; c < 3; ++c) vel[3 * i + c] += forces[3 * i + c] * k;
    }
    for (int half = 0; half < 2; ++half) {
        if (half == 1) {
            double one_minus = 1.0 - alpha * alpha;
            for (int i = 0; i < n; ++i) {
                double sigma = sqrt(one_minus * kbt * dof_correction * ACCEL_FACTOR / s->mass[i]);
                for (int c = 0; c < 3; ++c)
                    vel[3 * i + c] = alpha * vel[3 * i + c] + sigma * xi[3 * i + c];
            }
        }
        if (shake) memcpy(refb, pos, 3 * (size_t)n * sizeof(double));
        for (int k = 0; k < 3 * n; ++k) pos[k] += vel[k] * half_dt;
        if (shake) {
            if (orc_shake(s, pos, refb, 1e-6, 64) < 0) failures++;
            for (int k = 0; k < 3 * n; ++k) vel[k] = (pos[k] - refb[k]) / half_dt;
        }
    }
    orc_total_force(s, pos, cut_nb, cut_gb, cut_descreen, forces, NULL, NULL);
    for (int i = 0; i < n; ++i) {
        double k = ACCEL_FACTOR / s->mass[i] * half_dt;
        for (int c = 0; c < 3; ++c) vel[3 * i + c] += forces[3 * i + c] * k;
    }
    free(refb);
    return failures;
}

/* kinetic_energy_kj_mol, dynamics/src/langevin.rs:469-475 */
double orc_kinetic_energy(const orc_sys *s, const double *vel) {
    double sum = 0.0;
    for (int i = 0; i < s->n; ++i) {
        double v2 = vel[3 * i] * vel[3 * i] + vel[3 * i + 1] * vel[3 * i + 1] +
                    vel[3 * i + 2] * vel[3 * i + 2];
        sum += s->mass[i] * v2;
    }
    return 0.5 * sum / ACCEL_FACTOR;
}

/* ------------------------------------------------------------------------ */
/* RNG: xoshiro256++ seeded by SplitMix64, dynamics/src/rng.rs:18-98        */
/* ------------------------------------------------------------------------ */
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* SplitMix64::next, rng.rs:87-97: the seeding generator. */
void orc_splitmix64(uint64_t seed, int count, uint64_t *out) {
    uint64_t sm = seed;
    for (int k = 0; k < count; ++k) {
        sm += 0x9e3779b97f4a7c15ULL;
        uint64_t z = sm;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        out[k] = z ^ (z >> 31);
    }
}

/* Xoshiro256pp::next_u64, rng.rs:40-55, from an explicit state st[4]
 * (advanced in place). */
void orc_xoshiro_from_state(uint64_t *st, int count, uint64_t *out) {
    for (int k = 0; k < count; ++k) {
        uint64_t result = rotl64(st[0] + st[3], 23) + st[0];
        uint64_t t = st[1] << 17;
        st[2] ^= st[0];
        st[3] ^= st[1];
        st[1] ^= st[2];
        st[0] ^= st[3];
        st[2] ^= t;
        st[3] = rotl64(st[3], 45);
        out[k] = result;
    }
}

/* Xoshiro256pp::from_seed (rng.rs:27-37): state = four SplitMix64 outputs. */
void orc_xoshiro_u64(uint64_t seed, int count, uint64_t *out) {
    uint64_t st[4];
    orc_splitmix64(seed, 4, st);
    orc_xoshiro_from_state(st, count, out);
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based
 * generator the GPU integrator draws its O-step noise from.  The oracle
 * regenerates the same stream so both sides integrate with identical noise. */
void orc_philox4x32(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2,
                    uint32_t c3, uint32_t *out) {
    uint32_t c[4] = {c0, c1, c2, c3};
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        uint32_t n3 = lo0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

/* The O-step noise the GPU integrator draws for (system, atom, step): two
 * Philox4x32-10 blocks with counter (atom, step, system, sub), 53-bit
 * uniforms u = (bits + 0.5) 2^-53, Box-Muller.  Same stream as the engine. */
static double u53(uint32_t hi, uint32_t lo) {
    uint64_t v = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)v + 0.5) * (1.0 / 9007199254740992.0);
}

void orc_philox_normals(uint32_t k0, uint32_t k1, uint32_t sys, int n_atoms, uint32_t step,
                        double *out) {
    for (int a = 0; a < n_atoms; ++a) {
        uint32_t c[4];
        orc_philox4x32(k0, k1, (uint32_t)a, step, sys, 0u, c);
        double u1 = u53(c[0], c[1]), u2 = u53(c[2], c[3]);
        double r = sqrt(-2.0 * log(u1)), th = 2.0 * M_PI * u2;
        out[3 * a] = r * cos(th);
        out[3 * a + 1] = r * sin(th);
        orc_philox4x32(k0, k1, (uint32_t)a, step, sys, 1u, c);
        u1 = u53(c[0], c[1]);
        u2 = u53(c[2], c[3]);
        r = sqrt(-2.0 * log(u1));
        th = 2.0 * M_PI * u2;
        out[3 * a + 2] = r * cos(th);
    }
}

/* ------------------------------------------------------------------------ */
/* Cylindrical exit tunnel, dynamics/src/cotranslate.rs:131-155: atoms with  */
/* 0 <= (x - o).d <= length and radial distance beyond radius feel          */
/* -k (dist - radius) along the outward radial unit vector (added to f).    */
/* ------------------------------------------------------------------------ */
void orc_tunnel_force(int n, const double *pos, const double *origin, const double *dir, double radius,
                      double length, double k, double *f) {
    for (int i = 0; i < n; ++i) {
        const double rx = pos[3 * i] - origin[0], ry = pos[3 * i + 1] - origin[1], rz = pos[3 * i + 2] - origin[2];
        const double along = rx * dir[0] + ry * dir[1] + rz * dir[2];
        if (along < 0.0 || along > length) continue;
        const double px = rx - dir[0] * along, py = ry - dir[1] * along, pz = rz - dir[2] * along;
        const double d = sqrt(px * px + py * py + pz * pz);
        if (d > radius) {
            const double over = d - radius;
            f[3 * i] -= px / d * (k * over);
            f[3 * i + 1] -= py / d * (k * over);
            f[3 * i + 2] -= pz / d * (k * over);
        }
    }
}
