/*
 * FP64 CPU oracle, PowerSasa part: exact analytical SASA (spherical
 * Gauss-Bonnet over buried caps) and its analytical gradient.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.c).  Restates, in order:
 *   energy/src/powersasa/geometry.rs     SmallCircle, intersect_circles
 *   energy/src/powersasa/arrangement.rs  build_caps, find_boundary and helpers,
 *                                        count_accessible_components
 *   energy/src/powersasa/area.rs         bounded_area, walk_faces_summing_vertices
 *   energy/src/powersasa/derivatives.rs  vertex_point, identify_root_sign,
 *                                        cos_alpha_grad, vertex_point_jvp,
 *                                        arc_theta_jvp, vertex_epsilon_jvp
 *   energy/src/powersasa/analytical.rs   build_atom_cache, atom_area_gradient,
 *                                        directional_area_derivative,
 *                                        add_sasa_forces_analytical
 *   energy/src/powersasa/mod.rs          powersasa_energy
 * The force path follows the inline variant add_sasa_forces_analytical
 * (analytical.rs:583-649): neighbour lists by brute force, ascending index.
 * Loop count and accessible-component count are computed by origami's cache
 * builder but never read by the gradient (analytical.rs:285-421); the force
 * path here skips them, the area path computes them.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct { double x, y, z; } v3;
static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vmul(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vnorm2(v3 a) { return vdot(a, a); }
static inline double vnorm(v3 a) { return sqrt(vdot(a, a)); }
static inline v3 vnormalize(v3 a) { return vdiv(a, vnorm(a)); }
static inline v3 P(const double *p) { return mk(p[0], p[1], p[2]); }

typedef struct { v3 axis; double cos_alpha; } cap_t;

/* SmallCircle::new normalises the axis (geometry.rs:26-28). */
static cap_t small_circle_new(v3 axis, double cos_alpha) {
    cap_t c; c.axis = vnormalize(axis); c.cos_alpha = cos_alpha; return c;
}
static double cap_sin_alpha(cap_t c) {
    double t = 1.0 - c.cos_alpha * c.cos_alpha;
    return sqrt(t > 0.0 ? t : 0.0);
}
static int contains_strict(cap_t c, v3 p) { return vdot(p, c.axis) > c.cos_alpha + 1e-12; }

enum { INT_TWO, INT_TANGENT, INT_DISJOINT, INT_COINCIDENT };

/* intersect_circles, geometry.rs:67-100 */
static int intersect_circles(cap_t c1, cap_t c2, v3 *p, v3 *q) {
    double sigma = vdot(c1.axis, c2.axis);
    double denom = 1.0 - sigma * sigma;
    if (fabs(denom) < 1e-12) {
        return fabs(c1.cos_alpha - sigma * c2.cos_alpha) < 1e-9 ? INT_COINCIDENT : INT_DISJOINT;
    }
    double a = (c1.cos_alpha - sigma * c2.cos_alpha) / denom;
    double b = (c2.cos_alpha - sigma * c1.cos_alpha) / denom;
    double c_sq = 1.0 - a * a - b * b - 2.0 * a * b * sigma;
    if (c_sq < -1e-9) return INT_DISJOINT;
    v3 base = vadd(vmul(c1.axis, a), vmul(c2.axis, b));
    if (fabs(c_sq) <= 1e-9) { *p = base; return INT_TANGENT; }
    double c = sqrt(c_sq);
    v3 nc = vcross(c1.axis, c2.axis);
    v3 n_perp = vdiv(nc, vnorm(nc));
    v3 off = vmul(n_perp, c);
    *p = vadd(base, off);
    *q = vsub(base, off);
    return INT_TWO;
}

/* helper axis choice shared by angular_position / sample_circle_point */
static v3 helper_axis(v3 axis) {
    double ax = fabs(axis.x), ay = fabs(axis.y), az = fabs(axis.z);
    if (ax < ay && ax < az) return mk(1, 0, 0);
    if (ay < az) return mk(0, 1, 0);
    return mk(0, 0, 1);
}

/* angular_position, arrangement.rs:233-248 */
static double angular_position(v3 point, cap_t c) {
    v3 h = helper_axis(c.axis);
    v3 u = vnormalize(vsub(h, vmul(c.axis, vdot(h, c.axis))));
    v3 v = vcross(c.axis, u);
    v3 proj = vsub(point, vmul(c.axis, vdot(point, c.axis)));
    return atan2(vdot(proj, v), vdot(proj, u));
}

/* sample_circle_point, arrangement.rs:252-263 */
static v3 sample_circle_point(cap_t c) {
    v3 h = helper_axis(c.axis);
    v3 perp = vnormalize(vsub(h, vmul(c.axis, vdot(h, c.axis))));
    return vadd(vmul(c.axis, c.cos_alpha), vmul(perp, cap_sin_alpha(c)));
}

/* midpoint_on_circle, arrangement.rs:267-277 */
static v3 midpoint_on_circle(v3 point, cap_t c, double delta) {
    v3 axial = vmul(c.axis, vdot(point, c.axis));
    v3 radial = vsub(point, axial);
    double rn = vnorm(radial);
    if (rn < 1e-12) return point;
    v3 u = vdiv(radial, rn);
    v3 v = vcross(c.axis, u);
    return vadd(axial, vmul(vadd(vmul(u, cos(delta)), vmul(v, sin(delta))), rn));
}

/* wrap_2pi_positive, arrangement.rs:279-289 */
static double wrap_2pi_positive(double t) {
    const double two_pi = 2.0 * M_PI;
    while (t <= 0.0) t += two_pi;
    while (t > two_pi) t -= two_pi;
    return t;
}

/* every_point_buried, arrangement.rs:291-309 */
static int every_point_buried(const cap_t *caps, int n) {
    for (int i = 0; i < 200; ++i) {
        double phi = (double)i * 0.39 + 0.1;
        double th = (double)i * 0.71 + 0.2;
        v3 p = vnormalize(mk(sin(phi) * cos(th), sin(phi) * sin(th), cos(phi)));
        int any = 0;
        for (int k = 0; k < n && !any; ++k) any = contains_strict(caps[k], p);
        if (!any) return 0;
    }
    return 1;
}

/* count_accessible_components, arrangement.rs:327-389.  The golden-spiral
 * probes are filled once when the library loads (read-only afterwards, so
 * concurrent callers are safe). */
enum { PC = 500 };
static v3 probes[PC];

__attribute__((constructor)) static void init_probes(void) {
    double golden = M_PI * (3.0 - sqrt(5.0));
    for (int i = 0; i < PC; ++i) {
        double z = 1.0 - (2.0 * i + 1.0) / (double)PC;
        double t = 1.0 - z * z;
        double r = sqrt(t > 0.0 ? t : 0.0);
        double th = golden * i;
        probes[i] = mk(r * cos(th), r * sin(th), z);
    }
}

static int count_accessible_components(const cap_t *caps, int n) {
    if (n == 0) return 1;
    int acc[PC], parent[PC];
    int any_acc = 0;
    for (int i = 0; i < PC; ++i) {
        int buried = 0;
        for (int k = 0; k < n && !buried; ++k) buried = contains_strict(caps[k], probes[i]);
        acc[i] = !buried;
        any_acc |= acc[i];
        parent[i] = i;
    }
    if (!any_acc) return 0;
    double spacing = sqrt(4.0 * M_PI / PC);
    double thresh = 2.5 * spacing;
    if (thresh > M_PI) thresh = M_PI;
    double cos_thresh = cos(thresh);
    for (int i = 0; i < PC; ++i) {
        if (!acc[i]) continue;
        for (int j = i + 1; j < PC; ++j) {
            if (!acc[j]) continue;
            if (vdot(probes[i], probes[j]) > cos_thresh) {
                int ri = i, rj = j;
                while (parent[ri] != ri) { parent[ri] = parent[parent[ri]]; ri = parent[ri]; }
                while (parent[rj] != rj) { parent[rj] = parent[parent[rj]]; rj = parent[rj]; }
                if (ri != rj) parent[ri] = rj;
            }
        }
    }
    int roots = 0;
    int seen[PC];
    for (int i = 0; i < PC; ++i) {
        if (!acc[i]) continue;
        int r = i;
        while (parent[r] != r) { parent[r] = parent[parent[r]]; r = parent[r]; }
        int found = 0;
        for (int k = 0; k < roots; ++k) if (seen[k] == r) { found = 1; break; }
        if (!found) seen[roots++] = r;
    }
    return roots;
}

/* ------------------------------------------------------------------------ */
/* Boundary (find_boundary, arrangement.rs:100-228)                          */
/* ------------------------------------------------------------------------ */
typedef struct { double ang; v3 pt; int other; } trans_t;
typedef struct { int cap; v3 start, end; double theta; int full; } arc_t;
typedef struct { v3 point; int in_cap, out_cap; } vert_t;

enum { B_EXPOSED, B_BURIED, B_BOUNDED };

typedef struct {
    int kind;
    int n_arcs, n_verts;
    arc_t *arcs;
    vert_t *verts;
} boundary_t;

static void boundary_free(boundary_t *b) { free(b->arcs); free(b->verts); }

/* stable insertion sort on angle (Rust's sort_by is stable) */
static void sort_trans(trans_t *t, int m) {
    for (int i = 1; i < m; ++i) {
        trans_t x = t[i];
        int j = i - 1;
        while (j >= 0 && t[j].ang > x.ang) { t[j + 1] = t[j]; --j; }
        t[j + 1] = x;
    }
}

static boundary_t find_boundary(const cap_t *caps, int n) {
    boundary_t B;
    memset(&B, 0, sizeof(B));
    if (n == 0) { B.kind = B_EXPOSED; return B; }
    int cap_max = 2 * n;
    trans_t *tr = (trans_t *)malloc(sizeof(trans_t) * (size_t)(cap_max > 0 ? cap_max : 1));
    B.arcs = (arc_t *)malloc(sizeof(arc_t) * (size_t)(n * cap_max + n + 1));
    B.verts = (vert_t *)malloc(sizeof(vert_t) * (size_t)(n * cap_max + 1));
    for (int k = 0; k < n; ++k) {
        int m = 0;
        for (int o = 0; o < n; ++o) {
            if (o == k) continue;
            v3 p, q;
            int r = intersect_circles(caps[k], caps[o], &p, &q);
            if (r == INT_TWO) {
                tr[m].ang = angular_position(p, caps[k]); tr[m].pt = p; tr[m].other = o; ++m;
                tr[m].ang = angular_position(q, caps[k]); tr[m].pt = q; tr[m].other = o; ++m;
            } else if (r == INT_TANGENT) {
                tr[m].ang = angular_position(p, caps[k]); tr[m].pt = p; tr[m].other = o; ++m;
            }
        }
        sort_trans(tr, m);
        /* Vec::dedup_by: drop an element equal to the last one kept. */
        int w = 0;
        for (int i = 0; i < m; ++i) {
            if (w > 0 && vnorm2(vsub(tr[i].pt, tr[w - 1].pt)) < 1e-12) continue;
            tr[w++] = tr[i];
        }
        m = w;
        if (m == 0) {
            v3 probe = sample_circle_point(caps[k]);
            int buried = 0;
            for (int o = 0; o < n && !buried; ++o) buried = (o != k) && contains_strict(caps[o], probe);
            if (!buried) {
                arc_t *a = &B.arcs[B.n_arcs++];
                a->cap = k; a->start = probe; a->end = probe; a->theta = -2.0 * M_PI; a->full = 1;
            }
            continue;
        }
        for (int i = 0; i < m; ++i) {
            trans_t A = tr[i], Bt = tr[(i + 1) % m];
            double ccw = wrap_2pi_positive(Bt.ang - A.ang);
            if (ccw < 1e-9) continue;
            v3 mid = midpoint_on_circle(A.pt, caps[k], ccw / 2.0);
            int bad = 0;
            for (int o = 0; o < n && !bad; ++o) bad = (o != k) && contains_strict(caps[o], mid);
            if (bad) continue;
            arc_t *a = &B.arcs[B.n_arcs++];
            a->cap = k; a->start = Bt.pt; a->end = A.pt; a->theta = -ccw; a->full = 0;
            vert_t *v = &B.verts[B.n_verts++];
            v->point = A.pt; v->in_cap = k; v->out_cap = A.other;
        }
    }
    free(tr);
    if (B.n_arcs == 0 && every_point_buried(caps, n)) {
        B.kind = B_BURIED;
        return B;
    }
    B.kind = B_BOUNDED;
    return B;
}

/* walk_faces_summing_vertices, area.rs:140-222 */
static void walk_faces(const boundary_t *B, const cap_t *caps, int *loops_out, double *vsum_out) {
    int n = B->n_arcs;
    int *visited = (int *)calloc((size_t)(n ? n : 1), sizeof(int));
    int loops = 0;
    double vsum = 0.0;
    const double eps_sq = 1e-10, two_pi = 2.0 * M_PI;
    for (int start = 0; start < n; ++start) {
        if (visited[start]) continue;
        if (B->arcs[start].full) { visited[start] = 1; loops++; continue; }
        int cur = start;
        for (;;) {
            visited[cur] = 1;
            v3 end_pt = B->arcs[cur].end;
            v3 t_in = vnormalize(vcross(end_pt, caps[B->arcs[cur].cap].axis));
            int ncand = 0, first = -1;
            for (int i = 0; i < n; ++i) {
                if (!B->arcs[i].full && vnorm2(vsub(B->arcs[i].start, end_pt)) < eps_sq) {
                    if (first < 0) first = i;
                    ncand++;
                }
            }
            if (ncand == 0) break;
            int best = first;
            double best_angle = INFINITY;
            for (int c = 0; c < n; ++c) {
                if (B->arcs[c].full || !(vnorm2(vsub(B->arcs[c].start, end_pt)) < eps_sq)) continue;
                if (c == cur) continue;
                v3 t_out = vnormalize(vcross(end_pt, caps[B->arcs[c].cap].axis));
                double dot = vdot(t_in, t_out);
                if (dot < -1.0) dot = -1.0;
                if (dot > 1.0) dot = 1.0;
                double sg = vdot(end_pt, vcross(t_in, t_out));
                double ang = atan2(sg, dot);
                if (ang <= 1e-9) ang += two_pi;
                if (ang < best_angle) { best_angle = ang; best = c; }
            }
            v3 t_out = vnormalize(vcross(end_pt, caps[B->arcs[best].cap].axis));
            double dot = vdot(t_in, t_out);
            if (dot < -1.0) dot = -1.0;
            if (dot > 1.0) dot = 1.0;
            vsum += atan2(vdot(end_pt, vcross(t_in, t_out)), dot);
            if (best == start) break;
            cur = best;
        }
        loops++;
    }
    free(visited);
    *loops_out = loops;
    *vsum_out = vsum;
}

/* accessible_area / bounded_area, area.rs:19-126 */
static double accessible_area(double radius, const cap_t *caps, int n, const boundary_t *B) {
    double r_sq = radius * radius;
    double four_pi_r2 = 4.0 * M_PI * r_sq;
    if (B->kind == B_EXPOSED) return four_pi_r2;
    if (B->kind == B_BURIED) return 0.0;
    if (B->n_arcs == 0) return 0.0;
    int l;
    double vsum;
    walk_faces(B, caps, &l, &vsum);
    if (l == 0) return 0.0;
    double arc_sum = 0.0;
    for (int k = 0; k < B->n_arcs; ++k) arc_sum += caps[B->arcs[k].cap].cos_alpha * B->arcs[k].theta;
    int c_raw = count_accessible_components(caps, n);
    int c = c_raw < l ? c_raw : l;
    if (c < 1) c = 1;
    int chi = 2 * c - l;
    double area = r_sq * (2.0 * M_PI * (double)chi - arc_sum - vsum);
    if (area < 0.0) area = 0.0;
    if (area > four_pi_r2) area = four_pi_r2;
    return area;
}

/* build_caps, arrangement.rs:69-96.  nb_* are the candidate neighbours in
 * order.  Returns -1 if the atom is fully enclosed (None), else cap count. */
static int build_caps(v3 pi, double ri, int n_nb, const int *nb_idx, const double *pos,
                      const double *radii, cap_t *caps, int *owners) {
    int nc = 0;
    for (int t = 0; t < n_nb; ++t) {
        int j = nb_idx[t];
        v3 pj = P(pos + 3 * j);
        double rj = radii[j];
        double d = vnorm(vsub(pj, pi));
        if (d + ri <= rj) return -1;
        if (d + rj <= ri || d >= ri + rj) continue;
        v3 axis = vdiv(vsub(pj, pi), d);
        double ca = (d * d + ri * ri - rj * rj) / (2.0 * d * ri);
        if (!(ca >= -1.0 && ca <= 1.0)) continue;
        caps[nc].axis = axis;
        caps[nc].cos_alpha = ca;
        owners[nc] = j;
        nc++;
    }
    return nc;
}

/* ------------------------------------------------------------------------ */
/* Derivative primitives (derivatives.rs)                                    */
/* ------------------------------------------------------------------------ */
/* vertex_point, derivatives.rs:92-112.  sign = +1 / -1.  Returns 0 = None. */
static int vertex_point(cap_t c1, cap_t c2, double sign, v3 *out) {
    double sigma = vdot(c1.axis, c2.axis);
    double denom = 1.0 - sigma * sigma;
    if (fabs(denom) < 1e-12) return 0;
    double a = (c1.cos_alpha - sigma * c2.cos_alpha) / denom;
    double b = (c2.cos_alpha - sigma * c1.cos_alpha) / denom;
    double c_sq = 1.0 - a * a - b * b - 2.0 * a * b * sigma;
    if (c_sq <= 0.0) return 0;
    v3 base = vadd(vmul(c1.axis, a), vmul(c2.axis, b));
    v3 nc = vcross(c1.axis, c2.axis);
    double nn = vnorm(nc);
    if (nn < 1e-12) return 0;
    v3 off = vmul(vdiv(nc, nn), sqrt(c_sq));
    *out = vadd(base, vmul(off, sign));
    return 1;
}

/* identify_root_sign, derivatives.rs:117-127.  Returns +1/-1, 0 = None. */
static double identify_root_sign(cap_t c1, cap_t c2, v3 ref) {
    v3 vp, vm;
    if (!vertex_point(c1, c2, 1.0, &vp)) return 0.0;
    if (!vertex_point(c1, c2, -1.0, &vm)) return 0.0;
    return vnorm2(vsub(vp, ref)) <= vnorm2(vsub(vm, ref)) ? 1.0 : -1.0;
}

/* cos_alpha_directional via cos_alpha_grad, derivatives.rs:134-200 */
static double cos_alpha_directional(v3 pi, v3 pk, double ri, double rk, v3 dri, v3 drk) {
    v3 rv = vsub(pk, pi);
    double d2 = vnorm2(rv);
    double d = sqrt(d2);
    if (d < 1e-12) return 0.0;
    v3 u = vdiv(rv, d);
    double dcos_dd = (d2 - ri * ri + rk * rk) / (2.0 * d2 * ri);
    v3 gk = vmul(u, dcos_dd);
    v3 gi = vmul(gk, -1.0);
    return vdot(gi, dri) + vdot(gk, drk);
}

/* vertex_point_jvp, derivatives.rs:222-325 */
static int vertex_point_jvp(v3 pi, v3 pk, v3 pl, double ri, double rk, double rl, double sign,
                            v3 dri, v3 drk, v3 drl, v3 *dv_out) {
    v3 rkv = vsub(pk, pi);
    double dik = vnorm(rkv);
    if (dik < 1e-12) return 0;
    v3 wk = vdiv(rkv, dik);
    v3 rlv = vsub(pl, pi);
    double dil = vnorm(rlv);
    if (dil < 1e-12) return 0;
    v3 wl = vdiv(rlv, dil);
    double cak = (dik * dik + ri * ri - rk * rk) / (2.0 * dik * ri);
    double cal = (dil * dil + ri * ri - rl * rl) / (2.0 * dil * ri);
    double sigma = vdot(wk, wl);
    double denom = 1.0 - sigma * sigma;
    if (fabs(denom) < 1e-12) return 0;
    double a = (cak - sigma * cal) / denom;
    double b = (cal - sigma * cak) / denom;
    double c_sq = 1.0 - a * a - b * b - 2.0 * a * b * sigma;
    if (c_sq <= 0.0) return 0;
    double c = sqrt(c_sq);
    v3 m = vcross(wk, wl);
    double mn = sqrt(vnorm2(m));
    if (mn < 1e-12) return 0;
    v3 np = vdiv(m, mn);
    v3 dvk = vsub(drk, dri);
    v3 dwk = vdiv(vsub(dvk, vmul(wk, vdot(wk, dvk))), dik);
    v3 dvl = vsub(drl, dri);
    v3 dwl = vdiv(vsub(dvl, vmul(wl, vdot(wl, dvl))), dil);
    double dcos_k = cos_alpha_directional(pi, pk, ri, rk, dri, drk);
    double dcos_l = cos_alpha_directional(pi, pl, ri, rl, dri, drl);
    double dsig = vdot(dwk, wl) + vdot(wk, dwl);
    double dden = -2.0 * sigma * dsig;
    double num_a = cak - sigma * cal;
    double dnum_a = dcos_k - dsig * cal - sigma * dcos_l;
    double da = (dnum_a * denom - num_a * dden) / (denom * denom);
    double num_b = cal - sigma * cak;
    double dnum_b = dcos_l - dsig * cak - sigma * dcos_k;
    double db = (dnum_b * denom - num_b * dden) / (denom * denom);
    double dc_sq = -2.0 * a * da - 2.0 * b * db - 2.0 * (da * b + a * db) * sigma - 2.0 * a * b * dsig;
    double dc = dc_sq / (2.0 * c);
    v3 dbase = vadd(vadd(vadd(vmul(wk, da), vmul(dwk, a)), vmul(wl, db)), vmul(dwl, b));
    v3 dm = vadd(vcross(dwk, wl), vcross(wk, dwl));
    v3 dnp = vdiv(vsub(dm, vmul(np, vdot(np, dm))), mn);
    *dv_out = vadd(dbase, vmul(vadd(vmul(np, dc), vmul(dnp, c)), sign));
    return 1;
}

static cap_t cap_from(v3 pi, v3 pk, double ri, double rk, double *d_out) {
    v3 rv = vsub(pk, pi);
    double d = vnorm(rv);
    *d_out = d;
    v3 w = vdiv(rv, d);
    double ca = (d * d + ri * ri - rk * rk) / (2.0 * d * ri);
    return small_circle_new(w, ca);
}

/* arc_theta_jvp, derivatives.rs:346-444 */
static int arc_theta_jvp(v3 pi, v3 pk, v3 pl, v3 pm, double ri, double rk, double rl, double rm,
                         double sign_s, double sign_e, v3 dri, v3 drk, v3 drl, v3 drm,
                         double *out) {
    v3 rkv = vsub(pk, pi);
    double dik = vnorm(rkv);
    if (dik < 1e-12) return 0;
    v3 wk = vdiv(rkv, dik);
    v3 dv = vsub(drk, dri);
    v3 dwk = vdiv(vsub(dv, vmul(wk, vdot(wk, dv))), dik);
    double cak = (dik * dik + ri * ri - rk * rk) / (2.0 * dik * ri);
    cap_t ck = small_circle_new(wk, cak);
    double dil, dim;
    if (vnorm(vsub(pl, pi)) < 1e-12) return 0;
    cap_t cl = cap_from(pi, pl, ri, rl, &dil);
    if (vnorm(vsub(pm, pi)) < 1e-12) return 0;
    cap_t cm = cap_from(pi, pm, ri, rm, &dim);
    v3 vs, ve, dvs, dve;
    if (!vertex_point(ck, cl, sign_s, &vs)) return 0;
    if (!vertex_point(ck, cm, sign_e, &ve)) return 0;
    if (!vertex_point_jvp(pi, pk, pl, ri, rk, rl, sign_s, dri, drk, drl, &dvs)) return 0;
    if (!vertex_point_jvp(pi, pk, pm, ri, rk, rm, sign_e, dri, drk, drm, &dve)) return 0;
    v3 us, dus, ue, due;
    {
        double vdo = vdot(vs, wk);
        v3 proj = vsub(vs, vmul(wk, vdo));
        double dvdo = vdot(dvs, wk) + vdot(vs, dwk);
        v3 dproj = vsub(vsub(dvs, vmul(dwk, vdo)), vmul(wk, dvdo));
        double nn = vnorm(proj);
        if (nn < 1e-12) return 0;
        us = vdiv(proj, nn);
        dus = vdiv(vsub(dproj, vmul(us, vdot(us, dproj))), nn);
    }
    {
        double vdo = vdot(ve, wk);
        v3 proj = vsub(ve, vmul(wk, vdo));
        double dvdo = vdot(dve, wk) + vdot(ve, dwk);
        v3 dproj = vsub(vsub(dve, vmul(dwk, vdo)), vmul(wk, dvdo));
        double nn = vnorm(proj);
        if (nn < 1e-12) return 0;
        ue = vdiv(proj, nn);
        due = vdiv(vsub(dproj, vmul(ue, vdot(ue, dproj))), nn);
    }
    double cos_t = vdot(us, ue);
    double dcos_t = vdot(dus, ue) + vdot(us, due);
    v3 x = vcross(us, ue);
    double sin_t = vdot(wk, x);
    v3 dx = vadd(vcross(dus, ue), vcross(us, due));
    double dsin_t = vdot(dwk, x) + vdot(wk, dx);
    *out = cos_t * dsin_t - sin_t * dcos_t;
    return 1;
}

/* vertex_epsilon_jvp with incoming_is_k = true, derivatives.rs:466-545 */
static int vertex_epsilon_jvp(v3 pi, v3 pk, v3 pl, double ri, double rk, double rl, double sign,
                              v3 dri, v3 drk, v3 drl, double *out) {
    v3 rkv = vsub(pk, pi);
    double dik = vnorm(rkv);
    if (dik < 1e-12) return 0;
    v3 wk = vdiv(rkv, dik);
    v3 dvk = vsub(drk, dri);
    v3 dwk = vdiv(vsub(dvk, vmul(wk, vdot(wk, dvk))), dik);
    v3 rlv = vsub(pl, pi);
    double dil = vnorm(rlv);
    if (dil < 1e-12) return 0;
    v3 wl = vdiv(rlv, dil);
    v3 dvl = vsub(drl, dri);
    v3 dwl = vdiv(vsub(dvl, vmul(wl, vdot(wl, dvl))), dil);
    double cak = (dik * dik + ri * ri - rk * rk) / (2.0 * dik * ri);
    double cal = (dil * dil + ri * ri - rl * rl) / (2.0 * dil * ri);
    cap_t ck = small_circle_new(wk, cak), cl = small_circle_new(wl, cal);
    v3 v, dv;
    if (!vertex_point(ck, cl, sign, &v)) return 0;
    if (!vertex_point_jvp(pi, pk, pl, ri, rk, rl, sign, dri, drk, drl, &dv)) return 0;
    v3 tin, dtin, tout, dtout;
    {
        v3 cr = vcross(v, wk);
        double nn = vnorm(cr);
        if (nn < 1e-12) return 0;
        tin = vdiv(cr, nn);
        v3 dcr = vadd(vcross(dv, wk), vcross(v, dwk));
        dtin = vdiv(vsub(dcr, vmul(tin, vdot(tin, dcr))), nn);
    }
    {
        v3 cr = vcross(v, wl);
        double nn = vnorm(cr);
        if (nn < 1e-12) return 0;
        tout = vdiv(cr, nn);
        v3 dcr = vadd(vcross(dv, wl), vcross(v, dwl));
        dtout = vdiv(vsub(dcr, vmul(tout, vdot(tout, dcr))), nn);
    }
    double ce = vdot(tin, tout);
    double dce = vdot(dtin, tout) + vdot(tin, dtout);
    v3 x = vcross(tin, tout);
    double se = vdot(v, x);
    v3 dx = vadd(vcross(dtin, tout), vcross(tin, dtout));
    double dse = vdot(dv, x) + vdot(v, dx);
    *out = ce * dse - se * dce;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Atom cache and gradient (analytical.rs:101-421)                           */
/* ------------------------------------------------------------------------ */
typedef struct { int cap, sv, ev; double theta; int full; } carc_t;
typedef struct { int in_cap, out_cap; double sign; } cvert_t;

typedef struct {
    int atom;
    double ri;
    int nc;
    int *owners;
    cap_t *caps;
    int na, nv;
    carc_t *arcs;
    cvert_t *verts;
} cache_t;

static void cache_free(cache_t *c) { free(c->owners); free(c->caps); free(c->arcs); free(c->verts); }

/* build_atom_cache, analytical.rs:101-214 (without c / l: unused by forces).
 * Returns 0 for None (enclosed), 1 otherwise. */
static int build_atom_cache(int atom, const double *pos, const double *radii, int n_nb,
                            const int *nb, cache_t *out) {
    memset(out, 0, sizeof(*out));
    v3 pi = P(pos + 3 * atom);
    double ri = radii[atom];
    int *filt = (int *)malloc(sizeof(int) * (size_t)(n_nb ? n_nb : 1));
    int nf = 0;
    for (int t = 0; t < n_nb; ++t) {
        int j = nb[t];
        double d = vnorm(vsub(P(pos + 3 * j), pi));
        if (d <= ri + radii[j]) filt[nf++] = j;
    }
    cap_t *caps = (cap_t *)malloc(sizeof(cap_t) * (size_t)(nf ? nf : 1));
    int *owners = (int *)malloc(sizeof(int) * (size_t)(nf ? nf : 1));
    int nc = build_caps(pi, ri, nf, filt, pos, radii, caps, owners);
    free(filt);
    if (nc < 0) { free(caps); free(owners); return 0; }
    out->atom = atom; out->ri = ri; out->nc = nc; out->owners = owners; out->caps = caps;
    boundary_t B = find_boundary(caps, nc);
    if (B.kind != B_BOUNDED) { boundary_free(&B); return 1; }
    out->nv = B.n_verts;
    out->verts = (cvert_t *)malloc(sizeof(cvert_t) * (size_t)(B.n_verts ? B.n_verts : 1));
    for (int k = 0; k < B.n_verts; ++k) {
        double sg = identify_root_sign(caps[B.verts[k].in_cap], caps[B.verts[k].out_cap], B.verts[k].point);
        out->verts[k].in_cap = B.verts[k].in_cap;
        out->verts[k].out_cap = B.verts[k].out_cap;
        out->verts[k].sign = sg == 0.0 ? 1.0 : sg;
    }
    out->na = B.n_arcs;
    out->arcs = (carc_t *)malloc(sizeof(carc_t) * (size_t)(B.n_arcs ? B.n_arcs : 1));
    for (int k = 0; k < B.n_arcs; ++k) {
        carc_t *a = &out->arcs[k];
        a->cap = B.arcs[k].cap; a->theta = B.arcs[k].theta; a->full = B.arcs[k].full;
        a->sv = a->ev = -1;
        if (a->full) continue;
        for (int v = 0; v < B.n_verts; ++v)
            if (vnorm2(vsub(B.verts[v].point, B.arcs[k].start)) < 1e-10) { a->sv = v; break; }
        for (int v = 0; v < B.n_verts; ++v)
            if (vnorm2(vsub(B.verts[v].point, B.arcs[k].end)) < 1e-10) { a->ev = v; break; }
    }
    boundary_free(&B);
    return 1;
}

static v3 dr_of(int y, int x_atom, v3 basis) { return y == x_atom ? basis : mk(0, 0, 0); }

/* directional_area_derivative, analytical.rs:315-421 */
static double directional_area_derivative(const cache_t *c, const double *pos,
                                          const double *radii, int x_atom, v3 basis) {
    v3 pi = P(pos + 3 * c->atom);
    double ri = c->ri;
    v3 dri = dr_of(c->atom, x_atom, basis);
    double sum_arc = 0.0;
    for (int k = 0; k < c->na; ++k) {
        const carc_t *a = &c->arcs[k];
        int ka = c->owners[a->cap];
        v3 pk = P(pos + 3 * ka);
        double rk = radii[ka];
        v3 drk = dr_of(ka, x_atom, basis);
        double dcos = cos_alpha_directional(pi, pk, ri, rk, dri, drk);
        double dik = vnorm(vsub(pk, pi));
        double cak = (dik * dik + ri * ri - rk * rk) / (2.0 * dik * ri);
        if (a->full || a->sv < 0 || a->ev < 0 || a->sv >= c->nv || a->ev >= c->nv) {
            sum_arc += a->theta * dcos;
            continue;
        }
        cvert_t vs = c->verts[a->sv], ve = c->verts[a->ev];
        int vs_other, ve_other;
        double vs_sign, ve_sign;
        if (vs.in_cap == a->cap) { vs_other = vs.out_cap; vs_sign = vs.sign; }
        else { vs_other = vs.in_cap; vs_sign = -vs.sign; }
        if (ve.in_cap == a->cap) { ve_other = ve.out_cap; ve_sign = ve.sign; }
        else { ve_other = ve.in_cap; ve_sign = -ve.sign; }
        int la = c->owners[vs_other], ma = c->owners[ve_other];
        double dtheta = 0.0;
        if (!arc_theta_jvp(pi, pk, P(pos + 3 * la), P(pos + 3 * ma), ri, rk, radii[la], radii[ma],
                           vs_sign, ve_sign, dri, drk, dr_of(la, x_atom, basis),
                           dr_of(ma, x_atom, basis), &dtheta))
            dtheta = 0.0;
        sum_arc += a->theta * dcos + cak * dtheta;
    }
    double sum_vert = 0.0;
    for (int k = 0; k < c->nv; ++k) {
        int ia = c->owners[c->verts[k].in_cap], oa = c->owners[c->verts[k].out_cap];
        double de = 0.0;
        if (!vertex_epsilon_jvp(pi, P(pos + 3 * ia), P(pos + 3 * oa), ri, radii[ia], radii[oa],
                                c->verts[k].sign, dri, dr_of(ia, x_atom, basis),
                                dr_of(oa, x_atom, basis), &de))
            de = 0.0;
        sum_vert += de;
    }
    return -sum_arc - sum_vert;
}

/* atom_area_gradient, analytical.rs:285-311: grad[x] += dA_i/dr_x (R_i^2 folded in) */
static void atom_area_gradient(const cache_t *c, const double *pos, const double *radii,
                               double *grad) {
    int *aff = (int *)malloc(sizeof(int) * (size_t)(c->nc + 1));
    int naff = 0;
    aff[naff++] = c->atom;
    for (int k = 0; k < c->nc; ++k) {
        int o = c->owners[k], dup = 0;
        for (int t = 0; t < naff; ++t) if (aff[t] == o) { dup = 1; break; }
        if (!dup) aff[naff++] = o;
    }
    double r_sq = c->ri * c->ri;
    for (int t = 0; t < naff; ++t) {
        int x = aff[t];
        double g[3];
        for (int axis = 0; axis < 3; ++axis) {
            v3 basis = mk(axis == 0, axis == 1, axis == 2);
            g[axis] = directional_area_derivative(c, pos, radii, x, basis);
        }
        grad[3 * x] += g[0] * r_sq;
        grad[3 * x + 1] += g[1] * r_sq;
        grad[3 * x + 2] += g[2] * r_sq;
    }
    free(aff);
}

/* neighbour lists by brute force, ascending index (analytical.rs:616-625) */
static void neighbour_lists(int n, const double *pos, const double *radii, int **start_out,
                            int **list_out) {
    int *cnt = (int *)calloc((size_t)n + 1, sizeof(int));
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double d = vnorm(vsub(P(pos + 3 * j), P(pos + 3 * i)));
            if (d <= radii[i] + radii[j]) { cnt[i]++; cnt[j]++; }
        }
    int *start = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    start[0] = 0;
    for (int i = 0; i < n; ++i) start[i + 1] = start[i] + cnt[i];
    int *list = (int *)malloc(sizeof(int) * (size_t)(start[n] ? start[n] : 1));
    int *fill = (int *)calloc((size_t)n, sizeof(int));
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double d = vnorm(vsub(P(pos + 3 * j), P(pos + 3 * i)));
            if (d <= radii[i] + radii[j]) {
                list[start[i] + fill[i]++] = j;
                list[start[j] + fill[j]++] = i;
            }
        }
    /* entries were appended in i-then-j order, which leaves every list in
     * ascending order: smaller partners arrive first (as j's row), larger later. */
    free(cnt); free(fill);
    *start_out = start;
    *list_out = list;
}

/* add_sasa_forces_analytical, analytical.rs:583-649.  gamma[i] kJ/mol/Å^2
 * (already scaled).  stats (may be NULL): [0] work atoms, [1] atoms with
 * arcs, [2] total caps, [3] total arcs, [4] total vertices, [5] sum c^2. */
void orc_psa_forces(int n, const double *pos, const double *radii, const double *gamma,
                    double *f, double *stats) {
    int *start, *list;
    neighbour_lists(n, pos, radii, &start, &list);
    double *da = (double *)calloc(3 * (size_t)n, sizeof(double));
    double st[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        if (gamma[i] == 0.0) continue;
        st[0] += 1;
        cache_t c;
        if (!build_atom_cache(i, pos, radii, start[i + 1] - start[i], list + start[i], &c)) continue;
        st[2] += c.nc;
        st[5] += (double)c.nc * c.nc;
        if (c.na == 0) { cache_free(&c); continue; }
        st[1] += 1; st[3] += c.na; st[4] += c.nv;
        memset(da, 0, 3 * (size_t)n * sizeof(double));
        atom_area_gradient(&c, pos, radii, da);
        for (int x = 0; x < n; ++x) {
            f[3 * x] -= da[3 * x] * gamma[i];
            f[3 * x + 1] -= da[3 * x + 1] * gamma[i];
            f[3 * x + 2] -= da[3 * x + 2] * gamma[i];
        }
        cache_free(&c);
    }
    if (stats) memcpy(stats, st, sizeof(st));
    free(da); free(start); free(list);
}

/* ∂A_atom/∂r for one atom with an explicit neighbour list (the two-atom
 * known-answer case, analytical.rs:1000-1046).  grad: 3n, accumulated. */
int orc_psa_atom_gradient(int n, const double *pos, const double *radii, int atom, int n_nb,
                          const int *nb, double *grad) {
    (void)n;
    cache_t c;
    if (!build_atom_cache(atom, pos, radii, n_nb, nb, &c)) return 0;
    if (c.na > 0) atom_area_gradient(&c, pos, radii, grad);
    int na = c.na;
    cache_free(&c);
    return na;
}

/* powersasa_energy per-atom areas, mod.rs:55-111 (neighbours: overlap
 * d <= r_i + r_j, ascending order).  Returns total area (Å^2). */
double orc_psa_area(int n, const double *pos, const double *radii, double *per_atom) {
    int *start, *list;
    neighbour_lists(n, pos, radii, &start, &list);
    double total = 0.0;
    cap_t *caps = (cap_t *)malloc(sizeof(cap_t) * (size_t)(n ? n : 1));
    int *owners = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; ++i) {
        per_atom[i] = 0.0;
        int nc = build_caps(P(pos + 3 * i), radii[i], start[i + 1] - start[i], list + start[i],
                            pos, radii, caps, owners);
        if (nc < 0) continue;
        boundary_t B = find_boundary(caps, nc);
        double a = accessible_area(radii[i], caps, nc, &B);
        boundary_free(&B);
        per_atom[i] = a;
        total += a;
    }
    free(caps); free(owners); free(start); free(list);
    return total;
}

/* Area of one sphere under explicit caps (axes need not be unit; they are
 * normalised as SmallCircle::new does).  For the closed-form cases of
 * area.rs:418-447. */
double orc_psa_area_from_caps(int nc, const double *axes, const double *cos_a, double radius) {
    cap_t *caps = (cap_t *)malloc(sizeof(cap_t) * (size_t)(nc ? nc : 1));
    for (int k = 0; k < nc; ++k) caps[k] = small_circle_new(P(axes + 3 * k), cos_a[k]);
    boundary_t B = find_boundary(caps, nc);
    double a = accessible_area(radius, caps, nc, &B);
    boundary_free(&B);
    free(caps);
    return a;
}

/* ------------------------------------------------------------------------ */
/* Shrake-Rupley dot-density SASA energy, energy/src/sasa.rs:60-178         */
/* ------------------------------------------------------------------------ */
/* fibonacci_unit_sphere, sasa.rs:182-199: y = 1 - 2 i / (n - 1), r =
 * sqrt(1 - y^2), theta = pi (sqrt 3 + 1) i, (cos theta r, y, sin theta r). */
void orc_sr_dots(int n, double *out) {
    const double phi = M_PI * (sqrt(3.0) + 1.0);
    const double nf = (double)n;
    for (int i = 0; i < n; ++i) {
        const double y = 1.0 - 2.0 * (double)i / (nf - 1.0);
        const double r = sqrt(1.0 - y * y);
        const double th = phi * (double)i;
        out[3 * i] = cos(th) * r;
        out[3 * i + 1] = y;
        out[3 * i + 2] = sin(th) * r;
    }
}

/* sasa_energy_with_dots, sasa.rs:113-178: per atom, the dots of its expanded
 * sphere not strictly inside any neighbour's (neighbours: distance <=
 * R_i + R_j); area = accessible / n_dots * 4 pi R^2; energy = sum gamma_i A_i
 * (gamma in kJ/mol/Å^2 here; origami sums in kcal and converts once).
 * per_atom (may be NULL) receives the areas; returns the energy. */
double orc_sr_energy(int n, const double *pos, const double *radii, const double *gamma, int n_dots,
                     double *per_atom) {
    double *dots = (double *)malloc(sizeof(double) * 3 * (size_t)(n_dots > 0 ? n_dots : 1));
    orc_sr_dots(n_dots, dots);
    int *nb = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    double energy = 0.0;
    for (int i = 0; i < n; ++i) {
        const double *pi = pos + 3 * i;
        const double ri = radii[i];
        int m = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = pos[3 * j] - pi[0], dy = pos[3 * j + 1] - pi[1], dz = pos[3 * j + 2] - pi[2];
            if (sqrt(dx * dx + dy * dy + dz * dz) <= ri + radii[j]) nb[m++] = j;
        }
        int accessible = 0;
        for (int k = 0; k < n_dots; ++k) {
            const double tx = pi[0] + dots[3 * k] * ri, ty = pi[1] + dots[3 * k + 1] * ri,
                         tz = pi[2] + dots[3 * k + 2] * ri;
            int buried = 0;
            for (int q = 0; q < m; ++q) {
                const int j = nb[q];
                const double ex = tx - pos[3 * j], ey = ty - pos[3 * j + 1], ez = tz - pos[3 * j + 2];
                if (ex * ex + ey * ey + ez * ez < radii[j] * radii[j]) {
                    buried = 1;
                    break;
                }
            }
            accessible += !buried;
        }
        const double area = (double)accessible / (double)n_dots * 4.0 * M_PI * ri * ri;
        if (per_atom) per_atom[i] = area;
        energy += gamma[i] * area;
    }
    free(dots);
    free(nb);
    return energy;
}
