// Host build of the PowerSasa device helpers (sasa_core.h), for CPU tests of
// the shortcuts the arrangement kernel takes: each must give the same
// decision as the exact arithmetic it stands in for.
#include "../../origami_amd/csrc/sasa_core.h"

using namespace og_sasa;

extern "C" {

// Mismatches of pair_q against integer division over p < 65536, n in [1, 255].
long long host_pair_div_mismatches(void) {
    long long bad = 0;
    for (int n = 1; n < 256; ++n) {
        const PairDiv d = pair_div(n);
        for (int p = 0; p < 65536; ++p) bad += pair_q(d, p) != p / n;
    }
    return bad;
}

// counts[i] = intersect_count, fast[i] = intersect_count_fast for cap pairs
// (axis1, ca1, axis2, ca2) packed as 8 doubles each.
void host_intersect_counts(int n, const double* caps, int* counts, int* fast) {
    for (int i = 0; i < n; ++i) {
        const double* c = caps + 8 * i;
        const Cap a = {mk(c[0], c[1], c[2]), c[3]};
        const Cap b = {mk(c[4], c[5], c[6]), c[7]};
        counts[i] = intersect_count(a, b);
        fast[i] = intersect_count_fast(a, b);
    }
}

// Points intersect_circles returns for each pair (6 doubles, NaN when absent)
void host_intersect_points(int n, const double* caps, double* pts) {
    for (int i = 0; i < n; ++i) {
        const double* c = caps + 8 * i;
        const Cap a = {mk(c[0], c[1], c[2]), c[3]};
        const Cap b = {mk(c[4], c[5], c[6]), c[7]};
        v3 p = mk(NAN, NAN, NAN), q = mk(NAN, NAN, NAN);
        const int r = intersect_circles(a, b, p, q);
        if (r == 1) q = mk(NAN, NAN, NAN);
        if (r == 2) p = q = mk(NAN, NAN, NAN);
        double* o = pts + 6 * i;
        o[0] = p.x; o[1] = p.y; o[2] = p.z; o[3] = q.x; o[4] = q.y; o[5] = q.z;
    }
}

// Exact midpoints (midpoint_on_circle) of point a rotated by ccw/2 about the
// cap axis; input rows (ax, ay, az, nx, ny, nz, ca, ccw).
void host_midpoints(int n, const double* in, double* out) {
    for (int i = 0; i < n; ++i) {
        const double* r = in + 8 * i;
        const Cap c = {mk(r[3], r[4], r[5]), r[6]};
        const v3 m = midpoint_on_circle(mk(r[0], r[1], r[2]), c, r[7] / 2.0);
        out[3 * i] = m.x; out[3 * i + 1] = m.y; out[3 * i + 2] = m.z;
    }
}

// Union cover arcs of circle k by cap m for pairs (axis_k, ca_k, axis_m,
// ca_m) packed as 8 doubles; out: (start, width) per pair and the basis
// (u, v) of circle k used.
void host_cover_arcs(int n, const double* caps, float* arcs, double* basis) {
    for (int i = 0; i < n; ++i) {
        const double* c = caps + 8 * i;
        const Cap k = {mk(c[0], c[1], c[2]), c[3]};
        const Cap m = {mk(c[4], c[5], c[6]), c[7]};
        v3 u, v;
        circle_basis(k, u, v);
        float a, w2;
        cover_arc(k, m, u, v, a, w2);
        arcs[2 * i] = a;
        arcs[2 * i + 1] = w2;
        double* b = basis + 6 * i;
        b[0] = u.x; b[1] = u.y; b[2] = u.z; b[3] = v.x; b[4] = v.y; b[5] = v.z;
    }
}

// JVP terms of random arcs / vertices both ways: one full JVP per direction
// (the functions origami's derivatives are restated in) and the primal /
// tangent split the terms kernel uses.  rows: 4 atoms (12 doubles), 4 radii,
// 2 signs; out: 12 arc terms each way, 9 vertex terms each way.
void host_jvp_split(int n, const double* rows, double* arc_full, double* arc_split,
                    double* vtx_full, double* vtx_split) {
    for (int r = 0; r < n; ++r) {
        const double* q = rows + 18 * r;
        const v3 p[4] = {mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), mk(q[6], q[7], q[8]), mk(q[9], q[10], q[11])};
        const double rad[4] = {q[12], q[13], q[14], q[15]};
        const double ss = q[16], se = q[17];
        const v3 z = mk(0, 0, 0);
        const ArcThetaP tp = arc_theta_primal(p[0], p[1], p[2], p[3], rad[0], rad[1], rad[2], rad[3], ss, se);
        const VertexEpsP ep = vertex_eps_primal(p[0], p[1], p[2], rad[0], rad[1], rad[2], ss);
        for (int role = 0; role < 4; ++role) {
            for (int axis = 0; axis < 3; ++axis) {
                const v3 b = unit(axis);
                const v3 d[4] = {role == 0 ? b : z, role == 1 ? b : z, role == 2 ? b : z, role == 3 ? b : z};
                double full = 0.0;
                if (!arc_theta_jvp(p[0], p[1], p[2], p[3], rad[0], rad[1], rad[2], rad[3], ss, se,
                                   d[0], d[1], d[2], d[3], full))
                    full = 0.0;
                arc_full[12 * r + 3 * role + axis] = full;
                arc_split[12 * r + 3 * role + axis] = tp.ok ? arc_theta_tangent(tp, d[0], d[1], d[2], d[3]) : 0.0;
                if (role < 3) {
                    double vf = 0.0;
                    if (!vertex_epsilon_jvp(p[0], p[1], p[2], rad[0], rad[1], rad[2], ss, d[0], d[1], d[2], vf))
                        vf = 0.0;
                    vtx_full[9 * r + 3 * role + axis] = vf;
                    vtx_split[9 * r + 3 * role + axis] = ep.ok ? vertex_eps_tangent(ep, d[0], d[1], d[2]) : 0.0;
                }
            }
        }
    }
}

}
