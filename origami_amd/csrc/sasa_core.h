// PowerSasa per-atom force routine shared by the GPU kernel (sasa_kernel.hip)
// and its host build (used only to test the kernel logic on the CPU).
// See sasa_kernel.hip for the algorithm description and origami citations.
#pragma once
#include <math.h>
#include <stddef.h>

#ifdef __HIPCC__
#define OG_HD __host__ __device__
#else
#define OG_HD
#endif

// No FMA contraction here: the boundary topology is decided by tolerance
// tests (1e-12, 1e-9) that freshly built, exactly symmetric chains sit right
// on; origami's arithmetic is uncontracted, so this routine is too.
#ifdef __clang__
#pragma clang fp contract(off)
#endif

#define MAXC 128
#define MAXT (2 * MAXC)
#define MAXA 256
#define MAXV 256

// FP64 scratch layout (element offsets)
#define S_CAP 0                              // MAXC x (ax, ay, az, cos)
#define S_TR (S_CAP + 4 * MAXC)              // MAXT x (ang, px, py, pz)
#define S_ARC (S_TR + 4 * MAXT)              // MAXA x (sx, sy, sz, ex, ey, ez, theta)
#define S_VTX (S_ARC + 7 * MAXA)             // MAXV x (px, py, pz, sign)
#define S_GA (S_VTX + 4 * MAXV)              // (MAXC+1) x 3 arc-term sums
#define S_GV (S_GA + 3 * (MAXC + 1))         // (MAXC+1) x 3 vertex-term sums
#define S_TOTAL (S_GV + 3 * (MAXC + 1))
// int scratch layout
#define I_OWN 0                              // MAXC cap owners (global atom)
#define I_TRO (I_OWN + MAXC)                 // MAXT transition "other cap"
#define I_ARC (I_TRO + MAXT)                 // MAXA x (cap, full, sv, ev)
#define I_VTX (I_ARC + 4 * MAXA)             // MAXV x (in, out)
#define I_TOTAL (I_VTX + 2 * MAXV)

namespace og_sasa {

struct v3 { double x, y, z; };
OG_HD inline v3 mk(double x, double y, double z) { return {x, y, z}; }
OG_HD inline v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
OG_HD inline v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
OG_HD inline v3 operator*(v3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
OG_HD inline v3 operator/(v3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
OG_HD inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OG_HD inline v3 cross(v3 a, v3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
OG_HD inline double norm2(v3 a) { return dot(a, a); }
OG_HD inline double norm(v3 a) { return sqrt(dot(a, a)); }
OG_HD inline v3 normalize(v3 a) { return a / norm(a); }

struct Cap { v3 axis; double ca; };

struct Scr {
    double* d;
    int* i;
    int stride, t;
    OG_HD double& D(int k) const { return d[(size_t)k * stride + t]; }
    OG_HD int& I(int k) const { return i[(size_t)k * stride + t]; }
    OG_HD Cap cap(int k) const {
        return {mk(D(S_CAP + 4 * k), D(S_CAP + 4 * k + 1), D(S_CAP + 4 * k + 2)), D(S_CAP + 4 * k + 3)};
    }
    OG_HD void set_cap(int k, Cap c) const {
        D(S_CAP + 4 * k) = c.axis.x; D(S_CAP + 4 * k + 1) = c.axis.y; D(S_CAP + 4 * k + 2) = c.axis.z;
        D(S_CAP + 4 * k + 3) = c.ca;
    }
    OG_HD v3 v(int off) const { return mk(D(off), D(off + 1), D(off + 2)); }
    OG_HD void setv(int off, v3 p) const { D(off) = p.x; D(off + 1) = p.y; D(off + 2) = p.z; }
};

OG_HD inline Cap small_circle_new(v3 axis, double ca) { return {normalize(axis), ca}; }
OG_HD inline bool contains_strict(const Cap& c, v3 p) { return dot(p, c.axis) > c.ca + 1e-12; }

OG_HD inline v3 helper_axis(v3 a) {
    const double ax = fabs(a.x), ay = fabs(a.y), az = fabs(a.z);
    if (ax < ay && ax < az) return mk(1, 0, 0);
    if (ay < az) return mk(0, 1, 0);
    return mk(0, 0, 1);
}

// geometry.rs:67-100; returns 0 two points, 1 tangent, 2 none
OG_HD int intersect_circles(const Cap& c1, const Cap& c2, v3& p, v3& q) {
    const double sigma = dot(c1.axis, c2.axis);
    const double denom = 1.0 - sigma * sigma;
    if (fabs(denom) < 1e-12) return 2;
    const double a = (c1.ca - sigma * c2.ca) / denom;
    const double b = (c2.ca - sigma * c1.ca) / denom;
    const double c_sq = 1.0 - a * a - b * b - 2.0 * a * b * sigma;
    if (c_sq < -1e-9) return 2;
    const v3 base = c1.axis * a + c2.axis * b;
    if (fabs(c_sq) <= 1e-9) { p = base; return 1; }
    const double c = sqrt(c_sq);
    const v3 nc = cross(c1.axis, c2.axis);
    const v3 off = (nc / norm(nc)) * c;
    p = base + off;
    q = base - off;
    return 0;
}

// Number of points intersect_circles would return (same arithmetic, same
// decisions): 2, 1 (tangent) or 0.
OG_HD int intersect_count(const Cap& c1, const Cap& c2) {
    const double sigma = dot(c1.axis, c2.axis);
    const double denom = 1.0 - sigma * sigma;
    if (fabs(denom) < 1e-12) return 0;
    const double a = (c1.ca - sigma * c2.ca) / denom;
    const double b = (c2.ca - sigma * c1.ca) / denom;
    const double c_sq = 1.0 - a * a - b * b - 2.0 * a * b * sigma;
    if (c_sq < -1e-9) return 0;
    return fabs(c_sq) <= 1e-9 ? 1 : 2;
}

// intersect_count's answer, decided without its two divisions where it is
// clear.  intersect_circles forms c^2 = 1 - a^2 - b^2 - 2ab*sigma, which
// equals g / (1 - sigma^2) with g = (1 - ca1^2)(1 - ca2^2) - (sigma - ca1 ca2)^2.
// With 1 - sigma^2 >= 1e-2 and c^2 near its thresholds (|a|, |b| <= 15
// there), its rounding stays below 1e-10; the class is read from g when
// g / (1 - sigma^2) is at least 5e-10 away from both thresholds (+-1e-9),
// and by the exact arithmetic otherwise.
OG_HD int intersect_count_fast(const Cap& c1, const Cap& c2) {
    const double sigma = dot(c1.axis, c2.axis);
    const double denom = 1.0 - sigma * sigma;
    if (denom >= 1e-2) {
        const double u = sigma - c1.ca * c2.ca;
        const double g = (1.0 - c1.ca * c1.ca) * (1.0 - c2.ca * c2.ca) - u * u;
        if (g < -1.5e-9 * denom) return 0;
        if (g > 1.5e-9 * denom) return 2;
        if (g > -0.5e-9 * denom && g < 0.5e-9 * denom) return 1;
    }
    return intersect_count(c1, c2);
}

// Union cover test (no counterpart in origami; it only decides which
// circles can be left out, see the arrangement kernel): the arc of circle k
// (orthonormal in-plane basis u, v) on which cap m's containment value
// dot(X, n_m) - cos a_m is at least SA_COV_DELTA, as start angle a and width
// w2 (radians, FP32).  w2 < 0: no usable arc; w2 = SA_COV_FULL: the whole
// circle.  FP64 up to the final angles; acosf / atan2f only see arguments
// where they are well conditioned (|cos| <= 1 - 1e-3, amplitude >= 1e-3),
// nearly full arcs are cut to width 2 (pi - 0.05) and tiny ones dropped, so
// both ends lie within 2e-6 rad of the exact arc or inside it.
#define SA_COV_DELTA 1e-6
#define SA_COV_FULL 100.0f
OG_HD inline void cover_arc(const Cap& ck, const Cap& cm, v3 u, v3 v, float& a, float& w2) {
    const double sk = sqrt(fmax(1.0 - ck.ca * ck.ca, 0.0));
    const double ct = dot(ck.axis, cm.axis);
    const double x = dot(cm.axis, u), y = dot(cm.axis, v);
    const double amp = sk * sqrt(x * x + y * y);
    const double B = ck.ca * ct - cm.ca;
    a = 0.0f;
    w2 = -1.0f;
    if (B - amp >= SA_COV_DELTA) {
        w2 = SA_COV_FULL;
    } else if (amp >= 1e-3) {
        const double c = (SA_COV_DELTA - B) / amp;
        if (c < 1.0 - 1e-3) {
            const float w = c <= -1.0 + 1e-3 ? (float)(3.14159265358979323846 - 0.05) : acosf((float)c);
            a = atan2f((float)y, (float)x) - w;
            w2 = 2.0f * w;
        }
    }
}

// In-plane basis of a circle (arrangement.rs:233-248): u from the helper
// axis, v = axis x u.
OG_HD inline void circle_basis(const Cap& c, v3& u, v3& v) {
    const v3 h = helper_axis(c.axis);
    u = normalize(h - c.axis * dot(h, c.axis));
    v = cross(c.axis, u);
}

OG_HD inline double angular_position_b(v3 pt, const Cap& c, v3 u, v3 v) {
    const v3 proj = pt - c.axis * dot(pt, c.axis);
    return atan2(dot(proj, v), dot(proj, u));
}

OG_HD double angular_position(v3 pt, const Cap& c) {
    v3 u, v;
    circle_basis(c, u, v);
    return angular_position_b(pt, c, u, v);
}

OG_HD v3 sample_circle_point(const Cap& c) {
    const v3 h = helper_axis(c.axis);
    const v3 perp = normalize(h - c.axis * dot(h, c.axis));
    const double t = 1.0 - c.ca * c.ca;
    return c.axis * c.ca + perp * sqrt(t > 0.0 ? t : 0.0);
}

OG_HD v3 midpoint_on_circle(v3 pt, const Cap& c, double delta) {
    const v3 axial = c.axis * dot(pt, c.axis);
    const v3 radial = pt - axial;
    const double rn = norm(radial);
    if (rn < 1e-12) return pt;
    const v3 u = radial / rn;
    const v3 v = cross(c.axis, u);
    double sd, cd;
    sincos(delta, &sd, &cd);
    return axial + (u * cd + v * sd) * rn;
}

OG_HD inline double wrap_2pi_positive(double t) {
    const double two_pi = 2.0 * 3.14159265358979323846;
    while (t <= 0.0) t += two_pi;
    while (t > two_pi) t -= two_pi;
    return t;
}

// derivatives.rs:92-112
OG_HD bool vertex_point(const Cap& c1, const Cap& c2, double sign, v3& out) {
    const double sigma = dot(c1.axis, c2.axis);
    const double denom = 1.0 - sigma * sigma;
    if (fabs(denom) < 1e-12) return false;
    const double a = (c1.ca - sigma * c2.ca) / denom;
    const double b = (c2.ca - sigma * c1.ca) / denom;
    const double c_sq = 1.0 - a * a - b * b - 2.0 * a * b * sigma;
    if (c_sq <= 0.0) return false;
    const v3 base = c1.axis * a + c2.axis * b;
    const v3 nc = cross(c1.axis, c2.axis);
    const double nn = norm(nc);
    if (nn < 1e-12) return false;
    out = base + ((nc / nn) * sqrt(c_sq)) * sign;
    return true;
}

OG_HD double cos_alpha_directional(v3 pi, v3 pk, double ri, double rk, v3 dri, v3 drk) {
    const v3 rv = pk - pi;
    const double d2 = norm2(rv);
    const double d = sqrt(d2);
    if (d < 1e-12) return 0.0;
    const v3 u = rv / d;
    const double dcos_dd = (d2 - ri * ri + rk * rk) / (2.0 * d2 * ri);
    const v3 gk = u * dcos_dd;
    const v3 gi = gk * -1.0;
    return dot(gi, dri) + dot(gk, drk);
}

// derivatives.rs:222-325
OG_HD bool vertex_point_jvp(v3 pi, v3 pk, v3 pl, double ri, double rk, double rl, double sign,
                                 v3 dri, v3 drk, v3 drl, v3& dv) {
    const v3 rkv = pk - pi;
    const double dik = norm(rkv);
    if (dik < 1e-12) return false;
    const v3 wk = rkv / dik;
    const v3 rlv = pl - pi;
    const double dil = norm(rlv);
    if (dil < 1e-12) return false;
    const v3 wl = rlv / dil;
    const double cak = (dik * dik + ri * ri - rk * rk) / (2.0 * dik * ri);
    const double cal = (dil * dil + ri * ri - rl * rl) / (2.0 * dil * ri);
    const double sigma = dot(wk, wl);
    const double denom = 1.0 - sigma * sigma;
    if (fabs(denom) < 1e-12) return false;
    const double a = (cak - sigma * cal) / denom;
    const double b = (cal - sigma * cak) / denom;
    const double c_sq = 1.0 - a * a - b * b - 2.0 * a * b * sigma;
    if (c_sq <= 0.0) return false;
    const double c = sqrt(c_sq);
    const v3 m = cross(wk, wl);
    const double mn = sqrt(norm2(m));
    if (mn < 1e-12) return false;
    const v3 np = m / mn;
    const v3 dvk = drk - dri;
    const v3 dwk = (dvk - wk * dot(wk, dvk)) / dik;
    const v3 dvl = drl - dri;
    const v3 dwl = (dvl - wl * dot(wl, dvl)) / dil;
    const double dcos_k = cos_alpha_directional(pi, pk, ri, rk, dri, drk);
    const double dcos_l = cos_alpha_directional(pi, pl, ri, rl, dri, drl);
    const double dsig = dot(dwk, wl) + dot(wk, dwl);
    const double dden = -2.0 * sigma * dsig;
    const double num_a = cak - sigma * cal;
    const double dnum_a = dcos_k - dsig * cal - sigma * dcos_l;
    const double da = (dnum_a * denom - num_a * dden) / (denom * denom);
    const double num_b = cal - sigma * cak;
    const double dnum_b = dcos_l - dsig * cak - sigma * dcos_k;
    const double db = (dnum_b * denom - num_b * dden) / (denom * denom);
    const double dc_sq = -2.0 * a * da - 2.0 * b * db - 2.0 * (da * b + a * db) * sigma - 2.0 * a * b * dsig;
    const double dc = dc_sq / (2.0 * c);
    const v3 dbase = wk * da + dwk * a + wl * db + dwl * b;
    const v3 dm = cross(dwk, wl) + cross(wk, dwl);
    const v3 dnp = (dm - np * dot(np, dm)) / mn;
    dv = dbase + (np * dc + dnp * c) * sign;
    return true;
}

OG_HD Cap cap_from(v3 pi, v3 pk, double ri, double rk) {
    const v3 rv = pk - pi;
    const double d = norm(rv);
    return small_circle_new(rv / d, (d * d + ri * ri - rk * rk) / (2.0 * d * ri));
}

// derivatives.rs:346-444
OG_HD bool arc_theta_jvp(v3 pi, v3 pk, v3 pl, v3 pm, double ri, double rk, double rl, double rm,
                              double sign_s, double sign_e, v3 dri, v3 drk, v3 drl, v3 drm, double& out) {
    const v3 rkv = pk - pi;
    const double dik = norm(rkv);
    if (dik < 1e-12) return false;
    const v3 wk = rkv / dik;
    const v3 dvv = drk - dri;
    const v3 dwk = (dvv - wk * dot(wk, dvv)) / dik;
    const double cak = (dik * dik + ri * ri - rk * rk) / (2.0 * dik * ri);
    const Cap ck = small_circle_new(wk, cak);
    if (norm(pl - pi) < 1e-12) return false;
    const Cap cl = cap_from(pi, pl, ri, rl);
    if (norm(pm - pi) < 1e-12) return false;
    const Cap cm = cap_from(pi, pm, ri, rm);
    v3 vs, ve, dvs, dve;
    if (!vertex_point(ck, cl, sign_s, vs)) return false;
    if (!vertex_point(ck, cm, sign_e, ve)) return false;
    if (!vertex_point_jvp(pi, pk, pl, ri, rk, rl, sign_s, dri, drk, drl, dvs)) return false;
    if (!vertex_point_jvp(pi, pk, pm, ri, rk, rm, sign_e, dri, drk, drm, dve)) return false;
    v3 us, dus, ue, due;
    {
        const double vdo = dot(vs, wk);
        const v3 proj = vs - wk * vdo;
        const double dvdo = dot(dvs, wk) + dot(vs, dwk);
        const v3 dproj = dvs - dwk * vdo - wk * dvdo;
        const double nn = norm(proj);
        if (nn < 1e-12) return false;
        us = proj / nn;
        dus = (dproj - us * dot(us, dproj)) / nn;
    }
    {
        const double vdo = dot(ve, wk);
        const v3 proj = ve - wk * vdo;
        const double dvdo = dot(dve, wk) + dot(ve, dwk);
        const v3 dproj = dve - dwk * vdo - wk * dvdo;
        const double nn = norm(proj);
        if (nn < 1e-12) return false;
        ue = proj / nn;
        due = (dproj - ue * dot(ue, dproj)) / nn;
    }
    const double cos_t = dot(us, ue);
    const double dcos_t = dot(dus, ue) + dot(us, due);
    const v3 x = cross(us, ue);
    const double sin_t = dot(wk, x);
    const v3 dx = cross(dus, ue) + cross(us, due);
    const double dsin_t = dot(dwk, x) + dot(wk, dx);
    out = cos_t * dsin_t - sin_t * dcos_t;
    return true;
}

// derivatives.rs:466-545 with incoming_is_k = true
OG_HD bool vertex_epsilon_jvp(v3 pi, v3 pk, v3 pl, double ri, double rk, double rl, double sign,
                                   v3 dri, v3 drk, v3 drl, double& out) {
    const v3 rkv = pk - pi;
    const double dik = norm(rkv);
    if (dik < 1e-12) return false;
    const v3 wk = rkv / dik;
    const v3 dvk = drk - dri;
    const v3 dwk = (dvk - wk * dot(wk, dvk)) / dik;
    const v3 rlv = pl - pi;
    const double dil = norm(rlv);
    if (dil < 1e-12) return false;
    const v3 wl = rlv / dil;
    const v3 dvl = drl - dri;
    const v3 dwl = (dvl - wl * dot(wl, dvl)) / dil;
    const double cak = (dik * dik + ri * ri - rk * rk) / (2.0 * dik * ri);
    const double cal = (dil * dil + ri * ri - rl * rl) / (2.0 * dil * ri);
    const Cap ck = small_circle_new(wk, cak), cl = small_circle_new(wl, cal);
    v3 v, dv;
    if (!vertex_point(ck, cl, sign, v)) return false;
    if (!vertex_point_jvp(pi, pk, pl, ri, rk, rl, sign, dri, drk, drl, dv)) return false;
    v3 tin, dtin, tout, dtout;
    {
        const v3 cr = cross(v, wk);
        const double nn = norm(cr);
        if (nn < 1e-12) return false;
        tin = cr / nn;
        const v3 dcr = cross(dv, wk) + cross(v, dwk);
        dtin = (dcr - tin * dot(tin, dcr)) / nn;
    }
    {
        const v3 cr = cross(v, wl);
        const double nn = norm(cr);
        if (nn < 1e-12) return false;
        tout = cr / nn;
        const v3 dcr = cross(dv, wl) + cross(v, dwl);
        dtout = (dcr - tout * dot(tout, dcr)) / nn;
    }
    const double ce = dot(tin, tout);
    const double dce = dot(dtin, tout) + dot(tin, dtout);
    const v3 x = cross(tin, tout);
    const double se = dot(v, x);
    const v3 dx = cross(dtin, tout) + cross(tin, dtout);
    const double dse = dot(dv, x) + dot(v, dx);
    out = ce * dse - se * dce;
    return true;
}

OG_HD inline v3 ldp(const double* pos, int a) {
    return mk(pos[3 * (size_t)a], pos[3 * (size_t)a + 1], pos[3 * (size_t)a + 2]);
}

OG_HD inline v3 unit(int axis) { return mk(axis == 0, axis == 1, axis == 2); }

}  // namespace og_sasa


namespace og_sasa {

// Analytical SASA force contribution of one work atom.  emit(atom, axis, f)
// receives F_atom[axis] += f for every affected atom.  *overflow_flag is set
// (and nothing is emitted) when the arrangement exceeds MAXC/MAXA/MAXV.
template <class Emit>
OG_HD void atom_forces(const Scr& S, int work_atom, const double* pos, const double* radius,
                       int sys_begin, int sys_end, double gam, int* overflow_flag, Emit emit) {

    const int ia = work_atom;
    const v3 pi = ldp(pos, ia);
    const double ri = radius[ia];
    if (gam == 0.0) return;

    // ---- caps from overlapping neighbours (build_atom_cache + build_caps) ----
    int nc = 0;
    for (int j = sys_begin; j < sys_end; ++j) {
        if (j == ia) continue;
        const v3 pj = ldp(pos, j);
        const double rj = radius[j];
        const double dd = norm(pj - pi);
        if (!(dd <= ri + rj)) continue;
        if (dd + ri <= rj) return;               // enclosed: no accessible surface
        if (dd + rj <= ri || dd >= ri + rj) continue;
        const double ca = (dd * dd + ri * ri - rj * rj) / (2.0 * dd * ri);
        if (!(ca >= -1.0 && ca <= 1.0)) continue;
        if (nc >= MAXC) { *overflow_flag = 1; return; }
        S.set_cap(nc, {(pj - pi) / dd, ca});
        S.I(I_OWN + nc) = j;
        ++nc;
    }
    if (nc == 0) return;                         // fully exposed: constant area

    // ---- boundary arcs and vertices (find_boundary) ----
    int n_arcs = 0, n_verts = 0;
    for (int k = 0; k < nc; ++k) {
        const Cap ck = S.cap(k);
        int m = 0;
        for (int o = 0; o < nc; ++o) {
            if (o == k) continue;
            v3 p, q;
            const int r = intersect_circles(ck, S.cap(o), p, q);
            if (r == 2) continue;
            S.D(S_TR + 4 * m) = angular_position(p, ck);
            S.setv(S_TR + 4 * m + 1, p);
            S.I(I_TRO + m) = o;
            ++m;
            if (r == 0) {
                S.D(S_TR + 4 * m) = angular_position(q, ck);
                S.setv(S_TR + 4 * m + 1, q);
                S.I(I_TRO + m) = o;
                ++m;
            }
        }
        // stable insertion sort by angle
        for (int x = 1; x < m; ++x) {
            const double ang = S.D(S_TR + 4 * x);
            const v3 pt = S.v(S_TR + 4 * x + 1);
            const int oth = S.I(I_TRO + x);
            int y = x - 1;
            while (y >= 0 && S.D(S_TR + 4 * y) > ang) {
                S.D(S_TR + 4 * (y + 1)) = S.D(S_TR + 4 * y);
                S.setv(S_TR + 4 * (y + 1) + 1, S.v(S_TR + 4 * y + 1));
                S.I(I_TRO + y + 1) = S.I(I_TRO + y);
                --y;
            }
            S.D(S_TR + 4 * (y + 1)) = ang;
            S.setv(S_TR + 4 * (y + 1) + 1, pt);
            S.I(I_TRO + y + 1) = oth;
        }
        // dedup consecutive coincident points (keep the first)
        int w = 0;
        for (int x = 0; x < m; ++x) {
            const v3 pt = S.v(S_TR + 4 * x + 1);
            if (w > 0 && norm2(pt - S.v(S_TR + 4 * (w - 1) + 1)) < 1e-12) continue;
            if (w != x) {
                S.D(S_TR + 4 * w) = S.D(S_TR + 4 * x);
                S.setv(S_TR + 4 * w + 1, pt);
                S.I(I_TRO + w) = S.I(I_TRO + x);
            }
            ++w;
        }
        m = w;
        if (m == 0) {
            const v3 probe = sample_circle_point(ck);
            bool buried = false;
            for (int o = 0; o < nc && !buried; ++o) buried = (o != k) && contains_strict(S.cap(o), probe);
            if (!buried) {
                if (n_arcs >= MAXA) { *overflow_flag = 1; return; }
                const int A = n_arcs++;
                S.setv(S_ARC + 7 * A, probe);
                S.setv(S_ARC + 7 * A + 3, probe);
                S.D(S_ARC + 7 * A + 6) = -2.0 * 3.14159265358979323846;
                S.I(I_ARC + 4 * A) = k;
                S.I(I_ARC + 4 * A + 1) = 1;
            }
            continue;
        }
        for (int x = 0; x < m; ++x) {
            const int y = (x + 1) % m;
            const double ccw = wrap_2pi_positive(S.D(S_TR + 4 * y) - S.D(S_TR + 4 * x));
            if (ccw < 1e-9) continue;
            const v3 apt = S.v(S_TR + 4 * x + 1);
            const v3 mid = midpoint_on_circle(apt, ck, ccw / 2.0);
            bool bad = false;
            for (int o = 0; o < nc && !bad; ++o) bad = (o != k) && contains_strict(S.cap(o), mid);
            if (bad) continue;
            if (n_arcs >= MAXA || n_verts >= MAXV) { *overflow_flag = 1; return; }
            const int A = n_arcs++;
            S.setv(S_ARC + 7 * A, S.v(S_TR + 4 * y + 1));
            S.setv(S_ARC + 7 * A + 3, apt);
            S.D(S_ARC + 7 * A + 6) = -ccw;
            S.I(I_ARC + 4 * A) = k;
            S.I(I_ARC + 4 * A + 1) = 0;
            const int V = n_verts++;
            S.setv(S_VTX + 4 * V, apt);
            S.I(I_VTX + 2 * V) = k;
            S.I(I_VTX + 2 * V + 1) = S.I(I_TRO + x);
        }
    }
    // FullyBuried (no arcs and every probe buried) and the remaining
    // no-arc case both contribute nothing to the gradient.
    if (n_arcs == 0) return;

    // ---- cache: vertex root signs, arc -> vertex links ----
    for (int V = 0; V < n_verts; ++V) {
        const Cap cin = S.cap(S.I(I_VTX + 2 * V)), cout = S.cap(S.I(I_VTX + 2 * V + 1));
        const v3 ref = S.v(S_VTX + 4 * V);
        v3 vp, vm;
        double sg = 1.0;
        if (vertex_point(cin, cout, 1.0, vp) && vertex_point(cin, cout, -1.0, vm))
            sg = norm2(vp - ref) <= norm2(vm - ref) ? 1.0 : -1.0;
        S.D(S_VTX + 4 * V + 3) = sg;
    }
    for (int A = 0; A < n_arcs; ++A) {
        int sv = -1, ev = -1;
        if (!S.I(I_ARC + 4 * A + 1)) {
            const v3 st = S.v(S_ARC + 7 * A), en = S.v(S_ARC + 7 * A + 3);
            for (int V = 0; V < n_verts && sv < 0; ++V)
                if (norm2(S.v(S_VTX + 4 * V) - st) < 1e-10) sv = V;
            for (int V = 0; V < n_verts && ev < 0; ++V)
                if (norm2(S.v(S_VTX + 4 * V) - en) < 1e-10) ev = V;
        }
        S.I(I_ARC + 4 * A + 2) = sv;
        S.I(I_ARC + 4 * A + 3) = ev;
    }

    // ---- gradient: per (affected atom, axis) arc and vertex sums ----
    // affected local index: 0 = atom i, 1 + cap index = that cap's owner
    for (int k = 0; k < 3 * (nc + 1); ++k) {
        S.D(S_GA + k) = 0.0;
        S.D(S_GV + k) = 0.0;
    }
    const v3 zero = mk(0, 0, 0);
    for (int A = 0; A < n_arcs; ++A) {
        const int kc = S.I(I_ARC + 4 * A);
        const int ka = S.I(I_OWN + kc);
        const v3 pk = ldp(pos, ka);
        const double rk = radius[ka];
        const double theta = S.D(S_ARC + 7 * A + 6);
        const double dik = norm(pk - pi);
        const double cak = (dik * dik + ri * ri - rk * rk) / (2.0 * dik * ri);
        const int sv = S.I(I_ARC + 4 * A + 2), ev = S.I(I_ARC + 4 * A + 3);
        const bool simple = S.I(I_ARC + 4 * A + 1) || sv < 0 || ev < 0;
        int lc = -1, mc = -1;
        double ssg = 1.0, esg = 1.0;
        if (!simple) {
            const int vin = S.I(I_VTX + 2 * sv), vout = S.I(I_VTX + 2 * sv + 1);
            if (vin == kc) { lc = vout; ssg = S.D(S_VTX + 4 * sv + 3); }
            else { lc = vin; ssg = -S.D(S_VTX + 4 * sv + 3); }
            const int win = S.I(I_VTX + 2 * ev), wout = S.I(I_VTX + 2 * ev + 1);
            if (win == kc) { mc = wout; esg = S.D(S_VTX + 4 * ev + 3); }
            else { mc = win; esg = -S.D(S_VTX + 4 * ev + 3); }
        }
        // distinct atoms touched by this arc: i (local 0), k, l, m (local 1 + cap)
        int roles[4] = {0, 1 + kc, lc >= 0 ? 1 + lc : -1, mc >= 0 ? 1 + mc : -1};
        const int la = lc >= 0 ? S.I(I_OWN + lc) : -1, ma = mc >= 0 ? S.I(I_OWN + mc) : -1;
        const v3 pl = la >= 0 ? ldp(pos, la) : zero, pm = ma >= 0 ? ldp(pos, ma) : zero;
        const double rl = la >= 0 ? radius[la] : 0.0, rm = ma >= 0 ? radius[ma] : 0.0;
        for (int r = 0; r < 4; ++r) {
            const int x = roles[r];
            if (x < 0) continue;
            bool seen = false;
            for (int q = 0; q < r; ++q) seen = seen || (roles[q] == x);
            if (seen) continue;
            for (int axis = 0; axis < 3; ++axis) {
                const v3 b = unit(axis);
                const v3 dri = x == 0 ? b : zero;
                const v3 drk = x == 1 + kc ? b : zero;
                const double dcos = cos_alpha_directional(pi, pk, ri, rk, dri, drk);
                double term;
                if (simple) {
                    term = theta * dcos;
                } else {
                    const v3 drl = x == 1 + lc ? b : zero;
                    const v3 drm = x == 1 + mc ? b : zero;
                    double dth = 0.0;
                    if (!arc_theta_jvp(pi, pk, pl, pm, ri, rk, rl, rm, ssg, esg, dri, drk, drl, drm, dth))
                        dth = 0.0;
                    term = theta * dcos + cak * dth;
                }
                S.D(S_GA + 3 * x + axis) += term;
            }
        }
    }
    for (int V = 0; V < n_verts; ++V) {
        const int kcap = S.I(I_VTX + 2 * V), lcap = S.I(I_VTX + 2 * V + 1);
        const int ka = S.I(I_OWN + kcap), la = S.I(I_OWN + lcap);
        const v3 pk = ldp(pos, ka), pl = ldp(pos, la);
        const double rk = radius[ka], rl = radius[la];
        const double sg = S.D(S_VTX + 4 * V + 3);
        const int roles[3] = {0, 1 + kcap, 1 + lcap};
        for (int r = 0; r < 3; ++r) {
            const int x = roles[r];
            for (int axis = 0; axis < 3; ++axis) {
                const v3 b = unit(axis);
                const v3 dri = x == 0 ? b : zero;
                const v3 drk = x == 1 + kcap ? b : zero;
                const v3 drl = x == 1 + lcap ? b : zero;
                double de = 0.0;
                if (!vertex_epsilon_jvp(pi, pk, pl, ri, rk, rl, sg, dri, drk, drl, de)) de = 0.0;
                S.D(S_GV + 3 * x + axis) += de;
            }
        }
    }
    // ---- scatter F_x -= gamma_i * R_i^2 * (-sum_arc - sum_vert) ----
    const double r_sq = ri * ri;
    for (int x = 0; x <= nc; ++x) {
        const int atom = x == 0 ? ia : S.I(I_OWN + x - 1);
        for (int axis = 0; axis < 3; ++axis) {
            const double g = (-S.D(S_GA + 3 * x + axis) - S.D(S_GV + 3 * x + axis)) * r_sq;
            const double f = -(g * gam);
            if (f != 0.0) emit(atom, axis, f);
        }
    }
}

}  // namespace og_sasa
