// PowerSasa geometry and derivative primitives (FP64), shared by the GPU
// kernels (sasa_kernel.hip) and a host build that tests them on the CPU
// (tests/host).  See sasa_kernel.hip for the algorithm and origami citations.

// MAXC caps per atom (deeper-buried atoms are reported as overflow); MAXA
// arcs and MAXV vertices per atom.
#pragma once
#include <math.h>
#include <stddef.h>

// internal linkage: the header is included by more than one kernel file
#ifdef __HIPCC__
#define OG_HD static __host__ __device__
#else
#define OG_HD static
#endif

// No FMA contraction here: the boundary topology is decided by tolerance
// tests (1e-12, 1e-9) that freshly built, exactly symmetric chains sit right
// on; origami's arithmetic is uncontracted, so this routine is too.
#ifdef __clang__
#pragma clang fp contract(off)
#endif

#define MAXC 128
#define MAXA 256
#define MAXV 256

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

// Division of a pair index by the number of other caps without an integer
// divide: q = floor(p * ceil(2^24 / n) / 2^24).  The product overshoots
// p / n by less than p / 2^24 < 1/256, while the fractional part of p / n is
// at most 1 - 1/n < 1 - 1/256, so the floor is exact for p < 65536 and
// 1 <= n < 256 (checked exhaustively on the host).
struct PairDiv { int n; unsigned inv; };
OG_HD inline PairDiv pair_div(int n) { return {n, (unsigned)((((1u << 24) + (unsigned)n - 1u)) / (unsigned)n)}; }
OG_HD inline int pair_q(const PairDiv& d, int p) {
    return (int)(((unsigned long long)(unsigned)p * d.inv) >> 24);
}

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
// SA_COV_DELTA is also the depth the arrangement kernel requires before it
// leaves a covered circle's transitions out of the other circles' boundaries
// (it exceeds the 2.6e-4 that chains of deduplicated points can span).
#define SA_COV_DELTA 1e-3
#define SA_COV_FULL 100.0f
// The containment value on the circle is g + SA_COV_DELTA + amp cos(phi -
// phi_m) with g = B - SA_COV_DELTA; the whole-circle and no-arc cases are
// settled on amp^2 (no square root), the arc itself from amp.
OG_HD inline void cover_arc(const Cap& ck, const Cap& cm, v3 u, v3 v, float& a, float& w2) {
    const double ct = dot(ck.axis, cm.axis);
    const double x = dot(cm.axis, u), y = dot(cm.axis, v);
    const double amp2 = fmax(1.0 - ck.ca * ck.ca, 0.0) * (x * x + y * y);
    const double g = (ck.ca * ct - cm.ca) - SA_COV_DELTA;
    a = 0.0f;
    w2 = -1.0f;
    if (g >= 0.0 && g * g >= amp2) {
        w2 = SA_COV_FULL;                              // g >= amp
    } else if (amp2 >= 1e-6 && !(g <= 0.0 && g * g >= 0.998001 * amp2)) {
        // amp >= 1e-3, and not -g >= (1 - 1e-3) amp (no arc)
        const double c = -g / sqrt(amp2);
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

// ---------------------------------------------------------------------------
// The same JVPs split into a primal part, formed once per arc or vertex, and
// a tangent part per direction (role x axis).  Each tangent repeats the
// operations of the functions above on the same primal values, so every
// term is bit-identical to theirs; only the repeated primal work goes.
// ---------------------------------------------------------------------------
struct CosAlphaP {
    v3 gk;              // d cos(a) / d r_k (d cos(a) / d r_i = -gk)
    bool ok;
};

OG_HD inline CosAlphaP cos_alpha_primal(v3 pi, v3 pk, double ri, double rk) {
    const v3 rv = pk - pi;
    const double d2 = norm2(rv);
    const double d = sqrt(d2);
    if (d < 1e-12) return {mk(0, 0, 0), false};
    const v3 u = rv / d;
    const double dcos_dd = (d2 - ri * ri + rk * rk) / (2.0 * d2 * ri);
    return {u * dcos_dd, true};
}

OG_HD inline double cos_alpha_tangent(const CosAlphaP& P, v3 dri, v3 drk) {
    if (!P.ok) return 0.0;
    const v3 gi = P.gk * -1.0;
    return dot(gi, dri) + dot(P.gk, drk);
}

struct VertexJvpP {
    v3 wk, wl, np;
    double dik, dil, cak, cal, sigma, denom, denom2, a, b, c, mn, num_a, num_b, sign;
    CosAlphaP ck, cl;
    bool ok;
};

OG_HD inline VertexJvpP vertex_jvp_primal(v3 pi, v3 pk, v3 pl, double ri, double rk, double rl, double sign) {
    VertexJvpP P{};
    P.ok = false;
    const v3 rkv = pk - pi;
    P.dik = norm(rkv);
    if (P.dik < 1e-12) return P;
    P.wk = rkv / P.dik;
    const v3 rlv = pl - pi;
    P.dil = norm(rlv);
    if (P.dil < 1e-12) return P;
    P.wl = rlv / P.dil;
    P.cak = (P.dik * P.dik + ri * ri - rk * rk) / (2.0 * P.dik * ri);
    P.cal = (P.dil * P.dil + ri * ri - rl * rl) / (2.0 * P.dil * ri);
    P.sigma = dot(P.wk, P.wl);
    P.denom = 1.0 - P.sigma * P.sigma;
    if (fabs(P.denom) < 1e-12) return P;
    P.a = (P.cak - P.sigma * P.cal) / P.denom;
    P.b = (P.cal - P.sigma * P.cak) / P.denom;
    const double c_sq = 1.0 - P.a * P.a - P.b * P.b - 2.0 * P.a * P.b * P.sigma;
    if (c_sq <= 0.0) return P;
    P.c = sqrt(c_sq);
    const v3 m = cross(P.wk, P.wl);
    P.mn = sqrt(norm2(m));
    if (P.mn < 1e-12) return P;
    P.np = m / P.mn;
    P.ck = cos_alpha_primal(pi, pk, ri, rk);
    P.cl = cos_alpha_primal(pi, pl, ri, rl);
    P.num_a = P.cak - P.sigma * P.cal;
    P.num_b = P.cal - P.sigma * P.cak;
    P.denom2 = P.denom * P.denom;
    P.sign = sign;
    P.ok = true;
    return P;
}

OG_HD inline v3 vertex_jvp_tangent(const VertexJvpP& P, v3 dri, v3 drk, v3 drl) {
    const v3 dvk = drk - dri;
    const v3 dwk = (dvk - P.wk * dot(P.wk, dvk)) / P.dik;
    const v3 dvl = drl - dri;
    const v3 dwl = (dvl - P.wl * dot(P.wl, dvl)) / P.dil;
    const double dcos_k = cos_alpha_tangent(P.ck, dri, drk);
    const double dcos_l = cos_alpha_tangent(P.cl, dri, drl);
    const double dsig = dot(dwk, P.wl) + dot(P.wk, dwl);
    const double dden = -2.0 * P.sigma * dsig;
    const double dnum_a = dcos_k - dsig * P.cal - P.sigma * dcos_l;
    const double da = (dnum_a * P.denom - P.num_a * dden) / P.denom2;
    const double dnum_b = dcos_l - dsig * P.cak - P.sigma * dcos_k;
    const double db = (dnum_b * P.denom - P.num_b * dden) / P.denom2;
    const double dc_sq = -2.0 * P.a * da - 2.0 * P.b * db - 2.0 * (da * P.b + P.a * db) * P.sigma -
                         2.0 * P.a * P.b * dsig;
    const double dc = dc_sq / (2.0 * P.c);
    const v3 dbase = P.wk * da + dwk * P.a + P.wl * db + dwl * P.b;
    const v3 dm = cross(dwk, P.wl) + cross(P.wk, dwl);
    const v3 dnp = (dm - P.np * dot(P.np, dm)) / P.mn;
    return dbase + (P.np * dc + dnp * P.c) * P.sign;
}

// arc_theta_jvp (derivatives.rs:346-444) in primal / tangent form
struct ArcThetaP {
    VertexJvpP vs_p, ve_p;
    v3 wk, vs, ve, us, ue, x;
    double dik, vdo_s, vdo_e, nn_s, nn_e, cos_t, sin_t;
    bool ok;
};

OG_HD inline ArcThetaP arc_theta_primal(v3 pi, v3 pk, v3 pl, v3 pm, double ri, double rk, double rl, double rm,
                                        double sign_s, double sign_e) {
    ArcThetaP P{};
    P.ok = false;
    const v3 rkv = pk - pi;
    P.dik = norm(rkv);
    if (P.dik < 1e-12) return P;
    P.wk = rkv / P.dik;
    const double cak = (P.dik * P.dik + ri * ri - rk * rk) / (2.0 * P.dik * ri);
    const Cap ck = small_circle_new(P.wk, cak);
    if (norm(pl - pi) < 1e-12) return P;
    const Cap cl = cap_from(pi, pl, ri, rl);
    if (norm(pm - pi) < 1e-12) return P;
    const Cap cm = cap_from(pi, pm, ri, rm);
    if (!vertex_point(ck, cl, sign_s, P.vs)) return P;
    if (!vertex_point(ck, cm, sign_e, P.ve)) return P;
    P.vs_p = vertex_jvp_primal(pi, pk, pl, ri, rk, rl, sign_s);
    if (!P.vs_p.ok) return P;
    P.ve_p = vertex_jvp_primal(pi, pk, pm, ri, rk, rm, sign_e);
    if (!P.ve_p.ok) return P;
    {
        P.vdo_s = dot(P.vs, P.wk);
        const v3 proj = P.vs - P.wk * P.vdo_s;
        P.nn_s = norm(proj);
        if (P.nn_s < 1e-12) return P;
        P.us = proj / P.nn_s;
    }
    {
        P.vdo_e = dot(P.ve, P.wk);
        const v3 proj = P.ve - P.wk * P.vdo_e;
        P.nn_e = norm(proj);
        if (P.nn_e < 1e-12) return P;
        P.ue = proj / P.nn_e;
    }
    P.cos_t = dot(P.us, P.ue);
    P.x = cross(P.us, P.ue);
    P.sin_t = dot(P.wk, P.x);
    P.ok = true;
    return P;
}

OG_HD inline double arc_theta_tangent(const ArcThetaP& P, v3 dri, v3 drk, v3 drl, v3 drm) {
    const v3 dvv = drk - dri;
    const v3 dwk = (dvv - P.wk * dot(P.wk, dvv)) / P.dik;
    const v3 dvs = vertex_jvp_tangent(P.vs_p, dri, drk, drl);
    const v3 dve = vertex_jvp_tangent(P.ve_p, dri, drk, drm);
    v3 dus, due;
    {
        const double dvdo = dot(dvs, P.wk) + dot(P.vs, dwk);
        const v3 dproj = dvs - dwk * P.vdo_s - P.wk * dvdo;
        dus = (dproj - P.us * dot(P.us, dproj)) / P.nn_s;
    }
    {
        const double dvdo = dot(dve, P.wk) + dot(P.ve, dwk);
        const v3 dproj = dve - dwk * P.vdo_e - P.wk * dvdo;
        due = (dproj - P.ue * dot(P.ue, dproj)) / P.nn_e;
    }
    const double dcos_t = dot(dus, P.ue) + dot(P.us, due);
    const v3 dx = cross(dus, P.ue) + cross(P.us, due);
    const double dsin_t = dot(dwk, P.x) + dot(P.wk, dx);
    return P.cos_t * dsin_t - P.sin_t * dcos_t;
}

// vertex_epsilon_jvp (derivatives.rs:466-545) in primal / tangent form
struct VertexEpsP {
    VertexJvpP vp;
    v3 wk, wl, v, tin, tout, x;
    double dik, dil, nn_in, nn_out, ce, se;
    bool ok;
};

OG_HD inline VertexEpsP vertex_eps_primal(v3 pi, v3 pk, v3 pl, double ri, double rk, double rl, double sign) {
    VertexEpsP P{};
    P.ok = false;
    const v3 rkv = pk - pi;
    P.dik = norm(rkv);
    if (P.dik < 1e-12) return P;
    P.wk = rkv / P.dik;
    const v3 rlv = pl - pi;
    P.dil = norm(rlv);
    if (P.dil < 1e-12) return P;
    P.wl = rlv / P.dil;
    const double cak = (P.dik * P.dik + ri * ri - rk * rk) / (2.0 * P.dik * ri);
    const double cal = (P.dil * P.dil + ri * ri - rl * rl) / (2.0 * P.dil * ri);
    const Cap ck = small_circle_new(P.wk, cak), cl = small_circle_new(P.wl, cal);
    if (!vertex_point(ck, cl, sign, P.v)) return P;
    P.vp = vertex_jvp_primal(pi, pk, pl, ri, rk, rl, sign);
    if (!P.vp.ok) return P;
    {
        const v3 cr = cross(P.v, P.wk);
        P.nn_in = norm(cr);
        if (P.nn_in < 1e-12) return P;
        P.tin = cr / P.nn_in;
    }
    {
        const v3 cr = cross(P.v, P.wl);
        P.nn_out = norm(cr);
        if (P.nn_out < 1e-12) return P;
        P.tout = cr / P.nn_out;
    }
    P.ce = dot(P.tin, P.tout);
    P.x = cross(P.tin, P.tout);
    P.se = dot(P.v, P.x);
    P.ok = true;
    return P;
}

OG_HD inline double vertex_eps_tangent(const VertexEpsP& P, v3 dri, v3 drk, v3 drl) {
    const v3 dvk = drk - dri;
    const v3 dwk = (dvk - P.wk * dot(P.wk, dvk)) / P.dik;
    const v3 dvl = drl - dri;
    const v3 dwl = (dvl - P.wl * dot(P.wl, dvl)) / P.dil;
    const v3 dv = vertex_jvp_tangent(P.vp, dri, drk, drl);
    v3 dtin, dtout;
    {
        const v3 dcr = cross(dv, P.wk) + cross(P.v, dwk);
        dtin = (dcr - P.tin * dot(P.tin, dcr)) / P.nn_in;
    }
    {
        const v3 dcr = cross(dv, P.wl) + cross(P.v, dwl);
        dtout = (dcr - P.tout * dot(P.tout, dcr)) / P.nn_out;
    }
    const double dce = dot(dtin, P.tout) + dot(P.tin, dtout);
    const v3 dx = cross(dtin, P.tout) + cross(P.tin, dtout);
    const double dse = dot(dv, P.x) + dot(P.v, dx);
    return P.ce * dse - P.se * dce;
}

OG_HD inline v3 ldp(const double* pos, int a) {
    return mk(pos[3 * (size_t)a], pos[3 * (size_t)a + 1], pos[3 * (size_t)a + 2]);
}

OG_HD inline v3 unit(int axis) { return mk(axis == 0, axis == 1, axis == 2); }

}  // namespace og_sasa

