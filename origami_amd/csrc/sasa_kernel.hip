// Analytical hydrophobic-SASA forces (PowerSasa, Klenin et al. 2011 spherical
// Gauss-Bonnet), FP64.
//
// Per atom i with nonzero surface tension: overlapping neighbours -> buried
// caps on i's unit sphere -> accessible-region boundary (arcs on cap circles
// joined at vertices) -> forward-mode derivatives of
// A_i / R_i^2 = 2*pi*chi - sum_arcs cos(a) theta - sum_vertices eps,
// scattered as F_x -= gamma_i * dA_i/dr_x.
//
// Restates origami:
//   caps / boundary      energy/src/powersasa/arrangement.rs:69-289
//   circle intersection  energy/src/powersasa/geometry.rs:37-100
//   vertex sign / JVPs   energy/src/powersasa/derivatives.rs:92-545
//   cache + gradient     energy/src/powersasa/analytical.rs:101-214, 285-421, 583-649
// Topology decisions use origami's FP64 tolerances unchanged (1e-12, 1e-9,
// 1e-10) and no FMA contraction (sasa_core.h), which is why this stays FP64.
//
// Three kernels per force evaluation:
//
//  og_k_sasa_arrange  one wave64 per atom.  Caps from ballot-compacted
//      neighbour candidates (ascending atom order; far tiles culled by their
//      bounding spheres).  The boundary is built for many caps at once: the
//      caps are cut into chunks whose transitions fit in LDS; within a chunk
//      lanes run over (cap k, cap m) pairs in origami's push order (k, then m
//      ascending, p before q), a rank sort on (angle, push index) inside each
//      cap's segment reproduces the stable angular sort, each cap's
//      transitions are deduplicated against the last kept point, and lanes
//      run over "slots" (boundary segments, or the probe point of a cap with
//      no transitions) for the midpoint / probe burial tests.  Emission is
//      ballot-compacted in slot order, which is origami's (k, segment) order.
//      Vertex root signs and arc->vertex links follow, and the atom's arcs,
//      vertices and cap owners are appended to a pool.
//  og_k_sasa_terms    thread per arc or vertex: its JVP terms (role x axis),
//      the primal geometry formed once.
//  og_k_sasa_scatter  one wave per atom, lane per affected atom x: the terms
//      of every arc (in arc order) and then every vertex (in vertex order)
//      that involve x are summed, which is the order origami sums them in
//      (terms of arcs/vertices not involving x are exact zeros there), and
//      added to the int64 fixed-point force accumulator.
//
// The pool placement depends on which wave allocates first, but every
// value depends only on the atom's own arrangement, so results are
// bit-reproducible.
#include <cstddef>
#include <string>
#include <type_traits>

#include "og_device.h"
#include "og_kernels.h"
#include "og_sasa.h"
#include "sasa_core.h"

#define SA_WAVES 2                    // waves per block (arrangement, scatter)
#define SA_TMAX 256                   // transitions per chunk (>= 2 * (MAXC - 1))
#define SA_NC_MAIN 80                 // caps per atom in the main arrangement pass
#define SA_COV_CAPS (SA_TMAX / 6)     // union cover test: circles per batch (bases in ta)
#define SA_COV_PAIRS (3 * SA_TMAX)    // union cover test: arcs per batch (in tx..tz)
#define SA_COV_EPS 1e-5f              // cover margin on angles (rad)
#define SA_QUEUES 8                   // arrangement work queues (one per XCD)
#define SA_QSTRIDE 64                 // ints between queue counters (256 bytes)
#define SA_NCOUNTERS (SA_QSTRIDE * (SA_QUEUES + 1))

// per-wave global staging (arcs and vertices of the atom in progress)
#define WD_ARC 0                      // MAXA x (sx, sy, sz, ex, ey, ez, theta)
#define WD_VTX (WD_ARC + 7 * MAXA)    // MAXV x (px, py, pz)
#define WD_TOTAL (WD_VTX + 3 * MAXV)
#define WI_ARC 0                      // MAXA x (cap, full)
#define WI_VTX (WI_ARC + 2 * MAXA)    // MAXV x (cap in, cap out)
#define WI_TOTAL (WI_VTX + 2 * MAXV)

static std::string g_sasa_err;
const char* og_sasa_error() { return g_sasa_err.c_str(); }

using namespace og_sasa;

namespace {

// Per-wave LDS of the arrangement, for atoms with at most NC caps.  A grid
// of up to 5 blocks per CU fits NC = MAXC and runs one pass; wider grids use
// NC = SA_NC_MAIN (6 blocks per CU fit) and defer the atoms with more caps to
// a second pass with NC = MAXC.
template <int NC>
struct ArrLds {
    double4 cap4[NC];                                // caps: axis, cos(alpha)
    int own[NC];                                     // cap owner (global atom)
    unsigned char hint[NC];                          // a cap that likely buries circle k
    unsigned char skip[NC];                          // circle k provably contributes nothing
    unsigned char act[NC];                           // caps still to process, ascending
    unsigned char hl[NC];                            // ascending cap lists: hinted, cover candidates, not deep
    unsigned char cnt[NC];                           // transitions of cap k (<= 2 (NC - 1))
    unsigned short seg[NC + 1];                      // chunk: segment start per chunk cap
    alignas(16) double ta[SA_TMAX];                  // chunk: transition angles (push order)
    double tx[SA_TMAX], ty[SA_TMAX], tz[SA_TMAX];    // chunk: transition points
    unsigned char tk[SA_TMAX], tm[SA_TMAX], tq[SA_TMAX];  // chunk cap, other cap, p/q/probe
    unsigned char perm[SA_TMAX];                     // sorted order within segments
    unsigned char rm[SA_TMAX];                       // sorted position removed by dedup
    unsigned char est[SA_TMAX];                      // slot state: 0 no arc, 1 undecided, 2 arc
    unsigned char ulist[SA_TMAX];                    // undecided slots
};

// the transition pass lists its pairs as 16-bit words across rm and est
static_assert(offsetof(ArrLds<SA_NC_MAIN>, est) == offsetof(ArrLds<SA_NC_MAIN>, rm) + SA_TMAX &&
              offsetof(ArrLds<MAXC>, est) == offsetof(ArrLds<MAXC>, rm) + SA_TMAX &&
              offsetof(ArrLds<SA_NC_MAIN>, rm) % 2 == 0 && offsetof(ArrLds<MAXC>, rm) % 2 == 0,
              "rm, est must be contiguous and 2-aligned");
// the caps phase buffers atom indices in perm (64 ints)
static_assert(offsetof(ArrLds<SA_NC_MAIN>, perm) % 4 == 0 && offsetof(ArrLds<MAXC>, perm) % 4 == 0,
              "perm must be 4-aligned");
// the union cover test stores its arcs across tx, ty and tz
static_assert(offsetof(ArrLds<SA_NC_MAIN>, tz) == offsetof(ArrLds<SA_NC_MAIN>, tx) + 2 * SA_TMAX * sizeof(double),
              "tx, ty, tz must be contiguous");
static_assert(offsetof(ArrLds<MAXC>, tz) == offsetof(ArrLds<MAXC>, tx) + 2 * SA_TMAX * sizeof(double),
              "tx, ty, tz must be contiguous");

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ unsigned long long lanemask_lt(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ int popc(unsigned long long m) { return __popcll(m); }

__device__ __forceinline__ double rld(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <class LT>
__device__ __forceinline__ Cap cap_at(const LT& L, int k) {
    const double4 c = L.cap4[k];
    return {mk(c.x, c.y, c.z), c.w};
}

__device__ __forceinline__ Cap as_cap(double4 c) { return {mk(c.x, c.y, c.z), c.w}; }

// Lower bound of dot(x, axis_m) over the rotations x of point p about
// axis_k (ct, st: cosine and sine between the axes) clears cap m's
// containment test with margin; points near the axis are not trusted.
__device__ __forceinline__ bool bound_inside(v3 p, const Cap& ck, double ct, double st, double ca_m) {
    const double h = dot(p, ck.axis);
    const double rn = norm(p - ck.axis * h);
    return rn >= 1e-3 && h * ct - rn * st > ca_m + 1e-11;
}

// Segment midpoints for the first burial round.  The midpoint is the
// rotation of point a by ccw/2 about axis n, n*h + r*cos + (n x r)*sin with
// h = a.n and r = a - n*h, the same point midpoint_on_circle forms; here
// sine and cosine come from sincosf (a few FP32 ulp, plus the FP32 rounding
// of the angle), which keeps it within 1e-6 of the exact midpoint.  A cap
// that contains it with margin 1e-4 therefore also passes the exact test
// dot(mid, axis) > cos(a) + 1e-12 (the slot is buried with certainty), and
// one it lies outside of by 1e-4 fails it.  Slots that neither settles take
// the exact route.
__device__ __forceinline__ v3 approx_midpoint(v3 a, double4 n4, double ccw) {
    const v3 n = mk(n4.x, n4.y, n4.z);
    const double h = dot(a, n);
    const v3 r = a - n * h;
    const v3 t = cross(n, r);
    float sd, cd;
    sincosf((float)(0.5 * ccw), &sd, &cd);
    return n * h + r * (double)cd + t * (double)sd;
}

template <int CTRL, int ROWS, int BANKS>
__device__ __forceinline__ unsigned dpp_umax(unsigned acc, unsigned src) {
    const unsigned m = (unsigned)__builtin_amdgcn_update_dpp(0, (int)src, CTRL, ROWS, BANKS, false);
    return acc > m ? acc : m;
}

// Row-wise (16-lane) reductions over DPP: after the five row-local stages
// (shr 1, 2, 3 of the input, shr 4 and 8 within the row) lane 15 of each row
// holds the row's maximum (minimum); lanes a stage does not write read the
// identity.  All lanes must be active.
__device__ __forceinline__ unsigned row_umax(unsigned x) {
    unsigned y = dpp_umax<0x111, 0xF, 0xF>(x, x);
    y = dpp_umax<0x112, 0xF, 0xF>(y, x);
    y = dpp_umax<0x113, 0xF, 0xF>(y, x);
    y = dpp_umax<0x114, 0xF, 0xE>(y, y);
    return dpp_umax<0x118, 0xF, 0xC>(y, y);
}

template <int CTRL, int ROWS, int BANKS>
__device__ __forceinline__ unsigned dpp_umin(unsigned acc, unsigned src) {
    const unsigned m = (unsigned)__builtin_amdgcn_update_dpp(-1, (int)src, CTRL, ROWS, BANKS, false);
    return acc < m ? acc : m;
}

__device__ __forceinline__ unsigned row_umin(unsigned x) {
    unsigned y = dpp_umin<0x111, 0xF, 0xF>(x, x);
    y = dpp_umin<0x112, 0xF, 0xF>(y, x);
    y = dpp_umin<0x113, 0xF, 0xF>(y, x);
    y = dpp_umin<0x114, 0xF, 0xE>(y, y);
    return dpp_umin<0x118, 0xF, 0xC>(y, y);
}

// The value lane 15 of the caller's row holds.
__device__ __forceinline__ unsigned row_lane15(unsigned y, int row) {
    const unsigned r0 = (unsigned)__builtin_amdgcn_readlane((int)y, 15);
    const unsigned r1 = (unsigned)__builtin_amdgcn_readlane((int)y, 31);
    const unsigned r2 = (unsigned)__builtin_amdgcn_readlane((int)y, 47);
    const unsigned r3 = (unsigned)__builtin_amdgcn_readlane((int)y, 63);
    return row == 0 ? r0 : row == 1 ? r1 : row == 2 ? r2 : r3;
}

// Union cover walk: how far past angle p the arc (start e.x, width e.y) reaches
// when it contains p with SA_COV_EPS to spare, else 0.  The turn count comes
// from a multiply by 1/(2 pi): where it differs from the rounded quotient the
// reduced angle lies within 4e-6 of 0 or 2 pi, which both forms reject (arc
// widths stay below 2 pi - 0.08), so the result equals the divided form's.
// An empty slot (width -1) fails d < e.y - SA_COV_EPS for every reduced
// angle (d > -4e-6), so no branch is needed for it.
__device__ __forceinline__ float cover_ext(float p, float2 e) {
    const float two_pi = 6.28318530717958647692f;
    float d = p - e.x;
    d -= two_pi * floorf(d * (1.0f / two_pi));
    return (d > SA_COV_EPS && d < e.y - SA_COV_EPS) ? e.y - d : 0.0f;
}

__device__ __forceinline__ void add_sasa_energy(const og_sasa_data& d, int ia, double area) {
    // powersasa_energy (mod.rs:55-111): gamma * A, no gamma scale
    const int sys = d.atom_sys[ia];
    og_atomic_add_fix(d.eacc + (size_t)sys * OG_NE + OG_E_SASA,
                      og_to_fix_chk(d.gamma[ia] * area, d.sflag + sys));
}

// An atom whose arrangement does not fit the kernel's capacity: counted, and
// its system flagged so that dynamics stop it at this step.
__device__ __forceinline__ void sasa_capacity_exceeded(const og_sasa_data& d, int ia) {
    atomicAdd(d.overflow, 1);
    atomicOr(d.sflag + d.atom_sys[ia], OG_ST_SASA);
}

__device__ __forceinline__ double wave_min_d(double v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int wave_min_i(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}

// Accessible area of one atom from its boundary (area.rs:19-126):
// A = R^2 (2 pi (2c - L) - sum_arcs cos(a) theta - sum_vertices eps), clamped
// to [0, 4 pi R^2]; L and the vertex sum from the face walk
// (walk_faces_summing_vertices, area.rs:140-222), c from the probe
// components (count_accessible_components, arrangement.rs:327-389) capped
// at L and at least 1.  The chunk arrays of ArrLds are free here and hold
// the walk's visited flags, the probes' accessibility and their labels.
template <class LT>
__device__ void atom_area(LT& L, const og_sasa_data& d, int ia, double ri, int nc, int na, int lane,
                          const double* wd, const int* wi) {
    const double two_pi = 2.0 * 3.14159265358979323846;
    unsigned char* visited = L.perm;                                  // [na] (na <= 256)
    for (int a = lane; a < na; a += 64) visited[a] = 0;
    wave_sync();
    // ---- face walk: lanes over candidate arcs, the walk itself is uniform ----
    int loops = 0;
    double vsum = 0.0;
    for (int start = 0; start < na; ++start) {
        if (visited[start]) continue;
        if (wi[WI_ARC + 2 * start + 1]) {          // full circle: a loop, no vertices
            if (lane == 0) visited[start] = 1;
            wave_sync();
            ++loops;
            continue;
        }
        int cur = start;
        for (int guard = 0; guard <= na; ++guard) {   // origami's walk has no guard
            if (lane == 0) visited[cur] = 1;
            wave_sync();
            const double* Ac = wd + WD_ARC + 7 * cur;
            const v3 end_pt = mk(Ac[3], Ac[4], Ac[5]);
            const v3 t_in = normalize(cross(end_pt, cap_at(L, wi[WI_ARC + 2 * cur]).axis));
            int first = 0x7fffffff, best_i = 0x7fffffff;
            double best_a = INFINITY;
            for (int base = 0; base < na; base += 64) {
                const int c = base + lane;
                double ang = INFINITY;
                if (c < na && !wi[WI_ARC + 2 * c + 1]) {
                    const double* Cc = wd + WD_ARC + 7 * c;
                    if (norm2(mk(Cc[0], Cc[1], Cc[2]) - end_pt) < 1e-10) {
                        first = min(first, c);
                        if (c != cur) {
                            const v3 t_out = normalize(cross(end_pt, cap_at(L, wi[WI_ARC + 2 * c]).axis));
                            const double dt = fmin(fmax(dot(t_in, t_out), -1.0), 1.0);
                            ang = atan2(dot(end_pt, cross(t_in, t_out)), dt);
                            if (ang <= 1e-9) ang += two_pi;
                        }
                    }
                }
                // first strictly smaller angle in candidate order = argmin, ties to the lower index
                const double m = wave_min_d(ang);
                const int who = wave_min_i(ang == m && m < best_a ? c : 0x7fffffff);
                if (m < best_a) { best_a = m; best_i = who; }
            }
            first = wave_min_i(first);
            if (first == 0x7fffffff) break;           // open path
            const int best = best_i != 0x7fffffff ? best_i : first;
            const v3 t_out = normalize(cross(end_pt, cap_at(L, wi[WI_ARC + 2 * best]).axis));
            const double dt = fmin(fmax(dot(t_in, t_out), -1.0), 1.0);
            vsum += atan2(dot(end_pt, cross(t_in, t_out)), dt);
            if (best == start) break;
            cur = best;
        }
        ++loops;
    }
    wave_sync();
    // ---- accessible probe components ----
    unsigned char* acc = L.tk;                                        // [OG_SASA_PROBES] (tk..tq, 768 bytes)
    int* label = reinterpret_cast<int*>(L.ta);                        // [OG_SASA_PROBES] (ta, 2 KB)
    bool any_acc = false;
    for (int i = lane; i < OG_SASA_PROBES; i += 64) {
        const double4 pr = d.probes[i];
        bool buried = false;
        for (int k = 0; k < nc && !buried; ++k) buried = contains_strict(cap_at(L, k), mk(pr.x, pr.y, pr.z));
        acc[i] = buried ? 0 : 1;
        label[i] = buried ? 0x7fffffff : i;
        any_acc = any_acc || !buried;
    }
    int comps = 0;
    if (__ballot(any_acc)) {
        wave_sync();
        for (int it = 0; it < OG_SASA_PROBES; ++it) {
            bool changed = false;
            for (int i = lane; i < OG_SASA_PROBES; i += 64) {
                if (!acc[i]) continue;
                int m = label[i];
                for (int q = d.padj_start[i]; q < d.padj_start[i + 1]; ++q) {
                    const int j = d.padj[q];
                    if (acc[j]) m = min(m, label[j]);
                }
                if (m < label[i]) { label[i] = m; changed = true; }
            }
            wave_sync();
            if (!__ballot(changed)) break;
        }
        for (int i = lane; i < OG_SASA_PROBES; i += 64) comps += (acc[i] && label[i] == i) ? 1 : 0;
        for (int o = 32; o >= 1; o >>= 1) comps += __shfl_xor(comps, o, 64);
    }
    // ---- area ----
    if (lane == 0) {
        double arc_sum = 0.0;
        for (int a = 0; a < na; ++a) arc_sum += cap_at(L, wi[WI_ARC + 2 * a]).ca * wd[WD_ARC + 7 * a + 6];
        double area = 0.0;
        if (loops > 0) {
            int c = comps < loops ? comps : loops;
            if (c < 1) c = 1;
            const double r_sq = ri * ri;
            area = r_sq * (two_pi * (double)(2 * c - loops) - arc_sum - vsum);
            if (area < 0.0) area = 0.0;
            if (area > 2.0 * two_pi * r_sq) area = 2.0 * two_pi * r_sq;
        }
        add_sasa_energy(d, ia, area);
    }
    wave_sync();
}

// Arrangement of one work atom; returns its pool record (empty when the atom
// contributes nothing or exceeded capacity).
template <bool AREA, int NC>
__device__ int4 arrange_atom(ArrLds<NC>& L, const og_sasa_data& d, const double* pos, int w, int lane,
                             unsigned long long lt, double* wd, int* wi) {
    const int4 empty = make_int4(0, 0, 0, 0);
    const double two_pi = 2.0 * 3.14159265358979323846;
    const int ia = d.work[w];
    const v3 pi = ldp(pos, ia);
    const double ri = d.radius[ia];

    // ---------------- caps (build_atom_cache + build_caps) ----------------
    // Candidates in ascending atom order; an enclosing neighbour ends the
    // scan (no accessible surface) unless the cap count overflowed first.
    // A neighbour can only be a cap or enclose
    // atom i when |p_j - p_i| <= r_i + r_j (rounded); the scan keeps the
    // atoms whose squared distance is within (r_i + r_j)^2 (1 + 1e-12),
    // which covers the rounding of the square root, and buffers them in LDS
    // (ascending); origami's exact tests then run on full waves of buffered
    // atoms, so the square root and the divisions are formed only for them.
    int nc = 0, nb = 0;
    bool enclosed = false, overflow = false;
    int* cj = reinterpret_cast<int*>(L.perm);              // [64] buffered atoms (chunk arrays are free here)
    double* crj = L.ta;                                    // [64] their radii, positions in tx, ty, tz
    auto flush = [&]() {
        bool is_cap = false, enc = false;
        v3 axis = mk(0, 0, 0);
        double ca = 0.0;
        int j = 0;
        if (lane < nb) {
            j = cj[lane];
            const v3 pj = mk(L.tx[lane], L.ty[lane], L.tz[lane]);
            const double rj = crj[lane];
            const double dd = norm(pj - pi);
            if (dd <= ri + rj) {
                if (dd + ri <= rj) {
                    enc = true;
                } else if (!(dd + rj <= ri || dd >= ri + rj)) {
                    ca = (dd * dd + ri * ri - rj * rj) / (2.0 * dd * ri);
                    if (ca >= -1.0 && ca <= 1.0) {
                        is_cap = true;
                        axis = (pj - pi) / dd;
                    }
                }
            }
        }
        unsigned long long m = __ballot(is_cap);
        const unsigned long long em = __ballot(enc);
        if (em) m &= (em & (~em + 1)) - 1;     // caps before the first enclosing atom
        nb = 0;
        if (nc + popc(m) > NC) { overflow = true; return; }
        if ((m >> lane) & 1ull) {
            const int p = nc + popc(m & lt);
            L.cap4[p] = make_double4(axis.x, axis.y, axis.z, ca);
            L.own[p] = j;
        }
        nc += popc(m);
        if (em) enclosed = true;
    };
    // near tiles in ascending order (culled 64 at a time); the next tile's
    // atoms load while the current one is scanned; one flush site (after
    // the last tile, t < 0, it takes what is left)
    const int t_end = d.tile_end[w];
    int t_base = 0, t_next = d.tile_begin[w];
    unsigned long long near_m = 0;
    auto next_tile = [&]() -> int {
        while (near_m == 0 && t_next < t_end) {
            bool near = false;
            if (t_next + lane < t_end) {
                const double4 tb = d.tile_bound[t_next + lane];
                const double dxc = tb.x - pi.x, dyc = tb.y - pi.y, dzc = tb.z - pi.z;
                const double reach = ri + d.max_radius + tb.w;
                near = !(dxc * dxc + dyc * dyc + dzc * dzc > reach * reach);
            }
            near_m = __ballot(near);
            t_base = t_next;
            t_next += 64;
        }
        if (near_m == 0) return -1;
        const int t = t_base + __builtin_ctzll(near_m);
        near_m &= near_m - 1;
        return t;
    };
    int t = next_tile(), j = 0;
    bool valid = false;
    v3 pj = mk(0, 0, 0);
    double rj = 0.0;
    if (t >= 0) {
        j = d.tile_atom0[t] + lane;
        valid = lane < d.tile_cnt[t] && j != ia;
        if (valid) { pj = ldp(pos, j); rj = d.radius[j]; }
    }
    for (;;) {
        const int tn = t >= 0 ? next_tile() : -1;
        int jn = 0;
        bool validn = false;
        v3 pjn = mk(0, 0, 0);
        double rjn = 0.0;
        if (tn >= 0) {
            jn = d.tile_atom0[tn] + lane;
            validn = lane < d.tile_cnt[tn] && jn != ia;
            if (validn) { pjn = ldp(pos, jn); rjn = d.radius[jn]; }
        }
        bool keep = false;
        if (t >= 0 && valid) {
            const v3 dv = pj - pi;
            const double rs = ri + rj;
            keep = dot(dv, dv) <= rs * rs * (1.0 + 1e-12);
        }
        const unsigned long long km = __ballot(keep);
        if (t < 0 ? nb > 0 : nb + popc(km) > 64) {
            wave_sync();
            flush();
            wave_sync();
            if (enclosed || overflow) break;
        }
        if (t < 0) break;
        if (keep) {
            const int q = nb + popc(km & lt);
            cj[q] = j;
            crj[q] = rj;
            L.tx[q] = pj.x; L.ty[q] = pj.y; L.tz[q] = pj.z;
        }
        nb += popc(km);
        t = tn; j = jn; valid = validn; pj = pjn; rj = rjn;
    }
    wave_sync();
    if (overflow) {
        // more caps than this pass holds: the MAXC pass takes the atom
        if (lane == 0) {
            if (NC < MAXC) d.deferred[atomicAdd(d.counters + 3, 1)] = w;
            else sasa_capacity_exceeded(d, ia);
        }
        return empty;
    }
    if (AREA && !enclosed && nc == 0 && lane == 0)
        add_sasa_energy(d, ia, 4.0 * 3.14159265358979323846 * ri * ri);   // fully exposed
    if (enclosed || nc == 0) return empty;

    // ---------------- burial hints ----------------
    // hint[k]: the first cap m that (up to rounding) contains circle k
    // entirely.  It drives the exact skip test that follows (any cap that
    // contains the circle serves there; the test itself is exact).  The scan
    // has no branches: one cap per step, FMA arithmetic.  It also lists the
    // caps whose axis is within about 5.7 degrees of axis k (1 - cos^2 <
    // 1.0001e-2, a superset of the exact test's < 1e-2); those are then
    // checked with origami's arithmetic, and a nearly parallel circle that
    // meets circle k sets near_par exactly as a full exact scan would.
    // Transition counts are formed later, for the caps that stay active.
    for (int k = lane; k < nc; k += 64) {
        const Cap ck = cap_at(L, k);
        const double sk2 = fmax(1.0 - ck.ca * ck.ca, 0.0);
        int hint = 0xFF, np_m = 0, np_n = 0;
#pragma unroll 2
        for (int m = 0; m < nc; ++m) {
            const double4 c = L.cap4[m];
            const double sm = fma(ck.axis.z, c.z, fma(ck.axis.y, c.y, ck.axis.x * c.x));
            const double dm = fma(-sm, sm, 1.0);
            const double lm = fma(sm, ck.ca, -c.w);
            const bool o = m != k;
            // circle k inside cap m: cos(axes) cos(a_k) - sin(axes) sin(a_k) > cos(a_m)
            const bool in = o & (lm > 0.0) & (lm * lm > dm * sk2);
            hint = ((hint == 0xFF) & in) ? m : hint;
            const bool np = o & (dm < 1.0001e-2);
            np_m = (np & (np_n == 0)) ? m : np_m;
            np_n += np ? 1 : 0;
        }
        // a nearly parallel circle matters only where it meets circle k
        bool near_par = false;
        if (np_n == 1) {
            const Cap cm = cap_at(L, np_m);
            const double s0 = dot(ck.axis, cm.axis);
            near_par = 1.0 - s0 * s0 < 1e-2 && intersect_count_fast(ck, cm) > 0;
        } else if (np_n > 1) {
            for (int m = 0; m < nc; ++m) {
                if (m == k) continue;
                const Cap cm = cap_at(L, m);
                const double s0 = dot(ck.axis, cm.axis);
                if (1.0 - s0 * s0 < 1e-2 && intersect_count_fast(ck, cm) > 0) near_par = true;
            }
        }
        // skip[k] low bits: 0 keep, 1 test every point below, 2 proven (see
        // below); bit 2: a circle meeting circle k is nearly parallel to it;
        // bit 3: proven with depth SA_COV_DELTA (a deep circle, see the
        // transition counts).  ct cos a_k - st sin a_k - cos a_m is the least
        // containment value of cap m on circle k.
        int sk8 = near_par ? 4 : 0;
        if (hint != 0xFF) {
            const Cap ch = cap_at(L, hint);
            const double sk = sqrt(sk2);
            const double ct = dot(ck.axis, ch.axis);
            const double st = norm(cross(ck.axis, ch.axis));
            const double least = ct * ck.ca - st * sk - ch.ca;
            sk8 |= (!near_par && sk >= 0.05 && least > 1e-6) ? (least > SA_COV_DELTA ? 10 : 2) : 1;
        }
        L.hint[k] = (unsigned char)hint;
        L.skip[k] = (unsigned char)sk8;
    }
    wave_sync();

    // ---------------- exact skip of caps buried by their hint cap ---------
    // Every point origami tests on circle k is a segment midpoint (a rotation
    // of a transition point about axis_k, which keeps its height h and
    // radius rn) or, without transitions, the probe point.  Such a point x
    // satisfies dot(x, axis_m) >= h*ct - rn*st (ct, st: cosine and sine
    // between the axes).  When that bound clears cap m's containment test by
    // 1e-11 (>> rounding) for every transition point, or for the probe,
    // circle k contributes no arc and is dropped.
    // skip[k] = 2 settles this without forming the points: the whole circle
    // (h = cos a_k, rn = sin a_k >= 0.05) clears the test by 1e-6, and every
    // circle m that meets circle k is at least 5.7 degrees from parallel
    // (1 - cos^2 >= 1e-2), where intersect_circles keeps |a|, |b| <= 15 and
    // places its points within 1e-12 of circle k in h and rn (a tangent
    // point, |c^2| <= 1e-9, within 1e-8 in rn); the probe point is exact to
    // rounding.  The per-point bound then holds for every point with a
    // hundredfold margin.  skip[k] = 1 tests every point.
    int n_h = 0;
    for (int base = 0; base < nc; base += 64) {
        const int k = base + lane;
        const bool hh = k < nc && (L.skip[k] & 3) == 1;
        const unsigned long long hm = __ballot(hh);
        if (hh) L.hl[n_h + popc(hm & lt)] = (unsigned char)k;
        n_h += popc(hm);
    }
    // has[i]: circle hl[i] meets another circle (else its probe point is tested)
    unsigned char* has = L.act;                       // free until the active list
    for (int i = lane; i < n_h; i += 64) has[i] = 0;
    wave_sync();
    const int nper = nc > 1 ? nc - 1 : 1;
    const PairDiv pdiv = pair_div(nper);                  // pair indices stay below 65536 (NC <= 128)
    for (int base = 0; base < n_h * nper; base += 64) {
        const int p = base + lane;
        if (p < n_h * nper && nc > 1) {
            const int i = pair_q(pdiv, p), mi = p - i * nper;
            const int k = L.hl[i];
            const Cap ck = cap_at(L, k), chint = cap_at(L, L.hint[k]);
            const double ct = dot(ck.axis, chint.axis);
            const double st = norm(cross(ck.axis, chint.axis));
            const int m = mi < k ? mi : mi + 1;
            v3 P = mk(0, 0, 0), Q = mk(0, 0, 0);
            const int r = intersect_circles(ck, cap_at(L, m), P, Q);
            bool ok = true;
            if (r <= 1) {
                has[i] = 1;
                ok = bound_inside(P, ck, ct, st, chint.ca);
            }
            if (r == 0) ok = ok && bound_inside(Q, ck, ct, st, chint.ca);
            if (!ok) L.skip[k] &= 4;
        }
    }
    wave_sync();
    for (int i = lane; i < n_h; i += 64) {
        if (has[i]) continue;
        const int k = L.hl[i];
        const Cap ck = cap_at(L, k), chint = cap_at(L, L.hint[k]);
        const double ct = dot(ck.axis, chint.axis);
        const double st = norm(cross(ck.axis, chint.axis));
        if (!bound_inside(sample_circle_point(ck), ck, ct, st, chint.ca)) L.skip[k] &= 4;
    }
    wave_sync();
    // ---------------- union cover test ----------------
    // A circle k that lies, with margin, inside the union of the other caps
    // contributes no arc either: every point origami tests on it is then
    // inside some cap m != k.  On circle k (point X(phi) = cos a_k n_k +
    // sin a_k (cos(phi) u + sin(phi) v)), cap m's test value is
    // f_m(phi) = B + sin a_k sin(t) cos(phi - phi_m) with B = cos a_k cos(t)
    // - cos a_m (t: angle between the axes), so f_m >= SA_COV_DELTA on the
    // arc |phi - phi_m| <= w_m.  Those arcs are formed in FP64 up to the
    // angles, which come from atan2f/acosf on well-conditioned arguments
    // (|cos w_m| <= 1 - 1e-3, sin a_k sin t >= 1e-3; nearly full arcs are
    // cut to pi - 0.05, tiny ones dropped): each angle is within 2e-6 rad.
    // A greedy walk from the end of the widest arc then steps, each time, to
    // the end of the arc that contains the current angle with SA_COV_EPS
    // (1e-5 rad) to spare and reaches furthest; reaching back into the
    // widest arc proves the circle covered, every angle then lying inside
    // some exact arc.  As for the skip test, the points origami forms lie
    // within 1e-8 of circle k (no nearly parallel partner, sin a_k >= 0.05),
    // which moves f by less than 2e-8 << SA_COV_DELTA = 1e-3.  A covered
    // circle is also deep (bit 3 of skip): every point of it lies inside
    // another cap with containment value >= SA_COV_DELTA.
    int n_cand = 0;
    for (int base = 0; base < nc; base += 64) {
        const int k = base + lane;
        bool c = false;
        if (k < nc && L.skip[k] == 0) {
            const double ca = L.cap4[k].w;
            c = 1.0 - ca * ca >= 0.0025;
        }
        const unsigned long long cm = __ballot(c);
        if (c) L.hl[n_cand + popc(cm & lt)] = (unsigned char)k;
        n_cand += popc(cm);
    }
    wave_sync();
    if (n_cand > 0 && nc > 1) {
        double* bas = L.ta;                                      // [SA_COV_CAPS] x (u, v)
        float2* iv = reinterpret_cast<float2*>(L.tx);            // [SA_COV_PAIRS] (start, width)
        const int per_batch = SA_COV_CAPS < SA_COV_PAIRS / nper ? SA_COV_CAPS : SA_COV_PAIRS / nper;
        for (int c0 = 0; c0 < n_cand; c0 += per_batch) {
            const int nb = n_cand - c0 < per_batch ? n_cand - c0 : per_batch;
            for (int ci = lane; ci < nb; ci += 64) {
                v3 u, v;
                circle_basis(cap_at(L, L.hl[c0 + ci]), u, v);
                double* o = bas + 6 * ci;
                o[0] = u.x; o[1] = u.y; o[2] = u.z; o[3] = v.x; o[4] = v.y; o[5] = v.z;
            }
            wave_sync();
            for (int p = lane; p < nb * nper; p += 64) {
                const int ci = pair_q(pdiv, p), mi = p - ci * nper;
                const int k = L.hl[c0 + ci];
                const int m = mi < k ? mi : mi + 1;
                const double* o = bas + 6 * ci;
                float a, wd2;
                cover_arc(cap_at(L, k), cap_at(L, m), mk(o[0], o[1], o[2]), mk(o[3], o[4], o[5]), a, wd2);
                iv[p] = make_float2(a, wd2);
            }
            wave_sync();
            // The walk, four circles at a time, one per 16-lane row (lane l
            // of row r holds arcs l, l + 16, ... of circle g0 + r): each
            // circle takes exactly its own steps, from the end of its widest
            // arc (the first of equal widths), each step to the furthest
            // reach of the arcs that contain the current angle with
            // SA_COV_EPS to spare; rows that are done stay as they are.  The
            // slots per lane (NS) are a compile-time count, the smallest of
            // 1, 2, 3, 4 and 8 that holds the circle's nper arcs (the extra
            // slots are empty), so the step loop has no branches per slot.
            const int row = lane >> 4, lr = lane & 15;
            auto walk = [&](auto ns) {
                constexpr int NS = decltype(ns)::value;
                for (int g0 = 0; g0 < nb; g0 += 4) {
                    const int ci = g0 + row;
                    const bool rv = ci < nb;
                    float2 e[NS];
                    bool full = false;
                    float wl = 0.0f;
#pragma unroll
                    for (int q = 0; q < NS; ++q) {
                        const int mi = lr + 16 * q;
                        e[q] = (rv && mi < nper) ? iv[ci * nper + mi] : make_float2(0.0f, -1.0f);
                        full = full || e[q].y >= SA_COV_FULL;
                        wl = fmaxf(wl, e[q].y);
                    }
                    const bool rfull = ((__ballot(full) >> (16 * row)) & 0xFFFFull) != 0;
                    const float wm = __uint_as_float(row_lane15(row_umax(__float_as_uint(wl)), row));
                    unsigned first = 0xFFFFu;
#pragma unroll
                    for (int q = NS - 1; q >= 0; --q)
                        first = e[q].y == wm ? (unsigned)(lr + 16 * q) : first;
                    first = row_lane15(row_umin(first), row);
                    bool act = rv && !rfull && wm > 0.0f;
                    bool cov = rv && rfull;
                    float start = 0.0f, len = 0.0f, t = 0.0f;
                    if (act) {
                        start = iv[ci * nper + (int)first].x + wm;
                        len = 6.28318530717958647692f - wm;
                    }
                    for (int step = 0; step <= nper + 1; ++step) {
                        if (!__ballot(act)) break;
                        const float p = start + t;
                        float ext = 0.0f;
#pragma unroll
                        for (int q = 0; q < NS; ++q) ext = fmaxf(ext, cover_ext(p, e[q]));
                        const float best = __uint_as_float(row_lane15(row_umax(__float_as_uint(act ? ext : 0.0f)), row));
                        const bool adv = act && best > 0.0f;
                        t = adv ? t + best : t;
                        const bool done = adv && t >= len + SA_COV_EPS;
                        cov = cov || done;
                        act = adv && !done;
                    }
                    if (lr == 0 && cov) L.skip[L.hl[c0 + ci]] = 10;
                }
            };
            if (nper <= 16) walk(std::integral_constant<int, 1>{});
            else if (nper <= 32) walk(std::integral_constant<int, 2>{});
            else if (nper <= 48) walk(std::integral_constant<int, 3>{});
            else if (nper <= 64) walk(std::integral_constant<int, 4>{});
            else walk(std::integral_constant<int, 8>{});
            wave_sync();
        }
    }

    // active caps, ascending
    int n_act = 0;
    for (int base = 0; base < nc; base += 64) {
        const int k = base + lane;
        const bool a = k < nc && !(L.skip[k] & 3);
        const unsigned long long am = __ballot(a);
        if (a) L.act[n_act + popc(am & lt)] = (unsigned char)k;
        n_act += popc(am);
    }
    wave_sync();
    // Transition counts of the active caps, lanes over (active cap, other
    // cap).  The transitions of deep circles are left out.  A deep circle m
    // lies, at every point, inside another cap with containment value
    // >= SA_COV_DELTA (1e-3), which changes by at most the distance moved,
    // so the accessible region keeps that distance from circle m.  Leaving
    // m's points out of circle k's list merges segments that all lie inside
    // such a cap, which are buried either way; the exposed segments, their
    // ends and the caps there stay the same.  Deduplication (a point within
    // 1e-6 of the last kept one is dropped, so a run of dropped points spans
    // at most 2 (NC - 1) 1e-6 < 2.6e-4) cannot carry a left-out point's
    // effect to an exposed end.  A circle whose transitions are all left out
    // lies inside the accessible region's complement and takes the probe
    // test like a circle without transitions.  The arcs are origami's unless
    // the accessible boundary touches circle k within rounding of a tested
    // midpoint.  The caps that are not deep are listed once (ascending,
    // in hl, free after the cover test), and the pair loops here and in the
    // chunks below run over (active cap, listed cap).
    int n_nd = 0;
    for (int base = 0; base < nc; base += 64) {
        const int k = base + lane;
        const bool nd = k < nc && !(L.skip[k] & 8);
        const unsigned long long ndm = __ballot(nd);
        if (nd) L.hl[n_nd + popc(ndm & lt)] = (unsigned char)k;
        n_nd += popc(ndm);
    }
    const PairDiv ndiv = pair_div(n_nd);                  // n_nd >= n_act >= 1
    {
        unsigned* c32 = reinterpret_cast<unsigned*>(L.ta);      // [n_act], free until the first chunk
        for (int i = lane; i < n_act; i += 64) c32[i] = 0;
        wave_sync();
        for (int p = lane; p < n_act * n_nd; p += 64) {
            const int i = pair_q(ndiv, p), j = p - i * n_nd;
            const int k = L.act[i], m = L.hl[j];
            const int n = m == k ? 0 : intersect_count_fast(cap_at(L, k), cap_at(L, m));
            if (n) atomicAdd(c32 + i, (unsigned)n);
        }
        wave_sync();
        for (int i = lane; i < n_act; i += 64) L.cnt[L.act[i]] = (unsigned char)c32[i];
        wave_sync();
    }

    // ---------------- boundary (find_boundary), chunk by chunk ------------
    int na = 0, nv = 0;
    int a0 = 0;
    while (a0 < n_act && !overflow) {
        int a1 = a0, tb = 0;
        while (a1 < n_act) {
            const int c = L.cnt[L.act[a1]];
            const int e = c > 0 ? c : 1;      // a cap without transitions holds its probe entry
            if (tb + e > SA_TMAX) break;
            tb += e;
            ++a1;
        }
        if (a1 == a0) { overflow = true; break; }   // cannot happen: cnt <= 2 (NC - 1)
        const int nck = a1 - a0;

        // transitions in push order (cap, then m ascending, p before q); a cap
        // without transitions gets one probe entry (tq = 2).  Segment starts
        // come from the counts; the pairs that emit entries (the same fast
        // count decides, exactly as for cnt) are listed first, so that the
        // intersection points and their angles are formed on full waves.
        if (lane == 0) {
            int acc = 0;
            for (int kl = 0; kl < nck; ++kl) {
                L.seg[kl] = (unsigned short)acc;
                const int c = L.cnt[L.act[a0 + kl]];
                acc += c > 0 ? c : 1;
            }
            L.seg[nck] = (unsigned short)acc;
        }
        unsigned short* plist = reinterpret_cast<unsigned short*>(L.rm);   // [SA_TMAX], over rm and est
        const int npairs = nck * n_nd;
        int npl = 0;
        for (int base = 0; base < npairs; base += 64) {
            const int p = base + lane;
            bool emit = false;
            int code = 0;
            if (p < npairs) {
                const int kl = pair_q(ndiv, p), j = p - kl * n_nd;
                const int k = L.act[a0 + kl];
                const int m = L.hl[j];
                if (L.cnt[k] == 0) emit = j == 0;
                else emit = m != k && intersect_count_fast(cap_at(L, k), cap_at(L, m)) != 0;
                code = (kl << 7) | m;
            }
            const unsigned long long em = __ballot(emit);
            const int pos = npl + popc(em & lt);
            if (emit && pos < SA_TMAX) plist[pos] = (unsigned short)code;
            npl += popc(em);
        }
        if (npl > SA_TMAX) { overflow = true; break; }
        wave_sync();
        int T = 0;
        for (int base = 0; base < npl; base += 64) {
            const int q = base + lane;
            int cnt = 0, kl = 0, k = 0, m = 0, probe = 0;
            v3 P = mk(0, 0, 0), Q = mk(0, 0, 0);
            if (q < npl) {
                const int code = plist[q];
                kl = code >> 7;
                k = L.act[a0 + kl];
                if (L.cnt[k] == 0) {
                    probe = 1; cnt = 1; P = sample_circle_point(cap_at(L, k));
                } else {
                    m = code & 127;
                    const int r = intersect_circles(cap_at(L, k), cap_at(L, m), P, Q);
                    cnt = r == 0 ? 2 : (r == 1 ? 1 : 0);
                }
            }
            const unsigned long long b1 = __ballot(cnt >= 1), b2 = __ballot(cnt == 2);
            const int off = T + popc(b1 & lt) + popc(b2 & lt);
            if (cnt >= 1 && off < SA_TMAX) {
                const Cap ck = cap_at(L, k);
                v3 bu, bv;
                circle_basis(ck, bu, bv);
                L.ta[off] = probe ? 0.0 : angular_position_b(P, ck, bu, bv);
                L.tx[off] = P.x; L.ty[off] = P.y; L.tz[off] = P.z;
                L.tk[off] = (unsigned char)kl; L.tm[off] = (unsigned char)m;
                L.tq[off] = (unsigned char)(probe ? 2 : 0);
                if (cnt == 2 && off + 1 < SA_TMAX) {
                    L.ta[off + 1] = angular_position_b(Q, ck, bu, bv);
                    L.tx[off + 1] = Q.x; L.ty[off + 1] = Q.y; L.tz[off + 1] = Q.z;
                    L.tk[off + 1] = (unsigned char)kl; L.tm[off + 1] = (unsigned char)m; L.tq[off + 1] = 1;
                }
            }
            T += popc(b1) + popc(b2);
        }
        if (T > SA_TMAX) { overflow = true; break; }
        for (int x = lane; x < T; x += 64) L.perm[x] = (unsigned char)x;
        wave_sync();
        // every entry lies in its cap's segment (the counts and the points
        // come from the same arithmetic, so this holds; checked all the same)
        {
            bool bad = T != L.seg[nck];
            for (int x = lane; x < T; x += 64) {
                const int kl = L.tk[x];
                bad = bad || x < L.seg[kl] || x >= L.seg[kl + 1];
            }
            if (__ballot(bad)) { overflow = true; break; }
        }

        // stable angular sort inside each cap's segment: rank on (angle, push
        // index).  The fast pass counts smaller angles only; equal angles
        // then land on one position, which the check below detects, and the
        // chunk is ranked again with the push index as tie-break.
        int rk[SA_TMAX / 64];
#pragma unroll
        for (int i = 0; i < SA_TMAX / 64; ++i) {
            const int x = lane + 64 * i;
            rk[i] = 0;
            if (x < T) {
                const int kl = L.tk[x];
                const int s0 = L.seg[kl], s1 = L.seg[kl + 1];
                const double ax = L.ta[x];
                int n0 = 0, n1 = 0;
                int y = s0;
                if ((y & 1) && y < s1) {
                    n0 += L.ta[y] < ax ? 1 : 0;
                    ++y;
                }
                const double2* t2 = reinterpret_cast<const double2*>(L.ta);
                for (; y + 8 <= s1; y += 8) {
                    const double2 v0 = t2[y / 2], v1 = t2[y / 2 + 1], v2 = t2[y / 2 + 2], v3 = t2[y / 2 + 3];
                    n0 += (v0.x < ax ? 1 : 0) + (v0.y < ax ? 1 : 0) + (v1.x < ax ? 1 : 0) + (v1.y < ax ? 1 : 0);
                    n1 += (v2.x < ax ? 1 : 0) + (v2.y < ax ? 1 : 0) + (v3.x < ax ? 1 : 0) + (v3.y < ax ? 1 : 0);
                }
                for (; y + 2 <= s1; y += 2) {
                    const double2 v = t2[y / 2];
                    n0 += (v.x < ax ? 1 : 0) + (v.y < ax ? 1 : 0);
                }
                if (y < s1) n1 += L.ta[y] < ax ? 1 : 0;
                rk[i] = s0 + n0 + n1;
                L.perm[rk[i]] = (unsigned char)x;
            }
        }
        wave_sync();
        bool tied = false;
#pragma unroll
        for (int i = 0; i < SA_TMAX / 64; ++i) {
            const int x = lane + 64 * i;
            if (x < T) tied = tied || L.perm[rk[i]] != x;
        }
        if (__ballot(tied)) {
            wave_sync();
            for (int x = lane; x < T; x += 64) {
                const int kl = L.tk[x];
                const int s0 = L.seg[kl], s1 = L.seg[kl + 1];
                const double ax = L.ta[x];
                int r = 0;
                for (int y = s0; y < s1; ++y) {
                    const double ay = L.ta[y];
                    r += (ay < ax || (ay == ax && y < x)) ? 1 : 0;
                }
                L.perm[s0 + r] = (unsigned char)x;
            }
        }
        wave_sync();

        // dedup against the last kept point (Vec::dedup_by).  rm[y] marks a
        // removed sorted position.  While no removal happens before the last
        // position of a segment, "last kept" is the sorted predecessor and
        // lanes decide independently; otherwise one lane per cap redoes it.
        bool chain = false;
        for (int y = lane; y < T; y += 64) {
            const int kl = L.tk[L.perm[y]];
            const int s0 = L.seg[kl], s1 = L.seg[kl + 1];
            bool rem = false;
            if (y > s0) {
                const int a = L.perm[y], b = L.perm[y - 1];
                const v3 pa = mk(L.tx[a], L.ty[a], L.tz[a]), pb = mk(L.tx[b], L.ty[b], L.tz[b]);
                rem = norm2(pa - pb) < 1e-12;
            }
            L.rm[y] = rem ? 1 : 0;
            chain = chain || (rem && y + 1 < s1);
        }
        if (__ballot(chain)) {
            wave_sync();
            for (int kl = lane; kl < nck; kl += 64) {
                const int s0 = L.seg[kl], s1 = L.seg[kl + 1];
                int last = -1;
                for (int y = s0; y < s1; ++y) {
                    const int a = L.perm[y];
                    bool rem = false;
                    if (last >= 0) {
                        const v3 pa = mk(L.tx[a], L.ty[a], L.tz[a]), pb = mk(L.tx[last], L.ty[last], L.tz[last]);
                        rem = norm2(pa - pb) < 1e-12;
                    }
                    L.rm[y] = rem ? 1 : 0;
                    if (!rem) last = a;
                }
            }
        }
        wave_sync();

        // One slot per kept sorted position: the segment to the next kept
        // position (cyclic), or the probe test of a probe entry.  The burial
        // test is an any() over the other caps, run in two rounds:
        // (A) every slot tests an approximate midpoint against every other
        // cap: contained with margin 1e-4 by one of them, or outside all of
        // them by 1e-4, implies the exact test's answer (a probe entry takes
        // the exact test here, on its exact point);
        // (B) the slots within the margin of some cap form the exact
        // midpoint and scan all caps, lane-parallel over a compacted list.
        // est[y]: 0 no arc, 1 undecided, 2 arc.
        for (int y = lane; y < T; y += 64) {
            unsigned char st8 = 0;
            if (!L.rm[y]) {
                const int x = L.perm[y];
                const int kl = L.tk[x];
                const int kk = L.act[a0 + kl];
                const v3 apt = mk(L.tx[x], L.ty[x], L.tz[x]);
                if (L.tq[x] == 2) {
                    bool in = false;
                    for (int c = 0; c < nc; ++c) in = in | (c != kk && contains_strict(cap_at(L, c), apt));
                    st8 = in ? 0 : 2;
                } else {
                    const int s0 = L.seg[kl], s1 = L.seg[kl + 1];
                    int yn = y + 1 < s1 ? y + 1 : s0;
                    while (L.rm[yn]) yn = yn + 1 < s1 ? yn + 1 : s0;
                    const int bx = L.perm[yn];
                    const double ccw = wrap_2pi_positive(L.ta[bx] - L.ta[x]);
                    if (!(ccw < 1e-9)) {
                        // the largest containment value over the other caps
                        const v3 mid = approx_midpoint(apt, L.cap4[kk], ccw);
                        double vmax = -1.0e300;
                        for (int c = 0; c < nc; ++c) {
                            const double4 cc = L.cap4[c];
                            const double v = fma(cc.z, mid.z, fma(cc.y, mid.y, cc.x * mid.x)) - cc.w;
                            vmax = fmax(vmax, c != kk ? v : -1.0e300);
                        }
                        st8 = vmax > 1e-4 ? 0 : (vmax > -1e-4 ? 1 : 2);
                    }
                }
            }
            L.est[y] = st8;
        }
        wave_sync();
        int U = 0;
        for (int base = 0; base < T; base += 64) {
            const int y = base + lane;
            const bool u = y < T && L.est[y] == 1;
            const unsigned long long um = __ballot(u);
            if (u) L.ulist[U + popc(um & lt)] = (unsigned char)y;
            U += popc(um);
        }
        wave_sync();
        // Round B: exact midpoints lane-parallel, then one slot at a time with
        // the caps on the lanes (any() over the caps = one ballot).
        if (U > 0) {
            const Cap cl0 = cap_at(L, lane < nc ? lane : 0);
            const Cap cl1 = cap_at(L, lane + 64 < nc ? lane + 64 : 0);
            for (int qb = 0; qb < U; qb += 64) {
                const int q = qb + lane;
                v3 probe = mk(0, 0, 0);
                int kk = -1;
                if (q < U) {
                    const int y = L.ulist[q];
                    const int x = L.perm[y];
                    const int kl = L.tk[x];
                    kk = L.act[a0 + kl];
                    const v3 apt = mk(L.tx[x], L.ty[x], L.tz[x]);
                    probe = apt;
                    if (L.tq[x] != 2) {
                        const int s0 = L.seg[kl], s1 = L.seg[kl + 1];
                        int yn = y + 1 < s1 ? y + 1 : s0;
                        while (L.rm[yn]) yn = yn + 1 < s1 ? yn + 1 : s0;
                        const double ccw = wrap_2pi_positive(L.ta[L.perm[yn]] - L.ta[x]);
                        probe = midpoint_on_circle(apt, cap_at(L, kk), ccw / 2.0);
                    }
                }
                const int nq = U - qb < 64 ? U - qb : 64;
                for (int j = 0; j < nq; ++j) {
                    const v3 pj = mk(rld(probe.x, j), rld(probe.y, j), rld(probe.z, j));
                    const int kj = __builtin_amdgcn_readlane(kk, j);
                    const bool b = (lane < nc && lane != kj && contains_strict(cl0, pj)) ||
                                   (lane + 64 < nc && lane + 64 != kj && contains_strict(cl1, pj));
                    const bool buried = __ballot(b) != 0;
                    if (lane == j) L.est[L.ulist[q]] = buried ? 0 : 2;
                }
            }
        }
        wave_sync();
        // emission in slot order
        for (int base = 0; base < T; base += 64) {
            const int y = base + lane;
            const bool emit = y < T && L.est[y] == 2;
            bool segm = false;
            double theta = 0.0;
            v3 st = mk(0, 0, 0), en = mk(0, 0, 0);
            int kk = 0, oth = 0;
            if (emit) {
                const int x = L.perm[y];
                const int kl = L.tk[x];
                kk = L.act[a0 + kl];
                const v3 apt = mk(L.tx[x], L.ty[x], L.tz[x]);
                if (L.tq[x] == 2) {
                    st = apt; en = apt; theta = -two_pi;
                } else {
                    segm = true;
                    const int s0 = L.seg[kl], s1 = L.seg[kl + 1];
                    int yn = y + 1 < s1 ? y + 1 : s0;
                    while (L.rm[yn]) yn = yn + 1 < s1 ? yn + 1 : s0;
                    const int bx = L.perm[yn];
                    st = mk(L.tx[bx], L.ty[bx], L.tz[bx]);
                    en = apt;
                    theta = -wrap_2pi_positive(L.ta[bx] - L.ta[x]);
                    oth = L.tm[x];
                }
            }
            const unsigned long long em = __ballot(emit), es = __ballot(emit && segm);
            if (na + popc(em) > MAXA || nv + popc(es) > MAXV) { overflow = true; break; }
            if (emit) {
                const int A = na + popc(em & lt);
                double* Ad = wd + WD_ARC + 7 * A;
                Ad[0] = st.x; Ad[1] = st.y; Ad[2] = st.z;
                Ad[3] = en.x; Ad[4] = en.y; Ad[5] = en.z;
                Ad[6] = theta;
                wi[WI_ARC + 2 * A] = kk;
                wi[WI_ARC + 2 * A + 1] = segm ? 0 : 1;
                if (segm) {
                    const int V = nv + popc(es & lt);
                    double* Vd = wd + WD_VTX + 3 * V;
                    Vd[0] = en.x; Vd[1] = en.y; Vd[2] = en.z;
                    wi[WI_VTX + 2 * V] = kk;
                    wi[WI_VTX + 2 * V + 1] = oth;
                }
            }
            na += popc(em);
            nv += popc(es);
        }
        wave_sync();
        a0 = a1;
    }
    if (overflow) {
        if (lane == 0) sasa_capacity_exceeded(d, ia);
        return empty;
    }
    // FullyBuried and the other arc-less outcomes contribute no gradient
    if (na == 0) return empty;

    // ---------------- pool allocation ----------------
    // issued here; the offsets are first needed after the vertex signs
    int abase = 0, vbase = 0, obase = 0;
    if (lane == 0) {
        abase = atomicAdd(d.counters + 0, na);
        vbase = atomicAdd(d.counters + 1, nv);
        obase = atomicAdd(d.counters + 2, nc);
    }
    if (AREA) atom_area(L, d, ia, ri, nc, na, lane, wd, wi);

    // ---------------- vertex root signs ----------------
    // signs and the vertex caps stay in LDS (free chunk arrays) for the links
    double* vsg = L.ta;                                               // [nv] (MAXV = SA_TMAX)
    unsigned char* vin = L.tk;
    unsigned char* vout = L.tm;
    for (int V = lane; V < nv; V += 64) {
        const int cin = wi[WI_VTX + 2 * V], cout = wi[WI_VTX + 2 * V + 1];
        const double* Vd = wd + WD_VTX + 3 * V;
        const v3 ref = mk(Vd[0], Vd[1], Vd[2]);
        v3 vp, vm;
        double sg = 1.0;
        if (vertex_point(cap_at(L, cin), cap_at(L, cout), 1.0, vp) &&
            vertex_point(cap_at(L, cin), cap_at(L, cout), -1.0, vm))
            sg = norm2(vp - ref) <= norm2(vm - ref) ? 1.0 : -1.0;
        vsg[V] = sg;
        vin[V] = (unsigned char)cin;
        vout[V] = (unsigned char)cout;
    }
    abase = __shfl(abase, 0, 64);
    vbase = __shfl(vbase, 0, 64);
    obase = __shfl(obase, 0, 64);
    if (abase + na > d.arc_cap || vbase + nv > d.vtx_cap || obase + nc > d.own_cap) {
        if (lane == 0) sasa_capacity_exceeded(d, ia);
        return empty;
    }
    for (int c = lane; c < nc; c += 64) d.owners[obase + c] = L.own[c];
    for (int V = lane; V < nv; V += 64) {
        const int cin = vin[V], cout = vout[V];
        d.vtx_atoms[vbase + V] = make_int4(ia, L.own[cin], L.own[cout], 0);
        d.vtx_sign[vbase + V] = vsg[V];
        d.vtx_code[vbase + V] = 0xFF0000u | ((unsigned)(1 + cout) << 8) | (unsigned)(1 + cin);
    }
    wave_sync();

    // ---------------- arc -> vertex links, arc records ----------------
    for (int base = 0; base < na; base += 64) {
        const int A = base + lane;
        const bool valid = A < na;
        int kc = 0, full = 1;
        v3 st = mk(0, 0, 0), en = mk(0, 0, 0);
        double theta = 0.0;
        if (valid) {
            kc = wi[WI_ARC + 2 * A];
            full = wi[WI_ARC + 2 * A + 1];
            const double* Ad = wd + WD_ARC + 7 * A;
            st = mk(Ad[0], Ad[1], Ad[2]);
            en = mk(Ad[3], Ad[4], Ad[5]);
            theta = Ad[6];
        }
        int sv = -1, ev = -1;
        for (int vb = 0; vb < nv; vb += 64) {
            v3 pv = mk(0, 0, 0);
            if (vb + lane < nv) {
                const double* Vd = wd + WD_VTX + 3 * (vb + lane);
                pv = mk(Vd[0], Vd[1], Vd[2]);
            }
            const int lim = nv - vb < 64 ? nv - vb : 64;
            for (int j = 0; j < lim; ++j) {
                const v3 q = mk(rld(pv.x, j), rld(pv.y, j), rld(pv.z, j));
                if (sv < 0 && norm2(q - st) < 1e-10) sv = vb + j;
                if (ev < 0 && norm2(q - en) < 1e-10) ev = vb + j;
            }
        }
        if (valid) {
            int4 at = make_int4(ia, L.own[kc], -1, -1);
            double ssg = 1.0, esg = 1.0;
            unsigned code = 0xFFFF00u | (unsigned)(1 + kc);
            if (!full && sv >= 0 && ev >= 0) {
                const int s_in = vin[sv], s_out = vout[sv];
                const int e_in = vin[ev], e_out = vout[ev];
                const double svs = vsg[sv], evs = vsg[ev];
                int lc, mc;
                if (s_in == kc) { lc = s_out; ssg = svs; } else { lc = s_in; ssg = -svs; }
                if (e_in == kc) { mc = e_out; esg = evs; } else { mc = e_in; esg = -evs; }
                at.z = L.own[lc];
                at.w = L.own[mc];
                code = ((unsigned)(mc == lc ? 0xFF : 1 + mc) << 16) | ((unsigned)(1 + lc) << 8) |
                       (unsigned)(1 + kc);
            }
            d.arc_atoms[abase + A] = at;
            d.arc_par[abase + A] = make_double4(theta, ssg, esg, 0.0);
            d.arc_code[abase + A] = code;
        }
    }
    return make_int4(abase, vbase, obase, na | (nv << 9) | (nc << 18));
}

}  // namespace

// Main pass (SECOND = false): every SASA atom; with NC < MAXC the atoms with
// more caps are listed for the second pass (SECOND, NC = MAXC), which runs
// over that list.
template <bool AREA, int NC, bool SECOND>
__global__ void __launch_bounds__(64 * SA_WAVES) __attribute__((amdgpu_waves_per_eu(3)))
og_k_sasa_arrange(og_sasa_data d, const double* pos) {
    __shared__ ArrLds<NC> lds_all[SA_WAVES];
    const int lane = threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    ArrLds<NC>& L = lds_all[wib];
    const int slot = blockIdx.x * SA_WAVES + wib;
    const int n_slots = gridDim.x * SA_WAVES;
    double* wd = d.scratch + (size_t)slot * WD_TOTAL;
    int* wi = d.iscratch + (size_t)slot * WI_TOTAL;
    const unsigned long long lt = lanemask_lt(lane);
    static_assert(!SECOND || NC == MAXC, "the second pass holds MAXC caps");
    if (SECOND) {
        // the atoms the main pass deferred (usually few), split statically
        const int n_def = __builtin_amdgcn_readfirstlane(d.counters[3]);
        for (int it = slot; it < n_def; it += n_slots) {
            const int w = __builtin_amdgcn_readfirstlane(d.deferred[it]);
            const int4 rec = arrange_atom<AREA, NC>(L, d, pos, w, lane, lt, wd, wi);
            if (lane == 0) d.rec[w] = rec;
            wave_sync();
        }
        return;
    }
    // Work queues, one per XCD (blocks go to the XCDs in turn, so block b
    // runs on XCD b mod 8): queue g hands out the g-th eighth of the atom
    // list one atom at a time, which evens out the several-fold spread in
    // cost per atom and keeps each XCD's L2 on a contiguous part of the
    // batch (neighbouring atoms share their neighbours).  A wave whose
    // queue is spent takes atoms from the next ones in turn and leaves once
    // it has found all eight spent.  The counters sit in 256-byte lines of
    // their own.  One call site keeps a single copy of the arrangement code.
    const int n_items = d.n_work;
    const int g0 = blockIdx.x & (SA_QUEUES - 1);
    int g = g0, tried = 1;
    auto fetch = [&](int grp) {
        int q = 0;
        if (lane == 0) q = atomicAdd(d.counters + SA_QSTRIDE * (grp + 1), 1);
        return __builtin_amdgcn_readfirstlane(__shfl(q, 0, 64));
    };
    int q = fetch(g);
    for (;;) {
        const int b = (int)((long long)g * n_items / SA_QUEUES);
        const int e = (int)((long long)(g + 1) * n_items / SA_QUEUES);
        if (b + q >= e) {
            if (tried == SA_QUEUES) break;
            ++tried;
            g = (g + 1) & (SA_QUEUES - 1);
            q = fetch(g);
            continue;
        }
        const int w = b + q;
        const int4 rec = arrange_atom<AREA, NC>(L, d, pos, w, lane, lt, wd, wi);
        if (lane == 0) d.rec[w] = rec;
        wave_sync();
        q = fetch(g);
    }
}

__global__ void __launch_bounds__(64 * SA_WAVES) og_k_sasa_scatter(og_sasa_data d, long long* facc,
                                                                   double gamma_scale) {
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * SA_WAVES + (threadIdx.x >> 6);
    if (w >= d.n_work) return;
    const int4 rec = d.rec[w];
    const int na = rec.w & 511, nv = (rec.w >> 9) & 511, nc = rec.w >> 18;
    if (na == 0) return;
    const int ia = d.work[w];
    const double gam = d.gamma[ia] * gamma_scale;
    const double ri = d.radius[ia];
    const double r_sq = ri * ri;
    // The role codes of the first 64 arcs and vertices, one per lane, read
    // together up front (later ones are read where they are used).
    const unsigned acode = lane < na ? d.arc_code[rec.x + lane] : 0u;
    const unsigned vcode = lane < nv ? d.vtx_code[rec.y + lane] : 0u;
    auto arc_code = [&](int A) -> unsigned {
        return A < 64 ? (unsigned)__builtin_amdgcn_readlane((int)acode, A) : d.arc_code[rec.x + A];
    };
    auto vtx_code = [&](int V) -> unsigned {
        return V < 64 ? (unsigned)__builtin_amdgcn_readlane((int)vcode, V) : d.vtx_code[rec.y + V];
    };
    // Every lane reads the current arc's (vertex's) block of terms, at its
    // own role or at role 0 when atom x has none there (one 96- or 72-byte
    // block per step, no divergent loads), and adds the value or +0.0.  The
    // sums run in arc (vertex) order; a +0.0 can only turn a -0.0 sum into
    // +0.0, which the zero test below treats alike.
    for (int x = lane; x <= nc; x += 64) {
        double sa0 = 0.0, sa1 = 0.0, sa2 = 0.0, sv0 = 0.0, sv1 = 0.0, sv2 = 0.0;
        bool any = false;
        for (int A = 0; A < na; ++A) {
            const unsigned xr = arc_code(A);
            const int role = x == 0 ? 0 : x == (int)(xr & 0xFF) ? 1 : x == (int)((xr >> 8) & 0xFF) ? 2
                           : x == (int)((xr >> 16) & 0xFF) ? 3 : -1;
            const bool hit = role >= 0;
            const double* T = d.arc_terms + 12ll * (rec.x + A) + 3 * (hit ? role : 0);
            const double t0 = T[0], t1 = T[1], t2 = T[2];
            sa0 += hit ? t0 : 0.0;
            sa1 += hit ? t1 : 0.0;
            sa2 += hit ? t2 : 0.0;
            any = any || hit;
        }
        for (int V = 0; V < nv; ++V) {
            const unsigned xr = vtx_code(V);
            const int role = x == 0 ? 0 : x == (int)(xr & 0xFF) ? 1 : x == (int)((xr >> 8) & 0xFF) ? 2 : -1;
            const bool hit = role >= 0;
            const double* T = d.vtx_terms + 9ll * (rec.y + V) + 3 * (hit ? role : 0);
            const double t0 = T[0], t1 = T[1], t2 = T[2];
            sv0 += hit ? t0 : 0.0;
            sv1 += hit ? t1 : 0.0;
            sv2 += hit ? t2 : 0.0;
            any = any || hit;
        }
        if (!any) continue;
        const int atom = x == 0 ? ia : d.owners[rec.z + x - 1];
        const double g0 = (-sa0 - sv0) * r_sq, g1 = (-sa1 - sv1) * r_sq, g2 = (-sa2 - sv2) * r_sq;
        const double f0 = -(g0 * gam), f1 = -(g1 * gam), f2 = -(g2 * gam);
        int* fl = d.sflag + d.atom_sys[ia];
        if (f0 != 0.0) og_atomic_add_fix(facc + 3 * (size_t)atom + 0, og_to_fix_chk(f0, fl));
        if (f1 != 0.0) og_atomic_add_fix(facc + 3 * (size_t)atom + 1, og_to_fix_chk(f1, fl));
        if (f2 != 0.0) og_atomic_add_fix(facc + 3 * (size_t)atom + 2, og_to_fix_chk(f2, fl));
    }
}

// Per-tile bounding spheres (center = first atom, radius = max distance).
// counters != null: block 0 also clears the SASA launch's counters (the
// pool fill and deferred count, and the work queues; one launch fewer at
// the head of the SASA chain).
__global__ void og_k_tile_bounds(const double* pos, const int* tile_atom0, const int* tile_cnt,
                                 int n_tiles, double4* out, int* counters) {
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    if (counters && t == 0 && lane < 8) counters[lane] = 0;
    if (counters && t == 0 && lane < SA_QUEUES) counters[SA_QSTRIDE * (lane + 1)] = 0;
    if (t >= n_tiles) return;
    const int a0 = tile_atom0[t];
    const double cx = pos[3 * (size_t)a0], cy = pos[3 * (size_t)a0 + 1], cz = pos[3 * (size_t)a0 + 2];
    double r2 = 0.0;
    if (lane < tile_cnt[t]) {
        const size_t j = 3 * (size_t)(a0 + lane);
        const double dx = pos[j] - cx, dy = pos[j + 1] - cy, dz = pos[j + 2] - cz;
        r2 = dx * dx + dy * dy + dz * dz;
    }
    for (int off = 32; off >= 1; off >>= 1) r2 = fmax(r2, __shfl_xor(r2, off, 64));
    if (lane == 0) out[t] = make_double4(cx, cy, cz, sqrt(r2) * (1.0 + 1e-12) + 1e-9);
}

int og_sasa_build(og_sasa_data* d, int n_atoms, const std::vector<int>& atom_sys,
                  const std::vector<int>& atom_off, const std::vector<double>& radius,
                  const std::vector<double>& gamma, const std::vector<int>& sys_tile_off,
                  const int* tile_atom0_dev, const int* tile_cnt_dev, int n_tiles) {
    (void)atom_off;
    std::vector<int> work, tb, te;
    double rmax = 0.0;
    for (int i = 0; i < n_atoms; ++i) rmax = std::max(rmax, radius[i]);
    for (int i = 0; i < n_atoms; ++i) {
        if (gamma[i] == 0.0) continue;
        work.push_back(i);
        tb.push_back(sys_tile_off[atom_sys[i]]);
        te.push_back(sys_tile_off[atom_sys[i] + 1]);
    }
    d->n_work = (int)work.size();
    d->max_radius = rmax;
    d->n_tiles = n_tiles;
    d->tile_atom0 = tile_atom0_dev;
    d->tile_cnt = tile_cnt_dev;
    int dev = 0;
    d->n_cu = 256;
    d->lds_cu = 64 * 1024;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) {
            d->n_cu = prop.multiProcessorCount;
            if (prop.maxSharedMemoryPerMultiProcessor > 0) d->lds_cu = (long long)prop.maxSharedMemoryPerMultiProcessor;
        }
    }
    int blocks = (d->n_work + SA_WAVES - 1) / SA_WAVES;
    blocks = std::min(blocks, d->n_cu * 8);   // staging for up to 8 blocks per CU (autotuner range)
    d->blocks = std::max(blocks, 1);
    d->stride = d->blocks * SA_WAVES;
    og_sasa_set_blocks(d, d->blocks);
    // pool capacity: OG_SASA_POOL_ARCS arcs and vertices per SASA atom on average
    const long long per = OG_SASA_POOL_ARCS;
    d->arc_cap = (int)std::min<long long>(per * std::max(d->n_work, 1), 1ll << 30);
    d->vtx_cap = d->arc_cap;
    d->own_cap = (int)std::min<long long>((long long)MAXC * std::max(d->n_work, 1), 1ll << 30);

    auto fail = [](const char* m) { g_sasa_err = m; return -2; };
    auto up = [](auto** dst, const auto& v) -> bool {
        using T = typename std::remove_reference<decltype(v)>::type::value_type;
        size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
        if (hipMalloc((void**)dst, bytes) != hipSuccess) return false;
        if (!v.empty() && hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return false;
        return true;
    };
    auto zalloc = [](auto** dst, size_t count) -> bool {
        using T = typename std::remove_reference<decltype(**dst)>::type;
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if (hipMalloc((void**)dst, bytes) != hipSuccess) return false;
        return hipMemset(*dst, 0, bytes) == hipSuccess;
    };
    if (!up(&d->work, work) || !up(&d->tile_begin, tb) || !up(&d->tile_end, te) ||
        !up(&d->radius, radius) || !up(&d->gamma, gamma))
        return fail("device allocation failed");
    // area probes (count_accessible_components, arrangement.rs:327-389): golden-spiral
    // points and, per probe, the probes within 2.5 spacings
    {
        std::vector<double4> pr(OG_SASA_PROBES);
        const double golden = M_PI * (3.0 - std::sqrt(5.0));
        for (int i = 0; i < OG_SASA_PROBES; ++i) {
            const double z = 1.0 - (2.0 * i + 1.0) / (double)OG_SASA_PROBES;
            const double t = 1.0 - z * z;
            const double r = std::sqrt(t > 0.0 ? t : 0.0);
            const double th = golden * i;
            pr[i] = make_double4(r * std::cos(th), r * std::sin(th), z, 0.0);
        }
        const double spacing = std::sqrt(4.0 * M_PI / OG_SASA_PROBES);
        double thresh = 2.5 * spacing;
        if (thresh > M_PI) thresh = M_PI;
        const double cos_thresh = std::cos(thresh);
        std::vector<int> st{0}, adj;
        for (int i = 0; i < OG_SASA_PROBES; ++i) {
            for (int j = 0; j < OG_SASA_PROBES; ++j) {
                if (j == i) continue;
                const double dt = pr[i].x * pr[j].x + pr[i].y * pr[j].y + pr[i].z * pr[j].z;
                if (dt > cos_thresh) adj.push_back(j);
            }
            st.push_back((int)adj.size());
        }
        if (!up(&d->probes, pr) || !up(&d->padj_start, st) || !up(&d->padj, adj))
            return fail("device allocation failed");
    }
    if (!zalloc(&d->tile_bound, (size_t)n_tiles) || !zalloc(&d->overflow, 1) || !zalloc(&d->counters, SA_NCOUNTERS) ||
        !zalloc(&d->rec, (size_t)d->n_work) || !zalloc(&d->deferred, (size_t)d->n_work))
        return fail("device allocation failed");
    if (d->n_work > 0) {
        if (!zalloc(&d->scratch, (size_t)WD_TOTAL * d->stride) ||
            !zalloc(&d->iscratch, (size_t)WI_TOTAL * d->stride) ||
            !zalloc(&d->arc_atoms, (size_t)d->arc_cap) || !zalloc(&d->arc_par, (size_t)d->arc_cap) ||
            !zalloc(&d->arc_code, (size_t)d->arc_cap) || !zalloc(&d->arc_terms, 12 * (size_t)d->arc_cap) ||
            !zalloc(&d->vtx_atoms, (size_t)d->vtx_cap) || !zalloc(&d->vtx_sign, (size_t)d->vtx_cap) ||
            !zalloc(&d->vtx_code, (size_t)d->vtx_cap) || !zalloc(&d->vtx_terms, 9 * (size_t)d->vtx_cap) ||
            !zalloc(&d->owners, (size_t)d->own_cap))
            return fail("SASA pool allocation failed");
    }
    return 0;
}

void og_sasa_free(og_sasa_data* d) {
    void* p[] = {d->work, d->tile_begin, d->tile_end, d->radius, d->gamma, d->tile_bound,
                 d->scratch, d->iscratch, d->overflow, d->counters, d->rec, d->deferred,
                 d->arc_atoms, d->arc_par, d->arc_code, d->arc_terms,
                 d->vtx_atoms, d->vtx_sign, d->vtx_code, d->vtx_terms, d->owners,
                 (void*)d->probes, (void*)d->padj_start, (void*)d->padj};
    for (void* q : p)
        if (q) hipFree(q);
    *d = og_sasa_data{};
}

// ---------------------------------------------------------------------------
// Shrake-Rupley dot-density SASA energy (energy/src/sasa.rs:113-178): the
// swap energy of origami's REMD with SASA (remd.rs:529-539).  One wave per
// SASA atom (gamma != 0; the others add nothing to the energy): the atoms
// whose expanded spheres reach it (distance <= R_i + R_j, from the bounding
// spheres' near tiles, ballot-compacted into LDS), then lane k tests dots
// k, k + 64, ...: a dot is buried when it lies strictly inside a
// neighbour's expanded sphere.  Area = accessible / n_dots * 4 pi R^2, in
// origami's evaluation order and without contraction, so the dot counts are
// the reference arithmetic's.
// ---------------------------------------------------------------------------
#define OG_SR_MAX_NB 256
__global__ void __launch_bounds__(256) og_k_sasa_sr(og_sasa_data d, const double* pos, const double* dots,
                                                    int n_dots) {
#pragma clang fp contract(off)
    __shared__ double4 s_nb[4][OG_SR_MAX_NB];
    double4* nb = s_nb[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int n_waves = (gridDim.x * blockDim.x) >> 6;
    for (int w = wave; w < d.n_work; w += n_waves) {
        const int i = d.work[w];
        const int sys = d.atom_sys[i];
        const double ri = d.radius[i];
        const double xi = pos[3 * (size_t)i], yi = pos[3 * (size_t)i + 1], zi = pos[3 * (size_t)i + 2];
        int n = 0;
        bool over = false;
        for (int t = d.tile_begin[w]; t < d.tile_end[w]; ++t) {
            const double4 b = d.tile_bound[t];
            const double bx = b.x - xi, by = b.y - yi, bz = b.z - zi;
            const double reach = ri + d.max_radius + b.w;
            if (bx * bx + by * by + bz * bz > reach * reach) continue;
            const int j = d.tile_atom0[t] + lane;
            bool ok = lane < d.tile_cnt[t] && j != i;
            double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
            if (ok) {
                q.x = pos[3 * (size_t)j];
                q.y = pos[3 * (size_t)j + 1];
                q.z = pos[3 * (size_t)j + 2];
                const double rj = d.radius[j];
                const double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
                ok = sqrt(dx * dx + dy * dy + dz * dz) <= ri + rj;
                q.w = rj * rj;
            }
            const unsigned long long m = __ballot(ok);
            const int slot = n + __popcll(m & ((1ull << lane) - 1ull));
            if (ok && slot < OG_SR_MAX_NB) nb[slot] = q;
            n += __popcll(m);
        }
        if (n > OG_SR_MAX_NB) {
            over = true;
            n = OG_SR_MAX_NB;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        int acc = 0;
        for (int k = lane; k < n_dots; k += 64) {
            const double tx = xi + dots[3 * k] * ri, ty = yi + dots[3 * k + 1] * ri, tz = zi + dots[3 * k + 2] * ri;
            bool buried = false;
            for (int q = 0; q < n && !buried; ++q) {
                const double4 c = nb[q];
                const double ex = tx - c.x, ey = ty - c.y, ez = tz - c.z;
                buried = ex * ex + ey * ey + ez * ez < c.w;
            }
            acc += buried ? 0 : 1;
        }
        for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) {
            if (over) {
                atomicAdd(d.overflow, 1);
                atomicOr(d.sflag + sys, OG_ST_SASA);
            } else {
                const double area = (double)acc / (double)n_dots * 4.0 * 3.14159265358979323846 * ri * ri;
                og_atomic_add_fix(d.eacc + (size_t)sys * OG_NE + OG_E_SASA,
                                  og_to_fix_chk(d.gamma[i] * area, d.sflag + sys));
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

int og_sasa_sr_launch(og_sasa_data* d, const double* pos, const double* dots, int n_dots, hipStream_t s,
                      long long* eacc, const int* atom_sys, int* sflag) {
    if (d->n_work == 0) return 0;
    if (!eacc || !atom_sys || !sflag || !dots || n_dots < 2) {
        g_sasa_err = "Shrake-Rupley: eacc, atom_sys, sflag and at least two dots are required";
        return -1;
    }
    d->eacc = eacc;
    d->atom_sys = atom_sys;
    d->sflag = sflag;
    hipLaunchKernelGGL(og_k_tile_bounds, dim3(d->n_tiles > 0 ? d->n_tiles : 1), dim3(64), 0, s, pos,
                       d->tile_atom0, d->tile_cnt, d->n_tiles, d->tile_bound, nullptr);
    int blocks = (d->n_work + 3) / 4;
    if (blocks > 8 * d->n_cu) blocks = 8 * d->n_cu;
    hipLaunchKernelGGL(og_k_sasa_sr, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, s, *d, pos, dots, n_dots);
    if (hipGetLastError() != hipSuccess) {
        g_sasa_err = "kernel launch failed";
        return -2;
    }
    return 0;
}

int og_sasa_launch(og_sasa_data* d, const double* pos, long long* facc, double gamma_scale,
                   hipStream_t s, long long* eacc, const int* atom_sys, int* sflag) {
    const bool area = eacc != nullptr;
    if (d->n_work == 0 || (gamma_scale == 0.0 && !area)) return 0;
    if (!atom_sys || !sflag) {
        g_sasa_err = "atom_sys and sflag are required";
        return -1;
    }
    d->eacc = eacc;
    d->atom_sys = atom_sys;
    d->sflag = sflag;
    hipLaunchKernelGGL(og_k_tile_bounds, dim3(d->n_tiles > 0 ? d->n_tiles : 1), dim3(64), 0, s, pos,
                       d->tile_atom0, d->tile_cnt, d->n_tiles, d->tile_bound, d->counters);
    // One pass with MAXC caps when the grid fits its LDS; otherwise the
    // SA_NC_MAIN pass and the second pass over the atoms it deferred (usually
    // few or none), one block per CU.
    const int blocks2 = d->blocks < d->n_cu ? d->blocks : d->n_cu;
    const dim3 blk(64 * SA_WAVES);
    if (d->one_pass) {
        if (area) hipLaunchKernelGGL((og_k_sasa_arrange<true, MAXC, false>), dim3(d->blocks), blk, 0, s, *d, pos);
        else hipLaunchKernelGGL((og_k_sasa_arrange<false, MAXC, false>), dim3(d->blocks), blk, 0, s, *d, pos);
    } else if (area) {
        hipLaunchKernelGGL((og_k_sasa_arrange<true, SA_NC_MAIN, false>), dim3(d->blocks), blk, 0, s, *d, pos);
        hipLaunchKernelGGL((og_k_sasa_arrange<true, MAXC, true>), dim3(blocks2), blk, 0, s, *d, pos);
    } else {
        hipLaunchKernelGGL((og_k_sasa_arrange<false, SA_NC_MAIN, false>), dim3(d->blocks), blk, 0, s, *d, pos);
        hipLaunchKernelGGL((og_k_sasa_arrange<false, MAXC, true>), dim3(blocks2), blk, 0, s, *d, pos);
    }
    if (gamma_scale != 0.0) {                 // forces (not only areas)
        og_launch_sasa_terms(*d, pos, s);
        hipLaunchKernelGGL(og_k_sasa_scatter, dim3((d->n_work + SA_WAVES - 1) / SA_WAVES), dim3(64 * SA_WAVES),
                           0, s, *d, facc, gamma_scale);
    }
    if (hipGetLastError() != hipSuccess) {
        g_sasa_err = "kernel launch failed";
        return -2;
    }
    return 0;
}

int og_sasa_max_blocks(const og_sasa_data* d) { return d->stride / SA_WAVES; }

void og_sasa_set_blocks(og_sasa_data* d, int blocks) {
    const int mx = d->stride / SA_WAVES;
    d->blocks = blocks < 1 ? 1 : (blocks > mx ? mx : blocks);
    const long long per_block = (long long)SA_WAVES * sizeof(ArrLds<MAXC>);
    d->one_pass = d->passes == 1 ||
                  (d->passes == 0 && (long long)d->blocks <= (long long)d->n_cu * (d->lds_cu / per_block));
}

void og_sasa_set_passes(og_sasa_data* d, int passes) {
    d->passes = passes;
    og_sasa_set_blocks(d, d->blocks);
}

int og_sasa_overflow(og_sasa_data* d, int* count) {
    if (hipMemcpy(count, d->overflow, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return -2;
    if (*count != 0 && hipMemset(d->overflow, 0, sizeof(int)) != hipSuccess) return -2;
    return 0;
}
