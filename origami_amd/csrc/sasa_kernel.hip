// Analytical hydrophobic-SASA forces (PowerSasa, Klenin et al. 2011 spherical
// Gauss-Bonnet), FP64, one wave64 per atom with nonzero surface tension.
//
// Per atom i: overlapping neighbours -> buried caps on i's unit sphere ->
// accessible-region boundary (arcs on cap circles joined at vertices) ->
// forward-mode derivatives of A_i / R_i^2 = 2*pi*chi - sum_arcs cos(a) theta
// - sum_vertices eps, scattered as F_x -= gamma_i * dA_i/dr_x.
//
// Restates origami:
//   caps / boundary      energy/src/powersasa/arrangement.rs:69-289
//   circle intersection  energy/src/powersasa/geometry.rs:37-100
//   vertex sign / JVPs   energy/src/powersasa/derivatives.rs:92-545
//   cache + gradient     energy/src/powersasa/analytical.rs:101-214, 285-421, 583-649
// Topology decisions use origami's FP64 tolerances unchanged (1e-12, 1e-9,
// 1e-10) and no FMA contraction (sasa_core.h), which is why this stays FP64.
//
// Parallel structure (wave = one atom, lanes cooperate):
//   caps      lanes over the 64 atoms of each candidate tile (tiles culled by
//             bounding spheres), ballot-compacted in ascending atom order;
//   boundary  caps k in order; for each, lanes over the other caps m compute
//             the transitions (push order m, then p before q), a bitonic sort
//             on (angle, push index) reproduces the stable sort, duplicates
//             are dropped against the last kept point, lanes over segments do
//             the midpoint test, and ballot compaction emits arcs/vertices in
//             origami's (k, segment) order.  A cap whose whole circle lies
//             inside another cap by a margin of 1e-9 (>> the 1e-12 test
//             tolerance) emits nothing in origami and is skipped up front;
//   gradient  lanes over (arc|vertex, role, axis) items evaluate one JVP term
//             each; lanes over (atom, axis) then sum their terms in arc order
//             and vertex order, which is the order origami sums them in
//             (terms of arcs/vertices not involving an atom are exact zeros).
// Per-wave LDS holds the caps and the current cap's transitions; arcs,
// vertices and gradient terms live in a per-wave global scratch slot.
#include <string>

#include "og_device.h"
#include "og_sasa.h"
#include "sasa_core.h"

#define SW_WAVES 2                 // waves per block
#define SW_MAXA MAXA
#define SW_MAXV MAXV

static std::string g_sasa_err;
const char* og_sasa_error() { return g_sasa_err.c_str(); }

using namespace og_sasa;

namespace {

struct WaveLds {
    double cx[MAXC], cy[MAXC], cz[MAXC], cc[MAXC];   // caps: axis, cos(alpha)
    int own[MAXC];                                   // cap owner (global atom)
    double ta[MAXT], tx[MAXT], ty[MAXT], tz[MAXT];   // transitions of cap k (push order)
    int to[MAXT];                                    // other cap of each transition
    unsigned char perm[256];                         // sorted order
    unsigned char kept[MAXT];                        // kept (deduplicated) transitions, sorted
    unsigned int xarc[SW_MAXA];                      // arc roles k, l, m as local atom bytes
    unsigned int xvtx[SW_MAXV];                      // vertex roles in, out as local atom bytes
    int scalar;
};

// per-wave global scratch slot layout (doubles then ints)
#define SD_ARC 0                                     // MAXA x 7 (sx,sy,sz,ex,ey,ez,theta)
#define SD_VTX (SD_ARC + 7 * SW_MAXA)                // MAXV x 4 (px,py,pz,sign)
#define SD_TA (SD_VTX + 4 * SW_MAXV)                 // MAXA x 12 arc terms (role*3+axis)
#define SD_TV (SD_TA + 12 * SW_MAXA)                 // MAXV x 9 vertex terms
#define SD_TOTAL (SD_TV + 9 * SW_MAXV)
#define SI_ARC 0                                     // MAXA x 4 (cap, full, sv, ev)
#define SI_VTX (SI_ARC + 4 * SW_MAXA)                // MAXV x 2 (in, out)
#define SI_TOTAL (SI_VTX + 2 * SW_MAXV)

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ unsigned long long lanemask_lt(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ int popc(unsigned long long m) { return __popcll(m); }

__device__ __forceinline__ Cap cap_at(const WaveLds& L, int k) {
    return {mk(L.cx[k], L.cy[k], L.cz[k]), L.cc[k]};
}

// Sort key comparison (angle, push index): strict weak order that matches a
// stable sort by angle.
__device__ __forceinline__ bool key_less(const WaveLds& L, int a, int b, int M) {
    if (b >= M) return a < M;          // padding sorts last
    if (a >= M) return false;
    const double x = L.ta[a], y = L.ta[b];
    return x < y || (x == y && a < b);
}

}  // namespace

__global__ void __launch_bounds__(64 * SW_WAVES) og_k_sasa_wave(og_sasa_data d, const double* pos,
                                                                long long* facc, double gamma_scale) {
    __shared__ WaveLds lds_all[SW_WAVES];
    const int lane = threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    WaveLds& L = lds_all[wib];
    const int slot = blockIdx.x * SW_WAVES + wib;
    const int n_slots = gridDim.x * SW_WAVES;
    double* sd = d.scratch + (size_t)slot * SD_TOTAL;
    int* si = d.iscratch + (size_t)slot * SI_TOTAL;
    const unsigned long long lt = lanemask_lt(lane);
    const double two_pi = 2.0 * 3.14159265358979323846;

    for (int w = slot; w < d.n_work; w += n_slots) {
        const int ia = d.work[w];
        const double gam = d.gamma[ia] * gamma_scale;
        if (gam == 0.0) continue;
        const v3 pi = ldp(pos, ia);
        const double ri = d.radius[ia];

        // ---------------- caps (build_atom_cache + build_caps) ----------------
        // Candidates in ascending atom order; an enclosing neighbour ends the
        // scan (no accessible surface) unless the cap count overflowed first.
        int nc = 0;
        bool enclosed = false, overflow = false;
        for (int t = d.tile_begin[w]; t < d.tile_end[w]; ++t) {
            const double4 tb = d.tile_bound[t];
            const double dxc = tb.x - pi.x, dyc = tb.y - pi.y, dzc = tb.z - pi.z;
            const double reach = ri + d.max_radius + tb.w;
            if (dxc * dxc + dyc * dyc + dzc * dzc > reach * reach) continue;
            const int j = d.tile_atom0[t] + lane;
            const bool valid = lane < d.tile_cnt[t] && j != ia;
            bool is_cap = false, enc = false;
            v3 axis = mk(0, 0, 0);
            double ca = 0.0;
            if (valid) {
                const v3 pj = ldp(pos, j);
                const double rj = d.radius[j];
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
            if (nc + popc(m) > MAXC) { overflow = true; break; }
            if (is_cap && ((m >> lane) & 1ull)) {
                const int p = nc + popc(m & lt);
                L.cx[p] = axis.x; L.cy[p] = axis.y; L.cz[p] = axis.z; L.cc[p] = ca;
                L.own[p] = j;
            }
            nc += popc(m);
            if (em) { enclosed = true; break; }
        }
        wave_sync();
        if (overflow) {
            if (lane == 0) atomicAdd(d.overflow, 1);
            continue;
        }
        if (enclosed || nc == 0) continue;

        // ---------------- boundary (find_boundary) ----------------
        int na = 0, nv = 0;
        for (int k = 0; k < nc && !overflow; ++k) {
            const Cap ck = cap_at(L, k);
            // transitions of circle k, in origami's push order (cap m ascending, p then q)
            int M = 0;
            for (int base = 0; base < nc; base += 64) {
                const int m = base + lane;
                int cnt = 0;
                v3 p = mk(0, 0, 0), q = mk(0, 0, 0);
                if (m < nc && m != k) {
                    const int r = intersect_circles(ck, cap_at(L, m), p, q);
                    cnt = r == 0 ? 2 : (r == 1 ? 1 : 0);
                }
                const unsigned long long b1 = __ballot(cnt >= 1), b2 = __ballot(cnt == 2);
                const int off = M + popc(b1 & lt) + popc(b2 & lt);
                if (cnt >= 1) {
                    L.ta[off] = angular_position(p, ck);
                    L.tx[off] = p.x; L.ty[off] = p.y; L.tz[off] = p.z; L.to[off] = m;
                }
                if (cnt == 2) {
                    L.ta[off + 1] = angular_position(q, ck);
                    L.tx[off + 1] = q.x; L.ty[off + 1] = q.y; L.tz[off + 1] = q.z; L.to[off + 1] = m;
                }
                M += popc(b1) + popc(b2);
            }
            wave_sync();
            // Exact early-out: every point origami tests on circle k (segment
            // midpoints are rotations of transition points about axis_k, which
            // keep their height h and radius rn; the probe is on the circle)
            // satisfies dot(x, axis_m) >= h*ct - rn*st with ct, st the cosine
            // and sine between the axes.  When that bound clears cap m's
            // containment test for all of them, circle k contributes nothing.
            {
                double hmin = ck.ca, hmax = ck.ca;
                double rmx = sqrt(fmax(1.0 - ck.ca * ck.ca, 0.0));
                for (int x = lane; x < M; x += 64) {
                    const v3 p = mk(L.tx[x], L.ty[x], L.tz[x]);
                    const double h = dot(p, ck.axis);
                    const double rn = norm(p - ck.axis * h);
                    hmin = fmin(hmin, h);
                    hmax = fmax(hmax, h);
                    rmx = fmax(rmx, rn);
                }
                for (int o = 32; o >= 1; o >>= 1) {
                    hmin = fmin(hmin, __shfl_xor(hmin, o, 64));
                    hmax = fmax(hmax, __shfl_xor(hmax, o, 64));
                    rmx = fmax(rmx, __shfl_xor(rmx, o, 64));
                }
                bool inside = false;
                for (int base = 0; base < nc; base += 64) {
                    const int m = base + lane;
                    bool cont = false;
                    if (m < nc && m != k) {
                        const v3 am = mk(L.cx[m], L.cy[m], L.cz[m]);
                        const double ct = dot(ck.axis, am);
                        const double st = norm(cross(ck.axis, am));
                        const double lo = fmin(hmin * ct, hmax * ct) - rmx * st;
                        cont = lo > L.cc[m] + 1e-11;
                    }
                    if (__ballot(cont)) inside = true;
                }
                if (inside) continue;
            }
            // stable sort of the M transitions by angle (bitonic on (angle, index))
            int P = 1;
            while (P < M) P <<= 1;
            for (int x = lane; x < P; x += 64) L.perm[x] = (unsigned char)x;
            wave_sync();
            for (int k2 = 2; k2 <= P; k2 <<= 1) {
                for (int j2 = k2 >> 1; j2 > 0; j2 >>= 1) {
                    for (int x = lane; x < P; x += 64) {
                        const int y = x ^ j2;
                        if (y > x) {
                            const int a = L.perm[x], b = L.perm[y];
                            const bool asc = (x & k2) == 0;
                            const bool swap = asc ? key_less(L, b, a, M) : key_less(L, a, b, M);
                            if (swap) { L.perm[x] = (unsigned char)b; L.perm[y] = (unsigned char)a; }
                        }
                    }
                    wave_sync();
                }
            }
            // Dedup against the last kept point (Vec::dedup_by).  When no
            // removal happens before the last position, "last kept" is always
            // the sorted predecessor and lanes decide independently.
            bool chain = false;
            for (int base = 0; base < M; base += 64) {
                const int x = base + lane;
                bool c = false;
                if (x >= 1 && x + 1 < M) {
                    const int a = L.perm[x], b = L.perm[x - 1];
                    const v3 pa = mk(L.tx[a], L.ty[a], L.tz[a]), pb = mk(L.tx[b], L.ty[b], L.tz[b]);
                    c = norm2(pa - pb) < 1e-12;
                }
                if (__ballot(c)) chain = true;
            }
            int nk = 0;
            if (!chain) {
                for (int base = 0; base < M; base += 64) {
                    const int x = base + lane;
                    bool keep = false;
                    if (x < M) {
                        keep = true;
                        if (x > 0) {
                            const int a = L.perm[x], b = L.perm[x - 1];
                            const v3 pa = mk(L.tx[a], L.ty[a], L.tz[a]), pb = mk(L.tx[b], L.ty[b], L.tz[b]);
                            keep = !(norm2(pa - pb) < 1e-12);
                        }
                    }
                    const unsigned long long km = __ballot(keep);
                    if (keep) L.kept[nk + popc(km & lt)] = L.perm[x];
                    nk += popc(km);
                }
            } else {
                if (lane == 0) {
                    int wk = 0;
                    for (int x = 0; x < M; ++x) {
                        const int a = L.perm[x];
                        if (wk > 0) {
                            const int b = L.kept[wk - 1];
                            const v3 pa = mk(L.tx[a], L.ty[a], L.tz[a]), pb = mk(L.tx[b], L.ty[b], L.tz[b]);
                            if (norm2(pa - pb) < 1e-12) continue;
                        }
                        L.kept[wk++] = (unsigned char)a;
                    }
                    L.scalar = wk;
                }
                wave_sync();
                nk = L.scalar;
            }
            wave_sync();
            if (nk == 0) {
                // whole circle: boundary loop unless the probe point is buried
                const v3 probe = sample_circle_point(ck);
                bool buried = false;
                for (int base = 0; base < nc; base += 64) {
                    const int o = base + lane;
                    const bool b = o < nc && o != k && contains_strict(cap_at(L, o), probe);
                    if (__ballot(b)) buried = true;
                }
                if (!buried) {
                    if (na >= SW_MAXA) { overflow = true; break; }
                    if (lane == 0) {
                        double* A = sd + SD_ARC + 7 * na;
                        A[0] = probe.x; A[1] = probe.y; A[2] = probe.z;
                        A[3] = probe.x; A[4] = probe.y; A[5] = probe.z;
                        A[6] = -two_pi;
                        int* AI = si + SI_ARC + 4 * na;
                        AI[0] = k; AI[1] = 1; AI[2] = -1; AI[3] = -1;
                    }
                    ++na;
                }
                continue;
            }
            for (int base = 0; base < nk; base += 64) {
                const int x = base + lane;
                bool emit = false;
                double ccw = 0.0;
                int ax_ = 0, by_ = 0;
                if (x < nk) {
                    ax_ = L.kept[x];
                    by_ = L.kept[(x + 1) % nk];
                    ccw = wrap_2pi_positive(L.ta[by_] - L.ta[ax_]);
                    if (!(ccw < 1e-9)) {
                        const v3 apt = mk(L.tx[ax_], L.ty[ax_], L.tz[ax_]);
                        const v3 mid = midpoint_on_circle(apt, ck, ccw / 2.0);
                        // any() over the other caps; the two caps bounding the
                        // segment are the likeliest to contain it, test them first
                        const int oa = L.to[ax_], ob = L.to[by_];
                        bool bad = contains_strict(cap_at(L, oa), mid) || contains_strict(cap_at(L, ob), mid);
                        for (int o = 0; o < nc && !bad; ++o)
                            bad = (o != k) && contains_strict(cap_at(L, o), mid);
                        emit = !bad;
                    }
                }
                const unsigned long long em = __ballot(emit);
                if (na + popc(em) > SW_MAXA || nv + popc(em) > SW_MAXV) { overflow = true; break; }
                if (emit) {
                    const int A = na + popc(em & lt);
                    double* Ad = sd + SD_ARC + 7 * A;
                    Ad[0] = L.tx[by_]; Ad[1] = L.ty[by_]; Ad[2] = L.tz[by_];
                    Ad[3] = L.tx[ax_]; Ad[4] = L.ty[ax_]; Ad[5] = L.tz[ax_];
                    Ad[6] = -ccw;
                    int* AI = si + SI_ARC + 4 * A;
                    AI[0] = k; AI[1] = 0; AI[2] = -1; AI[3] = -1;
                    const int V = nv + popc(em & lt);
                    double* Vd = sd + SD_VTX + 4 * V;
                    Vd[0] = L.tx[ax_]; Vd[1] = L.ty[ax_]; Vd[2] = L.tz[ax_]; Vd[3] = 1.0;
                    int* VI = si + SI_VTX + 2 * V;
                    VI[0] = k; VI[1] = L.to[ax_];
                }
                na += popc(em);
                nv += popc(em);
            }
            wave_sync();
        }
        if (overflow) {
            if (lane == 0) atomicAdd(d.overflow, 1);
            continue;
        }
        if (na == 0) continue;
        wave_sync();

        // ---------------- cache: vertex signs, arc links ----------------
        for (int V = lane; V < nv; V += 64) {
            const int* VI = si + SI_VTX + 2 * V;
            double* Vd = sd + SD_VTX + 4 * V;
            const Cap cin = cap_at(L, VI[0]), cout = cap_at(L, VI[1]);
            const v3 ref = mk(Vd[0], Vd[1], Vd[2]);
            v3 vp, vm;
            double sg = 1.0;
            if (vertex_point(cin, cout, 1.0, vp) && vertex_point(cin, cout, -1.0, vm))
                sg = norm2(vp - ref) <= norm2(vm - ref) ? 1.0 : -1.0;
            Vd[3] = sg;
            L.xvtx[V] = (unsigned)(0xFF0000u | ((unsigned)(1 + VI[1]) << 8) | (unsigned)(1 + VI[0]));
        }
        for (int A = lane; A < na; A += 64) {
            int* AI = si + SI_ARC + 4 * A;
            int sv = -1, ev = -1;
            if (!AI[1]) {
                const double* Ad = sd + SD_ARC + 7 * A;
                const v3 st = mk(Ad[0], Ad[1], Ad[2]), en = mk(Ad[3], Ad[4], Ad[5]);
                for (int V = 0; V < nv && (sv < 0 || ev < 0); ++V) {
                    const double* Vd = sd + SD_VTX + 4 * V;
                    const v3 pv = mk(Vd[0], Vd[1], Vd[2]);
                    if (sv < 0 && norm2(pv - st) < 1e-10) sv = V;
                    if (ev < 0 && norm2(pv - en) < 1e-10) ev = V;
                }
                AI[2] = sv;
                AI[3] = ev;
            }
        }
        wave_sync();
        // arc roles as local atom indices (0 = atom i, 1 + c = owner of cap c);
        // byte 0xFF = role absent or already covered by an earlier role
        for (int A = lane; A < na; A += 64) {
            const int* AI = si + SI_ARC + 4 * A;
            const int kc = AI[0];
            unsigned xl = 0xFF, xm = 0xFF;
            if (!(AI[1] || AI[2] < 0 || AI[3] < 0)) {
                const int* SV = si + SI_VTX + 2 * AI[2];
                const int* EV = si + SI_VTX + 2 * AI[3];
                const int lc = SV[0] == kc ? SV[1] : SV[0];
                const int mc = EV[0] == kc ? EV[1] : EV[0];
                xl = 1 + lc;
                xm = mc == lc ? 0xFF : 1 + mc;
            }
            L.xarc[A] = (xm << 16) | (xl << 8) | (unsigned)(1 + kc);
        }
        wave_sync();

        // ---------------- gradient terms, one (item, role, axis) per lane ----
        const v3 zero = mk(0, 0, 0);
        for (int it = lane; it < 12 * na + 9 * nv; it += 64) {
            if (it < 12 * na) {
                const int A = it / 12, role = (it % 12) / 3, axis = it % 3;
                const unsigned xr = L.xarc[A];
                const int x = role == 0 ? 0 : (int)((xr >> (8 * (role - 1))) & 0xFF);
                double term = 0.0;
                if (x != 0xFF) {
                    const int* AI = si + SI_ARC + 4 * A;
                    const double theta = sd[SD_ARC + 7 * A + 6];
                    const int kc = AI[0];
                    const int ka = L.own[kc];
                    const v3 pk = ldp(pos, ka);
                    const double rk = d.radius[ka];
                    const v3 b = unit(axis);
                    const v3 dri = x == 0 ? b : zero;
                    const v3 drk = x == 1 + kc ? b : zero;
                    const double dcos = cos_alpha_directional(pi, pk, ri, rk, dri, drk);
                    if ((xr >> 8 & 0xFF) == 0xFF) {
                        term = theta * dcos;
                    } else {
                        const int sv = AI[2], ev = AI[3];
                        const int* SV = si + SI_VTX + 2 * sv;
                        const int* EV = si + SI_VTX + 2 * ev;
                        const double svs = sd[SD_VTX + 4 * sv + 3], evs = sd[SD_VTX + 4 * ev + 3];
                        int lc, mc;
                        double ssg, esg;
                        if (SV[0] == kc) { lc = SV[1]; ssg = svs; } else { lc = SV[0]; ssg = -svs; }
                        if (EV[0] == kc) { mc = EV[1]; esg = evs; } else { mc = EV[0]; esg = -evs; }
                        const int la = L.own[lc], ma = L.own[mc];
                        const v3 pl = ldp(pos, la), pm = ldp(pos, ma);
                        const double rl = d.radius[la], rm = d.radius[ma];
                        const v3 drl = x == 1 + lc ? b : zero;
                        const v3 drm = x == 1 + mc ? b : zero;
                        const double dik = norm(pk - pi);
                        const double cak = (dik * dik + ri * ri - rk * rk) / (2.0 * dik * ri);
                        double dth = 0.0;
                        if (!arc_theta_jvp(pi, pk, pl, pm, ri, rk, rl, rm, ssg, esg, dri, drk, drl, drm, dth))
                            dth = 0.0;
                        term = theta * dcos + cak * dth;
                    }
                }
                sd[SD_TA + 12 * A + 3 * role + axis] = term;
            } else {
                const int jt = it - 12 * na;
                const int V = jt / 9, role = (jt % 9) / 3, axis = jt % 3;
                const int* VI = si + SI_VTX + 2 * V;
                const int kcap = VI[0], lcap = VI[1];
                const int x = role == 0 ? 0 : role == 1 ? 1 + kcap : 1 + lcap;
                const int ka = L.own[kcap], la = L.own[lcap];
                const v3 pk = ldp(pos, ka), pl = ldp(pos, la);
                const v3 b = unit(axis);
                const v3 dri = x == 0 ? b : zero;
                const v3 drk = x == 1 + kcap ? b : zero;
                const v3 drl = x == 1 + lcap ? b : zero;
                double de = 0.0;
                if (!vertex_epsilon_jvp(pi, pk, pl, ri, d.radius[ka], d.radius[la], sd[SD_VTX + 4 * V + 3],
                                        dri, drk, drl, de))
                    de = 0.0;
                sd[SD_TV + 9 * V + 3 * role + axis] = de;
            }
        }
        wave_sync();

        // ---------------- per-atom sums in origami's order, scatter ----------
        // Lane x sums, for its three axes, the terms of every arc (in arc
        // order) and then every vertex (in vertex order) that involve it.
        const double r_sq = ri * ri;
        for (int x = lane; x <= nc; x += 64) {
            double sa0 = 0.0, sa1 = 0.0, sa2 = 0.0, sv0 = 0.0, sv1 = 0.0, sv2 = 0.0;
            bool any = false;
            for (int A = 0; A < na; ++A) {
                const unsigned xr = L.xarc[A];
                int role = -1;
                if (x == 0) role = 0;
                else if (x == (int)(xr & 0xFF)) role = 1;
                else if (x == (int)(xr >> 8 & 0xFF)) role = 2;
                else if (x == (int)(xr >> 16 & 0xFF)) role = 3;
                if (role >= 0) {
                    const double* T = sd + SD_TA + 12 * A + 3 * role;
                    sa0 += T[0]; sa1 += T[1]; sa2 += T[2];
                    any = true;
                }
            }
            for (int V = 0; V < nv; ++V) {
                const unsigned xr = L.xvtx[V];
                int role = -1;
                if (x == 0) role = 0;
                else if (x == (int)(xr & 0xFF)) role = 1;
                else if (x == (int)(xr >> 8 & 0xFF)) role = 2;
                if (role >= 0) {
                    const double* T = sd + SD_TV + 9 * V + 3 * role;
                    sv0 += T[0]; sv1 += T[1]; sv2 += T[2];
                    any = true;
                }
            }
            if (!any) continue;
            const int atom = x == 0 ? ia : L.own[x - 1];
            const double s_a[3] = {sa0, sa1, sa2}, s_v[3] = {sv0, sv1, sv2};
            for (int axis = 0; axis < 3; ++axis) {
                const double g = (-s_a[axis] - s_v[axis]) * r_sq;
                const double f = -(g * gam);
                if (f != 0.0) og_atomic_add_fix(facc + 3 * (size_t)atom + axis, og_to_fix(f));
            }
        }
        wave_sync();
    }
}

// Per-tile bounding spheres (center = first atom, radius = max distance).
__global__ void og_k_tile_bounds(const double* pos, const int* tile_atom0, const int* tile_cnt,
                                 int n_tiles, double4* out) {
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
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
    int dev = 0, n_cu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            n_cu = prop.multiProcessorCount;
    }
    int blocks = (d->n_work + SW_WAVES - 1) / SW_WAVES;
    blocks = std::min(blocks, n_cu * 8);
    d->blocks = std::max(blocks, 1);
    d->stride = d->blocks * SW_WAVES;
    auto up = [](auto** dst, const auto& v) -> bool {
        using T = typename std::remove_reference<decltype(v)>::type::value_type;
        size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
        if (hipMalloc((void**)dst, bytes) != hipSuccess) return false;
        if (!v.empty() && hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return false;
        return true;
    };
    if (!up(&d->work, work) || !up(&d->tile_begin, tb) || !up(&d->tile_end, te) ||
        !up(&d->radius, radius) || !up(&d->gamma, gamma)) {
        g_sasa_err = "device allocation failed";
        return -2;
    }
    if (hipMalloc((void**)&d->tile_bound, std::max(n_tiles, 1) * sizeof(double4)) != hipSuccess) {
        g_sasa_err = "device allocation failed";
        return -2;
    }
    if (d->n_work > 0) {
        if (hipMalloc((void**)&d->scratch, (size_t)SD_TOTAL * d->stride * sizeof(double)) != hipSuccess ||
            hipMalloc((void**)&d->iscratch, (size_t)SI_TOTAL * d->stride * sizeof(int)) != hipSuccess) {
            g_sasa_err = "SASA scratch allocation failed";
            return -2;
        }
    }
    if (hipMalloc((void**)&d->overflow, sizeof(int)) != hipSuccess ||
        hipMemset(d->overflow, 0, sizeof(int)) != hipSuccess) {
        g_sasa_err = "device allocation failed";
        return -2;
    }
    return 0;
}

void og_sasa_free(og_sasa_data* d) {
    void* p[] = {d->work, d->tile_begin, d->tile_end, d->radius, d->gamma, d->tile_bound,
                 d->scratch, d->iscratch, d->overflow};
    for (void* q : p)
        if (q) hipFree(q);
    *d = og_sasa_data{};
}

int og_sasa_launch(og_sasa_data* d, const double* pos, long long* facc, double gamma_scale,
                   hipStream_t s) {
    if (d->n_work == 0) return 0;
    hipLaunchKernelGGL(og_k_tile_bounds, dim3(d->n_tiles), dim3(64), 0, s, pos, d->tile_atom0,
                       d->tile_cnt, d->n_tiles, d->tile_bound);
    hipLaunchKernelGGL(og_k_sasa_wave, dim3(d->blocks), dim3(64 * SW_WAVES), 0, s, *d, pos, facc,
                       gamma_scale);
    if (hipGetLastError() != hipSuccess) {
        g_sasa_err = "kernel launch failed";
        return -2;
    }
    return 0;
}

int og_sasa_overflow(og_sasa_data* d, int* count) {
    if (hipMemcpy(count, d->overflow, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return -2;
    return 0;
}
