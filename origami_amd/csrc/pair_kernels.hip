// All-pairs tile kernels: GB descreening integral and the fused
// LJ + Coulomb + GB-pair force (optionally with energies).
//
// Work decomposition: one wave64 per tile pair (I <= J) of one system,
// grid-stride over a flattened work list that covers every system of the
// batch.  Lane l owns i-atom I*64+l.  The j atoms travel around the wave:
// lane l starts with j-atom l and, after every pair step, each lane passes
// its j data and its running j-side reaction (Newton's third law) to the
// neighbouring lane with one DPP wave rotate (v_mov_dpp wave_rol:1), so in
// 64 steps every lane meets every j atom and the j-side sums arrive back
// home.  On off-diagonal tiles without exclusions (most of the work) the j
// data are read from the wave's LDS instead, entry (l + k) mod 64 at step k,
// which is exactly what the rotation would deliver, so only the three (one)
// j-side sums rotate.  A diagonal tile (I = J) runs steps 1..32 (step 32 on
// lanes 0..31) and swaps the halves of the wave at the end.  Per tile pair every lane
// flushes one FP32 partial per component into the int64 fixed-point
// accumulators.  The FP64 energy kernel reads its j atoms from LDS (every lane
// the same entry: a broadcast) and accumulates the pair energies per lane.
//
// Physics follows origami's SoA kernels:
//   descreening   energy/src/gb.rs:271-283 (pair), :347-369 (h)
//   LJ + Coulomb  energy/src/forces_nonbonded.rs:196-232
//   GB pair force energy/src/forces_gb.rs:72-112
//   energies      energy/src/nonbonded.rs:97-174, energy/src/gb.rs:146-182
// Mode B (all pairs) is the fast path; Mode A adds origami's integrator
// cutoffs (10 Å LJ + reaction-field Coulomb, 10 Å GB force, 20 Å descreen).
#include "og_device.h"
#include "og_kernels.h"
#include "bonded_core.h"
#include "md_core.h"

namespace {

constexpr float kGbK = (float)(1.0 - 1.0 / OG_EPS_WATER);
constexpr float kHalfLn2 = 0.34657359027997264f;

struct LaneData {
    float x, y, z, q;
    float4 par;   // se, rm, se14, rm14
    float4 gb;    // rho~, S, R, invRq
};

__device__ __forceinline__ LaneData load_lane(const og_pair_args& a, int tile, int atom0, int cnt,
                                              int lane, double ox, double oy, double oz) {
    LaneData d;
    const int slot = tile * OG_TILE + lane;
    if (lane < cnt) {
        const double* p = a.pos + 3 * (size_t)(atom0 + lane);
        d.x = (float)(p[0] - ox);
        d.y = (float)(p[1] - oy);
        d.z = (float)(p[2] - oz);
    } else {
        // Padding slot: far away and (via zero parameters) inert.
        d.x = 1.0e6f + 97.0f * (float)lane;
        d.y = 1.0e6f;
        d.z = 1.0e6f;
    }
    d.q = a.qs[slot];
    d.par = a.par[slot];
    d.gb = a.gb[slot];
    return d;
}

__device__ __forceinline__ float rl(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// one-lane rotation of a value around the wave (DPP wave_rol:1)
// (mov_dpp: every lane is written, so no "old" value has to be materialised)
__device__ __forceinline__ float rot(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ int roti(int v) { return __builtin_amdgcn_mov_dpp(v, 0x134, 0xF, 0xF, false); }
// exchange with the lane 32 away (halves of the wave)
__device__ __forceinline__ float swap32(float v) { return __shfl_xor(v, 32, 64); }

// Descreening integral term h(r, rho~_i, S_j) of energy/src/gb.rs:347-369,
// written with one reciprocal and one log:  1/L - 1/U = (U-L)/(LU),
// 1/U^2 - 1/L^2 = -(1/L - 1/U)(U+L)/(LU),  ln(L/U) = ln(L^2/(LU)), with
// L = max(r - S, rho~) covering both the overlapping and the separate case.
__device__ __forceinline__ float descreen_h(float r, float inv_r, float rt, float S, float S2) {
    const float U = r + S;
    const float L = fmaxf(r - S, rt);
    const float UmL = U - L;
    const float UpL = U + L;
    const float w = __builtin_amdgcn_rcpf(L * U);
    const float t = UmL * w;
    const float s2q = S2 * (0.25f * inv_r);
    const float lg = __builtin_amdgcn_logf(L * L * w);   // log2
    float h = fmaf(0.5f, t, -(t * UpL * w) * s2q);
    h = fmaf(lg, kHalfLn2 * inv_r, h);
    return (U <= rt) ? 0.0f : h;
}

template <bool MODEA, bool DIAG>
__device__ __forceinline__ void descreen_tile(const og_pair_args& a, const LaneData& di,
                                              const LaneData& dj, int lane, float& acc_i, float& acc_j) {
    float xj = dj.x, yj = dj.y, zj = dj.z, rtj = dj.gb.x, Sj = dj.gb.y;
    float gj = 0.0f;
    const int k0 = DIAG ? 1 : 0, k1 = DIAG ? 32 : 63;
    if (DIAG) {
        xj = rot(xj); yj = rot(yj); zj = rot(zj); rtj = rot(rtj); Sj = rot(Sj);
    }
#pragma unroll 2
    for (int k = k0; k <= k1; ++k) {
        const float dx = xj - di.x, dy = yj - di.y, dz = zj - di.z;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float r = r2 * inv_r;
        float hij = descreen_h(r, inv_r, di.gb.x, Sj, Sj * Sj);
        float hji = descreen_h(r, inv_r, rtj, di.gb.y, di.gb.y * di.gb.y);
        bool ok = r2 >= 1e-18f;
        if (MODEA) ok = ok && (r2 <= a.rc2_ds);
        if (DIAG) ok = ok && (k < 32 || lane < 32);
        acc_i += ok ? hij : 0.0f;
        gj += ok ? hji : 0.0f;
        if (k < k1) {
            xj = rot(xj); yj = rot(yj); zj = rot(zj); rtj = rot(rtj); Sj = rot(Sj);
            gj = rot(gj);
        }
    }
    // home: one more step around the wave (64 in all), or half a turn on a diagonal tile
    acc_j += DIAG ? swap32(gj) : rot(gj);
}

// Orders a wave's own LDS accesses across lanes for the compiler; a wave's
// LDS operations are performed in program order, so no wait is needed.
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Off-diagonal descreening tile with the j data in LDS: P[e] = (x, y, z, rho~)
// and Q[e] = (S, S^2, -, -) of j atom e, entries 0..63 plus entry 64 = entry
// 0, so that the pair of steps k, k + 1 reads entries idx and idx + 1 without
// a wrap (Q follows P in the same array: one address for all four reads).
// Same arithmetic, in the same order, as the rotating form.
// PART (og_work::part): -1 all 64 steps, 0 steps 0-31, 1 steps 32-63.  After
// steps 0-31 the j-side sums sit 32 lanes from home, hence the swap.
template <bool MODEA, int PART>
__device__ __forceinline__ void descreen_tile_lds(const og_pair_args& a, const LaneData& di,
                                                  const float4* P, const float4* Q, int lane,
                                                  float& acc_i, float& acc_j) {
    float gj = 0.0f;
    const float Si2 = di.gb.y * di.gb.y;
    auto step = [&](float4 pj, float4 qj) {
        const float dx = pj.x - di.x, dy = pj.y - di.y, dz = pj.z - di.z;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float r = r2 * inv_r;
        const float hij = descreen_h(r, inv_r, di.gb.x, qj.x, qj.y);
        const float hji = descreen_h(r, inv_r, pj.w, di.gb.y, Si2);
        bool ok = r2 >= 1e-18f;
        if (MODEA) ok = ok && (r2 <= a.rc2_ds);
        acc_i += ok ? hij : 0.0f;
        gj += ok ? hji : 0.0f;
        gj = rot(gj);                  // after step 63 this is the way home
    };
    constexpr int kb = PART == 1 ? 32 : 0, ke = PART == 0 ? 32 : 64;
    for (int k = kb; k < ke; k += 2) {
        const int idx = (lane + k) & 63;
        const float4 p0 = P[idx], p1 = P[idx + 1];
        const float4 q0 = Q[idx], q1 = Q[idx + 1];
        step(p0, q0);
        step(p1, q1);
    }
    if (PART == 0) gj = swap32(gj);
    acc_j += gj;
}

// Far-field form of descreen_tile_lds for tile pairs whose atoms are all at
// least og_pair_args::ds_far apart: there S / r <= 0.1 and L = r - S, and
// the HCT term (1/r)[x/(1-x^2) - x^3/(1-x^2)^2 - atanh x], x = S/r, is the
// series (1/r) sum_n c_n x^(2n+1), c_n = 1 - n - 1/(2n+1):
//   h = -(S^3 / (3 r^4)) (1 + 3.6 y + (45/7) y^2 + (28/3) y^3),  y = x^2,
// whose next term, (135/11) y^4 relative, is below 1.3e-7 at x = 0.1 (FP32
// rounding is 6e-8); the closed form evaluated in FP32 there loses about
// 6e-8 / x^2 to cancellation.  Every term is negative, as h is.  Q[e].z holds
// -S^3/3 of j atom e.
template <bool MODEA, int PART>
__device__ __forceinline__ void descreen_tile_far(const og_pair_args& a, const LaneData& di,
                                                  const float4* P, const float4* Q, int lane,
                                                  float& acc_i, float& acc_j) {
    float gj = 0.0f;
    const float Si2 = di.gb.y * di.gb.y;
    const float ci3 = (-1.0f / 3.0f) * Si2 * di.gb.y;
    auto series = [](float y) { return fmaf(y, fmaf(y, fmaf(y, 28.0f / 3.0f, 45.0f / 7.0f), 3.6f), 1.0f); };
    auto step = [&](float4 pj, float4 qj) {
        const float dx = pj.x - di.x, dy = pj.y - di.y, dz = pj.z - di.z;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float inv_r2 = inv_r * inv_r;
        const float inv_r4 = inv_r2 * inv_r2;
        const float hij = (qj.z * inv_r4) * series(qj.y * inv_r2);
        const float hji = (ci3 * inv_r4) * series(Si2 * inv_r2);
        if (MODEA) {
            const bool ok = r2 <= a.rc2_ds;
            acc_i += ok ? hij : 0.0f;
            gj += ok ? hji : 0.0f;
        } else {
            acc_i += hij;
            gj += hji;
        }
        gj = rot(gj);                  // after step 63 this is the way home
    };
    constexpr int kb = PART == 1 ? 32 : 0, ke = PART == 0 ? 32 : 64;
    for (int k = kb; k < ke; k += 2) {
        const int idx = (lane + k) & 63;
        const float4 p0 = P[idx], p1 = P[idx + 1];
        const float4 q0 = Q[idx], q1 = Q[idx + 1];
        step(p0, q0);
        step(p1, q1);
    }
    if (PART == 0) gj = swap32(gj);
    acc_j += gj;
}

__device__ __forceinline__ float wave_max_f(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// One descreening work item (a tile pair, or half of one) of the wave; P, Q:
// the wave's LDS staging.
template <bool MODEA>
__device__ __forceinline__ void descreen_item(const og_pair_args& a, const og_work& it, int lane, float4* P,
                                              float4* Q) {
    {
        const int a0i = a.tile_atom0[it.ti], ci = a.tile_cnt[it.ti];
        const int a0j = a.tile_atom0[it.tj], cj = a.tile_cnt[it.tj];
        const double ox = a.pos[3 * (size_t)a0i], oy = a.pos[3 * (size_t)a0i + 1],
                     oz = a.pos[3 * (size_t)a0i + 2];
        const LaneData di = load_lane(a, it.ti, a0i, ci, lane, ox, oy, oz);
        float acc_i = 0.0f, acc_j = 0.0f;
        if (it.ti == it.tj) {
            descreen_tile<MODEA, true>(a, di, di, lane, acc_i, acc_j);
        } else {
            const LaneData dj = load_lane(a, it.tj, a0j, cj, lane, ox, oy, oz);
            const float4 pj = make_float4(dj.x, dj.y, dj.z, dj.gb.x);
            const float4 qj = make_float4(dj.gb.y, dj.gb.y * dj.gb.y, (-1.0f / 3.0f) * (dj.gb.y * dj.gb.y) * dj.gb.y,
                                          0.0f);
            wave_lds_order();                       // the previous tile's reads are done
            P[lane] = pj;
            Q[lane] = qj;
            if (lane == 0) { P[OG_TILE] = pj; Q[OG_TILE] = qj; }
            wave_lds_order();
            // far tile pair: the distance of the two first atoms less the
            // radii of both tiles about them (positions are relative to atom
            // a0i, so lane 0's i position is the origin) bounds every pair's
            // distance from below; 1e-3 A covers the FP32 rounding
            const float jx0 = rl(dj.x, 0), jy0 = rl(dj.y, 0), jz0 = rl(dj.z, 0);
            const float ri = lane < ci ? sqrtf(fmaf(di.x, di.x, fmaf(di.y, di.y, di.z * di.z))) : 0.0f;
            const float ex = dj.x - jx0, ey = dj.y - jy0, ez = dj.z - jz0;
            const float rj = lane < cj ? sqrtf(fmaf(ex, ex, fmaf(ey, ey, ez * ez))) : 0.0f;
            const float d0 = sqrtf(fmaf(jx0, jx0, fmaf(jy0, jy0, jz0 * jz0)));
            const bool far = a.ds_far > 0.0f && d0 - wave_max_f(ri) - wave_max_f(rj) - 1e-3f >= a.ds_far;
            if (far) {
                if (it.part < 0) descreen_tile_far<MODEA, -1>(a, di, P, Q, lane, acc_i, acc_j);
                else if (it.part == 0) descreen_tile_far<MODEA, 0>(a, di, P, Q, lane, acc_i, acc_j);
                else descreen_tile_far<MODEA, 1>(a, di, P, Q, lane, acc_i, acc_j);
            } else if (it.part < 0) descreen_tile_lds<MODEA, -1>(a, di, P, Q, lane, acc_i, acc_j);
            else if (it.part == 0) descreen_tile_lds<MODEA, 0>(a, di, P, Q, lane, acc_i, acc_j);
            else descreen_tile_lds<MODEA, 1>(a, di, P, Q, lane, acc_i, acc_j);
        }
        int* fl = a.sflag ? a.sflag + it.sys : nullptr;
        if (lane < ci) og_atomic_add_fix(a.iacc + a0i + lane, og_to_fix_chk((double)acc_i, fl));
        if (lane < cj) og_atomic_add_fix(a.iacc + a0j + lane, og_to_fix_chk((double)acc_j, fl));
    }
}

// blk / nblk: this block among the nblk blocks sharing the tile pairs.
template <bool MODEA>
__device__ __forceinline__ void descreen_body(const og_pair_args& a, int blk, int nblk) {
    __shared__ float4 s_pq[4][2 * (OG_TILE + 1)];       // per wave: P[65], then Q[65]
    float4* P = s_pq[threadIdx.x >> 6];
    float4* Q = P + (OG_TILE + 1);
    const int lane = threadIdx.x & 63;
    const int wave = (blk * blockDim.x + threadIdx.x) >> 6;
    const int n_waves = (nblk * blockDim.x) >> 6;
    for (int w = wave; w < a.n_work; w += n_waves) {
        const og_work it = a.work[w];
        if (a.skip_sys && a.skip_sys[it.sys]) continue;     // radii from the surrogate
        descreen_item<MODEA>(a, it, lane, P, Q);
    }
}

}  // namespace

__global__ void __launch_bounds__(256) og_k_descreen_b(og_pair_args a) {
    descreen_body<false>(a, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) og_k_descreen_a(og_pair_args a) {
    descreen_body<true>(a, blockIdx.x, gridDim.x);
}

// The descreening pass and the bonded terms (forces) in one launch, for small
// batches whose force evaluation is a chain of latency-bound launches: blocks
// [0, nd) take the tile pairs, the rest the bonded terms, side by side.
template <bool MODEA>
__global__ void __launch_bounds__(256) og_k_descreen_bonded(og_pair_args a, og_bonded_args b, int nd) {
    if ((int)blockIdx.x < nd) descreen_body<MODEA>(a, blockIdx.x, nd);
    else bonded_body<false>(b, blockIdx.x - nd, gridDim.x - nd);
}

namespace {

// ---------------------------------------------------------------------------
// Fused force tile.
// ---------------------------------------------------------------------------
struct ForceAcc {
    float fx, fy, fz;       // i-side
    float gx, gy, gz;       // j-side (lane j)
};

// PART (og_work::part, off-diagonal tiles): -1 all 64 steps, 0 steps 0-31,
// 1 steps 32-63 (the j data start half a turn round); after steps 0-31 the
// j-side sums sit 32 lanes from home, hence the swap.
template <bool MODEA, bool MASKED, bool DIAG, int PART>
__device__ __forceinline__ void force_tile(const og_pair_args& a, const LaneData& di,
                                           const LaneData& dj, int lane,
                                           unsigned long long xmask, unsigned long long m14,
                                           ForceAcc& acc) {
    float xj = dj.x, yj = dj.y, zj = dj.z, qj = dj.q;
    float sej = dj.par.x, rmj = dj.par.y, se14j = dj.par.z, rm14j = dj.par.w;
    float Rj = dj.gb.z, iRj = dj.gb.w;
    int jid = lane;                         // j index within tile J held by this lane
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    const float qki = kGbK * di.q;
    constexpr int k0 = DIAG ? 1 : (PART == 1 ? 32 : 0), k1 = DIAG ? 32 : (PART == 0 ? 31 : 63);
    auto advance = [&]() {
        xj = rot(xj); yj = rot(yj); zj = rot(zj); qj = rot(qj);
        sej = rot(sej); rmj = rot(rmj); Rj = rot(Rj); iRj = rot(iRj);
        if (MASKED) { se14j = rot(se14j); rm14j = rot(rm14j); jid = roti(jid); }
    };
    if (DIAG) advance();
    if (!DIAG && PART == 1) {
        xj = swap32(xj); yj = swap32(yj); zj = swap32(zj); qj = swap32(qj);
        sej = swap32(sej); rmj = swap32(rmj); Rj = swap32(Rj); iRj = swap32(iRj);
        if (MASKED) { se14j = swap32(se14j); rm14j = swap32(rm14j); jid = lane ^ 32; }
    }
#pragma unroll 2
    for (int k = k0; k <= k1; ++k) {
        const float dx = xj - di.x, dy = yj - di.y, dz = zj - di.z;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float inv_r2 = inv_r * inv_r;
        float rmin, eij;
        bool excl = false;
        if (MASKED) {
            excl = (xmask >> jid) & 1ull;
            const bool one4 = (m14 >> jid) & 1ull;
            rmin = one4 ? di.par.w + rm14j : di.par.y + rmj;
            eij = one4 ? di.par.z * se14j : di.par.x * sej;
        } else {
            rmin = di.par.y + rmj;
            eij = di.par.x * sej;
        }
        const float t2 = rmin * rmin * inv_r2;
        const float t6 = t2 * t2 * t2;
        const float lj = eij * inv_r2 * (t6 - t6 * t6);
        const float qq = di.q * qj;
        float coul;
        if (MODEA) coul = -qq * (inv_r2 * inv_r - a.inv_rc3);
        else coul = -qq * inv_r2 * inv_r;
        float cnb = lj + coul;
        if (MODEA) cnb = (r2 <= a.rc2_nb) ? cnb : 0.0f;
        if (MASKED) cnb = excl ? 0.0f : cnb;
        // GB pair (frozen radii): f^2 = r^2 + R_i R_j exp(-r^2 / (4 R_i R_j))
        const float Rp = di.gb.z * Rj;
        const float e = __builtin_amdgcn_exp2f(-r2 * (di.gb.w * iRj));
        const float f2 = fmaf(Rp, e, r2);
        const float inv_f = __builtin_amdgcn_rsqf(f2);
        const float qgb = qki * qj;
        float gbc = qgb * fmaf(-0.25f, e, 1.0f) * (inv_f * inv_f * inv_f);
        if (MODEA) gbc = (r2 <= a.rc2_gb) ? gbc : 0.0f;
        float c = cnb + gbc;
        bool ok = r2 >= 1e-18f;
        if (DIAG) ok = ok && (k < 32 || lane < 32);
        c = ok ? c : 0.0f;
        acc.fx = fmaf(c, dx, acc.fx);
        acc.fy = fmaf(c, dy, acc.fy);
        acc.fz = fmaf(c, dz, acc.fz);
        gx = fmaf(-c, dx, gx);
        gy = fmaf(-c, dy, gy);
        gz = fmaf(-c, dz, gz);
        if (k < k1) {
            advance();
            gx = rot(gx); gy = rot(gy); gz = rot(gz);
        }
    }
    // home: one more step around the wave (64 in all), or half a turn on a diagonal tile
    if (DIAG) {
        gx = swap32(gx); gy = swap32(gy); gz = swap32(gz);
    } else {
        gx = rot(gx); gy = rot(gy); gz = rot(gz);
        if (PART == 0) { gx = swap32(gx); gy = swap32(gy); gz = swap32(gz); }
    }
    acc.gx += gx;
    acc.gy += gy;
    acc.gz += gz;
}

// Off-diagonal force tile without exclusions, j data in LDS: P[e] = (x, y,
// z, q) and Q[e] = (se, rm, R, 1/R scaled) of j atom e, entries 0..63 plus
// entry 64 = entry 0 (steps k and k + 1 read entries idx and idx + 1; Q
// follows P in the same array).  The arithmetic is that of force_tile's
// unmasked path.
template <bool MODEA, int PART>
__device__ __forceinline__ void force_tile_lds(const og_pair_args& a, const LaneData& di,
                                               const float4* P, const float4* Q, int lane,
                                               ForceAcc& acc) {
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    const float qki = kGbK * di.q;
    auto step = [&](float4 pj, float4 qj4) {
        const float dx = pj.x - di.x, dy = pj.y - di.y, dz = pj.z - di.z;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float inv_r2 = inv_r * inv_r;
        const float rmin = di.par.y + qj4.y;
        const float eij = di.par.x * qj4.x;
        const float t2 = rmin * rmin * inv_r2;
        const float t6 = t2 * t2 * t2;
        const float lj = eij * inv_r2 * (t6 - t6 * t6);
        const float qq = di.q * pj.w;
        float coul;
        if (MODEA) coul = -qq * (inv_r2 * inv_r - a.inv_rc3);
        else coul = -qq * inv_r2 * inv_r;
        float cnb = lj + coul;
        if (MODEA) cnb = (r2 <= a.rc2_nb) ? cnb : 0.0f;
        const float Rp = di.gb.z * qj4.z;
        const float e = __builtin_amdgcn_exp2f(-r2 * (di.gb.w * qj4.w));
        const float f2 = fmaf(Rp, e, r2);
        const float inv_f = __builtin_amdgcn_rsqf(f2);
        const float qgb = qki * pj.w;
        float gbc = qgb * fmaf(-0.25f, e, 1.0f) * (inv_f * inv_f * inv_f);
        if (MODEA) gbc = (r2 <= a.rc2_gb) ? gbc : 0.0f;
        float c = cnb + gbc;
        c = r2 >= 1e-18f ? c : 0.0f;
        acc.fx = fmaf(c, dx, acc.fx);
        acc.fy = fmaf(c, dy, acc.fy);
        acc.fz = fmaf(c, dz, acc.fz);
        gx = rot(fmaf(-c, dx, gx));     // after step 63 this is the way home
        gy = rot(fmaf(-c, dy, gy));
        gz = rot(fmaf(-c, dz, gz));
    };
    constexpr int kb = PART == 1 ? 32 : 0, ke = PART == 0 ? 32 : 64;
    for (int k = kb; k < ke; k += 2) {
        const int idx = (lane + k) & 63;
        const float4 p0 = P[idx], p1 = P[idx + 1];
        const float4 q0 = Q[idx], q1 = Q[idx + 1];
        step(p0, q0);
        step(p1, q1);
    }
    if (PART == 0) { gx = swap32(gx); gy = swap32(gy); gz = swap32(gz); }
    acc.gx += gx;
    acc.gy += gy;
    acc.gz += gz;
}

// One i atom of a strip: position relative to the strip's origin, charge,
// LJ and GB parameters, and its force sum over the strip.
struct StripAtom {
    float x, y, z, q;
    float se, rm, R, gw;
    float fx, fy, fz;
};

__device__ __forceinline__ StripAtom strip_atom(const LaneData& d) {
    return {d.x, d.y, d.z, d.q, d.par.x, d.par.y, d.gb.z, d.gb.w, 0.0f, 0.0f, 0.0f};
}

// Pair coefficient c (force on i = c * d, d = x_j - x_i) of the fused
// LJ + Coulomb + GB-pair force for an unmasked pair; j data as in the LDS
// layout of force_tile_lds (P = x, y, z, q; Q = se, rm, R, 1/R scaled).
// Mode B groups the two charge terms: q_i q_j (k_GB (1 - e/4) / f^3 - 1/r^3).
// FAR: every pair of the j tile has r^2 gw_i gw_j >= 32 (gw = sqrt(log2 e / 4)
// / R), so e = exp2(-r^2 gw_i gw_j) <= 2^-32.  Then R_i R_j e is below 1e-11
// r^2 (r^2 = 4 R_i R_j x, x = r^2 gw_i gw_j / log2 e >= 22), far under half an
// FP32 ulp of r^2, so f^2 = fma(R_i R_j, e, r^2) rounds to r^2 and 1/f is the
// 1/r above; likewise k (1 - e/4) rounds to k.  The exponential and the
// second square root are left out, with the same products in the same order:
// the coefficient is bit-identical.
template <bool MODEA, bool FAR = false>
__device__ __forceinline__ float strip_coef(const og_pair_args& a, const StripAtom& I, float4 pj, float4 qj,
                                            float dx, float dy, float dz) {
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float inv_r2 = inv_r * inv_r;
    const float rmin = I.rm + qj.y;
    const float t2 = rmin * rmin * inv_r2;
    const float t6 = t2 * t2 * t2;
    const float lj = (I.se * qj.x) * inv_r2 * fmaf(-t6, t6, t6);
    const float qq = I.q * pj.w;
    float gb;
    if (FAR) {
        gb = kGbK * (inv_r * inv_r * inv_r);
    } else {
        const float e = __builtin_amdgcn_exp2f(-r2 * (I.gw * qj.w));
        const float f2 = fmaf(I.R * qj.z, e, r2);
        const float inv_f = __builtin_amdgcn_rsqf(f2);
        gb = fmaf(-0.25f * kGbK, e, kGbK) * (inv_f * inv_f * inv_f);
    }
    if (MODEA) {
        float cnb = fmaf(-qq, inv_r2 * inv_r - a.inv_rc3, lj);
        cnb = (r2 <= a.rc2_nb) ? cnb : 0.0f;
        const float gbc = (r2 <= a.rc2_gb) ? qq * gb : 0.0f;
        return cnb + gbc;
    }
    return fmaf(qq, gb - inv_r2 * inv_r, lj);
}

// A strip: two i tiles (64 i atoms each, two per lane) against up to OG_STRIP_J j
// tiles of the same system, j data in the wave's LDS as in force_tile_lds.
// Each step evaluates the lane's two i atoms against one j atom, adds both
// reactions into the rotating j-side sums and rotates them once, so the
// rotation and the LDS read serve two pairs.  The i-side sums stay in
// registers over the whole strip and are flushed once.  No pair of a strip
// is excluded or 1-4, and no i atom is a padding slot (the i tiles precede a
// j tile of their system, so both are full); distinct atoms at one position
// give a non-finite coefficient, which the fixed-point conversion reports.
template <bool MODEA>
__device__ __forceinline__ void force_strip(const og_pair_args& a, const og_strip* sp, int lane,
                                            float4* P, float4* Q) {
    // fields read through the pointer (the j tile list is indexed at run
    // time, which a local copy would keep in scratch memory)
    const int ta = __builtin_amdgcn_readfirstlane(sp->ti), tb = ta + 1;
    const int nj = __builtin_amdgcn_readfirstlane(sp->nj);
    const int a0 = a.tile_atom0[ta], b0 = a.tile_atom0[tb];
    const int ca = a.tile_cnt[ta], cb = a.tile_cnt[tb];
    const double ox = a.pos[3 * (size_t)a0], oy = a.pos[3 * (size_t)a0 + 1], oz = a.pos[3 * (size_t)a0 + 2];
    StripAtom A = strip_atom(load_lane(a, ta, a0, ca, lane, ox, oy, oz));
    StripAtom B = strip_atom(load_lane(a, tb, b0, cb, lane, ox, oy, oz));
    int* fl = a.sflag ? a.sflag + sp->sys : nullptr;
    // bounds of the two i tiles about the origin (atom a0; both tiles are
    // full) and their smallest gw, for the far test of each j tile
    const float rad_i = wave_max_f(fmaxf(sqrtf(fmaf(A.x, A.x, fmaf(A.y, A.y, A.z * A.z))),
                                         sqrtf(fmaf(B.x, B.x, fmaf(B.y, B.y, B.z * B.z)))));
    const float gw_i = -wave_max_f(-fminf(A.gw, B.gw));
    int tj = __builtin_amdgcn_readfirstlane(sp->tj[0]);
    LaneData dj = load_lane(a, tj, a.tile_atom0[tj], a.tile_cnt[tj], lane, ox, oy, oz);
    for (int q = 0; q < nj; ++q) {
        const int a0j = a.tile_atom0[tj], cj = a.tile_cnt[tj];
        const float4 pjl = make_float4(dj.x, dj.y, dj.z, dj.q);
        const float4 qjl = make_float4(dj.par.x, dj.par.y, dj.gb.z, dj.gb.w);
        wave_lds_order();                       // the previous j tile's reads are done
        P[lane] = pjl;
        Q[lane] = qjl;
        if (lane == 0) { P[OG_TILE] = pjl; Q[OG_TILE] = qjl; }
        wave_lds_order();
        // far j tile: the smallest pair distance (distance of the first atoms
        // less both tiles' radii, 1e-3 A for FP32 rounding) gives
        // r^2 gw_i gw_j >= 33 for every pair (padding slots excluded)
        bool far;
        {
            const float jx0 = rl(dj.x, 0), jy0 = rl(dj.y, 0), jz0 = rl(dj.z, 0);
            const float ex = dj.x - jx0, ey = dj.y - jy0, ez = dj.z - jz0;
            const float rad_j = wave_max_f(lane < cj ? sqrtf(fmaf(ex, ex, fmaf(ey, ey, ez * ez))) : 0.0f);
            const float gw_j = -wave_max_f(lane < cj ? -dj.gb.w : -1.0e30f);
            const float rmn = sqrtf(fmaf(jx0, jx0, fmaf(jy0, jy0, jz0 * jz0))) - rad_i - rad_j - 1e-3f;
            far = rmn > 0.0f && rmn * rmn * (gw_i * gw_j) >= 33.0f;
        }
        // the next j tile's data load while this one is evaluated
        const int tn = __builtin_amdgcn_readfirstlane(sp->tj[q + 1 < nj ? q + 1 : q]);
        dj = load_lane(a, tn, a.tile_atom0[tn], a.tile_cnt[tn], lane, ox, oy, oz);
        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
        auto step = [&](float4 pj, float4 qj, auto far_tag) {
            constexpr bool F = decltype(far_tag)::value;
            const float dxa = pj.x - A.x, dya = pj.y - A.y, dza = pj.z - A.z;
            const float dxb = pj.x - B.x, dyb = pj.y - B.y, dzb = pj.z - B.z;
            const float c_a = strip_coef<MODEA, F>(a, A, pj, qj, dxa, dya, dza);
            const float c_b = strip_coef<MODEA, F>(a, B, pj, qj, dxb, dyb, dzb);
            A.fx = fmaf(c_a, dxa, A.fx);
            A.fy = fmaf(c_a, dya, A.fy);
            A.fz = fmaf(c_a, dza, A.fz);
            B.fx = fmaf(c_b, dxb, B.fx);
            B.fy = fmaf(c_b, dyb, B.fy);
            B.fz = fmaf(c_b, dzb, B.fz);
            gx = rot(fmaf(-c_b, dxb, fmaf(-c_a, dxa, gx)));     // after step 63: home
            gy = rot(fmaf(-c_b, dyb, fmaf(-c_a, dya, gy)));
            gz = rot(fmaf(-c_b, dzb, fmaf(-c_a, dza, gz)));
        };
        if (far) {
            for (int k = 0; k < OG_TILE; k += 2) {
                const int idx = (lane + k) & 63;
                const float4 p0 = P[idx], q0 = Q[idx];
                const float4 p1 = P[idx + 1], q1 = Q[idx + 1];
                step(p0, q0, std::true_type{});
                step(p1, q1, std::true_type{});
            }
        } else {
            for (int k = 0; k < OG_TILE; k += 2) {
                const int idx = (lane + k) & 63;
                const float4 p0 = P[idx], q0 = Q[idx];
                const float4 p1 = P[idx + 1], q1 = Q[idx + 1];
                step(p0, q0, std::false_type{});
                step(p1, q1, std::false_type{});
            }
        }
        if (lane < cj) {
            long long* f = a.facc + 3 * (size_t)(a0j + lane);
            og_atomic_add_fix(f + 0, og_to_fix_chk((double)gx, fl));
            og_atomic_add_fix(f + 1, og_to_fix_chk((double)gy, fl));
            og_atomic_add_fix(f + 2, og_to_fix_chk((double)gz, fl));
        }
        tj = tn;
    }
    if (lane < ca) {
        long long* f = a.facc + 3 * (size_t)(a0 + lane);
        og_atomic_add_fix(f + 0, og_to_fix_chk((double)A.fx, fl));
        og_atomic_add_fix(f + 1, og_to_fix_chk((double)A.fy, fl));
        og_atomic_add_fix(f + 2, og_to_fix_chk((double)A.fz, fl));
    }
    if (lane < cb) {
        long long* f = a.facc + 3 * (size_t)(b0 + lane);
        og_atomic_add_fix(f + 0, og_to_fix_chk((double)B.fx, fl));
        og_atomic_add_fix(f + 1, og_to_fix_chk((double)B.fy, fl));
        og_atomic_add_fix(f + 2, og_to_fix_chk((double)B.fz, fl));
    }
}

// One force work item outside the strips (a diagonal, adjacent or masked
// tile pair, or a tile pair of a small system).
template <bool MODEA>
__device__ __forceinline__ void force_item(const og_pair_args& a, const og_work& it, int lane, float4* P,
                                           float4* Q) {
    {
        const int a0i = a.tile_atom0[it.ti], ci = a.tile_cnt[it.ti];
        const int a0j = a.tile_atom0[it.tj], cj = a.tile_cnt[it.tj];
        const double ox = a.pos[3 * (size_t)a0i], oy = a.pos[3 * (size_t)a0i + 1],
                     oz = a.pos[3 * (size_t)a0i + 2];
        const LaneData di = load_lane(a, it.ti, a0i, ci, lane, ox, oy, oz);
        ForceAcc acc = {0, 0, 0, 0, 0, 0};
        unsigned long long xm = 0, m14 = 0;
        if (it.mask >= 0) {
            xm = a.masks[(size_t)it.mask * 128 + lane];
            m14 = a.masks[(size_t)it.mask * 128 + 64 + lane];
        }
        if (it.ti == it.tj) {
            force_tile<MODEA, true, true, -1>(a, di, di, lane, xm, m14, acc);
        } else {
            const LaneData dj = load_lane(a, it.tj, a0j, cj, lane, ox, oy, oz);
            if (it.mask >= 0) {
                if (it.part < 0) force_tile<MODEA, true, false, -1>(a, di, dj, lane, xm, m14, acc);
                else if (it.part == 0) force_tile<MODEA, true, false, 0>(a, di, dj, lane, xm, m14, acc);
                else force_tile<MODEA, true, false, 1>(a, di, dj, lane, xm, m14, acc);
            } else {
                const float4 pj = make_float4(dj.x, dj.y, dj.z, dj.q);
                const float4 qj = make_float4(dj.par.x, dj.par.y, dj.gb.z, dj.gb.w);
                wave_lds_order();                   // the previous tile's reads are done
                P[lane] = pj;
                Q[lane] = qj;
                if (lane == 0) { P[OG_TILE] = pj; Q[OG_TILE] = qj; }
                wave_lds_order();
                if (it.part < 0) force_tile_lds<MODEA, -1>(a, di, P, Q, lane, acc);
                else if (it.part == 0) force_tile_lds<MODEA, 0>(a, di, P, Q, lane, acc);
                else force_tile_lds<MODEA, 1>(a, di, P, Q, lane, acc);
            }
        }
        int* fl = a.sflag ? a.sflag + it.sys : nullptr;
        if (lane < ci) {
            long long* f = a.facc + 3 * (size_t)(a0i + lane);
            og_atomic_add_fix(f + 0, og_to_fix_chk((double)acc.fx, fl));
            og_atomic_add_fix(f + 1, og_to_fix_chk((double)acc.fy, fl));
            og_atomic_add_fix(f + 2, og_to_fix_chk((double)acc.fz, fl));
        }
        if (lane < cj) {
            long long* f = a.facc + 3 * (size_t)(a0j + lane);
            og_atomic_add_fix(f + 0, og_to_fix_chk((double)acc.gx, fl));
            og_atomic_add_fix(f + 1, og_to_fix_chk((double)acc.gy, fl));
            og_atomic_add_fix(f + 2, og_to_fix_chk((double)acc.gz, fl));
        }
    }
}

// The force pass: the strips first (the heavier items), then the remaining
// tile pairs (fwork), grid-stride over both.
template <bool MODEA>
__device__ __forceinline__ void force_body(const og_pair_args& a) {
    __shared__ float4 s_pq[4][2 * (OG_TILE + 1)];       // per wave: P[65], then Q[65]
    float4* P = s_pq[threadIdx.x >> 6];
    float4* Q = P + (OG_TILE + 1);
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int n_waves = (gridDim.x * blockDim.x) >> 6;
    const int total = a.n_strip + a.n_fwork;
    for (int w = wave; w < total; w += n_waves) {
        if (w < a.n_strip) {
            force_strip<MODEA>(a, a.strips + w, lane, P, Q);
            continue;
        }
        const og_work it = a.fwork[w - a.n_strip];
        force_item<MODEA>(a, it, lane, P, Q);
    }
}

}  // namespace

__global__ void __launch_bounds__(256) og_k_force_b(og_pair_args a) { force_body<false>(a); }
__global__ void __launch_bounds__(256) og_k_force_a(og_pair_args a) { force_body<true>(a); }

static int pair_grid(int n_work) {
    int blocks = (n_work + 3) / 4;
    const int cap = 256 * 16;
    if (blocks > cap) blocks = cap;
    return blocks < 1 ? 1 : blocks;
}

void og_launch_pair_descreen(const og_pair_args& a, bool mode_a, hipStream_t s) {
    if (a.n_work <= 0) return;
    const int g = a.grid > 0 ? a.grid : pair_grid(a.n_work);
    if (mode_a) hipLaunchKernelGGL(og_k_descreen_a, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(og_k_descreen_b, dim3(g), dim3(256), 0, s, a);
}

void og_launch_descreen_bonded(const og_pair_args& a, const og_bonded_args& b, bool mode_a, hipStream_t s) {
    const int nd = a.n_work > 0 ? (a.grid > 0 ? a.grid : pair_grid(a.n_work)) : 0;
    const int total = b.nb + b.na + b.nd + b.ni;
    int nb = (total + 255) / 256;
    if (nb > 4096) nb = 4096;
    if (nd + nb <= 0) return;
    if (mode_a) hipLaunchKernelGGL(og_k_descreen_bonded<true>, dim3(nd + nb), dim3(256), 0, s, a, b, nd);
    else hipLaunchKernelGGL(og_k_descreen_bonded<false>, dim3(nd + nb), dim3(256), 0, s, a, b, nd);
}

void og_launch_pair_force(const og_pair_args& a, bool mode_a, hipStream_t s) {
    const int n = a.n_strip + a.n_fwork;
    if (n <= 0) return;
    // the force pass has fewer, larger work items than descreening (strips):
    // it runs one wave per item (measured fastest at every width tried:
    // 0.299 ms on the 157k-atom batch against 0.304-0.387 ms at 4096 to 512
    // blocks), whatever width the descreening pass takes
    const int g = (n + 3) / 4;
    if (mode_a) hipLaunchKernelGGL(og_k_force_a, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(og_k_force_b, dim3(g), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// FP64 pair energies (LJ, Coulomb, GB pair), used for energy totals, REMD
// exchange and the L-BFGS line search.  Same tile schedule; all arithmetic in
// double, evaluated as origami writes it:
//   LJ       eps_ij ((Rmin/r)^12 - 2 (Rmin/r)^6)         nonbonded.rs:410-415
//   Coulomb  k q_i q_j (1/r + kappa r^2 - c)  (Mode B: k q_i q_j / r)  nonbonded.rs:327-329
//   GB pair  2 (-1/2)(1 - 1/eps_w) k q_i q_j / f_GB, uncut  gb.rs:160-173
// ---------------------------------------------------------------------------
namespace {

// 1/sqrt(x) for x > 0 in FP64: the hardware estimate refined by two Newton
// steps (each doubles the number of correct bits), in place of a correctly
// rounded square root and division.
__device__ __forceinline__ double rsqrt64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = x * y;
    double e = fma(-h, y, 1.0);
    y = fma(0.5 * y, e, y);
    h = x * y;
    e = fma(-h, y, 1.0);
    return fma(0.5 * y, e, y);
}

// j atom data for the FP64 energy pass, one entry per j atom of the tile in
// the wave's LDS: every lane reads the same entry in a step (a broadcast).
struct EJ {
    double x, y, z, q;
    double rm, se, rm14, se14;     // Rmin/2, sqrt(eps), 1-4 variants
    double R, iR;                  // Born radius and its reciprocal
};

template <bool MODEA>
__device__ void energy_body(const og_pair_args& a, const og_energy64_args& e) {
    __shared__ EJ s_j[4][OG_TILE];
    EJ* J = s_j[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int n_waves = (gridDim.x * blockDim.x) >> 6;
    const double kc = OG_COULOMB_KCAL * OG_KCAL_TO_KJ;
    const double pref = -(1.0 - 1.0 / OG_EPS_WATER) * OG_COULOMB_KCAL * OG_KCAL_TO_KJ;
    for (int w = wave; w < a.n_work; w += n_waves) {
        const og_work it = a.work[w];
        if (it.part == 1) continue;                 // a split tile pair is evaluated whole once
        if (a.skip_sys && a.skip_sys[it.sys]) continue;
        const int a0i = a.tile_atom0[it.ti], ci = a.tile_cnt[it.ti];
        const int a0j = a.tile_atom0[it.tj], cj = a.tile_cnt[it.tj];
        const bool vi = lane < ci, vj = lane < cj;
        const int gi = a0i + (vi ? lane : 0), gj = a0j + (vj ? lane : 0);
        const double xi = a.pos[3 * (size_t)gi], yi = a.pos[3 * (size_t)gi + 1], zi = a.pos[3 * (size_t)gi + 2];
        const double qi = e.charge[gi];
        const double rmi = e.rmin[gi], sei = sqrt(e.eps[gi]);
        const double rm14i = e.rmin14[gi], se14i = sqrt(e.eps14[gi]);
        const double Ri = e.radii[gi], iRi = 1.0 / Ri;
        unsigned long long xm = 0, m14 = 0;
        if (it.mask >= 0) {
            xm = a.masks[(size_t)it.mask * 128 + lane];
            m14 = a.masks[(size_t)it.mask * 128 + 64 + lane];
        }
        EJ mine;
        mine.x = a.pos[3 * (size_t)gj]; mine.y = a.pos[3 * (size_t)gj + 1]; mine.z = a.pos[3 * (size_t)gj + 2];
        mine.q = e.charge[gj];
        mine.rm = e.rmin[gj]; mine.se = sqrt(e.eps[gj]);
        mine.rm14 = e.rmin14[gj]; mine.se14 = sqrt(e.eps14[gj]);
        mine.R = e.radii[gj]; mine.iR = 1.0 / mine.R;
        wave_lds_order();                           // the previous tile's reads are done
        J[lane] = mine;
        wave_lds_order();
        const bool diag = it.ti == it.tj;
        double elj = 0.0, ec = 0.0, eg = 0.0;
        if (!diag && it.mask < 0) {
            // off-diagonal tile pair without excluded or 1-4 pairs: the same
            // terms with the mask and 1-4 selects left out
            const double xo = vi ? xi : xi + 1.0e6;      // padding lanes: far away, added as 0
            for (int jj = 0; jj < cj; ++jj) {
                const EJ d = J[jj];
                const double dx = xo - d.x, dy = yi - d.y, dz = zi - d.z;
                const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                const double qq = qi * d.q;
                const double inv_r = rsqrt64(r2);
                const double t = (rmi + d.rm) * inv_r;
                const double t2 = t * t;
                const double t6 = t2 * t2 * t2;
                const double lj = (sei * d.se) * fma(t6, t6, -2.0 * t6);
                double c = qq * inv_r;
                bool nb = vi;
                if (MODEA) {
                    c = qq * (inv_r + e.kappa * r2 - e.crf);
                    nb = nb && (r2 <= e.rc2_nb);
                }
                elj += nb ? lj : 0.0;
                ec += nb ? c : 0.0;
                const double x = 0.25 * r2 * (iRi * d.iR);
                double inv_f = inv_r;
                if (!(x >= 60.0)) inv_f = rsqrt64(fma(Ri * d.R, exp(-x), r2));
                eg += vi ? qq * inv_f : 0.0;
            }
        } else
        for (int jj = 0; jj < cj; ++jj) {
            const EJ d = J[jj];
            const double dx = xi - d.x, dy = yi - d.y, dz = zi - d.z;
            const bool ok = vi && (!diag || jj > lane);
            const double r2 = ok ? fma(dx, dx, fma(dy, dy, dz * dz)) : 1.0;
            const double qq = qi * d.q;
            const bool one4 = (m14 >> jj) & 1ull;
            bool nb = ok && !((xm >> jj) & 1ull);
            if (MODEA) nb = nb && (r2 <= e.rc2_nb);
            const double inv_r = rsqrt64(r2);
            const double rmij = one4 ? rm14i + d.rm14 : rmi + d.rm;
            const double eij = one4 ? se14i * d.se14 : sei * d.se;
            const double t = rmij * inv_r;
            const double t2 = t * t;
            const double t6 = t2 * t2 * t2;
            const double lj = eij * fma(t6, t6, -2.0 * t6);
            double c = qq * inv_r;
            if (MODEA) c = qq * (inv_r + e.kappa * r2 - e.crf);
            elj += nb ? lj : 0.0;
            ec += nb ? c : 0.0;
            // f_GB^2 = r^2 + R_i R_j exp(-x), x = r^2 / (4 R_i R_j).  From
            // x >= 60 on, R_i R_j exp(-x) < 1e-12 x 2^-52 r^2 (r^2 = 4 x R_i R_j),
            // far below half an ulp of r^2, so the fma rounds to r^2 exactly
            // and 1/f_GB is the 1/r above: the exponential and the second
            // square root are left out there, bit for bit.
            const double x = 0.25 * r2 * (iRi * d.iR);
            double inv_f = inv_r;
            if (!(x >= 60.0)) inv_f = rsqrt64(fma(Ri * d.R, exp(-x), r2));
            eg += ok ? qq * inv_f : 0.0;
        }
        elj = og_wave_sum_d(elj * OG_KCAL_TO_KJ);
        ec = og_wave_sum_d(ec * kc);
        eg = og_wave_sum_d(eg * pref);
        if (lane == 0) {
            long long* acc = a.eacc + (size_t)it.sys * OG_NE;
            int* fl = a.sflag ? a.sflag + it.sys : nullptr;
            og_atomic_add_fix(acc + OG_E_LJ, og_to_fix_chk(elj, fl));
            og_atomic_add_fix(acc + OG_E_COUL, og_to_fix_chk(ec, fl));
            og_atomic_add_fix(acc + OG_E_GB_PAIR, og_to_fix_chk(eg, fl));
        }
    }
}

}  // namespace

__global__ void __launch_bounds__(256) og_k_energy64_b(og_pair_args a, og_energy64_args e) { energy_body<false>(a, e); }
__global__ void __launch_bounds__(256) og_k_energy64_a(og_pair_args a, og_energy64_args e) { energy_body<true>(a, e); }

void og_launch_pair_energy64(const og_pair_args& a, const og_energy64_args& e, bool mode_a, hipStream_t s) {
    if (a.n_work <= 0) return;
    const int g = pair_grid(a.n_work);
    if (mode_a) hipLaunchKernelGGL(og_k_energy64_a, dim3(g), dim3(256), 0, s, a, e);
    else hipLaunchKernelGGL(og_k_energy64_b, dim3(g), dim3(256), 0, s, a, e);
}

// ---------------------------------------------------------------------------
// Persistent replica steps: block s owns system s for n_steps steps.  Each
// phase runs the same per-item code as the step's separate launches (the
// same FP32 partial sums per work item, the same FP64 per-atom and per-term
// arithmetic; integer sums do not depend on order), so the trajectory is the
// one the launches give, bit for bit.  A system that stops (non-finite
// velocities, a range flag) leaves the loop at that step.
// ---------------------------------------------------------------------------
// threads per block of the persistent replica-step kernel: 8 waves, which
// take the work items of each phase (tile pairs, strips) in turn
#define OG_PERSIST_THREADS 512
template <bool MODEA, bool PRE>
__global__ void __launch_bounds__(OG_PERSIST_THREADS) og_k_replica_steps(og_pair_args pa, og_bonded_args ba, og_radii_args ra,
                                                          og_cluster_args ca, og_back_args bk, og_noise_args nz,
                                                          og_persist_args pp) {
    constexpr int NT = OG_PERSIST_THREADS, NW = OG_PERSIST_THREADS / 64;
    __shared__ float4 s_pq[NW][2 * (OG_TILE + 1)];
    __shared__ int s_stop;
    const int s = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float4* P = s_pq[wave];
    float4* Q = P + (OG_TILE + 1);
    const int a0 = pp.atom_off[s], a1 = pp.atom_off[s + 1];
    for (int k = 0; k < pp.n_steps; ++k) {
        // one thread reads the stop flag and the block follows it, so every
        // thread leaves the loop at the same step
        if (tid == 0) s_stop = __hip_atomic_load(ca.diverged + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (s_stop) break;
        const uint32_t step = pp.step0 + (uint32_t)k;
        if (PRE) {
            for (int i = a0 + tid; i < a1; i += NT) noise_atom_at(nz, i, step);
            __syncthreads();
        }
        for (int q = pp.cl_small_off[s] + tid; q < pp.cl_small_off[s + 1]; q += NT)
            cluster_body<OG_SMALL_CL, PRE>(ca, pp.cl_small[q], step);
        __syncthreads();
#pragma unroll
        for (int kind = 0; kind < 4; ++kind) {
            // global term index of this kind's first term (bonds, angles, dihedrals, impropers)
            const int base = kind == 0 ? 0 : kind == 1 ? ba.nb : kind == 2 ? ba.nb + ba.na : ba.nb + ba.na + ba.nd;
            const int* off = pp.term_off + kind * (pp.n_sys + 1);
            for (int t = off[s] + tid; t < off[s + 1]; t += NT) {
                int e_slot = -1, e_sys = 0;
                long long e_fix = 0;
                bonded_term<false>(ba, base + t, e_slot, e_sys, e_fix);
            }
        }
        for (int w = pp.work_off[s] + wave; w < pp.work_off[s + 1]; w += NW) {
            const og_work it = pa.work[w];
            descreen_item<MODEA>(pa, it, lane, P, Q);
        }
        __syncthreads();
        for (int i = a0 + tid; i < a1; i += NT) {
            radii_body<false>(ra, i);
            gather_atom(ba, i);
        }
        __syncthreads();
        for (int w = pp.strip_off[s] + wave; w < pp.strip_off[s + 1]; w += NW) {
            force_strip<MODEA>(pa, pa.strips + w, lane, P, Q);
        }
        for (int w = pp.fwork_off[s] + wave; w < pp.fwork_off[s + 1]; w += NW) {
            const og_work it = pa.fwork[w];
            force_item<MODEA>(pa, it, lane, P, Q);
        }
        __syncthreads();
        og_back_args b = bk;
        b.ke_acc = pp.ke_acc ? pp.ke_acc + (size_t)k * pp.n_sys : nullptr;
        for (int i = a0 + tid; i < a1; i += NT) back_atom(b, i);
        __syncthreads();
    }
}

void og_launch_replica_steps(const og_pair_args& pa, const og_bonded_args& ba, const og_radii_args& ra,
                             const og_cluster_args& ca, const og_back_args& bk, const og_noise_args& nz,
                             const og_persist_args& pp, bool mode_a, hipStream_t s) {
    if (pp.n_sys <= 0 || pp.n_steps <= 0) return;
    const bool pre = nz.noise != nullptr;
    if (mode_a) {
        if (pre) hipLaunchKernelGGL((og_k_replica_steps<true, true>), dim3(pp.n_sys), dim3(OG_PERSIST_THREADS), 0, s, pa, ba, ra, ca, bk, nz, pp);
        else hipLaunchKernelGGL((og_k_replica_steps<true, false>), dim3(pp.n_sys), dim3(OG_PERSIST_THREADS), 0, s, pa, ba, ra, ca, bk, nz, pp);
    } else {
        if (pre) hipLaunchKernelGGL((og_k_replica_steps<false, true>), dim3(pp.n_sys), dim3(OG_PERSIST_THREADS), 0, s, pa, ba, ra, ca, bk, nz, pp);
        else hipLaunchKernelGGL((og_k_replica_steps<false, false>), dim3(pp.n_sys), dim3(OG_PERSIST_THREADS), 0, s, pa, ba, ra, ca, bk, nz, pp);
    }
}

