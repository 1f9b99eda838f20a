// Shared device-side definitions for the origami MI355X engine (gfx950).
//
// Layout conventions used by every kernel:
//   * Atoms of all systems in a batch are concatenated ("atom index").
//   * Each system is cut into 64-atom tiles (one tile = one wave64).  The
//     last tile of a system may be partial; tiles never straddle systems.
//   * Tile-local FP32 coordinates are stored relative to the FP64 position
//     of the tile's first atom, so pair offsets keep ~1e-6 Å precision on
//     chains that span ~1000 Å.
//   * Forces, descreening integrals and energies are accumulated as int64
//     fixed point (2^32 units per kJ/mol/Å, per Å^-1, per kJ/mol): integer
//     addition is associative, so results are bit-reproducible regardless
//     of wave scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OG_TILE 64
#define OG_FIX_SCALE 4294967296.0        // 2^32
#define OG_FIX_INV (1.0 / 4294967296.0)

// Physical constants (origami units: Å, kJ/mol, fs, K, Da).
#define OG_KCAL_TO_KJ 4.184
#define OG_COULOMB_KCAL 332.0637
#define OG_EPS_WATER 78.5
#define OG_OBC_OFFSET 0.09
#define OG_ACCEL 1.0e-4
#define OG_KB 8.314462618e-3

// Flags shared with the host.
#define OG_F_MODE_A 1u        // origami integrator cutoffs (10 Å LJ/RF-Coulomb/GB, 20 Å descreen)
#define OG_F_SASA 2u          // add analytical SASA forces
#define OG_F_ENERGY 4u        // accumulate energies
#define OG_F_SURROGATE 8u     // Born radii from the learned surrogate
#define OG_F_SASA_SR 16u      // energies only: Shrake-Rupley SASA energy term

struct og_work {              // one tile pair (I <= J) of one system
    int32_t ti, tj;           // global tile indices
    int32_t mask;             // index into the exclusion-mask table, -1 if none
    int32_t sys;              // system index
    int32_t part;             // -1: all 64 steps; 0 / 1: steps 0-31 / 32-63 (I < J only)
};

// Force work item of a large system: the two i tiles (ti, ti + 1) against
// up to OG_STRIP_J j tiles of the same system (the engine's strip length, set
// by the autotuner), none of these tile pairs holding an excluded or 1-4
// pair.  One wave keeps the i-side sums of both i tiles in registers across
// the j tiles and flushes them once.
#define OG_STRIP_J 8
struct og_strip {
    int32_t ti;               // first i tile (global index); the second is ti + 1
    int32_t nj;               // j tiles in use (1..OG_STRIP_J)
    int32_t tj[OG_STRIP_J];   // j tiles (global indices), each > ti + 1
    int32_t sys;              // system index
    int32_t pad;
};

// Per-system status bits (the engine's diverged flags and the per-evaluation
// range flags share them).
#define OG_ST_NONFINITE 1     // non-finite velocities after the closing half-kick
#define OG_ST_RANGE 2         // a contribution outside the fixed-point range (or NaN)
#define OG_ST_SASA 4          // a SASA atom exceeded the arrangement's capacity

// Largest magnitude one contribution may have (kJ/mol/Å, kJ/mol or Å^-1):
// 2^30.  An int64 accumulator at 2^32 units holds sums below 2^31; a sum
// whose magnitude reaches 2^30 is reported as well when it is read (the
// force finalisation), so a wrapped sum goes unreported only when the
// contributions to one atom component add up to more than 3 * 2^30 in
// magnitude (four or more of them, each below 2^30).
#define OG_FIX_LIMIT 1073741824.0

__device__ __forceinline__ long long og_to_fix(double x) {
    return __double2ll_rn(x * OG_FIX_SCALE);
}

// og_to_fix for values that depend on the state: a NaN or a magnitude at or
// beyond OG_FIX_LIMIT sets OG_ST_RANGE in *flag (when flag is set) and
// converts as 0 (NaN) or the signed limit, so the integer sums stay defined
// and the system is reported instead of carrying a wrapped value.
__device__ __forceinline__ long long og_to_fix_chk(double x, int* flag) {
    if (!(fabs(x) < OG_FIX_LIMIT)) {
        if (flag) atomicOr(flag, OG_ST_RANGE);
        x = (x != x) ? 0.0 : (x > 0.0 ? OG_FIX_LIMIT : -OG_FIX_LIMIT);
    }
    return __double2ll_rn(x * OG_FIX_SCALE);
}

__device__ __forceinline__ void og_atomic_add_fix(long long* p, long long v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

// wave-wide int64 sum (device library reduction over DPP)
extern "C" __device__ __attribute__((const)) long long __ockl_wfred_add_i64(long long);

// Adds each active lane's fixed-point value v to acc[key * stride + slot].
// Lanes with the same key are summed first (int64, exact and order-free), so
// each distinct key in the wave costs one atomic instead of one per lane.
// Call from uniform control flow (every lane of the wave).
__device__ __forceinline__ void og_wave_add_fix_keyed(long long* acc, int key, int stride, int slot, long long v,
                                                      bool active) {
    unsigned long long pending = __ballot(active);
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const int k = __shfl(key, leader, 64);
        const bool mine = active && key == k;
        const long long part = __ockl_wfred_add_i64(mine ? v : 0ll);
        if ((int)(threadIdx.x & 63) == leader) og_atomic_add_fix(acc + (size_t)k * stride + slot, part);
        pending &= ~__ballot(mine);
    }
}

// ---------------------------------------------------------------------------
// Wave64 cross-lane sum.  Row (16-lane) stages use DPP; the two cross-row
// stages use ds_swizzle/bpermute through __shfl_xor.  Every lane ends with
// the full sum.  Deterministic: fixed combination order.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float og_dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ float og_wave_sum(float v) {
    v += og_dpp<0xB1>(v);   // quad_perm [1,0,3,2]
    v += og_dpp<0x4E>(v);   // quad_perm [2,3,0,1]
    v += og_dpp<0x124>(v);  // row_ror:4
    v += og_dpp<0x128>(v);  // row_ror:8
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

__device__ __forceinline__ double og_wave_sum_d(double v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
