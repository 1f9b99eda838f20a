// Learned Born-radius surrogate: descriptor, MLP forward, training (forward +
// backward + weight gradients + Adam), all hand-written for gfx950.
//
// There is no counterpart in origami: the surrogate stands in for the
// descreening pass of compute_born_radii_into_scratch (energy/src/gb.rs:203-342)
// and is trained on the radii that pass produces.  Plain GB stays the arbiter:
// systems whose surrogate radii drift more than a tolerance from the plain
// radii at a periodic check fall back to plain GB (engine.cpp).
//
// MFMA: v_mfma_f32_16x16x32_bf16.  A (16x32) lane l holds row l%16, k =
// 8*(l/16)..+7; B (32x16) lane l holds column l%16, the same k; D (16x16)
// lane l holds column l%16, rows 4*(l/16)..+3.
//
// Layouts (bf16 stored as raw 16-bit words):
//   X    [rows][192]  descriptors (row-major)       -> A operand, layer 1
//   W1t  [128][192]   layer-1 weights, output-major  -> B operand, layer 1
//   W2t  [128][128]   layer-2 weights, output-major  -> B operand, layer 2
//   W2i  [128][128]   layer-2 weights, input-major   -> B operand, dH1 = dZ2 W2^T
//   *T   [feature][batch] transposed activations      -> operands of the
//        weight-gradient GEMMs, whose reduction runs over the batch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "og_sur.h"

#pragma clang fp contract(off)

typedef __bf16 og_bf16x8 __attribute__((ext_vector_type(8)));
typedef float og_f32x4 __attribute__((ext_vector_type(4)));

static std::string g_sur_err;
const char* og_sur_error() { return g_sur_err.c_str(); }

// defined in sasa_kernel.hip
__global__ void og_k_tile_bounds(const double* pos, const int* tile_atom0, const int* tile_cnt,
                                 int n_tiles, double4* out, int* counters);

namespace {

__host__ __device__ inline unsigned short f2bf(float f) {
    unsigned u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);          // round to nearest even
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ og_bf16x8 ld8(const unsigned short* p) {
    const uint4 u = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(og_bf16x8, u);
}

__device__ __forceinline__ og_f32x4 mfma(og_bf16x8 a, og_bf16x8 b, og_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double rld(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

struct DescArgs {
    const double* pos;
    const double* rho;
    const int* cls;
    const int* tile_atom0;
    const int* tile_cnt;
    const int* tile_sys;
    const int* sys_tile_b;
    const int* sys_tile_e;
    const double4* tile_bound;
    const int* mode;       // null: all systems
    unsigned short* X;
};

}  // namespace

// Descriptor: one 256-thread block per 64-atom tile I.  The block lists the
// tiles J of the same system whose bounding spheres come within 12 Å of
// tile I, then its four waves take those tiles in turn: lane = atom i, the
// atoms j of tile J are broadcast one by one, and every j != i with r < 12 Å
// adds one count to i's histogram row in LDS (integer counts, so the result
// does not depend on the order).  The bin floor(r / 0.375 Å) is estimated
// from a single-precision square root and then settled exactly on r^2
// against the bin edges (0.375 k)^2 = 0.140625 k^2, which are exact in FP64.
// Counts are 16-bit, two bins to a word (25 KB of LDS per block, so that the
// kernel fits beside the SASA arrangement on a CU).
#define SUR_DESC_WAVES 4
#define SUR_HW ((OG_SUR_F + 1) / 2)                     // words per histogram row (81)
__global__ void __launch_bounds__(64 * SUR_DESC_WAVES) og_k_sur_desc(DescArgs a) {
    __shared__ unsigned hist[64 * SUR_HW];            // odd stride: conflict-free rows
    __shared__ int near_list[1024];
    __shared__ int n_near;
    const int t = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int sys = a.tile_sys[t];
    if (a.mode && !a.mode[sys]) return;
    for (int q = tid; q < 64 * SUR_HW; q += blockDim.x) hist[q] = 0u;
    if (tid == 0) n_near = 0;
    __syncthreads();
    const double4 bi = a.tile_bound[t];
    const int jb = a.sys_tile_b[sys], je = a.sys_tile_e[sys];
    for (int J = jb + tid; J < je; J += blockDim.x) {
        const double4 bj = a.tile_bound[J];
        const double cx = bj.x - bi.x, cy = bj.y - bi.y, cz = bj.z - bi.z;
        const double reach = bi.w + bj.w + 12.0;
        if (cx * cx + cy * cy + cz * cz <= reach * reach) {
            const int slot = atomicAdd(&n_near, 1);
            if (slot < 1024) near_list[slot] = J;
        }
    }
    __syncthreads();
    const int nn = n_near < 1024 ? n_near : 1024;
    const int a0 = a.tile_atom0[t], cnt = a.tile_cnt[t];
    const bool vi = lane < cnt;
    const int i = a0 + lane;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    if (vi) { xi = a.pos[3 * (size_t)i]; yi = a.pos[3 * (size_t)i + 1]; zi = a.pos[3 * (size_t)i + 2]; }
    unsigned* row = hist + lane * SUR_HW;
    for (int q = wv; q < nn; q += SUR_DESC_WAVES) {
        const int J = near_list[q];
        const int b0 = a.tile_atom0[J], cj = a.tile_cnt[J];
        double xj = 0.0, yj = 0.0, zj = 0.0;
        int cl = 0;
        if (lane < cj) {
            const size_t j = 3 * (size_t)(b0 + lane);
            xj = a.pos[j]; yj = a.pos[j + 1]; zj = a.pos[j + 2];
            cl = a.cls[b0 + lane];
        }
        for (int jj = 0; jj < cj; ++jj) {
            const double dx = rld(xj, jj) - xi, dy = rld(yj, jj) - yi, dz = rld(zj, jj) - zi;
            const int c = __builtin_amdgcn_readlane(cl, jj);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (vi && b0 + jj != i && r2 < 144.0) {
                int k = (int)(sqrtf((float)r2) * (1.0f / 0.375f));
                k = k > 31 ? 31 : (k < 0 ? 0 : k);
                while (k < 31 && 0.140625 * (double)((k + 1) * (k + 1)) <= r2) ++k;
                while (k > 0 && 0.140625 * (double)(k * k) > r2) --k;
                const int b = c * 32 + k;
                atomicAdd(row + (b >> 1), 1u << ((b & 1) * 16));
            }
        }
    }
    __syncthreads();
    // write the 64 rows: thread per (row, feature)
    for (int q = tid; q < cnt * OG_SUR_K1; q += blockDim.x) {
        const int rr = q / OG_SUR_K1, f = q % OG_SUR_K1;
        unsigned short v = 0;
        if (f < 160) v = f2bf((float)((hist[rr * SUR_HW + (f >> 1)] >> ((f & 1) * 16)) & 0xFFFFu) * 0.125f);
        else if (f == 160) v = f2bf((float)((a.rho[a0 + rr] - 1.5) * 4.0));
        a.X[(size_t)(a0 + rr) * OG_SUR_K1 + f] = v;
    }
}

// ---------------------------------------------------------------------------
// Forward pass for 16 rows per wave (shared by inference and training).
// z1, z2: pre-activations (bias included) in the D layout; h2: SiLU(z2).
// hs: this wave's LDS tile [16][136] (bf16).  Returns y for rows 4g..4g+3
// (every lane of a 16-lane group holds its group's four sums).
// ---------------------------------------------------------------------------
namespace {

struct MlpParams {
    const unsigned short* W1t;
    const unsigned short* W2t;
    const unsigned short* W2i;
    const float* p;            // FP32 master (b1, b2, w3, b3)
};

__device__ __forceinline__ void mlp_forward(const unsigned short* X, int row0, const MlpParams& P,
                                            unsigned short (*hs)[136], int lane, og_f32x4 (&z1)[8],
                                            og_f32x4 (&z2)[8], float (&y)[4]) {
    const int r = lane & 15, g = lane >> 4;
    og_bf16x8 a1[6];
#pragma unroll
    for (int kk = 0; kk < 6; ++kk) a1[kk] = ld8(X + (size_t)(row0 + r) * OG_SUR_K1 + kk * 32 + g * 8);
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
        og_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 6; ++kk)
            acc = mfma(a1[kk], ld8(P.W1t + (size_t)(nt * 16 + r) * OG_SUR_K1 + kk * 32 + g * 8), acc);
        const float b = P.p[OG_SUR_OFF_B1 + nt * 16 + r];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[i] += b;
            hs[g * 4 + i][nt * 16 + r] = f2bf(silu(acc[i]));
        }
        z1[nt] = acc;
    }
    wave_sync();
    og_bf16x8 a2[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) a2[kk] = ld8(&hs[r][kk * 32 + g * 8]);
    float ys[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
        og_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
            acc = mfma(a2[kk], ld8(P.W2t + (size_t)(nt * 16 + r) * OG_SUR_H + kk * 32 + g * 8), acc);
        const float b = P.p[OG_SUR_OFF_B2 + nt * 16 + r];
        const float w = P.p[OG_SUR_OFF_W3 + nt * 16 + r];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[i] += b;
            ys[i] += silu(acc[i]) * w;
        }
        z2[nt] = acc;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float v = ys[i];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        y[i] = v + P.p[OG_SUR_OFF_B3];
    }
    wave_sync();   // hs is reused by the caller
}

}  // namespace

// Inference, 64 rows per wave: four 16-row groups share every weight
// fragment load (the weights are read once per 64 rows instead of once per
// 16).  Each row goes through the 16-row forward's exact operations, so the
// radii are unchanged.
// Two waves per block (35 KB of LDS, so that the kernel fits beside the SASA
// arrangement on a CU).
__global__ void __launch_bounds__(128) og_k_sur_infer64(const unsigned short* X, int n_valid, int n_rows,
                                                        MlpParams P, double* r_out) {
    __shared__ __attribute__((aligned(16))) unsigned short hs[2][64][136];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int row0 = (blockIdx.x * 2 + w) * 64;
    if (row0 >= n_valid) return;
    unsigned short (*h)[136] = hs[w];
    og_bf16x8 a1[4][6];
#pragma unroll
    for (int rg = 0; rg < 4; ++rg) {
        const int rr = row0 + rg * 16 + r;
        const int rs = rr < n_rows ? rr : 0;
#pragma unroll
        for (int kk = 0; kk < 6; ++kk) a1[rg][kk] = ld8(X + (size_t)rs * OG_SUR_K1 + kk * 32 + g * 8);
    }
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
        og_f32x4 acc[4];
#pragma unroll
        for (int rg = 0; rg < 4; ++rg) acc[rg] = og_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 6; ++kk) {
            const og_bf16x8 bw = ld8(P.W1t + (size_t)(nt * 16 + r) * OG_SUR_K1 + kk * 32 + g * 8);
#pragma unroll
            for (int rg = 0; rg < 4; ++rg) acc[rg] = mfma(a1[rg][kk], bw, acc[rg]);
        }
        const float b = P.p[OG_SUR_OFF_B1 + nt * 16 + r];
#pragma unroll
        for (int rg = 0; rg < 4; ++rg)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float z = acc[rg][i] + b;
                h[rg * 16 + g * 4 + i][nt * 16 + r] = f2bf(silu(z));
            }
    }
    wave_sync();
    og_bf16x8 a2[4][4];
#pragma unroll
    for (int rg = 0; rg < 4; ++rg)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) a2[rg][kk] = ld8(&h[rg * 16 + r][kk * 32 + g * 8]);
    float ys[4][4];
#pragma unroll
    for (int rg = 0; rg < 4; ++rg)
#pragma unroll
        for (int i = 0; i < 4; ++i) ys[rg][i] = 0.f;
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
        og_f32x4 acc[4];
#pragma unroll
        for (int rg = 0; rg < 4; ++rg) acc[rg] = og_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const og_bf16x8 bw = ld8(P.W2t + (size_t)(nt * 16 + r) * OG_SUR_H + kk * 32 + g * 8);
#pragma unroll
            for (int rg = 0; rg < 4; ++rg) acc[rg] = mfma(a2[rg][kk], bw, acc[rg]);
        }
        const float b = P.p[OG_SUR_OFF_B2 + nt * 16 + r];
        const float wv = P.p[OG_SUR_OFF_W3 + nt * 16 + r];
#pragma unroll
        for (int rg = 0; rg < 4; ++rg)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float z = acc[rg][i] + b;
                ys[rg][i] += silu(z) * wv;
            }
    }
#pragma unroll
    for (int rg = 0; rg < 4; ++rg)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float v = ys[rg][i];
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            v += __shfl_xor(v, 8, 64);
            const float y = v + P.p[OG_SUR_OFF_B3];
            const int row = row0 + rg * 16 + g * 4 + i;
            if (r == 0 && row < n_valid) r_out[row] = OG_SUR_R0 + OG_SUR_RS * (double)y;
        }
}

// Training forward + backward for 16 batch rows per wave.
struct TrainArgs {
    const unsigned short* Xb;   // [B][K1]
    const float* tb;            // [B] normalised targets
    int B;
    unsigned short *H1T, *dZ1T, *dZ2T;   // [H][B] bf16
    float *dZ1Tf, *dZ2Tf, *H2Tf;         // [H][B] fp32
    float *dy, *lossr;                   // [B]
};

__global__ void __launch_bounds__(256) og_k_sur_fwdbwd(TrainArgs a, MlpParams P) {
    __shared__ __attribute__((aligned(16))) unsigned short hs[4][16][136];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int row0 = (blockIdx.x * 4 + w) * 16;
    if (row0 >= a.B) return;
    og_f32x4 z1[8], z2[8];
    float y[4];
    mlp_forward(a.Xb, row0, P, hs[w], lane, z1, z2, y);
    const float inv_b2 = 2.0f / (float)a.B;
    float dyv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = row0 + g * 4 + i;
        const float d = y[i] - a.tb[row];
        dyv[i] = d * inv_b2;
        if (r == 0) {
            a.dy[row] = dyv[i];
            a.lossr[row] = d * d;
        }
    }
    // dZ2 = dy w3 * silu'(z2); activations stored transposed for the weight GEMMs
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
        const int j = nt * 16 + r;
        const float w3 = P.p[OG_SUR_OFF_W3 + j];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = row0 + g * 4 + i;
            const size_t o = (size_t)j * a.B + row;
            const float dz2 = dyv[i] * w3 * dsilu(z2[nt][i]);
            a.dZ2Tf[o] = dz2;
            const unsigned short dz2b = f2bf(dz2);
            a.dZ2T[o] = dz2b;
            a.H2Tf[o] = silu(z2[nt][i]);
            a.H1T[o] = f2bf(silu(z1[nt][i]));
            hs[w][g * 4 + i][j] = dz2b;
        }
    }
    wave_sync();
    // dH1 = dZ2 W2^T, then dZ1 = dH1 * silu'(z1)
    og_bf16x8 a2[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) a2[kk] = ld8(&hs[w][r][kk * 32 + g * 8]);
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
        og_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
            acc = mfma(a2[kk], ld8(P.W2i + (size_t)(nt * 16 + r) * OG_SUR_H + kk * 32 + g * 8), acc);
        const int k = nt * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = row0 + g * 4 + i;
            const size_t o = (size_t)k * a.B + row;
            const float dz1 = acc[i] * dsilu(z1[nt][i]);
            a.dZ1Tf[o] = dz1;
            a.dZ1T[o] = f2bf(dz1);
        }
    }
}

// Weight gradients, one wave per task, fixed reduction order (deterministic):
//   tasks [0, 96)      dW1 tiles: D[j][k] = sum_b dZ1T[j][b] XbT[k][b]
//   tasks [96, 160)    dW2 tiles: D[j][k] = sum_b dZ2T[j][b] H1T[k][b]
//   tasks [160, 288)   db1[j] = sum_b dZ1[b][j]
//   tasks [288, 416)   db2[j] = sum_b dZ2[b][j]
//   tasks [416, 544)   dW3[j] = sum_b H2[b][j] dy[b]
//   task 544           db3 = sum_b dy[b];   task 545  loss = mean (y - t)^2
struct GradArgs {
    const unsigned short *dZ1T, *XbT, *dZ2T, *H1T;
    const float *dZ1Tf, *dZ2Tf, *H2Tf, *dy, *lossr;
    int B;
    float* grad;
    float* loss;
};

__global__ void __launch_bounds__(64) og_k_sur_wgrad(GradArgs a) {
    const int task = blockIdx.x;
    const int lane = threadIdx.x;
    const int r = lane & 15, g = lane >> 4;
    if (task < 160) {
        const bool first = task < 96;
        const int jt = first ? task / 12 : (task - 96) / 8;
        const int kt = first ? task % 12 : (task - 96) % 8;
        const unsigned short* A = first ? a.dZ1T : a.dZ2T;
        const unsigned short* Bm = first ? a.XbT : a.H1T;
        og_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kk = 0; kk < a.B / 32; ++kk)
            acc = mfma(ld8(A + (size_t)(jt * 16 + r) * a.B + kk * 32 + g * 8),
                       ld8(Bm + (size_t)(kt * 16 + r) * a.B + kk * 32 + g * 8), acc);
        const int k = kt * 16 + r;
        const int off = first ? OG_SUR_OFF_W1 : OG_SUR_OFF_W2;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = jt * 16 + g * 4 + i;
            a.grad[off + k * OG_SUR_H + j] = acc[i];
        }
        return;
    }
    float s = 0.0f;
    int out = -1;
    if (task < 288) {
        const int j = task - 160;
        for (int b = lane; b < a.B; b += 64) s += a.dZ1Tf[(size_t)j * a.B + b];
        out = OG_SUR_OFF_B1 + j;
    } else if (task < 416) {
        const int j = task - 288;
        for (int b = lane; b < a.B; b += 64) s += a.dZ2Tf[(size_t)j * a.B + b];
        out = OG_SUR_OFF_B2 + j;
    } else if (task < 544) {
        const int j = task - 416;
        for (int b = lane; b < a.B; b += 64) s += a.H2Tf[(size_t)j * a.B + b] * a.dy[b];
        out = OG_SUR_OFF_W3 + j;
    } else if (task == 544) {
        for (int b = lane; b < a.B; b += 64) s += a.dy[b];
        out = OG_SUR_OFF_B3;
    } else {
        for (int b = lane; b < a.B; b += 64) s += a.lossr[b];
    }
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
        if (out >= 0) a.grad[out] = s;
        else *a.loss = s / (float)a.B;
    }
}

// Adam update of the FP32 master parameters, then the BF16 operand copies.
__global__ void __launch_bounds__(256) og_k_sur_adam(float* p, float* m, float* v, const float* grad,
                                                     float lr, float c1, float c2, int update,
                                                     unsigned short* W1t, unsigned short* W2t,
                                                     unsigned short* W2i) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= OG_SUR_P) return;
    float w = p[q];
    if (update) {
        const float gq = grad[q];
        const float mq = 0.9f * m[q] + 0.1f * gq;
        const float vq = 0.999f * v[q] + 0.001f * gq * gq;
        m[q] = mq;
        v[q] = vq;
        w -= lr * (mq / c1) / (sqrtf(vq / c2) + 1e-8f);
        p[q] = w;
    }
    if (q < OG_SUR_OFF_B1) {
        const int k = q / OG_SUR_H, j = q % OG_SUR_H;
        W1t[j * OG_SUR_K1 + k] = f2bf(w);
    } else if (q >= OG_SUR_OFF_W2 && q < OG_SUR_OFF_B2) {
        const int e = q - OG_SUR_OFF_W2;
        const int k = e / OG_SUR_H, j = e % OG_SUR_H;
        W2t[j * OG_SUR_H + k] = f2bf(w);
        W2i[k * OG_SUR_H + j] = f2bf(w);
    }
}

__device__ __forceinline__ unsigned long long sur_mix(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Batch gather: row b of the batch = training row hash(seed, b) mod n_train.
__global__ void __launch_bounds__(256) og_k_sur_gather(const unsigned short* Xtr, const float* ttr,
                                                       long long n_train, int B, unsigned long long seed,
                                                       unsigned short* Xb, unsigned short* XbT, float* tb) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)B * OG_SUR_K1) return;
    const int b = (int)(q / OG_SUR_K1), f = (int)(q % OG_SUR_K1);
    const long long idx = (long long)(sur_mix(seed ^ sur_mix((unsigned long long)b)) % (unsigned long long)n_train);
    const unsigned short x = Xtr[idx * OG_SUR_K1 + f];
    Xb[(size_t)b * OG_SUR_K1 + f] = x;
    XbT[(size_t)f * B + b] = x;
    if (f == 0) tb[b] = ttr[idx];
}

// Append descriptors and normalised plain radii of atoms [0, n) at row base.
__global__ void __launch_bounds__(256) og_k_sur_append(const unsigned short* X, const double* radii, int n,
                                                       unsigned short* Xtr, float* ttr, long long base) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)n * OG_SUR_K1) return;
    const int i = (int)(q / OG_SUR_K1), f = (int)(q % OG_SUR_K1);
    Xtr[(base + i) * OG_SUR_K1 + f] = X[(size_t)i * OG_SUR_K1 + f];
    if (f == 0) ttr[base + i] = (float)((radii[i] - OG_SUR_R0) / OG_SUR_RS);
}

// Fallback check: dev[s] = max over the system's atoms of |R_sur - R_plain| / R_plain
// (float bits; non-negative floats order like their bit patterns).
__global__ void og_k_sur_zero(unsigned* dev, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dev[q] = 0u;
}

__global__ void og_k_sur_fallback(int* mode, const unsigned* dev, int n, float tol) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n && mode[q] && !(__uint_as_float(dev[q]) <= tol)) mode[q] = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

template <class T>
bool dalloc(T** p, size_t n) {
    if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return false;
    return hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess;
}

template <class T>
bool upload(T** p, const std::vector<T>& v) {
    if (!dalloc(p, v.size())) return false;
    return v.empty() || hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

int fail(const char* m) {
    g_sur_err = m;
    return -2;
}

MlpParams params_of(const og_sur* s) { return {s->W1t, s->W2t, s->W2i, s->p}; }

int refresh_copies(og_sur* s) {
    hipLaunchKernelGGL(og_k_sur_adam, dim3((OG_SUR_P + 255) / 256), dim3(256), 0, s->stream, s->p, s->m, s->v,
                       s->g, 0.0f, 1.0f, 1.0f, 0, s->W1t, s->W2t, s->W2i);
    return hipGetLastError() == hipSuccess ? 0 : fail("kernel launch failed");
}

// forward + backward + weight gradients on the batch already in Xb/XbT/tb
int batch_gradient(og_sur* s, int B) {
    TrainArgs t{s->Xb, s->tb, B, s->H1T, s->dZ1T, s->dZ2T, s->dZ1Tf, s->dZ2Tf, s->H2Tf, s->dy, s->lossr};
    hipLaunchKernelGGL(og_k_sur_fwdbwd, dim3((B / 16 + 3) / 4), dim3(256), 0, s->stream, t, params_of(s));
    GradArgs ga{s->dZ1T, s->XbT, s->dZ2T, s->H1T, s->dZ1Tf, s->dZ2Tf, s->H2Tf, s->dy, s->lossr, B, s->g, s->loss1};
    hipLaunchKernelGGL(og_k_sur_wgrad, dim3(546), dim3(64), 0, s->stream, ga);
    return hipGetLastError() == hipSuccess ? 0 : fail("kernel launch failed");
}

}  // namespace

int og_sur_create(og_sur* s, int n_atoms, int n_sys, const std::vector<double>& mass,
                  const std::vector<int>& sys_tile_off, const double* pos_dev, const double* rho_dev,
                  const int* tile_atom0_dev, const int* tile_cnt_dev, int n_tiles, long long capacity,
                  unsigned long long seed, hipStream_t stream) {
    s->n_atoms = n_atoms;
    s->n_pad = (n_atoms + 15) / 16 * 16;
    s->n_tiles = n_tiles;
    s->n_sys = n_sys;
    s->stream = stream;
    s->pos = pos_dev;
    s->rho = rho_dev;
    s->tile_atom0 = tile_atom0_dev;
    s->tile_cnt = tile_cnt_dev;
    std::vector<int> cls(n_atoms), tsys(n_tiles), tb(n_sys), te(n_sys);
    for (int i = 0; i < n_atoms; ++i) {
        const double mm = mass[i];
        cls[i] = mm < 2.0 ? 0 : mm < 13.0 ? 1 : mm < 15.0 ? 2 : mm < 17.0 ? 3 : 4;
    }
    for (int q = 0; q < n_sys; ++q) {
        tb[q] = sys_tile_off[q];
        te[q] = sys_tile_off[q + 1];
        if (te[q] - tb[q] > 1024) return fail("the descriptor kernel handles systems of up to 65536 atoms");
        for (int t = tb[q]; t < te[q]; ++t) tsys[t] = q;
    }
    if (!upload(&s->cls, cls) || !upload(&s->tile_sys, tsys) || !upload(&s->sys_tile_b, tb) ||
        !upload(&s->sys_tile_e, te) || !dalloc(&s->tile_bound, (size_t)n_tiles) ||
        !dalloc(&s->X, (size_t)s->n_pad * OG_SUR_K1) || !dalloc(&s->r_sur, (size_t)n_atoms) ||
        !dalloc(&s->mode, (size_t)n_sys) || !dalloc(&s->dev, (size_t)n_sys))
        return fail("device allocation failed");
    // parameters: W ~ N(0, 1/fan_in) from a seeded SplitMix64 + Box-Muller stream, biases 0
    std::vector<float> p(OG_SUR_P, 0.0f);
    unsigned long long st = seed;
    auto next = [&st]() {
        unsigned long long z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto gauss = [&]() {
        const double u1 = ((next() >> 11) + 0.5) * 0x1.0p-53, u2 = ((next() >> 11) + 0.5) * 0x1.0p-53;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    };
    for (int k = 0; k < OG_SUR_F; ++k)
        for (int j = 0; j < OG_SUR_H; ++j) p[OG_SUR_OFF_W1 + k * OG_SUR_H + j] = (float)(gauss() / std::sqrt((double)OG_SUR_F));
    for (int k = 0; k < OG_SUR_H; ++k)
        for (int j = 0; j < OG_SUR_H; ++j) p[OG_SUR_OFF_W2 + k * OG_SUR_H + j] = (float)(gauss() / std::sqrt((double)OG_SUR_H));
    for (int j = 0; j < OG_SUR_H; ++j) p[OG_SUR_OFF_W3 + j] = (float)(gauss() / std::sqrt((double)OG_SUR_H));
    if (!upload(&s->p, p) || !dalloc(&s->m, OG_SUR_P) || !dalloc(&s->v, OG_SUR_P) || !dalloc(&s->g, OG_SUR_P) ||
        !dalloc(&s->W1t, (size_t)OG_SUR_H * OG_SUR_K1) || !dalloc(&s->W2t, (size_t)OG_SUR_H * OG_SUR_H) ||
        !dalloc(&s->W2i, (size_t)OG_SUR_H * OG_SUR_H))
        return fail("device allocation failed");
    s->cap = capacity;
    s->n_train = 0;
    const size_t B = OG_SUR_BMAX;
    if (!dalloc(&s->Xtr, (size_t)std::max(capacity, 1ll) * OG_SUR_K1) || !dalloc(&s->ttr, (size_t)std::max(capacity, 1ll)) ||
        !dalloc(&s->Xb, B * OG_SUR_K1) || !dalloc(&s->XbT, B * OG_SUR_K1) || !dalloc(&s->H1T, B * OG_SUR_H) ||
        !dalloc(&s->dZ1T, B * OG_SUR_H) || !dalloc(&s->dZ2T, B * OG_SUR_H) || !dalloc(&s->dZ1Tf, B * OG_SUR_H) ||
        !dalloc(&s->dZ2Tf, B * OG_SUR_H) || !dalloc(&s->H2Tf, B * OG_SUR_H) || !dalloc(&s->tb, B) ||
        !dalloc(&s->dy, B) || !dalloc(&s->lossr, B) || !dalloc(&s->loss1, 1))
        return fail("device allocation failed (training buffers)");
    s->adam_t = 0;
    if (refresh_copies(s)) return -2;
    return hipStreamSynchronize(stream) == hipSuccess ? 0 : fail("surrogate set-up failed");
}

void og_sur_free(og_sur* s) {
    void* ptrs[] = {s->cls, s->tile_sys, s->sys_tile_b, s->sys_tile_e, s->tile_bound, s->X, s->r_sur, s->mode,
                    s->dev, s->p, s->m, s->v, s->g, s->W1t, s->W2t, s->W2i, s->Xtr, s->ttr, s->Xb, s->XbT,
                    s->H1T, s->dZ1T, s->dZ2T, s->dZ1Tf, s->dZ2Tf, s->H2Tf, s->tb, s->dy, s->lossr, s->loss1};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    *s = og_sur{};
}

int og_sur_descriptors(og_sur* s, bool all) {
    if (s->n_tiles == 0) return 0;
    hipLaunchKernelGGL(og_k_tile_bounds, dim3(s->n_tiles), dim3(64), 0, s->stream, s->pos, s->tile_atom0,
                       s->tile_cnt, s->n_tiles, s->tile_bound, nullptr);
    DescArgs a{s->pos, s->rho, s->cls, s->tile_atom0, s->tile_cnt, s->tile_sys, s->sys_tile_b, s->sys_tile_e,
               s->tile_bound, all ? nullptr : s->mode, s->X};
    hipLaunchKernelGGL(og_k_sur_desc, dim3(s->n_tiles), dim3(64 * SUR_DESC_WAVES), 0, s->stream, a);
    return hipGetLastError() == hipSuccess ? 0 : fail("kernel launch failed");
}

int og_sur_infer(og_sur* s) {
    if (s->n_atoms == 0) return 0;
    const int waves64 = (s->n_atoms + 63) / 64;
    hipLaunchKernelGGL(og_k_sur_infer64, dim3((waves64 + 1) / 2), dim3(128), 0, s->stream, s->X, s->n_atoms,
                       s->n_pad, params_of(s), s->r_sur);
    return hipGetLastError() == hipSuccess ? 0 : fail("kernel launch failed");
}

int og_sur_collect(og_sur* s, const double* radii_dev) {
    const long long room = s->cap - s->n_train;
    const int n = (int)std::min<long long>(room, s->n_atoms);
    if (n <= 0) return 0;
    const long long tot = (long long)n * OG_SUR_K1;
    hipLaunchKernelGGL(og_k_sur_append, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s->stream, s->X,
                       radii_dev, n, s->Xtr, s->ttr, s->n_train);
    if (hipGetLastError() != hipSuccess) return fail("kernel launch failed");
    s->n_train += n;
    return 0;
}

int og_sur_train_step(og_sur* s, int batch, double lr, unsigned long long seed, float* loss) {
    if (s->n_train <= 0) return fail("no training rows collected");
    if (batch < 32 || batch > OG_SUR_BMAX || batch % 32) return fail("batch must be a multiple of 32 in [32, 8192]");
    const long long tot = (long long)batch * OG_SUR_K1;
    hipLaunchKernelGGL(og_k_sur_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s->stream, s->Xtr,
                       s->ttr, s->n_train, batch, seed, s->Xb, s->XbT, s->tb);
    if (batch_gradient(s, batch)) return -2;
    s->adam_t += 1;
    const float c1 = (float)(1.0 - std::pow(0.9, s->adam_t)), c2 = (float)(1.0 - std::pow(0.999, s->adam_t));
    hipLaunchKernelGGL(og_k_sur_adam, dim3((OG_SUR_P + 255) / 256), dim3(256), 0, s->stream, s->p, s->m, s->v,
                       s->g, (float)lr, c1, c2, 1, s->W1t, s->W2t, s->W2i);
    if (hipGetLastError() != hipSuccess) return fail("kernel launch failed");
    if (loss && hipMemcpyAsync(loss, s->loss1, sizeof(float), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
        return fail("copy failed");
    return 0;
}

int og_sur_gradient(og_sur* s, const float* x, const float* r_target, int n, float* grad, double* loss) {
    if (n < 32 || n > OG_SUR_BMAX || n % 32) return fail("n must be a multiple of 32 in [32, 8192]");
    std::vector<unsigned short> xb((size_t)n * OG_SUR_K1, 0), xt((size_t)n * OG_SUR_K1, 0);
    std::vector<float> t(n);
    for (int b = 0; b < n; ++b) {
        for (int f = 0; f < OG_SUR_F; ++f) {
            const unsigned short h = f2bf(x[(size_t)b * OG_SUR_F + f]);
            xb[(size_t)b * OG_SUR_K1 + f] = h;
            xt[(size_t)f * n + b] = h;
        }
        t[b] = (float)((r_target[b] - OG_SUR_R0) / OG_SUR_RS);
    }
    if (hipMemcpyAsync(s->Xb, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(s->XbT, xt.data(), xt.size() * 2, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(s->tb, t.data(), t.size() * 4, hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return fail("copy failed");
    if (batch_gradient(s, n)) return -2;
    float l = 0.0f;
    if (hipMemcpyAsync(grad, s->g, OG_SUR_P * sizeof(float), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipMemcpyAsync(&l, s->loss1, sizeof(float), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return fail("copy failed");
    if (loss) *loss = l;
    return 0;
}

int og_sur_set_params(og_sur* s, const float* p) {
    if (hipMemcpyAsync(s->p, p, OG_SUR_P * sizeof(float), hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return fail("copy failed");
    if (refresh_copies(s)) return -2;
    return hipStreamSynchronize(s->stream) == hipSuccess ? 0 : fail("copy failed");
}

int og_sur_get_params(og_sur* s, float* p) {
    if (hipMemcpyAsync(p, s->p, OG_SUR_P * sizeof(float), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return fail("copy failed");
    return 0;
}

int og_sur_check_begin(og_sur* s) {
    hipLaunchKernelGGL(og_k_sur_zero, dim3((s->n_sys + 255) / 256), dim3(256), 0, s->stream, s->dev, s->n_sys);
    return hipGetLastError() == hipSuccess ? 0 : fail("kernel launch failed");
}

int og_sur_check_apply(og_sur* s) {
    hipLaunchKernelGGL(og_k_sur_fallback, dim3((s->n_sys + 255) / 256), dim3(256), 0, s->stream, s->mode, s->dev,
                       s->n_sys, (float)s->tol);
    return hipGetLastError() == hipSuccess ? 0 : fail("kernel launch failed");
}
