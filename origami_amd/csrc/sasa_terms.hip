// JVP terms of the PowerSasa arcs and vertices (derivatives.rs:134-545), the
// second of the three SASA kernels of a force evaluation (the arrangement and
// the ordered sums are in sasa_kernel.hip).
//
// Thread per arc (12 terms: role i, k, l, m x axis) or vertex (9: role i,
// in, out x axis): the primal geometry once, then one forward-mode tangent
// per direction of the roles other than i; role i's terms are minus their
// sum (the angles depend on the positions only through differences).
//
// Each tangent is one-hot: a unit vector on one role's atom.  When the roles'
// atoms are distinct (the usual case) the tangents are compile-time
// constants, and this file is compiled with -ffinite-math-only
// -fno-signed-zeros, which lets the compiler drop the products with the
// zero tangents and the additions of their zero results.  For finite values
// x * 0 = +-0, x + (+-0) = x and 0 / x = +-0 exactly, so every term is the
// value the full tangent gives, up to the sign of a zero (which the
// fixed-point sum does not see).  Roles that share an atom take the general
// form, where the tangents are selected at run time.
#include "og_device.h"
#include "og_sasa.h"
#include "sasa_core.h"

using namespace og_sasa;

namespace {

__device__ __forceinline__ void arc_terms(const og_sasa_data& d, const double* pos, int A) {
    const v3 zero = mk(0, 0, 0);
    const int4 at = d.arc_atoms[A];
    const double4 pr = d.arc_par[A];
    const v3 pi = ldp(pos, at.x), pk = ldp(pos, at.y);
    const double ri = d.radius[at.x], rk = d.radius[at.y];
    const CosAlphaP cp = cos_alpha_primal(pi, pk, ri, rk);
    const bool has_v = at.z >= 0;
    v3 pl = zero, pm = zero;
    double rl = 0.0, rm = 0.0, cak = 0.0;
    ArcThetaP tp{};
    tp.ok = false;
    if (has_v) {
        pl = ldp(pos, at.z);
        pm = ldp(pos, at.w);
        rl = d.radius[at.z];
        rm = d.radius[at.w];
        const double dik = norm(pk - pi);
        cak = (dik * dik + ri * ri - rk * rk) / (2.0 * dik * ri);
        tp = arc_theta_primal(pi, pk, pl, pm, ri, rk, rl, rm, pr.y, pr.z);
    }
    double out[12];
    if (!has_v || (at.y != at.z && at.y != at.w && at.z != at.w)) {
        // distinct roles: role k moves r_k only, role l r_l only, role m r_m only
#pragma unroll
        for (int axis = 0; axis < 3; ++axis) {
            const v3 b = unit(axis);
            const double dcos_k = cos_alpha_tangent(cp, zero, b);
            double tk, tl = 0.0, tm = 0.0;
            if (!has_v) {
                tk = pr.x * dcos_k;
            } else {
                const double dcos_0 = cos_alpha_tangent(cp, zero, zero);
                tk = pr.x * dcos_k + cak * (tp.ok ? arc_theta_tangent(tp, zero, b, zero, zero) : 0.0);
                tl = pr.x * dcos_0 + cak * (tp.ok ? arc_theta_tangent(tp, zero, zero, b, zero) : 0.0);
                tm = pr.x * dcos_0 + cak * (tp.ok ? arc_theta_tangent(tp, zero, zero, zero, b) : 0.0);
            }
            out[3 + axis] = tk;
            out[6 + axis] = tl;
            out[9 + axis] = tm;
            out[axis] = ((0.0 - tk) - tl) - tm;
        }
    } else {
        double ti[3] = {0.0, 0.0, 0.0};
        for (int role = 1; role < 4; ++role) {
            const int xa = role == 1 ? at.y : role == 2 ? at.z : (at.w == at.z ? -1 : at.w);
            for (int axis = 0; axis < 3; ++axis) {
                double term = 0.0;
                if (xa >= 0) {
                    const v3 b = unit(axis);
                    const v3 drk = xa == at.y ? b : zero;
                    const v3 drl = xa == at.z ? b : zero;
                    const v3 drm = xa == at.w ? b : zero;
                    const double dcos = cos_alpha_tangent(cp, zero, drk);
                    const double dth = tp.ok ? arc_theta_tangent(tp, zero, drk, drl, drm) : 0.0;
                    term = pr.x * dcos + cak * dth;
                }
                out[3 * role + axis] = term;
                ti[axis] -= term;
            }
        }
        out[0] = ti[0];
        out[1] = ti[1];
        out[2] = ti[2];
    }
    double* dst = d.arc_terms + 12ll * A;      // one pass, in order
#pragma unroll
    for (int q = 0; q < 12; ++q) dst[q] = out[q];
}

__device__ __forceinline__ void vertex_terms(const og_sasa_data& d, const double* pos, int V) {
    const v3 zero = mk(0, 0, 0);
    const int4 at = d.vtx_atoms[V];
    const v3 pi = ldp(pos, at.x), pk = ldp(pos, at.y), pl = ldp(pos, at.z);
    const VertexEpsP ep = vertex_eps_primal(pi, pk, pl, d.radius[at.x], d.radius[at.y], d.radius[at.z],
                                            d.vtx_sign[V]);
    double out[9];
    if (at.y != at.z) {
#pragma unroll
        for (int axis = 0; axis < 3; ++axis) {
            const v3 b = unit(axis);
            const double tk = ep.ok ? vertex_eps_tangent(ep, zero, b, zero) : 0.0;
            const double tl = ep.ok ? vertex_eps_tangent(ep, zero, zero, b) : 0.0;
            out[3 + axis] = tk;
            out[6 + axis] = tl;
            out[axis] = (0.0 - tk) - tl;
        }
    } else {
        double ti[3] = {0.0, 0.0, 0.0};
        for (int role = 1; role < 3; ++role) {
            const int xa = role == 1 ? at.y : at.z;
            for (int axis = 0; axis < 3; ++axis) {
                const v3 b = unit(axis);
                const v3 drk = xa == at.y ? b : zero;
                const v3 drl = xa == at.z ? b : zero;
                const double term = ep.ok ? vertex_eps_tangent(ep, zero, drk, drl) : 0.0;
                out[3 * role + axis] = term;
                ti[axis] -= term;
            }
        }
        out[0] = ti[0];
        out[1] = ti[1];
        out[2] = ti[2];
    }
    double* dst = d.vtx_terms + 9ll * V;
#pragma unroll
    for (int q = 0; q < 9; ++q) dst[q] = out[q];
}

}  // namespace

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
og_k_sasa_terms(og_sasa_data d, const double* pos) {
    const int n_a = min(d.counters[0], d.arc_cap);
    const int n_v = min(d.counters[1], d.vtx_cap);
    const int total = n_a + n_v;
    for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < total; it += gridDim.x * blockDim.x) {
        if (it < n_a) arc_terms(d, pos, it);
        else vertex_terms(d, pos, it - n_a);
    }
}

void og_launch_sasa_terms(const og_sasa_data& d, const double* pos, hipStream_t s) {
    hipLaunchKernelGGL(og_k_sasa_terms, dim3(d.n_cu * 8), dim3(256), 0, s, d, pos);
}
