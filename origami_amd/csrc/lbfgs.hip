// Device-resident batched L-BFGS with Armijo backtracking, every system of
// the engine independently (origami's lbfgs, dynamics/src/lbfgs.rs:61-254,
// and backtracking, line_search.rs:52-101).
//
// One block of 256 threads per system holds that system's vectors (3n FP64
// degrees of freedom, strided over the block) and its control state; the
// host only launches the phases and reads four counters between them.
// Reductions are block-level and deterministic (fixed wave order).
//
// Per iteration (the host's loop in engine.cpp, og_minimize):
//   begin    g = -F, |g|inf test, direction by the two-loop recursion over
//            the system's history ring (steepest descent when it is empty),
//            alpha0 = min(alpha0, max_step / |p|inf), p.g, x0 = x
//   probes   trial x = x0 + alpha p, FP64 energies, Armijo test, alpha *=
//            contraction (until accepted or alpha < min_alpha)
//   fallback a failed L-BFGS search retries along -g; a failed retry stops
//            the system at x0
//   update   forces at the accepted point, (s, y) pushed when s.y > 1e-12,
//            origami's convergence test
#include "og_device.h"
#include "og_kernels.h"

namespace {

constexpr int kThreads = 256;

// deterministic block sum / max (every thread gets the result)
__device__ double block_sum(double v, double* red) {
    v = og_wave_sum_d(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < kThreads / 64; ++k) t += red[k];
    return t;
}

__device__ double block_max(double v, double* red) {
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = red[0];
    for (int k = 1; k < kThreads / 64; ++k) t = fmax(t, red[k]);
    return t;
}

struct SysView {
    int base, n3;      // first degree of freedom, count
};

__device__ __forceinline__ SysView view(const og_lbfgs_args& a, int s) {
    const int off = a.atom_off[s];
    return {3 * off, 3 * (a.atom_off[s + 1] - off)};
}

__device__ __forceinline__ double* hist(double* h, const og_lbfgs_args& a, int slot, int base) {
    return h + (size_t)slot * 3 * (size_t)a.n_atoms + base;
}

}  // namespace

// e_prev = e_init = this evaluation's total per system; every system active.
__global__ void og_k_lb_init(og_lbfgs_args a) {
    for (int s = threadIdx.x; s < a.n_sys; s += blockDim.x) {
        og_lb_sys& st = a.sys[s];
        const double e = og_lb_total_energy(a.eacc + (size_t)s * OG_NE);
        st = og_lb_sys{};
        st.state = OG_LB_ACTIVE;
        st.e_init = e;
        st.e_prev = e;
        st.max_force = INFINITY;
        a.skip[s] = 1;
    }
}

__global__ void __launch_bounds__(kThreads) og_k_lb_begin(og_lbfgs_args a, int step) {
    __shared__ double red[kThreads / 64];
    __shared__ double alph[OG_LB_MAX_HISTORY];
    const int s = blockIdx.x;
    og_lb_sys& st = a.sys[s];
    if (st.state != OG_LB_ACTIVE) return;
    const SysView v = view(a, s);
    const int tid = threadIdx.x;
    double* g = a.g + v.base;
    double* p = a.p + v.base;
    const double* f = a.force + v.base;
    double mx = 0.0;
    for (int k = tid; k < v.n3; k += kThreads) {
        g[k] = -f[k];
        mx = fmax(mx, fabs(f[k]));
    }
    const double maxf = block_max(mx, red);
    if (tid == 0) {
        st.steps = step;
        st.max_force = maxf;
    }
    if (maxf < a.gradient_tol) {
        if (tid == 0) st.state = OG_LB_CONVERGED;
        return;
    }
    const int m = st.count;
    if (m == 0) {
        for (int k = tid; k < v.n3; k += kThreads) p[k] = -g[k];
    } else {
        // two_loop_recursion (lbfgs.rs:212-254), q and r in p; logical entry i
        // (0 = oldest) sits in ring slot (head + i) % history
        for (int k = tid; k < v.n3; k += kThreads) p[k] = g[k];
        double a_prev = 0.0;
        const double* y_prev = nullptr;
        for (int i = m - 1; i >= 0; --i) {
            const int slot = (st.head + i) % a.history;
            const double* si = hist(a.hs, a, slot, v.base);
            double acc = 0.0;
            for (int k = tid; k < v.n3; k += kThreads) {
                double q = p[k];
                if (y_prev) q -= a_prev * y_prev[k];
                p[k] = q;
                acc += si[k] * q;
            }
            const double ai = a.rho[s * OG_LB_MAX_HISTORY + slot] * block_sum(acc, red);
            if (tid == 0) alph[i] = ai;
            a_prev = ai;
            y_prev = hist(a.hy, a, slot, v.base);
        }
        const int last = (st.head + m - 1) % a.history;
        const double* sl = hist(a.hs, a, last, v.base);
        const double* yl = hist(a.hy, a, last, v.base);
        double sy = 0.0, yy = 0.0;
        for (int k = tid; k < v.n3; k += kThreads) {
            const double q = p[k] - a_prev * y_prev[k];
            p[k] = q;
            sy += sl[k] * yl[k];
            yy += yl[k] * yl[k];
        }
        sy = block_sum(sy, red);
        yy = block_sum(yy, red);
        const double gamma = yy > 0.0 ? sy / yy : 1.0;
        __syncthreads();
        double coef_prev = 0.0;
        const double* s_prev = nullptr;
        for (int i = 0; i < m; ++i) {
            const int slot = (st.head + i) % a.history;
            const double* yi = hist(a.hy, a, slot, v.base);
            double acc = 0.0;
            for (int k = tid; k < v.n3; k += kThreads) {
                double r = (i == 0) ? gamma * p[k] : p[k] + coef_prev * s_prev[k];
                p[k] = r;
                acc += yi[k] * r;
            }
            const double beta = a.rho[s * OG_LB_MAX_HISTORY + slot] * block_sum(acc, red);
            coef_prev = alph[i] - beta;
            s_prev = hist(a.hs, a, slot, v.base);
        }
        for (int k = tid; k < v.n3; k += kThreads) p[k] = -(p[k] + coef_prev * s_prev[k]);
    }
    __syncthreads();
    double mp = 0.0, gp = 0.0;
    for (int k = tid; k < v.n3; k += kThreads) {
        mp = fmax(mp, fabs(p[k]));
        gp += p[k] * g[k];
        a.x0[v.base + k] = a.pos[v.base + k];
    }
    const double maxp = block_max(mp, red);
    const double pdg = block_sum(gp, red);
    if (tid == 0) {
        st.alpha = maxp > 0.0 ? fmin(a.alpha0, a.max_step / maxp) : a.alpha0;
        st.gdotp = pdg;
        st.e_now = st.e_prev;
        st.phase = 0;
        st.state = pdg >= 0.0 ? OG_LB_FAILED : OG_LB_SEARCHING;   // not a descent direction
    }
}

// trial point of every searching system
__global__ void __launch_bounds__(kThreads) og_k_lb_trial(og_lbfgs_args a) {
    const int s = blockIdx.x;
    const og_lb_sys& st = a.sys[s];
    if (st.state != OG_LB_SEARCHING) return;
    const SysView v = view(a, s);
    const double al = st.alpha;
    for (int k = threadIdx.x; k < v.n3; k += kThreads)
        a.pos[v.base + k] = a.x0[v.base + k] + al * a.p[v.base + k];
}

// Armijo test of the energies just evaluated (one block); counters[0] =
// systems still searching, [1] = failed, [2] = active, [3] = accepted.
__global__ void __launch_bounds__(kThreads) og_k_lb_accept(og_lbfgs_args a, int test) {
    __shared__ double red[kThreads / 64];
    int n_search = 0, n_fail = 0, n_active = 0, n_acc = 0;
    for (int s = threadIdx.x; s < a.n_sys; s += blockDim.x) {
        og_lb_sys& st = a.sys[s];
        if (test && st.state == OG_LB_SEARCHING) {
            const double e = og_lb_total_energy(a.eacc + (size_t)s * OG_NE);
            if (e <= st.e_now + a.c1 * st.alpha * st.gdotp) {
                st.state = OG_LB_ACCEPTED;
                st.e_new = e;
            } else {
                st.alpha *= a.contraction;
                if (!(st.alpha >= a.min_alpha)) st.state = OG_LB_FAILED;
            }
        }
        a.skip[s] = st.state == OG_LB_SEARCHING ? 0 : 1;
        n_search += st.state == OG_LB_SEARCHING;
        n_fail += st.state == OG_LB_FAILED;
        n_active += st.state == OG_LB_ACTIVE;
        n_acc += st.state == OG_LB_ACCEPTED;
    }
    const double c0 = block_sum((double)n_search, red);
    const double c1 = block_sum((double)n_fail, red);
    const double c2 = block_sum((double)n_active, red);
    const double c3 = block_sum((double)n_acc, red);
    if (threadIdx.x == 0) {
        a.counters[0] = (int)c0;
        a.counters[1] = (int)c1;
        a.counters[2] = (int)c2;
        a.counters[3] = (int)c3;
    }
}

// A failed L-BFGS search retries along steepest descent (lbfgs.rs:128-160);
// a failed retry stops the system at x0 (lbfgs.rs:146-149).
__global__ void __launch_bounds__(kThreads) og_k_lb_fallback(og_lbfgs_args a) {
    __shared__ double red[kThreads / 64];
    const int s = blockIdx.x;
    og_lb_sys& st = a.sys[s];
    if (st.state != OG_LB_FAILED) return;
    const SysView v = view(a, s);
    const int tid = threadIdx.x;
    for (int k = tid; k < v.n3; k += kThreads) a.pos[v.base + k] = a.x0[v.base + k];
    if (st.phase == 1) {
        __syncthreads();
        if (tid == 0) {
            st.failures += 1;
            st.state = OG_LB_STOPPED;
        }
        return;
    }
    double mp = 0.0, gp = 0.0;
    for (int k = tid; k < v.n3; k += kThreads) {
        const double d = -a.g[v.base + k];
        a.p[v.base + k] = d;
        mp = fmax(mp, fabs(d));
        gp += d * a.g[v.base + k];
    }
    const double maxp = block_max(mp, red);
    const double pdg = block_sum(gp, red);
    if (tid == 0) {
        st.failures += 1;
        st.phase = 1;
        st.alpha = maxp > 0.0 ? fmin(a.alpha0, a.max_step / maxp) : a.alpha0;
        st.gdotp = pdg;
        if (pdg >= 0.0) {
            st.failures += 1;
            st.state = OG_LB_STOPPED;
        } else {
            st.state = OG_LB_SEARCHING;
        }
    }
}

// After the forces at the accepted points: history update and convergence
// test (lbfgs.rs:170-208), or the history reset of a steepest-descent step.
__global__ void __launch_bounds__(kThreads) og_k_lb_update(og_lbfgs_args a, int step) {
    __shared__ double red[kThreads / 64];
    const int s = blockIdx.x;
    og_lb_sys& st = a.sys[s];
    if (st.state != OG_LB_ACCEPTED) return;
    const int tid = threadIdx.x;
    if (st.phase == 1) {
        __syncthreads();
        if (tid == 0) {
            st.count = 0;
            st.head = 0;
            st.e_prev = st.e_new;
            st.phase = 0;
            st.state = OG_LB_ACTIVE;
        }
        return;
    }
    const SysView v = view(a, s);
    // slot the pair would take
    const int slot = st.count < a.history ? (st.head + st.count) % a.history : st.head;
    double* hs = hist(a.hs, a, slot, v.base);
    double* hy = hist(a.hy, a, slot, v.base);
    // s and y go to scratch first: the slot may hold the oldest pair, which
    // must survive when s.y fails the curvature test
    double* ts = a.x0 + v.base;            // x0 is not needed after the step
    double sy = 0.0;
    const double al = st.alpha;
    for (int k = tid; k < v.n3; k += kThreads) {
        const double sk = al * a.p[v.base + k];
        const double yk = -a.force[v.base + k] - a.g[v.base + k];
        ts[k] = sk;
        a.g[v.base + k] = yk;              // g is rebuilt from the forces at the next begin
        sy += sk * yk;
    }
    sy = block_sum(sy, red);
    if (sy > 1e-12) {
        for (int k = tid; k < v.n3; k += kThreads) {
            hs[k] = ts[k];
            hy[k] = a.g[v.base + k];
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (sy > 1e-12) {
            a.rho[s * OG_LB_MAX_HISTORY + slot] = 1.0 / sy;
            if (st.count < a.history) st.count += 1;
            else st.head = (st.head + 1) % a.history;
        }
        const double e_new = st.e_new;
        if (fabs(st.e_prev - e_new) < a.energy_tol && st.max_force < a.gradient_tol) {
            st.state = OG_LB_CONVERGED;
            st.steps = step + 1;
        } else {
            st.state = OG_LB_ACTIVE;
        }
        st.e_prev = e_new;
    }
}

void og_launch_lbfgs(int phase, const og_lbfgs_args& a, int step, hipStream_t s) {
    switch (phase) {
    case OG_LB_K_INIT:
        hipLaunchKernelGGL(og_k_lb_init, dim3(1), dim3(kThreads), 0, s, a);
        break;
    case OG_LB_K_BEGIN:
        hipLaunchKernelGGL(og_k_lb_begin, dim3(a.n_sys), dim3(kThreads), 0, s, a, step);
        break;
    case OG_LB_K_TRIAL:
        hipLaunchKernelGGL(og_k_lb_trial, dim3(a.n_sys), dim3(kThreads), 0, s, a);
        break;
    case OG_LB_K_ACCEPT:
        hipLaunchKernelGGL(og_k_lb_accept, dim3(1), dim3(kThreads), 0, s, a, 1);
        break;
    case OG_LB_K_COUNT:
        hipLaunchKernelGGL(og_k_lb_accept, dim3(1), dim3(kThreads), 0, s, a, 0);
        break;
    case OG_LB_K_FALLBACK:
        hipLaunchKernelGGL(og_k_lb_fallback, dim3(a.n_sys), dim3(kThreads), 0, s, a);
        break;
    case OG_LB_K_UPDATE:
        hipLaunchKernelGGL(og_k_lb_update, dim3(a.n_sys), dim3(kThreads), 0, s, a, step);
        break;
    default:
        break;
    }
}
