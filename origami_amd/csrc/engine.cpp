// origami MI355X engine: host side of the C ABI (include/origami_gpu.h).
//
// Owns the device-resident batch, builds the tile-pair work list, the
// exclusion-mask tiles and the SHAKE clusters once at create time
// (origami's ForceScratch::new, energy/src/scratch.rs:211-262), and drives
// the per-step kernel sequence on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/origami_gpu.h"
#include "og_device.h"
#include "og_kernels.h"
#include "og_sasa.h"
#include "og_sur.h"

// og_run's default: captured steps below this many atoms
#define OG_SPLIT_MAX_TILES 4           // systems up to 256 atoms: off-diagonal tile pairs in halves
#define OG_STRIP_DEFAULT 4             // j tiles per force strip until the autotuner picks
#define OG_SR_DOTS 960                 // Shrake-Rupley dots per atom (SHRAKE_RUPLEY_DOTS, sasa.rs:27)
#define OG_PERSIST_MAX_TILES 8         // persistent replica steps: systems of up to 512 atoms
#define OG_PERSIST_CHUNK 512           // steps per persistent launch

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t _e = (x);                                                                   \
        if (_e != hipSuccess)                                                                  \
            return fail(OG_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(_e),        \
                        __FILE__, __LINE__);                                                   \
    } while (0)

template <class T>
static hipError_t upload(T** dst, const std::vector<T>& src) {
    size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)dst, bytes);
    if (e != hipSuccess) return e;
    if (!src.empty()) e = hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

struct og_engine {
    int n_sys = 0, n_atoms = 0, n_tiles = 0, n_work = 0, n_masked = 0, n_clusters = 0, n_clusters_small = 0;
    int n_tile_pairs = 0;                  // n_work counts a split tile pair twice
    long long n_pairs = 0;
    int n_sasa_work = 0;
    std::vector<int> atom_off, sys_n;
    std::vector<double> temps, dofcorr;
    std::vector<int> n_cons_sys;
    hipStream_t stream = nullptr;
    // SASA runs on its own stream beside the all-pairs chain (fork after the
    // positions are final, join before the kick): the arrangement kernel is
    // latency bound and the pair kernels fill its idle issue slots
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // device buffers
    double *pos = nullptr, *vel = nullptr, *force = nullptr, *ref = nullptr;
    long long *facc = nullptr, *iacc = nullptr, *eacc = nullptr, *ke = nullptr;
    double *mass = nullptr, *rho = nullptr, *charge = nullptr, *radii = nullptr;
    double *sasa_r = nullptr, *sasa_g = nullptr;
    double *eps = nullptr, *rmin = nullptr, *eps14 = nullptr, *rmin14 = nullptr;
    int *atom_sys = nullptr, *atom_local = nullptr, *atom_slot = nullptr;
    float4 *par = nullptr, *gb = nullptr;
    float* qs = nullptr;
    int *tile_atom0 = nullptr, *tile_cnt = nullptr;
    og_work* work = nullptr;
    og_strip* strips = nullptr;            // force pass: strips of large systems
    og_work* fwork = nullptr;              // force pass: tile pairs outside the strips
    int n_strip = 0, n_fwork = 0;
    int strip_j = 0;                       // j tiles per strip
    std::vector<og_work> work_h;           // host copy of the tile-pair list (strip rebuilds)
    unsigned long long* masks = nullptr;
    int nb = 0, na = 0, nd = 0, ni = 0;
    int *bonds = nullptr, *angles = nullptr, *dih = nullptr, *imp = nullptr;
    double *bond_p = nullptr, *angle_p = nullptr, *dih_p = nullptr, *imp_p = nullptr;
    double* bonded_tf = nullptr;                    // per-(term, atom) force slots
    int *bonded_csr_start = nullptr, *bonded_csr_slot = nullptr;
    int *cl_start = nullptr, *cl_atoms = nullptr, *cn_start = nullptr;
    double* cn_len = nullptr;
    double *sys_temp = nullptr, *sys_dofc = nullptr;
    int *shake_fail = nullptr, *diverged = nullptr, *clamped = nullptr;
    int* sflag = nullptr;                  // [S] status bits of the current evaluation / run (OG_ST_*)
    std::vector<uint64_t> topo_hash;       // per system: parameters and term lists (swap check)
    og_sasa_data sasa{};
    og_sur* sur = nullptr;                 // learned Born-radius surrogate (optional)
    uint32_t* step_dev = nullptr;          // BAOAB step counter (Philox counter word)
    double* noise = nullptr;               // [2][3N] the O step's normals, step s in half s & 1
    int graph_mode = 2;                    // og_run: 0 direct launches, 1 captured steps, 3 persistent
                                           // replica steps, 2 the engine's choice (direct launches)
    bool graph_ok = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};   // step; step with surrogate check
    long long* noalloc_ke = nullptr;       // this device's og_run_noalloc scratch array
    int* eval_skip = nullptr;              // [S] systems an energy-only evaluation leaves out (minimiser probes)
    double* sr_dots = nullptr;             // [OG_SR_DOTS][3] Shrake-Rupley unit-sphere dots
    // persistent replica steps: per-system index ranges (og_persist_args)
    int *p_atom_off = nullptr, *p_cl_small_off = nullptr, *p_cl_small = nullptr, *p_cl_big_off = nullptr,
        *p_cl_big = nullptr, *p_term_off = nullptr, *p_work_off = nullptr, *p_fwork_off = nullptr,
        *p_strip_off = nullptr;
    int max_sys_tiles = 0;
    og_tunnel_dev tun{};                   // external tunnel of the dynamics (og_set_tunnel)
    // frame ring (og_frames_setup): every system's positions after every
    // frame_every-th step, kept on the device until og_frames_read
    double* frames = nullptr;              // [frame_cap][3N]
    int frame_cap = 0, frame_every = 0;
    long long frame_head = 0, frame_tail = 0;   // frames written / read
    std::vector<uint32_t> frame_step;      // step number held by each ring slot
    int pair_grid = 0;                     // autotuned pair-kernel width (0: default)
    float ds_far = 0.0f;                   // descreening far-field distance (og_pair_args::ds_far)
    std::vector<double> mass_h;            // host copies kept for the surrogate set-up
    std::vector<int> sys_tile_off_h;
    // dynamics
    bool dyn_ready = false;
    og_dynamics_params dyn{};
    double alpha = 1.0;
    uint32_t step = 0;
    // profiling
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double prof_ms[OG_N_KCLASS] = {0};
    long long prof_cnt[OG_N_KCLASS] = {0};
};

// ---------------------------------------------------------------------------
// profiling helpers
// ---------------------------------------------------------------------------
static hipEvent_t take_event(og_engine* e) {
    if (!e->ev_pool.empty()) {
        hipEvent_t ev = e->ev_pool.back();
        e->ev_pool.pop_back();
        return ev;
    }
    hipEvent_t ev;
    hipEventCreate(&ev);
    return ev;
}

struct ProfScope {
    og_engine* e;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(og_engine* e_, int c) : e(e_), cls(c) {
        if (e->prof) {
            a = take_event(e);
            b = take_event(e);
            hipEventRecord(a, e->stream);
        }
    }
    ~ProfScope() {
        if (e->prof) {
            hipEventRecord(b, e->stream);
            e->ev_pending.push_back({cls, {a, b}});
        }
    }
};

static void prof_collect(og_engine* e) {
    for (auto& p : e->ev_pending) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, p.second.first, p.second.second);
        e->prof_ms[p.first] += ms;
        e->prof_cnt[p.first] += 1;
        e->ev_pool.push_back(p.second.first);
        e->ev_pool.push_back(p.second.second);
    }
    e->ev_pending.clear();
}

// ---------------------------------------------------------------------------
// create / destroy
// ---------------------------------------------------------------------------
static void free_all(og_engine* e) {
    void* ptrs[] = {e->pos, e->vel, e->force, e->ref, e->facc, e->iacc, e->eacc, e->ke, e->mass,
                    e->rho, e->charge, e->radii, e->sasa_r, e->sasa_g, e->eps, e->rmin, e->eps14,
                    e->rmin14, e->atom_sys, e->atom_local,
                    e->atom_slot, e->par, e->gb, e->qs, e->tile_atom0, e->tile_cnt, e->work,
                    e->strips, e->fwork, e->masks, e->bonds, e->angles, e->dih, e->imp, e->bond_p, e->angle_p,
                    e->bonded_tf, e->bonded_csr_start, e->bonded_csr_slot,
                    e->dih_p, e->imp_p, e->cl_start, e->cl_atoms, e->cn_start,
                    e->cn_len, e->sys_temp, e->sys_dofc, e->shake_fail, e->diverged, e->clamped,
                    e->sflag};
    for (void* p : ptrs)
        if (p) hipFree(p);
    for (auto& g : e->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    if (e->step_dev) (void)hipFree(e->step_dev);
    if (e->frames) (void)hipFree(e->frames);
    if (e->sr_dots) (void)hipFree(e->sr_dots);
    for (int* q : {e->p_atom_off, e->p_cl_small_off, e->p_cl_small, e->p_cl_big_off, e->p_cl_big, e->p_term_off,
                   e->p_work_off, e->p_fwork_off, e->p_strip_off})
        if (q) (void)hipFree(q);
    if (e->noise) (void)hipFree(e->noise);
    if (e->sur) {
        og_sur_free(e->sur);
        delete e->sur;
        e->sur = nullptr;
    }
    og_sasa_free(&e->sasa);
    for (auto ev : e->ev_pool) hipEventDestroy(ev);
    for (auto& p : e->ev_pending) {
        hipEventDestroy(p.second.first);
        hipEventDestroy(p.second.second);
    }
    if (e->fork_ev) hipEventDestroy(e->fork_ev);
    if (e->join_ev) hipEventDestroy(e->join_ev);
    if (e->side) hipStreamDestroy(e->side);
    if (e->stream) hipStreamDestroy(e->stream);
}

static int check_pairs(const int32_t* p, int n, int width, int n_atoms, const char* what, int s) {
    if (n < 0) return fail(OG_ERR_ARG, "system %d: negative %s count", s, what);
    if (n > 0 && !p) return fail(OG_ERR_ARG, "system %d: %s array is NULL", s, what);
    for (int k = 0; k < n * width; ++k)
        if (p[k] < 0 || p[k] >= n_atoms)
            return fail(OG_ERR_ARG, "system %d: %s index %d out of range [0,%d)", s, what, p[k], n_atoms);
    return OG_OK;
}

extern "C" const char* og_last_error(void) { return g_err.c_str(); }

static int engine_build(og_engine* e, const og_system_desc* sys, int n_sys);

static int noalloc_setup(og_engine* e);

// The force pass's work lists for a strip length of strip_j j tiles.  Systems
// above OG_SPLIT_MAX_TILES tiles: tiles in pairs (I, I + 1) for even I; the
// pair's tile pairs with every later tile J go into strips of up to strip_j j
// tiles, except where either tile pair holds an excluded or 1-4 pair; the
// diagonal and adjacent tile pairs and those masked ones stay single work
// items.  Smaller systems keep their (split) tile pairs.  Every tile pair of
// a system is covered exactly once (checked).
static int build_force_lists(og_engine* e, int strip_j) {
    if (strip_j < 1 || strip_j > OG_STRIP_J) return fail(OG_ERR_ARG, "strip length must be 1..%d", OG_STRIP_J);
    std::vector<og_work> fwork;
    std::vector<og_strip> strips;
    std::vector<int> fwork_off{0}, strip_off{0};
    const std::vector<og_work>& work = e->work_h;
    size_t w = 0;
    for (int s = 0; s < e->n_sys; ++s) {
        const int t0 = e->sys_tile_off_h[s], nt = e->sys_tile_off_h[s + 1] - t0;
        const size_t w0 = w;
        while (w < work.size() && work[w].sys == s) ++w;
        if (nt <= OG_SPLIT_MAX_TILES) {
            fwork.insert(fwork.end(), work.begin() + w0, work.begin() + w);
            fwork_off.push_back((int)fwork.size());
            strip_off.push_back((int)strips.size());
            continue;
        }
        if ((long long)(w - w0) != (long long)nt * (nt + 1) / 2)
            return fail(OG_ERR_STATE, "internal: system %d holds %zu tile pairs", s, w - w0);
        auto tpi = [&](int I, int J) { return w0 + (size_t)I * nt - (size_t)I * (I - 1) / 2 + (J - I); };
        const size_t f0 = fwork.size(), s0 = strips.size();
        auto pushw = [&](int I, int J) { fwork.push_back(work[tpi(I, J)]); };
        for (int I = 0; I < nt; I += 2) {
            pushw(I, I);
            if (I + 1 >= nt) break;
            pushw(I + 1, I + 1);
            pushw(I, I + 1);
            og_strip cur{};
            cur.ti = t0 + I;
            cur.sys = s;
            for (int J = I + 2; J < nt; ++J) {
                if (work[tpi(I, J)].mask >= 0 || work[tpi(I + 1, J)].mask >= 0) {
                    pushw(I, J);
                    pushw(I + 1, J);
                    continue;
                }
                cur.tj[cur.nj++] = t0 + J;
                if (cur.nj == strip_j) {
                    strips.push_back(cur);
                    cur.nj = 0;
                }
            }
            if (cur.nj > 0) strips.push_back(cur);
        }
        std::vector<int> cov((size_t)nt * nt, 0);
        for (size_t k = f0; k < fwork.size(); ++k) cov[(size_t)(fwork[k].ti - t0) * nt + (fwork[k].tj - t0)]++;
        for (size_t k = s0; k < strips.size(); ++k)
            for (int q = 0; q < strips[k].nj; ++q)
                for (int h = 0; h < 2; ++h) cov[(size_t)(strips[k].ti + h - t0) * nt + (strips[k].tj[q] - t0)]++;
        for (int I = 0; I < nt; ++I)
            for (int J = 0; J < nt; ++J)
                if (cov[(size_t)I * nt + J] != (J >= I ? 1 : 0))
                    return fail(OG_ERR_STATE, "internal: force work list covers tile pair (%d, %d) of system %d %d times",
                                I, J, s, cov[(size_t)I * nt + J]);
        fwork_off.push_back((int)fwork.size());
        strip_off.push_back((int)strips.size());
    }
    if (w != work.size()) return fail(OG_ERR_STATE, "internal: tile-pair list out of system order");
    if (e->strips) { HIPCHK(hipFree(e->strips)); e->strips = nullptr; }
    if (e->fwork) { HIPCHK(hipFree(e->fwork)); e->fwork = nullptr; }
    if (e->p_fwork_off) { HIPCHK(hipFree(e->p_fwork_off)); e->p_fwork_off = nullptr; }
    if (e->p_strip_off) { HIPCHK(hipFree(e->p_strip_off)); e->p_strip_off = nullptr; }
    HIPCHK(upload(&e->strips, strips));
    HIPCHK(upload(&e->fwork, fwork));
    HIPCHK(upload(&e->p_fwork_off, fwork_off));
    HIPCHK(upload(&e->p_strip_off, strip_off));
    e->n_strip = (int)strips.size();
    e->n_fwork = (int)fwork.size();
    e->strip_j = strip_j;
    return OG_OK;
}

extern "C" int og_engine_create(const og_system_desc* systems, int32_t n_systems, og_engine** out) {
    if (!out) return fail(OG_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!systems || n_systems <= 0) return fail(OG_ERR_ARG, "need at least one system");
    for (int s = 0; s < n_systems; ++s) {
        const og_system_desc& d = systems[s];
        if (d.n_atoms <= 0) return fail(OG_ERR_ARG, "system %d has no atoms", s);
        const double* req[] = {d.pos, d.charge, d.rmin_half, d.epsilon, d.rmin_half_14, d.epsilon_14,
                               d.gb_rho, d.gb_scale, d.sasa_radius, d.sasa_gamma, d.mass};
        for (const double* p : req)
            if (!p) return fail(OG_ERR_ARG, "system %d: a per-atom array is NULL", s);
        int rc;
        if ((rc = check_pairs(d.excl, d.n_excl, 2, d.n_atoms, "exclusion", s))) return rc;
        if ((rc = check_pairs(d.pair14, d.n_14, 2, d.n_atoms, "1-4", s))) return rc;
        if ((rc = check_pairs(d.bonds, d.n_bonds, 2, d.n_atoms, "bond", s))) return rc;
        if ((rc = check_pairs(d.angles, d.n_angles, 3, d.n_atoms, "angle", s))) return rc;
        if ((rc = check_pairs(d.dihedrals, d.n_dihedrals, 4, d.n_atoms, "dihedral", s))) return rc;
        if ((rc = check_pairs(d.impropers, d.n_impropers, 4, d.n_atoms, "improper", s))) return rc;
        if ((rc = check_pairs(d.constraints, d.n_constraints, 2, d.n_atoms, "constraint", s))) return rc;
        for (int k = 0; k < d.n_atoms; ++k) {
            if (!(d.mass[k] > 0.0)) return fail(OG_ERR_ARG, "system %d atom %d: mass must be > 0", s, k);
            if (!(d.gb_rho[k] - OG_OBC_OFFSET > 0.0))
                return fail(OG_ERR_ARG, "system %d atom %d: gb_rho must exceed 0.09 Å", s, k);
            for (int c = 0; c < 3; ++c)
                if (!std::isfinite(d.pos[3 * k + c]))
                    return fail(OG_ERR_ARG, "system %d atom %d: non-finite position", s, k);
        }
        for (int k = 0; k < d.n_excl; ++k)
            if (d.excl[2 * k] >= d.excl[2 * k + 1])
                return fail(OG_ERR_ARG, "system %d: exclusion pairs must have i < j", s);
        for (int k = 0; k < d.n_14; ++k)
            if (d.pair14[2 * k] >= d.pair14[2 * k + 1])
                return fail(OG_ERR_ARG, "system %d: 1-4 pairs must have i < j", s);
    }
    og_engine* e = new og_engine();
    int rc = engine_build(e, systems, n_systems);
    if (rc == OG_OK) rc = noalloc_setup(e);
    if (rc != OG_OK) {
        std::string msg = g_err;
        free_all(e);
        delete e;
        g_err = msg;
        return rc;
    }
    *out = e;
    return OG_OK;
}

extern "C" void og_engine_destroy(og_engine* e) {
    if (!e) return;
    if (e->stream) hipStreamSynchronize(e->stream);
    if (e->side) hipStreamSynchronize(e->side);
    free_all(e);
    delete e;
}

// FNV-1a over everything that defines a system's energy function except the
// positions: per-atom parameters, exclusion and 1-4 lists, bonded terms with
// their parameters and the constraints.  Two systems may exchange
// configurations (REMD) only when these agree.
static uint64_t topology_hash(const og_system_desc& d) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t bytes) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t k = 0; k < bytes; ++k) {
            h ^= c[k];
            h *= 1099511628211ull;
        }
    };
    const size_t n = (size_t)d.n_atoms;
    mix(&d.n_atoms, sizeof d.n_atoms);
    const double* per_atom[] = {d.charge, d.rmin_half, d.epsilon, d.rmin_half_14, d.epsilon_14, d.gb_rho,
                                d.gb_scale, d.sasa_radius, d.sasa_gamma, d.mass};
    for (const double* p : per_atom) mix(p, n * sizeof(double));
    auto lst = [&](int32_t cnt, const int32_t* idx, int w, const double* par, int pw) {
        mix(&cnt, sizeof cnt);
        if (cnt > 0 && idx) mix(idx, (size_t)cnt * w * sizeof(int32_t));
        if (cnt > 0 && par) mix(par, (size_t)cnt * pw * sizeof(double));
    };
    lst(d.n_excl, d.excl, 2, nullptr, 0);
    lst(d.n_14, d.pair14, 2, nullptr, 0);
    lst(d.n_bonds, d.bonds, 2, d.bond_params, 2);
    lst(d.n_angles, d.angles, 3, d.angle_params, 2);
    lst(d.n_dihedrals, d.dihedrals, 4, d.dihedral_params, 3);
    lst(d.n_impropers, d.impropers, 4, d.improper_params, 2);
    lst(d.n_constraints, d.constraints, 2, d.constraint_len, 1);
    return h;
}

static int engine_build(og_engine* e, const og_system_desc* sys, int n_sys) {
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&e->fork_ev, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&e->join_ev, hipEventDisableTiming));
    e->n_sys = n_sys;
    e->atom_off.resize(n_sys + 1, 0);
    e->sys_n.resize(n_sys);
    for (int s = 0; s < n_sys; ++s) {
        e->sys_n[s] = sys[s].n_atoms;
        e->atom_off[s + 1] = e->atom_off[s] + sys[s].n_atoms;
    }
    const int N = e->atom_off[n_sys];
    e->n_atoms = N;

    std::vector<double> pos(3 * (size_t)N), mass(N), rho(N), charge(N), sr(N), sg(N);
    std::vector<double> epsv(N), rminv(N), eps14v(N), rmin14v(N);
    std::vector<int> a_sys(N), a_loc(N), a_slot(N);
    std::vector<int> tile_atom0, tile_cnt, sys_tile_off{0};
    std::vector<og_work> work;
    std::vector<int> term_cnt(4 * (size_t)n_sys, 0), sys_work_end;
    std::vector<unsigned long long> masks;
    std::vector<float4> par, gbv;
    std::vector<float> qs;
    std::vector<int> bonds, angles, dih, imp;
    std::vector<double> bond_p, angle_p, dih_p, imp_p;
    std::vector<int> cl_start{0}, cl_atoms, cn_start{0};
    std::vector<double> cn_len;
    e->n_cons_sys.resize(n_sys);
    e->topo_hash.assign(n_sys, 0);

    const double sqrt_lj = std::sqrt(12.0 * OG_KCAL_TO_KJ);
    const double sqrt_kc = std::sqrt(OG_COULOMB_KCAL * OG_KCAL_TO_KJ);
    for (int s = 0; s < n_sys; ++s) {
        const og_system_desc& d = sys[s];
        term_cnt[0 * (size_t)n_sys + s] = d.n_bonds;
        term_cnt[1 * (size_t)n_sys + s] = d.n_angles;
        term_cnt[2 * (size_t)n_sys + s] = d.n_dihedrals;
        term_cnt[3 * (size_t)n_sys + s] = d.n_impropers;
        const int n = d.n_atoms, off = e->atom_off[s];
        e->topo_hash[s] = topology_hash(d);
        const int t0 = (int)tile_atom0.size();
        const int nt = (n + OG_TILE - 1) / OG_TILE;
        e->n_pairs += (long long)n * (n - 1) / 2;
        for (int t = 0; t < nt; ++t) {
            tile_atom0.push_back(off + t * OG_TILE);
            tile_cnt.push_back(std::min(OG_TILE, n - t * OG_TILE));
        }
        sys_tile_off.push_back(t0 + nt);
        for (int k = 0; k < n; ++k) {
            const int g = off + k;
            for (int c = 0; c < 3; ++c) pos[3 * (size_t)g + c] = d.pos[3 * k + c];
            mass[g] = d.mass[k];
            rho[g] = d.gb_rho[k];
            charge[g] = d.charge[k];
            sr[g] = d.sasa_radius[k];
            sg[g] = d.sasa_gamma[k];
            epsv[g] = d.epsilon[k];
            rminv[g] = d.rmin_half[k];
            eps14v[g] = d.epsilon_14[k];
            rmin14v[g] = d.rmin_half_14[k];
            a_sys[g] = s;
            a_loc[g] = k;
            a_slot[g] = (t0 + k / OG_TILE) * OG_TILE + (k % OG_TILE);
            if (d.sasa_gamma[k] != 0.0) e->n_sasa_work++;
        }
        par.resize((size_t)(t0 + nt) * OG_TILE, make_float4(0.f, 0.f, 0.f, 0.f));
        gbv.resize((size_t)(t0 + nt) * OG_TILE, make_float4(1.f, 0.f, 1.f, 1.f));
        qs.resize((size_t)(t0 + nt) * OG_TILE, 0.f);
        for (int k = 0; k < n; ++k) {
            const size_t slot = (size_t)a_slot[off + k];
            par[slot] = make_float4((float)(std::sqrt(d.epsilon[k]) * sqrt_lj), (float)d.rmin_half[k],
                                    (float)(std::sqrt(d.epsilon_14[k]) * sqrt_lj), (float)d.rmin_half_14[k]);
            const double rt = d.gb_rho[k] - OG_OBC_OFFSET;
            gbv[slot] = make_float4((float)rt, (float)(d.gb_scale[k] * rt), (float)rt,
                                    (float)(0.60056120439322491 / rt));
            qs[slot] = (float)(d.charge[k] * sqrt_kc);
        }
        // Work list: every tile pair I <= J of this system.
        std::map<std::pair<int, int>, int> mask_of;
        const int w0 = (int)work.size();
        for (int I = 0; I < nt; ++I)
            for (int J = I; J < nt; ++J) work.push_back({t0 + I, t0 + J, -1, s, -1});
        auto tile_pair_index = [&](int I, int J) {  // I <= J local
            return w0 + I * nt - I * (I - 1) / 2 + (J - I);
        };
        // Lane l of a masked tile pair holds i atom l; bit m of its word marks
        // j atom m.  On a diagonal tile (I == J) the force kernel evaluates a
        // pair a < b on lane a when b - a <= 32 and on lane b (with j = a)
        // when b - a > 32, so both lanes carry the bit; the FP64 energy kernel
        // evaluates every pair on its lower lane and never reads the mirror.
        auto set_bit = [&](int a, int b, int which) {
            const int I = a / OG_TILE, J = b / OG_TILE;
            const int wi = tile_pair_index(I, J);
            if (work[wi].mask < 0) {
                work[wi].mask = (int)(masks.size() / 128);
                masks.resize(masks.size() + 128, 0ull);
            }
            unsigned long long* m = &masks[(size_t)work[wi].mask * 128 + which * 64];
            m[a % OG_TILE] |= 1ull << (b % OG_TILE);
            if (I == J) m[b % OG_TILE] |= 1ull << (a % OG_TILE);
        };
        for (int k = 0; k < d.n_excl; ++k) set_bit(d.excl[2 * k], d.excl[2 * k + 1], 0);
        for (int k = 0; k < d.n_14; ++k) set_bit(d.pair14[2 * k], d.pair14[2 * k + 1], 1);
        e->n_tile_pairs += (int)work.size() - w0;
        // Systems of at most OG_SPLIT_MAX_TILES tiles: every off-diagonal
        // tile pair becomes two work items of 32 steps each (the diagonal
        // ones already take 32), so that a batch of small systems, whose
        // pair kernels are bound by one wave's 64-step latency, runs on twice
        // the waves.  Decided per system, so a system's forces do not depend
        // on what else is in the batch.
        if (nt <= OG_SPLIT_MAX_TILES) {
            std::vector<og_work> seg(work.begin() + w0, work.end());
            work.resize(w0);
            for (const og_work& w : seg) {
                if (w.ti == w.tj) {
                    work.push_back(w);
                } else {
                    og_work h = w;
                    h.part = 0;
                    work.push_back(h);
                    h.part = 1;
                    work.push_back(h);
                }
            }
        }
        sys_work_end.push_back((int)work.size());
        e->max_sys_tiles = std::max(e->max_sys_tiles, nt);
        // Bonded terms (global atom indices).
        for (int k = 0; k < d.n_bonds; ++k) {
            bonds.push_back(off + d.bonds[2 * k]); bonds.push_back(off + d.bonds[2 * k + 1]);
            bond_p.push_back(d.bond_params[2 * k]); bond_p.push_back(d.bond_params[2 * k + 1]);
        }
        for (int k = 0; k < d.n_angles; ++k) {
            for (int c = 0; c < 3; ++c) angles.push_back(off + d.angles[3 * k + c]);
            angle_p.push_back(d.angle_params[2 * k]); angle_p.push_back(d.angle_params[2 * k + 1]);
        }
        for (int k = 0; k < d.n_dihedrals; ++k) {
            for (int c = 0; c < 4; ++c) dih.push_back(off + d.dihedrals[4 * k + c]);
            for (int c = 0; c < 3; ++c) dih_p.push_back(d.dihedral_params[3 * k + c]);
        }
        for (int k = 0; k < d.n_impropers; ++k) {
            for (int c = 0; c < 4; ++c) imp.push_back(off + d.impropers[4 * k + c]);
            imp_p.push_back(d.improper_params[2 * k]); imp_p.push_back(d.improper_params[2 * k + 1]);
        }
        // SHAKE clusters: each constraint joins its heavy atom's cluster; every
        // other atom is a singleton.  A cluster lists its heavy atom first,
        // then the hydrogen of each of its constraints, the constraints in the
        // caller's (origami's sorted bond) order.
        std::vector<int> owner(n, -1);  // heavy atom of an H, or itself
        std::vector<std::vector<int>> cons_of(n);
        for (int k = 0; k < d.n_constraints; ++k) {
            const int a = d.constraints[2 * k], b = d.constraints[2 * k + 1];
            // the hydrogen is the lighter atom
            const int h = d.mass[a] < d.mass[b] ? a : b;
            const int x = h == a ? b : a;
            if (owner[h] >= 0 && owner[h] != x)
                return fail(OG_ERR_ARG, "system %d: atom %d is constrained to two heavy atoms", s, h);
            if (owner[x] >= 0 && owner[x] != x)
                return fail(OG_ERR_ARG, "system %d: constraint graph is not X-H star shaped", s);
            owner[h] = x;
            owner[x] = x;
            cons_of[x].push_back(k);
        }
        e->n_cons_sys[s] = d.n_constraints;
        std::vector<std::vector<int>> members(n);
        for (int k = 0; k < n; ++k) {
            const int o = owner[k] < 0 ? k : owner[k];
            members[o].push_back(k);
        }
        for (int k = 0; k < n; ++k) {
            if (members[k].empty()) continue;
            if ((int)members[k].size() > 8)
                return fail(OG_ERR_ARG, "system %d: atom %d has more than 7 constrained partners", s, k);
            cl_atoms.push_back(off + k);
            for (int c : cons_of[k]) {
                const int a0 = d.constraints[2 * c], b0 = d.constraints[2 * c + 1];
                cl_atoms.push_back(off + (a0 == k ? b0 : a0));
                cn_len.push_back(d.constraint_len[c]);
            }
            if ((int)cl_atoms.size() - cl_start.back() != (int)members[k].size())
                return fail(OG_ERR_ARG, "system %d: atom %d carries a repeated constraint", s, k);
            cl_start.push_back((int)cl_atoms.size());
            cn_start.push_back((int)cn_len.size());
        }
    }
    e->n_tiles = (int)tile_atom0.size();
    e->n_work = (int)work.size();
    e->n_masked = (int)(masks.size() / 128);
    e->n_clusters = (int)cl_start.size() - 1;
    // clusters of at most OG_SMALL_CL atoms first (the narrow cluster kernel)
    {
        std::vector<int> order(e->n_clusters), s_start{0}, s_atoms, s_cstart{0};
        std::vector<double> s_len;
        for (int c = 0; c < e->n_clusters; ++c) order[c] = c;
        std::stable_partition(order.begin(), order.end(), [&](int c) {
            return cl_start[c + 1] - cl_start[c] <= OG_SMALL_CL;
        });
        e->n_clusters_small = 0;
        for (int c : order) {
            if (cl_start[c + 1] - cl_start[c] <= OG_SMALL_CL) e->n_clusters_small++;
            for (int q = cl_start[c]; q < cl_start[c + 1]; ++q) s_atoms.push_back(cl_atoms[q]);
            for (int q = cn_start[c]; q < cn_start[c + 1]; ++q) s_len.push_back(cn_len[q]);
            s_start.push_back((int)s_atoms.size());
            s_cstart.push_back((int)s_len.size());
        }
        cl_start.swap(s_start); cl_atoms.swap(s_atoms); cn_start.swap(s_cstart); cn_len.swap(s_len);
    }
    // persistent replica steps: each system's clusters, bonded terms and
    // descreening items as index ranges
    {
        std::vector<std::vector<int>> small(n_sys), big(n_sys);
        for (int c = 0; c < e->n_clusters; ++c) {
            const int sy = a_sys[cl_atoms[cl_start[c]]];
            (c < e->n_clusters_small ? small : big)[sy].push_back(c);
        }
        std::vector<int> so{0}, sl, bo{0}, bl;
        for (int sy = 0; sy < n_sys; ++sy) {
            sl.insert(sl.end(), small[sy].begin(), small[sy].end());
            so.push_back((int)sl.size());
            bl.insert(bl.end(), big[sy].begin(), big[sy].end());
            bo.push_back((int)bl.size());
        }
        std::vector<int> toff(4 * ((size_t)n_sys + 1), 0);
        for (int kind = 0; kind < 4; ++kind)
            for (int sy = 0; sy < n_sys; ++sy)
                toff[kind * ((size_t)n_sys + 1) + sy + 1] =
                    toff[kind * ((size_t)n_sys + 1) + sy] + term_cnt[kind * (size_t)n_sys + sy];
        std::vector<int> woff{0};
        woff.insert(woff.end(), sys_work_end.begin(), sys_work_end.end());
        std::vector<int> aoff(e->atom_off.begin(), e->atom_off.end());
        aoff.resize(n_sys);
        aoff.push_back(N);
        HIPCHK(upload(&e->p_atom_off, aoff));
        HIPCHK(upload(&e->p_cl_small_off, so));
        HIPCHK(upload(&e->p_cl_small, sl));
        HIPCHK(upload(&e->p_cl_big_off, bo));
        HIPCHK(upload(&e->p_cl_big, bl));
        HIPCHK(upload(&e->p_term_off, toff));
        HIPCHK(upload(&e->p_work_off, woff));
    }
    e->nb = (int)bond_p.size() / 2;
    e->na = (int)angle_p.size() / 2;
    e->nd = (int)dih_p.size() / 3;
    e->ni = (int)imp_p.size() / 2;
    // per-atom lists of bonded force slots (ascending slot order)
    {
        std::vector<int> slot_atom;
        slot_atom.reserve(2 * (size_t)e->nb + 3 * (size_t)e->na + 4 * (size_t)(e->nd + e->ni));
        slot_atom.insert(slot_atom.end(), bonds.begin(), bonds.end());
        slot_atom.insert(slot_atom.end(), angles.begin(), angles.end());
        slot_atom.insert(slot_atom.end(), dih.begin(), dih.end());
        slot_atom.insert(slot_atom.end(), imp.begin(), imp.end());
        std::vector<int> start(N + 1, 0), fillp(N, 0), slots(slot_atom.size());
        for (int a : slot_atom) start[a + 1]++;
        for (int i = 0; i < N; ++i) start[i + 1] += start[i];
        for (size_t q = 0; q < slot_atom.size(); ++q) {
            const int a = slot_atom[q];
            slots[start[a] + fillp[a]++] = (int)q;
        }
        HIPCHK(upload(&e->bonded_csr_start, start));
        HIPCHK(upload(&e->bonded_csr_slot, slots));
        HIPCHK(hipMalloc((void**)&e->bonded_tf, std::max<size_t>(slot_atom.size(), 1) * 3 * sizeof(double)));
    }
    e->temps.assign(n_sys, 0.0);
    e->dofcorr.assign(n_sys, 1.0);

    HIPCHK(upload(&e->pos, pos));
    std::vector<double> zeros3(3 * (size_t)N, 0.0);
    HIPCHK(upload(&e->vel, zeros3));
    HIPCHK(upload(&e->force, zeros3));
    HIPCHK(upload(&e->ref, zeros3));
    std::vector<long long> zl3(3 * (size_t)N, 0), zl(N, 0), ze((size_t)n_sys * OG_NE, 0), zs(n_sys, 0);
    HIPCHK(upload(&e->facc, zl3));
    HIPCHK(upload(&e->iacc, zl));
    HIPCHK(upload(&e->eacc, ze));
    HIPCHK(upload(&e->ke, zs));
    HIPCHK(upload(&e->mass, mass));
    HIPCHK(upload(&e->rho, rho));
    HIPCHK(upload(&e->charge, charge));
    std::vector<double> r0(N);
    for (int i = 0; i < N; ++i) r0[i] = rho[i] - OG_OBC_OFFSET;
    HIPCHK(upload(&e->radii, r0));
    HIPCHK(upload(&e->sasa_r, sr));
    HIPCHK(upload(&e->sasa_g, sg));
    HIPCHK(upload(&e->eps, epsv));
    HIPCHK(upload(&e->rmin, rminv));
    HIPCHK(upload(&e->eps14, eps14v));
    HIPCHK(upload(&e->rmin14, rmin14v));
    HIPCHK(upload(&e->atom_sys, a_sys));
    HIPCHK(upload(&e->atom_local, a_loc));
    HIPCHK(upload(&e->atom_slot, a_slot));
    HIPCHK(upload(&e->par, par));
    HIPCHK(upload(&e->gb, gbv));
    {
        // the series of the HCT term needs S / r <= 0.1 and the separated
        // branch (r - S >= rho~) for every pair of a far tile pair
        float s_max = 0.0f, rt_max = 0.0f;
        for (const float4& g : gbv) { s_max = std::max(s_max, g.y); rt_max = std::max(rt_max, g.x); }
        e->ds_far = std::max(10.0f * s_max, 1.25f * (rt_max + s_max));
    }
    HIPCHK(upload(&e->qs, qs));
    HIPCHK(upload(&e->tile_atom0, tile_atom0));
    HIPCHK(upload(&e->tile_cnt, tile_cnt));
    HIPCHK(upload(&e->work, work));
    HIPCHK(upload(&e->masks, masks));
    HIPCHK(upload(&e->bonds, bonds));
    HIPCHK(upload(&e->bond_p, bond_p));
    HIPCHK(upload(&e->angles, angles));
    HIPCHK(upload(&e->angle_p, angle_p));
    HIPCHK(upload(&e->dih, dih));
    HIPCHK(upload(&e->dih_p, dih_p));
    HIPCHK(upload(&e->imp, imp));
    HIPCHK(upload(&e->imp_p, imp_p));
    HIPCHK(upload(&e->cl_start, cl_start));
    HIPCHK(upload(&e->cl_atoms, cl_atoms));
    HIPCHK(upload(&e->cn_start, cn_start));
    HIPCHK(upload(&e->cn_len, cn_len));
    std::vector<double> z_sys(n_sys, 0.0), o_sys(n_sys, 1.0);
    HIPCHK(upload(&e->sys_temp, z_sys));
    HIPCHK(upload(&e->sys_dofc, o_sys));
    std::vector<int> zi(n_sys, 0);
    HIPCHK(upload(&e->shake_fail, zi));
    HIPCHK(upload(&e->diverged, zi));
    HIPCHK(upload(&e->sflag, zi));
    HIPCHK(hipMalloc((void**)&e->step_dev, sizeof(uint32_t)));
    HIPCHK(hipMemset(e->step_dev, 0, sizeof(uint32_t)));
    HIPCHK(hipMalloc((void**)&e->noise, 6 * (size_t)std::max(N, 1) * sizeof(double)));
    HIPCHK(upload(&e->clamped, zi));
    e->mass_h = mass;
    e->sys_tile_off_h = sys_tile_off;
    e->work_h = work;
    if (int rc = build_force_lists(e, OG_STRIP_DEFAULT)) return rc;
    int rc = og_sasa_build(&e->sasa, e->n_atoms, a_sys, e->atom_off, sr, sg, sys_tile_off,
                           e->tile_atom0, e->tile_cnt, e->n_tiles);
    if (rc != OG_OK) return fail(rc, "SASA set-up failed: %s", og_sasa_error());
    // Shrake-Rupley dots: the Fibonacci spiral of sasa.rs:182-199
    {
        std::vector<double> dots(3 * OG_SR_DOTS);
        const double phi = M_PI * (std::sqrt(3.0) + 1.0);
        for (int i = 0; i < OG_SR_DOTS; ++i) {
            const double y = 1.0 - 2.0 * (double)i / ((double)OG_SR_DOTS - 1.0);
            const double r = std::sqrt(1.0 - y * y);
            const double th = phi * (double)i;
            dots[3 * i] = std::cos(th) * r;
            dots[3 * i + 1] = y;
            dots[3 * i + 2] = std::sin(th) * r;
        }
        HIPCHK(upload(&e->sr_dots, dots));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipDeviceSynchronize());
    return OG_OK;
}

extern "C" int og_engine_info(const og_engine* e, int64_t* info) {
    if (!e || !info) return fail(OG_ERR_ARG, "NULL argument");
    info[0] = e->n_atoms;
    info[1] = e->n_sys;
    info[2] = e->n_tiles;
    info[3] = e->n_tile_pairs;
    info[4] = e->n_masked;
    info[5] = e->n_clusters;
    info[6] = e->n_pairs;
    info[7] = e->n_sasa_work;
    return OG_OK;
}

static int copy_state(og_engine* e, double* dev, const double* host, bool to_dev) {
    if (!e || !host) return fail(OG_ERR_ARG, "NULL argument");
    const size_t bytes = 3 * (size_t)e->n_atoms * sizeof(double);
    if (to_dev) {
        for (size_t k = 0; k < 3 * (size_t)e->n_atoms; ++k)
            if (!std::isfinite(host[k])) return fail(OG_ERR_ARG, "non-finite input at %zu", k);
        HIPCHK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, e->stream));
    } else {
        HIPCHK(hipMemcpyAsync((void*)host, dev, bytes, hipMemcpyDeviceToHost, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    return OG_OK;
}

extern "C" int og_set_positions(og_engine* e, const double* x) { return copy_state(e, e ? e->pos : nullptr, x, true); }
extern "C" int og_get_positions(og_engine* e, double* x) { return copy_state(e, e ? e->pos : nullptr, x, false); }
extern "C" int og_set_velocities(og_engine* e, const double* v) { return copy_state(e, e ? e->vel : nullptr, v, true); }
extern "C" int og_get_velocities(og_engine* e, double* v) { return copy_state(e, e ? e->vel : nullptr, v, false); }
extern "C" int og_get_forces(og_engine* e, double* f) { return copy_state(e, e ? e->force : nullptr, f, false); }

// ---------------------------------------------------------------------------
// Force pipeline (accumulators must be zero on entry).
// ---------------------------------------------------------------------------
// origami's integrator cutoffs under OG_MODE_A: the caller's cutoff_a for LJ
// and reaction-field Coulomb (forces_nonbonded.rs:125, nonbonded.rs:41-84),
// 10 Å for the GB pair force (GB_DEFAULT_CUTOFF_A_PUB, forces_gb.rs:33-36,
// forces.rs:115-120) and 20 Å for the descreening sum (gb.rs:37-40).
#define OG_GB_CUTOFF_A 10.0
#define OG_DESCREEN_CUTOFF_A 20.0

static og_pair_args pair_args(og_engine* e, double cutoff_a) {
    og_pair_args a{};
    a.pos = e->pos; a.par = e->par; a.gb = e->gb; a.qs = e->qs;
    a.tile_atom0 = e->tile_atom0; a.tile_cnt = e->tile_cnt;
    a.work = e->work; a.n_work = e->n_work; a.masks = e->masks;
    a.strips = e->strips; a.n_strip = e->n_strip; a.fwork = e->fwork; a.n_fwork = e->n_fwork;
    a.ds_far = e->ds_far;
    a.facc = e->facc; a.iacc = e->iacc; a.eacc = e->eacc;
    a.grid = e->pair_grid;
    a.sflag = e->sflag;
    const double rc = cutoff_a;
    a.rc2_nb = (float)(rc * rc);
    a.rc2_gb = (float)(OG_GB_CUTOFF_A * OG_GB_CUTOFF_A);
    a.rc2_ds = (float)(OG_DESCREEN_CUTOFF_A * OG_DESCREEN_CUTOFF_A);
    a.inv_rc3 = (float)(1.0 / (rc * rc * rc));
    a.kappa = (float)(1.0 / (2.0 * rc * rc * rc));
    a.crf = (float)(3.0 / (2.0 * rc));
    return a;
}

// The LJ / Coulomb cutoff of a call: cutoff_a under OG_MODE_A (finite and
// positive), none otherwise (the value is then not used).
static int check_cutoff(uint32_t flags, double cutoff_a) {
    if ((flags & OG_MODE_A) && !(cutoff_a > 0.0 && std::isfinite(cutoff_a)))
        return fail(OG_ERR_ARG, "cutoff_a must be finite and > 0 under OG_MODE_A (origami uses 10 Å)");
    return OG_OK;
}

static void invalidate_graphs(og_engine* e) {
    for (auto& g : e->gexec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    e->graph_ok = false;
}

// sur_check: -1 = decide from the call counter; 0 / 1 = forced (graph capture)
// forces = false (energies only): no pair-force kernel and no SASA terms or
// sums; the bonded kernel still adds its forces into the accumulator.
// next_noise (dynamics steps without SASA): the next step's O-step normals,
// formed one thread per atom by the bonded gather kernel, so that the cluster
// step only reads them.  They stay on the engine stream: a second branch in a
// captured step costs more than the work.
// step_mode (direct-launch dynamics with SASA): the cluster kernel already
// ran on the side stream, which keeps the SASA chain and, after this call,
// the final kick; the pair stream waits for the positions and the side
// stream for the pair forces, so no event wait sits on the SASA chain.
static int launch_forces(og_engine* e, uint32_t flags, double cutoff_a, double sasa_scale, bool energy,
                         int sur_check = -1, bool forces = true, const og_noise_args* next_noise = nullptr,
                         bool step_mode = false) {
    const bool mode_a = flags & OG_MODE_A;
    if (!mode_a) cutoff_a = OG_GB_CUTOFF_A;         // not used by the all-pairs kernels
    if (!forces) sasa_scale = 0.0;
    // SASA on the side stream, overlapping everything up to the join below.
    // Both branches add into the force and energy accumulators with integer
    // atomics only, so the sums do not depend on the interleaving.  The
    // per-kernel profile keeps one stream so its event pairs time each class.
    const bool fork = (flags & OG_SASA) && !e->prof;
    step_mode = step_mode && fork;
    if (fork) {
        if (step_mode) {
            HIPCHK(hipEventRecord(e->fork_ev, e->side));      // positions of this step
            HIPCHK(hipStreamWaitEvent(e->stream, e->fork_ev, 0));
        } else {
            HIPCHK(hipEventRecord(e->fork_ev, e->stream));
            HIPCHK(hipStreamWaitEvent(e->side, e->fork_ev, 0));
        }
        const int rc = og_sasa_launch(&e->sasa, e->pos, e->facc, sasa_scale, e->side,
                                      energy ? e->eacc : nullptr, e->atom_sys, e->sflag);
        if (rc != OG_OK) return fail(rc, "SASA launch failed: %s", og_sasa_error());
        if (!step_mode) HIPCHK(hipEventRecord(e->join_ev, e->side));
    }
    og_sur* sur = (flags & OG_SURROGATE) ? e->sur : nullptr;
    // Small batches (the LDS gather's range): the bonded terms share one
    // launch with the descreening pass, the bonded gather one with the Born
    // radii after it.
    const bool fuse = !energy && !sur && !e->prof && e->n_atoms < OG_GATHER_LDS_MAX_ATOMS;
    og_bonded_args b{};
    {
        ProfScope p(e, OG_K_BONDED);
        b.pos = e->pos; b.facc = e->facc; b.eacc = e->eacc; b.atom_sys = e->atom_sys;
        b.nb = e->nb; b.na = e->na; b.nd = e->nd; b.ni = e->ni;
        b.bonds = e->bonds; b.bond_p = e->bond_p; b.angles = e->angles; b.angle_p = e->angle_p;
        b.dih = e->dih; b.dih_p = e->dih_p; b.imp = e->imp; b.imp_p = e->imp_p;
        b.tf = e->bonded_tf; b.csr_start = e->bonded_csr_start; b.csr_slot = e->bonded_csr_slot;
        b.n_atoms = e->n_atoms;
        b.sflag = e->sflag;
        if (next_noise) b.noise = *next_noise;
        if (!fuse) og_launch_bonded(b, energy, e->stream);
    }
    og_pair_args a = pair_args(e, cutoff_a);
    // Surrogate radii: descriptors + MLP for the systems still in surrogate
    // mode; their tile pairs skip the descreening pass.  Every check_every
    // calls the plain pass runs for every system too and systems whose radii
    // deviate by more than the tolerance go back to plain GB.
    bool check = false;
    if (sur) {
        if (sur_check < 0) {
            check = sur->check_every > 0 && (sur->calls % sur->check_every) == 0;
            sur->calls++;
        } else {
            check = sur_check > 0;
        }
        ProfScope p(e, OG_K_SURROGATE);
        if (og_sur_descriptors(sur, false) || og_sur_infer(sur) || (check && og_sur_check_begin(sur))) {
            if (fork) (void)hipStreamWaitEvent(e->stream, e->join_ev, 0);   // SASA still joins
            return fail(OG_ERR_HIP, "surrogate launch failed: %s", og_sur_error());
        }
    }
    {
        ProfScope p(e, OG_K_DESCREEN);
        og_pair_args ad = a;
        ad.skip_sys = (sur && !check) ? sur->mode : (!forces ? e->eval_skip : nullptr);
        if (fuse) og_launch_descreen_bonded(ad, b, mode_a, e->stream);
        else og_launch_pair_descreen(ad, mode_a, e->stream);
    }
    {
        ProfScope p(e, OG_K_RADII);
        og_radii_args r{};
        r.iacc = e->iacc; r.atom_sys = e->atom_sys; r.atom_slot = e->atom_slot; r.rho = e->rho;
        r.charge = e->charge; r.gb = e->gb; r.radii = e->radii; r.clamped = e->clamped;
        r.eacc = e->eacc; r.n_atoms = e->n_atoms;
        r.sflag = e->sflag;
        if (sur) { r.sur_mode = sur->mode; r.r_sur = sur->r_sur; r.dev = sur->dev; r.check = check ? 1 : 0; }
        if (fuse) og_launch_radii_gather(r, b, e->stream);
        else og_launch_radii(r, energy, e->stream);
        if (sur && check && og_sur_check_apply(sur)) {
            if (fork) (void)hipStreamWaitEvent(e->stream, e->join_ev, 0);
            return fail(OG_ERR_HIP, "surrogate launch failed: %s", og_sur_error());
        }
    }
    if (forces) {
        ProfScope p(e, OG_K_FORCE);
        og_launch_pair_force(a, mode_a, e->stream);
    }
    if (energy) {
        ProfScope p(e, OG_K_OTHER);
        og_energy64_args en{};
        en.charge = e->charge; en.eps = e->eps; en.rmin = e->rmin; en.eps14 = e->eps14;
        en.rmin14 = e->rmin14; en.radii = e->radii;
        en.rc2_nb = cutoff_a * cutoff_a;
        en.kappa = 1.0 / (2.0 * cutoff_a * cutoff_a * cutoff_a);
        en.crf = 3.0 / (2.0 * cutoff_a);
        og_pair_args ae = a;
        ae.skip_sys = forces ? nullptr : e->eval_skip;
        og_launch_pair_energy64(ae, en, mode_a, e->stream);
        if (flags & OG_SASA_SR) {
            const int rc = og_sasa_sr_launch(&e->sasa, e->pos, e->sr_dots, OG_SR_DOTS, e->stream, e->eacc,
                                             e->atom_sys, e->sflag);
            if (rc != OG_OK) {
                if (fork) (void)hipStreamWaitEvent(e->stream, e->join_ev, 0);
                return fail(OG_ERR_HIP, "Shrake-Rupley launch failed: %s", og_sasa_error());
            }
        }
    }
    if (step_mode) {
        HIPCHK(hipEventRecord(e->join_ev, e->stream));         // pair forces of this step
        HIPCHK(hipStreamWaitEvent(e->side, e->join_ev, 0));
    } else if (fork) {
        HIPCHK(hipStreamWaitEvent(e->stream, e->join_ev, 0));
    }
    if (!fork && (flags & OG_SASA)) {
        ProfScope p(e, OG_K_SASA);
        int rc = og_sasa_launch(&e->sasa, e->pos, e->facc, sasa_scale, e->stream,
                                energy ? e->eacc : nullptr, e->atom_sys, e->sflag);
        if (rc != OG_OK) return fail(rc, "SASA launch failed: %s", og_sasa_error());
    }
    HIPCHK(hipGetLastError());
    return OG_OK;
}

static int check_sasa_overflow(og_engine* e, uint32_t flags) {
    if (!(flags & (OG_SASA | OG_SASA_SR))) return OG_OK;
    int ovf = 0;
    if (og_sasa_overflow(&e->sasa, &ovf) != 0) return fail(OG_ERR_HIP, "reading SASA status failed");
    if (ovf) return fail(OG_ERR_ARG, "%d SASA atom evaluations exceeded the per-atom capacity "
                         "(128 caps / 256 arcs / 256 vertices) and were skipped", ovf);
    return OG_OK;
}

static const char* status_reason(int bits) {
    if (bits & OG_ST_SASA) return "a SASA atom exceeded the arrangement's capacity (128 caps / 256 arcs / 256 vertices)";
    if (bits & OG_ST_RANGE) return "a force or energy contribution was NaN or beyond the fixed-point range (|x| >= 2^30)";
    return "non-finite velocities";
}

static int read_energies(og_engine* e, double* out) {
    std::vector<long long> acc((size_t)e->n_sys * OG_NE);
    HIPCHK(hipMemcpyAsync(acc.data(), e->eacc, acc.size() * sizeof(long long), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (size_t k = 0; k < acc.size(); ++k) out[k] = (double)acc[k] * OG_FIX_INV;
    return OG_OK;
}

// forces = false: energies only; the engine's force array keeps the forces of
// the last force evaluation (the next og_run step starts from them).
static int evaluate(og_engine* e, const double* xyz, double* f_out, uint32_t flags, double cutoff_a,
                    double sasa_gamma_scale, double* energies, bool forces, bool external = false) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if ((flags & OG_SURROGATE) && !(e->sur && e->sur->enabled))
        return fail(OG_ERR_STATE, "OG_SURROGATE needs og_surrogate_setup, training and og_surrogate_enable");
    int rc;
    if ((rc = check_cutoff(flags, cutoff_a))) return rc;
    if (flags & OG_SASA_SR) {
        if (forces) return fail(OG_ERR_ARG, "OG_SASA_SR is an energy term only (og_energies)");
        if (flags & OG_SASA) return fail(OG_ERR_ARG, "OG_SASA and OG_SASA_SR select two SASA energies; pass one");
    }
    if (xyz && (rc = og_set_positions(e, xyz))) return rc;
    const bool energy = (flags & OG_ENERGY) || energies;
    if (energy) HIPCHK(hipMemsetAsync(e->eacc, 0, (size_t)e->n_sys * OG_NE * sizeof(long long), e->stream));
    HIPCHK(hipMemsetAsync(e->clamped, 0, (size_t)e->n_sys * sizeof(int), e->stream));
    HIPCHK(hipMemsetAsync(e->sflag, 0, (size_t)e->n_sys * sizeof(int), e->stream));
    og_launch_zero_acc(e->facc, e->iacc, e->n_atoms, e->stream);
    if ((rc = launch_forces(e, flags, cutoff_a, sasa_gamma_scale, energy, -1, forces))) return rc;
    if (forces) {
        ProfScope p(e, OG_K_BACK);
        og_back_args b{};
        b.pos = e->pos; b.vel = e->vel; b.force = e->force; b.facc = e->facc; b.mass = e->mass;
        b.atom_sys = e->atom_sys; b.n_atoms = e->n_atoms; b.dt = 0.0; b.do_kick = 0;
        b.ke_acc = nullptr; b.diverged = e->diverged; b.sflag = e->sflag;
        if (external) b.tun = e->tun;
        og_launch_back(b, e->stream);
    }
    HIPCHK(hipGetLastError());
    if (f_out) {
        HIPCHK(hipMemcpyAsync(f_out, e->force, 3 * (size_t)e->n_atoms * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->prof) prof_collect(e);
    std::vector<int> st(e->n_sys);
    HIPCHK(hipMemcpy(st.data(), e->sflag, e->n_sys * sizeof(int), hipMemcpyDeviceToHost));
    if (energies) {
        if ((rc = read_energies(e, energies))) return rc;
        for (int s = 0; s < e->n_sys; ++s)     // a flagged system's sums are not its energies
            if (st[s])
                for (int k = 0; k < OG_NE; ++k) energies[(size_t)s * OG_NE + k] = NAN;
    }
    if ((rc = check_sasa_overflow(e, flags))) return rc;
    for (int s = 0; s < e->n_sys; ++s)
        if (st[s]) return fail(OG_ERR_NONFINITE, "system %d: %s", s, status_reason(st[s]));
    if (f_out) {
        for (size_t k = 0; k < 3 * (size_t)e->n_atoms; ++k)
            if (!std::isfinite(f_out[k])) return fail(OG_ERR_NONFINITE, "non-finite force at component %zu", k);
    }
    return OG_OK;
}

extern "C" int og_forces(og_engine* e, const double* xyz, double* f_out, uint32_t flags, double cutoff_a,
                         double sasa_gamma_scale, double* energies) {
    return evaluate(e, xyz, f_out, flags, cutoff_a, sasa_gamma_scale, energies, true);
}

extern "C" int og_energies(og_engine* e, const double* xyz, uint32_t flags, double cutoff_a, double* energies) {
    if (!energies) return fail(OG_ERR_ARG, "energies is NULL");
    return evaluate(e, xyz, nullptr, flags | OG_ENERGY, cutoff_a, 0.0, energies, false);
}

extern "C" int og_born_radii(og_engine* e, double* radii, double* integral) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if (radii) HIPCHK(hipMemcpyAsync(radii, e->radii, (size_t)e->n_atoms * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    std::vector<long long> acc;
    if (integral) {
        acc.resize(e->n_atoms);
        HIPCHK(hipMemcpyAsync(acc.data(), e->iacc, (size_t)e->n_atoms * sizeof(long long), hipMemcpyDeviceToHost, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    if (integral)
        for (int i = 0; i < e->n_atoms; ++i) integral[i] = (double)acc[i] * OG_FIX_INV;
    return OG_OK;
}

// ---------------------------------------------------------------------------
// Dynamics
// ---------------------------------------------------------------------------
extern "C" int og_dynamics_setup(og_engine* e, const og_dynamics_params* p, const double* temps) {
    if (!e || !p || !temps) return fail(OG_ERR_ARG, "NULL argument");
    if (!(p->dt_fs > 0.0) || !(p->friction_ps_inv >= 0.0))
        return fail(OG_ERR_ARG, "dt must be > 0 and friction >= 0");
    if (int rc = check_cutoff(p->flags, p->cutoff_a)) return rc;
    if ((p->flags & OG_SURROGATE) && !(e->sur && e->sur->enabled))
        return fail(OG_ERR_STATE, "OG_SURROGATE needs og_surrogate_setup, training and og_surrogate_enable");
    for (int s = 0; s < e->n_sys; ++s)
        if (!(temps[s] >= 0.0)) return fail(OG_ERR_ARG, "temperature of system %d must be >= 0", s);
    e->dyn = *p;
    e->alpha = std::exp(-p->friction_ps_inv * p->dt_fs * 1.0e-3);
    for (int s = 0; s < e->n_sys; ++s) {
        e->temps[s] = temps[s];
        const double total = 3.0 * e->sys_n[s];
        const int nc = p->constrain_h_bonds ? e->n_cons_sys[s] : 0;
        e->dofcorr[s] = nc > 0 ? total / std::max(total - nc, 1.0) : 1.0;
    }
    HIPCHK(hipMemcpyAsync(e->sys_temp, e->temps.data(), e->n_sys * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->sys_dofc, e->dofcorr.data(), e->n_sys * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemsetAsync(e->shake_fail, 0, e->n_sys * sizeof(int), e->stream));
    HIPCHK(hipMemsetAsync(e->diverged, 0, e->n_sys * sizeof(int), e->stream));
    e->step = 0;
    e->dyn_ready = true;
    invalidate_graphs(e);
    // initial forces at the current positions (with the external tunnel, as
    // every step's forces)
    return evaluate(e, nullptr, nullptr, p->flags & ~OG_ENERGY, p->cutoff_a, p->sasa_gamma_scale, nullptr, true,
                    true);
}

extern "C" int og_set_temperature(og_engine* e, int32_t s, double t) {
    if (!e || s < 0 || s >= e->n_sys || !(t >= 0.0)) return fail(OG_ERR_ARG, "bad system or temperature");
    e->temps[s] = t;
    HIPCHK(hipMemcpyAsync(e->sys_temp + s, &e->temps[s], sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return OG_OK;
}

// side_steps: the cluster kernel and the final kick on the side stream with
// the SASA chain between them (direct launches with SASA; the caller brackets
// the steps with a fork to and a join from the side stream).
static int record_step(og_engine* e, const og_cluster_args& c, const og_back_args& b, int sur_check,
                       const og_noise_args* next_noise, bool side_steps = false) {
    const og_dynamics_params& p = e->dyn;
    hipStream_t is = side_steps ? e->side : e->stream;
    {
        ProfScope ps(e, OG_K_CLUSTER);
        og_launch_cluster_step(c, is);
    }
    int rc = launch_forces(e, p.flags & ~OG_ENERGY, p.cutoff_a, p.sasa_gamma_scale, false, sur_check, true,
                           next_noise, side_steps);
    if (rc) return rc;
    {
        ProfScope ps(e, OG_K_BACK);
        og_launch_back(b, is);
    }
    return OG_OK;
}

// side_fork: the side stream waits for the work enqueued on the engine stream
// so far; side_join: the engine stream waits for the side stream's.
static int side_fork(og_engine* e) {
    HIPCHK(hipEventRecord(e->fork_ev, e->stream));
    HIPCHK(hipStreamWaitEvent(e->side, e->fork_ev, 0));
    return OG_OK;
}
static int side_join(og_engine* e) {
    HIPCHK(hipEventRecord(e->join_ev, e->side));
    HIPCHK(hipStreamWaitEvent(e->stream, e->join_ev, 0));
    return OG_OK;
}

static int capture_step(og_engine* e, const og_cluster_args& c, const og_back_args& b, int sur_check,
                        const og_noise_args* next_noise, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int rc = record_step(e, c, b, sur_check, next_noise);
    const hipError_t er = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (er != hipSuccess) return fail(OG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(er));
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(OG_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
    return OG_OK;
}

// Appends the current positions of every system to the frame ring (a
// device-to-device copy on the engine stream, in order with the steps).
static int save_frame(og_engine* e, uint32_t step) {
    const size_t n3 = 3 * (size_t)e->n_atoms;
    const int slot = (int)(e->frame_head % e->frame_cap);
    HIPCHK(hipMemcpyAsync(e->frames + (size_t)slot * n3, e->pos, n3 * sizeof(double), hipMemcpyDeviceToDevice,
                          e->stream));
    e->frame_step[slot] = step;
    e->frame_head++;
    return OG_OK;
}

// Frames the next n_steps steps add to the ring (steps that are multiples of
// frame_every), refused up front when the ring would overflow.
static int check_frames(og_engine* e, int32_t n_steps) {
    if (e->frame_every <= 0) return OG_OK;
    const long long s0 = e->step, s1 = (long long)e->step + n_steps;
    const long long n_new = s1 / e->frame_every - s0 / e->frame_every;
    if (e->frame_head - e->frame_tail + n_new > e->frame_cap)
        return fail(OG_ERR_ARG, "%lld frame(s) would exceed the ring of %d (%lld unread): read frames first",
                    n_new, e->frame_cap, e->frame_head - e->frame_tail);
    return OG_OK;
}

// The BAOAB step loop shared by og_run and og_run_noalloc.  ke_dev (may be
// null): n_steps x n_systems int64 kinetic-energy accumulators, zeroed by the
// caller; captured graphs only where allow_graphs.
static int run_steps(og_engine* e, int32_t n_steps, long long* ke_dev, bool allow_graphs) {
    const og_dynamics_params& p = e->dyn;
    if (int rc = check_frames(e, n_steps)) return rc;
    og_cluster_args c{};
    c.pos = e->pos; c.vel = e->vel; c.ref = e->ref; c.force = e->force; c.facc = e->facc;
    c.iacc = e->iacc; c.mass = e->mass; c.atom_sys = e->atom_sys; c.atom_local = e->atom_local;
    c.cl_start = e->cl_start; c.cl_atoms = e->cl_atoms; c.cn_start = e->cn_start;
    c.cn_len = e->cn_len; c.n_clusters = e->n_clusters; c.n_small = e->n_clusters_small;
    c.sys_temp = e->sys_temp; c.sys_dofcorr = e->sys_dofc; c.shake_fail = e->shake_fail;
    c.diverged = e->diverged; c.dt = p.dt_fs; c.alpha = e->alpha;
    c.use_shake = p.constrain_h_bonds ? 1 : 0;
    c.key0 = (uint32_t)(p.seed & 0xffffffffu);
    c.key1 = (uint32_t)(p.seed >> 32);
    c.step_dev = e->step_dev;
    og_back_args b{};
    b.pos = e->pos; b.vel = e->vel; b.force = e->force; b.facc = e->facc; b.mass = e->mass;
    b.atom_sys = e->atom_sys; b.n_atoms = e->n_atoms; b.dt = p.dt_fs; b.do_kick = 1;
    b.diverged = e->diverged;
    b.step_dev = e->step_dev;
    b.sflag = e->sflag;
    b.tun = e->tun;
    // the device step counter restarts from the host's count each call
    og_launch_set_u32(e->step_dev, e->step, e->stream);
    // status bits of this call's force evaluations (a system already stopped
    // keeps its diverged bits)
    HIPCHK(hipMemsetAsync(e->sflag, 0, (size_t)e->n_sys * sizeof(int), e->stream));
    // Without SASA the O-step normals come from a buffer: the first step's
    // here, each later step's during the force evaluation before it.  With
    // SASA the cluster step forms its own; the extra launch would share the
    // chip with the SASA branch, which sets the step time there.
    og_noise_args nz{};
    const og_noise_args* nzp = nullptr;
    if (!(p.flags & OG_SASA)) {
        nz.key0 = c.key0; nz.key1 = c.key1; nz.atom_sys = e->atom_sys; nz.atom_local = e->atom_local;
        nz.n_atoms = e->n_atoms; nz.step_dev = e->step_dev; nz.noise = e->noise;
        nz.ahead = 0;
        {
            ProfScope ps(e, OG_K_CLUSTER);
            og_launch_noise(nz, e->stream);
        }
        nz.ahead = 1;
        nzp = &nz;
        c.noise = e->noise;
        c.n_atoms = e->n_atoms;
    }
    og_sur* sur = (p.flags & OG_SURROGATE) ? e->sur : nullptr;
    // Persistent replica steps (og_use_graphs 3): one launch runs many whole
    // steps, a block per system (batches of small systems without SASA or the
    // surrogate), chunked at frame boundaries.  Measured on the REMD batch (8
    // and 64 replicas of 171 atoms): 56 us per step against 36.5 us with
    // direct launches and 39.4 us with captured steps: one CU's four SIMDs
    // take a replica's 18 tile-pair items in turn, where the launches spread
    // them over the chip.
    const bool persist_ok = !(p.flags & (OG_SASA | OG_SURROGATE)) && !e->prof &&
                            e->max_sys_tiles <= OG_PERSIST_MAX_TILES && e->n_clusters_small == e->n_clusters &&
                            n_steps > 0;
    if (persist_ok && e->graph_mode == 3) {
        og_pair_args pa = pair_args(e, p.cutoff_a > 0.0 && (p.flags & OG_MODE_A) ? p.cutoff_a : OG_GB_CUTOFF_A);
        og_bonded_args ba{};
        ba.pos = e->pos; ba.facc = e->facc; ba.eacc = e->eacc; ba.atom_sys = e->atom_sys;
        ba.nb = e->nb; ba.na = e->na; ba.nd = e->nd; ba.ni = e->ni;
        ba.bonds = e->bonds; ba.bond_p = e->bond_p; ba.angles = e->angles; ba.angle_p = e->angle_p;
        ba.dih = e->dih; ba.dih_p = e->dih_p; ba.imp = e->imp; ba.imp_p = e->imp_p;
        ba.tf = e->bonded_tf; ba.csr_start = e->bonded_csr_start; ba.csr_slot = e->bonded_csr_slot;
        ba.n_atoms = e->n_atoms;
        ba.sflag = e->sflag;
        og_radii_args ra{};
        ra.iacc = e->iacc; ra.atom_sys = e->atom_sys; ra.atom_slot = e->atom_slot; ra.rho = e->rho;
        ra.charge = e->charge; ra.gb = e->gb; ra.radii = e->radii; ra.clamped = e->clamped;
        ra.eacc = e->eacc; ra.n_atoms = e->n_atoms; ra.sflag = e->sflag;
        og_back_args bk = b;
        bk.step_dev = nullptr;
        og_persist_args pp{};
        pp.n_sys = e->n_sys;
        pp.atom_off = e->p_atom_off; pp.cl_small_off = e->p_cl_small_off; pp.cl_small = e->p_cl_small;
        pp.cl_big_off = e->p_cl_big_off; pp.cl_big = e->p_cl_big; pp.term_off = e->p_term_off;
        pp.work_off = e->p_work_off; pp.fwork_off = e->p_fwork_off; pp.strip_off = e->p_strip_off;
        og_noise_args pn{};
        if (nzp) pn = *nzp;
        int done = 0;
        while (done < n_steps) {
            int chunk = std::min(n_steps - done, OG_PERSIST_CHUNK);
            const uint32_t st0 = e->step + (uint32_t)done;
            if (e->frame_every > 0) {
                const int to_frame = e->frame_every - (int)(st0 % (uint32_t)e->frame_every);
                chunk = std::min(chunk, to_frame);
            }
            pp.n_steps = chunk;
            pp.step0 = st0;
            pp.ke_acc = ke_dev ? ke_dev + (size_t)done * e->n_sys : nullptr;
            og_launch_replica_steps(pa, ba, ra, c, bk, pn, pp, (p.flags & OG_MODE_A) != 0, e->stream);
            HIPCHK(hipGetLastError());
            done += chunk;
            const uint32_t st = e->step + (uint32_t)done;
            if (e->frame_every > 0 && st % (uint32_t)e->frame_every == 0)
                if (int rc = save_frame(e, st)) return rc;
        }
        e->step += (uint32_t)n_steps;
        return OG_OK;
    }
    // Captured steps (og_use_graphs 1): every launch of one BAOAB step
    // (cluster B-A-SHAKE-O-A, forces, final kick) is recorded once into a
    // hipGraph and replayed; the step counter lives in device memory so one
    // graph serves every step.  Direct launches measured faster on both the
    // REMD batch (36.5 against 39.4 us per step) and the 64-chain batch, so
    // they are the engine's choice.
    const bool want = e->graph_mode == 1;
    const bool graphs = allow_graphs && want && !e->prof && !ke_dev && n_steps >= 2;
    if (graphs) {
        if (!e->graph_ok) {
            invalidate_graphs(e);
            if (int rc = capture_step(e, c, b, 0, nzp, &e->gexec[0])) return rc;
            if (sur && sur->check_every > 0)
                if (int rc = capture_step(e, c, b, 1, nzp, &e->gexec[1])) return rc;
            e->graph_ok = true;
        }
        for (int k = 0; k < n_steps; ++k) {
            int chk = 0;
            if (sur) {
                chk = (sur->check_every > 0 && (sur->calls % sur->check_every) == 0) ? 1 : 0;
                sur->calls++;
            }
            HIPCHK(hipGraphLaunch(e->gexec[chk], e->stream));
            const uint32_t st = e->step + (uint32_t)k + 1u;
            if (e->frame_every > 0 && st % (uint32_t)e->frame_every == 0)
                if (int rc = save_frame(e, st)) return rc;
        }
    } else {
        // with SASA the integrator runs on the SASA (critical) stream
        const bool side_steps = (p.flags & OG_SASA) && !e->prof && n_steps > 0;
        if (side_steps)
            if (int rc = side_fork(e)) return rc;
        for (int k = 0; k < n_steps; ++k) {
            b.ke_acc = ke_dev ? ke_dev + (size_t)k * e->n_sys : nullptr;
            if (int rc = record_step(e, c, b, -1, nzp, side_steps)) {
                if (side_steps) (void)side_join(e);
                return rc;
            }
            const uint32_t st = e->step + (uint32_t)k + 1u;
            if (e->frame_every > 0 && st % (uint32_t)e->frame_every == 0) {
                if (side_steps)
                    if (int rc = side_join(e)) return rc;
                if (int rc = save_frame(e, st)) return rc;
                if (side_steps)
                    if (int rc = side_fork(e)) return rc;
            }
            if (e->prof && (k & 63) == 63) {
                HIPCHK(hipStreamSynchronize(e->stream));
                prof_collect(e);
            }
        }
        if (side_steps)
            if (int rc = side_join(e)) return rc;
    }
    e->step += (uint32_t)n_steps;
    HIPCHK(hipGetLastError());
    return OG_OK;
}

// After a run: SASA status, then OG_ERR_NONFINITE naming the first stopped
// system (the diverged flags are read in chunks on the stack).
static int run_status(og_engine* e) {
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->prof) prof_collect(e);
    int ovf = 0;
    if ((e->dyn.flags & OG_SASA) && og_sasa_overflow(&e->sasa, &ovf) != 0)
        return fail(OG_ERR_HIP, "reading SASA status failed");
    int n_div = 0, first = -1, first_bits = 0;
    int chunk[256];
    for (int s0 = 0; s0 < e->n_sys; s0 += 256) {
        const int m = std::min(256, e->n_sys - s0);
        HIPCHK(hipMemcpy(chunk, e->diverged + s0, (size_t)m * sizeof(int), hipMemcpyDeviceToHost));
        for (int k = 0; k < m; ++k)
            if (chunk[k]) {
                n_div++;
                if (first < 0) { first = s0 + k; first_bits = chunk[k]; }
            }
    }
    if (n_div)
        return fail(OG_ERR_NONFINITE, "%d system(s) stopped; system %d: %s (the others ran all steps)", n_div,
                    first, status_reason(first_bits));
    return OG_OK;
}

// int64 fixed-point kinetic energies, copied into the caller's buffer as is,
// converted in place to kJ/mol
static void ke_to_double(double* ke, size_t n) {
    for (size_t k = 0; k < n; ++k) {
        long long v;
        std::memcpy(&v, ke + k, sizeof v);
        ke[k] = (double)v * OG_FIX_INV;
    }
}

extern "C" int og_run(og_engine* e, int32_t n_steps, double* ke_per_step) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if (!e->dyn_ready) return fail(OG_ERR_STATE, "og_dynamics_setup has not been called");
    if (n_steps < 0) return fail(OG_ERR_ARG, "negative step count");
    if (int rc = check_frames(e, n_steps)) return rc;
    long long* ke_dev = nullptr;
    const size_t n_ke = (size_t)n_steps * e->n_sys;
    if (ke_per_step && n_steps > 0) {
        HIPCHK(hipMallocAsync((void**)&ke_dev, n_ke * sizeof(long long), e->stream));
        og_launch_zero_i64(ke_dev, (long long)n_ke, e->stream);
    }
    if (int rc = run_steps(e, n_steps, ke_dev, true)) {
        if (ke_dev) (void)hipFreeAsync(ke_dev, e->stream);
        return rc;
    }
    if (ke_dev) {
        HIPCHK(hipMemcpyAsync(ke_per_step, ke_dev, n_ke * sizeof(long long), hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipFreeAsync(ke_dev, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        ke_to_double(ke_per_step, n_ke);
    }
    return run_status(e);
}

// ---------------------------------------------------------------------------
// og_run_noalloc: the per-step kinetic-energy accumulators live in this
// file-scope device array instead of a per-call allocation.  Each GPU holds
// its own copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (g_noalloc_ke[device]),
// never through the array's name.  Every call on a GPU shares the array, so
// calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) long long scratch[OG_NOALLOC_KE_MAX];
#define OG_MAX_DEVICES 64
static long long* g_noalloc_ke[OG_MAX_DEVICES];
static std::mutex g_noalloc_mu;

// Resolves this device's copy of the scratch array (once per device, with
// the device current); called from og_engine_create.
static int noalloc_setup(og_engine* e) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= OG_MAX_DEVICES) return fail(OG_ERR_STATE, "device %d beyond the supported range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ke[dev]) {
        void* ptr = nullptr;
        const hipError_t er = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (er != hipSuccess || !ptr)
            return fail(OG_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(er));
        g_noalloc_ke[dev] = (long long*)ptr;
    }
    e->noalloc_ke = g_noalloc_ke[dev];
    return OG_OK;
}

extern "C" int og_run_noalloc(og_engine* e, int32_t n_steps, double* ke_per_step) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if (!e->dyn_ready) return fail(OG_ERR_STATE, "og_dynamics_setup has not been called");
    if (n_steps < 0) return fail(OG_ERR_ARG, "negative step count");
    if (!e->noalloc_ke) return fail(OG_ERR_STATE, "the engine's scratch array is not set up");
    long long* ke_dev = nullptr;
    const size_t n_ke = (size_t)n_steps * e->n_sys;
    if (ke_per_step && n_steps > 0) {
        if (n_ke > (size_t)OG_NOALLOC_KE_MAX)
            return fail(OG_ERR_ARG, "n_steps x n_systems = %zu exceeds the %d kinetic-energy entries of og_run_noalloc",
                        n_ke, OG_NOALLOC_KE_MAX);
        ke_dev = e->noalloc_ke;
        og_launch_zero_i64(ke_dev, (long long)n_ke, e->stream);
    }
    if (int rc = run_steps(e, n_steps, ke_dev, false)) return rc;
    if (ke_dev) {
        HIPCHK(hipMemcpyAsync(ke_per_step, ke_dev, n_ke * sizeof(long long), hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        ke_to_double(ke_per_step, n_ke);
    }
    return run_status(e);
}

extern "C" int og_use_graphs(og_engine* e, int32_t on) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if (on < 0 || on > 3) return fail(OG_ERR_ARG, "graph mode must be 0, 1, 2 or 3");
    e->graph_mode = on;
    invalidate_graphs(e);
    return OG_OK;
}

// Autotuner: times the candidate launch widths of the all-pairs kernels and
// of the SASA arrangement kernel on the engine's current positions and keeps
// the fastest.  Candidates: pair kernels 2..32 blocks of 256 threads per CU;
// SASA 2..8 blocks of 2 waves per CU.  With SASA the widths are timed on the
// whole force evaluation, SASA on its stream beside the pair chain, since the
// best split of the CUs between the two depends on both: first the SASA width
// with the pair width that is fastest alone, then the pair width again.
extern "C" int og_set_launch(og_engine* e, int32_t pair_blocks, int32_t sasa_blocks, int32_t sasa_passes) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if (sasa_passes < 0 || sasa_passes > 2) return fail(OG_ERR_ARG, "sasa_passes must be 0, 1 or 2");
    if (pair_blocks > 0) e->pair_grid = std::min(pair_blocks, std::max((e->n_work + 3) / 4, 1));
    if (sasa_blocks > 0) og_sasa_set_blocks(&e->sasa, sasa_blocks);
    og_sasa_set_passes(&e->sasa, sasa_passes);
    invalidate_graphs(e);
    return OG_OK;
}

extern "C" int og_autotune(og_engine* e, uint32_t flags, int32_t* choice, double* best_ms) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if ((flags & OG_SURROGATE) && !(e->sur && e->sur->enabled))
        return fail(OG_ERR_STATE, "OG_SURROGATE needs og_surrogate_setup, training and og_surrogate_enable");
    int dev = 0, n_cu = 256;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.multiProcessorCount > 0)
        n_cu = prop.multiProcessorCount;
    hipEvent_t t0, t1;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    auto timed = [&](auto&& launch, float& ms) -> int {
        launch();                                   // warm
        HIPCHK(hipEventRecord(t0, e->stream));
        for (int r = 0; r < 3; ++r) launch();
        HIPCHK(hipEventRecord(t1, e->stream));
        HIPCHK(hipEventSynchronize(t1));
        HIPCHK(hipEventElapsedTime(&ms, t0, t1));
        ms /= 3.0f;
        return OG_OK;
    };
    const bool mode_a = flags & OG_MODE_A;
    // the force pass's tile shape (j tiles per strip; the pass runs one wave
    // per work item), then the descreening pass's launch width; each
    // candidate timed twice, the faster kept
    int best_grid = 0, best_strip = e->strip_j;
    float best_force = 1e30f, best_desc = 1e30f;
    for (int sj : {2, 4, 8}) {
        if (int rc = build_force_lists(e, sj)) return rc;
        og_pair_args a = pair_args(e, 10.0);
        for (int rep = 0; rep < 2; ++rep) {
            float ms = 0.0f;
            if (int rc = timed([&] { og_launch_pair_force(a, mode_a, e->stream); }, ms)) return rc;
            if (ms < best_force) { best_force = ms; best_strip = sj; }
        }
    }
    if (int rc = build_force_lists(e, best_strip)) return rc;
    for (int per_cu : {2, 4, 8, 16, 32}) {
        const int g = std::min(per_cu * n_cu, std::max((e->n_work + 3) / 4, 1));
        og_pair_args a = pair_args(e, 10.0);
        a.grid = g;
        for (int rep = 0; rep < 2; ++rep) {
            float ms = 0.0f;
            if (int rc = timed([&] { og_launch_pair_descreen(a, mode_a, e->stream); }, ms)) return rc;
            if (ms < best_desc) { best_desc = ms; best_grid = g; }
        }
        if (g == std::max((e->n_work + 3) / 4, 1)) break;
    }
    const float best_pair = best_force + best_desc;
    e->pair_grid = best_grid;
    int best_sasa = og_sasa_max_blocks(&e->sasa);
    float best_s = 0.0f;
    if ((flags & OG_SASA) && e->n_sasa_work > 0) {
        og_sasa_set_passes(&e->sasa, 0);
        // 6 blocks of 2 waves per CU is the arrangement kernel's residency
        // (LDS and VGPRs); fewer leave room for the pair kernels beside it
        const uint32_t fl = (flags & (OG_MODE_A | OG_SURROGATE)) | OG_SASA;
        int lrc = OG_OK;
        // with the surrogate: its evaluations without the plain-GB check
        auto evaluation = [&] { if (!lrc) lrc = launch_forces(e, fl, 10.0, 1.0, false, 0); };
        best_s = 1e30f;
        for (int per_cu : {2, 3, 4, 5, 6, 8}) {
            const int g = std::min(per_cu * n_cu, og_sasa_max_blocks(&e->sasa));
            og_sasa_set_blocks(&e->sasa, g);
            float ms = 0.0f;
            if (int rc = timed(evaluation, ms)) return rc;
            if (lrc) return lrc;
            if (ms < best_s) { best_s = ms; best_sasa = g; }
        }
        og_sasa_set_blocks(&e->sasa, best_sasa);
        for (int per_cu : {2, 4, 8, 16, 32}) {
            const int g = std::min(per_cu * n_cu, std::max((e->n_work + 3) / 4, 1));
            const int keep = e->pair_grid;
            e->pair_grid = g;
            float ms = 0.0f;
            if (int rc = timed(evaluation, ms)) return rc;
            if (lrc) return lrc;
            if (ms < best_s) { best_s = ms; best_grid = g; } else { e->pair_grid = keep; }
            if (g == std::max((e->n_work + 3) / 4, 1)) break;
        }
        e->pair_grid = best_grid;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    og_launch_zero_acc(e->facc, e->iacc, e->n_atoms, e->stream);
    HIPCHK(hipStreamSynchronize(e->stream));
    invalidate_graphs(e);
    if (choice) { choice[0] = best_grid; choice[1] = best_sasa; }
    if (best_ms) { best_ms[0] = best_pair; best_ms[1] = best_s; }
    return OG_OK;
}

// ---------------------------------------------------------------------------
// og_minimize: the batched L-BFGS of lbfgs.hip driven phase by phase; the
// host reads four counters between phases (how many systems search, failed,
// are active, accepted) and launches only the phases some system needs.
// ---------------------------------------------------------------------------
extern "C" int og_minimize(og_engine* e, const og_minimize_params* prm, og_minimize_result* results) {
    if (!e || !prm) return fail(OG_ERR_ARG, "NULL argument");
    const og_minimize_params& P = *prm;
    if (P.max_steps < 0) return fail(OG_ERR_ARG, "max_steps must be >= 0");
    if (P.history < 1 || P.history > OG_LB_MAX_HISTORY)
        return fail(OG_ERR_ARG, "history must be 1..%d", OG_LB_MAX_HISTORY);
    if (!(P.alpha0 > 0.0) || !(P.contraction > 0.0 && P.contraction < 1.0) || !(P.max_step_a > 0.0) ||
        !(P.min_alpha > 0.0) || !(P.c1 >= 0.0) || !(P.gradient_tol >= 0.0) || !(P.energy_tol >= 0.0))
        return fail(OG_ERR_ARG, "minimiser parameters out of range");
    if (P.flags & ~(OG_MODE_A | OG_SASA)) return fail(OG_ERR_ARG, "og_minimize takes OG_MODE_A and OG_SASA only");
    if (int rc = check_cutoff(P.flags, P.cutoff_a)) return rc;
    const int S = e->n_sys;
    const size_t n3 = 3 * (size_t)e->n_atoms;
    // workspace: one allocation, carved
    const size_t hist_n = (size_t)P.history * n3;
    const size_t bytes_d = (3 * n3 + 2 * hist_n + (size_t)S * OG_LB_MAX_HISTORY) * sizeof(double);
    const size_t bytes_s = (size_t)S * sizeof(og_lb_sys);
    const size_t bytes_i = ((size_t)S + 1 + (size_t)S + 4) * sizeof(int);
    char* ws = nullptr;
    HIPCHK(hipMalloc((void**)&ws, bytes_d + bytes_s + bytes_i));
    og_lbfgs_args a{};
    a.n_sys = S;
    a.n_atoms = e->n_atoms;
    a.history = P.history;
    double* d = (double*)ws;
    a.x0 = d; a.g = d + n3; a.p = d + 2 * n3;
    a.hs = d + 3 * n3; a.hy = a.hs + hist_n;
    a.rho = a.hy + hist_n;
    a.sys = (og_lb_sys*)(ws + bytes_d);
    int* ip = (int*)(ws + bytes_d + bytes_s);
    a.atom_off = ip; a.skip = ip + S + 1; a.counters = ip + 2 * S + 1;
    a.pos = e->pos; a.force = e->force; a.eacc = e->eacc;
    a.gradient_tol = P.gradient_tol; a.energy_tol = P.energy_tol; a.max_step = P.max_step_a;
    a.alpha0 = P.alpha0; a.contraction = P.contraction; a.c1 = P.c1; a.min_alpha = P.min_alpha;
    std::vector<int> off(e->atom_off.begin(), e->atom_off.end());
    off.push_back(e->n_atoms);
    off.resize(S + 1);
    int rc = OG_OK;
    auto run = [&]() -> int {
        HIPCHK(hipMemcpy(ip, off.data(), (S + 1) * sizeof(int), hipMemcpyHostToDevice));
        auto energies = [&](bool probes) -> int {
            HIPCHK(hipMemsetAsync(e->eacc, 0, (size_t)S * OG_NE * sizeof(long long), e->stream));
            HIPCHK(hipMemsetAsync(e->sflag, 0, (size_t)S * sizeof(int), e->stream));
            og_launch_zero_acc(e->facc, e->iacc, e->n_atoms, e->stream);
            e->eval_skip = probes ? a.skip : nullptr;
            const int r = launch_forces(e, P.flags | OG_ENERGY, P.cutoff_a, P.sasa_gamma_scale, true, -1, false);
            e->eval_skip = nullptr;
            return r;
        };
        auto forces = [&]() -> int {
            HIPCHK(hipMemsetAsync(e->sflag, 0, (size_t)S * sizeof(int), e->stream));
            og_launch_zero_acc(e->facc, e->iacc, e->n_atoms, e->stream);
            if (int r = launch_forces(e, P.flags & ~OG_ENERGY, P.cutoff_a, P.sasa_gamma_scale, false, -1, true)) return r;
            og_back_args b{};
            b.pos = e->pos; b.vel = e->vel; b.force = e->force; b.facc = e->facc; b.mass = e->mass;
            b.atom_sys = e->atom_sys; b.n_atoms = e->n_atoms; b.dt = 0.0; b.do_kick = 0;
            b.diverged = e->diverged; b.sflag = e->sflag;
            og_launch_back(b, e->stream);
            return OG_OK;
        };
        int cnt[4] = {0, 0, 0, 0};
        auto read_counts = [&](int phase) -> int {
            og_launch_lbfgs(phase, a, 0, e->stream);
            HIPCHK(hipMemcpyAsync(cnt, a.counters, sizeof cnt, hipMemcpyDeviceToHost, e->stream));
            HIPCHK(hipStreamSynchronize(e->stream));
            return OG_OK;
        };
        auto probes = [&]() -> int {
            while (cnt[0] > 0) {
                og_launch_lbfgs(OG_LB_K_TRIAL, a, 0, e->stream);
                if (int r = energies(true)) return r;
                if (int r = read_counts(OG_LB_K_ACCEPT)) return r;
            }
            return OG_OK;
        };
        if (int r = energies(false)) return r;
        og_launch_lbfgs(OG_LB_K_INIT, a, 0, e->stream);
        if (P.max_steps > 0)
            if (int r = forces()) return r;
        for (int step = 0; step < P.max_steps; ++step) {
            og_launch_lbfgs(OG_LB_K_BEGIN, a, step, e->stream);
            if (int r = read_counts(OG_LB_K_COUNT)) return r;
            if (cnt[0] == 0 && cnt[1] == 0) break;          // every system converged or stopped
            if (int r = probes()) return r;
            if (cnt[1] > 0) {                               // steepest-descent retries
                og_launch_lbfgs(OG_LB_K_FALLBACK, a, 0, e->stream);
                if (int r = read_counts(OG_LB_K_COUNT)) return r;
                if (int r = probes()) return r;
                if (cnt[1] > 0) {
                    og_launch_lbfgs(OG_LB_K_FALLBACK, a, 0, e->stream);
                    if (int r = read_counts(OG_LB_K_COUNT)) return r;
                }
            }
            if (cnt[3] > 0) {
                if (int r = forces()) return r;
                og_launch_lbfgs(OG_LB_K_UPDATE, a, step, e->stream);
            }
        }
        HIPCHK(hipGetLastError());
        std::vector<og_lb_sys> st(S);
        HIPCHK(hipMemcpyAsync(st.data(), a.sys, bytes_s, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        if (results)
            for (int s = 0; s < S; ++s) {
                og_minimize_result& r = results[s];
                r.steps = st[s].steps;
                r.converged = st[s].state == OG_LB_CONVERGED ? 1 : 0;
                r.line_search_failures = st[s].failures;
                r.stopped = st[s].state == OG_LB_STOPPED ? 1 : 0;
                r.initial_energy = st[s].e_init;
                r.final_energy = st[s].e_prev;
                r.max_force = st[s].max_force;
            }
        return OG_OK;
    };
    rc = run();
    e->eval_skip = nullptr;
    (void)hipStreamSynchronize(e->stream);
    (void)hipFree(ws);
    invalidate_graphs(e);
    return rc;
}

extern "C" int og_strip_tiles(og_engine* e, int32_t set, int32_t* get) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if (set > 0) {
        HIPCHK(hipStreamSynchronize(e->stream));
        if (int rc = build_force_lists(e, set)) return rc;
        invalidate_graphs(e);
    }
    if (get) *get = e->strip_j;
    return OG_OK;
}

extern "C" int og_set_tunnel(og_engine* e, const og_tunnel* t) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    og_tunnel_dev d{};
    if (t) {
        const double n = std::sqrt(t->direction[0] * t->direction[0] + t->direction[1] * t->direction[1] +
                                   t->direction[2] * t->direction[2]);
        if (!(std::fabs(n - 1.0) < 1e-9)) return fail(OG_ERR_ARG, "tunnel direction must be a unit vector");
        if (!(t->radius_a >= 0.0) || !(t->length_a >= 0.0) || !(t->k_confine >= 0.0) ||
            !std::isfinite(t->origin[0] + t->origin[1] + t->origin[2]))
            return fail(OG_ERR_ARG, "tunnel radius, length and k must be >= 0 and the origin finite");
        d.ox = t->origin[0]; d.oy = t->origin[1]; d.oz = t->origin[2];
        d.dx = t->direction[0]; d.dy = t->direction[1]; d.dz = t->direction[2];
        d.radius = t->radius_a; d.length = t->length_a; d.k = t->k_confine;
        d.on = 1;
    }
    e->tun = d;
    invalidate_graphs(e);
    return OG_OK;
}

extern "C" int og_frames_setup(og_engine* e, int32_t capacity, int32_t every) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if (capacity < 0 || every < 0 || (capacity > 0) != (every > 0))
        return fail(OG_ERR_ARG, "capacity and every must both be > 0 (or both 0 to switch frames off)");
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->frames) {
        HIPCHK(hipFree(e->frames));
        e->frames = nullptr;
    }
    e->frame_cap = 0;
    e->frame_every = 0;
    e->frame_head = e->frame_tail = 0;
    e->frame_step.clear();
    if (capacity == 0) return OG_OK;
    HIPCHK(hipMalloc((void**)&e->frames, (size_t)capacity * 3 * (size_t)e->n_atoms * sizeof(double)));
    e->frame_cap = capacity;
    e->frame_every = every;
    e->frame_step.assign(capacity, 0u);
    // the frame at the current step, as langevin.rs:205-220 emits step 0
    if (int rc = save_frame(e, e->step)) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    return OG_OK;
}

extern "C" int og_frames_read(og_engine* e, double* xyz, uint32_t* steps, int32_t max_frames, int32_t* n_frames) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if (n_frames) *n_frames = 0;
    if (max_frames < 0) return fail(OG_ERR_ARG, "max_frames must be >= 0");
    if (e->frame_cap == 0) return OG_OK;
    HIPCHK(hipStreamSynchronize(e->stream));
    const size_t n3 = 3 * (size_t)e->n_atoms;
    int k = 0;
    for (; k < max_frames && e->frame_tail < e->frame_head; ++k, ++e->frame_tail) {
        const int slot = (int)(e->frame_tail % e->frame_cap);
        if (xyz) HIPCHK(hipMemcpy(xyz + (size_t)k * n3, e->frames + (size_t)slot * n3, n3 * sizeof(double),
                                  hipMemcpyDeviceToHost));
        if (steps) steps[k] = e->frame_step[slot];
    }
    if (n_frames) *n_frames = k;
    return OG_OK;
}

extern "C" int og_swap_systems(og_engine* e, int32_t a, int32_t b) {
    if (!e || a < 0 || b < 0 || a >= e->n_sys || b >= e->n_sys || a == b)
        return fail(OG_ERR_ARG, "bad system indices");
    if (e->sys_n[a] != e->sys_n[b]) return fail(OG_ERR_ARG, "systems %d and %d differ in size", a, b);
    if (e->topo_hash[a] != e->topo_hash[b])
        return fail(OG_ERR_ARG, "systems %d and %d differ in parameters or topology", a, b);
    if (!(e->temps[a] > 0.0) || !(e->temps[b] > 0.0)) return fail(OG_ERR_STATE, "temperatures not set");
    const double sa = std::sqrt(e->temps[a] / e->temps[b]);
    const double sb = std::sqrt(e->temps[b] / e->temps[a]);
    og_launch_swap(e->pos, e->vel, e->force, e->atom_off[a], e->atom_off[b], e->sys_n[a], sa, sb, e->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    return OG_OK;
}

extern "C" int og_get_status(og_engine* e, int32_t* div, int32_t* fails, int32_t* clamped) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if (div) HIPCHK(hipMemcpy(div, e->diverged, e->n_sys * sizeof(int), hipMemcpyDeviceToHost));
    if (fails) HIPCHK(hipMemcpy(fails, e->shake_fail, e->n_sys * sizeof(int), hipMemcpyDeviceToHost));
    if (clamped) HIPCHK(hipMemcpy(clamped, e->clamped, e->n_sys * sizeof(int), hipMemcpyDeviceToHost));
    return OG_OK;
}

extern "C" int og_profile_enable(og_engine* e, int32_t on) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    HIPCHK(hipStreamSynchronize(e->stream));
    prof_collect(e);
    e->prof = on != 0;
    for (int k = 0; k < OG_N_KCLASS; ++k) {
        e->prof_ms[k] = 0.0;
        e->prof_cnt[k] = 0;
    }
    return OG_OK;
}

extern "C" int og_profile_read(og_engine* e, double* ms, int64_t* count) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    HIPCHK(hipStreamSynchronize(e->stream));
    prof_collect(e);
    for (int k = 0; k < OG_N_KCLASS; ++k) {
        if (ms) ms[k] = e->prof_ms[k];
        if (count) count[k] = e->prof_cnt[k];
    }
    return OG_OK;
}

// ---------------------------------------------------------------------------
// Learned Born-radius surrogate
// ---------------------------------------------------------------------------
#define SUR_CHECK(e)                                                                   \
    do {                                                                               \
        if (!(e)) return fail(OG_ERR_ARG, "engine is NULL");                           \
        if (!(e)->sur) return fail(OG_ERR_STATE, "call og_surrogate_setup first");     \
    } while (0)
#define SUR_CALL(x)                                                                    \
    do {                                                                               \
        if (x) return fail(OG_ERR_HIP, "surrogate: %s", og_sur_error());               \
    } while (0)

extern "C" int og_surrogate_setup(og_engine* e, int64_t capacity_rows, uint64_t seed) {
    if (!e) return fail(OG_ERR_ARG, "engine is NULL");
    if (capacity_rows < 0) return fail(OG_ERR_ARG, "capacity_rows must be >= 0");
    if (e->sur) {
        og_sur_free(e->sur);
        delete e->sur;
    }
    invalidate_graphs(e);
    e->sur = new og_sur();
    if (og_sur_create(e->sur, e->n_atoms, e->n_sys, e->mass_h, e->sys_tile_off_h, e->pos, e->rho,
                      e->tile_atom0, e->tile_cnt, e->n_tiles, capacity_rows, seed, e->stream)) {
        og_sur_free(e->sur);
        delete e->sur;
        e->sur = nullptr;
        return fail(OG_ERR_HIP, "surrogate set-up failed: %s", og_sur_error());
    }
    return OG_OK;
}

// plain Born radii (descreening pass + radii) for every system at the current positions
static int plain_radii(og_engine* e, uint32_t flags) {
    og_launch_zero_acc(e->facc, e->iacc, e->n_atoms, e->stream);
    og_pair_args a = pair_args(e, 10.0);
    og_launch_pair_descreen(a, (flags & OG_MODE_A) != 0, e->stream);
    og_radii_args r{};
    r.iacc = e->iacc; r.atom_sys = e->atom_sys; r.atom_slot = e->atom_slot; r.rho = e->rho;
    r.charge = e->charge; r.gb = e->gb; r.radii = e->radii; r.clamped = e->clamped;
    r.eacc = e->eacc; r.n_atoms = e->n_atoms;
    r.sflag = e->sflag;
    og_launch_radii(r, false, e->stream);
    HIPCHK(hipGetLastError());
    return OG_OK;
}

extern "C" int og_surrogate_collect(og_engine* e, uint32_t flags, int64_t* rows_total) {
    SUR_CHECK(e);
    int rc;
    if ((rc = plain_radii(e, flags))) return rc;
    SUR_CALL(og_sur_descriptors(e->sur, true));
    SUR_CALL(og_sur_collect(e->sur, e->radii));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (rows_total) *rows_total = e->sur->n_train;
    return OG_OK;
}

extern "C" int og_surrogate_train(og_engine* e, int32_t n_steps, int32_t batch, double lr, uint64_t seed,
                                  float* loss_per_step) {
    SUR_CHECK(e);
    if (n_steps < 0 || !(lr > 0.0)) return fail(OG_ERR_ARG, "n_steps must be >= 0 and lr > 0");
    for (int t = 0; t < n_steps; ++t) {
        const uint64_t sd = seed * 0x9E3779B97F4A7C15ull + (uint64_t)t;
        SUR_CALL(og_sur_train_step(e->sur, batch, lr, sd, loss_per_step ? loss_per_step + t : nullptr));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    return OG_OK;
}

extern "C" int og_surrogate_enable(og_engine* e, int32_t check_every, double tolerance) {
    SUR_CHECK(e);
    if (check_every < 0 || !(tolerance >= 0.0)) return fail(OG_ERR_ARG, "check_every >= 0, tolerance >= 0");
    e->sur->check_every = check_every;
    e->sur->tol = tolerance;
    e->sur->calls = 0;
    e->sur->enabled = true;
    invalidate_graphs(e);
    std::vector<int> on(e->n_sys, 1);
    HIPCHK(hipMemcpyAsync(e->sur->mode, on.data(), on.size() * sizeof(int), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return OG_OK;
}

extern "C" int og_surrogate_radii(og_engine* e, const double* xyz, double* radii) {
    SUR_CHECK(e);
    int rc;
    if (xyz && (rc = og_set_positions(e, xyz))) return rc;
    SUR_CALL(og_sur_descriptors(e->sur, true));
    SUR_CALL(og_sur_infer(e->sur));
    if (radii)
        HIPCHK(hipMemcpyAsync(radii, e->sur->r_sur, (size_t)e->n_atoms * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return OG_OK;
}

extern "C" int og_surrogate_descriptors(og_engine* e, const double* xyz, float* x) {
    SUR_CHECK(e);
    if (!x) return fail(OG_ERR_ARG, "x is NULL");
    int rc;
    if (xyz && (rc = og_set_positions(e, xyz))) return rc;
    SUR_CALL(og_sur_descriptors(e->sur, true));
    std::vector<unsigned short> h((size_t)e->n_atoms * OG_SUR_K1);
    HIPCHK(hipMemcpyAsync(h.data(), e->sur->X, h.size() * 2, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int i = 0; i < e->n_atoms; ++i)
        for (int f = 0; f < OG_SUR_F; ++f) {
            const uint32_t u = (uint32_t)h[(size_t)i * OG_SUR_K1 + f] << 16;
            float v;
            std::memcpy(&v, &u, 4);
            x[(size_t)i * OG_SUR_F + f] = v;
        }
    return OG_OK;
}

extern "C" int og_surrogate_params(og_engine* e, float* get, const float* set) {
    SUR_CHECK(e);
    if (set) SUR_CALL(og_sur_set_params(e->sur, set));
    if (get) SUR_CALL(og_sur_get_params(e->sur, get));
    return OG_OK;
}

extern "C" int og_surrogate_gradient(og_engine* e, const float* x, const double* r_target, int32_t n,
                                     float* grad, double* loss) {
    SUR_CHECK(e);
    if (!x || !r_target || !grad) return fail(OG_ERR_ARG, "NULL argument");
    std::vector<float> t(n > 0 ? n : 1);
    for (int b = 0; b < n; ++b) t[b] = (float)r_target[b];
    SUR_CALL(og_sur_gradient(e->sur, x, t.data(), n, grad, loss));
    return OG_OK;
}

extern "C" int og_surrogate_status(og_engine* e, int32_t* surrogate_mode, float* deviation) {
    SUR_CHECK(e);
    if (surrogate_mode)
        HIPCHK(hipMemcpyAsync(surrogate_mode, e->sur->mode, (size_t)e->n_sys * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    if (deviation)
        HIPCHK(hipMemcpyAsync(deviation, e->sur->dev, (size_t)e->n_sys * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return OG_OK;
}
