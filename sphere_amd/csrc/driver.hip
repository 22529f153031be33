// driver.hip -- host driver and C ABI (include/sphere_hip.h).
//
// Replaces the reference's process boundary and time loop: DEM::DEM
// (src/sphere.cpp:16-95), DEM::startTime (src/device.cu:833-2834) and main()
// (src/main.cpp:28-225).  One context owns one device arena and one HIP
// stream; all kernels of a step are enqueued back to back on that stream
// (the reference synchronises after every launch).  The host waits only
// where the algorithm needs a device value: the Darcy convergence check every
// 10 sweeps, output events, and the end of each call.
#include <hip/hip_runtime.h>

#include "net.h"
#include "ns.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "host_state.h"
#include "kernels.h"

namespace {

thread_local std::string g_last_error;

int key_bits_for(uint32_t ncells)
{
    int bits = 1;
    while (bits < 32 && (ncells - 1) >> bits) ++bits;
    return bits;
}

}  // namespace

struct sph_ctx {
    int device = 0;
    sph_ns *ns = nullptr;  // an attached Navier-Stokes fluid (sph_ns_attach)
    hipStream_t stream = nullptr;
    // integrateWalls runs on `aux`, overlapped with the next step's contact
    // search (which does not read the walls); the contact forces wait for it
    hipStream_t aux = nullptr;
    hipEvent_t ev_integrated = nullptr, ev_walls = nullptr;
    bool walls_pending = false;
    std::string err;
    bool loaded = false;
    uint64_t gen = 0;  // incremented by every load (captured graphs hold the old buffers)
    HostState hs;
    DemBuffers db{};
    DarcyBuffers dd{};
    std::vector<void *> allocs;
    uint32_t np = 0, nb0 = 0, nw = 0, ncells = 0;
    uint32_t nc = 32;  // contact slots per particle
    int key_bits = 1;
    // neighbour lists
    double nl_frac = 0.1;    // requested skin / largest radius
    double nl_rmax = 0.0;    // largest radius of the loaded state
    uint32_t nl_cap = 128;   // list entries per particle
    bool rb_fused = true;    // list rebuild as one launch (k_rebuild)
    uint32_t rb_blocks = 0;  // its grid; 0: separate sort, reorder and list launches
    uint32_t rb_cap = 0;     // blocks its barrier words serve (the multi-box rebuild of fluid boxes too)
    uint32_t rb_spin = 0;    // its barrier wait bound in sleeps (0: the default)
    bool nl_on = false;
    NLArgs nla{};
    uint32_t nl_parity = 0;
    uint32_t nl_force = 1;
    bool hash_ready = false;  // the last integrate formed this step's hash
    // time loop host variables, src/device.cu:913-1036
    long iter = 0;
    double filetimeclock = 0.0;
    double host_wx0 = 0.0;
    unsigned int wall0_iz = 10000000u;
    double p_top_orig = 0.0;
    int velocity_state = 1;
    sph_darcy_stats stats{};
    std::vector<std::pair<long, unsigned>> conv;  // (iter+1, nijac)
    // pinned host scratch
    uint32_t *h_err = nullptr;
    unsigned long long *h_u64 = nullptr;
    unsigned long long *h_state = nullptr;  // Jacobi solve state (4 words)
    unsigned int jacobi_hint = 0;           // sweeps of the last solve
    double *h_d = nullptr;
    // learned Darcy correction: host parameters, and the network with its
    // device buffers for the loaded grid (made at the first solve)
    std::vector<float> net_host;
    int net_max_corr = 0;
    sph_net *net = nullptr;
    float *net_params = nullptr, *net_in = nullptr, *net_out = nullptr;
    // profiling
    bool prof = false;
    double ktime[SPH_K_COUNT] = {0};
    uint64_t klaunch[SPH_K_COUNT] = {0};
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::tuple<int, size_t, size_t>> ev_pending;
};

#define HIPCHK_E(call)                                                            \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ens_fail(e, SPH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define HIPCHK(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);         \
            return SPH_ERR_HIP;                                                   \
        }                                                                         \
    } while (0)

static int fail(sph_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    g_last_error = msg;
    return code;
}

// ------------------------------------------------------------------
// profiling with HIP events on the context stream
// ------------------------------------------------------------------

static size_t prof_event(sph_ctx *ctx, hipStream_t s)
{
    if (ctx->ev_used == ctx->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return (size_t)-1;
        ctx->ev_pool.push_back(e);
    }
    size_t i = ctx->ev_used++;
    (void)hipEventRecord(ctx->ev_pool[i], s);
    return i;
}

struct ProfScope {
    sph_ctx *ctx;
    int k;
    hipStream_t s;
    size_t e0;
    ProfScope(sph_ctx *c, int kid, hipStream_t st = nullptr)
        : ctx(c), k(kid), s(st ? st : c->stream), e0((size_t)-1)
    {
        if (ctx->prof) e0 = prof_event(ctx, s);
    }
    ~ProfScope()
    {
        if (ctx->prof && e0 != (size_t)-1) {
            size_t e1 = prof_event(ctx, s);
            if (e1 != (size_t)-1) ctx->ev_pending.emplace_back(k, e0, e1);
        }
    }
};

static void prof_flush(sph_ctx *ctx)
{
    for (auto &t : ctx->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev_pool[std::get<1>(t)], ctx->ev_pool[std::get<2>(t)]) ==
            hipSuccess) {
            ctx->ktime[std::get<0>(t)] += ms;
            ctx->klaunch[std::get<0>(t)] += 1;
        }
    }
    ctx->ev_pending.clear();
    ctx->ev_used = 0;
}

// The main stream waits for the pending wall update (work enqueued after
// this sees the new wall state).
static int wait_walls(sph_ctx *ctx)
{
    if (ctx->walls_pending) {
        HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->ev_walls, 0));
        ctx->walls_pending = false;
    }
    return SPH_OK;
}

static int sync(sph_ctx *ctx)
{
    int rc = wait_walls(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->aux));
    HIPCHK(hipGetLastError());
    if (ctx->prof) prof_flush(ctx);
    return SPH_OK;
}

// Reads and clears the kernel error flags.
static int check_flags(sph_ctx *ctx)
{
    if (!ctx->db.errflags) return SPH_OK;
    HIPCHK(hipMemcpyAsync(ctx->h_err, ctx->db.errflags, 4, hipMemcpyDeviceToHost, ctx->stream));
    int rc = sync(ctx);
    if (rc) return rc;
    uint32_t f = *ctx->h_err;
    if (!f) return SPH_OK;
    launch_fill_u32(ctx->stream, ctx->db.errflags, 1, 0u);
    if (f & ERRF_GRID)
        return fail(ctx, SPH_ERR_GRID, "One or more particles moved outside the grid (numerical "
                                       "instability? is the time step too large?)");
    if (f & ERRF_OVERFLOW)
        return fail(ctx, SPH_ERR_OVERFLOW, "more contacts on one particle than contact slots (" + std::to_string(ctx->nc) + ")");
    if (f & ERRF_BARRIER) {
        if (ctx->db.rb_bar) launch_fill_u32(ctx->stream, ctx->db.rb_bar, 2, 0u);
        return fail(ctx, SPH_ERR_HIP, "the list rebuild kernel's grid barrier timed out "
                                      "(its blocks were not all resident)");
    }
    return SPH_OK;
}

// ------------------------------------------------------------------
// device memory
// ------------------------------------------------------------------

template <class T>
static T *dalloc(sph_ctx *ctx, size_t n)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
    ctx->allocs.push_back(p);
    return (T *)p;
}

static void net_release(sph_ctx *ctx)
{
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    sph_net_destroy(ctx->net);
    ctx->net = nullptr;
    float *bufs[] = {ctx->net_params, ctx->net_in, ctx->net_out};
    for (float *b : bufs)
        if (b) (void)hipFree(b);
    ctx->net_params = ctx->net_in = ctx->net_out = nullptr;
}

static void free_device(sph_ctx *ctx)
{
    net_release(ctx);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->aux) (void)hipStreamSynchronize(ctx->aux);
    ctx->walls_pending = false;
    for (void *p : ctx->allocs) (void)hipFree(p);
    ctx->allocs.clear();
    ctx->db = DemBuffers{};
    ctx->dd = DarcyBuffers{};
    ctx->loaded = false;
}

static uint32_t dcells(const sph_state &h)
{
    return (h.num[0] + 2) * (h.num[1] + 2) * (h.num[2] + 2);
}

static inline uint32_t host_didx(const sph_state &h, uint32_t x, uint32_t y, uint32_t z)
{
    return (x + 1) + (h.num[0] + 2) * (y + 1) + (h.num[0] + 2) * (h.num[1] + 2) * (z + 1);
}

// ------------------------------------------------------------------
// value checks, src/sphere.cpp:145-345
// ------------------------------------------------------------------

static int check_values(sph_ctx *ctx, const HostState &hs)
{
    const sph_state &s = hs.h;
    char m[256];
    if (s.nd != 3) { snprintf(m, sizeof(m), "nd = %d, ND = 3", s.nd); return fail(ctx, SPH_ERR_VALUE, m); }
    if (s.np >= CT_MAX_NP) {  // contact slot words keep 28 bits for the partner
        snprintf(m, sizeof(m), "np = %u; at most %u particles", s.np, CT_MAX_NP - 1u);
        return fail(ctx, SPH_ERR_VALUE, m);
    }
    if (s.time_current > s.time_total || s.time_current < 0.0) {
        snprintf(m, sizeof(m), "time.current = %g s, time.total = %g s", s.time_current, s.time_total);
        return fail(ctx, SPH_ERR_VALUE, m);
    }
    if (s.origo[0] < 0.0 || s.origo[1] < 0.0 || s.origo[2] < 0.0)
        return fail(ctx, SPH_ERR_VALUE, "Negative values in grid.origo are known to cause problems");
    if (s.L[0] <= 0.0 || s.L[1] <= 0.0 || s.L[2] <= 0.0)
        return fail(ctx, SPH_ERR_VALUE, "grid.L must be positive");
    if (s.num[0] == 0 || s.num[1] == 0 || s.num[2] == 0)
        return fail(ctx, SPH_ERR_VALUE, "grid.num must be positive");
    if (s.periodic == 2 && s.num[0] < 3)
        return fail(ctx, SPH_ERR_VALUE, "When 1st dimension boundaries are periodic, there must be at least 3 cells");
    if (s.periodic == 1 && (s.num[0] < 3 || s.num[1] < 3))
        return fail(ctx, SPH_ERR_VALUE, "When 1st and 2nd dimension boundaries are periodic, there must be at least 3 cells");
    for (uint32_t i = 0; i < s.np; ++i) {
        const double *x = &hs.x[4 * (size_t)i];
        if (x[3] <= 0.0) {
            snprintf(m, sizeof(m), "Particle %u has a radius of %g m", i, x[3]);
            return fail(ctx, SPH_ERR_VALUE, m);
        }
        if (x[0] != x[0] || x[1] != x[1] || x[2] != x[2] || x[3] != x[3]) {
            snprintf(m, sizeof(m), "Particle %u position is NaN", i);
            return fail(ctx, SPH_ERR_VALUE, m);
        }
        if (x[0] < s.origo[0] || x[1] < s.origo[1] || x[2] < s.origo[2] || x[0] > s.L[0] ||
            x[1] > s.L[1] || x[2] > s.L[2]) {
            snprintf(m, sizeof(m), "Particle %u is outside of the computational grid", i);
            return fail(ctx, SPH_ERR_VALUE, m);
        }
    }
    if (s.nw > SPH_MAXWALLS) return fail(ctx, SPH_ERR_VALUE, "nw > MAXWALLS");
    if (s.nw > 0 && s.wmode[0] != 0 && s.np > 0) {
        double zmax = -1e300;
        for (uint32_t i = 0; i < s.np; ++i) zmax = std::max(zmax, hs.x[4 * (size_t)i + 2]);
        if (s.w_nx[3] < zmax)
            return fail(ctx, SPH_ERR_VALUE, "One or more particles have centres above the upper, dynamic wall");
    }
    if (s.k_n <= 0.0) { snprintf(m, sizeof(m), "k_n = %g N/m", s.k_n); return fail(ctx, SPH_ERR_VALUE, m); }
    if (s.contactmodel < 1 || s.contactmodel > 3) {  // reportValues, src/sphere.cpp:384-393
        snprintf(m, sizeof(m), "contact model value (%u) not understood", s.contactmodel);
        return fail(ctx, SPH_ERR_VALUE, m);
    }
    if (s.E < 0.0) { snprintf(m, sizeof(m), "E = %g", s.E); return fail(ctx, SPH_ERR_VALUE, m); }
    if (s.rho <= 0.0) { snprintf(m, sizeof(m), "rho = %g kg/m3", s.rho); return fail(ctx, SPH_ERR_VALUE, m); }
    if (s.fluid) {
        if (s.cfd_solver != 1) return fail(ctx, SPH_ERR_VALUE, "only the Darcy solver (cfd_solver = 1) is built");
        if (s.rho_f <= 0.0) return fail(ctx, SPH_ERR_VALUE, "rho_f <= 0");
        if (s.mu <= 0.0) return fail(ctx, SPH_ERR_VALUE, "mu <= 0");
        if (s.beta_f <= 0.0) return fail(ctx, SPH_ERR_VALUE, "beta_f <= 0");
        if (s.k_c <= 0.0) return fail(ctx, SPH_ERR_VALUE, "k_c <= 0");
        if (s.ndem == 0) return fail(ctx, SPH_ERR_VALUE, "ndem = 0");
    }
    return SPH_OK;
}

// ------------------------------------------------------------------
// upload / download
// ------------------------------------------------------------------

static void pack4(const std::vector<double> &src, size_t n, int w, std::vector<double> &dst)
{
    dst.assign(n * 4, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < w; ++c) dst[4 * i + c] = src[(size_t)w * i + c];
}

// Cells searched on each side at a list rebuild for this skin: 1 or 2 if a
// cell (or two) holds a diameter plus the skin and periodic axes keep at
// least 2 range + 1 cells; 0 if lists cannot be used.
static uint32_t nl_range(const sph_state &h, double rmax, double skin)
{
    double cmin = 1e300;
    for (int d = 0; d < 3; ++d) cmin = std::min(cmin, h.L[d] / h.num[d]);
    uint32_t range = 0;
    if (cmin >= (2.0 * rmax + skin) * (1.0 + 1e-9)) range = 1;
    else if (2.0 * cmin >= (2.0 * rmax + skin) * (1.0 + 1e-9)) range = 2;
    const uint32_t need = 2 * range + 1;
    bool ok = range > 0;
    if (h.periodic == 1 || h.periodic == 2) ok = ok && h.num[0] >= need;
    if (h.periodic == 1) ok = ok && h.num[1] >= need;
    return ok ? range : 0u;
}

static int upload(sph_ctx *ctx)
{
    HostState &hs = ctx->hs;
    sph_state &h = hs.h;
    const size_t np = h.np, nb = h.nb0;
    ctx->np = h.np;
    ctx->nb0 = h.nb0;
    ctx->nw = h.nw;
    ctx->ncells = h.num[0] * h.num[1] * h.num[2];
    ctx->key_bits = key_bits_for(ctx->ncells);
    DemBuffers &b = ctx->db;
    const size_t nblocks = (np + 255) / 256;
    const size_t ntiles = (np + 2047) / 2048;
#define AL(ptr, T, n)                                                     \
    do {                                                                  \
        ptr = dalloc<T>(ctx, n);                                          \
        if (!ptr) return fail(ctx, SPH_ERR_HIP, "hipMalloc failed for " #ptr); \
    } while (0)
    AL(b.x, double4, np); AL(b.vel, double4, np); AL(b.angvel, double3, np);
    AL(b.force, double3, np); AL(b.torque, double3, np); AL(b.angpos, double3, np);
    AL(b.acc, double3, np); AL(b.angacc, double3, np); AL(b.xyzsum, double3, np);
    AL(b.es_dot, double, np); AL(b.es, double, np); AL(b.ev_dot, double, np);
    AL(b.ev, double, np); AL(b.p, double, np);
    AL(b.keys, uint32_t, np); AL(b.vals, uint32_t, np);
    AL(b.keys_tmp, uint32_t, np); AL(b.vals_tmp, uint32_t, np);
    AL(b.hist, uint32_t, 256 * ntiles);
    AL(b.clock, StepClock, 1);
    AL(b.cell_start, uint32_t, ctx->ncells); AL(b.cell_end, uint32_t, ctx->ncells);
    AL(b.x_s, double4, np); AL(b.vel_s, double4, np); AL(b.angvel_s, double3, np);
    const size_t NCS = ctx->nc;
    const size_t NCE = ct_entries((uint32_t)np, (uint32_t)NCS);  // contact-array entries
    AL(b.contacts, uint32_t, NCE); AL(b.cmask, uint64_t, np);
    AL(b.delta_t, double, NCE * 3);
    AL(b.w_nx, double4, MAXW); AL(b.w_mvfd, double4, MAXW); AL(b.w_acc, double, MAXW);
    AL(b.walls_force_pp, double, np * (h.nw ? h.nw : 1));
    AL(b.wall_partial, double, nblocks);
    AL(b.tau_pp, double, np); AL(b.tau_partial, double, nblocks); AL(b.tau_eff, double, 1);
    AL(b.bonds, uint2, nb); AL(b.bond_broken, uint8_t, nb);
    AL(b.bond_delta, double4, nb); AL(b.bond_omega, double4, nb);
    AL(b.bond_f, double4, nb); AL(b.bond_t, double4, nb); AL(b.bond_applied, uint8_t, nb);
    AL(b.errflags, uint32_t, 1);
    hipStream_t s = ctx->stream;
    std::vector<double> t4;
    pack4(hs.x, np, 4, t4);
    HIPCHK(hipMemcpyAsync(b.x, t4.data(), np * 32, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    pack4(hs.vel, np, 4, t4);
    HIPCHK(hipMemcpyAsync(b.vel, t4.data(), np * 32, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    const std::vector<double> *v3[] = {&hs.angvel, &hs.force, &hs.torque, &hs.angpos, &hs.xyzsum};
    double3 *d3p[] = {b.angvel, b.force, b.torque, b.angpos, b.xyzsum};
    for (int k = 0; k < 5; ++k)  // three-component rows: the host layout as is
        HIPCHK(hipMemcpyAsync(d3p[k], v3[k]->data(), np * 24, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpyAsync(b.es_dot, hs.es_dot.data(), np * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b.es, hs.es.data(), np * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b.ev_dot, hs.ev_dot.data(), np * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b.ev, hs.ev.data(), np * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b.p, hs.p.data(), np * 8, hipMemcpyHostToDevice, s));
    launch_fill_f64(s, (double *)b.acc, np * 3, 0.0);
    launch_fill_f64(s, (double *)b.angacc, np * 3, 0.0);
    launch_fill_u32(s, b.contacts, NCE, (uint32_t)np);  // device.cu:471-473
    launch_fill_u32(s, (uint32_t *)b.cmask, 2 * np, 0u);
    launch_fill_f64(s, b.delta_t, NCE * 3, 0.0);
    launch_fill_f64(s, b.walls_force_pp, np * (h.nw ? h.nw : 1), 0.0);
    launch_fill_f64(s, b.w_acc, MAXW, 0.0);
    launch_fill_u32(s, b.errflags, 1, 0u);
    {
        double4 wn[MAXW], wm[MAXW];
        for (int w = 0; w < MAXW; ++w) {
            wn[w] = make_double4(h.w_nx[4 * w], h.w_nx[4 * w + 1], h.w_nx[4 * w + 2], h.w_nx[4 * w + 3]);
            wm[w] = make_double4(h.w_mvfd[4 * w], h.w_mvfd[4 * w + 1], h.w_mvfd[4 * w + 2], h.w_mvfd[4 * w + 3]);
        }
        HIPCHK(hipMemcpyAsync(b.w_nx, wn, sizeof(wn), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(b.w_mvfd, wm, sizeof(wm), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    // bonds: deltas as (t.x, t.y, t.z, n); CSR of bond endpoints per particle in
    // ascending bond order (the serial order of the scatter in cohesion.cuh:220-230)
    if (nb) {
        std::vector<uint32_t> bnd(hs.bonds);
        std::vector<uint8_t> broken(nb, 0);
        std::vector<double> bd(nb * 4), bo(nb * 4);
        for (size_t k = 0; k < nb; ++k) {
            if (bnd[2 * k] >= np || bnd[2 * k + 1] >= np) broken[k] = 1;  // cohesion.cuh:30
            bd[4 * k] = hs.bdt[3 * k]; bd[4 * k + 1] = hs.bdt[3 * k + 1];
            bd[4 * k + 2] = hs.bdt[3 * k + 2]; bd[4 * k + 3] = hs.bdn[k];
            bo[4 * k] = hs.bot[3 * k]; bo[4 * k + 1] = hs.bot[3 * k + 1];
            bo[4 * k + 2] = hs.bot[3 * k + 2]; bo[4 * k + 3] = hs.bon[k];
            if (broken[k]) { bnd[2 * k] = 0; bnd[2 * k + 1] = 0; }
        }
        std::vector<uint32_t> cnt(np + 1, 0);
        for (size_t k = 0; k < nb; ++k)
            if (!broken[k]) { cnt[bnd[2 * k]]++; cnt[bnd[2 * k + 1]]++; }
        std::vector<uint32_t> plist, off(1, 0), ent;
        std::vector<uint32_t> start(np, 0);
        for (size_t i = 0; i < np; ++i)
            if (cnt[i]) {
                start[i] = (uint32_t)plist.size();
                plist.push_back((uint32_t)i);
                off.push_back(off.back() + cnt[i]);
            }
        ent.assign(off.back(), 0u);
        std::vector<uint32_t> fillpos(plist.size(), 0);
        for (size_t k = 0; k < nb; ++k) {
            if (broken[k]) continue;
            for (int side = 0; side < 2; ++side) {
                uint32_t pi = bnd[2 * k + side];
                uint32_t slot = start[pi];
                ent[off[slot] + fillpos[slot]++] = (uint32_t)(k << 1) | (uint32_t)side;
            }
        }
        b.nbp = (uint32_t)plist.size();
        AL(b.bp_list, uint32_t, plist.size()); AL(b.bp_off, uint32_t, off.size());
        AL(b.bp_ent, uint32_t, ent.size());
        // the same entries indexed by every particle (empty ranges for the
        // unbonded ones), for the gather inside integrate
        std::vector<uint32_t> pp_off(np + 1, 0);
        for (size_t i = 0; i < np; ++i) pp_off[i + 1] = pp_off[i] + cnt[i];
        AL(b.pp_off, uint32_t, np + 1);
        HIPCHK(hipMemcpyAsync(b.pp_off, pp_off.data(), (np + 1) * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(b.bonds, bnd.data(), nb * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(b.bond_broken, broken.data(), nb, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(b.bond_delta, bd.data(), nb * 32, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(b.bond_omega, bo.data(), nb * 32, hipMemcpyHostToDevice, s));
        if (!plist.empty()) {
            HIPCHK(hipMemcpyAsync(b.bp_list, plist.data(), plist.size() * 4, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(b.bp_off, off.data(), off.size() * 4, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(b.bp_ent, ent.data(), ent.size() * 4, hipMemcpyHostToDevice, s));
        }
        HIPCHK(hipStreamSynchronize(s));
    }
    // Darcy
    if (h.fluid) {
        DarcyBuffers &d = ctx->dd;
        const size_t c = dcells(h);
        AL(d.p, double, c); AL(d.p_old, double, c); AL(d.p_new, double, c);
        AL(d.dp_expl, double, c); AL(d.phi, double, c); AL(d.dphi, double, c);
        AL(d.k, double, c);
        const size_t ci = (size_t)h.num[0] * h.num[1] * h.num[2];
        AL(d.jc, double, 9 * ci); AL(d.jfrozen, uint8_t, ci);
        AL(d.v, double, 3 * c); AL(d.vp_avg, double, 3 * c); AL(d.grad_p, double, 3 * c);
        AL(d.p_constant, int, c); AL(d.f_p, double4, np);
        AL(d.maxres, unsigned long long, DARCY_MAXRES_WORDS);
        std::vector<double> p(c, 0.0), phi(c, 0.0), dphi(c, 0.0), v(3 * c, 0.0);
        std::vector<int> pc(c, 0);
        for (uint32_t z = 0; z < h.num[2]; ++z)
            for (uint32_t y = 0; y < h.num[1]; ++y)
                for (uint32_t x = 0; x < h.num[0]; ++x) {
                    size_t f = x + (size_t)h.num[0] * (y + (size_t)h.num[1] * z);
                    uint32_t ci = host_didx(h, x, y, z);
                    p[ci] = hs.p_f[f];
                    phi[ci] = hs.phi[f];
                    dphi[ci] = hs.dphi[f];
                    pc[ci] = hs.p_constant[f];
                    v[3 * ci] = hs.v_f[3 * f];
                    v[3 * ci + 1] = hs.v_f[3 * f + 1];
                    v[3 * ci + 2] = hs.v_f[3 * f + 2];
                }
        std::vector<double> zero(3 * c, 0.0);
        HIPCHK(hipMemcpyAsync(d.p, p.data(), c * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.phi, phi.data(), c * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.dphi, dphi.data(), c * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.v, v.data(), 3 * c * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.p_constant, pc.data(), c * 4, hipMemcpyHostToDevice, s));
        double *zf[] = {d.p_old, d.p_new, d.dp_expl, d.k};
        for (double *q : zf) HIPCHK(hipMemcpyAsync(q, zero.data(), c * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.vp_avg, zero.data(), 3 * c * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.grad_p, zero.data(), 3 * c * 8, hipMemcpyHostToDevice, s));
        std::vector<double> fp4;
        pack4(hs.f_p, np, 3, fp4);
        HIPCHK(hipMemcpyAsync(d.f_p, fp4.data(), np * 32, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        // src/device.cu:1017-1021
        ctx->p_top_orig = p[host_didx(h, 0, 0, h.num[2] - 1)] -
                          h.p_mod_A * sin(2.0 * M_PI * h.p_mod_f * h.time_current + h.p_mod_phi);
    }
    // Neighbour lists: contact models 2 and 3 on a fixed grid whose cells
    // leave room for the skin within `range` cells, with at least 2 range + 1
    // cells along periodic axes and cell coordinates that pack into 32 bits.
    ctx->nl_on = false;
    if (np > 0 && ctx->nl_frac > 0.0 && (h.contactmodel == 2 || h.contactmodel == 3) &&
        !(h.adaptive == 1 && h.nw > 0) && h.num[0] <= 2048 && h.num[1] <= 2048 &&
        h.num[2] <= 1024) {
        double rmax = 0.0;
        for (size_t i = 0; i < np; ++i) rmax = std::max(rmax, hs.x[4 * i + 3]);
        ctx->nl_rmax = rmax;
        const double skin = ctx->nl_frac * rmax;
        const uint32_t range = nl_range(h, rmax, skin);
        if (range > 0) {
            ctx->nla.skin = skin;
            // both partners move at most 0.45 skin before a rebuild
            ctx->nla.disp2 = (0.45 * skin) * (0.45 * skin);
            ctx->nla.cap = ctx->nl_cap;
            ctx->nla.range = range;
            AL(b.cellc, uint32_t, np); AL(b.x_ref, double4, np);
            AL(b.nl, uint32_t, (size_t)ctx->nl_cap * np); AL(b.nl_cnt, uint32_t, np);
            AL(b.nl_pos, uint32_t, np); AL(b.nl_flags, uint32_t, NLF_WORDS);
            AL(b.nl_gap, float, (size_t)ctx->nl_cap * np);
            AL(b.nl_gmin, float, np); AL(b.nl_disp, float, np);
            AL(b.nl_slot, uint8_t, (size_t)ctx->nl_cap * np);
            AL(b.slot_cand, uint8_t, NCE);
            launch_fill_u32(s, b.nl_flags, NLF_WORDS, 0u);
            launch_fill_u32(s, (uint32_t *)b.slot_cand, (NCE + 3) / 4, 0xffffffffu);
            launch_fill_f64(s, (double *)b.x_ref, np * 4, 0.0);
            launch_fill_u32(s, (uint32_t *)b.nl_disp, np, 0u);
            ctx->nl_on = true;
            // one resident block per compute unit at most (the rebuild
            // kernel's grid barriers need every block running)
            ctx->rb_blocks = 0;
            ctx->rb_cap = 0;
            int cus = 0;
            // (a fluid run rebuilds every step: there the separate launches,
            // on full grids, are faster; its barrier words serve the
            // multi-box launches of an ensemble)
            if (ctx->rb_fused && rebuild_blocks_per_cu() > 0 &&
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                      ctx->device) == hipSuccess && cus > 0) {
                const uint32_t want = (uint32_t)((np + 255) / 256);
                const uint32_t cap = (uint32_t)(rebuild_blocks_per_cu() * cus / 4);
                const uint32_t blocks =
                    std::max<uint32_t>(1u, std::min<uint32_t>(std::min<uint32_t>((uint32_t)cus, cap), want));
                AL(b.rb_bar, uint32_t, 2);
                AL(b.rb_part, uint32_t, blocks);
                launch_fill_u32(s, b.rb_bar, 2, 0u);
                ctx->rb_cap = blocks;
                if (!h.fluid) ctx->rb_blocks = blocks;
            }
        }
    }
    ctx->nl_parity = 0;
    ctx->nl_force = 1;
    ctx->hash_ready = false;
#undef AL
    ctx->iter = 0;
    ctx->gen++;
    {
        StepClock c0{0, h.time_current};
        HIPCHK(hipMemcpyAsync(ctx->db.clock, &c0, sizeof(c0), hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    ctx->filetimeclock = 0.0;
    ctx->host_wx0 = h.nw > 0 ? h.w_nx[3] : 0.0;
    ctx->wall0_iz = 10000000u;
    ctx->velocity_state = 1;
    ctx->stats = sph_darcy_stats{};
    ctx->conv.clear();
    ctx->loaded = true;
    return SPH_OK;
}

static void unpack4(const std::vector<double> &src, size_t n, int w, double *dst)
{
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < w; ++c) dst[(size_t)w * i + c] = src[4 * i + c];
}

// Download the device state into ctx->hs (transferFromGlobalDeviceMemory).
static int download(sph_ctx *ctx, bool particles, bool fluid_fields)
{
    HostState &hs = ctx->hs;
    sph_state &h = hs.h;
    const size_t np = h.np, nb = h.nb0;
    DemBuffers &b = ctx->db;
    {
        int rc = sync(ctx);
        if (rc) return rc;
    }
    {
        double4 wn[MAXW], wm[MAXW];
        HIPCHK(hipMemcpy(wn, b.w_nx, sizeof(wn), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(wm, b.w_mvfd, sizeof(wm), hipMemcpyDeviceToHost));
        for (uint32_t w = 0; w < h.nw; ++w) {
            h.w_nx[4 * w] = wn[w].x; h.w_nx[4 * w + 1] = wn[w].y;
            h.w_nx[4 * w + 2] = wn[w].z; h.w_nx[4 * w + 3] = wn[w].w;
            h.w_mvfd[4 * w] = wm[w].x; h.w_mvfd[4 * w + 1] = wm[w].y;
            h.w_mvfd[4 * w + 2] = wm[w].z; h.w_mvfd[4 * w + 3] = wm[w].w;
        }
    }
    if (particles && np) {
        std::vector<double> t4(np * 4);
        HIPCHK(hipMemcpy(t4.data(), b.x, np * 32, hipMemcpyDeviceToHost));
        unpack4(t4, np, 4, hs.x.data());
        HIPCHK(hipMemcpy(t4.data(), b.vel, np * 32, hipMemcpyDeviceToHost));
        unpack4(t4, np, 4, hs.vel.data());
        std::vector<double> *v3[] = {&hs.angvel, &hs.force, &hs.torque, &hs.angpos, &hs.xyzsum};
        double3 *d3p[] = {b.angvel, b.force, b.torque, b.angpos, b.xyzsum};
        for (int k = 0; k < 5; ++k)
            HIPCHK(hipMemcpy(v3[k]->data(), d3p[k], np * 24, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hs.es_dot.data(), b.es_dot, np * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hs.es.data(), b.es, np * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hs.ev_dot.data(), b.ev_dot, np * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hs.ev.data(), b.ev, np * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hs.p.data(), b.p, np * 8, hipMemcpyDeviceToHost));
    }
    if (particles && nb) {
        std::vector<uint8_t> broken(nb);
        std::vector<double> bd(nb * 4), bo(nb * 4);
        HIPCHK(hipMemcpy(broken.data(), b.bond_broken, nb, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(bd.data(), b.bond_delta, nb * 32, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(bo.data(), b.bond_omega, nb * 32, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nb; ++k) {
            if (broken[k]) hs.bonds[2 * k] = (uint32_t)nb;  // cohesion.cuh:211
            hs.bdt[3 * k] = bd[4 * k]; hs.bdt[3 * k + 1] = bd[4 * k + 1];
            hs.bdt[3 * k + 2] = bd[4 * k + 2]; hs.bdn[k] = bd[4 * k + 3];
            hs.bot[3 * k] = bo[4 * k]; hs.bot[3 * k + 1] = bo[4 * k + 1];
            hs.bot[3 * k + 2] = bo[4 * k + 2]; hs.bon[k] = bo[4 * k + 3];
        }
    }
    if (fluid_fields && h.fluid) {
        DarcyBuffers &d = ctx->dd;
        const size_t c = dcells(h);
        std::vector<double> p(c), phi(c), dphi(c), v(3 * c);
        HIPCHK(hipMemcpy(p.data(), d.p, c * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(phi.data(), d.phi, c * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(dphi.data(), d.dphi, c * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(v.data(), d.v, 3 * c * 8, hipMemcpyDeviceToHost));
        for (uint32_t z = 0; z < h.num[2]; ++z)
            for (uint32_t y = 0; y < h.num[1]; ++y)
                for (uint32_t x = 0; x < h.num[0]; ++x) {
                    size_t f = x + (size_t)h.num[0] * (y + (size_t)h.num[1] * z);
                    uint32_t ci = host_didx(h, x, y, z);
                    hs.p_f[f] = p[ci];
                    hs.phi[f] = phi[ci];
                    hs.dphi[f] = dphi[ci];
                    hs.v_f[3 * f] = v[3 * ci];
                    hs.v_f[3 * f + 1] = v[3 * ci + 1];
                    hs.v_f[3 * f + 2] = v[3 * ci + 2];
                }
        if (np) {
            std::vector<double> t4(np * 4);
            HIPCHK(hipMemcpy(t4.data(), d.f_p, np * 32, hipMemcpyDeviceToHost));
            unpack4(t4, np, 3, hs.f_p.data());
        }
    }
    return SPH_OK;
}

// ------------------------------------------------------------------
// one time step
// ------------------------------------------------------------------

static StepArgs step_args(const sph_ctx *ctx)
{
    const sph_state &h = ctx->hs.h;
    StepArgs a{};
    for (int i = 0; i < 3; ++i) {
        a.grid.origo[i] = h.origo[i];
        a.grid.L[i] = h.L[i];
        a.grid.num[i] = h.num[i];
        a.prm.g[i] = h.g[i];
    }
    a.grid.periodic = h.periodic;
    a.prm.k_n = h.k_n; a.prm.k_t = h.k_t; a.prm.k_r = h.k_r; a.prm.E = h.E;
    a.prm.gamma_n = h.gamma_n; a.prm.gamma_t = h.gamma_t; a.prm.gamma_r = h.gamma_r;
    a.prm.mu_s = h.mu_s; a.prm.mu_d = h.mu_d; a.prm.mu_r = h.mu_r;
    a.prm.gamma_wn = h.gamma_wn; a.prm.gamma_wt = h.gamma_wt;
    a.prm.mu_ws = h.mu_ws; a.prm.mu_wd = h.mu_wd; a.prm.rho = h.rho;
    a.prm.contactmodel = h.contactmodel;
    a.prm.kappa = h.kappa; a.prm.db = h.db; a.prm.V_b = h.V_b;
    a.prm.lambda_bar = h.lambda_bar; a.prm.sigma_b = h.sigma_b; a.prm.tau_b = h.tau_b;
    a.prm.sigma0_A = h.sigma0_A; a.prm.sigma0_f = h.sigma0_f;
    {
        // the same IEEE operations as the per-contact form of the law
        // (volatile keeps the host compiler from reassociating or contracting)
        volatile double nu = h.k_r, E = h.E;
        volatile double E_star = E / (2.0 * (1.0 - nu * nu));
        volatile double G = E / (2.0 * (1.0 + nu));
        volatile double G_star = G / (2.0 * (2.0 - nu));
        a.prm.E_star = E_star;
        a.prm.G_star = G_star;
    }
    a.walls.nw = h.nw;
    for (int w = 0; w < MAXW; ++w) a.walls.wmode[w] = w < (int)h.nw ? h.wmode[w] : 0;
    a.np = h.np;
    a.nb0 = h.nb0;
    a.nc = ctx->nc;
    a.dt = h.time_dt;
    return a;
}

static DarcyArgs darcy_args(const sph_ctx *ctx)
{
    const sph_state &h = ctx->hs.h;
    DarcyArgs d{};
    d.nx = h.num[0]; d.ny = h.num[1]; d.nz = h.num[2];
    d.L0 = h.L[0]; d.L1 = h.L[1]; d.L2 = h.L[2];
    d.dx = h.L[0] / h.num[0]; d.dy = h.L[1] / h.num[1]; d.dz = h.L[2] / h.num[2];
    d.bc_xn = h.bc_xn; d.bc_xp = h.bc_xp; d.bc_yn = h.bc_yn; d.bc_yp = h.bc_yp;
    d.bc_bot = h.bc_bot; d.bc_top = h.bc_top;
    d.bc_bot_flux = h.bc_bot_flux; d.bc_top_flux = h.bc_top_flux;
    d.mu = h.mu; d.beta_f = h.beta_f; d.rho_f = h.rho_f; d.k_c = h.k_c; d.c_phi = h.c_phi;
    d.ndem = h.ndem; d.dt = h.time_dt;
    d.wall0_iz = ctx->wall0_iz;
    d.periodic = h.periodic;
    return d;
}

// The network of sph_set_darcy_net for the loaded grid, made on first use.
static int net_ready(sph_ctx *ctx)
{
    if (ctx->net || ctx->net_host.empty() || ctx->net_max_corr <= 0) return SPH_OK;
    const sph_state &h = ctx->hs.h;
    const uint32_t periodic = (h.bc_xn == 2 && h.bc_xp == 2 ? 1u : 0u) | (h.bc_yn == 2 && h.bc_yp == 2 ? 2u : 0u);
    ctx->net = sph_net_create(ctx->device, h.num[0], h.num[1], h.num[2], periodic);
    if (!ctx->net) return fail(ctx, SPH_ERR_VALUE, std::string("Darcy network: ") + sph_net_last_error());
    const size_t V = (size_t)h.num[0] * h.num[1] * h.num[2];
    if (hipMalloc((void **)&ctx->net_params, ctx->net_host.size() * 4) != hipSuccess ||
        hipMalloc((void **)&ctx->net_in, V * 3 * 4) != hipSuccess ||
        hipMalloc((void **)&ctx->net_out, V * 4) != hipSuccess) {
        net_release(ctx);
        return fail(ctx, SPH_ERR_HIP, "Darcy network: device allocation failed");
    }
    HIPCHK(hipMemcpyAsync(ctx->net_params, ctx->net_host.data(), ctx->net_host.size() * 4,
                          hipMemcpyHostToDevice, ctx->stream));
    net_pack(ctx->net, ctx->net_params, ctx->stream);
    return sync(ctx);
}

// Darcy branch of the loop body, src/device.cu:1901-2426
static int darcy_step(sph_ctx *ctx, std::ofstream *convlog)
{
    sph_state &h = ctx->hs.h;
    hipStream_t s = ctx->stream;
    DarcyBuffers &d = ctx->dd;
    if (h.nw > 0 && (h.wmode[0] == 1 || h.wmode[0] == 3))
        ctx->wall0_iz = (unsigned int)(ctx->host_wx0 / (h.L[2] / h.num[2]));
    const DarcyArgs da = darcy_args(ctx);
    const StepArgs a = step_args(ctx);
    if (h.np > 0) {
        ProfScope ps(ctx, SPH_K_DARCY_SETUP);
        launch_darcy_ghosts(s, da, d.p, 1);
        launch_darcy_pressure_gradient(s, da, d);
        launch_darcy_ghosts(s, da, d.grad_p, 3);
        launch_darcy_pressure_force(s, a, da, d, ctx->db);
    }
    if ((ctx->iter % h.ndem) != 0) return SPH_OK;
    {
        ProfScope ps(ctx, SPH_K_DARCY_SETUP);
        launch_darcy_porosities(s, a, da, d, ctx->db, ctx->iter);
        if (h.periodic == 2) launch_darcy_porosities_to_edges(s, da, d);
        if ((h.p_mod_A > 1.0e-5 || h.p_mod_A < -1.0e-5) && h.p_mod_f > 1.0e-7) {
            double new_pressure = ctx->p_top_orig +
                                  h.p_mod_A * sin(2.0 * M_PI * h.p_mod_f * h.time_current + h.p_mod_phi);
            launch_darcy_top_pressure(s, da, d, new_pressure);
        }
        launch_darcy_permeabilities(s, da, d);
        launch_darcy_ghosts(s, da, d.phi, 1);
        launch_darcy_ghosts(s, da, d.k, 1);
        if (ctx->iter == 0) launch_darcy_copy_interior(s, da, d.p, d.p_old);
        launch_darcy_ghosts_flux(s, da, d.p, d.k);
    }
    ctx->stats.fluid_steps++;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    (void)hipEventRecord(t0, s);
    // nijac == 0 (src/device.cu:2195-2240): p_old, the ghosts of p and the
    // explicit solution, with the constants every sweep of this solve shares
    {
        ProfScope ps(ctx, SPH_K_DARCY_SWEEP);
        launch_darcy_copy_interior(s, da, d.p, d.p_old);
        launch_darcy_ghosts(s, da, d.p, 1);
        launch_darcy_prepare(s, da, d);
        launch_fill_u32(s, (uint32_t *)d.maxres, 2 * DARCY_MAXRES_WORDS, 0u);
    }
    // The sweeps (src/device.cu:2184-2386) ping-pong between the two
    // buffers.  The convergence test of check sweep j (j % 10 == 0) runs on
    // the device at the start of sweep j + 1, and once it ends the solve
    // every later sweep returns at once.  The host enqueues sweeps in batches
    // (the first as long as the last solve, each ending with a deciding
    // sweep) and reads the solve state once per batch.
    double *buf[2] = {d.p, d.p_new};
    int rc0 = net_ready(ctx);
    if (rc0) return rc0;
    int corrections = 0;
    unsigned int nijac = 0;
    unsigned long long state = 0, last_nijac = 0, last_bits = 0;
    unsigned long long want = std::max(12u, (ctx->jacobi_hint + 9u) / 10u * 10u + 2u);
    while (nijac < h.maxiter && state == 0) {
        const unsigned int end = (unsigned int)std::min<unsigned long long>(h.maxiter, want);
        for (; nijac < end; ++nijac) {
            const bool check = (nijac % 10 == 0);
            // two check slots alternate: a check sweep zeroes the other one
            unsigned long long *slot = d.maxres + ((nijac / 10) & 1u) * RED_WORDS;
            unsigned long long *next = d.maxres + (((nijac / 10) & 1u) ^ 1u) * RED_WORDS;
            const bool decides = nijac > 0 && (nijac - 1) % 10 == 0;
            const unsigned long long *prev = decides ? d.maxres + (((nijac - 1) / 10) & 1u) * RED_WORDS : nullptr;
            ProfScope ps(ctx, SPH_K_DARCY_SWEEP);
            // setDarcyTopWallFixedFlow (bc_top == 1) writes dev_darcy_p, which
            // the copy of p_new into p overwrites: it has no effect
            // (device.cu:2284-2331)
            // learned correction of the iterate this deciding sweep wrote
            // (its kernels do nothing once the solve has ended); the sweep
            // clears the correction's scale slot
            const bool correct = decides && ctx->net && corrections < ctx->net_max_corr;
            launch_darcy_sweep(s, da, d, buf[nijac & 1u], buf[(nijac & 1u) ^ 1u],
                               check ? slot : nullptr, check ? next : nullptr, prev, h.tolerance,
                               nijac - 1u, false, correct);
            if (correct) {
                double *pc = buf[(nijac & 1u) ^ 1u];
                launch_darcy_residual_max(s, da, d, pc, false);
                // the input goes straight into the network's BF16 rows, and
                // the correction refreshes the ghosts of the cells it changes
                launch_darcy_net_input(s, da, d, pc, nullptr, false, net_input_bf16(ctx->net));
                int nrc = net_forward_packed(ctx->net, ctx->net_params, nullptr, ctx->net_out, s);
                if (nrc) return fail(ctx, nrc, sph_net_last_error());
                launch_darcy_net_apply(s, da, d, pc, ctx->net_out);
                ++corrections;
            }
        }
        if (nijac == h.maxiter && (h.maxiter - 1) % 10 == 0)
            launch_darcy_decide(s, d, d.maxres + (((h.maxiter - 1) / 10) & 1u) * RED_WORDS, h.tolerance,
                                h.maxiter - 1);
        HIPCHK(hipMemcpyAsync(ctx->h_state, d.maxres + DARCY_STATE, 32, hipMemcpyDeviceToHost, s));
        int rc = sync(ctx);
        if (rc) return rc;
        state = ctx->h_state[0];
        last_nijac = ctx->h_state[1];
        last_bits = ctx->h_state[2];
        want = (unsigned long long)nijac + 20u;
    }
    ctx->stats.corrections += ctx->h_state[3];
    // sweeps that ran: up to the deciding check, or all maxiter
    const unsigned int ran = state != 0 ? (unsigned int)last_nijac + 1u : h.maxiter;
    d.p = buf[ran & 1u];
    d.p_new = buf[(ran & 1u) ^ 1u];
    ctx->stats.sweeps += ran;
    ctx->jacobi_hint = ran;
    double max_norm_res;
    std::memcpy(&max_norm_res, &last_bits, 8);
    ctx->stats.last_nijac = (uint32_t)last_nijac;
    ctx->stats.last_max_residual = max_norm_res;
    if (state == 2) {
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        return fail(ctx, SPH_ERR_NAN, "normalized residual is NaN: the system has "
                                      "probably become unstable");
    }
    if (state == 1) {
        if (ctx->iter % 10 == 0) {
            ctx->conv.emplace_back(ctx->iter + 1, (unsigned int)last_nijac);
            if (convlog) *convlog << ctx->iter + 1 << '\t' << last_nijac << std::endl;
        }
    } else if (h.maxiter > 0) {
        ctx->conv.emplace_back(ctx->iter + 1, h.maxiter - 1);
        if (convlog) *convlog << ctx->iter + 1 << '\t' << h.maxiter - 1 << std::endl;
        ctx->stats.nonconverged++;
        fprintf(stderr, "\nIteration %ld, time %g s: Error, the pressure solution in the fluid "
                        "calculations did not converge. Try increasing the value of "
                        "'darcy.maxiter' (%u) or increase 'darcy.tolerance' (%g).\n",
                ctx->iter, ctx->iter * h.time_dt, h.maxiter, h.tolerance);
    }
    (void)hipEventRecord(t1, s);
    (void)hipEventSynchronize(t1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    ctx->stats.solve_ms += ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (ctx->filetimeclock + h.time_dt < h.time_file_dt) launch_darcy_zero_interior(s, da, d.dphi);
    launch_darcy_ghosts(s, da, d.p, 1);
    launch_darcy_velocities(s, da, d);
    return SPH_OK;
}

// ------------------------------------------------------------------
// Restart side car (no reference counterpart).  The v2.15 file holds the
// particle, wall, bond and fluid state but not the contact history, the
// previous accelerations of the TY3 integrator or the time loop's counters,
// so a run continued from an output file differs from the uninterrupted one
// (src/file_io.cpp:420-731; TY3 restarts with acc0 = 0,
// src/integration.cuh:74-81).  The side car, written beside each output
// file, carries exactly those, and a run loaded from that file continues
// bit for bit.
// ------------------------------------------------------------------
#define RESTART_VERSION 1u

struct RestartHeader {
    char magic[4];            // "SPHR"
    uint32_t version;
    double time_current;      // of the state file it belongs to
    uint32_t step_count, np, nc, nw;
    int64_t iter;             // steps since the run started (0: a fresh start)
    int32_t velocity_state, pad;
    double p_top_orig;        // the Darcy top pressure without its modulation
    uint64_t occupied;        // occupied contact slots that follow
};
// followed by: w_acc[nw], acc[np*3], angacc[np*3] (FP64), cmask[np] (u64),
// then per occupied slot in particle and slot order the contact word (u32),
// and then their tangential springs (3 FP64 each)

// The contact arrays as row-major host arrays (slot k of grain a at a nc +
// k, the order of the reference's dev_contacts), converted from and to the
// device layout (ct_slot).  Either array may be null.
static int rows_to_host(sph_ctx *ctx, uint32_t *rows, double *dts)
{
    const size_t np = ctx->np, nc = ctx->nc, ne = ct_entries(ctx->np, ctx->nc);
    if (!np) return SPH_OK;
    if (rows) {
        std::vector<uint32_t> d(ne);
        HIPCHK(hipMemcpy(d.data(), ctx->db.contacts, ne * 4, hipMemcpyDeviceToHost));
        for (size_t a = 0; a < np; ++a)
            for (size_t k = 0; k < nc; ++k) rows[a * nc + k] = d[ct_slot((uint32_t)a, (uint32_t)k, (uint32_t)nc)];
    }
    if (dts) {
        std::vector<double> d(ne * 3);
        HIPCHK(hipMemcpy(d.data(), ctx->db.delta_t, ne * 24, hipMemcpyDeviceToHost));
        for (size_t a = 0; a < np; ++a)
            for (size_t k = 0; k < nc; ++k) {
                const size_t q = ct_slot((uint32_t)a, (uint32_t)k, (uint32_t)nc);
                for (int c = 0; c < 3; ++c) dts[(a * nc + k) * 3 + c] = d[3 * q + c];
            }
    }
    return SPH_OK;
}

static int rows_to_device(sph_ctx *ctx, const uint32_t *rows, const double *dts)
{
    const size_t np = ctx->np, nc = ctx->nc, ne = ct_entries(ctx->np, ctx->nc);
    if (!np) return SPH_OK;
    if (rows) {
        std::vector<uint32_t> d(ne, (uint32_t)np);
        for (size_t a = 0; a < np; ++a)
            for (size_t k = 0; k < nc; ++k) d[ct_slot((uint32_t)a, (uint32_t)k, (uint32_t)nc)] = rows[a * nc + k];
        HIPCHK(hipMemcpy(ctx->db.contacts, d.data(), ne * 4, hipMemcpyHostToDevice));
    }
    if (dts) {
        std::vector<double> d(ne * 3, 0.0);
        for (size_t a = 0; a < np; ++a)
            for (size_t k = 0; k < nc; ++k) {
                const size_t q = ct_slot((uint32_t)a, (uint32_t)k, (uint32_t)nc);
                for (int c = 0; c < 3; ++c) d[3 * q + c] = dts[(a * nc + k) * 3 + c];
            }
        HIPCHK(hipMemcpy(ctx->db.delta_t, d.data(), ne * 24, hipMemcpyHostToDevice));
    }
    return SPH_OK;
}

static int write_restart(sph_ctx *ctx, const std::string &path)
{
    const sph_state &h = ctx->hs.h;
    const size_t np = ctx->np, nc = ctx->nc;
    int rc = sync(ctx);
    if (rc) return rc;
    RestartHeader H{};
    std::memcpy(H.magic, "SPHR", 4);
    H.version = RESTART_VERSION;
    H.time_current = h.time_current;
    H.step_count = h.time_step_count;
    H.np = (uint32_t)np;
    H.nc = (uint32_t)nc;
    H.nw = h.nw;
    H.iter = ctx->iter;
    H.velocity_state = ctx->velocity_state;
    H.p_top_orig = ctx->p_top_orig;
    std::vector<double> wacc(MAXW), acc(np * 3), angacc(np * 3);
    std::vector<uint64_t> mask(np);
    std::vector<uint32_t> rows(np * nc);
    std::vector<double> dts(np * nc * 3);
    HIPCHK(hipMemcpy(wacc.data(), ctx->db.w_acc, MAXW * 8, hipMemcpyDeviceToHost));
    if (np) {
        HIPCHK(hipMemcpy(acc.data(), ctx->db.acc, np * 24, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(angacc.data(), ctx->db.angacc, np * 24, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(mask.data(), ctx->db.cmask, np * 8, hipMemcpyDeviceToHost));
        const int rc = rows_to_host(ctx, rows.data(), dts.data());
        if (rc) return rc;
    }
    std::vector<uint32_t> words;
    std::vector<double> springs;
    for (size_t a = 0; a < np; ++a)
        for (size_t k = 0; k < nc; ++k)
            if ((mask[a] >> k) & 1ull) {
                words.push_back(rows[a * nc + k]);
                springs.insert(springs.end(), &dts[(a * nc + k) * 3], &dts[(a * nc + k) * 3] + 3);
            }
    H.occupied = words.size();
    FILE *fp = fopen(path.c_str(), "wb");
    if (!fp) return fail(ctx, SPH_ERR_FILE, "cannot write " + path);
    bool ok = fwrite(&H, sizeof(H), 1, fp) == 1;
    ok = ok && fwrite(wacc.data(), 8, h.nw, fp) == h.nw;
    ok = ok && fwrite(acc.data(), 8, np * 3, fp) == np * 3;
    ok = ok && fwrite(angacc.data(), 8, np * 3, fp) == np * 3;
    ok = ok && fwrite(mask.data(), 8, np, fp) == np;
    ok = ok && fwrite(words.data(), 4, words.size(), fp) == words.size();
    ok = ok && fwrite(springs.data(), 8, springs.size(), fp) == springs.size();
    ok = (fclose(fp) == 0) && ok;
    return ok ? SPH_OK : fail(ctx, SPH_ERR_FILE, "short write of " + path);
}

static int peek_restart(const std::string &path, RestartHeader &H)
{
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) return SPH_ERR_FILE;
    const bool ok = fread(&H, sizeof(H), 1, fp) == 1;
    fclose(fp);
    if (!ok || std::memcmp(H.magic, "SPHR", 4) != 0 || H.version != RESTART_VERSION)
        return SPH_ERR_FILE;
    return SPH_OK;
}

// Restores a side car into the loaded state, which must be the state file it
// was written with (same time, output number, grain and wall counts and slot
// count).
static int read_restart(sph_ctx *ctx, const std::string &path)
{
    const sph_state &h = ctx->hs.h;
    RestartHeader H{};
    if (peek_restart(path, H)) return fail(ctx, SPH_ERR_FILE, path + " is not a restart side car");
    if (H.np != ctx->np || H.nc != ctx->nc || H.nw != h.nw || H.time_current != h.time_current ||
        H.step_count != h.time_step_count)
        return fail(ctx, SPH_ERR_VALUE, path + " belongs to another state (time " +
                                            std::to_string(H.time_current) + ", output " +
                                            std::to_string(H.step_count) + ", " +
                                            std::to_string(H.np) + " grains, " +
                                            std::to_string(H.nc) + " slots)");
    const size_t np = ctx->np, nc = ctx->nc;
    std::vector<double> wacc(MAXW, 0.0), acc(np * 3), angacc(np * 3);
    std::vector<uint64_t> mask(np);
    std::vector<uint32_t> words(H.occupied);
    std::vector<double> springs(H.occupied * 3);
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) return fail(ctx, SPH_ERR_FILE, "cannot read " + path);
    bool ok = fseek(fp, (long)sizeof(H), SEEK_SET) == 0;
    ok = ok && fread(wacc.data(), 8, h.nw, fp) == h.nw;
    ok = ok && fread(acc.data(), 8, np * 3, fp) == np * 3;
    ok = ok && fread(angacc.data(), 8, np * 3, fp) == np * 3;
    ok = ok && fread(mask.data(), 8, np, fp) == np;
    ok = ok && fread(words.data(), 4, words.size(), fp) == words.size();
    ok = ok && fread(springs.data(), 8, springs.size(), fp) == springs.size();
    fclose(fp);
    if (!ok) return fail(ctx, SPH_ERR_FILE, path + " is truncated");
    std::vector<uint32_t> rows(np * nc, (uint32_t)np);
    std::vector<double> dts(np * nc * 3, 0.0);
    size_t j = 0;
    for (size_t a = 0; a < np; ++a)
        for (size_t k = 0; k < nc; ++k)
            if ((mask[a] >> k) & 1ull) {
                if (j >= words.size() || (words[j] & CT_IDX_MASK) >= np)
                    return fail(ctx, SPH_ERR_FILE, path + ": inconsistent contact slots");
                rows[a * nc + k] = words[j];
                dts[(a * nc + k) * 3] = springs[3 * j];
                dts[(a * nc + k) * 3 + 1] = springs[3 * j + 1];
                dts[(a * nc + k) * 3 + 2] = springs[3 * j + 2];
                ++j;
            }
    if (j != words.size()) return fail(ctx, SPH_ERR_FILE, path + ": inconsistent contact slots");
    hipStream_t s = ctx->stream;
    int rc = sync(ctx);
    if (rc) return rc;
    HIPCHK(hipMemcpy(ctx->db.w_acc, wacc.data(), MAXW * 8, hipMemcpyHostToDevice));
    if (np) {
        HIPCHK(hipMemcpy(ctx->db.acc, acc.data(), np * 24, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->db.angacc, angacc.data(), np * 24, hipMemcpyHostToDevice));
        rc = rows_to_device(ctx, rows.data(), dts.data());
        if (rc) return rc;
        HIPCHK(hipMemcpy(ctx->db.cmask, mask.data(), np * 8, hipMemcpyHostToDevice));
    }
    // the device copy of the step counter and time (read by the kernels)
    const StepClock c{(long long)H.iter, h.time_current};
    HIPCHK(hipMemcpyAsync(ctx->db.clock, &c, sizeof(c), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    ctx->iter = (long)H.iter;
    ctx->velocity_state = H.velocity_state;
    ctx->p_top_orig = H.p_top_orig;
    ctx->nl_force = 1;  // list and slot maps are rebuilt from the restored rows
    ctx->hash_ready = false;
    return SPH_OK;
}

// Output event (src/device.cu:2576-2680).  With dir == nullptr the state
// effects are applied without writing files.
static int output_event(sph_ctx *ctx, const std::string *dir, const std::string &sid)
{
    sph_state &h = ctx->hs.h;
    int rc = download(ctx, dir != nullptr, dir != nullptr);
    if (rc) return rc;
    ctx->host_wx0 = h.nw > 0 ? h.w_nx[3] : 0.0;
    if (h.fluid) launch_darcy_zero_interior(ctx->stream, darcy_args(ctx), ctx->dd.dphi);
    if (ctx->ns) {
        rc = ns_output_event(ctx->ns, ctx->stream);
        if (rc) return fail(ctx, rc, sph_ns_last_error());
    }
    h.time_step_count += 1;
    if (dir) {
        char name[64];
        snprintf(name, sizeof(name), ".output%05u.bin", h.time_step_count);
        std::string err;
        rc = write_state_file((*dir + "/" + sid + name).c_str(), ctx->hs, err);
        if (rc) return fail(ctx, rc, err);
        snprintf(name, sizeof(name), ".contacts.%05u.bin", h.time_step_count);
        rc = write_restart(ctx, *dir + "/" + sid + name);
        if (rc) return rc;
        FILE *fp = fopen((*dir + "/" + sid + ".status.dat").c_str(), "w");
        if (fp) {
            fprintf(fp, "%2.4e %2.4e %d\n", h.time_current, 100.0 * h.time_current / h.time_total,
                    (int)h.time_step_count);
            fclose(fp);
        }
    }
    ctx->filetimeclock = 0.0;
    return SPH_OK;
}

// The bond forces are added inside integrate unless something between the
// bonds and integrate reads or adds to the forces: the Darcy pressure force
// (added after the bonds, src/device.cu:1970-1990) or the wmode-3 shear
// stress (reads the forces, src/integration.cuh:457-484).
static bool gather_in_integrate(const sph_ctx *ctx)
{
    const sph_state &h = ctx->hs.h;
    return !h.fluid && !ctx->ns && !(h.nw > 0 && h.wmode[0] == 3);
}

static int step_once(sph_ctx *ctx, std::ofstream *convlog, const std::string *dir,
                     const std::string &sid)
{
    sph_state &h = ctx->hs.h;
    hipStream_t s = ctx->stream;
    int rc;
    if (h.np > 0) {
        if (h.adaptive == 1 && h.nw > 0) {  // updateGridSize, device.cu:370-400
            rc = wait_walls(ctx);
            if (rc) return rc;
            HIPCHK(hipMemcpyAsync(ctx->h_d, (const char *)ctx->db.w_nx + 24, 8,
                                  hipMemcpyDeviceToHost, s));
            rc = sync(ctx);
            if (rc) return rc;
            h.L[2] = *ctx->h_d;
        }
        const StepArgs a = step_args(ctx);
        if (ctx->nl_on) {
            // Lists are rebuilt (sort, reorder, list build) only on steps
            // whose device-side flag is set; the Darcy solve reads the
            // sorted arrays, so a fluid run rebuilds every step.
            const uint32_t cur = ctx->nl_parity;
            const uint32_t *cond = ctx->db.nl_flags + cur;
            const uint32_t force = ctx->nl_force || h.fluid || ctx->ns ? 1u : 0u;
            if (force || !ctx->hash_ready) {  // else the last integrate formed the hash
                ProfScope ps(ctx, SPH_K_HASH);
                launch_hash_nl(s, a, ctx->nla, ctx->db, cur, force);
            }
            if (ctx->rb_blocks) {
                ProfScope ps(ctx, SPH_K_NLIST);
                launch_rebuild(s, a, ctx->nla, ctx->db, ctx->ncells, ctx->key_bits,
                               ctx->rb_blocks, cond, ctx->rb_spin);
            } else {
                { ProfScope ps(ctx, SPH_K_SORT); launch_sort(s, a.np, ctx->key_bits, ctx->db, cond); }
                { ProfScope ps(ctx, SPH_K_REORDER); launch_reorder(s, a, ctx->db, ctx->ncells, cond); }
                { ProfScope ps(ctx, SPH_K_NLIST); launch_nl_build(s, a, ctx->nla, ctx->db, cond); }
            }
            { ProfScope ps(ctx, SPH_K_TOPOLOGY);
              launch_topology(s, a, ctx->db, true, ctx->nla.range == 1, cur); }
            rc = wait_walls(ctx);
            if (rc) return rc;
            { ProfScope ps(ctx, SPH_K_CONTACT); launch_contact(s, a, ctx->db, true); }
            ctx->nl_parity ^= 1u;
            ctx->nl_force = 0;
        } else {
            { ProfScope ps(ctx, SPH_K_HASH); launch_hash(s, a, ctx->db); }
            { ProfScope ps(ctx, SPH_K_SORT); launch_sort(s, a.np, ctx->key_bits, ctx->db, nullptr); }
            { ProfScope ps(ctx, SPH_K_REORDER); launch_reorder(s, a, ctx->db, ctx->ncells, nullptr); }
            if (h.contactmodel != 1) {
                ProfScope ps(ctx, SPH_K_TOPOLOGY);
                launch_topology(s, a, ctx->db, false, false, 0u);
            }
            rc = wait_walls(ctx);
            if (rc) return rc;
            { ProfScope ps(ctx, SPH_K_CONTACT); launch_contact(s, a, ctx->db, false); }
        }
        if (h.nb0 > 0) {
            ProfScope ps(ctx, SPH_K_BONDS);
            launch_bonds(s, a, ctx->db, !gather_in_integrate(ctx));
        }
    }
    if (h.fluid) {
        rc = darcy_step(ctx, convlog);
        if (rc) return rc;
    }
    if (ctx->ns && h.np > 0) {
        // the Navier-Stokes fluid between the grains' forces and the
        // integration, its force on each grain added (src/device.cu:1213-1898)
        rc = ns_coupled_step(ctx->ns, h.np, (const double *)ctx->db.x, (const double *)ctx->db.vel,
                             (const double *)ctx->db.x_s, ctx->db.cell_start, ctx->db.cell_end, ctx->db.vals,
                             h.origo, h.periodic, (double *)ctx->db.force,
                             h.nw > 0 && (h.wmode[0] == 1 || h.wmode[0] == 3) ? ctx->host_wx0 : -1.0, s);
        if (rc) return fail(ctx, rc, sph_ns_last_error());
    }
    if (h.np > 0) {
        const StepArgs a = step_args(ctx);
        if (h.nw > 0 && h.wmode[0] == 3) launch_shear_stress(s, a, ctx->db);
        int change_velocity_state = 0;  // device.cu:2461-2469
        if (h.time_current >= 10.0 && h.time_current < 15.0 && ctx->velocity_state == 1) {
            change_velocity_state = 1;
            ctx->velocity_state = 2;
        } else if (h.time_current >= 15.0 && ctx->velocity_state == 2) {
            change_velocity_state = -1;
            ctx->velocity_state = 1;
        }
        // Without fluid the next step's cell hash is formed by integrate (list
        // mode); the fluid step rebuilds and hashes every step.
        NextHash nh{};
        const bool fuse_hash = ctx->nl_on && !h.fluid;
        if (fuse_hash) {
            nh.keys = ctx->db.keys;
            nh.vals = ctx->db.vals;
            nh.cellc = ctx->db.cellc;
            nh.x_ref = ctx->db.x_ref;
            nh.disp = ctx->db.nl_disp;
            nh.flags = ctx->db.nl_flags;
            nh.cur = ctx->nl_parity;  // already the next step's word
            nh.disp2 = ctx->nla.disp2;
        }
        { ProfScope ps(ctx, SPH_K_INTEGRATE);
          launch_integrate(s, a, ctx->db, change_velocity_state, 10.0, h.tau_x,
                           gather_in_integrate(ctx), nh); }
        ctx->hash_ready = fuse_hash;
        if (h.nw > 0) {
            HIPCHK(hipEventRecord(ctx->ev_integrated, s));
            HIPCHK(hipStreamWaitEvent(ctx->aux, ctx->ev_integrated, 0));
            {
                ProfScope ps(ctx, SPH_K_WALLS, ctx->aux);
                launch_integrate_walls(ctx->aux, a, ctx->db);
            }
            // the clock's last readers this step are integrate and the walls
            launch_clock_advance(ctx->aux, ctx->db.clock, h.time_dt);
            HIPCHK(hipEventRecord(ctx->ev_walls, ctx->aux));
            ctx->walls_pending = true;
        } else {
            launch_clock_advance(s, ctx->db.clock, h.time_dt);
        }
    } else if (ctx->db.clock) {
        launch_clock_advance(s, ctx->db.clock, h.time_dt);
    }
    h.time_current += h.time_dt;
    ctx->filetimeclock += h.time_dt;
    ctx->iter++;
    if (ctx->filetimeclock >= h.time_file_dt) {
        rc = output_event(ctx, dir, sid);
        if (rc) return rc;
    }
    return SPH_OK;
}

// ------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------

extern "C" {

sph_ctx *sph_create(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        g_last_error = "no HIP device available";
        return nullptr;
    }
    if (device < 0) device = 0;
    if (device >= n) {
        g_last_error = "device index out of range";
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        g_last_error = "hipSetDevice failed";
        return nullptr;
    }
    sph_ctx *ctx = new sph_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_integrated, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_walls, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void **)&ctx->h_err, 128, hipHostMallocDefault) != hipSuccess) {
        g_last_error = "stream or pinned allocation failed";
        delete ctx;
        return nullptr;
    }
    ctx->h_u64 = (unsigned long long *)((char *)ctx->h_err + 16);
    ctx->h_d = (double *)((char *)ctx->h_err + 32);
    ctx->h_state = (unsigned long long *)((char *)ctx->h_err + 64);
    return ctx;
}

int sph_ns_attach(sph_ns *ns, sph_ctx *ctx)
{
    if (!ctx) return fail(ctx, SPH_ERR_ARG, "null context");
    if (!ns) {
        if (ctx->ns) ns_set_attached(ctx->ns, 0);
        ctx->ns = nullptr;
        return SPH_OK;
    }
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    const sph_state &h = ctx->hs.h;
    if (h.fluid) return fail(ctx, SPH_ERR_VALUE, "the context already has a Darcy fluid");
    uint32_t num[3];
    double L[3];
    ns_grid(ns, num, L);
    for (int k = 0; k < 3; ++k)
        if (num[k] != h.num[k] || L[k] != h.L[k])
            return fail(ctx, SPH_ERR_VALUE, "the fluid grid must be the context's grid (num and L)");
    if (h.contactmodel == 1) return fail(ctx, SPH_ERR_VALUE, "contact model 1 is not coupled to a fluid here");
    if (ctx->ns && ctx->ns != ns) ns_set_attached(ctx->ns, 0);
    ctx->ns = ns;
    ns_set_attached(ns, 1);
    return SPH_OK;
}

void sph_destroy(sph_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->ns) ns_set_attached(ctx->ns, 0);
    (void)hipSetDevice(ctx->device);
    free_device(ctx);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->ev_integrated) (void)hipEventDestroy(ctx->ev_integrated);
    if (ctx->ev_walls) (void)hipEventDestroy(ctx->ev_walls);
    if (ctx->h_err) (void)hipHostFree(ctx->h_err);
    if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *sph_last_error(const sph_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

static int copy_in(sph_ctx *ctx, const sph_state *s)
{
    HostState &hs = ctx->hs;
    hs.h = *s;
    const size_t np = s->np, nb = s->nb0;
    if (np && (!s->x || !s->vel)) return fail(ctx, SPH_ERR_ARG, "x and vel are required");
    hs.alloc_particles(np);
    auto cp = [](std::vector<double> &dst, const double *src, size_t n) {
        if (src && n) std::memcpy(dst.data(), src, n * 8);
    };
    cp(hs.x, s->x, np * 4); cp(hs.vel, s->vel, np * 4); cp(hs.xyzsum, s->xyzsum, np * 3);
    cp(hs.force, s->force, np * 3); cp(hs.angpos, s->angpos, np * 3);
    cp(hs.angvel, s->angvel, np * 3); cp(hs.torque, s->torque, np * 3);
    cp(hs.es_dot, s->es_dot, np); cp(hs.es, s->es, np); cp(hs.ev_dot, s->ev_dot, np);
    cp(hs.ev, s->ev, np); cp(hs.p, s->p, np);
    if (s->color && np) std::memcpy(hs.color.data(), s->color, np * 4);
    hs.alloc_bonds(nb);
    if (nb) {
        if (!s->bonds) return fail(ctx, SPH_ERR_ARG, "bonds array missing");
        std::memcpy(hs.bonds.data(), s->bonds, nb * 8);
        cp(hs.bdn, s->bonds_delta_n, nb); cp(hs.bdt, s->bonds_delta_t, nb * 3);
        cp(hs.bon, s->bonds_omega_n, nb); cp(hs.bot, s->bonds_omega_t, nb * 3);
    }
    if (s->fluid) {
        const size_t cells = (size_t)s->num[0] * s->num[1] * s->num[2];
        hs.alloc_cells(cells);
        cp(hs.v_f, s->v_f, cells * 3); cp(hs.p_f, s->p_f, cells);
        cp(hs.phi, s->phi, cells); cp(hs.dphi, s->dphi, cells);
        if (s->p_constant) std::memcpy(hs.p_constant.data(), s->p_constant, cells * 4);
        hs.f_p.assign(np * 3, 0.0);
        cp(hs.f_p, s->f_p, np * 3);
    }
    return SPH_OK;
}

int sph_set_contact_slots(sph_ctx *ctx, int nc)
{
    if (!ctx) return fail(ctx, SPH_ERR_ARG, "null context");
    if (nc != 32 && nc != 64) return fail(ctx, SPH_ERR_ARG, "contact slots must be 32 or 64");
    if (ctx->loaded && (uint32_t)nc != ctx->nc)
        return fail(ctx, SPH_ERR_STATE, "set the contact slot count before loading a state");
    ctx->nc = (uint32_t)nc;
    return SPH_OK;
}

int sph_set_neighbor_lists(sph_ctx *ctx, double skin_fraction, int cap)
{
    if (!ctx) return fail(ctx, SPH_ERR_ARG, "null context");
    if (!(skin_fraction >= 0.0 && skin_fraction <= 1.0))
        return fail(ctx, SPH_ERR_ARG, "skin fraction must lie in [0, 1]");
    if (cap < 1 || cap > 128) return fail(ctx, SPH_ERR_ARG, "list capacity must lie in [1, 128]");
    ctx->nl_frac = skin_fraction;
    ctx->nl_cap = (uint32_t)cap;
    return SPH_OK;
}

int sph_set_rebuild_kernel(sph_ctx *ctx, int fused)
{
    if (!ctx) return fail(ctx, SPH_ERR_ARG, "null context");
    ctx->rb_fused = fused != 0;
    return SPH_OK;
}

int sph_set_barrier_spin_limit(sph_ctx *ctx, uint32_t sleeps)
{
    if (!ctx) return fail(ctx, SPH_ERR_ARG, "null context");
    ctx->rb_spin = sleeps;
    return SPH_OK;
}

int sph_get_neighbor_stats(sph_ctx *ctx, sph_neighbor_stats *out)
{
    if (!ctx || !out) return fail(ctx, SPH_ERR_ARG, "null argument");
    *out = sph_neighbor_stats{};
    if (!ctx->loaded || !ctx->nl_on) return SPH_OK;
    (void)hipSetDevice(ctx->device);
    uint32_t f[NLF_DMAX];
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(f, ctx->db.nl_flags, sizeof(f), hipMemcpyDeviceToHost));
    if (f[3] > 0) {  // lists exist
        unsigned long long *tmp = nullptr;
        HIPCHK(hipMalloc((void **)&tmp, 16));
        unsigned long long h2[2] = {0ull, 0ull};
        hipError_t e = hipMemcpy(tmp, h2, 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            launch_nl_stats(ctx->stream, ctx->db, ctx->np, ctx->nla.cap, tmp);
            e = hipStreamSynchronize(ctx->stream);
        }
        if (e == hipSuccess) e = hipMemcpy(h2, tmp, 16, hipMemcpyDeviceToHost);
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(ctx, SPH_ERR_HIP, std::string("list statistics: ") + hipGetErrorString(e));
        out->entries = h2[0];
        out->longest = (uint32_t)h2[1];
    }
    out->enabled = 1;
    out->range = ctx->nla.range;
    out->cap = ctx->nla.cap;
    out->overflow = f[NLF_STICKY] ? 1u : 0u;
    out->skin = ctx->nla.skin;
    out->rebuilds = f[NLF_COUNT];
    out->full_steps = f[NLF_FULL];
    return SPH_OK;
}

int sph_load_state(sph_ctx *ctx, const sph_state *s)
{
    if (!ctx || !s) return fail(ctx, SPH_ERR_ARG, "null argument");
    (void)hipSetDevice(ctx->device);
    free_device(ctx);
    int rc = copy_in(ctx, s);
    if (rc) return rc;
    rc = check_values(ctx, ctx->hs);
    if (rc) return rc;
    return upload(ctx);
}

int sph_load_file(sph_ctx *ctx, const char *path, int fluid)
{
    if (!ctx || !path) return fail(ctx, SPH_ERR_ARG, "null argument");
    (void)hipSetDevice(ctx->device);
    free_device(ctx);
    std::string err;
    int rc = read_state_file(path, fluid, ctx->hs, err);
    if (rc) return fail(ctx, rc, err);
    rc = check_values(ctx, ctx->hs);
    if (rc) return rc;
    return upload(ctx);
}

int sph_get_header(sph_ctx *ctx, sph_state *hdr)
{
    if (!ctx || !hdr) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    (void)hipSetDevice(ctx->device);
    int rc = download(ctx, false, false);
    if (rc) return rc;
    *hdr = ctx->hs.h;
    hdr->x = hdr->xyzsum = hdr->vel = hdr->force = hdr->angpos = hdr->angvel = hdr->torque = nullptr;
    hdr->es_dot = hdr->es = hdr->ev_dot = hdr->ev = hdr->p = nullptr;
    hdr->bonds = nullptr;
    hdr->bonds_delta_n = hdr->bonds_delta_t = hdr->bonds_omega_n = hdr->bonds_omega_t = nullptr;
    hdr->v_f = hdr->p_f = hdr->phi = hdr->dphi = hdr->f_p = nullptr;
    hdr->p_constant = nullptr;
    hdr->color = nullptr;
    return SPH_OK;
}

int sph_read_state(sph_ctx *ctx, sph_state *out)
{
    if (!ctx || !out) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    (void)hipSetDevice(ctx->device);
    int rc = download(ctx, true, true);
    if (rc) return rc;
    const HostState &hs = ctx->hs;
    sph_state keep = *out;
    *out = hs.h;
    const size_t np = hs.h.np, nb = hs.h.nb0;
    auto cp = [](double *dst, const std::vector<double> &src, size_t n) {
        if (dst && n) std::memcpy(dst, src.data(), n * 8);
    };
    cp(keep.x, hs.x, np * 4); cp(keep.xyzsum, hs.xyzsum, np * 3); cp(keep.vel, hs.vel, np * 4);
    cp(keep.force, hs.force, np * 3); cp(keep.angpos, hs.angpos, np * 3);
    cp(keep.angvel, hs.angvel, np * 3); cp(keep.torque, hs.torque, np * 3);
    cp(keep.es_dot, hs.es_dot, np); cp(keep.es, hs.es, np); cp(keep.ev_dot, hs.ev_dot, np);
    cp(keep.ev, hs.ev, np); cp(keep.p, hs.p, np);
    if (keep.color && np) std::memcpy(keep.color, hs.color.data(), np * 4);
    if (keep.bonds && nb) std::memcpy(keep.bonds, hs.bonds.data(), nb * 8);
    cp(keep.bonds_delta_n, hs.bdn, nb); cp(keep.bonds_delta_t, hs.bdt, nb * 3);
    cp(keep.bonds_omega_n, hs.bon, nb); cp(keep.bonds_omega_t, hs.bot, nb * 3);
    if (hs.h.fluid) {
        const size_t cells = (size_t)hs.h.num[0] * hs.h.num[1] * hs.h.num[2];
        cp(keep.v_f, hs.v_f, cells * 3); cp(keep.p_f, hs.p_f, cells);
        cp(keep.phi, hs.phi, cells); cp(keep.dphi, hs.dphi, cells);
        if (keep.p_constant) std::memcpy(keep.p_constant, hs.p_constant.data(), cells * 4);
        cp(keep.f_p, hs.f_p, np * 3);
    }
    out->x = keep.x; out->xyzsum = keep.xyzsum; out->vel = keep.vel; out->force = keep.force;
    out->angpos = keep.angpos; out->angvel = keep.angvel; out->torque = keep.torque;
    out->es_dot = keep.es_dot; out->es = keep.es; out->ev_dot = keep.ev_dot; out->ev = keep.ev;
    out->p = keep.p; out->color = keep.color; out->bonds = keep.bonds;
    out->bonds_delta_n = keep.bonds_delta_n; out->bonds_delta_t = keep.bonds_delta_t;
    out->bonds_omega_n = keep.bonds_omega_n; out->bonds_omega_t = keep.bonds_omega_t;
    out->v_f = keep.v_f; out->p_f = keep.p_f; out->phi = keep.phi; out->dphi = keep.dphi;
    out->p_constant = keep.p_constant; out->f_p = keep.f_p;
    return SPH_OK;
}

int sph_write_file(sph_ctx *ctx, const char *path)
{
    if (!ctx || !path) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    (void)hipSetDevice(ctx->device);
    int rc = download(ctx, true, true);
    if (rc) return rc;
    std::string err;
    rc = write_state_file(path, ctx->hs, err);
    return rc ? fail(ctx, rc, err) : SPH_OK;
}

int sph_step(sph_ctx *ctx, int nsteps)
{
    if (!ctx) return fail(ctx, SPH_ERR_ARG, "null context");
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i < nsteps; ++i) {
        int rc = step_once(ctx, nullptr, nullptr, std::string());
        if (rc) return rc;
    }
    int rc = check_flags(ctx);
    if (rc) return rc;
    return sync(ctx);
}

int sph_run(const sph_run_opts *opts)
{
    if (!opts || !opts->input_path) return fail(nullptr, SPH_ERR_ARG, "null options");
    std::string in(opts->input_path);
    size_t dotpos = in.rfind('.');
    size_t slashpos = in.rfind('/');
    size_t from = slashpos == std::string::npos ? 0 : slashpos + 1;
    std::string sid = in.substr(from, (dotpos == std::string::npos || dotpos < from) ? std::string::npos
                                                                                     : dotpos - from);
    std::string dir = opts->output_dir ? opts->output_dir : "output";
    sph_ctx *ctx = sph_create(opts->device);
    if (!ctx) return SPH_ERR_HIP;
    // a restart side car beside the input (<input>.bin -> <input>.contacts.bin)
    std::string side;
    if (in.size() > 4 && in.compare(in.size() - 4, 4, ".bin") == 0)
        side = in.substr(0, in.size() - 4) + ".contacts.bin";
    RestartHeader rh{};
    const bool have_side = !side.empty() && peek_restart(side, rh) == SPH_OK;
    if (have_side && (rh.nc == 32 || rh.nc == 64)) ctx->nc = rh.nc;
    int rc = sph_load_file(ctx, opts->input_path, opts->fluid);
    if (rc) {
        g_last_error = ctx->err;
        sph_destroy(ctx);
        return rc;
    }
    sph_state &h = ctx->hs.h;
    if (have_side) {
        rc = read_restart(ctx, side);
        if (rc == SPH_ERR_VALUE) {
            fprintf(stderr, "Warning: ignoring %s: %s\n", side.c_str(), ctx->err.c_str());
        } else if (rc) {
            g_last_error = ctx->err;
            sph_destroy(ctx);
            return rc;
        }
    }
    if (opts->dry) {  // reportValues, src/sphere.cpp:349-465
        printf("  - Number of particles: %u\n  - Time step length: %g s\n  - Total time: %g s\n"
               "  - File output interval: %g s\n  - Grid: %u*%u*%u, periodic %d\n  - Contact model: %u\n",
               h.np, h.time_dt, h.time_total, h.time_file_dt, h.num[0], h.num[1], h.num[2],
               h.periodic, h.contactmodel);
        sph_destroy(ctx);
        return SPH_OK;
    }
    std::string err;
    if (h.time_step_count == 0) {
        rc = write_state_file((dir + "/" + sid + ".output00000.bin").c_str(), ctx->hs, err);
        if (rc) {
            g_last_error = err;
            sph_destroy(ctx);
            return rc;
        }
    }
    {
        FILE *fp = fopen((dir + "/" + sid + ".status.dat").c_str(), "w");
        if (fp) {
            fprintf(fp, "%2.4e %2.4e %d\n", h.time_current, 100.0 * h.time_current / h.time_total,
                    (int)h.time_step_count);
            fclose(fp);
        }
    }
    std::ofstream convlog((dir + "/" + sid + "-conv.log").c_str());
    long nstep = 0;
    while (h.time_current <= h.time_total) {
        rc = step_once(ctx, &convlog, &dir, sid);
        if (!rc && (++nstep % 64 == 0)) rc = check_flags(ctx);
        if (rc) break;
        if (opts->verbose && ctx->iter % 10000 == 0) {
            printf("\r  Current time: %g/%g s.", h.time_current, h.time_total);
            fflush(stdout);
        }
    }
    if (!rc) rc = check_flags(ctx);
    if (!rc) rc = sync(ctx);
    if (opts->verbose) printf("\n");
    if (rc) g_last_error = ctx->err;
    sph_destroy(ctx);
    return rc;
}

int sph_write_restart(sph_ctx *ctx, const char *path)
{
    if (!ctx || !path) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    (void)hipSetDevice(ctx->device);
    int rc = download(ctx, false, false);  // the header's time and output number
    if (rc) return rc;
    return write_restart(ctx, path);
}

int sph_read_restart(sph_ctx *ctx, const char *path)
{
    if (!ctx || !path) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    (void)hipSetDevice(ctx->device);
    return read_restart(ctx, path);
}

int sph_read_sort(sph_ctx *ctx, uint32_t *cellid, uint32_t *index, uint32_t *cell_start,
                  uint32_t *cell_end)
{
    if (!ctx) return fail(ctx, SPH_ERR_ARG, "null context");
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    (void)hipSetDevice(ctx->device);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const size_t np = ctx->np;
    if (cellid && np) HIPCHK(hipMemcpy(cellid, ctx->db.keys, np * 4, hipMemcpyDeviceToHost));
    if (index && np) HIPCHK(hipMemcpy(index, ctx->db.vals, np * 4, hipMemcpyDeviceToHost));
    if (cell_start) HIPCHK(hipMemcpy(cell_start, ctx->db.cell_start, ctx->ncells * 4, hipMemcpyDeviceToHost));
    if (cell_end) HIPCHK(hipMemcpy(cell_end, ctx->db.cell_end, ctx->ncells * 4, hipMemcpyDeviceToHost));
    return SPH_OK;
}

int sph_read_contacts(sph_ctx *ctx, uint32_t *contacts, double *distmod, double *delta_t)
{
    if (!ctx) return fail(ctx, SPH_ERR_ARG, "null context");
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    (void)hipSetDevice(ctx->device);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const size_t n = (size_t)ctx->np * ctx->nc;
    if (!n) return SPH_OK;
    std::vector<uint32_t> c(n), code(n);
    int rc = rows_to_host(ctx, c.data(), delta_t);
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
        code[i] = c[i] >> CT_CODE_SHIFT;
        c[i] &= CT_IDX_MASK;
    }
    if (contacts) std::memcpy(contacts, c.data(), n * 4);
    if (distmod) {
        const sph_state &h = ctx->hs.h;
        for (size_t i = 0; i < n; ++i) {
            double dx = 0.0, dy = 0.0;
            if (c[i] != ctx->np) {
                if (code[i] & 1u) dx = dx + h.L[0];
                if (code[i] & 2u) dx = dx - h.L[0];
                if (code[i] & 4u) dy = dy + h.L[1];
                if (code[i] & 8u) dy = dy - h.L[1];
            }
            distmod[3 * i] = dx;
            distmod[3 * i + 1] = dy;
            distmod[3 * i + 2] = 0.0;
        }
    }
    return SPH_OK;
}

int sph_write_contacts(sph_ctx *ctx, const uint32_t *contacts, const double *distmod,
                       const double *delta_t)
{
    if (!ctx || !contacts) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    (void)hipSetDevice(ctx->device);
    const size_t n = (size_t)ctx->np * ctx->nc;
    if (!n) return SPH_OK;
    const sph_state &h = ctx->hs.h;
    std::vector<uint32_t> words(contacts, contacts + n);
    for (size_t i = 0; i < n; ++i) {
        if (contacts[i] > ctx->np) return fail(ctx, SPH_ERR_ARG, "contact index out of range");
        if (distmod && contacts[i] != ctx->np) {
            double dx = distmod[3 * i], dy = distmod[3 * i + 1];
            uint8_t cd = 0;
            if (dx == h.L[0]) cd |= 1; else if (dx == -h.L[0]) cd |= 2; else if (dx != 0.0)
                return fail(ctx, SPH_ERR_ARG, "distmod x is not 0 or +-L[0]");
            if (dy == h.L[1]) cd |= 4; else if (dy == -h.L[1]) cd |= 8; else if (dy != 0.0)
                return fail(ctx, SPH_ERR_ARG, "distmod y is not 0 or +-L[1]");
            words[i] = ct_word(contacts[i], cd);
        }
    }
    hipStream_t s = ctx->stream;
    HIPCHK(hipStreamSynchronize(s));
    int rc = rows_to_device(ctx, words.data(), delta_t);
    if (rc) return rc;
    launch_rebuild_mask(s, ctx->db, ctx->np, ctx->nc);
    ctx->nl_force = 1;  // the list/slot maps are rebuilt from the new rows
    return sync(ctx);
}

int sph_count_contacts(sph_ctx *ctx, uint64_t *ncontacts)
{
    if (!ctx || !ncontacts) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!ctx->loaded) return fail(ctx, SPH_ERR_STATE, "no state loaded");
    (void)hipSetDevice(ctx->device);
    unsigned long long *d = (unsigned long long *)ctx->dd.maxres;
    unsigned long long *tmp = nullptr;
    HIPCHK(hipMalloc((void **)&tmp, 8));
    (void)d;
    launch_fill_u32(ctx->stream, (uint32_t *)tmp, 2, 0u);
    launch_count_contacts(ctx->stream, ctx->db.cmask, ctx->np, tmp);
    HIPCHK(hipMemcpyAsync(ctx->h_u64, tmp, 8, hipMemcpyDeviceToHost, ctx->stream));
    int rc = sync(ctx);
    (void)hipFree(tmp);
    if (rc) return rc;
    *ncontacts = *ctx->h_u64;
    return SPH_OK;
}

int sph_set_darcy_net(sph_ctx *ctx, const float *params, uint64_t count, int max_corrections)
{
    if (!ctx) return fail(nullptr, SPH_ERR_ARG, "null context");
    (void)hipSetDevice(ctx->device);
    net_release(ctx);
    if (!params || max_corrections <= 0) {
        ctx->net_host.clear();
        ctx->net_max_corr = 0;
        return SPH_OK;
    }
    if (count != sph_net_param_count())
        return fail(ctx, SPH_ERR_ARG, "Darcy network: expected " + std::to_string(sph_net_param_count()) +
                                          " parameters, got " + std::to_string(count));
    ctx->net_host.assign(params, params + count);
    ctx->net_max_corr = max_corrections;
    return SPH_OK;
}

static int darcy_system_ready(sph_ctx *ctx)
{
    if (!ctx) return fail(nullptr, SPH_ERR_ARG, "null context");
    if (!ctx->loaded || !ctx->hs.h.fluid) return fail(ctx, SPH_ERR_STATE, "no fluid state loaded");
    if (ctx->stats.fluid_steps == 0) return fail(ctx, SPH_ERR_STATE, "no Darcy solve has run yet");
    (void)hipSetDevice(ctx->device);
    return SPH_OK;
}

int sph_darcy_export(sph_ctx *ctx, int which, double *dst)
{
    int rc = darcy_system_ready(ctx);
    if (rc) return rc;
    if (!dst || (which != SPH_DARCY_P_START && which != SPH_DARCY_P_FINAL))
        return fail(ctx, SPH_ERR_ARG, "bad field or destination");
    const double *src = which == SPH_DARCY_P_START ? ctx->dd.p_old : ctx->dd.p;
    HIPCHK(hipMemcpyAsync(dst, src, (size_t)dcells(ctx->hs.h) * 8, hipMemcpyDeviceToDevice, ctx->stream));
    return sync(ctx);
}

int sph_darcy_sweeps(sph_ctx *ctx, double *p, int nsweeps)
{
    int rc = darcy_system_ready(ctx);
    if (rc) return rc;
    if (!p || nsweeps < 0) return fail(ctx, SPH_ERR_ARG, "bad iterate or sweep count");
    const DarcyArgs da = darcy_args(ctx);
    DarcyBuffers &d = ctx->dd;
    hipStream_t s = ctx->stream;
    launch_darcy_ghosts(s, da, p, 1);
    double *buf[2] = {p, d.p_new};
    for (int k = 0; k < nsweeps; ++k)
        launch_darcy_sweep(s, da, d, buf[k & 1], buf[(k & 1) ^ 1], nullptr, nullptr, nullptr, 0.0, 0u, true);
    if (nsweeps & 1)
        HIPCHK(hipMemcpyAsync(p, d.p_new, (size_t)dcells(ctx->hs.h) * 8, hipMemcpyDeviceToDevice, s));
    return sync(ctx);
}

int sph_darcy_net_input(sph_ctx *ctx, const double *p, float *input, double *scale)
{
    int rc = darcy_system_ready(ctx);
    if (rc) return rc;
    if (!p || !input || !scale) return fail(ctx, SPH_ERR_ARG, "null argument");
    const DarcyArgs da = darcy_args(ctx);
    DarcyBuffers &d = ctx->dd;
    hipStream_t s = ctx->stream;
    launch_fill_u32(s, (uint32_t *)(d.maxres + DARCY_SCALE), 2 * RED_WORDS, 0u);
    launch_darcy_residual_max(s, da, d, p, true);
    launch_darcy_net_input(s, da, d, p, input, true);
    rc = sync(ctx);
    if (rc) return rc;
    std::vector<unsigned long long> w(RED_WORDS);
    HIPCHK(hipMemcpy(w.data(), d.maxres + DARCY_SCALE, 8 * RED_WORDS, hipMemcpyDeviceToHost));
    unsigned long long m = 0;
    for (int j = 0; j < RED_LANES; ++j) m = std::max(m, w[(size_t)j * RED_STRIDE]);
    std::memcpy(scale, &m, 8);
    return SPH_OK;
}

int sph_get_darcy_stats(sph_ctx *ctx, sph_darcy_stats *out)
{
    if (!ctx || !out) return fail(ctx, SPH_ERR_ARG, "null argument");
    *out = ctx->stats;
    return SPH_OK;
}

int sph_set_profiling(sph_ctx *ctx, int on)
{
    if (!ctx) return fail(ctx, SPH_ERR_ARG, "null context");
    int rc = sync(ctx);
    if (rc) return rc;
    ctx->prof = on != 0;
    for (int k = 0; k < SPH_K_COUNT; ++k) {
        ctx->ktime[k] = 0.0;
        ctx->klaunch[k] = 0;
    }
    return SPH_OK;
}

int sph_kernel_time(sph_ctx *ctx, int kernel_id, double *total_ms, uint64_t *launches)
{
    if (!ctx || kernel_id < 0 || kernel_id >= SPH_K_COUNT) return fail(ctx, SPH_ERR_ARG, "bad argument");
    int rc = sync(ctx);
    if (rc) return rc;
    if (total_ms) *total_ms = ctx->ktime[kernel_id];
    if (launches) *launches = ctx->klaunch[kernel_id];
    return SPH_OK;
}

// ------------------------------------------------------------------
// Ensembles of independent boxes (no reference counterpart)
//
// Boxes on the list path without fluid (the shear and consolidation boxes
// of a DEM ensemble) are stepped together: every stage of a step is one
// launch for all of them (dem_kernels.hip, "multi-box launches"), whose
// blocks are divided among the boxes and whose per-box arguments are task
// records in device memory.  Each box runs the body of its single-box
// kernel, so its results equal those of sph_step on that box alone.  The
// task records depend on two host values that alternate from step to step
// (the rebuild-flag word and, for an odd number of radix passes, which key
// buffer is current), so a plan holds the records of two consecutive
// steps, and two steps are captured in one HIP graph that is replayed.
// Other boxes (fluid, cell search, shear-stress walls, ...) are stepped
// box by box on their own streams beside the multi-box launches.
// ------------------------------------------------------------------

}  // extern "C"

struct EnsStage {
    void *tasks = nullptr;      // device task records
    uint32_t *map = nullptr;    // device: block starts, then block counts
    uint32_t nbox = 0, blocks = 0;
};

struct EnsPhase {  // the launches of one step
    EnsStage rebuild, nlbuild, topo[2], forces[2], bonds, integ;
    void *walls = nullptr;
    uint32_t nwalls = 0;
    void *clocks = nullptr;
    uint32_t nclocks = 0;
};

struct EnsFluidPhase {  // the particle-side launches of one step of the fluid boxes
    EnsStage hash, rebuild, nlbuild, topo[2], forces[2], bonds, gather, integ;
    void *walls = nullptr;
    uint32_t nwalls = 0;
    void *clocks = nullptr;
    uint32_t nclocks = 0;
};

struct sph_ensemble {
    std::vector<sph_ctx *> boxes;
    int device = 0;
    std::string err;
    uint64_t steps = 0;
    hipStream_t stream = nullptr;
    // plan of the multi-box boxes
    std::vector<int> multi;             // indices of the boxes stepped by multi-box launches
    std::vector<uint64_t> plan_gen;     // their load generation at planning
    std::vector<uint32_t> plan_parity;  // their rebuild-flag word at planning
    std::vector<uint32_t *> plan_keys;  // their current key buffer at planning
    bool planned = false;
    uint32_t phase = 0;                 // steps since planning, mod 2
    EnsPhase ph[2];
    std::vector<void *> dmem;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = true;
    uint64_t graph_replays = 0, multi_steps = 0;
    uint32_t launches_per_step = 0;
    // fluid boxes stepped by multi-box launches (section of their own: the
    // Jacobi loop's batches end in a host read of every box's solve state)
    std::vector<int> fluid;
    std::vector<char> fl_elig;  // which boxes were eligible at planning
    std::vector<uint64_t> fl_gen;
    std::vector<uint32_t> fl_parity;
    std::vector<uint32_t *> fl_keys;
    bool fl_planned = false;
    uint32_t fl_phase = 0;
    EnsFluidPhase fph[2];
    // Darcy stages: every box (cells, particles, sweep blocks, cells with
    // ghosts), and the boxes with a flux face or periodic = 2
    EnsStage dm_cells, dm_parts, dm_sweep, dm_all, dm_flux, dm_per2, dm_net;
    void *dm_tasks = nullptr;  // every box's record, one thread or block per box
    NetMulti *netm = nullptr;  // the forward passes of the boxes with a network
    int net_max_corr = 0;      // their corrections per solve at most
    uint32_t fl_maxiter = 0;
    unsigned long long *dm_states = nullptr;    // device: 4 state words per box
    unsigned long long *dm_states_h = nullptr;  // pinned host copy
    // the sweep launches cover only the boxes whose solve has not ended by
    // the last host read: their block starts, counts and box indices
    std::vector<uint32_t> fl_sweep_blocks;   // sweep blocks of each box
    uint32_t *fl_sweep_map = nullptr;        // device, 3 words per box
    uint32_t *fl_sweep_map_h = nullptr;      // pinned host staging
    std::vector<void *> fl_dmem;
    uint64_t fluid_steps = 0;
};

static int ens_fail(sph_ensemble *e, int code, const std::string &msg)
{
    e->err = msg;
    g_last_error = msg;
    return code;
}

static void ens_drop_plan(sph_ensemble *e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    if (!e->dmem.empty() && e->stream) (void)hipStreamSynchronize(e->stream);
    for (void *p : e->dmem) (void)hipFree(p);
    e->dmem.clear();
    for (EnsPhase &p : e->ph) p = EnsPhase{};
    e->planned = false;
    e->multi.clear();
}

// A box the multi-box launches can step: the list path with the fused
// rebuild, contact model 2 or 3, no fluid, no adaptive grid, no wmode-3
// wall, no profiling.
static bool ens_multi_eligible(const sph_ctx *c)
{
    const sph_state &h = c->hs.h;
    return c->loaded && h.np > 0 && !h.fluid && c->nl_on && c->rb_cap > 0 &&
           (h.contactmodel == 2 || h.contactmodel == 3) && !(h.adaptive == 1 && h.nw > 0) &&
           !(h.nw > 0 && h.wmode[0] == 3) && !c->prof && c->db.rb_bar && c->db.rb_part;
}

static int radix_passes(const sph_ctx *c)
{
    const int p = (c->key_bits + 7) / 8;
    return p < 1 ? 1 : p;
}

// The device copy of one stage: task records in box order and the map of
// block starts (multiples of 8) and counts.
template <class T>
static int ens_stage(sph_ensemble *e, EnsStage &st, const std::vector<T> &tasks,
                     const std::vector<uint32_t> &counts, std::vector<void *> *mem = nullptr)
{
    std::vector<void *> &keep = mem ? *mem : e->dmem;
    st.nbox = (uint32_t)tasks.size();
    st.blocks = 0;
    if (tasks.empty()) return SPH_OK;
    std::vector<uint32_t> map(2 * tasks.size());
    uint32_t at = 0;
    for (size_t k = 0; k < tasks.size(); ++k) {
        map[k] = at;
        map[tasks.size() + k] = counts[k];
        at += (counts[k] + 7u) & ~7u;
    }
    st.blocks = at;
    void *dt = nullptr, *dm = nullptr;
    if (hipMalloc(&dt, sizeof(T) * tasks.size()) != hipSuccess)
        return ens_fail(e, SPH_ERR_HIP, "ensemble: device allocation failed");
    keep.push_back(dt);
    if (hipMalloc(&dm, 4 * map.size()) != hipSuccess)
        return ens_fail(e, SPH_ERR_HIP, "ensemble: device allocation failed");
    keep.push_back(dm);
    if (hipMemcpy(dt, tasks.data(), sizeof(T) * tasks.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dm, map.data(), 4 * map.size(), hipMemcpyHostToDevice) != hipSuccess)
        return ens_fail(e, SPH_ERR_HIP, "ensemble: copy of the task records failed");
    st.tasks = dt;
    st.map = (uint32_t *)dm;
    return SPH_OK;
}

template <class T>
static int ens_array(sph_ensemble *e, void *&dst, const std::vector<T> &v,
                     std::vector<void *> *mem = nullptr)
{
    dst = nullptr;
    if (v.empty()) return SPH_OK;
    if (hipMalloc(&dst, sizeof(T) * v.size()) != hipSuccess)
        return ens_fail(e, SPH_ERR_HIP, "ensemble: device allocation failed");
    (mem ? *mem : e->dmem).push_back(dst);
    if (hipMemcpy(dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
        return ens_fail(e, SPH_ERR_HIP, "ensemble: copy of the task records failed");
    return SPH_OK;
}

// Rebuild blocks of each multi-box box: its single-box grid, scaled down so
// that the whole launch stays within a quarter of the device's resident
// capacity for the kernel (every box's grid barrier needs its blocks
// running together).
static std::vector<uint32_t> ens_rebuild_blocks(const sph_ensemble *e, const std::vector<int> &ids)
{
    std::vector<uint32_t> nb(ids.size(), 1u);
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device);
    const uint64_t budget = std::max<uint64_t>(ids.size(),
                                               (uint64_t)rebuild_blocks_per_cu() * std::max(cus, 1) / 4);
    uint64_t total_np = 0, want = 0;
    for (int k : ids) {
        total_np += e->boxes[k]->hs.h.np;
        want += e->boxes[k]->rb_cap;
    }
    for (size_t i = 0; i < ids.size(); ++i) {
        const sph_ctx *c = e->boxes[ids[i]];
        uint64_t b = c->rb_cap;
        if (want > budget) b = std::max<uint64_t>(1, (budget - ids.size()) * c->hs.h.np / std::max<uint64_t>(total_np, 1));
        nb[i] = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, c->rb_cap));
    }
    return nb;
}

static int ens_plan(sph_ensemble *e)
{
    ens_drop_plan(e);
    for (size_t k = 0; k < e->boxes.size(); ++k)
        if (ens_multi_eligible(e->boxes[k])) e->multi.push_back((int)k);
    e->plan_gen.clear();
    e->plan_parity.clear();
    e->plan_keys.clear();
    for (int k : e->multi) {
        const sph_ctx *c = e->boxes[k];
        e->plan_gen.push_back(c->gen);
        e->plan_parity.push_back(c->nl_parity);
        e->plan_keys.push_back(c->db.keys);
    }
    e->phase = 0;
    const std::vector<uint32_t> rbn = ens_rebuild_blocks(e, e->multi);
    e->launches_per_step = 0;
    for (int j = 0; j < 2; ++j) {
        std::vector<RebuildTask> rb;
        std::vector<NlBuildTask> nlb;
        std::vector<TopoTask> topo[2];
        std::vector<ForcesTask> frc[2];
        std::vector<BondTask> bnd;
        std::vector<IntegrateTask> itg;
        std::vector<WallsTask> wls;
        std::vector<ClockTask> clk;
        std::vector<uint32_t> n_rb, n_nlb, n_topo[2], n_frc[2], n_bnd, n_itg;
        for (size_t i = 0; i < e->multi.size(); ++i) {
            const sph_ctx *c = e->boxes[e->multi[i]];
            const sph_state &h = c->hs.h;
            const StepArgs a = step_args(c);
            const uint32_t cur = c->nl_parity ^ (uint32_t)j;
            DemBuffers b = c->db;
            const bool odd = (radix_passes(c) & 1) != 0;
            if (odd && j == 1) {
                std::swap(b.keys, b.keys_tmp);
                std::swap(b.vals, b.vals_tmp);
            }
            const uint32_t *cond = c->db.nl_flags + cur;
            rb.push_back(rebuild_task(a, b, c->ncells, c->key_bits, cond, c->rb_spin));
            n_rb.push_back(rbn[i]);
            if (odd) {  // the sorted keys end in the other buffer (launch_rebuild)
                std::swap(b.keys, b.keys_tmp);
                std::swap(b.vals, b.vals_tmp);
            }
            nlb.push_back(nl_build_task(a, c->nla, b, cond));
            n_nlb.push_back((h.np + 255) / 256);
            const int near = c->nla.range == 1 ? 1 : 0;
            topo[near].push_back(topology_task(a, b, cur));
            n_topo[near].push_back(topology_blocks(h.np));
            const int hz = h.contactmodel == 3 ? 1 : 0;
            frc[hz].push_back(forces_task(a, b, true));
            n_frc[hz].push_back(forces_blocks(h.np));
            if (h.nb0 > 0) {
                bnd.push_back(bond_task(a, b));
                n_bnd.push_back((h.nb0 + 255) / 256);
            }
            NextHash nh{};
            nh.keys = b.keys;
            nh.vals = b.vals;
            nh.cellc = b.cellc;
            nh.x_ref = b.x_ref;
            nh.disp = b.nl_disp;
            nh.flags = b.nl_flags;
            nh.cur = cur ^ 1u;  // the next step's word
            nh.disp2 = c->nla.disp2;
            itg.push_back(integrate_task(a, b, 0, 10.0, h.tau_x, true, nh));
            n_itg.push_back((h.np + 255) / 256);
            if (h.nw > 0) wls.push_back(walls_task(a, b));
            clk.push_back(ClockTask{b.clock, h.time_dt});
        }
        EnsPhase &P = e->ph[j];
        int rc;
        if ((rc = ens_stage(e, P.rebuild, rb, n_rb)) || (rc = ens_stage(e, P.nlbuild, nlb, n_nlb)) ||
            (rc = ens_stage(e, P.topo[0], topo[0], n_topo[0])) ||
            (rc = ens_stage(e, P.topo[1], topo[1], n_topo[1])) ||
            (rc = ens_stage(e, P.forces[0], frc[0], n_frc[0])) ||
            (rc = ens_stage(e, P.forces[1], frc[1], n_frc[1])) ||
            (rc = ens_stage(e, P.bonds, bnd, n_bnd)) || (rc = ens_stage(e, P.integ, itg, n_itg)) ||
            (rc = ens_array(e, P.walls, wls)) || (rc = ens_array(e, P.clocks, clk)))
            return rc;
        P.nwalls = (uint32_t)wls.size();
        P.nclocks = (uint32_t)clk.size();
        if (j == 0) {
            const EnsStage *all[] = {&P.rebuild, &P.nlbuild, &P.topo[0], &P.topo[1], &P.forces[0],
                                     &P.forces[1], &P.bonds, &P.integ};
            for (const EnsStage *st : all) e->launches_per_step += st->blocks ? 1u : 0u;
            e->launches_per_step += (P.nwalls ? 1u : 0u) + (P.nclocks ? 1u : 0u);
        }
    }
    e->planned = true;
    return SPH_OK;
}

// The plan still describes the boxes: same boxes eligible, same loads, and
// the host values it assumed for the current phase.
static bool ens_plan_current(const sph_ensemble *e)
{
    if (!e->planned) return false;
    size_t i = 0;
    for (size_t k = 0; k < e->boxes.size(); ++k) {
        const sph_ctx *c = e->boxes[k];
        const bool in = i < e->multi.size() && e->multi[i] == (int)k;
        if (ens_multi_eligible(c) != in) return false;
        if (!in) continue;
        if (c->gen != e->plan_gen[i]) return false;
        if (c->nl_parity != (e->plan_parity[i] ^ e->phase)) return false;
        const bool odd = (radix_passes(c) & 1) != 0;
        const bool swapped = odd && e->phase == 1;
        if ((c->db.keys == e->plan_keys[i]) == swapped) return false;
        ++i;
    }
    return i == e->multi.size();
}

static void ens_enqueue_phase(sph_ensemble *e, const EnsPhase &P)
{
    hipStream_t s = e->stream;
    auto mm = [](const EnsStage &st) { return MultiMap{st.map, st.nbox}; };
    launch_rebuild_multi(s, (const RebuildTask *)P.rebuild.tasks, mm(P.rebuild), P.rebuild.blocks);
    launch_nl_build_multi(s, (const NlBuildTask *)P.nlbuild.tasks, mm(P.nlbuild), P.nlbuild.blocks);
    launch_topology_multi(s, (const TopoTask *)P.topo[0].tasks, mm(P.topo[0]), P.topo[0].blocks, false);
    launch_topology_multi(s, (const TopoTask *)P.topo[1].tasks, mm(P.topo[1]), P.topo[1].blocks, true);
    launch_forces_multi(s, (const ForcesTask *)P.forces[0].tasks, mm(P.forces[0]), P.forces[0].blocks, false);
    launch_forces_multi(s, (const ForcesTask *)P.forces[1].tasks, mm(P.forces[1]), P.forces[1].blocks, true);
    launch_bonds_multi(s, (const BondTask *)P.bonds.tasks, mm(P.bonds), P.bonds.blocks);
    launch_integrate_multi(s, (const IntegrateTask *)P.integ.tasks, mm(P.integ), P.integ.blocks);
    launch_walls_multi(s, (const WallsTask *)P.walls, P.nwalls);
    launch_clock_multi(s, (const ClockTask *)P.clocks, P.nclocks);
}

// The host side of one step of a multi-box box (what step_once changes).
static void ens_host_advance(sph_ctx *c)
{
    sph_state &h = c->hs.h;
    c->nl_parity ^= 1u;
    if (radix_passes(c) & 1) {
        std::swap(c->db.keys, c->db.keys_tmp);
        std::swap(c->db.vals, c->db.vals_tmp);
    }
    c->nl_force = 0;
    c->hash_ready = true;
    h.time_current += h.time_dt;
    c->filetimeclock += h.time_dt;
    c->iter++;
}

static void ens_advance_all(sph_ensemble *e)
{
    for (int k : e->multi) ens_host_advance(e->boxes[k]);
    e->phase ^= 1u;
    e->multi_steps++;
}

// Steps whose launches need no host decision between them: no output
// event and no change of the velocity-step state (src/device.cu:2461-2469)
// inside the steps, and no adaptive grid.
static int ens_check(sph_ensemble *e, int nsteps)
{
    for (size_t i = 0; i < e->boxes.size(); ++i) {
        sph_ctx *c = e->boxes[i];
        const sph_state &h = c->hs.h;
        const std::string b = "box " + std::to_string(i) + ": ";
        if (!c->loaded) return ens_fail(e, SPH_ERR_STATE, b + "no state loaded");
        if (h.adaptive == 1 && h.nw > 0) return ens_fail(e, SPH_ERR_VALUE, b + "adaptive grid");
        double clk = c->filetimeclock, t = h.time_current;
        int vs = c->velocity_state;
        for (int k = 0; k < nsteps; ++k) {
            if ((t >= 10.0 && t < 15.0 && vs == 1) || (t >= 15.0 && vs == 2))
                return ens_fail(e, SPH_ERR_VALUE, b + "the velocity step at t = 10 s or 15 s falls in the steps");
            clk += h.time_dt;
            t += h.time_dt;
        }
        if (clk >= h.time_file_dt) return ens_fail(e, SPH_ERR_VALUE, b + "an output event falls in the steps");
    }
    return SPH_OK;
}

// Two steps of every multi-box box as one graph.  Nothing is pending on
// the stream when capture begins, and the captured work is kernel launches
// on this one stream only (no events, memsets or host calls).
static int ens_capture(sph_ensemble *e)
{
    HIPCHK_E(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    ens_enqueue_phase(e, e->ph[0]);
    ens_enqueue_phase(e, e->ph[1]);
    hipGraph_t g = nullptr;
    HIPCHK_E(hipStreamEndCapture(e->stream, &g));
    e->graph = g;
    HIPCHK_E(hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0));
    return SPH_OK;
}

extern "C" {

static void ens_fluid_drop(sph_ensemble *e);

// A new list skin for a loaded context on the list path: the next step
// rebuilds the lists (exact for any skin, so the results do not change).
static int nl_set_skin(sph_ctx *c, double frac)
{
    const sph_state &h = c->hs.h;
    const double skin = frac * c->nl_rmax;
    const uint32_t range = frac > 0.0 ? nl_range(h, c->nl_rmax, skin) : 0u;
    if (range == 0) return SPH_ERR_VALUE;
    c->nl_frac = frac;
    c->nla.skin = skin;
    c->nla.disp2 = (0.45 * skin) * (0.45 * skin);
    c->nla.range = range;
    c->nl_force = 1;
    return SPH_OK;
}

int sph_ensemble_autotune(sph_ensemble *e, int trial_steps, const double *fractions, int nfrac,
                          double *chosen, double *ms_per_step)
{
    if (!e || trial_steps <= 0 || !fractions || nfrac <= 0) return SPH_ERR_ARG;
    (void)hipSetDevice(e->device);
    for (int f = 0; f < nfrac; ++f)
        if (!(fractions[f] > 0.0 && fractions[f] <= 1.0))
            return ens_fail(e, SPH_ERR_ARG, "autotune: skin fractions must lie in (0, 1]");
    int rc = ens_check(e, trial_steps * nfrac);
    if (rc) return rc;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIPCHK_E(hipEventCreate(&t0));
    HIPCHK_E(hipEventCreate(&t1));
    for (size_t k = 0; k < e->boxes.size(); ++k) {
        sph_ctx *c = e->boxes[k];
        double best = 1e300, best_f = c->nl_frac;
        const bool tunable = c->loaded && c->nl_on;
        for (int f = 0; f < nfrac; ++f) {
            double ms = -1.0;
            if (tunable && nl_set_skin(c, fractions[f]) == SPH_OK) {
                // the first step rebuilds for the new skin; it is timed with
                // the rest, as a run with this skin would pay it once
                if ((rc = sync(c))) break;
                if (hipEventRecord(t0, c->stream) != hipSuccess) { rc = SPH_ERR_HIP; break; }
                for (int s = 0; s < trial_steps && !rc; ++s)
                    rc = step_once(c, nullptr, nullptr, std::string());
                if (rc) break;
                if ((rc = wait_walls(c))) break;
                float el = 0.0f;
                if (hipEventRecord(t1, c->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                    hipEventElapsedTime(&el, t0, t1) != hipSuccess) { rc = SPH_ERR_HIP; break; }
                ms = el / trial_steps;
                if (ms < best) {
                    best = ms;
                    best_f = fractions[f];
                }
            } else {
                // not tunable: the steps still run, so that every box keeps
                // the same step count
                for (int s = 0; s < trial_steps && !rc; ++s)
                    rc = step_once(c, nullptr, nullptr, std::string());
                if (rc) break;
            }
            if (ms_per_step) ms_per_step[k * nfrac + f] = ms;
        }
        if (rc) {
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
            return ens_fail(e, rc, c->err.empty() ? "autotune: step failed" : c->err);
        }
        if (tunable && best_f != c->nl_frac) (void)nl_set_skin(c, best_f);
        if (chosen) chosen[k] = tunable ? best_f : 0.0;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    for (sph_ctx *c : e->boxes) {
        if ((rc = check_flags(c))) return ens_fail(e, rc, c->err);
        if ((rc = sync(c))) return ens_fail(e, rc, c->err);
    }
    e->steps += (uint64_t)trial_steps * (uint64_t)nfrac;
    ens_drop_plan(e);  // the task records hold the old list parameters
    ens_fluid_drop(e);
    return SPH_OK;
}

// ------------------------------------------------------------------
// Fluid boxes in the multi-box launches: the list path with a Darcy grid,
// one DEM step per fluid step (ndem = 1), no modulated top pressure.  Each
// stage of darcy_step is one launch for all of them; the Jacobi sweeps run
// in batches like darcy_step's (each box's device-side test ends its own
// sweeps), and after each batch the host reads the solve state of every box
// at once.  The boxes with a learned correction (the same number of
// corrections per solve) get it after the same deciding sweeps as in
// darcy_step, their networks' forward passes as multi-box launches.  After a solve the converged
// iterate is copied into the first iterate buffer where a single-box solve
// would swap the two buffers, so the task records stay the same from step to
// step (the other buffer's content is never read before it is written).
// ------------------------------------------------------------------

static bool ens_fluid_eligible(const sph_ctx *c)
{
    const sph_state &h = c->hs.h;
    return c->loaded && h.np > 0 && h.fluid && c->nl_on && c->rb_cap > 0 &&
           (h.contactmodel == 2 || h.contactmodel == 3) && !(h.adaptive == 1 && h.nw > 0) &&
           !(h.nw > 0 && h.wmode[0] == 3) && h.ndem == 1 &&
           !((h.p_mod_A > 1.0e-5 || h.p_mod_A < -1.0e-5) && h.p_mod_f > 1.0e-7) && !c->prof &&
           c->db.rb_bar && c->db.rb_part &&
           (c->net_host.empty() || c->net_max_corr <= 0 ||
            (h.num[0] % 8 == 0 && h.num[1] % 8 == 0 && h.num[2] % 8 == 0));
}

static bool ens_learned(const sph_ctx *c) { return !c->net_host.empty() && c->net_max_corr > 0; }

static void ens_fluid_drop(sph_ensemble *e)
{
    if (!e->fl_dmem.empty() && e->stream) (void)hipStreamSynchronize(e->stream);
    for (void *p : e->fl_dmem) (void)hipFree(p);
    e->fl_dmem.clear();
    if (e->dm_states_h) (void)hipHostFree(e->dm_states_h);
    e->dm_states_h = nullptr;
    if (e->fl_sweep_map_h) (void)hipHostFree(e->fl_sweep_map_h);
    e->fl_sweep_map_h = nullptr;
    e->fl_sweep_map = nullptr;
    e->fl_sweep_blocks.clear();
    net_multi_destroy(e->netm);
    e->netm = nullptr;
    e->net_max_corr = 0;
    e->dm_states = nullptr;
    e->dm_tasks = nullptr;
    for (EnsFluidPhase &p : e->fph) p = EnsFluidPhase{};
    e->dm_cells = e->dm_parts = e->dm_sweep = e->dm_all = e->dm_flux = e->dm_per2 = e->dm_net = EnsStage{};
    e->fl_planned = false;
    e->fluid.clear();
}

static int ens_fluid_plan(sph_ensemble *e)
{
    ens_fluid_drop(e);
    e->fl_elig.assign(e->boxes.size(), 0);
    for (size_t k = 0; k < e->boxes.size(); ++k) {
        sph_ctx *c = e->boxes[k];
        if (!ens_fluid_eligible(c)) continue;
        e->fl_elig[k] = 1;
        if (!e->fluid.empty() && c->hs.h.maxiter != e->boxes[e->fluid[0]]->hs.h.maxiter) continue;
        if (ens_learned(c)) {
            if (e->net_max_corr == 0) e->net_max_corr = c->net_max_corr;
            else if (c->net_max_corr != e->net_max_corr) continue;
            if (net_ready(c) != SPH_OK || !c->net) continue;  // the network and its buffers
        }
        e->fluid.push_back((int)k);
    }
    e->fl_gen.clear();
    e->fl_parity.clear();
    e->fl_keys.clear();
    if (e->fluid.empty()) {
        e->fl_planned = true;
        return SPH_OK;
    }
    std::vector<void *> *mem = &e->fl_dmem;
    for (int k : e->fluid) {
        sph_ctx *c = e->boxes[k];
        const sph_state &h = c->hs.h;
        // darcy_step's top-wall cell (from the wall position at the load)
        if (h.nw > 0 && (h.wmode[0] == 1 || h.wmode[0] == 3))
            c->wall0_iz = (unsigned int)(c->host_wx0 / (h.L[2] / h.num[2]));
        e->fl_gen.push_back(c->gen);
        e->fl_parity.push_back(c->nl_parity);
        e->fl_keys.push_back(c->db.keys);
    }
    e->fl_maxiter = e->boxes[e->fluid[0]]->hs.h.maxiter;
    e->fl_phase = 0;
    const std::vector<uint32_t> rbn = ens_rebuild_blocks(e, e->fluid);
    int rc;
    for (int j = 0; j < 2; ++j) {
        std::vector<HashTask> hs;
        std::vector<RebuildTask> rb;
        std::vector<NlBuildTask> nlb;
        std::vector<TopoTask> topo[2];
        std::vector<ForcesTask> frc[2];
        std::vector<BondTask> bnd;
        std::vector<BondGatherTask> bgt;
        std::vector<IntegrateTask> itg;
        std::vector<WallsTask> wls;
        std::vector<ClockTask> clk;
        std::vector<uint32_t> n_hs, n_rb, n_nlb, n_topo[2], n_frc[2], n_bnd, n_bgt, n_itg;
        for (size_t i = 0; i < e->fluid.size(); ++i) {
            const sph_ctx *c = e->boxes[e->fluid[i]];
            const sph_state &h = c->hs.h;
            const StepArgs a = step_args(c);
            const uint32_t cur = c->nl_parity ^ (uint32_t)j;
            DemBuffers b = c->db;
            const bool odd = (radix_passes(c) & 1) != 0;
            if (odd && j == 1) {
                std::swap(b.keys, b.keys_tmp);
                std::swap(b.vals, b.vals_tmp);
            }
            hs.push_back(hash_task(a, c->nla, b, cur, 1u));  // a fluid step rebuilds every step
            n_hs.push_back((h.np + 255) / 256);
            const uint32_t *cond = c->db.nl_flags + cur;
            rb.push_back(rebuild_task(a, b, c->ncells, c->key_bits, cond, c->rb_spin));
            n_rb.push_back(rbn[i]);
            if (odd) {
                std::swap(b.keys, b.keys_tmp);
                std::swap(b.vals, b.vals_tmp);
            }
            nlb.push_back(nl_build_task(a, c->nla, b, cond));
            n_nlb.push_back((h.np + 255) / 256);
            const int near = c->nla.range == 1 ? 1 : 0;
            topo[near].push_back(topology_task(a, b, cur));
            n_topo[near].push_back(topology_blocks(h.np));
            const int hz = h.contactmodel == 3 ? 1 : 0;
            frc[hz].push_back(forces_task(a, b, true));
            n_frc[hz].push_back(forces_blocks(h.np));
            if (h.nb0 > 0) {
                bnd.push_back(bond_task(a, b));
                n_bnd.push_back((h.nb0 + 255) / 256);
                if (b.nbp) {
                    bgt.push_back(bond_gather_task(b));
                    n_bgt.push_back((b.nbp + 255) / 256);
                }
            }
            itg.push_back(integrate_task(a, b, 0, 10.0, h.tau_x, false, NextHash{}));
            n_itg.push_back((h.np + 255) / 256);
            if (h.nw > 0) wls.push_back(walls_task(a, b));
            clk.push_back(ClockTask{b.clock, h.time_dt});
        }
        EnsFluidPhase &P = e->fph[j];
        if ((rc = ens_stage(e, P.hash, hs, n_hs, mem)) || (rc = ens_stage(e, P.rebuild, rb, n_rb, mem)) ||
            (rc = ens_stage(e, P.nlbuild, nlb, n_nlb, mem)) ||
            (rc = ens_stage(e, P.topo[0], topo[0], n_topo[0], mem)) ||
            (rc = ens_stage(e, P.topo[1], topo[1], n_topo[1], mem)) ||
            (rc = ens_stage(e, P.forces[0], frc[0], n_frc[0], mem)) ||
            (rc = ens_stage(e, P.forces[1], frc[1], n_frc[1], mem)) ||
            (rc = ens_stage(e, P.bonds, bnd, n_bnd, mem)) || (rc = ens_stage(e, P.gather, bgt, n_bgt, mem)) ||
            (rc = ens_stage(e, P.integ, itg, n_itg, mem)) || (rc = ens_array(e, P.walls, wls, mem)) ||
            (rc = ens_array(e, P.clocks, clk, mem)))
            return rc;
        P.nwalls = (uint32_t)wls.size();
        P.nclocks = (uint32_t)clk.size();
    }
    // Darcy records (the same for every step)
    std::vector<DarcyTask> all, flux, per2, learned;
    std::vector<uint32_t> n_cells, n_parts, n_sweep, n_all, n_flux, n_per2, n_learned;
    std::vector<sph_net *> nets;
    std::vector<const float *> nparams;
    std::vector<float *> nouts;
    for (int k : e->fluid) {
        const sph_ctx *c = e->boxes[k];
        const sph_state &h = c->hs.h;
        DarcyTask t{};
        t.a = step_args(c);
        t.d = darcy_args(c);
        t.b = c->dd;
        t.x = c->db.x;
        t.force = c->db.force;
        t.cell_start = c->db.cell_start;
        t.cell_end = c->db.cell_end;
        t.x_s = c->db.x_s;
        t.vel_s = c->db.vel_s;
        t.err = c->db.errflags;
        t.clock = c->db.clock;
        t.tolerance = h.tolerance;
        t.maxiter = h.maxiter;
        if (ens_learned(c) && c->net) {
            t.net_in4 = net_input_bf16(c->net);
            t.net_out = c->net_out;
            learned.push_back(t);
            n_learned.push_back(darcy_cell_blocks(t.d));
            nets.push_back(c->net);
            nparams.push_back(c->net_params);
            nouts.push_back(c->net_out);
        }
        all.push_back(t);
        n_cells.push_back(darcy_cell_blocks(t.d));
        n_parts.push_back((h.np + 255) / 256);
        n_sweep.push_back(darcy_sweep_blocks(t.d));
        n_all.push_back(darcy_all_cell_blocks(t.d));
        if (h.bc_bot == 4 || h.bc_top == 4) {
            flux.push_back(t);
            n_flux.push_back(darcy_cell_blocks(t.d));
        }
        if (h.periodic == 2) {
            per2.push_back(t);
            n_per2.push_back(darcy_cell_blocks(t.d));
        }
    }
    if ((rc = ens_stage(e, e->dm_cells, all, n_cells, mem)) || (rc = ens_stage(e, e->dm_parts, all, n_parts, mem)) ||
        (rc = ens_stage(e, e->dm_sweep, all, n_sweep, mem)) || (rc = ens_stage(e, e->dm_all, all, n_all, mem)) ||
        (rc = ens_stage(e, e->dm_flux, flux, n_flux, mem)) || (rc = ens_stage(e, e->dm_per2, per2, n_per2, mem)) ||
        (rc = ens_stage(e, e->dm_net, learned, n_learned, mem)))
        return rc;
    if (!nets.empty()) {
        e->netm = net_multi_create(nets.data(), nparams.data(), nouts.data(), (int)nets.size());
        if (!e->netm) return ens_fail(e, SPH_ERR_HIP, "ensemble: the networks' launch plan failed");
    }
    e->dm_tasks = e->dm_cells.tasks;
    void *st = nullptr, *sm = nullptr;
    if (hipMalloc(&st, 32 * e->fluid.size()) != hipSuccess)
        return ens_fail(e, SPH_ERR_HIP, "ensemble: device allocation failed");
    e->fl_dmem.push_back(st);
    if (hipMalloc(&sm, 12 * e->fluid.size()) != hipSuccess)
        return ens_fail(e, SPH_ERR_HIP, "ensemble: device allocation failed");
    e->fl_dmem.push_back(sm);
    if (hipHostMalloc((void **)&e->dm_states_h, 32 * e->fluid.size()) != hipSuccess ||
        hipHostMalloc((void **)&e->fl_sweep_map_h, 12 * e->fluid.size()) != hipSuccess)
        return ens_fail(e, SPH_ERR_HIP, "ensemble: host allocation failed");
    e->dm_states = (unsigned long long *)st;
    e->fl_sweep_map = (uint32_t *)sm;
    e->fl_sweep_blocks = n_sweep;
    e->fl_planned = true;
    return SPH_OK;
}

static bool ens_fluid_current(const sph_ensemble *e)
{
    if (!e->fl_planned || e->fl_elig.size() != e->boxes.size()) return false;
    size_t i = 0;
    for (size_t k = 0; k < e->boxes.size(); ++k) {
        const sph_ctx *c = e->boxes[k];
        if (ens_fluid_eligible(c) != (e->fl_elig[k] != 0)) return false;
        const bool in = i < e->fluid.size() && e->fluid[i] == (int)k;
        if (!in) continue;
        if (c->gen != e->fl_gen[i]) return false;
        if (c->nl_parity != (e->fl_parity[i] ^ e->fl_phase)) return false;
        const bool swapped = (radix_passes(c) & 1) != 0 && e->fl_phase == 1;
        if ((c->db.keys == e->fl_keys[i]) == swapped) return false;
        ++i;
    }
    return i == e->fluid.size();
}

static MultiMap stage_map(const EnsStage &st) { return MultiMap{st.map, st.nbox}; }

// One step of every fluid box (step_once with darcy_step, multi-box).
static int ens_fluid_step(sph_ensemble *e)
{
    hipStream_t s = e->stream;
    int rc;
    const EnsFluidPhase &P = e->fph[e->fl_phase];
    const DarcyTask *dt = (const DarcyTask *)e->dm_tasks;
    const uint32_t nb = (uint32_t)e->fluid.size();
    // particles: hash, list rebuild, contacts, bonds
    launch_hash_multi(s, (const HashTask *)P.hash.tasks, stage_map(P.hash), P.hash.blocks);
    launch_rebuild_multi(s, (const RebuildTask *)P.rebuild.tasks, stage_map(P.rebuild), P.rebuild.blocks);
    launch_nl_build_multi(s, (const NlBuildTask *)P.nlbuild.tasks, stage_map(P.nlbuild), P.nlbuild.blocks);
    launch_topology_multi(s, (const TopoTask *)P.topo[0].tasks, stage_map(P.topo[0]), P.topo[0].blocks, false);
    launch_topology_multi(s, (const TopoTask *)P.topo[1].tasks, stage_map(P.topo[1]), P.topo[1].blocks, true);
    launch_forces_multi(s, (const ForcesTask *)P.forces[0].tasks, stage_map(P.forces[0]), P.forces[0].blocks, false);
    launch_forces_multi(s, (const ForcesTask *)P.forces[1].tasks, stage_map(P.forces[1]), P.forces[1].blocks, true);
    launch_bonds_multi(s, (const BondTask *)P.bonds.tasks, stage_map(P.bonds), P.bonds.blocks);
    launch_bond_gather_multi(s, (const BondGatherTask *)P.gather.tasks, stage_map(P.gather), P.gather.blocks);
    // darcy_step before the solve
    const DarcyTask *tc = (const DarcyTask *)e->dm_cells.tasks;
    const MultiMap mc = stage_map(e->dm_cells);
    const uint32_t bc = e->dm_cells.blocks;
    launch_dm_ghosts(s, tc, mc, bc, DM_P);
    launch_dm_pressure_gradient(s, tc, mc, bc);
    launch_dm_ghosts(s, tc, mc, bc, DM_GRAD_P);
    launch_dm_pressure_force(s, (const DarcyTask *)e->dm_parts.tasks, stage_map(e->dm_parts), e->dm_parts.blocks);
    launch_dm_porosities(s, tc, mc, bc);
    launch_dm_por_edges(s, (const DarcyTask *)e->dm_per2.tasks, stage_map(e->dm_per2), e->dm_per2.blocks);
    launch_dm_permeabilities(s, tc, mc, bc);
    launch_dm_ghosts(s, tc, mc, bc, DM_PHI);
    launch_dm_ghosts(s, tc, mc, bc, DM_K);
    launch_dm_ghosts_flux(s, (const DarcyTask *)e->dm_flux.tasks, stage_map(e->dm_flux), e->dm_flux.blocks);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    (void)hipEventRecord(t0, s);
    launch_dm_solve_start(s, tc, mc, bc);
    launch_dm_ghosts(s, tc, mc, bc, DM_P);
    launch_dm_prepare(s, tc, mc, bc);
    launch_dm_reset(s, dt, nb);
    // The sweeps, in batches, until every box's solve ended.  A batch ends
    // with the deciding sweep after the next sweep count that a running box
    // needed in its last solve (at least 12 sweeps in, then at least 20 more
    // per batch), and the next batch sweeps only the boxes whose solve had
    // not ended by then: a box with a long solve does not keep the others'
    // blocks in every launch.  The sweeps an ended box skips would return
    // at once, so the results are the same.
    const DarcyTask *tw = (const DarcyTask *)e->dm_sweep.tasks;
    const uint32_t maxiter = e->fl_maxiter;
    std::vector<unsigned long long> target(nb);
    for (uint32_t i = 0; i < nb; ++i)
        target[i] = (e->boxes[e->fluid[i]]->jacobi_hint + 9ull) / 10ull * 10ull + 2ull;
    std::vector<char> running(nb, 1);
    auto next_end = [&](unsigned long long from) {
        unsigned long long best = ~0ull;
        for (uint32_t i = 0; i < nb; ++i)
            if (running[i] && target[i] >= from && target[i] < best) best = target[i];
        return best == ~0ull ? from : best;
    };
    MultiMap mw{};
    uint32_t wblocks = 0;
    auto remap = [&]() {
        uint32_t n = 0, at = 0;
        for (uint32_t i = 0; i < nb; ++i) n += running[i] ? 1u : 0u;
        uint32_t *h = e->fl_sweep_map_h;
        for (uint32_t i = 0, k = 0; i < nb; ++i) {
            if (!running[i]) continue;
            h[k] = at;
            h[n + k] = e->fl_sweep_blocks[i];
            h[2 * n + k] = i;
            at += (e->fl_sweep_blocks[i] + 7u) & ~7u;
            ++k;
        }
        mw = MultiMap{e->fl_sweep_map, n, e->fl_sweep_map + 2 * n};
        wblocks = at;
        return hipMemcpyAsync(e->fl_sweep_map, h, 12 * (size_t)n, hipMemcpyHostToDevice, s);
    };
    HIPCHK_E(remap());
    unsigned int nijac = 0;
    bool all_ended = false;
    unsigned long long want = next_end(12);
    int corrections = 0;
    const DarcyTask *tn = (const DarcyTask *)e->dm_net.tasks;
    const MultiMap mn = stage_map(e->dm_net);
    while (nijac < maxiter && !all_ended) {
        const unsigned int end = (unsigned int)std::min<unsigned long long>(maxiter, want);
        for (; nijac < end; ++nijac) {
            const bool decides = nijac > 0 && (nijac - 1) % 10 == 0;
            const bool correct = decides && e->netm && corrections < e->net_max_corr;
            launch_dm_sweep(s, tw, mw, wblocks, nijac, correct);
            if (correct) {
                launch_dm_net_input(s, tn, mn, e->dm_net.blocks, nijac);
                if ((rc = net_multi_forward(e->netm, s)))
                    return ens_fail(e, rc, std::string("ensemble: ") + sph_net_last_error());
                launch_dm_net_apply(s, tn, mn, e->dm_net.blocks, nijac);
                ++corrections;
            }
        }
        if (nijac == maxiter && (maxiter - 1) % 10 == 0) launch_dm_decide(s, dt, nb, maxiter - 1);
        launch_dm_states(s, dt, nb, e->dm_states);
        HIPCHK_E(hipMemcpyAsync(e->dm_states_h, e->dm_states, 32 * nb, hipMemcpyDeviceToHost, s));
        HIPCHK_E(hipStreamSynchronize(s));
        all_ended = true;
        bool changed = false;
        for (uint32_t i = 0; i < nb; ++i) {
            const bool run = e->dm_states_h[4 * i] == 0ull;
            changed = changed || (run != (running[i] != 0));
            running[i] = run ? 1 : 0;
            all_ended = all_ended && !run;
        }
        if (changed && !all_ended) HIPCHK_E(remap());
        want = next_end((unsigned long long)nijac + 20u);
    }
    launch_dm_normalize(s, (const DarcyTask *)e->dm_all.tasks, stage_map(e->dm_all), e->dm_all.blocks);
    (void)hipEventRecord(t1, s);
    // after the solve (the step holds no output event: dphi is zeroed)
    launch_dm_finish(s, tc, mc, bc);
    launch_dm_ghosts(s, tc, mc, bc, DM_P);
    launch_dm_velocities(s, tc, mc, bc);
    // integration
    launch_integrate_multi(s, (const IntegrateTask *)P.integ.tasks, stage_map(P.integ), P.integ.blocks);
    launch_walls_multi(s, (const WallsTask *)P.walls, P.nwalls);
    launch_clock_multi(s, (const ClockTask *)P.clocks, P.nclocks);
    (void)hipEventSynchronize(t1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    // the host side of the step, box by box (darcy_step's statistics and
    // conv.log lines, step_once's counters)
    for (uint32_t i = 0; i < nb; ++i) {
        sph_ctx *c = e->boxes[e->fluid[i]];
        const unsigned long long *st = e->dm_states_h + 4 * i;
        c->stats.fluid_steps++;
        c->stats.corrections += st[3];
        const unsigned int ran = st[0] != 0 ? (unsigned int)st[1] + 1u : maxiter;
        c->stats.sweeps += ran;
        c->jacobi_hint = ran;
        double max_norm_res;
        std::memcpy(&max_norm_res, &st[2], 8);
        c->stats.last_nijac = (uint32_t)st[1];
        c->stats.last_max_residual = max_norm_res;
        c->stats.solve_ms += ms;
        if (st[0] == 2)
            return ens_fail(e, SPH_ERR_NAN, "box " + std::to_string(e->fluid[i]) +
                                                ": normalized residual is NaN: the system has "
                                                "probably become unstable");
        if (st[0] == 1) {
            if (c->iter % 10 == 0) c->conv.emplace_back(c->iter + 1, (unsigned int)st[1]);
        } else if (maxiter > 0) {
            c->conv.emplace_back(c->iter + 1, maxiter - 1);
            c->stats.nonconverged++;
            fprintf(stderr, "\nIteration %ld, time %g s: Error, the pressure solution in the fluid "
                            "calculations did not converge. Try increasing the value of "
                            "'darcy.maxiter' (%u) or increase 'darcy.tolerance' (%g).\n",
                    c->iter, c->iter * c->hs.h.time_dt, maxiter, c->hs.h.tolerance);
        }
        ens_host_advance(c);
        c->hash_ready = false;  // a fluid step hashes every step
    }
    e->fl_phase ^= 1u;
    e->fluid_steps++;
    return SPH_OK;
}

sph_ensemble *sph_ensemble_create(sph_ctx **boxes, int n)
{
    if (!boxes || n <= 0) {
        g_last_error = "no boxes";
        return nullptr;
    }
    for (int i = 0; i < n; ++i)
        if (!boxes[i] || boxes[i]->device != boxes[0]->device) {
            g_last_error = "null context or contexts on different devices";
            return nullptr;
        }
    sph_ensemble *e = new sph_ensemble();
    e->boxes.assign(boxes, boxes + n);
    e->device = boxes[0]->device;
    (void)hipSetDevice(e->device);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        g_last_error = "ensemble: stream creation failed";
        delete e;
        return nullptr;
    }
    return e;
}

void sph_ensemble_destroy(sph_ensemble *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    ens_drop_plan(e);
    ens_fluid_drop(e);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char *sph_ensemble_last_error(const sph_ensemble *e)
{
    return e ? e->err.c_str() : g_last_error.c_str();
}

int sph_ensemble_stats(const sph_ensemble *e, uint64_t *steps)
{
    if (!e) return SPH_ERR_ARG;
    if (steps) *steps = e->steps;
    return SPH_OK;
}

int sph_ensemble_info(const sph_ensemble *e, sph_ensemble_info_t *out)
{
    if (!e || !out) return SPH_ERR_ARG;
    *out = sph_ensemble_info_t{};
    out->boxes = (uint32_t)e->boxes.size();
    out->multi_boxes = (uint32_t)e->multi.size();
    out->launches_per_step = e->launches_per_step;
    out->graph = e->exec ? 1 : 0;
    out->multi_steps = e->multi_steps;
    out->graph_replays = e->graph_replays;
    out->fluid_boxes = (uint32_t)e->fluid.size();
    out->fluid_steps = e->fluid_steps;
    return SPH_OK;
}

int sph_ensemble_set_graph(sph_ensemble *e, int on)
{
    if (!e) return SPH_ERR_ARG;
    e->use_graph = on != 0;
    if (!e->use_graph && e->exec) {
        (void)hipStreamSynchronize(e->stream);
        (void)hipGraphExecDestroy(e->exec);
        (void)hipGraphDestroy(e->graph);
        e->exec = nullptr;
        e->graph = nullptr;
    }
    return SPH_OK;
}

int sph_ensemble_step(sph_ensemble *e, int nsteps)
{
    if (!e) return SPH_ERR_ARG;
    if (nsteps <= 0) return SPH_OK;
    (void)hipSetDevice(e->device);
    int rc = ens_check(e, nsteps);
    if (rc) return rc;
    // nothing of the boxes' own streams may be pending under the
    // multi-box launches
    for (sph_ctx *c : e->boxes)
        if ((rc = sync(c))) return ens_fail(e, rc, c->err);
    if (!ens_plan_current(e) && (rc = ens_plan(e))) return rc;
    if (!ens_fluid_current(e) && (rc = ens_fluid_plan(e))) return rc;
    std::vector<sph_ctx *> single;
    {
        size_t i = 0, f = 0;
        for (size_t k = 0; k < e->boxes.size(); ++k) {
            if (i < e->multi.size() && e->multi[i] == (int)k) ++i;
            else if (f < e->fluid.size() && e->fluid[f] == (int)k) ++f;
            else single.push_back(e->boxes[k]);
        }
    }
    int left = nsteps;
    if (!e->multi.empty()) {
        // the hash of a box's first step after a load (later ones are
        // formed by its integrate)
        bool hashed = false;
        for (int k : e->multi) {
            sph_ctx *c = e->boxes[k];
            if (c->nl_force || !c->hash_ready) {
                launch_hash_nl(e->stream, step_args(c), c->nla, c->db, c->nl_parity, c->nl_force ? 1u : 0u);
                hashed = true;
            }
        }
        if (hashed || e->phase == 1) {
            ens_enqueue_phase(e, e->ph[e->phase]);
            ens_advance_all(e);
            --left;
        }
        if (left >= 2 && e->use_graph && !e->exec) {
            HIPCHK_E(hipStreamSynchronize(e->stream));
            if ((rc = ens_capture(e))) {
                // capture refused: run the launches without the graph
                if (e->exec) (void)hipGraphExecDestroy(e->exec);
                if (e->graph) (void)hipGraphDestroy(e->graph);
                e->exec = nullptr;
                e->graph = nullptr;
                e->use_graph = false;
                (void)hipGetLastError();
            }
        }
        while (left >= 2) {
            if (e->exec) {
                HIPCHK_E(hipGraphLaunch(e->exec, e->stream));
                e->graph_replays++;
            } else {
                ens_enqueue_phase(e, e->ph[0]);
                ens_enqueue_phase(e, e->ph[1]);
            }
            ens_advance_all(e);
            ens_advance_all(e);
            left -= 2;
        }
        if (left == 1) {
            ens_enqueue_phase(e, e->ph[e->phase]);
            ens_advance_all(e);
        }
    }
    // the fluid boxes, step by step
    if (!e->fluid.empty())
        for (int k = 0; k < nsteps; ++k)
            if ((rc = ens_fluid_step(e))) return rc;
    // the other boxes, box by box on their own streams
    for (int k = 0; k < nsteps; ++k)
        for (sph_ctx *c : single)
            if ((rc = step_once(c, nullptr, nullptr, std::string()))) return ens_fail(e, rc, c->err);
    HIPCHK_E(hipStreamSynchronize(e->stream));
    for (sph_ctx *c : e->boxes) {
        if ((rc = check_flags(c))) return ens_fail(e, rc, c->err);
        if ((rc = sync(c))) return ens_fail(e, rc, c->err);
    }
    e->steps += (uint64_t)nsteps;
    return SPH_OK;
}

}  // extern "C"
