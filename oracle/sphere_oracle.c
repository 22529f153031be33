/*
 * sphere_oracle.c -- TEST INFRASTRUCTURE ONLY (see sphere_oracle.h).
 *
 * FP64 CPU restatement of the reference DEM step and Darcy solve.  Every
 * function cites the reference file:line it follows.  Arithmetic is written
 * in the reference's operation order (CUDA vector operators are component
 * wise, src/vector_arithmetic.h) and is built with -ffp-contract=off, so it
 * is IEEE FP64 with no fused multiply-add.  Parallel loops only split work
 * whose results do not depend on the split (each particle or cell writes its
 * own outputs), so the result is independent of the thread count.
 *
 * Declared deviations from the reference (the reference is not FP64 or is
 * ill-defined at these points):
 *   D1  length() uses FP64 sqrt (reference: sqrtf on a double, fast math).
 *   D2  contactmodel 3 is the Hertz-Mindlin law specified below (the
 *       reference calls powf on a negative overlap and produces NaN).
 *   D3  bonds accumulate into force/torque in ascending bond order
 *       (reference: unsynchronised scatter, cohesion.cuh:220-226); broken
 *       bonds are tracked by a flag and written as nb0.
 *   D4  a 33rd contact on one particle is counted as overflow and not
 *       written (reference writes past the slot row, contactsearch.cuh:298).
 *   D5  wall0mode is 0 when nw == 0 (reference reads it uninitialised).
 *   D6  sphere volume uses r*r*r and permeability phi*phi*phi and
 *       (1-phi)*(1-phi) instead of pow().
 *   D7  tangential displacement keeps delta_t0 when neither the old spring
 *       nor the sliding velocity is non-zero (reference stores an
 *       uninitialised register, contactmodels.cuh:311-393).
 *   D8  setDarcyTopPressure writes only x < nx, y < ny, z < nz cells
 *       (reference operator precedence, darcy.cuh:1206-1208).
 *   D9  cellEnd of empty cells is 0xffffffff (reference leaves it stale).
 *   D10 the Jacobi residual field is reduced to its maximum, not stored.
 *   D11 the ghost permeability of a bc-4 (prescribed flux) face is 0: the
 *       reference never writes it (setDarcyGhostNodes handles bc 0-2,
 *       darcy.cuh:194-264) in memory it does not clear (darcy.cuh:32).
 */
#include "sphere_oracle.h"

#include <math.h>
#include <time.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EMPTY 0xffffffffu

/* src/constants.h:15 */
static const double PI_REF = 3.14159265358979;
/* PI_M as used by sphereVolume, src/darcy.cuh:317 */
static const double PI_M = 3.14159265358979323846; /* M_PI */

typedef struct { double x, y, z; } v3;
static inline v3 mk(double x, double y, double z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 smul(double s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 sdiv(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
/* src/vector_arithmetic.h:307-310: x*x + y*y + z*z, left to right */
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* src/vector_arithmetic.h:317-320 (D1: FP64 sqrt) */
static inline double len(v3 a) { return sqrt(dot(a, a)); }
/* src/vector_arithmetic.h:410-413 */
static inline v3 cross(v3 a, v3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 ld4(const double *p) { return mk(p[0], p[1], p[2]); }
static inline v3 ld3(const double *p) { return mk(p[0], p[1], p[2]); }
static inline void st3(double *p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* ns_oracle.c: one fluid iteration with the DEM oracle's grains */
int nso_step_coupled(void *h, uint32_t np, double *x, double *vel, double *xs, uint32_t *cstart,
                     uint32_t *cend, uint32_t *gidx, const double *origo, int periodic, double *force,
                     double wall_x);
void nso_output_event(void *h);

int orc_get_threads(void) { return g_threads; }
/* contact slots per particle of the next orc_create (the reference's NC) */
static int g_nc = SPH_NC;
void orc_set_contact_slots(int nc) { g_nc = (nc == 64) ? 64 : 32; }

struct orc {
    sph_state h;            /* header scalars; array pointers unused */
    void *ns;               /* an attached Navier-Stokes fluid (ns_oracle.c) */
    uint32_t np, nb0, nw, ncells;
    /* particle state, orig order; x/vel stride 4 (w = radius / fixvel) */
    double *x, *vel, *angvel4;
    double *xyzsum, *force, *angpos, *torque, *acc, *angacc;
    double *es_dot, *es, *ev_dot, *ev, *p;
    int32_t *color;
    /* bonds */
    uint32_t *bonds;
    uint8_t *broken;
    double *bdn, *bdt, *bon, *bot;
    /* walls */
    double w_acc[SPH_MAXWALLS];
    double *walls_force_pp, *wall_partial, *tau_pp, *tau_partial;
    uint32_t nblocks;
    /* sorting, src/datatypes.h:52-60 */
    uint32_t *cellid, *index, *cell_start, *cell_end, *cnt;
    double *x_s, *vel_s, *angvel_s;
    /* contact history, src/device.cu:436-439 */
    uint32_t *contacts;
    double *distmod, *delta_t;
    int nc;                 /* contact slots per particle */
    long overflow;
    /* Darcy (cells with one ghost layer) */
    uint32_t dcells;
    double *d_p, *d_p_old, *d_p_new, *d_dp_expl, *d_phi, *d_dphi, *d_k, *d_norm;
    double *d_v, *d_vp_avg, *d_grad_p;   /* 3 per cell */
    int32_t *d_pconst;
    double *f_p;                          /* np*3 */
    sph_darcy_stats stats;
    uint32_t *conv;
    int nconv, cap_conv;
    /* time loop host variables, src/device.cu:913-1036 */
    long iter;
    double filetimeclock;
    double host_wx0;                      /* host copy of walls.nx[0].w */
    unsigned int wall0_iz;
    double p_top_orig;
    int velocity_state;
    char err[256];
};

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

const char *orc_error(const orc *o) { return o ? o->err : "null oracle"; }

static inline unsigned int didx(const orc *o, int x, int y, int z)
{
    /* src/darcy.cuh:134-146 */
    return (unsigned)(x + 1) + (o->h.num[0] + 2) * (unsigned)(y + 1) +
           (o->h.num[0] + 2) * (o->h.num[1] + 2) * (unsigned)(z + 1);
}

/* ------------------------------------------------------------------ */
/* creation                                                            */
/* ------------------------------------------------------------------ */

static int check_values(const sph_state *s, char *err, int errlen)
{
    /* src/sphere.cpp:145-345 */
    if (s->nd != 3) { snprintf(err, errlen, "nd = %d, ND = 3", s->nd); return SPH_ERR_VALUE; }
    if (s->time_current > s->time_total || s->time_current < 0.0) {
        snprintf(err, errlen, "time.current = %g s, time.total = %g s", s->time_current, s->time_total);
        return SPH_ERR_VALUE;
    }
    if (s->origo[0] < 0.0 || s->origo[1] < 0.0 || s->origo[2] < 0.0) {
        snprintf(err, errlen, "Negative values in grid.origo"); return SPH_ERR_VALUE;
    }
    if (s->L[0] <= 0.0 || s->L[1] <= 0.0 || s->L[2] <= 0.0) {
        snprintf(err, errlen, "grid.L must be positive"); return SPH_ERR_VALUE;
    }
    if (s->num[0] == 0 || s->num[1] == 0 || s->num[2] == 0) {
        snprintf(err, errlen, "grid.num must be positive"); return SPH_ERR_VALUE;
    }
    if (s->periodic == 2 && s->num[0] < 3) { snprintf(err, errlen, "periodic x needs >= 3 cells"); return SPH_ERR_VALUE; }
    if (s->periodic == 1 && (s->num[0] < 3 || s->num[1] < 3)) {
        snprintf(err, errlen, "periodic xy needs >= 3 cells"); return SPH_ERR_VALUE;
    }
    for (uint32_t i = 0; i < s->np; ++i) {
        const double *x = s->x + 4 * (size_t)i;
        if (x[3] <= 0.0) { snprintf(err, errlen, "Particle %u has a radius of %g m", i, x[3]); return SPH_ERR_VALUE; }
        if (x[0] != x[0] || x[1] != x[1] || x[2] != x[2] || x[3] != x[3]) {
            snprintf(err, errlen, "Particle %u position is NaN", i); return SPH_ERR_VALUE;
        }
        if (x[0] < s->origo[0] || x[1] < s->origo[1] || x[2] < s->origo[2] ||
            x[0] > s->L[0] || x[1] > s->L[1] || x[2] > s->L[2]) {
            snprintf(err, errlen, "Particle %u is outside of the computational grid", i);
            return SPH_ERR_VALUE;
        }
    }
    if (s->nw > SPH_MAXWALLS) { snprintf(err, errlen, "nw = %u > MAXWALLS", s->nw); return SPH_ERR_VALUE; }
    if (s->nw > 0 && s->wmode[0] != 0 && s->np > 0) {
        double zmax = -1e300;
        for (uint32_t i = 0; i < s->np; ++i) if (s->x[4 * (size_t)i + 2] > zmax) zmax = s->x[4 * (size_t)i + 2];
        if (s->w_nx[3] < zmax) { snprintf(err, errlen, "particles above the upper, dynamic wall"); return SPH_ERR_VALUE; }
    }
    if (s->k_n <= 0.0) { snprintf(err, errlen, "k_n = %g N/m", s->k_n); return SPH_ERR_VALUE; }
    if (s->contactmodel < 1 || s->contactmodel > 3) { /* src/sphere.cpp:384-393 */
        snprintf(err, errlen, "contact model value (%u) not understood", s->contactmodel);
        return SPH_ERR_VALUE;
    }
    if (s->E < 0.0) { snprintf(err, errlen, "E = %g", s->E); return SPH_ERR_VALUE; }
    if (s->rho <= 0.0) { snprintf(err, errlen, "rho = %g kg/m3", s->rho); return SPH_ERR_VALUE; }
    if (s->fluid) {
        if (s->cfd_solver != 1) { snprintf(err, errlen, "cfd_solver = %d (only Darcy, 1)", s->cfd_solver); return SPH_ERR_VALUE; }
        if (s->rho_f <= 0.0) { snprintf(err, errlen, "rho_f = %g", s->rho_f); return SPH_ERR_VALUE; }
        if (s->mu <= 0.0) { snprintf(err, errlen, "mu = %g", s->mu); return SPH_ERR_VALUE; }
        if (s->beta_f <= 0.0) { snprintf(err, errlen, "beta_f = %g", s->beta_f); return SPH_ERR_VALUE; }
        if (s->k_c <= 0.0) { snprintf(err, errlen, "k_c = %g", s->k_c); return SPH_ERR_VALUE; }
        if (s->ndem == 0) { snprintf(err, errlen, "ndem = 0"); return SPH_ERR_VALUE; }
    }
    return SPH_OK;
}

orc *orc_create(const sph_state *s, char *err, int errlen)
{
    int rc = check_values(s, err, errlen);
    if (rc != SPH_OK) return NULL;
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->h = *s;
    uint32_t np = s->np;
    o->np = np; o->nb0 = s->nb0; o->nw = s->nw;
    o->ncells = s->num[0] * s->num[1] * s->num[2];
    o->x = xcalloc((size_t)np * 4, 8); o->vel = xcalloc((size_t)np * 4, 8);
    o->angvel4 = xcalloc((size_t)np * 4, 8);
    o->xyzsum = xcalloc((size_t)np * 3, 8); o->force = xcalloc((size_t)np * 3, 8);
    o->angpos = xcalloc((size_t)np * 3, 8); o->torque = xcalloc((size_t)np * 3, 8);
    o->acc = xcalloc((size_t)np * 3, 8); o->angacc = xcalloc((size_t)np * 3, 8);
    o->es_dot = xcalloc(np, 8); o->es = xcalloc(np, 8); o->ev_dot = xcalloc(np, 8);
    o->ev = xcalloc(np, 8); o->p = xcalloc(np, 8);
    o->color = xcalloc(np, 4);
    if (np) {
        memcpy(o->x, s->x, (size_t)np * 32);
        memcpy(o->vel, s->vel, (size_t)np * 32);
        for (uint32_t i = 0; i < np; ++i) {
            o->angvel4[4 * (size_t)i + 0] = s->angvel[3 * (size_t)i + 0];
            o->angvel4[4 * (size_t)i + 1] = s->angvel[3 * (size_t)i + 1];
            o->angvel4[4 * (size_t)i + 2] = s->angvel[3 * (size_t)i + 2];
        }
        memcpy(o->xyzsum, s->xyzsum, (size_t)np * 24);
        memcpy(o->force, s->force, (size_t)np * 24);
        memcpy(o->angpos, s->angpos, (size_t)np * 24);
        memcpy(o->torque, s->torque, (size_t)np * 24);
        memcpy(o->es_dot, s->es_dot, (size_t)np * 8);
        memcpy(o->es, s->es, (size_t)np * 8);
        memcpy(o->ev_dot, s->ev_dot, (size_t)np * 8);
        memcpy(o->ev, s->ev, (size_t)np * 8);
        memcpy(o->p, s->p, (size_t)np * 8);
        if (s->color) memcpy(o->color, s->color, (size_t)np * 4);
    }
    uint32_t nb = s->nb0;
    o->bonds = xcalloc((size_t)nb * 2, 4); o->broken = xcalloc(nb, 1);
    o->bdn = xcalloc(nb, 8); o->bdt = xcalloc((size_t)nb * 3, 8);
    o->bon = xcalloc(nb, 8); o->bot = xcalloc((size_t)nb * 3, 8);
    if (nb) {
        memcpy(o->bonds, s->bonds, (size_t)nb * 8);
        memcpy(o->bdn, s->bonds_delta_n, (size_t)nb * 8);
        memcpy(o->bdt, s->bonds_delta_t, (size_t)nb * 24);
        memcpy(o->bon, s->bonds_omega_n, (size_t)nb * 8);
        memcpy(o->bot, s->bonds_omega_t, (size_t)nb * 24);
        /* cohesion.cuh:30: bonds with bond.x >= np are erased */
        for (uint32_t b = 0; b < nb; ++b)
            if (o->bonds[2 * (size_t)b] >= np || o->bonds[2 * (size_t)b + 1] >= np) o->broken[b] = 1;
    }
    o->nblocks = (np + 255) / 256;
    o->walls_force_pp = xcalloc((size_t)np * (o->nw ? o->nw : 1), 8);
    o->wall_partial = xcalloc(o->nblocks, 8);
    o->tau_pp = xcalloc(np, 8);
    o->tau_partial = xcalloc(o->nblocks, 8);
    o->cellid = xcalloc(np, 4); o->index = xcalloc(np, 4);
    o->cell_start = xcalloc(o->ncells, 4); o->cell_end = xcalloc(o->ncells, 4);
    o->cnt = xcalloc(o->ncells + 1, 4);
    o->x_s = xcalloc((size_t)np * 4, 8); o->vel_s = xcalloc((size_t)np * 4, 8);
    o->angvel_s = xcalloc((size_t)np * 4, 8);
    o->nc = g_nc;
    const int NC = o->nc;
    o->contacts = xcalloc((size_t)np * NC, 4);
    for (size_t i = 0; i < (size_t)np * NC; ++i) o->contacts[i] = np; /* device.cu:471-473 */
    o->distmod = xcalloc((size_t)np * NC * 3, 8);
    o->delta_t = xcalloc((size_t)np * NC * 3, 8);
    /* Darcy */
    if (s->fluid) {
        uint32_t nx = s->num[0], ny = s->num[1], nz = s->num[2];
        o->dcells = (nx + 2) * (ny + 2) * (nz + 2);
        size_t c = o->dcells;
        o->d_p = xcalloc(c, 8); o->d_p_old = xcalloc(c, 8); o->d_p_new = xcalloc(c, 8);
        o->d_dp_expl = xcalloc(c, 8); o->d_phi = xcalloc(c, 8); o->d_dphi = xcalloc(c, 8);
        o->d_k = xcalloc(c, 8); o->d_norm = xcalloc(c, 8);
        o->d_v = xcalloc(c * 3, 8); o->d_vp_avg = xcalloc(c * 3, 8); o->d_grad_p = xcalloc(c * 3, 8);
        o->d_pconst = xcalloc(c, 4);
        for (uint32_t z = 0; z < nz; ++z)
            for (uint32_t y = 0; y < ny; ++y)
                for (uint32_t x = 0; x < nx; ++x) {
                    size_t f = x + (size_t)nx * (y + (size_t)ny * z);
                    unsigned int ci = didx(o, x, y, z);
                    o->d_v[3 * ci + 0] = s->v_f[3 * f + 0];
                    o->d_v[3 * ci + 1] = s->v_f[3 * f + 1];
                    o->d_v[3 * ci + 2] = s->v_f[3 * f + 2];
                    o->d_p[ci] = s->p_f[f];
                    o->d_phi[ci] = s->phi[f];
                    o->d_dphi[ci] = s->dphi[f];
                    o->d_pconst[ci] = s->p_constant ? s->p_constant[f] : 0;
                }
        o->f_p = xcalloc((size_t)np * 3, 8);
        if (np && s->f_p) memcpy(o->f_p, s->f_p, (size_t)np * 24);
        /* src/device.cu:1017-1021 */
        o->p_top_orig = o->d_p[didx(o, 0, 0, nz - 1)] -
                        s->p_mod_A * sin(2.0 * PI_M * s->p_mod_f * s->time_current + s->p_mod_phi);
    }
    o->host_wx0 = o->nw > 0 ? s->w_nx[3] : 0.0;
    o->wall0_iz = 10000000u;  /* src/device.cu:1008 */
    o->velocity_state = 1;
    o->cap_conv = 1024;
    o->conv = xcalloc((size_t)o->cap_conv * 2, 4);
    return o;
}

void orc_destroy(orc *o)
{
    if (!o) return;
    void *ptrs[] = {o->x, o->vel, o->angvel4, o->xyzsum, o->force, o->angpos, o->torque,
                    o->acc, o->angacc, o->es_dot, o->es, o->ev_dot, o->ev, o->p, o->color,
                    o->bonds, o->broken, o->bdn, o->bdt, o->bon, o->bot, o->walls_force_pp,
                    o->wall_partial, o->tau_pp, o->tau_partial, o->cellid, o->index,
                    o->cell_start, o->cell_end, o->cnt, o->x_s, o->vel_s, o->angvel_s,
                    o->contacts, o->distmod, o->delta_t, o->d_p, o->d_p_old, o->d_p_new,
                    o->d_dp_expl, o->d_phi, o->d_dphi, o->d_k, o->d_norm, o->d_v,
                    o->d_vp_avg, o->d_grad_p, o->d_pconst, o->f_p, o->conv};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(o);
}

/* ------------------------------------------------------------------ */
/* A1-A3: hash, stable sort, reorder                                  */
/* ------------------------------------------------------------------ */

/* src/sorting.cuh:7-20 and 25-51.  Returns 0 if the particle is outside the
 * grid (the reference's unsigned conversion is undefined there). */
static int calc_cell(const orc *o, const double *x, unsigned int *cell)
{
    const sph_state *h = &o->h;
    double fx = floor((x[0] - h->origo[0]) / (h->L[0] / h->num[0]));
    double fy = floor((x[1] - h->origo[1]) / (h->L[1] / h->num[1]));
    double fz = floor((x[2] - h->origo[2]) / (h->L[2] / h->num[2]));
    if (!(fx >= 0.0 && fx < h->num[0] && fy >= 0.0 && fy < h->num[1] && fz >= 0.0 && fz < h->num[2]))
        return 0;
    unsigned int ix = (unsigned int)fx, iy = (unsigned int)fy, iz = (unsigned int)fz;
    *cell = (iz * h->num[1]) * h->num[0] + iy * h->num[0] + ix;
    return 1;
}

static int hash_sort_reorder(orc *o)
{
    uint32_t np = o->np, nc = o->ncells;
    int bad = 0;
    /* calcParticleCellID, sorting.cuh:25-51 */
#pragma omp parallel for num_threads(g_threads) reduction(| : bad) schedule(static)
    for (long i = 0; i < (long)np; ++i) {
        unsigned int c;
        if (!calc_cell(o, o->x + 4 * i, &c)) { bad |= 1; c = 0; }
        o->cellid[i] = c;
    }
    if (bad) { snprintf(o->err, sizeof(o->err), "a particle moved outside the grid"); return SPH_ERR_GRID; }
    /* thrust::sort_by_key, device.cu:1082-1085: a stable sort by cell with
     * ties in original-index order (counting sort) */
    memset(o->cnt, 0, (size_t)(nc + 1) * 4);
    for (uint32_t i = 0; i < np; ++i) o->cnt[o->cellid[i] + 1]++;
    for (uint32_t c = 0; c < nc; ++c) o->cnt[c + 1] += o->cnt[c];
    uint32_t *keys = (uint32_t *)malloc((size_t)(np ? np : 1) * 4);
    for (uint32_t i = 0; i < np; ++i) {
        uint32_t pos = o->cnt[o->cellid[i]]++;
        keys[pos] = o->cellid[i];
        o->index[pos] = i;
    }
    memcpy(o->cellid, keys, (size_t)np * 4);
    free(keys);
    /* cudaMemset(cellStart, 0xff) + reorderArrays, device.cu:1096, sorting.cuh:56-132 */
    for (uint32_t c = 0; c < nc; ++c) { o->cell_start[c] = EMPTY; o->cell_end[c] = EMPTY; }
    for (uint32_t idx = 0; idx < np; ++idx) {
        uint32_t c = o->cellid[idx];
        if (idx == 0 || c != o->cellid[idx - 1]) {
            o->cell_start[c] = idx;
            if (idx > 0) o->cell_end[o->cellid[idx - 1]] = idx;
        }
        if (idx == np - 1) o->cell_end[c] = idx + 1;
    }
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long idx = 0; idx < (long)np; ++idx) {
        uint32_t s = o->index[idx];
        memcpy(o->x_s + 4 * idx, o->x + 4 * (size_t)s, 32);
        memcpy(o->vel_s + 4 * idx, o->vel + 4 * (size_t)s, 32);
        memcpy(o->angvel_s + 4 * idx, o->angvel4 + 4 * (size_t)s, 32);
    }
    return SPH_OK;
}

/* ------------------------------------------------------------------ */
/* A4: periodic images                                                 */
/* ------------------------------------------------------------------ */

/* src/contactsearch.cuh:11-83 */
static int find_dist_mod(const orc *o, int *t, v3 *dm)
{
    const sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    if (h->periodic == 1) {
        if (t[0] < 0) { t[0] += nx; *dm = add(*dm, mk(h->L[0], 0.0, 0.0)); }
        if (t[0] >= nx) { t[0] -= nx; *dm = sub(*dm, mk(h->L[0], 0.0, 0.0)); }
        if (t[1] < 0) { t[1] += ny; *dm = add(*dm, mk(0.0, h->L[1], 0.0)); }
        if (t[1] >= ny) { t[1] -= ny; *dm = sub(*dm, mk(0.0, h->L[1], 0.0)); }
    }
    if (h->periodic == 2) {
        if (t[0] < 0) { t[0] += nx; *dm = add(*dm, mk(h->L[0], 0.0, 0.0)); }
        if (t[0] >= nx) { t[0] -= nx; *dm = sub(*dm, mk(h->L[0], 0.0, 0.0)); }
        if (t[1] < 0 || t[1] >= ny) return -1;
    }
    if (h->periodic > 2 || h->periodic < 1) {
        if (t[0] < 0 || t[0] >= nx) return -1;
        if (t[1] < 0 || t[1] >= ny) return -1;
    }
    if (t[2] < 0 || t[2] >= nz) return -1;
    return 0;
}

static inline unsigned int cell_of_target(const orc *o, const int *t)
{
    /* contactsearch.cuh:217-218 */
    return (unsigned)t[0] + (unsigned)t[1] * o->h.num[0] + (o->h.num[0] * o->h.num[1]) * (unsigned)t[2];
}

static void grid_pos(const orc *o, const double *x, int *g)
{
    /* contactsearch.cuh:352-354 */
    const sph_state *h = &o->h;
    g[0] = (int)floor((x[0] - h->origo[0]) / (h->L[0] / h->num[0]));
    g[1] = (int)floor((x[1] - h->origo[1]) / (h->L[1] / h->num[1]));
    g[2] = (int)floor((x[2] - h->origo[2]) / (h->L[2] / h->num[2]));
}

/* ------------------------------------------------------------------ */
/* A5: topology (contact slot bookkeeping)                             */
/* ------------------------------------------------------------------ */

/* src/contactsearch.cuh:194-322 and 329-371 */
static long topology_one(orc *o, uint32_t idx_a)
{
    const int NC = o->nc;
    long overflow = 0;
    const double *xa4 = o->x_s + 4 * (size_t)idx_a;
    v3 x_a = ld4(xa4);
    double r_a = xa4[3];
    int nc = 0;
    int gp[3];
    grid_pos(o, xa4, gp);
    uint32_t np = o->np;
    uint32_t a_orig = o->index[idx_a];
    uint32_t *row = o->contacts + (size_t)a_orig * NC;
    for (int zd = -1; zd < 2; ++zd)
        for (int yd = -1; yd < 2; ++yd)
            for (int xd = -1; xd < 2; ++xd) {
                int t[3] = {gp[0] + xd, gp[1] + yd, gp[2] + zd};
                v3 dm = mk(0.0, 0.0, 0.0);
                if (find_dist_mod(o, t, &dm) == -1) continue;
                unsigned int cid = cell_of_target(o, t);
                uint32_t start = o->cell_start[cid];
                if (start == EMPTY) continue;
                uint32_t end = o->cell_end[cid];
                for (uint32_t idx_b = start; idx_b < end; ++idx_b) {
                    if (idx_b == idx_a) continue;
                    const double *xb4 = o->x_s + 4 * (size_t)idx_b;
                    double r_b = xb4[3];
                    uint32_t b_orig = o->index[idx_b];
                    v3 x_ab = sub(x_a, ld4(xb4));
                    x_ab = add(x_ab, dm);
                    double l = len(x_ab);
                    double delta = l - (r_a + r_b);
                    if (delta < 0.0) {
                        int cpos = nc;
                        for (int i = 0; i < NC; ++i) {
                            uint32_t cidx = row[i];
                            if (cidx == np) cpos = i;
                            else if (cidx == b_orig) { cpos = i; break; }
                        }
                        if (cpos < NC) {
                            row[cpos] = b_orig;
                            st3(o->distmod + 3 * ((size_t)a_orig * NC + cpos), dm);
                        } else {
                            overflow++; /* D4 */
                        }
                        ++nc;
                        if (nc > NC) goto next_cell;
                    }
                }
            next_cell:;
            }
    return overflow;
}

/* ------------------------------------------------------------------ */
/* A7: contactLinear (model 2), A8: Hertz-Mindlin (model 3)           */
/* ------------------------------------------------------------------ */

typedef struct {
    v3 F, T;
    double es_dot, ev_dot, p;
} acc_t;

/* src/contactmodels.cuh:241-408 */
static void contact_linear(const orc *o, acc_t *A, uint32_t b_orig, v3 vel_a, v3 angvel_a,
                           double r_a, double r_b, v3 x, double x_length, double delta,
                           double *dt_slot)
{
    const sph_state *h = &o->h;
    const double dt = h->time_dt;
    v3 vel_b = ld4(o->vel + 4 * (size_t)b_orig);
    v3 angvel_b = ld4(o->angvel4 + 4 * (size_t)b_orig);
    v3 delta_t0_uncor = ld3(dt_slot);
    v3 n = sdiv(x, x_length);
    v3 vel_linear = sub(vel_a, vel_b);
    v3 vel = add(add(vel_linear, smul(r_a + delta / 2.0, cross(n, angvel_a))),
                 smul(r_b + delta / 2.0, cross(n, angvel_b)));
    double vel_n = -dot(vel_linear, n);
    v3 vel_t = sub(vel, smul(dot(n, vel), n));
    double vel_t_length = len(vel_t);
    v3 delta_t0 = sub(delta_t0_uncor, smul(dot(n, delta_t0_uncor), n));
    double delta_t0_length = len(delta_t0);
    v3 delta_t = delta_t0; /* D7 */
    double k_n = h->k_n, k_t = h->k_t;
    if (h->E > .001) {
        k_n = 3.141592654 / 2.0 * h->E * (r_a + r_b) / 2.;
        k_t = k_n;
    }
    v3 f_n = smul(fmax(0.0, -k_n * delta + h->gamma_n * vel_n), n);
    double f_n_length = len(f_n);
    A->ev_dot += 0.5 * h->gamma_n * vel_n * vel_n;
    v3 f_c = smul(-h->kappa * sqrt(r_a * r_b), n);
    v3 f_t = mk(0.0, 0.0, 0.0);
    if (delta_t0_length > 0.0 || vel_t_length > 0.0) {
        delta_t = add(delta_t0, smul(dt, vel_t));
        v3 f_t_elast = smul(-k_t, delta_t);
        v3 f_t_visc = smul(-h->gamma_t, vel_t);
        f_t = add(f_t_elast, f_t_visc);
        double f_t_length = len(f_t);
        double f_t_limit = h->mu_s * len(sub(f_n, f_c));
        A->ev_dot += 0.5 * h->gamma_t * vel_t_length * vel_t_length;
        if (f_t_length > f_t_limit) {
            v3 t = sdiv(f_t, len(f_t));
            f_t = smul(f_t_limit, t);
            delta_t = smul(-1.0 / k_t, add(smul(h->mu_d * len(sub(f_n, f_c)), t), smul(h->gamma_t, vel_t)));
            A->es_dot += 0.5 * len(smul(dt, smul(len(f_t), vel_t))) / dt;
        }
    }
    A->F = add(A->F, add(add(f_n, f_t), f_c));
    A->T = add(A->T, cross(smul(-(r_a + delta * 0.5), n), f_t));
    A->p += f_n_length / (4.0 * PI_REF * r_a * r_a);
    st3(dt_slot, delta_t);
}

/*
 * contactmodel 3: Hertz-Mindlin (D2).  Structure of contactHertz
 * (src/contactmodels.cuh:414-591) with the overlap magnitude d = -delta and
 * powers written as d^(3/2) = d*sqrt(d), d^(1/4) = sqrt(sqrt(d)).
 * If E > 0.001 the coefficients follow Hertz (1882) and Mindlin &
 * Deresiewicz (1953) for two spheres of one material, with Poisson's ratio
 * nu carried in the k_r field:
 *   E* = E / (2 (1 - nu^2)),  R* = r_a r_b / (r_a + r_b),
 *   G  = E / (2 (1 + nu)),    G* = G / (2 (2 - nu)),
 *   k_n = 4/3 E* sqrt(R*)  (normal force k_n d^(3/2)),
 *   k_t = 8 G* sqrt(R*)    (tangential stiffness k_t sqrt(d)).
 * Otherwise k_n, k_t are taken from the parameters.  The viscous terms are
 * gamma_n d^(1/4) v_n and gamma_t d^(1/4) v_t.  The energy rates are the
 * reference's (contactmodels.cuh:506,560): ev_dot += gamma_n v_n^2, and on a
 * slip es_dot += |delta_t0 - delta_t| k_t / dt with this law's k_t.
 */
static void contact_hertz(const orc *o, acc_t *A, uint32_t b_orig, v3 vel_a, v3 angvel_a,
                          double r_a, double r_b, v3 x_ab, double x_ab_length, double delta_ab,
                          double *dt_slot)
{
    const sph_state *h = &o->h;
    const double dt = h->time_dt;
    v3 vel_b = ld4(o->vel + 4 * (size_t)b_orig);
    v3 angvel_b = ld4(o->angvel4 + 4 * (size_t)b_orig);
    v3 delta_t0_uncor = ld3(dt_slot);
    v3 n_ab = sdiv(x_ab, x_ab_length);
    v3 vel_ab_linear = sub(vel_a, vel_b);
    v3 vel_ab = add(add(vel_ab_linear, smul(r_a + delta_ab / 2.0, cross(n_ab, angvel_a))),
                    smul(r_b + delta_ab / 2.0, cross(n_ab, angvel_b)));
    v3 angvel_ab = sub(angvel_a, angvel_b);
    double angvel_ab_length = len(angvel_ab);
    double vel_n_ab = dot(vel_ab_linear, n_ab);
    v3 vel_t_ab = sub(vel_ab, smul(dot(vel_ab, n_ab), n_ab));
    double vel_t_ab_length = len(vel_t_ab);
    v3 delta_t0 = sub(delta_t0_uncor, smul(dot(delta_t0_uncor, n_ab), n_ab));
    double delta_t0_length = len(delta_t0);
    v3 delta_t = delta_t0;
    double d = -delta_ab;
    double sd = sqrt(d);
    double qd = sqrt(sd);
    double k_n = h->k_n, k_t = h->k_t;
    if (h->E > .001) {
        double nu = h->k_r;
        double E_star = h->E / (2.0 * (1.0 - nu * nu));
        double R_star = r_a * r_b / (r_a + r_b);
        double G = h->E / (2.0 * (1.0 + nu));
        double G_star = G / (2.0 * (2.0 - nu));
        double sR = sqrt(R_star);
        k_n = 4.0 / 3.0 * E_star * sR;
        k_t = 8.0 * G_star * sR;
    }
    v3 f_n = smul(fmax(0.0, k_n * d * sd - h->gamma_n * qd * vel_n_ab), n_ab);
    /* contactmodels.cuh:506: the normal damping power gamma_n v_n^2 */
    A->ev_dot += h->gamma_n * vel_n_ab * vel_n_ab;
    double f_n_length = len(f_n);
    v3 f_c = smul(-h->kappa * sqrt(r_a * r_b), n_ab);
    v3 f_t = mk(0.0, 0.0, 0.0);
    v3 T_res = mk(0.0, 0.0, 0.0);
    if (delta_t0_length > 0.0 || vel_t_ab_length > 0.0) {
        v3 f_t_elast = smul(-k_t * sd, delta_t0);
        v3 f_t_visc = smul(-h->gamma_t * qd, vel_t_ab);
        double f_t_limit;
        if (vel_t_ab_length > (double)0.001f)
            f_t_limit = h->mu_d * len(sub(f_n, f_c));
        else
            f_t_limit = h->mu_s * len(sub(f_n, f_c));
        f_t = add(f_t_elast, f_t_visc);
        double f_t_length = len(f_t);
        if (f_t_length > f_t_limit) {
            f_t = smul(f_t_limit / f_t_length, f_t);
            delta_t = sdiv(add(f_t, smul(h->gamma_t * qd, vel_t_ab)), -k_t * sd);
            /* contactmodels.cuh:560: the spring energy released by the slip */
            A->es_dot += len(sub(delta_t0, delta_t)) * k_t / dt;
        } else {
            delta_t = add(delta_t0, smul(dt, vel_t_ab));
        }
    }
    if (angvel_ab_length > 0.0) {
        double m = fmin(h->gamma_r * r_a * angvel_ab_length, h->mu_r * r_a * f_n_length);
        T_res = smul(-1.0 * m, sdiv(angvel_ab, angvel_ab_length));
    }
    A->F = add(A->F, add(add(f_n, f_t), f_c));
    A->T = add(A->T, add(smul(-(r_a + delta_ab / 2.0), cross(n_ab, f_t)), T_res));
    A->p += f_n_length / (4.0 * PI_REF * r_a * r_a);
    st3(dt_slot, delta_t);
}

/* ------------------------------------------------------------------ */
/* A10: model 1 (contactLinearViscous + capillary cohesion)            */
/* ------------------------------------------------------------------ */

/* src/contactmodels.cuh:119-237 */
static void contact_linear_viscous(const orc *o, acc_t *A, uint32_t idx_a, uint32_t idx_b,
                                   double r_a, double r_b, v3 x_ab, double x_ab_length,
                                   double delta_ab, double kappa)
{
    const sph_state *h = &o->h;
    v3 vel_a = ld4(o->vel_s + 4 * (size_t)idx_a), vel_b = ld4(o->vel_s + 4 * (size_t)idx_b);
    v3 angvel_a = ld4(o->angvel_s + 4 * (size_t)idx_a), angvel_b = ld4(o->angvel_s + 4 * (size_t)idx_b);
    v3 n_ab = sdiv(x_ab, x_ab_length);
    v3 vel_ab_linear = sub(vel_a, vel_b);
    v3 vel_ab = add(add(vel_ab_linear, smul(r_a + delta_ab / 2.0, cross(n_ab, angvel_a))),
                    smul(r_b + delta_ab / 2.0, cross(n_ab, angvel_b)));
    double vel_n_ab = dot(vel_ab_linear, n_ab);
    v3 vel_t_ab = sub(vel_ab, smul(dot(vel_ab, n_ab), n_ab));
    double vel_t_ab_length = len(vel_t_ab);
    double k_n = h->k_n;
    if (h->E > .001) k_n = 3.141592654 / 2.0 * h->E * (r_a + r_b) / 2.;
    v3 f_n = smul(fmax(0.0, -k_n * delta_ab - h->gamma_n * vel_n_ab), n_ab);
    if (dot(f_n, n_ab) < 0.0) f_n = mk(0.0, 0.0, 0.0);
    double f_n_length = len(f_n);
    v3 f_c = smul(-kappa * sqrt(r_a * r_b), n_ab);
    v3 f_t = mk(0.0, 0.0, 0.0);
    if (vel_t_ab_length > 0.0) {
        double f_t_visc = h->gamma_t * vel_t_ab_length;
        double f_t_limit;
        if (vel_t_ab_length > (double)0.001f)
            f_t_limit = h->mu_d * len(sub(f_n, f_c));
        else
            f_t_limit = h->mu_s * len(sub(f_n, f_c));
        if (f_t_visc < f_t_limit)
            f_t = sdiv(smul(-f_t_visc, vel_t_ab), vel_t_ab_length);
        else
            f_t = sdiv(smul(-f_t_limit, vel_t_ab), vel_t_ab_length);
    }
    A->F = add(A->F, add(add(f_n, f_t), f_c));
    A->T = add(A->T, smul(-(r_a + delta_ab / 2.0), cross(n_ab, f_t)));
    A->p += f_n_length / (4.0 * PI_REF * r_a * r_a);
}

/* src/cohesion.cuh:237-277 */
static void capillary_cohesion_exp(const orc *o, acc_t *A, double r_a, double r_b,
                                   double delta_ab, v3 x_ab, double x_ab_length, double kappa)
{
    v3 n_ab = sdiv(x_ab, x_ab_length);
    double r;
    if ((r_a / r_b) > (r_b / r_a)) r = r_a / r_b;
    else r = r_b / r_a;
    double hh = -sqrt(r);
    double R_har = (2.0 * r_a * r_b) / (r_a + r_b);
    double R_geo = sqrt(r_a * r_b);
    double lambda = (double)0.9f * hh * sqrt(o->h.V_b / R_har);
    v3 f_c = smul(-kappa * R_geo * exp(-delta_ab / lambda), n_ab);
    A->F = add(A->F, f_c);
}

/* ------------------------------------------------------------------ */
/* A9: wall contact                                                    */
/* ------------------------------------------------------------------ */

/* src/contactmodels.cuh:10-114 */
static double contact_linear_wall(const orc *o, acc_t *A, uint32_t idx_a, double radius_a,
                                  v3 n, double delta, double wvel)
{
    const sph_state *h = &o->h;
    const double dt = h->time_dt;
    v3 vel_linear = ld4(o->vel_s + 4 * (size_t)idx_a);
    v3 angvel = ld4(o->angvel_s + 4 * (size_t)idx_a);
    v3 vel = add(vel_linear, smul(radius_a + delta / 2.0, cross(n, angvel)));
    vel = mk(vel.x + wvel, vel.y + wvel, vel.z + wvel);
    double vel_n = dot(vel_linear, n);
    v3 vel_t = sub(vel, smul(dot(n, vel), n));
    double vel_t_length = len(vel_t);
    double k_n = h->k_n;
    if (h->E > .001) k_n = 3.141592654 / 2.0 * h->E * radius_a;
    v3 f_n = smul(fmax(0.0, -k_n * delta - h->gamma_wn * vel_n), n);
    double f_n_length = len(f_n);
    A->ev_dot += h->gamma_wn * vel_n * vel_n;
    v3 f_t = mk(0.0, 0.0, 0.0);
    if (vel_t_length > 0.0 && h->gamma_wt > 0.0) {
        v3 f_t_visc = smul(h->gamma_wt, vel_t);
        double f_t_visc_length = len(f_t_visc);
        double f_t_limit;
        if (vel_t_length > 0.0) f_t_limit = h->mu_wd * f_n_length;
        else f_t_limit = h->mu_ws * f_n_length;
        if (f_t_visc_length < f_t_limit) {
            f_t = smul(-1.0, f_t_visc);
        } else {
            f_t = sdiv(smul(-f_t_limit, vel_t), vel_t_length);
            A->es_dot += len(smul(dt, smul(len(f_t), vel_t))) / dt;
        }
    }
    A->F = add(A->F, add(f_n, f_t));
    A->T = add(A->T, cross(smul(-(radius_a + delta * 0.5), n), f_t));
    A->p += f_n_length / (4.0 * PI_REF * radius_a * radius_a);
    return dot(f_n, n);
}

/* ------------------------------------------------------------------ */
/* A6: interact                                                        */
/* ------------------------------------------------------------------ */

/* src/contactsearch.cuh:385-726 */
static void interact_one(orc *o, uint32_t idx_a)
{
    const int NC = o->nc;
    const sph_state *h = &o->h;
    uint32_t np = o->np;
    uint32_t a_orig = o->index[idx_a];
    const double *xa4 = o->x_s + 4 * (size_t)idx_a;
    v3 x_a = ld4(xa4);
    double radius_a = xa4[3];
    v3 origo = mk(h->origo[0], h->origo[1], h->origo[2]);
    v3 L = mk(h->L[0], h->L[1], h->L[2]);
    double wnx[5][4] = {{0.0, 0.0, -1.0, L.z}, {-1.0, 0.0, 0.0, L.x}, {1.0, 0.0, 0.0, 0.0},
                        {0.0, -1.0, 0.0, L.y}, {0.0, 1.0, 0.0, 0.0}};
    double wmv[5][4] = {{0}};
    for (uint32_t w = 0; w < o->nw && w < 5; ++w)
        for (int c = 0; c < 4; ++c) { wnx[w][c] = h->w_nx[4 * w + c]; wmv[w][c] = h->w_mvfd[4 * w + c]; }
    acc_t A;
    A.F = mk(0.0, 0.0, 0.0); A.T = mk(0.0, 0.0, 0.0);
    A.es_dot = 0.0; A.ev_dot = 0.0; A.p = 0.0;
    if (h->contactmodel == 2 || h->contactmodel == 3) {
        v3 vel_a = ld4(o->vel_s + 4 * (size_t)idx_a);
        v3 angvel_a = ld4(o->angvel_s + 4 * (size_t)idx_a);
        for (int i = 0; i < NC; ++i) {
            size_t mempos = (size_t)a_orig * NC + i;
            uint32_t b = o->contacts[mempos];
            double *dts = o->delta_t + 3 * mempos;
            if (b != np) {
                const double *dm = o->distmod + 3 * mempos;
                const double *xb4 = o->x + 4 * (size_t)b;
                double radius_b = xb4[3];
                v3 x_ab = mk(x_a.x - xb4[0] + dm[0], x_a.y - xb4[1] + dm[1], x_a.z - xb4[2] + dm[2]);
                double l = len(x_ab);
                double delta_n = l - (radius_a + radius_b);
                if (delta_n < 0.0) {
                    if (h->contactmodel == 2)
                        contact_linear(o, &A, b, vel_a, angvel_a, radius_a, radius_b, x_ab, l, delta_n, dts);
                    else
                        contact_hertz(o, &A, b, vel_a, angvel_a, radius_a, radius_b, x_ab, l, delta_n, dts);
                } else {
                    o->contacts[mempos] = np;
                    dts[0] = 0.0; dts[1] = 0.0; dts[2] = 0.0;
                }
            } else {
                dts[0] = 0.0; dts[1] = 0.0; dts[2] = 0.0;
            }
        }
    } else if (h->contactmodel == 1) {
        int gp[3];
        grid_pos(o, xa4, gp);
        for (int zd = -1; zd < 2; ++zd)
            for (int yd = -1; yd < 2; ++yd)
                for (int xd = -1; xd < 2; ++xd) {
                    /* findAndProcessContactsInCell, contactsearch.cuh:91-186 */
                    int t[3] = {gp[0] + xd, gp[1] + yd, gp[2] + zd};
                    v3 dm = mk(0.0, 0.0, 0.0);
                    if (find_dist_mod(o, t, &dm) == -1) continue;
                    unsigned int cid = cell_of_target(o, t);
                    uint32_t start = o->cell_start[cid];
                    if (start == EMPTY) continue;
                    uint32_t end = o->cell_end[cid];
                    for (uint32_t idx_b = start; idx_b < end; ++idx_b) {
                        if (idx_b == idx_a) continue;
                        const double *xb4 = o->x_s + 4 * (size_t)idx_b;
                        double radius_b = xb4[3];
                        double kappa = h->kappa;
                        v3 x_ab = add(sub(x_a, ld4(xb4)), dm);
                        double l = len(x_ab);
                        double delta_ab = l - (radius_a + radius_b);
                        if (delta_ab < 0.0)
                            contact_linear_viscous(o, &A, idx_a, idx_b, radius_a, radius_b, x_ab, l, delta_ab, kappa);
                        else if (delta_ab < h->db)
                            capillary_cohesion_exp(o, &A, radius_a, radius_b, delta_ab, x_ab, l, kappa);
                    }
                }
    }
    /* walls, contactsearch.cuh:602-701 */
    double wf[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double delta_w;
    delta_w = wnx[0][3] - (x_a.z + radius_a);
    if (delta_w < 0.0) wf[0] = contact_linear_wall(o, &A, idx_a, radius_a, mk(wnx[0][0], wnx[0][1], wnx[0][2]), delta_w, wmv[0][1]);
    delta_w = x_a.z - radius_a - origo.z;
    if (delta_w < 0.0) (void)contact_linear_wall(o, &A, idx_a, radius_a, mk(0.0, 0.0, 1.0), delta_w, 0.0);
    if (h->periodic == 0) {
        delta_w = wnx[1][3] - (x_a.x + radius_a);
        if (delta_w < 0.0) wf[1] = contact_linear_wall(o, &A, idx_a, radius_a, mk(wnx[1][0], wnx[1][1], wnx[1][2]), delta_w, wmv[1][1]);
        delta_w = x_a.x - radius_a - wnx[2][3];
        if (delta_w < 0.0) wf[2] = contact_linear_wall(o, &A, idx_a, radius_a, mk(wnx[2][0], wnx[2][1], wnx[2][2]), delta_w, wmv[2][1]);
        delta_w = wnx[3][3] - (x_a.y + radius_a);
        if (delta_w < 0.0) wf[3] = contact_linear_wall(o, &A, idx_a, radius_a, mk(wnx[3][0], wnx[3][1], wnx[3][2]), delta_w, wmv[3][1]);
        delta_w = x_a.y - radius_a - wnx[4][3];
        if (delta_w < 0.0) wf[4] = contact_linear_wall(o, &A, idx_a, radius_a, mk(wnx[4][0], wnx[4][1], wnx[4][2]), delta_w, wmv[4][1]);
    } else if (h->periodic == 2) {
        delta_w = wnx[3][3] - (x_a.y + radius_a);
        if (delta_w < 0.0) wf[3] = contact_linear_wall(o, &A, idx_a, radius_a, mk(wnx[3][0], wnx[3][1], wnx[3][2]), delta_w, wmv[3][1]);
        delta_w = x_a.y - radius_a - wnx[4][3];
        if (delta_w < 0.0) wf[4] = contact_linear_wall(o, &A, idx_a, radius_a, mk(wnx[4][0], wnx[4][1], wnx[4][2]), delta_w, wmv[4][1]);
    }
    /* contactsearch.cuh:706-724 */
    st3(o->force + 3 * (size_t)a_orig, A.F);
    st3(o->torque + 3 * (size_t)a_orig, A.T);
    o->es_dot[a_orig] = A.es_dot;
    o->ev_dot[a_orig] = A.ev_dot;
    o->es[a_orig] += A.es_dot * h->time_dt;
    o->ev[a_orig] += A.ev_dot * h->time_dt;
    o->p[a_orig] = A.p;
    for (uint32_t w = 0; w < o->nw && w < 5; ++w) o->walls_force_pp[a_orig + (size_t)np * w] = wf[w];
}

/* ------------------------------------------------------------------ */
/* A11: bonds                                                          */
/* ------------------------------------------------------------------ */

/* src/cohesion.cuh:8-233, serialised in bond order (D3) */
static void bonds_linear(orc *o)
{
    const sph_state *h = &o->h;
    const double dt = h->time_dt;
    for (uint32_t idx = 0; idx < o->nb0; ++idx) {
        if (o->broken[idx]) continue;
        uint32_t bi = o->bonds[2 * (size_t)idx], bj = o->bonds[2 * (size_t)idx + 1];
        v3 delta0_t = ld3(o->bdt + 3 * (size_t)idx);
        double delta0_n = o->bdn[idx];
        v3 omega0_t = ld3(o->bot + 3 * (size_t)idx);
        double omega0_n = o->bon[idx];
        const double *xi = o->x + 4 * (size_t)bi, *xj = o->x + 4 * (size_t)bj;
        v3 vel_i = ld4(o->vel + 4 * (size_t)bi), vel_j = ld4(o->vel + 4 * (size_t)bj);
        v3 angvel_i = ld4(o->angvel4 + 4 * (size_t)bi), angvel_j = ld4(o->angvel4 + 4 * (size_t)bj);
        double R_bar = h->lambda_bar * fmin(xi[3], xj[3]);
        double A = PI_REF * R_bar * R_bar;
        double I = 0.25 * PI_REF * R_bar * R_bar * R_bar * R_bar;
        double J = I * 2.0;
        v3 x = mk(xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]);
        double x_length = len(x);
        double overlap = fmin(0.0, x_length - (xi[3] + xj[3]));
        v3 n = sdiv(x, x_length);
        delta0_t = sub(delta0_t, smul(dot(n, delta0_t), n));
        v3 ddelta = smul(dt, add(add(sub(vel_i, vel_j), smul(xi[3] + overlap / 2.0, cross(n, angvel_i))),
                                 smul(xj[3] + overlap / 2.0, cross(n, angvel_j))));
        double ddelta_n = -dot(ddelta, n);
        double delta_n = delta0_n + ddelta_n;
        v3 ddelta_t = sub(ddelta, smul(dot(n, ddelta), n));
        v3 delta_t = add(delta0_t, ddelta_t);
        v3 f_n = smul(h->k_n * A * delta_n + h->gamma_n * ddelta_n / dt, n);
        v3 f_t = sub(smul(-h->k_t * A, delta_t), sdiv(smul(h->gamma_t, ddelta_t), dt));
        v3 f = add(f_n, f_t);
        v3 mn = neg(n);
        omega0_t = sub(omega0_t, smul(dot(omega0_t, mn), mn));
        v3 domega = smul(dt, sub(angvel_j, angvel_i));
        double domega_n = dot(domega, mn);
        double omega_n = omega0_n + domega_n;
        v3 domega_t = sub(domega, smul(domega_n, mn));
        v3 omega_t = add(omega0_t, domega_t);
        v3 t_n = smul(-h->k_t * J * omega_n, mn);
        v3 t_t = smul(-h->k_n * I, omega_t);
        v3 t_j = add(t_n, t_t);
        double sigma_max = len(f_n) / A + len(t_t) * R_bar / I;
        double tau_max = len(f_t) / A + len(t_n) * R_bar / J;
        if (sigma_max >= h->sigma_b || tau_max >= h->tau_b) {
            o->broken[idx] = 1;
            continue;
        }
        st3(o->bdt + 3 * (size_t)idx, delta_t);
        o->bdn[idx] = delta_n;
        st3(o->bot + 3 * (size_t)idx, omega_t);
        o->bon[idx] = omega_n;
        double *fi = o->force + 3 * (size_t)bi, *fj = o->force + 3 * (size_t)bj;
        double *ti = o->torque + 3 * (size_t)bi, *tj = o->torque + 3 * (size_t)bj;
        fi[0] = fi[0] + f.x; fi[1] = fi[1] + f.y; fi[2] = fi[2] + f.z;
        fj[0] = fj[0] - f.x; fj[1] = fj[1] - f.y; fj[2] = fj[2] - f.z;
        ti[0] = ti[0] - t_j.x; ti[1] = ti[1] - t_j.y; ti[2] = ti[2] - t_j.z;
        tj[0] = tj[0] + t_j.x; tj[1] = tj[1] + t_j.y; tj[2] = tj[2] + t_j.z;
    }
}

/* ------------------------------------------------------------------ */
/* A12-A13: integrate, summation, integrateWalls                       */
/* ------------------------------------------------------------------ */

/* src/integration.cuh:309-341 with 256-thread blocks over np elements */
static void summation(const orc *o, const double *in, double *out)
{
    uint32_t np = o->np;
    double cache[256];
    for (uint32_t b = 0; b < o->nblocks; ++b) {
        for (uint32_t t = 0; t < 256; ++t) {
            uint32_t idx = b * 256 + t;
            double temp = 0.0;
            if (idx < np) temp += in[idx];
            cache[t] = temp;
        }
        for (uint32_t i = 128; i != 0; i /= 2)
            for (uint32_t t = 0; t < i; ++t) cache[t] += cache[t + i];
        out[b] = cache[0];
    }
}

/* src/integration.cuh:457-484 */
static void shear_stress_pp(orc *o)
{
    const sph_state *h = &o->h;
    for (uint32_t i = 0; i < o->np; ++i) {
        double z = o->x[4 * (size_t)i + 2], fixvel = o->vel[4 * (size_t)i + 3];
        double f_x = 0.0;
        if (fixvel > 0.0 && z > h->L[2] * 0.5) f_x = o->force[3 * (size_t)i];
        o->tau_pp[i] = f_x / (h->L[0] * h->L[1]);
    }
}

/* src/integration.cuh:14-305 (TY3).  Rows are in original order: the
 * reference reads the *_sorted copies, which hold the same values. */
static void integrate(orc *o, int change_velocity_state, double velocity_factor)
{
    const sph_state *h = &o->h;
    const double dt = h->time_dt;
    int wall0mode = 0; /* D5 */
    double wall0mass = 0.0;
    if (o->nw > 0) { wall0mode = h->wmode[0]; wall0mass = h->w_mvfd[0]; }
    double tau_eff_x = 0.0;
    if (o->nw > 0 && wall0mode == 3)
        for (uint32_t i = 0; i < o->nblocks; ++i) tau_eff_x += o->tau_partial[i];
    const double tau_x = h->tau_x;
    const long iter = o->iter;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long i = 0; i < (long)o->np; ++i) {
        const double *force = o->force + 3 * i, *torque = o->torque + 3 * i;
        const double *angpos = o->angpos + 3 * i;
        const double *x = o->x + 4 * i, *vel = o->vel + 4 * i, *angvel = o->angvel4 + 4 * i;
        double *xyzsum = o->xyzsum + 3 * i;
        double acc0[3] = {0.0, 0.0, 0.0}, angacc0[3] = {0.0, 0.0, 0.0};
        if (iter != 0) {
            for (int c = 0; c < 3; ++c) { acc0[c] = o->acc[3 * i + c]; angacc0[c] = o->angacc[3 * i + c]; }
        }
        double radius = x[3];
        double m = 4.0 / 3.0 * PI_REF * radius * radius * radius * h->rho;
        double acc[3], angacc[3];
        acc[0] = force[0] / m + h->g[0];
        acc[1] = force[1] / m + h->g[1];
        acc[2] = force[2] / m + h->g[2];
        angacc[0] = torque[0] * 1.0 / (2.0 / 5.0 * m * radius * radius);
        angacc[1] = torque[1] * 1.0 / (2.0 / 5.0 * m * radius * radius);
        angacc[2] = torque[2] * 1.0 / (2.0 / 5.0 * m * radius * radius);
        if (wall0mode == 3 && vel[3] > 0.0001 && x[2] > h->L[2] * 0.5) {
            double f_tau_x = (tau_x + tau_eff_x) * h->L[0] * h->L[1];
            acc[0] = f_tau_x / wall0mass;
            acc[1] = 0.0;
            acc[2] -= h->g[2];
            angacc[0] = angacc[1] = angacc[2] = 0.0;
        } else if (vel[3] > 0.0001 && vel[3] < 10.0) {
            acc[0] = 0.0;
            acc[1] = 0.0;
            acc[2] -= h->g[2];
            angacc[0] = angacc[1] = angacc[2] = 0.0;
        }
        if (vel[3] >= 10.0) {
            acc[1] = 0.0;
            angacc[0] = 0.0;
            angacc[2] = 0.0;
        }
        if (vel[3] < -0.0001) {
            acc[0] = acc[1] = acc[2] = 0.0;
            angacc[0] = angacc[1] = angacc[2] = 0.0;
        }
        double xn[4], vn[4], apn[3], avn[3];
        for (int c = 0; c < 3; ++c) {
            double dacc = (acc[c] - acc0[c]) / dt;
            double dang = (angacc[c] - angacc0[c]) / dt;
            xn[c] = x[c] + vel[c] * dt + 0.5 * acc[c] * dt * dt + 1.0 / 6.0 * dacc * dt * dt * dt;
            vn[c] = vel[c] + acc[c] * dt + 0.5 * dacc * dt * dt;
            apn[c] = angpos[c] + angvel[c] * dt + 0.5 * angacc[c] * dt * dt + 1.0 / 6.0 * dang * dt * dt * dt;
            avn[c] = angvel[c] + angacc[c] * dt + 0.5 * dang * dt * dt;
            xyzsum[c] += vel[c] * dt + 0.5 * acc[c] * dt * dt + 1.0 / 6.0 * dacc * dt * dt * dt;
        }
        xn[3] = x[3];
        vn[3] = vel[3];
        if (h->periodic == 1) {
            if (xn[0] < h->origo[0]) xn[0] += h->L[0];
            if (xn[0] > h->L[0]) xn[0] -= h->L[0];
            if (xn[1] < h->origo[1]) xn[1] += h->L[1];
            if (xn[1] > h->L[1]) xn[1] -= h->L[1];
        } else if (h->periodic == 2) {
            if (xn[0] < h->origo[0]) xn[0] += h->L[0];
            if (xn[0] > h->L[0]) xn[0] -= h->L[0];
        }
        if (vel[3] > 5.0 && vel[3] < 10.0) {
            if (change_velocity_state == 1) vn[0] *= velocity_factor;
            else if (change_velocity_state == -1) vn[0] /= velocity_factor;
        }
        for (int c = 0; c < 3; ++c) {
            o->acc[3 * i + c] = acc[c];
            o->angacc[3 * i + c] = angacc[c];
            o->angvel4[4 * i + c] = avn[c];
            o->angpos[3 * i + c] = apn[c];
        }
        memcpy(o->vel + 4 * i, vn, 32);
        memcpy(o->x + 4 * i, xn, 32);
    }
}

/* src/integration.cuh:344-450 */
static void integrate_walls(orc *o, double t_current)
{
    sph_state *h = &o->h;
    const double dt = h->time_dt;
    for (uint32_t idx = 0; idx < o->nw; ++idx) {
        double *w_nx = h->w_nx + 4 * idx, *w_mvfd = h->w_mvfd + 4 * idx;
        int wmode = h->wmode[idx];
        if (wmode == 0) continue;
        double acc0 = o->iter == 0 ? 0.0 : o->w_acc[idx];
        w_mvfd[2] = 0.0;
        for (uint32_t i = 0; i < o->nblocks; ++i) w_mvfd[2] += o->wall_partial[i];
        double sigma_0 = w_mvfd[3] + h->sigma0_A * sin(2.0 * PI_REF * h->sigma0_f * t_current);
        double N = -sigma_0 * h->L[0] * h->L[1];
        double acc = (w_mvfd[2] + N) / w_mvfd[0];
        if (wmode == 2) acc = 0.0;
        double dacc_dt = (acc - acc0) / dt;
        w_nx[3] += w_mvfd[1] * dt + 0.5 * acc * dt * dt + 1.0 / 6.0 * dacc_dt * dt * dt * dt;
        w_mvfd[1] += acc * dt + 0.5 * dacc_dt * dt * dt;
        o->w_acc[idx] = acc;
    }
}

/* ------------------------------------------------------------------ */
/* A14-A21: Darcy                                                      */
/* ------------------------------------------------------------------ */

/* setDarcyGhostNodes<T>, src/darcy.cuh:194-264 (ncomp = 1 or 3) */
static void set_ghosts(orc *o, double *f, int ncomp)
{
    const sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const double *val = f + (size_t)ncomp * didx(o, x, y, z);
#define GW(X, Y, Z) memcpy(f + (size_t)ncomp * didx(o, X, Y, Z), val, 8 * (size_t)ncomp)
                if (x == 0 && h->bc_xn == 0) GW(-1, y, z);
                if (x == 1 && h->bc_xn == 1) GW(-1, y, z);
                if (x == 0 && h->bc_xn == 2) GW(nx, y, z);
                if (x == nx - 1 && h->bc_xp == 0) GW(nx, y, z);
                if (x == nx - 2 && h->bc_xp == 1) GW(nx, y, z);
                if (x == nx - 1 && h->bc_xp == 2) GW(-1, y, z);
                if (y == 0 && h->bc_yn == 0) GW(x, -1, z);
                if (y == 1 && h->bc_yn == 1) GW(x, -1, z);
                if (y == 0 && h->bc_yn == 2) GW(x, ny, z);
                if (y == ny - 1 && h->bc_yp == 0) GW(x, ny, z);
                if (y == ny - 2 && h->bc_yp == 1) GW(x, ny, z);
                if (y == ny - 1 && h->bc_yp == 2) GW(x, -1, z);
                if (z == 0 && h->bc_bot == 0) GW(x, y, -1);
                if (z == 1 && h->bc_bot == 1) GW(x, y, -1);
                if (z == 0 && h->bc_bot == 2) GW(x, y, nz);
                if (z == nz - 1 && h->bc_top == 0) GW(x, y, nz);
                if (z == nz - 2 && h->bc_top == 1) GW(x, y, nz);
                if (z == nz - 1 && h->bc_top == 2) GW(x, y, -1);
#undef GW
            }
}

/* setDarcyGhostNodesFlux, src/darcy.cuh:270-312: ghosts from the interior of
 * src, written into dst (the same array in the reference) */
static void set_ghosts_flux(orc *o, double *f)
{
    const sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    if (!(h->bc_bot == 4 || h->bc_top == 4)) return;
    const double dz = h->L[2] / nz;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                double pv = f[didx(o, x, y, z)];
                double k = o->d_k[didx(o, x, y, z)];
                double q_z = 0.;
                if (z == 0) q_z = h->bc_bot_flux;
                else if (z == nz - 1) q_z = h->bc_top_flux;
                double p_ghost = -h->mu / k * q_z * dz + pv;
                if (z == 0 && h->bc_bot == 4) f[didx(o, x, y, -1)] = p_ghost;
                if (z == nz - 1 && h->bc_top == 4) f[didx(o, x, y, nz)] = p_ghost;
            }
}

/* src/darcy.cuh:889-943 */
static void pressure_gradient(orc *o)
{
    const sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    const double dx = h->L[0] / nx, dy = h->L[1] / ny, dz = h->L[2] / nz;
#pragma omp parallel for num_threads(g_threads) collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const double *p = o->d_p;
                double p_xn = p[didx(o, x - 1, y, z)], p_xp = p[didx(o, x + 1, y, z)];
                double p_yn = p[didx(o, x, y - 1, z)], p_yp = p[didx(o, x, y + 1, z)];
                double p_zn = p[didx(o, x, y, z - 1)], p_zp = p[didx(o, x, y, z + 1)];
                double *g = o->d_grad_p + 3 * (size_t)didx(o, x, y, z);
                g[0] = (p_xp - p_xn) / (dx + dx);
                g[1] = (p_yp - p_yn) / (dy + dy);
                g[2] = (p_zp - p_zn) / (dz + dz);
            }
}

/* weight(), src/darcy.cuh:324-337 (x_f = X + n) */
static inline double weight(v3 x_p, v3 x_f, double dx, double dy, double dz)
{
    v3 d = sub(x_p, x_f);
    d = mk(fabs(d.x), fabs(d.y), fabs(d.z));
    if (d.x < dx && d.y < dy && d.z < dz)
        return (1.0 - d.x / dx) * (1.0 - d.y / dy) * (1.0 - d.z / dz);
    return 0.0;
}

/* sphereVolume, src/darcy.cuh:315-318 (D6) */
static inline double sphere_volume(double r) { return 4.0 / 3.0 * PI_M * (r * r * r); }

/* findDarcyPressureForceLinear, src/darcy.cuh:948-1102 */
static void pressure_force(orc *o, unsigned int wall0_iz)
{
    const sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    const double dx = h->L[0] / nx, dy = h->L[1] / ny, dz = h->L[2] / nz;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long i = 0; i < (long)o->np; ++i) {
        const double *x4 = o->x + 4 * i;
        v3 x3 = ld4(x4);
        unsigned int i_x = (unsigned int)floor((x4[0] - h->origo[0]) / (h->L[0] / h->num[0]));
        unsigned int i_y = (unsigned int)floor((x4[1] - h->origo[1]) / (h->L[1] / h->num[1]));
        unsigned int i_z = (unsigned int)floor((x4[2] - h->origo[2]) / (h->L[2] / h->num[2]));
        v3 X = mk(i_x * dx + 0.5 * dx, i_y * dy + 0.5 * dy, i_z * dz + 0.5 * dz);
        v3 grad_p = mk(0., 0., 0.);
        for (int d_iz = -1; d_iz < 2; d_iz++)
            for (int d_iy = -1; d_iy < 2; d_iy++)
                for (int d_ix = -1; d_ix < 2; d_ix++) {
                    int ix_n = (int)i_x + d_ix, iy_n = (int)i_y + d_iy, iz_n = (int)i_z + d_iz;
                    v3 gpi = ld3(o->d_grad_p + 3 * (size_t)didx(o, ix_n, iy_n, iz_n));
                    if ((ix_n == -1 && iy_n == -1) || (ix_n == -1 && iz_n == -1) ||
                        (iy_n == -1 && iz_n == -1) || (ix_n == nx && iy_n == ny) ||
                        (ix_n == nx && iz_n == nz) || (iy_n == ny && iz_n == nz) ||
                        (ix_n == nx && iy_n == -1) || (ix_n == nx && iz_n == -1) ||
                        (iy_n == ny && ix_n == -1) || (iy_n == ny && iz_n == -1) ||
                        (iz_n == nz && ix_n == -1) || (iz_n == nz && iy_n == -1))
                        gpi = mk(0., 0., 0.);
                    /* unsigned comparison, darcy.cuh:1033 */
                    if (h->bc_top == 0 && i_z + (unsigned int)d_iz >= wall0_iz - 1u) gpi.z = 0.0;
                    v3 nn = mk(dx * d_ix, dy * d_iy, dz * d_iz);
                    grad_p = add(grad_p, smul(weight(x3, add(X, nn), dx, dy, dz), gpi));
                }
        double v = sphere_volume(x4[3]);
        v3 f_p = sub(smul(v, smul(-1.0, grad_p)), smul(h->rho_f * v, mk(h->g[0], h->g[1], h->g[2])));
        if (h->bc_top == 0 && i_z >= wall0_iz) f_p.z = 0.0;
        double *F = o->force + 3 * i;
        F[0] = F[0] + f_p.x; F[1] = F[1] + f_p.y; F[2] = F[2] + f_p.z;
        st3(o->f_p + 3 * i, f_p);
    }
}

/* findDarcyPorositiesLinear, src/darcy.cuh:359-550 */
static void porosities(orc *o)
{
    const sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    const double dx = h->L[0] / nx, dy = h->L[1] / ny, dz = h->L[2] / nz;
    const double dt = h->time_dt;
    const long iteration = o->iter;
#pragma omp parallel for num_threads(g_threads) collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                unsigned int cellidx = didx(o, x, y, z);
                if (o->np == 0) {
                    o->d_phi[cellidx] = 0.9;
                    o->d_dphi[cellidx] = 0.0;
                    continue;
                }
                double solid_volume = 0.0, solid_volume_new = 0.0;
                v3 X = mk(x * dx + 0.5 * dx, y * dy + 0.5 * dy, z * dz + 0.5 * dz);
                v3 vp_avg_num = mk(0.0, 0.0, 0.0);
                double vp_avg_denum = 0.0;
                double phi_0 = o->d_phi[cellidx];
                for (int zd = -1; zd < 2; ++zd)
                    for (int yd = -1; yd < 2; ++yd)
                        for (int xd = -1; xd < 2; ++xd) {
                            int t[3] = {x + xd, y + yd, z + zd};
                            if (t[2] == -1) t[2] = 1;
                            if (h->periodic == 2) {
                                if (t[1] == -1) t[1] = 1;
                                if (t[1] == ny) t[1] = ny - 2;
                            }
                            v3 dm = mk(0.0, 0.0, 0.0);
                            if (find_dist_mod(o, t, &dm) == -1) continue;
                            unsigned int cid = cell_of_target(o, t);
                            uint32_t start = o->cell_start[cid];
                            if (start == EMPTY) continue;
                            uint32_t end = o->cell_end[cid];
                            for (uint32_t i = start; i < end; ++i) {
                                const double *xr = o->x_s + 4 * (size_t)i;
                                const double *v = o->vel_s + 4 * (size_t)i;
                                v3 dist = add(sub(X, ld4(xr)), dm);
                                double s = 0.0;
                                {
                                    v3 ad = mk(fabs(dist.x), fabs(dist.y), fabs(dist.z));
                                    if (ad.x < dx && ad.y < dy && ad.z < dz)
                                        s = (1.0 - ad.x / dx) * (1.0 - ad.y / dy) * (1.0 - ad.z / dz);
                                }
                                double vol_p = sphere_volume(xr[3]);
                                solid_volume += s * vol_p;
                                vp_avg_num = add(vp_avg_num, smul(s * vol_p, ld4(v)));
                                vp_avg_denum += s * vol_p;
                                v3 xrn = mk(xr[0] + v[0] * dt, xr[1] + v[1] * dt, xr[2] + v[2] * dt);
                                v3 dist2 = add(sub(X, xrn), dm);
                                double s2 = 0.0;
                                {
                                    v3 ad = mk(fabs(dist2.x), fabs(dist2.y), fabs(dist2.z));
                                    if (ad.x < dx && ad.y < dy && ad.z < dz)
                                        s2 = (1.0 - ad.x / dx) * (1.0 - ad.y / dy) * (1.0 - ad.z / dz);
                                }
                                solid_volume_new += s2 * vol_p;
                            }
                        }
                double cell_volume = dx * dy * dz;
                if (z == nz - 1) cell_volume *= 0.875;
                double phi = fmin(0.9, fmax(0.1, 1.0 - solid_volume / cell_volume));
                double phi_new = fmin(0.9, fmax(0.1, 1.0 - solid_volume_new / cell_volume));
                double dphi;
                v3 vp_avg;
                if (iteration == 0) {
                    dphi = 0.0;
                    vp_avg = mk(0.0, 0.0, 0.0);
                } else {
                    dphi = 0.5 * (phi_new - phi_0);
                    vp_avg = sdiv(vp_avg_num, fmax(1.0e-16, vp_avg_denum));
                }
                o->d_phi[cellidx] = phi * h->c_phi;
                o->d_dphi[cellidx] = dphi * h->c_phi;
                st3(o->d_vp_avg + 3 * (size_t)cellidx, vp_avg);
            }
}

/* copyDarcyPorositiesToEdges / ToBottom, src/darcy.cuh:556-628 */
static void porosities_to_edges(orc *o)
{
    const sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int side = 0; side < 2; ++side) {
                int y = side == 0 ? 0 : ny - 1;
                int y_read = side == 0 ? 1 : ny - 2;
                unsigned int r = didx(o, x, y_read, z), w = didx(o, x, y, z);
                o->d_phi[w] = o->d_phi[r];
                o->d_dphi[w] = o->d_dphi[r];
                memcpy(o->d_vp_avg + 3 * (size_t)w, o->d_vp_avg + 3 * (size_t)r, 24);
            }
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            unsigned int r = didx(o, x, y, 1), w = didx(o, x, y, 0);
            o->d_phi[w] = o->d_phi[r];
            o->d_dphi[w] = o->d_dphi[r];
            memcpy(o->d_vp_avg + 3 * (size_t)w, o->d_vp_avg + 3 * (size_t)r, 24);
        }
}

/* findDarcyPermeabilities, src/darcy.cuh:1266-1313 (D6) */
static void permeabilities(orc *o)
{
    const sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                unsigned int c = didx(o, x, y, z);
                double phi = o->d_phi[c];
                if (phi > 0.9999) phi = 0.9999;
                double k = h->k_c * (phi * phi * phi) / ((1.0 - phi) * (1.0 - phi));
                k = fmin(2.7e-10, k);
                o->d_k[c] = k;
            }
}

/* firstDarcySolution / updateDarcySolution, src/darcy.cuh:1425-1799.
 * first != 0: writes dp_expl; otherwise writes p_new and norm. */
static void darcy_sweep(orc *o, int first, unsigned int wall0_iz)
{
    const sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    const double dx = h->L[0] / nx, dy = h->L[1] / ny, dz = h->L[2] / nz;
    const double dt = h->time_dt;
    const unsigned int ndem = h->ndem;
    const double beta_f = h->beta_f, mu = h->mu;
#pragma omp parallel for num_threads(g_threads) collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                unsigned int c = didx(o, x, y, z);
                const double *phiA = o->d_phi, *pA = o->d_p, *kA = o->d_k;
                double phi_xn = phiA[didx(o, x - 1, y, z)], phi = phiA[c], phi_xp = phiA[didx(o, x + 1, y, z)];
                double phi_yn = phiA[didx(o, x, y - 1, z)], phi_yp = phiA[didx(o, x, y + 1, z)];
                double phi_zn = phiA[didx(o, x, y, z - 1)], phi_zp = phiA[didx(o, x, y, z + 1)];
                double dphi = o->d_dphi[c];
                v3 vp_avg = ld3(o->d_vp_avg + 3 * (size_t)c);
                int p_constant = o->d_pconst[c];
                double p_xn = pA[didx(o, x - 1, y, z)], p = pA[c], p_xp = pA[didx(o, x + 1, y, z)];
                double p_yn = pA[didx(o, x, y - 1, z)], p_yp = pA[didx(o, x, y + 1, z)];
                double p_zn = pA[didx(o, x, y, z - 1)], p_zp = pA[didx(o, x, y, z + 1)];
                double k_xn = kA[didx(o, x - 1, y, z)], k = kA[c], k_xp = kA[didx(o, x + 1, y, z)];
                double k_yn = kA[didx(o, x, y - 1, z)], k_yp = kA[didx(o, x, y + 1, z)];
                double k_zn = kA[didx(o, x, y, z - 1)], k_zp = kA[didx(o, x, y, z + 1)];
                if (x == 0 && h->bc_xn == 1) p_xn = p;
                if (x == nx - 1 && h->bc_xp == 1) p_xp = p;
                if (y == 0 && h->bc_yn == 1) p_yn = p;
                if (y == ny - 1 && h->bc_yp == 1) p_yp = p;
                if (z == 0 && h->bc_bot == 1) p_zn = p;
                if (z == nz - 1 && h->bc_top == 1) p_zp = p;
                v3 grad_phi = mk((phi_xp - phi_xn) / (dx + dx), (phi_yp - phi_yn) / (dy + dy),
                                 (phi_zp - phi_zn) / (dz + dz));
                double div_k_grad_p =
                    (2. * k_xp * k / (k_xp + k) * (p_xp - p) / dx - 2. * k_xn * k / (k_xn + k) * (p - p_xn) / dx) / dx +
                    (2. * k_yp * k / (k_yp + k) * (p_yp - p) / dy - 2. * k_yn * k / (k_yn + k) * (p - p_yn) / dy) / dy +
                    (2. * k_zp * k / (k_zp + k) * (p_zp - p) / dz - 2. * k_zn * k / (k_zn + k) * (p - p_zn) / dz) / dz;
                double dp = ndem * dt / (beta_f * phi * mu) * div_k_grad_p -
                            (ndem * dt / (beta_f * phi * (1.0 - phi))) * (dphi / (ndem * dt) + dot(vp_avg, grad_phi));
                /* src/darcy.cuh:1575-1580 (y == nx-1 as in the reference) */
                if ((h->bc_bot == 0 && z == 0) || (h->bc_top == 0 && z == nz - 1) ||
                    ((unsigned)z >= wall0_iz && h->bc_top == 0) || (h->bc_xn == 0 && x == 0) ||
                    (h->bc_xp == 0 && x == nx - 1) || (h->bc_yn == 0 && y == 0) ||
                    (h->bc_yp == 0 && y == nx - 1) || p_constant == 1)
                    dp = 0.0;
                if (first) {
                    o->d_dp_expl[c] = dp;
                } else {
                    const double epsilon = 0.5;
                    double p_new = o->d_p_old[c] + (1.0 - epsilon) * o->d_dp_expl[c] + epsilon * dp;
                    const double theta = 0.05;
                    p_new = p * (1.0 - theta) + p_new * theta;
                    double res_norm = (p_new - p) / (p + 1.0e-16);
                    o->d_p_new[c] = p_new;
                    o->d_norm[c] = res_norm;
                }
            }
}

/* findDarcyVelocities, src/darcy.cuh:1839-1897 */
static void darcy_velocities(orc *o)
{
    const sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    const double dx = h->L[0] / nx, dy = h->L[1] / ny, dz = h->L[2] / nz;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                unsigned int c = didx(o, x, y, z);
                const double *p = o->d_p;
                double p_xn = p[didx(o, x - 1, y, z)], p_xp = p[didx(o, x + 1, y, z)];
                double p_yn = p[didx(o, x, y - 1, z)], p_yp = p[didx(o, x, y + 1, z)];
                double p_zn = p[didx(o, x, y, z - 1)], p_zp = p[didx(o, x, y, z + 1)];
                double k = o->d_k[c], phi = o->d_phi[c];
                v3 g = mk((p_xp - p_xn) / (dx + dx), (p_yp - p_yn) / (dy + dy), (p_zp - p_zn) / (dz + dz));
                v3 v = sdiv(smul(-k / h->mu, g), phi);
                st3(o->d_v + 3 * (size_t)c, v);
            }
}

static void copy_interior(orc *o, const double *src, double *dst)
{
    const sph_state *h = &o->h;
    for (uint32_t z = 0; z < h->num[2]; ++z)
        for (uint32_t y = 0; y < h->num[1]; ++y)
            for (uint32_t x = 0; x < h->num[0]; ++x) {
                unsigned int c = didx(o, x, y, z);
                dst[c] = src[c];
            }
}

static void conv_push(orc *o, long iterp1, unsigned int nijac)
{
    if (o->nconv == o->cap_conv) {
        o->cap_conv *= 2;
        o->conv = (uint32_t *)realloc(o->conv, (size_t)o->cap_conv * 8);
    }
    o->conv[2 * o->nconv] = (uint32_t)iterp1;
    o->conv[2 * o->nconv + 1] = nijac;
    o->nconv++;
}

/* Darcy branch of the time loop, src/device.cu:1901-2426 */
static int darcy_step(orc *o)
{
    sph_state *h = &o->h;
    int nx = (int)h->num[0], ny = (int)h->num[1], nz = (int)h->num[2];
    if (o->nw > 0 && (h->wmode[0] == 1 || h->wmode[0] == 3))
        o->wall0_iz = (unsigned int)(o->host_wx0 / (h->L[2] / h->num[2]));
    unsigned int wall0_iz = o->wall0_iz;
    if (o->np > 0) {
        set_ghosts(o, o->d_p, 1);
        pressure_gradient(o);
        set_ghosts(o, o->d_grad_p, 3);
        pressure_force(o, wall0_iz);
    }
    if ((o->iter % h->ndem) == 0) {
        porosities(o);
        if (h->periodic == 2) porosities_to_edges(o);
        if ((h->p_mod_A > 1.0e-5 || h->p_mod_A < -1.0e-5) && h->p_mod_f > 1.0e-7) {
            double new_pressure = o->p_top_orig + h->p_mod_A * sin(2.0 * PI_M * h->p_mod_f * h->time_current + h->p_mod_phi);
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    o->d_p[didx(o, x, y, nz - 1)] = new_pressure;                       /* D8 */
                    if (wall0_iz < (unsigned)nz) o->d_p[didx(o, x, y, (int)wall0_iz)] = new_pressure;
                }
        }
        permeabilities(o);
        set_ghosts(o, o->d_phi, 1);
        set_ghosts(o, o->d_k, 1);
        if (o->iter == 0) copy_interior(o, o->d_p, o->d_p_old);
        if (h->bc_bot == 4 || h->bc_top == 4) set_ghosts_flux(o, o->d_p);
        o->stats.fluid_steps++;
        struct timespec t_solve0;
        clock_gettime(CLOCK_MONOTONIC, &t_solve0);
        for (unsigned int nijac = 0; nijac < h->maxiter; ++nijac) {
            if (nijac == 0) copy_interior(o, o->d_p, o->d_p_old);
            set_ghosts(o, o->d_p, 1);
            if (nijac == 0) darcy_sweep(o, 1, wall0_iz);
            darcy_sweep(o, 0, wall0_iz);
            /* setDarcyTopWallFixedFlow writes dev_darcy_p, which the copy below
             * overwrites: no observable effect (src/device.cu:2284-2297) */
            if (h->bc_bot == 4 || h->bc_top == 4) set_ghosts_flux(o, o->d_p);
            copy_interior(o, o->d_p_new, o->d_p);
            o->stats.sweeps++;
            if (nijac % 10 == 0) {
                /* maxNormResDarcy, src/darcy.cpp:266-292 */
                double max_norm_res = -1.0e9;
                for (int z = 0; z < nz; ++z)
                    for (int y = 0; y < ny; ++y)
                        for (int x = 0; x < nx; ++x) {
                            double r = fabs(o->d_norm[didx(o, x, y, z)]);
                            if (r != r) {
                                snprintf(o->err, sizeof(o->err), "normalized residual is NaN in cell %d,%d,%d", x, y, z);
                                return SPH_ERR_NAN;
                            }
                            if (r > max_norm_res) max_norm_res = r;
                        }
                o->stats.last_nijac = nijac;
                o->stats.last_max_residual = max_norm_res;
                if (max_norm_res <= h->tolerance) {
                    if (o->iter % 10 == 0) conv_push(o, o->iter + 1, nijac);
                    break;
                }
            }
            if (nijac == h->maxiter - 1) {
                conv_push(o, o->iter + 1, nijac);
                o->stats.nonconverged++;
            }
        }
        {
            /* wall time of the Jacobi loop (the solve time the GPU reports) */
            struct timespec t_solve1;
            clock_gettime(CLOCK_MONOTONIC, &t_solve1);
            o->stats.solve_ms += 1e3 * (double)(t_solve1.tv_sec - t_solve0.tv_sec) +
                                 1e-6 * (double)(t_solve1.tv_nsec - t_solve0.tv_nsec);
        }
        if (o->filetimeclock + h->time_dt < h->time_file_dt) {
            for (int z = 0; z < nz; ++z)
                for (int y = 0; y < ny; ++y)
                    for (int x = 0; x < nx; ++x) o->d_dphi[didx(o, x, y, z)] = 0.0;
        }
        set_ghosts(o, o->d_p, 1);
        darcy_velocities(o);
    }
    return SPH_OK;
}

/* ------------------------------------------------------------------ */
/* one loop body of DEM::startTime, src/device.cu:1038-2680            */
/* ------------------------------------------------------------------ */

static int step_once(orc *o)
{
    sph_state *h = &o->h;
    int rc;
    if (o->np > 0) {
        if (h->adaptive == 1 && o->nw > 0) h->L[2] = h->w_nx[3]; /* updateGridSize */
        rc = hash_sort_reorder(o);
        if (rc) return rc;
        long ovf = 0;
        if (h->contactmodel == 2 || h->contactmodel == 3) {
#pragma omp parallel for num_threads(g_threads) reduction(+ : ovf) schedule(dynamic, 256)
            for (long i = 0; i < (long)o->np; ++i) ovf += topology_one(o, (uint32_t)i);
        }
        o->overflow += ovf;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 256)
        for (long i = 0; i < (long)o->np; ++i) interact_one(o, (uint32_t)i);
        if (o->nb0 > 0) bonds_linear(o);
    }
    if (h->fluid) {
        rc = darcy_step(o);
        if (rc) return rc;
    }
    if (o->ns && o->np > 0) {
        const int wall = o->nw > 0 && (h->wmode[0] == 1 || h->wmode[0] == 3);
        rc = nso_step_coupled(o->ns, o->np, o->x, o->vel, o->x_s, o->cell_start, o->cell_end, o->index,
                              h->origo, h->periodic, o->force, wall ? o->host_wx0 : -1.0);
        if (rc) return rc;
    }
    if (o->np > 0) {
        if (o->nw > 0 && h->wmode[0] == 3) {
            shear_stress_pp(o);
            summation(o, o->tau_pp, o->tau_partial);
        }
        int change_velocity_state = 0;
        if (h->time_current >= 10.0 && h->time_current < 15.0 && o->velocity_state == 1) {
            change_velocity_state = 1;
            o->velocity_state = 2;
        } else if (h->time_current >= 15.0 && o->velocity_state == 2) {
            change_velocity_state = -1;
            o->velocity_state = 1;
        }
        integrate(o, change_velocity_state, 10.0);
        if (o->nw > 0) {
            summation(o, o->walls_force_pp, o->wall_partial);
            integrate_walls(o, h->time_current);
        }
    }
    h->time_current += h->time_dt;
    o->filetimeclock += h->time_dt;
    o->iter++;
    if (o->filetimeclock >= h->time_file_dt) {
        /* CS4: host copies refresh, dphi zeroed after the transfer, file
         * number advanced (src/device.cu:2576-2680) */
        o->host_wx0 = o->nw > 0 ? h->w_nx[3] : 0.0;
        if (o->ns) nso_output_event(o->ns);
        if (h->fluid) {
            for (uint32_t z = 0; z < h->num[2]; ++z)
                for (uint32_t y = 0; y < h->num[1]; ++y)
                    for (uint32_t x = 0; x < h->num[0]; ++x) o->d_dphi[didx(o, x, y, z)] = 0.0;
        }
        h->time_step_count += 1;
        o->filetimeclock = 0.0;
    }
    if (o->overflow) {
        snprintf(o->err, sizeof(o->err), "more than %d contacts on a particle (%ld overflows)", o->nc, o->overflow);
        return SPH_ERR_OVERFLOW;
    }
    return SPH_OK;
}

int orc_step(orc *o, int n)
{
    for (int i = 0; i < n; ++i) {
        int rc = step_once(o);
        if (rc) return rc;
    }
    return SPH_OK;
}

/* ------------------------------------------------------------------ */
/* read-out                                                            */
/* ------------------------------------------------------------------ */

int orc_get_header(orc *o, sph_state *hdr)
{
    *hdr = o->h;
    hdr->x = hdr->xyzsum = hdr->vel = hdr->force = hdr->angpos = hdr->angvel = hdr->torque = NULL;
    hdr->es_dot = hdr->es = hdr->ev_dot = hdr->ev = hdr->p = NULL;
    hdr->bonds = NULL;
    hdr->bonds_delta_n = hdr->bonds_delta_t = hdr->bonds_omega_n = hdr->bonds_omega_t = NULL;
    hdr->v_f = hdr->p_f = hdr->phi = hdr->dphi = hdr->f_p = NULL;
    hdr->p_constant = NULL;
    hdr->color = NULL;
    return SPH_OK;
}

#define CP(dst, src, bytes) do { if (dst) memcpy(dst, src, bytes); } while (0)

int orc_read_state(orc *o, sph_state *out)
{
    uint32_t np = o->np, nb = o->nb0;
    sph_state keep = *out;
    orc_get_header(o, out);
    out->x = keep.x; out->xyzsum = keep.xyzsum; out->vel = keep.vel; out->force = keep.force;
    out->angpos = keep.angpos; out->angvel = keep.angvel; out->torque = keep.torque;
    out->es_dot = keep.es_dot; out->es = keep.es; out->ev_dot = keep.ev_dot; out->ev = keep.ev;
    out->p = keep.p; out->bonds = keep.bonds; out->bonds_delta_n = keep.bonds_delta_n;
    out->bonds_delta_t = keep.bonds_delta_t; out->bonds_omega_n = keep.bonds_omega_n;
    out->bonds_omega_t = keep.bonds_omega_t; out->v_f = keep.v_f; out->p_f = keep.p_f;
    out->phi = keep.phi; out->dphi = keep.dphi; out->f_p = keep.f_p;
    out->p_constant = keep.p_constant; out->color = keep.color;
    CP(out->x, o->x, (size_t)np * 32);
    CP(out->xyzsum, o->xyzsum, (size_t)np * 24);
    CP(out->vel, o->vel, (size_t)np * 32);
    CP(out->force, o->force, (size_t)np * 24);
    CP(out->angpos, o->angpos, (size_t)np * 24);
    if (out->angvel)
        for (uint32_t i = 0; i < np; ++i)
            for (int c = 0; c < 3; ++c) out->angvel[3 * (size_t)i + c] = o->angvel4[4 * (size_t)i + c];
    CP(out->torque, o->torque, (size_t)np * 24);
    CP(out->es_dot, o->es_dot, (size_t)np * 8);
    CP(out->es, o->es, (size_t)np * 8);
    CP(out->ev_dot, o->ev_dot, (size_t)np * 8);
    CP(out->ev, o->ev, (size_t)np * 8);
    CP(out->p, o->p, (size_t)np * 8);
    CP(out->color, o->color, (size_t)np * 4);
    if (out->bonds)
        for (uint32_t b = 0; b < nb; ++b) {
            out->bonds[2 * (size_t)b] = o->broken[b] ? nb : o->bonds[2 * (size_t)b];
            out->bonds[2 * (size_t)b + 1] = o->bonds[2 * (size_t)b + 1];
        }
    CP(out->bonds_delta_n, o->bdn, (size_t)nb * 8);
    CP(out->bonds_delta_t, o->bdt, (size_t)nb * 24);
    CP(out->bonds_omega_n, o->bon, (size_t)nb * 8);
    CP(out->bonds_omega_t, o->bot, (size_t)nb * 24);
    if (o->h.fluid) {
        uint32_t nx = o->h.num[0], ny = o->h.num[1], nz = o->h.num[2];
        for (uint32_t z = 0; z < nz; ++z)
            for (uint32_t y = 0; y < ny; ++y)
                for (uint32_t x = 0; x < nx; ++x) {
                    size_t f = x + (size_t)nx * (y + (size_t)ny * z);
                    unsigned int c = didx(o, x, y, z);
                    if (out->v_f) memcpy(out->v_f + 3 * f, o->d_v + 3 * (size_t)c, 24);
                    if (out->p_f) out->p_f[f] = o->d_p[c];
                    if (out->phi) out->phi[f] = o->d_phi[c];
                    if (out->dphi) out->dphi[f] = o->d_dphi[c];
                    if (out->p_constant) out->p_constant[f] = o->d_pconst[c];
                }
        CP(out->f_p, o->f_p, (size_t)np * 24);
    }
    return SPH_OK;
}

int orc_read_sort(orc *o, uint32_t *cellid, uint32_t *index, uint32_t *cs, uint32_t *ce)
{
    CP(cellid, o->cellid, (size_t)o->np * 4);
    CP(index, o->index, (size_t)o->np * 4);
    CP(cs, o->cell_start, (size_t)o->ncells * 4);
    CP(ce, o->cell_end, (size_t)o->ncells * 4);
    return SPH_OK;
}

int orc_read_contacts(orc *o, uint32_t *contacts, double *distmod, double *delta_t)
{
    size_t n = (size_t)o->np * o->nc;
    CP(contacts, o->contacts, n * 4);
    CP(distmod, o->distmod, n * 24);
    CP(delta_t, o->delta_t, n * 24);
    return SPH_OK;
}

int orc_write_contacts(orc *o, const uint32_t *contacts, const double *distmod, const double *delta_t)
{
    size_t n = (size_t)o->np * o->nc;
    if (contacts) memcpy(o->contacts, contacts, n * 4);
    if (distmod) memcpy(o->distmod, distmod, n * 24);
    if (delta_t) memcpy(o->delta_t, delta_t, n * 24);
    return SPH_OK;
}

int orc_darcy_stats(orc *o, sph_darcy_stats *out)
{
    *out = o->stats;
    return SPH_OK;
}

int orc_conv_log(orc *o, uint32_t *pairs, int max_pairs)
{
    int n = o->nconv < max_pairs ? o->nconv : max_pairs;
    if (pairs && n > 0) memcpy(pairs, o->conv, (size_t)n * 8);
    return o->nconv;
}

int orc_read_darcy_field(orc *o, int which, double *out)
{
    if (!o->h.fluid) return SPH_ERR_STATE;
    const double *src = NULL;
    switch (which) {
    case 0: src = o->d_p; break;
    case 1: src = o->d_phi; break;
    case 2: src = o->d_dphi; break;
    case 3: src = o->d_k; break;
    case 4: src = o->d_norm; break;
    case 5: src = o->d_dp_expl; break;
    case 6: src = o->d_p_old; break;
    default: return SPH_ERR_ARG;
    }
    memcpy(out, src, (size_t)o->dcells * 8);
    return SPH_OK;
}

/* Couples a Navier-Stokes fluid of ns_oracle.c to the oracle (the
 * reference's fluid block between the bond forces and the integration,
 * src/device.cu:1213-1898); NULL detaches. */
void orc_attach_ns(orc *o, void *ns) { o->ns = ns; }
