/*
 * ns_oracle.c -- test-only CPU restatement of the reference's Navier-Stokes
 * fluid step (cfd_solver = 0) without grains, the checker of the sph_ns_*
 * entry points.  Not part of the product.
 *
 * Each function below restates one reference kernel as a plain loop over
 * the cells it covers, in the reference's arithmetic order (strict IEEE
 * FP64, no contraction); nso_step follows the per-step sequence of
 * src/device.cu:1213-1898 (np = 0, no walls, SET_2 of src/debug.h:66).
 *
 * Where the reference's kernels race (two threads writing one ghost cell
 * with different values, or a kernel reading cells that other threads of
 * it overwrite), this restatement fixes one outcome:
 *   - the ghost layer below z = 0 takes the bottom rule's value and the one
 *     above z = nz-1 the top rule's; the other side's write to it counts
 *     only where the own side writes nothing (setNSghostNodes,
 *     src/navierstokes.cuh:660-713, with bc_bot = 2 or bc_top = 2);
 *   - smoothing reads the field as it was before the launch (the reference
 *     updates it in place, src/navierstokes.cuh:2619-2677).
 * Fields the reference leaves uninitialized on the host (epsilon, the
 * ghost corners; src/navierstokes.cpp:20-44) start at zero.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/sphere_hip.h"

static const double PI_NS = 3.14159265358979323846; /* M_PI */

typedef struct {
    sph_ns_params P;
    int nx, ny, nz;
    double dx, dy, dz;
    size_t nc, nf;
    /* cell arrays with one ghost layer */
    double *p, *ph, *phi, *dphi, *v[3], *vp[3], *dpvv[3], *eps, *eps_new, *norm, *f, *f1, *f2[3];
    /* face arrays (staggered grid) with ghosts */
    double *vf[3], *vpf[3], *dtau[3];
    /* grains held in place (nso_set_grains): original order x (x, y, z, r)
     * and vel rows, cell lists of the stable (cell, index) sort, the
     * interaction forces per grain and their sum per cell */
    uint32_t np;
    double *gx, *gv, *gxs, origo[3];
    int periodic;
    uint32_t *cstart, *cend, *gidx;
    double *fd, *fp, *fv, *fsum, *Fpf[3];
    double *force;  /* coupled to the DEM oracle: its grains' force rows */
    /* the top wall (coupled, wmode 1 or 3): its host position (< 0: none),
     * the layer it is in and the hydrostatic step per layer
     * (src/device.cu:1008-1015, 1380-1392) */
    double wall_x, dp_dz;
    unsigned wall0_iz;
    long iter;
    double t;
    sph_ns_stats st;
} nso;

#define IDX(o, x, y, z) ((size_t)((x) + 1) + (size_t)((o)->nx + 2) * (size_t)((y) + 1) + \
                         (size_t)((o)->nx + 2) * (size_t)((o)->ny + 2) * (size_t)((z) + 1))
#define VIDX(o, x, y, z) ((size_t)((x) + 1) + (size_t)((o)->nx + 3) * (size_t)((y) + 1) + \
                          (size_t)((o)->nx + 3) * (size_t)((o)->ny + 3) * (size_t)((z) + 1))
#define FOR_CELLS(o) for (int z = 0; z < (o)->nz; ++z) for (int y = 0; y < (o)->ny; ++y) \
                         for (int x = 0; x < (o)->nx; ++x)

static double amean(double a, double b) { return (a + b) * 0.5; }  /* navierstokes.cuh:17-20 */

/* findDistMod, contactsearch.cuh:12-84: the periodic image of a target
 * cell; -1 when it lies outside a non-periodic boundary */
static int find_dist_mod(const nso *o, int *t, double *dm)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const double *L = o->P.L;
    if (o->periodic == 1 || o->periodic == 2) {
        if (t[0] < 0) { t[0] += nx; dm[0] += L[0]; }
        if (t[0] >= nx) { t[0] -= nx; dm[0] -= L[0]; }
    }
    if (o->periodic == 1) {
        if (t[1] < 0) { t[1] += ny; dm[1] += L[1]; }
        if (t[1] >= ny) { t[1] -= ny; dm[1] -= L[1]; }
    }
    if (o->periodic == 2 && (t[1] < 0 || t[1] >= ny)) return -1;
    if (o->periodic > 2 || o->periodic < 1) {
        if (t[0] < 0 || t[0] >= nx) return -1;
        if (t[1] < 0 || t[1] >= ny) return -1;
    }
    if (t[2] < 0 || t[2] >= nz) return -1;
    return 0;
}

/* findPorositiesVelocitiesDiametersSpherical with grains,
 * navierstokes.cuh:1136-1318: the void volume of a sphere of radius
 * min(dx, dy, dz) at each cell centre, over the grains of the 5 x 5 x 5
 * cells around it (the averaged grain velocities and diameters it also
 * forms are read by no kernel of the fluid step) */
static void porosity_grains(nso *o)
{
    const double dx = o->dx, dy = o->dy, dz = o->dz;
    const double R = fmin(dx, fmin(dy, dz));
    const double cell_volume = 4.0 / 3.0 * PI_NS * R * R * R;
    FOR_CELLS(o) {
        const double X[3] = {x * dx + 0.5 * dx, y * dy + 0.5 * dy, z * dz + 0.5 * dz};
        double void_volume = cell_volume;
        const size_t c = IDX(o, x, y, z);
        const double phi_0 = o->phi[c];
        for (int zd = -2; zd < 3; ++zd)
            for (int yd = -2; yd < 3; ++yd)
                for (int xd = -2; xd < 3; ++xd) {
                    int t[3] = {x + xd, y + yd, z + zd};
                    double dm[3] = {0.0, 0.0, 0.0};
                    if (find_dist_mod(o, t, dm) == -1) continue;
                    const uint32_t cell = (uint32_t)t[0] + (uint32_t)t[1] * (uint32_t)o->nx +
                                          (uint32_t)(o->nx * o->ny) * (uint32_t)t[2];
                    const uint32_t start = o->cstart[cell];
                    if (start == 0xffffffffu) continue;
                    for (uint32_t i = start; i < o->cend[cell]; ++i) {
                        const double *xr = o->gxs + 4 * (size_t)i;
                        const double r = xr[3];
                        double dist[3] = {X[0] - xr[0], X[1] - xr[1], X[2] - xr[2]};
                        for (int k = 0; k < 3; ++k) dist[k] = dist[k] + dm[k];
                        const double d = sqrt(dist[0] * dist[0] + dist[1] * dist[1] + dist[2] * dist[2]);
                        if ((R - r) < d && d < (R + r))
                            void_volume -= 1.0 / (12.0 * d) *
                                           (PI_NS * (R + r - d) * (R + r - d) *
                                            (d * d + 2.0 * d * r - 3.0 * r * r + 2.0 * d * R + 6.0 * r * R -
                                             3.0 * R * R));
                        if (d <= R - r) void_volume -= 4.0 / 3.0 * PI_NS * r * r * r;
                    }
                }
        const double phi = fmin(1.00, fmax(0.00, void_volume / cell_volume));
        double dphi = phi - phi_0;
        if (o->iter == 0) dphi = 0.0;
        o->phi[c] = phi * o->P.c_phi;
        o->dphi[c] = dphi * o->P.c_phi;
    }
}

/* findInteractionForce, navierstokes.cuh:3141-3282 (SET_2): drag (with
 * the radius where the reference's formula names a diameter), pressure and
 * viscous force on each grain from the fluid of its cell */
static void interaction_force(nso *o)
{
    const double dx = o->dx, dy = o->dy, dz = o->dz, mu = o->P.mu, rho_f = o->P.rho_f;
    for (uint32_t i = 0; i < o->np; ++i) {
        const double *x = o->gx + 4 * (size_t)i, *vel = o->gv + 4 * (size_t)i;
        const double d = x[3];
        const int ix = (int)(unsigned)floor((x[0] - o->origo[0]) / (o->P.L[0] / o->nx));
        const int iy = (int)(unsigned)floor((x[1] - o->origo[1]) / (o->P.L[1] / o->ny));
        const int iz = (int)(unsigned)floor((x[2] - o->origo[2]) / (o->P.L[2] / o->nz));
        const size_t c = IDX(o, ix, iy, iz);
        const double phi = o->phi[c];
        const size_t f0 = VIDX(o, ix, iy, iz);
        const size_t fp[3] = {VIDX(o, ix + 1, iy, iz), VIDX(o, ix, iy + 1, iz), VIDX(o, ix, iy, iz + 1)};
        double v_rel[3], div_tau[3];
        for (int k = 0; k < 3; ++k) {
            const double v_f = amean(o->vf[k][f0], o->vf[k][fp[k]]);
            div_tau[k] = amean(o->dtau[k][f0], o->dtau[k][fp[k]]);
            v_rel[k] = v_f - vel[k];
        }
        const double v_rel_length = sqrt(v_rel[0] * v_rel[0] + v_rel[1] * v_rel[1] + v_rel[2] * v_rel[2]);
        const double V_p = dx * dy * dz - phi * dx * dy * dz;
        const double Re = rho_f * d * phi * v_rel_length / mu;
        double Cd = pow(0.63 + 4.8 / pow(Re, 0.5), 2.0);
        double chi = 3.7 - 0.65 * exp(-pow(1.5 - log10(Re), 2.0) / 2.0);
        if (v_rel_length < 1.0e-6) {
            Cd = 0.0;
            chi = 0.0;
        }
        const double sd = 0.125 * Cd * rho_f * PI_NS * d * d * phi * phi * v_rel_length;
        const double pw = pow(phi, -chi);
        const double gp[3] = {(o->p[IDX(o, ix + 1, iy, iz)] - o->p[IDX(o, ix - 1, iy, iz)]) / (2.0 * dx),
                              (o->p[IDX(o, ix, iy + 1, iz)] - o->p[IDX(o, ix, iy - 1, iz)]) / (2.0 * dy),
                              (o->p[IDX(o, ix, iy, iz + 1)] - o->p[IDX(o, ix, iy, iz - 1)]) / (2.0 * dz)};
        for (int k = 0; k < 3; ++k) {
            const double f_d = sd * v_rel[k] * pw;
            const double f_p = -1.0 * gp[k] * V_p;
            const double f_v = -1.0 * div_tau[k] * V_p;
            o->fd[3 * i + k] = f_d;
            o->fp[3 * i + k] = f_p;
            o->fv[3 * i + k] = f_v;
            o->fsum[3 * i + k] = f_d + f_p + f_v;
            if (o->force) o->force[3 * i + k] += f_d + f_p + f_v;  /* dev_force[i] += f_pf */
        }
    }
}

/* applyInteractionForceToFluid, navierstokes.cuh:3284-3345: the grains'
 * drag summed per cell in sorted order over the cell volume (the
 * reference's sum starts uninitialised; here at zero) */
static void apply_force_to_fluid(nso *o)
{
    const double vol = o->dx * o->dy * o->dz;
    FOR_CELLS(o) {
        const uint32_t cell = (uint32_t)x + (uint32_t)y * (uint32_t)o->nx + (uint32_t)(o->nx * o->ny) * (uint32_t)z;
        double fi[3] = {0.0, 0.0, 0.0};
        const uint32_t start = o->cstart[cell];
        if (start != 0xffffffffu)
            for (uint32_t i = start; i < o->cend[cell]; ++i)
                for (int k = 0; k < 3; ++k) fi[k] += o->fd[3 * (size_t)o->gidx[i] + k];
        for (int k = 0; k < 3; ++k) o->Fpf[k][IDX(o, x, y, z)] = fi[k] / vol;
    }
}

/* setNSghostNodes<T>(field, bc_bot, bc_top), navierstokes.cuh:660-713;
 * the z ghost writes as stated in the header comment. */
static void ghost_cells(nso *o, double *a)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz, bb = o->P.bc_bot, bt = o->P.bc_top;
    FOR_CELLS(o) {
        const double val = a[IDX(o, x, y, z)];
        if (x == 0) a[IDX(o, nx, y, z)] = val;
        if (x == nx - 1) a[IDX(o, -1, y, z)] = val;
        if (y == 0) a[IDX(o, x, ny, z)] = val;
        if (y == ny - 1) a[IDX(o, x, -1, z)] = val;
        if (z == 0 && (bb == 0 || bb == 1)) a[IDX(o, x, y, -1)] = val;
        if (z == 0 && bb == 2 && bt == 2) a[IDX(o, x, y, nz)] = val;
        if (z == nz - 1 && bt == 0) a[IDX(o, x, y, nz)] = val;
        if (z == nz - 2 && bt == 1) a[IDX(o, x, y, nz)] = val;
        if (z == nz - 1 && bt == 2 && bb == 2) a[IDX(o, x, y, -1)] = val;
    }
}

static void ghost_cells3(nso *o, double **a)
{
    for (int c = 0; c < 3; ++c) ghost_cells(o, a[c]);
}

/* setNSghostNodesFace<T>, navierstokes.cuh:715-911 (bc 0-2) */
static void ghost_faces(nso *o, double **F)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz, bb = o->P.bc_bot, bt = o->P.bc_top;
    double *X = F[0], *Y = F[1], *Z = F[2];
    FOR_CELLS(o) {
        const double vx = X[VIDX(o, x, y, z)], vy = Y[VIDX(o, x, y, z)], vz = Z[VIDX(o, x, y, z)];
        if (x == 0) {
            X[VIDX(o, nx, y, z)] = vx;
            Y[VIDX(o, nx, y, z)] = vy;
            Z[VIDX(o, nx, y, z)] = vz;
        }
        if (x == 1) X[VIDX(o, nx + 1, y, z)] = vx;
        if (x == nx - 1) {
            X[VIDX(o, -1, y, z)] = vx;
            Y[VIDX(o, -1, y, z)] = vy;
            Z[VIDX(o, -1, y, z)] = vz;
        }
        if (z == nz - 1 && x == nx - 1 && bt == 0) Z[VIDX(o, -1, y, nz)] = vz;
        if (z == nz - 1 && x == nx - 1 && (bt == 1 || bt == 2)) Z[VIDX(o, -1, y, nz)] = 0.0;
        if (z == nz - 1 && y == ny - 1 && bt == 0) Z[VIDX(o, x, -1, nz)] = vz;
        if (z == nz - 1 && y == ny - 1 && (bt == 1 || bt == 2)) Z[VIDX(o, x, -1, nz)] = 0.0;
        if (z == 0 && x == 0 && (bb == 0 || bb == 1)) X[VIDX(o, nx, y, -1)] = vx;
        if (z == 0 && x == 0 && bb == 2) X[VIDX(o, nx, y, -1)] = -vx;
        if (z == 0 && y == 0 && (bb == 0 || bb == 1)) Y[VIDX(o, x, ny, -1)] = vy;
        if (z == 0 && y == 0 && bb == 2) Y[VIDX(o, x, ny, -1)] = -vy;
        if (z == nz - 1 && x == 0 && bt == 0) Z[VIDX(o, nx, y, nz)] = vz;
        if (z == nz - 1 && x == 0 && (bt == 1 || bt == 2)) Z[VIDX(o, nx, y, nz)] = 0.0;
        if (z == nz - 1 && y == 0 && bt == 0) Z[VIDX(o, x, ny, nz)] = vz;
        if (z == nz - 1 && y == 0 && (bt == 1 || bt == 2)) Z[VIDX(o, x, ny, nz)] = 0.0;
        /* the reference tests bc_bot for these two top corners */
        if (z == nz - 1 && x == 0 && (bb == 0 || bb == 1)) X[VIDX(o, nx, y, nz)] = vx;
        if (z == nz - 1 && x == 0 && bb == 2) X[VIDX(o, nx, y, nz)] = -vx;
        if (z == nz - 1 && y == 0 && (bb == 0 || bb == 1)) Y[VIDX(o, x, ny, nz)] = vy;
        if (z == nz - 1 && y == 0 && bb == 2) Y[VIDX(o, x, ny, nz)] = -vy;
        if (y == 0) {
            X[VIDX(o, x, ny, z)] = vx;
            Y[VIDX(o, x, ny, z)] = vy;
            Z[VIDX(o, x, ny, z)] = vz;
        }
        if (y == 1) Y[VIDX(o, x, ny + 1, z)] = vy;
        if (y == ny - 1) {
            X[VIDX(o, x, -1, z)] = vx;
            Y[VIDX(o, x, -1, z)] = vy;
            Z[VIDX(o, x, -1, z)] = vz;
        }
        if (z == 0 && bb == 0) {  /* the reference swaps x and y here */
            X[VIDX(o, x, y, -1)] = vy;
            Y[VIDX(o, x, y, -1)] = vx;
            Z[VIDX(o, x, y, -1)] = vz;
        }
        if (z == 0 && bb == 1) {
            X[VIDX(o, x, y, -1)] = vx;
            Y[VIDX(o, x, y, -1)] = vy;
            Z[VIDX(o, x, y, -1)] = 0.0;
        }
        if (z == 0 && bb == 2) {
            X[VIDX(o, x, y, -1)] = -vx;
            Y[VIDX(o, x, y, -1)] = -vy;
            Z[VIDX(o, x, y, -1)] = 0.0;
        }
        if (z == nz - 1 && bt == 0) Z[VIDX(o, x, y, nz)] = vz;
        if (z == nz - 1 && bt == 1) {
            X[VIDX(o, x, y, nz)] = vx;
            Y[VIDX(o, x, y, nz)] = vy;
            Z[VIDX(o, x, y, nz)] = 0.0;
            Z[VIDX(o, x, y, nz + 1)] = 0.0;
        }
        if (z == nz - 1 && bt == 2) {
            X[VIDX(o, x, y, nz)] = -vx;
            Y[VIDX(o, x, y, nz)] = -vy;
            Z[VIDX(o, x, y, nz)] = 0.0;
            Z[VIDX(o, x, y, nz + 1)] = 0.0;
        }
    }
}

/* interpolateCenterToFace, navierstokes.cuh:3347-3382 */
static void center_to_face(nso *o, double **c, double **F)
{
    for (int z = 0; z <= o->nz; ++z)
        for (int y = 0; y <= o->ny; ++y)
            for (int x = 0; x <= o->nx; ++x) {
                const size_t f = VIDX(o, x, y, z);
                F[0][f] = amean(c[0][IDX(o, x, y, z)], c[0][IDX(o, x - 1, y, z)]);
                F[1][f] = amean(c[1][IDX(o, x, y, z)], c[1][IDX(o, x, y - 1, z)]);
                F[2][f] = amean(c[2][IDX(o, x, y, z)], c[2][IDX(o, x, y, z - 1)]);
            }
}

/* interpolateFaceToCenter, navierstokes.cuh:3384-3425 */
static void face_to_center(nso *o, double **F, double **c)
{
    FOR_CELLS(o) {
        const size_t i = IDX(o, x, y, z);
        c[0][i] = amean(F[0][VIDX(o, x, y, z)], F[0][VIDX(o, x + 1, y, z)]);
        c[1][i] = amean(F[1][VIDX(o, x, y, z)], F[1][VIDX(o, x, y + 1, z)]);
        c[2][i] = amean(F[2][VIDX(o, x, y, z)], F[2][VIDX(o, x, y, z + 1)]);
    }
}

/* findFaceDivTau, navierstokes.cuh:3427-3505 */
static void face_div_tau(nso *o)
{
    const double dx = o->dx, dy = o->dy, dz = o->dz, mu = o->P.mu;
    FOR_CELLS(o) {
        const size_t f = VIDX(o, x, y, z);
        for (int c = 0; c < 3; ++c) {
            const double *F = o->vf[c];
            const double v = F[f];
            o->dtau[c][f] = mu * ((F[VIDX(o, x + 1, y, z)] - 2.0 * v + F[VIDX(o, x - 1, y, z)]) / (dx * dx) +
                                  (F[VIDX(o, x, y + 1, z)] - 2.0 * v + F[VIDX(o, x, y - 1, z)]) / (dy * dy) +
                                  (F[VIDX(o, x, y, z + 1)] - 2.0 * v + F[VIDX(o, x, y, z - 1)]) / (dz * dz));
        }
    }
}

/* findNSdivphiviv, navierstokes.cuh:1909-2044: half upwind, half central
 * differences of div(phi v_i v) */
static void div_phi_vi_v(nso *o)
{
    const double dx = o->dx, dy = o->dy, dz = o->dz;
    const double *ph = o->phi;
    FOR_CELLS(o) {
        const size_t c = IDX(o, x, y, z);
        const size_t n[3] = {IDX(o, x - 1, y, z), IDX(o, x, y - 1, z), IDX(o, x, y, z - 1)};
        const size_t pp[3] = {IDX(o, x + 1, y, z), IDX(o, x, y + 1, z), IDX(o, x, y, z + 1)};
        const double h[3] = {dx, dy, dz};
        double v[3], a[3];
        for (int k = 0; k < 3; ++k) {
            v[k] = o->v[k][c];
            a[k] = copysign(1.0, v[k]);
        }
        const double phi = ph[c];
        for (int i = 0; i < 3; ++i) {  /* component i: sum over directions j */
            double uw = 0.0, cd = 0.0;
            for (int j = 0; j < 3; ++j) {
                const double vn_i = o->v[i][n[j]], vn_j = o->v[j][n[j]];
                const double vp_i = o->v[i][pp[j]], vp_j = o->v[j][pp[j]];
                const double tu = ((1.0 + a[j]) * (phi * v[i] * v[j] - ph[n[j]] * vn_i * vn_j) +
                                   (1.0 - a[j]) * (ph[pp[j]] * vp_i * vp_j - phi * v[i] * v[j])) / (2.0 * h[j]);
                const double tc = (ph[pp[j]] * vp_i * vp_j - ph[n[j]] * vn_i * vn_j) / (2.0 * h[j]);
                uw = j == 0 ? tu : uw + tu;
                cd = j == 0 ? tc : cd + tc;
            }
            const double tau = 0.5;
            o->dpvv[i][c] = tau * uw + (1.0 - tau) * cd;
        }
    }
}

/* findPredNSvelocities, navierstokes.cuh:2276-2495 (SET_2, np = 0) */
static void predict(nso *o)
{
    const double dx = o->dx, dy = o->dy, dz = o->dz;
    const sph_ns_params *P = &o->P;
    const double dt = P->ndem * P->dt;
    FOR_CELLS(o) {
        const size_t f = VIDX(o, x, y, z);
        const size_t c = IDX(o, x, y, z);
        const size_t n[3] = {IDX(o, x - 1, y, z), IDX(o, x, y - 1, z), IDX(o, x, y, z - 1)};
        const double h[3] = {dx, dy, dz};
        double vp[3];
        for (int k = 0; k < 3; ++k) {
            const double v = o->vf[k][f];
            const double div_tau = P->mu > 0.0 ? o->dtau[k][f] : 0.0;
            const double phi = amean(o->phi[c], o->phi[n[k]]);
            const double dphi = amean(o->dphi[c], o->dphi[n[k]]);
            const int coupled = P->mu > 0.0 && o->np > 0;
            const double f_i = coupled ? amean(o->Fpf[k][c], o->Fpf[k][n[k]]) : amean(0.0, 0.0);
            double pressure_term = 0.0;
            if (P->beta > 0.0) {
                const double grad_p = (o->p[c] - o->p[n[k]]) / h[k];
                pressure_term = -P->beta * dt / P->rho_f * grad_p;
            }
            /* the reference averages the x component of the lower
             * neighbour's div(phi v v) for every direction */
            const double dpvv = amean(o->dpvv[0][n[k]], o->dpvv[k][c]);
            const double interaction_term = -dt / (P->rho_f * phi) * f_i;
            const double gravity_term = P->g[k] * dt;
            const double advection_term = -1.0 * dpvv * dt / phi;
            const double porosity_term = -1.0 * v * dphi / phi;
            const double diffusion_term = dt / P->rho_f * div_tau;
            vp[k] = v + P->c_v * (pressure_term + interaction_term + diffusion_term + gravity_term +
                                  porosity_term + advection_term);
        }
        if (z == 0 && P->bc_bot == 1) vp[2] = 0.0;
        if (z == 0 && P->bc_bot == 2) vp[0] = vp[1] = vp[2] = 0.0;
        if ((unsigned)z >= o->wall0_iz) vp[0] = vp[1] = vp[2] = 0.0;
        for (int k = 0; k < 3; ++k) o->vpf[k][f] = vp[k];
    }
}

/* findNSforcing, navierstokes.cuh:2497-2617 (SET_2) with gradient and
 * divergence, navierstokes.cuh:1590-1640 */
static void forcing(nso *o, unsigned nijac)
{
    const double dx = o->dx, dy = o->dy, dz = o->dz;
    const sph_ns_params *P = &o->P;
    FOR_CELLS(o) {
        const size_t c = IDX(o, x, y, z);
        if (nijac == 0) {
            const double phi = o->phi[c], dphi = o->dphi[c];
            const double div_v_p = (o->vpf[0][VIDX(o, x + 1, y, z)] - o->vpf[0][VIDX(o, x, y, z)]) / dx +
                                   (o->vpf[1][VIDX(o, x, y + 1, z)] - o->vpf[1][VIDX(o, x, y, z)]) / dy +
                                   (o->vpf[2][VIDX(o, x, y, z + 1)] - o->vpf[2][VIDX(o, x, y, z)]) / dz;
            const double gp[3] = {(o->phi[IDX(o, x + 1, y, z)] - o->phi[IDX(o, x - 1, y, z)]) / (2.0 * dx),
                                  (o->phi[IDX(o, x, y + 1, z)] - o->phi[IDX(o, x, y - 1, z)]) / (2.0 * dy),
                                  (o->phi[IDX(o, x, y, z + 1)] - o->phi[IDX(o, x, y, z - 1)]) / (2.0 * dz)};
            const double dt = P->dt * P->ndem;
            const double dot_gv = gp[0] * o->vp[0][c] + gp[1] * o->vp[1][c] + gp[2] * o->vp[2][c];
            const double t1 = P->rho_f * div_v_p / (P->c_v * dt);
            const double t2 = P->rho_f * dot_gv / (P->c_v * dt * phi);
            const double t4 = P->rho_f * dphi / (dt * dt * phi * P->c_v);
            o->f1[c] = t1 + t2 + t4;
            for (int k = 0; k < 3; ++k) o->f2[k][c] = gp[k] / phi;
        }
        const double ge[3] = {(o->eps[IDX(o, x + 1, y, z)] - o->eps[IDX(o, x - 1, y, z)]) / (2.0 * dx),
                              (o->eps[IDX(o, x, y + 1, z)] - o->eps[IDX(o, x, y - 1, z)]) / (2.0 * dy),
                              (o->eps[IDX(o, x, y, z + 1)] - o->eps[IDX(o, x, y, z - 1)]) / (2.0 * dz)};
        o->f[c] = o->f1[c] - (ge[0] * o->f2[0][c] + ge[1] * o->f2[1][c] + ge[2] * o->f2[2][c]);
    }
}

static void z_range(const nso *o, int *zmin, int *zmax)
{
    *zmin = o->P.bc_bot == 0 ? 1 : 0;
    *zmax = o->P.bc_top == 0 ? o->nz - 2 : o->nz - 1;
}

/* jacobiIterationNS, navierstokes.cuh:2679-2792 (no wall: wall0_iz is
 * 10000000, src/device.cu:1008) */
static void jacobi(nso *o)
{
    const double dxdx = o->dx * o->dx, dydy = o->dy * o->dy, dzdz = o->dz * o->dz;
    const double theta = o->P.theta;
    int zmin, zmax;
    z_range(o, &zmin, &zmax);
    FOR_CELLS(o) {
        if (z < zmin || z > zmax) continue;
        const size_t c = IDX(o, x, y, z);
        const double *E = o->eps;
        const double e = E[c];
        double e_new = (-dxdx * dydy * dzdz * o->f[c] +
                        dydy * dzdz * (E[IDX(o, x - 1, y, z)] + E[IDX(o, x + 1, y, z)]) +
                        dxdx * dzdz * (E[IDX(o, x, y - 1, z)] + E[IDX(o, x, y + 1, z)]) +
                        dxdx * dydy * (E[IDX(o, x, y, z - 1)] + E[IDX(o, x, y, z + 1)])) /
                       (2.0 * (dxdx * dydy + dxdx * dzdz + dydy * dzdz));
        if ((unsigned)z == o->wall0_iz) e_new = e;
        if ((unsigned)z == o->wall0_iz + 1u) e_new = E[IDX(o, x, y, z - 1)] + o->dp_dz;
        o->norm[c] = (e_new - e) * (e_new - e) / (e_new * e_new + 1.0e-16);
        o->eps_new[c] = e * (1.0 - theta) + e_new * theta;
    }
}

/* smoothing<Float>, navierstokes.cuh:2619-2677, reading the field as it was
 * before the pass */
static void smooth(nso *o)
{
    const double g = o->P.gamma;
    int zmin, zmax;
    z_range(o, &zmin, &zmax);
    double *E = o->eps, *S = o->eps_new;
    memcpy(S, E, o->nc * sizeof(double));
    FOR_CELLS(o) {
        if (z < zmin || z > zmax) continue;
        const size_t c = IDX(o, x, y, z);
        const double avg = 1.0 / 6.0 * (E[IDX(o, x - 1, y, z)] + E[IDX(o, x + 1, y, z)] +
                                        E[IDX(o, x, y - 1, z)] + E[IDX(o, x, y + 1, z)] +
                                        E[IDX(o, x, y, z - 1)] + E[IDX(o, x, y, z + 1)]);
        S[c] = (1.0 - g) * E[c] + g * avg;
    }
    FOR_CELLS(o) E[IDX(o, x, y, z)] = S[IDX(o, x, y, z)];
}

/* updateNSvelocity, navierstokes.cuh:2915-3057 (SET_2) */
static void update_velocity(nso *o)
{
    const sph_ns_params *P = &o->P;
    const double h[3] = {o->dx, o->dy, o->dz};
    FOR_CELLS(o) {
        const size_t f = VIDX(o, x, y, z);
        const size_t c = IDX(o, x, y, z);
        const size_t n[3] = {IDX(o, x - 1, y, z), IDX(o, x, y - 1, z), IDX(o, x, y, z - 1)};
        double v[3];
        for (int k = 0; k < 3; ++k) {
            const double ge = (o->eps[c] - o->eps[n[k]]) / h[k];
            v[k] = o->vpf[k][f] - P->c_v * P->ndem * P->dt / P->rho_f * ge;
        }
        if (z == 0 && P->bc_bot == 1) v[2] = 0.0;
        if (z == 0 && P->bc_bot == 2) v[0] = v[1] = v[2] = 0.0;
        if ((unsigned)z >= o->wall0_iz) v[0] = v[1] = v[2] = 0.0;
        for (int k = 0; k < 3; ++k) o->vf[k][f] = v[k];
    }
}

/* setNSepsilonTop / setNSepsilonBottom / setNSepsilonInterior,
 * navierstokes.cuh:356-446 */
static void set_eps_layer(nso *o, int z0, int z1, double value, int both)
{
    for (int z = z0; z <= z1; ++z)
        for (int y = 0; y < o->ny; ++y)
            for (int x = 0; x < o->nx; ++x) {
                o->eps[IDX(o, x, y, z)] = value;
                if (both) o->eps_new[IDX(o, x, y, z)] = value;
            }
}

/* One fluid solve: src/device.cu:1354-1898 */
static int ns_solve(nso *o)
{
    const sph_ns_params *P = &o->P;
    const int nz = o->nz;
    const double ptop = o->ph[IDX(o, 0, 0, nz - 1)];
    const double eps_top = ptop - P->beta * ptop;
    set_eps_layer(o, nz - 1, nz - 1, eps_top, 1);
    if (o->wall_x >= 0.0) {
        /* setNSepsilonAtTopWall, navierstokes.cuh:448-516 */
        o->wall0_iz = (unsigned)(o->wall_x / (P->L[2] / P->nz));
        if (o->wall0_iz < (unsigned)nz) set_eps_layer(o, (int)o->wall0_iz, (int)o->wall0_iz, eps_top, 1);
        if (o->wall0_iz + 1u < (unsigned)nz)
            set_eps_layer(o, (int)o->wall0_iz + 1, (int)o->wall0_iz + 1, eps_top - o->dp_dz, 1);
    }
    if ((P->p_mod_A > 1.0e-5 || P->p_mod_A < -1.0e-5) && P->p_mod_f > 1.0e-7) {
        /* setUpperPressureNS, navierstokes.cuh:1550-1588 */
        const double newp = o->ph[IDX(o, 0, 0, nz - 1)] +
                            P->p_mod_A * sin(2.0 * PI_NS * P->p_mod_f * o->t + P->p_mod_phi);
        for (int y = 0; y < o->ny; ++y)
            for (int x = 0; x < o->nx; ++x) {
                const size_t c = IDX(o, x, y, nz - 1);
                const double e = newp - P->beta * o->p[c];
                o->eps[c] = e;
                o->eps_new[c] = e;
                o->p[c] = newp;
            }
    }
    ghost_cells(o, o->p);
    ghost_faces(o, o->vpf);
    ghost_cells(o, o->phi);
    ghost_cells(o, o->dphi);
    face_to_center(o, o->vf, o->v);
    ghost_cells3(o, o->v);
    div_phi_vi_v(o);
    ghost_cells3(o, o->dpvv);
    predict(o);
    ghost_faces(o, o->vpf);
    face_to_center(o, o->vpf, o->vp);
    if (o->iter == 0) {
        const double p2 = o->ph[IDX(o, 2, 2, 2)];
        set_eps_layer(o, 1, nz - 2, p2 - P->beta * p2, 0);
        FOR_CELLS(o) o->norm[IDX(o, x, y, z)] = 0.0;
        const double p0 = o->ph[IDX(o, 0, 0, 0)];
        set_eps_layer(o, 0, 0, p0 - P->beta * p0, 1);
    }
    double max_norm = 0.0;
    unsigned nijac;
    int converged = 0;
    for (nijac = 0; nijac < P->maxiter; ++nijac) {
        forcing(o, nijac);
        ghost_cells(o, o->eps);
        jacobi(o);
        FOR_CELLS(o) o->eps[IDX(o, x, y, z)] = o->eps_new[IDX(o, x, y, z)];
        if (nijac % 10 == 0) {
            /* maxNormResNS, src/navierstokes.cpp:275-297 */
            max_norm = -1.0e9;
            FOR_CELLS(o) {
                const double r = fabs(o->norm[IDX(o, x, y, z)]);
                if (r != r) return SPH_ERR_NAN;
                if (r > max_norm) max_norm = r;
            }
        }
        if (max_norm < P->tolerance) {
            ghost_cells(o, o->eps);
            if (P->gamma > 0.0) {
                smooth(o);
                ghost_cells(o, o->eps);
            }
            converged = 1;
            break;
        }
    }
    o->st.solves++;
    o->st.iterations += converged ? nijac + 1 : nijac;
    o->st.last_nijac = converged ? nijac : nijac - 1;
    o->st.last_max_norm = max_norm;
    if (!converged) o->st.nonconverged++;
    /* updateNSpressure, navierstokes.cuh:2881-2913 */
    FOR_CELLS(o) {
        const size_t c = IDX(o, x, y, z);
        o->p[c] = P->beta * o->p[c] + o->eps[c];
    }
    update_velocity(o);
    ghost_faces(o, o->vpf);
    face_to_center(o, o->vf, o->v);
    return SPH_OK;
}

/* One DEM iteration of the fluid, src/device.cu:1213-1898 */
static int step_body(nso *o)
{
    const sph_ns_params *P = &o->P;
    /* findPorositiesVelocitiesDiametersSpherical, with np = 0
     * navierstokes.cuh:1319-1328 */
    if (o->np > 0) {
        porosity_grains(o);
    } else {
        FOR_CELLS(o) {
            o->phi[IDX(o, x, y, z)] = 1.0;
            o->dphi[IDX(o, x, y, z)] = 0.0;
        }
    }
    if (o->iter == 0) {
        ghost_cells3(o, o->v);
        center_to_face(o, o->v, o->vf);
    }
    ghost_faces(o, o->vf);
    face_div_tau(o);
    ghost_faces(o, o->dtau);
    ghost_cells(o, o->p);
    ghost_cells(o, o->phi);
    if (o->np > 0) {  /* src/device.cu:1310-1351 */
        interaction_force(o);
        ghost_cells(o, o->p);
        apply_force_to_fluid(o);
        ghost_cells3(o, o->Fpf);
    }
    if (o->iter % P->ndem == 0) {
        const int rc = ns_solve(o);
        if (rc) return rc;
    }
    o->t += P->dt;
    o->iter++;
    o->st.steps++;
    return SPH_OK;
}

int nso_step(void *h, int nsteps)
{
    nso *o = (nso *)h;
    for (int s = 0; s < nsteps; ++s) {
        const int rc = step_body(o);
        if (rc) return rc;
    }
    return SPH_OK;
}

/* One iteration coupled to the DEM oracle's grains (its original-order x
 * and vel rows of 4, the cell sort's sorted x, cell starts and ends and
 * original indices), the fluid's force on each grain added to force
 * (rows of 3). */
int nso_step_coupled(void *h, uint32_t np, double *x, double *vel, double *xs, uint32_t *cstart,
                     uint32_t *cend, uint32_t *gidx, const double *origo, int periodic, double *force,
                     double wall_x)
{
    nso *o = (nso *)h;
    o->wall_x = wall_x;
    if (np != o->np || !o->fd) {
        free(o->fd); free(o->fp); free(o->fv); free(o->fsum);
        o->fd = (double *)calloc(3 * (size_t)(np ? np : 1), sizeof(double));
        o->fp = (double *)calloc(3 * (size_t)(np ? np : 1), sizeof(double));
        o->fv = (double *)calloc(3 * (size_t)(np ? np : 1), sizeof(double));
        o->fsum = (double *)calloc(3 * (size_t)(np ? np : 1), sizeof(double));
        for (int k = 0; k < 3; ++k)
            if (!o->Fpf[k]) o->Fpf[k] = (double *)calloc(o->nc, sizeof(double));
    }
    double *keep_x = o->gx, *keep_v = o->gv, *keep_xs = o->gxs;
    uint32_t *keep_cs = o->cstart, *keep_ce = o->cend, *keep_gi = o->gidx;
    o->np = np;
    o->gx = x;
    o->gv = vel;
    o->gxs = xs;
    o->cstart = cstart;
    o->cend = cend;
    o->gidx = gidx;
    o->periodic = periodic;
    for (int k = 0; k < 3; ++k) o->origo[k] = origo[k];
    o->force = force;
    const int rc = step_body(o);
    o->force = NULL;
    o->gx = keep_x;
    o->gv = keep_v;
    o->gxs = keep_xs;
    o->cstart = keep_cs;
    o->cend = keep_ce;
    o->gidx = keep_gi;
    return rc;
}

void *nso_create(const sph_ns_params *p, const double *v, const double *pf)
{
    if (p->nx < 3 || p->ny < 3 || p->nz < 3 || p->ndem == 0 || p->bc_bot < 0 || p->bc_bot > 2 ||
        p->bc_top < 0 || p->bc_top > 2)
        return NULL;
    nso *o = (nso *)calloc(1, sizeof(nso));
    o->P = *p;
    o->nx = (int)p->nx;
    o->ny = (int)p->ny;
    o->nz = (int)p->nz;
    o->dx = p->L[0] / p->nx;
    o->dy = p->L[1] / p->ny;
    o->dz = p->L[2] / p->nz;
    o->nc = (size_t)(o->nx + 2) * (o->ny + 2) * (o->nz + 2);
    o->nf = (size_t)(o->nx + 3) * (o->ny + 3) * (o->nz + 3);
    double **cells[] = {&o->p, &o->ph, &o->phi, &o->dphi, &o->v[0], &o->v[1], &o->v[2], &o->vp[0],
                        &o->vp[1], &o->vp[2], &o->dpvv[0], &o->dpvv[1], &o->dpvv[2], &o->eps,
                        &o->eps_new, &o->norm, &o->f, &o->f1, &o->f2[0], &o->f2[1], &o->f2[2]};
    for (size_t i = 0; i < sizeof(cells) / sizeof(cells[0]); ++i) *cells[i] = calloc(o->nc, sizeof(double));
    for (int k = 0; k < 3; ++k) {
        o->vf[k] = calloc(o->nf, sizeof(double));
        o->vpf[k] = calloc(o->nf, sizeof(double));
        o->dtau[k] = calloc(o->nf, sizeof(double));
    }
    FOR_CELLS(o) {
        const size_t i = (size_t)x + (size_t)o->nx * ((size_t)y + (size_t)o->ny * z);
        const size_t c = IDX(o, x, y, z);
        for (int k = 0; k < 3; ++k) o->v[k][c] = v[3 * i + k];
        o->p[c] = pf[i];
    }
    memcpy(o->ph, o->p, o->nc * sizeof(double));
    o->t = p->t0;
    o->wall_x = -1.0;
    o->wall0_iz = 10000000u;
    o->dp_dz = fabs(p->rho_f * p->g[2] * p->L[2] / p->nz);
    return o;
}

void nso_read(void *h, double *v, double *pf, double *phi, double *dphi)
{
    nso *o = (nso *)h;
    FOR_CELLS(o) {
        const size_t i = (size_t)x + (size_t)o->nx * ((size_t)y + (size_t)o->ny * z);
        const size_t c = IDX(o, x, y, z);
        if (v)
            for (int k = 0; k < 3; ++k) v[3 * i + k] = o->v[k][c];
        if (pf) pf[i] = o->p[c];
        if (phi) phi[i] = o->phi[c];
        if (dphi) dphi[i] = o->dphi[c];
    }
    memcpy(o->ph, o->p, o->nc * sizeof(double));
}

/* which: 0-2 face velocities, 3 epsilon, 4-6 predicted face velocities,
 * 7-9 div(tau) on faces, 10-12 div(phi v v) at cells, 13 forcing f */
void nso_field(void *h, int which, double *dst)
{
    nso *o = (nso *)h;
    const double *src = NULL;
    size_t n = o->nc;
    if (which >= 0 && which < 3) src = o->vf[which], n = o->nf;
    else if (which == 3) src = o->eps;
    else if (which >= 4 && which < 7) src = o->vpf[which - 4], n = o->nf;
    else if (which >= 7 && which < 10) src = o->dtau[which - 7], n = o->nf;
    else if (which >= 10 && which < 13) src = o->dpvv[which - 10];
    else if (which == 13) src = o->f;
    if (src) memcpy(dst, src, n * sizeof(double));
}

void nso_stats(void *h, sph_ns_stats *out) { *out = ((nso *)h)->st; }

/* Grains held in place: x and vel rows of 4 (x, y, z, radius; vx, vy, vz,
 * unused), the grid's origin and periodic mode.  The cell lists are the
 * reference's sort (cell ID of calcCellID, sorting.cuh:3-14; stable on the
 * index).  -1 when a grain lies outside the grid. */
int nso_set_grains(void *h, uint32_t np, const double *x, const double *vel, const double *origo,
                   int periodic)
{
    nso *o = (nso *)h;
    const uint32_t ncells = (uint32_t)(o->nx * o->ny * o->nz);
    uint32_t *cid = (uint32_t *)malloc(sizeof(uint32_t) * (np ? np : 1));
    for (uint32_t i = 0; i < np; ++i) {
        long ic[3];
        for (int k = 0; k < 3; ++k) {
            const double q = floor((x[4 * (size_t)i + k] - origo[k]) / (o->P.L[k] / (k == 0 ? o->nx : k == 1 ? o->ny : o->nz)));
            const long lim = k == 0 ? o->nx : k == 1 ? o->ny : o->nz;
            if (!(q >= 0.0) || q >= (double)lim) {
                free(cid);
                return -1;
            }
            ic[k] = (long)q;
        }
        cid[i] = (uint32_t)((ic[2] * o->ny) * o->nx + ic[1] * o->nx + ic[0]);
    }
    free(o->gx); free(o->gv); free(o->gxs); free(o->cstart); free(o->cend); free(o->gidx);
    free(o->fd); free(o->fp); free(o->fv); free(o->fsum);
    o->np = np;
    o->periodic = periodic;
    for (int k = 0; k < 3; ++k) o->origo[k] = origo[k];
    o->gx = (double *)malloc(sizeof(double) * 4 * (np ? np : 1));
    o->gv = (double *)malloc(sizeof(double) * 4 * (np ? np : 1));
    o->gxs = (double *)malloc(sizeof(double) * 4 * (np ? np : 1));
    o->gidx = (uint32_t *)malloc(sizeof(uint32_t) * (np ? np : 1));
    o->cstart = (uint32_t *)malloc(sizeof(uint32_t) * ncells);
    o->cend = (uint32_t *)malloc(sizeof(uint32_t) * ncells);
    o->fd = (double *)calloc(3 * (size_t)(np ? np : 1), sizeof(double));
    o->fp = (double *)calloc(3 * (size_t)(np ? np : 1), sizeof(double));
    o->fv = (double *)calloc(3 * (size_t)(np ? np : 1), sizeof(double));
    o->fsum = (double *)calloc(3 * (size_t)(np ? np : 1), sizeof(double));
    for (int k = 0; k < 3; ++k)
        if (!o->Fpf[k]) o->Fpf[k] = (double *)calloc(o->nc, sizeof(double));
    memcpy(o->gx, x, sizeof(double) * 4 * np);
    memcpy(o->gv, vel, sizeof(double) * 4 * np);
    /* counting sort by cell, stable on the index */
    uint32_t *count = (uint32_t *)calloc(ncells + 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < np; ++i) count[cid[i] + 1]++;
    for (uint32_t c = 0; c < ncells; ++c) count[c + 1] += count[c];
    for (uint32_t c = 0; c < ncells; ++c) {
        o->cstart[c] = count[c] == count[c + 1] ? 0xffffffffu : count[c];
        o->cend[c] = count[c + 1];
    }
    for (uint32_t i = 0; i < np; ++i) {
        const uint32_t pos = count[cid[i]]++;
        o->gidx[pos] = i;
        memcpy(o->gxs + 4 * (size_t)pos, x + 4 * (size_t)i, sizeof(double) * 4);
    }
    free(count);
    free(cid);
    return 0;
}

/* An output event of the coupled DEM oracle: the host copy of p refreshed
 * (src/device.cu:2576-2596) */
void nso_output_event(void *h)
{
    nso *o = (nso *)h;
    memcpy(o->ph, o->p, o->nc * sizeof(double));
}

/* The last step's forces on the grains (np rows of 3 each) */
void nso_grain_forces(void *h, double *f_d, double *f_p, double *f_v, double *f_sum)
{
    nso *o = (nso *)h;
    const size_t n = 3 * (size_t)o->np;
    if (f_d) memcpy(f_d, o->fd, n * sizeof(double));
    if (f_p) memcpy(f_p, o->fp, n * sizeof(double));
    if (f_v) memcpy(f_v, o->fv, n * sizeof(double));
    if (f_sum) memcpy(f_sum, o->fsum, n * sizeof(double));
}

void nso_destroy(void *h)
{
    nso *o = (nso *)h;
    if (!o) return;
    double *cells[] = {o->p, o->ph, o->phi, o->dphi, o->v[0], o->v[1], o->v[2], o->vp[0], o->vp[1],
                       o->vp[2], o->dpvv[0], o->dpvv[1], o->dpvv[2], o->eps, o->eps_new, o->norm, o->f,
                       o->f1, o->f2[0], o->f2[1], o->f2[2]};
    for (size_t i = 0; i < sizeof(cells) / sizeof(cells[0]); ++i) free(cells[i]);
    for (int k = 0; k < 3; ++k) {
        free(o->vf[k]);
        free(o->vpf[k]);
        free(o->dtau[k]);
        free(o->Fpf[k]);
    }
    free(o->gx);
    free(o->gv);
    free(o->gxs);
    free(o->cstart);
    free(o->cend);
    free(o->gidx);
    free(o->fd);
    free(o->fp);
    free(o->fv);
    free(o->fsum);
    free(o);
}
