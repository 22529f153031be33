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
            const double f_i = amean(0.0, 0.0);
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
        const double e_new = (-dxdx * dydy * dzdz * o->f[c] +
                              dydy * dzdz * (E[IDX(o, x - 1, y, z)] + E[IDX(o, x + 1, y, z)]) +
                              dxdx * dzdz * (E[IDX(o, x, y - 1, z)] + E[IDX(o, x, y + 1, z)]) +
                              dxdx * dydy * (E[IDX(o, x, y, z - 1)] + E[IDX(o, x, y, z + 1)])) /
                             (2.0 * (dxdx * dydy + dxdx * dzdz + dydy * dzdz));
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
    set_eps_layer(o, nz - 1, nz - 1, ptop - P->beta * ptop, 1);
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

int nso_step(void *h, int nsteps)
{
    nso *o = (nso *)h;
    const sph_ns_params *P = &o->P;
    for (int s = 0; s < nsteps; ++s) {
        /* findPorositiesVelocitiesDiametersSpherical with np = 0,
         * navierstokes.cuh:1319-1328 */
        FOR_CELLS(o) {
            o->phi[IDX(o, x, y, z)] = 1.0;
            o->dphi[IDX(o, x, y, z)] = 0.0;
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
        if (o->iter % P->ndem == 0) {
            const int rc = ns_solve(o);
            if (rc) return rc;
        }
        o->t += P->dt;
        o->iter++;
        o->st.steps++;
    }
    return SPH_OK;
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
    }
    free(o);
}
