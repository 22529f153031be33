// ns.hip -- the reference's Navier-Stokes fluid step (cfd_solver = 0) for a
// fluid without grains and walls, behind sph_ns_* (include/sphere_hip.h).
//
// The per-step sequence is that of src/device.cu:1213-1898; each kernel
// below restates the reference kernel it names (src/navierstokes.cuh) with
// the same FP64 operations in the same order, one thread per cell or face.
// Two things differ in form, not in result:
//   - one Jacobi iteration is one launch.  The reference's forcing, ghost,
//     Jacobi and copy kernels (src/device.cu:1636-1748) become one kernel
//     reading buffer A and writing buffer B: the forcing uses A's ghost
//     layer as it stands (the reference computes it before that
//     iteration's ghost update), the Jacobi stencil takes each ghost
//     neighbour from the interior cell the ghost rules copy it from, and
//     the kernel writes B's ghost layer from A's interior, which is what
//     the reference's eps holds at the next iteration's forcing;
//   - the max normalized residual of every tenth iteration is reduced on
//     the device (the bit patterns of non-negative doubles order like their
//     values; a NaN orders above every number) and tested there against the
//     tolerance, where the reference copies the whole field to the host
//     (src/navierstokes.cpp:275-297).  Once a test passes, the iterations
//     already enqueued return at once; the host reads the outcome after
//     each batch of about 100 iterations.
// Where the reference's ghost kernel races (bc_bot = 2 or bc_top = 2), the
// ghost layer below z = 0 keeps the bottom rule's value and the layer above
// z = nz-1 the top rule's; smoothing reads the field as it was before the
// pass (the reference smooths in place).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sphere_hip.h"
#include "multi.h"
#include "ns.h"

namespace {

constexpr int NS_THREADS = 256;
constexpr double NS_PI = 3.14159265358979323846;  // M_PI

struct NsGrid {
    int nx, ny, nz;
    double dx, dy, dz;
    int bb, bt;  // bc_bot, bc_top
};

__host__ __device__ inline size_t cidx(const NsGrid &g, int x, int y, int z)
{
    return (size_t)(x + 1) + (size_t)(g.nx + 2) * (size_t)(y + 1) +
           (size_t)(g.nx + 2) * (size_t)(g.ny + 2) * (size_t)(z + 1);
}

__host__ __device__ inline size_t fidx(const NsGrid &g, int x, int y, int z)
{
    return (size_t)(x + 1) + (size_t)(g.nx + 3) * (size_t)(y + 1) +
           (size_t)(g.nx + 3) * (size_t)(g.ny + 3) * (size_t)(z + 1);
}

__device__ inline double amean(double a, double b) { return (a + b) * 0.5; }

// interior cell of thread i (x fastest; each XCD takes one contiguous
// range of blocks); false past the last cell
__device__ inline bool cell_of(const NsGrid &g, int &x, int &y, int &z)
{
    const unsigned i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const unsigned nxy = (unsigned)g.nx * (unsigned)g.ny;
    if (i >= nxy * (unsigned)g.nz) return false;
    z = (int)(i / nxy);
    const unsigned r = i - (unsigned)z * nxy;
    y = (int)(r / (unsigned)g.nx);
    x = (int)(r - (unsigned)y * (unsigned)g.nx);
    return true;
}

struct Vec3 {
    double *c[3];
};
struct CVec3 {
    const double *c[3];
};

// findPorositiesVelocitiesDiametersSpherical with np = 0,
// src/navierstokes.cuh:1319-1328
__global__ void k_ns_porosity(NsGrid g, double *phi, double *dphi)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    phi[cidx(g, x, y, z)] = 1.0;
    dphi[cidx(g, x, y, z)] = 0.0;
}

// setNSghostNodes<T>(a, bc_bot, bc_top), src/navierstokes.cuh:660-713, on
// up to three arrays; the z ghost writes as stated at the top of this file
__device__ inline void ghost_cell_writes(const NsGrid &g, int x, int y, int z, double val, double *a)
{
    if (x == 0) a[cidx(g, g.nx, y, z)] = val;
    if (x == g.nx - 1) a[cidx(g, -1, y, z)] = val;
    if (y == 0) a[cidx(g, x, g.ny, z)] = val;
    if (y == g.ny - 1) a[cidx(g, x, -1, z)] = val;
    if (z == 0 && (g.bb == 0 || g.bb == 1)) a[cidx(g, x, y, -1)] = val;
    if (z == 0 && g.bb == 2 && g.bt == 2) a[cidx(g, x, y, g.nz)] = val;
    if (z == g.nz - 1 && g.bt == 0) a[cidx(g, x, y, g.nz)] = val;
    if (z == g.nz - 2 && g.bt == 1) a[cidx(g, x, y, g.nz)] = val;
    if (z == g.nz - 1 && g.bt == 2 && g.bb == 2) a[cidx(g, x, y, -1)] = val;
}

__global__ void k_ns_ghost_cells(NsGrid g, Vec3 a, int n)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    for (int k = 0; k < n; ++k) ghost_cell_writes(g, x, y, z, a.c[k][cidx(g, x, y, z)], a.c[k]);
}

// setNSghostNodesFace<T>, src/navierstokes.cuh:715-911 (bc 0-2)
__global__ void k_ns_ghost_faces(NsGrid g, Vec3 F)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz, bb = g.bb, bt = g.bt;
    double *X = F.c[0], *Y = F.c[1], *Z = F.c[2];
    const size_t f = fidx(g, x, y, z);
    const double vx = X[f], vy = Y[f], vz = Z[f];
    if (x == 0) {
        X[fidx(g, nx, y, z)] = vx;
        Y[fidx(g, nx, y, z)] = vy;
        Z[fidx(g, nx, y, z)] = vz;
    }
    if (x == 1) X[fidx(g, nx + 1, y, z)] = vx;
    if (x == nx - 1) {
        X[fidx(g, -1, y, z)] = vx;
        Y[fidx(g, -1, y, z)] = vy;
        Z[fidx(g, -1, y, z)] = vz;
    }
    if (z == nz - 1 && x == nx - 1 && bt == 0) Z[fidx(g, -1, y, nz)] = vz;
    if (z == nz - 1 && x == nx - 1 && (bt == 1 || bt == 2)) Z[fidx(g, -1, y, nz)] = 0.0;
    if (z == nz - 1 && y == ny - 1 && bt == 0) Z[fidx(g, x, -1, nz)] = vz;
    if (z == nz - 1 && y == ny - 1 && (bt == 1 || bt == 2)) Z[fidx(g, x, -1, nz)] = 0.0;
    if (z == 0 && x == 0 && (bb == 0 || bb == 1)) X[fidx(g, nx, y, -1)] = vx;
    if (z == 0 && x == 0 && bb == 2) X[fidx(g, nx, y, -1)] = -vx;
    if (z == 0 && y == 0 && (bb == 0 || bb == 1)) Y[fidx(g, x, ny, -1)] = vy;
    if (z == 0 && y == 0 && bb == 2) Y[fidx(g, x, ny, -1)] = -vy;
    if (z == nz - 1 && x == 0 && bt == 0) Z[fidx(g, nx, y, nz)] = vz;
    if (z == nz - 1 && x == 0 && (bt == 1 || bt == 2)) Z[fidx(g, nx, y, nz)] = 0.0;
    if (z == nz - 1 && y == 0 && bt == 0) Z[fidx(g, x, ny, nz)] = vz;
    if (z == nz - 1 && y == 0 && (bt == 1 || bt == 2)) Z[fidx(g, x, ny, nz)] = 0.0;
    // the reference tests bc_bot for these two top corners
    if (z == nz - 1 && x == 0 && (bb == 0 || bb == 1)) X[fidx(g, nx, y, nz)] = vx;
    if (z == nz - 1 && x == 0 && bb == 2) X[fidx(g, nx, y, nz)] = -vx;
    if (z == nz - 1 && y == 0 && (bb == 0 || bb == 1)) Y[fidx(g, x, ny, nz)] = vy;
    if (z == nz - 1 && y == 0 && bb == 2) Y[fidx(g, x, ny, nz)] = -vy;
    if (y == 0) {
        X[fidx(g, x, ny, z)] = vx;
        Y[fidx(g, x, ny, z)] = vy;
        Z[fidx(g, x, ny, z)] = vz;
    }
    if (y == 1) Y[fidx(g, x, ny + 1, z)] = vy;
    if (y == ny - 1) {
        X[fidx(g, x, -1, z)] = vx;
        Y[fidx(g, x, -1, z)] = vy;
        Z[fidx(g, x, -1, z)] = vz;
    }
    if (z == 0 && bb == 0) {  // the reference swaps x and y here
        X[fidx(g, x, y, -1)] = vy;
        Y[fidx(g, x, y, -1)] = vx;
        Z[fidx(g, x, y, -1)] = vz;
    }
    if (z == 0 && bb == 1) {
        X[fidx(g, x, y, -1)] = vx;
        Y[fidx(g, x, y, -1)] = vy;
        Z[fidx(g, x, y, -1)] = 0.0;
    }
    if (z == 0 && bb == 2) {
        X[fidx(g, x, y, -1)] = -vx;
        Y[fidx(g, x, y, -1)] = -vy;
        Z[fidx(g, x, y, -1)] = 0.0;
    }
    if (z == nz - 1 && bt == 0) Z[fidx(g, x, y, nz)] = vz;
    if (z == nz - 1 && bt == 1) {
        X[fidx(g, x, y, nz)] = vx;
        Y[fidx(g, x, y, nz)] = vy;
        Z[fidx(g, x, y, nz)] = 0.0;
        Z[fidx(g, x, y, nz + 1)] = 0.0;
    }
    if (z == nz - 1 && bt == 2) {
        X[fidx(g, x, y, nz)] = -vx;
        Y[fidx(g, x, y, nz)] = -vy;
        Z[fidx(g, x, y, nz)] = 0.0;
        Z[fidx(g, x, y, nz + 1)] = 0.0;
    }
}

// interpolateCenterToFace, src/navierstokes.cuh:3347-3382, over the
// (nx+1)(ny+1)(nz+1) faces
__global__ void k_ns_center_to_face(NsGrid g, CVec3 c, Vec3 F)
{
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned ax = (unsigned)g.nx + 1, ay = (unsigned)g.ny + 1;
    if (i >= ax * ay * ((unsigned)g.nz + 1)) return;
    const int z = (int)(i / (ax * ay));
    const unsigned r = i - (unsigned)z * ax * ay;
    const int y = (int)(r / ax);
    const int x = (int)(r - (unsigned)y * ax);
    const size_t f = fidx(g, x, y, z);
    F.c[0][f] = amean(c.c[0][cidx(g, x, y, z)], c.c[0][cidx(g, x - 1, y, z)]);
    F.c[1][f] = amean(c.c[1][cidx(g, x, y, z)], c.c[1][cidx(g, x, y - 1, z)]);
    F.c[2][f] = amean(c.c[2][cidx(g, x, y, z)], c.c[2][cidx(g, x, y, z - 1)]);
}

// interpolateFaceToCenter, src/navierstokes.cuh:3384-3425
__global__ void k_ns_face_to_center(NsGrid g, CVec3 F, Vec3 c)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    const size_t i = cidx(g, x, y, z);
    c.c[0][i] = amean(F.c[0][fidx(g, x, y, z)], F.c[0][fidx(g, x + 1, y, z)]);
    c.c[1][i] = amean(F.c[1][fidx(g, x, y, z)], F.c[1][fidx(g, x, y + 1, z)]);
    c.c[2][i] = amean(F.c[2][fidx(g, x, y, z)], F.c[2][fidx(g, x, y, z + 1)]);
}

// findFaceDivTau, src/navierstokes.cuh:3427-3505
__global__ void k_ns_div_tau(NsGrid g, CVec3 V, double mu, Vec3 D)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    const size_t f = fidx(g, x, y, z);
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    for (int k = 0; k < 3; ++k) {
        const double *F = V.c[k];
        const double v = F[f];
        D.c[k][f] = mu * ((F[fidx(g, x + 1, y, z)] - 2.0 * v + F[fidx(g, x - 1, y, z)]) / (dx * dx) +
                          (F[fidx(g, x, y + 1, z)] - 2.0 * v + F[fidx(g, x, y - 1, z)]) / (dy * dy) +
                          (F[fidx(g, x, y, z + 1)] - 2.0 * v + F[fidx(g, x, y, z - 1)]) / (dz * dz));
    }
}

// findNSdivphiviv, src/navierstokes.cuh:1909-2044: half upwind, half
// central differences of div(phi v_i v)
__global__ void k_ns_div_phi_vv(NsGrid g, const double *ph, CVec3 V, Vec3 D)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    const size_t c = cidx(g, x, y, z);
    const size_t n[3] = {cidx(g, x - 1, y, z), cidx(g, x, y - 1, z), cidx(g, x, y, z - 1)};
    const size_t pp[3] = {cidx(g, x + 1, y, z), cidx(g, x, y + 1, z), cidx(g, x, y, z + 1)};
    const double h[3] = {g.dx, g.dy, g.dz};
    double v[3], a[3];
    for (int k = 0; k < 3; ++k) {
        v[k] = V.c[k][c];
        a[k] = copysign(1.0, v[k]);
    }
    const double phi = ph[c];
    for (int i = 0; i < 3; ++i) {
        double uw = 0.0, cd = 0.0;
        for (int j = 0; j < 3; ++j) {
            const double vn_i = V.c[i][n[j]], vn_j = V.c[j][n[j]];
            const double vp_i = V.c[i][pp[j]], vp_j = V.c[j][pp[j]];
            const double tu = ((1.0 + a[j]) * (phi * v[i] * v[j] - ph[n[j]] * vn_i * vn_j) +
                               (1.0 - a[j]) * (ph[pp[j]] * vp_i * vp_j - phi * v[i] * v[j])) / (2.0 * h[j]);
            const double tc = (ph[pp[j]] * vp_i * vp_j - ph[n[j]] * vn_i * vn_j) / (2.0 * h[j]);
            uw = j == 0 ? tu : uw + tu;
            cd = j == 0 ? tc : cd + tc;
        }
        const double tau = 0.5;
        D.c[i][c] = tau * uw + (1.0 - tau) * cd;
    }
}

struct NsPhys {
    double dt;  // ndem * time.dt
    double mu, rho_f, beta, theta, c_v;
    double g[3];
};

// findPredNSvelocities, src/navierstokes.cuh:2276-2495 (SET_2, np = 0)
__global__ void k_ns_predict(NsGrid g, NsPhys q, const double *p, CVec3 VF, const double *phi_a,
                             const double *dphi_a, CVec3 DT, CVec3 DPVV, CVec3 FPF, int coupled,
                             unsigned wall0_iz, Vec3 VP)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    const size_t f = fidx(g, x, y, z);
    const size_t c = cidx(g, x, y, z);
    const size_t n[3] = {cidx(g, x - 1, y, z), cidx(g, x, y - 1, z), cidx(g, x, y, z - 1)};
    const double h[3] = {g.dx, g.dy, g.dz};
    const double dt = q.dt;
    double vp[3];
    for (int k = 0; k < 3; ++k) {
        const double v = VF.c[k][f];
        const double div_tau = q.mu > 0.0 ? DT.c[k][f] : 0.0;
        const double phi = amean(phi_a[c], phi_a[n[k]]);
        const double dphi = amean(dphi_a[c], dphi_a[n[k]]);
        const double f_i = coupled ? amean(FPF.c[k][c], FPF.c[k][n[k]]) : amean(0.0, 0.0);
        double pressure_term = 0.0;
        if (q.beta > 0.0) {
            const double grad_p = (p[c] - p[n[k]]) / h[k];
            pressure_term = -q.beta * dt / q.rho_f * grad_p;
        }
        // the x component of the lower neighbour's div(phi v v) for every
        // direction, as the reference averages it
        const double dpvv = amean(DPVV.c[0][n[k]], DPVV.c[k][c]);
        const double interaction_term = -dt / (q.rho_f * phi) * f_i;
        const double gravity_term = q.g[k] * dt;
        const double advection_term = -1.0 * dpvv * dt / phi;
        const double porosity_term = -1.0 * v * dphi / phi;
        const double diffusion_term = dt / q.rho_f * div_tau;
        vp[k] = v + q.c_v * (pressure_term + interaction_term + diffusion_term + gravity_term +
                             porosity_term + advection_term);
    }
    if (z == 0 && g.bb == 1) vp[2] = 0.0;
    if (z == 0 && g.bb == 2) vp[0] = vp[1] = vp[2] = 0.0;
    if ((unsigned)z >= wall0_iz) vp[0] = vp[1] = vp[2] = 0.0;
    for (int k = 0; k < 3; ++k) VP.c[k][f] = vp[k];
}

struct NsGrains {
    uint32_t np;
    int periodic;
    double origo[3], L[3];
    double c_phi;
    const double *x, *vel, *xs;       // rows of 4: original order, sorted
    const uint32_t *cstart, *cend, *gidx;
    double *fd, *fp, *fv, *fsum;      // rows of 3
    double *force;                    // rows of 3: the grains' forces (attached fluid), or null
};

// findDistMod, src/contactsearch.cuh:12-84
__device__ inline int find_dist_mod(const NsGrid &g, const NsGrains &G, int *t, double *dm)
{
    if (G.periodic == 1 || G.periodic == 2) {
        if (t[0] < 0) { t[0] += g.nx; dm[0] += G.L[0]; }
        if (t[0] >= g.nx) { t[0] -= g.nx; dm[0] -= G.L[0]; }
    }
    if (G.periodic == 1) {
        if (t[1] < 0) { t[1] += g.ny; dm[1] += G.L[1]; }
        if (t[1] >= g.ny) { t[1] -= g.ny; dm[1] -= G.L[1]; }
    }
    if (G.periodic == 2 && (t[1] < 0 || t[1] >= g.ny)) return -1;
    if (G.periodic > 2 || G.periodic < 1) {
        if (t[0] < 0 || t[0] >= g.nx) return -1;
        if (t[1] < 0 || t[1] >= g.ny) return -1;
    }
    if (t[2] < 0 || t[2] >= g.nz) return -1;
    return 0;
}

// findPorositiesVelocitiesDiametersSpherical with grains,
// src/navierstokes.cuh:1136-1318 (the grain velocity and diameter averages
// it also forms are read by no kernel of the fluid step)
__global__ void k_ns_porosity_grains(NsGrid g, NsGrains G, int first, double *phi_a, double *dphi_a)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    const double R = fmin(dx, fmin(dy, dz));
    const double cell_volume = 4.0 / 3.0 * NS_PI * R * R * R;
    const double X[3] = {x * dx + 0.5 * dx, y * dy + 0.5 * dy, z * dz + 0.5 * dz};
    double void_volume = cell_volume;
    const size_t c = cidx(g, x, y, z);
    const double phi_0 = phi_a[c];
    for (int zd = -2; zd < 3; ++zd)
        for (int yd = -2; yd < 3; ++yd)
            for (int xd = -2; xd < 3; ++xd) {
                int t[3] = {x + xd, y + yd, z + zd};
                double dm[3] = {0.0, 0.0, 0.0};
                if (find_dist_mod(g, G, t, dm) == -1) continue;
                const uint32_t cell = (uint32_t)t[0] + (uint32_t)t[1] * (uint32_t)g.nx +
                                      (uint32_t)(g.nx * g.ny) * (uint32_t)t[2];
                const uint32_t start = G.cstart[cell];
                if (start == 0xffffffffu) continue;
                const uint32_t end = G.cend[cell];
                for (uint32_t i = start; i < end; ++i) {
                    const double *xr = G.xs + 4 * (size_t)i;
                    const double r = xr[3];
                    double dist[3] = {X[0] - xr[0], X[1] - xr[1], X[2] - xr[2]};
                    for (int k = 0; k < 3; ++k) dist[k] = dist[k] + dm[k];
                    const double d = sqrt(dist[0] * dist[0] + dist[1] * dist[1] + dist[2] * dist[2]);
                    if ((R - r) < d && d < (R + r))
                        void_volume -= 1.0 / (12.0 * d) *
                                       (NS_PI * (R + r - d) * (R + r - d) *
                                        (d * d + 2.0 * d * r - 3.0 * r * r + 2.0 * d * R + 6.0 * r * R -
                                         3.0 * R * R));
                    if (d <= R - r) void_volume -= 4.0 / 3.0 * NS_PI * r * r * r;
                }
            }
    const double phi = fmin(1.00, fmax(0.00, void_volume / cell_volume));
    double dphi = phi - phi_0;
    if (first) dphi = 0.0;
    phi_a[c] = phi * G.c_phi;
    dphi_a[c] = dphi * G.c_phi;
}

// findInteractionForce, src/navierstokes.cuh:3141-3282 (SET_2): drag (the
// reference's formula takes the radius where it names the diameter),
// pressure and viscous force on each grain from its cell's fluid
__global__ void k_ns_interaction(NsGrid g, NsPhys q, NsGrains G, const double *phi_a, const double *p,
                                 CVec3 VF, CVec3 DT)
{
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.np) return;
    const double *x = G.x + 4 * (size_t)i, *vel = G.vel + 4 * (size_t)i;
    const double dx = g.dx, dy = g.dy, dz = g.dz, mu = q.mu, rho_f = q.rho_f;
    const double d = x[3];
    const int ix = (int)(unsigned)floor((x[0] - G.origo[0]) / (G.L[0] / g.nx));
    const int iy = (int)(unsigned)floor((x[1] - G.origo[1]) / (G.L[1] / g.ny));
    const int iz = (int)(unsigned)floor((x[2] - G.origo[2]) / (G.L[2] / g.nz));
    const size_t c = cidx(g, ix, iy, iz);
    const double phi = phi_a[c];
    const size_t f0 = fidx(g, ix, iy, iz);
    const size_t fp[3] = {fidx(g, ix + 1, iy, iz), fidx(g, ix, iy + 1, iz), fidx(g, ix, iy, iz + 1)};
    double v_rel[3], div_tau[3];
    for (int k = 0; k < 3; ++k) {
        const double v_f = amean(VF.c[k][f0], VF.c[k][fp[k]]);
        div_tau[k] = amean(DT.c[k][f0], DT.c[k][fp[k]]);
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
    const double sd = 0.125 * Cd * rho_f * NS_PI * d * d * phi * phi * v_rel_length;
    const double pw = pow(phi, -chi);
    const double gp[3] = {(p[cidx(g, ix + 1, iy, iz)] - p[cidx(g, ix - 1, iy, iz)]) / (2.0 * dx),
                          (p[cidx(g, ix, iy + 1, iz)] - p[cidx(g, ix, iy - 1, iz)]) / (2.0 * dy),
                          (p[cidx(g, ix, iy, iz + 1)] - p[cidx(g, ix, iy, iz - 1)]) / (2.0 * dz)};
    for (int k = 0; k < 3; ++k) {
        const double f_d = sd * v_rel[k] * pw;
        const double f_p = -1.0 * gp[k] * V_p;
        const double f_v = -1.0 * div_tau[k] * V_p;
        G.fd[3 * i + k] = f_d;
        G.fp[3 * i + k] = f_p;
        G.fv[3 * i + k] = f_v;
        G.fsum[3 * i + k] = f_d + f_p + f_v;
        if (G.force) G.force[3 * i + k] += f_d + f_p + f_v;  // dev_force[i] += f_pf
    }
}

// applyInteractionForceToFluid, src/navierstokes.cuh:3284-3345: the drag
// summed per cell in sorted order over the cell volume (the reference's sum
// starts uninitialised; here at zero)
__global__ void k_ns_apply_force(NsGrid g, NsGrains G, Vec3 F)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    const double vol = g.dx * g.dy * g.dz;
    const uint32_t cell = (uint32_t)x + (uint32_t)y * (uint32_t)g.nx + (uint32_t)(g.nx * g.ny) * (uint32_t)z;
    double fi[3] = {0.0, 0.0, 0.0};
    const uint32_t start = G.cstart[cell];
    if (start != 0xffffffffu) {
        const uint32_t end = G.cend[cell];
        for (uint32_t i = start; i < end; ++i)
            for (int k = 0; k < 3; ++k) fi[k] += G.fd[3 * (size_t)G.gidx[i] + k];
    }
    for (int k = 0; k < 3; ++k) F.c[k][cidx(g, x, y, z)] = fi[k] / vol;
}

// setNSepsilonTop / Bottom / Interior, src/navierstokes.cuh:356-446, on
// layers z0..z1 of the current epsilon buffer
__global__ void k_ns_set_layers(NsGrid g, double *eps, int z0, int z1, double value)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    if (z >= z0 && z <= z1) eps[cidx(g, x, y, z)] = value;
}

// setUpperPressureNS, src/navierstokes.cuh:1550-1588
__global__ void k_ns_upper_pressure(NsGrid g, double *p, double *eps, double beta, double newp)
{
    int x, y, z;
    if (!cell_of(g, x, y, z) || z != g.nz - 1) return;
    const size_t c = cidx(g, x, y, z);
    const double e = newp - beta * p[c];
    eps[c] = e;
    p[c] = newp;
}

// The interior cell whose value the ghost rules copy into ghost (x, y, z),
// for the ghost neighbours of the Jacobi stencil; a ghost no rule writes
// keeps its own value.
__device__ inline size_t ghost_source(const NsGrid &g, int x, int y, int z)
{
    if (x < 0) return cidx(g, g.nx - 1, y, z);
    if (x >= g.nx) return cidx(g, 0, y, z);
    if (y < 0) return cidx(g, x, g.ny - 1, z);
    if (y >= g.ny) return cidx(g, x, 0, z);
    if (z < 0) {
        if (g.bb == 0 || g.bb == 1) return cidx(g, x, y, 0);
        if (g.bb == 2 && g.bt == 2) return cidx(g, x, y, g.nz - 1);
        return cidx(g, x, y, -1);
    }
    if (z >= g.nz) {
        if (g.bt == 0) return cidx(g, x, y, g.nz - 1);
        if (g.bt == 1) return cidx(g, x, y, g.nz - 2);
        if (g.bt == 2 && g.bb == 2) return cidx(g, x, y, 0);
        return cidx(g, x, y, g.nz);
    }
    return cidx(g, x, y, z);
}

__device__ inline unsigned long long wave_max_u64(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Solve state on the device: done (0 running, 1 converged, 2 NaN), the
// converged iteration, the running max of the current test and the last
// tested max (bit patterns).
struct NsSolveState {
    unsigned done, conv_nijac;
    unsigned long long maxw, lastmax;
};

// One Jacobi iteration: findNSforcing (src/navierstokes.cuh:2497-2617),
// setNSghostNodes(epsilon), jacobiIterationNS (:2679-2792, no wall) and
// copyValues (:2795-2827), A -> B, as described at the top of this file.
// check: the iteration's max |norm| goes into st->maxw (bit pattern).
__global__ void __launch_bounds__(NS_THREADS)
k_ns_jacobi(NsGrid g, NsPhys q, unsigned nijac, const double *A, double *B, double *f1a, Vec3 f2a,
            CVec3 VPc, CVec3 VPF, const double *phi_a, const double *dphi_a, NsSolveState *st, int check,
            unsigned wall0_iz, double dp_dz)
{
    if (*(volatile unsigned *)&st->done) return;
    int x, y, z;
    unsigned long long bits = 0ull;
    if (cell_of(g, x, y, z)) {
        const double dx = g.dx, dy = g.dy, dz = g.dz;
        const size_t c = cidx(g, x, y, z);
        double f1, f2[3];
        if (nijac == 0) {
            const double phi = phi_a[c], dphi = dphi_a[c];
            const double div_v_p = (VPF.c[0][fidx(g, x + 1, y, z)] - VPF.c[0][fidx(g, x, y, z)]) / dx +
                                   (VPF.c[1][fidx(g, x, y + 1, z)] - VPF.c[1][fidx(g, x, y, z)]) / dy +
                                   (VPF.c[2][fidx(g, x, y, z + 1)] - VPF.c[2][fidx(g, x, y, z)]) / dz;
            const double gp[3] = {(phi_a[cidx(g, x + 1, y, z)] - phi_a[cidx(g, x - 1, y, z)]) / (2.0 * dx),
                                  (phi_a[cidx(g, x, y + 1, z)] - phi_a[cidx(g, x, y - 1, z)]) / (2.0 * dy),
                                  (phi_a[cidx(g, x, y, z + 1)] - phi_a[cidx(g, x, y, z - 1)]) / (2.0 * dz)};
            const double dt = q.dt;
            const double dot_gv = gp[0] * VPc.c[0][c] + gp[1] * VPc.c[1][c] + gp[2] * VPc.c[2][c];
            const double t1 = q.rho_f * div_v_p / (q.c_v * dt);
            const double t2 = q.rho_f * dot_gv / (q.c_v * dt * phi);
            const double t4 = q.rho_f * dphi / (dt * dt * phi * q.c_v);
            f1 = t1 + t2 + t4;
            for (int k = 0; k < 3; ++k) f2[k] = gp[k] / phi;
            f1a[c] = f1;
            for (int k = 0; k < 3; ++k) f2a.c[k][c] = f2[k];
        } else {
            f1 = f1a[c];
            for (int k = 0; k < 3; ++k) f2[k] = f2a.c[k][c];
        }
        // forcing from A as it stands (its ghost layer from the previous
        // iteration)
        const double ge[3] = {(A[cidx(g, x + 1, y, z)] - A[cidx(g, x - 1, y, z)]) / (2.0 * dx),
                              (A[cidx(g, x, y + 1, z)] - A[cidx(g, x, y - 1, z)]) / (2.0 * dy),
                              (A[cidx(g, x, y, z + 1)] - A[cidx(g, x, y, z - 1)]) / (2.0 * dz)};
        const double f = f1 - (ge[0] * f2[0] + ge[1] * f2[1] + ge[2] * f2[2]);
        const double e = A[c];
        const int zmin = g.bb == 0 ? 1 : 0;
        const int zmax = g.bt == 0 ? g.nz - 2 : g.nz - 1;
        double out = e;
        if (z >= zmin && z <= zmax) {
            const double dxdx = dx * dx, dydy = dy * dy, dzdz = dz * dz;
            const double e_xn = A[ghost_source(g, x - 1, y, z)], e_xp = A[ghost_source(g, x + 1, y, z)];
            const double e_yn = A[ghost_source(g, x, y - 1, z)], e_yp = A[ghost_source(g, x, y + 1, z)];
            const double e_zn = A[ghost_source(g, x, y, z - 1)], e_zp = A[ghost_source(g, x, y, z + 1)];
            double e_new = (-dxdx * dydy * dzdz * f + dydy * dzdz * (e_xn + e_xp) +
                            dxdx * dzdz * (e_yn + e_yp) + dxdx * dydy * (e_zn + e_zp)) /
                           (2.0 * (dxdx * dydy + dxdx * dzdz + dydy * dzdz));
            if ((unsigned)z == wall0_iz) e_new = e;
            if ((unsigned)z == wall0_iz + 1u) e_new = e_zn + dp_dz;
            const double norm = (e_new - e) * (e_new - e) / (e_new * e_new + 1.0e-16);
            out = e * (1.0 - q.theta) + e_new * q.theta;
            bits = (unsigned long long)__double_as_longlong(fabs(norm));
        }
        B[c] = out;
        ghost_cell_writes(g, x, y, z, e, B);
    }
    if (check) {
        __shared__ unsigned long long part[NS_THREADS / 64];
        bits = wave_max_u64(bits);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0) part[wv] = bits;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = part[0];
            for (int w = 1; w < NS_THREADS / 64; ++w) m = part[w] > m ? part[w] : m;
            atomicMax(&st->maxw, m);
        }
    }
}

// The test after iteration nijac (src/device.cu:1763-1790): converged when
// the max is below the tolerance; a NaN ends the solve with an error.
__global__ void k_ns_decide(NsSolveState *st, unsigned nijac, double tolerance)
{
    if (threadIdx.x != 0 || st->done) return;
    const unsigned long long m = st->maxw;
    st->lastmax = m;
    st->maxw = 0ull;
    const double v = __longlong_as_double((long long)m);
    if (v != v) {
        st->done = 2u;
    } else if (v < tolerance) {
        st->done = 1u;
        st->conv_nijac = nijac;
    }
}

// smoothing<Float>, src/navierstokes.cuh:2619-2677, A -> B (all interior
// cells written; ghosts follow from a ghost pass)
__global__ void k_ns_smooth(NsGrid g, double gamma, const double *A, double *B)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    const size_t c = cidx(g, x, y, z);
    const int zmin = g.bb == 0 ? 1 : 0;
    const int zmax = g.bt == 0 ? g.nz - 2 : g.nz - 1;
    double out = A[c];
    if (z >= zmin && z <= zmax) {
        const double avg = 1.0 / 6.0 * (A[cidx(g, x - 1, y, z)] + A[cidx(g, x + 1, y, z)] +
                                        A[cidx(g, x, y - 1, z)] + A[cidx(g, x, y + 1, z)] +
                                        A[cidx(g, x, y, z - 1)] + A[cidx(g, x, y, z + 1)]);
        out = (1.0 - gamma) * A[c] + gamma * avg;
    }
    B[c] = out;
}

// updateNSpressure, src/navierstokes.cuh:2881-2913
__global__ void k_ns_update_p(NsGrid g, const double *eps, double beta, double *p)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    const size_t c = cidx(g, x, y, z);
    p[c] = beta * p[c] + eps[c];
}

// updateNSvelocity, src/navierstokes.cuh:2915-3057 (SET_2, no wall)
__global__ void k_ns_update_v(NsGrid g, NsPhys q, double c_v_ndem_dt, CVec3 VPF, const double *eps,
                              unsigned wall0_iz, Vec3 VF)
{
    int x, y, z;
    if (!cell_of(g, x, y, z)) return;
    const size_t f = fidx(g, x, y, z);
    const size_t c = cidx(g, x, y, z);
    const size_t n[3] = {cidx(g, x - 1, y, z), cidx(g, x, y - 1, z), cidx(g, x, y, z - 1)};
    const double h[3] = {g.dx, g.dy, g.dz};
    double v[3];
    for (int k = 0; k < 3; ++k) {
        const double ge = (eps[c] - eps[n[k]]) / h[k];
        v[k] = VPF.c[k][f] - c_v_ndem_dt / q.rho_f * ge;
    }
    if (z == 0 && g.bb == 1) v[2] = 0.0;
    if (z == 0 && g.bb == 2) v[0] = v[1] = v[2] = 0.0;
    if ((unsigned)z >= wall0_iz) v[0] = v[1] = v[2] = 0.0;
    for (int k = 0; k < 3; ++k) VF.c[k][f] = v[k];
}

thread_local std::string g_ns_error;

}  // namespace

struct sph_ns {
    int device = 0;
    hipStream_t s = nullptr;
    sph_ns_params P{};
    NsGrid g{};
    NsPhys q{};
    size_t nc = 0, nf = 0;
    void *arena = nullptr;
    double *p = nullptr, *phi = nullptr, *dphi = nullptr, *eps[2] = {nullptr, nullptr}, *f1 = nullptr;
    Vec3 v{}, vp{}, dpvv{}, f2{}, vf{}, vpf{}, dtau{};
    NsSolveState *dst = nullptr;
    NsSolveState *hst = nullptr;  // pinned
    std::vector<double> ph;              // the host copy of p
    // grains held in place (sph_ns_set_grains)
    NsGrains G{};
    void *grains = nullptr;
    Vec3 Fpf{};
    // attached to a context (sph_ns_attach): force rows of its grains
    int attached = 0;
    double *own_forces = nullptr;
    uint32_t own_cap = 0;
    // the context's top wall (wmode 1 or 3): its host position (< 0: none),
    // its layer and the hydrostatic step per layer (src/device.cu:1008-1015)
    double wall_x = -1.0, dp_dz = 0.0;
    unsigned wall0_iz = 10000000u;
    int cur = 0;
    long iter = 0;
    double t = 0.0;
    sph_ns_stats st{};
};

namespace {

int fail(int code, const std::string &msg)
{
    g_ns_error = msg;
    return code;
}

int hip_ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return SPH_OK;
    return fail(SPH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

unsigned cell_blocks(const sph_ns *n)
{
    const unsigned cells = (unsigned)n->g.nx * (unsigned)n->g.ny * (unsigned)n->g.nz;
    return (cells + NS_THREADS - 1) / NS_THREADS;
}

CVec3 cv(const Vec3 &a) { return CVec3{{a.c[0], a.c[1], a.c[2]}}; }

void ghost_cells(sph_ns *n, double *a0, double *a1 = nullptr, double *a2 = nullptr)
{
    Vec3 a{{a0, a1, a2}};
    hipLaunchKernelGGL(k_ns_ghost_cells, dim3(cell_blocks(n)), dim3(NS_THREADS), 0, n->s, n->g, a,
                       a1 ? 3 : 1);
}

void ghost_faces(sph_ns *n, const Vec3 &F)
{
    hipLaunchKernelGGL(k_ns_ghost_faces, dim3(cell_blocks(n)), dim3(NS_THREADS), 0, n->s, n->g, F);
}

void face_to_center(sph_ns *n, const Vec3 &F, const Vec3 &c)
{
    hipLaunchKernelGGL(k_ns_face_to_center, dim3(cell_blocks(n)), dim3(NS_THREADS), 0, n->s, n->g, cv(F), c);
}

void set_layers(sph_ns *n, int z0, int z1, double value)
{
    hipLaunchKernelGGL(k_ns_set_layers, dim3(cell_blocks(n)), dim3(NS_THREADS), 0, n->s, n->g,
                       n->eps[n->cur], z0, z1, value);
}

// the device's solve state after the iterations enqueued so far
int read_state(sph_ns *n)
{
    int rc = hip_ok(hipMemcpyAsync(n->hst, n->dst, sizeof(NsSolveState), hipMemcpyDeviceToHost, n->s),
                    "hipMemcpyAsync");
    if (rc) return rc;
    return hip_ok(hipStreamSynchronize(n->s), "Navier-Stokes Jacobi iterations");
}

// One fluid solve, src/device.cu:1354-1898
int ns_solve(sph_ns *n)
{
    const sph_ns_params &P = n->P;
    const NsGrid &g = n->g;
    const unsigned cb = cell_blocks(n);
    const double ptop = n->ph[cidx(g, 0, 0, g.nz - 1)];
    const double eps_top = ptop - P.beta * ptop;
    set_layers(n, g.nz - 1, g.nz - 1, eps_top);
    if (n->wall_x >= 0.0) {
        // setNSepsilonAtTopWall, src/navierstokes.cuh:448-516
        n->wall0_iz = (unsigned)(n->wall_x / (P.L[2] / P.nz));
        if (n->wall0_iz < (unsigned)g.nz) set_layers(n, (int)n->wall0_iz, (int)n->wall0_iz, eps_top);
        if (n->wall0_iz + 1u < (unsigned)g.nz)
            set_layers(n, (int)n->wall0_iz + 1, (int)n->wall0_iz + 1, eps_top - n->dp_dz);
    }
    if ((P.p_mod_A > 1.0e-5 || P.p_mod_A < -1.0e-5) && P.p_mod_f > 1.0e-7) {
        // the reference forms the new top pressure on the host
        // (src/device.cu:1400-1412)
        const double newp = n->ph[cidx(g, 0, 0, g.nz - 1)] +
                            P.p_mod_A * std::sin(2.0 * NS_PI * P.p_mod_f * n->t + P.p_mod_phi);
        hipLaunchKernelGGL(k_ns_upper_pressure, dim3(cb), dim3(NS_THREADS), 0, n->s, g, n->p, n->eps[n->cur],
                           P.beta, newp);
    }
    ghost_cells(n, n->p);
    ghost_faces(n, n->vpf);
    ghost_cells(n, n->phi);
    ghost_cells(n, n->dphi);
    face_to_center(n, n->vf, n->v);
    ghost_cells(n, n->v.c[0], n->v.c[1], n->v.c[2]);
    hipLaunchKernelGGL(k_ns_div_phi_vv, dim3(cb), dim3(NS_THREADS), 0, n->s, g, n->phi, cv(n->v), n->dpvv);
    ghost_cells(n, n->dpvv.c[0], n->dpvv.c[1], n->dpvv.c[2]);
    const int coupled = P.mu > 0.0 && n->G.np > 0;
    hipLaunchKernelGGL(k_ns_predict, dim3(cb), dim3(NS_THREADS), 0, n->s, g, n->q, n->p, cv(n->vf), n->phi,
                       n->dphi, cv(n->dtau), cv(n->dpvv), cv(n->Fpf), coupled, n->wall0_iz, n->vpf);
    ghost_faces(n, n->vpf);
    face_to_center(n, n->vpf, n->vp);
    if (n->iter == 0) {
        const double p2 = n->ph[cidx(g, 2, 2, 2)];
        set_layers(n, 1, g.nz - 2, p2 - P.beta * p2);
        const double p0 = n->ph[cidx(g, 0, 0, 0)];
        set_layers(n, 0, 0, p0 - P.beta * p0);
    }
    // Jacobi iterations in batches of about 100, each ending at a test
    // (src/device.cu:1763-1790); the device skips what follows a passed test
    int rc = hip_ok(hipMemsetAsync(n->dst, 0, sizeof(NsSolveState), n->s), "hipMemsetAsync");
    if (rc) return rc;
    const int cur0 = n->cur;
    unsigned next = 0;
    bool converged = false;
    while (next < P.maxiter) {
        const unsigned last = std::min<unsigned>(P.maxiter - 1, (next + 99) / 10 * 10);
        for (unsigned it = next; it <= last; ++it) {
            const int check = it % 10 == 0;
            const int a = cur0 ^ (int)(it & 1u);
            hipLaunchKernelGGL(k_ns_jacobi, dim3(cb), dim3(NS_THREADS), 0, n->s, g, n->q, it, n->eps[a],
                               n->eps[1 - a], n->f1, n->f2, cv(n->vp), cv(n->vpf), n->phi, n->dphi, n->dst,
                               check, n->wall0_iz, n->dp_dz);
            if (check)
                hipLaunchKernelGGL(k_ns_decide, dim3(1), dim3(64), 0, n->s, n->dst, it, P.tolerance);
        }
        next = last + 1;
        rc = read_state(n);
        if (rc) return rc;
        if (n->hst->done == 2u) return fail(SPH_ERR_NAN, "normalized residual is NaN");
        if (n->hst->done == 1u) {
            converged = true;
            break;
        }
    }
    const unsigned nijac = converged ? n->hst->conv_nijac + 1 : P.maxiter;
    n->cur = cur0 ^ (int)(nijac & 1u);
    double max_norm;
    std::memcpy(&max_norm, &n->hst->lastmax, sizeof max_norm);
    n->st.solves++;
    n->st.iterations += nijac;
    n->st.last_nijac = nijac - 1;
    n->st.last_max_norm = max_norm;
    if (converged) {
        ghost_cells(n, n->eps[n->cur]);
        if (P.gamma > 0.0) {
            hipLaunchKernelGGL(k_ns_smooth, dim3(cb), dim3(NS_THREADS), 0, n->s, g, P.gamma, n->eps[n->cur],
                               n->eps[1 - n->cur]);
            n->cur = 1 - n->cur;
            ghost_cells(n, n->eps[n->cur]);
        }
    } else {
        n->st.nonconverged++;
    }
    hipLaunchKernelGGL(k_ns_update_p, dim3(cb), dim3(NS_THREADS), 0, n->s, g, n->eps[n->cur], P.beta, n->p);
    const double c_v_ndem_dt = P.c_v * P.ndem * P.dt;
    hipLaunchKernelGGL(k_ns_update_v, dim3(cb), dim3(NS_THREADS), 0, n->s, g, n->q, c_v_ndem_dt, cv(n->vpf),
                       n->eps[n->cur], n->wall0_iz, n->vf);
    ghost_faces(n, n->vpf);
    face_to_center(n, n->vf, n->v);
    return hip_ok(hipGetLastError(), "Navier-Stokes kernels");
}

// One DEM iteration of the fluid, src/device.cu:1213-1898
int ns_step_body(sph_ns *n)
{
    const NsGrid &g = n->g;
    const unsigned cb = cell_blocks(n);
    if (n->G.np > 0)
        hipLaunchKernelGGL(k_ns_porosity_grains, dim3(cb), dim3(NS_THREADS), 0, n->s, g, n->G,
                           (int)(n->iter == 0), n->phi, n->dphi);
    else
        hipLaunchKernelGGL(k_ns_porosity, dim3(cb), dim3(NS_THREADS), 0, n->s, g, n->phi, n->dphi);
    if (n->iter == 0) {
        ghost_cells(n, n->v.c[0], n->v.c[1], n->v.c[2]);
        const unsigned faces = (unsigned)(g.nx + 1) * (unsigned)(g.ny + 1) * (unsigned)(g.nz + 1);
        hipLaunchKernelGGL(k_ns_center_to_face, dim3((faces + NS_THREADS - 1) / NS_THREADS), dim3(NS_THREADS), 0,
                           n->s, g, cv(n->v), n->vf);
    }
    ghost_faces(n, n->vf);
    hipLaunchKernelGGL(k_ns_div_tau, dim3(cb), dim3(NS_THREADS), 0, n->s, g, cv(n->vf), n->P.mu, n->dtau);
    ghost_faces(n, n->dtau);
    ghost_cells(n, n->p);
    ghost_cells(n, n->phi);
    if (n->G.np > 0) {  // src/device.cu:1310-1351
        hipLaunchKernelGGL(k_ns_interaction, dim3((n->G.np + NS_THREADS - 1) / NS_THREADS), dim3(NS_THREADS), 0,
                           n->s, g, n->q, n->G, n->phi, n->p, cv(n->vf), cv(n->dtau));
        ghost_cells(n, n->p);
        hipLaunchKernelGGL(k_ns_apply_force, dim3(cb), dim3(NS_THREADS), 0, n->s, g, n->G, n->Fpf);
        ghost_cells(n, n->Fpf.c[0], n->Fpf.c[1], n->Fpf.c[2]);
    }
    if (n->iter % n->P.ndem == 0) {
        const int rc = ns_solve(n);
        if (rc) return rc;
    }
    n->t += n->P.dt;
    n->iter++;
    n->st.steps++;
    return SPH_OK;
}

// grain force rows of an attached fluid, grown to np
int ensure_force_rows(sph_ns *n, uint32_t np)
{
    if (np <= n->own_cap) return SPH_OK;
    if (n->own_forces) (void)hipFree(n->own_forces);
    n->own_forces = nullptr;
    n->own_cap = 0;
    if (hipMalloc(&n->own_forces, 12 * (size_t)np * sizeof(double)) != hipSuccess)
        return fail(SPH_ERR_HIP, "device allocation failed");
    if (hipMemsetAsync(n->own_forces, 0, 12 * (size_t)np * sizeof(double), n->s) != hipSuccess)
        return fail(SPH_ERR_HIP, "hipMemsetAsync failed");
    n->own_cap = np;
    return SPH_OK;
}

}  // namespace

int ns_coupled_step(sph_ns *n, uint32_t np, const double *x, const double *vel, const double *xs,
                    const uint32_t *cstart, const uint32_t *cend, const uint32_t *gidx,
                    const double *origo, int periodic, double *force, double wall_x, hipStream_t s)
{
    n->wall_x = wall_x;
    // the fluid's launches go on the context's stream, after its forces
    hipStream_t own = n->s;
    n->s = s;
    int rc = ensure_force_rows(n, np);
    if (!rc) {
        n->G = NsGrains{};
        n->G.np = np;
        n->G.periodic = periodic;
        n->G.c_phi = n->P.c_phi;
        for (int k = 0; k < 3; ++k) {
            n->G.origo[k] = origo[k];
            n->G.L[k] = n->P.L[k];
        }
        n->G.x = x;
        n->G.vel = vel;
        n->G.xs = xs;
        n->G.cstart = cstart;
        n->G.cend = cend;
        n->G.gidx = gidx;
        n->G.fd = n->own_forces;
        n->G.fp = n->G.fd + 3 * (size_t)np;
        n->G.fv = n->G.fp + 3 * (size_t)np;
        n->G.fsum = n->G.fv + 3 * (size_t)np;
        n->G.force = force;
        rc = ns_step_body(n);
    }
    n->s = own;
    return rc;
}

void ns_set_attached(sph_ns *n, int on) { n->attached = on; }

int ns_output_event(sph_ns *n, hipStream_t s)
{
    if (hipMemcpyAsync(n->ph.data(), n->p, n->nc * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(SPH_ERR_HIP, "download of the pressure failed");
    return SPH_OK;
}

void ns_grid(const sph_ns *n, uint32_t num[3], double L[3])
{
    num[0] = n->P.nx;
    num[1] = n->P.ny;
    num[2] = n->P.nz;
    for (int k = 0; k < 3; ++k) L[k] = n->P.L[k];
}

extern "C" {

const char *sph_ns_last_error(void) { return g_ns_error.c_str(); }

sph_ns *sph_ns_create(int device, const sph_ns_params *params, const double *v, const double *p)
{
    if (!params || !v || !p) {
        fail(SPH_ERR_ARG, "null argument");
        return nullptr;
    }
    const sph_ns_params &P = *params;
    if (P.nx < 3 || P.ny < 3 || P.nz < 3 || P.nx > 4096 || P.ny > 4096 || P.nz > 4096 ||
        (uint64_t)P.nx * P.ny * P.nz > (1ull << 28)) {
        fail(SPH_ERR_ARG, "each grid dimension must be 3..4096 (at most 2^28 cells)");
        return nullptr;
    }
    if (P.ndem == 0 || P.maxiter == 0 || P.bc_bot < 0 || P.bc_bot > 2 || P.bc_top < 0 || P.bc_top > 2 ||
        !(P.dt > 0.0) || !(P.rho_f > 0.0) || !(P.L[0] > 0.0) || !(P.L[1] > 0.0) || !(P.L[2] > 0.0)) {
        fail(SPH_ERR_ARG, "ndem and maxiter must be positive, bc_bot and bc_top 0, 1 or 2, "
                          "dt, rho_f and L positive");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        fail(SPH_ERR_HIP, "no HIP device available");
        return nullptr;
    }
    if (device < 0) device = 0;
    if (device >= ndev || hipSetDevice(device) != hipSuccess) {
        fail(SPH_ERR_HIP, "device index out of range");
        return nullptr;
    }
    sph_ns *n = new sph_ns();
    n->device = device;
    n->P = P;
    n->g = NsGrid{(int)P.nx, (int)P.ny, (int)P.nz, P.L[0] / P.nx, P.L[1] / P.ny, P.L[2] / P.nz,
                  P.bc_bot, P.bc_top};
    n->q = NsPhys{P.ndem * P.dt, P.mu, P.rho_f, P.beta, P.theta, P.c_v, {P.g[0], P.g[1], P.g[2]}};
    n->nc = (size_t)(P.nx + 2) * (P.ny + 2) * (P.nz + 2);
    n->nf = (size_t)(P.nx + 3) * (P.ny + 3) * (P.nz + 3);
    const size_t ncell_arrays = 21, nface_arrays = 9;
    const size_t bytes = (ncell_arrays * n->nc + nface_arrays * n->nf) * sizeof(double) + 256;
    if (hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&n->arena, bytes) != hipSuccess ||
        hipHostMalloc((void **)&n->hst, sizeof(NsSolveState), hipHostMallocDefault) != hipSuccess) {
        fail(SPH_ERR_HIP, "device allocation failed");
        sph_ns_destroy(n);
        return nullptr;
    }
    if (hipMemsetAsync(n->arena, 0, bytes, n->s) != hipSuccess) {
        fail(SPH_ERR_HIP, "hipMemsetAsync failed");
        sph_ns_destroy(n);
        return nullptr;
    }
    double *a = (double *)n->arena;
    auto take_c = [&]() { double *r = a; a += n->nc; return r; };
    auto take_f = [&]() { double *r = a; a += n->nf; return r; };
    n->p = take_c();
    n->phi = take_c();
    n->dphi = take_c();
    n->eps[0] = take_c();
    n->eps[1] = take_c();
    n->f1 = take_c();
    for (int k = 0; k < 3; ++k) {
        n->v.c[k] = take_c();
        n->vp.c[k] = take_c();
        n->dpvv.c[k] = take_c();
        n->f2.c[k] = take_c();
        n->Fpf.c[k] = take_c();
    }
    for (int k = 0; k < 3; ++k) {
        n->vf.c[k] = take_f();
        n->vpf.c[k] = take_f();
        n->dtau.c[k] = take_f();
    }
    n->dst = (NsSolveState *)a;
    // interior values: p into the host copy and the device, v into the
    // cell-centred velocity
    n->ph.assign(n->nc, 0.0);
    std::vector<double> vh[3];
    for (int k = 0; k < 3; ++k) vh[k].assign(n->nc, 0.0);
    for (uint32_t z = 0; z < P.nz; ++z)
        for (uint32_t y = 0; y < P.ny; ++y)
            for (uint32_t x = 0; x < P.nx; ++x) {
                const size_t i = x + (size_t)P.nx * (y + (size_t)P.ny * z);
                const size_t c = cidx(n->g, (int)x, (int)y, (int)z);
                n->ph[c] = p[i];
                for (int k = 0; k < 3; ++k) vh[k][c] = v[3 * i + k];
            }
    bool ok = hipMemcpyAsync(n->p, n->ph.data(), n->nc * sizeof(double), hipMemcpyHostToDevice, n->s) == hipSuccess;
    for (int k = 0; k < 3 && ok; ++k)
        ok = hipMemcpyAsync(n->v.c[k], vh[k].data(), n->nc * sizeof(double), hipMemcpyHostToDevice, n->s) ==
             hipSuccess;
    if (!ok || hipStreamSynchronize(n->s) != hipSuccess) {
        fail(SPH_ERR_HIP, "upload failed");
        sph_ns_destroy(n);
        return nullptr;
    }
    n->t = P.t0;
    n->dp_dz = std::fabs(P.rho_f * P.g[2] * P.L[2] / P.nz);
    return n;
}

void sph_ns_destroy(sph_ns *n)
{
    if (!n) return;
    if (n->s) hipStreamSynchronize(n->s);
    if (n->arena) hipFree(n->arena);
    if (n->grains) hipFree(n->grains);
    if (n->own_forces) hipFree(n->own_forces);
    if (n->hst) hipHostFree(n->hst);
    if (n->s) hipStreamDestroy(n->s);
    delete n;
}

int sph_ns_step(sph_ns *n, int nsteps)
{
    if (!n || nsteps < 0) return fail(SPH_ERR_ARG, "bad argument");
    if (hipSetDevice(n->device) != hipSuccess) return fail(SPH_ERR_HIP, "hipSetDevice failed");
    if (n->attached) return fail(SPH_ERR_STATE, "the fluid is attached to a context: step the context");
    for (int s = 0; s < nsteps; ++s) {
        const int rc = ns_step_body(n);
        if (rc) return rc;
    }
    const int rc = hip_ok(hipGetLastError(), "Navier-Stokes kernels");
    if (rc) return rc;
    return hip_ok(hipStreamSynchronize(n->s), "Navier-Stokes step");
}

int sph_ns_read(sph_ns *n, double *v, double *p, double *phi, double *dphi)
{
    if (!n) return fail(SPH_ERR_ARG, "null context");
    if (hipSetDevice(n->device) != hipSuccess) return fail(SPH_ERR_HIP, "hipSetDevice failed");
    std::vector<double> vh[3], phih(n->nc), dphih(n->nc);
    bool ok = hipMemcpyAsync(n->ph.data(), n->p, n->nc * sizeof(double), hipMemcpyDeviceToHost, n->s) ==
              hipSuccess;
    for (int k = 0; k < 3 && ok; ++k) {
        vh[k].resize(n->nc);
        ok = hipMemcpyAsync(vh[k].data(), n->v.c[k], n->nc * sizeof(double), hipMemcpyDeviceToHost, n->s) ==
             hipSuccess;
    }
    ok = ok && hipMemcpyAsync(phih.data(), n->phi, n->nc * sizeof(double), hipMemcpyDeviceToHost, n->s) ==
                   hipSuccess;
    ok = ok && hipMemcpyAsync(dphih.data(), n->dphi, n->nc * sizeof(double), hipMemcpyDeviceToHost, n->s) ==
                   hipSuccess;
    if (!ok || hipStreamSynchronize(n->s) != hipSuccess) return fail(SPH_ERR_HIP, "download failed");
    const sph_ns_params &P = n->P;
    for (uint32_t z = 0; z < P.nz; ++z)
        for (uint32_t y = 0; y < P.ny; ++y)
            for (uint32_t x = 0; x < P.nx; ++x) {
                const size_t i = x + (size_t)P.nx * (y + (size_t)P.ny * z);
                const size_t c = cidx(n->g, (int)x, (int)y, (int)z);
                if (v)
                    for (int k = 0; k < 3; ++k) v[3 * i + k] = vh[k][c];
                if (p) p[i] = n->ph[c];
                if (phi) phi[i] = phih[c];
                if (dphi) dphi[i] = dphih[c];
            }
    return SPH_OK;
}

int sph_ns_read_field(sph_ns *n, int which, double *dst)
{
    if (!n || !dst || which < 0 || which > 3) return fail(SPH_ERR_ARG, "bad argument");
    if (hipSetDevice(n->device) != hipSuccess) return fail(SPH_ERR_HIP, "hipSetDevice failed");
    const double *src = which < 3 ? n->vf.c[which] : n->eps[n->cur];
    const size_t count = which < 3 ? n->nf : n->nc;
    if (hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, n->s) != hipSuccess ||
        hipStreamSynchronize(n->s) != hipSuccess)
        return fail(SPH_ERR_HIP, "download failed");
    return SPH_OK;
}

int sph_ns_set_grains(sph_ns *n, uint32_t np, const double *x, const double *vel, const double *origo,
                      int32_t periodic)
{
    if (!n || (np && (!x || !vel)) || !origo) return fail(SPH_ERR_ARG, "null argument");
    if (n->attached) return fail(SPH_ERR_STATE, "the fluid is attached to a context and moves its grains");
    if (hipSetDevice(n->device) != hipSuccess) return fail(SPH_ERR_HIP, "hipSetDevice failed");
    const NsGrid &g = n->g;
    const uint32_t ncells = (uint32_t)(g.nx * g.ny * g.nz);
    const int num[3] = {g.nx, g.ny, g.nz};
    // cell IDs of calcCellID (src/sorting.cuh:3-14), each grain inside the grid
    std::vector<uint32_t> cid(np);
    for (uint32_t i = 0; i < np; ++i) {
        long ic[3];
        for (int k = 0; k < 3; ++k) {
            const double q = std::floor((x[4 * (size_t)i + k] - origo[k]) / (n->P.L[k] / num[k]));
            if (!(q >= 0.0) || q >= (double)num[k]) return fail(SPH_ERR_GRID, "a grain lies outside the grid");
            ic[k] = (long)q;
        }
        cid[i] = (uint32_t)((ic[2] * g.ny) * g.nx + ic[1] * g.nx + ic[0]);
    }
    // the stable (cell, index) sort and the cell lists
    std::vector<uint32_t> count(ncells + 1, 0u), cstart(ncells), cend(ncells), gidx(np);
    std::vector<double> xs(4 * (size_t)np);
    for (uint32_t i = 0; i < np; ++i) count[cid[i] + 1]++;
    for (uint32_t c = 0; c < ncells; ++c) count[c + 1] += count[c];
    for (uint32_t c = 0; c < ncells; ++c) {
        cstart[c] = count[c] == count[c + 1] ? 0xffffffffu : count[c];
        cend[c] = count[c + 1];
    }
    for (uint32_t i = 0; i < np; ++i) {
        const uint32_t pos = count[cid[i]]++;
        gidx[pos] = i;
        std::memcpy(&xs[4 * (size_t)pos], x + 4 * (size_t)i, 4 * sizeof(double));
    }
    if (hipStreamSynchronize(n->s) != hipSuccess) return fail(SPH_ERR_HIP, "hipStreamSynchronize failed");
    if (n->grains) {
        (void)hipFree(n->grains);
        n->grains = nullptr;
    }
    n->G = NsGrains{};
    n->G.periodic = periodic;
    n->G.c_phi = n->P.c_phi;
    for (int k = 0; k < 3; ++k) {
        n->G.origo[k] = origo[k];
        n->G.L[k] = n->P.L[k];
    }
    if (np == 0) return SPH_OK;
    const size_t dbl = (size_t)np * (4 + 4 + 4 + 3 * 4), u32 = 2 * (size_t)ncells + np;
    if (hipMalloc(&n->grains, dbl * sizeof(double) + u32 * sizeof(uint32_t)) != hipSuccess)
        return fail(SPH_ERR_HIP, "device allocation failed");
    double *a = (double *)n->grains;
    double *dx = a, *dv = a + 4 * (size_t)np, *dxs = a + 8 * (size_t)np;
    n->G.fd = a + 12 * (size_t)np;
    n->G.fp = n->G.fd + 3 * (size_t)np;
    n->G.fv = n->G.fp + 3 * (size_t)np;
    n->G.fsum = n->G.fv + 3 * (size_t)np;
    uint32_t *u = (uint32_t *)(a + dbl);
    uint32_t *dcs = u, *dce = u + ncells, *dgi = u + 2 * (size_t)ncells;
    bool ok = hipMemcpyAsync(dx, x, 4 * (size_t)np * sizeof(double), hipMemcpyHostToDevice, n->s) == hipSuccess;
    ok = ok && hipMemcpyAsync(dv, vel, 4 * (size_t)np * sizeof(double), hipMemcpyHostToDevice, n->s) == hipSuccess;
    ok = ok && hipMemcpyAsync(dxs, xs.data(), xs.size() * sizeof(double), hipMemcpyHostToDevice, n->s) ==
                   hipSuccess;
    ok = ok && hipMemcpyAsync(dcs, cstart.data(), ncells * 4, hipMemcpyHostToDevice, n->s) == hipSuccess;
    ok = ok && hipMemcpyAsync(dce, cend.data(), ncells * 4, hipMemcpyHostToDevice, n->s) == hipSuccess;
    ok = ok && hipMemcpyAsync(dgi, gidx.data(), (size_t)np * 4, hipMemcpyHostToDevice, n->s) == hipSuccess;
    ok = ok && hipMemsetAsync(n->G.fd, 0, 12 * (size_t)np * sizeof(double), n->s) == hipSuccess;
    if (!ok || hipStreamSynchronize(n->s) != hipSuccess) return fail(SPH_ERR_HIP, "upload failed");
    n->G.np = np;
    n->G.x = dx;
    n->G.vel = dv;
    n->G.xs = dxs;
    n->G.cstart = dcs;
    n->G.cend = dce;
    n->G.gidx = dgi;
    return SPH_OK;
}

int sph_ns_read_grain_forces(sph_ns *n, double *f_d, double *f_p, double *f_v, double *f_sum)
{
    if (!n) return fail(SPH_ERR_ARG, "null context");
    if (hipSetDevice(n->device) != hipSuccess) return fail(SPH_ERR_HIP, "hipSetDevice failed");
    const size_t bytes = 3 * (size_t)n->G.np * sizeof(double);
    if (!bytes) return SPH_OK;
    double *dst[4] = {f_d, f_p, f_v, f_sum};
    const double *src[4] = {n->G.fd, n->G.fp, n->G.fv, n->G.fsum};
    for (int k = 0; k < 4; ++k)
        if (dst[k] && hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, n->s) != hipSuccess)
            return fail(SPH_ERR_HIP, "download failed");
    if (hipStreamSynchronize(n->s) != hipSuccess) return fail(SPH_ERR_HIP, "download failed");
    return SPH_OK;
}

int sph_ns_get_stats(const sph_ns *n, sph_ns_stats *out)
{
    if (!n || !out) return fail(SPH_ERR_ARG, "bad argument");
    *out = n->st;
    return SPH_OK;
}

}  // extern "C"
