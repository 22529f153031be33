// darcy_kernels.hip -- gfx950 kernels of the porosity-coupled Darcy pressure
// solve (src/darcy.cuh).  Cell fields carry one ghost layer,
// index (x+1) + (nx+2)(y+1) + (nx+2)(ny+2)(z+1) (src/darcy.cuh:134-146);
// kernels run one thread per interior cell with x fastest, so a wave reads
// consecutive doubles of a row.
#include "kernels.h"
#include "net_kernels.h"

__device__ __forceinline__ uint32_t didx(const DarcyArgs &d, int x, int y, int z)
{
    return (uint32_t)(x + 1) + (d.nx + 2) * (uint32_t)(y + 1) + (d.nx + 2) * (d.ny + 2) * (uint32_t)(z + 1);
}

__device__ __forceinline__ bool cell_xyz(const DarcyArgs &d, int &x, int &y, int &z)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = d.nx * d.ny * d.nz;
    if (c >= n) return false;
    x = (int)(c % d.nx);
    y = (int)((c / d.nx) % d.ny);
    z = (int)(c / (d.nx * d.ny));
    return true;
}

// cell of thread threadIdx.x in block bid of a launch over the interior cells
__device__ __forceinline__ bool cell_xyz_b(const DarcyArgs &d, uint32_t bid, int &x, int &y, int &z)
{
    uint32_t c = bid * blockDim.x + threadIdx.x;
    uint32_t n = d.nx * d.ny * d.nz;
    if (c >= n) return false;
    x = (int)(c % d.nx);
    y = (int)((c / d.nx) % d.ny);
    z = (int)(c / (d.nx * d.ny));
    return true;
}

static inline dim3 cell_grid(const DarcyArgs &d)
{
    uint32_t n = d.nx * d.ny * d.nz;
    return dim3((n + 255) / 256);
}

// setDarcyGhostNodes<T>, src/darcy.cuh:194-264
__device__ __forceinline__ void ghosts_body(DarcyArgs d, double *__restrict__ f, int ncomp, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
    const double *val = f + (size_t)ncomp * didx(d, x, y, z);
    double v0 = val[0], v1 = 0.0, v2 = 0.0;
    if (ncomp == 3) { v1 = val[1]; v2 = val[2]; }
#define GW(X, Y, Z)                                                   \
    do {                                                              \
        double *q = f + (size_t)ncomp * didx(d, X, Y, Z);             \
        q[0] = v0;                                                    \
        if (ncomp == 3) { q[1] = v1; q[2] = v2; }                     \
    } while (0)
    if (x == 0 && d.bc_xn == 0) GW(-1, y, z);
    if (x == 1 && d.bc_xn == 1) GW(-1, y, z);
    if (x == 0 && d.bc_xn == 2) GW(nx, y, z);
    if (x == nx - 1 && d.bc_xp == 0) GW(nx, y, z);
    if (x == nx - 2 && d.bc_xp == 1) GW(nx, y, z);
    if (x == nx - 1 && d.bc_xp == 2) GW(-1, y, z);
    if (y == 0 && d.bc_yn == 0) GW(x, -1, z);
    if (y == 1 && d.bc_yn == 1) GW(x, -1, z);
    if (y == 0 && d.bc_yn == 2) GW(x, ny, z);
    if (y == ny - 1 && d.bc_yp == 0) GW(x, ny, z);
    if (y == ny - 2 && d.bc_yp == 1) GW(x, ny, z);
    if (y == ny - 1 && d.bc_yp == 2) GW(x, -1, z);
    if (z == 0 && d.bc_bot == 0) GW(x, y, -1);
    if (z == 1 && d.bc_bot == 1) GW(x, y, -1);
    if (z == 0 && d.bc_bot == 2) GW(x, y, nz);
    if (z == nz - 1 && d.bc_top == 0) GW(x, y, nz);
    if (z == nz - 2 && d.bc_top == 1) GW(x, y, nz);
    if (z == nz - 1 && d.bc_top == 2) GW(x, y, -1);
#undef GW
}

__global__ void __launch_bounds__(256) k_ghosts(DarcyArgs d, double *__restrict__ f, int ncomp)
{
    ghosts_body(d, f, ncomp, blockIdx.x);
}

void launch_darcy_ghosts(hipStream_t s, const DarcyArgs &d, double *f, int ncomp)
{
    hipLaunchKernelGGL(k_ghosts, cell_grid(d), dim3(256), 0, s, d, f, ncomp);
}

// setDarcyGhostNodesFlux, src/darcy.cuh:270-312
__device__ __forceinline__ void ghosts_flux_body(DarcyArgs d, double *__restrict__ f,
                                                     const double *__restrict__ k, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    const int nz = (int)d.nz;
    const double pv = f[didx(d, x, y, z)];
    const double kv = k[didx(d, x, y, z)];
    const double dz = d.L2 / nz;
    double q_z = 0.;
    if (z == 0) q_z = d.bc_bot_flux;
    else if (z == nz - 1) q_z = d.bc_top_flux;
    const double p_ghost = -d.mu / kv * q_z * dz + pv;
    if (z == 0 && d.bc_bot == 4) f[didx(d, x, y, -1)] = p_ghost;
    if (z == nz - 1 && d.bc_top == 4) f[didx(d, x, y, nz)] = p_ghost;
}

__global__ void __launch_bounds__(256) k_ghosts_flux(DarcyArgs d, double *__restrict__ f,
                                                     const double *__restrict__ k)
{
    ghosts_flux_body(d, f, k, blockIdx.x);
}

void launch_darcy_ghosts_flux(hipStream_t s, const DarcyArgs &d, double *f, const double *k)
{
    if (!(d.bc_bot == 4 || d.bc_top == 4)) return;
    hipLaunchKernelGGL(k_ghosts_flux, cell_grid(d), dim3(256), 0, s, d, f, k);
}

// findDarcyPressureGradient, src/darcy.cuh:889-943
__device__ __forceinline__ void pressure_gradient_body(DarcyArgs d, const double *__restrict__ p,
                                                           double *__restrict__ grad_p, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    const double p_xn = p[didx(d, x - 1, y, z)], p_xp = p[didx(d, x + 1, y, z)];
    const double p_yn = p[didx(d, x, y - 1, z)], p_yp = p[didx(d, x, y + 1, z)];
    const double p_zn = p[didx(d, x, y, z - 1)], p_zp = p[didx(d, x, y, z + 1)];
    double *g = grad_p + 3 * (size_t)didx(d, x, y, z);
    g[0] = (p_xp - p_xn) / (d.dx + d.dx);
    g[1] = (p_yp - p_yn) / (d.dy + d.dy);
    g[2] = (p_zp - p_zn) / (d.dz + d.dz);
}

__global__ void __launch_bounds__(256) k_pressure_gradient(DarcyArgs d, const double *__restrict__ p,
                                                           double *__restrict__ grad_p)
{
    pressure_gradient_body(d, p, grad_p, blockIdx.x);
}

void launch_darcy_pressure_gradient(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b)
{
    hipLaunchKernelGGL(k_pressure_gradient, cell_grid(d), dim3(256), 0, s, d, b.p, b.grad_p);
}

// weight(), src/darcy.cuh:324-337
__device__ __forceinline__ double weight(d3 x_p, d3 x_f, double dx, double dy, double dz)
{
    const d3 dd = x_p - x_f;
    const double ax = fabs(dd.x), ay = fabs(dd.y), az = fabs(dd.z);
    if (ax < dx && ay < dy && az < dz) return (1.0 - ax / dx) * (1.0 - ay / dy) * (1.0 - az / dz);
    return 0.0;
}

__device__ __forceinline__ double sphere_volume(double r) { return 4.0 / 3.0 * PI_M * (r * r * r); }

// findDarcyPressureForceLinear, src/darcy.cuh:948-1102
__device__ __forceinline__ void pressure_force_body(StepArgs a, DarcyArgs d,
                                                        const double4 *__restrict__ xarr,
                                                        const double *__restrict__ grad_parr,
                                                        double3 *__restrict__ force,
                                                        double4 *__restrict__ f_parr,
                                                        uint32_t *__restrict__ err, uint32_t bid)
{
    uint32_t i = bid * blockDim.x + threadIdx.x;
    if (i >= a.np) return;
    const Grid &g = a.grid;
    const double4 x4 = xarr[i];
    const d3 x3 = xyz(x4);
    const double fx = floor((x4.x - g.origo[0]) / (g.L[0] / g.num[0]));
    const double fy = floor((x4.y - g.origo[1]) / (g.L[1] / g.num[1]));
    const double fz = floor((x4.z - g.origo[2]) / (g.L[2] / g.num[2]));
    if (!(fx >= 0.0 && fx < (double)d.nx && fy >= 0.0 && fy < (double)d.ny && fz >= 0.0 &&
          fz < (double)d.nz)) {
        atomicOr(err, ERRF_GRID);  // outside the grid: the step is invalid
        return;
    }
    const uint32_t i_x = (uint32_t)fx, i_y = (uint32_t)fy, i_z = (uint32_t)fz;
    const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
    const double dx = d.dx, dy = d.dy, dz = d.dz;
    const d3 X = mk3(i_x * dx + 0.5 * dx, i_y * dy + 0.5 * dy, i_z * dz + 0.5 * dz);
    d3 grad_p = mk3(0., 0., 0.);
    for (int d_iz = -1; d_iz < 2; d_iz++)
        for (int d_iy = -1; d_iy < 2; d_iy++)
            for (int d_ix = -1; d_ix < 2; d_ix++) {
                const int ix_n = (int)i_x + d_ix, iy_n = (int)i_y + d_iy, iz_n = (int)i_z + d_iz;
                const double *gp = grad_parr + 3 * (size_t)didx(d, ix_n, iy_n, iz_n);
                d3 gpi = mk3(gp[0], gp[1], gp[2]);
                if ((ix_n == -1 && iy_n == -1) || (ix_n == -1 && iz_n == -1) ||
                    (iy_n == -1 && iz_n == -1) || (ix_n == nx && iy_n == ny) ||
                    (ix_n == nx && iz_n == nz) || (iy_n == ny && iz_n == nz) ||
                    (ix_n == nx && iy_n == -1) || (ix_n == nx && iz_n == -1) ||
                    (iy_n == ny && ix_n == -1) || (iy_n == ny && iz_n == -1) ||
                    (iz_n == nz && ix_n == -1) || (iz_n == nz && iy_n == -1))
                    gpi = mk3(0., 0., 0.);
                // unsigned comparison as in src/darcy.cuh:1033
                if (d.bc_top == 0 && i_z + (uint32_t)d_iz >= d.wall0_iz - 1u) gpi.z = 0.0;
                const d3 nn = mk3(dx * d_ix, dy * d_iy, dz * d_iz);
                grad_p = grad_p + weight(x3, X + nn, dx, dy, dz) * gpi;
            }
    const double v = sphere_volume(x4.w);
    d3 f_p = v * ((-1.0) * grad_p) - (d.rho_f * v) * mk3(a.prm.g[0], a.prm.g[1], a.prm.g[2]);
    if (d.bc_top == 0 && i_z >= d.wall0_iz) f_p.z = 0.0;
    double3 F = force[i];
    F.x = F.x + f_p.x;
    F.y = F.y + f_p.y;
    F.z = F.z + f_p.z;
    force[i] = F;
    f_parr[i] = make_double4(f_p.x, f_p.y, f_p.z, 0.0);
}

__global__ void __launch_bounds__(256) k_pressure_force(StepArgs a, DarcyArgs d,
                                                        const double4 *__restrict__ xarr,
                                                        const double *__restrict__ grad_parr,
                                                        double3 *__restrict__ force,
                                                        double4 *__restrict__ f_parr,
                                                        uint32_t *__restrict__ err)
{
    pressure_force_body(a, d, xarr, grad_parr, force, f_parr, err, blockIdx.x);
}

void launch_darcy_pressure_force(hipStream_t s, const StepArgs &a, const DarcyArgs &d,
                                 const DarcyBuffers &b, const DemBuffers &db)
{
    if (!a.np) return;
    hipLaunchKernelGGL(k_pressure_force, dim3((a.np + 255) / 256), dim3(256), 0, s, a, d, db.x,
                       b.grad_p, db.force, b.f_p, db.errflags);
}

// findDistMod for the porosity loop (src/contactsearch.cuh:11-83)
__device__ __forceinline__ int dist_mod_p(const Grid &g, int &tx, int &ty, int tz, d3 &dm)
{
    int nx = (int)g.num[0], ny = (int)g.num[1], nz = (int)g.num[2];
    if (g.periodic == 1) {
        if (tx < 0) { tx += nx; dm = dm + mk3(g.L[0], 0.0, 0.0); }
        if (tx >= nx) { tx -= nx; dm = dm - mk3(g.L[0], 0.0, 0.0); }
        if (ty < 0) { ty += ny; dm = dm + mk3(0.0, g.L[1], 0.0); }
        if (ty >= ny) { ty -= ny; dm = dm - mk3(0.0, g.L[1], 0.0); }
    }
    if (g.periodic == 2) {
        if (tx < 0) { tx += nx; dm = dm + mk3(g.L[0], 0.0, 0.0); }
        if (tx >= nx) { tx -= nx; dm = dm - mk3(g.L[0], 0.0, 0.0); }
        if (ty < 0 || ty >= ny) return -1;
    }
    if (g.periodic > 2 || g.periodic < 1) {
        if (tx < 0 || tx >= nx) return -1;
        if (ty < 0 || ty >= ny) return -1;
    }
    if (tz < 0 || tz >= nz) return -1;
    return 0;
}

// findDarcyPorositiesLinear, src/darcy.cuh:359-550
__device__ __forceinline__ void porosities_body(StepArgs a, DarcyArgs d, long iteration,
                                                    const uint32_t *__restrict__ cell_start,
                                                    const uint32_t *__restrict__ cell_end,
                                                    const double4 *__restrict__ x_s,
                                                    const double4 *__restrict__ vel_s,
                                                    double *__restrict__ phi_a,
                                                    double *__restrict__ dphi_a,
                                                    double *__restrict__ vp_avg_a, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    const uint32_t cellidx = didx(d, x, y, z);
    if (a.np == 0) {
        phi_a[cellidx] = 0.9;
        dphi_a[cellidx] = 0.0;
        return;
    }
    const Grid &g = a.grid;
    const int ny = (int)d.ny, nz = (int)d.nz;
    const double dx = d.dx, dy = d.dy, dz = d.dz;
    const double dt = a.dt;
    double solid_volume = 0.0, solid_volume_new = 0.0;
    const d3 X = mk3(x * dx + 0.5 * dx, y * dy + 0.5 * dy, z * dz + 0.5 * dz);
    d3 vp_avg_num = mk3(0.0, 0.0, 0.0);
    double vp_avg_denum = 0.0;
    const double phi_0 = phi_a[cellidx];
    for (int zd = -1; zd < 2; ++zd)
        for (int yd = -1; yd < 2; ++yd)
            for (int xd = -1; xd < 2; ++xd) {
                int tx = x + xd, ty = y + yd, tz = z + zd;
                if (tz == -1) tz = 1;
                if (g.periodic == 2) {
                    if (ty == -1) ty = 1;
                    if (ty == ny) ty = ny - 2;
                }
                d3 dm = mk3(0.0, 0.0, 0.0);
                if (dist_mod_p(g, tx, ty, tz, dm) == -1) continue;
                const uint32_t cid = (uint32_t)tx + (uint32_t)ty * g.num[0] +
                                     (g.num[0] * g.num[1]) * (uint32_t)tz;
                const uint32_t start = cell_start[cid];
                if (start == EMPTY_CELL) continue;
                const uint32_t end = cell_end[cid];
                for (uint32_t i = start; i < end; ++i) {
                    const double4 xr = x_s[i];
                    const double4 v = vel_s[i];
                    const d3 dist = (X - xyz(xr)) + dm;
                    double s = 0.0;
                    {
                        const double ax = fabs(dist.x), ay = fabs(dist.y), az = fabs(dist.z);
                        if (ax < dx && ay < dy && az < dz)
                            s = (1.0 - ax / dx) * (1.0 - ay / dy) * (1.0 - az / dz);
                    }
                    const double vol_p = sphere_volume(xr.w);
                    solid_volume += s * vol_p;
                    vp_avg_num = vp_avg_num + (s * vol_p) * xyz(v);
                    vp_avg_denum += s * vol_p;
                    const d3 xrn = mk3(xr.x + v.x * dt, xr.y + v.y * dt, xr.z + v.z * dt);
                    const d3 dist2 = (X - xrn) + dm;
                    double s2 = 0.0;
                    {
                        const double ax = fabs(dist2.x), ay = fabs(dist2.y), az = fabs(dist2.z);
                        if (ax < dx && ay < dy && az < dz)
                            s2 = (1.0 - ax / dx) * (1.0 - ay / dy) * (1.0 - az / dz);
                    }
                    solid_volume_new += s2 * vol_p;
                }
            }
    double cell_volume = dx * dy * dz;
    if (z == nz - 1) cell_volume *= 0.875;
    const double phi = fmin(0.9, fmax(0.1, 1.0 - solid_volume / cell_volume));
    const double phi_new = fmin(0.9, fmax(0.1, 1.0 - solid_volume_new / cell_volume));
    double dphi;
    d3 vp_avg;
    if (iteration == 0) {
        dphi = 0.0;
        vp_avg = mk3(0.0, 0.0, 0.0);
    } else {
        dphi = 0.5 * (phi_new - phi_0);
        vp_avg = vp_avg_num / fmax(1.0e-16, vp_avg_denum);
    }
    phi_a[cellidx] = phi * d.c_phi;
    dphi_a[cellidx] = dphi * d.c_phi;
    vp_avg_a[3 * (size_t)cellidx] = vp_avg.x;
    vp_avg_a[3 * (size_t)cellidx + 1] = vp_avg.y;
    vp_avg_a[3 * (size_t)cellidx + 2] = vp_avg.z;
}

__global__ void __launch_bounds__(256) k_porosities(StepArgs a, DarcyArgs d, long iteration,
                                                    const uint32_t *__restrict__ cell_start,
                                                    const uint32_t *__restrict__ cell_end,
                                                    const double4 *__restrict__ x_s,
                                                    const double4 *__restrict__ vel_s,
                                                    double *__restrict__ phi_a,
                                                    double *__restrict__ dphi_a,
                                                    double *__restrict__ vp_avg_a)
{
    porosities_body(a, d, iteration, cell_start, cell_end, x_s, vel_s, phi_a, dphi_a, vp_avg_a, blockIdx.x);
}

void launch_darcy_porosities(hipStream_t s, const StepArgs &a, const DarcyArgs &d,
                             const DarcyBuffers &b, const DemBuffers &db, long iter)
{
    hipLaunchKernelGGL(k_porosities, cell_grid(d), dim3(256), 0, s, a, d, iter, db.cell_start,
                       db.cell_end, db.x_s, db.vel_s, b.phi, b.dphi, b.vp_avg);
}

// copyDarcyPorositiesToEdges, src/darcy.cuh:556-590
__device__ __forceinline__ void por_edges_body(DarcyArgs d, double *__restrict__ phi,
                                                   double *__restrict__ dphi,
                                                   double *__restrict__ vp, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    const int ny = (int)d.ny;
    if (!(y == 0 || y == ny - 1)) return;
    const int y_read = y == 0 ? 1 : ny - 2;
    const uint32_t r = didx(d, x, y_read, z), w = didx(d, x, y, z);
    phi[w] = phi[r];
    dphi[w] = dphi[r];
    vp[3 * (size_t)w] = vp[3 * (size_t)r];
    vp[3 * (size_t)w + 1] = vp[3 * (size_t)r + 1];
    vp[3 * (size_t)w + 2] = vp[3 * (size_t)r + 2];
}

__global__ void __launch_bounds__(256) k_por_edges(DarcyArgs d, double *__restrict__ phi,
                                                   double *__restrict__ dphi,
                                                   double *__restrict__ vp)
{
    por_edges_body(d, phi, dphi, vp, blockIdx.x);
}

// copyDarcyPorositiesToBottom, src/darcy.cuh:595-628
__device__ __forceinline__ void por_bottom_body(DarcyArgs d, double *__restrict__ phi,
                                                    double *__restrict__ dphi,
                                                    double *__restrict__ vp, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    if (z != 0) return;
    const uint32_t r = didx(d, x, y, 1), w = didx(d, x, y, 0);
    phi[w] = phi[r];
    dphi[w] = dphi[r];
    vp[3 * (size_t)w] = vp[3 * (size_t)r];
    vp[3 * (size_t)w + 1] = vp[3 * (size_t)r + 1];
    vp[3 * (size_t)w + 2] = vp[3 * (size_t)r + 2];
}

__global__ void __launch_bounds__(256) k_por_bottom(DarcyArgs d, double *__restrict__ phi,
                                                    double *__restrict__ dphi,
                                                    double *__restrict__ vp)
{
    por_bottom_body(d, phi, dphi, vp, blockIdx.x);
}

void launch_darcy_porosities_to_edges(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b)
{
    hipLaunchKernelGGL(k_por_edges, cell_grid(d), dim3(256), 0, s, d, b.phi, b.dphi, b.vp_avg);
    hipLaunchKernelGGL(k_por_bottom, cell_grid(d), dim3(256), 0, s, d, b.phi, b.dphi, b.vp_avg);
}

// setDarcyTopPressure + setDarcyTopWallPressure, src/darcy.cuh:1195-1240,
// restricted to x < nx, y < ny, z < nz
__global__ void __launch_bounds__(256) k_top_pressure(DarcyArgs d, double *__restrict__ p,
                                                      double new_pressure)
{
    int x, y, z;
    if (!cell_xyz(d, x, y, z)) return;
    if (z == (int)d.nz - 1 || (uint32_t)z == d.wall0_iz) p[didx(d, x, y, z)] = new_pressure;
}

void launch_darcy_top_pressure(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b,
                               double new_pressure)
{
    hipLaunchKernelGGL(k_top_pressure, cell_grid(d), dim3(256), 0, s, d, b.p, new_pressure);
}

// findDarcyPermeabilities, src/darcy.cuh:1266-1313 (phi^3, (1-phi)^2 as products)
__device__ __forceinline__ void permeabilities_body(DarcyArgs d, const double *__restrict__ phi_a,
                                                        double *__restrict__ k_a, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    const uint32_t c = didx(d, x, y, z);
    double phi = phi_a[c];
    if (phi > 0.9999) phi = 0.9999;
    double k = d.k_c * (phi * phi * phi) / ((1.0 - phi) * (1.0 - phi));
    k = fmin(2.7e-10, k);
    k_a[c] = k;
}

__global__ void __launch_bounds__(256) k_permeabilities(DarcyArgs d, const double *__restrict__ phi_a,
                                                        double *__restrict__ k_a)
{
    permeabilities_body(d, phi_a, k_a, blockIdx.x);
}

void launch_darcy_permeabilities(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b)
{
    hipLaunchKernelGGL(k_permeabilities, cell_grid(d), dim3(256), 0, s, d, b.phi, b.k);
}

// copyValues, src/navierstokes.cuh:2794-2827
__device__ __forceinline__ void copy_interior_body(DarcyArgs d, const double *__restrict__ src,
                                                       double *__restrict__ dst, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    const uint32_t c = didx(d, x, y, z);
    dst[c] = src[c];
}

__global__ void __launch_bounds__(256) k_copy_interior(DarcyArgs d, const double *__restrict__ src,
                                                       double *__restrict__ dst)
{
    copy_interior_body(d, src, dst, blockIdx.x);
}

void launch_darcy_copy_interior(hipStream_t s, const DarcyArgs &d, const double *src, double *dst)
{
    hipLaunchKernelGGL(k_copy_interior, cell_grid(d), dim3(256), 0, s, d, src, dst);
}

// setDarcyZeros, src/darcy.cuh:172-187
__device__ __forceinline__ void zero_interior_body(DarcyArgs d, double *__restrict__ f, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    f[didx(d, x, y, z)] = 0.0;
}

__global__ void __launch_bounds__(256) k_zero_interior(DarcyArgs d, double *__restrict__ f)
{
    zero_interior_body(d, f, blockIdx.x);
}

void launch_darcy_zero_interior(hipStream_t s, const DarcyArgs &d, double *f)
{
    hipLaunchKernelGGL(k_zero_interior, cell_grid(d), dim3(256), 0, s, d, f);
}

// Pressure increment of firstDarcySolution / updateDarcySolution
// (src/darcy.cuh:1464-1580, 1646-1733) for cell (x, y, z) and iterate p.
__device__ __forceinline__ double darcy_dp(const DarcyArgs &d, int x, int y, int z,
                                           const double *__restrict__ p_a,
                                           const double *__restrict__ k_a,
                                           const double *__restrict__ phi_a,
                                           const double *__restrict__ dphi_a,
                                           const double *__restrict__ vp_a,
                                           const int *__restrict__ pconst, double &p_out)
{
    const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
    const double dx = d.dx, dy = d.dy, dz = d.dz;
    const uint32_t c = didx(d, x, y, z);
    const uint32_t cxn = didx(d, x - 1, y, z), cxp = didx(d, x + 1, y, z);
    const uint32_t cyn = didx(d, x, y - 1, z), cyp = didx(d, x, y + 1, z);
    const uint32_t czn = didx(d, x, y, z - 1), czp = didx(d, x, y, z + 1);
    const double phi_xn = phi_a[cxn], phi = phi_a[c], phi_xp = phi_a[cxp];
    const double phi_yn = phi_a[cyn], phi_yp = phi_a[cyp];
    const double phi_zn = phi_a[czn], phi_zp = phi_a[czp];
    const double dphi = dphi_a[c];
    const d3 vp_avg = mk3(vp_a[3 * (size_t)c], vp_a[3 * (size_t)c + 1], vp_a[3 * (size_t)c + 2]);
    const int p_constant = pconst[c];
    double p_xn = p_a[cxn];
    const double p = p_a[c];
    double p_xp = p_a[cxp], p_yn = p_a[cyn], p_yp = p_a[cyp], p_zn = p_a[czn], p_zp = p_a[czp];
    const double k_xn = k_a[cxn], k = k_a[c], k_xp = k_a[cxp];
    const double k_yn = k_a[cyn], k_yp = k_a[cyp], k_zn = k_a[czn], k_zp = k_a[czp];
    if (x == 0 && d.bc_xn == 1) p_xn = p;
    if (x == nx - 1 && d.bc_xp == 1) p_xp = p;
    if (y == 0 && d.bc_yn == 1) p_yn = p;
    if (y == ny - 1 && d.bc_yp == 1) p_yp = p;
    if (z == 0 && d.bc_bot == 1) p_zn = p;
    if (z == nz - 1 && d.bc_top == 1) p_zp = p;
    const d3 grad_phi = mk3((phi_xp - phi_xn) / (dx + dx), (phi_yp - phi_yn) / (dy + dy),
                            (phi_zp - phi_zn) / (dz + dz));
    const double div_k_grad_p =
        (2. * k_xp * k / (k_xp + k) * (p_xp - p) / dx - 2. * k_xn * k / (k_xn + k) * (p - p_xn) / dx) / dx +
        (2. * k_yp * k / (k_yp + k) * (p_yp - p) / dy - 2. * k_yn * k / (k_yn + k) * (p - p_yn) / dy) / dy +
        (2. * k_zp * k / (k_zp + k) * (p_zp - p) / dz - 2. * k_zn * k / (k_zn + k) * (p - p_zn) / dz) / dz;
    const double ndt = d.ndem * d.dt;
    double dp = ndt / (d.beta_f * phi * d.mu) * div_k_grad_p -
                (ndt / (d.beta_f * phi * (1.0 - phi))) * (dphi / ndt + dot3(vp_avg, grad_phi));
    // Dirichlet faces and the dynamic top wall (y == nx-1 as in the reference)
    if ((d.bc_bot == 0 && z == 0) || (d.bc_top == 0 && z == nz - 1) ||
        ((uint32_t)z >= d.wall0_iz && d.bc_top == 0) || (d.bc_xn == 0 && x == 0) ||
        (d.bc_xp == 0 && x == nx - 1) || (d.bc_yn == 0 && y == 0) ||
        (d.bc_yp == 0 && y == nx - 1) || p_constant == 1)
        dp = 0.0;
    p_out = p;
    return dp;
}

// Layout of DarcyBuffers::jc: face conductances on the x, y and z faces
// ((nx+1) ny nz, nx (ny+1) nz and nx ny (nz+1) values, face f of a row lying
// between cells f-1 and f, the ghost cells at -1 and n), then the per-cell
// A, F and base planes (nx ny nz each, x fastest).
struct JLayout {
    size_t hy, hz, A, F, base;
    uint32_t sxf, syf;  // row strides of the x- and y-face planes
};

__device__ __forceinline__ JLayout jlayout(const DarcyArgs &d)
{
    JLayout j;
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    j.hy = (size_t)(d.nx + 1) * d.ny * d.nz;
    j.hz = j.hy + (size_t)d.nx * (d.ny + 1) * d.nz;
    j.A = j.hz + (size_t)d.nx * d.ny * (d.nz + 1);
    j.F = j.A + n;
    j.base = j.F + n;
    j.sxf = d.nx + 1;
    j.syf = d.ny + 1;
    return j;
}

// firstDarcySolution (src/darcy.cuh:1425-1598) plus the constants of the
// solve's sweeps.  Within one solve k, phi, dphi, vp_avg, p_old and dp_expl
// are fixed, so each sweep of updateDarcySolution only needs the iterate
// p.  The sweep's face factor 2 k_nb k / (k_nb + k) is the same number seen
// from either cell of a face (2x is exact, and * and + commute in IEEE
// arithmetic), so it is stored once per face.  With
//   A = ndt / (beta_f phi mu),
//   F = ndt / (beta_f phi (1 - phi)) (dphi / ndt + vp_avg . grad phi),
//   base = p_old + (1 - epsilon) dp_expl,
// the sweep's dp_impl = A div - F and p* = base + epsilon dp_impl are the
// reference's expressions, operation for operation.
__device__ __forceinline__ void darcy_prepare_body(DarcyArgs d, const double *__restrict__ p_a,
                                                       const double *__restrict__ k_a,
                                                       const double *__restrict__ phi_a,
                                                       const double *__restrict__ dphi_a,
                                                       const double *__restrict__ vp_a,
                                                       const int *__restrict__ pconst,
                                                       const double *__restrict__ p_old,
                                                       double *__restrict__ dp_expl,
                                                       double *__restrict__ jc,
                                                       uint8_t *__restrict__ frozen, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    const uint32_t i = bid * blockDim.x + threadIdx.x;
    const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
    const double dx = d.dx, dy = d.dy, dz = d.dz;
    const JLayout j = jlayout(d);
    double pc;
    const double dpe = darcy_dp(d, x, y, z, p_a, k_a, phi_a, dphi_a, vp_a, pconst, pc);
    const uint32_t c = didx(d, x, y, z);
    dp_expl[c] = dpe;
    const uint32_t cxn = didx(d, x - 1, y, z), cxp = didx(d, x + 1, y, z);
    const uint32_t cyn = didx(d, x, y - 1, z), cyp = didx(d, x, y + 1, z);
    const uint32_t czn = didx(d, x, y, z - 1), czp = didx(d, x, y, z + 1);
    const double k = k_a[c];
    const size_t fx = (size_t)x + j.sxf * ((size_t)y + d.ny * (size_t)z);
    const size_t fy = (size_t)x + d.nx * ((size_t)y + j.syf * (size_t)z);
    jc[fx] = 2. * k_a[cxn] * k / (k_a[cxn] + k);
    if (x == nx - 1) jc[fx + 1] = 2. * k_a[cxp] * k / (k_a[cxp] + k);
    jc[j.hy + fy] = 2. * k_a[cyn] * k / (k_a[cyn] + k);
    if (y == ny - 1) jc[j.hy + fy + d.nx] = 2. * k_a[cyp] * k / (k_a[cyp] + k);
    jc[j.hz + i] = 2. * k_a[czn] * k / (k_a[czn] + k);
    if (z == nz - 1) jc[j.hz + i + (size_t)d.nx * d.ny] = 2. * k_a[czp] * k / (k_a[czp] + k);
    const double phi = phi_a[c];
    const d3 grad_phi = mk3((phi_a[cxp] - phi_a[cxn]) / (dx + dx), (phi_a[cyp] - phi_a[cyn]) / (dy + dy),
                            (phi_a[czp] - phi_a[czn]) / (dz + dz));
    const d3 vp_avg = mk3(vp_a[3 * (size_t)c], vp_a[3 * (size_t)c + 1], vp_a[3 * (size_t)c + 2]);
    const double ndt = d.ndem * d.dt;
    jc[j.A + i] = ndt / (d.beta_f * phi * d.mu);
    jc[j.F + i] = (ndt / (d.beta_f * phi * (1.0 - phi))) * (dphi_a[c] / ndt + dot3(vp_avg, grad_phi));
    const double epsilon = 0.5;
    jc[j.base + i] = p_old[c] + (1.0 - epsilon) * dpe;
    frozen[i] = ((d.bc_bot == 0 && z == 0) || (d.bc_top == 0 && z == nz - 1) ||
                 ((uint32_t)z >= d.wall0_iz && d.bc_top == 0) || (d.bc_xn == 0 && x == 0) ||
                 (d.bc_xp == 0 && x == nx - 1) || (d.bc_yn == 0 && y == 0) ||
                 (d.bc_yp == 0 && y == nx - 1) || pconst[c] == 1) ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_darcy_prepare(DarcyArgs d, const double *__restrict__ p_a,
                                                       const double *__restrict__ k_a,
                                                       const double *__restrict__ phi_a,
                                                       const double *__restrict__ dphi_a,
                                                       const double *__restrict__ vp_a,
                                                       const int *__restrict__ pconst,
                                                       const double *__restrict__ p_old,
                                                       double *__restrict__ dp_expl,
                                                       double *__restrict__ jc,
                                                       uint8_t *__restrict__ frozen)
{
    darcy_prepare_body(d, p_a, k_a, phi_a, dphi_a, vp_a, pconst, p_old, dp_expl, jc, frozen, blockIdx.x);
}

void launch_darcy_prepare(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b)
{
    hipLaunchKernelGGL(k_darcy_prepare, cell_grid(d), dim3(256), 0, s, d, b.p, b.k, b.phi, b.dphi,
                       b.vp_avg, b.p_constant, b.p_old, b.dp_expl, b.jc, b.jfrozen);
}

// The convergence test of the Jacobi loop (src/device.cu:2342-2386,
// maxNormResDarcy src/darcy.cpp:266-292) for check sweep nijac: a NaN max
// ends the solve with an error, max <= tolerance ends it converged.
// state: [0] 0 running, 1 converged, 2 NaN; [1] nijac of the last check;
// [2] bit pattern of its max |norm|.  Returns whether the solve ends.
// The max of a reduction slot's RED_LANES words.  Every lane of the wave
// calls red_max_wave (one word per lane, RED_LANES being the wave size).
__device__ __forceinline__ unsigned long long red_max_wave(const unsigned long long *__restrict__ slot)
{
    unsigned long long v = slot[(threadIdx.x & 63u) * RED_STRIDE];
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long red_max_serial(const unsigned long long *__restrict__ slot)
{
    unsigned long long v = 0ull;
    for (int j = 0; j < RED_LANES; ++j) {
        const unsigned long long o = slot[j * RED_STRIDE];
        v = o > v ? o : v;
    }
    return v;
}

// block bid's share of a launch-wide max
__device__ __forceinline__ void red_fold(unsigned long long *__restrict__ slot, uint32_t bid, unsigned long long m)
{
    atomicMax(slot + (bid % RED_LANES) * RED_STRIDE, m);
}

__device__ __forceinline__ bool darcy_decide_bits(unsigned long long bits, double tolerance, uint32_t nijac,
                                                  unsigned long long *__restrict__ state, bool writer)
{
    const double m = __longlong_as_double((long long)bits);
    const bool nan = m != m;
    const bool stop = nan || m <= tolerance;
    if (writer) {
        state[1] = nijac;
        state[2] = bits;
        if (stop) state[0] = nan ? 2ull : 1ull;
    }
    return stop;
}

// the test of a check sweep from its slot; every lane of the wave calls it
__device__ __forceinline__ bool darcy_decide(const unsigned long long *__restrict__ slot, double tolerance,
                                             uint32_t nijac, unsigned long long *__restrict__ state, bool writer)
{
    return darcy_decide_bits(red_max_wave(slot), tolerance, nijac, state, writer);
}

// p* of updateDarcySolution (src/darcy.cuh:1646-1739) for interior cell i
// = (x, y, z) (ghost-layer index c) of iterate p_in, with the solve's
// constants: p* = base + epsilon (A div(k grad p) - F), dp_impl held at 0
// where frozen.  The sweep relaxes p toward it by theta.
__device__ __forceinline__ double jacobi_target(const DarcyArgs &d, const double *__restrict__ jc,
                                                const uint8_t *__restrict__ frozen,
                                                const double *__restrict__ p_in, uint32_t i, int x, int y,
                                                int z, uint32_t c)
{
    const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
    const JLayout j = jlayout(d);
    const uint32_t sy = d.nx + 2, sz = (d.nx + 2) * (d.ny + 2);
    const double p = p_in[c];
    double p_xn = p_in[c - 1], p_xp = p_in[c + 1];
    double p_yn = p_in[c - sy], p_yp = p_in[c + sy];
    double p_zn = p_in[c - sz], p_zp = p_in[c + sz];
    if (x == 0 && d.bc_xn == 1) p_xn = p;
    if (x == nx - 1 && d.bc_xp == 1) p_xp = p;
    if (y == 0 && d.bc_yn == 1) p_yn = p;
    if (y == ny - 1 && d.bc_yp == 1) p_yp = p;
    if (z == 0 && d.bc_bot == 1) p_zn = p;
    if (z == nz - 1 && d.bc_top == 1) p_zp = p;
    const double dx = d.dx, dy = d.dy, dz = d.dz;
    const size_t fx = (size_t)x + j.sxf * ((size_t)y + d.ny * (size_t)z);
    const size_t fy = j.hy + (size_t)x + d.nx * ((size_t)y + j.syf * (size_t)z);
    const double h_xn = jc[fx], h_xp = jc[fx + 1];
    const double h_yn = jc[fy], h_yp = jc[fy + d.nx];
    const double h_zn = jc[j.hz + i], h_zp = jc[j.hz + i + (size_t)d.nx * d.ny];
    const double div_k_grad_p = (h_xp * (p_xp - p) / dx - h_xn * (p - p_xn) / dx) / dx +
                                (h_yp * (p_yp - p) / dy - h_yn * (p - p_yn) / dy) / dy +
                                (h_zp * (p_zp - p) / dz - h_zn * (p - p_zn) / dz) / dz;
    double dp_impl = jc[j.A + i] * div_k_grad_p - jc[j.F + i];
    if (frozen[i]) dp_impl = 0.0;
    const double epsilon = 0.5;
    return jc[j.base + i] + epsilon * dp_impl;
}

// One sweep of the Jacobi loop body (src/device.cu:2184-2331) in a single
// launch: updateDarcySolution on iterate p_in into p_out, then, into p_out,
// the ghost values that the next sweep's setDarcyGhostNodes would copy from
// p_out's interior (src/darcy.cuh:194-264) and the flux ghosts that
// setDarcyGhostNodesFlux computes from p_in's interior
// (src/darcy.cuh:270-312; the reference writes them after the update and
// before copying p_new into p, so they lag one sweep).  Ping-ponging p_in
// and p_out therefore leaves exactly the state the reference's
// ghosts / update / flux ghosts / copy sequence leaves.  The residual
// field itself is not stored, only its max |.| for the convergence check
// (check != null): a per-block max of the bit pattern of |norm| (monotone
// for non-negative doubles; NaN patterns sort above +inf, so a NaN residual
// survives the max) folded with one atomic max per block; a check sweep
// also zeroes the slot of the next check (clear).  The sweep after a check
// sweep first applies that check's test (prev != null; every block reaches
// the same verdict from the same slot) and returns if the solve has ended;
// once state[0] is set, every later sweep returns at once, so the host may
// enqueue sweeps beyond the convergence point.
#define SWEEP_THREADS 512
__device__ __forceinline__ void darcy_sweep_body(DarcyArgs d, const double *__restrict__ jc,
                                                     const uint8_t *__restrict__ frozen,
                                                     const double *__restrict__ p_in,
                                                     const double *__restrict__ k_a,
                                                     double *__restrict__ p_out,
                                                     unsigned long long *__restrict__ check,
                                                     unsigned long long *__restrict__ clear,
                                                     const unsigned long long *__restrict__ prev,
                                                     double tolerance, uint32_t prev_nijac,
                                                     unsigned long long *__restrict__ state,
                                                     unsigned long long *__restrict__ zero_scale, uint32_t bid)
{
    __shared__ unsigned long long wmax[SWEEP_THREADS / 64];
    const uint32_t n = d.nx * d.ny * d.nz;
    // each XCD sweeps one contiguous slab of cells, so that the stencil's
    // neighbouring rows and planes are mostly in that XCD's L2
    const uint32_t i = xcd_tile(bid, (n + SWEEP_THREADS - 1) / SWEEP_THREADS) * SWEEP_THREADS + threadIdx.x;
    // the scale slot of the correction that follows
    if (zero_scale && i < RED_LANES) zero_scale[i * RED_STRIDE] = 0ull;
    if (state[0] != 0ull) return;
    if (prev && darcy_decide(prev, tolerance, prev_nijac, state, i == 0)) return;
    unsigned long long bits = 0;
    if (clear && i < RED_LANES) clear[i * RED_STRIDE] = 0ull;
    if (i < n) {
        const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
        const int x = (int)(i % d.nx), y = (int)((i / d.nx) % d.ny), z = (int)(i / (d.nx * d.ny));
        const uint32_t c = didx(d, x, y, z);
        const double p = p_in[c];
        double p_new = jacobi_target(d, jc, frozen, p_in, i, x, y, z, c);
        const double theta = 0.05;
        p_new = p * (1.0 - theta) + p_new * theta;
        const double res_norm = (p_new - p) / (p + 1.0e-16);
        p_out[c] = p_new;
#define GW(X, Y, Z) p_out[didx(d, X, Y, Z)] = p_new
        if (x == 0 && d.bc_xn == 0) GW(-1, y, z);
        if (x == 1 && d.bc_xn == 1) GW(-1, y, z);
        if (x == 0 && d.bc_xn == 2) GW(nx, y, z);
        if (x == nx - 1 && d.bc_xp == 0) GW(nx, y, z);
        if (x == nx - 2 && d.bc_xp == 1) GW(nx, y, z);
        if (x == nx - 1 && d.bc_xp == 2) GW(-1, y, z);
        if (y == 0 && d.bc_yn == 0) GW(x, -1, z);
        if (y == 1 && d.bc_yn == 1) GW(x, -1, z);
        if (y == 0 && d.bc_yn == 2) GW(x, ny, z);
        if (y == ny - 1 && d.bc_yp == 0) GW(x, ny, z);
        if (y == ny - 2 && d.bc_yp == 1) GW(x, ny, z);
        if (y == ny - 1 && d.bc_yp == 2) GW(x, -1, z);
        if (z == 0 && d.bc_bot == 0) GW(x, y, -1);
        if (z == 1 && d.bc_bot == 1) GW(x, y, -1);
        if (z == 0 && d.bc_bot == 2) GW(x, y, nz);
        if (z == nz - 1 && d.bc_top == 0) GW(x, y, nz);
        if (z == nz - 2 && d.bc_top == 1) GW(x, y, nz);
        if (z == nz - 1 && d.bc_top == 2) GW(x, y, -1);
#undef GW
        if ((z == 0 && d.bc_bot == 4) || (z == nz - 1 && d.bc_top == 4)) {
            const double q_z = z == 0 ? d.bc_bot_flux : d.bc_top_flux;
            const double p_ghost = -d.mu / k_a[c] * q_z * (d.L2 / nz) + p;
            if (z == 0 && d.bc_bot == 4) p_out[didx(d, x, y, -1)] = p_ghost;
            if (z == nz - 1 && d.bc_top == 4) p_out[didx(d, x, y, nz)] = p_ghost;
        }
        bits = (unsigned long long)__double_as_longlong(fabs(res_norm));
    }
    if (check) {
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_down(bits, off);
            bits = o > bits ? o : bits;
        }
        if ((threadIdx.x & 63u) == 0) wmax[threadIdx.x >> 6] = bits;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = wmax[0];
            for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = wmax[w] > m ? wmax[w] : m;
            red_fold(check, bid, m);
        }
    }
}

__global__ void __launch_bounds__(SWEEP_THREADS) k_darcy_sweep(DarcyArgs d, const double *__restrict__ jc,
                                                     const uint8_t *__restrict__ frozen,
                                                     const double *__restrict__ p_in,
                                                     const double *__restrict__ k_a,
                                                     double *__restrict__ p_out,
                                                     unsigned long long *__restrict__ check,
                                                     unsigned long long *__restrict__ clear,
                                                     const unsigned long long *__restrict__ prev,
                                                     double tolerance, uint32_t prev_nijac,
                                                     unsigned long long *__restrict__ state,
                                                     unsigned long long *__restrict__ zero_scale)
{
    darcy_sweep_body(d, jc, frozen, p_in, k_a, p_out, check, clear, prev, tolerance, prev_nijac, state, zero_scale, blockIdx.x);
}

void launch_darcy_sweep(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b,
                        const double *p_in, double *p_out, unsigned long long *check,
                        unsigned long long *clear, const unsigned long long *prev,
                        double tolerance, uint32_t prev_nijac, bool free_running, bool zero_scale)
{
    // 512-thread blocks, each XCD on one slab: 5.7-5.9 us per sweep on the
    // 48 x 48 x 96 grid (6.7 with the blocks dealt round-robin), against
    // 6.2 with 256 threads, 5.8 with 1024, and 6.5-8.7 with 2 or 4 cells per
    // thread
    const uint32_t n = d.nx * d.ny * d.nz;
    hipLaunchKernelGGL(k_darcy_sweep, dim3((n + SWEEP_THREADS - 1) / SWEEP_THREADS), dim3(SWEEP_THREADS), 0, s,
                       d, b.jc, b.jfrozen, p_in,
                       b.k, p_out, check, clear, prev, tolerance, prev_nijac,
                       b.maxres + (free_running ? DARCY_ZERO : DARCY_STATE),
                       zero_scale ? b.maxres + DARCY_SCALE : (unsigned long long *)nullptr);
}

// The test of a check sweep that no sweep follows (the last of maxiter).
__global__ void k_darcy_decide(const unsigned long long *__restrict__ slot, double tolerance,
                               uint32_t nijac, unsigned long long *__restrict__ state)
{
    if (state[0] != 0ull) return;
    (void)darcy_decide(slot, tolerance, nijac, state, threadIdx.x == 0);
}

void launch_darcy_decide(hipStream_t s, const DarcyBuffers &b, const unsigned long long *slot,
                         double tolerance, uint32_t nijac)
{
    hipLaunchKernelGGL(k_darcy_decide, dim3(1), dim3(64), 0, s, slot, tolerance, nijac,
                       b.maxres + DARCY_STATE);
}

// ---------------------------------------------------------------------------
// Learned correction between sweeps (no counterpart in the reference).  For
// iterate p the fixed-point residual r = p*(p) - p, with p* as above, and
// the error e = p_fixed - p satisfy (I - epsilon A L) e = r, L the
// conductance-weighted Laplacian; the network maps (r / s, log(1 + D),
// frozen) to e / s, s = max |r|, D = epsilon A sum_faces h / dx^2 the
// cell's coupling.  Every kernel below returns at once when state != null
// and state[0] != 0 (the solve has ended), so corrections can be enqueued
// ahead of the convergence decision.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void darcy_residual_max_body(DarcyArgs d, const double *__restrict__ jc,
                                                            const uint8_t *__restrict__ frozen,
                                                            const double *__restrict__ p,
                                                            const unsigned long long *__restrict__ state,
                                                            unsigned long long *__restrict__ smax, uint32_t bid)
{
    __shared__ unsigned long long wmax[4];
    if (state && state[0] != 0ull) return;
    const uint32_t n = d.nx * d.ny * d.nz;
    const uint32_t i = bid * blockDim.x + threadIdx.x;
    unsigned long long bits = 0;
    if (i < n) {
        const int x = (int)(i % d.nx), y = (int)((i / d.nx) % d.ny), z = (int)(i / (d.nx * d.ny));
        const uint32_t c = didx(d, x, y, z);
        const double r = jacobi_target(d, jc, frozen, p, i, x, y, z, c) - p[c];
        bits = (unsigned long long)__double_as_longlong(fabs(r));
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_down(bits, off);
        bits = o > bits ? o : bits;
    }
    if ((threadIdx.x & 63u) == 0) wmax[threadIdx.x >> 6] = bits;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = wmax[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = wmax[w] > m ? wmax[w] : m;
        red_fold(smax, bid, m);
    }
}

__global__ void __launch_bounds__(256) k_darcy_residual_max(DarcyArgs d, const double *__restrict__ jc,
                                                            const uint8_t *__restrict__ frozen,
                                                            const double *__restrict__ p,
                                                            const unsigned long long *__restrict__ state,
                                                            unsigned long long *__restrict__ smax)
{
    darcy_residual_max_body(d, jc, frozen, p, state, smax, blockIdx.x);
}

__device__ __forceinline__ bool usable_scale(unsigned long long bits, double &s)
{
    s = __longlong_as_double((long long)bits);
    return s > 0.0 && s < 1.0e300;  // false for 0, inf and NaN
}

// network input [V][3] (x fastest) for iterate p
__device__ __forceinline__ void darcy_net_input_body(DarcyArgs d, const double *__restrict__ jc,
                                                         const uint8_t *__restrict__ frozen,
                                                         const double *__restrict__ p,
                                                         const unsigned long long *__restrict__ state,
                                                         const unsigned long long *__restrict__ smax,
                                                         float *__restrict__ input,
                                                         uint16_t *__restrict__ in4, uint32_t bid)
{
    if (state && state[0] != 0ull) return;
    const unsigned long long sbits = red_max_wave(smax);
    const uint32_t n = d.nx * d.ny * d.nz;
    const uint32_t i = bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % d.nx), y = (int)((i / d.nx) % d.ny), z = (int)(i / (d.nx * d.ny));
    const uint32_t c = didx(d, x, y, z);
    double s;
    const bool ok = usable_scale(sbits, s);
    const double r = jacobi_target(d, jc, frozen, p, i, x, y, z, c) - p[c];
    const JLayout j = jlayout(d);
    const size_t fx = (size_t)x + j.sxf * ((size_t)y + d.ny * (size_t)z);
    const size_t fy = j.hy + (size_t)x + d.nx * ((size_t)y + j.syf * (size_t)z);
    const double hsum = (jc[fx] + jc[fx + 1]) / (d.dx * d.dx) + (jc[fy] + jc[fy + d.nx]) / (d.dy * d.dy) +
                        (jc[j.hz + i] + jc[j.hz + i + (size_t)d.nx * d.ny]) / (d.dz * d.dz);
    const double D = 0.5 * jc[j.A + i] * hsum;
    const bool fz = frozen[i] != 0;
    const float v0 = ok ? (float)(r / s) : 0.f;
    const float v1 = (fz || !(D >= 0.0)) ? 0.f : (float)log1p(D);
    const float v2 = fz ? 1.f : 0.f;
    if (input) {
        input[3 * (size_t)i] = v0;
        input[3 * (size_t)i + 1] = v1;
        input[3 * (size_t)i + 2] = v2;
    }
    if (in4) {  // the network's own input layout: BF16, padded to 4 channels
        in4[4 * (size_t)i] = f2bf(v0);
        in4[4 * (size_t)i + 1] = f2bf(v1);
        in4[4 * (size_t)i + 2] = f2bf(v2);
        in4[4 * (size_t)i + 3] = 0;
    }
}

__global__ void __launch_bounds__(256) k_darcy_net_input(DarcyArgs d, const double *__restrict__ jc,
                                                         const uint8_t *__restrict__ frozen,
                                                         const double *__restrict__ p,
                                                         const unsigned long long *__restrict__ state,
                                                         const unsigned long long *__restrict__ smax,
                                                         float *__restrict__ input,
                                                         uint16_t *__restrict__ in4)
{
    darcy_net_input_body(d, jc, frozen, p, state, smax, input, in4, blockIdx.x);
}

// p += s e on the cells that are not held fixed; counts applied corrections
// in state[3]
__device__ __forceinline__ void darcy_net_apply_body(DarcyArgs d, const uint8_t *__restrict__ frozen,
                                                         double *__restrict__ p,
                                                         const float *__restrict__ e,
                                                         const unsigned long long *__restrict__ smax,
                                                         unsigned long long *__restrict__ state, uint32_t bid)
{
    if (state[0] != 0ull) return;
    double s;
    if (!usable_scale(red_max_wave(smax), s)) return;
    const uint32_t n = d.nx * d.ny * d.nz;
    const uint32_t i = bid * blockDim.x + threadIdx.x;
    if (i == 0) state[3] += 1ull;
    if (i >= n || frozen[i]) return;
    const float ev = e[i];
    if (!(ev == ev)) return;
    const int x = (int)(i % d.nx), y = (int)((i / d.nx) % d.ny), z = (int)(i / (d.nx * d.ny));
    double &pc = p[didx(d, x, y, z)];
    pc += s * (double)ev;
    // the ghost cells that copy this cell (k_ghosts' rules for a scalar
    // field); the other ghosts already hold the unchanged cells' values
    const double v = pc;
    const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
#define GW(X, Y, Z) p[didx(d, X, Y, Z)] = v
    if (x == 0 && d.bc_xn == 0) GW(-1, y, z);
    if (x == 1 && d.bc_xn == 1) GW(-1, y, z);
    if (x == 0 && d.bc_xn == 2) GW(nx, y, z);
    if (x == nx - 1 && d.bc_xp == 0) GW(nx, y, z);
    if (x == nx - 2 && d.bc_xp == 1) GW(nx, y, z);
    if (x == nx - 1 && d.bc_xp == 2) GW(-1, y, z);
    if (y == 0 && d.bc_yn == 0) GW(x, -1, z);
    if (y == 1 && d.bc_yn == 1) GW(x, -1, z);
    if (y == 0 && d.bc_yn == 2) GW(x, ny, z);
    if (y == ny - 1 && d.bc_yp == 0) GW(x, ny, z);
    if (y == ny - 2 && d.bc_yp == 1) GW(x, ny, z);
    if (y == ny - 1 && d.bc_yp == 2) GW(x, -1, z);
    if (z == 0 && d.bc_bot == 0) GW(x, y, -1);
    if (z == 1 && d.bc_bot == 1) GW(x, y, -1);
    if (z == 0 && d.bc_bot == 2) GW(x, y, nz);
    if (z == nz - 1 && d.bc_top == 0) GW(x, y, nz);
    if (z == nz - 2 && d.bc_top == 1) GW(x, y, nz);
    if (z == nz - 1 && d.bc_top == 2) GW(x, y, -1);
#undef GW
}

__global__ void __launch_bounds__(256) k_darcy_net_apply(DarcyArgs d, const uint8_t *__restrict__ frozen,
                                                         double *__restrict__ p,
                                                         const float *__restrict__ e,
                                                         const unsigned long long *__restrict__ smax,
                                                         unsigned long long *__restrict__ state)
{
    darcy_net_apply_body(d, frozen, p, e, smax, state, blockIdx.x);
}

void launch_darcy_residual_max(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b, const double *p,
                               bool free_running)
{
    hipLaunchKernelGGL(k_darcy_residual_max, cell_grid(d), dim3(256), 0, s, d, b.jc, b.jfrozen, p,
                       free_running ? (const unsigned long long *)nullptr : b.maxres + DARCY_STATE,
                       b.maxres + DARCY_SCALE);
}

void launch_darcy_net_input(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b, const double *p,
                            float *input, bool free_running, uint16_t *in4)
{
    hipLaunchKernelGGL(k_darcy_net_input, cell_grid(d), dim3(256), 0, s, d, b.jc, b.jfrozen, p,
                       free_running ? (const unsigned long long *)nullptr : b.maxres + DARCY_STATE,
                       b.maxres + DARCY_SCALE, input, in4);
}

void launch_darcy_net_apply(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b, double *p,
                            const float *e)
{
    hipLaunchKernelGGL(k_darcy_net_apply, cell_grid(d), dim3(256), 0, s, d, b.jfrozen, p, e,
                       b.maxres + DARCY_SCALE, b.maxres + DARCY_STATE);
}

// findDarcyVelocities, src/darcy.cuh:1839-1897
__device__ __forceinline__ void darcy_velocities_body(DarcyArgs d, const double *__restrict__ p,
                                                          const double *__restrict__ phi_a,
                                                          const double *__restrict__ k_a,
                                                          double *__restrict__ v_a, uint32_t bid)
{
    int x, y, z;
    if (!cell_xyz_b(d, bid, x, y, z)) return;
    const uint32_t c = didx(d, x, y, z);
    const double p_xn = p[didx(d, x - 1, y, z)], p_xp = p[didx(d, x + 1, y, z)];
    const double p_yn = p[didx(d, x, y - 1, z)], p_yp = p[didx(d, x, y + 1, z)];
    const double p_zn = p[didx(d, x, y, z - 1)], p_zp = p[didx(d, x, y, z + 1)];
    const double k = k_a[c], phi = phi_a[c];
    const d3 gp = mk3((p_xp - p_xn) / (d.dx + d.dx), (p_yp - p_yn) / (d.dy + d.dy),
                      (p_zp - p_zn) / (d.dz + d.dz));
    const d3 v = ((-k / d.mu) * gp) / phi;
    v_a[3 * (size_t)c] = v.x;
    v_a[3 * (size_t)c + 1] = v.y;
    v_a[3 * (size_t)c + 2] = v.z;
}

__global__ void __launch_bounds__(256) k_darcy_velocities(DarcyArgs d, const double *__restrict__ p,
                                                          const double *__restrict__ phi_a,
                                                          const double *__restrict__ k_a,
                                                          double *__restrict__ v_a)
{
    darcy_velocities_body(d, p, phi_a, k_a, v_a, blockIdx.x);
}

void launch_darcy_velocities(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b)
{
    hipLaunchKernelGGL(k_darcy_velocities, cell_grid(d), dim3(256), 0, s, d, b.p, b.phi, b.k, b.v);
}

// ---------------------------------------------------------------------------
// Multi-box launches (ensembles of fluid boxes): every stage of a fluid step
// is one launch for all boxes; a block finds its box (multi_block) and runs
// the body of the single-box kernel on that box's task record.
// ---------------------------------------------------------------------------
uint32_t darcy_cell_blocks(const DarcyArgs &d) { return (d.nx * d.ny * d.nz + 255u) / 256u; }

uint32_t darcy_sweep_blocks(const DarcyArgs &d)
{
    return (d.nx * d.ny * d.nz + SWEEP_THREADS - 1u) / SWEEP_THREADS;
}

uint32_t darcy_all_cell_blocks(const DarcyArgs &d)
{
    return ((d.nx + 2u) * (d.ny + 2u) * (d.nz + 2u) + 255u) / 256u;
}

#define DM_BOX                                   \
    uint32_t k_, bid, nblk_;                     \
    if (!multi_block(m, k_, bid, nblk_)) return; \
    const DarcyTask &T = t[k_]

__global__ void __launch_bounds__(256) k_dm_ghosts(const DarcyTask *__restrict__ t, MultiMap m, int field)
{
    DM_BOX;
    double *f = field == DM_P ? T.b.p : field == DM_GRAD_P ? T.b.grad_p : field == DM_PHI ? T.b.phi : T.b.k;
    ghosts_body(T.d, f, field == DM_GRAD_P ? 3 : 1, bid);
}

__global__ void __launch_bounds__(256) k_dm_ghosts_flux(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    ghosts_flux_body(T.d, T.b.p, T.b.k, bid);
}

__global__ void __launch_bounds__(256) k_dm_pressure_gradient(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    pressure_gradient_body(T.d, T.b.p, T.b.grad_p, bid);
}

__global__ void __launch_bounds__(256) k_dm_pressure_force(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    pressure_force_body(T.a, T.d, T.x, T.b.grad_p, T.force, T.b.f_p, T.err, bid);
}

// the iteration count of the step from the box's device clock (the host's
// value in a single-box step)
__global__ void __launch_bounds__(256) k_dm_porosities(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    porosities_body(T.a, T.d, (long)T.clock->iter, T.cell_start, T.cell_end, T.x_s, T.vel_s, T.b.phi,
                    T.b.dphi, T.b.vp_avg, bid);
}

__global__ void __launch_bounds__(256) k_dm_por_edges(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    por_edges_body(T.d, T.b.phi, T.b.dphi, T.b.vp_avg, bid);
}

__global__ void __launch_bounds__(256) k_dm_por_bottom(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    por_bottom_body(T.d, T.b.phi, T.b.dphi, T.b.vp_avg, bid);
}

__global__ void __launch_bounds__(256) k_dm_permeabilities(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    permeabilities_body(T.d, T.b.phi, T.b.k, bid);
}

__global__ void __launch_bounds__(256) k_dm_solve_start(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    copy_interior_body(T.d, T.b.p, T.b.p_old, bid);
}

__global__ void __launch_bounds__(256) k_dm_prepare(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    darcy_prepare_body(T.d, T.b.p, T.b.k, T.b.phi, T.b.dphi, T.b.vp_avg, T.b.p_constant, T.b.p_old,
                       T.b.dp_expl, T.b.jc, T.b.jfrozen, bid);
}

__global__ void k_dm_reset(const DarcyTask *__restrict__ t)
{
    for (uint32_t w = threadIdx.x; w < DARCY_MAXRES_WORDS; w += blockDim.x) t[blockIdx.x].b.maxres[w] = 0ull;
}

// The iterate buffers and the check slots of sweep nijac, as the host loop
// of a single-box solve passes them (driver.hip, darcy_step).
__global__ void __launch_bounds__(SWEEP_THREADS) k_dm_sweep(const DarcyTask *__restrict__ t, MultiMap m,
                                                            uint32_t nijac, int correct)
{
    DM_BOX;
    const bool even = (nijac & 1u) == 0u;
    unsigned long long *mr = T.b.maxres;
    const bool check = nijac % 10u == 0u;
    const bool decides = nijac > 0u && (nijac - 1u) % 10u == 0u;
    darcy_sweep_body(T.d, T.b.jc, T.b.jfrozen, even ? T.b.p : T.b.p_new, T.b.k, even ? T.b.p_new : T.b.p,
                     check ? mr + ((nijac / 10u) & 1u) * RED_WORDS : nullptr,
                     check ? mr + (((nijac / 10u) & 1u) ^ 1u) * RED_WORDS : nullptr,
                     decides ? mr + (((nijac - 1u) / 10u) & 1u) * RED_WORDS : nullptr, T.tolerance, nijac - 1u,
                     mr + DARCY_STATE, (correct && T.net_in4) ? mr + DARCY_SCALE : nullptr, bid);
}

__global__ void k_dm_decide(const DarcyTask *__restrict__ t, uint32_t n, uint32_t nijac)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const DarcyTask &T = t[k];
    unsigned long long *state = T.b.maxres + DARCY_STATE;
    if (state[0] != 0ull) return;
    (void)darcy_decide_bits(red_max_serial(T.b.maxres + ((nijac / 10u) & 1u) * RED_WORDS), T.tolerance, nijac,
                            state, true);
}

// the iterate that deciding sweep nijac wrote
__device__ __forceinline__ double *dm_iterate(const DarcyTask &T, uint32_t nijac)
{
    return (nijac & 1u) ? T.b.p : T.b.p_new;
}

__global__ void __launch_bounds__(256) k_dm_residual_max(const DarcyTask *__restrict__ t, MultiMap m,
                                                         uint32_t nijac)
{
    DM_BOX;
    darcy_residual_max_body(T.d, T.b.jc, T.b.jfrozen, dm_iterate(T, nijac), T.b.maxres + DARCY_STATE,
                            T.b.maxres + DARCY_SCALE, bid);
}

__global__ void __launch_bounds__(256) k_dm_net_input(const DarcyTask *__restrict__ t, MultiMap m,
                                                      uint32_t nijac)
{
    DM_BOX;
    darcy_net_input_body(T.d, T.b.jc, T.b.jfrozen, dm_iterate(T, nijac), T.b.maxres + DARCY_STATE,
                         T.b.maxres + DARCY_SCALE, nullptr, T.net_in4, bid);
}

__global__ void __launch_bounds__(256) k_dm_net_apply(const DarcyTask *__restrict__ t, MultiMap m,
                                                      uint32_t nijac)
{
    DM_BOX;
    darcy_net_apply_body(T.d, T.b.jfrozen, dm_iterate(T, nijac), T.net_out, T.b.maxres + DARCY_SCALE,
                         T.b.maxres + DARCY_STATE, bid);
}

__global__ void k_dm_states(const DarcyTask *__restrict__ t, uint32_t n, unsigned long long *__restrict__ out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    for (int j = 0; j < 4; ++j) out[4 * k + j] = t[k].b.maxres[DARCY_STATE + j];
}

// The sweeps that ran (up to the deciding check, or all maxiter) leave the
// solution in b.p_new when their number is odd; it is copied into b.p (every
// cell, ghosts included), so that every solve starts from b.p.
__global__ void __launch_bounds__(256) k_dm_normalize(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    const unsigned long long *state = T.b.maxres + DARCY_STATE;
    const unsigned long long ran = state[0] != 0ull ? state[1] + 1ull : (unsigned long long)T.maxiter;
    if ((ran & 1ull) == 0ull) return;
    const uint32_t c = bid * blockDim.x + threadIdx.x;
    if (c < (T.d.nx + 2u) * (T.d.ny + 2u) * (T.d.nz + 2u)) T.b.p[c] = T.b.p_new[c];
}

__global__ void __launch_bounds__(256) k_dm_finish(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    zero_interior_body(T.d, T.b.dphi, bid);
}

__global__ void __launch_bounds__(256) k_dm_velocities(const DarcyTask *__restrict__ t, MultiMap m)
{
    DM_BOX;
    darcy_velocities_body(T.d, T.b.p, T.b.phi, T.b.k, T.b.v, bid);
}

#define DM_LAUNCH(K, BS, ...) \
    if (blocks) hipLaunchKernelGGL(K, dim3(blocks), dim3(BS), 0, s, t, m, ##__VA_ARGS__)

void launch_dm_ghosts(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks, int field)
{
    DM_LAUNCH(k_dm_ghosts, 256, field);
}
void launch_dm_ghosts_flux(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_ghosts_flux, 256);
}
void launch_dm_pressure_gradient(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_pressure_gradient, 256);
}
void launch_dm_pressure_force(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_pressure_force, 256);
}
void launch_dm_porosities(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_porosities, 256);
}
void launch_dm_por_edges(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_por_edges, 256);
    DM_LAUNCH(k_dm_por_bottom, 256);
}
void launch_dm_permeabilities(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_permeabilities, 256);
}
void launch_dm_solve_start(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_solve_start, 256);
}
void launch_dm_prepare(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_prepare, 256);
}
void launch_dm_reset(hipStream_t s, const DarcyTask *t, uint32_t n)
{
    if (n) hipLaunchKernelGGL(k_dm_reset, dim3(n), dim3(64), 0, s, t);
}
void launch_dm_sweep(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks, uint32_t nijac,
                     bool correct)
{
    DM_LAUNCH(k_dm_sweep, SWEEP_THREADS, nijac, correct ? 1 : 0);
}
void launch_dm_decide(hipStream_t s, const DarcyTask *t, uint32_t n, uint32_t nijac)
{
    if (n) hipLaunchKernelGGL(k_dm_decide, dim3((n + 63u) / 64u), dim3(64), 0, s, t, n, nijac);
}
void launch_dm_net_input(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks, uint32_t nijac)
{
    DM_LAUNCH(k_dm_residual_max, 256, nijac);
    DM_LAUNCH(k_dm_net_input, 256, nijac);
}
void launch_dm_net_apply(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks, uint32_t nijac)
{
    DM_LAUNCH(k_dm_net_apply, 256, nijac);
}
void launch_dm_states(hipStream_t s, const DarcyTask *t, uint32_t n, unsigned long long *out)
{
    if (n) hipLaunchKernelGGL(k_dm_states, dim3((n + 63u) / 64u), dim3(64), 0, s, t, n, out);
}
void launch_dm_normalize(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_normalize, 256);
}
void launch_dm_finish(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_finish, 256);
}
void launch_dm_velocities(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks)
{
    DM_LAUNCH(k_dm_velocities, 256);
}
