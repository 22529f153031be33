// dem_kernels.hip -- gfx950 kernels of one DEM time step.
//
// Reference pipeline (src/device.cu:1038-2545): calcParticleCellID ->
// thrust::sort_by_key -> reorderArrays -> topology -> interact -> bondsLinear
// -> integrate -> summation -> integrateWalls.  Here:
//   k_hash            calcParticleCellID (src/sorting.cuh:25-51)
//   k_radix_*         stable LSD radix sort, 8-bit digits, wave64 ballot ranking
//   k_reorder         reorderArrays (src/sorting.cuh:56-132)
//   k_contact         topology + interact fused per particle
//                     (src/contactsearch.cuh:194-371, 385-726)
//   k_bond_compute /  bondsLinear (src/cohesion.cuh:8-233) as a per-bond pass
//   k_bond_gather     and a per-particle gather in ascending bond order
//   k_integrate       integrate (src/integration.cuh:14-305) fused with the
//                     256-wide block reduction of summation (:309-341)
//   k_integrate_walls integrateWalls (src/integration.cuh:344-450)
#include "kernels.h"

// ------------------------------------------------------------------
// A1: cell hash
// ------------------------------------------------------------------

// src/sorting.cuh:7-20.  Out-of-grid particles raise ERRF_GRID (the
// reference converts a negative double to unsigned there).
__device__ __forceinline__ bool cell_of(const Grid &g, double4 x, uint32_t &cell)
{
    double fx = floor((x.x - g.origo[0]) / (g.L[0] / g.num[0]));
    double fy = floor((x.y - g.origo[1]) / (g.L[1] / g.num[1]));
    double fz = floor((x.z - g.origo[2]) / (g.L[2] / g.num[2]));
    if (!(fx >= 0.0 && fx < (double)g.num[0] && fy >= 0.0 && fy < (double)g.num[1] &&
          fz >= 0.0 && fz < (double)g.num[2]))
        return false;
    uint32_t ix = (uint32_t)fx, iy = (uint32_t)fy, iz = (uint32_t)fz;
    cell = (iz * g.num[1]) * g.num[0] + iy * g.num[0] + ix;
    return true;
}

__global__ void __launch_bounds__(256) k_hash(StepArgs a, const double4 *__restrict__ x,
                                              uint32_t *__restrict__ keys,
                                              uint32_t *__restrict__ vals,
                                              uint32_t *__restrict__ err)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.np) return;
    uint32_t c;
    if (!cell_of(a.grid, x[i], c)) {
        atomicOr(err, ERRF_GRID);
        c = 0;
    }
    keys[i] = c;
    vals[i] = i;
}

void launch_hash(hipStream_t s, const StepArgs &a, const DemBuffers &b)
{
    uint32_t nb = (a.np + 255) / 256;
    hipLaunchKernelGGL(k_hash, dim3(nb), dim3(256), 0, s, a, b.x, b.keys, b.vals, b.errflags);
}

// Squared displacement of a grain since the last rebuild, along periodic
// axes to the nearest image (a grain that wraps around has not moved far).
__device__ __forceinline__ double disp2_since_rebuild(const Grid &g, double4 xi, double4 xr)
{
    double dx = xi.x - xr.x, dy = xi.y - xr.y;
    const double dz = xi.z - xr.z;
    if (g.periodic == 1 || g.periodic == 2) {
        if (dx > 0.5 * g.L[0]) dx -= g.L[0];
        else if (dx < -0.5 * g.L[0]) dx += g.L[0];
    }
    if (g.periodic == 1) {
        if (dy > 0.5 * g.L[1]) dy -= g.L[1];
        else if (dy < -0.5 * g.L[1]) dy += g.L[1];
    }
    return dx * dx + dy * dy + dz * dz;
}

// One wave's share of the step's list bookkeeping: the rebuild bit (any lane
// moved too far), the wrap bit (any lane crossed a periodic boundary; the
// search then tests every list entry) and the largest squared displacement
// since the rebuild (atomic max on the bit pattern, monotonic for d2 >= 0).
__device__ __forceinline__ void record_displacement(uint32_t *flags, uint32_t cur, bool need,
                                                    bool wrapped, double d2)
{
    const uint64_t any_need = __ballot(need);
    const uint64_t any_wrap = __ballot(wrapped);
    double m = d2 == d2 ? d2 : 0.0;  // NaN already asks for a rebuild
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    const uint64_t active = __ballot(true);
    const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)active) - 1);
    if ((threadIdx.x & 63u) == leader) {
        const uint32_t bits = (any_need ? NLF_REBUILD : 0u) | (any_wrap ? NLF_WRAP : 0u);
        if (bits) atomicOr(&flags[cur], bits);
        const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        unsigned long long *slot =
            (unsigned long long *)(flags + NLF_DMAX) + (wave % NLF_DSLOTS);
        const unsigned long long v = (unsigned long long)__double_as_longlong(m);
        if (v > *slot) atomicMax(slot, v);  // most waves see a larger value already
    }
}

// Largest squared displacement since the last rebuild: maximum of the
// NLF_DSLOTS words, formed by one wave (every lane gets it).
__device__ __forceinline__ double dmax_since_rebuild(const uint32_t *flags)
{
    const uint32_t lane = threadIdx.x & 63u;
    double m = __longlong_as_double(((const long long *)(flags + NLF_DMAX))[lane % NLF_DSLOTS]);
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    return m;
}

// Neighbour-list mode: the same cell IDs plus the packed cell coordinates of
// every particle (x | y << 11 | z << 22, all ones outside the grid) and the
// rebuild decision of this step.  flags[cur] becomes non-zero when any
// particle has moved more than sqrt(disp2) since the last rebuild, when the
// host forces a rebuild, or when the sticky list-overflow bit flags[2] is
// set; flags[cur ^ 1] (read by the previous step's kernels) is cleared here.
__device__ __forceinline__ void hash_nl_body(StepArgs a, NLArgs n,
                                                 const double4 *__restrict__ x,
                                                 uint32_t *__restrict__ keys,
                                                 uint32_t *__restrict__ vals,
                                                 uint32_t *__restrict__ cellc,
                                                 const double4 *__restrict__ x_ref,
                                                 float *__restrict__ disp,
                                                 uint32_t *__restrict__ flags, uint32_t cur,
                                                 uint32_t force, uint32_t *__restrict__ err, uint32_t bid)
{
    const uint32_t i = bid * blockDim.x + threadIdx.x;
    if (i == 0) flags[cur ^ 1u] = 0u;
    if (i >= a.np) return;
    const Grid &g = a.grid;
    const double4 xi = x[i];
    const double fx = floor((xi.x - g.origo[0]) / (g.L[0] / g.num[0]));
    const double fy = floor((xi.y - g.origo[1]) / (g.L[1] / g.num[1]));
    const double fz = floor((xi.z - g.origo[2]) / (g.L[2] / g.num[2]));
    uint32_t c = 0, packed = 0xffffffffu;
    if (fx >= 0.0 && fx < (double)g.num[0] && fy >= 0.0 && fy < (double)g.num[1] && fz >= 0.0 &&
        fz < (double)g.num[2]) {
        const uint32_t ix = (uint32_t)fx, iy = (uint32_t)fy, iz = (uint32_t)fz;
        c = (iz * g.num[1]) * g.num[0] + iy * g.num[0] + ix;
        packed = ix | (iy << 11) | (iz << 22);
    } else {
        atomicOr(err, ERRF_GRID);
    }
    keys[i] = c;
    vals[i] = i;
    cellc[i] = packed;
    const double d2 = disp2_since_rebuild(g, xi, x_ref[i]);
    disp[i] = __double2float_ru(sqrt(d2));
    bool need = !(d2 <= n.disp2);
    if (i == 0 && (force || flags[NLF_STICKY] != 0u)) need = true;
    record_displacement(flags, cur, need, false, d2);
}

__global__ void __launch_bounds__(256) k_hash_nl(StepArgs a, NLArgs n,
                                                 const double4 *__restrict__ x,
                                                 uint32_t *__restrict__ keys,
                                                 uint32_t *__restrict__ vals,
                                                 uint32_t *__restrict__ cellc,
                                                 const double4 *__restrict__ x_ref,
                                                 float *__restrict__ disp,
                                                 uint32_t *__restrict__ flags, uint32_t cur,
                                                 uint32_t force, uint32_t *__restrict__ err)
{
    hash_nl_body(a, n, x, keys, vals, cellc, x_ref, disp, flags, cur, force, err, blockIdx.x);
}

void launch_hash_nl(hipStream_t s, const StepArgs &a, const NLArgs &n, const DemBuffers &b,
                    uint32_t cur, uint32_t force)
{
    uint32_t nb = (a.np + 255) / 256;
    hipLaunchKernelGGL(k_hash_nl, dim3(nb), dim3(256), 0, s, a, n, b.x, b.keys, b.vals, b.cellc,
                       b.x_ref, b.nl_disp, b.nl_flags, cur, force, b.errflags);
}

// ------------------------------------------------------------------
// A2: stable LSD radix sort (replaces thrust::sort_by_key)
// ------------------------------------------------------------------
// Tiles of RS_THREADS * RS_ITEMS keys, item j of thread t at tile position
// j * RS_THREADS + t.  Per pass: (1) per-tile digit histogram, (2) one
// exclusive scan over (digit, tile) in digit-major order, (3) a stable
// scatter that ranks equal digits with wave ballots in position order.

#define RS_THREADS 256
#define RS_ITEMS 8
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define RS_WAVES (RS_THREADS / 64)

// Kernels that take `cond` return at once when it is non-null and its
// NLF_REBUILD bit is clear (neighbour-list mode: the cell sort only runs on
// rebuild steps).
#define SKIP_UNLESS(cond) \
    if ((cond) != nullptr && (*(cond) & NLF_REBUILD) == 0u) return

// Digit histogram of one tile of RS_TILE keys (RS_THREADS threads), written
// digit-major: hist[digit * ntiles + tile].
__device__ __forceinline__ void radix_hist_tile(const uint32_t *__restrict__ keys, uint32_t n,
                                                int shift, uint32_t *__restrict__ hist,
                                                uint32_t ntiles, uint32_t tile, uint32_t *h)
{
    const uint32_t t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const uint32_t base = tile * RS_TILE;
    for (int j = 0; j < RS_ITEMS; ++j) {
        const uint32_t i = base + j * RS_THREADS + t;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[t * ntiles + tile] = h[t];
    __syncthreads();
}

__global__ void __launch_bounds__(RS_THREADS) k_radix_hist(const uint32_t *__restrict__ keys,
                                                           uint32_t n, int shift,
                                                           uint32_t *__restrict__ hist,
                                                           uint32_t ntiles,
                                                           const uint32_t *__restrict__ cond)
{
    SKIP_UNLESS(cond);
    __shared__ uint32_t h[256];
    radix_hist_tile(keys, n, shift, hist, ntiles, blockIdx.x, h);
}

// Exclusive scan of m entries in place, one block of 1024 threads, row by
// row of 1024 consecutive entries (coalesced): a wave scan by shuffles, the
// 16 wave totals scanned by one wave, and the running total of the rows
// before; the next row's loads are issued before the current row is
// scanned.
__global__ void __launch_bounds__(1024) k_scan_exclusive(uint32_t *__restrict__ a, uint32_t m,
                                                         const uint32_t *__restrict__ cond)
{
    SKIP_UNLESS(cond);
    __shared__ uint32_t wsum[16];
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    uint32_t carry = 0;
    uint32_t nxt = t < m ? a[t] : 0u;
    for (uint32_t base = 0; base < m; base += 1024u) {
        const uint32_t v = nxt;
        if (base + 1024u + t < m) nxt = a[base + 1024u + t];
        uint32_t x = v;  // inclusive scan over the wave
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63u) wsum[w] = x;
        __syncthreads();
        if (w == 0) {  // inclusive scan of the 16 wave totals
            uint32_t sw = lane < 16u ? wsum[lane] : 0u;
            for (uint32_t off = 1; off < 16u; off <<= 1) {
                const uint32_t y = __shfl_up(sw, off, 64);
                if (lane >= off) sw += y;
            }
            if (lane < 16u) wsum[lane] = sw;
        }
        __syncthreads();
        if (base + t < m) a[base + t] = carry + (w ? wsum[w - 1] : 0u) + x - v;
        carry += wsum[15];
        __syncthreads();  // wsum is rewritten by the next row
    }
}

// Stable scatter of one tile (RS_THREADS threads): a key's place is its
// tile's digit offset plus its rank among the tile's earlier keys of the
// same digit (wave ballots, then the counts of the earlier waves).
__device__ __forceinline__ void radix_scatter_tile(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, uint32_t n, int shift,
    const uint32_t *__restrict__ offs, uint32_t ntiles, uint32_t tile, uint32_t *run,
    uint32_t (*wcnt)[256])
{
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u, wave = t >> 6;
    run[t] = offs[t * ntiles + tile];
    for (int w = 0; w < RS_WAVES; ++w) wcnt[w][t] = 0;
    __syncthreads();
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint32_t base = tile * RS_TILE;
    for (int j = 0; j < RS_ITEMS; ++j) {
        const uint32_t i = base + j * RS_THREADS + t;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t val = valid ? vin[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const uint64_t bb = __ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        const uint32_t rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (uint32_t w = 0; w < wave; ++w) pos += wcnt[w][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < RS_WAVES; ++w) {
            add += wcnt[w][t];
            wcnt[w][t] = 0;
        }
        run[t] += add;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(RS_THREADS) k_radix_scatter(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, uint32_t n, int shift,
    const uint32_t *__restrict__ offs, uint32_t ntiles, const uint32_t *__restrict__ cond)
{
    SKIP_UNLESS(cond);
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_WAVES][256];
    radix_scatter_tile(kin, vin, kout, vout, n, shift, offs, ntiles, blockIdx.x, run, wcnt);
}

void launch_sort(hipStream_t s, uint32_t n, int bits, DemBuffers &b, const uint32_t *cond)
{
    if (n == 0) return;
    uint32_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    for (int pass = 0; pass < passes; ++pass) {
        int shift = pass * 8;
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(RS_THREADS), 0, s, b.keys, n, shift,
                           b.hist, ntiles, cond);
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, b.hist, 256u * ntiles,
                           cond);
        hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(RS_THREADS), 0, s, b.keys, b.vals,
                           b.keys_tmp, b.vals_tmp, n, shift, b.hist, ntiles, cond);
        uint32_t *tk = b.keys;
        b.keys = b.keys_tmp;
        b.keys_tmp = tk;
        uint32_t *tv = b.vals;
        b.vals = b.vals_tmp;
        b.vals_tmp = tv;
    }
}

// ------------------------------------------------------------------
// A3: reorder + cellStart/cellEnd
// ------------------------------------------------------------------

__global__ void k_fill_u32_if(uint32_t *__restrict__ p, size_t n, uint32_t v,
                              const uint32_t *__restrict__ cond)
{
    SKIP_UNLESS(cond);
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double *__restrict__ p, size_t n, double v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_u32(hipStream_t s, uint32_t *p, size_t n, uint32_t v)
{
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

void launch_fill_f64(hipStream_t s, double *p, size_t n, double v)
{
    if (!n) return;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

// src/sorting.cuh:56-132
__global__ void __launch_bounds__(256) k_reorder(
    uint32_t np, const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
    const double4 *__restrict__ x, const double4 *__restrict__ vel,
    const double3 *__restrict__ angvel, double4 *__restrict__ x_s, double4 *__restrict__ vel_s,
    double3 *__restrict__ angvel_s, uint32_t *__restrict__ cell_start,
    uint32_t *__restrict__ cell_end, const uint32_t *__restrict__ cond)
{
    SKIP_UNLESS(cond);
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < np;
         idx += gridDim.x * blockDim.x) {
        uint32_t c = keys[idx];
        if (idx == 0 || c != keys[idx - 1]) {
            cell_start[c] = idx;
            if (idx > 0) cell_end[keys[idx - 1]] = idx;
        }
        if (idx == np - 1) cell_end[c] = idx + 1;
        uint32_t s = vals[idx];
        x_s[idx] = x[s];
        vel_s[idx] = vel[s];
        angvel_s[idx] = angvel[s];
    }
}

// cell_start = cell_end = EMPTY_CELL for every cell
__global__ void __launch_bounds__(256) k_clear_cells(uint32_t *__restrict__ cell_start,
                                                     uint32_t *__restrict__ cell_end,
                                                     uint32_t ncells,
                                                     const uint32_t *__restrict__ cond)
{
    SKIP_UNLESS(cond);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells;
         c += gridDim.x * blockDim.x) {
        cell_start[c] = EMPTY_CELL;
        cell_end[c] = EMPTY_CELL;
    }
}

// Kernels that may be skipped run on at most this many blocks (grid-stride
// loops): a skipped launch then costs little more than its dispatch.
#define COND_BLOCKS 1024u

void launch_reorder(hipStream_t s, const StepArgs &a, const DemBuffers &b, uint32_t ncells,
                    const uint32_t *cond)
{
    uint32_t fb = (ncells + 255) / 256;
    if (fb > COND_BLOCKS) fb = COND_BLOCKS;
    hipLaunchKernelGGL(k_clear_cells, dim3(fb), dim3(256), 0, s, b.cell_start, b.cell_end,
                       ncells, cond);
    uint32_t nb = (a.np + 255) / 256;
    if (nb > COND_BLOCKS) nb = COND_BLOCKS;
    hipLaunchKernelGGL(k_reorder, dim3(nb), dim3(256), 0, s, a.np, b.keys, b.vals, b.x, b.vel,
                       b.angvel, b.x_s, b.vel_s, b.angvel_s, b.cell_start, b.cell_end, cond);
}

// ------------------------------------------------------------------
// A4: periodic images (src/contactsearch.cuh:11-83)
// ------------------------------------------------------------------
// code bits 0-1: x image (1: +L0, 2: -L0), bits 2-3: y image.

__device__ __forceinline__ int dist_mod(const Grid &g, int &tx, int &ty, int tz, d3 &dm,
                                        uint32_t &code)
{
    int nx = (int)g.num[0], ny = (int)g.num[1], nz = (int)g.num[2];
    if (g.periodic == 1) {
        if (tx < 0) { tx += nx; dm = dm + mk3(g.L[0], 0.0, 0.0); code |= 1u; }
        if (tx >= nx) { tx -= nx; dm = dm - mk3(g.L[0], 0.0, 0.0); code |= 2u; }
        if (ty < 0) { ty += ny; dm = dm + mk3(0.0, g.L[1], 0.0); code |= 4u; }
        if (ty >= ny) { ty -= ny; dm = dm - mk3(0.0, g.L[1], 0.0); code |= 8u; }
    }
    if (g.periodic == 2) {
        if (tx < 0) { tx += nx; dm = dm + mk3(g.L[0], 0.0, 0.0); code |= 1u; }
        if (tx >= nx) { tx -= nx; dm = dm - mk3(g.L[0], 0.0, 0.0); code |= 2u; }
        if (ty < 0 || ty >= ny) return -1;
    }
    if (g.periodic > 2 || g.periodic < 1) {
        if (tx < 0 || tx >= nx) return -1;
        if (ty < 0 || ty >= ny) return -1;
    }
    if (tz < 0 || tz >= nz) return -1;
    return 0;
}

__device__ __forceinline__ d3 decode_dm(const Grid &g, uint32_t code)
{
    d3 dm = mk3(0.0, 0.0, 0.0);
    if (code & 1u) dm = dm + mk3(g.L[0], 0.0, 0.0);
    if (code & 2u) dm = dm - mk3(g.L[0], 0.0, 0.0);
    if (code & 4u) dm = dm + mk3(0.0, g.L[1], 0.0);
    if (code & 8u) dm = dm - mk3(0.0, g.L[1], 0.0);
    return dm;
}

// ------------------------------------------------------------------
// A7-A10: contact laws
// ------------------------------------------------------------------

// Slot holding b among the occupied slots of a row, or -1.
__device__ __forceinline__ int find_slot(const uint32_t *row, uint64_t mask, uint32_t b)
{
    while (mask) {
        const int i = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1;
        if ((row[RS(i)] & CT_IDX_MASK) == b) return i;
    }
    return -1;
}

// ------------------------------------------------------------------
// Neighbour (Verlet) lists
// ------------------------------------------------------------------
// A rebuild step lists, for every particle, each other particle whose
// surface gap was below `skin` over the (2R+1)^3 cells around it.  Until a
// particle moves more than skin/2 (less a margin) no pair outside the lists
// can come into contact, so the per-step contact search only tests list
// entries.  The list is column-major, entry k of particle a (original
// index) at nl[k * np + a], so a wave over consecutive particles reads
// consecutive words.  nl_pos[a] is a's position in the cell sort of the
// rebuild.

__device__ void nl_build_one(
    const StepArgs &a, const NLArgs &n, uint32_t pos, const uint32_t *__restrict__ vals,
    const uint32_t *__restrict__ cell_start, const uint32_t *__restrict__ cell_end,
    const double4 *__restrict__ x_s, uint32_t *__restrict__ nl, uint32_t *__restrict__ nl_cnt,
    uint32_t *__restrict__ nl_pos, double4 *__restrict__ x_ref, uint32_t *__restrict__ flags,
    const uint32_t *__restrict__ contacts, const uint64_t *__restrict__ cmask,
    uint8_t *__restrict__ nl_slot, uint8_t *__restrict__ slot_cand, float *__restrict__ nl_gap,
    float *__restrict__ nl_gmin)
{
    const uint32_t np = a.np;
    const Grid &g = a.grid;
    const uint32_t a_orig = vals[pos];
    const double4 xa = x_s[pos];
    nl_pos[a_orig] = pos;
    x_ref[a_orig] = xa;
    const int gx = (int)floor((xa.x - g.origo[0]) / (g.L[0] / g.num[0]));
    const int gy = (int)floor((xa.y - g.origo[1]) / (g.L[1] / g.num[1]));
    const int gz = (int)floor((xa.z - g.origo[2]) / (g.L[2] / g.num[2]));
    const bool inside = gx >= 0 && gx < (int)g.num[0] && gy >= 0 && gy < (int)g.num[1] &&
                        gz >= 0 && gz < (int)g.num[2];
    uint32_t cnt = 0;
    const int R = (int)n.range;
    if (inside) {
        for (int zd = -R; zd <= R; ++zd)
            for (int yd = -R; yd <= R; ++yd)
                for (int xd = -R; xd <= R; ++xd) {
                    int tx = gx + xd, ty = gy + yd, tz = gz + zd;
                    d3 dm = mk3(0.0, 0.0, 0.0);
                    uint32_t code = 0;
                    if (dist_mod(g, tx, ty, tz, dm, code) == -1) continue;
                    const uint32_t cid = (uint32_t)tx + (uint32_t)ty * g.num[0] +
                                         (g.num[0] * g.num[1]) * (uint32_t)tz;
                    const uint32_t start = cell_start[cid];
                    if (start == EMPTY_CELL) continue;
                    const uint32_t end = cell_end[cid];
                    for (uint32_t idx_b = start; idx_b < end; ++idx_b) {
                        if (idx_b == pos) continue;
                        const double4 xb = x_s[idx_b];
                        const d3 x_ab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z) + dm;
                        const double rr = xa.w + xb.w + n.skin;
                        const double s2 = dot3(x_ab, x_ab);
                        if (s2 < rr * rr) {
                            if (cnt < n.cap) {
                                nl[(size_t)cnt * np + a_orig] = vals[idx_b];
                                // surface gap, rounded down
                                nl_gap[(size_t)cnt * np + a_orig] =
                                    __double2float_rd(sqrt(s2) - (xa.w + xb.w));
                            }
                            ++cnt;
                        }
                    }
                }
    }
    // maps between list entries and the slots of the particle's row
    const int NCS = (int)a.nc;
    uint8_t *scand = slot_cand + ct_base(a_orig, NCS);
    for (int sl = 0; sl < NCS; ++sl) scand[RS(sl)] = 0xffu;
    if (cnt > n.cap) {
        nl_cnt[a_orig] = NL_OVF;
        atomicOr(&flags[NLF_STICKY], 1u);
    } else {
        nl_cnt[a_orig] = cnt;
        const uint32_t *row = contacts + ct_base(a_orig, NCS);
        const uint64_t mask = cmask[a_orig];
        float gmin = INFINITY;
        for (uint32_t k = 0; k < cnt; ++k) {
            const int sl = find_slot(row, mask, nl[(size_t)k * np + a_orig]);
            nl_slot[(size_t)k * np + a_orig] = sl < 0 ? 0xffu : (uint8_t)sl;
            if (sl >= 0) scand[RS(sl)] = (uint8_t)k;
            else gmin = fminf(gmin, nl_gap[(size_t)k * np + a_orig]);
        }
        nl_gmin[a_orig] = gmin;
    }
}


__device__ __forceinline__ void nl_build_body(
    const StepArgs &a, const NLArgs &n, const uint32_t *__restrict__ cond,
    const uint32_t *__restrict__ vals, const uint32_t *__restrict__ cell_start,
    const uint32_t *__restrict__ cell_end, const double4 *__restrict__ x_s,
    uint32_t *__restrict__ nl, uint32_t *__restrict__ nl_cnt, uint32_t *__restrict__ nl_pos,
    double4 *__restrict__ x_ref, uint32_t *__restrict__ flags,
    const uint32_t *__restrict__ contacts, const uint64_t *__restrict__ cmask,
    uint8_t *__restrict__ nl_slot, uint8_t *__restrict__ slot_cand, float *__restrict__ nl_gap,
    float *__restrict__ nl_gmin, const uint32_t *__restrict__ err, uint32_t bid, uint32_t nblk)
{
    SKIP_UNLESS(cond);
    if ((*err & ERRF_BARRIER) != 0u) return;  // the sort before it did not finish
    const uint32_t np = a.np;
    if (bid == 0 && threadIdx.x == 0) {
        flags[NLF_COUNT] += 1u;  // rebuild counter
    }
    if (bid == 0 && threadIdx.x < NLF_DSLOTS)  // displacements restart
        ((unsigned long long *)(flags + NLF_DMAX))[threadIdx.x] = 0ull;
    for (uint32_t pos = bid * blockDim.x + threadIdx.x; pos < np; pos += nblk * blockDim.x)
        nl_build_one(a, n, pos, vals, cell_start, cell_end, x_s, nl, nl_cnt, nl_pos, x_ref, flags,
                     contacts, cmask, nl_slot, slot_cand, nl_gap, nl_gmin);
}

__global__ void __launch_bounds__(256) k_nl_build(
    StepArgs a, NLArgs n, const uint32_t *__restrict__ cond, const uint32_t *__restrict__ vals,
    const uint32_t *__restrict__ cell_start, const uint32_t *__restrict__ cell_end,
    const double4 *__restrict__ x_s, uint32_t *__restrict__ nl, uint32_t *__restrict__ nl_cnt,
    uint32_t *__restrict__ nl_pos, double4 *__restrict__ x_ref, uint32_t *__restrict__ flags,
    const uint32_t *__restrict__ contacts, const uint64_t *__restrict__ cmask,
    uint8_t *__restrict__ nl_slot, uint8_t *__restrict__ slot_cand, float *__restrict__ nl_gap,
    float *__restrict__ nl_gmin, const uint32_t *__restrict__ err)
{
    nl_build_body(a, n, cond, vals, cell_start, cell_end, x_s, nl, nl_cnt, nl_pos, x_ref, flags,
                  contacts, cmask, nl_slot, slot_cand, nl_gap, nl_gmin, err, blockIdx.x,
                  gridDim.x);
}

void launch_nl_build(hipStream_t s, const StepArgs &a, const NLArgs &n, const DemBuffers &b,
                     const uint32_t *cond)
{
    const NlBuildTask l = nl_build_task(a, n, b, cond);
    hipLaunchKernelGGL(k_nl_build, dim3(nl_build_blocks(a.np)), dim3(256), 0, s, l.a, l.n, l.cond,
                       l.vals, l.cell_start, l.cell_end, l.x_s, l.nl, l.nl_cnt, l.nl_pos, l.x_ref,
                       l.flags, l.contacts, l.cmask, l.nl_slot, l.slot_cand, l.nl_gap, l.nl_gmin,
                       l.err);
}

// ------------------------------------------------------------------
// The whole list rebuild (cell sort, reorder, list build) in one launch
// ------------------------------------------------------------------
// Between rebuilds the sort, reorder and list kernels only read the rebuild
// flag and return, and each such launch still costs its dispatch.  On the
// list path the radix passes, the cell clear and the reorder run as one
// kernel whose phases are separated by grid barriers, and the list build
// follows as one more launch; a step without a rebuild then costs two
// dispatches instead of nine.  The grid is at most one block per compute
// unit, so every block is resident.

// Grid barrier.  bar[0] counts arrivals, bar[1] is the generation; `gen` is
// the generation this block waits to see end.  A waiter that spins longer
// than `limit` sleeps, or that sees ERRF_BARRIER raised by another block,
// stops waiting, raises ERRF_BARRIER and returns false; the caller then
// returns, so every block of the kernel ends and the step fails with an
// error.  Kernels launched after it skip their work while the flag is set.
#define RB_SPIN_LIMIT (1u << 22)
__device__ __forceinline__ bool grid_barrier(uint32_t *bar, uint32_t nblocks, uint32_t &gen,
                                             uint32_t *err, uint32_t limit)
{
    __shared__ uint32_t s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ok = 1u;
        __threadfence();
        const uint32_t arrived =
            __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks - 1u) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (++spins > limit ||
                    (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
                     ERRF_BARRIER) != 0u) {
                    atomicOr(err, ERRF_BARRIER);
                    ok = 0u;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __threadfence();
        s_ok = ok;
    }
    gen += 1u;
    __syncthreads();
    return s_ok != 0u;
}

// Exclusive scan of 256 values over the block; returns the prefix of this
// thread's value and the block total in *total.
__device__ __forceinline__ uint32_t block_scan_256(uint32_t v, uint32_t *lds, uint32_t *total)
{
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    uint32_t incl = v;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
    }
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w) before += lds[w];
    *total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return before + incl - v;
}

// Exclusive scan of a[0, m) in place by all blocks: block sums, one
// barrier, then each block scans its chunk from the sum of the chunks before
// it.
__device__ bool grid_scan(uint32_t *__restrict__ a, uint32_t m, uint32_t *__restrict__ part,
                          uint32_t *bar, uint32_t bid, uint32_t nblocks, uint32_t &gen,
                          uint32_t *err, uint32_t limit, uint32_t *lds)
{
    const uint32_t t = threadIdx.x;
    const uint32_t chunk = (m + nblocks - 1u) / nblocks;
    const uint32_t lo = min(bid * chunk, m), hi = min(lo + chunk, m);
    uint32_t sum = 0, tot;
    for (uint32_t i = lo + t; i < hi; i += 256u) sum += a[i];
    (void)block_scan_256(sum, lds, &tot);
    if (t == 0) part[bid] = tot;
    if (!grid_barrier(bar, nblocks, gen, err, limit)) return false;
    uint32_t off = 0;
    for (uint32_t b = t; b < bid; b += 256u) off += part[b];
    (void)block_scan_256(off, lds, &off);
    // each thread takes `per` consecutive entries of the chunk
    const uint32_t per = (chunk + 255u) / 256u;
    const uint32_t s0 = min(lo + t * per, hi), s1 = min(s0 + per, hi);
    uint32_t mine = 0;
    for (uint32_t i = s0; i < s1; ++i) mine += a[i];
    uint32_t run = off + block_scan_256(mine, lds, &tot);
    for (uint32_t i = s0; i < s1; ++i) {
        const uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
    return true;
}

// Block `bid` of the G blocks that rebuild one box.
__device__ __forceinline__ void rebuild_body(const StepArgs &a, const RebuildArgs &r, uint32_t bid,
                                             uint32_t G)
{
    if ((*r.cond & NLF_REBUILD) == 0u) return;
    if ((*r.err & ERRF_BARRIER) != 0u) return;  // an earlier barrier timed out
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_WAVES][256];
    uint32_t gen = 0;
    if (threadIdx.x == 0)
        gen = __hip_atomic_load(&r.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t gtid = bid * 256u + threadIdx.x, gsz = G * 256u;
    for (uint32_t c = gtid; c < r.ncells; c += gsz) {
        r.cell_start[c] = EMPTY_CELL;
        r.cell_end[c] = EMPTY_CELL;
    }
    // LSD radix passes, keys and values ping-ponged between the two buffers
    for (uint32_t pass = 0; pass < r.passes; ++pass) {
        const int shift = (int)(8u * pass);
        const uint32_t src = pass & 1u;
        for (uint32_t tile = bid; tile < r.ntiles; tile += G)
            radix_hist_tile(r.keys[src], r.n, shift, r.hist, r.ntiles, tile, run);
        if (!grid_barrier(r.bar, G, gen, r.err, r.spin_limit)) return;
        if (!grid_scan(r.hist, 256u * r.ntiles, r.part, r.bar, bid, G, gen, r.err, r.spin_limit,
                       run))
            return;
        if (!grid_barrier(r.bar, G, gen, r.err, r.spin_limit)) return;
        for (uint32_t tile = bid; tile < r.ntiles; tile += G)
            radix_scatter_tile(r.keys[src], r.vals[src], r.keys[src ^ 1u], r.vals[src ^ 1u], r.n,
                               shift, r.hist, r.ntiles, tile, run, wcnt);
        if (!grid_barrier(r.bar, G, gen, r.err, r.spin_limit)) return;
    }
    const uint32_t *keys = r.keys[r.passes & 1u], *vals = r.vals[r.passes & 1u];
    const uint32_t np = a.np;
    for (uint32_t idx = gtid; idx < np; idx += gsz) {
        const uint32_t c = keys[idx];
        if (idx == 0 || c != keys[idx - 1]) {
            r.cell_start[c] = idx;
            if (idx > 0) r.cell_end[keys[idx - 1]] = idx;
        }
        if (idx == np - 1) r.cell_end[c] = idx + 1;
        const uint32_t sidx = vals[idx];
        r.x_s[idx] = r.x[sidx];
        r.vel_s[idx] = r.vel[sidx];
        r.angvel_s[idx] = r.angvel[sidx];
    }
}

__global__ void __launch_bounds__(256) k_rebuild(StepArgs a, RebuildArgs r)
{
    rebuild_body(a, r, blockIdx.x, gridDim.x);
}

int rebuild_blocks_per_cu()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_rebuild, 256, 0) != hipSuccess)
        return 0;
    return nb;
}

void launch_rebuild(hipStream_t s, const StepArgs &a, const NLArgs &n, DemBuffers &b,
                    uint32_t ncells, int bits, uint32_t blocks, const uint32_t *cond,
                    uint32_t spin_limit)
{
    if (a.np == 0) return;
    const RebuildTask t = rebuild_task(a, b, ncells, bits, cond, spin_limit);
    hipLaunchKernelGGL(k_rebuild, dim3(blocks), dim3(256), 0, s, t.a, t.r);
    // the sorted keys and values end in the buffer the last pass wrote
    if (t.r.passes & 1u) {
        uint32_t *tk = b.keys;
        b.keys = b.keys_tmp;
        b.keys_tmp = tk;
        uint32_t *tv = b.vals;
        b.vals = b.vals_tmp;
        b.vals_tmp = tv;
    }
    // the list build, the heaviest phase, on one thread per grain (it
    // needs no barrier after it, so its grid is not bounded by residency)
    const NlBuildTask l = nl_build_task(a, n, b, cond);
    hipLaunchKernelGGL(k_nl_build, dim3((a.np + 255) / 256), dim3(256), 0, s, l.a, l.n, l.cond,
                       l.vals, l.cell_start, l.cell_end, l.x_s, l.nl, l.nl_cnt, l.nl_pos, l.x_ref,
                       l.flags, l.contacts, l.cmask, l.nl_slot, l.slot_cand, l.nl_gap, l.nl_gmin,
                       l.err);
}

// Sum and maximum of the list lengths (overflowed lists count as cap + 1).
__global__ void __launch_bounds__(256) k_nl_stats(const uint32_t *__restrict__ nl_cnt, uint32_t np,
                                                  uint32_t cap, unsigned long long *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = 0;
    if (i < np) c = nl_cnt[i] == NL_OVF ? cap + 1u : nl_cnt[i];
    unsigned long long sum = c;
    uint32_t mx = c;
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off, 64);
        const uint32_t o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63u) == 0) {
        atomicAdd(&out[0], sum);
        atomicMax(&out[1], (unsigned long long)mx);
    }
}

void launch_nl_stats(hipStream_t s, const DemBuffers &b, uint32_t np, uint32_t cap,
                     unsigned long long *out)
{
    hipLaunchKernelGGL(k_nl_stats, dim3((np + 255) / 256), dim3(256), 0, s, b.nl_cnt, np, cap,
                       out);
}

// Image of list entry b as the reference's 27-cell search would reach it
// this step: b counts only if its cell is one of the 27 around the cell
// (gx, gy, gz) (periodic axes wrapped to the nearest image; valid for 3 or
// more cells along a periodic axis).  offc = (dz+1)*9 + (dy+1)*3 + (dx+1) is
// the position of that cell in the search order of
// src/contactsearch.cuh:358-369.
__device__ __forceinline__ bool nl_image(const Grid &g, int gx, int gy, int gz, uint32_t cb,
                                         uint32_t &offc, uint32_t &code, d3 &dm)
{
    if (cb == 0xffffffffu) return false;  // outside the grid (the step reports it)
    int dx = (int)(cb & 2047u) - gx;
    int dy = (int)((cb >> 11) & 2047u) - gy;
    const int dz = (int)(cb >> 22) - gz;
    if (g.periodic == 1 || g.periodic == 2) {
        if (dx > 1) dx -= (int)g.num[0];
        else if (dx < -1) dx += (int)g.num[0];
    }
    if (g.periodic == 1) {
        if (dy > 1) dy -= (int)g.num[1];
        else if (dy < -1) dy += (int)g.num[1];
    }
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return false;
    int tx = gx + dx, ty = gy + dy;
    dm = mk3(0.0, 0.0, 0.0);
    code = 0;
    if (dist_mod(g, tx, ty, gz + dz, dm, code) == -1) return false;
    offc = (uint32_t)((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1));
    return true;
}

// 128-bit set over list positions
struct Set128 {
    uint64_t lo, hi;
};
__device__ __forceinline__ void set_bit(Set128 &m, uint32_t k)
{
    if (k < 64) m.lo |= 1ull << k;
    else m.hi |= 1ull << (k - 64);
}
__device__ __forceinline__ void clear_bit(Set128 &m, uint32_t k)
{
    if (k < 64) m.lo &= ~(1ull << k);
    else m.hi &= ~(1ull << (k - 64));
}
__device__ __forceinline__ bool any_bit(const Set128 &m) { return (m.lo | m.hi) != 0ull; }

// List entry of `m` with the smallest discovery key (offc, b): the order in
// which the reference's cell search meets the contacts (cells in search
// order, then ascending original index inside a cell, which is the order of
// the stable cell sort).
__device__ __forceinline__ uint32_t nl_first(const Grid &g, int gx, int gy, int gz, Set128 m,
                                             const uint32_t *__restrict__ nl, uint32_t np,
                                             uint32_t pos, const uint32_t *__restrict__ cellc,
                                             uint32_t &b_out, uint32_t &offc_out,
                                             uint32_t &code_out)
{
    uint64_t best = ~0ull;
    uint32_t bk = 0;
    for (int part = 0; part < 2; ++part) {
        uint64_t w = part ? m.hi : m.lo;
        while (w) {
            const uint32_t k = (uint32_t)(__ffsll((unsigned long long)w) - 1) + 64u * part;
            w &= w - 1;
            const uint32_t b = nl[(size_t)k * np + pos];
            uint32_t offc = 0, code = 0;
            d3 dm;
            (void)nl_image(g, gx, gy, gz, cellc[b], offc, code, dm);
            const uint64_t key = ((uint64_t)offc << 32) | b;
            if (key < best) {
                best = key;
                bk = k;
                b_out = b;
                offc_out = offc;
                code_out = code;
            }
        }
    }
    return bk;
}

// Slot rule of src/contactsearch.cuh:267-286 applied to the contacts found
// this step from the neighbour list.  found: list positions of every contact
// found; fresh: those not in the row yet; recode: those in the row whose
// periodic image changed.  The maps nl_slot (slot of each list entry, 0xff
// if none) and scand (list entry of each slot of the row, 0xff if none) are
// kept exact.  Returns true on slot overflow.
__device__ bool slot_update(const Grid &g, uint32_t np, uint32_t pos, Set128 found, Set128 fresh,
                            Set128 recode, const uint32_t *__restrict__ nl,
                            uint8_t *__restrict__ nl_slot, const uint32_t *__restrict__ cellc,
                            int gx, int gy, int gz, uint32_t *row, uint8_t *scand,
                            uint64_t &mask, int NCS, uint64_t full)
{
    const int n_found = __popcll(found.lo) + __popcll(found.hi);
    const int n_new = __popcll(fresh.lo) + __popcll(fresh.hi);
    if (n_found <= NCS && __popcll(mask) + n_new <= NCS) {
        // No slot runs out: contacts already in the row keep their slots
        // (a changed image code is rewritten) and the new ones take the
        // highest empty slot in discovery order.
        for (int part = 0; part < 2; ++part) {
            uint64_t w = part ? recode.hi : recode.lo;
            while (w) {
                const uint32_t k = (uint32_t)(__ffsll((unsigned long long)w) - 1) + 64u * part;
                w &= w - 1;
                const size_t li = (size_t)k * np + pos;
                uint32_t offc = 0, code = 0;
                d3 dm;
                (void)nl_image(g, gx, gy, gz, cellc[nl[li]], offc, code, dm);
                const uint8_t sl = nl_slot[li];
                row[RS(sl)] = ct_word(row[RS(sl)] & CT_IDX_MASK, code);
            }
        }
        while (any_bit(fresh)) {
            uint32_t b = 0, offc = 0, code = 0;
            const uint32_t k = nl_first(g, gx, gy, gz, fresh, nl, np, pos, cellc, b, offc, code);
            clear_bit(fresh, k);
            const int cpos = 63 - __clzll((long long)(~mask & full));
            row[RS(cpos)] = ct_word(b, code);
            mask |= 1ull << cpos;
            nl_slot[(size_t)k * np + pos] = (uint8_t)cpos;
            scand[RS(cpos)] = (uint8_t)k;
        }
        return false;
    }
    // General case: every contact in discovery order through the full rule.
    int nc = 0;
    int broken = -1;  // search cell left early (more than NCS contacts)
    bool ovf = false;
    while (any_bit(found)) {
        uint32_t b = 0, offc = 0, code = 0;
        const uint32_t k = nl_first(g, gx, gy, gz, found, nl, np, pos, cellc, b, offc, code);
        clear_bit(found, k);
        if ((int)offc == broken) continue;
        int cpos = find_slot(row, mask, b);
        if (cpos < 0) {
            const uint64_t emp = ~mask & full;
            cpos = emp ? 63 - __clzll((long long)emp) : nc;
        }
        if (cpos < NCS) {
            const uint8_t prev = scand[RS(cpos)];
            if (prev != 0xffu && prev != (uint8_t)k) nl_slot[(size_t)prev * np + pos] = 0xffu;
            row[RS(cpos)] = ct_word(b, code);
            mask |= 1ull << cpos;
            nl_slot[(size_t)k * np + pos] = (uint8_t)cpos;
            scand[RS(cpos)] = (uint8_t)k;
        } else {
            ovf = true;
        }
        ++nc;
        if (nc > NCS) broken = (int)offc;
    }
    return ovf;
}

// Overlap test of list entry b for a particle at xa in cell (gx, gy, gz),
// with the reference's arithmetic (contactsearch.cuh:225-236).
__device__ __forceinline__ bool nl_overlap(const Grid &g, double4 xa, int gx, int gy, int gz,
                                           double4 xb, uint32_t cb)
{
    uint32_t offc, code;
    d3 dm;
    if (!nl_image(g, gx, gy, gz, cb, offc, code, dm)) return false;
    const d3 x_ab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z) + dm;
    const double s2 = dot3(x_ab, x_ab);
    const double rsum = xa.w + xb.w;
    if (s2 >= rsum * rsum * (1.0 + 1e-9)) return false;
    return sqrt(s2) - rsum < 0.0;
}

struct Acc {
    d3 F, T;
    double es_dot, ev_dot, p;
};

// What one contact adds to its particle's sums.  The reference adds these
// terms to the running sums inside the contact law; the sums see the same
// additions in the same order when they are applied afterwards by
// add_contrib (a term a law does not add is +0.0 here, and adding +0.0
// changes no sum: every sum starts at +0.0 and so is never -0.0).
struct Contrib {
    d3 F, T;
    double p;
    double ev_n;  // normal damping
    double ev_t;  // tangential damping (static friction)
    double es;    // frictional slip
};

__device__ __forceinline__ void add_contrib(Acc &A, const Contrib &C)
{
    A.ev_dot += C.ev_n;
    A.ev_dot += C.ev_t;
    A.es_dot += C.es;
    A.F = A.F + C.F;
    A.T = A.T + C.T;
    A.p += C.p;
}

// -kappa sqrt(r_a r_b): for kappa = +0 it is -0 for any positive radii, so
// the square root is skipped without changing a bit of f_c = factor * n.
__device__ __forceinline__ double cohesion_factor(double kappa, double r_a, double r_b)
{
    return __double_as_longlong(kappa) == 0ll ? -0.0 : -kappa * sqrt(r_a * r_b);
}

// contactLinear, src/contactmodels.cuh:241-408
__device__ __forceinline__ void law_linear(const StepArgs &a, Contrib &C, d3 vel_a, d3 angvel_a,
                                           double4 vb, double3 wb, double r_a, double r_b,
                                           d3 x, double x_length, double delta, double *dts)
{
    const Params &P = a.prm;
    const double dt = a.dt;
    d3 vel_b = xyz(vb), angvel_b = xyz(wb);
    d3 delta_t0_uncor = mk3(dts[0], dts[1], dts[2]);
    d3 n = x / x_length;
    d3 vel_linear = vel_a - vel_b;
    d3 vel = (vel_linear + (r_a + delta / 2.0) * cross3(n, angvel_a)) +
             (r_b + delta / 2.0) * cross3(n, angvel_b);
    double vel_n = -dot3(vel_linear, n);
    d3 vel_t = vel - dot3(n, vel) * n;
    double vel_t_length = len3(vel_t);
    d3 delta_t0 = delta_t0_uncor - dot3(n, delta_t0_uncor) * n;
    const double delta_t0_sq = dot3(delta_t0, delta_t0);  // only tested > 0
    d3 delta_t = delta_t0;
    double k_n = P.k_n, k_t = P.k_t;
    if (P.E > .001) {
        k_n = 3.141592654 / 2.0 * P.E * (r_a + r_b) / 2.;
        k_t = k_n;
    }
    d3 f_n = fmax(0.0, -k_n * delta + P.gamma_n * vel_n) * n;
    double f_n_length = len3(f_n);
    C.ev_n = 0.5 * P.gamma_n * vel_n * vel_n;
    d3 f_c = cohesion_factor(P.kappa, r_a, r_b) * n;
    d3 f_t = mk3(0.0, 0.0, 0.0);
    if (delta_t0_sq > 0.0 || vel_t_length > 0.0) {
        delta_t = delta_t0 + dt * vel_t;
        d3 f_t_elast = (-k_t) * delta_t;
        d3 f_t_visc = (-P.gamma_t) * vel_t;
        f_t = f_t_elast + f_t_visc;
        double f_t_length = len3(f_t);
        const bool no_cohesion = __double_as_longlong(P.kappa) == 0ll;
        const double fnc = no_cohesion ? f_n_length : len3(f_n - f_c);
        double f_t_limit = P.mu_s * fnc;
        C.ev_t = 0.5 * P.gamma_t * vel_t_length * vel_t_length;
        if (f_t_length > f_t_limit) {
            d3 t = f_t / len3(f_t);
            f_t = f_t_limit * t;
            delta_t = (-1.0 / k_t) * ((P.mu_d * fnc) * t + P.gamma_t * vel_t);
            C.es = 0.5 * len3(dt * (len3(f_t) * vel_t)) / dt;
        }
    }
    C.F = (f_n + f_t) + f_c;
    C.T = cross3((-(r_a + delta * 0.5)) * n, f_t);
    C.p = f_n_length / (4.0 * PI_REF * r_a * r_a);
    dts[0] = delta_t.x;
    dts[1] = delta_t.y;
    dts[2] = delta_t.z;
}

// contactmodel 3, Hertz-Mindlin: the structure of contactHertz
// (src/contactmodels.cuh:414-591) on the overlap magnitude d = -delta with
// d^(3/2) = d*sqrt(d) and d^(1/4) = sqrt(sqrt(d)).  With E > 0.001 (Poisson's
// ratio nu in the k_r field): E* = E/(2(1-nu^2)), R* = r_a r_b/(r_a+r_b),
// G* = E/(2(1+nu)) / (2(2-nu)), k_n = 4/3 E* sqrt(R*), k_t = 8 G* sqrt(R*).
__device__ __forceinline__ void law_hertz(const StepArgs &a, Contrib &C, d3 vel_a, d3 angvel_a,
                                          double4 vb, double3 wb, double r_a, double r_b,
                                          d3 x_ab, double x_ab_length, double delta_ab,
                                          double *dts)
{
    const Params &P = a.prm;
    const double dt = a.dt;
    d3 vel_b = xyz(vb), angvel_b = xyz(wb);
    d3 delta_t0_uncor = mk3(dts[0], dts[1], dts[2]);
    d3 n_ab = x_ab / x_ab_length;
    d3 vel_ab_linear = vel_a - vel_b;
    d3 vel_ab = (vel_ab_linear + (r_a + delta_ab / 2.0) * cross3(n_ab, angvel_a)) +
                (r_b + delta_ab / 2.0) * cross3(n_ab, angvel_b);
    d3 angvel_ab = angvel_a - angvel_b;
    double vel_n_ab = dot3(vel_ab_linear, n_ab);
    d3 vel_t_ab = vel_ab - dot3(vel_ab, n_ab) * n_ab;
    double vel_t_ab_length = len3(vel_t_ab);
    d3 delta_t0 = delta_t0_uncor - dot3(delta_t0_uncor, n_ab) * n_ab;
    // |delta_t0| only enters as |delta_t0| > 0, which equals dot > 0
    const double delta_t0_sq = dot3(delta_t0, delta_t0);
    d3 delta_t = delta_t0;
    double d = -delta_ab;
    double sd = sqrt(d);
    double qd = sqrt(sd);
    double k_n = P.k_n, k_t = P.k_t;
    if (P.E > .001) {
        // E* = E/(2(1-nu^2)) and G* = (E/(2(1+nu)))/(2(2-nu)) are formed on
        // the host with the same IEEE operations
        double R_star = r_a * r_b / (r_a + r_b);
        double sR = sqrt(R_star);
        k_n = 4.0 / 3.0 * P.E_star * sR;
        k_t = 8.0 * P.G_star * sR;
    }
    d3 f_n = fmax(0.0, k_n * d * sd - P.gamma_n * qd * vel_n_ab) * n_ab;
    C.ev_n = P.gamma_n * vel_n_ab * vel_n_ab;  // contactmodels.cuh:506
    double f_n_length = len3(f_n);
    d3 f_c = cohesion_factor(P.kappa, r_a, r_b) * n_ab;
    d3 f_t = mk3(0.0, 0.0, 0.0);
    d3 T_res = mk3(0.0, 0.0, 0.0);
    if (delta_t0_sq > 0.0 || vel_t_ab_length > 0.0) {
        d3 f_t_elast = (-k_t * sd) * delta_t0;
        d3 f_t_visc = (-P.gamma_t * qd) * vel_t_ab;
        // with kappa = +0, f_c is a vector of zeros and |f_n - f_c| = |f_n|
        const double fnc = __double_as_longlong(P.kappa) == 0ll ? f_n_length : len3(f_n - f_c);
        double f_t_limit;
        if (vel_t_ab_length > (double)0.001f)
            f_t_limit = P.mu_d * fnc;
        else
            f_t_limit = P.mu_s * fnc;
        f_t = f_t_elast + f_t_visc;
        double f_t_length = len3(f_t);
        if (f_t_length > f_t_limit) {
            f_t = (f_t_limit / f_t_length) * f_t;
            delta_t = (f_t + (P.gamma_t * qd) * vel_t_ab) / (-k_t * sd);
            // the spring energy released by the slip, contactmodels.cuh:560
            C.es = len3(delta_t0 - delta_t) * k_t / dt;
        } else {
            delta_t = delta_t0 + dt * vel_t_ab;
        }
    }
    const double w2 = dot3(angvel_ab, angvel_ab);
    if (__double_as_longlong(P.gamma_r) == 0ll && __double_as_longlong(P.mu_r) == 0ll &&
        w2 < INFINITY && f_n_length < INFINITY) {
        // gamma_r = mu_r = +0: m = fmin(+0, +0) = +0 and T_res = -0 (w/|w|),
        // whose components have the signs of -0 * w (|w| = sqrt(w2) > 0 is
        // finite), so no square root or division is needed
        if (w2 > 0.0) T_res = (-0.0) * angvel_ab;
    } else {
        const double angvel_ab_length = sqrt(w2);
        if (angvel_ab_length > 0.0) {
            double m = fmin(P.gamma_r * r_a * angvel_ab_length, P.mu_r * r_a * f_n_length);
            T_res = (-1.0 * m) * (angvel_ab / angvel_ab_length);
        }
    }
    C.F = (f_n + f_t) + f_c;
    C.T = (-(r_a + delta_ab / 2.0)) * cross3(n_ab, f_t) + T_res;
    C.p = f_n_length / (4.0 * PI_REF * r_a * r_a);
    dts[0] = delta_t.x;
    dts[1] = delta_t.y;
    dts[2] = delta_t.z;
}

// contactLinearViscous, src/contactmodels.cuh:119-237 (model 1)
__device__ __forceinline__ void law_linear_viscous(const StepArgs &a, Acc &A, double4 va4,
                                                   double4 vb4, double3 wa4, double3 wb4,
                                                   double r_a, double r_b, d3 x_ab,
                                                   double x_ab_length, double delta_ab,
                                                   double kappa)
{
    const Params &P = a.prm;
    d3 vel_a = xyz(va4), vel_b = xyz(vb4), angvel_a = xyz(wa4), angvel_b = xyz(wb4);
    d3 n_ab = x_ab / x_ab_length;
    d3 vel_ab_linear = vel_a - vel_b;
    d3 vel_ab = (vel_ab_linear + (r_a + delta_ab / 2.0) * cross3(n_ab, angvel_a)) +
                (r_b + delta_ab / 2.0) * cross3(n_ab, angvel_b);
    double vel_n_ab = dot3(vel_ab_linear, n_ab);
    d3 vel_t_ab = vel_ab - dot3(vel_ab, n_ab) * n_ab;
    double vel_t_ab_length = len3(vel_t_ab);
    double k_n = P.k_n;
    if (P.E > .001) k_n = 3.141592654 / 2.0 * P.E * (r_a + r_b) / 2.;
    d3 f_n = fmax(0.0, -k_n * delta_ab - P.gamma_n * vel_n_ab) * n_ab;
    if (dot3(f_n, n_ab) < 0.0) f_n = mk3(0.0, 0.0, 0.0);
    double f_n_length = len3(f_n);
    d3 f_c = (-kappa * sqrt(r_a * r_b)) * n_ab;
    d3 f_t = mk3(0.0, 0.0, 0.0);
    if (vel_t_ab_length > 0.0) {
        double f_t_visc = P.gamma_t * vel_t_ab_length;
        double f_t_limit;
        if (vel_t_ab_length > (double)0.001f)
            f_t_limit = P.mu_d * len3(f_n - f_c);
        else
            f_t_limit = P.mu_s * len3(f_n - f_c);
        if (f_t_visc < f_t_limit)
            f_t = ((-f_t_visc) * vel_t_ab) / vel_t_ab_length;
        else
            f_t = ((-f_t_limit) * vel_t_ab) / vel_t_ab_length;
    }
    A.F = A.F + ((f_n + f_t) + f_c);
    A.T = A.T + (-(r_a + delta_ab / 2.0)) * cross3(n_ab, f_t);
    A.p += f_n_length / (4.0 * PI_REF * r_a * r_a);
}

// capillaryCohesion_exp, src/cohesion.cuh:237-277
__device__ __forceinline__ void law_capillary(const StepArgs &a, Acc &A, double r_a, double r_b,
                                              double delta_ab, d3 x_ab, double x_ab_length,
                                              double kappa)
{
    d3 n_ab = x_ab / x_ab_length;
    double r;
    if ((r_a / r_b) > (r_b / r_a)) r = r_a / r_b;
    else r = r_b / r_a;
    double hh = -sqrt(r);
    double R_har = (2.0 * r_a * r_b) / (r_a + r_b);
    double R_geo = sqrt(r_a * r_b);
    double lambda = (double)0.9f * hh * sqrt(a.prm.V_b / R_har);
    d3 f_c = (-kappa * R_geo * exp(-delta_ab / lambda)) * n_ab;
    A.F = A.F + f_c;
}

// contactLinear_wall, src/contactmodels.cuh:10-114
__device__ __forceinline__ double law_wall(const StepArgs &a, Acc &A, double4 va4, double3 wa4,
                                           double radius_a, d3 n, double delta, double wvel)
{
    const Params &P = a.prm;
    const double dt = a.dt;
    d3 vel_linear = xyz(va4), angvel = xyz(wa4);
    d3 vel = vel_linear + (radius_a + delta / 2.0) * cross3(n, angvel);
    vel = mk3(vel.x + wvel, vel.y + wvel, vel.z + wvel);
    double vel_n = dot3(vel_linear, n);
    d3 vel_t = vel - dot3(n, vel) * n;
    double vel_t_length = len3(vel_t);
    double k_n = P.k_n;
    if (P.E > .001) k_n = 3.141592654 / 2.0 * P.E * radius_a;
    d3 f_n = fmax(0.0, -k_n * delta - P.gamma_wn * vel_n) * n;
    double f_n_length = len3(f_n);
    A.ev_dot += P.gamma_wn * vel_n * vel_n;
    d3 f_t = mk3(0.0, 0.0, 0.0);
    if (vel_t_length > 0.0 && P.gamma_wt > 0.0) {
        d3 f_t_visc = P.gamma_wt * vel_t;
        double f_t_visc_length = len3(f_t_visc);
        double f_t_limit;
        if (vel_t_length > 0.0) f_t_limit = P.mu_wd * f_n_length;
        else f_t_limit = P.mu_ws * f_n_length;
        if (f_t_visc_length < f_t_limit) {
            f_t = (-1.0) * f_t_visc;
        } else {
            f_t = ((-f_t_limit) * vel_t) / vel_t_length;
            A.es_dot += len3(dt * (len3(f_t) * vel_t)) / dt;
        }
    }
    A.F = A.F + (f_n + f_t);
    A.T = A.T + cross3((-(radius_a + delta * 0.5)) * n, f_t);
    A.p += f_n_length / (4.0 * PI_REF * radius_a * radius_a);
    return dot3(f_n, n);
}

// ------------------------------------------------------------------
// A5 + A6: contact search (topology) and contact forces (interact)
// ------------------------------------------------------------------
// Contact models 2 and 3 run as two kernels, like the reference's topology
// (src/contactsearch.cuh:194-371) and interact (:385-726):
//   k_topology  one thread per particle: overlaps found this step take
//               slots of the particle's row by the reference's rule.
//   k_forces    one wave per 64 particles: the occupied slots of the 64 rows
//               are dealt out to the 64 lanes, one contact per lane per
//               round; each lane evaluates its contact law, and the lane
//               owning the row adds the contributions in slot order.  The
//               wave stays busy however unevenly its 64 rows are filled.
// Model 1 (no contact history) runs as one fused thread per particle.
//
// cmask holds one bit per occupied slot; empty slots keep delta_t == 0,
// which the reference re-establishes every step.  The search skips the sqrt
// of a candidate whose squared distance s is at least (r_a+r_b)^2 (1+1e-9):
// then sqrt(s) > r_a+r_b and the reference's test sqrt(s) - (r_a+r_b) < 0 is
// false, so the skip never changes it.

// contactsearch.cuh:352-354; false outside the grid
__device__ __forceinline__ bool grid_pos(const Grid &g, double4 xa, int &gx, int &gy, int &gz)
{
    gx = (int)floor((xa.x - g.origo[0]) / (g.L[0] / g.num[0]));
    gy = (int)floor((xa.y - g.origo[1]) / (g.L[1] / g.num[1]));
    gz = (int)floor((xa.z - g.origo[2]) / (g.L[2] / g.num[2]));
    return gx >= 0 && gx < (int)g.num[0] && gy >= 0 && gy < (int)g.num[1] && gz >= 0 &&
           gz < (int)g.num[2];
}

// Topology over the 27 cells (z, y, x order, contactsearch.cuh:358-369)
// with the slot rule of contactsearch.cuh:267-286.  Returns true on slot
// overflow.
__device__ bool topology_cells(const Grid &g, uint32_t idx_a, double4 xa, int gx, int gy,
                               int gz, const uint32_t *__restrict__ vals,
                               const uint32_t *__restrict__ cell_start,
                               const uint32_t *__restrict__ cell_end,
                               const double4 *__restrict__ x_s, uint32_t *row,
                               uint64_t &mask, int NCS, uint64_t full)
{
    int nc = 0;
    bool ovf = false;
    for (int zd = -1; zd < 2; ++zd)
        for (int yd = -1; yd < 2; ++yd)
            for (int xd = -1; xd < 2; ++xd) {
                int tx = gx + xd, ty = gy + yd, tz = gz + zd;
                d3 dm = mk3(0.0, 0.0, 0.0);
                uint32_t code = 0;
                if (dist_mod(g, tx, ty, tz, dm, code) == -1) continue;
                const uint32_t cid = (uint32_t)tx + (uint32_t)ty * g.num[0] +
                                     (g.num[0] * g.num[1]) * (uint32_t)tz;
                const uint32_t start = cell_start[cid];
                if (start == EMPTY_CELL) continue;
                const uint32_t end = cell_end[cid];
                for (uint32_t idx_b = start; idx_b < end; ++idx_b) {
                    if (idx_b == idx_a) continue;
                    const double4 xb = x_s[idx_b];
                    const d3 x_ab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z) + dm;
                    const double s2 = dot3(x_ab, x_ab);
                    const double rsum = xa.w + xb.w;
                    if (s2 >= rsum * rsum * (1.0 + 1e-9)) continue;
                    const double delta = sqrt(s2) - rsum;
                    if (!(delta < 0.0)) continue;
                    const uint32_t b_orig = vals[idx_b];
                    int cpos = find_slot(row, mask, b_orig);
                    if (cpos < 0) {
                        const uint64_t emp = ~mask & full;
                        cpos = emp ? 63 - __clzll((long long)emp) : nc;
                    }
                    if (cpos < NCS) {
                        row[RS(cpos)] = ct_word(b_orig, code);
                        mask |= 1ull << cpos;
                    } else {
                        ovf = true;
                    }
                    ++nc;
                    if (nc > NCS) break;
                }
            }
    return ovf;
}

// Topology with the 27-cell search, one thread per sorted position (the
// reference's pipeline).
__global__ void __launch_bounds__(256) k_topology_cells(
    StepArgs a, const uint32_t *__restrict__ vals, const uint32_t *__restrict__ cell_start,
    const uint32_t *__restrict__ cell_end, const double4 *__restrict__ x_s,
    uint32_t *__restrict__ contacts, uint64_t *__restrict__ cmask,
    uint32_t *__restrict__ err)
{
    const uint32_t idx_a = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx_a >= a.np) return;
    const Grid &g = a.grid;
    const uint32_t a_orig = vals[idx_a];
    const double4 xa = x_s[idx_a];
    int gx, gy, gz;
    if (!grid_pos(g, xa, gx, gy, gz)) {
        // the step is invalid (k_hash reports it too); no cell index may
        // leave the grid
        atomicOr(err, ERRF_GRID);
        return;
    }
    const int NCS = (int)a.nc;
    const uint64_t full = NCS >= 64 ? ~0ull : ((1ull << NCS) - 1ull);
    const uint64_t mask0 = cmask[a_orig];
    uint64_t mask = mask0;
    const bool ovf = topology_cells(g, idx_a, xa, gx, gy, gz, vals, cell_start, cell_end, x_s,
                                    contacts + ct_base(a_orig, NCS), mask, NCS, full);
    if (ovf) atomicOr(err, ERRF_OVERFLOW);
    if (mask != mask0) cmask[a_orig] = mask;
}

// Topology from the neighbour lists, one wave per 64 consecutive particles.
// The list entries of the 64 particles are dealt out to the lanes, TU
// overlap tests per lane per round with their loads issued together (so
// long lists of large grains do not hold up the wave); a test also tells
// from nl_slot whether the contact is new and from the row's image code
// whether its periodic image changed.  Each owner then applies the slot rule
// to its new contacts.  A particle whose list overflowed searches its 27
// cells instead (the sort then runs every step, so the sorted arrays are
// current).
// Periodic image of b for a grid searched one cell deep (cells at least as
// wide as a diameter plus the skin, 3 or more cells along periodic axes):
// every pair that can overlap lies in neighbouring cells, and the image the
// reference's cell search uses for such a pair is the nearest one, so it
// follows from the coordinate difference alone.  code as in dist_mod.
__device__ __forceinline__ void nearest_image(const Grid &g, double4 xa, double4 xb, d3 &dm,
                                              uint32_t &code)
{
    dm = mk3(0.0, 0.0, 0.0);
    code = 0;
    if (g.periodic == 1 || g.periodic == 2) {
        const double dx = xa.x - xb.x;
        if (dx < -0.5 * g.L[0]) { dm.x = g.L[0]; code |= 1u; }
        else if (dx > 0.5 * g.L[0]) { dm.x = -g.L[0]; code |= 2u; }
    }
    if (g.periodic == 1) {
        const double dy = xa.y - xb.y;
        if (dy < -0.5 * g.L[1]) { dm.y = g.L[1]; code |= 4u; }
        else if (dy > 0.5 * g.L[1]) { dm.y = -g.L[1]; code |= 8u; }
    }
}

// Every list entry of one particle tested by its owner lane (the rare case
// in which new contacts find no empty slot on a step that tested only the
// candidates for new contacts).
template <bool NEAR>
__device__ void owner_test_all(const Grid &g, uint32_t np, uint32_t pos, uint32_t lcnt,
                               const uint32_t *__restrict__ nl,
                               const uint8_t *__restrict__ nl_slot,
                               const uint32_t *__restrict__ cellc, const double4 *__restrict__ x,
                               double4 xa, int gx, int gy, int gz, const uint32_t *row,
                               Set128 &found, Set128 &fresh, Set128 &recode)
{
    found = fresh = recode = Set128{0ull, 0ull};
    for (uint32_t k = 0; k < lcnt; ++k) {
        const size_t li = (size_t)k * np + pos;
        const uint32_t b = nl[li];
        const double4 xb = x[b];
        uint32_t offc = 0, code = 0;
        d3 dm;
        bool adjacent;
        if (NEAR) {
            nearest_image(g, xa, xb, dm, code);
            adjacent = true;
        } else {
            adjacent = nl_image(g, gx, gy, gz, cellc[b], offc, code, dm);
        }
        if (!adjacent) continue;
        const d3 x_ab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z) + dm;
        const double s2 = dot3(x_ab, x_ab);
        const double rsum = xa.w + xb.w;
        if (!(s2 < rsum * rsum * (1.0 + 1e-9) && sqrt(s2) - rsum < 0.0)) continue;
        set_bit(found, k);
        const uint8_t sl = nl_slot[li];
        if (sl == 0xffu) set_bit(fresh, k);
        else if ((row[RS(sl)] >> CT_CODE_SHIFT) != code) set_bit(recode, k);
    }
}

#define TW 64
#define TU 4
// NEAR: grid searched one cell deep (nearest_image applies; the cell code of
// list entries is only read for new contacts)
template <bool NEAR>
__device__ __forceinline__ void topology_lists_body(
    const StepArgs &a, const uint32_t *__restrict__ vals, const uint32_t *__restrict__ cell_start,
    const uint32_t *__restrict__ cell_end, const double4 *__restrict__ x,
    const double4 *__restrict__ x_s, uint32_t *__restrict__ contacts,
    uint64_t *__restrict__ cmask, uint32_t *__restrict__ err,
    const uint32_t *__restrict__ nl, uint8_t *__restrict__ nl_slot,
    uint8_t *__restrict__ slot_cand, const uint32_t *__restrict__ nl_cnt,
    const uint32_t *__restrict__ nl_pos, const uint32_t *__restrict__ cellc,
    const float *__restrict__ nl_gap, const float *__restrict__ nl_gmin,
    const float *__restrict__ nl_disp, uint32_t *__restrict__ flags_rw, uint32_t cur,
    uint32_t bid, uint32_t nblk)
{
    // the list rebuild before it did not finish (its grid barrier timed out)
    if ((*err & ERRF_BARRIER) != 0u) return;
    const uint32_t *flags = flags_rw;
    __shared__ double4 s_x[TW];
    __shared__ int s_g[TW][3];
    __shared__ uint32_t s_a[TW];
    __shared__ uint8_t s_own[TW * TU], s_k[TW * TU], s_fl[TW * TU];
    const uint32_t lane = threadIdx.x;
    const uint32_t tile = xcd_tile(bid, nblk);
    const uint32_t pos = tile * TW + lane;
    const uint32_t np = a.np;
    const Grid &g = a.grid;
    const int NCS = (int)a.nc;
    const bool valid = pos < np;
    const uint32_t a_orig = valid ? pos : 0u;  // lists are by original index
    // this step's cell from the hash (all ones outside the grid)
    const uint32_t cc = valid ? cellc[pos] : 0xffffffffu;
    const bool inside = cc != 0xffffffffu;
    if (valid && !inside) atomicOr(err, ERRF_GRID);
    const int gx = (int)(cc & 2047u), gy = (int)((cc >> 11) & 2047u), gz = (int)(cc >> 22);
    const uint32_t lcnt = inside ? nl_cnt[pos] : 0u;
    const bool list_ovf = lcnt == NL_OVF;
    // Entries to test.  On steps that rebuilt the lists, or on which a grain
    // wrapped around a periodic boundary (images may change), every entry.
    // Otherwise the row can only change by a new contact: an entry not in the
    // row whose surface gap at the rebuild is at most twice the largest
    // displacement since then (plus a rounding margin).
    const bool full = (flags[cur] & (NLF_REBUILD | NLF_WRAP)) != 0u || flags[NLF_STICKY] != 0u;
    if (full && bid == 0 && lane == 0) flags_rw[NLF_FULL] += 1u;
    const double dmax = dmax_since_rebuild(flags);
    Set128 q = {0ull, 0ull};
    uint32_t n = 0;
    if (!list_ovf && lcnt > 0) {
        if (full) {
            n = lcnt;
            q.lo = lcnt >= 64 ? ~0ull : ((1ull << lcnt) - 1ull);
            q.hi = lcnt > 64 ? (lcnt >= 128 ? ~0ull : ((1ull << (lcnt - 64)) - 1ull)) : 0ull;
        } else {
            // An entry's gap can have closed by at most the two grains'
            // displacements since the rebuild: this grain's (nl_disp, rounded
            // up) and at most sqrt(dmax) for the partner.  Bounds are rounded
            // up to float, as the stored gaps are rounded down, with a margin
            // for the rounding of the displacements themselves.
            const double marg = 1e-11 * (g.L[0] + g.L[1] + g.L[2]);
            const double da = (double)nl_disp[pos];
            const float thr = __double2float_ru((da + sqrt(dmax)) * (1.0 + 1e-9) + marg);
            // nl_gmin bounds the gaps of the entries not in the row
            if (!(nl_gmin[pos] > thr)) {
                // eight entries per iteration, their loads issued together
                for (uint32_t k0 = 0; k0 < lcnt; k0 += 8) {
                    uint8_t sl[8];
                    float gp[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        sl[u] = 0u;
                        gp[u] = INFINITY;
                        if (k0 + u < lcnt) {
                            const size_t li = (size_t)(k0 + u) * np + pos;
                            sl[u] = nl_slot[li];
                            gp[u] = nl_gap[li];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (sl[u] == 0xffu && !(gp[u] > thr)) {
                            set_bit(q, k0 + u);
                            ++n;
                        }
                    }
                }
            }
        }
    }
    const bool busy = inside && (full || list_ovf || n > 0);
    const double4 xa = busy ? x[a_orig] : make_double4(0.0, 0.0, 0.0, 0.0);
    s_x[lane] = xa;
    s_g[lane][0] = gx;
    s_g[lane][1] = gy;
    s_g[lane][2] = gz;
    s_a[lane] = a_orig;
    uint32_t incl = n;
    for (uint32_t off = 1; off < TW; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, TW);
        if (lane >= off) incl += t;
    }
    const uint32_t excl = incl - n;
    const uint32_t total = __shfl(incl, TW - 1, TW);
    Set128 found = {0ull, 0ull}, fresh = {0ull, 0ull}, recode = {0ull, 0ull};
    Set128 deal = q;  // entries not yet dealt, in list order
    for (uint32_t base = 0; base < total; base += TW * TU) {
        const uint32_t lo = excl > base ? excl : base;
        const uint32_t hi = excl + n < base + TW * TU ? excl + n : base + TW * TU;
        for (uint32_t c = lo; c < hi; ++c) {
            uint32_t k;
            if (deal.lo) {
                k = (uint32_t)(__ffsll((unsigned long long)deal.lo) - 1);
                deal.lo &= deal.lo - 1;
            } else {
                k = 64u + (uint32_t)(__ffsll((unsigned long long)deal.hi) - 1);
                deal.hi &= deal.hi - 1;
            }
            s_own[c - base] = (uint8_t)lane;
            s_k[c - base] = (uint8_t)k;
        }
        __syncthreads();
        uint32_t bq[TU];
        uint8_t sq[TU];
        double4 xb[TU];
        uint32_t cb[TU];
#pragma unroll
        for (int u = 0; u < TU; ++u) {
            const uint32_t t = lane + (uint32_t)u * TW;
            bq[u] = 0u;
            sq[u] = 0xffu;
            if (base + t < total) {
                const size_t li = (size_t)s_k[t] * np + tile * TW + s_own[t];
                bq[u] = nl[li];
                sq[u] = nl_slot[li];
            }
        }
#pragma unroll
        for (int u = 0; u < TU; ++u) {
            const uint32_t t = lane + (uint32_t)u * TW;
            if (base + t < total) {
                xb[u] = x[bq[u]];
                cb[u] = NEAR ? 0u : cellc[bq[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < TU; ++u) {
            const uint32_t t = lane + (uint32_t)u * TW;
            if (base + t < total) {
                const uint32_t o = s_own[t];
                const double4 xo = s_x[o];
                uint32_t offc = 0, code = 0;
                d3 dm;
                uint8_t fl = 0;
                bool adjacent;
                if (NEAR) {
                    nearest_image(g, xo, xb[u], dm, code);
                    adjacent = true;
                } else {
                    adjacent = nl_image(g, s_g[o][0], s_g[o][1], s_g[o][2], cb[u], offc, code, dm);
                }
                if (adjacent) {
                    const d3 x_ab = mk3(xo.x - xb[u].x, xo.y - xb[u].y, xo.z - xb[u].z) + dm;
                    const double s2 = dot3(x_ab, x_ab);
                    const double rsum = xo.w + xb[u].w;
                    if (s2 < rsum * rsum * (1.0 + 1e-9) && sqrt(s2) - rsum < 0.0) {
                        fl = 1;
                        if (sq[u] == 0xffu)
                            fl |= 2;
                        else if ((contacts[ct_slot(s_a[o], sq[u], NCS)] >> CT_CODE_SHIFT) !=
                                 code)
                            fl |= 4;
                    }
                }
                s_fl[t] = fl;
            }
        }
        __syncthreads();
        for (uint32_t c = lo; c < hi; ++c) {
            const uint8_t fl = s_fl[c - base];
            const uint32_t k = s_k[c - base];
            if (fl & 1) set_bit(found, k);
            if (fl & 2) set_bit(fresh, k);
            if (fl & 4) set_bit(recode, k);
        }
        __syncthreads();
    }
    if (!busy) return;
    const uint64_t fullmask = NCS >= 64 ? ~0ull : ((1ull << NCS) - 1ull);
    uint32_t *row = contacts + ct_base(a_orig, NCS);
    const uint64_t mask0 = cmask[a_orig];
    uint64_t mask = mask0;
    bool ovf = false;
    if (list_ovf) {
        ovf = topology_cells(g, nl_pos[a_orig], xa, gx, gy, gz, vals, cell_start, cell_end, x_s,
                             row, mask, NCS, fullmask);
    } else if (full) {
        if (any_bit(fresh) || any_bit(recode) || __popcll(found.lo) + __popcll(found.hi) > NCS)
            ovf = slot_update(g, np, pos, found, fresh, recode, nl, nl_slot, cellc, gx, gy, gz,
                              row, slot_cand + ct_base(a_orig, NCS), mask, NCS, fullmask);
    } else if (any_bit(fresh)) {
        // only new contacts were tested; if they do not all fit into empty
        // slots the full rule needs every contact found: test all entries
        const int n_new = __popcll(fresh.lo) + __popcll(fresh.hi);
        if (__popcll(mask) + n_new > NCS)
            owner_test_all<NEAR>(g, np, pos, lcnt, nl, nl_slot, cellc, x, xa, gx, gy, gz, row,
                                 found, fresh, recode);
        else
            found = fresh;
        ovf = slot_update(g, np, pos, found, fresh, recode, nl, nl_slot, cellc, gx, gy, gz, row,
                          slot_cand + ct_base(a_orig, NCS), mask, NCS, fullmask);
    }
    if (ovf) atomicOr(err, ERRF_OVERFLOW);
    if (mask != mask0) cmask[a_orig] = mask;
}


template <bool NEAR>
__global__ void __launch_bounds__(TW) k_topology_lists(
    StepArgs a, const uint32_t *__restrict__ vals, const uint32_t *__restrict__ cell_start,
    const uint32_t *__restrict__ cell_end, const double4 *__restrict__ x,
    const double4 *__restrict__ x_s, uint32_t *__restrict__ contacts,
    uint64_t *__restrict__ cmask, uint32_t *__restrict__ err,
    const uint32_t *__restrict__ nl, uint8_t *__restrict__ nl_slot,
    uint8_t *__restrict__ slot_cand, const uint32_t *__restrict__ nl_cnt,
    const uint32_t *__restrict__ nl_pos, const uint32_t *__restrict__ cellc,
    const float *__restrict__ nl_gap, const float *__restrict__ nl_gmin,
    const float *__restrict__ nl_disp, uint32_t *__restrict__ flags_rw, uint32_t cur)
{
    topology_lists_body<NEAR>(a, vals, cell_start, cell_end, x, x_s, contacts, cmask, err, nl,
                              nl_slot, slot_cand, nl_cnt, nl_pos, cellc, nl_gap, nl_gmin, nl_disp,
                              flags_rw, cur, blockIdx.x, gridDim.x);
}

// Wall contacts of one particle, contactsearch.cuh:420-701: the top wall
// (wall 0, or the plane z = L2 without walls), the fixed bottom z = origo2,
// and for non-periodic axes walls 1-4.  Returns the normal forces on walls
// 0-4 in wf0..wf4.
// The grain's velocities are read only by a grain that touches a wall.
__device__ __forceinline__ void walls_contact(const StepArgs &a, Acc &A, double4 xa,
                                              const double4 *__restrict__ vel,
                                              const double3 *__restrict__ angvel, uint32_t i,
                                              const double4 *__restrict__ w_nx,
                                              const double4 *__restrict__ w_mvfd, double &wf0,
                                              double &wf1, double &wf2, double &wf3,
                                              double &wf4)
{
    const Grid &g = a.grid;
    const uint32_t nw = a.walls.nw;
    const double radius_a = xa.w;
    wf0 = wf1 = wf2 = wf3 = wf4 = 0.0;
    const double4 wn0 = nw > 0 ? w_nx[0] : make_double4(0.0, 0.0, -1.0, g.L[2]);
    const double wv0 = nw > 0 ? w_mvfd[0].y : 0.0;
    double delta_w = wn0.w - (xa.z + radius_a);
    if (delta_w < 0.0) wf0 = law_wall(a, A, vel[i], angvel[i], radius_a, xyz(wn0), delta_w, wv0);
    delta_w = xa.z - radius_a - g.origo[2];
    if (delta_w < 0.0) (void)law_wall(a, A, vel[i], angvel[i], radius_a, mk3(0.0, 0.0, 1.0), delta_w, 0.0);
    if (g.periodic == 0) {
        const double4 wn1 = nw > 1 ? w_nx[1] : make_double4(-1.0, 0.0, 0.0, g.L[0]);
        const double4 wn2 = nw > 2 ? w_nx[2] : make_double4(1.0, 0.0, 0.0, 0.0);
        const double wv1 = nw > 1 ? w_mvfd[1].y : 0.0, wv2 = nw > 2 ? w_mvfd[2].y : 0.0;
        delta_w = wn1.w - (xa.x + radius_a);
        if (delta_w < 0.0) wf1 = law_wall(a, A, vel[i], angvel[i], radius_a, xyz(wn1), delta_w, wv1);
        delta_w = xa.x - radius_a - wn2.w;
        if (delta_w < 0.0) wf2 = law_wall(a, A, vel[i], angvel[i], radius_a, xyz(wn2), delta_w, wv2);
    }
    if (g.periodic == 0 || g.periodic == 2) {
        const double4 wn3 = nw > 3 ? w_nx[3] : make_double4(0.0, -1.0, 0.0, g.L[1]);
        const double4 wn4 = nw > 4 ? w_nx[4] : make_double4(0.0, 1.0, 0.0, 0.0);
        const double wv3 = nw > 3 ? w_mvfd[3].y : 0.0, wv4 = nw > 4 ? w_mvfd[4].y : 0.0;
        delta_w = wn3.w - (xa.y + radius_a);
        if (delta_w < 0.0) wf3 = law_wall(a, A, vel[i], angvel[i], radius_a, xyz(wn3), delta_w, wv3);
        delta_w = xa.y - radius_a - wn4.w;
        if (delta_w < 0.0) wf4 = law_wall(a, A, vel[i], angvel[i], radius_a, xyz(wn4), delta_w, wv4);
    }
}

// contactsearch.cuh:706-724
__device__ __forceinline__ void write_sums(const StepArgs &a, const Acc &A, uint32_t a_orig,
                                           double3 *__restrict__ force,
                                           double3 *__restrict__ torque,
                                           double *__restrict__ es_dot_out,
                                           double *__restrict__ ev_dot_out,
                                           double *__restrict__ es, double *__restrict__ ev,
                                           double *__restrict__ p_out,
                                           double *__restrict__ walls_force_pp, double wf0,
                                           double wf1, double wf2, double wf3, double wf4)
{
    const size_t np = a.np;
    force[a_orig] = make_double3(A.F.x, A.F.y, A.F.z);
    torque[a_orig] = make_double3(A.T.x, A.T.y, A.T.z);
    es_dot_out[a_orig] = A.es_dot;
    ev_dot_out[a_orig] = A.ev_dot;
    es[a_orig] += A.es_dot * a.dt;
    ev[a_orig] += A.ev_dot * a.dt;
    p_out[a_orig] = A.p;
    const uint32_t nw = a.walls.nw;
    if (nw > 0) walls_force_pp[a_orig] = wf0;
    if (nw > 1) walls_force_pp[a_orig + np] = wf1;
    if (nw > 2) walls_force_pp[a_orig + 2 * np] = wf2;
    if (nw > 3) walls_force_pp[a_orig + 3 * np] = wf3;
    if (nw > 4) walls_force_pp[a_orig + 4 * np] = wf4;
}

__device__ __forceinline__ Acc load_acc(const double (*s)[64], uint32_t lane)
{
    Acc A;
    A.F = mk3(s[0][lane], s[1][lane], s[2][lane]);
    A.T = mk3(s[3][lane], s[4][lane], s[5][lane]);
    A.es_dot = s[6][lane];
    A.ev_dot = s[7][lane];
    A.p = s[8][lane];
    return A;
}

__device__ __forceinline__ void store_acc(double (*s)[64], uint32_t lane, const Acc &A)
{
    s[0][lane] = A.F.x;
    s[1][lane] = A.F.y;
    s[2][lane] = A.F.z;
    s[3][lane] = A.T.x;
    s[4][lane] = A.T.y;
    s[5][lane] = A.T.z;
    s[6][lane] = A.es_dot;
    s[7][lane] = A.ev_dot;
    s[8][lane] = A.p;
}

#define FW 64  // k_forces: one wave per block
#define NREC 10

// HERTZ: contact model 3, else 2
template <bool HERTZ>
__device__ __forceinline__ void forces_body(
    const StepArgs &a, const uint32_t *__restrict__ order, const double4 *__restrict__ x,
    const double4 *__restrict__ vel, const double3 *__restrict__ angvel,
    double3 *__restrict__ force, double3 *__restrict__ torque, double *__restrict__ es_dot_out,
    double *__restrict__ ev_dot_out, double *__restrict__ es, double *__restrict__ ev,
    double *__restrict__ p_out, const double4 *__restrict__ w_nx,
    const double4 *__restrict__ w_mvfd, double *__restrict__ walls_force_pp,
    uint32_t *__restrict__ contacts, uint64_t *__restrict__ cmask,
    double *__restrict__ delta_t,
    uint8_t *__restrict__ nl_slot, uint8_t *__restrict__ slot_cand,
    const float *__restrict__ nl_gap, float *__restrict__ nl_gmin,
    const uint32_t *__restrict__ err, uint32_t bid, uint32_t nblk)
{
    // list mode: the list rebuild of this step did not finish
    if (err && (*err & ERRF_BARRIER) != 0u) return;
    __shared__ uint32_t s_a[FW];
    __shared__ uint8_t s_own[FW], s_slot[FW], s_rm[FW], s_rmk[FW];
    __shared__ double s_rec[NREC][FW];
    __shared__ double s_acc[9][FW];  // the owners' running sums (F, T, es_dot, ev_dot, p)
    const uint32_t lane = threadIdx.x;
    const uint32_t tile = xcd_tile(bid, nblk);
    const uint32_t pos = tile * FW + lane;
    const uint32_t np = a.np;
    const bool valid = pos < np;
    const uint32_t a_orig = valid ? (order ? order[pos] : pos) : 0u;
    const uint64_t mask = valid ? cmask[a_orig] : 0ull;
    const uint32_t n = (uint32_t)__popcll(mask);
    // exclusive scan of the row sizes over the wave
    uint32_t incl = n;
    for (uint32_t off = 1; off < FW; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, FW);
        if (lane >= off) incl += t;
    }
    const uint32_t excl = incl - n;
    const uint32_t total = __shfl(incl, FW - 1, FW);
    s_a[lane] = a_orig;
    const Grid &g = a.grid;
    const int NCS = (int)a.nc;

    // The sums live in LDS between rounds, which keeps the contact law's
    // registers free of them.
#pragma unroll
    for (int c = 0; c < 9; ++c) s_acc[c][lane] = 0.0;
    uint64_t deal = mask, keep = mask;
    float gfree = INFINITY;  // smallest rebuild gap among list entries freed here
    for (uint32_t base = 0; base < total; base += FW) {
        // this row's contacts with wave index in [base, base + FW)
        const uint32_t lo = excl > base ? excl : base;
        const uint32_t hi = excl + n < base + FW ? excl + n : base + FW;
        for (uint32_t c = lo; c < hi; ++c) {
            const int sl = __ffsll((unsigned long long)deal) - 1;
            deal &= deal - 1;
            s_own[c - base] = (uint8_t)lane;
            s_slot[c - base] = (uint8_t)sl;
        }
        __syncthreads();
        if (base + lane < total) {
            // interact body for one slot, contactsearch.cuh:484-559
            const uint32_t ao = s_a[s_own[lane]];
            const uint32_t sl = s_slot[lane];
            const size_t mempos = ct_slot(ao, sl, NCS);
            const double4 xa = x[ao];
            const uint32_t cw = contacts[mempos];
            const uint32_t b = cw & CT_IDX_MASK;
            const d3 dm = decode_dm(g, cw >> CT_CODE_SHIFT);
            const double4 xb = x[b];
            const d3 x_ab = mk3(xa.x - xb.x + dm.x, xa.y - xb.y + dm.y, xa.z - xb.z + dm.z);
            const double l = len3(x_ab);
            const double delta_n = l - (xa.w + xb.w);
            double *dts = delta_t + 3 * mempos;
            Contrib C;
            C.F = mk3(0.0, 0.0, 0.0);
            C.T = mk3(0.0, 0.0, 0.0);
            C.p = C.ev_n = C.ev_t = C.es = 0.0;
            uint8_t rm = 0, rmk = 0xffu;
            if (delta_n < 0.0) {
                const double4 va4 = vel[ao];
                const double3 wa4 = angvel[ao];
                double loc[3] = {dts[0], dts[1], dts[2]};
                if (HERTZ)
                    law_hertz(a, C, xyz(va4), xyz(wa4), vel[b], angvel[b], xa.w, xb.w, x_ab, l,
                              delta_n, loc);
                else
                    law_linear(a, C, xyz(va4), xyz(wa4), vel[b], angvel[b], xa.w, xb.w, x_ab, l,
                               delta_n, loc);
                dts[0] = loc[0];
                dts[1] = loc[1];
                dts[2] = loc[2];
            } else {
                contacts[mempos] = np;
                dts[0] = 0.0;
                dts[1] = 0.0;
                dts[2] = 0.0;
                rm = 1;
                if (slot_cand) {  // list mode: the list entry loses its slot
                    const uint8_t kk = slot_cand[mempos];
                    if (kk != 0xffu) {
                        nl_slot[(size_t)kk * np + tile * FW + s_own[lane]] = 0xffu;
                        slot_cand[mempos] = 0xffu;
                        rmk = kk;
                    }
                }
            }
            s_rec[0][lane] = C.F.x;
            s_rec[1][lane] = C.F.y;
            s_rec[2][lane] = C.F.z;
            s_rec[3][lane] = C.T.x;
            s_rec[4][lane] = C.T.y;
            s_rec[5][lane] = C.T.z;
            s_rec[6][lane] = C.p;
            s_rec[7][lane] = C.ev_n;
            s_rec[8][lane] = C.ev_t;
            s_rec[9][lane] = C.es;
            s_rm[lane] = rm;
            s_rmk[lane] = rmk;
        }
        __syncthreads();
        if (lo < hi) {
            Acc A = load_acc(s_acc, lane);
            for (uint32_t c = lo; c < hi; ++c) {
                const uint32_t k = c - base;
                if (s_rm[k]) {
                    keep &= ~(1ull << s_slot[k]);
                    if (s_rmk[k] != 0xffu)  // the freed list entry may touch again
                        gfree = fminf(gfree, nl_gap[(size_t)s_rmk[k] * np + pos]);
                } else {
                    Contrib C;
                    C.F = mk3(s_rec[0][k], s_rec[1][k], s_rec[2][k]);
                    C.T = mk3(s_rec[3][k], s_rec[4][k], s_rec[5][k]);
                    C.p = s_rec[6][k];
                    C.ev_n = s_rec[7][k];
                    C.ev_t = s_rec[8][k];
                    C.es = s_rec[9][k];
                    add_contrib(A, C);
                }
            }
            store_acc(s_acc, lane, A);
        }
        __syncthreads();
    }
    if (!valid) return;
    Acc A = load_acc(s_acc, lane);
    const double4 xa = x[a_orig];
    double wf0, wf1, wf2, wf3, wf4;
    walls_contact(a, A, xa, vel, angvel, a_orig, w_nx, w_mvfd, wf0, wf1, wf2, wf3, wf4);
    write_sums(a, A, a_orig, force, torque, es_dot_out, ev_dot_out, es, ev, p_out,
               walls_force_pp, wf0, wf1, wf2, wf3, wf4);
    if (keep != mask) cmask[a_orig] = keep;
    if (gfree < INFINITY && nl_gmin[a_orig] > gfree) nl_gmin[a_orig] = gfree;
}

template <bool HERTZ>
__global__ void __launch_bounds__(FW) k_forces(
    StepArgs a, const uint32_t *__restrict__ order, const double4 *__restrict__ x,
    const double4 *__restrict__ vel, const double3 *__restrict__ angvel,
    double3 *__restrict__ force, double3 *__restrict__ torque, double *__restrict__ es_dot_out,
    double *__restrict__ ev_dot_out, double *__restrict__ es, double *__restrict__ ev,
    double *__restrict__ p_out, const double4 *__restrict__ w_nx,
    const double4 *__restrict__ w_mvfd, double *__restrict__ walls_force_pp,
    uint32_t *__restrict__ contacts, uint64_t *__restrict__ cmask,
    double *__restrict__ delta_t,
    uint8_t *__restrict__ nl_slot, uint8_t *__restrict__ slot_cand,
    const float *__restrict__ nl_gap, float *__restrict__ nl_gmin,
    const uint32_t *__restrict__ err)
{
    forces_body<HERTZ>(a, order, x, vel, angvel, force, torque, es_dot_out, ev_dot_out, es, ev,
                       p_out, w_nx, w_mvfd, walls_force_pp, contacts, cmask, delta_t, nl_slot,
                       slot_cand, nl_gap, nl_gmin, err, blockIdx.x, gridDim.x);
}

// Model 1 (contactLinearViscous + capillary cohesion, no contact history):
// one thread per sorted position, cells searched every step.
__global__ void __launch_bounds__(256) k_contact_viscous(
    StepArgs a, const uint32_t *__restrict__ vals, const uint32_t *__restrict__ cell_start,
    const uint32_t *__restrict__ cell_end, const double4 *__restrict__ x_s,
    const double4 *__restrict__ vel_s, const double3 *__restrict__ angvel_s,
    double3 *__restrict__ force, double3 *__restrict__ torque, double *__restrict__ es_dot_out,
    double *__restrict__ ev_dot_out, double *__restrict__ es, double *__restrict__ ev,
    double *__restrict__ p_out, const double4 *__restrict__ w_nx,
    const double4 *__restrict__ w_mvfd, double *__restrict__ walls_force_pp,
    uint32_t *__restrict__ err)
{
    const uint32_t idx_a = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx_a >= a.np) return;
    const Grid &g = a.grid;
    const uint32_t a_orig = vals[idx_a];
    const double4 xa = x_s[idx_a];
    const double radius_a = xa.w;
    const double4 va4 = vel_s[idx_a];
    const double3 wa4 = angvel_s[idx_a];
    Acc A;
    A.F = mk3(0.0, 0.0, 0.0);
    A.T = mk3(0.0, 0.0, 0.0);
    A.es_dot = 0.0;
    A.ev_dot = 0.0;
    A.p = 0.0;
    int gx, gy, gz;
    const bool inside = grid_pos(g, xa, gx, gy, gz);
    if (!inside) atomicOr(err, ERRF_GRID);
    if (inside) {
        // findAndProcessContactsInCell, contactsearch.cuh:91-186
        for (int zd = -1; zd < 2; ++zd)
            for (int yd = -1; yd < 2; ++yd)
                for (int xd = -1; xd < 2; ++xd) {
                    int tx = gx + xd, ty = gy + yd, tz = gz + zd;
                    d3 dm = mk3(0.0, 0.0, 0.0);
                    uint32_t code = 0;
                    if (dist_mod(g, tx, ty, tz, dm, code) == -1) continue;
                    uint32_t cid = (uint32_t)tx + (uint32_t)ty * g.num[0] +
                                   (g.num[0] * g.num[1]) * (uint32_t)tz;
                    uint32_t start = cell_start[cid];
                    if (start == EMPTY_CELL) continue;
                    uint32_t end = cell_end[cid];
                    for (uint32_t idx_b = start; idx_b < end; ++idx_b) {
                        if (idx_b == idx_a) continue;
                        double4 xb = x_s[idx_b];
                        double radius_b = xb.w;
                        double kappa = a.prm.kappa;
                        d3 x_ab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z) + dm;
                        double l = len3(x_ab);
                        double delta_ab = l - (radius_a + radius_b);
                        if (delta_ab < 0.0)
                            law_linear_viscous(a, A, va4, vel_s[idx_b], wa4, angvel_s[idx_b],
                                               radius_a, radius_b, x_ab, l, delta_ab, kappa);
                        else if (delta_ab < a.prm.db)
                            law_capillary(a, A, radius_a, radius_b, delta_ab, x_ab, l, kappa);
                    }
                }
    }
    double wf0, wf1, wf2, wf3, wf4;
    walls_contact(a, A, xa, vel_s, angvel_s, idx_a, w_nx, w_mvfd, wf0, wf1, wf2, wf3, wf4);
    write_sums(a, A, a_orig, force, torque, es_dot_out, ev_dot_out, es, ev, p_out,
               walls_force_pp, wf0, wf1, wf2, wf3, wf4);
}

void launch_topology(hipStream_t s, const StepArgs &a, const DemBuffers &b, bool nlist,
                     bool near, uint32_t cur)
{
    if (a.prm.contactmodel == 1) return;  // model 1 keeps no contact history
    if (nlist) {
        const TopoTask t = topology_task(a, b, cur);
        if (near)
            hipLaunchKernelGGL(k_topology_lists<true>, dim3(topology_blocks(a.np)), dim3(TW), 0, s,
                               t.a, t.vals, t.cell_start, t.cell_end, t.x, t.x_s, t.contacts,
                               t.cmask, t.err, t.nl, t.nl_slot, t.slot_cand, t.nl_cnt, t.nl_pos,
                               t.cellc, t.nl_gap, t.nl_gmin, t.nl_disp, t.flags, t.cur);
        else
            hipLaunchKernelGGL(k_topology_lists<false>, dim3(topology_blocks(a.np)), dim3(TW), 0,
                               s, t.a, t.vals, t.cell_start, t.cell_end, t.x, t.x_s, t.contacts,
                               t.cmask, t.err, t.nl, t.nl_slot, t.slot_cand, t.nl_cnt, t.nl_pos,
                               t.cellc, t.nl_gap, t.nl_gmin, t.nl_disp, t.flags, t.cur);
    } else
        hipLaunchKernelGGL(k_topology_cells, dim3((a.np + 255) / 256), dim3(256), 0, s, a,
                           b.vals, b.cell_start, b.cell_end, b.x_s, b.contacts, b.cmask,
                           b.errflags);
}

void launch_contact(hipStream_t s, const StepArgs &a, const DemBuffers &b, bool nlist)
{
    if (a.prm.contactmodel == 1) {
        hipLaunchKernelGGL(k_contact_viscous, dim3((a.np + 255) / 256), dim3(256), 0, s, a,
                           b.vals, b.cell_start, b.cell_end, b.x_s, b.vel_s, b.angvel_s, b.force,
                           b.torque, b.es_dot, b.ev_dot, b.es, b.ev, b.p, b.w_nx, b.w_mvfd,
                           b.walls_force_pp, b.errflags);
        return;
    }
    const ForcesTask t = forces_task(a, b, nlist);
    const uint32_t nw = forces_blocks(a.np);
    if (a.prm.contactmodel == 3)
        hipLaunchKernelGGL(k_forces<true>, dim3(nw), dim3(FW), 0, s, t.a, t.order, t.x, t.vel,
                           t.angvel, t.force, t.torque, t.es_dot, t.ev_dot, t.es, t.ev, t.p,
                           t.w_nx, t.w_mvfd, t.walls_force_pp, t.contacts, t.cmask, t.delta_t,
                           t.nl_slot, t.slot_cand, t.nl_gap, t.nl_gmin, t.err);
    else
        hipLaunchKernelGGL(k_forces<false>, dim3(nw), dim3(FW), 0, s, t.a, t.order, t.x, t.vel,
                           t.angvel, t.force, t.torque, t.es_dot, t.ev_dot, t.es, t.ev, t.p,
                           t.w_nx, t.w_mvfd, t.walls_force_pp, t.contacts, t.cmask, t.delta_t,
                           t.nl_slot, t.slot_cand, t.nl_gap, t.nl_gmin, t.err);
}

// ------------------------------------------------------------------
// A11: bonds
// ------------------------------------------------------------------

// bondsLinear body, src/cohesion.cuh:24-217, one thread per bond
__device__ __forceinline__ void bond_compute_body(
    const StepArgs &a, const uint2 *__restrict__ bonds, uint8_t *__restrict__ broken,
    double4 *__restrict__ bdelta, double4 *__restrict__ bomega, const double4 *__restrict__ x,
    const double4 *__restrict__ vel, const double3 *__restrict__ angvel,
    double4 *__restrict__ bf, double4 *__restrict__ bt, uint8_t *__restrict__ applied,
    uint32_t bid)
{
    uint32_t idx = bid * blockDim.x + threadIdx.x;
    if (idx >= a.nb0) return;
    applied[idx] = 0;
    if (broken[idx]) return;
    const Params &P = a.prm;
    const double dt = a.dt;
    const uint2 bond = bonds[idx];
    const double4 d0 = bdelta[idx], o0 = bomega[idx];
    d3 delta0_t = xyz(d0);
    const double delta0_n = d0.w;
    d3 omega0_t = xyz(o0);
    const double omega0_n = o0.w;
    const double4 x_i = x[bond.x], x_j = x[bond.y];
    const d3 vel_i = xyz(vel[bond.x]), vel_j = xyz(vel[bond.y]);
    const d3 angvel_i = xyz(angvel[bond.x]), angvel_j = xyz(angvel[bond.y]);
    const double R_bar = P.lambda_bar * fmin(x_i.w, x_j.w);
    const double Ab = PI_REF * R_bar * R_bar;
    const double I = 0.25 * PI_REF * R_bar * R_bar * R_bar * R_bar;
    const double J = I * 2.0;
    const d3 xv = mk3(x_i.x - x_j.x, x_i.y - x_j.y, x_i.z - x_j.z);
    const double x_length = len3(xv);
    const double overlap = fmin(0.0, x_length - (x_i.w + x_j.w));
    const d3 n = xv / x_length;
    delta0_t = delta0_t - dot3(n, delta0_t) * n;
    const d3 ddelta = dt * (((vel_i - vel_j) + (x_i.w + overlap / 2.0) * cross3(n, angvel_i)) +
                            (x_j.w + overlap / 2.0) * cross3(n, angvel_j));
    const double ddelta_n = -dot3(ddelta, n);
    const double delta_n = delta0_n + ddelta_n;
    const d3 ddelta_t = ddelta - dot3(n, ddelta) * n;
    const d3 delta_t = delta0_t + ddelta_t;
    const d3 f_n = (P.k_n * Ab * delta_n + P.gamma_n * ddelta_n / dt) * n;
    const d3 f_t = (-P.k_t * Ab) * delta_t - (P.gamma_t * ddelta_t) / dt;
    const d3 f = f_n + f_t;
    const d3 mn = -n;
    omega0_t = omega0_t - dot3(omega0_t, mn) * mn;
    const d3 domega = dt * (angvel_j - angvel_i);
    const double domega_n = dot3(domega, mn);
    const double omega_n = omega0_n + domega_n;
    const d3 domega_t = domega - domega_n * mn;
    const d3 omega_t = omega0_t + domega_t;
    const d3 t_n = (-P.k_t * J * omega_n) * mn;
    const d3 t_t = (-P.k_n * I) * omega_t;
    const d3 t_j = t_n + t_t;
    const double sigma_max = len3(f_n) / Ab + len3(t_t) * R_bar / I;
    const double tau_max = len3(f_t) / Ab + len3(t_n) * R_bar / J;
    if (sigma_max >= P.sigma_b || tau_max >= P.tau_b) {
        broken[idx] = 1;
        return;
    }
    bdelta[idx] = make_double4(delta_t.x, delta_t.y, delta_t.z, delta_n);
    bomega[idx] = make_double4(omega_t.x, omega_t.y, omega_t.z, omega_n);
    bf[idx] = make_double4(f.x, f.y, f.z, 0.0);
    bt[idx] = make_double4(t_j.x, t_j.y, t_j.z, 0.0);
    applied[idx] = 1;
}

__global__ void __launch_bounds__(256) k_bond_compute(
    StepArgs a, const uint2 *__restrict__ bonds, uint8_t *__restrict__ broken,
    double4 *__restrict__ bdelta, double4 *__restrict__ bomega, const double4 *__restrict__ x,
    const double4 *__restrict__ vel, const double3 *__restrict__ angvel,
    double4 *__restrict__ bf, double4 *__restrict__ bt, uint8_t *__restrict__ applied)
{
    bond_compute_body(a, bonds, broken, bdelta, bomega, x, vel, angvel, bf, bt, applied,
                      blockIdx.x);
}

// Per bonded particle: force/torque updates of cohesion.cuh:220-230 applied in
// ascending bond order.  ent = bond index << 1 | (1 if the particle is bond.y).
__device__ __forceinline__ void bond_gather_body(
    uint32_t nbp, const uint32_t *__restrict__ plist, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ ent, const double4 *__restrict__ bf,
    const double4 *__restrict__ bt, const uint8_t *__restrict__ applied,
    double3 *__restrict__ force, double3 *__restrict__ torque, uint32_t bid)
{
    uint32_t k = bid * blockDim.x + threadIdx.x;
    if (k >= nbp) return;
    uint32_t pi = plist[k];
    double3 F = force[pi], T = torque[pi];
    for (uint32_t e = off[k]; e < off[k + 1]; ++e) {
        uint32_t v = ent[e];
        uint32_t b = v >> 1;
        if (!applied[b]) continue;
        double4 f = bf[b], t = bt[b];
        if ((v & 1u) == 0) {
            F.x = F.x + f.x; F.y = F.y + f.y; F.z = F.z + f.z;
            T.x = T.x - t.x; T.y = T.y - t.y; T.z = T.z - t.z;
        } else {
            F.x = F.x - f.x; F.y = F.y - f.y; F.z = F.z - f.z;
            T.x = T.x + t.x; T.y = T.y + t.y; T.z = T.z + t.z;
        }
    }
    force[pi] = F;
    torque[pi] = T;
}

__global__ void __launch_bounds__(256) k_bond_gather(
    uint32_t nbp, const uint32_t *__restrict__ plist, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ ent, const double4 *__restrict__ bf,
    const double4 *__restrict__ bt, const uint8_t *__restrict__ applied,
    double3 *__restrict__ force, double3 *__restrict__ torque)
{
    bond_gather_body(nbp, plist, off, ent, bf, bt, applied, force, torque, blockIdx.x);
}

void launch_bonds(hipStream_t s, const StepArgs &a, const DemBuffers &b, bool gather)
{
    if (a.nb0 == 0) return;
    const BondTask t = bond_task(a, b);
    hipLaunchKernelGGL(k_bond_compute, dim3((a.nb0 + 255) / 256), dim3(256), 0, s, t.a, t.bonds,
                       t.broken, t.bdelta, t.bomega, t.x, t.vel, t.angvel, t.bf, t.bt, t.applied);
    if (gather && b.nbp)
        hipLaunchKernelGGL(k_bond_gather, dim3((b.nbp + 255) / 256), dim3(256), 0, s, b.nbp,
                           b.bp_list, b.bp_off, b.bp_ent, b.bond_f, b.bond_t, b.bond_applied,
                           b.force, b.torque);
}

// ------------------------------------------------------------------
// A12-A13: integrate, wall force reduction, walls
// ------------------------------------------------------------------

// Block tree reduction of summation(), src/integration.cuh:309-341 (256 wide).
__device__ __forceinline__ double block_sum_256(double v, double *cache)
{
    uint32_t t = threadIdx.x;
    cache[t] = v;
    __syncthreads();
    for (uint32_t i = 128; i != 0; i /= 2) {
        if (t < i) cache[t] += cache[t + i];
        __syncthreads();
    }
    return cache[0];
}

// findShearStressOnFixedMovingParticles + summation, src/integration.cuh:457-484
__global__ void __launch_bounds__(256) k_shear_stress(StepArgs a, const double4 *__restrict__ x,
                                                      const double4 *__restrict__ vel,
                                                      const double3 *__restrict__ force,
                                                      double *__restrict__ tau_partial)
{
    __shared__ double cache[256];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double temp = 0.0;
    if (i < a.np) {
        double z = x[i].z, fixvel = vel[i].w;
        double f_x = 0.0;
        if (fixvel > 0.0 && z > a.grid.L[2] * 0.5) f_x = force[i].x;
        temp += f_x / (a.grid.L[0] * a.grid.L[1]);
    }
    double s = block_sum_256(temp, cache);
    if (threadIdx.x == 0) tau_partial[blockIdx.x] = s;
}

// Serial sum over block partials (the order of integration.cuh:64-66).
// Left-to-right sum part[0] + part[1] + ... (the reference's serial loop,
// src/integration.cuh:361-364), one dependent add after another.  The
// block stages chunks of the partials in LDS with coalesced loads so that
// thread 0 only waits on the adds, not on global-memory latency.  Result
// valid in thread 0; every thread of the block must call it.
#define SUM_THREADS 256
#define SUM_CHUNK (SUM_THREADS * 8)
__device__ double serial_sum_staged(const double *__restrict__ part, uint32_t n, double *lds)
{
    // Terms equal to +0.0 are skipped: the running sum starts at +0.0 and can
    // never become -0.0 (a sum is -0.0 only when both addends are), and
    // s + (+0.0) == s for every other s, so the result is the left-to-right
    // sum bit for bit.  Most partials of a wall force are zero (only grains
    // touching the wall contribute), which leaves lane 0 a short chain.
    __shared__ uint32_t nzm[SUM_THREADS / 4];  // one byte per thread
    double s = 0.0;
    const uint32_t t = threadIdx.x;
    for (uint32_t base = 0; base < n; base += SUM_CHUNK) {
        const uint32_t m = n - base < SUM_CHUNK ? n - base : SUM_CHUNK;
        for (uint32_t j = t; j < m; j += SUM_THREADS) lds[j] = part[base + j];
        __syncthreads();
        uint32_t bits = 0;
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t j = 8u * t + k;
            if (j < m && __double_as_longlong(lds[j]) != 0ll) bits |= 1u << k;
        }
        ((uint8_t *)nzm)[t] = (uint8_t)bits;
        __syncthreads();
        if (t == 0) {
            for (uint32_t q = 0; q < SUM_THREADS; q += 4) {
                uint32_t w = nzm[q / 4u];
                while (w) {
                    const uint32_t bit = (uint32_t)__ffs(w) - 1u;
                    w &= w - 1u;
                    s += lds[8u * q + bit];  // bit = 8 (thread - q) + k
                }
            }
        }
        __syncthreads();
    }
    return s;
}

__global__ void __launch_bounds__(SUM_THREADS) k_serial_sum(const double *__restrict__ part,
                                                            uint32_t n, double *__restrict__ out)
{
    __shared__ double lds[SUM_CHUNK];
    const double s = serial_sum_staged(part, n, lds);
    if (threadIdx.x == 0) *out = s;
}

void launch_shear_stress(hipStream_t s, const StepArgs &a, const DemBuffers &b)
{
    uint32_t nb = (a.np + 255) / 256;
    hipLaunchKernelGGL(k_shear_stress, dim3(nb), dim3(256), 0, s, a, b.x, b.vel, b.force,
                       b.tau_partial);
    hipLaunchKernelGGL(k_serial_sum, dim3(1), dim3(SUM_THREADS), 0, s, b.tau_partial, nb,
                       b.tau_eff);
}

// Bond contributions of entries [e0, e1) of one particle, in ascending bond
// order (what k_bond_gather adds).
__device__ __forceinline__ void bond_sum(const BondGather &bg, uint32_t e0, uint32_t e1,
                                         double3 &F, double3 &T)
{
    for (uint32_t e = e0; e < e1; ++e) {
        const uint32_t v = bg.ent[e];
        const uint32_t b = v >> 1;
        if (!bg.applied[b]) continue;
        const double4 f = bg.f[b], t = bg.t[b];
        if ((v & 1u) == 0) {
            F.x = F.x + f.x; F.y = F.y + f.y; F.z = F.z + f.z;
            T.x = T.x - t.x; T.y = T.y - t.y; T.z = T.z - t.z;
        } else {
            F.x = F.x - f.x; F.y = F.y - f.y; F.z = F.z - f.z;
            T.x = T.x + t.x; T.y = T.y + t.y; T.z = T.z + t.z;
        }
    }
}

// The next step's cell hash of a new position (k_hash_nl's work done where
// the position is formed).  Thread 0 clears the current step's rebuild flag,
// which no later kernel of this step reads.  A grain outside the grid gets
// the all-ones cell code; the next step's topology reports it.
__device__ __forceinline__ void next_hash(const StepArgs &a, const NextHash &nh, uint32_t i,
                                          double4 xi, bool wrapped)
{
    const Grid &g = a.grid;
    if (i == 0) nh.flags[nh.cur ^ 1u] = 0u;
    const double fx = floor((xi.x - g.origo[0]) / (g.L[0] / g.num[0]));
    const double fy = floor((xi.y - g.origo[1]) / (g.L[1] / g.num[1]));
    const double fz = floor((xi.z - g.origo[2]) / (g.L[2] / g.num[2]));
    uint32_t c = 0, packed = 0xffffffffu;
    if (fx >= 0.0 && fx < (double)g.num[0] && fy >= 0.0 && fy < (double)g.num[1] && fz >= 0.0 &&
        fz < (double)g.num[2]) {
        const uint32_t ix = (uint32_t)fx, iy = (uint32_t)fy, iz = (uint32_t)fz;
        c = (iz * g.num[1]) * g.num[0] + iy * g.num[0] + ix;
        packed = ix | (iy << 11) | (iz << 22);
    }
    nh.keys[i] = c;
    nh.vals[i] = i;
    nh.cellc[i] = packed;
    const double d2 = disp2_since_rebuild(g, xi, nh.x_ref[i]);
    nh.disp[i] = __double2float_ru(sqrt(d2));
    bool need = !(d2 <= nh.disp2);
    if (i == 0 && nh.flags[NLF_STICKY] != 0u) need = true;
    record_displacement(nh.flags, nh.cur, need, wrapped, d2);
}

// integrate (TY3), src/integration.cuh:14-305, in original order (the
// reference reads the *_sorted copies, which hold the same values), with the
// wall-0 force partial of summation() reduced over the same 256-particle block.
__device__ __forceinline__ void integrate_body(
    const StepArgs &a, const StepClock *__restrict__ clock, int change_velocity_state,
    double velocity_factor, double tau_x,
    double4 *__restrict__ x, double4 *__restrict__ vel, double3 *__restrict__ angvel,
    double3 *__restrict__ force, double3 *__restrict__ torque,
    double3 *__restrict__ angpos, double3 *__restrict__ acc_a, double3 *__restrict__ angacc_a,
    double3 *__restrict__ xyzsum_a, const double4 *__restrict__ w_mvfd,
    const double *__restrict__ tau_eff, const double *__restrict__ walls_force_pp,
    double *__restrict__ wall_partial, const BondGather &bg, const NextHash &nh, uint32_t bid)
{
    __shared__ double cache[256];
    const uint32_t i = bid * blockDim.x + threadIdx.x;
    if (a.walls.nw > 0) {
        double temp = 0.0;
        if (i < a.np) temp += walls_force_pp[i];
        double s = block_sum_256(temp, cache);
        if (threadIdx.x == 0) wall_partial[bid] = s;
    }
    if (i >= a.np) return;
    const Grid &g = a.grid;
    const Params &P = a.prm;
    const double dt = a.dt;
    int wall0mode = 0;  // 0 when nw == 0 (the reference reads it uninitialised)
    double wall0mass = 0.0;
    if (a.walls.nw > 0) {
        wall0mode = a.walls.wmode[0];
        wall0mass = w_mvfd[0].x;
    }
    double tau_eff_x = 0.0;
    if (a.walls.nw > 0 && wall0mode == 3) tau_eff_x = *tau_eff;
    double3 f4 = force[i], t4 = torque[i];
    if (bg.off) {
        // bond forces in ascending bond order (cohesion.cuh:220-226, D3),
        // added before integration as the separate gather would
        const uint32_t e0 = bg.off[i], e1 = bg.off[i + 1];
        if (e0 < e1) {
            bond_sum(bg, e0, e1, f4, t4);
            force[i] = f4;
            torque[i] = t4;
        }
    }
    const double3 ap = angpos[i];
    const double4 xo = x[i], vo = vel[i];
    const double3 wo = angvel[i];
    const double3 xs = xyzsum_a[i];
    double acc0[3] = {0.0, 0.0, 0.0}, angacc0[3] = {0.0, 0.0, 0.0};
    if (clock->iter != 0) {
        const double3 a0 = acc_a[i], w0 = angacc_a[i];
        acc0[0] = a0.x; acc0[1] = a0.y; acc0[2] = a0.z;
        angacc0[0] = w0.x; angacc0[1] = w0.y; angacc0[2] = w0.z;
    }
    const double radius = xo.w;
    const double m = 4.0 / 3.0 * PI_REF * radius * radius * radius * P.rho;
    double acc[3], angacc[3];
    acc[0] = f4.x / m + P.g[0];
    acc[1] = f4.y / m + P.g[1];
    acc[2] = f4.z / m + P.g[2];
    angacc[0] = t4.x * 1.0 / (2.0 / 5.0 * m * radius * radius);
    angacc[1] = t4.y * 1.0 / (2.0 / 5.0 * m * radius * radius);
    angacc[2] = t4.z * 1.0 / (2.0 / 5.0 * m * radius * radius);
    if (wall0mode == 3 && vo.w > 0.0001 && xo.z > g.L[2] * 0.5) {
        const double f_tau_x = (tau_x + tau_eff_x) * g.L[0] * g.L[1];
        acc[0] = f_tau_x / wall0mass;
        acc[1] = 0.0;
        acc[2] -= P.g[2];
        angacc[0] = angacc[1] = angacc[2] = 0.0;
    } else if (vo.w > 0.0001 && vo.w < 10.0) {
        acc[0] = 0.0;
        acc[1] = 0.0;
        acc[2] -= P.g[2];
        angacc[0] = angacc[1] = angacc[2] = 0.0;
    }
    if (vo.w >= 10.0) {
        acc[1] = 0.0;
        angacc[0] = 0.0;
        angacc[2] = 0.0;
    }
    if (vo.w < -0.0001) {
        acc[0] = acc[1] = acc[2] = 0.0;
        angacc[0] = angacc[1] = angacc[2] = 0.0;
    }
    const double pos[3] = {xo.x, xo.y, xo.z}, v[3] = {vo.x, vo.y, vo.z};
    const double w[3] = {wo.x, wo.y, wo.z}, apo[3] = {ap.x, ap.y, ap.z};
    double xsum[3] = {xs.x, xs.y, xs.z};
    double xn[3], vn[3], apn[3], avn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double dacc = (acc[c] - acc0[c]) / dt;
        const double dang = (angacc[c] - angacc0[c]) / dt;
        xn[c] = pos[c] + v[c] * dt + 0.5 * acc[c] * dt * dt + 1.0 / 6.0 * dacc * dt * dt * dt;
        vn[c] = v[c] + acc[c] * dt + 0.5 * dacc * dt * dt;
        apn[c] = apo[c] + w[c] * dt + 0.5 * angacc[c] * dt * dt + 1.0 / 6.0 * dang * dt * dt * dt;
        avn[c] = w[c] + angacc[c] * dt + 0.5 * dang * dt * dt;
        xsum[c] += v[c] * dt + 0.5 * acc[c] * dt * dt + 1.0 / 6.0 * dacc * dt * dt * dt;
    }
    bool wrapped = false;
    if (g.periodic == 1) {
        if (xn[0] < g.origo[0]) { xn[0] += g.L[0]; wrapped = true; }
        if (xn[0] > g.L[0]) { xn[0] -= g.L[0]; wrapped = true; }
        if (xn[1] < g.origo[1]) { xn[1] += g.L[1]; wrapped = true; }
        if (xn[1] > g.L[1]) { xn[1] -= g.L[1]; wrapped = true; }
    } else if (g.periodic == 2) {
        if (xn[0] < g.origo[0]) { xn[0] += g.L[0]; wrapped = true; }
        if (xn[0] > g.L[0]) { xn[0] -= g.L[0]; wrapped = true; }
    }
    if (vo.w > 5.0 && vo.w < 10.0) {
        if (change_velocity_state == 1) vn[0] *= velocity_factor;
        else if (change_velocity_state == -1) vn[0] /= velocity_factor;
    }
    xyzsum_a[i] = make_double3(xsum[0], xsum[1], xsum[2]);
    acc_a[i] = make_double3(acc[0], acc[1], acc[2]);
    angacc_a[i] = make_double3(angacc[0], angacc[1], angacc[2]);
    angvel[i] = make_double3(avn[0], avn[1], avn[2]);
    vel[i] = make_double4(vn[0], vn[1], vn[2], vo.w);
    angpos[i] = make_double3(apn[0], apn[1], apn[2]);
    const double4 x_new = make_double4(xn[0], xn[1], xn[2], xo.w);
    x[i] = x_new;
    if (nh.keys) next_hash(a, nh, i, x_new, wrapped);
}

__global__ void __launch_bounds__(256) k_integrate(
    StepArgs a, const StepClock *__restrict__ clock, int change_velocity_state, double velocity_factor,
    double tau_x,
    double4 *__restrict__ x, double4 *__restrict__ vel, double3 *__restrict__ angvel,
    double3 *__restrict__ force, double3 *__restrict__ torque,
    double3 *__restrict__ angpos, double3 *__restrict__ acc_a, double3 *__restrict__ angacc_a,
    double3 *__restrict__ xyzsum_a, const double4 *__restrict__ w_mvfd,
    const double *__restrict__ tau_eff, const double *__restrict__ walls_force_pp,
    double *__restrict__ wall_partial, BondGather bg, NextHash nh)
{
    integrate_body(a, clock, change_velocity_state, velocity_factor, tau_x, x, vel, angvel, force,
                   torque, angpos, acc_a, angacc_a, xyzsum_a, w_mvfd, tau_eff, walls_force_pp,
                   wall_partial, bg, nh, blockIdx.x);
}

void launch_integrate(hipStream_t s, const StepArgs &a, const DemBuffers &b,
                      int change_velocity_state, double velocity_factor, double tau_x,
                      bool gather_bonds, const NextHash &nh)
{
    const uint32_t nb = (a.np + 255) / 256;
    const IntegrateTask t = integrate_task(a, b, change_velocity_state, velocity_factor, tau_x,
                                           gather_bonds, nh);
    hipLaunchKernelGGL(k_integrate, dim3(nb), dim3(256), 0, s, t.a, t.clock,
                       t.change_velocity_state, t.velocity_factor, t.tau_x, t.x, t.vel, t.angvel,
                       t.force, t.torque, t.angpos, t.acc, t.angacc, t.xyzsum, t.w_mvfd,
                       t.tau_eff, t.walls_force_pp, t.wall_partial, t.bg, t.nh);
}

// integrateWalls, src/integration.cuh:344-450.  One wave: lane 0 forms the
// serial sum over the block partials (the reference's order) once and every
// wall thread uses it (the reference sums wall 0's partials for each wall).
__device__ __forceinline__ void integrate_walls_body(
    const StepArgs &a, const StepClock *__restrict__ clock, double4 *__restrict__ w_nx,
    double4 *__restrict__ w_mvfd, double *__restrict__ w_acc,
    const double *__restrict__ partial, uint32_t nparts)
{
    const long long iter = clock->iter;
    const double t_current = clock->time;
    __shared__ double lds[SUM_CHUNK];
    __shared__ double total;
    const double s = serial_sum_staged(partial, nparts, lds);
    if (threadIdx.x == 0) total = s;
    __syncthreads();
    uint32_t idx = threadIdx.x;
    if (idx >= a.walls.nw) return;
    int wmode = a.walls.wmode[idx];
    if (wmode == 0) return;
    double4 nx = w_nx[idx], mvfd = w_mvfd[idx];
    const double dt = a.dt;
    double acc0 = iter == 0 ? 0.0 : w_acc[idx];
    mvfd.z = total;
    const double sigma_0 = mvfd.w + a.prm.sigma0_A * sin(2.0 * PI_REF * a.prm.sigma0_f * t_current);
    const double N = -sigma_0 * a.grid.L[0] * a.grid.L[1];
    double acc = (mvfd.z + N) / mvfd.x;
    if (wmode == 2) acc = 0.0;
    const double dacc_dt = (acc - acc0) / dt;
    nx.w += mvfd.y * dt + 0.5 * acc * dt * dt + 1.0 / 6.0 * dacc_dt * dt * dt * dt;
    mvfd.y += acc * dt + 0.5 * dacc_dt * dt * dt;
    w_nx[idx] = nx;
    w_mvfd[idx] = mvfd;
    w_acc[idx] = acc;
}

__global__ void __launch_bounds__(SUM_THREADS) k_integrate_walls(
    StepArgs a, const StepClock *__restrict__ clock, double4 *__restrict__ w_nx,
    double4 *__restrict__ w_mvfd, double *__restrict__ w_acc,
    const double *__restrict__ partial, uint32_t nparts)
{
    integrate_walls_body(a, clock, w_nx, w_mvfd, w_acc, partial, nparts);
}

void launch_integrate_walls(hipStream_t s, const StepArgs &a, const DemBuffers &b)
{
    const WallsTask t = walls_task(a, b);
    hipLaunchKernelGGL(k_integrate_walls, dim3(1), dim3(SUM_THREADS), 0, s, t.a, t.clock, t.w_nx,
                       t.w_mvfd, t.w_acc, t.partial, t.nparts);
}

// iter += 1, time += dt: the host's time loop update (src/device.cu:2541-2545)
__global__ void k_clock_advance(StepClock *__restrict__ clock, double dt)
{
    if (threadIdx.x == 0) {
        clock->iter += 1;
        clock->time += dt;
    }
}

void launch_clock_advance(hipStream_t s, StepClock *clock, double dt)
{
    hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, s, clock, dt);
}

// ------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_count_contacts(const uint64_t *__restrict__ cmask,
                                                        uint32_t np,
                                                        unsigned long long *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c = i < np ? (unsigned long long)__popcll(cmask[i]) : 0ull;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    if ((threadIdx.x & 63u) == 0 && c) atomicAdd(out, c);
}

void launch_count_contacts(hipStream_t s, const uint64_t *cmask, uint32_t np,
                           unsigned long long *out)
{
    if (!np) return;
    hipLaunchKernelGGL(k_count_contacts, dim3((np + 255) / 256), dim3(256), 0, s, cmask, np, out);
}

// Occupancy bits from a contact row (after a host write of contact history);
// empty slots get delta_t = 0, the state the reference's interact leaves them in.
__global__ void __launch_bounds__(256) k_rebuild_mask(uint32_t np, uint32_t nc,
                                                      const uint32_t *__restrict__ contacts,
                                                      uint64_t *__restrict__ cmask,
                                                      double *__restrict__ delta_t)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    uint64_t m = 0;
    for (uint32_t k = 0; k < nc; ++k) {
        size_t pos = ct_slot(i, k, nc);
        if (contacts[pos] != np) {
            m |= 1ull << k;
        } else {
            delta_t[3 * pos] = 0.0;
            delta_t[3 * pos + 1] = 0.0;
            delta_t[3 * pos + 2] = 0.0;
        }
    }
    cmask[i] = m;
}

void launch_rebuild_mask(hipStream_t s, const DemBuffers &b, uint32_t np, uint32_t nc)
{
    if (!np) return;
    hipLaunchKernelGGL(k_rebuild_mask, dim3((np + 255) / 256), dim3(256), 0, s, np, nc, b.contacts,
                       b.cmask, b.delta_t);
}

// ------------------------------------------------------------------
// Multi-box launches (ensembles): task records and kernels
// ------------------------------------------------------------------

RebuildTask rebuild_task(const StepArgs &a, const DemBuffers &b, uint32_t ncells, int bits,
                         const uint32_t *cond, uint32_t spin_limit)
{
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    RebuildTask t;
    t.a = a;
    RebuildArgs &r = t.r;
    r.cond = cond;
    r.keys[0] = b.keys;
    r.keys[1] = b.keys_tmp;
    r.vals[0] = b.vals;
    r.vals[1] = b.vals_tmp;
    r.hist = b.hist;
    r.part = b.rb_part;
    r.n = a.np;
    r.ntiles = (a.np + RS_TILE - 1) / RS_TILE;
    r.passes = (uint32_t)passes;
    r.ncells = ncells;
    r.cell_start = b.cell_start;
    r.cell_end = b.cell_end;
    r.x = b.x;
    r.vel = b.vel;
    r.angvel = b.angvel;
    r.x_s = b.x_s;
    r.vel_s = b.vel_s;
    r.angvel_s = b.angvel_s;
    r.bar = b.rb_bar;
    r.err = b.errflags;
    r.spin_limit = spin_limit ? spin_limit : RB_SPIN_LIMIT;
    return t;
}

NlBuildTask nl_build_task(const StepArgs &a, const NLArgs &n, const DemBuffers &b,
                          const uint32_t *cond)
{
    NlBuildTask t;
    t.a = a;
    t.n = n;
    t.cond = cond;
    t.vals = b.vals;
    t.cell_start = b.cell_start;
    t.cell_end = b.cell_end;
    t.x_s = b.x_s;
    t.nl = b.nl;
    t.nl_cnt = b.nl_cnt;
    t.nl_pos = b.nl_pos;
    t.x_ref = b.x_ref;
    t.flags = b.nl_flags;
    t.contacts = b.contacts;
    t.cmask = b.cmask;
    t.nl_slot = b.nl_slot;
    t.slot_cand = b.slot_cand;
    t.nl_gap = b.nl_gap;
    t.nl_gmin = b.nl_gmin;
    t.err = b.errflags;
    return t;
}

TopoTask topology_task(const StepArgs &a, const DemBuffers &b, uint32_t cur)
{
    TopoTask t;
    t.a = a;
    t.vals = b.vals;
    t.cell_start = b.cell_start;
    t.cell_end = b.cell_end;
    t.x = b.x;
    t.x_s = b.x_s;
    t.contacts = b.contacts;
    t.cmask = b.cmask;
    t.err = b.errflags;
    t.nl = b.nl;
    t.nl_slot = b.nl_slot;
    t.slot_cand = b.slot_cand;
    t.nl_cnt = b.nl_cnt;
    t.nl_pos = b.nl_pos;
    t.cellc = b.cellc;
    t.nl_gap = b.nl_gap;
    t.nl_gmin = b.nl_gmin;
    t.nl_disp = b.nl_disp;
    t.flags = b.nl_flags;
    t.cur = cur;
    return t;
}

ForcesTask forces_task(const StepArgs &a, const DemBuffers &b, bool nlist)
{
    ForcesTask t;
    t.a = a;
    t.order = nlist ? nullptr : b.vals;  // list mode runs in original order
    t.x = b.x;
    t.vel = b.vel;
    t.angvel = b.angvel;
    t.force = b.force;
    t.torque = b.torque;
    t.es_dot = b.es_dot;
    t.ev_dot = b.ev_dot;
    t.es = b.es;
    t.ev = b.ev;
    t.p = b.p;
    t.w_nx = b.w_nx;
    t.w_mvfd = b.w_mvfd;
    t.walls_force_pp = b.walls_force_pp;
    t.contacts = b.contacts;
    t.cmask = b.cmask;
    t.delta_t = b.delta_t;
    t.nl_slot = nlist ? b.nl_slot : nullptr;
    t.slot_cand = nlist ? b.slot_cand : nullptr;
    t.nl_gap = nlist ? b.nl_gap : nullptr;
    t.nl_gmin = nlist ? b.nl_gmin : nullptr;
    t.err = nlist ? b.errflags : nullptr;
    return t;
}

BondTask bond_task(const StepArgs &a, const DemBuffers &b)
{
    BondTask t;
    t.a = a;
    t.bonds = b.bonds;
    t.broken = b.bond_broken;
    t.bdelta = b.bond_delta;
    t.bomega = b.bond_omega;
    t.x = b.x;
    t.vel = b.vel;
    t.angvel = b.angvel;
    t.bf = b.bond_f;
    t.bt = b.bond_t;
    t.applied = b.bond_applied;
    return t;
}

IntegrateTask integrate_task(const StepArgs &a, const DemBuffers &b, int change_velocity_state,
                             double velocity_factor, double tau_x, bool gather_bonds,
                             const NextHash &nh)
{
    IntegrateTask t;
    t.a = a;
    t.clock = b.clock;
    t.change_velocity_state = change_velocity_state;
    t.velocity_factor = velocity_factor;
    t.tau_x = tau_x;
    t.x = b.x;
    t.vel = b.vel;
    t.angvel = b.angvel;
    t.force = b.force;
    t.torque = b.torque;
    t.angpos = b.angpos;
    t.acc = b.acc;
    t.angacc = b.angacc;
    t.xyzsum = b.xyzsum;
    t.w_mvfd = b.w_mvfd;
    t.tau_eff = b.tau_eff;
    t.walls_force_pp = b.walls_force_pp;
    t.wall_partial = b.wall_partial;
    t.bg = BondGather{};
    if (gather_bonds && a.nb0 > 0) {
        t.bg.off = b.pp_off;
        t.bg.ent = b.bp_ent;
        t.bg.f = b.bond_f;
        t.bg.t = b.bond_t;
        t.bg.applied = b.bond_applied;
    }
    t.nh = nh;
    return t;
}

WallsTask walls_task(const StepArgs &a, const DemBuffers &b)
{
    WallsTask t;
    t.a = a;
    t.clock = b.clock;
    t.w_nx = b.w_nx;
    t.w_mvfd = b.w_mvfd;
    t.w_acc = b.w_acc;
    t.partial = b.wall_partial;
    t.nparts = (a.np + 255) / 256;
    return t;
}

uint32_t topology_blocks(uint32_t np) { return (np + TW - 1) / TW; }
uint32_t forces_blocks(uint32_t np) { return (np + FW - 1) / FW; }
uint32_t nl_build_blocks(uint32_t np)
{
    uint32_t nb = (np + 255) / 256;
    return nb > COND_BLOCKS ? COND_BLOCKS : nb;
}

__global__ void __launch_bounds__(256) k_rebuild_multi(const RebuildTask *__restrict__ t, MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    rebuild_body(t[k].a, t[k].r, bid, nblk);
}

__global__ void __launch_bounds__(256) k_nl_build_multi(const NlBuildTask *__restrict__ t, MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    const NlBuildTask &T = t[k];
    nl_build_body(T.a, T.n, T.cond, T.vals, T.cell_start, T.cell_end, T.x_s, T.nl, T.nl_cnt,
                  T.nl_pos, T.x_ref, T.flags, T.contacts, T.cmask, T.nl_slot, T.slot_cand,
                  T.nl_gap, T.nl_gmin, T.err, bid, nblk);
}

template <bool NEAR>
__global__ void __launch_bounds__(TW) k_topology_multi(const TopoTask *__restrict__ t, MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    const TopoTask &T = t[k];
    topology_lists_body<NEAR>(T.a, T.vals, T.cell_start, T.cell_end, T.x, T.x_s, T.contacts,
                              T.cmask, T.err, T.nl, T.nl_slot, T.slot_cand, T.nl_cnt, T.nl_pos,
                              T.cellc, T.nl_gap, T.nl_gmin, T.nl_disp, T.flags, T.cur, bid, nblk);
}

template <bool HERTZ>
__global__ void __launch_bounds__(FW) k_forces_multi(const ForcesTask *__restrict__ t, MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    const ForcesTask &T = t[k];
    forces_body<HERTZ>(T.a, T.order, T.x, T.vel, T.angvel, T.force, T.torque, T.es_dot, T.ev_dot,
                       T.es, T.ev, T.p, T.w_nx, T.w_mvfd, T.walls_force_pp, T.contacts, T.cmask,
                       T.delta_t, T.nl_slot, T.slot_cand, T.nl_gap, T.nl_gmin, T.err, bid, nblk);
}

__global__ void __launch_bounds__(256) k_bonds_multi(const BondTask *__restrict__ t, MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    const BondTask &T = t[k];
    bond_compute_body(T.a, T.bonds, T.broken, T.bdelta, T.bomega, T.x, T.vel, T.angvel, T.bf,
                      T.bt, T.applied, bid);
}

__global__ void __launch_bounds__(256) k_integrate_multi(const IntegrateTask *__restrict__ t,
                                                         MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    const IntegrateTask &T = t[k];
    integrate_body(T.a, T.clock, T.change_velocity_state, T.velocity_factor, T.tau_x, T.x, T.vel,
                   T.angvel, T.force, T.torque, T.angpos, T.acc, T.angacc, T.xyzsum, T.w_mvfd,
                   T.tau_eff, T.walls_force_pp, T.wall_partial, T.bg, T.nh, bid);
}

__global__ void __launch_bounds__(SUM_THREADS) k_walls_multi(const WallsTask *__restrict__ t)
{
    const WallsTask &T = t[blockIdx.x];
    integrate_walls_body(T.a, T.clock, T.w_nx, T.w_mvfd, T.w_acc, T.partial, T.nparts);
}

__global__ void k_clock_multi(const ClockTask *__restrict__ t, uint32_t n)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        t[k].clock->iter += 1;
        t[k].clock->time += t[k].dt;
    }
}

void launch_rebuild_multi(hipStream_t s, const RebuildTask *t, MultiMap m, uint32_t blocks)
{
    if (blocks) hipLaunchKernelGGL(k_rebuild_multi, dim3(blocks), dim3(256), 0, s, t, m);
}

void launch_nl_build_multi(hipStream_t s, const NlBuildTask *t, MultiMap m, uint32_t blocks)
{
    if (blocks) hipLaunchKernelGGL(k_nl_build_multi, dim3(blocks), dim3(256), 0, s, t, m);
}

void launch_topology_multi(hipStream_t s, const TopoTask *t, MultiMap m, uint32_t blocks,
                           bool near)
{
    if (!blocks) return;
    if (near) hipLaunchKernelGGL(k_topology_multi<true>, dim3(blocks), dim3(TW), 0, s, t, m);
    else hipLaunchKernelGGL(k_topology_multi<false>, dim3(blocks), dim3(TW), 0, s, t, m);
}

void launch_forces_multi(hipStream_t s, const ForcesTask *t, MultiMap m, uint32_t blocks,
                         bool hertz)
{
    if (!blocks) return;
    if (hertz) hipLaunchKernelGGL(k_forces_multi<true>, dim3(blocks), dim3(FW), 0, s, t, m);
    else hipLaunchKernelGGL(k_forces_multi<false>, dim3(blocks), dim3(FW), 0, s, t, m);
}

void launch_bonds_multi(hipStream_t s, const BondTask *t, MultiMap m, uint32_t blocks)
{
    if (blocks) hipLaunchKernelGGL(k_bonds_multi, dim3(blocks), dim3(256), 0, s, t, m);
}

void launch_integrate_multi(hipStream_t s, const IntegrateTask *t, MultiMap m, uint32_t blocks)
{
    if (blocks) hipLaunchKernelGGL(k_integrate_multi, dim3(blocks), dim3(256), 0, s, t, m);
}

HashTask hash_task(const StepArgs &a, const NLArgs &n, const DemBuffers &b, uint32_t cur,
                   uint32_t force)
{
    HashTask t;
    t.a = a;
    t.n = n;
    t.x = b.x;
    t.keys = b.keys;
    t.vals = b.vals;
    t.cellc = b.cellc;
    t.x_ref = b.x_ref;
    t.disp = b.nl_disp;
    t.flags = b.nl_flags;
    t.cur = cur;
    t.force = force;
    t.err = b.errflags;
    return t;
}

BondGatherTask bond_gather_task(const DemBuffers &b)
{
    BondGatherTask t;
    t.nbp = b.nbp;
    t.plist = b.bp_list;
    t.off = b.bp_off;
    t.ent = b.bp_ent;
    t.bf = b.bond_f;
    t.bt = b.bond_t;
    t.applied = b.bond_applied;
    t.force = b.force;
    t.torque = b.torque;
    return t;
}

__global__ void __launch_bounds__(256) k_hash_multi(const HashTask *__restrict__ t, MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    const HashTask &T = t[k];
    hash_nl_body(T.a, T.n, T.x, T.keys, T.vals, T.cellc, T.x_ref, T.disp, T.flags, T.cur, T.force,
                 T.err, bid);
}

__global__ void __launch_bounds__(256) k_bond_gather_multi(const BondGatherTask *__restrict__ t,
                                                           MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    const BondGatherTask &T = t[k];
    bond_gather_body(T.nbp, T.plist, T.off, T.ent, T.bf, T.bt, T.applied, T.force, T.torque, bid);
}

void launch_hash_multi(hipStream_t s, const HashTask *t, MultiMap m, uint32_t blocks)
{
    if (blocks) hipLaunchKernelGGL(k_hash_multi, dim3(blocks), dim3(256), 0, s, t, m);
}

void launch_bond_gather_multi(hipStream_t s, const BondGatherTask *t, MultiMap m, uint32_t blocks)
{
    if (blocks) hipLaunchKernelGGL(k_bond_gather_multi, dim3(blocks), dim3(256), 0, s, t, m);
}

void launch_walls_multi(hipStream_t s, const WallsTask *t, uint32_t n)
{
    if (n) hipLaunchKernelGGL(k_walls_multi, dim3(n), dim3(SUM_THREADS), 0, s, t);
}

void launch_clock_multi(hipStream_t s, const ClockTask *t, uint32_t n)
{
    if (n) hipLaunchKernelGGL(k_clock_multi, dim3((n + 63) / 64), dim3(64), 0, s, t, n);
}
