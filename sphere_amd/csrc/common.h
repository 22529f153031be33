// common.h -- device-side types of the gfx950 DEM + Darcy library.
//
// Run parameters travel by value in the kernel-argument segment (the
// reference keeps them in __constant__ memory, src/constants.cuh:10-19).
// All particle rows are double4 (32 B): one row is two 16-byte loads and a
// wave reads 2 KiB of consecutive rows as fully used cache lines.
//
// Every floating-point expression that has a counterpart in the reference
// keeps the reference's operation order; the library is compiled with
// -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define EMPTY_CELL 0xffffffffu
#define MAXW 6

// Error bits raised by kernels (read by the host after a step).
#define ERRF_GRID 1u       // particle outside the grid
#define ERRF_OVERFLOW 2u   // more than NCS contacts
#define ERRF_NAN 4u        // NaN Darcy residual
#define ERRF_BARRIER 8u    // a grid barrier wait ran past its bound

// A contact slot holds one 32-bit word: the partner's index in bits 0-27
// and the periodic image code in bits 28-31 (bit 0: +L0, 1: -L0, 2: +L1,
// 3: -L1).  An empty slot holds np with code 0.  Particle counts stay below
// 2^28.
#define CT_IDX_MASK 0x0fffffffu
#define CT_CODE_SHIFT 28
#define CT_MAX_NP (1u << 28)
__host__ __device__ inline uint32_t ct_word(uint32_t b, uint32_t code)
{
    return b | (code << CT_CODE_SHIFT);
}

// Contact rows (slot words, tangential springs, slot-to-list maps): slot s
// of grain a lives at ct_slot(a, s, nc).  With CT_G = 1 (the default) that
// is the row-major a * nc + s.  CT_G > 1 interleaves the rows of each CT_G
// consecutive grains slot by slot, so that one slot of those grains is one
// run of CT_G entries; occupied slots gather at the top of the rows (new
// contacts take the highest empty slot), so a wave's reads of them fill
// whole cache lines.  On the 1M-grain shear box CT_G = 64 (the grains of
// one wave) made the contact kernel slower, 188.7 against 186.7 us.
#ifndef CT_G
#define CT_G 1u
#endif
__host__ __device__ inline size_t ct_base(uint32_t a, uint32_t nc)
{
    return (size_t)(a / CT_G) * nc * CT_G + (a % CT_G);
}
__host__ __device__ inline size_t ct_slot(uint32_t a, uint32_t s, uint32_t nc)
{
    return ct_base(a, nc) + (size_t)s * CT_G;
}
// entries of the contact arrays for np grains (whole groups of CT_G)
__host__ __device__ inline size_t ct_entries(uint32_t np, uint32_t nc)
{
    return (size_t)((np + CT_G - 1) / CT_G) * CT_G * nc;
}
#define RS(s) ((size_t)(s) * CT_G)  // offset of slot s from a row's base

struct Grid {
    double origo[3];
    double L[3];
    uint32_t num[3];
    int periodic;
};

struct Params {
    double g[3];
    double k_n, k_t, k_r, E, gamma_n, gamma_t, gamma_r;
    double mu_s, mu_d, mu_r, gamma_wn, gamma_wt, mu_ws, mu_wd, rho;
    uint32_t contactmodel;
    double kappa, db, V_b;
    double lambda_bar, sigma_b, tau_b;
    double sigma0_A, sigma0_f;
    // Hertz-Mindlin moduli E* = E/(2(1-nu^2)), G* = (E/(2(1+nu)))/(2(2-nu)),
    // nu = k_r, formed once on the host
    double E_star, G_star;
};

struct Walls {
    uint32_t nw;
    int wmode[MAXW];
};

// Everything a DEM kernel needs that is constant within one step.
struct StepArgs {
    Grid grid;
    Params prm;
    Walls walls;
    uint32_t np;
    uint32_t nb0;
    uint32_t nc;  // contact slots per particle (32 or 64)
    double dt;
};

// Neighbour-list parameters (no counterpart in the reference, which searches
// the 27 cells around every particle each step).
#define NL_OVF 0xffffffffu  // list count of a particle whose list overflowed
// nl_flags words: [0], [1] per step (alternating), [2] sticky list overflow,
// [3] rebuild counter, [4] steps whose search tested every list entry,
// [5] padding, then NLF_DSLOTS 64-bit words whose maximum is the largest
// squared displacement since the last rebuild (FP64 bit patterns; waves
// spread their atomic maxima over the words)
#define NLF_STICKY 2
#define NLF_COUNT 3
#define NLF_FULL 4
#define NLF_DMAX 6
#define NLF_DSLOTS 64
#define NLF_WORDS (NLF_DMAX + 2 * NLF_DSLOTS)
#define NLF_REBUILD 1u  // per-step bit: rebuild the lists
#define NLF_WRAP 2u     // per-step bit: a grain crossed a periodic boundary
struct NLArgs {
    double skin;     // listed pairs: surface gap below skin at the rebuild
    double disp2;    // rebuild once a particle moved more than sqrt(disp2)
    uint32_t cap;    // list entries per particle
    uint32_t range;  // cells searched on each side at a rebuild (1 or 2)
};

struct DarcyArgs {
    uint32_t nx, ny, nz;
    double dx, dy, dz;
    double L0, L1, L2;
    int bc_xn, bc_xp, bc_yn, bc_yp, bc_bot, bc_top;
    double bc_bot_flux, bc_top_flux;
    double mu, beta_f, rho_f, k_c, c_phi;
    uint32_t ndem;
    double dt;
    unsigned int wall0_iz;
    int periodic;
};

// ------------------------------------------------------------------
// FP64 vector helpers (component-wise, like src/vector_arithmetic.h)
// ------------------------------------------------------------------
struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 mk3(double x, double y, double z) { d3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 operator*(double s, d3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ d3 operator/(d3 a, double s) { return mk3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ d3 operator-(d3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double len3(d3 a) { return sqrt(dot3(a, a)); }
__device__ __forceinline__ d3 cross3(d3 a, d3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ d3 xyz(double4 v) { return mk3(v.x, v.y, v.z); }
__device__ __forceinline__ d3 xyz(double3 v) { return mk3(v.x, v.y, v.z); }

// Reference constants
#define PI_REF 3.14159265358979          // src/constants.h:15
#define PI_M 3.14159265358979323846      // M_PI, src/darcy.cuh:317
