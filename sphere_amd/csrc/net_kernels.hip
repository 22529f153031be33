// net_kernels.hip -- gfx950 kernels of the Darcy correction network: 3x3x3
// convolutions (stride 1, zero padding 1) as implicit GEMMs on the matrix
// cores (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation), their
// data and weight gradients, 2x average pooling, nearest upsampling with a
// skip add, and the ReLU mask.
//
// Activations are channels-last, act[v * C + c] with v = (z * ny + y) * nx
// + x, BF16 bit patterns (uint16_t) for the matrix-core operands and FP32
// for gradients that are summed from several paths.  Weights are
// w[tap][co][ci], tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1), so that four
// consecutive input channels of one tap and output channel are one 8-byte
// load.
#include "net_kernels.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b)
{
    return __uint_as_float((uint32_t)b << 16);
}

// round to nearest even; NaN stays NaN
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ void vox_xyz(const NetGeom &g, uint32_t v, int &x, int &y, int &z)
{
    x = (int)(v % g.nx);
    const uint32_t r = v / g.nx;
    y = (int)(r % g.ny);
    z = (int)(r / g.ny);
}

// ---------------------------------------------------------------------------
// Forward convolution.  One wave computes 64 voxels (four 16-voxel M tiles)
// by all 16*NT output channels; K runs over taps and 16-channel chunks of
// the input (CIN4: inputs of 4 channels, four taps per K step, one per
// 16-lane group).  MFMA 16x16x16 lane layout: A[m][k] with m = lane % 16,
// k = 4 (lane / 16) + 0..3; B[k][n] with n = lane % 16, k likewise; the
// accumulator holds D[4 (lane / 16) + i][lane % 16].
// ---------------------------------------------------------------------------
template <int NT, bool CIN4>
__global__ void __launch_bounds__(256) k_conv_fwd(NetGeom g, const uint16_t *__restrict__ in, int cin,
                                                  const uint16_t *__restrict__ w,
                                                  const float *__restrict__ bias, int cout_real,
                                                  void *__restrict__ out, int mode)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t V = g.nx * g.ny * g.nz;
    const uint32_t v0 = wave * 64u;
    if (v0 >= V) return;
    const uint32_t m = lane & 15u, grp = lane >> 4;
    const int cout = 16 * NT;
    int vx[4], vy[4], vz[4];
    bool vin[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t v = v0 + 16u * j + m;
        vin[j] = v < V;
        vox_xyz(g, vin[j] ? v : 0u, vx[j], vy[j], vz[j]);
    }
    f32x4 acc[4][NT];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[j][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    if (CIN4) {
        // K step s covers taps 4s .. 4s+3 (tap 27 is padding), 4 channels each
        for (int s = 0; s < 7; ++s) {
            const int tap = 4 * s + (int)grp;
            const bool tok = tap < 27;
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            bf16x4 b[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                b[t] = bf16x4{0, 0, 0, 0};
                if (tok) b[t] = *(const bf16x4 *)(w + ((size_t)(tap * cout + t * 16 + (int)m) * 4));
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int x = vx[j] + dx, y = vy[j] + dy, z = vz[j] + dz;
                bf16x4 a = bf16x4{0, 0, 0, 0};
                if (tok && vin[j] && x >= 0 && x < (int)g.nx && y >= 0 && y < (int)g.ny && z >= 0 &&
                    z < (int)g.nz)
                    a = *(const bf16x4 *)(in + ((size_t)((z * (int)g.ny + y) * (int)g.nx + x) * 4));
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    acc[j][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[t], acc[j][t], 0, 0, 0);
            }
        }
    } else {
        const int chunks = cin >> 4;
        for (int tap = 0; tap < 27; ++tap) {
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            size_t nb[4];
            bool ok[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int x = vx[j] + dx, y = vy[j] + dy, z = vz[j] + dz;
                ok[j] = vin[j] && x >= 0 && x < (int)g.nx && y >= 0 && y < (int)g.ny && z >= 0 &&
                        z < (int)g.nz;
                nb[j] = ok[j] ? (size_t)((z * (int)g.ny + y) * (int)g.nx + x) * (size_t)cin : 0;
            }
            for (int c = 0; c < chunks; ++c) {
                const int ci = c * 16 + 4 * (int)grp;
                bf16x4 b[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    b[t] = *(const bf16x4 *)(w + ((size_t)(tap * cout + t * 16 + (int)m) * cin + ci));
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    bf16x4 a = bf16x4{0, 0, 0, 0};
                    if (ok[j]) a = *(const bf16x4 *)(in + nb[j] + ci);
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        acc[j][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[t], acc[j][t], 0, 0, 0);
                }
            }
        }
    }
    // epilogue: lane holds voxels v0 + 16 j + 4 grp + i, channel 16 t + m
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = t * 16 + (int)m;
            if (co >= cout_real) continue;
            const float bv = bias ? bias[co] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t v = v0 + 16u * j + 4u * grp + (uint32_t)i;
                if (v >= V) continue;
                float r = acc[j][t][i] + bv;
                const size_t o = (size_t)v * cout_real + co;
                if (mode == NET_OUT_BF16_RELU) {
                    ((uint16_t *)out)[o] = f2bf(r > 0.f ? r : 0.f);
                } else if (mode == NET_OUT_F32) {
                    ((float *)out)[o] = r;
                } else {
                    ((float *)out)[o] += r;
                }
            }
        }
}

template <int NT, bool CIN4>
static void conv_fwd_nt(hipStream_t s, const NetGeom &g, const uint16_t *in, int cin, const uint16_t *w,
                        const float *bias, int cout_real, void *out, int mode)
{
    const uint32_t V = g.nx * g.ny * g.nz;
    const uint32_t waves = (V + 63u) / 64u;
    hipLaunchKernelGGL((k_conv_fwd<NT, CIN4>), dim3((waves + 3u) / 4u), dim3(256), 0, s, g, in, cin, w,
                       bias, cout_real, out, mode);
}

int launch_conv_fwd(hipStream_t s, const NetGeom &g, const uint16_t *in, int cin, const uint16_t *w,
                    int cout_pad, const float *bias, int cout_real, void *out, int mode)
{
    const bool c4 = cin == 4;
    if (!c4 && (cin % 16) != 0) return -1;
    switch (cout_pad) {
    case 16: c4 ? conv_fwd_nt<1, true>(s, g, in, cin, w, bias, cout_real, out, mode)
                : conv_fwd_nt<1, false>(s, g, in, cin, w, bias, cout_real, out, mode); break;
    case 32: c4 ? conv_fwd_nt<2, true>(s, g, in, cin, w, bias, cout_real, out, mode)
                : conv_fwd_nt<2, false>(s, g, in, cin, w, bias, cout_real, out, mode); break;
    case 48: c4 ? conv_fwd_nt<3, true>(s, g, in, cin, w, bias, cout_real, out, mode)
                : conv_fwd_nt<3, false>(s, g, in, cin, w, bias, cout_real, out, mode); break;
    case 64: c4 ? conv_fwd_nt<4, true>(s, g, in, cin, w, bias, cout_real, out, mode)
                : conv_fwd_nt<4, false>(s, g, in, cin, w, bias, cout_real, out, mode); break;
    default: return -1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Weight gradient dW[tap][co][ci] += sum_v dz[v][co] x[v + tap][ci] as
// MFMA tiles with M = 16 output channels, N = 16 input channels (CIN4: four
// taps of 4 channels) and K = 16 voxels per step.  One wave sums a chunk of
// WG_VOX voxels for one tap (CIN4: one group of four taps) and one (co, ci)
// tile, then adds its 16x16 partial to dW with FP32 atomics.
// ---------------------------------------------------------------------------
#define WG_VOX 1024

template <bool CIN4>
__global__ void __launch_bounds__(256) k_conv_wgrad(NetGeom g, const uint16_t *__restrict__ x, int cin,
                                                    const uint16_t *__restrict__ dz, int cout,
                                                    float *__restrict__ dw, uint32_t ntaps,
                                                    uint32_t nco, uint32_t nci, uint32_t nchunk)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t job = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t total = ntaps * nco * nci * nchunk;
    if (job >= total) return;
    const uint32_t chunk = job % nchunk;
    job /= nchunk;
    const uint32_t cit = job % nci;
    job /= nci;
    const uint32_t cot = job % nco;
    const uint32_t tg = job / nco;  // tap (CIN4: group of four taps)
    const uint32_t m = lane & 15u, grp = lane >> 4;
    const uint32_t V = g.nx * g.ny * g.nz;
    // the lane's B column: input channel (CIN4: tap of the group, channel)
    int tap, ci;
    if (CIN4) {
        tap = (int)(tg * 4u + (m >> 2));
        ci = (int)(m & 3u);
    } else {
        tap = (int)tg;
        ci = (int)(cit * 16u + m);
    }
    const bool tok = tap < 27;
    const int dzz = tap / 9 - 1, dyy = (tap / 3) % 3 - 1, dxx = tap % 3 - 1;
    const int co = (int)(cot * 16u + m);
    const bool cok = co < cout;
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint32_t vb = chunk * WG_VOX;
    const uint32_t ve = vb + WG_VOX < V ? vb + WG_VOX : V;
    for (uint32_t k0 = vb; k0 < ve; k0 += 16u) {
        bf16x4 a, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t v = k0 + 4u * grp + (uint32_t)i;
            short av = 0, bv = 0;
            if (v < ve) {
                if (cok) av = (short)dz[(size_t)v * cout + co];
                if (tok) {
                    int vx, vy, vz;
                    vox_xyz(g, v, vx, vy, vz);
                    const int xx = vx + dxx, yy = vy + dyy, zz = vz + dzz;
                    if (xx >= 0 && xx < (int)g.nx && yy >= 0 && yy < (int)g.ny && zz >= 0 && zz < (int)g.nz)
                        bv = (short)x[(size_t)((zz * (int)g.ny + yy) * (int)g.nx + xx) * cin + ci];
                }
            }
            a[i] = av;
            b[i] = bv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    // D[row = co tile row 4 grp + i][col = lane % 16]
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = (int)(cot * 16u + 4u * grp + (uint32_t)i);
        if (r >= cout || !tok) continue;
        const size_t o = ((size_t)tap * cout + r) * cin + ci;
        atomicAdd(dw + o, acc[i]);
    }
}

int launch_conv_wgrad(hipStream_t s, const NetGeom &g, const uint16_t *x, int cin, const uint16_t *dz,
                      int cout, float *dw)
{
    const uint32_t V = g.nx * g.ny * g.nz;
    const uint32_t nchunk = (V + WG_VOX - 1) / WG_VOX;
    const uint32_t nco = (uint32_t)(cout + 15) / 16u;
    if (cin == 4) {
        const uint32_t jobs = 7u * nco * nchunk;
        hipLaunchKernelGGL(k_conv_wgrad<true>, dim3((jobs + 3u) / 4u), dim3(256), 0, s, g, x, cin, dz, cout,
                           dw, 7u, nco, 1u, nchunk);
    } else {
        if (cin % 16) return -1;
        const uint32_t nci = (uint32_t)cin / 16u;
        const uint32_t jobs = 27u * nco * nci * nchunk;
        hipLaunchKernelGGL(k_conv_wgrad<false>, dim3((jobs + 3u) / 4u), dim3(256), 0, s, g, x, cin, dz,
                           cout, dw, 27u, nco, nci, nchunk);
    }
    return 0;
}

// db[c] += sum_v dz[v][c] for c < c_real (dz has c_stride channels); one
// block per 1024 voxels
__global__ void __launch_bounds__(256) k_bias_grad(uint32_t V, const uint16_t *__restrict__ dz, int c_stride,
                                                   int c_real, float *__restrict__ db)
{
    const uint32_t v0 = blockIdx.x * 1024u;
    const uint32_t ve = v0 + 1024u < V ? v0 + 1024u : V;
    for (int ch = (int)threadIdx.x; ch < c_real; ch += 256) {
        float sum = 0.f;
        for (uint32_t v = v0; v < ve; ++v) sum += bf2f(dz[(size_t)v * c_stride + ch]);
        atomicAdd(db + ch, sum);
    }
}

void launch_bias_grad(hipStream_t s, uint32_t V, const uint16_t *dz, int c_stride, int c_real, float *db)
{
    hipLaunchKernelGGL(k_bias_grad, dim3((V + 1023u) / 1024u), dim3(256), 0, s, V, dz, c_stride, c_real, db);
}

// ---------------------------------------------------------------------------
// Elementwise kernels; fine grid g, coarse grid g/2 (even sizes).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_pool_fwd(NetGeom gc, const uint16_t *__restrict__ in, int c,
                                                  uint16_t *__restrict__ out)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    const uint32_t Vc = gc.nx * gc.ny * gc.nz;
    if (idx >= Vc * (uint32_t)c) return;
    const uint32_t vc = idx / (uint32_t)c, ch = idx % (uint32_t)c;
    int x, y, z;
    vox_xyz(gc, vc, x, y, z);
    const uint32_t fnx = 2u * gc.nx, fny = 2u * gc.ny;
    float sum = 0.f;
    for (int k = 0; k < 8; ++k) {
        const uint32_t fx = 2u * x + (k & 1), fy = 2u * y + ((k >> 1) & 1), fz = 2u * z + (k >> 2);
        sum += bf2f(in[((size_t)(fz * fny + fy) * fnx + fx) * c + ch]);
    }
    out[idx] = f2bf(0.125f * sum);
}

void launch_pool_fwd(hipStream_t s, const NetGeom &gc, const uint16_t *in, int c, uint16_t *out)
{
    const uint32_t n = gc.nx * gc.ny * gc.nz * (uint32_t)c;
    hipLaunchKernelGGL(k_pool_fwd, dim3((n + 255u) / 256u), dim3(256), 0, s, gc, in, c, out);
}

// d_fine[v][c] += d_coarse[parent][c] / 8
__global__ void __launch_bounds__(256) k_pool_bwd(NetGeom gf, const float *__restrict__ dcoarse, int c,
                                                  float *__restrict__ dfine)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    const uint32_t Vf = gf.nx * gf.ny * gf.nz;
    if (idx >= Vf * (uint32_t)c) return;
    const uint32_t vf = idx / (uint32_t)c, ch = idx % (uint32_t)c;
    int x, y, z;
    vox_xyz(gf, vf, x, y, z);
    const uint32_t cnx = gf.nx / 2u, cny = gf.ny / 2u;
    const size_t p = ((size_t)((z >> 1) * cny + (y >> 1)) * cnx + (x >> 1)) * c + ch;
    dfine[idx] += 0.125f * dcoarse[p];
}

void launch_pool_bwd(hipStream_t s, const NetGeom &gf, const float *dcoarse, int c, float *dfine)
{
    const uint32_t n = gf.nx * gf.ny * gf.nz * (uint32_t)c;
    hipLaunchKernelGGL(k_pool_bwd, dim3((n + 255u) / 256u), dim3(256), 0, s, gf, dcoarse, c, dfine);
}

// out[v][c] = coarse[parent(v)][c] + skip[v][c]
__global__ void __launch_bounds__(256) k_up_add(NetGeom gf, const uint16_t *__restrict__ coarse,
                                                const uint16_t *__restrict__ skip, int c,
                                                uint16_t *__restrict__ out)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    const uint32_t Vf = gf.nx * gf.ny * gf.nz;
    if (idx >= Vf * (uint32_t)c) return;
    const uint32_t vf = idx / (uint32_t)c, ch = idx % (uint32_t)c;
    int x, y, z;
    vox_xyz(gf, vf, x, y, z);
    const uint32_t cnx = gf.nx / 2u, cny = gf.ny / 2u;
    const size_t p = ((size_t)((z >> 1) * cny + (y >> 1)) * cnx + (x >> 1)) * c + ch;
    out[idx] = f2bf(bf2f(coarse[p]) + bf2f(skip[idx]));
}

void launch_up_add(hipStream_t s, const NetGeom &gf, const uint16_t *coarse, const uint16_t *skip, int c,
                   uint16_t *out)
{
    const uint32_t n = gf.nx * gf.ny * gf.nz * (uint32_t)c;
    hipLaunchKernelGGL(k_up_add, dim3((n + 255u) / 256u), dim3(256), 0, s, gf, coarse, skip, c, out);
}

// d_coarse[v][c] = sum of the 8 children's d_fine
__global__ void __launch_bounds__(256) k_up_bwd(NetGeom gc, const float *__restrict__ dfine, int c,
                                                float *__restrict__ dcoarse)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    const uint32_t Vc = gc.nx * gc.ny * gc.nz;
    if (idx >= Vc * (uint32_t)c) return;
    const uint32_t vc = idx / (uint32_t)c, ch = idx % (uint32_t)c;
    int x, y, z;
    vox_xyz(gc, vc, x, y, z);
    const uint32_t fnx = 2u * gc.nx, fny = 2u * gc.ny;
    float sum = 0.f;
    for (int k = 0; k < 8; ++k) {
        const uint32_t fx = 2u * x + (k & 1), fy = 2u * y + ((k >> 1) & 1), fz = 2u * z + (k >> 2);
        sum += dfine[((size_t)(fz * fny + fy) * fnx + fx) * c + ch];
    }
    dcoarse[idx] = sum;
}

void launch_up_bwd(hipStream_t s, const NetGeom &gc, const float *dfine, int c, float *dcoarse)
{
    const uint32_t n = gc.nx * gc.ny * gc.nz * (uint32_t)c;
    hipLaunchKernelGGL(k_up_bwd, dim3((n + 255u) / 256u), dim3(256), 0, s, gc, dfine, c, dcoarse);
}

// dz = dy where the ReLU output y > 0, else 0, as BF16 with cpad channels
// (channels c .. cpad-1 are zero)
__global__ void __launch_bounds__(256) k_relu_mask(uint32_t V, const float *__restrict__ dy,
                                                   const uint16_t *__restrict__ y, int c, int cpad,
                                                   uint16_t *__restrict__ dz)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= V * (uint32_t)cpad) return;
    const uint32_t v = idx / (uint32_t)cpad, ch = idx % (uint32_t)cpad;
    float r = 0.f;
    if ((int)ch < c) {
        const size_t i = (size_t)v * c + ch;
        r = (y == nullptr || bf2f(y[i]) > 0.f) ? dy[i] : 0.f;
    }
    dz[idx] = f2bf(r);
}

void launch_relu_mask(hipStream_t s, uint32_t V, const float *dy, const uint16_t *y, int c, int cpad,
                      uint16_t *dz)
{
    const uint32_t n = V * (uint32_t)cpad;
    hipLaunchKernelGGL(k_relu_mask, dim3((n + 255u) / 256u), dim3(256), 0, s, V, dy, y, c, cpad, dz);
}

// FP32 -> BF16 with channel padding: out[v][cpad] from in[v][c]
__global__ void __launch_bounds__(256) k_to_bf16_pad(uint32_t V, const float *__restrict__ in, int c,
                                                     int cpad, uint16_t *__restrict__ out)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= V * (uint32_t)cpad) return;
    const uint32_t v = idx / (uint32_t)cpad, ch = idx % (uint32_t)cpad;
    out[idx] = (int)ch < c ? f2bf(in[(size_t)v * c + ch]) : (uint16_t)0;
}

void launch_to_bf16_pad(hipStream_t s, uint32_t V, const float *in, int c, int cpad, uint16_t *out)
{
    const uint32_t n = V * (uint32_t)cpad;
    hipLaunchKernelGGL(k_to_bf16_pad, dim3((n + 255u) / 256u), dim3(256), 0, s, V, in, c, cpad, out);
}

// Pack one layer's FP32 parameters (w[27][cout][cin] then b[cout]) into
// the BF16 operand layouts: forward wf[27][cout_pad][cin_pad] and, for the
// data gradient (the adjoint: taps mirrored, channels swapped),
// wt[27][cin_pad2][cout_pad2] with wt[tap][ci][co] = w[26 - tap][co][ci].
__global__ void __launch_bounds__(256) k_pack_weights(const float *__restrict__ w, int cin, int cout,
                                                      uint16_t *__restrict__ wf, int cin_pad, int cout_pad,
                                                      uint16_t *__restrict__ wt, int tin_pad, int tout_pad)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    const uint32_t nf = 27u * (uint32_t)cout_pad * (uint32_t)cin_pad;
    if (idx < nf) {
        const uint32_t ci = idx % (uint32_t)cin_pad;
        const uint32_t r = idx / (uint32_t)cin_pad;
        const uint32_t co = r % (uint32_t)cout_pad, tap = r / (uint32_t)cout_pad;
        wf[idx] = ((int)ci < cin && (int)co < cout) ? f2bf(w[((size_t)tap * cout + co) * cin + ci]) : (uint16_t)0;
    }
    if (wt == nullptr) return;
    const uint32_t nt = 27u * (uint32_t)tout_pad * (uint32_t)tin_pad;
    if (idx < nt) {
        // wt layout [tap][new co = ci][new ci = co]
        const uint32_t nci = idx % (uint32_t)tin_pad;      // original output channel
        const uint32_t r = idx / (uint32_t)tin_pad;
        const uint32_t nco = r % (uint32_t)tout_pad, tap = r / (uint32_t)tout_pad;  // nco: original input channel
        wt[idx] = ((int)nci < cout && (int)nco < cin)
                      ? f2bf(w[((size_t)(26u - tap) * cout + nci) * cin + nco])
                      : (uint16_t)0;
    }
}

void launch_pack_weights(hipStream_t s, const float *w, int cin, int cout, uint16_t *wf, int cin_pad,
                         int cout_pad, uint16_t *wt, int tin_pad, int tout_pad)
{
    uint32_t n = 27u * (uint32_t)cout_pad * (uint32_t)cin_pad;
    const uint32_t n2 = 27u * (uint32_t)tout_pad * (uint32_t)tin_pad;
    if (wt && n2 > n) n = n2;
    hipLaunchKernelGGL(k_pack_weights, dim3((n + 255u) / 256u), dim3(256), 0, s, w, cin, cout, wf, cin_pad,
                       cout_pad, wt, tin_pad, tout_pad);
}

// Add a padded FP32 weight gradient gw[27][cout_pad][cin_pad] into the
// flat layout g[27][cout][cin].
__global__ void __launch_bounds__(256) k_unpack_wgrad(const float *__restrict__ gw, int cin, int cin_pad,
                                                      int cout, int cout_pad, float *__restrict__ g)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    const uint32_t n = 27u * (uint32_t)cout * (uint32_t)cin;
    if (idx >= n) return;
    const uint32_t ci = idx % (uint32_t)cin;
    const uint32_t r = idx / (uint32_t)cin;
    const uint32_t co = r % (uint32_t)cout, tap = r / (uint32_t)cout;
    g[idx] += gw[((size_t)tap * cout_pad + co) * cin_pad + ci];
}

void launch_unpack_wgrad(hipStream_t s, const float *gw, int cin, int cin_pad, int cout, int cout_pad,
                         float *g)
{
    const uint32_t n = 27u * (uint32_t)cout * (uint32_t)cin;
    hipLaunchKernelGGL(k_unpack_wgrad, dim3((n + 255u) / 256u), dim3(256), 0, s, gw, cin, cin_pad, cout,
                       cout_pad, g);
}

__global__ void __launch_bounds__(256) k_fill_f32(float *__restrict__ p, size_t n, float v)
{
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_f32(hipStream_t s, float *p, size_t n, float v)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}
