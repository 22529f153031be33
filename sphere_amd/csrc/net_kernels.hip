// net_kernels.hip -- gfx950 kernels of the Darcy correction network: 3x3x3
// convolutions (stride 1, zero padding 1) as implicit GEMMs on the matrix
// cores (v_mfma_f32_16x16x32_bf16 for the forward pass, 16x16x16 for the
// weight gradients: BF16 operands, FP32 accumulation), their
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
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// gfx950's v_mfma_f32_16x16x32_bf16: A[m][k] with m = lane % 16 and k = 8
// (lane / 16) + 0..7 (eight BF16 per lane), B[k][n] with n = lane % 16 and k
// likewise, D as for the 16x16x16 shape.  Half the instructions, and half
// the dependent steps of a K loop, of the 16x16x16 shape.
__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

__device__ __forceinline__ s16x8 cat8(bf16x4 lo, bf16x4 hi)
{
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}


// neighbour (x, y, z) of the 3x3x3 stencil: wrapped along periodic axes,
// false where it falls in the zero padding
__device__ __forceinline__ bool nb_in(const NetGeom &g, int &x, int &y, int z)
{
    const int nx = (int)g.nx, ny = (int)g.ny;
    if (g.wrap & 1u) x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
    if (g.wrap & 2u) y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
    return x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < (int)g.nz;
}

__device__ __forceinline__ void vox_xyz(const NetGeom &g, uint32_t v, int &x, int &y, int &z)
{
    x = (int)(v % g.nx);
    const uint32_t r = v / g.nx;
    y = (int)(r % g.ny);
    z = (int)(r / g.ny);
}

// ---------------------------------------------------------------------------
// Forward convolution.  A wave computes MT 16-voxel M tiles by all 16*NT
// output channels; K runs over 32-wide steps: two (tap, 16-channel chunk)
// pairs of the input (CIN4: inputs of 4 channels, eight taps per step, two
// per 16-lane group).  With KS > 1 the KS waves of a block share the M
// tiles and take every KS-th step, and their partial sums meet in LDS
// (small grids: more waves, shorter chains).  Each wave issues the loads of
// G steps before their MFMAs.  MFMA 16x16x32 lane layout: A[m][k] with
// m = lane % 16, k = 8 (lane / 16) + 0..7; B[k][n] with n = lane % 16, k
// likewise; the accumulator holds D[4 (lane / 16) + i][lane % 16].
// ---------------------------------------------------------------------------
#define CONV_G 4
// K split of the grids of at most CONV_SMALL_V voxels: CONV_KS_SMALL waves
// per 16-voxel tile (4 on larger grids below 65536 voxels).  On the network's
// 48 x 48 x 96 grid, 8 waves on its two coarsest levels took one forward pass
// from 154 to 144 us; 16 waves, or 8 on the 27,648-voxel level too, were
// slower, except for inputs of 64 channels (143.4 us).
#define CONV_SMALL_V 8192u
#define CONV_KS_SMALL 8

// Epilogue of the convolutions: acc[j][t] holds, for M tile j (16
// voxels from vbase[j] along x), voxels vbase[j] + 4 grp + i (i = 0..3) and
// output channel 16 t + m.
// The children of coarse voxel v on the grid twice as fine, channel co:
// yfine = a + skip, for children k0, k0 + dk, ... (k < 8)
__device__ __forceinline__ void up_children(const NetGeom &g, uint32_t v, int co, int cout_real, float av,
                                            int k0, int dk, const uint16_t *__restrict__ skip,
                                            uint16_t *__restrict__ yfine)
{
    int x, y, z;
    vox_xyz(g, v, x, y, z);
    const uint32_t fnx = 2u * g.nx, fny = 2u * g.ny;
    for (int k = k0; k < 8; k += dk) {
        const uint32_t fv = ((2u * z + (k >> 2)) * fny + 2u * y + ((k >> 1) & 1)) * fnx + 2u * x + (k & 1);
        const size_t fo = (size_t)fv * cout_real + co;
        yfine[fo] = f2bf(av + bf2f(skip[fo]));
    }
}

template <int NT, int MT>
__device__ __forceinline__ void conv_store(const f32x4 (&acc)[MT][NT], const uint32_t (&vbase)[MT], uint32_t V,
                                           uint32_t m, uint32_t grp, const float *__restrict__ bias,
                                           int cout_real, void *__restrict__ out, int mode,
                                           const float *__restrict__ proj, float *__restrict__ out2,
                                           const NetGeom &g, const uint16_t *__restrict__ skip,
                                           uint16_t *__restrict__ yfine)
{
    if (NT == 1 && mode == NET_OUT_PROJ) {
        // h = bf16(relu(acc + bias)) stored, and out2[v] = sum_c h[v][c]
        // proj[c] + proj[16]: the 16 channels of a voxel sit in the 16 lanes
        // of one lane group, summed by xor shuffles within the group
        const float bv = bias[m];
        const float wp = proj[m];
#pragma unroll
        for (int j = 0; j < MT; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t v = vbase[j] + 4u * grp + (uint32_t)i;
                const float r = acc[j][0][i] + bv;
                const uint16_t h = f2bf(r > 0.f ? r : 0.f);
                float sum = bf2f(h) * wp;
                sum += __shfl_xor(sum, 1);
                sum += __shfl_xor(sum, 2);
                sum += __shfl_xor(sum, 4);
                sum += __shfl_xor(sum, 8);
                if (v < V) {
                    ((uint16_t *)out)[(size_t)v * 16 + m] = h;
                    if (m == 0) out2[v] = sum + proj[16];
                }
            }
        return;
    }
#pragma unroll
    for (int j = 0; j < MT; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = t * 16 + (int)m;
            if (co >= cout_real) continue;
            const float bv = bias ? bias[co] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t v = vbase[j] + 4u * grp + (uint32_t)i;
                if (v >= V) continue;
                float r = acc[j][t][i] + bv;
                const size_t o = (size_t)v * cout_real + co;
                if (mode == NET_OUT_BF16_RELU || mode == NET_OUT_RELU_UP) {
                    const uint16_t a = f2bf(r > 0.f ? r : 0.f);
                    ((uint16_t *)out)[o] = a;
                    if (mode == NET_OUT_RELU_UP && yfine)  // the 8 children: a + skip
                        up_children(g, v, co, cout_real, bf2f(a), 0, 1, skip, yfine);
                } else if (mode == NET_OUT_F32) {
                    ((float *)out)[o] = r;
                } else {
                    ((float *)out)[o] += r;
                }
            }
        }
}


template <int NT, int MT, int KS, bool CIN4>
__device__ __forceinline__ void conv_fwd_body(const NetGeom &g, const uint16_t *__restrict__ in, int cin,
                                              const uint16_t *__restrict__ w,
                                              const float *__restrict__ bias, int cout_real,
                                              void *__restrict__ out, int mode,
                                              const float *__restrict__ proj, float *__restrict__ out2,
                                              const uint16_t *__restrict__ skip,
                                              uint16_t *__restrict__ yfine, uint32_t bid, uint32_t nblk)
{
    __shared__ float red[KS > 1 ? (KS - 1) * MT * NT * 4 * 64 : 1];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    // each XCD works through one contiguous range of tiles, whose stencil
    // reads then mostly hit that XCD's L2
    const uint32_t b = xcd_tile(bid, nblk);
    const uint32_t tile = KS == 1 ? b * 4u + wv : b;
    const int kpart = KS == 1 ? 0 : (int)wv;
    const uint32_t V = g.nx * g.ny * g.nz;
    const uint32_t v0 = tile * 16u * MT;
    if (v0 >= V) return;  // uniform over the block when KS > 1
    const uint32_t m = lane & 15u, grp = lane >> 4;
    const int cout = 16 * NT;
    const int nx = (int)g.nx, ny = (int)g.ny;
    int vx[MT], vy[MT], vz[MT];
    bool vin[MT];
#pragma unroll
    for (int j = 0; j < MT; ++j) {
        const uint32_t v = v0 + 16u * j + m;
        vin[j] = v < V;
        vox_xyz(g, vin[j] ? v : 0u, vx[j], vy[j], vz[j]);
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int j = 0; j < MT; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[j][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // K runs over 32-wide steps u.  Inputs of 16k channels: 16-channel steps
    // st = (tap, chunk) taken in pairs, lane group grp reading channels
    // 8 (grp & 1) .. + 7 of step 2u + (grp >> 1).  Inputs of 4 channels: taps
    // 8u + 2 grp and 8u + 2 grp + 1, four channels each.
    const int chunks = CIN4 ? 1 : (cin >> 4);
    const int S = CIN4 ? 27 : 27 * chunks;
    const int U = CIN4 ? 4 : (S + 1) / 2;
    const int half = (int)(grp & 1u), sub = (int)(grp >> 1);
    const int cs = CIN4 ? 4 : cin;
    const s16x8 zero8 = s16x8{0, 0, 0, 0, 0, 0, 0, 0};
    for (int u0 = kpart; u0 < U; u0 += KS * CONV_G) {
        s16x8 a[CONV_G][MT], b[CONV_G][NT];
#pragma unroll
        for (int q = 0; q < CONV_G; ++q) {
            const int u = u0 + q * KS;
            if (CIN4) {
                bf16x4 bl[NT][2], al[MT][2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int tap = 8 * u + 2 * (int)grp + e;
                    const bool tok = u < U && tap < 27;
                    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        bl[t][e] = bf16x4{0, 0, 0, 0};
                        if (tok) bl[t][e] = *(const bf16x4 *)(w + ((size_t)(tap * cout + t * 16 + (int)m) * 4));
                    }
#pragma unroll
                    for (int j = 0; j < MT; ++j) {
                        int x = vx[j] + dx, y = vy[j] + dy;
                        const int z = vz[j] + dz;
                        al[j][e] = bf16x4{0, 0, 0, 0};
                        if (tok && vin[j] && nb_in(g, x, y, z))
                            al[j][e] = *(const bf16x4 *)(in + ((size_t)((z * ny + y) * nx + x) * 4));
                    }
                }
#pragma unroll
                for (int t = 0; t < NT; ++t) b[q][t] = cat8(bl[t][0], bl[t][1]);
#pragma unroll
                for (int j = 0; j < MT; ++j) a[q][j] = cat8(al[j][0], al[j][1]);
            } else {
                const int st = 2 * u + sub;
                const bool tok = u < U && st < S;
                const int tap = st / chunks;
                const int ci = (st - tap * chunks) * 16 + 8 * half;
                const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    b[q][t] = zero8;
                    if (tok) b[q][t] = *(const s16x8 *)(w + ((size_t)(tap * cout + t * 16 + (int)m) * cs + ci));
                }
#pragma unroll
                for (int j = 0; j < MT; ++j) {
                    int x = vx[j] + dx, y = vy[j] + dy;
                    const int z = vz[j] + dz;
                    a[q][j] = zero8;
                    if (tok && vin[j] && nb_in(g, x, y, z))
                        a[q][j] = *(const s16x8 *)(in + ((size_t)((z * ny + y) * nx + x) * cs + ci));
                }
            }
        }
#pragma unroll
        for (int q = 0; q < CONV_G; ++q) {
            if (u0 + q * KS >= U) break;
#pragma unroll
            for (int j = 0; j < MT; ++j)
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[j][t] = mfma32(a[q][j], b[q][t], acc[j][t]);
        }
    }
    if (KS > 1) {
        if (wv > 0) {
#pragma unroll
            for (int j = 0; j < MT; ++j)
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        red[((((wv - 1) * MT + j) * NT + t) * 4 + i) * 64 + lane] = acc[j][t][i];
        }
        __syncthreads();
        if (mode != NET_OUT_RELU_UP && wv > 0) return;
        if (wv == 0) {
#pragma unroll 1
            for (int k = 0; k < KS - 1; ++k)
#pragma unroll
                for (int j = 0; j < MT; ++j)
#pragma unroll
                    for (int t = 0; t < NT; ++t)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            acc[j][t][i] += red[(((k * MT + j) * NT + t) * 4 + i) * 64 + lane];
        }
        if (mode == NET_OUT_RELU_UP) {
            // wave 0 stores the coarse output and puts a = bf16(relu(acc +
            // bias)) in LDS, from which all waves write the 8 children
            __shared__ float s_a[MT * NT * 4][64];
            uint32_t vb[MT];
#pragma unroll
            for (int j = 0; j < MT; ++j) vb[j] = v0 + 16u * j;
            if (wv == 0) {
                conv_store<NT, MT>(acc, vb, V, m, grp, bias, cout_real, out, NET_OUT_RELU_UP, proj, out2, g,
                                   skip, nullptr);
#pragma unroll
                for (int j = 0; j < MT; ++j)
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const int co = t * 16 + (int)m;
                        const float bv = (bias && co < cout_real) ? bias[co] : 0.f;
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const float r = acc[j][t][i] + bv;
                            s_a[(j * NT + t) * 4 + i][lane] = bf2f(f2bf(r > 0.f ? r : 0.f));
                        }
                    }
            }
            __syncthreads();
            // then each lane writes 8 channels (16 B) of one child: item =
            // ((voxel * 8 + child) * cout_real / 8 + channel octet), so
            // neighbouring lanes fill whole fine-voxel rows (cout_real % 8 == 0)
            const int octets = cout_real >> 3;
            const int items = 16 * MT * 8 * octets;
            const uint32_t fnx = 2u * g.nx, fny = 2u * g.ny;
            for (int it = (int)threadIdx.x; it < items; it += 64 * KS) {
                const int h = it % octets, r = it / octets;
                const int k = r & 7, vv = r >> 3;
                const uint32_t v = v0 + (uint32_t)vv;
                if (v >= V) continue;
                int x, y, z;
                vox_xyz(g, v, x, y, z);
                const uint32_t fv = ((2u * z + (k >> 2)) * fny + 2u * y + ((k >> 1) & 1)) * fnx + 2u * x + (k & 1);
                const size_t fo = (size_t)fv * cout_real + 8 * h;
                const uint4 sk = *(const uint4 *)(skip + fo);
                const uint32_t sw[4] = {sk.x, sk.y, sk.z, sk.w};
                uint32_t ow[4];
                const int j = vv >> 4, vl = vv & 15;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint16_t o[2];
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int co = 8 * h + 2 * q + e;
                        const float av = s_a[(j * NT + (co >> 4)) * 4 + (vl & 3)][(vl >> 2) * 16 + (co & 15)];
                        o[e] = f2bf(av + bf2f((uint16_t)(sw[q] >> (16 * e))));
                    }
                    ow[q] = (uint32_t)o[0] | ((uint32_t)o[1] << 16);
                }
                *(uint4 *)(yfine + fo) = make_uint4(ow[0], ow[1], ow[2], ow[3]);
            }
            return;
        }
    }
    uint32_t vbase[MT];
#pragma unroll
    for (int j = 0; j < MT; ++j) vbase[j] = v0 + 16u * j;
    conv_store<NT, MT>(acc, vbase, V, m, grp, bias, cout_real, out, mode, proj, out2, g, skip, yfine);
}

template <int NT, int MT, int KS, bool CIN4>
__global__ void __launch_bounds__(KS > 1 ? 64 * KS : 256) k_conv_fwd(NetGeom g, const uint16_t *__restrict__ in, int cin,
                                                  const uint16_t *__restrict__ w,
                                                  const float *__restrict__ bias, int cout_real,
                                                  void *__restrict__ out, int mode,
                                                  const float *__restrict__ proj, float *__restrict__ out2,
                                                  const uint16_t *__restrict__ skip, uint16_t *__restrict__ yfine)
{
    conv_fwd_body<NT, MT, KS, CIN4>(g, in, cin, w, bias, cout_real, out, mode, proj, out2, skip, yfine,
                                    blockIdx.x, gridDim.x);
}

// Full-resolution convolution with its operands staged in LDS: a block
// computes a 16 x 4 x 4 voxel tile (wave w: y = w, M tiles z = 0..3) by 16
// output channels; the input tile with its one-voxel halo (18 x 6 x 6
// voxels) and the layer's weights are loaded once, so each input voxel is
// read from HBM/L2 about once instead of 27 times.  Requires nx % 16 == 0,
// ny % 4 == 0, nz % 4 == 0 and 16 output channels; cin 4 or 16.
template <bool CIN4>
__device__ __forceinline__ void conv_fwd_lds_body(const NetGeom &g, const uint16_t *__restrict__ in, int cin,
                                                  const uint16_t *__restrict__ w,
                                                  const float *__restrict__ bias, int cout_real,
                                                  void *__restrict__ out, int mode,
                                                  const float *__restrict__ proj, float *__restrict__ out2,
                                                  const uint16_t *__restrict__ skip,
                                                  uint16_t *__restrict__ yfine, uint32_t bid, uint32_t nblk)
{
    constexpr int CS = CIN4 ? 4 : 16;         // channels per voxel
    constexpr int HX = 18, HY = 6, HZ = 6;    // tile with halo
    __shared__ bf16x4 s_in[HX * HY * HZ * CS / 4];
    __shared__ bf16x4 s_w[27 * 16 * CS / 4];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wv = tid >> 6;
    const uint32_t tx = g.nx / 16u, ty = g.ny / 4u;
    const uint32_t t = xcd_tile(bid, nblk);  // one contiguous slab of tiles per XCD
    const uint32_t bx = t % tx, by = (t / tx) % ty, bz = t / (tx * ty);
    const int x0 = (int)bx * 16, y0 = (int)by * 4, z0 = (int)bz * 4;
    // stage the input tile (zeros or wrap-around outside the grid) and the
    // weights: every load of a thread is issued before its LDS stores
    constexpr int NIN = HX * HY * HZ * (CS / 4), NW = 27 * 16 * CS / 4;
    constexpr int UIN = (NIN + 255) / 256, UW = (NW + 255) / 256;
    bf16x4 vin[UIN], vw[UW];
#pragma unroll
    for (int u = 0; u < UIN; ++u) {
        const int q = (int)tid + 256 * u;
        vin[u] = bf16x4{0, 0, 0, 0};
        if (q < NIN) {
            const int vox = q / (CS / 4), part = q % (CS / 4);
            const int lx = vox % HX, ly = (vox / HX) % HY, lz = vox / (HX * HY);
            int x = x0 - 1 + lx, y = y0 - 1 + ly;
            const int z = z0 - 1 + lz;
            if (nb_in(g, x, y, z))
                vin[u] = *(const bf16x4 *)(in + ((size_t)((z * (int)g.ny + y) * (int)g.nx + x) * CS + part * 4));
        }
    }
#pragma unroll
    for (int u = 0; u < UW; ++u) {
        const int q = (int)tid + 256 * u;
        vw[u] = bf16x4{0, 0, 0, 0};
        if (q < NW) vw[u] = *(const bf16x4 *)(w + (size_t)q * 4);
    }
#pragma unroll
    for (int u = 0; u < UIN; ++u) {
        const int q = (int)tid + 256 * u;
        if (q < NIN) s_in[q] = vin[u];
    }
#pragma unroll
    for (int u = 0; u < UW; ++u) {
        const int q = (int)tid + 256 * u;
        if (q < NW) s_w[q] = vw[u];
    }
    __syncthreads();
    const uint32_t m = lane & 15u, grp = lane >> 4;
    f32x4 acc[4][1];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j][0] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x4 zero4 = bf16x4{0, 0, 0, 0};
    if (CIN4) {
        // 32-wide steps of eight taps: lane group grp takes taps 8 st + 2 grp
        // and + 1 (four channels each); the padding taps 27..31 of the last
        // step read tap 26's operands and zero them
        for (int st = 0; st < 4; ++st) {
            bf16x4 al[4][2], bl[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int tap = 8 * st + 2 * (int)grp + e;
                const bool tok = tap < 27;
                const int tc = tok ? tap : 26;
                const int dz = tc / 9, dy = (tc / 3) % 3, dx = tc % 3;  // halo offsets 0..2
                bl[e] = tok ? s_w[tc * 16 + (int)m] : zero4;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const bf16x4 a = s_in[((j + dz) * HY + ((int)wv + dy)) * HX + ((int)m + dx)];
                    al[j][e] = tok ? a : zero4;
                }
            }
            const s16x8 b = cat8(bl[0], bl[1]);
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j][0] = mfma32(cat8(al[j][0], al[j][1]), b, acc[j][0]);
        }
    } else {
        // 32-wide steps of two taps: lane group grp reads channels
        // 8 (grp & 1) .. + 7 of tap 2 st + (grp >> 1); the padding tap 27
        // reads zeros
        const int half = (int)(grp & 1u), sub = (int)(grp >> 1);
        for (int st = 0; st < 14; ++st) {
            const int tap = 2 * st + sub;
            const bool tok = tap < 27;
            const int tc = tok ? tap : 26;
            const int dz = tc / 9, dy = (tc / 3) % 3, dx = tc % 3;
            const bf16x4 *wp = &s_w[(tc * 16 + (int)m) * 4 + 2 * half];
            const s16x8 b = tok ? cat8(wp[0], wp[1]) : cat8(zero4, zero4);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bf16x4 *ap = &s_in[(((j + dz) * HY + ((int)wv + dy)) * HX + ((int)m + dx)) * 4 + 2 * half];
                const s16x8 a = tok ? cat8(ap[0], ap[1]) : cat8(zero4, zero4);
                acc[j][0] = mfma32(a, b, acc[j][0]);
            }
        }
    }
    (void)cin;
    const uint32_t V = g.nx * g.ny * g.nz;
    uint32_t vbase[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) vbase[j] = ((uint32_t)(z0 + j) * g.ny + (uint32_t)(y0 + (int)wv)) * g.nx + (uint32_t)x0;
    conv_store<1, 4>(acc, vbase, V, m, grp, bias, cout_real, out, mode, proj, out2, g, skip, yfine);
}

template <bool CIN4>
__global__ void __launch_bounds__(256) k_conv_fwd_lds(NetGeom g, const uint16_t *__restrict__ in, int cin,
                                                      const uint16_t *__restrict__ w,
                                                      const float *__restrict__ bias, int cout_real,
                                                      void *__restrict__ out, int mode,
                                                      const float *__restrict__ proj, float *__restrict__ out2,
                                                      const uint16_t *__restrict__ skip, uint16_t *__restrict__ yfine)
{
    conv_fwd_lds_body<CIN4>(g, in, cin, w, bias, cout_real, out, mode, proj, out2, skip, yfine, blockIdx.x,
                            gridDim.x);
}

template <int NT, bool CIN4>
static void conv_fwd_nt(hipStream_t s, const NetGeom &g, const uint16_t *in, int cin, const uint16_t *w,
                        const float *bias, int cout_real, void *out, int mode, const float *proj,
                        float *out2, const uint16_t *skip, uint16_t *yfine)
{
    const uint32_t V = g.nx * g.ny * g.nz;
    if (NT == 1 && V >= 65536u && g.nx % 16u == 0 && g.ny % 4u == 0 && g.nz % 4u == 0 &&
        (cin == 4 || cin == 16)) {
        const uint32_t blocks = (g.nx / 16u) * (g.ny / 4u) * (g.nz / 4u);
        hipLaunchKernelGGL((k_conv_fwd_lds<CIN4>), dim3(blocks), dim3(256), 0, s, g, in, cin, w, bias,
                           cout_real, out, mode, proj, out2, skip, yfine);
    } else if (V >= 65536u) {
        const uint32_t tiles = (V + 63u) / 64u;
        hipLaunchKernelGGL((k_conv_fwd<NT, 4, 1, CIN4>), dim3((tiles + 3u) / 4u), dim3(256), 0, s, g, in, cin,
                           w, bias, cout_real, out, mode, proj, out2, skip, yfine);
    } else if (V > CONV_SMALL_V) {
        // below 65536 voxels the 4 waves of a block split K for one 16-voxel
        // tile (and share an upsampling epilogue): four times the waves of
        // one tile per wave, which these latency-bound grids need
        const uint32_t tiles = (V + 15u) / 16u;
        hipLaunchKernelGGL((k_conv_fwd<NT, 1, 4, CIN4>), dim3(tiles), dim3(256), 0, s, g, in, cin, w, bias,
                           cout_real, out, mode, proj, out2, skip, yfine);
    } else if (cin >= 64) {
        // the longest K (27 x 64) of the small grids: 16 waves per tile
        const uint32_t tiles = (V + 15u) / 16u;
        hipLaunchKernelGGL((k_conv_fwd<NT, 1, 16, CIN4>), dim3(tiles), dim3(1024), 0, s, g, in, cin, w, bias,
                           cout_real, out, mode, proj, out2, skip, yfine);
    } else {
        const uint32_t tiles = (V + 15u) / 16u;
        hipLaunchKernelGGL((k_conv_fwd<NT, 1, CONV_KS_SMALL, CIN4>), dim3(tiles), dim3(64 * CONV_KS_SMALL), 0, s,
                           g, in, cin, w, bias, cout_real, out, mode, proj, out2, skip, yfine);
    }
}

// The kernel variants of conv_fwd_nt, in its order of choice.
#define CV_LDS 0
#define CV_BIG 1
#define CV_KS4 2
#define CV_KS16 3
#define CV_KS8 4

int conv_variant(const NetGeom &g, int cin, int cout_pad)
{
    const uint32_t V = g.nx * g.ny * g.nz;
    if (cout_pad == 16 && V >= 65536u && g.nx % 16u == 0 && g.ny % 4u == 0 && g.nz % 4u == 0 &&
        (cin == 4 || cin == 16))
        return CV_LDS;
    if (V >= 65536u) return CV_BIG;
    if (V > CONV_SMALL_V) return CV_KS4;
    if (cin >= 64) return CV_KS16;
    return CV_KS8;
}

uint32_t conv_variant_blocks(int variant, const NetGeom &g)
{
    const uint32_t V = g.nx * g.ny * g.nz;
    if (variant == CV_LDS) return (g.nx / 16u) * (g.ny / 4u) * (g.nz / 4u);
    if (variant == CV_BIG) return ((V + 63u) / 64u + 3u) / 4u;
    return (V + 15u) / 16u;
}

template <int NT, int MT, int KS, bool CIN4>
__global__ void __launch_bounds__(KS > 1 ? 64 * KS : 256) k_conv_fwd_multi(const ConvTask *__restrict__ t,
                                                                         MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    const ConvTask &T = t[k];
    conv_fwd_body<NT, MT, KS, CIN4>(T.g, T.in, T.cin, T.w, T.bias, T.cout_real, T.out, T.mode, T.proj, T.out2,
                                    T.skip, T.yfine, bid, nblk);
}

template <bool CIN4>
__global__ void __launch_bounds__(256) k_conv_fwd_lds_multi(const ConvTask *__restrict__ t, MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    const ConvTask &T = t[k];
    conv_fwd_lds_body<CIN4>(T.g, T.in, T.cin, T.w, T.bias, T.cout_real, T.out, T.mode, T.proj, T.out2,
                            T.skip, T.yfine, bid, nblk);
}

template <int NT, bool CIN4>
static void conv_multi_nt(hipStream_t s, int variant, const ConvTask *t, MultiMap m, uint32_t blocks)
{
    switch (variant) {
    case CV_LDS:
        if (NT == 1)
            hipLaunchKernelGGL((k_conv_fwd_lds_multi<CIN4>), dim3(blocks), dim3(256), 0, s, t, m);
        break;
    case CV_BIG:
        hipLaunchKernelGGL((k_conv_fwd_multi<NT, 4, 1, CIN4>), dim3(blocks), dim3(256), 0, s, t, m);
        break;
    case CV_KS4:
        hipLaunchKernelGGL((k_conv_fwd_multi<NT, 1, 4, CIN4>), dim3(blocks), dim3(256), 0, s, t, m);
        break;
    case CV_KS16:
        hipLaunchKernelGGL((k_conv_fwd_multi<NT, 1, 16, CIN4>), dim3(blocks), dim3(1024), 0, s, t, m);
        break;
    default:
        hipLaunchKernelGGL((k_conv_fwd_multi<NT, 1, CONV_KS_SMALL, CIN4>), dim3(blocks),
                           dim3(64 * CONV_KS_SMALL), 0, s, t, m);
    }
}

int launch_conv_fwd_multi(hipStream_t s, int variant, int cin, int cout_pad, const ConvTask *t, MultiMap m,
                          uint32_t blocks)
{
    if (!blocks) return 0;
    const bool c4 = cin == 4;
    if (!c4 && (cin % 16) != 0) return -1;
    if (variant == CV_LDS && cout_pad != 16) return -1;
#define CONV_NT_MULTI(NTV)                                                      \
    (c4 ? conv_multi_nt<NTV, true>(s, variant, t, m, blocks)                    \
        : conv_multi_nt<NTV, false>(s, variant, t, m, blocks))
    switch (cout_pad) {
    case 16: CONV_NT_MULTI(1); break;
    case 32: CONV_NT_MULTI(2); break;
    case 48: CONV_NT_MULTI(3); break;
    case 64: CONV_NT_MULTI(4); break;
    default: return -1;
    }
#undef CONV_NT_MULTI
    return 0;
}

int launch_conv_fwd(hipStream_t s, const NetGeom &g, const uint16_t *in, int cin, const uint16_t *w,
                    int cout_pad, const float *bias, int cout_real, void *out, int mode, const float *proj,
                    float *out2, const uint16_t *skip, uint16_t *yfine)
{
    const bool c4 = cin == 4;
    if (!c4 && (cin % 16) != 0) return -1;
    if (mode == NET_OUT_PROJ && (cout_pad != 16 || cout_real != 16 || !bias || !proj || !out2)) return -1;
    if (mode == NET_OUT_RELU_UP && (!skip || !yfine || (cout_real & 7) != 0)) return -1;
#define CONV_NT(NTV)                                                                                         \
    (c4 ? conv_fwd_nt<NTV, true>(s, g, in, cin, w, bias, cout_real, out, mode, proj, out2, skip, yfine)       \
        : conv_fwd_nt<NTV, false>(s, g, in, cin, w, bias, cout_real, out, mode, proj, out2, skip, yfine))
    switch (cout_pad) {
    case 16: CONV_NT(1); break;
    case 32: CONV_NT(2); break;
    case 48: CONV_NT(3); break;
    case 64: CONV_NT(4); break;
    default: return -1;
    }
#undef CONV_NT
    return 0;
}

// Backward of the pointwise head out[v] = sum_c h[v][c] w[c] + w[16]:
// dh[v][c] = dout[v] w[c], g[c] += sum_v dout[v] h[v][c], g[16] += sum_v
// dout[v].  One block per 1024 voxels, 16 voxels x 16 channels per pass.
__global__ void __launch_bounds__(256) k_proj_bwd(uint32_t V, const float *__restrict__ dout,
                                                  const uint16_t *__restrict__ h, const float *__restrict__ w,
                                                  float *__restrict__ dh, float *__restrict__ g)
{
    __shared__ float part[256], pb[16];
    const uint32_t c = threadIdx.x & 15u, vl = threadIdx.x >> 4;
    const uint32_t v0 = blockIdx.x * 1024u;
    const uint32_t ve = v0 + 1024u < V ? v0 + 1024u : V;
    const float wc = w[c];
    float sw = 0.f, sb = 0.f;
    for (uint32_t v = v0 + vl; v < ve; v += 16u) {
        const float d = dout[v];
        dh[(size_t)v * 16 + c] = d * wc;
        sw += d * bf2f(h[(size_t)v * 16 + c]);
        if (c == 0) sb += d;
    }
    part[threadIdx.x] = sw;
    if (c == 0) pb[vl] = sb;
    __syncthreads();
    if (threadIdx.x < 16) {
        float t = 0.f;
        for (int k = 0; k < 16; ++k) t += part[k * 16 + threadIdx.x];
        atomicAdd(g + threadIdx.x, t);
    } else if (threadIdx.x == 16) {
        float t = 0.f;
        for (int k = 0; k < 16; ++k) t += pb[k];
        atomicAdd(g + 16, t);
    }
}

void launch_proj_bwd(hipStream_t s, uint32_t V, const float *dout, const uint16_t *h, const float *w, float *dh,
                     float *g)
{
    hipLaunchKernelGGL(k_proj_bwd, dim3((V + 1023u) / 1024u), dim3(256), 0, s, V, dout, h, w, dh, g);
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
                    int xx = vx + dxx, yy = vy + dyy;
                    const int zz = vz + dzz;
                    if (nb_in(g, xx, yy, zz))
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

// db[c] += sum_v dz[v][c] for c < c_real (dz has c_stride <= 256
// channels); one block per 1024 voxels, threads over (voxel lane, channel)
// so that a block reads dz rows contiguously, partial sums folded in LDS
__global__ void __launch_bounds__(256) k_bias_grad(uint32_t V, const uint16_t *__restrict__ dz, int c_stride,
                                                   int c_real, float *__restrict__ db)
{
    __shared__ float part[256];
    const uint32_t C = (uint32_t)c_stride;
    const uint32_t lanes = 256u / C;
    const uint32_t ch = threadIdx.x % C, vl = threadIdx.x / C;
    const uint32_t v0 = blockIdx.x * 1024u;
    const uint32_t ve = v0 + 1024u < V ? v0 + 1024u : V;
    float sum = 0.f;
    if (vl < lanes)
        for (uint32_t v = v0 + vl; v < ve; v += lanes) sum += bf2f(dz[(size_t)v * C + ch]);
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x < C && (int)threadIdx.x < c_real) {
        float t = 0.f;
        for (uint32_t k = 0; k < lanes; ++k) t += part[k * C + threadIdx.x];
        atomicAdd(db + threadIdx.x, t);
    }
}

void launch_bias_grad(hipStream_t s, uint32_t V, const uint16_t *dz, int c_stride, int c_real, float *db)
{
    hipLaunchKernelGGL(k_bias_grad, dim3((V + 1023u) / 1024u), dim3(256), 0, s, V, dz, c_stride, c_real, db);
}

// ---------------------------------------------------------------------------
// Elementwise kernels; fine grid g, coarse grid g/2 (even sizes).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pool_fwd_body(const NetGeom &gc, const uint16_t *__restrict__ in, int c,
                                              uint16_t *__restrict__ out, uint32_t bid, uint32_t nblk)
{
    const uint32_t idx = xcd_tile(bid, nblk) * 256u + threadIdx.x;
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

__global__ void __launch_bounds__(256) k_pool_fwd(NetGeom gc, const uint16_t *__restrict__ in, int c,
                                                  uint16_t *__restrict__ out)
{
    pool_fwd_body(gc, in, c, out, blockIdx.x, gridDim.x);
}

uint32_t pool_blocks(const NetGeom &gc, int c) { return (gc.nx * gc.ny * gc.nz * (uint32_t)c + 255u) / 256u; }

__global__ void __launch_bounds__(256) k_pool_fwd_multi(const PoolTask *__restrict__ t, MultiMap m)
{
    uint32_t k, bid, nblk;
    if (!multi_block(m, k, bid, nblk)) return;
    const PoolTask &T = t[k];
    pool_fwd_body(T.gc, T.in, T.c, T.out, bid, nblk);
}

void launch_pool_fwd_multi(hipStream_t s, const PoolTask *t, MultiMap m, uint32_t blocks)
{
    if (blocks) hipLaunchKernelGGL(k_pool_fwd_multi, dim3(blocks), dim3(256), 0, s, t, m);
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
