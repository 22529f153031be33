// net_kernels.h -- launchers of the Darcy correction network's kernels
// (net_kernels.hip).  Device pointers; every launcher enqueues on stream s.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "multi.h"

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

struct NetGeom {
    uint32_t nx, ny, nz;
    uint32_t wrap;  // bit 0: x periodic, bit 1: y periodic (else zero padding)
};

#define NET_OUT_BF16_RELU 0  // out = bf16(relu(acc + bias))
#define NET_OUT_F32 1        // out = acc + bias
#define NET_OUT_F32_ACC 2    // out += acc + bias
#define NET_OUT_PROJ 3       // 16 channels: out = bf16(relu(acc + bias)) and
                             // out2[v] = sum_c out[v][c] proj[c] + proj[16]
#define NET_OUT_RELU_UP 4    // out = bf16(relu(acc + bias)) and, on the grid
                             // twice as fine, yfine = bf16(up(out) + skip)

// in[V][cin] (cin 4 or a multiple of 16), w[27][cout_pad][cin]; out has
// cout_real channels.  Returns -1 for an unsupported shape.
int launch_conv_fwd(hipStream_t s, const NetGeom &g, const uint16_t *in, int cin, const uint16_t *w,
                    int cout_pad, const float *bias, int cout_real, void *out, int mode,
                    const float *proj = nullptr, float *out2 = nullptr,
                    const uint16_t *skip = nullptr, uint16_t *yfine = nullptr);
void launch_proj_bwd(hipStream_t s, uint32_t V, const float *dout, const uint16_t *h, const float *w, float *dh,
                     float *g);
// dw[27][cout][cin] += sum_v dz[v][co] x[v + tap][ci]; x[V][cin] (cin 4 or
// a multiple of 16), dz[V][cout]
int launch_conv_wgrad(hipStream_t s, const NetGeom &g, const uint16_t *x, int cin, const uint16_t *dz,
                      int cout, float *dw);
void launch_bias_grad(hipStream_t s, uint32_t V, const uint16_t *dz, int c_stride, int c_real, float *db);
void launch_pool_fwd(hipStream_t s, const NetGeom &gc, const uint16_t *in, int c, uint16_t *out);
void launch_pool_bwd(hipStream_t s, const NetGeom &gf, const float *dcoarse, int c, float *dfine);
void launch_up_bwd(hipStream_t s, const NetGeom &gc, const float *dfine, int c, float *dcoarse);
void launch_relu_mask(hipStream_t s, uint32_t V, const float *dy, const uint16_t *y, int c, int cpad,
                      uint16_t *dz);
void launch_to_bf16_pad(hipStream_t s, uint32_t V, const float *in, int c, int cpad, uint16_t *out);
void launch_pack_weights(hipStream_t s, const float *w, int cin, int cout, uint16_t *wf, int cin_pad,
                         int cout_pad, uint16_t *wt, int tin_pad, int tout_pad);
void launch_unpack_wgrad(hipStream_t s, const float *gw, int cin, int cin_pad, int cout, int cout_pad,
                         float *g);
void launch_fill_f32(hipStream_t s, float *p, size_t n, float v);

// Multi-box launches of the forward pass (ensembles): one task record per
// box and layer, the arguments of launch_conv_fwd / launch_pool_fwd.  The
// kernel variant a single-box launch picks depends on the grid and the
// channels (conv_variant); a multi-box launch runs one variant for all its
// boxes, so that every box gets the same sums as in a launch of its own.
struct ConvTask {
    NetGeom g;
    const uint16_t *in;
    const uint16_t *w;
    const float *bias;
    void *out;
    const float *proj;
    float *out2;
    const uint16_t *skip;
    uint16_t *yfine;
    int cin, cout_real, mode;
};
struct PoolTask {
    NetGeom gc;
    const uint16_t *in;
    int c;
    uint16_t *out;
};
int conv_variant(const NetGeom &g, int cin, int cout_pad);
uint32_t conv_variant_blocks(int variant, const NetGeom &g);
int launch_conv_fwd_multi(hipStream_t s, int variant, int cin, int cout_pad, const ConvTask *t, MultiMap m,
                          uint32_t blocks);
uint32_t pool_blocks(const NetGeom &gc, int c);
void launch_pool_fwd_multi(hipStream_t s, const PoolTask *t, MultiMap m, uint32_t blocks);
