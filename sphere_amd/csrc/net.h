// net.h -- the Darcy correction network (net.hip) as used by the driver.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/sphere_hip.h"

// BF16 copies of the FP32 parameters for the passes that follow
void net_pack(sph_net *n, const float *params, hipStream_t s);
// forward pass (input [V][3] -> output [V]) with the last packed weights
int net_forward_packed(sph_net *n, const float *params, const float *input, float *output, hipStream_t s);
uint16_t *net_input_bf16(sph_net *n);

// The forward passes of several networks (the boxes of an ensemble, each
// with its own grid, packed weights and buffers) as multi-box launches: per
// layer one launch for each kernel variant the boxes' grids need, every box
// getting the sums of net_forward_packed.  The input is each network's BF16
// input rows (net_input_bf16).
struct NetMulti;
NetMulti *net_multi_create(sph_net *const *nets, const float *const *params, float *const *outputs, int n);
int net_multi_forward(NetMulti *m, hipStream_t s);
void net_multi_destroy(NetMulti *m);
