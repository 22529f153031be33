// net.h -- the Darcy correction network (net.hip) as used by the driver.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/sphere_hip.h"

// BF16 copies of the FP32 parameters for the passes that follow
void net_pack(sph_net *n, const float *params, hipStream_t s);
// forward pass (input [V][3] -> output [V]) with the last packed weights
int net_forward_packed(sph_net *n, const float *params, const float *input, float *output, hipStream_t s);
uint16_t *net_input_bf16(sph_net *n);
