// conv_types.h -- shapes shared by the 3D convolution kernels of the flow
// surrogate and their host side.
//
// Activations are channels-last BF16, [sample][z][y][x][channel], with a
// per-voxel row stride that may exceed the channel count (a layer can read or
// write a channel range of a wider concatenation buffer).  Every spatial side
// is a power of two and the volume is periodic, as the crops are.
// Weights of a 3x3x3 layer are [Cout][KP] BF16 with k = tap * Cin + ci,
// tap = (dz+1)*9 + (dy+1)*3 + (dx+1), zero-padded to KP = roundup(27 Cin, 32).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbf {

constexpr int CONV_K = 32;          // K of one MFMA step (v_mfma_f32_16x16x32_bf16)

__host__ __device__ constexpr int conv_kp(int cin, int taps) { return (taps * cin + CONV_K - 1) / CONV_K * CONV_K; }

struct ConvGeom {
    uint32_t log_s;          // log2 of the spatial side
    uint32_t nvox;           // voxels over all samples (samples * side^3)
    uint32_t in_stride;      // channels per voxel row of the input buffer
    uint32_t out_stride;     // channels per voxel row of the output buffer
    const __bf16* gate;      // ACT_RELU_GATE: the tensor whose sign gates the output (same voxels, channels)
    uint32_t gate_stride;    // channels per voxel row of gate
    float* kpart;            // forward split over input channels: FP32 partials [split][voxel][Cout]
                             // (null: no split)
    uint32_t kpart_floats;   // floats available at kpart
};

// ACT_RELU_GATE: out = gate > 0 ? x : 0, the ReLU gradient of the layer
// whose output is gate, applied to an input gradient as it is stored
enum ConvAct { ACT_NONE = 0, ACT_RELU = 1, ACT_RELU_GATE = 2 };

struct Chan {           // a channel range of a channels-last BF16 buffer
    __bf16* p;
    uint32_t stride, off;
};

}  // namespace lbf
