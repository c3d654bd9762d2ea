// batch_types.h -- structures shared by the batched-crop kernels and their host side.
#pragma once
#include "collide_f32.h"
#include "device.h"

namespace lbf {

constexpr uint32_t TILE = 256;
constexpr uint32_t NO_SITE = 0xffffffffu;

struct CropDev {
    GridDev g;            // crop grid (nx, ny, nz, fast divisions)
    uint32_t slot_base;   // first slot of the crop (multiple of TILE)
    uint32_t rank_base;   // first word of the crop's rank index
    uint32_t nfluid;
    uint32_t tile_begin, tile_end;
    uint32_t pad;
};

struct BatchParams {
    uint64_t S;           // slots per population plane (all crops, padded)
    uint32_t flags;       // bit0: write velocity, bit1: write adjacent velocity,
                          // bit2: collide with the prescribed velocity upres
    uint32_t pad;
    unsigned long long t;
    float inv_rho0;
    RelaxF32 relax;
};

// Launch shape of the persistent step: BS threads per workgroup walk work
// units of UNIT slots (inside one tile); the autotuner picks it with the grid.
struct BatchShape {
    uint32_t block = 256, unit = 256;
};

struct BatchBuffers {
    const void* fin;
    void* fout;
    const uint32_t* tiles;      // active tiles
    const uint32_t* ntiles;     // number of active tiles (device)
    const uint32_t* tile_crop;  // crop of every tile
    const CropDev* crops;
    const uint32_t* site;       // crop-local site of every slot, NO_SITE on padding
    const uint8_t* lab8;        // geometry label of every slot
    const uint32_t* desc;       // per slot: label << DESC_LABEL_SHIFT | bit op(q) where x - e_q is solid
    const uint2* rank;          // per crop, per 32 sites: {fluid sites before, fluid bit mask}
    const LabelF* lab;
    float* ucur;                // velocity of the last fields step [d][slot]
    float* uprev;               // velocity of the previous report
    float* uadj;                // velocity of the step before a report
    float* upres;               // prescribed velocity of the consistent-initialisation sweeps
    DevStatus* status;
};

}  // namespace lbf
