/*
 * lbflow.h -- C ABI of the MI355X lattice Boltzmann engine.
 *
 * The reference (zhou1800/lbflow) has no plugin API: its time-step loop is
 * the virtual Simulator::Run of a monolithic executable
 * (solver/inc/lbm/simulator.h:35, :118; solver/apps/LBSim/lbsim.cpp:40-61).
 * This header is the seam under that loop.  The drop-in executable
 * (lbflow_amd/bin/LBFlow) keeps the Run orchestration, file formats and
 * logging on the host and calls these entry points for every per-site
 * operator; any other host language binds the same symbols (ctypes, cgo,
 * JNI ...).  Plain C types only: pointers, sizes, doubles.  All device
 * memory is owned by the context; host buffers are owned by the caller and
 * copied.  One host thread per context; calls are stream-ordered and not
 * re-entrant.  Errors are int return codes (0 = success) with a message in
 * lbf_last_error(); nothing throws across the ABI.
 *
 * Index order everywhere is the reference's: z,y,x row-major, x fastest
 * (solver/inc/grid/grid.h:153-177); sizes are given slowest first, i.e.
 * sz = {Nz, Ny, Nx} for D3Q19 and {Ny, Nx} for D2Q9.
 */
#ifndef LBFLOW_H
#define LBFLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBF_ABI_VERSION 2

/* Label types: same bit values as FilePrm::IDXTYPE (solver/inc/file/fileprm.h:44-53).
 * lbf_create accepts VD, GV, PB, VB, BB, SD, and HF in the colour model
 * (half-permeable film: the site holds the label's phase only and reflects
 * the other phase, simulator.cpp:418-421, colorfluid.cpp:156-159, 752-764). */
enum lbf_label_type {
    LBF_VD = 0x1, LBF_GV = 0x2, LBF_PB = 0x4, LBF_VB = 0x8,
    LBF_AB = 0x10, LBF_BB = 0x20, LBF_HF = 0x40, LBF_SD = 0x80
};

/* One geometry label, FilePrm::IDXVAL (solver/inc/file/fileprm.h:55-90,
 * solver/src/file/fileprm.cpp:633-670).  type 0 = label not defined. */
typedef struct lbf_label {
    int32_t type;      /* lbf_label_type */
    int32_t dir;       /* lattice direction of a GV force; PB/VB: axial direction
                          1..2D of the unknown populations (scsp.cpp:188-256) */
    int32_t phase;     /* phase of VD/GV/PB/VB/HF sites */
    int32_t reserved;
    double  value;     /* GV force magnitude, SD contact angle in degrees,
                          PB density, VB speed */
} lbf_label;

enum lbf_model { LBF_MODEL_SCSP = 0, LBF_MODEL_COLOR = 1 };
enum lbf_equil { LBF_SRT = 0, LBF_MRT = 1 };
enum lbf_precision {
    LBF_FP64 = 0,          /* reference arithmetic, operation order kept: bit-exact */
    LBF_FP32 = 1           /* FP32 populations shifted by their rest equilibrium;
                              macro fields and reductions in FP64 */
};
enum lbf_create_flags {
    LBF_STREAM_SOLIDS = 0x1, /* stream populations through solid sites as the reference
                                does, so checkpoints are byte-identical (dense layout);
                                without it only fluid sites are stored and updated and
                                checkpoints hold 0 on solid sites */
    LBF_LAYOUT_DENSE = 0x2   /* without LBF_STREAM_SOLIDS: keep the dense layout and skip
                                solid sites instead of storing fluid sites only */
};

/* Colour-gradient treatment at wall sites (input.prm wallGradient,
 * colorfluid.cpp:458-492): the one-sided stencil as is, or renormalized by
 * the inverse of its fluid-link second moment (calcWallGradientInverse,
 * colorfluid.cpp:917-957); both followed by the contact-angle rotation. */
enum lbf_wall_gradient { LBF_WALL_LEGACY = 0, LBF_WALL_RENORMALIZED = 1, LBF_WALL_TANH = 2 };
/* color capillaryModel (colorfluid.cpp:292-433, 572-634): the perturbation
 * operator, or the experimental continuum-surface force (closed flow:
 * VD/GV/SD/BB labels only, as the reference requires without pressure
 * reservoirs; in FP32 the normal field is cut at |g| <= 1e-5, not 1e-9). */
enum lbf_capillary { LBF_CAPILLARY_PERTURBATION = 0, LBF_CAPILLARY_CSF = 1 };
/* CSF curvatureStencil: curvature from the unit normal field, or from the gradient */
enum lbf_curvature { LBF_CURVATURE_NORMAL = 0, LBF_CURVATURE_GRADIENT = 1 };

/* Simulation description; the fields of input.prm that the step needs
 * (solver/src/file/fileprm.cpp:364-456). */
typedef struct lbf_desc {
    int32_t  D, Q;             /* 2,9 or 3,19 */
    int32_t  model;            /* lbf_model */
    int32_t  equil;            /* lbf_equil */
    int32_t  precision;        /* lbf_precision */
    int32_t  flags;            /* lbf_create_flags */
    int32_t  harmonic;         /* color viscosityInterpolation: 0 arithmetic, 1 harmonic */
    int32_t  wall_gradient;    /* color wallGradient: lbf_wall_gradient */
    uint64_t sz[3];            /* grid sizes, slowest first */
    double   tau[2];           /* rexlationTime per phase */
    double   density[2];       /* reference density per phase */
    double   alpha0;           /* color alpha0 */
    double   sigma;            /* color surfaceTension */
    double   beta;             /* color interfaceThickness (recoloring beta) */
    int32_t  capillary;        /* color capillaryModel: lbf_capillary */
    int32_t  curvature;        /* CSF curvatureStencil: lbf_curvature */
} lbf_desc;

typedef struct lbf_ctx lbf_ctx;

/* lbf_step flags */
enum lbf_step_flags {
    LBF_STEP_FIELDS   = 0x1,   /* keep this step's macro fields (rho, u) for output/report */
    LBF_STEP_ADJACENT = 0x2    /* keep this step's velocity (and color indicator) as the
                                  adjacent history of the next report (scsp.cpp:82-83,
                                  colorfluid.cpp:166-171) */
};

/* Failure bits of lbf_status, one per reference exception text. */
enum lbf_status_bits {
    LBF_FAIL_DENSITY    = 0x1,  /* "Non-finite or non-positive fluid density." (scsp.cpp:145) /
                                   "Non-finite or non-positive total fluid density." (colorfluid.cpp:257) */
    LBF_FAIL_NEGATIVE   = 0x2,  /* "Negative component density outside roundoff tolerance." */
    LBF_FAIL_TAU        = 0x4,  /* "Invalid local relaxation time." */
    LBF_FAIL_VELOCITY   = 0x8,  /* "Non-finite fluid velocity." */
    LBF_FAIL_POPULATION = 0x10, /* "Non-finite post-collision population; no restart will be written." */
    LBF_FAIL_CAPILLARY  = 0x20  /* "Non-finite capillary force." (colorfluid.cpp:616) */
};

/* Report request: the reduction behind SCSP::report / ColorFluid::report and
 * Simulator::updateConvergence (scsp.cpp:304-372, colorfluid.cpp:814-915,
 * simulator.cpp:248-325). */
typedef struct lbf_report_req {
    int32_t  axis;             /* report axis in reference order (Qaxis[reportDir]) */
    int32_t  convergence;      /* 1: also compute the convergence sums vs. the previous report */
    int32_t  adjacent;         /* 1: the adjacent history was kept by LBF_STEP_ADJACENT */
    int32_t  reserved;
    uint64_t first, last;      /* report slab [first, last] along axis */
} lbf_report_req;

typedef struct lbf_report_out {
    double   vel_range;        /* sum of u[axis] over fluid sites in the slab */
    double   rho_range[2];     /* per-phase density sums over the slab (color) */
    double   rho_slice[2][2];  /* [first|last slice][phase] density sums */
    double   count_slice[2];   /* fluid sites in the first/last slice */
    double   conv[4];          /* sum (u-u_prev)^2, sum u^2, count, sum (u-u_adj)^2 */
    double   phase_change[2];  /* max |chi-chi_prev|, max |chi-chi_adj| (color) */
    int32_t  has_previous;     /* 0 on the first report after create/init/set_populations */
    int32_t  nonfinite;        /* 1 if a reported velocity or indicator was not finite */
} lbf_report_out;

/* --- lifetime ----------------------------------------------------------
 * Replaces the Simulator/SCSP/ColorFluid constructors
 * (solver/src/lbm/simulator.cpp:17-38, scsp.cpp:22-44, colorfluid.cpp:20-107):
 * validates labels, builds the boundary links (Simulator::buildIb,
 * simulator.cpp:386-436), the MRT operators (simulator.cpp:162-216) and the
 * wall rotations (colorfluid.cpp:1111-1129), and allocates device memory. */
int  lbf_create(const lbf_desc* desc, const lbf_label* labels /* 256 entries */,
                const uint8_t* geometry /* prod(sz) bytes */, lbf_ctx** out);
void lbf_destroy(lbf_ctx* ctx);
const char* lbf_last_error(const lbf_ctx* ctx);   /* ctx may be NULL: last create error */

/* --- time stepping -----------------------------------------------------
 * lbf_init: the t = 0 state, SCSP::Run / ColorFluid::Run initial branch
 * (scsp.cpp:60-72, colorfluid.cpp:123-138): initialize, macro, body force
 * (fields kept), collision.  lbf_step: one step t -> t+1, the loop body
 * (scsp.cpp:75-104, colorfluid.cpp:142-198): stream + bounce-back + macro
 * + body force + collision fused.  lbf_advance: n steps without fields. */
int lbf_init(lbf_ctx* ctx);
int lbf_step(lbf_ctx* ctx, uint32_t flags);
int lbf_advance(lbf_ctx* ctx, uint64_t n);
uint64_t lbf_time(const lbf_ctx* ctx);
int lbf_synchronize(lbf_ctx* ctx);

/* Failure bits raised by the device since create/init/set_populations, and
 * the first step at which each bit was raised (6 entries, one per bit in bit
 * order, UINT64_MAX = never). */
int lbf_status(lbf_ctx* ctx, uint32_t* bits, uint64_t* first_step /* 6 */);
const char* lbf_status_message(uint32_t bit);

/* --- data exchange -----------------------------------------------------
 * lbf_fields: the macro fields kept by the last LBF_STEP_FIELDS step (or
 * init), as the reference's VTI arrays (filevti.cpp:113-172): rho per phase
 * and u[d][site] with d in reference order; rho1 may be NULL for scsp.
 * lbf_get/set_populations: post-collision populations in the checkpoint
 * layout [phase][q][site] of SCSP::saveRst / ColorFluid::saveRst
 * (scsp.cpp:290-302, colorfluid.cpp:789-812). */
int lbf_fields(lbf_ctx* ctx, double* rho0, double* rho1, double* u);
int lbf_get_populations(lbf_ctx* ctx, double* out);
int lbf_set_populations(lbf_ctx* ctx, const double* in, uint64_t t);

/* Device reductions of one report (see lbf_report_req). */
int lbf_report(lbf_ctx* ctx, const lbf_report_req* req, lbf_report_out* out);

/* lbf_report without the context's reduction buffers: no counterpart in the
 * reference.  For callers that cannot allocate device memory at call time
 * (serving loops, harnesses that forbid hipMalloc): the block partials and
 * the sums go to a static device array of 4096 x 16 + 16 doubles (512 KiB
 * per GPU, held until the process ends), which covers the at most 4096
 * reduction blocks of any lattice; a larger need is refused with an error.
 * It calls no allocating routine, and gives bitwise the sums lbf_report
 * gives on the same state.  Every call on one GPU shares that array and
 * overwrites it, so calls that overlap on one GPU (from several threads or
 * contexts) race; the context's GPU must be the current device. */
int lbf_report_static(lbf_ctx* ctx, const lbf_report_req* req, lbf_report_out* out);

/* --- introspection (parity tests) --------------------------------------
 * linkmask[site]: bit q set iff the site is in ibS[q]; wall[site] = ibI;
 * normal/sin/cos per wall entry (entry 0 unused).  Returns the number of
 * wall entries including entry 0, or 0 on error. */
uint64_t lbf_links(lbf_ctx* ctx, uint32_t* linkmask, int32_t* wall,
                   double* normal, double* sinca, double* cosca);
uint64_t lbf_sites(const lbf_ctx* ctx);
/* Device bytes the context holds. */
uint64_t lbf_device_bytes(const lbf_ctx* ctx);

/* Name of the kernel that lbf_step launches for this context, and the
 * algorithmic HBM bytes one launch moves (for roofline accounting). */
const char* lbf_step_kernel(const lbf_ctx* ctx);
double lbf_step_bytes(const lbf_ctx* ctx);

/* --- measurement (no reference counterpart) ----------------------------
 * lbf_timed_advance: n steps (no fields) bracketed by two HIP events on the
 * context's stream; *ms = elapsed device time.  lbf_copy_bandwidth: a plain
 * device copy kernel over `bytes` bytes, `iters` launches, timed with HIP
 * events; returns (read + write bytes) / time in GB/s, or a negative value
 * on error. */
int lbf_timed_advance(lbf_ctx* ctx, uint64_t n, double* ms);
double lbf_copy_bandwidth(uint64_t bytes, int iters);

/* --- batched crops (no reference counterpart) ---------------------------
 * Many independent single-phase crops (VD/GV fluid, BB/SD solids, periodic
 * per crop) advanced by one persistent FP32 kernel.  A crop's trajectory is
 * the one lbf_create(FP32, fluid-only layout) computes for that crop alone:
 * same operator (SCSP::equilSRT/equilMRT, scsp.cpp:374-464), same FP32
 * arithmetic.  The replaced reference workflow is one LBFlow process per
 * crop (scripts/rock_workflow.py:63-69 loops over crops).  lbf_batch_run
 * applies the stop rule of Simulator::updateConvergence (simulator.cpp:
 * 248-325: velocity and adjacent-velocity residual below tol, or their RMS
 * change <= 1e-12, for 3 consecutive reports) to every crop independently;
 * converged crops stop stepping. */
typedef struct lbf_batch lbf_batch;

typedef struct lbf_batch_desc {
    int32_t  D, Q;             /* 2,9 or 3,19 */
    int32_t  equil;            /* lbf_equil */
    uint32_t ncrops;
    double   tau;              /* relaxation time */
    double   density;          /* reference density rho0 */
} lbf_batch_desc;

typedef struct lbf_crop {
    uint64_t       sz[3];      /* crop sizes, slowest first */
    const uint8_t* geometry;   /* prod(sz) label bytes, z,y,x order */
} lbf_crop;

typedef struct lbf_crop_result {
    uint64_t steps;            /* step of the crop's last report */
    int32_t  converged;
    int32_t  reserved;
    double   residual;         /* velocity residual of the last report */
    double   adjacent_residual;
    double   k_darcy;          /* nu <u_axis> / F over all crop voxels, lattice units
                                  (scripts/rock_workflow.py:63-69) */
} lbf_crop_result;

int  lbf_batch_create(const lbf_batch_desc* desc, const lbf_label* labels /* 256 */,
                      const lbf_crop* crops /* desc->ncrops */, lbf_batch** out);
void lbf_batch_destroy(lbf_batch* batch);
const char* lbf_batch_error(const lbf_batch* batch);  /* batch may be NULL: last create error */
/* Total voxels of all crops; *fluid (if not NULL) = fluid sites. */
uint64_t lbf_batch_sites(const lbf_batch* batch, uint64_t* fluid);
/* Times `steps` steps for each launch shape of the persistent step (64, 128
 * or 256 threads per workgroup, work units of 64, 128 or 256 slots) and each
 * persistent grid size (1..128 workgroups per CU), and keeps the fastest.
 * Steps the current state: call before lbf_batch_init.  Every shape gives
 * bitwise the same trajectory. */
int lbf_batch_tune(lbf_batch* batch, uint32_t steps, uint32_t* per_cu, double* ms_per_step);
/* The step's launch shape in use: threads per workgroup, slots per work unit
 * and workgroups of the persistent grid (any pointer may be NULL). */
int lbf_batch_shape(const lbf_batch* batch, uint32_t* threads, uint32_t* unit, uint32_t* workgroups);
/* t = 0 state of every crop (SCSP::Run initial branch, scsp.cpp:60-72). */
int lbf_batch_init(lbf_batch* batch);
/* n steps of every crop, no convergence bookkeeping. */
int lbf_batch_advance(lbf_batch* batch, uint64_t n);
/* Runs every crop to convergence along `axis` (reference order), reporting
 * every `interval` steps (rounded up to an even number >= 2), at most
 * max_steps; results[ncrops]; *seconds = wall time of the call. */
int lbf_batch_run(lbf_batch* batch, int axis, double tol, uint32_t interval, uint64_t max_steps,
                  lbf_crop_result* results, double* seconds);
/* Crop c's populations, [q][site] over all crop voxels (0 on solids), as FP64. */
int lbf_batch_crop_populations(lbf_batch* batch, uint32_t c, double* out);
/* Current step of the batch. */
uint64_t lbf_batch_time(const lbf_batch* batch);
/* Macro fields of crop c after its last report: fields[site][0..D-1] = u
 * (reference axis order), fields[site][D] = rho, over all crop voxels
 * (solids 0).  `fields` is DEVICE memory of prod(sz) * (D+1) floats. */
int lbf_batch_crop_fields(lbf_batch* batch, uint32_t c, float* fields);
/* Sizes of crop c (slowest axis first) into sz[0..D-1]; returns D, or 0 for
 * a crop that does not exist. */
int lbf_batch_crop_size(const lbf_batch* batch, uint32_t c, uint64_t* sz);
/* Warm start: set crop c to the equilibrium of the given fields (same layout,
 * DEVICE memory), shifted by half the body force so that the first macro
 * step returns u; the FP32 counterpart of SCSP::initialize (scsp.cpp:116-134)
 * with a velocity.  Fluid sites whose cluster does not wrap around the box
 * along the force (no steady flow: isolated pockets, dead clusters) keep the
 * rest state of lbf_batch_init.  Call after lbf_batch_init and before
 * lbf_batch_run. */
int lbf_batch_set_fields(lbf_batch* batch, uint32_t c, const float* fields);
/* Consistent initialisation after lbf_batch_set_fields: `steps` sweeps that
 * stream, give the streamed populations the momentum of the prescribed
 * velocity and collide (after Mei, Luo, Lallemand & d'Humieres, Computers &
 * Fluids 35 (2006) 855), so density and non-equilibrium populations settle
 * to the given flow; then two free steps, each averaged with the state before
 * it, which cancels the non-decaying mode that flips sign every step (a
 * prediction that is not smooth on the lattice scale excites it).  All of
 * these count as steps. */
int lbf_batch_relax(lbf_batch* batch, uint32_t steps);

/* --- learned warm-start network (no reference counterpart) -------------
 * A periodic 3D U-Net (4 levels, 32-256 channels, 5.8 M parameters) mapping
 * a crop's fluid mask to its steady (u, rho); BF16 MFMA convolutions with
 * FP32 accumulation, FP32 master weights, Adam.  Masks, targets and outputs
 * are DEVICE memory: mask[sample][site] (1 = fluid), target/out
 * [sample][site][4] FP32 in the channel order the caller chose
 * (lbflow_amd/surrogate.py uses u_z, u_y, u_x, rho, normalised).  Sites are
 * z,y,x with side `side`. */
typedef struct lbf_net lbf_net;
typedef struct lbf_adam {
    float lr, beta1, beta2, eps;
    float weights[4];          /* loss weight per output channel */
} lbf_adam;

int  lbf_net_create(uint32_t side, uint32_t capacity /* samples per call */, uint64_t seed, lbf_net** out);
void lbf_net_destroy(lbf_net* net);
const char* lbf_net_error(const lbf_net* net);   /* net may be NULL: last create error */
uint64_t lbf_net_param_count(const lbf_net* net);
/* Parameters in layer order; per layer W[cout][tap][cin] then bias[cout]
 * (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)), FP32, host memory. */
int lbf_net_get_params(lbf_net* net, float* host);
int lbf_net_set_params(lbf_net* net, const float* host);
int lbf_net_forward(lbf_net* net, const uint8_t* mask, uint32_t count, float* out);
/* Warm start of the first `count` crops of a batch from the network (no
 * reference counterpart).  masks: HOST memory, count x side^3 bytes, crop
 * order, z,y,x with x fastest, 1 = pore; every one of those crops must be a
 * side^3 D3Q19 crop.  `capacity` crops at a time go through the network;
 * each prediction p becomes fields f_c = p_c scale[c] + offset[c] (the
 * velocity channels 0 on solid voxels), loaded with lbf_batch_set_fields,
 * then `relax` sweeps of lbf_batch_relax.  Call after lbf_batch_init. */
int lbf_batch_warm_start(lbf_batch* batch, lbf_net* net, const uint8_t* masks, uint32_t count,
                         const float* scale /* 4 */, const float* offset /* 4 */, uint32_t relax);
/* Loss (mean over sites of sum_c w_c (out_c - target_c)^2) and its gradient
 * in the parameter layout (host memory, may be NULL); no update. */
int lbf_net_gradients(lbf_net* net, const uint8_t* mask, const float* target, uint32_t count,
                      const float* weights /* 4 */, double* loss, float* grad_host);
int lbf_net_train_step(lbf_net* net, const uint8_t* mask, const float* target, uint32_t count,
                       const lbf_adam* opt, double* loss);
/* `steps` training steps timed with HIP events; *ms per step and the
 * convolution FLOPs of one step (forward, input and weight gradients). */
int lbf_net_timed_train(lbf_net* net, const uint8_t* mask, const float* target, uint32_t count,
                        const lbf_adam* opt, uint32_t steps, double* ms, double* flops_per_step);
/* One periodic convolution on DEVICE buffers (kernel tests): channels-last
 * BF16 in/out with row strides, weights [cout][roundup(taps*cin, 32)] BF16,
 * optional FP32 bias, ReLU, FP32 output; wgrad writes dW [cout][kp] FP32. */
int lbf_conv3d_forward(int cin, int cout, int taps, uint32_t side, uint32_t samples, const void* in,
                       uint32_t in_stride, const void* w, const float* bias, void* out, uint32_t out_stride,
                       int relu, int out_f32);
int lbf_conv3d_wgrad(int cin, int cout, int taps, uint32_t side, uint32_t samples, const void* in,
                     uint32_t in_stride, const void* dz, uint32_t dz_stride, uint32_t splits, float* dw);
/* Cubic side the network was built for. */
uint32_t lbf_net_side(const lbf_net* net);

/* --- device buffers (no reference counterpart) ----------------------------
 * Plain device memory on the library's own HIP runtime, for callers that
 * have no GPU runtime of their own and must hand DEVICE pointers to the
 * entry points above (lbf_net_forward, lbf_conv3d_*, lbf_batch_crop_fields).
 * lbf_buffer_alloc zero-fills; all calls are synchronous. */
int  lbf_buffer_alloc(uint64_t bytes, void** ptr);
void lbf_buffer_free(void* ptr);
int  lbf_buffer_upload(void* dst, const void* host, uint64_t bytes);
int  lbf_buffer_download(void* host, const void* src, uint64_t bytes);

/* --- training set of the warm-start network (no reference counterpart) ----
 * count cubic samples of `side` voxels (a power of two) held on the device:
 * per sample a pore mask byte and (u_z, u_y, u_x, rho) floats per voxel.
 * The reference workflow has no learned warm start; the training data are
 * steady fields of the batched engine (lbf_batch_run), taken over without
 * leaving the device, scaled as targets u nu / F and (rho - 1) / F over
 * their RMS on pore voxels (0 on solid voxels), and fed to the network in
 * minibatches mapped by one of 16 lattice symmetries: bit 0 flips z, bit 1
 * flips y, bit 3 flips x (u_z, u_y, u_x as vectors; u_y, u_z and the
 * density deviation change sign under the x flip, the force staying +x),
 * bit 2 swaps y and z after the flips (u_y and u_z swapped, carried across
 * their RMS).  Flips of y, z and the swap map steady fields of the periodic
 * D3Q19 lattice exactly; the x flip holds in the Stokes limit. */
typedef struct lbf_dataset lbf_dataset;
int  lbf_dataset_create(uint32_t side, uint32_t count, lbf_dataset** out);
void lbf_dataset_destroy(lbf_dataset* data);
const char* lbf_dataset_error(const lbf_dataset* data);   /* data may be NULL: last create error */
/* Masks of samples first..first+n-1 from host memory, n * side^3 bytes. */
int lbf_dataset_set_masks(lbf_dataset* data, uint32_t first, uint32_t n, const uint8_t* masks);
/* Targets (or raw fields) of samples first..first+n-1 from host memory. */
int lbf_dataset_set_targets(lbf_dataset* data, uint32_t first, uint32_t n, const float* targets);
/* Steady fields (lbf_batch_crop_fields) of crops crop0..crop0+n-1 of a batch
 * of side^3 crops into samples first..first+n-1. */
int lbf_dataset_take_fields(lbf_dataset* data, uint32_t first, lbf_batch* batch, uint32_t crop0, uint32_t n);
/* Raw fields -> targets, once, in place.  fit = 1: rms[4] receives the RMS
 * of each raw target channel over all pore voxels (FP64, fixed order);
 * fit = 0: rms[4] is given. */
int lbf_dataset_normalise(lbf_dataset* data, double tau, double force, int fit, double* rms);
/* Sample i mapped by symmetry sym (0..15) into host memory (mask side^3
 * bytes, target side^3 * 4 floats; either may be NULL). */
int lbf_dataset_sample(lbf_dataset* data, uint32_t i, uint32_t sym, uint8_t* mask, float* target);
/* Device addresses of the masks and targets of sample 0 (samples follow
 * each other), for lbf_net_forward / lbf_net_timed_train. */
int lbf_dataset_device(const lbf_dataset* data, uint8_t** masks, float** targets);
/* One Adam step (lbf_net_train_step) on samples idx[0..count) mapped by
 * symmetry sym; idx is host memory; count <= the network's capacity. */
int lbf_net_train_samples(lbf_net* net, lbf_dataset* data, const uint32_t* idx, uint32_t count, uint32_t sym,
                          const lbf_adam* opt, double* loss);

#ifdef __cplusplus
}
#endif
#endif
