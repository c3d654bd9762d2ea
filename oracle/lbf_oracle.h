/*
 * lbf_oracle.h -- CPU oracle for the LBFlow time-step loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a sequential restatement of the
 * reference solver's algorithm (zhou1800/lbflow, solver/src/lbm/ sources),
 * written in plain C so that the HIP engine can be checked against it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the timed CPU baseline; the product
 * (lbflow_amd/) never links, loads or calls it.
 *
 * Every operation keeps the reference's order of floating-point operations
 * (no FMA contraction: build with -ffp-contract=off), so an FP64 engine that
 * follows the same order can be compared bit for bit.
 *
 * Pinning (see tests/test_oracle_golden.py):
 *  - uniform periodic force, 16 cases: absolute errors recorded by the
 *    reference (research/drlab-publication/paper/data/measurements.json);
 *  - 24 Poiseuille channels (scsp D3Q19 + color D2Q9, SRT/MRT): recorded
 *    velocity profiles;
 *  - 12 two-phase layered channels (color D2Q9 with interface, perturbation,
 *    recoloring and the 90-degree wall rotation): recorded profiles;
 *  - slit bundles (scsp/color, D2Q9/D3Q19, MRT): recorded Darcy histories;
 *  - the 8x10x4 SRT channel checkpoint SHA-256 recorded for the reference;
 *  - moment identities of the operators (M M^-1 = I, sum w = 1, ...);
 *  - 3D contact-angle rotation: angle/norm invariants (parity of the 3D
 *    curved-wall trajectory itself is otherwise unpinned).
 */
#ifndef LBF_ORACLE_H
#define LBF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Label types, same bit values as the reference (solver/inc/file/fileprm.h:44-53). */
enum { ORC_VD = 0x1, ORC_GV = 0x2, ORC_PB = 0x4, ORC_VB = 0x8,
       ORC_AB = 0x10, ORC_BB = 0x20, ORC_HF = 0x40, ORC_SD = 0x80 };

typedef struct orc_label {
    int32_t type;      /* ORC_* bit, 0 = label undefined */
    int32_t dir;       /* lattice direction (GV/PB/VB) */
    int32_t phase;     /* phase index (VD/GV/PB/VB/HF) */
    int32_t pad;
    double  value;     /* GV force magnitude, SD contact angle (deg) ... */
} orc_label;

typedef struct orc_desc {
    int32_t D, Q;          /* 2/9 or 3/19 */
    int32_t model;         /* 0 = scsp, 1 = color */
    int32_t equil;         /* 0 = SRT, 1 = MRT */
    int32_t harmonic;      /* color viscosity interpolation: 0 arithmetic, 1 harmonic */
    int32_t wall_gradient; /* color wallGradient: 0 legacy, 1 renormalized, 2 tanh */
    uint64_t sz[3];        /* grid size in reference order: slowest first (z,y,x / y,x) */
    double tau[2];         /* relaxation time per phase */
    double density[2];     /* reference density per phase */
    double alpha0;         /* color phase weight */
    double sigma;          /* surface tension */
    double beta;           /* recoloring beta */
    int32_t capillary;     /* color capillaryModel: 0 perturbation, 1 csf */
    int32_t curvature;     /* CSF curvatureStencil: 0 normal, 1 gradient */
} orc_desc;

typedef struct orc_sim orc_sim;

/* Returns NULL and writes a message into err on invalid input. */
orc_sim* orc_new(const orc_desc* desc, const orc_label* labels /*256*/,
                 const uint8_t* geometry, char* err, size_t errlen);
void orc_free(orc_sim* s);

/* t = 0 state: initialize, macro, body force; then collision. Returns 0 or an error code. */
int orc_init(orc_sim* s);
/* One time step t: stream, bounce back, macro, body force, collision. */
int orc_step(orc_sim* s);
/* Test hook: the streaming half of a step (streaming, bounce-back, the
 * half-permeable film and the Zou-He boundaries) applied to the current
 * populations, which then hold the pre-collision state; time is unchanged. */
void orc_stream_only(orc_sim* s);
/* Test hook: the macro variables (densities, colour indicator, velocity) of
 * the current populations, without streaming.  Returns 0 or an error code. */
int orc_macro_only(orc_sim* s);
/* Run n steps; stops at the first error. */
int orc_run(orc_sim* s, uint64_t n);
const char* orc_error(const orc_sim* s);
uint64_t orc_time(const orc_sim* s);

uint64_t orc_sites(const orc_sim* s);
/* Macro fields of the last step (after body force, before collision):
 * rho per phase (scsp: rho1 ignored) and velocity [d][site], d in reference order. */
void orc_fields(const orc_sim* s, double* rho0, double* rho1, double* u);
/* Adjacent report history of the last step: the velocity (and colour
 * indicator (rho0 - rho1) / rhoT, 0 where rhoT <= 0) as the reference copies
 * them after the boundaries and before the macro variables (scsp.cpp:82-83,
 * colorfluid.cpp:166-171). */
void orc_adjacent(const orc_sim* s, double* u, double* chi);
/* Post-collision populations in checkpoint layout [phase][q][site]. */
void orc_get_populations(const orc_sim* s, double* out);
void orc_set_populations(orc_sim* s, const double* in, uint64_t t);

/* Boundary-link tables of buildIb: linkmask[site] bit q set iff site is in ibS[q];
 * wall[site] = ibI (0 = no wall); per wall index c (1..nwall): normal[c*D+d],
 * sinca[c], cosca[c].  Returns the number of wall entries incl. the unused 0. */
uint64_t orc_links(const orc_sim* s, uint32_t* linkmask, int32_t* wall,
                   double* normal, double* sinca, double* cosca);

/* Host operators: A = MInv S M and B = MInv (I - S/2) M for relaxation time tau. */
void orc_mrt_matrices(int Q, double tau, double* A, double* B, double* M, double* MInv);

/* Colour gradient of one site after the wall treatment (diagnostic). */
void orc_color_gradient(const orc_sim* s, uint64_t site, double* g, double* norm);
/* The geometric contact-angle rotation of one wall gradient g (D entries,
 * reference component order) for the integer wall normal n and the contact
 * angle (sA, cA) = (sin, cos); out receives the rotated gradient.  plant > 0
 * plants a known mistake (1 sign of sin, 2 frame not inverted, 3 sin/cos
 * swapped, 4 gradient reversed) for the test that must catch it. */
int orc_contact_angle(int D, const double* n, const double* g, double sA, double cA, double* out, int plant);

/* Report sums (scsp and color) over the report range/slices, as in
 * SCSP::report / ColorFluid::report.  out: [velR, rhoR0, rhoR1, rhoS0_0, rhoS0_1,
 * rhoS1_0, rhoS1_1, cnt0, cnt1]. axis in reference order. */
void orc_report_sums(const orc_sim* s, int axis, uint64_t first, uint64_t last, double* out);
/* Convergence sums of Simulator::updateConvergence:
 * out[0]=sum (u-prev)^2, out[1]=sum u^2, out[2]=count, out[3]=sum (u-adj)^2. */
void orc_convergence_sums(const orc_sim* s, const double* prev, const double* adj, double* out);

#ifdef __cplusplus
}
#endif
#endif
