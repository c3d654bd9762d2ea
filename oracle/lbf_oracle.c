/*
 * lbf_oracle.c -- CPU oracle for the LBFlow time-step loop.
 *
 * TEST INFRASTRUCTURE ONLY (see lbf_oracle.h).  A sequential C restatement
 * of the reference algorithm; each function names the reference lines it
 * follows.  Streaming is done the reference's way (push over every site,
 * then halfway bounce-back from the per-direction link lists), which is a
 * different formulation from the engine's fused pull kernel.
 *
 * Build: gcc -O2 -std=c99 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 */
#include "lbf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXQ 19
#define SOLID_DENS_N 86.0   /* colorfluid.cpp:49 */

/* ---------------------------------------------------------------- lattice */
/* gridspace.cpp:82-96 (D2Q9) and :106-143 (D3Q19); components in reference
 * order (y,x) / (z,y,x). */
static const int QI9[9][2] = {{0,0},{0,1},{1,0},{0,-1},{-1,0},{1,1},{1,-1},{-1,-1},{-1,1}};
static const int OP9[9] = {0,3,4,1,2,7,8,5,6};
static const int RELD9[2][6] = {{2,5,6,4,7,8},{1,5,8,3,6,7}};
static const int QI19[19][3] = {{0,0,0},{0,0,1},{0,0,-1},{0,1,0},{0,-1,0},{1,0,0},{-1,0,0},
    {0,1,1},{0,1,-1},{0,-1,1},{0,-1,-1},{1,0,1},{1,0,-1},{-1,0,1},{-1,0,-1},
    {1,1,0},{1,-1,0},{-1,1,0},{-1,-1,0}};
static const int OP19[19] = {0,2,1,4,3,6,5,10,9,8,7,14,13,12,11,18,17,16,15};
static const int RELD19[3][10] = {{5,11,12,15,16,6,13,14,17,18},
                                  {3,8,7,15,17,4,9,10,16,18},
                                  {1,7,9,11,13,2,8,10,12,14}};

/* Qrel / Qprep / Qaxis / Qsign of the axial directions, in the reference's
 * list order (gridspace.cpp:79-85 D2Q9, :123-137 D3Q19); used by the
 * Zou-He boundaries. */
static const int REL9[5][3] = {{0,0,0},{1,5,8},{2,5,6},{3,6,7},{4,7,8}};
static const int PERP9[5][3] = {{0,0,0},{0,2,4},{0,1,3},{0,2,4},{0,1,3}};
static const int AXIS9[5] = {0,1,0,1,0};
static const int SIGN9[5] = {0,1,1,-1,-1};
static const int REL19[7][5] = {{0,0,0,0,0},
    {1,7,9,11,13},{2,8,10,12,14},{3,8,7,15,17},{4,9,10,16,18},{5,11,12,15,16},{6,13,14,17,18}};
static const int PERP19[7][9] = {{0,0,0,0,0,0,0,0,0},
    {0,3,4,5,6,15,16,17,18},{0,3,4,5,6,15,16,17,18},
    {0,1,2,5,6,11,12,13,14},{0,1,2,5,6,11,12,13,14},
    {0,1,2,3,4,7,8,9,10},{0,1,2,3,4,7,8,9,10}};
static const int AXIS19[7] = {0,2,2,1,1,0,0};
static const int SIGN19[7] = {0,1,-1,1,-1,1,-1};

/* simulator.cpp:162-197 */
static const double M9[81] = {
     1, 1, 1, 1, 1, 1, 1, 1, 1,
    -4,-1,-1,-1,-1, 2, 2, 2, 2,
     4,-2,-2,-2,-2, 1, 1, 1, 1,
     0, 1, 0,-1, 0, 1,-1,-1, 1,
     0,-2, 0, 2, 0, 1,-1,-1, 1,
     0, 0, 1, 0,-1, 1, 1,-1,-1,
     0, 0,-2, 0, 2, 1, 1,-1,-1,
     0, 1,-1, 1,-1, 0, 0, 0, 0,
     0, 0, 0, 0, 0, 1,-1, 1,-1};
static const double M19[361] = {
      1,  1,  1,  1,  1,  1,  1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
    -30,-11,-11,-11,-11,-11,-11, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8,
     12, -4, -4, -4, -4, -4, -4, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
      0,  1, -1,  0,  0,  0,  0, 1,-1, 1,-1, 1,-1, 1,-1, 0, 0, 0, 0,
      0, -4,  4,  0,  0,  0,  0, 1,-1, 1,-1, 1,-1, 1,-1, 0, 0, 0, 0,
      0,  0,  0,  1, -1,  0,  0, 1, 1,-1,-1, 0, 0, 0, 0, 1,-1, 1,-1,
      0,  0,  0, -4,  4,  0,  0, 1, 1,-1,-1, 0, 0, 0, 0, 1,-1, 1,-1,
      0,  0,  0,  0,  0,  1, -1, 0, 0, 0, 0, 1, 1,-1,-1, 1, 1,-1,-1,
      0,  0,  0,  0,  0, -4,  4, 0, 0, 0, 0, 1, 1,-1,-1, 1, 1,-1,-1,
      0,  2,  2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1,-2,-2,-2,-2,
      0, -4, -4,  2,  2,  2,  2, 1, 1, 1, 1, 1, 1, 1, 1,-2,-2,-2,-2,
      0,  0,  0,  1,  1, -1, -1, 1, 1, 1, 1,-1,-1,-1,-1, 0, 0, 0, 0,
      0,  0,  0, -2, -2,  2,  2, 1, 1, 1, 1,-1,-1,-1,-1, 0, 0, 0, 0,
      0,  0,  0,  0,  0,  0,  0, 1,-1,-1, 1, 0, 0, 0, 0, 0, 0, 0, 0,
      0,  0,  0,  0,  0,  0,  0, 0, 0, 0, 0, 0, 0, 0, 0, 1,-1,-1, 1,
      0,  0,  0,  0,  0,  0,  0, 0, 0, 0, 0, 1,-1,-1, 1, 0, 0, 0, 0,
      0,  0,  0,  0,  0,  0,  0, 1,-1, 1,-1,-1, 1,-1, 1, 0, 0, 0, 0,
      0,  0,  0,  0,  0,  0,  0,-1,-1, 1, 1, 0, 0, 0, 0, 1,-1, 1,-1,
      0,  0,  0,  0,  0,  0,  0, 0, 0, 0, 0, 1, 1,-1,-1,-1,-1, 1, 1};

/* ------------------------------------------------------------- dense ops */
/* fn.cpp:29-39, inner products from 0.0 in index order. */
static void mat_vec(const double* m, const double* v, int rows, int cols, double* r)
{
    for (int i = 0; i < rows; ++i) {
        double s = 0.0;
        for (int j = 0; j < cols; ++j) s = s + v[j] * m[i * cols + j];
        r[i] = s;
    }
}
/* fn.cpp:71-83 (MatTranspose(mat, m, n)) */
static void mat_transpose(const double* mat, int m, int n, double* r)
{
    int k = 0;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) r[k++] = mat[i + j * n];
}
/* fn.cpp:53-69 (MatMulMat(mat1, mat2, m, p, n)) */
static void mat_mul(const double* a, const double* b, int m, int p, int n, double* r)
{
    double bt[MAXQ * MAXQ];
    mat_transpose(b, p, n, bt);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < p; ++k) s = s + a[i * p + k] * bt[j * p + k];
            r[i * n + j] = s;
        }
}
/* fn.cpp:41-51 */
static void mat_diag(const double* a, const double* d, int m, double* r)
{
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) r[i * m + j] = a[i * m + j] * d[j];
}
/* fn.cpp:102-116 */
static void inv_ortho_row(const double* mat, int m, int n, double* r)
{
    for (int j = 0; j < m; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s = s + mat[j * n + k] * mat[j * n + k];
        s = 1 / s;
        for (int i = 0; i < n; ++i) r[m * i + j] = mat[j * n + i] * s;
    }
}
/* fn.cpp:118-164 */
static void rotate_t2d(double o, double a, double* r)
{
    double c = 1, s = 0, h = sqrt(o * o + a * a);
    if (h != 0) { c = a / h; s = o / h; }
    r[0] = c; r[1] = -s; r[2] = s; r[3] = c;
}
static void rotate_tx3d(double o, double a, double* r)
{
    double c = 1, s = 0, h = sqrt(o * o + a * a);
    if (h != 0) { c = a / h; s = o / h; }
    double t[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
    memcpy(r, t, sizeof t);
}
static void rotate_tz3d(double o, double a, double* r)
{
    double c = 1, s = 0, h = sqrt(o * o + a * a);
    if (h != 0) { c = a / h; s = o / h; }
    double t[9] = {1, 0, 0, 0, c, -s, 0, s, c};
    memcpy(r, t, sizeof t);
}

/* simulator.cpp:199-216 (legacy spectrum) */
static void mrt_rates(int Q, double rt, double* S)
{
    const double su = 1.0 / rt;
    const double sq = 8.0 * (2.0 - su) / (8.0 - su);
    if (Q == 9) {
        double s[9] = {0, su, su, 0, sq, 0, sq, su, su};
        memcpy(S, s, sizeof s);
    } else {
        double s[19] = {0, su, su, 0, sq, 0, sq, 0, sq, su, su, su, su, su, su, su, sq, sq, sq};
        memcpy(S, s, sizeof s);
    }
}

void orc_mrt_matrices(int Q, double tau, double* A, double* B, double* Mout, double* MInvOut)
{
    const double* M = Q == 9 ? M9 : M19;
    double MInv[MAXQ * MAXQ], S[MAXQ], H[MAXQ], T[MAXQ * MAXQ] = {0};
    inv_ortho_row(M, Q, Q, MInv);
    mrt_rates(Q, tau, S);
    for (int k = 0; k < Q; ++k) H[k] = 1.0 - 0.5 * S[k];
    mat_diag(MInv, S, Q, T);
    if (A) mat_mul(T, M, Q, Q, Q, A);
    mat_diag(MInv, H, Q, T);
    if (B) mat_mul(T, M, Q, Q, Q, B);
    if (Mout) memcpy(Mout, M, sizeof(double) * Q * Q);
    if (MInvOut) memcpy(MInvOut, MInv, sizeof(double) * Q * Q);
}

/* ------------------------------------------------------------- simulator */
struct orc_sim {
    orc_desc d;
    int D, Q, NP;
    uint64_t N, sz[3], stride[3];
    int qi[MAXQ][3];
    int op[MAXQ];
    double wi[MAXQ], qnorm[MAXQ], qp[MAXQ][3];
    orc_label lab[256];
    uint8_t* raw;
    /* sites grouped by label, ascending (ROI::getRawIdx) */
    uint64_t* bylabel; uint64_t start[257];
    uint64_t* adj;          /* N*Q periodic neighbours (Grid::calcAdjI) */
    uint64_t* ibs[MAXQ]; uint64_t nibs[MAXQ];
    uint64_t* ibf[2][MAXQ]; uint64_t nibf[2][MAXQ];   /* ibF[phase][q]: q-neighbour is an HF site of that phase */
    int32_t* ibI;
    double* ibNorm; double* ibSin; double* ibCos; uint64_t nwall;
    double* tib; double* tibInv;   /* per wall index D*D */
    double* ginv;                  /* per wall index D*D: renormalized wall gradient */
    /* CSF (capillaryModel csf): colour gradient, curvature field and force, [d][site] */
    double* capGrad; double* capField; double* capForce;
    double* f[2]; double* tmp[2];  /* [q][N] per phase */
    double* dens[2]; double* densT; double* densN; double* tau; double* v;
    double* adjV; double* adjChi;  /* adjacent report history of the last step */
    double A[MAXQ * MAXQ], AT[MAXQ * MAXQ], B[MAXQ * MAXQ];
    double M[MAXQ * MAXQ], MInv[MAXQ * MAXQ];
    double ai[2][MAXQ], bi[MAXQ];
    int uniformTau;
    uint64_t t;
    char err[256];
};

static int is_solid(int type) { return type == ORC_BB || type == ORC_SD; }
static int calc_capillary_force(orc_sim* s);

#define F(s, ph, q, i) ((s)->f[ph][(uint64_t)(q) * (s)->N + (i)])
#define V(s, d, i) ((s)->v[(uint64_t)(d) * (s)->N + (i)])

static int fail(orc_sim* s, const char* msg)
{
    snprintf(s->err, sizeof s->err, "%s", msg);
    return 1;
}

/* Grid::calcAdjI (grid.h:179-189): periodic neighbour of i along q. */
static uint64_t calc_adj(const orc_sim* s, uint64_t i, int q)
{
    uint64_t a = i;
    for (int d = 0; d < s->D; ++d) {
        uint64_t block = s->stride[d] * s->sz[d];
        int64_t e = s->qi[q][d];
        uint64_t shifted = (uint64_t)((int64_t)i + e * (int64_t)s->stride[d] + (int64_t)block) % block;
        a += shifted - i % block;
    }
    return a;
}

static void force_for_label(const orc_sim* s, int label, double* force)
{
    /* simulator.cpp:218-228 */
    for (int d = 0; d < s->D; ++d) force[d] = 0.0;
    if (s->lab[label].type != ORC_GV) return;
    for (int d = 0; d < s->D; ++d) force[d] = s->qp[s->lab[label].dir][d] * s->lab[label].value;
}

static int label_in(const orc_sim* s, int label, int mask) { return (s->lab[label].type & mask) != 0; }

/* simulator.cpp:230-246 */
static void force_source(const orc_sim* s, uint64_t i, const double* force, double* src)
{
    double uf = 0.0;
    for (int d = 0; d < s->D; ++d) uf += V(s, d, i) * force[d];
    for (int q = 0; q < s->Q; ++q) {
        double eu = 0.0, ef = 0.0;
        for (int d = 0; d < s->D; ++d) {
            eu += s->qi[q][d] * V(s, d, i);
            ef += s->qi[q][d] * force[d];
        }
        src[q] = s->wi[q] * (3.0 * (ef - uf) + 9.0 * eu * ef);
    }
}

static void setup_lattice(orc_sim* s)
{
    const double ns = sqrt(2.0), sq = ns / 2;
    s->NP = s->d.model == 1 ? 2 : 1;
    for (int q = 0; q < s->Q; ++q) {
        for (int d = 0; d < 3; ++d) s->qi[q][d] = 0;
        for (int d = 0; d < s->D; ++d) {
            s->qi[q][d] = s->Q == 9 ? QI9[q][d] : QI19[q][d];
            s->qp[q][d] = s->qi[q][d] == 0 ? 0.0 : (q <= 2 * s->D ? (double)s->qi[q][d]
                                                               : (s->qi[q][d] > 0 ? sq : -sq));
        }
        s->op[q] = s->Q == 9 ? OP9[q] : OP19[q];
        s->qnorm[q] = q == 0 ? 0 : (q <= 2 * s->D ? 1 : ns);
    }
    /* simulator.cpp:438-463 */
    if (s->Q == 9) {
        double w[9] = {4.0/9, 1.0/9, 1.0/9, 1.0/9, 1.0/9, 1.0/36, 1.0/36, 1.0/36, 1.0/36};
        memcpy(s->wi, w, sizeof w);
    } else {
        double w[19] = {1.0/3, 1.0/18, 1.0/18, 1.0/18, 1.0/18, 1.0/18, 1.0/18,
                        1.0/36, 1.0/36, 1.0/36, 1.0/36, 1.0/36, 1.0/36,
                        1.0/36, 1.0/36, 1.0/36, 1.0/36, 1.0/36, 1.0/36};
        memcpy(s->wi, w, sizeof w);
    }
}

/* Simulator::buildIb, simulator.cpp:386-436 */
static void build_ib(orc_sim* s)
{
    static const double pi = 3.1415926535897932384626433832795;
    const int set = ORC_VD | ORC_PB | ORC_VB | ORC_GV | ORC_HF;
    uint64_t c = 1;
    for (int q = 0; q < s->Q; ++q) { s->ibs[q] = malloc(sizeof(uint64_t) * (s->N + 1)); s->nibs[q] = 0; }
    for (int ph = 0; ph < 2; ++ph)
        for (int q = 0; q < s->Q; ++q) { s->ibf[ph][q] = malloc(sizeof(uint64_t) * (s->N + 1)); s->nibf[ph][q] = 0; }
    s->ibNorm = calloc((s->N + 1) * 3, sizeof(double));
    s->ibSin = calloc(s->N + 1, sizeof(double));
    s->ibCos = calloc(s->N + 1, sizeof(double));
    for (int label = 0; label < 256; ++label) {
        if (!label_in(s, label, set)) continue;
        for (uint64_t k = s->start[label]; k < s->start[label + 1]; ++k) {
            uint64_t i = s->bylabel[k];
            int isIB = 0, nCA = 0;
            double norm[3] = {0.0, 0.0, 0.0}, ca = 0;
            for (int q = 1; q < s->Q; ++q) {
                uint64_t a = s->adj[i * s->Q + q];
                int type = s->lab[s->raw[a]].type;
                if (is_solid(type)) {
                    s->ibs[q][s->nibs[q]++] = i;
                    isIB = 1;
                    for (int d = 0; d < s->D; ++d) norm[d] -= s->qi[q][d];
                    ca += s->lab[s->raw[a]].value;
                    nCA++;
                } else if (type == ORC_HF) {   /* half-permeable film, simulator.cpp:418-421 */
                    const int phs = s->lab[s->raw[a]].phase;
                    s->ibf[phs][q][s->nibf[phs][q]++] = i;
                }
            }
            if (isIB) {
                ca /= nCA;
                s->ibI[i] = (int32_t)c;
                for (int d = 0; d < s->D; ++d) s->ibNorm[c * 3 + d] = norm[d];
                s->ibSin[c] = sin(ca / 180.0 * pi);
                s->ibCos[c] = cos(ca / 180.0 * pi);
                c++;
            }
        }
    }
    s->nwall = c;
}

/* One wall site's frame rotation and its inverse from its integer normal
 * (ColorFluid::calcTibNorm, colorfluid.cpp:1111-1129). */
static void wall_frame(int D, const double* n, double* tib, double* tibInv)
{
    if (D == 3) {
        double tx[9], txInv[9], tz[9], tzInv[9];
        rotate_tx3d(n[0], n[1], tx);
        rotate_tx3d(-n[0], n[1], txInv);
        double val = sqrt(n[0] * n[0] + n[1] * n[1]);
        rotate_tz3d(val, n[2], tz);
        rotate_tz3d(-val, n[2], tzInv);
        mat_mul(tz, tx, 3, 3, 3, tib);
        mat_mul(txInv, tzInv, 3, 3, 3, tibInv);
    } else {
        rotate_t2d(n[0], n[1], tib);
        rotate_t2d(-n[0], n[1], tibInv);
    }
}

static void calc_tib(orc_sim* s)
{
    const int D = s->D;
    s->tib = calloc(s->nwall * D * D, sizeof(double));
    s->tibInv = calloc(s->nwall * D * D, sizeof(double));
    for (uint64_t c = 1; c < s->nwall; ++c)
        wall_frame(D, s->ibNorm + c * 3, s->tib + c * D * D, s->tibInv + c * D * D);
}

/* The geometric contact angle at a wall site (colorfluid.cpp:550-567): the
 * gradient g, of length norm, is replaced by the vector of the same length
 * that makes the contact angle (sin sA, cos cA) with the wall normal in the
 * plane of the normal and g.  3D: rotate g into the wall frame (normal along
 * +z), turn its tangential part onto +y (RotateTx3D), set
 * (0, norm sA, norm cA) and rotate back.  2D: the tangential sign of g in
 * the wall frame is kept. */
static void contact_angle(int D, const double* tib, const double* tibInv, double sA, double cA, double norm,
                          double* g)
{
    if (D == 3) {
        double gn[3], txInv[9], tInv[9], h[3];
        mat_mul(tib, g, 3, 3, 1, gn);
        rotate_tx3d(-gn[0], gn[1], txInv);
        mat_mul(tibInv, txInv, 3, 3, 3, tInv);
        h[0] = 0; h[1] = norm * sA; h[2] = norm * cA;
        mat_mul(tInv, h, 3, 3, 1, g);
    } else {
        double gn[2], h[2];
        mat_mul(tib, g, 2, 2, 1, gn);
        if (gn[0] >= 0) { h[0] = norm * sA; h[1] = norm * cA; }
        else { h[0] = -norm * sA; h[1] = norm * cA; }
        mat_mul(tibInv, h, 2, 2, 1, g);
    }
}

int orc_contact_angle(int D, const double* n, const double* g, double sA, double cA, double* out, int plant)
{
    if (D != 2 && D != 3) return -1;
    double tib[9], tibInv[9], norm2 = 0.0;
    wall_frame(D, n, tib, tibInv);
    for (int d = 0; d < D; ++d) { out[d] = g[d]; norm2 += g[d] * g[d]; }
    switch (plant) {   /* planted mistakes, for the test that must catch them */
    case 0: break;
    case 1: sA = -sA; break;                                   /* sign of the tangential part */
    case 2: memcpy(tibInv, tib, sizeof tib); break;            /* frame not inverted */
    case 3: { double t = sA; sA = cA; cA = t; } break;         /* sin and cos swapped */
    case 4: for (int d = 0; d < D; ++d) out[d] = -g[d]; break;  /* gradient reversed before the rotation */
    default: return -1;
    }
    contact_angle(D, tib, tibInv, sA, cA, sqrt(norm2), out);
    return 0;
}

/* ColorFluid::calcWallGradientInverse (colorfluid.cpp:917-957): per wall
 * site the inverse of sum over fluid links of w e e^T (Gauss-Jordan with
 * partial pivoting), for the renormalized wall gradient. */
static int calc_wall_gradient_inverse(orc_sim* s)
{
    const int D = s->D;
    s->ginv = calloc(s->nwall * D * D, sizeof(double));
    for (uint64_t i = 0; i < s->N; ++i) {
        if (!s->ibI[i]) continue;
        double aug[3][6] = {{0}};
        for (int q = 1; q < s->Q; ++q) {
            const int type = s->lab[s->raw[s->adj[i * s->Q + q]]].type;
            if (type == ORC_SD || type == ORC_BB) continue;
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) aug[a][b] += s->wi[q] * s->qi[q][a] * s->qi[q][b];
        }
        for (int a = 0; a < D; ++a) aug[a][D + a] = 1.0;
        for (int col = 0; col < D; ++col) {
            int pivot = col;
            for (int row = col + 1; row < D; ++row)
                if (fabs(aug[row][col]) > fabs(aug[pivot][col])) pivot = row;
            if (fabs(aug[pivot][col]) < 1e-12)
                return fail(s, "Rank-deficient fluid wall stencil; reconstructed derivatives are unsupported here.");
            for (int k = 0; k < 2 * D; ++k) { double t = aug[col][k]; aug[col][k] = aug[pivot][k]; aug[pivot][k] = t; }
            const double diagonal = aug[col][col];
            for (int k = 0; k < 2 * D; ++k) aug[col][k] /= diagonal;
            for (int row = 0; row < D; ++row) {
                if (row == col) continue;
                const double scale = aug[row][col];
                for (int k = 0; k < 2 * D; ++k) aug[row][k] -= scale * aug[col][k];
            }
        }
        double* inv = s->ginv + (uint64_t)s->ibI[i] * D * D;
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) inv[a * D + b] = aug[a][D + b];
    }
    return 0;
}

orc_sim* orc_new(const orc_desc* desc, const orc_label* labels, const uint8_t* geometry,
                 char* err, size_t errlen)
{
    orc_sim* s = calloc(1, sizeof *s);
    s->d = *desc;
    s->D = desc->D; s->Q = desc->Q;
    memcpy(s->lab, labels, sizeof(orc_label) * 256);
    if (!((s->D == 2 && s->Q == 9) || (s->D == 3 && s->Q == 19))) {
        snprintf(err, errlen, "Only D2Q9 and D3Q19 are validated for simulation.");
        free(s); return NULL;
    }
    setup_lattice(s);
    s->N = 1;
    for (int d = 0; d < s->D; ++d) { s->sz[d] = desc->sz[d]; s->N *= s->sz[d]; }
    s->stride[s->D - 1] = 1;
    for (int d = s->D - 1; d > 0; --d) s->stride[d - 1] = s->stride[d] * s->sz[d];
    s->raw = malloc(s->N);
    memcpy(s->raw, geometry, s->N);
    /* every label used by the geometry must be defined (simulator.cpp:30) */
    for (uint64_t i = 0; i < s->N; ++i)
        if (!s->lab[s->raw[i]].type) {
            snprintf(err, errlen, "map::at");
            free(s->raw); free(s); return NULL;
        }
    for (int l = 0; l < 256; ++l) {
        int t = s->lab[l].type;
        if (t == ORC_AB || (t == ORC_HF && desc->model == 0)) {
            snprintf(err, errlen, "oracle: label type %d is outside the restated scope", t);
            free(s->raw); free(s); return NULL;
        }
    }
    /* ROI::getRawIdx: sites per label, ascending */
    memset(s->start, 0, sizeof s->start);
    for (uint64_t i = 0; i < s->N; ++i) s->start[s->raw[i] + 1]++;
    for (int l = 0; l < 256; ++l) s->start[l + 1] += s->start[l];
    s->bylabel = malloc(sizeof(uint64_t) * s->N);
    {
        uint64_t fill[256];
        for (int l = 0; l < 256; ++l) fill[l] = s->start[l];
        for (uint64_t i = 0; i < s->N; ++i) s->bylabel[fill[s->raw[i]]++] = i;
    }
    s->adj = malloc(sizeof(uint64_t) * s->N * s->Q);
    for (uint64_t i = 0; i < s->N; ++i)
        for (int q = 0; q < s->Q; ++q) s->adj[i * s->Q + q] = calc_adj(s, i, q);
    s->ibI = calloc(s->N, sizeof(int32_t));
    build_ib(s);
    for (int ph = 0; ph < s->NP; ++ph) {
        s->f[ph] = calloc(s->N * s->Q, sizeof(double));
        s->tmp[ph] = calloc(s->N * s->Q, sizeof(double));
        s->dens[ph] = calloc(s->N, sizeof(double));
    }
    s->v = calloc(s->N * s->D, sizeof(double));
    s->densT = calloc(s->N, sizeof(double));
    s->adjV = calloc(s->N * s->D, sizeof(double));
    s->adjChi = calloc(s->N, sizeof(double));
    s->tau = calloc(s->N, sizeof(double));
    s->densN = malloc(sizeof(double) * s->N);
    for (uint64_t i = 0; i < s->N; ++i) s->densN[i] = SOLID_DENS_N;
    /* operators */
    if (s->d.model == 0) {
        orc_mrt_matrices(s->Q, s->d.tau[0], s->A, s->B, s->M, s->MInv);
        mat_transpose(s->A, s->Q, s->Q, s->AT);
    } else {
        /* colorfluid.cpp:315-322 */
        orc_mrt_matrices(s->Q, s->d.tau[0], s->A, s->B, s->M, s->MInv);
        s->uniformTau = s->d.tau[0] == s->d.tau[1];
        /* initPhi colorfluid.cpp:1131-1150 */
        double alpha[2];
        alpha[0] = s->d.alpha0;
        alpha[1] = 1 - s->d.density[0] / s->d.density[1] * (1 - alpha[0]);
        for (int ph = 0; ph < 2; ++ph) {
            if (s->Q == 9) {
                double a = (1 - alpha[ph]) / 5, b = (1 - alpha[ph]) / 20;
                double t[9] = {alpha[ph], a, a, a, a, b, b, b, b};
                memcpy(s->ai[ph], t, sizeof t);
            } else {
                double a = (1 - alpha[ph]) / 12, b = (1 - alpha[ph]) / 24;
                for (int q = 0; q < 19; ++q) s->ai[ph][q] = q == 0 ? alpha[ph] : (q <= 6 ? a : b);
            }
        }
        /* getBi colorfluid.cpp:1152-1168 */
        const double x = 2, c2 = 1.0;
        const double b0 = -(2.0 + 2.0 * x) / (3.0 * x + 12.0) * c2;
        const double b1 = x / (6.0 * x + 24.0) * c2;
        const double b7 = 1.0 / (6.0 * x + 24.0) * c2;
        if (s->Q == 9) {
            double t[9] = {-4.0/27, 2.0/27, 2.0/27, 2.0/27, 2.0/27, 5.0/108, 5.0/108, 5.0/108, 5.0/108};
            memcpy(s->bi, t, sizeof t);
        } else {
            for (int q = 0; q < 19; ++q) s->bi[q] = q == 0 ? b0 : (q <= 6 ? b1 : b7);
        }
        calc_tib(s);
        if (s->d.capillary) {
            for (int l = 0; l < 256; ++l)
                if (s->lab[l].type && !label_in(s, l, ORC_VD | ORC_GV | ORC_SD | ORC_BB)) {
                    snprintf(err, errlen, "oracle: CSF is restated for closed flow (VD/GV/SD/BB) only");
                    orc_free(s);
                    return NULL;
                }
            s->capGrad = calloc(s->N * s->D, sizeof(double));
            s->capField = calloc(s->N * s->D, sizeof(double));
            s->capForce = calloc(s->N * s->D, sizeof(double));
        }
        /* calcWallGradientInverse runs for renormalized and tanh walls and for CSF
         * (colorfluid.cpp:917-920) */
        if ((s->d.wall_gradient == 1 || s->d.wall_gradient == 2 || s->d.capillary) &&
            calc_wall_gradient_inverse(s)) {
            snprintf(err, errlen, "%s", s->err);
            orc_free(s);
            return NULL;
        }
    }
    return s;
}

void orc_free(orc_sim* s)
{
    if (!s) return;
    free(s->raw); free(s->bylabel); free(s->adj); free(s->ibI);
    for (int q = 0; q < MAXQ; ++q) { free(s->ibs[q]); free(s->ibf[0][q]); free(s->ibf[1][q]); }
    free(s->ibNorm); free(s->ibSin); free(s->ibCos); free(s->tib); free(s->tibInv); free(s->ginv);
    free(s->capGrad); free(s->capField); free(s->capForce);
    for (int ph = 0; ph < 2; ++ph) { free(s->f[ph]); free(s->tmp[ph]); free(s->dens[ph]); }
    free(s->v); free(s->densT); free(s->densN); free(s->tau); free(s->adjV); free(s->adjChi);
    free(s);
}

/* ---------------------------------------------------------------- scsp */
/* scsp.cpp:116-134 */
static void scsp_initialize(orc_sim* s)
{
    const double rho = s->d.density[0];
    double iDist[MAXQ];
    for (int q = 0; q < s->Q; ++q) iDist[q] = s->wi[q] * rho;
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_VD | ORC_GV | ORC_VB | ORC_PB)) continue;
        double force[3];
        force_for_label(s, l, force);
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            uint64_t i = s->bylabel[k];
            for (int q = 0; q < s->Q; ++q) {
                F(s, 0, q, i) = iDist[q];
                double ef = 0.0;
                for (int d = 0; d < s->D; ++d) ef += s->qi[q][d] * force[d];
                F(s, 0, q, i) -= 1.5 * s->wi[q] * ef;
            }
            s->dens[0][i] = rho;
        }
    }
}

/* scsp.cpp:136-160 (ISCOMPRESSABLE 0, lbcfg.h:12) */
static int scsp_macro(orc_sim* s)
{
    const double dens0 = s->d.density[0];
    int bad = 0;
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_VD | ORC_GV)) continue;
        /* sites are independent: threads change no result, only which sites
         * are updated before a failure is reported */
#pragma omp parallel for schedule(static) reduction(| : bad)
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            uint64_t i = s->bylabel[k];
            double r = F(s, 0, 0, i);
            for (int q = 1; q < s->Q; ++q) r += F(s, 0, q, i);
            s->dens[0][i] = r;
            if (!isfinite(r) || r <= 0.0) bad = 1;
            for (int d = 0; d < s->D; ++d) {
                double u = 0;
                for (int q = 1; q < s->Q; ++q) {
                    if (s->qi[q][d] == 1) u += F(s, 0, q, i);
                    else if (s->qi[q][d] == -1) u -= F(s, 0, q, i);
                }
                u /= dens0;
                V(s, d, i) = u;
            }
        }
        if (bad) return fail(s, "Non-finite or non-positive fluid density.");
    }
    return 0;
}

/* scsp.cpp:162-175 */
static void scsp_body_force(orc_sim* s)
{
    const double rhoRef = s->d.density[0];
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_GV)) continue;
        double force[3], fi[3];
        force_for_label(s, l, force);
        for (int d = 0; d < s->D; ++d) fi[d] = force[d] * 0.5;
        for (int d = 0; d < s->D; ++d)
#pragma omp parallel for schedule(static)
            for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k)
                V(s, d, s->bylabel[k]) += fi[d] / rhoRef;
    }
}

/* Zou-He pressure (PB) and velocity (VB) boundaries, SCSP::zouHeBoundary
 * (scsp.cpp:188-256): density/velocity of the boundary site from the known
 * populations, then the unknown populations along the boundary direction.
 * wiT = w * 6 / latticeSpeed (simulator.cpp:28) with latticeSpeed 1. */
static int zh_rel(const orc_sim* s, int dir, int k) { return s->Q == 9 ? REL9[dir][k] : REL19[dir][k]; }
static int zh_perp(const orc_sim* s, int dir, int k) { return s->Q == 9 ? PERP9[dir][k] : PERP19[dir][k]; }

static void scsp_zouhe(orc_sim* s)
{
    const int nrel = s->Q == 9 ? 3 : 5, nperp = s->Q == 9 ? 3 : 9;
    for (int pass = 0; pass < 2; ++pass)      /* VB labels first, then PB */
        for (int l = 0; l < 256; ++l) {
            if (!label_in(s, l, pass == 0 ? ORC_VB : ORC_PB)) continue;
            const int dir = s->lab[l].dir;
            const int axis = s->Q == 9 ? AXIS9[dir] : AXIS19[dir];
            const int sign = s->Q == 9 ? SIGN9[dir] : SIGN19[dir];
            const int opdir = s->op[dir];
            double velocity = 0.0, vel_1 = 0.0;
            const double density = s->lab[l].value;
            if (pass == 0) {
                velocity = sign == 1 ? s->lab[l].value : -s->lab[l].value;
                vel_1 = 1.0 / (1.0 - velocity);
            }
            for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
                const uint64_t i = s->bylabel[k];
                double sumF = 0;
                for (int r = 0; r < nrel; ++r) { int q = zh_rel(s, opdir, r); sumF += F(s, 0, q, i) + F(s, 0, q, i); }
                for (int r = 0; r < nperp; ++r) sumF += F(s, 0, zh_perp(s, dir, r), i);
                if (pass == 0) {
                    for (int d = 0; d < s->D; ++d)
                        if (d != axis) V(s, d, i) = 0;
                        else { V(s, d, i) = velocity; s->dens[0][i] = sumF * vel_1; }
                } else {
                    s->dens[0][i] = density;
                    for (int d = 0; d < s->D; ++d)
                        if (d != axis) V(s, d, i) = 0;
                        else {
                            sumF /= density; sumF--;
                            V(s, d, i) = sign == 1 ? -sumF : sumF;
                        }
                }
            }
        }
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_PB | ORC_VB)) continue;
        const int dir = s->lab[l].dir;
        const int axis = s->Q == 9 ? AXIS9[dir] : AXIS19[dir];
        const int sign = s->Q == 9 ? SIGN9[dir] : SIGN19[dir];
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            const uint64_t i = s->bylabel[k];
            for (int r = 0; r < nrel; ++r) {
                const int q = zh_rel(s, dir, r);
                F(s, 0, q, i) = F(s, 0, s->op[q], i);
                F(s, 0, q, i) += sign * (s->wi[q] * 6.0 / 1.0) * s->dens[0][i] * V(s, axis, i);
                double sumF = 0;
                for (int n = 0; n < nperp; ++n) {
                    const int qn = zh_perp(s, dir, n);
                    for (int d = 0; d < s->D; ++d) {
                        if (s->qi[qn][d] == 0 || s->qi[q][d] == 0) continue;
                        else if (s->qi[qn][d] == s->qi[q][d]) sumF += F(s, 0, qn, i);
                        else sumF -= F(s, 0, qn, i);
                    }
                }
                F(s, 0, q, i) -= sumF * 0.5;
            }
        }
    }
}

/* simulator.cpp:327-333 */
static int validate(orc_sim* s, int ph)
{
    const uint64_t n = s->N * s->Q;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (uint64_t k = 0; k < n; ++k)
        if (!isfinite(s->f[ph][k])) bad = 1;
    if (bad) return fail(s, "Non-finite post-collision population; no restart will be written.");
    return 0;
}

/* scsp.cpp:374-464 */
static int scsp_collision(orc_sim* s)
{
    const int Q = s->Q;
    const double c1 = 3.0, c2 = 4.5, c3 = 1.5;        /* c = latticeSpeed = 1 */
    const double omega = 1.0 / s->d.tau[0];
    const double dens0 = s->d.density[0];
    const int mrt = s->d.equil == 1;
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_VD | ORC_GV | ORC_VB | ORC_PB)) continue;
        double force[3];
        force_for_label(s, l, force);
        double ff2 = 0.0;
        for (int d = 0; d < s->D; ++d) ff2 = ff2 + force[d] * force[d];
        const int forced = ff2 != 0.0;
#pragma omp parallel for schedule(static)
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            uint64_t i = s->bylabel[k];
            double usq = 0, src[MAXQ];
            for (int d = 0; d < s->D; ++d) usq += V(s, d, i) * V(s, d, i);
            if (forced) force_source(s, i, force, src);
            else for (int q = 0; q < Q; ++q) src[q] = 0.0;
            if (!mrt) {
                for (int q = 0; q < Q; ++q) {
                    double ui = 0;
                    for (int d = 0; d < s->D; ++d) ui += V(s, d, i) * s->qi[q][d];
                    double feq = s->wi[q] * (s->dens[0][i] + dens0 * (c1 * ui + c2 * ui * ui - c3 * usq));
                    F(s, 0, q, i) += (feq - F(s, 0, q, i)) * omega;
                    if (forced) F(s, 0, q, i) += (1.0 - 0.5 * omega) * src[q];
                }
            } else {
                double ff[MAXQ];
                for (int q = 0; q < Q; ++q) ff[q] = 0.0;
                const double* im = s->AT;
                for (int q = 0; q < Q; ++q) {
                    double ui = 0;
                    for (int d = 0; d < s->D; ++d) ui += V(s, d, i) * s->qi[q][d];
                    double feq = s->wi[q] * (s->dens[0][i] + dens0 * (c1 * ui + c2 * ui * ui - c3 * usq));
                    for (int qq = 0; qq < Q; ++qq) ff[qq] += *im++ * (feq - F(s, 0, q, i));
                }
                for (int q = 0; q < Q; ++q) F(s, 0, q, i) += ff[q];
                if (forced) {
                    double corr[MAXQ];
                    mat_vec(s->B, src, Q, Q, corr);
                    for (int q = 0; q < Q; ++q) F(s, 0, q, i) += corr[q];
                }
            }
        }
    }
    return validate(s, 0);
}

/* --------------------------------------------------------------- color */
/* colorfluid.cpp:214-236 */
static void color_initialize(orc_sim* s)
{
    const double rho[2] = {s->d.density[0], s->d.density[1]};
    double dist[2][MAXQ];
    for (int ph = 0; ph < 2; ++ph)
        for (int q = 0; q < s->Q; ++q) dist[ph][q] = s->ai[ph][q] * rho[ph];
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_VD | ORC_GV | ORC_PB | ORC_VB | ORC_AB | ORC_HF)) continue;
        const int phs = s->lab[l].phase;
        double force[3];
        force_for_label(s, l, force);
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            uint64_t i = s->bylabel[k];
            for (int q = 0; q < s->Q; ++q) {
                F(s, phs, q, i) = dist[phs][q];
                double ef = 0.0;
                for (int d = 0; d < s->D; ++d) ef += s->qi[q][d] * force[d];
                F(s, phs, q, i) -= 1.5 * s->wi[q] * ef;
            }
            s->densT[i] = s->dens[phs][i] = rho[phs];
            s->densN[i] = phs == 0 ? 1 : -1;
        }
    }
}

/* ColorFluid::zouHeBoundary (colorfluid.cpp:638-734): the Zou-He boundary on
 * the total population fT = f0 + f1; the unknowns go to the label's phase,
 * the other phase gets 0.  densT and v are set here but calcMacroVar
 * recomputes them at every PB/VB site right after; only the unknown update
 * and the adjacent report history (colorfluid.cpp:166-171) see them. */
static void color_zouhe(orc_sim* s)
{
    const int nrel = s->Q == 9 ? 3 : 5, nperp = s->Q == 9 ? 3 : 9;
    for (int pass = 0; pass < 2; ++pass)      /* VB labels first, then PB */
        for (int l = 0; l < 256; ++l) {
            if (!label_in(s, l, pass == 0 ? ORC_VB : ORC_PB)) continue;
            const int dir = s->lab[l].dir;
            const int axis = s->Q == 9 ? AXIS9[dir] : AXIS19[dir];
            const int sign = s->Q == 9 ? SIGN9[dir] : SIGN19[dir];
            const int opdir = s->op[dir];
            double velocity = 0.0, vel_1 = 0.0;
            const double density = s->lab[l].value;
            if (pass == 0) {
                velocity = sign == 1 ? s->lab[l].value : -s->lab[l].value;
                vel_1 = 1.0 / (1.0 - velocity);
            }
            for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
                const uint64_t i = s->bylabel[k];
                double sumF = 0;
                for (int r = 0; r < nrel; ++r) {
                    const int q = zh_rel(s, opdir, r);
                    sumF += F(s, 0, q, i) + F(s, 0, q, i) + F(s, 1, q, i) + F(s, 1, q, i);
                }
                for (int r = 0; r < nperp; ++r) {
                    const int q = zh_perp(s, dir, r);
                    sumF += F(s, 0, q, i) + F(s, 1, q, i);
                }
                if (pass == 0) {
                    for (int d = 0; d < s->D; ++d)
                        if (d != axis) V(s, d, i) = 0;
                        else { V(s, d, i) = velocity; s->densT[i] = sumF * vel_1; }
                } else {
                    s->densT[i] = density;
                    for (int d = 0; d < s->D; ++d)
                        if (d != axis) V(s, d, i) = 0;
                        else {
                            sumF /= density; sumF--;
                            V(s, d, i) = sign == 1 ? -sumF : sumF;
                        }
                }
            }
        }
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_PB | ORC_VB)) continue;
        const int dir = s->lab[l].dir;
        const int axis = s->Q == 9 ? AXIS9[dir] : AXIS19[dir];
        const int sign = s->Q == 9 ? SIGN9[dir] : SIGN19[dir];
        const int phs = s->lab[l].phase, phsOp = 1 - phs;
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            const uint64_t i = s->bylabel[k];
            for (int r = 0; r < nrel; ++r) {
                const int q = zh_rel(s, dir, r);
                double fT = F(s, 0, s->op[q], i) + F(s, 1, s->op[q], i);
                fT += sign * (s->wi[q] * 6.0 / 1.0) * s->densT[i] * V(s, axis, i);
                double sumF = 0;
                for (int n = 0; n < nperp; ++n) {
                    const int qn = zh_perp(s, dir, n);
                    for (int d = 0; d < s->D; ++d) {
                        if (s->qi[qn][d] == 0 || s->qi[q][d] == 0) continue;
                        else if (s->qi[qn][d] == s->qi[q][d]) sumF += F(s, 0, qn, i) + F(s, 1, qn, i);
                        else sumF -= F(s, 0, qn, i) + F(s, 1, qn, i);
                    }
                }
                fT -= sumF * 0.5;
                F(s, phs, q, i) = fT;
                F(s, phsOp, q, i) = 0;
            }
        }
    }
}

/* Adjacent report history, copied after the boundaries and before the macro
 * variables of each step (scsp.cpp:82-83, colorfluid.cpp:166-171). */
static void record_adjacent(orc_sim* s)
{
    memcpy(s->adjV, s->v, sizeof(double) * s->N * s->D);
    if (s->NP == 2)
        for (uint64_t i = 0; i < s->N; ++i)
            s->adjChi[i] = s->densT[i] > 0.0 ? (s->dens[0][i] - s->dens[1][i]) / s->densT[i] : 0.0;
}

/* colorfluid.cpp:238-290 (arithmetic and harmonic viscosity) */
static int color_macro(orc_sim* s)
{
    const double relax[2] = {s->d.tau[0] / 2.0, s->d.tau[1] / 2.0};
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_VD | ORC_GV | ORC_PB | ORC_VB | ORC_HF)) continue;
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            uint64_t i = s->bylabel[k];
            double r0 = F(s, 0, 0, i), r1 = F(s, 1, 0, i);
            for (int q = 1; q < s->Q; ++q) { r0 += F(s, 0, q, i); r1 += F(s, 1, q, i); }
            s->dens[0][i] = r0; s->dens[1][i] = r1;
            const double rt = r0 + r1;
            s->densT[i] = rt;
            if (!isfinite(r0) || !isfinite(r1) || !isfinite(rt) || rt <= 0.0)
                return fail(s, "Non-finite or non-positive total fluid density.");
            if (r0 < -1e-10 * rt || r1 < -1e-10 * rt)
                return fail(s, "Negative component density outside roundoff tolerance.");
            const double dn = (r0 - r1) / rt;
            s->densN[i] = dn;
            double tau = (1 + dn) * relax[0] + (1 - dn) * relax[1];
            if (s->d.harmonic) {
                const double phi = r0 / rt;
                const double nu0 = (s->d.tau[0] - 0.5) / 3.0;
                const double nu1 = (s->d.tau[1] - 0.5) / 3.0;
                tau = 0.5 + 3.0 / (phi / nu0 + (1.0 - phi) / nu1);
            }
            s->tau[i] = tau;
            if (!isfinite(tau) || tau <= 0.5) return fail(s, "Invalid local relaxation time.");
            for (int d = 0; d < s->D; ++d) {
                double u = 0;
                for (int q = 1; q < s->Q; ++q) {
                    if (s->qi[q][d] == 1) u += F(s, 0, q, i) + F(s, 1, q, i);
                    else if (s->qi[q][d] == -1) u -= F(s, 0, q, i) + F(s, 1, q, i);
                }
                u /= rt;
                V(s, d, i) = u;
                if (!isfinite(u)) return fail(s, "Non-finite fluid velocity.");
            }
        }
    }
    return 0;
}

/* colorfluid.cpp:736-750 (no CSF) */
static int color_body_force(orc_sim* s)
{
    if (s->d.capillary) {   /* colorfluid.cpp:738-743 */
        if (calc_capillary_force(s)) return 1;
        for (int l = 0; l < 256; ++l) {
            if (!label_in(s, l, ORC_VD | ORC_GV)) continue;
            for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
                const uint64_t i = s->bylabel[k];
                for (int d = 0; d < s->D; ++d) V(s, d, i) += .5 * s->capForce[d * s->N + i] / s->densT[i];
            }
        }
    }
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_GV)) continue;
        double force[3], fi[3];
        force_for_label(s, l, force);
        for (int d = 0; d < s->D; ++d) fi[d] = force[d] * 0.5;
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            uint64_t i = s->bylabel[k];
            for (int d = 0; d < s->D; ++d) V(s, d, i) += fi[d] / s->densT[i];
        }
    }
    return 0;
}

/* ColorFluid::colorGradient, legacy wall branch, colorfluid.cpp:435-570 */
/* atanh of the colour clamped to (-1, 1) by 1e-8 (colorfluid.cpp:497-499) */
static double tanh_wall_transform(double color)
{
    const double hi = 1.0 - 1e-8, lo = -1.0 + 1e-8;
    const double c = color < hi ? color : hi;   /* std::min(1 - 1e-8, color) */
    return atanh(lo < c ? c : lo);              /* std::max(-1 + 1e-8, .) */
}

static void color_gradient(const orc_sim* s, uint64_t i, double* g, double* normOut)
{
    const int D = s->D;
    double norm = 0;
    for (int d = 0; d < D; ++d) g[d] = 0.0;
    const int nrel = s->Q == 9 ? 6 : 10;
    if (!s->ibI[i]) {
        for (int d = 0; d < D; ++d)
            for (int k = 0; k < nrel; ++k) {
                const int q = s->Q == 9 ? RELD9[d][k] : RELD19[d][k];
                const double dn = s->densN[s->adj[i * s->Q + q]];
                if (s->qi[q][d] == 1) g[d] += s->wi[q] * dn;
                else g[d] -= s->wi[q] * dn;
            }
        for (int d = 0; d < D; ++d) g[d] *= 3.0;
        double n2 = 0.0;
        for (int d = 0; d < D; ++d) n2 = n2 + g[d] * g[d];
        norm = sqrt(n2);
    } else {
        for (int d = 0; d < D; ++d)
            for (int k = 0; k < nrel; ++k) {
                const int q = s->Q == 9 ? RELD9[d][k] : RELD19[d][k];
                const double dn0 = s->densN[i];
                const double dnq = s->densN[s->adj[i * s->Q + q]];
                if (dnq == SOLID_DENS_N) continue;
                if (s->qi[q][d] == 1) g[d] += s->wi[q] * (dnq - dn0);
                else g[d] -= s->wi[q] * (dnq - dn0);
            }
        for (int d = 0; d < D; ++d) g[d] *= 3.0;
        double n2 = 0.0;
        for (int d = 0; d < D; ++d) n2 = n2 + g[d] * g[d];
        norm = sqrt(n2);
        const int c = s->ibI[i];
        const double cA = s->ibCos[c], sA = s->ibSin[c];
        if (s->d.wall_gradient == 1) {
            /* colorfluid.cpp:487-492: g = MatMulVec(inverse, g) / 3 (fn.cpp:29-39) */
            const double* inv = s->ginv + (uint64_t)c * D * D;
            double r[3];
            for (int a = 0; a < D; ++a) {
                double t = 0.0;
                for (int b = 0; b < D; ++b) t = t + g[b] * inv[a * D + b];
                r[a] = t;
            }
            for (int a = 0; a < D; ++a) g[a] = r[a] / 3.0;
            n2 = 0.0;
            for (int d = 0; d < D; ++d) n2 = n2 + g[d] * g[d];
            norm = sqrt(n2);
        }
        if (s->d.wall_gradient == 2) {
            /* colorfluid.cpp:494-518: fit atanh(chi) as affine over the fluid
             * neighbours (slope = inverse moment), continue tanh over the full
             * stencil and take the bulk gradient of that */
            const double* inv = s->ginv + (uint64_t)c * D * D;
            const double center = tanh_wall_transform(s->densN[i]);
            double moment[3] = {0.0, 0.0, 0.0}, slope[3];
            for (int q = 1; q < s->Q; ++q) {
                const double nb = s->densN[s->adj[i * s->Q + q]];
                if (nb == SOLID_DENS_N) continue;
                const double difference = tanh_wall_transform(nb) - center;
                for (int d = 0; d < D; ++d) moment[d] += s->wi[q] * s->qi[q][d] * difference;
            }
            for (int a = 0; a < D; ++a) {
                double t = 0.0;
                for (int b = 0; b < D; ++b) t = t + moment[b] * inv[a * D + b];
                slope[a] = t;
            }
            for (int d = 0; d < D; ++d) g[d] = 0.0;
            for (int q = 1; q < s->Q; ++q) {
                double position = center;
                for (int d = 0; d < D; ++d) position += s->qi[q][d] * slope[d];
                const double color = tanh(position);
                for (int d = 0; d < D; ++d) g[d] += 3.0 * s->wi[q] * s->qi[q][d] * color;
            }
            n2 = 0.0;
            for (int d = 0; d < D; ++d) n2 = n2 + g[d] * g[d];
            norm = sqrt(n2);
        }
        contact_angle(D, s->tib + c * D * D, s->tibInv + c * D * D, sA, cA, norm, g);
    }
    *normOut = norm;
}

void orc_color_gradient(const orc_sim* s, uint64_t site, double* g, double* norm)
{
    color_gradient(s, site, g, norm);
}

/* ColorFluid::collision, colorfluid.cpp:292-433 (perturbation, cosine recoloring) */
/* ColorFluid::calcCapillaryForce, colorfluid.cpp:572-617: on VD/GV sites the
 * colour gradient (with the wall treatment) and the curvature field (unit
 * normal, or the gradient itself for curvatureStencil "gradient"); then
 * kappa = sum_ab (n_a n_b - delta_ab) d_b field_a from the fluid neighbours
 * (inverse moment at wall sites), and F = sigma/2 kappa direction. */
static int calc_capillary_force(orc_sim* s)
{
    const int D = s->D, Q = s->Q;
    const uint64_t N = s->N;
    const int gc = s->d.curvature == 1;
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_VD | ORC_GV)) continue;
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            const uint64_t i = s->bylabel[k];
            double g[3], norm;
            color_gradient(s, i, g, &norm);
            for (int d = 0; d < D; ++d) {
                s->capGrad[d * N + i] = g[d];
                s->capField[d * N + i] = gc ? g[d] : (norm > 1e-9 ? g[d] / norm : 0.0);
            }
        }
    }
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_VD | ORC_GV)) continue;
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            const uint64_t i = s->bylabel[k];
            double curvature = 0.0, norm2 = 0.0;
            for (int d = 0; d < D; ++d) norm2 += s->capGrad[d * N + i] * s->capGrad[d * N + i];
            if (norm2 > 1e-18) {
                for (int a = 0; a < D; ++a) {
                    double moment[3] = {0.0, 0.0, 0.0}, derivative[3];
                    const double center = s->capField[a * N + i];
                    for (int q = 1; q < Q; ++q) {
                        const uint64_t j = s->adj[i * Q + q];
                        if (s->densN[j] == SOLID_DENS_N) continue;
                        const double difference = s->capField[a * N + j] - center;
                        for (int b = 0; b < D; ++b) moment[b] += s->wi[q] * s->qi[q][b] * difference;
                    }
                    if (s->ibI[i]) {
                        const double* inv = s->ginv + (uint64_t)s->ibI[i] * D * D;
                        for (int r = 0; r < D; ++r) {
                            double t = 0.0;
                            for (int b = 0; b < D; ++b) t = t + moment[b] * inv[r * D + b];
                            derivative[r] = t;
                        }
                    } else {
                        for (int b = 0; b < D; ++b) derivative[b] = moment[b] * 3.0;
                    }
                    for (int b = 0; b < D; ++b) {
                        const double na = gc ? s->capGrad[a * N + i] / sqrt(norm2) : center;
                        const double nb = gc ? s->capGrad[b * N + i] / sqrt(norm2) : s->capField[b * N + i];
                        curvature += (na * nb - (a == b ? 1.0 : 0.0)) * derivative[b];
                    }
                }
            }
            for (int d = 0; d < D; ++d) {
                const double direction = gc && norm2 > 1e-18 ? s->capGrad[d * N + i] / sqrt(norm2)
                                                             : s->capGrad[d * N + i];
                const double force = .5 * s->d.sigma * curvature * direction;
                s->capForce[d * N + i] = force;
                if (!isfinite(force)) return fail(s, "Non-finite capillary force.");
            }
        }
    }
    return 0;
}

/* ColorFluid::initializeCapillaryForce, colorfluid.cpp:619-634 */
static int initialize_capillary_force(orc_sim* s)
{
    const int D = s->D, Q = s->Q;
    const uint64_t N = s->N;
    if (calc_capillary_force(s)) return 1;
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_VD | ORC_GV)) continue;
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            const uint64_t i = s->bylabel[k];
            for (int q = 0; q < Q; ++q) {
                double ef = 0.0;
                for (int d = 0; d < D; ++d) ef += s->qi[q][d] * s->capForce[d * N + i];
                for (int ph = 0; ph < s->NP; ++ph)
                    F(s, ph, q, i) -= 1.5 * s->wi[q] * ef * s->dens[ph][i] / s->densT[i];
            }
        }
    }
    return color_macro(s);
}

static int color_collision(orc_sim* s)
{
    const int Q = s->Q, D = s->D;
    const double c1 = 3.0, c2 = 4.5, c3 = 1.5;
    const double ak = 2.25 * s->d.sigma;
    const double beta = s->d.beta;
    const int mrt = s->d.equil == 1;
    for (int l = 0; l < 256; ++l) {
        if (!label_in(s, l, ORC_VD | ORC_GV | ORC_VB | ORC_PB | ORC_HF)) continue;
        const int csf = s->d.capillary;
        double external[3];
        force_for_label(s, l, external);
        double ff2 = 0.0;
        for (int d = 0; d < D; ++d) ff2 = ff2 + external[d] * external[d];
        const int externally_forced = ff2 != 0.0;
        for (uint64_t k = s->start[l]; k < s->start[l + 1]; ++k) {
            const uint64_t i = s->bylabel[k];
            /* CSF: the site's force is the label's plus the capillary force
             * (colorfluid.cpp:326-333) */
            double force[3] = {external[0], external[1], external[2]};
            int forced = externally_forced;
            if (csf) {
                for (int d = 0; d < D; ++d) force[d] += s->capForce[d * s->N + i];
                double f2 = 0.0;
                for (int d = 0; d < D; ++d) f2 = f2 + force[d] * force[d];
                forced = f2 != 0.0;
            }
            double usq = 0;
            for (int d = 0; d < D; ++d) usq += V(s, d, i) * V(s, d, i);
            double feq_u0[MAXQ], fT[MAXQ], src[MAXQ];
            if (forced) force_source(s, i, force, src);
            else for (int q = 0; q < Q; ++q) src[q] = 0.0;
            const double tau = s->tau[i];
            if (!mrt) {
                for (int q = 0; q < Q; ++q) {
                    fT[q] = F(s, 0, q, i) + F(s, 1, q, i);
                    double ui = 0;
                    for (int d = 0; d < D; ++d) ui += V(s, d, i) * s->qi[q][d];
                    feq_u0[q] = s->dens[0][i] * s->ai[0][q] + s->dens[1][i] * s->ai[1][q];
                    double feq = feq_u0[q] + s->densT[i] * s->wi[q] * (c1 * ui + c2 * ui * ui - c3 * usq);
                    fT[q] += (feq - fT[q]) / tau;
                    if (forced) fT[q] += (1.0 - 0.5 / tau) * src[q];
                }
            } else {
                double res[MAXQ], corr[MAXQ];
                for (int q = 0; q < Q; ++q) {
                    fT[q] = F(s, 0, q, i) + F(s, 1, q, i);
                    double ui = 0;
                    for (int d = 0; d < D; ++d) ui += V(s, d, i) * s->qi[q][d];
                    feq_u0[q] = s->dens[0][i] * s->ai[0][q] + s->dens[1][i] * s->ai[1][q];
                    double feq = feq_u0[q] + s->densT[i] * s->wi[q] * (c1 * ui + c2 * ui * ui - c3 * usq);
                    res[q] = feq - fT[q];
                }
                if (s->uniformTau) {
                    mat_vec(s->A, res, Q, Q, corr);
                    if (forced) {
                        double c2v[MAXQ];
                        mat_vec(s->B, src, Q, Q, c2v);
                        for (int q = 0; q < Q; ++q) corr[q] += c2v[q];
                    }
                } else {
                    double rates[MAXQ], m[MAXQ];
                    if (!isfinite(tau) || tau <= 0.5)
                        return fail(s, "Relaxation time must be finite and greater than 0.5.");
                    mrt_rates(Q, tau, rates);
                    mat_vec(s->M, res, Q, Q, m);
                    for (int q = 0; q < Q; ++q) m[q] = m[q] * rates[q];
                    if (forced) {
                        double mg[MAXQ];
                        mat_vec(s->M, src, Q, Q, mg);
                        for (int q = 0; q < Q; ++q) m[q] += mg[q] * (1.0 - 0.5 * rates[q]);
                    }
                    mat_vec(s->MInv, m, Q, Q, corr);
                }
                for (int q = 0; q < Q; ++q) fT[q] += corr[q];
            }
            double g[3], norm;
            if (csf) {   /* the stored capillary gradient, no perturbation (colorfluid.cpp:389-411) */
                double n2 = 0.0;
                for (int d = 0; d < D; ++d) {
                    g[d] = s->capGrad[d * s->N + i];
                    n2 = n2 + g[d] * g[d];
                }
                norm = sqrt(n2);
            } else {
                color_gradient(s, i, g, &norm);
            }
            const double dens0T = s->dens[0][i] / s->densT[i];
            if (norm > 1e-9) {
                const double b01 = beta * dens0T * s->dens[1][i] / s->densT[i];
                const double akn = ak * norm / tau;
                for (int q = 0; q < Q; ++q) {
                    double angle = 0.0;
                    for (int d = 0; d < D; ++d) {
                        if (s->qi[q][d] == 1) angle += g[d];
                        else if (s->qi[q][d] == -1) angle -= g[d];
                    }
                    angle /= norm;
                    if (!csf) fT[q] += akn * (s->wi[q] * angle * angle - s->bi[q]);
                    if (q != 0) angle /= s->qnorm[q];
                    F(s, 0, q, i) = dens0T * fT[q] + b01 * angle * feq_u0[q];
                    F(s, 1, q, i) = fT[q] - F(s, 0, q, i);
                }
            } else {
                for (int q = 0; q < Q; ++q) {
                    F(s, 0, q, i) = dens0T * fT[q];
                    F(s, 1, q, i) = fT[q] - F(s, 0, q, i);
                }
            }
        }
    }
    if (validate(s, 0)) return 1;
    return validate(s, 1);
}

/* ---------------------------------------------------------------- steps */
/* Simulator::streaming (push with periodic wrap) + bounceBack, simulator.cpp:73-111 */
static void stream_and_bounce(orc_sim* s, int ph)
{
    const uint64_t N = s->N;
    for (int q = 1; q < s->Q; ++q) {
        double* src = s->f[ph] + (uint64_t)q * N;
        double* dst = s->tmp[ph] + (uint64_t)q * N;
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < N; ++i) dst[s->adj[i * s->Q + q]] = src[i];
    }
    for (int q = 1; q < s->Q; ++q) {   /* swap f[q] <-> fOld[q] */
        double* a = s->f[ph] + (uint64_t)q * N;
        double* b = s->tmp[ph] + (uint64_t)q * N;
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < N; ++i) { double t = a[i]; a[i] = b[i]; b[i] = t; }
    }
    for (int q = 1; q < s->Q; ++q) {
        const int opq = s->op[q];
#pragma omp parallel for schedule(static)
        for (uint64_t k = 0; k < s->nibs[q]; ++k) {
            const uint64_t i = s->ibs[q][k];
            s->f[ph][(uint64_t)opq * N + i] = s->tmp[ph][(uint64_t)q * N + i];
        }
    }
}

/* Half-permeable film (ColorFluid::Run, colorfluid.cpp:156-159): each phase
 * bounces back at links into a film site of the other phase (bounceBack with
 * ibF, simulator.cpp:103-111), then clearHalfFilm (colorfluid.cpp:752-764)
 * empties the other phase at every film site. */
static void film_bounce_and_clear(orc_sim* s)
{
    const uint64_t N = s->N;
    for (int ph = 0; ph < 2; ++ph) {
        const int other = 1 - ph;
        for (int q = 1; q < s->Q; ++q) {
            const int opq = s->op[q];
            for (uint64_t k = 0; k < s->nibf[other][q]; ++k) {
                const uint64_t i = s->ibf[other][q][k];
                s->f[ph][(uint64_t)opq * N + i] = s->tmp[ph][(uint64_t)q * N + i];
            }
        }
    }
    for (int label = 0; label < 256; ++label) {
        if (!label_in(s, label, ORC_HF)) continue;
        const int clear = s->lab[label].phase == 0 ? 1 : 0;
        for (uint64_t k = s->start[label]; k < s->start[label + 1]; ++k) {
            const uint64_t i = s->bylabel[k];
            for (int q = 0; q < s->Q; ++q) s->f[clear][(uint64_t)q * N + i] = 0.0;
        }
    }
}

/* The streaming half of a time step: streaming, bounce-back, the film and
 * the Zou-He boundaries, leaving the pre-collision populations in place. */
static void stream_half(orc_sim* s)
{
    for (int ph = 0; ph < s->NP; ++ph) stream_and_bounce(s, ph);
    if (s->d.model == 0) {
        scsp_zouhe(s);
        return;
    }
    film_bounce_and_clear(s);
    color_zouhe(s);
}

void orc_stream_only(orc_sim* s) { stream_half(s); }

int orc_macro_only(orc_sim* s) { return s->d.model == 0 ? scsp_macro(s) : color_macro(s); }

int orc_init(orc_sim* s)
{
    s->t = 0;
    if (s->d.model == 0) {
        scsp_initialize(s);
        if (scsp_macro(s)) return 1;
        scsp_body_force(s);
        return scsp_collision(s);
    }
    color_initialize(s);
    if (color_macro(s)) return 1;
    if (s->d.capillary && initialize_capillary_force(s)) return 1;
    if (color_body_force(s)) return 1;
    return color_collision(s);
}

int orc_step(orc_sim* s)
{
    s->t++;
    stream_half(s);
    if (s->d.model == 0) {
        record_adjacent(s);
        if (scsp_macro(s)) return 1;
        scsp_body_force(s);
        return scsp_collision(s);
    }
    record_adjacent(s);
    if (color_macro(s)) return 1;
    if (color_body_force(s)) return 1;
    return color_collision(s);
}

int orc_run(orc_sim* s, uint64_t n)
{
    for (uint64_t k = 0; k < n; ++k)
        if (orc_step(s)) return 1;
    return 0;
}

const char* orc_error(const orc_sim* s) { return s->err; }
uint64_t orc_time(const orc_sim* s) { return s->t; }
uint64_t orc_sites(const orc_sim* s) { return s->N; }

void orc_fields(const orc_sim* s, double* rho0, double* rho1, double* u)
{
    if (rho0) memcpy(rho0, s->dens[0], sizeof(double) * s->N);
    if (rho1 && s->NP == 2) memcpy(rho1, s->dens[1], sizeof(double) * s->N);
    if (u) memcpy(u, s->v, sizeof(double) * s->N * s->D);
}

void orc_adjacent(const orc_sim* s, double* u, double* chi)
{
    if (u) memcpy(u, s->adjV, sizeof(double) * s->N * s->D);
    if (chi && s->NP == 2) memcpy(chi, s->adjChi, sizeof(double) * s->N);
}

void orc_get_populations(const orc_sim* s, double* out)
{
    for (int ph = 0; ph < s->NP; ++ph)
        memcpy(out + (uint64_t)ph * s->N * s->Q, s->f[ph], sizeof(double) * s->N * s->Q);
}

void orc_set_populations(orc_sim* s, const double* in, uint64_t t)
{
    for (int ph = 0; ph < s->NP; ++ph)
        memcpy(s->f[ph], in + (uint64_t)ph * s->N * s->Q, sizeof(double) * s->N * s->Q);
    s->t = t;
}

uint64_t orc_links(const orc_sim* s, uint32_t* linkmask, int32_t* wall,
                   double* normal, double* sinca, double* cosca)
{
    if (linkmask) {
        memset(linkmask, 0, sizeof(uint32_t) * s->N);
        for (int q = 1; q < s->Q; ++q)
            for (uint64_t k = 0; k < s->nibs[q]; ++k) linkmask[s->ibs[q][k]] |= 1u << q;
    }
    if (wall) memcpy(wall, s->ibI, sizeof(int32_t) * s->N);
    for (uint64_t c = 0; c < s->nwall; ++c) {
        if (normal) for (int d = 0; d < s->D; ++d) normal[c * s->D + d] = s->ibNorm[c * 3 + d];
        if (sinca) sinca[c] = s->ibSin[c];
        if (cosca) cosca[c] = s->ibCos[c];
    }
    return s->nwall;
}

/* Report sums over the sites of getSpace() (VD/GV/VB/PB/HF) inside the slab
 * [first,last] of the report axis, site order ascending (simulator.cpp:129-141,
 * scsp.cpp:325-341, colorfluid.cpp:841-867). */
void orc_report_sums(const orc_sim* s, int axis, uint64_t first, uint64_t last, double* out)
{
    for (int k = 0; k < 9; ++k) out[k] = 0.0;
    const int space = ORC_VD | ORC_GV | ORC_VB | ORC_PB | ORC_HF;
    for (uint64_t i = 0; i < s->N; ++i) {
        if (!label_in(s, s->raw[i], space)) continue;
        const uint64_t c = (i / s->stride[axis]) % s->sz[axis];
        if (c >= first && c <= last) {
            out[0] += V(s, axis, i);
            out[1] += s->dens[0][i];
            if (s->NP == 2) out[2] += s->dens[1][i];
        }
        if (c == first) { out[3] += s->dens[0][i]; if (s->NP == 2) out[4] += s->dens[1][i]; out[7] += 1.0; }
        if (c == last) { out[5] += s->dens[0][i]; if (s->NP == 2) out[6] += s->dens[1][i]; out[8] += 1.0; }
    }
}

/* simulator.cpp:263-282: d outer, all sites inner (solids included). */
void orc_convergence_sums(const orc_sim* s, const double* prev, const double* adj, double* out)
{
    double l[4] = {0.0, 0.0, 0.0, 0.0};
    for (int d = 0; d < s->D; ++d)
        for (uint64_t i = 0; i < s->N; ++i) {
            const double value = V(s, d, i);
            const double delta = value - prev[(uint64_t)d * s->N + i];
            l[0] += delta * delta;
            l[1] += value * value;
            l[2] += 1.0;
            if (adj) {
                const double a = adj[(uint64_t)d * s->N + i];
                l[3] += (value - a) * (value - a);
            }
        }
    memcpy(out, l, sizeof l);
}
