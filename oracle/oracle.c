/*
 * oracle.c -- CPU restatement of TCLB's node kernel (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * Lattice traversal follows the reference CPU build: every node of the
 * nx*ny*nz box, x fastest (src/LatticeContainer.inc.cpp.Rt:206-222); pull
 * streaming f_i(x) = in_i(x - e_i) with periodic wrap on all three axes
 * (src/LatticeAccess.inc.cpp.Rt:319-335 with self-aliased margins,
 * src/Lattice.cu.Rt:445-453); A/B snapshots swapped every step
 * (src/Lattice.cu.Rt:958-964).
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------ storage conversions --- */
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* IEEE binary32 -> binary16, round to nearest even (as __float2half_rn). */
uint16_t orc_f32_to_f16(float v)
{
    uint32_t x = f2u(v);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) /* inf or nan */
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0));
    if (ax >= 0x477ff000u) /* rounds to >= 65520 -> inf */
        return (uint16_t)(sign | 0x7c00u);
    if (ax < 0x38800000u) { /* below 2^-14: subnormal half or zero */
        if (ax < 0x33000000u) return (uint16_t)sign; /* < 2^-25 rounds to 0 */
        {
            uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
            int e = (int)(ax >> 23);           /* biased exponent 102..112 */
            int shift = 126 - e;               /* 14..24 */
            uint32_t q = mant >> shift;
            uint32_t rem = mant & ((1u << shift) - 1);
            uint32_t half = 1u << (shift - 1);
            if (rem > half || (rem == half && (q & 1))) q++;
            return (uint16_t)(sign | q);
        }
    }
    {
        uint32_t e = (ax >> 23) - 112;         /* rebias 127 -> 15 */
        uint32_t mant = ax & 0x7fffffu;
        uint32_t q = (e << 10) | (mant >> 13);
        uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (q & 1))) q++;
        return (uint16_t)(sign | q);
    }
}

float orc_f16_to_f32(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1f, m = h & 0x3ff;
    if (e == 0) {
        /* zero or subnormal: m * 2^-24 is exact in float */
        float v = (float)m * 5.9604644775390625e-08f;
        return sign ? -v : v;
    }
    if (e == 31) return u2f(sign | 0x7f800000u | (m << 13));
    return u2f(sign | ((e + 112) << 23) | (m << 13));
}

/* binary32 -> bfloat16, round to nearest even (as __float2bfloat16_rn). */
uint16_t orc_f32_to_bf16(float v)
{
    uint32_t x = f2u(v);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}

float orc_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

/* ----------------------------------------------- stochastic rounding ---
 * The engine's optional stochastic rounding of the 16-bit stored values
 * (no counterpart in the reference, which rounds to nearest): a value x
 * between the neighbouring representable values lo and hi (|lo| <= |x| <
 * |hi|) is stored as hi with probability (|x| - |lo|)/(|hi| - |lo|), so its
 * expectation is x.  The rounding restates gfx950's v_cvt_sr_f16_f32 and
 * v_cvt_sr_bf16_f32, which the engine uses, from their measured
 * input/output behaviour (every finite float of exponent -60 to 17 against
 * random and structured 32-bit r).
 *
 * Random words: MurmurHash3's 32-bit finalizer (fmix32, public domain) of
 * (node, step, seed), forced odd, is the node's first word; Marsaglia's
 * xorshift32 (13, 17, 5; J. Stat. Softw. 8(14), 2003) gives the next ones.
 * Direction 2j rounds with word j, direction 2j + 1 with word j << 16, so
 * the two take disjoint bits (binary16 uses the top 13, bfloat16 the top
 * 16).  A run is reproducible. */
static uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16; h *= 0x85ebca6bu;
    h ^= h >> 13; h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

static uint32_t xorshift32(uint32_t x)
{
    x ^= x << 13; x ^= x >> 17; x ^= x << 5;
    return x;
}

uint32_t orc_sr_bits(uint32_t node, uint32_t q, uint32_t step, uint32_t seed)
{
    uint32_t w = fmix32(node * 0x9e3779b1u ^ fmix32(step * 0x85ebca77u + seed)) | 1u, j;
    for (j = 0; j < q / 2; j++) w = xorshift32(w);
    return (q & 1u) ? w << 16 : w;
}

/* x -> binary16, stochastically: the float's significand aligned to
 * binary16's spacing 2^L at |x| (L = e - 10 where binary16 is normal, -24
 * where it is subnormal), cut to the 13 bits below the kept ones (lower
 * bits dropped), plus the top 13 bits of r; the carry into the kept bits
 * rounds up with probability (those 13 bits)/2^13.  |x| < 2^-37 (no bit in
 * the window) becomes +-0; a result of 2^16 or more is +-inf. */
uint16_t orc_f32_to_f16_sr(float x, uint32_t r)
{
    uint32_t u = f2u(x), a = u & 0x7fffffffu, s = (u >> 16) & 0x8000u, sig, q;
    int e;
    if (a >= 0x7f800000u) return orc_f32_to_f16(x);
    if (a < 0x2d000000u) return (uint16_t)s;
    e = (int)(a >> 23) - 127;
    if (e >= 16) return (uint16_t)(s | 0x7c00u);
    sig = (a & 0x7fffffu) | 0x800000u;
    if (e >= -14) {                    /* normal: 13 bits below the 10 kept */
        q = (sig + (r >> 19)) >> 13;   /* in [2^10, 2^11] */
        return (uint16_t)(s | (((uint32_t)(e + 14) << 10) + q));
    }
    /* subnormal: the window is 2^-25 .. 2^-37, sig's bit 23 is 2^e */
    q = ((sig >> (uint32_t)(-14 - e)) + (r >> 19)) >> 13;   /* in [0, 2^10] */
    return (uint16_t)(s | q);
}

/* x -> bfloat16, stochastically: the low 16 bits of x plus the top 16 bits
 * of r carry into the kept half with probability (low bits)/2^16 */
uint16_t orc_f32_to_bf16_sr(float x, uint32_t r)
{
    uint32_t u = f2u(x);
    if ((u & 0x7fffffffu) >= 0x7f800000u) return orc_f32_to_bf16(x);
    return (uint16_t)((u + (r >> 16)) >> 16);
}

/* ----------------------------------------------------- node arithmetic --- */
#define REAL double
#define SFX(n) n##_d
#include "oracle_core.inc"
#undef REAL
#undef SFX
#define REAL float
#define SFX(n) n##_f
#include "oracle_core.inc"
#undef REAL
#undef SFX

/* ------------------------------------------------------------ lattice --- */
static const int d2q9_e[9][3] = {{0,0,0},{1,0,0},{0,1,0},{-1,0,0},{0,-1,0},
                                 {1,1,0},{-1,1,0},{-1,-1,0},{1,-1,0}};
static const double d2q9_w[9] = {4./9, 1./9, 1./9, 1./9, 1./9,
                                 1./36, 1./36, 1./36, 1./36};
static const double d3q19_w[19] = {1./3, 1./18, 1./18, 1./18, 1./18, 1./18, 1./18,
                                   1./36, 1./36, 1./36, 1./36, 1./36, 1./36,
                                   1./36, 1./36, 1./36, 1./36, 1./36, 1./36};

struct orc_lattice {
    int model, nx, ny, nz, storage, Q, nglob;
    size_t N;
    void* buf[2];               /* A/B snapshots, storage format */
    int cur;                    /* snapshot holding the current state */
    uint16_t* flags;
    uint16_t* cuts;             /* [26][N] or NULL: no cuts */
    double* extra;              /* densities not streamed, [nextra][N]: AVG averages or bc's BC[0..1] */
    int nextra, navg;           /* how many, and how many of them are averages */
    int iter, reset_iter;       /* iterations since Init; of the last average reset */
    int sr;                     /* stochastic rounding of the 16-bit formats' stores in steps */
    uint32_t sr_seed;
    double params[16];
    double* zone;               /* [nzones][ncols] */
    int nzones, ncols;
    int e[27][3];
    double w[27];
};

static size_t esize(int storage)
{
    switch (storage) {
    case ORC_F64: return 8;
    case ORC_F32: case ORC_F32S: return 4;
    default: return 2;
    }
}

static int shifted(int storage) { return storage == ORC_F32S || storage == ORC_F16S || storage == ORC_BF16S; }

/* storage -> real (as double; exact for every format) */
static double ld(const orc_lattice* L, const void* b, size_t i, int q)
{
    switch (L->storage) {
    case ORC_F64: return ((const double*)b)[i];
    case ORC_F32: return (double)((const float*)b)[i];
    case ORC_F32S: return (double)(((const float*)b)[i] + (float)L->w[q]);
    case ORC_F16S: return (double)(orc_f16_to_f32(((const uint16_t*)b)[i]) + (float)L->w[q]);
    default: return (double)(orc_bf16_to_f32(((const uint16_t*)b)[i]) + (float)L->w[q]);
    }
}

/* real -> storage; for reduced formats v is a float value */
static void st(const orc_lattice* L, void* b, size_t i, int q, double v)
{
    switch (L->storage) {
    case ORC_F64: ((double*)b)[i] = v; break;
    case ORC_F32: ((float*)b)[i] = (float)v; break;
    case ORC_F32S: ((float*)b)[i] = (float)v - (float)L->w[q]; break;
    case ORC_F16S: ((uint16_t*)b)[i] = orc_f32_to_f16((float)v - (float)L->w[q]); break;
    default: ((uint16_t*)b)[i] = orc_f32_to_bf16((float)v - (float)L->w[q]); break;
    }
}

/* the store of a step: stochastic in the 16-bit formats when asked */
static void st_step(const orc_lattice* L, void* b, size_t n, int q, double v)
{
    size_t i = (size_t)q * L->N + n;
    if (L->sr && L->storage == ORC_F16S)
        ((uint16_t*)b)[i] = orc_f32_to_f16_sr((float)v - (float)L->w[q],
                                              orc_sr_bits((uint32_t)n, (uint32_t)q, (uint32_t)L->iter, L->sr_seed));
    else if (L->sr && L->storage == ORC_BF16S)
        ((uint16_t*)b)[i] = orc_f32_to_bf16_sr((float)v - (float)L->w[q],
                                               orc_sr_bits((uint32_t)n, (uint32_t)q, (uint32_t)L->iter, L->sr_seed));
    else
        st(L, b, i, q, v);
}

void orc_set_stochastic(orc_lattice* L, int on, uint32_t seed)
{
    L->sr = on;
    L->sr_seed = seed;
}

orc_lattice* orc_create(int model, int nx, int ny, int nz, int storage)
{
    orc_lattice* L = (orc_lattice*)calloc(1, sizeof(orc_lattice));
    int q, a;
    if (!L) return NULL;
    d3q19_sparse_init_d();
    d3q19_sparse_init_f();
    L->model = model; L->nx = nx; L->ny = ny; L->nz = nz; L->storage = storage;
    L->N = (size_t)nx * ny * nz;
    if (model == ORC_D3Q27 || model == ORC_D3Q27_CUMULANT || model == ORC_D3Q27_CUMULANT_SMAG ||
        (model >= ORC_D3Q27_CUMULANT_IB && model <= ORC_D3Q27_CUMULANT_AVG_IB_SMAG)) {
        int cum = model != ORC_D3Q27;
        cum_init_d();
        cum_init_f();
        L->Q = 27; L->nglob = cum ? 10 : 7;
        for (q = 0; q < 27; q++) {
            int j = cum ? cum_map_d[q] : q;
            for (a = 0; a < 3; a++) L->e[q][a] = q27_e_d[j][a];
            L->w[q] = q27_w_d[j];
        }
    } else if (model == ORC_D3Q19 || model == ORC_D3Q19_LID) {
        L->Q = 19; L->nglob = model == ORC_D3Q19_LID ? 30 : 26;
        for (q = 0; q < 19; q++) {
            L->e[q][0] = M_d[3][q]; L->e[q][1] = M_d[5][q]; L->e[q][2] = M_d[7][q];
            L->w[q] = d3q19_w[q];
        }
    } else {
        L->Q = 9;
        L->nglob = model == ORC_D2Q9_SRT_DRAG ? 3 : (model == ORC_D2Q9_DRAG ? 6 : 1);
        if (model == ORC_D2Q9 || model == ORC_D2Q9_AUTOSYM || model == ORC_D2Q9_AUTOSYM2 || model >= ORC_D2Q9_BC)
            L->nglob = 4;
        for (q = 0; q < 9; q++) {
            for (a = 0; a < 3; a++) L->e[q][a] = d2q9_e[q][a];
            L->w[q] = d2q9_w[q];
        }
    }
    if (model >= ORC_D3Q27_CUMULANT_AVG && model <= ORC_D3Q27_CUMULANT_AVG_IB_SMAG) L->nextra = L->navg = 13;
    if (model >= ORC_D2Q9_BC) L->nextra = 2;
    if (L->nextra) L->extra = (double*)calloc((size_t)L->nextra * L->N, sizeof(double));
    L->buf[0] = calloc(L->N * L->Q, esize(storage));
    L->buf[1] = calloc(L->N * L->Q, esize(storage));
    L->flags = (uint16_t*)calloc(L->N, sizeof(uint16_t));
    L->nzones = 1; L->ncols = 32;
    L->zone = (double*)calloc(32, sizeof(double));
    return L;
}

void orc_destroy(orc_lattice* L)
{
    if (!L) return;
    free(L->buf[0]); free(L->buf[1]); free(L->flags); free(L->cuts); free(L->extra); free(L->zone); free(L);
}

int orc_q(const orc_lattice* L) { return L->Q; }
int orc_nglobals(const orc_lattice* L) { return L->nglob; }

void orc_set_params(orc_lattice* L, const double* p, int n)
{
    int i;
    for (i = 0; i < 16; i++) L->params[i] = i < n ? p[i] : 0.0;
}

void orc_set_zone_table(orc_lattice* L, const double* table, int nzones, int ncols)
{
    int z, c;
    free(L->zone);
    /* pad the columns to 32 so every model can index its InObj weights */
    L->ncols = ncols < 32 ? 32 : ncols;
    L->nzones = nzones;
    L->zone = (double*)calloc((size_t)nzones * L->ncols, sizeof(double));
    for (z = 0; z < nzones; z++)
        for (c = 0; c < ncols; c++) L->zone[(size_t)z * L->ncols + c] = table[(size_t)z * ncols + c];
}

void orc_set_flags(orc_lattice* L, const uint16_t* flags)
{
    memcpy(L->flags, flags, L->N * sizeof(uint16_t));
}

void orc_set_cuts(orc_lattice* L, const uint16_t* cuts)
{
    if (!L->cuts) L->cuts = (uint16_t*)malloc(26 * L->N * sizeof(uint16_t));
    memcpy(L->cuts, cuts, 26 * L->N * sizeof(uint16_t));
}

static int is3d(const orc_lattice* L) { return L->model == ORC_D3Q19 || L->model == ORC_D3Q19_LID; }
static int is_d2q9m(const orc_lattice* L)
{
    return L->model == ORC_D2Q9 || L->model == ORC_D2Q9_DRAG || L->model == ORC_D2Q9_AUTOSYM ||
           L->model == ORC_D2Q9_AUTOSYM2 || L->model >= ORC_D2Q9_BC;
}
static int is_autosym(const orc_lattice* L)
{
    return L->model == ORC_D2Q9_AUTOSYM || L->model == ORC_D2Q9_AUTOSYM2 || L->model == ORC_D2Q9_BC_AUTOSYM ||
           L->model == ORC_D2Q9_BC_AUTOSYM2;
}
static int is_bc(const orc_lattice* L) { return L->model >= ORC_D2Q9_BC; }
static int is_q27(const orc_lattice* L) { return L->model == ORC_D3Q27; }
static int is_cum(const orc_lattice* L)
{
    return L->model == ORC_D3Q27_CUMULANT || L->model == ORC_D3Q27_CUMULANT_SMAG ||
           (L->model >= ORC_D3Q27_CUMULANT_IB && L->model <= ORC_D3Q27_CUMULANT_AVG_IB_SMAG);
}
static int is_ib(const orc_lattice* L)
{
    return L->model == ORC_D3Q27_CUMULANT_IB || L->model == ORC_D3Q27_CUMULANT_IB_SMAG ||
           L->model == ORC_D3Q27_CUMULANT_AVG_IB || L->model == ORC_D3Q27_CUMULANT_AVG_IB_SMAG;
}
static int is_smag(const orc_lattice* L)
{
    return L->model == ORC_D3Q27_CUMULANT_SMAG || L->model == ORC_D3Q27_CUMULANT_IB_SMAG ||
           L->model == ORC_D3Q27_CUMULANT_AVG_SMAG || L->model == ORC_D3Q27_CUMULANT_AVG_IB_SMAG;
}
/* the IB builds add group HO_BOUNDARY (IB = bit 8) before the zone */
static int zone_shift(const orc_lattice* L)
{
    return is3d(L) ? 6 : (is_cum(L) ? (is_ib(L) ? 9 : 8) : (is_autosym(L) ? 11 : (is_d2q9m(L) || is_q27(L) ? 7 : 4)));
}
/* the cuts of node n when it is an IB node of an IB build that has cuts */
static const uint16_t* ib_cuts(const orc_lattice* L, size_t n, uint16_t nt, uint16_t* tmp)
{
    int d;
    if (!is_ib(L) || !(nt & 0x100) || !L->cuts) return NULL;
    for (d = 0; d < 26; d++) tmp[d] = L->cuts[(size_t)d * L->N + n];
    return tmp;
}

static const double* zone_row(const orc_lattice* L, uint16_t nt)
{
    int z = nt >> zone_shift(L);
    if (z >= L->nzones) z = 0;
    return L->zone + (size_t)z * L->ncols;
}

void orc_init(orc_lattice* L)
{
    size_t n;
    int q;
    /* Lattice::Init (src/Lattice.cu.Rt:799-821): iter = 0, reset_iter kept;
     * the averaged densities start at their default 0 (src/conf.R:70) */
    L->iter = 0;
    if (L->navg) memset(L->extra, 0, (size_t)L->navg * L->N * sizeof(double));
    for (n = 0; n < L->N; n++) {
        const double* zr = zone_row(L, L->flags[n]);
        if (is_cum(L)) {
            /* Init(), models/flow/d3q27_cumulant/Dynamics.c.Rt:266-275 */
            int wall = ((L->flags[n] >> 3) & 0xF) == 10;
            if (L->storage == ORC_F64) {
                double f[27];
                if (wall) cum_feq_d(f, 1.0, 0., 0., 0.);
                else cum_feq_d(f, 1.0 + zr[1] * 3.0, zr[0], 0., 0.);
                for (q = 0; q < 27; q++) st(L, L->buf[0], (size_t)q * L->N + n, q, f[q]);
            } else {
                float f[27];
                if (wall) cum_feq_f(f, 1.0f, 0.f, 0.f, 0.f);
                else cum_feq_f(f, 1.0f + (float)zr[1] * 3.0f, (float)zr[0], 0.f, 0.f);
                for (q = 0; q < 27; q++) st(L, L->buf[0], (size_t)q * L->N + n, q, f[q]);
            }
        } else if (is_q27(L)) {
            /* Init(), models/flow/d3q27/Dynamics.c.Rt:184-190, no synthetic
             * turbulence modes: SetEquilibrum(1 + 3 Pressure, Velocity, 0, 0) */
            if (L->storage == ORC_F64) {
                double f[27];
                q27_feq_d(f, 1.0 + zr[1] * 3.0, zr[0], 0., 0.);
                for (q = 0; q < 27; q++) st(L, L->buf[0], (size_t)q * L->N + n, q, f[q]);
            } else {
                float f[27];
                q27_feq_f(f, 1.0f + (float)zr[1] * 3.0f, (float)zr[0], 0.f, 0.f);
                for (q = 0; q < 27; q++) st(L, L->buf[0], (size_t)q * L->N + n, q, f[q]);
            }
        } else if (is3d(L)) {
            /* Init(): SetEquilibrum(1.0, InletVelocity, 0, 0), Dynamics.c.Rt:219-221 */
            if (L->storage == ORC_F64) {
                double f[19];
                d3q19_feq_d(f, 1.0, L->params[1], 0., 0.);
                for (q = 0; q < 19; q++) st(L, L->buf[0], (size_t)q * L->N + n, q, f[q]);
            } else {
                float f[19];
                d3q19_feq_f(f, 1.0f, (float)L->params[1], 0.f, 0.f);
                for (q = 0; q < 19; q++) st(L, L->buf[0], (size_t)q * L->N + n, q, f[q]);
            }
        } else if (is_d2q9m(L)) {
            /* Init(), models/flow/d2q9/Dynamics.c.Rt:89-102: rho = 1 + 3 Pressure,
             * u = (VelocityX, VelocityY); Wall nodes at rest with rho = 1 */
            int wall = (L->flags[n] & 0xF) == 10;
            if (L->storage == ORC_F64) {
                double f[9], rho = 1 + zr[2] * 3, ux = zr[0], uy = zr[1];
                if (wall) { rho = 1; ux = 0; uy = 0; }
                d2q9m_feq_d(f, rho, ux * rho, uy * rho);
                for (q = 0; q < 9; q++) st(L, L->buf[0], (size_t)q * L->N + n, q, f[q]);
                if (is_bc(L)) {   /* SetEquilibrum: BC = J/rho on BOUNDARY nodes, else 0 */
                    int bnd = (L->flags[n] & 0xF) != 0;
                    L->extra[n] = bnd ? ux * rho / rho : 0.0;
                    L->extra[L->N + n] = bnd ? uy * rho / rho : 0.0;
                }
            } else {
                float f[9], rho = 1 + (float)zr[2] * 3, ux = (float)zr[0], uy = (float)zr[1];
                if (wall) { rho = 1; ux = 0; uy = 0; }
                d2q9m_feq_f(f, rho, ux * rho, uy * rho);
                for (q = 0; q < 9; q++) st(L, L->buf[0], (size_t)q * L->N + n, q, f[q]);
                if (is_bc(L)) {
                    int bnd = (L->flags[n] & 0xF) != 0;
                    L->extra[n] = bnd ? (double)(ux * rho / rho) : 0.0;
                    L->extra[L->N + n] = bnd ? (double)(uy * rho / rho) : 0.0;
                }
            }
        } else {
            /* Init(): SetEquilibrium(Density, Velocity_x, Velocity_y), Dynamics.c:26-38 */
            if (L->storage == ORC_F64) {
                double f[9];
                d2q9_feq_d(f, L->params[1], zr[1], zr[2]);
                for (q = 0; q < 9; q++) st(L, L->buf[0], (size_t)q * L->N + n, q, f[q]);
            } else {
                float f[9];
                d2q9_feq_f(f, (float)L->params[1], (float)zr[1], (float)zr[2]);
                for (q = 0; q < 9; q++) st(L, L->buf[0], (size_t)q * L->N + n, q, f[q]);
            }
        }
    }
    L->cur = 0;
}

static int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

/* pull the Q populations of node (x,y,z) from snapshot b */
static void pull(const orc_lattice* L, const void* b, int x, int y, int z, double* f)
{
    int q;
    for (q = 0; q < L->Q; q++) {
        int xs = wrap(x - L->e[q][0], L->nx);
        int ys = wrap(y - L->e[q][1], L->ny);
        int zs = wrap(z - L->e[q][2], L->nz);
        size_t i = (size_t)q * L->N + (size_t)xs + (size_t)L->nx * ((size_t)ys + (size_t)L->ny * zs);
        f[q] = L->storage == ORC_F64 ? ((const double*)b)[i] : ld(L, b, i, q);
    }
}

/* SymmetryAccess of the autosym builds (src/conf.R:373-445,
 * src/LatticeAccess.inc.cpp.Rt:233-260,376-470) on the pulled f of node
 * (x, y): SYMX = bits 7-8, SYMY = bits 9-10, 1 = _plus, 2 = _minus. */
static int d2q9_mirror(int q, int axis)
{
    int j;
    for (j = 0; j < 9; j++)
        if (d2q9_e[j][axis] == -d2q9_e[q][axis] && d2q9_e[j][1 - axis] == d2q9_e[q][1 - axis]) return j;
    return q;
}

static void autosym_access(const orc_lattice* L, const void* b, int x, int y, uint16_t nt, double* f)
{
    int sym[2], axis, q;
    sym[0] = (nt >> 7) & 3;
    sym[1] = (nt >> 9) & 3;
    if (L->model == ORC_D2Q9_AUTOSYM || L->model == ORC_D2Q9_BC_AUTOSYM) {
        /* resolve.symmetries after the pull: on a _plus node the densities
         * with e_axis < 0 take their mirrors, on a _minus node those with
         * e_axis > 0; X, then Y */
        for (axis = 0; axis < 2; axis++) {
            double g[9];
            for (q = 0; q < 9; q++) g[q] = f[q];
            for (q = 0; q < 9; q++) {
                if (sym[axis] == 1 && d2q9_e[q][axis] < 0) f[q] = g[d2q9_mirror(q, axis)];
                if (sym[axis] == 2 && d2q9_e[q][axis] > 0) f[q] = g[d2q9_mirror(q, axis)];
            }
        }
        return;
    }
    if (!sym[0] && !sym[1]) return;
    /* SymmetryEdge: load_ -> loadX_ -> loadY_ -> plain load, each step
     * replacing (field, offset) when the offset crosses its plane */
    for (q = 0; q < 9; q++) {
        int fld = q, d[2], i;
        size_t n;
        d[0] = -d2q9_e[q][0];
        d[1] = -d2q9_e[q][1];
        for (axis = 0; axis < 2; axis++) {
            if (d[axis] > 0 && sym[axis] == 1) { fld = d2q9_mirror(fld, axis); d[axis] = 1 - d[axis]; }
            else if (d[axis] < 0 && sym[axis] == 2) { fld = d2q9_mirror(fld, axis); d[axis] = -1 - d[axis]; }
        }
        n = (size_t)wrap(x + d[0], L->nx) + (size_t)L->nx * (size_t)wrap(y + d[1], L->ny);
        i = fld;
        f[q] = L->storage == ORC_F64 ? ((const double*)b)[(size_t)i * L->N + n] : ld(L, b, (size_t)i * L->N + n, i);
    }
}

/* rows r0 <= y + ny*z < r1 of one step */
static void step_rows(orc_lattice* L, long double* glob, long r0, long r1)
{
    const void* in = L->buf[L->cur];
    void* out = L->buf[1 - L->cur];
    int x, y, z, q;
    long r;
    int drag = L->model == ORC_D2Q9_SRT_DRAG;
    int lid = L->model == ORC_D3Q19_LID;
    for (r = r0; r < r1; r++)
        for (z = (int)(r / L->ny), y = (int)(r % L->ny), x = 0; x < L->nx; x++) {
                size_t n = (size_t)x + (size_t)L->nx * ((size_t)y + (size_t)L->ny * z);
                uint16_t nt = L->flags[n];
                const double* zr = zone_row(L, nt);
                double fd[27];
                uint16_t cbuf[26];
                const uint16_t* cut = ib_cuts(L, n, nt, cbuf);
                double old[27];
                pull(L, in, x, y, z, fd);
                if (is_autosym(L)) autosym_access(L, in, x, y, nt, fd);
                if (cut)   /* the node's own populations in the snapshot being read */
                    for (q = 0; q < 27; q++) old[q] = ld(L, in, (size_t)q * L->N + n, q);
                if (L->storage == ORC_F64) {
                    if (cut) cum_ib_d(fd, nt, cut, old, zr, glob);
                    if (is_cum(L)) cum_node_d(fd, nt, L->params, zr, glob, is_smag(L), L->navg ? L->extra + n : NULL, L->N);
                    else if (is_q27(L)) d3q27_node_d(fd, nt, L->params, zr, glob);
                    else if (is3d(L)) d3q19_node_d(fd, nt, L->params, zr, glob, lid, L->flags, x, y, z, L->nx, L->ny, L->nz);
                    else if (is_d2q9m(L)) d2q9m_node_d(fd, nt, L->params, zr, glob, L->model == ORC_D2Q9_DRAG,
                                                       is_bc(L) ? L->extra + n : NULL, L->N);
                    else d2q9_node_d(fd, nt, L->params[0], L->params[1], zr, drag, glob);
                    for (q = 0; q < L->Q; q++) st(L, out, (size_t)q * L->N + n, q, fd[q]);
                } else {
                    float ff[27];
                    for (q = 0; q < L->Q; q++) ff[q] = (float)fd[q];
                    if (cut) {
                        float oldf[27];
                        for (q = 0; q < 27; q++) oldf[q] = (float)old[q];
                        cum_ib_f(ff, nt, cut, oldf, zr, glob);
                    }
                    if (is_cum(L)) cum_node_f(ff, nt, L->params, zr, glob, is_smag(L), L->navg ? L->extra + n : NULL, L->N);
                    else if (is_q27(L)) d3q27_node_f(ff, nt, L->params, zr, glob);
                    else if (is3d(L)) d3q19_node_f(ff, nt, L->params, zr, glob, lid, L->flags, x, y, z, L->nx, L->ny, L->nz);
                    else if (is_d2q9m(L)) d2q9m_node_f(ff, nt, L->params, zr, glob, L->model == ORC_D2Q9_DRAG,
                                                       is_bc(L) ? L->extra + n : NULL, L->N);
                    else d2q9_node_f(ff, nt, (float)L->params[0], (float)L->params[1], zr, drag, glob);
                    for (q = 0; q < L->Q; q++) st_step(L, out, n, q, ff[q]);
                }
        }
}

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

/* One step.  With several threads the (y,z) rows are split into contiguous
 * chunks (OpenMP, as the reference's --with-openmp CPU build runs its
 * kernel loops); the Globals of the chunks are added in chunk order. */
static void step(orc_lattice* L, long double* glob)
{
    long rows = (long)L->ny * L->nz;
    int nt = g_threads < rows ? g_threads : (int)rows;
    int lid = L->model == ORC_D3Q19_LID;
    if (nt <= 1) {
        step_rows(L, glob, 0, rows);
    } else {
        long double part[64][32];
        int t, g;
        if (nt > 64) nt = 64;
        #pragma omp parallel for num_threads(nt) schedule(static, 1) private(g)
        for (t = 0; t < nt; t++) {
            long r0 = rows * t / nt, r1 = rows * (t + 1) / nt;
            for (g = 0; g < 32; g++) part[t][g] = 0.0L;
            step_rows(L, glob ? part[t] : NULL, r0, r1);
        }
        if (glob)
            for (t = 0; t < nt; t++)
                for (g = 0; g < 32; g++) {
                    if (is3d(L) && g == G_MAXV_OF(lid)) { if (part[t][g] > glob[g]) glob[g] = part[t][g]; }
                    else glob[g] += part[t][g];
                }
    }
    L->cur = 1 - L->cur;
}

void orc_iterate(orc_lattice* L, int niter, int glob_last, double* globals_out)
{
    int i, g;
    long double glob[32];
    for (i = 0; i < niter; i++) {
        int last = glob_last && i == niter - 1;
        if (last) for (g = 0; g < 32; g++) glob[g] = 0.0L;
        step(L, last ? glob : NULL);
        L->iter++;
    }
    if (glob_last && niter > 0 && globals_out)
        for (g = 0; g < L->nglob; g++) globals_out[g] = (double)glob[g];
}

void orc_get_f(const orc_lattice* L, double* out)
{
    size_t i;
    int q;
    for (q = 0; q < L->Q; q++)
        for (i = 0; i < L->N; i++) out[(size_t)q * L->N + i] = ld(L, L->buf[L->cur], (size_t)q * L->N + i, q);
}

void orc_set_f(orc_lattice* L, const double* in)
{
    size_t i;
    int q;
    for (q = 0; q < L->Q; q++)
        for (i = 0; i < L->N; i++) {
            double v = in[(size_t)q * L->N + i];
            if (L->storage != ORC_F64) v = (double)(float)v;
            st(L, L->buf[L->cur], (size_t)q * L->N + i, q, v);
        }
}

/* Lattice::resetAverage (src/Lattice.cu.Rt:1360-1365) */
void orc_reset_average(orc_lattice* L)
{
    L->reset_iter = L->iter;
    if (L->navg) memset(L->extra, 0, (size_t)L->navg * L->N * sizeof(double));
}

int orc_nextra(const orc_lattice* L) { return L->nextra; }

void orc_get_extra(const orc_lattice* L, double* out)
{
    if (L->nextra) memcpy(out, L->extra, (size_t)L->nextra * L->N * sizeof(double));
}

void orc_set_extra(orc_lattice* L, const double* in)
{
    if (L->nextra) memcpy(L->extra, in, (size_t)L->nextra * L->N * sizeof(double));
}

/* the AVG quantities (models/flow/d3q27_cumulant/Dynamics.c.Rt:62-99) of
 * node (x,y,z): which = 5 KinE, 6 ReStr, 7 Dissipation, 8 avgU, 9 varU,
 * 10 averageP, over average_iter = iter - reset_iter iterations */
static void avg_quantity(const orc_lattice* L, int which, int x, int y, int z, double* out)
{
    const double* A = L->extra;
    const int it = L->iter - L->reset_iter;
    const double N = (double)it;
#define AV(k, xx, yy, zz) A[(size_t)(k) * L->N + (size_t)(((xx) + L->nx) % L->nx) + \
                             (size_t)L->nx * ((size_t)(((yy) + L->ny) % L->ny) + (size_t)L->ny * (((zz) + L->nz) % L->nz))]
    double mu[3], var[3];
    int k;
    for (k = 0; k < 3; k++) mu[k] = AV(10 + k, x, y, z) / N;
    for (k = 0; k < 3; k++) var[k] = AV(1 + k, x, y, z) / N - mu[k] * mu[k];
    switch (which) {
    case 5: out[0] = 0.5 * (var[0] + var[1] + var[2]); break;
    case 6:
        out[0] = AV(6, x, y, z) / N - mu[1] * mu[2];
        out[1] = AV(5, x, y, z) / N - mu[0] * mu[2];
        out[2] = AV(4, x, y, z) / N - mu[0] * mu[1];
        break;
    case 7: {
        /* 4*average_iter*average_iter is int arithmetic in the reference */
        const double q4 = (double)(int32_t)(4u * (uint32_t)it * (uint32_t)it);
        const double nu = L->params[0];
        double gx = AV(10, x + 1, y, z) - AV(10, x - 1, y, z);
        double gy = AV(11, x, y + 1, z) - AV(11, x, y - 1, z);
        double gz = AV(12, x, y, z + 1) - AV(12, x, y, z - 1);
        double d = nu * (AV(7, x, y, z) / N - gx * gx / q4);
        d = d + nu * (AV(8, x, y, z) / N - gy * gy / q4);
        d = d + nu * (AV(9, x, y, z) / N - gz * gz / q4);
        out[0] = d;
        break; }
    case 8: for (k = 0; k < 3; k++) out[k] = mu[k]; break;
    case 9: for (k = 0; k < 3; k++) out[k] = var[k]; break;
    default: out[0] = AV(0, x, y, z) / N; break;
    }
#undef AV
}

void orc_get_quantity(const orc_lattice* L, int which, double* out)
{
    int x, y, z, q;
    for (z = 0; z < L->nz; z++)
        for (y = 0; y < L->ny; y++)
            for (x = 0; x < L->nx; x++) {
                size_t n = (size_t)x + (size_t)L->nx * ((size_t)y + (size_t)L->ny * z);
                uint16_t nt = L->flags[n];
                double f[27], rho = 0, jx = 0, jy = 0, jz = 0;
                pull(L, L->buf[L->cur], x, y, z, f);
                if (is_autosym(L)) autosym_access(L, L->buf[L->cur], x, y, nt, f);
                if (which >= 5) {
                    const int vec = which == 6 || which == 8 || which == 9;
                    avg_quantity(L, which, x, y, z, vec ? out + 3 * n : out + n);
                    continue;
                }
                if (is_cum(L)) {
                    /* getP on the pulled f; getU after RunBoundaries (0 on
                     * Wall, else (J + F/2)/rho); getSolid (1 on Wall),
                     * models/flow/d3q27_cumulant/Dynamics.c.Rt:23-62 */
                    int wall = ((nt >> 3) & 0xF) == 10;
                    for (q = 0; q < 27; q++) rho += f[q];
                    if (which == 0) out[n] = rho;
                    else if (which == 2) out[n] = (rho - 1.0) / 3.0;
                    else if (which == 4) out[n] = wall ? 1.0 : 0.0;
                    else if (which == 1) {
                        uint16_t cbuf[26];
                        const uint16_t* cut = ib_cuts(L, n, nt, cbuf);
                        if (cut) {
                            double old[27];
                            for (q = 0; q < 27; q++) old[q] = ld(L, L->buf[L->cur], (size_t)q * L->N + n, q);
                            cum_ib_d(f, nt, cut, old, zone_row(L, nt), NULL);
                        }
                        cum_boundaries_d(f, nt, zone_row(L, nt), NULL);
                        rho = 0;
                        for (q = 0; q < 27; q++) {
                            rho += f[q];
                            jx += f[q] * L->e[q][0]; jy += f[q] * L->e[q][1]; jz += f[q] * L->e[q][2];
                        }
                        out[3*n] = wall ? 0.0 : (jx + L->params[3] / 2.) / rho;
                        out[3*n+1] = wall ? 0.0 : (jy + L->params[4] / 2.) / rho;
                        out[3*n+2] = wall ? 0.0 : (jz + L->params[5] / 2.) / rho;
                    } else out[3*n] = out[3*n+1] = out[3*n+2] = 0;
                } else if (is_q27(L)) {
                    /* getP, getU, getFd, models/flow/d3q27/Dynamics.c.Rt:33-73 */
                    for (q = 0; q < 27; q++) {
                        rho += f[q];
                        jx += f[q] * L->e[q][0]; jy += f[q] * L->e[q][1]; jz += f[q] * L->e[q][2];
                    }
                    if (which == 0) out[n] = rho;
                    else if (which == 2) out[n] = (rho - 1.0) / 3.0;
                    else if (which == 1) {
                        out[3*n] = (jx + L->params[2] * 0.5) / rho;
                        out[3*n+1] = (jy + L->params[3] * 0.5) / rho;
                        out[3*n+2] = (jz + L->params[4] * 0.5) / rho;
                    } else {
                        int wall = (nt & 0xF) == 6;
                        out[3*n] = wall ? -2.0 * jx : 0.0;
                        out[3*n+1] = wall ? -2.0 * jy : 0.0;
                        out[3*n+2] = wall ? -2.0 * jz : 0.0;
                    }
                } else if (is3d(L)) {
                    /* getRho / getU / getP, Dynamics.c.Rt:23-40 */
                    for (q = 0; q < 19; q++) {
                        rho += f[q];
                        jx += f[q] * L->e[q][0]; jy += f[q] * L->e[q][1]; jz += f[q] * L->e[q][2];
                    }
                    if (which == 0) out[n] = rho;
                    else if (which == 2) out[n] = (rho - 1.0) / 3.0;
                    else {
                        double ux = jx / rho, uy = jy / rho, uz = jz / rho;
                        if ((nt & 0x1C) == 0x0C) ux = uy = uz = 0;
                        if (L->model == ORC_D3Q19_LID && (nt & 0x1C) == 0x1C) {
                            ux = L->params[11]; uy = L->params[12]; uz = L->params[13];
                        }
                        out[3*n] = ux; out[3*n+1] = uy; out[3*n+2] = uz;
                    }
                } else if (is_d2q9m(L)) {
                    /* getRho / getU, models/flow/d2q9/Dynamics.c.Rt:46-63 */
                    double d = 0, jx = 0, jy = 0;
                    for (q = 0; q < 9; q++) { d += f[q]; jx += f[q] * L->e[q][0]; jy += f[q] * L->e[q][1]; }
                    if (which == 0) out[n] = d;
                    else if (which == 1) {
                        double ux = jx / d, uy = jy / d;
                        if (is_bc(L) && (nt & 0xF) == 0) {   /* + BC/2 off the BOUNDARY nodes */
                            ux += L->extra[n] * 0.5;
                            uy += L->extra[L->N + n] * 0.5;
                        }
                        out[3*n] = ux + L->params[3] * 0.5;
                        out[3*n+1] = uy + L->params[4] * 0.5;
                        out[3*n+2] = 0;
                    } else out[n] = 0;
                } else {
                    /* getRho / getU, Dynamics.c:102-116 */
                    const double* zr = zone_row(L, nt);
                    double d = f[8]+f[7]+f[6]+f[5]+f[4]+f[3]+f[2]+f[1]+f[0];
                    if (which == 0) out[n] = d;
                    else if (which == 1) {
                        out[3*n]   = (( f[8]-f[7]-f[6]+f[5]-f[3]+f[1] )/d + zr[3]*0.5 );
                        out[3*n+1] = ((-f[8]-f[7]+f[6]+f[5]-f[4]+f[2] )/d + zr[4]*0.5 );
                        out[3*n+2] = 0;
                    } else out[n] = 0;
                }
            }
}

/* ---------------------------------------------- single-node exports --- */
void orc_d2q9_feq(double feq[9], double d, double ux, double uy) { d2q9_feq_d(feq, d, ux, uy); }
void orc_d3q19_feq(double feq[19], double rho, double jx, double jy, double jz) { d3q19_feq_d(feq, rho, jx, jy, jz); }
void orc_d3q27_feq(double feq[27], double rho, double jx, double jy, double jz)
{
    q27_init_d();
    q27_feq_d(feq, rho, jx, jy, jz);
}
void orc_d3q27_node(double f[27], uint16_t nt, const double* params, const double* zone, double* globals)
{
    long double g[8] = {0};
    int i;
    q27_init_d();
    d3q27_node_d(f, nt, params, zone, globals ? g : NULL);
    if (globals) for (i = 0; i < 7; i++) globals[i] = (double)g[i];
}
void orc_cum_node(double f[27], uint16_t nt, const double* params, const double* zone, double* globals, int smag)
{
    long double g[10] = {0};
    int i;
    cum_init_d();
    cum_node_d(f, nt, params, zone, globals ? g : NULL, smag, NULL, 0);
    if (globals) for (i = 0; i < 10; i++) globals[i] = (double)g[i];
}
void orc_cum_ib(double f[27], uint16_t nt, const uint16_t cut[26], const double old[27], const double* zone,
                double* globals)
{
    long double g[10] = {0};
    int i;
    cum_ib_d(f, nt, cut, old, zone, globals ? g : NULL);
    if (globals) for (i = 0; i < 10; i++) globals[i] = (double)g[i];
}
void orc_d3q27_basis(double M[27 * 27], int order[27], int p[27 * 3])
{
    int i, j;
    q27_init_d();
    for (i = 0; i < 27; i++) {
        order[i] = q27_ord_d[i];
        for (j = 0; j < 3; j++) p[3 * i + j] = q27_p_d[i][j];
        for (j = 0; j < 27; j++) M[27 * i + j] = q27_M_d[i][j];
    }
}
void orc_d3q19_collide(double f[19], const double* params)
{
    static const double zone[32] = {0};
    d3q19_mrt_d(f, 0x20, params, zone, NULL, 0);
}
