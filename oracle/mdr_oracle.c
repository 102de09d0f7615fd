/*
 * mdr_oracle.c -- CPU oracle for the MDR hot path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this code; the product library never does. Each function cites the ProDM
 * source it restates. Type-generic parts live in mdr_oracle_impl.inc and are
 * instantiated for (float, uint32_t) and (double, uint64_t).
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math, so
 * every floating-point operation rounds exactly as written).
 */
#include "mdr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* geometry: RefactorUtils.hpp:21-34 (compute_level_dims) and :41-59          */
/* ------------------------------------------------------------------------ */
void orc_level_dims(const uint32_t *dims, int nd, int L, uint32_t *out) {
    for (int d = 0; d < nd; d++) {
        uint32_t n = dims[d];
        for (int j = 0; j <= L; j++) {
            out[(L - j) * nd + d] = n;
            n = (n >> 1) + 1;
        }
    }
}

void orc_level_elements(const uint32_t *ld, int nd, int L, uint64_t *out) {
    uint64_t prev = 1;
    for (int d = 0; d < nd; d++) prev *= ld[d];
    out[0] = prev;
    for (int l = 1; l <= L; l++) {
        uint64_t cur = 1;
        for (int d = 0; d < nd; d++) cur *= ld[l * nd + d];
        out[l] = cur - prev;
        prev = cur;
    }
}

int orc_geometry_ok(const uint32_t *dims, int nd, int L) {
    if (nd < 1 || nd > 3 || L < 0 || L > 30) return 0;
    for (int d = 0; d < nd; d++) {
        uint32_t n = dims[d];
        if (n == 0) return 0;
        if (n == 1) continue;
        for (int l = L; l >= 1; l--) {
            if (n < 3 || (n & 1u) == 0) return 0;
            n = (n >> 1) + 1;
        }
    }
    return 1;
}

/* pad nd dims to 3 with leading 1s */
static void pad3(const uint32_t *v, int nd, uint32_t fill, uint32_t *o) {
    for (int d = 0; d < 3; d++) o[d] = fill;
    for (int d = 0; d < nd; d++) o[3 - nd + d] = v[d];
}

/* strides of the (padded) array: DirectInterleaver.hpp:20-21 semantics (the
 * innermost stride is 1; padded outer dims have extent 1 so their stride is
 * never used).                                                              */
static void strides3(const uint32_t *dims3, const uint32_t *strides, int nd, uint64_t *s) {
    s[2] = 1;
    s[1] = dims3[2];
    s[0] = (uint64_t)dims3[1] * dims3[2];
    if (strides)
        for (int d = 0; d < nd - 1; d++) s[3 - nd + d] = strides[d];
}

/* ------------------------------------------------------------------------ */
/* exact accumulation helpers for the squared-error collector                 */
/* A term t >= 0 (float or double) is added exactly into per-exponent buckets */
/* of mantissa sums; the bucket array is then rounded once to double.         */
/* ------------------------------------------------------------------------ */
typedef unsigned __int128 u128;

/* big unsigned integer, little-endian 64-bit limbs */
#define BIG_LIMBS 40
typedef struct { uint64_t w[BIG_LIMBS]; } big_t;

static void big_add_shifted(big_t *b, u128 v, int shift) {
    /* adds v << shift (shift >= 0) */
    int limb = shift / 64, off = shift % 64;
    uint64_t parts[4] = {0, 0, 0, 0};
    uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
    if (off == 0) {
        parts[0] = lo; parts[1] = hi; parts[2] = 0;
    } else {
        parts[0] = lo << off;
        parts[1] = (lo >> (64 - off)) | (hi << off);
        parts[2] = hi >> (64 - off);
    }
    uint64_t carry = 0;
    for (int i = 0; limb + i < BIG_LIMBS; i++) {
        uint64_t add = (i < 3) ? parts[i] : 0;
        if (i >= 3 && carry == 0) break;
        u128 s = (u128)b->w[limb + i] + add + carry;
        b->w[limb + i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
}

static int big_bit(const big_t *b, int bit) { return (int)((b->w[bit / 64] >> (bit % 64)) & 1u); }

/* value = big * 2^scale, rounded to nearest-even double (normal range) */
static double big_to_double(const big_t *b, int scale) {
    int top = -1;
    for (int i = BIG_LIMBS - 1; i >= 0; i--)
        if (b->w[i]) { top = i * 64 + 63 - __builtin_clzll(b->w[i]); break; }
    if (top < 0) return 0.0;
    if (top <= 52) return ldexp((double)b->w[0], scale); /* exact */
    int lsb = top - 52;
    uint64_t mant = 0;
    for (int bit = top; bit >= lsb; bit--) mant = (mant << 1) | (uint64_t)big_bit(b, bit);
    int rbit = big_bit(b, lsb - 1);
    int sticky = 0;
    for (int bit = lsb - 2; bit >= 0 && !sticky; bit--) sticky = big_bit(b, bit);
    if (rbit && (sticky || (mant & 1u))) mant += 1;
    return ldexp((double)mant, scale + lsb);
}

/* ---- MGARD helpers shared by both precisions (restated from the math) --- */
/* Load-vector stencil (R * M_f) on a uniform line, scaled by 3/h: interior
 * [1/4, 3/2, 5/2, 3/2, 1/4], ends [5/4, 3/2, 1/4]. c == 1 means the dim is
 * inactive (extent 1) and the operator is the identity.                   */
static double rweight(int j, int k, uint32_t c) {
    if (c == 1) return k == 0 ? 1.0 : 0.0;
    if (k == 0) return (j == 0 || j == (int)c - 1) ? 1.25 : 2.5;
    if (k == 1 || k == -1) return 1.5;
    return 0.25;
}

/* Thomas factors of the coarse mass matrix scaled by 3/h:
 * tridiag(1, 4, 1) with end diagonals 2.                                  */
static void thomas_factors(uint32_t m, double *inv, double *cp) {
    for (uint32_t i = 0; i < m; i++) {
        double b = (i == 0 || i == m - 1) ? 2.0 : 4.0;
        double w = (i == 0) ? b : b - cp[i - 1];
        inv[i] = 1.0 / w;
        cp[i] = inv[i];
    }
}

static void thomas_line(double *x, uint64_t stride, uint32_t m, const double *inv, const double *cp) {
    double y = x[0] * inv[0];
    x[0] = y;
    for (uint32_t i = 1; i < m; i++) {
        y = (x[i * stride] - y) * inv[i];
        x[i * stride] = y;
    }
    double z = x[(uint64_t)(m - 1) * stride];
    for (int i = (int)m - 2; i >= 0; i--) {
        z = x[(uint64_t)i * stride] - cp[i] * z;
        x[(uint64_t)i * stride] = z;
    }
}

/* solve along dims 2, 1, 0 of a dense c0 x c1 x c2 array */
static void solve3(double *f, const uint32_t c[3]) {
    uint32_t mx = c[0] > c[1] ? c[0] : c[1];
    if (c[2] > mx) mx = c[2];
    double *inv = (double *)malloc(sizeof(double) * mx), *cp = (double *)malloc(sizeof(double) * mx);
    if (c[2] > 1) {
        thomas_factors(c[2], inv, cp);
#pragma omp parallel for schedule(static)
        for (uint64_t l = 0; l < (uint64_t)c[0] * c[1]; l++) thomas_line(f + l * c[2], 1, c[2], inv, cp);
    }
    if (c[1] > 1) {
        thomas_factors(c[1], inv, cp);
#pragma omp parallel for schedule(static)
        for (uint32_t a = 0; a < c[0]; a++)
            for (uint32_t b = 0; b < c[2]; b++)
                thomas_line(f + (uint64_t)a * c[1] * c[2] + b, c[2], c[1], inv, cp);
    }
    if (c[0] > 1) {
        thomas_factors(c[0], inv, cp);
#pragma omp parallel for schedule(static)
        for (uint64_t l = 0; l < (uint64_t)c[1] * c[2]; l++)
            thomas_line(f + l, (uint64_t)c[1] * c[2], c[0], inv, cp);
    }
    free(inv);
    free(cp);
}

/* position of fine index i after the coarse-first reorder of one dim */
static inline uint32_t rpos(uint32_t i, uint32_t c) { return (i & 1u) ? c + (i >> 1) : (i >> 1); }

/* GroupedBP progressive state (GroupedBPEncoder.hpp:379-380), keyed by level */
struct orc_gb_state {
    int max_levels;
    uint8_t **rb;      /* recording bitplanes per block, NULL until seen */
    uint32_t *nblocks;
    uint8_t **signs;   /* one byte per element */
};

/* BitDecoder::decode (PerBitBPEncoder.hpp:41-60): next bit, LSB first */
static inline int pb_next_bit(const uint64_t *stream, uint64_t *pos) {
    const int b = (int)((stream[*pos / 64] >> (*pos % 64)) & 1u);
    (*pos)++;
    return b;
}

/* PerBit progressive state (PerBitBPEncoder.hpp:400-401: signs and "sign
 * recorded" flags per element), keyed by level                              */
struct orc_pb_state {
    int max_levels;
    uint32_t *n;
    uint8_t **signs, **flags;
};

/* ------------------------------------------------------------------------ */
/* type-generic part                                                          */
/* ------------------------------------------------------------------------ */
#define SFX f32
#define T float
#define TS uint32_t
#define TFP uint32_t
#define TFPS int32_t
#define BLOCK 32
#define PREC 23
#define NB_MASK 0xaaaaaaaau
#define IS_DOUBLE 0
#include "mdr_oracle_impl.inc"
#undef SFX
#undef T
#undef TS
#undef TFP
#undef TFPS
#undef BLOCK
#undef PREC
#undef NB_MASK
#undef IS_DOUBLE

#define SFX f64
#define T double
#define TS uint64_t
#define TFP uint64_t
#define TFPS int64_t
#define BLOCK 64
#define PREC 52
#define NB_MASK 0xaaaaaaaaaaaaaaaaull
#define IS_DOUBLE 1
#include "mdr_oracle_impl.inc"

/* ---- GroupedBP progressive state lifetime ---------------------------- */
orc_gb_state *orc_gb_state_new(int max_levels) {
    orc_gb_state *s = (orc_gb_state *)calloc(1, sizeof(orc_gb_state));
    s->max_levels = max_levels;
    s->rb = (uint8_t **)calloc((size_t)max_levels, sizeof(uint8_t *));
    s->signs = (uint8_t **)calloc((size_t)max_levels, sizeof(uint8_t *));
    s->nblocks = (uint32_t *)calloc((size_t)max_levels, sizeof(uint32_t));
    return s;
}

void orc_gb_state_free(orc_gb_state *s) {
    if (!s) return;
    for (int i = 0; i < s->max_levels; i++) { free(s->rb[i]); free(s->signs[i]); }
    free(s->rb); free(s->signs); free(s->nblocks); free(s);
}

/* shared by both instantiations: fetch or parse the per-level header */
int orc_gb_state_level(orc_gb_state *s, int level, const uint8_t **stream0,
                       uint32_t n, uint32_t block, int start) {
    if (level < 0 || level >= s->max_levels) return -1;
    if (!s->rb[level]) {
        if (start != 0) return -2; /* header lives in plane 0 */
        uint32_t nblk;
        memcpy(&nblk, *stream0, 4);
        if (nblk != (n - 1) / block + 1) return -3;
        s->rb[level] = (uint8_t *)malloc(nblk);
        memcpy(s->rb[level], *stream0 + 4, nblk);
        s->nblocks[level] = nblk;
        s->signs[level] = (uint8_t *)calloc(n, 1);
        *stream0 += 4 + nblk;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* encoder-fused level errors (fp32 data)                                     */
/* NegaBinaryBPEncoder::encode(..., level_errors) (NegaBinaryBPEncoder.hpp:   */
/* 65-125) with collect_level_errors (:231-239), and GroupedBPEncoder::encode */
/* (..., level_errors) (GroupedBPEncoder.hpp:76-145) with collect_level_errors*/
/* (:324-335). Per element, nb + 1 squared truncation errors in the shifted   */
/* units (value * 2^(nb - e), e + 2 for NegaBinary), each computed with the   */
/* reference's types (float products where it multiplies floats, double       */
/* otherwise); the table is scaled by 2^(2 (e - nb)) at the end. Both         */
/* collectors take float / uint32 arguments, so only fp32 data is defined.    */
/* ------------------------------------------------------------------------ */
typedef void (*fused_emit_fn)(void *ctx, int bin, double term);

static uint32_t fused_mask(int k) { return k >= 31 ? 0x7fffffffu : ((1u << k) - 1u); }

static void nb_fused_terms(float x, int exp2, int nb, fused_emit_fn emit, void *ctx) {
    const float s = ldexpf(x, nb - exp2);
    const int32_t q = (int32_t)s;
    const uint32_t nbq = ((uint32_t)q + 0xaaaaaaaau) ^ 0xaaaaaaaau;
    const float m = s - (float)q;
    const float mm = m * m;
    emit(ctx, nb, (double)mm);
    for (int k = 1; k < nb; k++) {
        const int32_t v = (int32_t)(((nbq & fused_mask(k)) ^ 0xaaaaaaaau) - 0xaaaaaaaau);
        const double d = (double)v + (double)m;
        emit(ctx, nb - k, d * d);
    }
    const float ss = s * s;
    emit(ctx, 0, (double)ss);
}

static void gb_fused_terms(float x, int exp, int nb, fused_emit_fn emit, void *ctx) {
    const float s = ldexpf(x, nb - exp);
    const float data = fabsf(s);
    const uint32_t fp = (uint32_t)data;
    const float mf = data - (float)fp;
    const double m = (double)mf;
    emit(ctx, nb, m * m);
    for (int k = 1; k < nb; k++) {
        const double d = (double)(fp & fused_mask(k)) + m;
        emit(ctx, nb - k, d * d);
    }
    const float dd = data * data;
    emit(ctx, 0, (double)dd);
}

static void fused_emit_seq(void *ctx, int bin, double term) { ((double *)ctx)[bin] += term; }

typedef struct { u128 *bucket; } fused_exact_ctx;
static void fused_emit_exact(void *ctx, int bin, double term) {
    fused_exact_ctx *c = (fused_exact_ctx *)ctx;
    uint64_t b;
    memcpy(&b, &term, 8);
    const unsigned be = (unsigned)((b >> 52) & 0x7ffu);
    const uint64_t mant = b & ((1ull << 52) - 1);
    const unsigned slot = be ? be : 1;
    const uint64_t M = be ? (mant | (1ull << 52)) : mant;
    c->bucket[(size_t)bin * 2048 + slot] += (u128)M;
}

typedef void (*fused_terms_fn)(float x, int e, int nb, fused_emit_fn emit, void *ctx);

static void fused_level_errors(fused_terms_fn terms, const float *x, uint64_t n, int e, int nb, int exact,
                               double *out) {
    for (int i = 0; i <= nb; i++) out[i] = 0;
    if (!exact) {
        for (uint64_t i = 0; i < n; i++) terms(x[i], e, nb, fused_emit_seq, out);
    } else {
        fused_exact_ctx c;
        c.bucket = (u128 *)calloc((size_t)(nb + 1) * 2048, sizeof(u128));
        for (uint64_t i = 0; i < n; i++) terms(x[i], e, nb, fused_emit_exact, &c);
        for (int bin = 0; bin <= nb; bin++) {
            big_t big;
            memset(&big, 0, sizeof(big));
            for (unsigned slot = 1; slot < 2048; slot++) {
                const u128 v = c.bucket[(size_t)bin * 2048 + slot];
                if (v) big_add_shifted(&big, v, (int)slot - 1);
            }
            out[bin] = big_to_double(&big, -1074);
        }
        free(c.bucket);
    }
    for (int i = 0; i <= nb; i++) out[i] = ldexp(out[i], 2 * (e - nb));
}

void orc_nb_level_errors_f32(const float *x, uint64_t n, int32_t exp, int nb, double *out) {
    fused_level_errors(nb_fused_terms, x, n, exp + 2, nb, 1, out);
}
void orc_nb_level_errors_seq_f32(const float *x, uint64_t n, int32_t exp, int nb, double *out) {
    fused_level_errors(nb_fused_terms, x, n, exp + 2, nb, 0, out);
}
void orc_gb_level_errors_f32(const float *x, uint64_t n, int32_t exp, int nb, double *out) {
    fused_level_errors(gb_fused_terms, x, n, exp, nb, 1, out);
}
void orc_gb_level_errors_seq_f32(const float *x, uint64_t n, int32_t exp, int nb, double *out) {
    fused_level_errors(gb_fused_terms, x, n, exp, nb, 0, out);
}

orc_pb_state *orc_pb_state_new(int max_levels) {
    orc_pb_state *s = (orc_pb_state *)calloc(1, sizeof(*s));
    s->max_levels = max_levels;
    s->n = (uint32_t *)calloc((size_t)max_levels, sizeof(uint32_t));
    s->signs = (uint8_t **)calloc((size_t)max_levels, sizeof(uint8_t *));
    s->flags = (uint8_t **)calloc((size_t)max_levels, sizeof(uint8_t *));
    return s;
}
void orc_pb_state_free(orc_pb_state *s) {
    if (!s) return;
    for (int i = 0; i < s->max_levels; i++) { free(s->signs[i]); free(s->flags[i]); }
    free(s->signs); free(s->flags); free(s->n); free(s);
}
int orc_pb_state_level(orc_pb_state *s, int level, uint32_t n, uint8_t **signs, uint8_t **flags) {
    if (level < 0 || level >= s->max_levels) return -1;
    if (!s->signs[level]) {
        s->signs[level] = (uint8_t *)calloc(n ? n : 1, 1);
        s->flags[level] = (uint8_t *)calloc(n ? n : 1, 1);
        s->n[level] = n;
    } else if (s->n[level] != n) {
        return -3;
    }
    *signs = s->signs[level];
    *flags = s->flags[level];
    return 0;
}
