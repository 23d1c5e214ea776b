/* Test infrastructure only: the oracle.  A CPU restatement of the reference's dense-CRF
 * mean-field path (lucasb-eyer/pydensecrf, pydensecrf/densecrf/src; "permutohedral.cpp",
 * "pairwise.cpp", ... below are files of that directory).  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and only as the
 * checker; the product path never calls it.
 *
 * Pinning (tests/test_oracle_*.py):
 *   orc_lattice_build  - bit-exact against the reference build (oracle/_ref) on golden
 *                        fixtures (offset, rank, barycentric, neighbours, V); invariants
 *                        (barycentric sum 1, keys = remainder mod d+1, sum of key coords
 *                        = 0, sum b_r key_r = elevated point, neighbour symmetry).
 *   orc_compute_*      - against the reference build; linearity; reverse == exact
 *                        transpose (dense matrix on tiny N); K*1 = norm.
 *   softmax/compat     - against the reference build; closed forms.
 *   orc_crf_kl         - against the reference build; closed form on tiny N.
 *   orc_crf_gradient   - against the reference DenseCRF::gradient; FP64 central
 *                        differences of orc_crf_ce_fp64.
 *   orc_exp_packet / orc_sum_packet - bit-exact through Q of the reference build.
 */
#include "dcrf_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_L 4096
#define ORC_MAX_TERMS 32

/* ------------------------------------------------------------------ */
/* Eigen 3.3.9 arithmetic the reference's Q depends on                 */
/* ------------------------------------------------------------------ */

/* Cephes single-precision exp (S. L. Moshier, Cephes Math Library) in the form the
 * reference's vendored Eigen evaluates it for SSE2 packets without FMA
 * (Eigen/src/Core/arch/SSE/MathFunctions.h:107-170): clamp, n = floor(x*log2e + 1/2),
 * two-step Cody-Waite reduction, degree-5 polynomial, scale by 2^n, max with x.  Every
 * multiply-add is a separate multiply and add, as the reference's SSE2 build has no FMA. */
float orc_exp_packet(float x0)
{
    float x = x0;
    x = x < 88.3762626647950f ? x : 88.3762626647950f;
    x = x > -88.3762626647949f ? x : -88.3762626647949f;
    float fx = x * 1.44269504088896341f;
    fx = fx + 0.5f;
    /* floor via truncation and a conditional subtract */
    float t = (float)(int32_t)fx;
    fx = t > fx ? t - 1.0f : t;
    const float a = fx * 0.693359375f;
    const float b = fx * -2.12194440e-4f;
    x = x - a;
    x = x - b;
    const float z = x * x;
    float y = 1.9875691500E-4f;
    y = y * x; y = y + 1.3981999507E-3f;
    y = y * x; y = y + 8.3334519073E-3f;
    y = y * x; y = y + 4.1665795894E-2f;
    y = y * x; y = y + 1.6666665459E-1f;
    y = y * x; y = y + 5.0000001201E-1f;
    y = y * z; y = y + x;
    y = y + 1.0f;
    int32_t e = ((int32_t)fx + 0x7f) << 23;
    float p2;
    memcpy(&p2, &e, 4);
    const float r = y * p2;
    return r > x0 ? r : x0; /* _mm_max_ps(r, x0) */
}

/* Sum in the order of Eigen's vectorised reduction over a 16-byte aligned float array
 * (Eigen/src/Core/Redux.h:210-256, SSE predux PacketMath.h:540-551): two 4-wide
 * accumulators over blocks of 8, combined, one more 4-wide block if present, the
 * horizontal sum (a0+a2)+(a1+a3), then the scalar tail in order. */
float orc_sum_packet(const float *x, long n)
{
    if (n < 4) {
        float r = n > 0 ? x[0] : 0.0f;
        for (long i = 1; i < n; i++) r = r + x[i];
        return r;
    }
    const long end2 = (n / 8) * 8, end = (n / 4) * 4;
    float p0[4], p1[4];
    for (int k = 0; k < 4; k++) p0[k] = x[k];
    if (end > 4) {
        for (int k = 0; k < 4; k++) p1[k] = x[4 + k];
        for (long i = 8; i < end2; i += 8)
            for (int k = 0; k < 4; k++) {
                p0[k] = p0[k] + x[i + k];
                p1[k] = p1[k] + x[i + 4 + k];
            }
        for (int k = 0; k < 4; k++) p0[k] = p0[k] + p1[k];
        if (end > end2)
            for (int k = 0; k < 4; k++) p0[k] = p0[k] + x[end2 + k];
    }
    const float h0 = p0[0] + p0[2], h1 = p0[1] + p0[3];
    float r = h0 + h1;
    for (long i = end; i < n; i++) r = r + x[i];
    return r;
}

/* expAndNormalize (densecrf.cpp:98-106) on one column of L values: subtract the max,
 * exponentiate (packets over the first 4*floor(L/4) entries, libm expf for the tail, as
 * Eigen's assignment loop does), divide by the packet-ordered sum. */
static void softmax_col_f(const float *in, float *out, int L)
{
    float b[ORC_MAX_L];
    float mx = in[0];
    for (int k = 1; k < L; k++) mx = in[k] > mx ? in[k] : mx;
    const int vec_end = (L / 4) * 4;
    for (int k = 0; k < L; k++) {
        const float v = in[k] - mx;
        b[k] = k < vec_end ? orc_exp_packet(v) : expf(v);
    }
    const float s = orc_sum_packet(b, L);
    for (int k = 0; k < L; k++) out[k] = b[k] / s;
}

static void softmax_col_d(const double *in, double *out, int L)
{
    double mx = in[0], s = 0;
    for (int k = 1; k < L; k++) mx = in[k] > mx ? in[k] : mx;
    for (int k = 0; k < L; k++) { out[k] = exp(in[k] - mx); s += out[k]; }
    for (int k = 0; k < L; k++) out[k] /= s;
}

static float alpha_f(int d) { return 1.0f / (1 + powf(2, (float)-d)); }
static double alpha_d(int d) { return 1.0 / (1.0 + pow(2.0, (double)-d)); }

#define T float
#define SFX(name) name##_f
#include "dcrf_oracle_impl.inc"
#undef T
#undef SFX
#define T double
#define SFX(name) name##_d
#include "dcrf_oracle_impl.inc"
#undef T
#undef SFX

/* ------------------------------------------------------------------ */
/* Lattice build                                                        */
/* ------------------------------------------------------------------ */

/* Open-addressing table from a d-short key to its vertex id; ids are handed out in
 * insertion order, which is what fixes the reference's vertex numbering
 * (permutohedral.cpp:103-111).  The hash function itself does not affect results. */
typedef struct {
    int d, cap, size;
    int *slot;    /* [cap] vertex id or -1 */
    short *keys;  /* [room*d] */
    int room;
} keymap;

static uint64_t key_hash(const short *k, int d)
{
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < d; i++) {
        h ^= (uint16_t)k[i];
        h *= 0xff51afd7ed558ccdull;
        h ^= h >> 29;
    }
    return h;
}

static void keymap_init(keymap *m, int d, int expected)
{
    m->d = d;
    m->cap = 64;
    while (m->cap < 2 * expected) m->cap *= 2;
    m->size = 0;
    m->slot = (int *)malloc(sizeof(int) * (size_t)m->cap);
    for (int i = 0; i < m->cap; i++) m->slot[i] = -1;
    m->room = m->cap / 2 + 1;
    m->keys = (short *)malloc(sizeof(short) * (size_t)m->room * (d > 0 ? d : 1));
}

static void keymap_free(keymap *m)
{
    free(m->slot);
    free(m->keys);
}

static int keymap_find(const keymap *m, const short *k)
{
    size_t h = key_hash(k, m->d) & (size_t)(m->cap - 1);
    for (;;) {
        const int e = m->slot[h];
        if (e < 0) return -1;
        if (memcmp(m->keys + (size_t)e * m->d, k, sizeof(short) * m->d) == 0) return e;
        h = (h + 1) & (size_t)(m->cap - 1);
    }
}

static void keymap_grow(keymap *m)
{
    const int ncap = m->cap * 2;
    int *ns = (int *)malloc(sizeof(int) * (size_t)ncap);
    for (int i = 0; i < ncap; i++) ns[i] = -1;
    for (int e = 0; e < m->size; e++) {
        size_t h = key_hash(m->keys + (size_t)e * m->d, m->d) & (size_t)(ncap - 1);
        while (ns[h] >= 0) h = (h + 1) & (size_t)(ncap - 1);
        ns[h] = e;
    }
    free(m->slot);
    m->slot = ns;
    m->cap = ncap;
    m->room = ncap / 2 + 1;
    m->keys = (short *)realloc(m->keys, sizeof(short) * (size_t)m->room * (m->d > 0 ? m->d : 1));
}

static int keymap_insert(keymap *m, const short *k)
{
    const int f = keymap_find(m, k);
    if (f >= 0) return f;
    if (2 * (m->size + 1) > m->cap) keymap_grow(m);
    size_t h = key_hash(k, m->d) & (size_t)(m->cap - 1);
    while (m->slot[h] >= 0) h = (h + 1) & (size_t)(m->cap - 1);
    m->slot[h] = m->size;
    memcpy(m->keys + (size_t)m->size * m->d, k, sizeof(short) * m->d);
    return m->size++;
}

/* Permutohedral::init, SSE2 form (permutohedral.cpp:140-321), one point at a time.
 * Points are processed in blocks of four lanes; lanes past N load feature 0.0 and are
 * hashed too (:193-196, :261-276), so npad = ceil(N/4)*4 points insert vertices.
 * Rounding is round-half-even (_mm_cvtps_epi32 under round-to-nearest, :211-220). */
int orc_lattice_build(orc_lattice *lat, const float *feat, int N, int d)
{
    const int d1 = d + 1;
    const int npad = (N + 3) / 4 * 4;
    memset(lat, 0, sizeof(*lat));
    lat->N = N;
    lat->npad = npad;
    lat->d = d;
    lat->offset = (int *)calloc((size_t)npad * d1, sizeof(int));
    lat->rank = (int *)calloc((size_t)npad * d1, sizeof(int));
    lat->bary = (float *)calloc((size_t)npad * d1, sizeof(float));

    float scale[64], f[64], elev[65], rem0[65], rank[65], bc[66];
    short canonical[65 * 65], key[65];
    if (d > 63) return -1;
    /* canonical simplex (:168-173) */
    for (int i = 0; i <= d; i++) {
        for (int j = 0; j <= d - i; j++) canonical[i * d1 + j] = (short)i;
        for (int j = d - i + 1; j <= d; j++) canonical[i * d1 + j] = (short)(i - d1);
    }
    /* diagonal of E, computed in double then stored as float (:176-179) */
    const float inv_std_dev = (float)(sqrt(2.0 / 3.0) * d1);
    for (int i = 0; i < d; i++) scale[i] = (float)(1.0 / sqrt((double)((i + 2) * (i + 1))) * inv_std_dev);
    const float inv_d1 = 1.0f / (float)d1;
    const float fd1 = (float)d1;

    keymap map;
    keymap_init(&map, d, N);
    for (int p = 0; p < npad; p++) {
        for (int j = 0; j < d; j++) f[j] = p < N ? feat[(size_t)p * d + j] : 0.0f;
        /* elevate (:199-207) */
        float sm = 0.0f;
        for (int j = d; j > 0; j--) {
            const float cf = f[j - 1] * scale[j - 1];
            const float jc = (float)j * cf;
            elev[j] = sm - jc;
            sm = sm + cf;
        }
        elev[0] = sm;
        /* closest remainder-0 point (:209-221) */
        float sum = 0.0f;
        for (int i = 0; i <= d; i++) {
            float v = inv_d1 * elev[i];
            v = rintf(v);
            rem0[i] = v * fd1;
            sum = sum + v;
        }
        /* rank by pairwise comparison (:223-233) */
        for (int i = 0; i <= d; i++) rank[i] = 0.0f;
        for (int i = 0; i < d; i++) {
            const float di = elev[i] - rem0[i];
            for (int j = i + 1; j <= d; j++) {
                const float dj = elev[j] - rem0[j];
                const float c = di < dj ? 1.0f : 0.0f;
                rank[i] = rank[i] + c;
                rank[j] = rank[j] + (1.0f - c);
            }
        }
        /* bring the point back onto the plane (:235-242) */
        for (int i = 0; i <= d; i++) {
            rank[i] = rank[i] + sum;
            const float add = rank[i] < 0.0f ? fd1 : 0.0f;
            const float sub = rank[i] >= fd1 ? fd1 : 0.0f;
            rank[i] = rank[i] + (add - sub);
            rem0[i] = rem0[i] + (add - sub);
        }
        /* barycentric coordinates (:244-263) */
        for (int i = 0; i < d + 2; i++) bc[i] = 0.0f;
        for (int i = 0; i <= d; i++) {
            const float e = elev[i] - rem0[i];
            const float v = e * inv_d1;
            const int q = d - (int)rank[i];
            bc[q] = bc[q] + v;
            bc[q + 1] = bc[q + 1] - v;
        }
        bc[0] = bc[0] + (1.0f + bc[d + 1]);
        /* vertices of the enclosing simplex (:266-276) */
        for (int r = 0; r <= d; r++) {
            for (int i = 0; i < d; i++) {
                const float kv = rem0[i] + (float)canonical[r * d1 + (int)rank[i]];
                const int ki = (int)kv;
                if (ki > 32767 || ki < -32768) lat->overflow = 1;
                key[i] = (short)ki;
            }
            lat->offset[(size_t)p * d1 + r] = keymap_insert(&map, key);
            lat->rank[(size_t)p * d1 + r] = (int)rank[r];
            lat->bary[(size_t)p * d1 + r] = bc[r];
        }
    }
    lat->V = map.size;
    const int V = lat->V;
    lat->keys = (short *)malloc(sizeof(short) * (size_t)(V > 0 ? V : 1) * (d > 0 ? d : 1));
    memcpy(lat->keys, map.keys, sizeof(short) * (size_t)V * d);
    /* blur neighbours along each of the d+1 axes (:296-318) */
    lat->nbr = (int *)malloc(sizeof(int) * 2 * (size_t)d1 * (V > 0 ? V : 1));
    short n1[65], n2[65];
    for (int j = 0; j <= d; j++)
        for (int v = 0; v < V; v++) {
            const short *k = map.keys + (size_t)v * d;
            for (int i = 0; i < d; i++) {
                n1[i] = (short)(k[i] - 1);
                n2[i] = (short)(k[i] + 1);
            }
            if (j < d) { /* axis d only moves every stored coordinate */
                n1[j] = (short)(k[j] + d);
                n2[j] = (short)(k[j] - d);
            }
            lat->nbr[2 * ((size_t)j * V + v)] = keymap_find(&map, n1);
            lat->nbr[2 * ((size_t)j * V + v) + 1] = keymap_find(&map, n2);
        }
    keymap_free(&map);
    return 0;
}

void orc_lattice_free(orc_lattice *lat)
{
    free(lat->offset);
    free(lat->rank);
    free(lat->bary);
    free(lat->nbr);
    free(lat->keys);
    memset(lat, 0, sizeof(*lat));
}

void orc_compute_f(const orc_lattice *lat, const float *in, float *out, int L, int reverse)
{
    compute_f(lat, in, out, L, reverse, L <= 2);
}

void orc_compute_d(const orc_lattice *lat, const double *in, double *out, int L, int reverse)
{
    compute_d(lat, in, out, L, reverse, 0);
}

/* Permutohedral::gradient (permutohedral.cpp:612-695).  df [N][d] is zeroed, then for
 * dir 0 the lattice splats a and blurs axes 0..d, for dir 1 it splats b and blurs d..0
 * (its own scalar splat and blur: the old value plus a double 0.5*(n1+n2), :645-659);
 * the slicing gradient rotates each point's vertex rows by rank into differences
 * (alpha = 1/(1+2^-d)/(d+1), :628) and pairs their elevation with b (dir 0) or a
 * (dir 1), summed over the vs values (:662-691). */
void orc_lattice_gradient(const orc_lattice *lat, float *df, const float *a, const float *b, int vs)
{
    const int N = lat->N, d = lat->d, V = lat->V;
    float *values = (float *)malloc(sizeof(float) * (size_t)(V + 2) * vs);
    float *nv = (float *)malloc(sizeof(float) * (size_t)(V + 2) * vs);
    float *r_a = (float *)malloc(sizeof(float) * (size_t)(d + 1) * vs);
    float *sm = (float *)malloc(sizeof(float) * (size_t)vs);
    memset(df, 0, sizeof(float) * (size_t)N * d);
    float sf[32];
    const float inv_std_dev = (float)(sqrt(2.0 / 3.0) * (d + 1));
    for (int i = 0; i < d; i++) sf[i] = (float)(1.0 / sqrt((double)((i + 2) * (i + 1))) * inv_std_dev);
    const float alpha = 1.0f / (1 + powf(2, (float)-d)) / (float)(d + 1);
    for (int dir = 0; dir < 2; dir++) {
        memset(values, 0, sizeof(float) * (size_t)(V + 2) * vs);
        memset(nv, 0, sizeof(float) * (size_t)(V + 2) * vs);
        const float *src = dir ? b : a, *other = dir ? a : b;
        for (int i = 0; i < N; i++)
            for (int j = 0; j <= d; j++) {
                const int o = lat->offset[i * (d + 1) + j] + 1;
                const float w = lat->bary[i * (d + 1) + j];
                for (int k = 0; k < vs; k++) {
                    const float p = w * src[(size_t)i * vs + k];
                    values[(size_t)o * vs + k] = values[(size_t)o * vs + k] + p;
                }
            }
        for (int j = dir ? d : 0; j <= d && j >= 0; dir ? j-- : j++) {
            for (int i = 0; i < V; i++) {
                const float *ov = values + (size_t)(i + 1) * vs;
                float *w = nv + (size_t)(i + 1) * vs;
                const int n1 = lat->nbr[2 * ((size_t)j * V + i)] + 1;
                const int n2 = lat->nbr[2 * ((size_t)j * V + i) + 1] + 1;
                const float *v1 = values + (size_t)n1 * vs, *v2 = values + (size_t)n2 * vs;
                for (int k = 0; k < vs; k++) {
                    const float s2 = v1[k] + v2[k];
                    w[k] = (float)((double)ov[k] + 0.5 * (double)s2);
                }
            }
            float *t = values;
            values = nv;
            nv = t;
        }
        for (int i = 0; i < N; i++) {
            memset(r_a, 0, sizeof(float) * (size_t)(d + 1) * vs);
            for (int j = 0; j <= d; j++) {
                const int r0 = d - lat->rank[i * (d + 1) + j];
                const int r1 = r0 + 1 > d ? 0 : r0 + 1;
                const int o0 = lat->offset[i * (d + 1) + r0] + 1;
                const int o1 = lat->offset[i * (d + 1) + r1] + 1;
                for (int k = 0; k < vs; k++) {
                    const float p0 = alpha * values[(size_t)o0 * vs + k];
                    r_a[j * vs + k] = r_a[j * vs + k] + p0;
                    const float p1 = alpha * values[(size_t)o1 * vs + k];
                    r_a[j * vs + k] = r_a[j * vs + k] - p1;
                }
            }
            memcpy(sm, r_a, sizeof(float) * (size_t)vs);
            for (int j = 1; j <= d; j++) {
                float grad = 0;
                for (int k = 0; k < vs; k++) {
                    const float jr = (float)j * r_a[j * vs + k];
                    const float diff = sm[k] - jr;
                    const float v = sf[j - 1] * diff;
                    const float p = other[(size_t)i * vs + k] * v;
                    grad = grad + p;
                    sm[k] = sm[k] + r_a[j * vs + k];
                }
                df[(size_t)i * d + j - 1] = df[(size_t)i * d + j - 1] + grad;
            }
        }
    }
    free(values);
    free(nv);
    free(r_a);
    free(sm);
}

/* ------------------------------------------------------------------ */
/* DenseCRF                                                             */
/* ------------------------------------------------------------------ */

typedef struct {
    orc_lattice lat;
    float *feat;     /* [N][d] the term's features (DenseKernel::f_) */
    float *kparams;  /* DIAG [d] scales, FULL [d*d] column-major matrix (parameters_) */
    int ktype, ntype, ckind, ncompat;
    float *compat;   /* Potts [1], Diagonal [L], Matrix [L*L] symmetrised W */
    float *norm;     /* [N] */
    double *norm_d;  /* [N] FP64 normaliser on the same lattice */
} orc_term;

struct orc_crf {
    int N, L, W, H, nterms;
    float *unary;    /* [N][L] or NULL */
    orc_term terms[ORC_MAX_TERMS];
};

orc_crf *orc_crf_new(int N, int L)
{
    orc_crf *c = (orc_crf *)calloc(1, sizeof(orc_crf));
    c->N = N;
    c->L = L;
    c->W = N;
    c->H = 1;
    return c;
}

orc_crf *orc_crf_new_2d(int W, int H, int L)
{
    orc_crf *c = orc_crf_new(W * H, L);
    c->W = W;
    c->H = H;
    return c;
}

void orc_crf_free(orc_crf *c)
{
    if (!c) return;
    for (int k = 0; k < c->nterms; k++) {
        orc_lattice_free(&c->terms[k].lat);
        free(c->terms[k].feat);
        free(c->terms[k].kparams);
        free(c->terms[k].compat);
        free(c->terms[k].norm);
        free(c->terms[k].norm_d);
    }
    free(c->unary);
    free(c);
}

int orc_crf_num_terms(const orc_crf *c) { return c->nterms; }
const orc_lattice *orc_crf_lattice(const orc_crf *c, int term) { return &c->terms[term].lat; }
const float *orc_crf_norm(const orc_crf *c, int term) { return c->terms[term].norm; }

void orc_crf_set_unary(orc_crf *c, const float *U)
{
    if (!c->unary) c->unary = (float *)malloc(sizeof(float) * (size_t)c->N * c->L);
    memcpy(c->unary, U, sizeof(float) * (size_t)c->N * c->L);
}

static int compat_len(int ckind, int L) { return ckind == ORC_POTTS ? 1 : ckind == ORC_DIAGONAL ? L : L * L; }

/* MatrixCompatibility stores W = 0.5*(m + m^T) (labelcompatibility.cpp:79). */
static void set_compat(orc_term *t, int L, const float *raw)
{
    const int n = compat_len(t->ckind, L);
    if (t->ckind != ORC_MATRIX) {
        memcpy(t->compat, raw, sizeof(float) * n);
        return;
    }
    for (int a = 0; a < L; a++)
        for (int b = 0; b < L; b++) {
            const float s = raw[a * L + b] + raw[b * L + a];
            t->compat[a * L + b] = 0.5f * s;
        }
}

/* DenseCRF::addPairwiseEnergy -> PairwisePotential -> DenseKernel (densecrf.cpp:54-60,
 * pairwise.cpp:116-122, 40-62): build the lattice now, then the normaliser. */
int orc_crf_add_pairwise(orc_crf *c, const float *feat, int d, int ckind, const float *compat,
                         int ktype, int ntype)
{
    if (c->nterms >= ORC_MAX_TERMS || c->L > ORC_MAX_L) return -1;
    orc_term *t = &c->terms[c->nterms];
    memset(t, 0, sizeof(*t));
    if (orc_lattice_build(&t->lat, feat, c->N, d) != 0) return -1;
    t->feat = (float *)malloc(sizeof(float) * (size_t)c->N * d);
    memcpy(t->feat, feat, sizeof(float) * (size_t)c->N * d);
    /* DenseKernel constructor (pairwise.cpp:116-122): ones, resp. the identity */
    t->kparams = (float *)calloc((size_t)d * d, sizeof(float));
    for (int j = 0; j < d; j++) t->kparams[ktype == ORC_FULL_KERNEL ? j * d + j : j] = 1.0f;
    t->ktype = ktype;
    t->ntype = ntype;
    t->ckind = ckind;
    t->ncompat = compat_len(ckind, c->L);
    t->compat = (float *)malloc(sizeof(float) * t->ncompat);
    set_compat(t, c->L, compat);
    t->norm = (float *)malloc(sizeof(float) * (size_t)c->N);
    t->norm_d = (double *)malloc(sizeof(double) * (size_t)c->N);
    make_norm_f(&t->lat, ntype, t->norm);
    make_norm_d(&t->lat, ntype, t->norm_d);
    c->nterms++;
    return 0;
}

/* DenseCRF2D::addPairwiseGaussian (densecrf.cpp:61-69): features (x/sx, y/sy). */
int orc_crf_add_gaussian_2d(orc_crf *c, float sx, float sy, int ckind, const float *compat,
                            int ktype, int ntype)
{
    float *f = (float *)malloc(sizeof(float) * 2 * (size_t)c->N);
    for (int j = 0; j < c->H; j++)
        for (int i = 0; i < c->W; i++) {
            f[2 * ((size_t)j * c->W + i)] = (float)i / sx;
            f[2 * ((size_t)j * c->W + i) + 1] = (float)j / sy;
        }
    const int r = orc_crf_add_pairwise(c, f, 2, ckind, compat, ktype, ntype);
    free(f);
    return r;
}

/* DenseCRF2D::addPairwiseBilateral (densecrf.cpp:70-81): (x/sx, y/sy, r/sr, g/sg, b/sb)
 * from an HWC uint8 image. */
int orc_crf_add_bilateral_2d(orc_crf *c, float sx, float sy, float sr, float sg, float sb,
                             const unsigned char *im, int ckind, const float *compat, int ktype,
                             int ntype)
{
    float *f = (float *)malloc(sizeof(float) * 5 * (size_t)c->N);
    for (int j = 0; j < c->H; j++)
        for (int i = 0; i < c->W; i++) {
            const size_t p = (size_t)j * c->W + i;
            f[5 * p + 0] = (float)i / sx;
            f[5 * p + 1] = (float)j / sy;
            f[5 * p + 2] = (float)im[p * 3 + 0] / sr;
            f[5 * p + 3] = (float)im[p * 3 + 1] / sg;
            f[5 * p + 4] = (float)im[p * 3 + 2] / sb;
        }
    const int r = orc_crf_add_pairwise(c, f, 5, ckind, compat, ktype, ntype);
    free(f);
    return r;
}

/* PairwisePotential::apply / applyTranspose (pairwise.cpp:173-183): kernel filter, then
 * the compatibility in place. */
static void term_apply_f(const orc_crf *c, const orc_term *t, const float *Q, float *out, int transpose)
{
    filter_f(&t->lat, t->norm, t->ntype, Q, out, c->L, transpose, c->L <= 2);
    compat_apply_f(t->ckind, t->compat, c->N, c->L, out, out, transpose);
}

void orc_crf_apply(orc_crf *c, int term, const float *Q, float *out, int transpose)
{
    term_apply_f(c, &c->terms[term], Q, out, transpose);
}

static void softmax_f(const float *in, float *out, int N, int L)
{
    for (int i = 0; i < N; i++) softmax_col_f(in + (size_t)i * L, out + (size_t)i * L, L);
}

static void neg_unary_f(const orc_crf *c, float *dst)
{
    const size_t n = (size_t)c->N * c->L;
    for (size_t i = 0; i < n; i++) dst[i] = c->unary ? -c->unary[i] : 0.0f;
}

/* DenseCRF::inference (densecrf.cpp:115-131): Q0 = softmax(-U); per iteration
 * tmp1 = -U - sum_k pairwise_k(Q); Q = softmax(tmp1). */
void orc_crf_inference(orc_crf *c, int n_iter, float *Q)
{
    const size_t n = (size_t)c->N * c->L;
    float *tmp1 = (float *)malloc(sizeof(float) * n), *tmp2 = (float *)malloc(sizeof(float) * n);
    neg_unary_f(c, tmp1);
    softmax_f(tmp1, Q, c->N, c->L);
    for (int it = 0; it < n_iter; it++) {
        neg_unary_f(c, tmp1);
        for (int k = 0; k < c->nterms; k++) {
            term_apply_f(c, &c->terms[k], Q, tmp2, 0);
            for (size_t i = 0; i < n; i++) tmp1[i] = tmp1[i] - tmp2[i];
        }
        softmax_f(tmp1, Q, c->N, c->L);
    }
    free(tmp1);
    free(tmp2);
}

/* DenseCRF::startInference (densecrf.cpp:178-186): softmax(-U), or zeros without a unary. */
void orc_crf_start(orc_crf *c, float *Q)
{
    const size_t n = (size_t)c->N * c->L;
    if (!c->unary) {
        memset(Q, 0, sizeof(float) * n);
        return;
    }
    float *tmp = (float *)malloc(sizeof(float) * n);
    neg_unary_f(c, tmp);
    softmax_f(tmp, Q, c->N, c->L);
    free(tmp);
}

/* DenseCRF::stepInference (densecrf.cpp:187-201): Q updated in place; tmp1 holds
 * -U - sum_k pairwise_k(Q), tmp2 the last term's message (zeros with no terms). */
void orc_crf_step(orc_crf *c, float *Q, float *tmp1, float *tmp2)
{
    const size_t n = (size_t)c->N * c->L;
    for (size_t i = 0; i < n; i++) tmp1[i] = c->unary ? 0.0f - c->unary[i] : 0.0f;
    for (int k = 0; k < c->nterms; k++) {
        term_apply_f(c, &c->terms[k], Q, tmp2, 0);
        for (size_t i = 0; i < n; i++) tmp1[i] = tmp1[i] - tmp2[i];
    }
    softmax_f(tmp1, Q, c->N, c->L);
}

/* DenseCRF::klDivergence (densecrf.cpp:214-235): entropy and unary terms accumulate in
 * double (the unary product is a float product); each pairwise term adds the float
 * packet-ordered sum of Q .* pairwise_k(Q). */
double orc_crf_kl(orc_crf *c, const float *Q)
{
    const size_t n = (size_t)c->N * c->L;
    double kl = 0;
    for (size_t i = 0; i < n; i++) {
        const float q = Q[i] > 1e-20f ? Q[i] : 1e-20f;
        kl += (double)Q[i] * log((double)q);
    }
    if (c->unary)
        for (size_t i = 0; i < n; i++) {
            const float p = c->unary[i] * Q[i];
            kl += (double)p;
        }
    float *tmp = (float *)malloc(sizeof(float) * n);
    for (int k = 0; k < c->nterms; k++) {
        term_apply_f(c, &c->terms[k], Q, tmp, 0);
        for (size_t i = 0; i < n; i++) tmp[i] = Q[i] * tmp[i];
        kl += (double)orc_sum_packet(tmp, (long)n);
    }
    free(tmp);
    return kl;
}

int orc_crf_num_compat_params(const orc_crf *c)
{
    int n = 0;
    for (int k = 0; k < c->nterms; k++) {
        const orc_term *t = &c->terms[k];
        n += t->ckind == ORC_MATRIX ? c->L * (c->L + 1) / 2 : t->ncompat;
    }
    return n;
}

/* LabelCompatibility::parameters (labelcompatibility.cpp:49-53, 70-72, 88-94): Matrix
 * packs the upper triangle W(i, j>=i) row by row. */
void orc_crf_get_compat_params(const orc_crf *c, float *p)
{
    const int L = c->L;
    for (int k = 0; k < c->nterms; k++) {
        const orc_term *t = &c->terms[k];
        if (t->ckind == ORC_MATRIX) {
            for (int i = 0; i < L; i++)
                for (int j = i; j < L; j++) *p++ = t->compat[i * L + j];
        } else {
            for (int i = 0; i < t->ncompat; i++) *p++ = t->compat[i];
        }
    }
}

void orc_crf_set_compat_params(orc_crf *c, const float *p)
{
    const int L = c->L;
    for (int k = 0; k < c->nterms; k++) {
        orc_term *t = &c->terms[k];
        if (t->ckind == ORC_MATRIX) {
            for (int i = 0; i < L; i++)
                for (int j = i; j < L; j++) {
                    t->compat[i * L + j] = *p;
                    t->compat[j * L + i] = *p++;
                }
        } else {
            for (int i = 0; i < t->ncompat; i++) t->compat[i] = *p++;
        }
    }
}

/* sumAndNormalize (densecrf.cpp:107-114): out = sum(b)*q - b per pixel. */
static void sum_and_normalize_f(const float *b, const float *Q, float *out, int N, int L)
{
    for (int i = 0; i < N; i++) {
        const float s = orc_sum_packet(b + (size_t)i * L, L);
        for (int k = 0; k < L; k++) {
            const float sq = s * Q[(size_t)i * L + k];
            out[(size_t)i * L + k] = sq - b[(size_t)i * L + k];
        }
    }
}

/* Compatibility gradient of one term given b and KQ (labelcompatibility.cpp:57-61,
 * 76-78, 101-108).  Potts: minus the packet-ordered float sum of b.*KQ; Diagonal: a
 * sequential float sum per label (Eigen's rowwise sum over a strided row); Matrix:
 * G = b*KQ^T accumulated in double (the reference's GEMM blocking order depends on the
 * host's cache sizes, so tests compare it with a relative tolerance). */
static void compat_gradient_f(const orc_term *t, int N, int L, const float *b, const float *KQ, float *g)
{
    if (t->ckind == ORC_POTTS) {
        const size_t n = (size_t)N * L;
        float *p = (float *)malloc(sizeof(float) * n);
        for (size_t i = 0; i < n; i++) p[i] = b[i] * KQ[i];
        g[0] = -orc_sum_packet(p, (long)n);
        free(p);
    } else if (t->ckind == ORC_DIAGONAL) {
        for (int k = 0; k < L; k++) {
            float s = b[k] * KQ[k];
            for (int i = 1; i < N; i++) {
                const float p = b[(size_t)i * L + k] * KQ[(size_t)i * L + k];
                s = s + p;
            }
            g[k] = s;
        }
    } else {
        double *G = (double *)calloc((size_t)L * L, sizeof(double));
        for (int i = 0; i < N; i++)
            for (int a = 0; a < L; a++)
                for (int bb = 0; bb < L; bb++)
                    G[a * L + bb] += (double)b[(size_t)i * L + a] * (double)KQ[(size_t)i * L + bb];
        int k = 0;
        for (int i = 0; i < L; i++)
            for (int j = i; j < L; j++, k++)
                g[k] = (float)(G[i * L + j] + (i != j ? G[j * L + i] : 0.0));
        free(G);
    }
}

/* DenseCRF::kernelParameters (densecrf.cpp:334-347, pairwise.cpp:129-140). */
void orc_crf_get_kernel_params(const orc_crf *c, float *p)
{
    for (int k = 0; k < c->nterms; k++) {
        const orc_term *t = &c->terms[k];
        const int d = t->lat.d;
        const int n = t->ktype == ORC_CONST_KERNEL ? 0 : (t->ktype == ORC_DIAG_KERNEL ? d : d * d);
        memcpy(p, t->kparams, sizeof(float) * (size_t)n);
        p += n;
    }
}

/* DenseCRF::setKernelParameters (densecrf.cpp:348-360) -> DenseKernel::setParameters
 * (pairwise.cpp:141-151): the lattice and normaliser are rebuilt from diag(p) f (DIAG)
 * or P f (FULL, P column-major); f itself is kept. */
int orc_crf_set_kernel_params(orc_crf *c, const float *p)
{
    for (int k = 0; k < c->nterms; k++) {
        orc_term *t = &c->terms[k];
        const int d = t->lat.d, N = c->N;
        if (t->ktype == ORC_CONST_KERNEL) continue;
        const int n = t->ktype == ORC_DIAG_KERNEL ? d : d * d;
        memcpy(t->kparams, p, sizeof(float) * (size_t)n);
        p += n;
        float *g = (float *)malloc(sizeof(float) * (size_t)N * d);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < d; j++) {
                const float *f = t->feat + (size_t)i * d;
                if (t->ktype == ORC_DIAG_KERNEL) {
                    g[(size_t)i * d + j] = t->kparams[j] * f[j];
                } else {
                    float s = 0.0f;
                    for (int q = 0; q < d; q++) {
                        const float m = t->kparams[q * d + j] * f[q];
                        s = s + m;
                    }
                    g[(size_t)i * d + j] = s;
                }
            }
        orc_lattice_free(&t->lat);
        const int r = orc_lattice_build(&t->lat, g, N, d);
        free(g);
        if (r != 0) return -1;
        make_norm_f(&t->lat, t->ntype, t->norm);
        make_norm_d(&t->lat, t->ntype, t->norm_d);
    }
    return 0;
}

int orc_crf_num_kernel_params(const orc_crf *c)
{
    int n = 0;
    for (int k = 0; k < c->nterms; k++) {
        const orc_term *t = &c->terms[k];
        const int d = t->lat.d;
        n += t->ktype == ORC_CONST_KERNEL ? 0 : (t->ktype == ORC_DIAG_KERNEL ? d : d * d);
    }
    return n;
}

/* DenseKernel::featureGradient (pairwise.cpp:88-114): fg [N][d] for the term's
 * normalisation, from lattice gradients (above) and plain lattice filters. */
static void feature_gradient_f(const orc_crf *c, const orc_term *t, const float *a, const float *b, float *fg)
{
    const int N = c->N, L = c->L, d = t->lat.d;
    const size_t n = (size_t)N * L;
    if (t->ntype == ORC_NO_NORMALIZATION) {
        orc_lattice_gradient(&t->lat, fg, a, b, L);
        return;
    }
    const float *nm = t->norm;
    float *x = (float *)malloc(sizeof(float) * n), *ones = (float *)malloc(sizeof(float) * n);
    float *u = (float *)malloc(sizeof(float) * n), *v = (float *)malloc(sizeof(float) * n);
    float *r = (float *)malloc(sizeof(float) * (size_t)N * d);
    for (size_t i = 0; i < n; i++) ones[i] = 1.0f;
    if (t->ntype == ORC_NORMALIZE_SYMMETRIC) {
        /* fa = K^T(a n), fb = K(b n); r = grad(0.5 (a fb + fa b) n^3, 1); g = grad(a n, b n) */
        float *an = (float *)malloc(sizeof(float) * n), *bn = (float *)malloc(sizeof(float) * n);
        for (int i = 0; i < N; i++)
            for (int k = 0; k < L; k++) {
                an[(size_t)i * L + k] = a[(size_t)i * L + k] * nm[i];
                bn[(size_t)i * L + k] = b[(size_t)i * L + k] * nm[i];
            }
        orc_compute_f(&t->lat, an, u, L, 1); /* fa */
        orc_compute_f(&t->lat, bn, v, L, 0); /* fb */
        for (int i = 0; i < N; i++) {
            const float n2 = nm[i] * nm[i];
            const float n3 = n2 * nm[i];
            for (int k = 0; k < L; k++) {
                const size_t q = (size_t)i * L + k;
                const float p1 = a[q] * v[q], p2 = u[q] * b[q];
                const float sum = p1 + p2;
                const float h = 0.5f * sum;
                x[q] = h * n3;
            }
        }
        orc_lattice_gradient(&t->lat, r, x, ones, L);
        orc_lattice_gradient(&t->lat, fg, an, bn, L);
        free(an);
        free(bn);
    } else if (t->ntype == ORC_NORMALIZE_AFTER) {
        /* fb = K(b); r = grad(a fb n^2, 1); g = grad(a n, b) */
        orc_compute_f(&t->lat, b, v, L, 0);
        for (int i = 0; i < N; i++) {
            const float n2 = nm[i] * nm[i];
            for (int k = 0; k < L; k++) {
                const size_t q = (size_t)i * L + k;
                const float p = a[q] * v[q];
                x[q] = p * n2;
                u[q] = a[q] * nm[i];
            }
        }
        orc_lattice_gradient(&t->lat, r, x, ones, L);
        orc_lattice_gradient(&t->lat, fg, u, b, L);
    } else {
        /* NORMALIZE_BEFORE: fa = K^T(a); r = grad(fa b n^2, 1); g = grad(a, b n) */
        orc_compute_f(&t->lat, a, u, L, 1);
        for (int i = 0; i < N; i++) {
            const float n2 = nm[i] * nm[i];
            for (int k = 0; k < L; k++) {
                const size_t q = (size_t)i * L + k;
                const float p = u[q] * b[q];
                x[q] = p * n2;
                v[q] = b[q] * nm[i];
            }
        }
        orc_lattice_gradient(&t->lat, r, x, ones, L);
        orc_lattice_gradient(&t->lat, fg, a, v, L);
    }
    for (size_t i = 0; i < (size_t)N * d; i++) fg[i] = -r[i] + fg[i];
    free(x);
    free(ones);
    free(u);
    free(v);
    free(r);
}

/* PairwisePotential::kernelGradient (pairwise.cpp:202-207) -> DenseKernel::gradient
 * (:152-163): lbl = compatibility applied to Q; fg = featureGradient(b, lbl); DIAG
 * returns the row sums of f .* fg (a sequential float sum per dimension), FULL the
 * d x d product fg f^T packed column-major (accumulated in double: the reference's
 * GEMM order depends on the host's cache blocking). */
static void kernel_gradient_f(const orc_crf *c, const orc_term *t, const float *b, const float *Q, float *g)
{
    const int N = c->N, L = c->L, d = t->lat.d;
    if (t->ktype == ORC_CONST_KERNEL) return;
    float *lbl = (float *)malloc(sizeof(float) * (size_t)N * L);
    float *fg = (float *)malloc(sizeof(float) * (size_t)N * d);
    compat_apply_f(t->ckind, t->compat, N, L, Q, lbl, 0);
    feature_gradient_f(c, t, b, lbl, fg);
    if (t->ktype == ORC_DIAG_KERNEL) {
        for (int j = 0; j < d; j++) {
            float s = t->feat[j] * fg[j];
            for (int i = 1; i < N; i++) {
                const float p = t->feat[(size_t)i * d + j] * fg[(size_t)i * d + j];
                s = s + p;
            }
            g[j] = s;
        }
    } else {
        for (int col = 0; col < d; col++)
            for (int row = 0; row < d; row++) {
                double s = 0;
                for (int i = 0; i < N; i++) s += (double)fg[(size_t)i * d + row] * (double)t->feat[(size_t)i * d + col];
                g[col * d + row] = (float)s;
            }
    }
    free(lbl);
    free(fg);
}

double orc_crf_gradient(orc_crf *c, int n_iter, const short *gt, float robust, float *unary_grad,
                        float *compat_grad)
{
    return orc_crf_gradient_k(c, n_iter, gt, robust, unary_grad, compat_grad, NULL);
}

double orc_crf_gradient_k(orc_crf *c, int n_iter, const short *gt, float robust, float *unary_grad,
                          float *compat_grad, float *kernel_grad)
{
    return orc_crf_gradient_obj(c, n_iter, ORC_LOG_LIKELIHOOD, gt, robust, NULL, 0, unary_grad, compat_grad,
                                kernel_grad);
}

/* ObjectiveFunction::evaluate (objective.cpp:37-108): returns the objective, b = d_mul_Q
 * (zeroed first).  LogLikelihood (:37-50, param = robust); Hamming (:51-80, class weights
 * cw[ncw] if given, else from param = class_weight_pow as in the constructor :51-63);
 * IntersectionOverUnion (:81-108). */
static double objective_f(int kind, const short *gt, float param, const float *cw_in, int ncw, const float *Q,
                          int N, int L, float *b)
{
    memset(b, 0, sizeof(float) * (size_t)N * L);
    double r = 0;
    if (kind == ORC_LOG_LIKELIHOOD) {
        for (int i = 0; i < N; i++)
            if (gt[i] >= 0 && gt[i] < L) {
                const float q = Q[(size_t)i * L + gt[i]];
                const float qr = q + param;
                const float QQ = qr > 1e-20f ? qr : 1e-20f;
                r += log((double)QQ) / N;
                const float dq = q / QQ;
                b[(size_t)i * L + gt[i]] = b[(size_t)i * L + gt[i]] + dq / (float)N;
            }
    } else if (kind == ORC_HAMMING) {
        float *cw;
        int M;
        if (cw_in) {
            M = ncw;
            cw = (float *)malloc(sizeof(float) * (size_t)(M > 0 ? M : 1));
            memcpy(cw, cw_in, sizeof(float) * (size_t)M);
        } else {
            M = 0;
            for (int i = 0; i < N; i++)
                if (gt[i] >= M) M = gt[i] + 1;
            float *cnt = (float *)calloc((size_t)(M > 0 ? M : 1), sizeof(float));
            cw = (float *)malloc(sizeof(float) * (size_t)(M > 0 ? M : 1));
            for (int i = 0; i < N; i++)
                if (gt[i] >= 0) cnt[gt[i]] = cnt[gt[i]] + 1.0f;
            const float total = orc_sum_packet(cnt, M);
            for (int l = 0; l < M; l++) cw[l] = powf(cnt[l] / total, -param);
            float *prod = (float *)malloc(sizeof(float) * (size_t)(M > 0 ? M : 1));
            for (int l = 0; l < M; l++) prod[l] = cnt[l] * cw[l];
            const float z = orc_sum_packet(prod, M);
            for (int l = 0; l < M; l++) cw[l] = cw[l] / z;
            free(cnt);
            free(prod);
        }
        for (int i = 0; i < N; i++)
            if (gt[i] >= 0 && gt[i] < L) {
                const float QQ = cw[gt[i]] * Q[(size_t)i * L + gt[i]];
                r += QQ;
                b[(size_t)i * L + gt[i]] = b[(size_t)i * L + gt[i]] + QQ;
            }
        free(cw);
    } else {
        double *in = (double *)calloc((size_t)L, sizeof(double)), *un = (double *)malloc(sizeof(double) * L);
        for (int l = 0; l < L; l++) un[l] = 1e-20;
        for (int i = 0; i < N; i++)
            if (gt[i] >= 0 && gt[i] < L) {
                in[gt[i]] += Q[(size_t)i * L + gt[i]];
                un[gt[i]] += 1;
                for (int l = 0; l < L; l++)
                    if (l != gt[i]) un[l] += Q[(size_t)i * L + l];
            }
        for (int i = 0; i < N; i++)
            if (gt[i] >= 0 && gt[i] < L)
                for (int l = 0; l < L; l++) {
                    const double q = Q[(size_t)i * L + l];
                    b[(size_t)i * L + l] = l == gt[i] ? (float)(q / (un[l] * L)) : (float)(-q * in[l] / (un[l] * un[l] * L));
                }
        for (int l = 0; l < L; l++) r += in[l] / un[l];
        r /= L;
        free(in);
        free(un);
    }
    return r;
}

/* DenseCRF::gradient (densecrf.cpp:238-297) with the objectives above. */
double orc_crf_gradient_obj(orc_crf *c, int n_iter, int objective, const short *gt, float param,
                            const float *class_weight, int n_class_weight, float *unary_grad, float *compat_grad,
                            float *kernel_grad)
{
    const int N = c->N, L = c->L;
    const size_t n = (size_t)N * L;
    /* Hamming with given class weights: a labelled pixel whose label has no weight has no
     * defined objective (the reference reads past its weight vector, objective.cpp:73):
     * NaN, outputs untouched (the library rejects the call the same way) */
    if (objective == ORC_HAMMING && class_weight)
        for (int i = 0; i < N; i++)
            if (gt[i] >= 0 && gt[i] < L && gt[i] >= n_class_weight) return NAN;
    float **Q = (float **)malloc(sizeof(float *) * (n_iter + 1));
    float *tmp1 = (float *)malloc(sizeof(float) * n), *tmp2 = (float *)malloc(sizeof(float) * n);
    float *b = (float *)malloc(sizeof(float) * n), *kq = (float *)malloc(sizeof(float) * n);
    for (int it = 0; it <= n_iter; it++) Q[it] = (float *)malloc(sizeof(float) * n);
    neg_unary_f(c, tmp1);
    softmax_f(tmp1, Q[0], N, L);
    for (int it = 0; it < n_iter; it++) {
        neg_unary_f(c, tmp1);
        for (int k = 0; k < c->nterms; k++) {
            term_apply_f(c, &c->terms[k], Q[it], tmp2, 0);
            for (size_t i = 0; i < n; i++) tmp1[i] = tmp1[i] - tmp2[i];
        }
        softmax_f(tmp1, Q[it + 1], N, L);
    }
    const float *Qn = Q[n_iter];
    const double r = objective_f(objective, gt, param, class_weight, n_class_weight, Qn, N, L, b);
    sum_and_normalize_f(b, Qn, b, N, L);
    if (unary_grad) memcpy(unary_grad, b, sizeof(float) * n);
    const int ncp = orc_crf_num_compat_params(c);
    const int nkp = orc_crf_num_kernel_params(c);
    if (compat_grad) memset(compat_grad, 0, sizeof(float) * ncp);
    if (kernel_grad) memset(kernel_grad, 0, sizeof(float) * nkp);
    float *pg = (float *)malloc(sizeof(float) * (ncp > 0 ? ncp : 1));
    float *kg = (float *)malloc(sizeof(float) * (nkp > 0 ? nkp : 1));
    for (int it = n_iter - 1; it >= 0; it--) {
        memset(tmp1, 0, sizeof(float) * n);
        int ip = 0, ik = 0;
        for (int k = 0; k < c->nterms; k++) {
            const orc_term *t = &c->terms[k];
            const int np = t->ckind == ORC_MATRIX ? L * (L + 1) / 2 : t->ncompat;
            if (compat_grad) {
                /* PairwisePotential::gradient re-filters Q[it] (pairwise.cpp:190-195) */
                filter_f(&t->lat, t->norm, t->ntype, Q[it], kq, L, 0, L <= 2);
                compat_gradient_f(t, N, L, b, kq, pg);
                for (int j = 0; j < np; j++) compat_grad[ip + j] = compat_grad[ip + j] + pg[j];
            }
            ip += np;
            const int nk = t->ktype == ORC_CONST_KERNEL ? 0 : (t->ktype == ORC_DIAG_KERNEL ? t->lat.d : t->lat.d * t->lat.d);
            if (kernel_grad && nk) {
                kernel_gradient_f(c, t, b, Q[it], kg);
                for (int j = 0; j < nk; j++) kernel_grad[ik + j] = kernel_grad[ik + j] + kg[j];
            }
            ik += nk;
            term_apply_f(c, t, b, tmp2, 1);
            for (size_t i = 0; i < n; i++) tmp1[i] = tmp1[i] + tmp2[i];
        }
        for (size_t i = 0; i < n; i++) tmp2[i] = tmp1[i] * Q[it][i];
        sum_and_normalize_f(tmp2, Q[it], b, N, L);
        if (unary_grad)
            for (size_t i = 0; i < n; i++) unary_grad[i] = unary_grad[i] + b[i];
    }
    for (int it = 0; it <= n_iter; it++) free(Q[it]);
    free(Q);
    free(tmp1);
    free(tmp2);
    free(b);
    free(kq);
    free(pg);
    free(kg);
    return r;
}

/* ------------------------------------------------------------------ */
/* FP64 re-run                                                          */
/* ------------------------------------------------------------------ */

typedef struct {
    double *compat[ORC_MAX_TERMS]; /* per term, symmetrised */
    double kw[ORC_MAX_TERMS];
} compat64;

static void compat64_init(const orc_crf *c, const double *raw, const double *kweights, compat64 *out)
{
    const int L = c->L;
    for (int k = 0; k < c->nterms; k++) {
        const orc_term *t = &c->terms[k];
        out->compat[k] = (double *)malloc(sizeof(double) * t->ncompat);
        out->kw[k] = kweights ? kweights[k] : 1.0;
        if (raw) {
            if (t->ckind == ORC_MATRIX) {
                for (int a = 0; a < L; a++)
                    for (int b = 0; b < L; b++) out->compat[k][a * L + b] = 0.5 * (raw[a * L + b] + raw[b * L + a]);
            } else {
                for (int i = 0; i < t->ncompat; i++) out->compat[k][i] = raw[i];
            }
            raw += t->ncompat;
        } else {
            for (int i = 0; i < t->ncompat; i++) out->compat[k][i] = (double)t->compat[i];
        }
    }
}

static void compat64_free(const orc_crf *c, compat64 *cc)
{
    for (int k = 0; k < c->nterms; k++) free(cc->compat[k]);
}

static void inference64(orc_crf *c, int n_iter, const double *U, const compat64 *cc, double *Q)
{
    const int N = c->N, L = c->L;
    const size_t n = (size_t)N * L;
    double *e = (double *)malloc(sizeof(double) * n), *m = (double *)malloc(sizeof(double) * n);
    for (size_t i = 0; i < n; i++) e[i] = U ? -U[i] : (c->unary ? -(double)c->unary[i] : 0.0);
    for (int i = 0; i < N; i++) softmax_col_d(e + (size_t)i * L, Q + (size_t)i * L, L);
    for (int it = 0; it < n_iter; it++) {
        for (size_t i = 0; i < n; i++) e[i] = U ? -U[i] : (c->unary ? -(double)c->unary[i] : 0.0);
        for (int k = 0; k < c->nterms; k++) {
            const orc_term *t = &c->terms[k];
            filter_d(&t->lat, t->norm_d, t->ntype, Q, m, L, 0, 0);
            compat_apply_d(t->ckind, cc->compat[k], N, L, m, m, 0);
            for (size_t i = 0; i < n; i++) e[i] -= cc->kw[k] * m[i];
        }
        for (int i = 0; i < N; i++) softmax_col_d(e + (size_t)i * L, Q + (size_t)i * L, L);
    }
    free(e);
    free(m);
}

void orc_crf_inference_fp64(orc_crf *c, int n_iter, double *Q)
{
    compat64 cc;
    compat64_init(c, NULL, NULL, &cc);
    inference64(c, n_iter, NULL, &cc, Q);
    compat64_free(c, &cc);
}

double orc_crf_ce_fp64(orc_crf *c, int n_iter, const short *gt, const double *U,
                       const double *compat_raw, const double *kweights)
{
    const int N = c->N, L = c->L;
    double *Q = (double *)malloc(sizeof(double) * (size_t)N * L);
    compat64 cc;
    compat64_init(c, compat_raw, kweights, &cc);
    inference64(c, n_iter, U, &cc, Q);
    double s = 0;
    for (int i = 0; i < N; i++)
        if (gt[i] >= 0 && gt[i] < L) s -= log(Q[(size_t)i * L + gt[i]]);
    compat64_free(c, &cc);
    free(Q);
    return s / N;
}
