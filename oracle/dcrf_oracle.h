/* Test infrastructure only.  CPU restatement ("oracle") of the reference's dense-CRF
 * mean-field path (lucasb-eyer/pydensecrf, pydensecrf/densecrf/src).  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and only as
 * the checker; the product path never calls it.
 *
 * Layouts follow the reference's in-memory (Eigen column-major) storage:
 *   features  [N][d]   point-major, d contiguous
 *   values    [N][L]   label-fastest (Q, unary, messages)
 *
 * Pinning: see the header of dcrf_oracle.c for which check pins which function.
 */
#ifndef DCRF_ORACLE_H
#define DCRF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_NO_NORMALIZATION = 0, ORC_NORMALIZE_BEFORE = 1, ORC_NORMALIZE_AFTER = 2, ORC_NORMALIZE_SYMMETRIC = 3 };
enum { ORC_CONST_KERNEL = 0, ORC_DIAG_KERNEL = 1, ORC_FULL_KERNEL = 2 };
enum { ORC_POTTS = 0, ORC_DIAGONAL = 1, ORC_MATRIX = 2 };

typedef struct orc_lattice {
    int N;        /* real points */
    int npad;     /* points hashed, including the zero pad lanes: ceil(N/4)*4 */
    int d;        /* feature dimension */
    int V;        /* lattice vertices */
    int overflow; /* 1 if a key coordinate left the int16 range (wrapped, like the reference) */
    int *offset;  /* [npad*(d+1)] vertex id per (point, remainder) */
    int *rank;    /* [npad*(d+1)] */
    float *bary;  /* [npad*(d+1)] barycentric weight per (point, remainder) */
    int *nbr;     /* [(d+1)*V*2] blur neighbours (n1, n2), axis-major, -1 = absent */
    short *keys;  /* [V*d] lattice key of each vertex, vertex-id order */
} orc_lattice;

int orc_lattice_build(orc_lattice *lat, const float *feat, int N, int d);
void orc_lattice_free(orc_lattice *lat);
/* Filter with the lattice: splat, d+1 blurs (reverse order if reverse), slice.
 * in/out may alias.  The float form dispatches like the reference: L <= 2 uses the
 * scalar association, L > 2 the vector one. */
void orc_compute_f(const orc_lattice *lat, const float *in, float *out, int L, int reverse);
void orc_compute_d(const orc_lattice *lat, const double *in, double *out, int L, int reverse);

/* Exponential as the reference's Eigen build evaluates it on 4-wide packets. */
float orc_exp_packet(float x);
/* Sum of n floats in the reduction order of the reference's Eigen build. */
float orc_sum_packet(const float *x, long n);

typedef struct orc_crf orc_crf;
orc_crf *orc_crf_new(int N, int L);
orc_crf *orc_crf_new_2d(int W, int H, int L);
void orc_crf_free(orc_crf *c);
int orc_crf_num_terms(const orc_crf *c);
const orc_lattice *orc_crf_lattice(const orc_crf *c, int term);
const float *orc_crf_norm(const orc_crf *c, int term);
void orc_crf_set_unary(orc_crf *c, const float *U);       /* [N][L] */
/* compat: Potts [1], Diagonal [L], Matrix [L*L] row-major m (symmetrised as (m+m^T)/2) */
int orc_crf_add_pairwise(orc_crf *c, const float *feat, int d, int ckind, const float *compat,
                         int ktype, int ntype);
int orc_crf_add_gaussian_2d(orc_crf *c, float sx, float sy, int ckind, const float *compat,
                            int ktype, int ntype);
int orc_crf_add_bilateral_2d(orc_crf *c, float sx, float sy, float sr, float sg, float sb,
                             const unsigned char *im, int ckind, const float *compat, int ktype,
                             int ntype);
void orc_crf_inference(orc_crf *c, int n_iter, float *Q);
void orc_crf_start(orc_crf *c, float *Q);
void orc_crf_step(orc_crf *c, float *Q, float *tmp1, float *tmp2);
double orc_crf_kl(orc_crf *c, const float *Q);
void orc_crf_apply(orc_crf *c, int term, const float *Q, float *out, int transpose);
/* Compatibility parameter vector of all terms, packed like the reference. */
int orc_crf_num_compat_params(const orc_crf *c);
void orc_crf_get_compat_params(const orc_crf *c, float *p);
void orc_crf_set_compat_params(orc_crf *c, const float *p);
/* DenseCRF::gradient with the LogLikelihood objective.  Returns the objective (+mean
 * log-likelihood).  unary_grad [N][L] (sum of b over iterations), compat_grad packed. */
double orc_crf_gradient(orc_crf *c, int n_iter, const short *gt, float robust, float *unary_grad,
                        float *compat_grad);
/* The same with the kernel-parameter gradient (DenseKernel::gradient, pairwise.cpp:
 * 152-163): per term none (CONST_KERNEL), d values (DIAG_KERNEL) or d*d column-major
 * (FULL_KERNEL), at the current kernel parameters. */
double orc_crf_gradient_k(orc_crf *c, int n_iter, const short *gt, float robust, float *unary_grad,
                          float *compat_grad, float *kernel_grad);
int orc_crf_num_kernel_params(const orc_crf *c);
/* The same with another objective (objective.cpp:37-108): ORC_LOG_LIKELIHOOD (param =
 * robust), ORC_HAMMING (class_weight[n_class_weight], or NULL and param =
 * class_weight_pow), ORC_IOU.  Hamming with class weights and a labelled pixel whose
 * label is n_class_weight or more: NaN, outputs untouched. */
enum { ORC_LOG_LIKELIHOOD = 0, ORC_HAMMING = 1, ORC_IOU = 2 };
double orc_crf_gradient_obj(orc_crf *c, int n_iter, int objective, const short *gt, float param,
                            const float *class_weight, int n_class_weight, float *unary_grad, float *compat_grad,
                            float *kernel_grad);
/* Kernel parameters of all terms, packed like the reference (ones / identity at first);
 * setting them rebuilds each DIAG/FULL term's lattice from the scaled features. */
void orc_crf_get_kernel_params(const orc_crf *c, float *p);
int orc_crf_set_kernel_params(orc_crf *c, const float *p);
/* Permutohedral::gradient: df [N][d] from a, b [N][vs]. */
void orc_lattice_gradient(const orc_lattice *lat, float *df, const float *a, const float *b, int vs);

/* FP64 re-run on the same lattices (normalisers recomputed in double). */
void orc_crf_inference_fp64(orc_crf *c, int n_iter, double *Q);
/* FP64 cross-entropy loss -mean log Q_n(gt) for finite differences.  Overrides may be
 * NULL: U [N][L]; per-term raw compat parameters concatenated (Potts 1, Diagonal L,
 * Matrix L*L raw m, symmetrised inside); per-term scalar weights multiplying the term. */
double orc_crf_ce_fp64(orc_crf *c, int n_iter, const short *gt, const double *U,
                       const double *compat_raw, const double *kweights);

#ifdef __cplusplus
}
#endif
#endif
