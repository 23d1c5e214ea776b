/* Memory and undefined-behaviour check of the oracle (test infrastructure): built with
 * gcc -fsanitize=address,undefined (oracle/Makefile, target `sanitize`) and run by
 * tests/test_oracle_sanitize.py on the CPU.  It drives every entry point of
 * dcrf_oracle.h that the tests use, on small synthetic problems covering the shapes
 * they take: 2-D and generic models, Potts / Diagonal / Matrix compatibilities, the
 * four normalisations, CONST / DIAG / FULL kernels, d = 1..7, L = 1, 2, 3 and 21,
 * odd sizes (pad lanes of the SSE lattice build).  A sanitizer report makes the run
 * fail; the program prints "ok" at the end otherwise. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dcrf_oracle.h"

static unsigned rng_state = 12345u;
static float frand(void) {
    rng_state = rng_state * 1664525u + 1013904223u;
    return (float)(rng_state >> 8) / (float)(1u << 24);
}

static float *rand_array(size_t n, float lo, float hi) {
    float *a = (float *)malloc(sizeof(float) * (n ? n : 1));
    for (size_t i = 0; i < n; i++) a[i] = lo + (hi - lo) * frand();
    return a;
}

/* One model through inference, stepping, KL, apply, the three objectives, the kernel
 * gradient, parameter get/set and the FP64 re-run. */
static void exercise(orc_crf *c, int N, int L, int n_iter) {
    float *U = rand_array((size_t)N * L, 0.0f, 3.0f);
    orc_crf_set_unary(c, U);
    float *Q = (float *)malloc(sizeof(float) * N * L);
    float *t1 = (float *)malloc(sizeof(float) * N * L);
    float *t2 = (float *)malloc(sizeof(float) * N * L);
    float *out = (float *)malloc(sizeof(float) * N * L);
    orc_crf_inference(c, n_iter, Q);
    orc_crf_start(c, Q);
    orc_crf_step(c, Q, t1, t2);
    (void)orc_crf_kl(c, Q);
    for (int k = 0; k < orc_crf_num_terms(c); k++) {
        orc_crf_apply(c, k, Q, out, 0);
        orc_crf_apply(c, k, Q, out, 1);
    }
    short *gt = (short *)malloc(sizeof(short) * N);
    for (int i = 0; i < N; i++) gt[i] = (short)(i % 7 == 3 ? -1 : (i * 5) % L);
    const int ncp = orc_crf_num_compat_params(c), nkp = orc_crf_num_kernel_params(c);
    float *cg = (float *)malloc(sizeof(float) * (ncp + 1));
    float *kg = (float *)malloc(sizeof(float) * (nkp + 1));
    float *ug = (float *)malloc(sizeof(float) * N * L);
    (void)orc_crf_gradient(c, n_iter, gt, 0.0f, ug, cg);
    (void)orc_crf_gradient_k(c, n_iter, gt, 0.1f, ug, cg, kg);
    float *cw = rand_array((size_t)L, 0.5f, 2.0f);
    (void)orc_crf_gradient_obj(c, n_iter, ORC_HAMMING, gt, 1.0f, NULL, 0, ug, cg, kg);
    (void)orc_crf_gradient_obj(c, n_iter, ORC_HAMMING, gt, 0.0f, cw, L, ug, cg, kg);
    int above0 = 0;  /* a labelled pixel without a weight among the first one only */
    for (int i = 0; i < N; i++) above0 |= gt[i] >= 1;
    if (above0) {  /* weights for fewer labels than the ground truth uses: rejected */
        const double r = orc_crf_gradient_obj(c, n_iter, ORC_HAMMING, gt, 0.0f, cw, 1, ug, cg, kg);
        if (r == r) {
            fprintf(stderr, "short class weights were not rejected\n");
            exit(1);
        }
    }
    free(cw);
    (void)orc_crf_gradient_obj(c, n_iter, ORC_IOU, gt, 0.0f, NULL, 0, ug, cg, kg);
    float *cp = (float *)malloc(sizeof(float) * (ncp + 1));
    orc_crf_get_compat_params(c, cp);
    orc_crf_set_compat_params(c, cp);
    if (nkp) {
        float *kp = (float *)malloc(sizeof(float) * nkp);
        orc_crf_get_kernel_params(c, kp);
        for (int i = 0; i < nkp; i++) kp[i] *= 1.1f;
        if (orc_crf_set_kernel_params(c, kp) == 0) orc_crf_inference(c, 1, Q);
        free(kp);
    }
    double *Qd = (double *)malloc(sizeof(double) * N * L);
    orc_crf_inference_fp64(c, n_iter, Qd);
    (void)orc_crf_ce_fp64(c, n_iter, gt, NULL, NULL, NULL);
    free(Qd);
    free(cp);
    free(ug);
    free(kg);
    free(cg);
    free(gt);
    free(out);
    free(t2);
    free(t1);
    free(Q);
    free(U);
}

int main(void) {
    const int shapes[][3] = {{13, 7, 3}, {17, 9, 1}, {10, 10, 2}, {24, 20, 21}, {1, 1, 3}, {29, 1, 2}};
    for (size_t s = 0; s < sizeof(shapes) / sizeof(shapes[0]); s++) {
        const int W = shapes[s][0], H = shapes[s][1], L = shapes[s][2], N = W * H;
        unsigned char *img = (unsigned char *)malloc((size_t)N * 3);
        for (int i = 0; i < N * 3; i++) img[i] = (unsigned char)(frand() * 255.0f);
        for (int ntype = 0; ntype < 4; ntype++) {
            /* 2-D model: Potts Gaussian (DIAG), bilateral with the compatibility kind
             * cycling over Potts / Diagonal / Matrix and the kernel over CONST / DIAG / FULL */
            orc_crf *c = orc_crf_new_2d(W, H, L);
            const float potts = 3.0f;
            orc_crf_add_gaussian_2d(c, 3.0f, 2.0f, 0, &potts, 1, ntype);
            const int ck = ntype % 3, kt = (ntype + 1) % 3;
            float *cmp = rand_array((size_t)L * L, -1.0f, 1.0f);
            orc_crf_add_bilateral_2d(c, 20.0f, 30.0f, 13.0f, 17.0f, 11.0f, img, ck, cmp, kt, ntype);
            exercise(c, N, L, 2);
            orc_crf_free(c);
            /* generic model: features of d = 1 + ntype and 7 */
            orc_crf *g = orc_crf_new(N, L);
            const int d = 1 + ntype;
            float *f = rand_array((size_t)d * N, 0.0f, 4.0f);
            orc_crf_add_pairwise(g, f, d, 1, cmp, 1, ntype);
            float *f7 = rand_array((size_t)7 * N, 0.0f, 2.0f);
            orc_crf_add_pairwise(g, f7, 7, 2, cmp, 2, ntype);
            exercise(g, N, L, 2);
            orc_crf_free(g);
            free(f7);
            free(f);
            free(cmp);
        }
        /* the lattice on its own, both compute precisions and its gradient */
        float *feat = rand_array((size_t)N * 5, 0.0f, 3.0f);
        orc_lattice lat;
        memset(&lat, 0, sizeof(lat));
        if (orc_lattice_build(&lat, feat, N, 5) == 0) {
            float *in = rand_array((size_t)N * L, 0.0f, 1.0f), *o = rand_array((size_t)N * L, 0.0f, 1.0f);
            double *ind = (double *)calloc((size_t)N * L, sizeof(double)), *od = (double *)calloc((size_t)N * L, sizeof(double));
            orc_compute_f(&lat, in, o, L, 0);
            orc_compute_f(&lat, in, o, L, 1);
            orc_compute_d(&lat, ind, od, L, 1);
            float *df = (float *)malloc(sizeof(float) * N * 5);
            orc_lattice_gradient(&lat, df, in, o, L);
            free(df);
            free(od);
            free(ind);
            free(o);
            free(in);
            orc_lattice_free(&lat);
        }
        free(feat);
        free(img);
    }
    printf("ok\n");
    return 0;
}
