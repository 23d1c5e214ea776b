/* dcrf_hip.h - C ABI of the MI355X (gfx950) dense-CRF mean-field library.
 *
 * Every entry point stands in for one operation of the reference's DenseCRF /
 * DenseCRF2D interface (lucasb-eyer/pydensecrf: pydensecrf/densecrf.pyx and
 * pydensecrf/densecrf/src), cited beside it.  The host front end
 * (pydensecrf_amd/densecrf.pyx) binds these with `cdef extern`.
 *
 * Conventions
 *  - Plain pointers and sizes; no C++ types cross the boundary.
 *  - Every function returns an int status: DCRF_OK (0) or a negative code.  The
 *    message of the last failure on the calling thread is dcrf_last_error().
 *  - Host arrays use the reference's Python layouts: features (d, N) row-major, unary
 *    (L, N) row-major, rgb images (H, W, 3) uint8.  Q and the step temporaries use the
 *    reference's in-memory layout [N][L] (label fastest), which is what the reference's
 *    MatrixXf buffers expose (pydensecrf/eigen.pyx:68-95).
 *  - Inputs are copied; the caller keeps ownership.  A model owns its device memory,
 *    lattices and stream; dcrf_destroy frees them.  Models are not thread-safe.
 *  - Host-pointer calls are synchronous at return.  *_device calls take device
 *    pointers and a hipStream_t (as void*), enqueue work on it and return.
 */
#ifndef DCRF_HIP_H
#define DCRF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DCRF_OK 0
#define DCRF_ERR_VALUE (-1)    /* bad argument: shape, range, unsupported size */
#define DCRF_ERR_MEMORY (-2)   /* device or host allocation failed */
#define DCRF_ERR_DEVICE (-3)   /* HIP runtime / kernel failure, no device */
#define DCRF_ERR_OVERFLOW (-4) /* a lattice key coordinate left the int16 range */
#define DCRF_ERR_STATE (-5)    /* call not valid in the model's current state */

/* NormalizationType (pydensecrf/densecrf/include/pairwise.h:32-37) */
#define DCRF_NO_NORMALIZATION 0
#define DCRF_NORMALIZE_BEFORE 1
#define DCRF_NORMALIZE_AFTER 2
#define DCRF_NORMALIZE_SYMMETRIC 3
/* KernelType (pairwise.h:38-42) */
#define DCRF_CONST_KERNEL 0
#define DCRF_DIAG_KERNEL 1
#define DCRF_FULL_KERNEL 2
/* LabelCompatibility kind (pydensecrf/densecrf.pyx:11-20: number / 1-D / 2-D) */
#define DCRF_COMPAT_POTTS 0    /* compat[1]: mu(a,b) = -w [a==b]  (labelcompatibility.cpp:46-48) */
#define DCRF_COMPAT_DIAGONAL 1 /* compat[L]: out = diag(w) Q          (labelcompatibility.cpp:66-69) */
#define DCRF_COMPAT_MATRIX 2   /* compat[L*L] row-major m; W = (m+m^T)/2 (labelcompatibility.cpp:79-87) */

/* Largest supported feature dimension, label count and pairwise-term count. */
#define DCRF_MAX_DIM 15
#define DCRF_MAX_LABELS 1024
#define DCRF_MAX_TERMS 8

typedef struct dcrf_model dcrf_model;

/* Message of the last failed call on this thread ("" if none). */
const char *dcrf_last_error(void);
/* Number of visible HIP devices (0 without a GPU); never fails. */
int dcrf_device_count(void);
/* No counterpart in the reference: the library keeps freed device buffers for reuse
 * (up to 4 GiB per process) and one page-locked host staging area for large transfers
 * (as large as the largest one); this returns both to the runtime.  bytes may be NULL. */
int dcrf_release_cached_memory(size_t *bytes);

/* ---- construction --------------------------------------------------------------- */
/* DenseCRF(N, M): densecrf.cpp:38-39, densecrf.pyx:53-64.  Touches no device. */
int dcrf_create(int N, int L, dcrf_model **out);
/* DenseCRF2D(W, H, M): densecrf.cpp:47-48, densecrf.pyx:107-118. */
int dcrf_create_2d(int W, int H, int L, dcrf_model **out);
/* Batched form: n instances of N[i] variables each, one set of terms applied to every
 * instance, lattices concatenated with per-instance offset tables.  No reference
 * counterpart as one call: equivalent to n independent DenseCRF(N[i], L) models. */
int dcrf_create_batch(int n, const int *N, int L, dcrf_model **out);
/* Batched DenseCRF2D: instance i is W[i] x H[i]. */
int dcrf_create_batch_2d(int n, const int *W, const int *H, int L, dcrf_model **out);
/* ~DenseCRF: densecrf.cpp:40-45. */
void dcrf_destroy(dcrf_model *m);

/* ---- unary ---------------------------------------------------------------------- */
/* setUnaryEnergy(u): densecrf.cpp:90-92 (ConstUnaryEnergy, unary.cpp:42-46),
 * densecrf.pyx:80-89.  U is (L, N) row-major (N = sum of instance sizes). */
int dcrf_set_unary(dcrf_model *m, const float *U_LxN);
/* Batched form: the unary of instance i, (L, N[i]) row-major. */
int dcrf_set_unary_instance(dcrf_model *m, int i, const float *U_LxNi);

/* ---- pairwise terms (lattice built on the GPU at call time) -------------------- */
/* addPairwiseEnergy(features, compat, kernel, normalization): densecrf.cpp:54-57,
 * pairwise.cpp:116-122 + 40-62, densecrf.pyx:71-75.  feat is (d, N) row-major. */
int dcrf_add_pairwise(dcrf_model *m, const float *feat_dxN, int d, int compat_kind,
                      const float *compat, int ktype, int ntype);
/* addPairwiseGaussian(sx, sy, ...): densecrf.cpp:61-69, densecrf.pyx:120-124.
 * Features (x/sx, y/sy) are generated on the device (2-D models only). */
int dcrf_add_gaussian_2d(dcrf_model *m, float sx, float sy, int compat_kind,
                         const float *compat, int ktype, int ntype);
/* addPairwiseBilateral(sx, sy, sr, sg, sb, im, ...): densecrf.cpp:70-81,
 * densecrf.pyx:126-140.  rgb is (H, W, 3) uint8; for a batch, the images of all
 * instances concatenated in instance order. */
int dcrf_add_bilateral_2d(dcrf_model *m, float sx, float sy, float sr, float sg, float sb,
                          const uint8_t *rgb_hwc, int compat_kind, const float *compat,
                          int ktype, int ntype);

/* ---- inference (host buffers) -------------------------------------------------- */
/* inference(n): densecrf.cpp:115-131, densecrf.pyx:91-92.  Q out [N][L]. */
int dcrf_inference(dcrf_model *m, int n_iterations, float *Q);
/* startInference(): densecrf.cpp:178-186, densecrf.pyx:94-95.  Q out [N][L]. */
int dcrf_start_inference(dcrf_model *m, float *Q);
/* stepInference(Q, tmp1, tmp2): densecrf.cpp:187-201, densecrf.pyx:97-98.  Q is
 * updated in place; tmp1 receives -U - sum_k pairwise_k(Q), tmp2 the last term's
 * message (either may be NULL).  All [N][L]. */
int dcrf_step_inference(dcrf_model *m, float *Q, float *tmp1, float *tmp2);
/* klDivergence(Q): densecrf.cpp:214-235, densecrf.pyx:100-101. */
int dcrf_kl_divergence(dcrf_model *m, const float *Q, double *kl);
/* DenseCRF::map(n) (densecrf.cpp:132-137, densecrf.h:75; the reference's C++ interface,
 * not its Python binding): n iterations of inference, then labels[p] = the first label
 * of the largest Q[p][l] (currentMap).  labels out [N]. */
int dcrf_map(dcrf_model *m, int n_iterations, int16_t *labels);
/* DenseCRF::currentMap(Q) (densecrf.cpp:202-211, densecrf.h:80): labels[p] = the first
 * label of the largest Q[p][l], Q [N][L]; labels out [N]. */
int dcrf_current_map(dcrf_model *m, const float *Q, int16_t *labels);
/* currentMap on device buffers: Q_dev [N][L] float32 -> labels_dev [N] int16, on stream. */
int dcrf_current_map_device(dcrf_model *m, const float *Q_dev, int16_t *labels_dev, void *stream);
/* PairwisePotential::apply / applyTranspose of one term (pairwise.cpp:173-183):
 * out = compat(K~ Q) or its transpose.  Q, out [N][L]. */
int dcrf_apply_pairwise(dcrf_model *m, int term, const float *Q, float *out, int transpose);

/* ---- inference (device buffers, stream-ordered) ------------------------------- */
/* inference(n) on device buffers: U_dev [N][L] (NULL = the model's unary), Q_dev out
 * [N][L].  Enqueued on `stream` (a hipStream_t; NULL is the null stream).  Returns
 * once the work is enqueued; the model's internal buffers are in use until it ends. */
int dcrf_inference_device(dcrf_model *m, int n_iterations, const float *U_dev, float *Q_dev,
                          void *stream);

/* Record inference(n) on device buffers as a hipGraph (kernels only: no allocation,
 * copy or memset inside) for repeated launches.  U_dev NULL = the model's unary.  The
 * graph reads and writes the given pointers and the model's buffers; adding a term or
 * changing the splat chunk invalidates it.  No reference counterpart (the reference
 * re-runs DenseCRF::inference, densecrf.cpp:115-131). */
int dcrf_capture_inference(dcrf_model *m, int n_iterations, const float *U_dev, float *Q_dev);
/* Launch the captured graph on `stream` (NULL is the null stream). */
int dcrf_launch_graph(dcrf_model *m, void *stream);

/* dcrf_inference_device without any device allocation, for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The filter's vertex rows live in one
 * static device array of 64 MiB per GPU (held until the process ends) instead of the
 * model's buffers; the model must have run dcrf_inference_device (or dcrf_inference)
 * once, so that its layout and unary are on the device, and its vertex rows must fit
 * the array (DCRF_ERR_VALUE otherwise).  Every call on one GPU shares that array, which
 * keeps the last call's contents: calls that overlap on one GPU race.  Same kernels and
 * results as dcrf_inference_device. */
int dcrf_inference_device_noalloc(dcrf_model *m, int n_iterations, const float *U_dev, float *Q_dev,
                                  void *stream);

/* utils.unary_from_softmax (pydensecrf/utils.py:55-79) on device buffers: sm [L][N]
 * (class first, as the reference takes it) -> U [N][L] (the layout of
 * dcrf_inference_device); scale <= 0: none, clip <= 0: none.  No model needed. */
int dcrf_unary_from_softmax_device(const float *sm, int L, int N, double scale, double clip, float *U,
                                   void *stream);
/* utils.unary_from_labels (pydensecrf/utils.py:6-45) on device buffers: labels [N] int32
 * -> U [N][L]; with zero_unsure label 0 means unsure and label l means class l-1 (a
 * label past the classes gets the off-label energy everywhere, where numpy raises). */
int dcrf_unary_from_labels_device(const int32_t *labels, int N, int L, double gt_prob, int zero_unsure, float *U,
                                  void *stream);
/* DenseCRF::unaryEnergy / pairwiseEnergy (densecrf.cpp:141-177), the per-variable
 * energies of a labelling: labels [N] int16 (outside [0, L): 0); out [N].  The pairwise
 * energy of term k is -0.5 (C_k K_k onehot(labels))(l_i, i) (times the term weight when
 * one is set); term -1 sums all terms. */
int dcrf_unary_energy(dcrf_model *m, const int16_t *labels, float *out);
int dcrf_pairwise_energy(dcrf_model *m, const int16_t *labels, int term, float *out);

/* ---- parameters and gradients ------------------------------------------------- */
/* Replace the compatibility of one term; params in the raw layout of dcrf_add_pairwise
 * (Potts [1], Diagonal [L], Matrix [L*L] row-major, symmetrised as (m+m^T)/2).  The
 * reference's setLabelCompatibilityParameters (densecrf.cpp:329-341) for one term. */
int dcrf_set_compat(dcrf_model *m, int term, const float *params);
/* Per-term weights w_k multiplying each pairwise message, E = U + sum_k C_k(w_k K~_k Q)
 * ([num_terms] host floats; NULL restores 1).  No reference counterpart: the
 * reference's terms have unit weight; w_k = 1 reproduces it exactly. */
int dcrf_set_kernel_weights(dcrf_model *m, const float *w);
/* labelCompatibilityParameters (densecrf.cpp:314-328): count and packed values (Potts 1,
 * Diagonal L, Matrix upper triangle W(i, j>=i) row by row). */
int dcrf_num_compat_params(const dcrf_model *m);
int dcrf_get_compat_params(const dcrf_model *m, float *params);
/* setLabelCompatibilityParameters (densecrf.cpp:329-341): all terms from the packed
 * vector of dcrf_get_compat_params (a Matrix entry sets w(i,j) and w(j,i)). */
int dcrf_set_compat_params(dcrf_model *m, const float *params);
/* LogisticUnaryEnergy (unary.cpp:48-68; DenseCRF::setUnaryEnergy(L, f), densecrf.cpp:92-94):
 * U = Lmat feat with Lmat (num_labels, F) and feat (F, N), both row-major; the product
 * runs on the device (matrix cores) and is redone when the parameters change.  Its
 * parameters are Lmat packed column-major (unaryParameters / setUnaryParameters,
 * densecrf.cpp:298-306, unary.cpp:53-62); a plain unary (dcrf_set_unary) has none.  After dcrf_gradient,
 * dcrf_unary_param_gradient returns the reference's unary_grad for this unary,
 * sum over iterations of b feat^T, packed like the parameters. */
int dcrf_set_logistic_unary(dcrf_model *m, const float *Lmat, int F, const float *feat);
int dcrf_num_unary_params(const dcrf_model *m);
int dcrf_get_unary_params(const dcrf_model *m, float *params);
int dcrf_set_unary_params(dcrf_model *m, const float *params);
int dcrf_unary_param_gradient(const dcrf_model *m, float *grad);
/* kernelParameters / setKernelParameters (densecrf.cpp:334-360, pairwise.cpp:129-151):
 * per term none (CONST_KERNEL), d scales (DIAG_KERNEL, ones at first) or a d x d matrix
 * stored column-major (FULL_KERNEL, the identity at first), concatenated over terms.
 * Setting them rebuilds each DIAG/FULL term's lattice and normaliser from diag(p) f or
 * P f, f the features the term was added with. */
int dcrf_num_kernel_params(const dcrf_model *m);
int dcrf_get_kernel_params(const dcrf_model *m, float *params);
int dcrf_set_kernel_params(dcrf_model *m, const float *params);
/* DenseCRF::gradient with the LogLikelihood objective (densecrf.cpp:238-297,
 * objective.cpp:37-50).  gt [N] int16 (labels outside [0, L) ignored).  Outputs (any
 * may be NULL): objective = mean log-likelihood; unary_grad [N][L] = d objective / dU
 * (the unary gradient an identity-parameterised unary would report); compat_grad
 * packed like dcrf_get_compat_params; kweight_grad [num_terms] = d objective / d w_k
 * (no counterpart in the reference: its terms have no weight); kernel_grad packed like
 * dcrf_get_kernel_params (the reference's kernel_grad, pairwise.cpp:152-163, 202-207).
 * Q_t and the messages K~_k Q_t are kept from the forward pass (no re-filtering). */
int dcrf_gradient(dcrf_model *m, int n_iterations, const int16_t *gt, float robust, double *objective,
                  float *unary_grad, float *compat_grad, float *kweight_grad, float *kernel_grad);
/* The same with the reference's other learning objectives (objective.h:34-75,
 * objective.cpp:37-108): DCRF_OBJECTIVE_LOG_LIKELIHOOD (param = robust),
 * DCRF_OBJECTIVE_HAMMING (class_weight[n_class_weight], or NULL and param =
 * class_weight_pow: weights from the label counts of gt), DCRF_OBJECTIVE_IOU (mean
 * intersection over union of the soft labelling).  objective = the objective's value.
 * The objective and dL/dQ are evaluated on the device (gt is uploaded; Q_n stays in
 * device memory); sums over pixels are double in a fixed order. */
#define DCRF_OBJECTIVE_LOG_LIKELIHOOD 0
#define DCRF_OBJECTIVE_HAMMING 1
#define DCRF_OBJECTIVE_IOU 2
int dcrf_gradient_objective(dcrf_model *m, int n_iterations, int objective_kind, const int16_t *gt, float param,
                            const float *class_weight, int n_class_weight, double *objective, float *unary_grad,
                            float *compat_grad, float *kweight_grad, float *kernel_grad);
/* Training forward on device buffers: like dcrf_inference_device, but keeps Q_0..Q_n
 * and the messages for dcrf_backward_device.  compat_dev: raw compatibility parameters
 * of all terms concatenated (NULL = the model's); kw_dev: [num_terms] term weights (NULL
 * = the model's); both must stay valid until the backward.  Q_dev may be NULL. */
int dcrf_forward_device(dcrf_model *m, int n_iterations, const float *U_dev, const float *compat_dev,
                        const float *kw_dev, float *Q_dev, void *stream);
/* Adjoint of the last dcrf_forward_device: dQ_dev = dL/dQ_n [N][L].  Writes dU_dev
 * [N][L], dcompat_dev (raw layout, may be NULL), dkw_dev ([num_terms], may be NULL) and
 * dkernel_dev (packed like dcrf_get_kernel_params, may be NULL: the reference's
 * kernel_grad recursion, densecrf.cpp:280-284, which costs two extra filters and four
 * lattice-gradient passes per term and iteration). */
int dcrf_backward_device(dcrf_model *m, const float *dQ_dev, float *dU_dev, float *dcompat_dev, float *dkw_dev,
                         float *dkernel_dev, void *stream);

/* ---- introspection (parity tests, metrics) ------------------------------------- */
int dcrf_num_variables(const dcrf_model *m);
int dcrf_num_labels(const dcrf_model *m);
int dcrf_num_terms(const dcrf_model *m);
int dcrf_num_instances(const dcrf_model *m);
/* Dimension and vertex count of a term's lattice (Permutohedral d_, M_:
 * permutohedral.h:53-55); for a batch, V is summed over instances. */
int dcrf_term_info(const dcrf_model *m, int term, int *d, int *V);
/* Copy a term's lattice to host: offset/rank [N][d+1] int32, bary [N][d+1] float
 * (Permutohedral offset_/rank_/barycentric_ for the real points, permutohedral.cpp:
 * 272-274), nbr [(d+1)][V][2] int32 with -1 for an absent neighbour (blur_neighbors_,
 * :305-318), keys [V][d] int16 (the lattice key of each vertex), norm [N] (DenseKernel
 * norm_, pairwise.cpp:40-62).  Any pointer may be NULL.  Vertex ids are per instance,
 * in first-occurrence order like the reference's hash table. */
int dcrf_get_lattice(const dcrf_model *m, int term, int instance, int *offset, int *rank,
                     float *bary, int *nbr, int16_t *keys, float *norm);
/* Vertex count of one instance of a term. */
int dcrf_term_instance_vertices(const dcrf_model *m, int term, int instance, int *V);
/* Set the chunk length of the splat's segmented reduction (0 = the default, 32):
 * vertices with more contributors are summed in pieces that are recombined in order. */
int dcrf_set_splat_chunk(dcrf_model *m, int chunk);
/* Side T of the splat tiles (0 = the default, 128; call before adding terms).  Pixels
 * are grouped into tiles (T x T squares of a 2-D image, runs of T*T pixels otherwise);
 * each vertex's contributors are cut at tile boundaries and the splat processes the
 * segments tile by tile, so a tile's input rows are fetched from memory about once.
 * When a tile's rows and entries fit a workgroup's LDS (T*T*L*4 bytes plus 8 per
 * entry, at most 64 KiB) the splat stages them there (one workgroup per tile).  No
 * counterpart in the reference (a scheduling knob). */
int dcrf_set_splat_tile(dcrf_model *m, int side);
/* Inference mean-field kernel: 0 = automatic (the fused kernel over interleaved slice
 * pairs where the term layout has one, else the generic kernel), 1 = generic, 2 =
 * fused, 3 = the fused kernel over pixel runs whose distinct vertex rows are staged in
 * LDS (where a run's rows and pairs fit 64 KiB; measured slower than 2 at 1024x768 with
 * 21 labels).  run_rows (0 = 16) is the height of a run rectangle, whose width is two
 * waves of pixels.  Every choice computes the same arithmetic in the same order.  No
 * counterpart in the reference (a kernel choice for testing and benchmarking). */
int dcrf_set_meanfield_kernel(dcrf_model *m, int mode, int run_rows);
/* No counterpart in the reference (a scheduling choice; results are bit-identical):
 * on (default) blurs two lattice axes per pass through a table of the 9 rows each
 * vertex's two steps read, off runs one pass per axis. */
int dcrf_set_blur_pairs(dcrf_model *m, int on);
/* No counterpart in the reference (a scheduling choice; results are bit-identical):
 * the most blur steps per multi-step pass, 2 (default) or 3 (27 rows per vertex), each
 * pass one launch over all terms that still have passes.  Applies while
 * dcrf_set_blur_pairs is on. */
int dcrf_set_blur_steps(dcrf_model *m, int steps);
/* Inference only, 2-D single-image models with the fused mean-field kernel: keep Q
 * between iterations in 8 x 8 pixel blocks (the splat then gathers the rows of a
 * vertex's contributors from a few contiguous blocks instead of several image rows).
 * Same results, bit for bit; 0 (off) by default.  No reference counterpart. */
int dcrf_set_q_blocks(dcrf_model *m, int on);
/* Blur passes over fewer than below_vertices vertices (all passes when <= 0) with
 * labels_per_lane labels per lane (0: the default for the label count), i.e. more lanes
 * per vertex for small, latency-bound passes.  The blur is elementwise per label: the
 * same bits for any setting.  No reference counterpart. */
int dcrf_set_blur_lanes(dcrf_model *m, int labels_per_lane, int below_vertices);
/* Hash-table load factor of later lattice builds, entries / slots in (0, 0.5]
 * (default 0.5): lower trades memory for shorter probe sequences. */
int dcrf_set_hash_load(dcrf_model *m, float load);
/* Kernel timing: when enabled, the library records HIP events around every launch of
 * the inference kernels on the stream they run on (enabling resets the counters).
 * dcrf_kernel_stats returns, for one kernel class, the number of timed launches, their
 * total milliseconds and the algorithmic bytes they moved.  Classes:
 *   0 splat     N*L*4 + sum_k [N*(d_k+1)*8 + N*4 + V_k*L*4]
 *   1 combine   0 (recombination of split splat segments)
 *   2 blur      per step s: sum over terms with d_k >= s of V_k*(2*L*4 + 8)
 *   3 meanfield 2*N*L*4 + sum_k [N*(d_k+1)*8 + N*4 + V_k*L*4]
 * One iteration's classes add up to dcrf_iteration_bytes. */
int dcrf_enable_timing(dcrf_model *m, int on);
int dcrf_kernel_stats(dcrf_model *m, int kernel, long long *launches, double *ms, double *bytes);
/* Algorithmic bytes of one mean-field iteration (Q/U streams, lattice entry arrays,
 * normalisers, vertex rows, neighbour tables). */
int dcrf_iteration_bytes(const dcrf_model *m, double *bytes);
/* Filter layout of the model (built on first use): whether the LDS-tiled splat runs
 * (else the global form), the tile side, tile count, splat segments, partial rows, and
 * the inference mean-field kernel (0 generic, 1 fused over interleaved slice pairs,
 * 2 fused over LDS-staged pixel runs).  No counterpart in the reference
 * (introspection for tests and benchmarks). */
int dcrf_splat_layout(dcrf_model *m, int *tiled, int *tile_side, int *n_tiles, int *segments, int *partials,
                      int *meanfield);

#ifdef __cplusplus
}
#endif
#endif /* DCRF_HIP_H */
