// Test infrastructure only: a C-callable probe around the reference's own C++
// mean-field code (lucasb-eyer/pydensecrf, pydensecrf/densecrf/src), compiled from
// the reference sources where they lie by oracle/Makefile into oracle/_ref/.
// It lets the tests compare the oracle's restatement (oracle/dcrf_oracle.c) with
// the real reference on identical inputs, and lets bench.py time the reference
// CPU path.  Nothing in the product links or loads this file.
//
// Layout conventions at this boundary (all numpy-friendly):
//   features  (d, N) row-major  -> Eigen MatrixXf d x N (the reference's own copy,
//             pydensecrf/eigen_impl.cpp:21-25)
//   unary     (L, N) row-major  -> Eigen MatrixXf L x N
//   Q / msgs  [N][L] memory (Eigen col-major L x N), i.e. the reference's storage.
#include "densecrf.h"
#include "optimization.h"
#include "permutohedral.h"
#include "pairwise.h"
#include "labelcompatibility.h"
#include "objective.h"
#include "unary.h"
#include <cstring>
#include <memory>
#include <vector>

namespace {

typedef Eigen::Matrix<float, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor> RowMatF;

// Exposes Permutohedral's protected lattice arrays (permutohedral.h:44-55).
struct LatticeProbe : public Permutohedral {
    using Permutohedral::offset_;
    using Permutohedral::rank_;
    using Permutohedral::barycentric_;
    using Permutohedral::blur_neighbors_;
    using Permutohedral::N_;
    using Permutohedral::M_;
    using Permutohedral::d_;
};

// Exposes DenseCRF's protected pairwise list (densecrf.h) so single terms can be
// applied, and keeps N/L for the buffer copies.
struct CRFProbe : public DenseCRF2D {
    CRFProbe(int W, int H, int L) : DenseCRF2D(W, H, L) {}
    using DenseCRF::pairwise_;
    using DenseCRF::N_;
    using DenseCRF::M_;
};

// A unary whose gradient is the identity map (b -> vec(b)), so that
// DenseCRF::gradient (densecrf.cpp:261-262, 293-294) returns dL/dU summed
// over the unrolled iterations.  Uses the virtual API of unary.h:31-40.
struct IdentityUnary : public UnaryEnergy {
    MatrixXf u_;
    explicit IdentityUnary(const MatrixXf& u) : u_(u) {}
    MatrixXf get() const override { return u_; }
    VectorXf gradient(const MatrixXf& b) const override {
        return Eigen::Map<const VectorXf>(b.data(), b.size());
    }
};

MatrixXf from_rowmajor(const float* p, int rows, int cols) {
    return Eigen::Map<const RowMatF>(p, rows, cols);
}

LabelCompatibility* make_compat(int kind, const float* c, int L) {
    if (kind == 0) return new PottsCompatibility(c[0]);
    if (kind == 1) return new DiagonalCompatibility(Eigen::Map<const VectorXf>(c, L));
    return new MatrixCompatibility(from_rowmajor(c, L, L));
}

void copy_out(const MatrixXf& m, float* dst) {
    std::memcpy(dst, m.data(), sizeof(float) * m.size());
}

// An EnergyFunction (optimization.h:31-35) whose value and gradient come from a C
// callback, so that minimizeLBFGS can be driven on any energy from Python.
typedef double (*energy_cb)(const float* x, float* g, int n, void* ctx);
struct CallbackEnergy : public EnergyFunction {
    VectorXf x0_;
    energy_cb cb_;
    void* ctx_;
    CallbackEnergy(const float* x0, int n, energy_cb cb, void* ctx) : x0_(n), cb_(cb), ctx_(ctx) {
        std::memcpy(x0_.data(), x0, sizeof(float) * n);
    }
    VectorXf initialValue() override { return x0_; }
    double gradient(const VectorXf& x, VectorXf& dx) override {
        dx.resize(x.rows());
        return cb_(x.data(), dx.data(), (int)x.rows(), ctx_);
    }
};

}  // namespace

extern "C" {

// minimizeLBFGS (optimization.cpp:68-103) from x0 on the callback's energy; x_out gets
// the result (n floats).
void ref_minimize_lbfgs(int n, const float* x0, energy_cb cb, void* ctx, int restart, float* x_out) {
    CallbackEnergy e(x0, n, cb, ctx);
    VectorXf r = minimizeLBFGS(e, restart, false);
    std::memcpy(x_out, r.data(), sizeof(float) * n);
}

// ---------------- lattice probe ----------------
void* ref_lattice_new(const float* feat_dxN, int d, int N) {
    LatticeProbe* p = new LatticeProbe();
    p->init(from_rowmajor(feat_dxN, d, N));
    return p;
}
void ref_lattice_free(void* h) { delete static_cast<LatticeProbe*>(h); }
int ref_lattice_vertices(void* h) { return static_cast<LatticeProbe*>(h)->M_; }
int ref_lattice_entries(void* h) { return (int)static_cast<LatticeProbe*>(h)->offset_.size(); }
// offset/rank/bary: the full (N+16)*(d+1) arrays; nbr: 2*(d+1)*V ints (n1,n2 axis-major)
void ref_lattice_get(void* h, int* offset, int* rank, float* bary, int* nbr) {
    LatticeProbe* p = static_cast<LatticeProbe*>(h);
    std::memcpy(offset, p->offset_.data(), sizeof(int) * p->offset_.size());
    std::memcpy(rank, p->rank_.data(), sizeof(int) * p->rank_.size());
    std::memcpy(bary, p->barycentric_.data(), sizeof(float) * p->barycentric_.size());
    for (size_t i = 0; i < p->blur_neighbors_.size(); i++) {
        nbr[2 * i] = p->blur_neighbors_[i].n1;
        nbr[2 * i + 1] = p->blur_neighbors_[i].n2;
    }
}
// in/out: [N][L] memory
void ref_lattice_compute(void* h, const float* in, float* out, int L, int reverse) {
    LatticeProbe* p = static_cast<LatticeProbe*>(h);
    MatrixXf mi = Eigen::Map<const MatrixXf>(in, L, p->N_);
    MatrixXf mo;
    p->compute(mo, mi, reverse != 0);
    copy_out(mo, out);
}
// df: [N][d]; a, b: [N][L]
void ref_lattice_gradient(void* h, float* df, const float* a, const float* b, int L) {
    static_cast<LatticeProbe*>(h)->gradient(df, a, b, L);
}

// ---------------- CRF ----------------
void* ref_crf_new(int W, int H, int L) { return new CRFProbe(W, H, L); }
void ref_crf_free(void* h) { delete static_cast<CRFProbe*>(h); }
void ref_crf_set_unary(void* h, const float* U_LxN) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    c->setUnaryEnergy(from_rowmajor(U_LxN, c->M_, c->N_));
}
// Identity-gradient unary (see IdentityUnary) for gradient checks.
void ref_crf_set_unary_identity(void* h, const float* U_LxN) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    c->setUnaryEnergy(new IdentityUnary(from_rowmajor(U_LxN, c->M_, c->N_)));
}
void ref_crf_add_pairwise(void* h, const float* feat_dxN, int d, int ckind, const float* compat,
                          int ktype, int ntype) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    c->addPairwiseEnergy(from_rowmajor(feat_dxN, d, c->N_), make_compat(ckind, compat, c->M_),
                         (KernelType)ktype, (NormalizationType)ntype);
}
void ref_crf_add_gaussian(void* h, float sx, float sy, int ckind, const float* compat, int ktype,
                          int ntype) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    c->addPairwiseGaussian(sx, sy, make_compat(ckind, compat, c->M_), (KernelType)ktype,
                           (NormalizationType)ntype);
}
void ref_crf_add_bilateral(void* h, float sx, float sy, float sr, float sg, float sb,
                           const unsigned char* im, int ckind, const float* compat, int ktype,
                           int ntype) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    c->addPairwiseBilateral(sx, sy, sr, sg, sb, im, make_compat(ckind, compat, c->M_),
                            (KernelType)ktype, (NormalizationType)ntype);
}
void ref_crf_inference(void* h, int n, float* Q) { copy_out(static_cast<CRFProbe*>(h)->inference(n), Q); }
void ref_crf_start(void* h, float* Q) { copy_out(static_cast<CRFProbe*>(h)->startInference(), Q); }
// Q in/out; tmp1/tmp2 out.  All [N][L].
void ref_crf_step(void* h, float* Q, float* tmp1, float* tmp2) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    MatrixXf q = Eigen::Map<const MatrixXf>(Q, c->M_, c->N_), t1, t2;
    c->stepInference(q, t1, t2);
    copy_out(q, Q);
    copy_out(t1, tmp1);
    copy_out(t2, tmp2);
}
double ref_crf_kl(void* h, const float* Q) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    return c->klDivergence(Eigen::Map<const MatrixXf>(Q, c->M_, c->N_));
}
// PairwisePotential::apply / applyTranspose of one term (pairwise.cpp:173-183).
void ref_crf_apply(void* h, int term, const float* Q, float* out, int transpose) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    MatrixXf q = Eigen::Map<const MatrixXf>(Q, c->M_, c->N_), o;
    if (transpose) c->pairwise_[term]->applyTranspose(o, q);
    else c->pairwise_[term]->apply(o, q);
    copy_out(o, out);
}
int ref_crf_num_compat_params(void* h) { return (int)static_cast<CRFProbe*>(h)->labelCompatibilityParameters().size(); }
int ref_crf_num_kernel_params(void* h) { return (int)static_cast<CRFProbe*>(h)->kernelParameters().size(); }
void ref_crf_get_compat_params(void* h, float* p) {
    VectorXf v = static_cast<CRFProbe*>(h)->labelCompatibilityParameters();
    std::memcpy(p, v.data(), sizeof(float) * v.size());
}
void ref_crf_set_compat_params(void* h, const float* p, int n) {
    static_cast<CRFProbe*>(h)->setLabelCompatibilityParameters(Eigen::Map<const VectorXf>(p, n));
}
void ref_crf_get_kernel_params(void* h, float* p) {
    VectorXf v = static_cast<CRFProbe*>(h)->kernelParameters();
    std::memcpy(p, v.data(), sizeof(float) * v.size());
}
void ref_crf_set_kernel_params(void* h, const float* p, int n) {
    static_cast<CRFProbe*>(h)->setKernelParameters(Eigen::Map<const VectorXf>(p, n));
}
// DenseCRF::gradient with the LogLikelihood objective (densecrf.cpp:238-297,
// objective.cpp:37-50).  gt: int16 labels [N].  Any gradient pointer may be null.
double ref_crf_gradient(void* h, int n_iter, const short* gt, float robust, float* unary_grad,
                        float* compat_grad, float* kernel_grad) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    VectorXs g = Eigen::Map<const VectorXs>(gt, c->N_);
    LogLikelihood obj(g, robust);
    VectorXf ug, cg, kg;
    double r = c->gradient(n_iter, obj, unary_grad ? &ug : nullptr, compat_grad ? &cg : nullptr,
                           kernel_grad ? &kg : nullptr);
    if (unary_grad) std::memcpy(unary_grad, ug.data(), sizeof(float) * ug.size());
    if (compat_grad) std::memcpy(compat_grad, cg.data(), sizeof(float) * cg.size());
    if (kernel_grad) std::memcpy(kernel_grad, kg.data(), sizeof(float) * kg.size());
    return r;
}

// DenseCRF::setUnaryEnergy(L, f) (densecrf.cpp:92-94): LogisticUnaryEnergy with
// L (M x F) and f (F x N), both row-major.
void ref_crf_set_unary_logistic(void* h, const float* Lm, int F, const float* f) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    c->setUnaryEnergy(from_rowmajor(Lm, c->M_, F), from_rowmajor(f, F, c->N_));
}
int ref_crf_num_unary_params(void* h) { return (int)static_cast<CRFProbe*>(h)->unaryParameters().size(); }
void ref_crf_get_unary_params(void* h, float* p) {
    VectorXf v = static_cast<CRFProbe*>(h)->unaryParameters();
    std::memcpy(p, v.data(), sizeof(float) * v.size());
}
void ref_crf_set_unary_params(void* h, const float* p, int n) {
    static_cast<CRFProbe*>(h)->setUnaryParameters(Eigen::Map<const VectorXf>(p, n));
}
// DenseCRF::gradient with objective 0 LogLikelihood(gt, param), 1 Hamming(gt, param =
// class_weight_pow) or Hamming(gt, class_weight) when given, 2 IntersectionOverUnion(gt).
double ref_crf_gradient_obj(void* h, int n_iter, int objective, const short* gt, float param,
                            const float* class_weight, int ncw, float* unary_grad, float* compat_grad,
                            float* kernel_grad) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    VectorXs g = Eigen::Map<const VectorXs>(gt, c->N_);
    std::unique_ptr<ObjectiveFunction> obj;
    if (objective == 0)
        obj.reset(new LogLikelihood(g, param));
    else if (objective == 1 && class_weight)
        obj.reset(new Hamming(g, VectorXf(Eigen::Map<const VectorXf>(class_weight, ncw))));
    else if (objective == 1)
        obj.reset(new Hamming(g, param));
    else
        obj.reset(new IntersectionOverUnion(g));
    VectorXf ug, cg, kg;
    double r = c->gradient(n_iter, *obj, unary_grad ? &ug : nullptr, compat_grad ? &cg : nullptr,
                           kernel_grad ? &kg : nullptr);
    if (unary_grad) std::memcpy(unary_grad, ug.data(), sizeof(float) * ug.size());
    if (compat_grad) std::memcpy(compat_grad, cg.data(), sizeof(float) * cg.size());
    if (kernel_grad) std::memcpy(kernel_grad, kg.data(), sizeof(float) * kg.size());
    return r;
}

// DenseCRF::unaryEnergy / pairwiseEnergy (densecrf.cpp:141-177); labels int16 [N].
void ref_crf_unary_energy(void* h, const short* l, float* out) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    VectorXf r = c->unaryEnergy(Eigen::Map<const VectorXs>(l, c->N_));
    std::memcpy(out, r.data(), sizeof(float) * r.size());
}
void ref_crf_pairwise_energy(void* h, const short* l, int term, float* out) {
    CRFProbe* c = static_cast<CRFProbe*>(h);
    VectorXf r = c->pairwiseEnergy(Eigen::Map<const VectorXs>(l, c->N_), term);
    std::memcpy(out, r.data(), sizeof(float) * r.size());
}

}  // extern "C"
