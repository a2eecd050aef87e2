// kern_interp_amd.hpp -- C++ facade of the reference's solve-path classes on
// top of the MI355X C-ABI (ki_abi.h).  Header-only; link with libki_hip.so.
//
// Same namespace, names and argument meaning as the reference, so that the
// callers of the solve path (boundary_integral_solve and the experiment
// drivers) compile against it unchanged:
//   ki_Mat             kern_interp/ki_mat.h:19-103 (column-major doubles)
//   Boundary           kern_interp/boundaries/boundary.h:47-64 (the data
//                      members the solver reads)
//   Kernel             kern_interp/kernel/kernel.h:17-110
//   QuadTree           kern_interp/quadtree/quadtree.h:111-179
//   SkelFactorization  kern_interp/skel_factorization/skel_factorization.h:
//                      20-60
//   boundary_integral_solve  kern_interp/linear_solve.h:17-19
// The factorization lives on the GPU in the context that the Kernel owns;
// the QuadTree binds to it when it is skeletonized.  Failures throw
// std::runtime_error with the library's message (the reference asserts).
#ifndef KERN_INTERP_AMD_HPP_
#define KERN_INTERP_AMD_HPP_

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ki_abi.h"

namespace kern_interp {

inline void ki_check(int rc) {
  if (rc != KI_OK)
    throw std::runtime_error(std::string("kern_interp_amd: ") +
                             ki_last_error());
}

// Shared device context (geometry, tree, factorization).
class KiContext {
 public:
  explicit KiContext(int device = 0) { ki_check(ki_ctx_create(device, &h_)); }
  ~KiContext() { ki_ctx_destroy(h_); }
  KiContext(const KiContext&) = delete;
  KiContext& operator=(const KiContext&) = delete;
  ki_ctx* get() const { return h_; }
  // A context for matrix arithmetic that has no Kernel (device 0).
  static KiContext& scratch() {
    static KiContext c(0);
    return c;
  }

 private:
  ki_ctx* h_ = nullptr;
};

struct ki_Mat {
  std::vector<double> mat;  // column-major, lda = height
  int lda_ = 0, height_ = 0, width_ = 0;

  ki_Mat() {}
  ki_Mat(int h, int w) : mat((size_t)h * w, 0.0), lda_(h), height_(h),
                         width_(w) {}

  double get(int i, int j) const { return mat[i + (size_t)lda_ * j]; }
  void set(int i, int j, double a) { mat[i + (size_t)lda_ * j] = a; }
  void addset(int i, int j, double a) { mat[i + (size_t)lda_ * j] += a; }
  int height() const { return height_; }
  int width() const { return width_; }
  double* data() { return mat.data(); }
  const double* data() const { return mat.data(); }

  void eye(int n) {
    *this = ki_Mat(n, n);
    for (int i = 0; i < n; i++) set(i, i, 1.0);
  }
  ki_Mat transpose() const {
    ki_Mat t(width_, height_);
    for (int j = 0; j < width_; j++)
      for (int i = 0; i < height_; i++) t.set(j, i, get(i, j));
    return t;
  }
  ki_Mat operator()(const std::vector<int>& I, const std::vector<int>& J)
      const {
    ki_Mat s(I.size(), J.size());
    for (size_t j = 0; j < J.size(); j++)
      for (size_t i = 0; i < I.size(); i++) s.set(i, j, get(I[i], J[j]));
    return s;
  }
  ki_Mat operator()(int r0, int r1, int c0, int c1) const {
    ki_Mat s(r1 - r0, c1 - c0);
    for (int j = c0; j < c1; j++)
      for (int i = r0; i < r1; i++) s.set(i - r0, j - c0, get(i, j));
    return s;
  }
  void set_submatrix(const std::vector<int>& I, const std::vector<int>& J,
                     const ki_Mat& A, bool transpose_A = false) {
    for (size_t j = 0; j < J.size(); j++)
      for (size_t i = 0; i < I.size(); i++)
        set(I[i], J[j], transpose_A ? A.get(j, i) : A.get(i, j));
  }

  ki_Mat& operator+=(const ki_Mat& o) { return axpy(1.0, o); }
  ki_Mat& operator-=(const ki_Mat& o) { return axpy(-1.0, o); }
  ki_Mat& operator*=(double a) {
    for (double& v : mat) v *= a;
    return *this;
  }
  ki_Mat& operator/=(double a) { return *this *= 1.0 / a; }
  ki_Mat operator-() const { return *this * -1.0; }
  ki_Mat operator+(const ki_Mat& o) const { return ki_Mat(*this) += o; }
  ki_Mat operator-(const ki_Mat& o) const { return ki_Mat(*this) -= o; }
  ki_Mat operator*(double a) const { return ki_Mat(*this) *= a; }
  ki_Mat operator/(double a) const { return ki_Mat(*this) /= a; }

  // Matrix product on the device matrix cores (ki_mat.cpp:328-335).
  ki_Mat operator*(const ki_Mat& o) const {
    if (width_ != o.height_) throw std::runtime_error("ki_Mat: shape");
    ki_Mat c(height_, o.width_);
    if (height_ == 0 || o.width_ == 0) return c;
    ki_check(ki_gemm(KiContext::scratch().get(), 64, 0, 0, height_, o.width_,
                     width_, 1.0, data(), std::max(lda_, 1), o.data(),
                     std::max(o.lda_, 1), 0.0, c.data(),
                     std::max(c.lda_, 1)));
    return c;
  }

  double vec_two_norm() const {
    double s = 0;
    for (double v : mat) s += v * v;
    return std::sqrt(s);
  }
  double frob_norm() const { return vec_two_norm(); }
  double max_entry_magnitude() const {
    double m = 0;
    for (double v : mat) m = std::max(m, std::fabs(v));
    return m;
  }

  // LU with partial pivoting on the device (ki_mat.cpp:445-453); piv is
  // 1-based like LAPACK.
  void LU_factorize(ki_Mat* K_LU, std::vector<int>* piv) const {
    *K_LU = ki_Mat(height_, width_);
    piv->assign(height_, 0);
    ki_check(ki_lu_factorize(KiContext::scratch().get(), data(), height_,
                             K_LU->data(), piv->data()));
  }
  // U <- K^{-1} U with the factors in *this (ki_mat.cpp:487-499).
  void left_multiply_inverse(const ki_Mat& K, const std::vector<int>& piv,
                             ki_Mat* U) const {
    (void)K;
    ki_check(ki_lu_solve(KiContext::scratch().get(), data(), piv.data(),
                         height_, U->data(), U->width(), 0));
  }

 private:
  ki_Mat& axpy(double a, const ki_Mat& o) {
    if (o.height_ != height_ || o.width_ != width_)
      throw std::runtime_error("ki_Mat: shape");
    for (size_t e = 0; e < mat.size(); e++) mat[e] += a * o.mat[e];
    return *this;
  }
};

// The data members of the reference's Boundary that the solver reads;
// callers derive their geometries from it.
class Boundary {
 public:
  std::vector<double> points, normals, curvatures, weights;
  ki_Mat boundary_values;
  virtual ~Boundary() {}
};

struct Kernel {
  enum Pde { LAPLACE, LAPLACE_NEUMANN, GAUSS, STOKES };

  int solution_dimension = 0, domain_dimension = 0;
  Kernel::Pde pde = LAPLACE;
  std::vector<double> domain_points;
  std::shared_ptr<KiContext> ctx;

  Kernel() {}
  Kernel(int solution_dimension_, int domain_dimension_, Kernel::Pde pde_,
         Boundary* boundary, std::vector<double> domain_points_,
         int device = 0)
      : solution_dimension(solution_dimension_),
        domain_dimension(domain_dimension_),
        pde(pde_),
        domain_points(std::move(domain_points_)),
        ctx(std::make_shared<KiContext>(device)) {
    if (pde != LAPLACE && pde != STOKES)
      throw std::runtime_error("kern_interp_amd: LAPLACE or STOKES only");
    update_data(boundary);
  }

  // kernel.cpp:24-29; the 3D row-sum diagonals are recomputed on the GPU.
  void update_data(Boundary* boundary) {
    const int n = boundary->weights.size();
    ki_check(ki_set_geometry(
        ctx->get(), domain_dimension, pde, boundary->points.data(),
        boundary->normals.data(), boundary->weights.data(),
        boundary->curvatures.empty() ? nullptr : boundary->curvatures.data(),
        n));
    if (domain_dimension == 3) ki_check(ki_compute_diag3d(ctx->get()));
  }
  void compute_diag_entries_3dlaplace(Boundary*) {
    ki_check(ki_compute_diag3d(ctx->get()));
  }
  void compute_diag_entries_3dstokes(Boundary*) {
    ki_check(ki_compute_diag3d(ctx->get()));
  }

  // K(tgt, src) (kernel.cpp:304-429).
  ki_Mat operator()(const std::vector<int>& tgt_inds,
                    const std::vector<int>& src_inds, bool forward = false,
                    bool parallel = false) const {
    (void)forward;
    (void)parallel;
    ki_Mat out(tgt_inds.size(), src_inds.size());
    ki_check(ki_kernel_block(ctx->get(), tgt_inds.data(), tgt_inds.size(),
                             src_inds.data(), src_inds.size(), out.data()));
    return out;
  }

  // Field at the domain points (kernel.cpp:809-825).
  ki_Mat apply_forward(const ki_Mat& mu) const {
    const int nd = domain_points.size() / domain_dimension;
    ki_Mat out(nd * solution_dimension, 1);
    ki_check(ki_apply_forward(ctx->get(), domain_points.data(), nd, mu.data(),
                              out.data()));
    return out;
  }
};

class QuadTree {
 public:
  int solution_dimension = 0, domain_dimension = 0;
  int max_leaf_dofs = 64;  // MAX_LEAF_DOFS (quadtree.h:11)
  std::shared_ptr<KiContext> ctx;

  // quadtree.cpp:240-287.  The tree is built in the Kernel's context when
  // the factorization binds them (SkelFactorization::skeletonize).
  void initialize_tree(Boundary* boundary, int solution_dimension_,
                       int domain_dimension_) {
    (void)boundary;
    solution_dimension = solution_dimension_;
    domain_dimension = domain_dimension_;
  }
  int num_levels() const { return ctx ? ki_tree_num_levels(ctx->get()) : 0; }
  int level_size(int l) const {
    return ctx ? ki_tree_level_size(ctx->get(), l) : 0;
  }
};

class SkelFactorization {
 public:
  double id_tol = 1e-6;
  int fact_threads = 1;  // accepted; the GPU path does not use it
  bool using_hif = false;
  ki_params params;

  SkelFactorization() { ki_params_default(&params); }
  SkelFactorization(double id_tol_, int fact_threads_, bool using_hif_ = false)
      : id_tol(id_tol_), fact_threads(fact_threads_), using_hif(using_hif_) {
    if (using_hif) throw std::runtime_error("kern_interp_amd: no HIF levels");
    ki_params_default(&params);
    params.id_tol = id_tol;
  }

  // skel_factorization.cpp:173-310.
  void skeletonize(const Kernel& K, QuadTree* tree) {
    tree->ctx = K.ctx;
    ki_check(ki_build_tree(K.ctx->get(), tree->max_leaf_dofs));
    params.id_tol = id_tol;
    ki_check(ki_skeletonize(K.ctx->get(), &params));
  }

  // x = A^{-1} b with the factorization (skel_factorization.cpp:595-778).
  void solve(const QuadTree& quadtree, ki_Mat* x, const ki_Mat& b) const {
    *x = ki_Mat(b.height(), 1);
    ki_check(ki_solve(quadtree.ctx->get(), b.data(), x->data()));
  }

  // FP64 GMRES preconditioned by the factorization to relative residual
  // `tol` (no reference counterpart).  Returns the iteration count.
  int refined_solve(const QuadTree& quadtree, ki_Mat* x, const ki_Mat& b,
                    double tol = 1e-10, int maxit = 50) const {
    *x = ki_Mat(b.height(), 1);
    int it = 0;
    double rel = 0;
    ki_check(ki_gmres(quadtree.ctx->get(), b.data(), x->data(), tol, maxit,
                      30, &it, &rel));
    return it;
  }
};

// linear_solve.cpp:257-303 without the is_in_domain masking: factor, solve
// for the density, evaluate the field at the kernel's domain points.
inline ki_Mat boundary_integral_solve(const Kernel& kernel,
                                      const Boundary& boundary,
                                      QuadTree* quadtree, double id_tol,
                                      int fact_threads,
                                      const std::vector<double>& domain_points) {
  (void)domain_points;
  SkelFactorization fac(id_tol, fact_threads);
  fac.skeletonize(kernel, quadtree);
  ki_Mat mu;
  fac.solve(*quadtree, &mu, boundary.boundary_values);
  return kernel.apply_forward(mu);
}

}  // namespace kern_interp

#endif  // KERN_INTERP_AMD_HPP_
