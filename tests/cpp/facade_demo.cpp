// The reference's caller sequence (experiments/experiment_two.cpp:30-50,
// linear_solve.cpp:257-303) against the C++ facade: 2D Laplace interior
// Dirichlet problem on an ellipse, boundary data from exterior log charges,
// factor + FP64-refined solve on the GPU, field at interior points compared
// with the exact harmonic function.  Prints the errors; exit status 0 when
// the field error is below 1e-8.
#include <cmath>
#include <cstdio>
#include <vector>

#include "kern_interp_amd.hpp"

using namespace kern_interp;

namespace {

struct Ellipse : public Boundary {
  // Trapezoid rule in the parameter t (spectrally accurate for this smooth
  // periodic curve): x = (a cos t, b sin t), outward normal, curvature ab/s^3.
  Ellipse(int n, double a, double b) {
    for (int i = 0; i < n; i++) {
      const double t = 2 * M_PI * i / n;
      const double dx = -a * std::sin(t), dy = b * std::cos(t);
      const double sp = std::sqrt(dx * dx + dy * dy);
      points.push_back(a * std::cos(t));
      points.push_back(b * std::sin(t));
      normals.push_back(dy / sp);
      normals.push_back(-dx / sp);
      weights.push_back(sp * 2 * M_PI / n);
      curvatures.push_back(a * b / (sp * sp * sp));
    }
  }
};

double field(const double* x, const std::vector<double>& c,
             const std::vector<double>& q) {
  double u = 0;
  for (size_t j = 0; j < q.size(); j++)
    u += q[j] * std::log(std::hypot(x[0] - c[2 * j], x[1] - c[2 * j + 1]));
  return u;
}

}  // namespace

int main() {
  const int n = 2048;
  Ellipse boundary(n, 1.0, 0.5);
  const std::vector<double> c = {2.1, 0.4, -1.7, 1.9, 0.3, -2.6};
  const std::vector<double> q = {1.0, -0.6, 0.8};
  boundary.boundary_values = ki_Mat(n, 1);
  for (int i = 0; i < n; i++)
    boundary.boundary_values.set(i, 0, field(&boundary.points[2 * i], c, q));
  std::vector<double> domain_points;
  for (int i = 0; i < 10; i++) {
    const double r = 0.1 + 0.07 * i, t = 0.6 * i;
    domain_points.push_back(r * std::cos(t));
    domain_points.push_back(0.5 * r * std::sin(t));
  }

  QuadTree quadtree;
  quadtree.initialize_tree(&boundary, 1, 2);
  Kernel kernel(1, 2, Kernel::LAPLACE, &boundary, domain_points);

  // direct solve, as the reference does
  ki_Mat direct = boundary_integral_solve(kernel, boundary, &quadtree, 1e-10,
                                          1, domain_points);
  // the same factorization, refined by FP64 GMRES
  SkelFactorization fac(1e-10, 1);
  fac.skeletonize(kernel, &quadtree);
  ki_Mat mu;
  const int it = fac.refined_solve(quadtree, &mu, boundary.boundary_values);
  ki_Mat refined = kernel.apply_forward(mu);

  double e_direct = 0, e_refined = 0;
  for (int i = 0; i < 10; i++) {
    const double u = field(&domain_points[2 * i], c, q);
    e_direct = std::max(e_direct, std::fabs(direct.get(i, 0) - u));
    e_refined = std::max(e_refined, std::fabs(refined.get(i, 0) - u));
  }
  // ki_Mat arithmetic on the device: (K mu) from the kernel block and a
  // product, against the boundary data
  std::vector<int> all(n);
  for (int i = 0; i < n; i++) all[i] = i;
  ki_Mat K = kernel(all, all);
  ki_Mat r = K * mu - boundary.boundary_values;
  const double rel = r.vec_two_norm() / boundary.boundary_values.vec_two_norm();
  std::printf("levels %d gmres_iters %d field_err_direct %.3e "
              "field_err_refined %.3e dense_residual %.3e\n",
              quadtree.num_levels(), it, e_direct, e_refined, rel);
  return (e_direct < 1e-8 && e_refined < 1e-8 && rel < 1e-10) ? 0 : 1;
}
