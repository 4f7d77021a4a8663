// Reference-style programs against the C++ facade (include/bembel_amd/Bembel.hpp), written the
// way Bembel's own examples and tests use the library.  Without arguments beyond the geometry
// file only the host part runs (geometry, ansatz space); with --gpu it also runs:
//   1. examples/LaplaceSingleLayerH2.cpp: harmonic Dirichlet data through DiscreteLinearForm with
//      a lambda, H2 assembly, ConjugateGradient, DiscretePotential, and the rate assert
//      (>= 0.9 (2p+3) over the last two levels);
//   2. tests/test_LazyEigenSum.cpp: H2 vs dense products, lazy sums and scalings, get_dense();
//   3. the dense path of config C1: DiscreteOperator<Eigen::MatrixXd, Laplace> with c^T A c = 4 pi;
//   4. RotatedTangentialTrace with a lambda equals the built-in dipole trace.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "bembel_amd/Bembel.hpp"

using namespace Bembel;

static int failures = 0;
#define EXPECT(cond, ...)              \
  do {                                 \
    if (!(cond)) {                     \
      std::printf("FAILED: " __VA_ARGS__); \
      std::printf("\n");               \
      ++failures;                      \
    }                                  \
  } while (0)

template <typename V>
static double norm(const V& v) {
  double s = 0;
  for (const auto& x : v) s += std::norm(x);
  return std::sqrt(s);
}
template <typename V>
static double diff_norm(const V& a, const V& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += std::norm(a[i] - b[i]);
  return std::sqrt(s);
}

// examples/Error.hpp:102-109: least-squares slope of log2 errors
static double rate(const std::vector<double>& e) {
  const int n = static_cast<int>(e.size());
  double sx = 0, sy = 0, sxx = 0, sxy = 0;
  for (int i = 0; i < n; ++i) {
    const double y = std::log(e[i]) / std::log(2.0);
    sx += i, sy += y, sxx += i * i, sxy += i * y;
  }
  return -(n * sxy - sx * sy) / (n * sxx - sx * sx);
}

static void laplace_example(const Geometry& geometry) {
  Eigen::MatrixXd gridpoints(1000, 3);
  for (int a = 0; a < 10; ++a)
    for (int b = 0; b < 10; ++b)
      for (int c = 0; c < 10; ++c) {
        const int k = (a * 10 + b) * 10 + c;
        gridpoints(k, 0) = -0.25 + 0.5 * a / 9, gridpoints(k, 1) = -0.25 + 0.5 * b / 9;
        gridpoints(k, 2) = -0.25 + 0.5 * c / 9;
      }
  std::function<double(const Vector3d&)> fun = [](const Vector3d& in) {
    return 4 * in[0] * in[0] - 3 * in[1] * in[1] - in[2] * in[2];
  };
  for (int polynomial_degree = 1; polynomial_degree <= 2; ++polynomial_degree) {
    std::vector<double> error;
    for (int refinement_level = 2; refinement_level <= 3; ++refinement_level) {
      AnsatzSpace<LaplaceSingleLayerOperator> ansatz_space(geometry, refinement_level, polynomial_degree);
      DiscreteLinearForm<DirichletTrace<double>, LaplaceSingleLayerOperator> disc_lf(ansatz_space);
      disc_lf.get_linear_form().set_function(fun);
      disc_lf.compute();
      DiscreteOperator<Eigen::H2Matrix<double>, LaplaceSingleLayerOperator> disc_op(ansatz_space);
      disc_op.compute();
      Eigen::ConjugateGradient<Eigen::H2Matrix<double>, Eigen::Lower | Eigen::Upper, Eigen::IdentityPreconditioner>
          cg;
      cg.compute(disc_op.get_discrete_operator());
      auto rho = cg.solve(disc_lf.get_discrete_linear_form());
      DiscretePotential<LaplaceSingleLayerPotential<LaplaceSingleLayerOperator>, LaplaceSingleLayerOperator>
          disc_pot(ansatz_space);
      disc_pot.set_cauchy_data(rho);
      auto pot = disc_pot.evaluate(gridpoints);
      double err = 0;
      for (int i = 0; i < gridpoints.rows(); ++i) {
        Vector3d x{gridpoints(i, 0), gridpoints(i, 1), gridpoints(i, 2)};
        err = std::max(err, std::fabs(pot(i, 0) - fun(x)));
      }
      std::printf("laplace P %d M %d cg iterations %d error %.3e\n", polynomial_degree, refinement_level,
                  cg.iterations(), err);
      error.push_back(err);
    }
    const double r = rate(error);
    std::printf("laplace P %d rate %.3f (needs > %.2f)\n", polynomial_degree, r, 0.9 * (2 * polynomial_degree + 3));
    EXPECT(r > 0.9 * (2 * polynomial_degree + 3), "Laplace rate P=%d", polynomial_degree);
  }
}

static void lazy_sum_test(const Geometry& geometry) {
  using S = std::complex<double>;
  using Op = HelmholtzSingleLayerOperator;
  AnsatzSpace<Op> ansatz_space(geometry, 2, 1);
  DiscreteOperator<Eigen::MatrixXcd, Op> disc_op_dense(ansatz_space);
  disc_op_dense.compute();
  const auto& D = disc_op_dense.get_discrete_operator();
  DiscreteOperator<Eigen::H2Matrix<S>, Op> disc_op_H2(ansatz_space);
  disc_op_H2.compute();
  const auto& H2 = disc_op_H2.get_discrete_operator();
  std::vector<S> a(static_cast<size_t>(H2.cols()));
  for (size_t i = 0; i < a.size(); ++i) a[i] = S(std::sin(1.0 + i), std::cos(3.0 * i));
  using LC = LinearCombination<S>;
  auto test_equal = [&](const LC& m1, const LC& m2, const char* what) {
    const double e = diff_norm(m1 * a, m2 * a);
    EXPECT(e < 1e-8, "lazy sum %s: %.3e", what, e);
  };
  const S b = 3.;
  test_equal(LC(D) + D, H2 + H2, "D+D vs H+H");
  test_equal(LC(D) - D, H2 - H2, "D-D vs H-H");
  test_equal(LC(D) + D + D, H2 + H2 + H2, "D+D+D vs H+H+H");
  test_equal(LC(D) - D + D, H2 - H2 + H2, "D-D+D vs H-H+H");
  test_equal(LC(D) + D, D + H2, "D+D vs D+H");
  test_equal(LC(D) + D, H2 + D, "D+D vs H+D");
  test_equal(LC(D) - D + D, H2 - D + H2, "D-D+D vs H-D+H");
  test_equal(LC(D) + D - D, H2 + H2 - D, "D+D-D vs H+H-D");
  test_equal(LC(D), LC(H2), "D vs H");
  test_equal(b * LC(D), b * H2, "bD vs bH");
  test_equal(b * LC(D) + D, b * H2 + H2, "bD+D vs bH+H");
  // H2Matrix::get_dense
  const auto G = H2.get_dense();
  double num = 0, den = 0;
  for (int64_t j = 0; j < D.cols(); ++j)
    for (int64_t i = 0; i < D.rows(); ++i) num += std::norm(G(i, j) - D(i, j)), den += std::norm(D(i, j));
  std::printf("get_dense vs dense: %.3e\n", std::sqrt(num / den));
  EXPECT(std::sqrt(num / den) < 1e-8, "get_dense");
}

static void dense_c1(const Geometry& geometry) {
  const int M = 3;
  AnsatzSpace<LaplaceSingleLayerOperator> ansatz_space(geometry, M, 1);
  DiscreteOperator<Eigen::MatrixXd, LaplaceSingleLayerOperator> disc_op(ansatz_space);
  disc_op.compute();
  const auto& A = disc_op.get_discrete_operator();
  const double c = std::ldexp(1.0, -M);
  double s = 0;
  for (int64_t j = 0; j < A.cols(); ++j)
    for (int64_t i = 0; i < A.rows(); ++i) s += c * A(i, j) * c;
  std::printf("dense C1: %lld dofs, c^T A c = %.10f (4 pi = %.10f)\n", static_cast<long long>(A.rows()), s, 4 * M_PI);
  EXPECT(A.rows() == 486 && std::fabs(s - 4 * M_PI) < 1e-5 * 4 * M_PI, "dense C1 closed form");
}

static void tangential_trace(const Geometry& geometry) {
  using S = std::complex<double>;
  const double kappa = 2.0;
  const double pos[3] = {0.1, -0.05, 0.08}, mom[3] = {0.02, 0.05, -0.07};
  AnsatzSpace<MaxwellSingleLayerOperator> ansatz_space(geometry, 2, 2);
  // Data::Dipole (examples/Data.hpp:76-90)
  std::function<Vector3<S>(const Vector3d&)> fun = [&](const Vector3d& x) {
    double c[3] = {x[0] - pos[0], x[1] - pos[1], x[2] - pos[2]};
    const double r = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const double n[3] = {c[0] / r, c[1] / r, c[2] / r};
    const S e = std::exp(S(0, -kappa * r));
    const double nl = n[0] * mom[0] + n[1] * mom[1] + n[2] * mom[2];
    const S ec = (1 / (r * r * r) + S(0, kappa / (r * r))) * e;
    Vector3<S> E;
    for (int k = 0; k < 3; ++k) E[k] = kappa * kappa * e / r * (mom[k] - nl * n[k]) + ec * (3 * nl * n[k] - mom[k]);
    return E;
  };
  DiscreteLinearForm<RotatedTangentialTrace<S>, MaxwellSingleLayerOperator> disc_lf(ansatz_space);
  disc_lf.get_linear_form().set_function(fun);
  disc_lf.compute();
  const auto& f = disc_lf.get_discrete_linear_form();
  std::vector<S> g(f.size());
  const double prm[6] = {pos[0], pos[1], pos[2], mom[0], mom[1], mom[2]};
  check(bmi_rhs_assemble(ansatz_space.handle(), 2, prm, 1, kappa, reinterpret_cast<double*>(g.data()), BMI_MEM_HOST,
                         nullptr));
  const double e = diff_norm(f, g) / norm(g);
  std::printf("rotated tangential trace: lambda vs built-in %.3e\n", e);
  EXPECT(e < 1e-12, "tangential trace");
}

int main(int argc, char** argv) {
  const char* geo_file = argc > 1 ? argv[1] : "sphere.dat";
  const bool gpu = argc > 2 && std::strcmp(argv[2], "--gpu") == 0;
  Geometry geometry(geo_file);
  AnsatzSpace<LaplaceSingleLayerOperator> ansatz_space(geometry, 3, 2);
  std::printf("patches %d elements %d dofs %d\n", geometry.get_number_of_patches(),
              ansatz_space.get_number_of_elements(), ansatz_space.get_number_of_dofs());
  if (ansatz_space.get_number_of_dofs() != 6 * (8 + 2) * (8 + 2)) return 1;
  if (!gpu) return 0;
  laplace_example(geometry);
  lazy_sum_test(geometry);
  dense_c1(geometry);
  tangential_trace(geometry);
  std::printf(failures ? "FAILURES %d\n" : "all facade checks passed\n", failures);
  return failures ? 2 : 0;
}
