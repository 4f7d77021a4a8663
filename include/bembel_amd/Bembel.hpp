// Bembel.hpp -- C++ facade with the reference's class names over the bembel_amd C ABI.
//
// A Bembel program that uses the single-layer H2 path switches by replacing its includes
// (<Bembel/AnsatzSpace>, <Bembel/Geometry>, <Bembel/H2Matrix>, <Bembel/Laplace>, ...) with
// this header; the calls keep their shape:
//
//   Bembel::Geometry geometry("sphere.dat");                         // Geometry.hpp:37
//   Bembel::AnsatzSpace<Bembel::HelmholtzSingleLayerOperator>
//       ansatz_space(geometry, refinement_level, polynomial_degree); // AnsatzSpace.hpp:80
//   Bembel::DiscreteOperator<Eigen::H2Matrix<std::complex<double>>,
//                            Bembel::HelmholtzSingleLayerOperator> disc_op(ansatz_space);
//   disc_op.get_linear_operator().set_wavenumber(kappa);
//   disc_op.compute();                                                // DiscreteOperator.hpp:127
//   const auto& H = disc_op.get_discrete_operator();
//   y = H * x;   (Eigen vectors, when Eigen is available)            // H2DenseProduct.hpp
//
// Everything runs on the GPU through libbembel_amd.so; host vectors are copied per product,
// device pointers can be passed to H2Matrix::apply directly.
#pragma once
#include <complex>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../bembel_amd.h"

#if defined(__has_include)
#if __has_include(<Eigen/Core>)
#include <Eigen/Core>
#define BEMBEL_AMD_HAVE_EIGEN 1
#endif
#endif

namespace Bembel {

class Error : public std::runtime_error {
 public:
  Error(bmi_status s, const std::string& m) : std::runtime_error(m), status(s) {}
  bmi_status status;
};

inline void check(bmi_status s) {
  if (s != BMI_OK) throw Error(s, bmi_last_error());
}

// ------------------------------------------------------------------ operators
struct LaplaceSingleLayerOperator {
  using Scalar = double;
  static constexpr int kind = BMI_LAPLACE_SL;
  static constexpr int form = BMI_FORM_DISCONTINUOUS;
  std::complex<double> wavenumber_ = 0.0;
};
struct HelmholtzSingleLayerOperator {
  using Scalar = std::complex<double>;
  static constexpr int kind = BMI_HELMHOLTZ_SL;
  static constexpr int form = BMI_FORM_DISCONTINUOUS;
  std::complex<double> wavenumber_ = 0.0;
  void set_wavenumber(std::complex<double> k) { wavenumber_ = k; }  // SingleLayerOperator.hpp:124
  std::complex<double> get_wavenumber() const { return wavenumber_; }
};
struct MaxwellSingleLayerOperator {
  using Scalar = std::complex<double>;
  static constexpr int kind = BMI_MAXWELL_SL;
  static constexpr int form = BMI_FORM_DIVCONFORMING;
  std::complex<double> wavenumber_ = 0.0;
  void set_wavenumber(std::complex<double> k) { wavenumber_ = k; }  // Maxwell/...:134-137
  std::complex<double> get_wavenumber() const { return wavenumber_; }
};

// ------------------------------------------------------------------ geometry
class Geometry {
 public:
  Geometry() = default;
  explicit Geometry(const std::string& filename, const double* scale = nullptr) {
    bmi_geometry* g = nullptr;
    check(bmi_geometry_load(filename.c_str(), scale, &g));
    g_ = std::shared_ptr<bmi_geometry>(g, [](bmi_geometry* p) { bmi_geometry_destroy(p); });
  }
  int get_number_of_patches() const {
    int n = 0;
    check(bmi_geometry_num_patches(g_.get(), &n));
    return n;
  }
  const bmi_geometry* handle() const { return g_.get(); }

 private:
  std::shared_ptr<bmi_geometry> g_;
};

// ------------------------------------------------------------------ space
template <typename Op>
class AnsatzSpace {
 public:
  AnsatzSpace() = default;
  AnsatzSpace(const Geometry& geometry, int refinement_level, int polynomial_degree, int knot_repetition = 1)
      : geometry_(geometry) {
    if (knot_repetition != 1) throw Error(BMI_UNSUPPORTED, "knot_repetition != 1 is not supported");
    bmi_space* s = nullptr;
    check(bmi_space_create(geometry.handle(), refinement_level, polynomial_degree, Op::form, &s));
    s_ = std::shared_ptr<bmi_space>(s, [](bmi_space* p) { bmi_space_destroy(p); });
    check(bmi_space_info(s_.get(), info_));
    polynomial_degree_ = polynomial_degree;
  }
  int get_number_of_dofs() const { return static_cast<int>(info_[5]); }
  int get_number_of_elements() const { return static_cast<int>(info_[0]); }
  int get_refinement_level() const { return static_cast<int>(info_[7]); }
  int get_polynomial_degree() const { return polynomial_degree_; }
  const bmi_space* handle() const { return s_.get(); }

 private:
  Geometry geometry_;  // keeps the geometry alive as long as the space
  std::shared_ptr<bmi_space> s_;
  int64_t info_[8] = {0};
  int polynomial_degree_ = 0;
};

}  // namespace Bembel

namespace Eigen {

// Device-resident H2 matrix (Eigen::H2Matrix, H2Matrix/H2Matrix.hpp:59-296)
template <typename ScalarT>
class H2Matrix {
 public:
  using Scalar = ScalarT;
  H2Matrix() = default;
  template <typename Op>
  void init_H2Matrix(const Op& op, const Bembel::AnsatzSpace<Op>& space, int number_of_points = 9,
                     double eta = 1.6, int min_cluster_level = 1, int precision = BMI_PRECISION_FP64) {
    bmi_op_desc d;
    Bembel::check(bmi_op_desc_default(Op::kind, &d));
    d.kappa_re = op.wavenumber_.real();
    d.kappa_im = op.wavenumber_.imag();
    d.n_cheb = number_of_points;
    d.eta = eta;
    d.min_cluster_level = min_cluster_level;
    d.precision = precision;
    bmi_h2* h = nullptr;
    Bembel::check(bmi_h2_assemble(space.handle(), &d, nullptr, &h));
    h_ = std::shared_ptr<bmi_h2>(h, [](bmi_h2* p) { bmi_h2_destroy(p); });
    n_ = space.get_number_of_dofs();
  }
  int64_t rows() const { return n_; }
  int64_t cols() const { return n_; }
  // Y = alpha * H * X + beta * Y; X, Y column-major host or device (mem) arrays of Scalar
  void apply(int nrhs, const Scalar* X, int64_t ldx, Scalar* Y, int64_t ldy, Scalar alpha = Scalar(1),
             Scalar beta = Scalar(0), int mem = BMI_MEM_HOST, void* stream = nullptr) const {
    const double a[2] = {std::real(std::complex<double>(alpha)), std::imag(std::complex<double>(alpha))};
    const double b[2] = {std::real(std::complex<double>(beta)), std::imag(std::complex<double>(beta))};
    Bembel::check(bmi_h2_matvec(h_.get(), nrhs, reinterpret_cast<const double*>(X), ldx,
                                reinterpret_cast<double*>(Y), ldy, a, b, mem, stream));
  }
  std::vector<Scalar> operator*(const std::vector<Scalar>& x) const {
    std::vector<Scalar> y(x.size());
    apply(1, x.data(), static_cast<int64_t>(x.size()), y.data(), static_cast<int64_t>(y.size()));
    return y;
  }
#ifdef BEMBEL_AMD_HAVE_EIGEN
  // Eigen-compatible matrix-free product consumed by ConjugateGradient / GMRES through a
  // wrapper that forwards operator* (the reference's scaleAndAddTo computes res += alpha H x)
  template <typename Derived>
  Eigen::Matrix<Scalar, Eigen::Dynamic, Eigen::Dynamic> operator*(const Eigen::MatrixBase<Derived>& x) const {
    Eigen::Matrix<Scalar, Eigen::Dynamic, Eigen::Dynamic> X = x;
    Eigen::Matrix<Scalar, Eigen::Dynamic, Eigen::Dynamic> Y(X.rows(), X.cols());
    apply(static_cast<int>(X.cols()), X.data(), X.rows(), Y.data(), Y.rows());
    return Y;
  }
#endif
  const bmi_h2* handle() const { return h_.get(); }

 private:
  std::shared_ptr<bmi_h2> h_;
  int64_t n_ = 0;
};

}  // namespace Eigen

namespace Bembel {

// DiscreteOperator<Format, Op> (LinearOperator/DiscreteOperator.hpp:107-145)
template <typename Format, typename Op>
class DiscreteOperator {
 public:
  DiscreteOperator() = default;
  explicit DiscreteOperator(const AnsatzSpace<Op>& space) : space_(space) {}
  void init_DiscreteOperator(const AnsatzSpace<Op>& space) { space_ = space; }
  void compute() { disc_op_.init_H2Matrix(lin_op_, space_); }
  // the H2 parameters of H2Matrix::init_H2Matrix, which the reference reaches only by calling
  // init_H2Matrix directly
  void compute(int number_of_points, double eta = 1.6, int min_cluster_level = 1,
               int precision = BMI_PRECISION_FP64) {
    disc_op_.init_H2Matrix(lin_op_, space_, number_of_points, eta, min_cluster_level, precision);
  }
  const Format& get_discrete_operator() const { return disc_op_; }
  const Op& get_linear_operator() const { return lin_op_; }
  Op& get_linear_operator() { return lin_op_; }

 private:
  Op lin_op_;
  Format disc_op_;
  AnsatzSpace<Op> space_;
};

}  // namespace Bembel
