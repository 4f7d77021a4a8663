// Bembel.hpp -- C++ facade with the reference's class names over the bembel_amd C ABI.
//
// A Bembel program that uses the single-layer path switches by replacing its includes
// (<Bembel/AnsatzSpace>, <Bembel/Geometry>, <Bembel/H2Matrix>, <Bembel/Laplace>,
// <Bembel/LinearForm>, ...) with this header; the calls keep their shape:
//
//   Bembel::Geometry geometry("sphere.dat");                          // Geometry.hpp:37
//   Bembel::AnsatzSpace<Bembel::HelmholtzSingleLayerOperator>
//       ansatz_space(geometry, refinement_level, polynomial_degree);  // AnsatzSpace.hpp:80
//   Bembel::DiscreteLinearForm<Bembel::DirichletTrace<std::complex<double>>,
//                              Bembel::HelmholtzSingleLayerOperator> disc_lf(ansatz_space);
//   disc_lf.get_linear_form().set_function(fun);                      // DirichletTrace.hpp:38
//   disc_lf.compute();                                                // DiscreteLinearForm.hpp:44
//   Bembel::DiscreteOperator<Eigen::H2Matrix<std::complex<double>>,
//                            Bembel::HelmholtzSingleLayerOperator> disc_op(ansatz_space);
//   disc_op.get_linear_operator().set_wavenumber(kappa);
//   disc_op.compute();                                                // DiscreteOperator.hpp:127
//   Eigen::GMRES<Eigen::H2Matrix<std::complex<double>>, Eigen::IdentityPreconditioner> gmres;
//   gmres.compute(disc_op.get_discrete_operator());
//   auto rho = gmres.solve(disc_lf.get_discrete_linear_form());
//   Bembel::DiscretePotential<Bembel::HelmholtzSingleLayerPotential<Op>, Op> pot(ansatz_space);
//   pot.set_cauchy_data(rho);
//   auto u = pot.evaluate(points);                                    // DiscretePotential.hpp:49
//
// Everything runs on the GPU through libbembel_amd.so.  With Eigen available (<Eigen/Core> and
// <Eigen/Sparse> found), Eigen::H2Matrix is an Eigen matrix-free operator: H * x is an Eigen
// product whose scaleAndAddTo is the device product (res += alpha H x, all columns at once), so
// Eigen's own iterative solvers consume it.  Without Eigen the header brings the few Eigen names
// a Bembel program needs (MatrixXd, VectorXd, Vector3d, ConjugateGradient, GMRES,
// IdentityPreconditioner, Lower, Upper) as small host containers and device solvers.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../bembel_amd.h"

#if defined(__has_include)
#if __has_include(<Eigen/Core>) && __has_include(<Eigen/Sparse>)
#include <Eigen/Core>
#include <Eigen/Sparse>
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

template <typename S>
struct is_complex : std::false_type {};
template <typename T>
struct is_complex<std::complex<T>> : std::true_type {};

// ------------------------------------------------------------------ host containers
// Column-major dense matrix (used when Eigen is absent, and as the storage of dense operators)
template <typename S>
class DenseMatrix {
 public:
  using Scalar = S;
  DenseMatrix() = default;
  DenseMatrix(int64_t r, int64_t c) : r_(r), c_(c), d_(static_cast<size_t>(r * c), S(0)) {}
  void resize(int64_t r, int64_t c) {
    r_ = r, c_ = c;
    d_.assign(static_cast<size_t>(r * c), S(0));
  }
  void setZero() { std::fill(d_.begin(), d_.end(), S(0)); }
  int64_t rows() const { return r_; }
  int64_t cols() const { return c_; }
  S* data() { return d_.data(); }
  const S* data() const { return d_.data(); }
  S& operator()(int64_t i, int64_t j) { return d_[static_cast<size_t>(j * r_ + i)]; }
  const S& operator()(int64_t i, int64_t j) const { return d_[static_cast<size_t>(j * r_ + i)]; }
  // y = A x
  std::vector<S> operator*(const std::vector<S>& x) const {
    std::vector<S> y(static_cast<size_t>(r_), S(0));
    for (int64_t j = 0; j < c_; ++j)
      for (int64_t i = 0; i < r_; ++i) y[static_cast<size_t>(i)] += (*this)(i, j) * x[static_cast<size_t>(j)];
    return y;
  }

 private:
  int64_t r_ = 0, c_ = 0;
  std::vector<S> d_;
};

#ifdef BEMBEL_AMD_HAVE_EIGEN
using Vector3d = Eigen::Vector3d;
template <typename S>
using Vector3 = Eigen::Matrix<S, 3, 1>;
template <typename S>
using MatrixX = Eigen::Matrix<S, Eigen::Dynamic, Eigen::Dynamic>;
template <typename S>
using VectorX = Eigen::Matrix<S, Eigen::Dynamic, 1>;
#else
using Vector3d = std::array<double, 3>;
template <typename S>
using Vector3 = std::array<S, 3>;
template <typename S>
using MatrixX = DenseMatrix<S>;
template <typename S>
using VectorX = std::vector<S>;
#endif

namespace detail {
template <typename V>
int64_t size_of(const V& v) {
  return static_cast<int64_t>(v.size());
}
template <typename M>
const double* points_data(const M& pts, std::vector<double>& tmp) {
  // n x 3 points -> 3 doubles per point
  const int64_t n = pts.rows();
  tmp.resize(static_cast<size_t>(3 * n));
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) tmp[static_cast<size_t>(3 * i + k)] = pts(i, k);
  return tmp.data();
}
}  // namespace detail

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

template <typename Op>
bmi_op_desc make_desc(const Op& op) {
  bmi_op_desc d;
  check(bmi_op_desc_default(Op::kind, &d));
  d.kappa_re = op.wavenumber_.real();
  d.kappa_im = op.wavenumber_.imag();
  return d;
}

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

// ------------------------------------------------------------------ H2 matrix
namespace Eigen {
template <typename ScalarT>
class H2Matrix;
}

#ifdef BEMBEL_AMD_HAVE_EIGEN
namespace Eigen {
namespace internal {
// matrix-free operator: the traits of a sparse matrix give it SparseShape, so products with
// dense vectors and matrices go through generic_product_impl below
template <typename ScalarT>
struct traits<Eigen::H2Matrix<ScalarT>> : public Eigen::internal::traits<Eigen::SparseMatrix<ScalarT>> {};
}  // namespace internal
}  // namespace Eigen
#endif

namespace Eigen {

// Device-resident H2 matrix (Eigen::H2Matrix, H2Matrix/H2Matrix.hpp:59-296)
template <typename ScalarT>
class H2Matrix
#ifdef BEMBEL_AMD_HAVE_EIGEN
    : public Eigen::EigenBase<H2Matrix<ScalarT>>
#endif
{
 public:
  using Scalar = ScalarT;
  using RealScalar = double;
  using StorageIndex = int;
#ifdef BEMBEL_AMD_HAVE_EIGEN
  enum { ColsAtCompileTime = Eigen::Dynamic, MaxColsAtCompileTime = Eigen::Dynamic, IsRowMajor = false };
#endif
  H2Matrix() = default;
  // H2Matrix::init_H2Matrix (H2Matrix.hpp:93-96); eta, min_cluster_level and precision are the
  // block-tree parameters and the storage precision of this implementation
  template <typename Op>
  void init_H2Matrix(const Op& op, const Bembel::AnsatzSpace<Op>& space, int number_of_points = 9,
                     double eta = 1.6, int min_cluster_level = 1, int precision = BMI_PRECISION_FP64) {
    static_assert(std::is_same<typename Op::Scalar, ScalarT>::value, "H2Matrix scalar must match the operator");
    bmi_op_desc d = Bembel::make_desc(op);
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
  // H2Matrix::get_dense (H2Matrix.hpp:254-264): the product with the identity
  Bembel::MatrixX<Scalar> get_dense() const {
    Bembel::MatrixX<Scalar> I(n_, n_), D(n_, n_);
    I.setZero();
    for (int64_t i = 0; i < n_; ++i) I(i, i) = Scalar(1);
    apply(static_cast<int>(n_), I.data(), n_, D.data(), n_);
    return D;
  }
#ifdef BEMBEL_AMD_HAVE_EIGEN
  template <typename Rhs>
  Eigen::Product<H2Matrix, Rhs, Eigen::AliasFreeProduct> operator*(const Eigen::MatrixBase<Rhs>& x) const {
    return Eigen::Product<H2Matrix, Rhs, Eigen::AliasFreeProduct>(*this, x.derived());
  }
#endif
  const bmi_h2* handle() const { return h_.get(); }

 private:
  std::shared_ptr<bmi_h2> h_;
  int64_t n_ = 0;
};

}  // namespace Eigen

#ifdef BEMBEL_AMD_HAVE_EIGEN
namespace Eigen {
namespace internal {
// res += alpha * H * rhs for every column at once (the reference's scaleAndAddTo,
// H2Matrix/EigenHelper/H2DenseProduct.hpp:268-290, maps to the same device product)
template <typename ScalarT, typename Rhs, int ProductType>
struct generic_product_impl<Eigen::H2Matrix<ScalarT>, Rhs, SparseShape, DenseShape, ProductType>
    : generic_product_impl_base<Eigen::H2Matrix<ScalarT>, Rhs,
                                generic_product_impl<Eigen::H2Matrix<ScalarT>, Rhs, SparseShape, DenseShape,
                                                     ProductType>> {
  typedef typename Product<Eigen::H2Matrix<ScalarT>, Rhs>::Scalar Scalar;
  template <typename Dest>
  static void scaleAndAddTo(Dest& dst, const Eigen::H2Matrix<ScalarT>& lhs, const Rhs& rhs, const Scalar& alpha) {
    const Eigen::Matrix<ScalarT, Eigen::Dynamic, Eigen::Dynamic> X = rhs;
    Eigen::Matrix<ScalarT, Eigen::Dynamic, Eigen::Dynamic> Y = dst;
    lhs.apply(static_cast<int>(X.cols()), X.data(), X.rows(), Y.data(), Y.rows(), ScalarT(alpha), ScalarT(1));
    dst = Y;
  }
};
}  // namespace internal
}  // namespace Eigen
#endif

namespace Bembel {

// ------------------------------------------------------------------ lazy sums
// (A ± B) x and (a A) x by the distributive law, as the reference's product glue does
// (H2Matrix/EigenHelper/H2DenseProduct.hpp:179-258): a list of (coefficient, operator) terms,
// applied term by term.  Terms are H2 matrices or dense matrices (MatrixX<S>); the operands
// must outlive the combination.
template <typename S>
class LinearCombination {
 public:
  struct Term {
    S coef;
    const Eigen::H2Matrix<S>* h2;
    const MatrixX<S>* dense;
  };
  LinearCombination() = default;
  explicit LinearCombination(const Eigen::H2Matrix<S>& h, S c = S(1)) { terms_.push_back({c, &h, nullptr}); }
  explicit LinearCombination(const MatrixX<S>& d, S c = S(1)) { terms_.push_back({c, nullptr, &d}); }
  LinearCombination& add(const LinearCombination& o, S sign) {
    for (const Term& t : o.terms_) terms_.push_back({sign * t.coef, t.h2, t.dense});
    return *this;
  }
  LinearCombination scaled(S a) const {
    LinearCombination r = *this;
    for (Term& t : r.terms_) t.coef *= a;
    return r;
  }
  int64_t rows() const { return terms_.empty() ? 0 : (terms_[0].h2 ? terms_[0].h2->rows() : terms_[0].dense->rows()); }
  int64_t cols() const { return rows(); }
  template <typename V>
  V operator*(const V& x) const {
    const int64_t n = detail::size_of(x);
    V y(x);
    for (int64_t i = 0; i < n; ++i) y[i] = S(0);
    for (const Term& t : terms_) {
      if (t.h2) {
        V hx(x);
        t.h2->apply(1, &x[0], n, &hx[0], n);
        for (int64_t i = 0; i < n; ++i) y[i] += t.coef * hx[i];
      } else {
        for (int64_t j = 0; j < n; ++j)
          for (int64_t i = 0; i < n; ++i) y[i] += t.coef * (*t.dense)(i, j) * x[j];
      }
    }
    return y;
  }

 private:
  std::vector<Term> terms_;
};

template <typename S>
LinearCombination<S> as_combination(const Eigen::H2Matrix<S>& h) {
  return LinearCombination<S>(h);
}
template <typename S>
LinearCombination<S> as_combination(const MatrixX<S>& d) {
  return LinearCombination<S>(d);
}
template <typename S>
LinearCombination<S> as_combination(const LinearCombination<S>& c) {
  return c;
}

}  // namespace Bembel

namespace Eigen {
// H2 ± H2, H2 ± dense, dense ± H2, a H2 and their chains (test_LazyEigenSum.cpp)
template <typename S>
Bembel::LinearCombination<S> operator+(const H2Matrix<S>& a, const H2Matrix<S>& b) {
  return Bembel::LinearCombination<S>(a).add(Bembel::LinearCombination<S>(b), S(1));
}
template <typename S>
Bembel::LinearCombination<S> operator-(const H2Matrix<S>& a, const H2Matrix<S>& b) {
  return Bembel::LinearCombination<S>(a).add(Bembel::LinearCombination<S>(b), S(-1));
}
template <typename S>
Bembel::LinearCombination<S> operator+(const H2Matrix<S>& a, const Bembel::MatrixX<S>& b) {
  return Bembel::LinearCombination<S>(a).add(Bembel::LinearCombination<S>(b), S(1));
}
template <typename S>
Bembel::LinearCombination<S> operator-(const H2Matrix<S>& a, const Bembel::MatrixX<S>& b) {
  return Bembel::LinearCombination<S>(a).add(Bembel::LinearCombination<S>(b), S(-1));
}
template <typename S>
Bembel::LinearCombination<S> operator+(const Bembel::MatrixX<S>& a, const H2Matrix<S>& b) {
  return Bembel::LinearCombination<S>(a).add(Bembel::LinearCombination<S>(b), S(1));
}
template <typename S>
Bembel::LinearCombination<S> operator-(const Bembel::MatrixX<S>& a, const H2Matrix<S>& b) {
  return Bembel::LinearCombination<S>(a).add(Bembel::LinearCombination<S>(b), S(-1));
}
template <typename S>
Bembel::LinearCombination<S> operator*(const S& a, const H2Matrix<S>& h) {
  return Bembel::LinearCombination<S>(h, a);
}
}  // namespace Eigen

namespace Bembel {
template <typename S, typename T>
LinearCombination<S> operator+(const LinearCombination<S>& a, const T& b) {
  return LinearCombination<S>(a).add(LinearCombination<S>(b), S(1));
}
template <typename S, typename T>
LinearCombination<S> operator-(const LinearCombination<S>& a, const T& b) {
  return LinearCombination<S>(a).add(LinearCombination<S>(b), S(-1));
}
template <typename S>
LinearCombination<S> operator+(const LinearCombination<S>& a, const LinearCombination<S>& b) {
  return LinearCombination<S>(a).add(b, S(1));
}
template <typename S>
LinearCombination<S> operator-(const LinearCombination<S>& a, const LinearCombination<S>& b) {
  return LinearCombination<S>(a).add(b, S(-1));
}
template <typename S>
LinearCombination<S> operator*(const S& a, const LinearCombination<S>& c) {
  return c.scaled(a);
}

// ------------------------------------------------------------------ discrete operators
// DiscreteOperatorComputer (LinearOperator/DiscreteOperator.hpp:27-102): how compute() fills
// each format
template <typename Format, typename Op>
struct DiscreteOperatorComputer;

// H2 (DiscreteOperator.hpp:92-102)
template <typename S, typename Op>
struct DiscreteOperatorComputer<Eigen::H2Matrix<S>, Op> {
  static void compute(Eigen::H2Matrix<S>* disc_op, const Op& op, const AnsatzSpace<Op>& space) {
    disc_op->init_H2Matrix(op, space);
  }
};

// dense Galerkin matrix T^T A T (DiscreteOperator.hpp:27-87), assembled on the device
template <typename S, typename Op, typename Mat>
void compute_dense(Mat* disc_op, const Op& op, const AnsatzSpace<Op>& space) {
  static_assert(std::is_same<typename Op::Scalar, S>::value, "matrix scalar must match the operator");
  const bmi_op_desc d = make_desc(op);
  const int64_t n = space.get_number_of_dofs();
  disc_op->resize(n, n);
  check(bmi_dense_assemble(space.handle(), &d, 0, reinterpret_cast<double*>(disc_op->data()), BMI_MEM_HOST, nullptr));
}

template <typename S, typename Op>
struct DiscreteOperatorComputer<DenseMatrix<S>, Op> {
  static void compute(DenseMatrix<S>* disc_op, const Op& op, const AnsatzSpace<Op>& space) {
    compute_dense<S>(disc_op, op, space);
  }
};

#ifdef BEMBEL_AMD_HAVE_EIGEN
template <typename S, typename Op>
struct DiscreteOperatorComputer<Eigen::Matrix<S, Eigen::Dynamic, Eigen::Dynamic>, Op> {
  static void compute(Eigen::Matrix<S, Eigen::Dynamic, Eigen::Dynamic>* disc_op, const Op& op,
                      const AnsatzSpace<Op>& space) {
    compute_dense<S>(disc_op, op, space);
  }
};
#endif

// DiscreteOperator<Format, Op> (LinearOperator/DiscreteOperator.hpp:107-145)
template <typename Format, typename Op>
class DiscreteOperator {
 public:
  DiscreteOperator() = default;
  explicit DiscreteOperator(const AnsatzSpace<Op>& space) : space_(space) {}
  void init_DiscreteOperator(const AnsatzSpace<Op>& space) { space_ = space; }
  void compute() { DiscreteOperatorComputer<Format, Op>::compute(&disc_op_, lin_op_, space_); }
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

// ------------------------------------------------------------------ linear forms
// DirichletTrace<Scalar> (LinearForm/DirichletTrace.hpp:31-70): trace of a scalar function
template <typename S>
class DirichletTrace {
 public:
  using Scalar = S;
  static constexpr int value_kind = is_complex<S>::value ? 1 : 0;
  void set_function(const std::function<S(const Vector3d&)>& f) { f_ = f; }
  const std::function<S(const Vector3d&)>& get_function() const { return f_; }

 private:
  std::function<S(const Vector3d&)> f_;
};

// RotatedTangentialTrace<Scalar> (LinearForm/RotatedTangentialTrace.hpp:30-87): rotated
// tangential trace of a vector field, DivConforming spaces
template <typename S>
class RotatedTangentialTrace {
 public:
  using Scalar = S;
  static constexpr int value_kind = 2;
  void set_function(const std::function<Vector3<S>(const Vector3d&)>& f) { f_ = f; }
  const std::function<Vector3<S>(const Vector3d&)>& get_function() const { return f_; }

 private:
  std::function<Vector3<S>(const Vector3d&)> f_;
};

// DiscreteLinearForm<LinearForm, Op> (LinearForm/DiscreteLinearForm.hpp:44-83): the function
// is evaluated on the host at the points of the (P+2)^2-point rule, the Galerkin projection
// (surface measure or tangential trace, basis functions, T^T) runs on the device
template <typename LF, typename Op>
class DiscreteLinearForm {
 public:
  using Scalar = typename LF::Scalar;
  DiscreteLinearForm() = default;
  explicit DiscreteLinearForm(const AnsatzSpace<Op>& space) : space_(space) {}
  void init_DiscreteLinearForm(const AnsatzSpace<Op>& space) { space_ = space; }
  LF& get_linear_form() { return lf_; }
  const LF& get_linear_form() const { return lf_; }
  void compute() {
    int64_t npts = 0;
    check(bmi_linear_form_points(space_.handle(), &npts, nullptr));
    std::vector<double> pts(static_cast<size_t>(3 * npts));
    check(bmi_linear_form_points(space_.handle(), &npts, pts.data()));
    const int per = LF::value_kind == 2 ? 3 : 1;
    std::vector<Scalar> vals(static_cast<size_t>(per * npts));
    const auto& f = lf_.get_function();
    if (!f) throw Error(BMI_INVALID_ARG, "DiscreteLinearForm: set_function() has not been called");
    for (int64_t k = 0; k < npts; ++k) {
      Vector3d x;
      for (int c = 0; c < 3; ++c) x[c] = pts[static_cast<size_t>(3 * k + c)];
      if constexpr (LF::value_kind == 2) {
        const Vector3<Scalar> v = f(x);
        for (int c = 0; c < 3; ++c) vals[static_cast<size_t>(3 * k + c)] = v[c];
      } else {
        vals[static_cast<size_t>(k)] = f(x);
      }
    }
    const int64_t n = space_.get_number_of_dofs();
    disc_lf_ = VectorX<Scalar>(static_cast<size_t>(n));
    check(bmi_rhs_from_values(space_.handle(), LF::value_kind, reinterpret_cast<const double*>(vals.data()), 1,
                              reinterpret_cast<double*>(&disc_lf_[0]), BMI_MEM_HOST, nullptr));
  }
  const VectorX<Scalar>& get_discrete_linear_form() const { return disc_lf_; }

 private:
  AnsatzSpace<Op> space_;
  LF lf_;
  VectorX<Scalar> disc_lf_;
};

// ------------------------------------------------------------------ potentials
// {Laplace,Helmholtz,Maxwell}SingleLayerPotential<Op> (Laplace/SingleLayerPotential.hpp,
// Helmholtz/SingleLayerPotential.hpp, Maxwell/SingleLayerPotential.hpp)
template <typename Op>
struct LaplaceSingleLayerPotential {
  using Scalar = double;
  static constexpr int kind = BMI_LAPLACE_SL;
  static constexpr int dim = 1;
  std::complex<double> wavenumber_ = 0.0;
};
template <typename Op>
struct HelmholtzSingleLayerPotential {
  using Scalar = std::complex<double>;
  static constexpr int kind = BMI_HELMHOLTZ_SL;
  static constexpr int dim = 1;
  std::complex<double> wavenumber_ = 0.0;
  void set_wavenumber(std::complex<double> k) { wavenumber_ = k; }
  std::complex<double> get_wavenumber() const { return wavenumber_; }
};
template <typename Op>
struct MaxwellSingleLayerPotential {
  using Scalar = std::complex<double>;
  static constexpr int kind = BMI_MAXWELL_SL;
  static constexpr int dim = 3;
  std::complex<double> wavenumber_ = 0.0;
  void set_wavenumber(std::complex<double> k) { wavenumber_ = k; }
  std::complex<double> get_wavenumber() const { return wavenumber_; }
};

// DiscretePotential<Potential, Op> (Potential/DiscretePotential.hpp:20-120)
template <typename Pot, typename Op>
class DiscretePotential {
 public:
  using Scalar = typename Pot::Scalar;
  DiscretePotential() = default;
  explicit DiscretePotential(const AnsatzSpace<Op>& space) : space_(space) {}
  void init_DiscretePotential(const AnsatzSpace<Op>& space) { space_ = space; }
  Pot& get_potential() { return pot_; }
  void set_cauchy_data(const VectorX<typename Op::Scalar>& rho) { rho_ = rho; }
  // points: n x 3 (MatrixX<double>); result: n x dim
  template <typename M>
  MatrixX<Scalar> evaluate(const M& points) const {
    std::vector<double> tmp;
    const double* p = detail::points_data(points, tmp);
    const int n = static_cast<int>(points.rows());
    bmi_op_desc d;
    check(bmi_op_desc_default(Pot::kind, &d));
    d.kappa_re = pot_.wavenumber_.real();
    d.kappa_im = pot_.wavenumber_.imag();
    MatrixX<Scalar> out(n, Pot::dim);
    check(bmi_potential_evaluate(space_.handle(), &d, reinterpret_cast<const double*>(&rho_[0]), n, p,
                                 reinterpret_cast<double*>(out.data()), BMI_MEM_HOST, nullptr));
    return out;
  }

 private:
  AnsatzSpace<Op> space_;
  Pot pot_;
  VectorX<typename Op::Scalar> rho_;
};

// ------------------------------------------------------------------ device solvers
// Restarted GMRES (FP64 Arnoldi) and CG on the device: the consumers of the reference
// (Eigen::GMRES, examples/HelmholtzSingleLayerH2.cpp:82-85; Eigen::ConjugateGradient,
// examples/LaplaceSingleLayerH2.cpp:78-81) with Eigen's method names
template <typename Matrix, bool UseCG>
class DeviceSolver {
 public:
  using Scalar = typename Matrix::Scalar;
  void compute(const Matrix& A) { A_ = &A; }
  void setTolerance(double t) { tol_ = t; }
  void setMaxIterations(int n) { maxit_ = n; }
  void set_restart(int r) { restart_ = r; }
  template <typename V>
  V solve(const V& b) {
    const int64_t n = detail::size_of(b);
    V x(b);
    for (int64_t i = 0; i < n; ++i) x[i] = Scalar(0);
    int it = 0;
    double res = 0;
    const int maxit = maxit_ > 0 ? maxit_ : static_cast<int>(2 * n);  // Eigen's default: 2 n
    const bmi_status st =
        UseCG ? bmi_solve_cg(A_->handle(), 1, reinterpret_cast<const double*>(&b[0]), reinterpret_cast<double*>(&x[0]),
                             tol_, maxit, 1, &it, &res, BMI_MEM_HOST, nullptr)
              : bmi_solve_gmres(A_->handle(), 1, reinterpret_cast<const double*>(&b[0]),
                                reinterpret_cast<double*>(&x[0]), tol_, restart_, maxit, 1, &it, &res, BMI_MEM_HOST,
                                nullptr);
    if (st != BMI_OK && st != BMI_NOT_CONVERGED) check(st);
    info_ = st;
    iters_ = it, err_ = res;
    return x;
  }
  int iterations() const { return iters_; }
  double error() const { return err_; }
  bmi_status info() const { return info_; }

 private:
  const Matrix* A_ = nullptr;
  double tol_ = 1e-12;
  int maxit_ = -1, restart_ = 30, iters_ = 0;
  double err_ = 0;
  bmi_status info_ = BMI_OK;
};

}  // namespace Bembel

#ifndef BEMBEL_AMD_HAVE_EIGEN
// The Eigen names a Bembel program uses, when Eigen itself is absent
namespace Eigen {
using MatrixXd = Bembel::DenseMatrix<double>;
using MatrixXcd = Bembel::DenseMatrix<std::complex<double>>;
using VectorXd = std::vector<double>;
using VectorXcd = std::vector<std::complex<double>>;
using Vector3d = Bembel::Vector3d;
using Vector3cd = Bembel::Vector3<std::complex<double>>;
enum { Lower = 1, Upper = 2 };
struct IdentityPreconditioner {};
template <typename Matrix, int UpLo = Lower, typename Preconditioner = IdentityPreconditioner>
class ConjugateGradient : public Bembel::DeviceSolver<Matrix, true> {};
template <typename Matrix, typename Preconditioner = IdentityPreconditioner>
class GMRES : public Bembel::DeviceSolver<Matrix, false> {};
}  // namespace Eigen
#endif
