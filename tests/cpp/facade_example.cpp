// Reference-style program against the C++ facade (include/bembel_amd/Bembel.hpp).
// Host-only part (geometry, ansatz space) runs everywhere; with --gpu it also assembles the H2
// matrix of the Laplace single layer and applies it to the constant function.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "bembel_amd/Bembel.hpp"

int main(int argc, char** argv) {
  const char* geo_file = argc > 1 ? argv[1] : "sphere.dat";
  const bool gpu = argc > 2 && std::strcmp(argv[2], "--gpu") == 0;
  Bembel::Geometry geometry(geo_file);
  const int M = 3, P = 2;
  Bembel::AnsatzSpace<Bembel::LaplaceSingleLayerOperator> ansatz_space(geometry, M, P);
  std::printf("patches %d elements %d dofs %d\n", geometry.get_number_of_patches(),
              ansatz_space.get_number_of_elements(), ansatz_space.get_number_of_dofs());
  if (ansatz_space.get_number_of_dofs() != 6 * (8 + 2) * (8 + 2)) return 1;
  if (!gpu) return 0;
  Bembel::DiscreteOperator<Eigen::H2Matrix<double>, Bembel::LaplaceSingleLayerOperator> disc_op(ansatz_space);
  disc_op.compute(4);
  const auto& H = disc_op.get_discrete_operator();
  std::vector<double> c(H.cols(), std::ldexp(1.0, -M));
  std::vector<double> y = H * c;
  double s = 0;
  for (size_t i = 0; i < c.size(); ++i) s += c[i] * y[i];
  std::printf("c^T H c = %.12f (4 pi = %.12f)\n", s, 4 * M_PI);
  return std::fabs(s - 4 * M_PI) < 1e-3 * 4 * M_PI ? 0 : 2;
}
