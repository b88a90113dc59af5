// Host side of the numerical MCSE attachment: the covariance of the
// replayed fold scores and of the target draws, and its PSD-safe finalize
// (rust/crates/vckss-core/src/numerical_mc.rs:180-354).  3 x 3 matrices are
// row-major in primitive order worker, firm, covariance.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace vk {

using Matrix3 = std::array<double, 9>;

// Scaled centred cross product of two score sets (n x 3 each, n >= 2):
// returns false for bad counts or nonfinite scores.
bool cross_covariance(const std::vector<double> &a, const std::vector<double> &b, Matrix3 &out);

struct Covariance {
  Matrix3 conditional{}, leverage{}, raw{}, usable{};
  std::array<double, 4> mcse{};
  uint32_t status = 6;  // 1 exact zero, 2 ok, 3 PSD adjusted, 4 unstable non-PSD, 6 nonfinite
  double psd_adjustment = 0.0;
};

Covariance finalize_covariance(const Matrix3 &conditional, const Matrix3 &leverage);

// Means of n x 3 values, each added as v / n with compensation.
std::array<double, 3> score_means(const std::vector<double> &x);

}  // namespace vk
