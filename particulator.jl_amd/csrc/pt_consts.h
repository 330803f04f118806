// pt_consts.h -- physical constants of the reference (src/constants.jl:6-166) and the
// derived constants in its evaluation order (:157-164).  Shared by the device code and
// the host table builder; plain C++, no HIP.
#pragma once

namespace pt {

constexpr double PI = 3.141592653589793;
constexpr double C = 299792458.0;
constexpr double C2 = C * C;
constexpr double ME = 9.10938356e-31;
constexpr double QE = 1.6021766208e-19;
constexpr double ALPHA = 0.0072973525664;
constexpr double HBAR = 1.0545718001391127e-34;
constexpr double EPS0 = 8.854187817620389e-12;
constexpr double EV = 1.6021766208e-19;
constexpr double MC2 = ME * C2;
constexpr double RE = QE * QE / (ME * C2) / (4.0 * PI * EPS0);
constexpr double A0 = HBAR / (ME * C * ALPHA);
constexpr double TWO_PI = 2.0 * PI;
constexpr double EPS64 = 2.220446049250313e-16;  // eps(Float64), src/mixed_population.jl:65

}  // namespace pt
