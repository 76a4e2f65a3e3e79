// Initial-field functions of the reference (src/math/Function.h) that need
// Bessel functions: the 2D cylindrical and 3D spherical J6 cavity modes.
#pragma once

#include <cmath>

namespace dgtd {

constexpr double kAlpha6 = 13.589290170541217;  // first zero of J_6 (Function.h:182, 208)

// J_6(a6 r) cos(6 phi) at t = 0 (BesselJ6::eval, Function.h:180-192); jn is
// the POSIX C library's integer-order Bessel function, as there.
inline double besselj6_2d(double x, double y)
{
    return ::jn(6, kAlpha6 * std::sqrt(x * x + y * y)) * std::cos(6.0 * std::atan2(y, x));
}

// Spherical Bessel j_6: the power series
//   j_n(x) = x^n / (2n+1)!! * sum_k (-x^2/2)^k / (k! (2n+3)(2n+5)...(2n+2k+1))
// below x = 8, upward recurrence j_{n+1} = (2n+1)/x j_n - j_{n-1} from
// j_0 = sin x / x, j_1 = sin x / x^2 - cos x / x above (stable for x > n).
inline double spherical_j6(double x)
{
    if (x < 8.0) {
        double pref = 1.0;
        for (int i = 0; i < 6; ++i) pref *= x / (2 * i + 3);  // x^6 / 13!!
        double term = 1.0, sum = 1.0;
        for (int k = 1; k < 80; ++k) {
            term *= -0.5 * x * x / (k * (2.0 * k + 13.0));
            sum += term;
            if (std::fabs(term) <= 1e-18 * std::fabs(sum)) break;
        }
        return pref * sum;
    }
    double jm = std::sin(x) / x, j = std::sin(x) / (x * x) - std::cos(x) / x;
    for (int n = 1; n < 6; ++n) {
        const double jp = (2 * n + 1) / x * j - jm;
        jm = j;
        j = jp;
    }
    return j;
}

// j_6(a6 r) Re Y_6^6(theta, phi) at t = 0 (SphericalBesselJ6::eval,
// Function.h:206-222): the normalised associated Legendre function
// sqrt(13 / 4pi) sqrt(0! / 12!) P_6^6(cos theta), P_6^6(c) = 10395 (1 - c^2)^3
// (the Condon-Shortley sign (-1)^6 is +1), times cos(6 phi).
inline double besselj6_3d(double x, double y, double z)
{
    const double r = std::sqrt(x * x + y * y + z * z);
    if (r == 0.0) return 0.0;
    const double c = z / r, s2 = 1.0 - c * c;
    const double norm = std::sqrt(13.0 / (4.0 * M_PI) / 479001600.0) * 10395.0;
    return spherical_j6(kAlpha6 * r) * norm * s2 * s2 * s2 * std::cos(6.0 * std::atan2(y, x));
}

}  // namespace dgtd
