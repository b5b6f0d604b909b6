// Range check of the 64-bit fixed-point terms behind the order-independent
// sums (IBM spreading at 2^40, membrane volume/area/MAD at 2^30).  Plain C++
// (constexpr, so HIP device code calls it too) with no HIP dependency, so a
// host compiler can test it on its own.
#pragma once

namespace lbfs {

// A scaled term must stay below 2^52 in magnitude: then 2^11 of them cannot
// wrap a signed 64-bit sum, and the double -> integer rounding is exact.
// At the IBM scale 2^40 this is |F w| < 4096 per node-stencil term, at the
// membrane scale 2^30 about 4.2e6 per triangle or edge term: values far
// beyond any stable run's, so reaching the limit means the run has blown up.
constexpr double FIX_TERM_LIMIT = 4503599627370496.0;  // 2^52

// true when the scaled term s can be converted: finite and |s| < 2^52
// (NaN fails both comparisons)
constexpr bool fix_in_range(double s) { return s < FIX_TERM_LIMIT && s > -FIX_TERM_LIMIT; }

// bits of the context's device error word
constexpr int ERR_FIX_IBM = 1;  // a spread force term out of range or not finite
constexpr int ERR_FIX_MEM = 2;  // a membrane volume / area / MAD term out of range or not finite
constexpr int ERR_FIX_ADJ = 4;  // an interpolation-adjoint term not finite

}  // namespace lbfs
