// fft_roots.h -- exp(-2 pi i e / R) for the radix-3 and radix-5 butterflies,
// as exact-to-rounding double literals (generated with 30-digit arithmetic).
#pragma once

namespace skl {
template <int R> struct RootTab;
template <> struct RootTab<3> {
    static constexpr double c[3] = {1, -0.5, -0.5};
    static constexpr double s[3] = {0, -0.8660254037844386, 0.8660254037844386};
};
template <> struct RootTab<5> {
    static constexpr double c[5] = {1, 0.30901699437494745, -0.80901699437494745, -0.80901699437494745, 0.30901699437494745};
    static constexpr double s[5] = {0, -0.95105651629515353, -0.58778525229247314, 0.58778525229247314, 0.95105651629515353};
};
} // namespace skl
