// Acklam's rational approximation to the inverse normal CDF: the published coefficients of
// the central (|u - 1/2| <= 0.47575) and tail rationals (reference src/core/normal.cpp:13-48).
// Shared by the device FP64 inverse (device_common.hpp) and the host one (host/core.cpp).
#pragma once

#define MCDX_ACKLAM_CENTRAL(T, S)                                                           \
    constexpr T A1 = S(-3.969683028665376e+01), A2 = S(2.209460984245205e+02),              \
                A3 = S(-2.759285104469687e+02), A4 = S(1.383577518672690e+02),              \
                A5 = S(-3.066479806614716e+01), A6 = S(2.506628277459239e+00);              \
    constexpr T B1 = S(-5.447609879822406e+01), B2 = S(1.615858368580409e+02),              \
                B3 = S(-1.556989798598866e+02), B4 = S(6.680131188771972e+01),              \
                B5 = S(-1.328068155288572e+01);
#define MCDX_ACKLAM_TAIL(T, S)                                                              \
    constexpr T C1 = S(-7.784894002430293e-03), C2 = S(-3.223964580411365e-01),             \
                C3 = S(-2.400758277161838e+00), C4 = S(-2.549732539343734e+00),             \
                C5 = S(4.374664141464968e+00), C6 = S(2.938163982698783e+00);               \
    constexpr T D1 = S(7.784695709041462e-03), D2 = S(3.224671290700398e-01),               \
                D3 = S(2.445134137142996e+00), D4 = S(3.754408661907416e+00);

