// Rational approximation of the Ewald real-space gradient factor
//
//   H(u) = h(a) / a^3,   u = a^2,   h(a) = erf(a) - (2/sqrt(pi)) a exp(-a^2)
//
// so that NAMD's slow-force factor (1 - corr_grad)/r^3 - 1/rc^3
// (src/CudaComputeNonbondedInteractions.h:131-159) becomes
//
//   fSlow / qq = beta^3 H(beta^2 r^2) - 1/rc^3.
//
// H is smooth and completely monotone on u >= 0 (H(0) = 4/(3 sqrt(pi)),
// H ~ u^-3/2 for large u), so a rational P6(u)/Q6(u) fits it without the
// cancellation that 1 - erfc(a) - ... has at short range and without an
// exp() or a branch.  Coefficients: Lawson-reweighted linear least squares
// for the relative error on u in [0, 16] (beta*rc <= 4), fitted in this
// repository from the closed form above; maximum relative error 6.0e-9 in
// exact arithmetic, 3.5e-7 (RMS 7.9e-8) when evaluated in FP32 by Horner's
// rule as below.
#pragma once

namespace mi355nb {

constexpr float HFIT_UMAX = 16.0f;

constexpr float HP0 = 7.52252773574053513e-01f;
constexpr float HP1 = -1.84793401771936316e-02f;
constexpr float HP2 = 1.91025334554565940e-02f;
constexpr float HP3 = 1.52724524478513025e-04f;
constexpr float HP4 = 9.11055002804028545e-05f;
constexpr float HP5 = 4.14862931853304791e-07f;
constexpr float HP6 = 2.38911116000539242e-09f;

constexpr float HQ1 = 5.75434429035597050e-01f;
constexpr float HQ2 = 1.56370292386496801e-01f;
constexpr float HQ3 = 2.62691793108253448e-02f;
constexpr float HQ4 = 2.98560768772017371e-03f;
constexpr float HQ5 = 2.28815603143183235e-04f;
constexpr float HQ6 = 1.20003380559213302e-05f;

}  // namespace mi355nb
