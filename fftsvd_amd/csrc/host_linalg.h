#pragma once
#include <complex>

namespace fsvd {
void pinv(const double* X, int m, int n, double* XI);
void dft3_r2c(int N, const double* in, std::complex<double>* out);
void translation_spectra(int n, double h, int kernel, double kappa, std::complex<double>* out);
}  // namespace fsvd
