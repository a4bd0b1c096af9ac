// Host-side table builders of the MI355X transport library.
//
// Mie amplitudes and efficiencies for the S1/S2 tables and the theta-sampling
// PDF (reference: src/mie/dmiev.cpp:6-81 calling the Fortran MIEV0, used by
// src/medium.cpp:203-236 and src/phase.cpp:270-304), and the normalised
// inverse-CDF table of SamplingTable::initialize (src/table.cpp:6-48).
//
// The Mie series follows Wiscombe's formulation (NCAR/TN-140+STR, 1979): the
// coefficients a_n, b_n from the logarithmic derivative A_n(mx) of psi_n(mx)
// and zeta_n = psi_n + i chi_n, with A_n obtained by downward recurrence
// started from Lentz's continued fraction (Lentz 1976, Appl. Opt. 15, 668),
// and the amplitudes S1, S2 from the angular functions pi_n, tau_n.  Sign
// conventions are MIEV0's: the refractive index is used with Im(m) <= 0 and
// results are conjugated when the caller passes Im(m) > 0
// (src/mie/MIEV0noP.f:160-161, 472-485).  Term count: MIEV0noP.f:190-206.

#include <cmath>
#include <complex>
#include <vector>

#include "../../include/luminis_hip.h"

namespace {

using cd = std::complex<double>;

int mie_terms(double x) {
  if (x <= 8.0) return static_cast<int>(x + 4.0 * std::pow(x, 1.0 / 3.0) + 1.0);
  if (x < 4200.0) return static_cast<int>(x + 4.05 * std::pow(x, 1.0 / 3.0) + 2.0);
  return static_cast<int>(x + 4.0 * std::pow(x, 1.0 / 3.0) + 2.0);
}

// A_N(z) = psi_N'(z)/psi_N(z) by Lentz's continued fraction
//   A_N = -N/z + 1/( (2N+1)/z - 1/( (2N+3)/z - ... ) ).
cd lentz_logderiv(int N, cd z) {
  const double tiny = 1e-300;
  auto a = [&](int k) -> cd {  // k-th partial denominator, k >= 1
    const double s = (k % 2 == 1) ? 1.0 : -1.0;
    return s * (2.0 * (N + k) - 1.0) / z;
  };
  cd f = a(1);
  if (std::abs(f) < tiny) f = tiny;
  cd C = f, D = 0.0;
  for (int k = 2; k < 100000; ++k) {
    const cd ak = a(k);
    D = ak + D;
    if (std::abs(D) < tiny) D = tiny;
    C = ak + 1.0 / C;
    if (std::abs(C) < tiny) C = tiny;
    D = 1.0 / D;
    const cd delta = C * D;
    f *= delta;
    if (std::abs(delta - 1.0) < 1e-16) break;
  }
  return -static_cast<double>(N) / z + f;
}

// a_n, b_n (n = 1..ntrm) in MIEV0's convention for m with Im(m) <= 0.
void mie_coefficients(double x, cd m, int ntrm, std::vector<cd>& an, std::vector<cd>& bn) {
  const cd mx = m * x;
  std::vector<cd> A(ntrm + 1);
  A[ntrm] = lentz_logderiv(ntrm, mx);
  for (int n = ntrm; n >= 2; --n) {
    const cd r = static_cast<double>(n) / mx;
    A[n - 1] = r - 1.0 / (A[n] + r);
  }
  an.assign(ntrm + 1, 0.0);
  bn.assign(ntrm + 1, 0.0);
  // Riccati-Bessel psi_n and chi_n by upward recurrence from n = 0, 1.
  double psi_prev = std::sin(x), chi_prev = std::cos(x);
  double psi = psi_prev / x - chi_prev, chi = chi_prev / x + psi_prev;
  const cd minv = 1.0 / m;
  for (int n = 1; n <= ntrm; ++n) {
    const double nx = n / x;
    const cd zeta(psi, chi), zeta_prev(psi_prev, chi_prev);
    const cd ta = A[n] * minv + nx;
    const cd tb = A[n] * m + nx;
    an[n] = (ta * psi - psi_prev) / (ta * zeta - zeta_prev);
    bn[n] = (tb * psi - psi_prev) / (tb * zeta - zeta_prev);
    const double twon1 = 2.0 * n + 1.0;
    const double psi_next = twon1 / x * psi - psi_prev;
    const double chi_next = twon1 / x * chi - chi_prev;
    psi_prev = psi;
    chi_prev = chi;
    psi = psi_next;
    chi = chi_next;
  }
}

}  // namespace

extern "C" {

int lmc_mie_amplitudes(double x, double m_re, double m_im, int32_t n_mu, const double* mu,
                       double* s1, double* s2) {
  if (!(x > 0.0) || !(m_re > 0.0) || n_mu < 0 || (n_mu > 0 && (!mu || !s1 || !s2)))
    return LMC_ERR_INVALID;
  const bool flip = m_im > 0.0;
  const cd m(m_re, flip ? -m_im : m_im);
  const int ntrm = mie_terms(x);
  std::vector<cd> an, bn;
  mie_coefficients(x, m, ntrm, an, bn);
  for (int j = 0; j < n_mu; ++j) {
    const double c = mu[j];
    double pi_nm1 = 0.0, pi_n = 1.0;
    cd S1 = 0.0, S2 = 0.0;
    for (int n = 1; n <= ntrm; ++n) {
      const double tau_n = n * c * pi_n - (n + 1) * pi_nm1;
      const double f = (2.0 * n + 1.0) / (n * (n + 1.0));
      S1 += f * (an[n] * pi_n + bn[n] * tau_n);
      S2 += f * (an[n] * tau_n + bn[n] * pi_n);
      const double pi_np1 = ((2.0 * n + 1.0) * c * pi_n - (n + 1.0) * pi_nm1) / n;
      pi_nm1 = pi_n;
      pi_n = pi_np1;
    }
    if (flip) {
      S1 = std::conj(S1);
      S2 = std::conj(S2);
    }
    s1[2 * j] = S1.real();
    s1[2 * j + 1] = S1.imag();
    s2[2 * j] = S2.real();
    s2[2 * j + 1] = S2.imag();
  }
  return LMC_OK;
}

int lmc_mie_info(double x, double m_re, double m_im, double* qext, double* qsca, double* g) {
  if (!(x > 0.0) || !(m_re > 0.0)) return LMC_ERR_INVALID;
  const cd m(m_re, m_im > 0.0 ? -m_im : m_im);
  const int ntrm = mie_terms(x);
  std::vector<cd> an, bn;
  mie_coefficients(x, m, ntrm, an, bn);
  double sca = 0.0, ext = 0.0, gq = 0.0;
  for (int n = 1; n <= ntrm; ++n) {
    const double t = 2.0 * n + 1.0;
    sca += t * (std::norm(an[n]) + std::norm(bn[n]));
    ext += t * (an[n].real() + bn[n].real());
    gq += t / (n * (n + 1.0)) * (an[n] * std::conj(bn[n])).real();
    if (n < ntrm)
      gq += n * (n + 2.0) / (n + 1.0) *
            (an[n] * std::conj(an[n + 1]) + bn[n] * std::conj(bn[n + 1])).real();
  }
  const double k = 2.0 / (x * x);
  if (qsca) *qsca = k * sca;
  if (qext) *qext = k * ext;
  if (g) *g = (sca > 0.0) ? (2.0 * gq) / sca : 0.0;  // gqsc/qsca, dmiev.cpp:80
  return LMC_OK;
}

// SamplingTable::initialize (src/table.cpp:6-48) on pdf values at the nDiv+1
// nodes min + i*step: normalise by the left-Riemann sum, accumulate, and
// divide by the last entry (node 0 carries pdf[0]*step).
int lmc_sampling_cdf(int32_t n_div, double min_val, double max_val, const double* pdf,
                     double* cdf) {
  if (n_div <= 0 || !(min_val < max_val) || !pdf || !cdf) return LMC_ERR_INVALID;
  const double step = (max_val - min_val) / n_div;
  double sum = 0.0;
  for (int i = 0; i <= n_div; ++i) sum += pdf[i] * step;
  if (!(sum > 0.0) || !std::isfinite(sum)) return LMC_ERR_INVALID;
  double acc = 0.0;
  for (int i = 0; i <= n_div; ++i) {
    acc += (pdf[i] / sum) * step;
    cdf[i] = acc;
  }
  const double last = cdf[n_div];
  for (int i = 0; i <= n_div; ++i) cdf[i] /= last;
  return LMC_OK;
}

}  // extern "C"
