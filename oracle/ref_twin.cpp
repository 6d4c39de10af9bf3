// ref_twin.cpp -- C-ABI wrapper over the reference's own CPU twin of the
// per-Gaussian update equations, compiled from the reference sources where
// they lie (see oracle/Makefile target `ref`). TEST INFRASTRUCTURE ONLY: it
// pins the oracle's scalar functions against the reference's code:
//   gsl_sf_psi                           src/external/gsl/psi.c:834
//   GaussianMixtureModel::expectedLogDetCovariance   GaussianMixtureModel.h:229-238
//   GaussianMixtureModel::expectedLogResponsivity    GaussianMixtureModel.h:241-243
//   GaussianMixtureModel::updateGaussianParameters   GaussianMixtureModel.h:252-272
//   GaussianMixtureModel::regularizePrecisionMatrix  GaussianMixtureModel.h:283-368
//   GaussianMixtureModel::updatePriors               GaussianMixtureModel.cpp:892-910
// The output library lives in oracle/_ref/ and is never shipped.
#include "GaussianMixtureModel.h"

static void load_W(Matrix<double, 3, 3>& M, const double* w) {
  M(0, 0) = w[0];
  M(0, 1) = M(1, 0) = w[1];
  M(0, 2) = M(2, 0) = w[2];
  M(1, 1) = w[3];
  M(1, 2) = M(2, 1) = w[4];
  M(2, 2) = w[5];
}
static void store_W(const Matrix<double, 3, 3>& M, double* w) {
  w[0] = M(0, 0);
  w[1] = M(0, 1);
  w[2] = M(0, 2);
  w[3] = M(1, 1);
  w[4] = M(1, 2);
  w[5] = M(2, 2);
}

extern "C" {

double ref_psi(double x) { return gsl_sf_psi(x); }

double ref_expected_logdet(const double* W, double nu) {
  GaussianMixtureModel g(0);
  load_W(g.W_k, W);
  g.nu_k = nu;
  return g.expectedLogDetCovariance();
}

double ref_expected_logresp(double alpha, double total_alpha) {
  GaussianMixtureModel g(0);
  g.alpha_k = alpha;
  return g.expectedLogResponsivity(total_alpha);
}

// prior: [beta_o, nu_o, alpha_o, m_o(3), W_o(6)] ; in Nk, xk(3), Sk (packed
// 6, normalised covariance: sum w (x-xk)(x-xk)^T / Nk); out: [beta_k, nu_k,
// alpha_k, m_k(3), W_k(6), W_o(6)]
void ref_update_regularize(const double* prior, double Nk, const double* xk,
                           const double* Sk, int do_regularize, int w4dof,
                           const float* scale, double lmin, double lmax,
                           double emax, double* out) {
  regularizePrecisionMatrixConstants::setConstants(lmin, lmax, emax);
  GaussianMixtureModel::setScale(scale);
  GaussianMixtureModel g(0);
  g.beta_o = prior[0];
  g.nu_o = prior[1];
  g.alpha_o = prior[2];
  for (int i = 0; i < 3; i++) g.m_o(i) = prior[3 + i];
  load_W(g.W_o, prior + 6);
  g.N_k = Nk;
  g.fixed = false;
  Matrix<double, 3, 1> x;
  for (int i = 0; i < 3; i++) x(i) = xk[i];
  Matrix<double, 3, 3> S;
  load_W(S, Sk);
  g.updateGaussianParameters(x, S);
  if (do_regularize) g.regularizePrecisionMatrix(w4dof != 0);
  out[0] = g.beta_k;
  out[1] = g.nu_k;
  out[2] = g.alpha_k;
  for (int i = 0; i < 3; i++) out[3 + i] = g.m_k(i);
  store_W(g.W_k, out + 6);
  store_W(g.W_o, out + 12);
}

void ref_regularize(double* W, double nu, int w4dof, const float* scale,
                    double lmin, double lmax, double emax) {
  regularizePrecisionMatrixConstants::setConstants(lmin, lmax, emax);
  GaussianMixtureModel::setScale(scale);
  GaussianMixtureModel g(0);
  load_W(g.W_k, W);
  g.nu_k = nu;
  g.regularizePrecisionMatrix(w4dof != 0);
  store_W(g.W_k, W);
}

// state: [beta_k, nu_k, alpha_k, m_k(3), W_k(6), N_k, alpha_o]; out prior
// [beta_o, nu_o, alpha_o, m_o(3), W_o(6)]
void ref_update_priors(const double* state, double beta_pct, double nu_pct,
                       double alpha_pct, double alpha_total, double* out) {
  GaussianMixtureModel g(0);
  g.beta_k = state[0];
  g.nu_k = state[1];
  g.alpha_k = state[2];
  for (int i = 0; i < 3; i++) g.m_k(i) = state[3 + i];
  load_W(g.W_k, state + 6);
  g.N_k = state[12];
  g.alpha_o = state[13];
  g.updatePriors(beta_pct, nu_pct, alpha_pct, alpha_total);
  out[0] = g.beta_o;
  out[1] = g.nu_o;
  out[2] = g.alpha_o;
  for (int i = 0; i < 3; i++) out[3 + i] = g.m_o(i);
  store_W(g.W_o, out + 6);
}
}
