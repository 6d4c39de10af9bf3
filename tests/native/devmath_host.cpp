// Host build of the kernels' per-Gaussian math (tgmm_amd/csrc/tgmm_device.h),
// exported for CPU tests that compare it with the oracle and the reference's
// CPU twin without a GPU. Test infrastructure only.
#include "../../tgmm_amd/csrc/tgmm_device.h"

extern "C" {
double dm_digamma(double x) { return tgmm::digamma(x); }
void dm_eig3(const double* w, double* d, double* v) { tgmm::eig3sym(w, d, v); }
void dm_regularize(double* W, double nu, int w4dof, const float* scale, double lmin, double lmax,
                   double emax) {
  tgmm::regularize_W(W, nu, w4dof, scale, lmin, lmax, emax);
}
void dm_inv3(const double* w, double* o) { tgmm::inv3(w, o); }
int dm_pair_significant(float r) { return tgmm::pair_significant(r) ? 1 : 0; }
}
