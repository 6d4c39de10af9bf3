// tgmm_device.h -- per-Gaussian FP64 math for the gfx950 kernels: digamma,
// 3x3 symmetric determinant / inverse, closed-form symmetric eigensystem,
// precision-matrix regularisation and the Normal-Wishart-Dirichlet posterior.
// Packed symmetric layout [xx, xy, xz, yy, yz, zz] throughout.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace tgmm {

// Digamma by Bernardo's AS 103 with Minka's refinements, as the reference
// evaluates it (GMEMupdateCUDA.cu:494-525): series below 1e-6, upward
// recurrence to 12, asymptotic expansion beyond; reflection for x < 0;
// -1e300 at non-positive integers.
__host__ __device__ inline double digamma_pos(double x) {
  if (x <= 1e-6) return -0.57721566490153286 - 1.0 / x + 1.6449340668482264365 * x;
  double acc = 0.0;
  while (x < 12.0) {
    acc -= 1.0 / x;
    x += 1.0;
  }
  const double r = 1.0 / x;
  acc += log(x) - 0.5 * r;
  const double r2 = r * r;
  acc -= r2 * (1.0 / 12.0 -
               r2 * (1.0 / 120.0 - r2 * (1.0 / 252.0 - r2 * (1.0 / 240.0 - r2 * (1.0 / 132.0)))));
  return acc;
}

__host__ __device__ inline double digamma(double x) {
  const double pi = 3.14159265358979323846;
  if (x <= 0.0 && floor(x) == x) return -1e300;
  if (x < 0.0) return digamma_pos(1.0 - x) + pi / tan(-pi * x);
  return digamma_pos(x);
}

// Whether a stored responsibility enters the M-step. The reference compares
// the float value with the double literal 1e-3 (GMEM.cu:1350, and :1195 in
// voxel mode), so the comparison is in double and r == 1e-3f (0.00100000005)
// is kept.
__host__ __device__ inline bool pair_significant(float r) { return (double)r > 1e-3; }

__host__ __device__ inline double det3(const double* w) {
  return w[0] * (w[3] * w[5] - w[4] * w[4]) - w[1] * (w[1] * w[5] - w[2] * w[4]) +
         w[2] * (w[1] * w[4] - w[2] * w[3]);
}

// Adjugate inverse; |det| < 1e-16 -> every entry 1e300 (GMEM.cu:170-192).
__host__ __device__ inline void inv3(const double* w, double* o) {
  const double det = det3(w);
  if (fabs(det) < 1e-16) {
    for (int i = 0; i < 6; i++) o[i] = 1e300;
    return;
  }
  const double id = 1.0 / det;
  o[0] = (w[3] * w[5] - w[4] * w[4]) * id;
  o[1] = (w[2] * w[4] - w[1] * w[5]) * id;
  o[2] = (w[1] * w[4] - w[2] * w[3]) * id;
  o[3] = (w[0] * w[5] - w[2] * w[2]) * id;
  o[4] = (w[1] * w[2] - w[0] * w[4]) * id;
  o[5] = (w[0] * w[3] - w[1] * w[1]) * id;
}

__host__ __device__ inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

__host__ __device__ inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Unit eigenvector of symmetric w for eigenvalue lam: the longest cross
// product of two rows of (w - lam I). Returns its squared length before
// normalisation (0 when the rows are all parallel / zero).
__host__ __device__ inline double eigvec_cross(const double* w, double lam, double* v) {
  const double r0[3] = {w[0] - lam, w[1], w[2]};
  const double r1[3] = {w[1], w[3] - lam, w[4]};
  const double r2[3] = {w[2], w[4], w[5] - lam};
  double c0[3], c1[3], c2[3];
  cross3(r0, r1, c0);
  cross3(r0, r2, c1);
  cross3(r1, r2, c2);
  const double n0 = dot3(c0, c0), n1 = dot3(c1, c1), n2 = dot3(c2, c2);
  const double* best = c0;
  double nb = n0;
  if (n1 > nb) {
    best = c1;
    nb = n1;
  }
  if (n2 > nb) {
    best = c2;
    nb = n2;
  }
  if (nb > 0.0) {
    const double s = 1.0 / sqrt(nb);
    v[0] = best[0] * s;
    v[1] = best[1] * s;
    v[2] = best[2] * s;
  }
  return nb;
}

// Any unit vector orthogonal to unit a.
__host__ __device__ inline void any_orthogonal(const double* a, double* o) {
  const double e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  int i = 0;
  if (fabs(a[1]) < fabs(a[i])) i = 1;
  if (fabs(a[2]) < fabs(a[i])) i = 2;
  double c[3];
  cross3(a, e[i], c);
  const double s = 1.0 / sqrt(dot3(c, c));
  o[0] = c[0] * s;
  o[1] = c[1] * s;
  o[2] = c[2] * s;
}

// Symmetric 3x3 eigensystem. Eigenvalues in the reference's eig3 order
// (GMEM.cu:198-262): the trigonometric closed form gives (largest, smallest,
// middle); when |det| < 1e-24 the first is exactly 0 and the other two are
// the roots of the remaining quadratic, larger first. Eigenvectors: the one
// of the eigenvalue farthest from the other two comes from row cross
// products, the second is made orthogonal to it, the third completes a
// right-handed basis (any orthonormal basis of a repeated eigenspace gives
// the same V diag(d) V^T).
__host__ __device__ inline void eig3sym(const double* w, double* d, double* v) {
  const double det = det3(w);
  if (fabs(det) < 1e-24) {
    const double tr = w[0] + w[3] + w[5];
    const double pm = w[0] * w[3] + w[0] * w[5] + w[3] * w[5] - w[1] * w[1] - w[2] * w[2] -
                      w[4] * w[4];
    double disc = tr * tr - 4.0 * pm;
    disc = disc < 0.0 ? 0.0 : sqrt(disc);
    d[0] = 0.0;
    d[1] = 0.5 * (tr + disc);
    d[2] = 0.5 * (tr - disc);
  } else {
    const double m = (w[0] + w[3] + w[5]) / 3.0;
    const double a = w[0] - m, b = w[3] - m, c = w[5] - m;
    const double q = 0.5 * (a * (b * c - w[4] * w[4]) - w[1] * (w[1] * c - w[2] * w[4]) +
                            w[2] * (w[1] * w[4] - w[2] * b));
    const double p = (2.0 * (w[1] * w[1] + w[2] * w[2] + w[4] * w[4]) + a * a + b * b + c * c) / 6.0;
    double phi;
    const double sp = sqrt(p);
    const double x = (p > 0.0) ? q / (p * sp) : 0.0;
    if (x <= -1.0)
      phi = 3.14159265358979323846 / 3.0;
    else if (x >= 1.0)
      phi = 0.0;
    else
      phi = acos(x) / 3.0;
    const double cs = cos(phi), sn = sin(phi);
    d[0] = m + 2.0 * sp * cs;
    d[1] = m - sp * (cs + 1.7320508075688772 * sn);
    d[2] = m - sp * (cs - 1.7320508075688772 * sn);
  }
  // most isolated eigenvalue first
  const double g0 = fmin(fabs(d[0] - d[1]), fabs(d[0] - d[2]));
  const double g1 = fmin(fabs(d[1] - d[0]), fabs(d[1] - d[2]));
  const double g2 = fmin(fabs(d[2] - d[0]), fabs(d[2] - d[1]));
  int i0 = 0;
  if (g1 > g0) i0 = 1;
  if (g2 > (i0 == 0 ? g0 : g1)) i0 = 2;
  const double l0 = (i0 == 0) ? d[0] : (i0 == 1 ? d[1] : d[2]);
  const double l1 = (i0 == 0) ? d[1] : (i0 == 1 ? d[2] : d[0]);
  double a0[3], a1[3], a2[3];
  if (!(eigvec_cross(w, l0, a0) > 0.0)) {
    // all three equal (w is a multiple of I): identity basis
    for (int i = 0; i < 9; i++) v[i] = (i % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  double n1 = eigvec_cross(w, l1, a1);
  if (n1 > 0.0) {
    const double pr = dot3(a1, a0);
    for (int k = 0; k < 3; k++) a1[k] -= pr * a0[k];
    const double nn = dot3(a1, a1);
    if (nn > 1e-20) {
      const double s = 1.0 / sqrt(nn);
      for (int k = 0; k < 3; k++) a1[k] *= s;
    } else {
      n1 = 0.0;
    }
  }
  if (!(n1 > 0.0)) any_orthogonal(a0, a1);
  cross3(a0, a1, a2);
  // vectors of eigenvalues i0, i0+1, i0+2 (mod 3); constant indices only
  for (int k = 0; k < 3; k++) {
    const double x0 = a0[k], x1 = a1[k], x2 = a2[k];
    v[k] = (i0 == 0) ? x0 : (i0 == 1 ? x2 : x1);
    v[3 + k] = (i0 == 0) ? x1 : (i0 == 1 ? x0 : x2);
    v[6 + k] = (i0 == 0) ? x2 : (i0 == 1 ? x1 : x0);
  }
}

// Eigenvalue clamping of the scale-normalised precision (GMEM.cu:1693-1809).
__host__ __device__ inline void regularize_W(double* W, double nu, int w4dof, const float* scale,
                                             double lmin, double lmax, double emax) {
  const double hi = lmax / nu, lo = lmin / nu;
  double ss[6];
  {
    int c = 0;
    for (int i = 0; i < 3; i++)
      for (int j = i; j < 3; j++) ss[c++] = (double)(scale[i] * scale[j]);
  }
  double w[6];
  for (int i = 0; i < 6; i++) w[i] = W[i] / ss[i];
  if (w4dof) {
    w[2] = 0.0;
    w[4] = 0.0;
  }
  double d[3], v[9];
  eig3sym(w, d, v);
  for (int i = 0; i < 3; i++) d[i] = d[i] > hi ? hi : d[i];
  for (int i = 0; i < 3; i++) d[i] = d[i] < lo ? lo : d[i];
  {
    double r = d[0] / d[1];
    if (r > emax)
      d[0] = emax * d[1];
    else if (1.0 / r > emax)
      d[1] = emax * d[0];
    r = d[0] / d[2];
    if (r > emax)
      d[0] = emax * d[2];
    else if (1.0 / r > emax)
      d[2] = emax * d[0];
    r = d[1] / d[2];
    if (r > emax)
      d[1] = emax * d[2];
    else if (1.0 / r > emax)
      d[2] = emax * d[1];
  }
  double rc[6];
  {
    int c = 0;
    for (int a = 0; a < 3; a++)
      for (int b = a; b < 3; b++) {
        rc[c++] = d[0] * v[a] * v[b] + d[1] * v[3 + a] * v[3 + b] + d[2] * v[6 + a] * v[6 + b];
      }
  }
  w[0] = rc[0];
  w[1] = rc[1];
  w[3] = rc[3];
  w[5] = rc[5];
  if (!w4dof) {
    w[2] = rc[2];
    w[4] = rc[4];
  }
  for (int i = 0; i < 6; i++) W[i] = w[i] * ss[i];
}

}  // namespace tgmm
