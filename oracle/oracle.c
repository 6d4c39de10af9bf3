/*
 * oracle.c -- FP64 CPU restatement of the reference's supervoxel-mode VI-EM.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library. Written from the reference's equations, independently of
 * the HIP code:
 *   kNN            knnCuda.cu:61-144 (the scan in index order, early exits,
 *                  slot replacement rule), GMEM.cu:527-539
 *   Sigma alpha    GMEM.cu:579-604, 2765-2799
 *   expectations   GMEM.cu:553-628 (K7, K8)
 *   E-step         GMEM.cu:2244-2336 (K9), per voxel, FP64
 *   M-step         GMEM.cu:1286-1455 (K13), per voxel, FP64, r > 1e-3 rule
 *   posterior      GMEM.cu:1601-1677 (K14), FP64 throughout
 *   regularise     GMEM.cu:1693-1809 (K15), eig3 eigenvalue order
 *   death          GMEM.cu:1813-1843 (K16)
 *   log-likelihood GMEM.cu:1913-1978 (K17), per voxel, FP64
 *   convergence    GMEM.cu:4143-4148; supervoxel assignment GMEM.cu:4328-4350
 *   local LL       GMEM.cu:2952-3050 with the kernel :1989-2103 (lists of
 *                  supervoxels, the cell-division likelihood-ratio input)
 *   supervoxel kNN temporalLogicalRules/knnCUDA/knnCuda.cu:43-232 (the
 *                  scan in reference order, slot replacement rule)
 *   marshalling    mainTrackingGaussianMixtureModel.cpp:598-669, 774-870,
 *                  supervoxel.cpp:1410-1449, mylib array.p:557-576,
 *                  2708-2820 (raw uint16 voxels to the EM's float inputs)
 * and, with tgmm_em_opts.voxel_mode, the per-voxel EM of
 * GMEMvariationalInferenceCUDA (GMEM.cu:3420-3973):
 *   kNN            each voxel is a query (knnCUDAinPlace on queryCUDA, :3576)
 *   E-step         GMEM.cu:2177-2242 (GMEMcomputeRnkKernel): cut at -100,
 *                  stored r multiplied by the voxel's intensity
 *   M-step         GMEM.cu:1142-1284 (GMEMcomputeXkNkSkKernelTr): the stored
 *                  value is the weight, pairs with weight > 1e-3
 *   log-likelihood GMEM.cu:1845-1912 (GMEMcomputeLikelihoodKernel)
 *   no supervoxel assignment (the records keep supervoxelIdx/Num)
 *
 * Declared deviations from the reference (identical in the HIP code):
 *   - kNN slots never filled (fewer than K living Gaussians) hold index -1
 *     and r = 0 (the reference leaves them uninitialised, knnCuda.cu:74).
 *   - E-step normalisation is max-subtracted log-sum-exp (same value as
 *     1/sum_j exp(L_j - L_k), GMEM.cu:2327-2335, without overflow).
 *   - M-step statistics, N_k and the posterior are FP64 (the reference
 *     accumulates in FP32 shared memory and narrows N_k, beta_o, x_k, m_o to
 *     float, GMEM.cu:1381-1394, 1614-1629).
 *   - log-likelihood sums every voxel of a supervoxel (the reference's
 *     `cache[tid] =` keeps only the last term per thread for supervoxels of
 *     more than 1024 voxels, GMEM.cu:1961-1962; the same in the local
 *     kernel, :2067-2068), and sums in FP64 (the reference's per-block and
 *     per-list sums are FP32).
 *   - death NaN test reads W_k[0..5] and E[ln det Lambda] (the reference's
 *     7-element loop over W_k reaches that field, GMEM.cu:1823-1824).
 *
 * Eigen-decomposition uses cyclic Jacobi rotations (not the reference's
 * closed form), then orders the eigenvalues as eig3 does (GMEM.cu:198-262):
 * non-singular -> (largest, smallest, middle); |det| < 1e-24 -> (0, larger,
 * smaller). Digamma uses upward recurrence to x >= 10 and the asymptotic
 * Bernoulli series (Abramowitz & Stegun 6.3.6 / 6.3.18), with the reference's
 * special value -1e300 at non-positive integers (GMEM.cu:515-518).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;

void orc_set_threads(int n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static int is_dead(const GaussianMixtureModelCUDA* g) { return g->m_o[0] < -1e31; }

/* ------------------------------------------------------------------------ */
double orc_digamma(double x) {
  static const double PI = 3.14159265358979323846;
  if (x <= 0.0 && floor(x) == x) return -1e300;
  if (x < 0.0) return orc_digamma(1.0 - x) + PI / tan(-PI * x);
  double acc = 0.0;
  while (x < 10.0) {
    acc -= 1.0 / x;
    x += 1.0;
  }
  /* psi(x) ~ ln x - 1/(2x) - sum_n B_2n / (2n x^2n) */
  const double x2 = 1.0 / (x * x);
  double series =
      x2 * (1.0 / 12 -
            x2 * (1.0 / 120 -
                  x2 * (1.0 / 252 -
                        x2 * (1.0 / 240 -
                              x2 * (1.0 / 132 -
                                    x2 * (691.0 / 32760 - x2 * (1.0 / 12)))))));
  return acc + log(x) - 0.5 / x - series;
}

/* packed symmetric [xx, xy, xz, yy, yz, zz] */
double orc_det3(const double w[6]) {
  const double a = w[0], b = w[1], c = w[2], d = w[3], e = w[4], f = w[5];
  return a * (d * f - e * e) - b * (b * f - c * e) + c * (b * e - c * d);
}

/* adjugate / det with the reference's singular sentinel (GMEM.cu:170-192) */
int orc_inv3(const double w[6], double o[6]) {
  const double det = orc_det3(w);
  if (fabs(det) < 1e-16) {
    for (int i = 0; i < 6; i++) o[i] = 1e300;
    return 1;
  }
  const double a = w[0], b = w[1], c = w[2], d = w[3], e = w[4], f = w[5];
  o[0] = (d * f - e * e) / det;
  o[1] = (c * e - b * f) / det;
  o[2] = (b * e - c * d) / det;
  o[3] = (a * f - c * c) / det;
  o[4] = (b * c - a * e) / det;
  o[5] = (a * d - b * b) / det;
  return 0;
}

/* cyclic Jacobi; A full 3x3 row-major in, eigenvalues in ev[], vectors as
 * columns of V (row-major) */
static void jacobi3(double A[3][3], double ev[3], double V[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) V[i][j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 64; sweep++) {
    double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
    double scale = fabs(A[0][0]) + fabs(A[1][1]) + fabs(A[2][2]);
    if (off == 0.0 || off <= 1e-18 * scale) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        if (A[p][q] == 0.0) continue;
        double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) /
                   (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; k++) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; k++) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; k++) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 3; i++) ev[i] = A[i][i];
}

/* d[i] with eigenvector v[3i..3i+2], ordered like eig3 (GMEM.cu:198-262) */
void orc_eig3_ordered(const double w[6], double d[3], double v[9]) {
  double A[3][3] = {{w[0], w[1], w[2]}, {w[1], w[3], w[4]}, {w[2], w[4], w[5]}};
  double ev[3], V[3][3];
  jacobi3(A, ev, V);
  int order[3];
  const int singular = fabs(orc_det3(w)) < 1e-24;
  if (!singular) {
    int imax = 0, imin = 0;
    for (int i = 1; i < 3; i++) {
      if (ev[i] > ev[imax]) imax = i;
      if (ev[i] < ev[imin]) imin = i;
    }
    if (imax == imin) imin = (imax + 1) % 3;
    order[0] = imax;
    order[1] = imin;
    order[2] = 3 - imax - imin;
  } else {
    int iz = 0;
    for (int i = 1; i < 3; i++)
      if (fabs(ev[i]) < fabs(ev[iz])) iz = i;
    int a = (iz + 1) % 3, b = (iz + 2) % 3;
    order[0] = iz;
    order[1] = ev[a] >= ev[b] ? a : b;
    order[2] = ev[a] >= ev[b] ? b : a;
  }
  for (int i = 0; i < 3; i++) {
    d[i] = ev[order[i]];
    for (int k = 0; k < 3; k++) v[3 * i + k] = V[k][order[i]];
  }
  if (singular) d[0] = 0.0;
}

/* GMEM.cu:1693-1809 */
void orc_regularize_W(double W[6], double nu, int w4dof, const float scale[3],
                      double lambda_min, double lambda_max, double max_exc) {
  const double auxMax = lambda_max / nu, auxMin = lambda_min / nu;
  double w[6];
  double ss[6];
  int c = 0;
  for (int i = 0; i < 3; i++)
    for (int j = i; j < 3; j++) ss[c++] = (double)(scale[i] * scale[j]);
  for (int i = 0; i < 6; i++) w[i] = W[i] / ss[i];
  if (w4dof) {
    w[2] = 0.0;
    w[4] = 0.0;
  }
  double d[3], v[9];
  orc_eig3_ordered(w, d, v);
  for (int i = 0; i < 3; i++)
    if (d[i] > auxMax) d[i] = auxMax;
  for (int i = 0; i < 3; i++)
    if (d[i] < auxMin) d[i] = auxMin;
  static const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
  for (int p = 0; p < 3; p++) {
    const int a = pa[p], b = pb[p];
    const double r = d[a] / d[b];
    if (r > max_exc)
      d[a] = max_exc * d[b];
    else if (1.0 / r > max_exc)
      d[b] = max_exc * d[a];
  }
  /* reconstruct V diag(d) V^T */
  double r6[6];
  c = 0;
  for (int a = 0; a < 3; a++)
    for (int b = a; b < 3; b++) {
      double s = 0.0;
      for (int i = 0; i < 3; i++) s += d[i] * v[3 * i + a] * v[3 * i + b];
      r6[c++] = s;
    }
  w[0] = r6[0];
  w[1] = r6[1];
  w[3] = r6[3];
  w[5] = r6[5];
  if (!w4dof) {
    w[2] = r6[2];
    w[4] = r6[4];
  }
  for (int i = 0; i < 6; i++) W[i] = w[i] * ss[i];
}

/* GMEM.cu:1601-1677 (FP64 throughout) */
void orc_posterior(GaussianMixtureModelCUDA* g, double Nk, const double xk[3],
                   const double Ck[6]) {
  const double beta_o = g->beta_o;
  g->beta_k = beta_o + Nk;
  g->nu_k = g->nu_o + Nk;
  g->alpha_k = fmax(g->alpha_o + Nk, 0.0);
  for (int i = 0; i < 3; i++) g->m_k[i] = (beta_o * g->m_o[i] + Nk * xk[i]) / g->beta_k;
  if (g->W_o[5] < 1e-8) g->W_o[5] = 0.5 * (g->W_o[0] + g->W_o[3]);
  double A[6];
  orc_inv3(g->W_o, A);
  const double dx[3] = {xk[0] - g->m_o[0], xk[1] - g->m_o[1], xk[2] - g->m_o[2]};
  const double f = beta_o / (beta_o + Nk);
  int c = 0;
  for (int a = 0; a < 3; a++)
    for (int b = a; b < 3; b++) {
      A[c] += Ck[c] + Nk * f * dx[a] * dx[b];
      c++;
    }
  orc_inv3(A, g->W_k);
}

/* ------------------------------------------------------------------------ */
/* The EM's kNN, knnKernel / knnKernelNoConstantMemory (knnCuda.cu:61-144,
 * 147-246) on the float means copied out of the records (GMEM.cu:527-539),
 * replayed per query: the Gaussians are scanned in index order; K slots of
 * (distance, index) start at 1e32 and the threshold at 1e32; the distance is
 * ((c0 - m0) s0)^2 + ((c1 - m1) s1)^2 + ((c2 - m2) s2)^2 in float, added in
 * that order, and the reference leaves a Gaussian as soon as the x term, the
 * x + y partial sum or the full sum is above the threshold (:116-124); a
 * Gaussian that passes replaces the slot at minPos, then the threshold
 * becomes the largest slot distance and minPos the first slot holding it
 * (findMaxPosition, :49-59). A dead Gaussian (m = -1e32) has an infinite
 * distance and is never taken. Output slot-major ind [K * S] in the
 * reference's slot layout; slots never filled (fewer than K living
 * Gaussians) hold -1 (the reference leaves them uninitialised, :74).
 *
 * Each block of 256 Gaussians first forms every full distance without
 * branches; the scan then applies the exact test only where the full
 * distance is not above the current threshold or is NaN (for finite values
 * the partial sums are non-decreasing, so no other Gaussian can pass). */
void orc_knn(long long S, const float* cent, int G,
             const GaussianMixtureModelCUDA* gm, const float scale[3], int K,
             int* ind) {
  float* mf = (float*)malloc(sizeof(float) * 3 * (size_t)(G > 0 ? G : 1));
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < G; k++) mf[(size_t)i * G + k] = (float)gm[k].m_k[i];
  const float s0 = scale[0], s1 = scale[1], s2 = scale[2];
#pragma omp parallel for schedule(dynamic, 64)
  for (long long s = 0; s < S; s++) {
    float md[TGMM_KMAX];
    int mi[TGMM_KMAX];
    float d[256];
    for (int j = 0; j < K; j++) {
      md[j] = 1e32f;
      mi[j] = -1;
    }
    float thr = 1e32f;
    int pos = 0;
    const float q0 = cent[s], q1 = cent[s + S], q2 = cent[s + 2 * S];
    for (int b0 = 0; b0 < G; b0 += 256) {
      const int n = G - b0 < 256 ? G - b0 : 256;
      int any = 0;
      for (int u = 0; u < n; u++) {
        const float a = (q0 - mf[b0 + u]) * s0;
        const float b = (q1 - mf[(size_t)G + b0 + u]) * s1;
        const float e = (q2 - mf[2 * (size_t)G + b0 + u]) * s2;
        float t = a * a;
        t = t + b * b;
        d[u] = t + e * e;
        any |= !(d[u] > thr);
      }
      if (!any) continue; /* nothing in this block can pass */
      for (int u = 0; u < n; u++) {
        if (d[u] > thr) continue; /* finite: the partial sums are not above it either */
        const int k = b0 + u;
        float a = (q0 - mf[k]) * s0;
        float dist = a * a;
        if (dist > thr) continue;
        a = (q1 - mf[(size_t)G + k]) * s1;
        dist += a * a;
        if (dist > thr) continue;
        a = (q2 - mf[2 * (size_t)G + k]) * s2;
        dist += a * a;
        if (dist > thr) continue;
        md[pos] = dist;
        mi[pos] = k;
        thr = md[0];
        pos = 0;
        for (int j = 1; j < K; j++)
          if (thr < md[j]) {
            thr = md[j];
            pos = j;
          }
      }
    }
    for (int j = 0; j < K; j++) ind[s + (long long)j * S] = mi[j];
  }
  free(mf);
}

double orc_sum_alpha(int G, const GaussianMixtureModelCUDA* gm) {
  double s = 0.0;
  for (int k = 0; k < G; k++) s += gm[k].alpha_k;
  return s;
}

/* GMEM.cu:553-628 */
void orc_expectations(int G, GaussianMixtureModelCUDA* gm, double sum_alpha) {
  const double psiA = orc_digamma(sum_alpha);
#pragma omp parallel for
  for (int k = 0; k < G; k++) {
    GaussianMixtureModelCUDA* g = &gm[k];
    if (is_dead(g)) {
      g->expectedLogDetCovarianceCUDA = -1e32;
      g->expectedLogResponsivityCUDA = -1e32;
    } else {
      double out = 3.0 * log(2.0) + log(orc_det3(g->W_k));
      for (int i = 0; i < 3; i++) out += orc_digamma(0.5 * (g->nu_k + 1.0 - i));
      g->expectedLogDetCovarianceCUDA = out;
      g->expectedLogResponsivityCUDA = orc_digamma(g->alpha_k) - psiA;
    }
  }
}

static double quad(const double W[6], double d0, double d1, double d2) {
  return W[0] * d0 * d0 + W[3] * d1 * d1 + W[5] * d2 * d2 + 2.0 * W[1] * d0 * d1 +
         2.0 * W[2] * d0 * d2 + 2.0 * W[4] * d1 * d2;
}

/* GMEM.cu:2244-2336, per voxel in FP64 */
void orc_estep(long long N, long long S, const float* query,
               const long long* ptr, int G, const GaussianMixtureModelCUDA* gm,
               int K, const int* ind, float* rnk) {
  (void)G;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long s = 0; s < S; s++) {
    double L[TGMM_KMAX];
    int valid[TGMM_KMAX];
    for (int j = 0; j < K; j++) {
      const int k = ind[s + (long long)j * S];
      valid[j] = k >= 0;
      L[j] = 0.0;
      if (!valid[j]) continue;
      const GaussianMixtureModelCUDA* g = &gm[k];
      double acc = 0.0;
      for (long long v = ptr[s]; v < ptr[s + 1]; v++) {
        const double d0 = (double)query[v] - g->m_k[0];
        const double d1 = (double)query[v + N] - g->m_k[1];
        const double d2 = (double)query[v + 2 * N] - g->m_k[2];
        const double maha = 3.0 / g->beta_k + g->nu_k * quad(g->W_k, d0, d1, d2);
        acc += -0.5 * maha + g->expectedLogResponsivityCUDA +
               0.5 * g->expectedLogDetCovarianceCUDA;
      }
      L[j] = acc;
    }
    double mx = -INFINITY;
    for (int j = 0; j < K; j++)
      if (valid[j] && L[j] > mx) mx = L[j];
    double den = 0.0;
    if (mx > -INFINITY)
      for (int j = 0; j < K; j++)
        if (valid[j]) den += exp(L[j] - mx);
    for (int j = 0; j < K; j++) {
      double r = 0.0;
      if (valid[j] && mx > -INFINITY) r = exp(L[j] - mx) / den;
      rnk[s + (long long)j * S] = (float)r;
    }
  }
}

/* GMEM.cu:2177-2242, per voxel in FP64. A candidate whose log score
 * aux = -0.5 (3/beta + nu d^T W d) + E[ln pi] + 0.5 E[ln det] is <= -100
 * counts as zero (:2208); the others are normalised among themselves (the
 * reference's exp(aux) / sum exp(aux), evaluated max-subtracted) and the
 * stored value is r * img (:2216-2222). A voxel with no candidate above the
 * cut keeps r = 0 (the reference's memset, :3104-3106). */
void orc_estep_voxel(long long N, const float* query, const float* img, int G,
                     const GaussianMixtureModelCUDA* gm, int K, const int* ind,
                     float* rnk) {
  (void)G;
#pragma omp parallel for schedule(static, 256)
  for (long long v = 0; v < N; v++) {
    double aux[TGMM_KMAX];
    int keep[TGMM_KMAX];
    double mx = -INFINITY;
    for (int j = 0; j < K; j++) {
      const int k = ind[v + (long long)j * N];
      keep[j] = 0;
      if (k < 0) continue;
      const GaussianMixtureModelCUDA* g = &gm[k];
      const double d0 = (double)query[v] - g->m_k[0];
      const double d1 = (double)query[v + N] - g->m_k[1];
      const double d2 = (double)query[v + 2 * N] - g->m_k[2];
      aux[j] = -0.5 * (3.0 / g->beta_k + g->nu_k * quad(g->W_k, d0, d1, d2)) +
               g->expectedLogResponsivityCUDA + 0.5 * g->expectedLogDetCovarianceCUDA;
      keep[j] = aux[j] > -100.0;
      if (keep[j] && aux[j] > mx) mx = aux[j];
    }
    double den = 0.0;
    for (int j = 0; j < K; j++)
      if (keep[j]) den += exp(aux[j] - mx);
    for (int j = 0; j < K; j++) {
      double r = 0.0;
      if (keep[j] && den > 0.0) r = (double)img[v] * (exp(aux[j] - mx) / den);
      rnk[v + (long long)j * N] = (float)r;
    }
  }
}

/* GMEM.cu:1286-1455: weights r_sk * img_v over pairs with r > 1e-3; centred
 * second pass in FP64 (Ck = sum w (x - xk)(x - xk)^T, unnormalised). The
 * reference compares the float r with the double literal (`auxOrig > 1e-3`,
 * :1350; voxel mode `aux > 1e-3`, :1195), so the comparison is in double: a
 * pair with r == 1e-3f (0.00100000005) is kept. */
void orc_mstep(long long N, long long S, const float* query, const float* img,
               const long long* ptr, int G, int K, const int* ind,
               const float* rnk, double* Nk, double* xk, double* Ck) {
  /* pair lists per Gaussian, slot-major scan order */
  int* cnt = (int*)calloc((size_t)G + 1, sizeof(int));
  for (int j = 0; j < K; j++)
    for (long long s = 0; s < S; s++) {
      const long long p = s + (long long)j * S;
      if (ind[p] >= 0 && (double)rnk[p] > 1e-3) cnt[ind[p] + 1]++;
    }
  for (int k = 0; k < G; k++) cnt[k + 1] += cnt[k];
  long long* list = (long long*)malloc(sizeof(long long) * (size_t)(cnt[G] > 0 ? cnt[G] : 1));
  int* fill = (int*)calloc((size_t)G + 1, sizeof(int));
  for (int j = 0; j < K; j++)
    for (long long s = 0; s < S; s++) {
      const long long p = s + (long long)j * S;
      if (ind[p] >= 0 && (double)rnk[p] > 1e-3) list[cnt[ind[p]] + fill[ind[p]]++] = p;
    }
#pragma omp parallel for schedule(dynamic, 16)
  for (int k = 0; k < G; k++) {
    double n = 0.0, sx[3] = {0, 0, 0};
    for (int e = cnt[k]; e < cnt[k + 1]; e++) {
      const long long p = list[e];
      const long long s = p % S;
      const double r = (double)rnk[p];
      for (long long v = ptr[s]; v < ptr[s + 1]; v++) {
        const double w = r * (double)img[v];
        n += w;
        sx[0] += w * query[v];
        sx[1] += w * query[v + N];
        sx[2] += w * query[v + 2 * N];
      }
    }
    double m[3] = {0, 0, 0}, C[6] = {0, 0, 0, 0, 0, 0};
    if (n > 0.0) {
      for (int i = 0; i < 3; i++) m[i] = sx[i] / n;
      for (int e = cnt[k]; e < cnt[k + 1]; e++) {
        const long long p = list[e];
        const long long s = p % S;
        const double r = (double)rnk[p];
        for (long long v = ptr[s]; v < ptr[s + 1]; v++) {
          const double w = r * (double)img[v];
          const double d[3] = {query[v] - m[0], query[v + N] - m[1], query[v + 2 * N] - m[2]};
          int c = 0;
          for (int a = 0; a < 3; a++)
            for (int b = a; b < 3; b++) C[c++] += w * d[a] * d[b];
        }
      }
    }
    Nk[k] = n;
    for (int i = 0; i < 3; i++) xk[3 * k + i] = m[i];
    for (int i = 0; i < 6; i++) Ck[6 * k + i] = C[i];
  }
  free(cnt);
  free(fill);
  free(list);
}

void orc_update(int G, GaussianMixtureModelCUDA* gm, const double* Nk,
                const double* xk, const double* Ck, const tgmm_em_opts* o) {
#pragma omp parallel for
  for (int k = 0; k < G; k++) {
    GaussianMixtureModelCUDA* g = &gm[k];
    if (is_dead(g) || g->fixed) continue;
    orc_posterior(g, Nk[k], xk + 3 * k, Ck + 6 * k);
    orc_regularize_W(g->W_k, g->nu_k, o->w4dof, o->scale, o->lambda_min,
                     o->lambda_max, o->max_excentricity);
  }
}

/* GMEM.cu:1813-1843 */
void orc_death(int G, GaussianMixtureModelCUDA* gm, double sum_alpha) {
  for (int k = 0; k < G; k++) {
    GaussianMixtureModelCUDA* g = &gm[k];
    int nanW = isnan(g->expectedLogDetCovarianceCUDA);
    for (int i = 0; i < 6; i++) nanW |= isnan(g->W_k[i]);
    if (nanW || (g->alpha_k / sum_alpha < 1e-5 && !is_dead(g) && !g->fixed)) {
      g->alpha_k = 0.0;
      for (int i = 0; i < 3; i++) {
        g->m_o[i] = -1e32;
        g->m_k[i] = -1e32;
      }
      g->W_k[0] = 1.0;
      g->W_k[1] = 0.0;
      g->W_k[2] = 0.0;
      g->W_k[3] = 1.0;
      g->W_k[4] = 0.0;
      g->W_k[5] = 1.0;
    }
  }
}

/* One supervoxel's term of the log-likelihood (GMEM.cu:1940-1959, the same
 * per-voxel expression in the local kernel :2038-2078): the intensity-weighted
 * log of the mixture density over the supervoxel's candidates, each term
 * sqrt(det W nu^3) / (2 pi)^1.5 * exp(-nu d^T W d / 2) * alpha / sum_alpha,
 * evaluated in that order. Slots with no candidate (ind < 0 or >= G) are
 * skipped; a voxel whose density is not > 0 (all terms underflow, or NaN from
 * an indefinite precision) adds nothing. */
static double sv_loglik(long long N, long long S, long long s, const float* query,
                        const float* img, const long long* ptr, int G,
                        const GaussianMixtureModelCUDA* gm, int K, const int* ind,
                        double sum_alpha) {
  static const double two_pi_3_2 = 15.749609945722419; /* (2 pi)^1.5 */
  double coef[TGMM_KMAX], ratio[TGMM_KMAX];
  int kk[TGMM_KMAX];
  for (int j = 0; j < K; j++) {
    kk[j] = ind[s + (long long)j * S];
    if (kk[j] < 0 || kk[j] >= G) {
      kk[j] = -1;
      continue;
    }
    const GaussianMixtureModelCUDA* g = &gm[kk[j]];
    coef[j] = sqrt(orc_det3(g->W_k) * pow(g->nu_k, 3.0)) / two_pi_3_2;
    ratio[j] = g->alpha_k / sum_alpha;
  }
  double acc = 0.0;
  for (long long v = ptr[s]; v < ptr[s + 1]; v++) {
    double like = 0.0;
    for (int j = 0; j < K; j++) {
      if (kk[j] < 0) continue;
      const GaussianMixtureModelCUDA* g = &gm[kk[j]];
      const double d0 = (double)query[v] - g->m_k[0];
      const double d1 = (double)query[v + N] - g->m_k[1];
      const double d2 = (double)query[v + 2 * N] - g->m_k[2];
      double aux = coef[j];
      aux *= exp(-0.5 * g->nu_k * quad(g->W_k, d0, d1, d2));
      aux *= ratio[j];
      like += aux;
    }
    if (like > 0.0) acc += (double)img[v] * log(like);
  }
  return acc;
}

/* GMEM.cu:1913-1978 + 2834-2873 */
double orc_loglik(long long N, long long S, const float* query,
                  const float* img, const long long* ptr, int G,
                  const GaussianMixtureModelCUDA* gm, int K, const int* ind,
                  double sum_alpha) {
  double* part = (double*)calloc((size_t)(S > 0 ? S : 1), sizeof(double));
#pragma omp parallel for schedule(dynamic, 16)
  for (long long s = 0; s < S; s++)
    part[s] = sv_loglik(N, S, s, query, img, ptr, G, gm, K, ind, sum_alpha);
  double ll = 0.0;
  for (long long s = 0; s < S; s++) ll += part[s];
  free(part);
  return ll;
}

/* calculateLocalLikelihood (GMEM.cu:2952-3050) with
 * GMEMcomputeLocalLikelihoodKernelWithSupervoxels (:1989-2103) and its sum
 * addLocalLikelihoodWithSupervoxels (:2876-2945): sum_alpha over all G
 * records (calculateTotalAlpha, :2765-2799), then for each non-empty list
 * the sum of its supervoxels' terms, a supervoxel listed twice counted
 * twice; an empty list leaves out[i] as given (the reference skips it). */
void orc_local_loglik(long long N, long long S, const float* query,
                      const float* img, const long long* ptr, int G,
                      const GaussianMixtureModelCUDA* gm, int K, const int* ind,
                      int n_lists, const long long* list_ptr,
                      const int* list_idx, double* out) {
  const double sum_alpha = orc_sum_alpha(G, gm);
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < n_lists; i++) {
    if (list_ptr[i + 1] == list_ptr[i]) continue;
    double ll = 0.0;
    for (long long p = list_ptr[i]; p < list_ptr[i + 1]; p++)
      ll += sv_loglik(N, S, list_idx[p], query, img, ptr, G, gm, K, ind, sum_alpha);
    out[i] = ll;
  }
}

/* Input marshalling of one frame, from the supervoxels' raw uint16 voxels to
 * the EM's inputs (mainTrackingGaussianMixtureModel.cpp:598-669, 774-870;
 * supervoxel.cpp:1410-1449). Supervoxel s holds voxels
 * sv_ptr[s] .. sv_ptr[s+1] with linear indices pixel_idx (x fastest, dims
 * d0 d1 d2) and raw intensities raw. Per voxel:
 *   query = (float) coordinates (coord4idx in mylib array.p:557-576);
 *   img   = (float)(uint16)(raw - thr[s]) (local background subtraction,
 *           main.cpp:618-627; Convert_Array_Inplace to FLOAT32, :663-664),
 *           then Scale_Array_To_Range(img, 0, 1) over the whole volume
 *           (mylib array.p:2708-2820): with M, m the volume's max and min
 *           (voxels outside every supervoxel are 0, main.cpp:644-649),
 *           img = (float)((1.0 / (M - m)) * (img - m)) in double, unless M == m.
 * Per supervoxel the centroid of the RAW intensities, accumulated in float in
 * list order (weightedGaussianStatistics, computed before the subtraction):
 * sum(raw * coord) / sum(raw), or 0 when sum(raw) < 1e-5.
 * vol_size = d0 d1 d2; supervoxels are disjoint, so a voxel outside every
 * supervoxel exists iff N < vol_size. Returns 0, or -1 for an index outside
 * the volume (mylib exits there). */
int orc_marshal_frame(long long N, long long S, const unsigned long long* pixel_idx,
                      const unsigned short* raw, const long long* sv_ptr,
                      const unsigned short* thr, const long long dims[3], float* query,
                      float* img, float* cent) {
  const unsigned long long d0 = (unsigned long long)dims[0], d1 = (unsigned long long)dims[1],
                           d2 = (unsigned long long)dims[2];
  const unsigned long long vol = d0 * d1 * d2;
  unsigned hi = 0, lo = 0xFFFFu;
  for (long long s = 0; s < S; s++) {
    float nk = 0.0f, c[3] = {0.0f, 0.0f, 0.0f};
    for (long long v = sv_ptr[s]; v < sv_ptr[s + 1]; v++) {
      unsigned long long id = pixel_idx[v];
      const unsigned long long x = id % d0;
      id /= d0;
      const unsigned long long y = id % d1;
      id /= d1;
      const unsigned long long z = id % d2;
      id /= d2;
      if (id > 0) return -1;
      query[v] = (float)x;
      query[v + N] = (float)y;
      query[v + 2 * N] = (float)z;
      const unsigned short sub = (unsigned short)(raw[v] - thr[s]);
      img[v] = (float)sub;
      if (sub > hi) hi = sub;
      if (sub < lo) lo = sub;
      const float w = (float)raw[v];
      nk += w;
      c[0] += w * (float)x;
      c[1] += w * (float)y;
      c[2] += w * (float)z;
    }
    for (int a = 0; a < 3; a++) cent[s + a * S] = (nk < 1e-5) ? 0.0f : c[a] / nk;
  }
  if (N == 0) return 0;
  if ((unsigned long long)N < vol) lo = 0; /* the zeroed background */
  if (hi == lo) return 0;
  const double f = 1.0 / ((double)hi - (double)lo);
  for (long long v = 0; v < N; v++) img[v] = (float)(f * ((double)img[v] - (double)lo));
  return 0;
}

/* Supervoxel nearest neighbours: knnCUDA_ (temporalLogicalRules/knnCUDA/
 * knnCuda.cu:142-232) with knnKernelNoConstantMemory (:56-140). For each
 * query the references are scanned in index order; K slots of (distance,
 * index) start at (1e32, -1) and the threshold at 1e32. The scaled squared
 * distance is ((q0 - r0) s0)^2 + ((q1 - r1) s1)^2 + ((q2 - r2) s2)^2 in float,
 * added in that order, and the reference leaves a reference as soon as the
 * x term, or the x + y partial sum, is above the threshold (:103-113). For
 * finite coordinates the partial sums are non-decreasing, so only the full
 * sum decides; with a NaN coordinate the full sum is NaN, which is never
 * above the threshold, and the partial sums before it still skip. A
 * reference that passes all three tests replaces the slot at minPos; then
 * the threshold becomes the largest slot distance and minPos the first slot
 * holding it (findMaxPosition, :43-53: a NaN in slot 0 makes the threshold
 * NaN, so the next reference is inserted there). Output slot-major, ind/dist
 * [K * query_nb]; unfilled slots keep index -1 (the reference leaves them
 * uninitialised) and distance 1e32. ref_nb <= 0: dist set to 1e32, ind left
 * as given (:165-171). dist may be NULL. */
void orc_sv_knn(int* ind, float* dist, const float* query, const float* ref,
                long long query_nb, int ref_nb, int K, const float scale[3]) {
  if (ref_nb <= 0) {
    if (dist)
      for (long long i = 0; i < query_nb * K; i++) dist[i] = 1e32f;
    return;
  }
#pragma omp parallel for schedule(static)
  for (long long q = 0; q < query_nb; q++) {
    float md[TGMM_SV_KNN_MAX];
    int mi[TGMM_SV_KNN_MAX];
    for (int j = 0; j < K; j++) {
      md[j] = 1e32f;
      mi[j] = -1;
    }
    float thr = 1e32f;
    int pos = 0;
    const float q0 = query[q], q1 = query[q + query_nb], q2 = query[q + 2 * query_nb];
    for (int r = 0; r < ref_nb; r++) {
      float a = (q0 - ref[r]) * scale[0];
      float d = a * a;
      if (d > thr) continue;
      a = (q1 - ref[r + ref_nb]) * scale[1];
      d += a * a;
      if (d > thr) continue;
      a = (q2 - ref[r + 2 * ref_nb]) * scale[2];
      d += a * a;
      if (d > thr) continue;
      md[pos] = d;
      mi[pos] = r;
      thr = md[0];
      pos = 0;
      for (int j = 1; j < K; j++)
        if (thr < md[j]) {
          thr = md[j];
          pos = j;
        }
    }
    for (int j = 0; j < K; j++) {
      ind[q + j * query_nb] = mi[j];
      if (dist) dist[q + j * query_nb] = md[j];
    }
  }
}

/* loop condition GMEM.cu:4145-4148: 1 = keep iterating */
int orc_converged(double ll, double ll1, double ll2, double ll3, double tol) {
  return !(fabs(ll - ll1) / fabs(ll1) > tol && fabs(ll - ll2) / fabs(ll2) > tol &&
           fabs(ll - ll3) / fabs(ll3) > tol);
}

/* GMEM.cu:4328-4350 */
void orc_assign_supervoxels(long long S, int G, GaussianMixtureModelCUDA* gm,
                            int K, const float* rnk, const int* ind, int cap) {
  for (int k = 0; k < G; k++) gm[k].supervoxelNum = 0;
  for (int j = 0; j < K; j++)
    for (long long s = 0; s < S; s++) {
      const long long p = s + (long long)j * S;
      if (rnk[p] > 0.5f && ind[p] >= 0) {
        GaussianMixtureModelCUDA* g = &gm[ind[p]];
        if (g->supervoxelNum < cap) g->supervoxelIdx[g->supervoxelNum++] = (int)s;
      }
    }
}

static int em_loop(long long N, long long S, int G, const float* query,
                   const float* img, const float* cent, const long long* ptr,
                   GaussianMixtureModelCUDA* gm, const tgmm_em_opts* o,
                   int stop_after, float* rnk, int* ind, double* Nk_out,
                   double* xk_out, double* Ck_out, double* ll_hist,
                   int* iters, double* ll_final) {
  const int K = o->K;
  if (K < 1 || K > TGMM_KMAX || G < 0 || S < 0 || N < 0) return TGMM_ERR_ARG;
  /* voxel mode: every voxel is its own query and its own "supervoxel" for
   * the kNN and the log-likelihood; the M-step weight is the stored value */
  const int vox = o->voxel_mode != 0;
  long long* vptr = NULL;
  float* ones = NULL;
  if (vox) {
    vptr = (long long*)malloc(sizeof(long long) * (size_t)(N + 1));
    ones = (float*)malloc(sizeof(float) * (size_t)(N > 0 ? N : 1));
    for (long long v = 0; v <= N; v++) vptr[v] = v;
    for (long long v = 0; v < N; v++) ones[v] = 1.0f;
    S = N;
    cent = query;
    ptr = vptr;
  }
  double* Nk = (double*)malloc(sizeof(double) * (size_t)(G > 0 ? G : 1));
  double* xk = (double*)malloc(sizeof(double) * 3 * (size_t)(G > 0 ? G : 1));
  double* Ck = (double*)malloc(sizeof(double) * 6 * (size_t)(G > 0 ? G : 1));
  double ll = -1.0, ll1 = -10.0, ll2 = -100.0, ll3 = -1000.0;
  int it = 1;
  while (!orc_converged(ll, ll1, ll2, ll3, o->tol) && it <= o->max_iter) {
    orc_knn(S, cent, G, gm, o->scale, K, ind);
    orc_expectations(G, gm, orc_sum_alpha(G, gm));
    if (vox) {
      orc_estep_voxel(N, query, img, G, gm, K, ind, rnk);
      orc_mstep(N, S, query, ones, ptr, G, K, ind, rnk, Nk, xk, Ck);
    } else {
      orc_estep(N, S, query, ptr, G, gm, K, ind, rnk);
      orc_mstep(N, S, query, img, ptr, G, K, ind, rnk, Nk, xk, Ck);
    }
    if (stop_after > 0 && it == stop_after) {
      memcpy(Nk_out, Nk, sizeof(double) * G);
      memcpy(xk_out, xk, sizeof(double) * 3 * G);
      memcpy(Ck_out, Ck, sizeof(double) * 6 * G);
      break;
    }
    orc_update(G, gm, Nk, xk, Ck, o);
    orc_death(G, gm, orc_sum_alpha(G, gm));
    ll3 = ll2;
    ll2 = ll1;
    ll1 = ll;
    ll = orc_loglik(N, S, query, img, ptr, G, gm, K, ind, orc_sum_alpha(G, gm));
    if (ll_hist) ll_hist[it - 1] = ll;
    it++;
  }
  if (iters) *iters = it - 1;
  if (ll_final) *ll_final = ll;
  free(vptr);
  free(ones);
  free(Nk);
  free(xk);
  free(Ck);
  return TGMM_OK;
}

int orc_em_run(long long N, long long S, int G, const float* query,
               const float* img, const float* cent, const long long* ptr,
               GaussianMixtureModelCUDA* gm, const tgmm_em_opts* o, float* rnk,
               int* ind, double* ll_hist, int* iters, double* ll_final) {
  int rc = em_loop(N, S, G, query, img, cent, ptr, gm, o, 0, rnk, ind, NULL,
                   NULL, NULL, ll_hist, iters, ll_final);
  if (rc == TGMM_OK && !o->voxel_mode) {
    const int cap = o->max_sv_per_gaussian > 0 ? o->max_sv_per_gaussian : 0;
    orc_assign_supervoxels(S, G, gm, o->K, rnk, ind, cap);
  }
  return rc;
}

int orc_em_partial(long long N, long long S, int G, const float* query,
                   const float* img, const float* cent, const long long* ptr,
                   GaussianMixtureModelCUDA* gm, const tgmm_em_opts* o,
                   int iters, float* rnk, int* ind, double* Nk, double* xk,
                   double* Ck) {
  return em_loop(N, S, G, query, img, cent, ptr, gm, o, iters, rnk, ind, Nk,
                 xk, Ck, NULL, NULL, NULL);
}
