/* Exercises the oracle (oracle/oracle.c) on small random frames for a build
 * with AddressSanitizer and UndefinedBehaviorSanitizer: supervoxel- and
 * voxel-mode EM runs (convergence on and off), the supervoxel kNN, the input
 * marshalling and the local log-likelihood. Prints "ok" on success. Test
 * infrastructure only. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/oracle.h"

static unsigned long long g_state = 88172645463325252ULL;
static double urand(void) {
  g_state ^= g_state << 13;
  g_state ^= g_state >> 7;
  g_state ^= g_state << 17;
  return (double)(g_state >> 11) / 9007199254740992.0;
}

#define CHECK(c)                                               \
  do {                                                         \
    if (!(c)) {                                                \
      printf("FAILED %s at line %d\n", #c, __LINE__);         \
      return 1;                                                \
    }                                                          \
  } while (0)

int main(void) {
  const long long S = 60, per = 9, N = S * per;
  const int G = 14;
  float* query = malloc(sizeof(float) * 3 * N);
  float* img = malloc(sizeof(float) * N);
  float* cent = malloc(sizeof(float) * 3 * S);
  long long* ptr = malloc(sizeof(long long) * (S + 1));
  for (long long s = 0; s <= S; s++) ptr[s] = s * per;
  for (long long s = 0; s < S; s++) {
    const double c[3] = {40 * urand(), 40 * urand(), 8 * urand()};
    double m[3] = {0, 0, 0};
    for (long long v = ptr[s]; v < ptr[s + 1]; v++) {
      for (int a = 0; a < 3; a++) {
        query[v + a * N] = (float)floor(c[a] + 3 * urand());
        m[a] += query[v + a * N];
      }
      img[v] = (float)(0.05 + urand());
    }
    for (int a = 0; a < 3; a++) cent[s + a * S] = (float)(m[a] / per);
  }
  GaussianMixtureModelCUDA* gm0 = calloc(G, sizeof(GaussianMixtureModelCUDA));
  for (int k = 0; k < G; k++) {
    GaussianMixtureModelCUDA* g = &gm0[k];
    g->m_k[0] = 40 * urand();
    g->m_k[1] = 40 * urand();
    g->m_k[2] = 8 * urand();
    memcpy(g->m_o, g->m_k, sizeof(g->m_k));
    const double W[6] = {0.04, 0, 0, 0.04, 0, 0.4};
    memcpy(g->W_k, W, sizeof(W));
    memcpy(g->W_o, W, sizeof(W));
    g->nu_k = g->nu_o = 4;
    g->beta_k = g->beta_o = 1;
    g->alpha_k = g->alpha_o = 1;
  }
  gm0[G - 1].m_o[0] = gm0[G - 1].m_k[0] = -1e32; /* one dead Gaussian */
  for (int mode = 0; mode < 2; mode++)
    for (int conv = 0; conv < 2; conv++) {
      tgmm_em_opts o;
      memset(&o, 0, sizeof(o));
      o.K = 5;
      o.max_iter = conv ? 60 : 6;
      o.tol = conv ? 1e-6 : -1.0;
      o.lambda_min = 0.02;
      o.lambda_max = 0.2222;
      o.max_excentricity = 9.0;
      o.w4dof = 1;
      o.scale[0] = o.scale[1] = 1.0f;
      o.scale[2] = 5.0f;
      o.max_sv_per_gaussian = 16;
      o.voxel_mode = mode;
      const long long Q = mode ? N : S;
      GaussianMixtureModelCUDA* gm = malloc(sizeof(GaussianMixtureModelCUDA) * G);
      memcpy(gm, gm0, sizeof(GaussianMixtureModelCUDA) * G);
      float* rnk = malloc(sizeof(float) * 5 * Q);
      int* ind = malloc(sizeof(int) * 5 * Q);
      double* hist = malloc(sizeof(double) * o.max_iter);
      int iters = 0;
      double ll = 0;
      CHECK(orc_em_run(N, S, G, query, img, cent, ptr, gm, &o, rnk, ind, hist, &iters, &ll) ==
            TGMM_OK);
      CHECK(iters >= 1 && iters <= o.max_iter && isfinite(ll));
      for (long long i = 0; i < 5 * Q; i++) CHECK(ind[i] >= -1 && ind[i] < G);
      if (!mode) {
        const long long lp[3] = {0, 3, 5};
        const int li[5] = {0, 1, 2, 7, 7};
        double out[2] = {0, 0};
        orc_local_loglik(N, S, query, img, ptr, G, gm, 5, ind, 2, lp, li, out);
        CHECK(isfinite(out[0]) && isfinite(out[1]));
      }
      free(gm);
      free(rnk);
      free(ind);
      free(hist);
    }
  /* supervoxel kNN with a NaN reference */
  int sind[11 * 60];
  float sdist[11 * 60];
  float ref[3 * 40];
  for (int i = 0; i < 120; i++) ref[i] = (float)(40 * urand());
  ref[5] = NAN;
  orc_sv_knn(sind, sdist, cent, ref, S, 40, 11, (const float[3]){1, 1, 5});
  for (int i = 0; i < 11 * 60; i++) CHECK(sind[i] >= -1 && sind[i] < 40);
  /* marshalling of a 12x10x6 volume */
  const long long dims[3] = {12, 10, 6}, MN = 50, MS = 5;
  unsigned long long pix[50];
  unsigned short raw[50], thr[5] = {3, 0, 9, 1, 4};
  long long sp[6] = {0, 10, 20, 30, 40, 50};
  for (int i = 0; i < 50; i++) {
    pix[i] = (unsigned long long)(i * 13 % 720);
    raw[i] = (unsigned short)(urand() * 4000);
  }
  float mq[150], mi[50], mc[15];
  CHECK(orc_marshal_frame(MN, MS, pix, raw, sp, thr, dims, mq, mi, mc) == 0);
  pix[7] = 720; /* outside the volume */
  CHECK(orc_marshal_frame(MN, MS, pix, raw, sp, thr, dims, mq, mi, mc) == -1);
  free(query);
  free(img);
  free(cent);
  free(ptr);
  free(gm0);
  printf("ok\n");
  return 0;
}
