/*
 * oracle.h -- CPU restatement (FP64, plain C, OpenMP) of the reference's
 * supervoxel-mode variational-inference EM
 * (nmeth.3036-S2/src/UtilsCUDA/GMEMupdateCUDA.cu:3985-4421 and the kernels it
 * launches, knnCuda.cu:61-246) and, with tgmm_em_opts.voxel_mode, of the
 * per-voxel EM (GMEMupdateCUDA.cu:3420-3973).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg, and by
 * nothing that ships. It shares no code with the HIP implementation; both are
 * written independently from the reference's equations so that one checks
 * the other. Every step follows the reference per voxel, in FP64, with the
 * documented deviations listed in the .c file header.
 */
#ifndef TGMM_ORACLE_H_
#define TGMM_ORACLE_H_

#include "../include/tgmm_em.h"

#ifdef __cplusplus
extern "C" {
#endif

void orc_set_threads(int n);
int orc_get_threads(void);

/* scalar pieces */
double orc_digamma(double x);
double orc_det3(const double w[6]);
int orc_inv3(const double w[6], double out[6]);
void orc_eig3_ordered(const double w[6], double d[3], double v[9]);
void orc_regularize_W(double W[6], double nu, int w4dof, const float scale[3],
                      double lambda_min, double lambda_max, double max_exc);
void orc_posterior(GaussianMixtureModelCUDA* g, double Nk, const double xk[3],
                   const double Ck[6]);

/* per-step functions over a whole frame */
void orc_knn(long long S, const float* cent, int G,
             const GaussianMixtureModelCUDA* gm, const float scale[3], int K,
             int* ind);
double orc_sum_alpha(int G, const GaussianMixtureModelCUDA* gm);
void orc_expectations(int G, GaussianMixtureModelCUDA* gm, double sum_alpha);
void orc_estep(long long N, long long S, const float* query,
               const long long* ptr, int G, const GaussianMixtureModelCUDA* gm,
               int K, const int* ind, float* rnk);
/* voxel mode: N voxels, rnk/ind [K*N] slot-major */
void orc_estep_voxel(long long N, const float* query, const float* img, int G,
                     const GaussianMixtureModelCUDA* gm, int K, const int* ind,
                     float* rnk);
void orc_mstep(long long N, long long S, const float* query, const float* img,
               const long long* ptr, int G, int K, const int* ind,
               const float* rnk, double* Nk, double* xk, double* Ck);
void orc_update(int G, GaussianMixtureModelCUDA* gm, const double* Nk,
                const double* xk, const double* Ck, const tgmm_em_opts* o);
void orc_death(int G, GaussianMixtureModelCUDA* gm, double sum_alpha);
double orc_loglik(long long N, long long S, const float* query,
                  const float* img, const long long* ptr, int G,
                  const GaussianMixtureModelCUDA* gm, int K, const int* ind,
                  double sum_alpha);
/* local log-likelihood over lists of supervoxels: list i is
 * list_idx[list_ptr[i] .. list_ptr[i+1]); empty lists leave out[i] as given */
void orc_local_loglik(long long N, long long S, const float* query,
                      const float* img, const long long* ptr, int G,
                      const GaussianMixtureModelCUDA* gm, int K, const int* ind,
                      int n_lists, const long long* list_ptr,
                      const int* list_idx, double* out);
/* input marshalling of one frame from raw uint16 voxels: query [3N] SoA,
 * img [N], centroids [3S] SoA; 0, or -1 for an index outside the volume */
int orc_marshal_frame(long long N, long long S, const unsigned long long* pixel_idx,
                      const unsigned short* raw, const long long* sv_ptr,
                      const unsigned short* thr, const long long dims[3], float* query,
                      float* img, float* cent);
/* supervoxel kNN (knnCUDA_): slot-major ind/dist [K * query_nb], K <= 11 */
void orc_sv_knn(int* ind, float* dist, const float* query, const float* ref,
                long long query_nb, int ref_nb, int K, const float scale[3]);
int orc_converged(double ll, double ll1, double ll2, double ll3, double tol);
void orc_assign_supervoxels(long long S, int G, GaussianMixtureModelCUDA* gm,
                            int K, const float* rnk, const int* ind, int cap);

/* one full solve; ll_hist may be NULL (else room for max_iter entries) */
int orc_em_run(long long N, long long S, int G, const float* query,
               const float* img, const float* cent, const long long* ptr,
               GaussianMixtureModelCUDA* gm, const tgmm_em_opts* o, float* rnk,
               int* ind, double* ll_hist, int* iters, double* ll_final);

/* The same solve, stopped after `iters` iterations have completed kNN,
 * E-step and M-step statistics of iteration `iters` (used to check the
 * per-iteration intermediates). */
int orc_em_partial(long long N, long long S, int G, const float* query,
                   const float* img, const float* cent, const long long* ptr,
                   GaussianMixtureModelCUDA* gm, const tgmm_em_opts* o,
                   int iters, float* rnk, int* ind, double* Nk, double* xk,
                   double* Ck);

#ifdef __cplusplus
}
#endif
#endif
