// tgmm_em_compat.hpp -- drop-in replacement, for the tracker's EM path
// (supervoxel and voxel mode), of the reference header
// nmeth.3036-S2/src/UtilsCUDA/GMEMupdateCUDA.h.
//
// The tracker includes this file where it included GMEMupdateCUDA.h (and
// tgmm_utils_knn_compat.hpp where it included UtilsCUDA/knnCuda.h) and links
// libtgmm_em.so instead of the CUDA objects. The inline functions below carry
// the reference's names and parameter lists (GMEMupdateCUDA.h:82, 87-98,
// 116-144) and its error behaviour: print a message and exit(), with the
// exit codes the reference uses (0 for the size limits, GMEM.cu:3992-4007;
// 3 when the regularisation constants were never set, GMEM.cu:4217-4222;
// EXIT_FAILURE for runtime errors, book.h:20-25).
//
// The regularisation constants are read from the tracker's own host statics
// regularizePrecisionMatrixConstants (constants.h:19-45), so the tracker's
// constants.h must be included before this header.
#ifndef TGMM_EM_COMPAT_HPP_
#define TGMM_EM_COMPAT_HPP_

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tgmm_em.h"

typedef float pxi;  // CSparse.h:44

namespace tgmm_compat {
inline void check(int rc, const char* where) {
  if (rc == TGMM_OK) return;
  std::printf("ERROR: at %s: %s\n", where, tgmm_last_error());
  if (rc == TGMM_ERR_LIMIT) std::exit(0);
  if (rc == TGMM_ERR_NOT_SET) std::exit(3);
  std::exit(EXIT_FAILURE);
}
}  // namespace tgmm_compat

inline void GMEMinitializeMemoryWithSupervoxels(
    float** queryCUDA, float* query, float** imgDataCUDA, float* imgData, float* scale,
    long long int query_nb, float** rnkCUDA, int** indCUDA, int** indCUDAtr,
    unsigned short int numLabels, float** centroidLabelPositionCUDA,
    float* centroidLabelPositionlong, long long int** labelListPtrCUDA,
    long long int* labelListPtrHOST) {
  tgmm_compat::check(
      tgmm_GMEMinitializeMemoryWithSupervoxels(
          queryCUDA, query, imgDataCUDA, imgData, scale, query_nb, rnkCUDA, indCUDA, indCUDAtr,
          numLabels, centroidLabelPositionCUDA, centroidLabelPositionlong, labelListPtrCUDA,
          labelListPtrHOST),
      "GMEMinitializeMemoryWithSupervoxels");
}

inline void copyScaleToDvice(float* scale) {
  tgmm_compat::check(tgmm_copyScaleToDvice(scale), "copyScaleToDvice");
}

inline void GMEMvariationalInferenceCUDAWithSupervoxels(
    float* queryCUDA, float* imgDataCUDA, pxi* rnkCUDA, pxi* rnkCUDAtr, int* indCUDA,
    int* indCUDAtr, float* centroidLabelPositionCUDA, long long int* labelListPtrCUDA,
    GaussianMixtureModelCUDA* vecGMtemp, long long int query_nb, int ref_nb,
    unsigned short int maxLabels, int maxIterEM, double tolLikelihood, int devCUDA, int frame,
    bool W4DOF, std::string debugPath = "") {
  if (regularizePrecisionMatrixConstants::lambdaMin >= 0.0)
    tgmm_setRegularizePrecisionMatrixConstants(regularizePrecisionMatrixConstants::lambdaMin,
                                               regularizePrecisionMatrixConstants::lambdaMax,
                                               regularizePrecisionMatrixConstants::maxExcentricity);
  tgmm_compat::check(
      tgmm_GMEMvariationalInferenceCUDAWithSupervoxels(
          queryCUDA, imgDataCUDA, rnkCUDA, rnkCUDAtr, indCUDA, indCUDAtr,
          centroidLabelPositionCUDA, labelListPtrCUDA, vecGMtemp, query_nb, ref_nb, maxLabels,
          maxIterEM, tolLikelihood, devCUDA, frame, W4DOF ? 1 : 0, debugPath.c_str()),
      "GMEMvariationalInferenceCUDAWithSupervoxels");
}

inline void GMEMvariationalInferenceCUDA(float* queryCUDA, float* imgDataCUDA, pxi* rnkCUDA,
                                         pxi* rnkCUDAtr, int* indCUDA, int* indCUDAtr,
                                         GaussianMixtureModelCUDA* vecGMtemp,
                                         long long int query_nb, int ref_nb, int maxIterEM,
                                         double tolLikelihood, int devCUDA, int frame, bool W4DOF,
                                         std::string debugPath = "") {
  if (regularizePrecisionMatrixConstants::lambdaMin >= 0.0)
    tgmm_setRegularizePrecisionMatrixConstants(regularizePrecisionMatrixConstants::lambdaMin,
                                               regularizePrecisionMatrixConstants::lambdaMax,
                                               regularizePrecisionMatrixConstants::maxExcentricity);
  tgmm_compat::check(
      tgmm_GMEMvariationalInferenceCUDA(queryCUDA, imgDataCUDA, rnkCUDA, rnkCUDAtr, indCUDA,
                                        indCUDAtr, vecGMtemp, query_nb, ref_nb, maxIterEM,
                                        tolLikelihood, devCUDA, frame, W4DOF ? 1 : 0,
                                        debugPath.c_str()),
      "GMEMvariationalInferenceCUDA");
}

inline void calculateLocalLikelihood(std::vector<double>& localLikelihood,
                                     const std::vector<std::vector<int> >& listSupervoxels,
                                     float* queryCUDA, float* imgDataCUDA, pxi* rnkCUDA,
                                     int* indCUDA, GaussianMixtureModelCUDA* vecGMtemp,
                                     long long int* labelListPtrCUDA, long long int query_nb,
                                     int ref_nb, int numLabels) {
  localLikelihood.resize(listSupervoxels.size());
  std::vector<long long> ptr(listSupervoxels.size() + 1, 0);
  std::vector<int> idx;
  for (size_t i = 0; i < listSupervoxels.size(); i++) {
    ptr[i + 1] = ptr[i] + (long long)listSupervoxels[i].size();
    idx.insert(idx.end(), listSupervoxels[i].begin(), listSupervoxels[i].end());
  }
  tgmm_compat::check(
      tgmm_calculateLocalLikelihood(localLikelihood.data(), ptr.data(), idx.data(),
                                    (int)listSupervoxels.size(), queryCUDA, imgDataCUDA, rnkCUDA,
                                    indCUDA, vecGMtemp, labelListPtrCUDA, query_nb, ref_nb,
                                    numLabels),
      "calculateLocalLikelihood");
}

inline void GMEMinitializeMemory(float** queryCUDA, float* query, float** imgDataCUDA,
                                 float* imgData, float* scale, long long int query_nb,
                                 float** rnkCUDA, int** indCUDA, int** indCUDAtr) {
  tgmm_compat::check(tgmm_GMEMinitializeMemory(queryCUDA, query, imgDataCUDA, imgData, scale,
                                               query_nb, rnkCUDA, indCUDA, indCUDAtr),
                     "GMEMinitializeMemory");
}

inline void GMEMreleaseMemory(float** queryCUDA, float** imgDataCUDA, float** rnkCUDA,
                              int** indCUDA, int** indCUDAtr) {
  tgmm_compat::check(tgmm_GMEMreleaseMemory(queryCUDA, imgDataCUDA, rnkCUDA, indCUDA, indCUDAtr),
                     "GMEMreleaseMemory");
}

inline void GMEMreleaseMemoryWithSupervoxels(float** queryCUDA, float** imgDataCUDA,
                                             float** rnkCUDA, int** indCUDA, int** indCUDAtr,
                                             float** centroidLabelPositionCUDA,
                                             long long int** labelListPtrCUDA) {
  tgmm_compat::check(
      tgmm_GMEMreleaseMemoryWithSupervoxels(queryCUDA, imgDataCUDA, rnkCUDA, indCUDA, indCUDAtr,
                                            centroidLabelPositionCUDA, labelListPtrCUDA),
      "GMEMreleaseMemoryWithSupervoxels");
}

#endif  // TGMM_EM_COMPAT_HPP_
