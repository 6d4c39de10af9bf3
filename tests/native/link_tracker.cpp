// link_tracker.cpp -- one translation unit that calls every function the
// tracker takes from GMEMupdateCUDA.h and UtilsCUDA/knnCuda.h
// (mainTrackingGaussianMixtureModel.cpp:207, 778, 1064-1107, 1336-1375,
// 1497, 1935-1937; variationalInference.h:11-12 includes both headers) and
// from temporalLogicalRules/knnCUDA/knnCuda.h (lineageHyperTree.cpp:2568,
// supervoxel.cpp:700-1290), through the compat headers, linked against
// libtgmm_em.so alone. TEST PROGRAM for tests/test_abi_cpu.py: building it
// proves the drop-in links; run with no argument it calls nothing.
#include <string>
#include <vector>

struct regularizePrecisionMatrixConstants {
  static double lambdaMin, lambdaMax, maxExcentricity;
};
double regularizePrecisionMatrixConstants::lambdaMin = -1.0;
double regularizePrecisionMatrixConstants::lambdaMax = -1.0;
double regularizePrecisionMatrixConstants::maxExcentricity = -1.0;

#include "tgmm_em_compat.hpp"
#include "tgmm_knn_compat.hpp"
#include "tgmm_utils_knn_compat.hpp"

int main(int argc, char** argv) {
  (void)argv;
  if (argc < 2) return 0;
  float *q = nullptr, *im = nullptr, *rnk = nullptr, *cent = nullptr, sc[3] = {1, 1, 5};
  int *ind = nullptr, *indtr = nullptr;
  long long* ptr = nullptr;
  long long ptrh[1] = {0};
  std::vector<GaussianMixtureModelCUDA> gm(1);
  const int dev = selectCUDAdeviceMaxMemory();
  setDeviceCUDA(dev);
  uploadScaleCUDA(sc);
  GMEMinitializeMemoryWithSupervoxels(&q, nullptr, &im, nullptr, sc, 0, &rnk, &ind, &indtr, 0,
                                      &cent, nullptr, &ptr, ptrh);
  copyScaleToDvice(sc);
  GMEMvariationalInferenceCUDAWithSupervoxels(q, im, rnk, nullptr, ind, indtr, cent, ptr,
                                              gm.data(), 0, 1, 0, 1, 1e-6, dev, 0, true,
                                              std::string(""));
  std::vector<double> ll;
  std::vector<std::vector<int> > lists(1);
  calculateLocalLikelihood(ll, lists, q, im, rnk, ind, gm.data(), ptr, 0, 1, 0);
  GMEMreleaseMemoryWithSupervoxels(&q, &im, &rnk, &ind, &indtr, &cent, &ptr);
  GMEMinitializeMemory(&q, nullptr, &im, nullptr, sc, 0, &rnk, &ind, &indtr);
  GMEMvariationalInferenceCUDA(q, im, rnk, nullptr, ind, indtr, gm.data(), 0, 1, 1, 1e-6, dev,
                               0, true);
  GMEMreleaseMemory(&q, &im, &rnk, &ind, &indtr);
  int kind[1];
  float kdist[1];
  return knnCUDA_(kind, kdist, sc, sc, 0, 0, 1, sc, dev);
}
