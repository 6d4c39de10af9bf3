// compat_frame.cpp -- one frame through the reference's own C++ names, in
// the tracker's order (mainTrackingGaussianMixtureModel.cpp:207 device
// selection, :778 setDeviceCUDA, :1065-1105 EM init and solve, :1336 local
// likelihood, :1935-1937 release; lineageHyperTree.cpp:2568 supervoxel
// neighbours), then the voxel-mode EM (GMEMinitializeMemory,
// GMEMvariationalInferenceCUDA, GMEMreleaseMemory, GMEMupdateCUDA.h:87-91,
// 129-139), with include/tgmm_em_compat.hpp, include/tgmm_utils_knn_compat.hpp
// and include/tgmm_knn_compat.hpp in place of GMEMupdateCUDA.h,
// UtilsCUDA/knnCuda.h and knnCUDA/knnCuda.h.
// TEST PROGRAM: tests/test_gpu_compat.py writes the input file, runs this
// program as a child process and compares its output with the Python path.
//
//   compat_frame <input> <output>
//
// input:  int64 N, S, G, n_lists, n_entries, max_iter;
//         float query[3N], img[N], centroids[3S], scale[3];
//         int64 ptr[S+1]; GaussianMixtureModelCUDA records[G];
//         int64 list_ptr[n_lists+1]; int32 list_idx[n_entries];
//         double lambda_min, lambda_max, max_excentricity
// output: records[G] after the EM; double local_ll[n_lists];
//         int32 knn_ind[11 S]; float knn_dist[11 S];
//         records[G] after the voxel-mode EM from the same inputs;
//         int32 device (selectCUDAdeviceMaxMemory)
#include <cstdio>
#include <cstdlib>
#include <vector>

// The tracker's host statics (constants.h:19-45), which the compat header
// reads; this program plays the tracker.
struct regularizePrecisionMatrixConstants {
  static double lambdaMin, lambdaMax, maxExcentricity;
};
double regularizePrecisionMatrixConstants::lambdaMin = -1.0;
double regularizePrecisionMatrixConstants::lambdaMax = -1.0;
double regularizePrecisionMatrixConstants::maxExcentricity = -1.0;

#include "tgmm_em_compat.hpp"
#include "tgmm_knn_compat.hpp"
#include "tgmm_utils_knn_compat.hpp"

template <typename T>
static void rd(FILE* f, T* p, long long n) {
  if (n > 0 && fread(p, sizeof(T), (size_t)n, f) != (size_t)n) {
    std::fprintf(stderr, "short read\n");
    std::exit(2);
  }
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  long long hdr[6];
  rd(in, hdr, 6);
  const long long N = hdr[0], S = hdr[1], G = hdr[2], L = hdr[3], E = hdr[4];
  const int max_iter = (int)hdr[5];
  std::vector<float> query(3 * N), img(N), cent(3 * S), scale(3);
  std::vector<long long> ptr(S + 1), lptr(L + 1);
  std::vector<GaussianMixtureModelCUDA> gm(G);
  std::vector<int> lidx(E);
  double lam[3];
  rd(in, query.data(), 3 * N);
  rd(in, img.data(), N);
  rd(in, cent.data(), 3 * S);
  rd(in, scale.data(), 3);
  rd(in, ptr.data(), S + 1);
  rd(in, gm.data(), G);
  rd(in, lptr.data(), L + 1);
  rd(in, lidx.data(), E);
  rd(in, lam, 3);
  std::fclose(in);

  regularizePrecisionMatrixConstants::lambdaMin = lam[0];
  regularizePrecisionMatrixConstants::lambdaMax = lam[1];
  regularizePrecisionMatrixConstants::maxExcentricity = lam[2];

  // main.cpp:207, 778
  const int devCUDA = selectCUDAdeviceMaxMemory();
  if (devCUDA < 0) return 4;
  setDeviceCUDA(devCUDA);
  const std::vector<GaussianMixtureModelCUDA> gm_in = gm;
  // main.cpp:1065-1070
  float *queryCUDA = nullptr, *imgDataCUDA = nullptr, *rnkCUDA = nullptr,
        *centroidLabelPositionCUDA = nullptr;
  int *indCUDA = nullptr, *indCUDAtr = nullptr;
  long long* labelListPtrCUDA = nullptr;
  GMEMinitializeMemoryWithSupervoxels(&queryCUDA, query.data(), &imgDataCUDA, img.data(),
                                      scale.data(), N, &rnkCUDA, &indCUDA, &indCUDAtr,
                                      (unsigned short)S, &centroidLabelPositionCUDA, cent.data(),
                                      &labelListPtrCUDA, ptr.data());
  copyScaleToDvice(scale.data());
  // main.cpp:1100-1105
  GMEMvariationalInferenceCUDAWithSupervoxels(queryCUDA, imgDataCUDA, rnkCUDA, nullptr, indCUDA,
                                              indCUDAtr, centroidLabelPositionCUDA,
                                              labelListPtrCUDA, gm.data(), N, (int)G,
                                              (unsigned short)S, max_iter, -1.0, devCUDA, 0,
                                              true);
  // main.cpp:1336-1339
  std::vector<std::vector<int> > lists(L);
  for (long long i = 0; i < L; i++) lists[i].assign(lidx.begin() + lptr[i], lidx.begin() + lptr[i + 1]);
  std::vector<double> local;
  calculateLocalLikelihood(local, lists, queryCUDA, imgDataCUDA, rnkCUDA, indCUDA, gm.data(),
                           labelListPtrCUDA, N, (int)G, (int)S);
  // lineageHyperTree.cpp:2568 (in space: the centroids against themselves)
  std::vector<int> kind(11 * S);
  std::vector<float> kdist(11 * S);
  if (knnCUDA_(kind.data(), kdist.data(), cent.data(), cent.data(), S, (int)S, 11, scale.data(),
               0) != 0)
    return 3;
  // main.cpp:1935-1937
  GMEMreleaseMemoryWithSupervoxels(&queryCUDA, &imgDataCUDA, &rnkCUDA, &indCUDA, &indCUDAtr,
                                   &centroidLabelPositionCUDA, &labelListPtrCUDA);

  // the voxel-mode EM on the same voxels and initial records
  std::vector<GaussianMixtureModelCUDA> gv = gm_in;
  GMEMinitializeMemory(&queryCUDA, query.data(), &imgDataCUDA, img.data(), scale.data(), N,
                       &rnkCUDA, &indCUDA, &indCUDAtr);
  uploadScaleCUDA(scale.data());
  copyScaleToDvice(scale.data());
  GMEMvariationalInferenceCUDA(queryCUDA, imgDataCUDA, rnkCUDA, nullptr, indCUDA, indCUDAtr,
                               gv.data(), N, (int)G, max_iter, -1.0, devCUDA, 0, true);
  GMEMreleaseMemory(&queryCUDA, &imgDataCUDA, &rnkCUDA, &indCUDA, &indCUDAtr);
  if (queryCUDA || imgDataCUDA || rnkCUDA || indCUDA || indCUDAtr) return 5;

  FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 2;
  std::fwrite(gm.data(), sizeof(GaussianMixtureModelCUDA), gm.size(), out);
  std::fwrite(local.data(), sizeof(double), local.size(), out);
  std::fwrite(kind.data(), sizeof(int), kind.size(), out);
  std::fwrite(kdist.data(), sizeof(float), kdist.size(), out);
  std::fwrite(gv.data(), sizeof(GaussianMixtureModelCUDA), gv.size(), out);
  std::fwrite(&devCUDA, sizeof(int), 1, out);
  std::fclose(out);
  return 0;
}
