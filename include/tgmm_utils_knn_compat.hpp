// tgmm_utils_knn_compat.hpp -- drop-in replacement, for the tracker, of the
// reference header nmeth.3036-S2/src/UtilsCUDA/knnCuda.h, which the tracker
// includes directly (mainTrackingGaussianMixtureModel.cpp:19) and through
// variationalInference.h:12. The tracker calls three of its functions:
// selectCUDAdeviceMaxMemory at start-up (main.cpp:207), setDeviceCUDA before
// each frame (main.cpp:778), and uploadScaleCUDA through the EM's init
// (GMEMupdateCUDA.cu:4578, 4634); they keep the reference's names, parameter
// lists (knnCuda.h:20-23) and error behaviour (HANDLE_ERROR prints and exits,
// book.h:20-25). The EM's own kNN (knnCUDAinPlace and the other functions the
// header declares) runs inside libtgmm_em.so; the tracker never calls it.
#ifndef TGMM_UTILS_KNN_COMPAT_HPP_
#define TGMM_UTILS_KNN_COMPAT_HPP_

#include <cstdio>
#include <cstdlib>

#include "tgmm_em.h"

inline int selectCUDAdeviceMaxMemory() {
  int device = -1;
  if (tgmm_selectCUDAdeviceMaxMemory(&device) != TGMM_OK) {
    std::printf("ERROR: at selectCUDAdeviceMaxMemory(): %s\n", tgmm_last_error());
    std::exit(EXIT_FAILURE);
  }
  return device;
}

inline void setDeviceCUDA(int devCUDA) {
  if (tgmm_setDeviceCUDA(devCUDA) != TGMM_OK) {
    std::printf("ERROR: at setDeviceCUDA(): %s\n", tgmm_last_error());
    std::exit(EXIT_FAILURE);
  }
}

inline void uploadScaleCUDA(float* scale) {
  if (tgmm_uploadScaleCUDA(scale) != TGMM_OK) {
    std::printf("ERROR: at uploadScaleCUDA(): %s\n", tgmm_last_error());
    std::exit(EXIT_FAILURE);
  }
}

#endif  // TGMM_UTILS_KNN_COMPAT_HPP_
