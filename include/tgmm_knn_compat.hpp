// tgmm_knn_compat.hpp -- drop-in replacement, for the tracker's supervoxel
// neighbour searches, of the reference header
// nmeth.3036-S2/src/temporalLogicalRules/knnCUDA/knnCuda.h.
//
// The tracker includes this file where it included knnCUDA/knnCuda.h and
// links libtgmm_em.so instead of knnCuda.cu. knnCUDA_ keeps the reference's
// name, parameter list (knnCuda.h:56-58) and return convention (0 on
// success); a device or argument error prints a message and exits as the
// reference's HANDLE_ERROR does (book.h:20-25), and a KNN above maxKNN
// prints a message and returns 2 (the callers check that bound first,
// lineageHyperTree.cpp:2536-2542).
#ifndef TGMM_KNN_COMPAT_HPP_
#define TGMM_KNN_COMPAT_HPP_

#include <cstdio>
#include <cstdlib>

#include "tgmm_em.h"

static const int maxKNN = TGMM_SV_KNN_MAX;  // knnCuda.h:28

inline int knnCUDA_(int* ind, float* dist, float* query, float* ref, long long int query_nb,
                    int ref_nb, int KNN, float* scale, int devCUDA) {
  const int rc = tgmm_knnCUDA_(ind, dist, query, ref, query_nb, ref_nb, KNN, scale, devCUDA);
  if (rc == TGMM_OK) return 0;
  std::printf("ERROR: at knnCUDA_: %s\n", tgmm_last_error());
  if (rc == TGMM_ERR_LIMIT) return 2;
  std::exit(EXIT_FAILURE);
}

#endif  // TGMM_KNN_COMPAT_HPP_
