// ref_twin_sv.cpp -- C-ABI wrapper over the reference's own supervoxel
// statistics, compiled from the reference sources where they lie (see
// oracle/Makefile target `ref`). TEST INFRASTRUCTURE ONLY: it pins the
// oracle's input marshalling against
//   supervoxel::weightedGaussianStatistics<uint16>
//       temporalLogicalRules/supervoxel.cpp:1410-1449
// the intensity-weighted centroid that the tracker passes to the EM as the
// supervoxel's kNN query (mainTrackingGaussianMixtureModel.cpp:616-617,
// 852-857). The output library lives in oracle/_ref/ and is never shipped.
#include "temporalLogicalRules/supervoxel.h"

extern "C" {

// One supervoxel over a uint16 volume of dimensions dims (x fastest): its
// centroid[3] and total intensity, as the reference computes them.
int ref_sv_centroid(const unsigned short* volume, const unsigned long long* dims,
                    const unsigned long long* idx, long long n, float* centroid,
                    float* intensity) {
  uint64 d[dimsImage];
  for (int a = 0; a < dimsImage; a++) d[a] = dims[a];
  supervoxel::setDataDims(d);
  supervoxel sv;
  sv.dataPtr = (void*)volume;
  sv.PixelIdxList.assign(idx, idx + n);
  const int rc = sv.weightedGaussianStatistics<unsigned short>(false);
  for (int a = 0; a < dimsImage; a++) centroid[a] = sv.centroid[a];
  *intensity = sv.intensity;
  return rc;
}

}  // extern "C"
