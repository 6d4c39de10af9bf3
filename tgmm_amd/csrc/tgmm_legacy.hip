// tgmm_legacy.hip -- the reference host API (GMEMupdateCUDA.h, and the device
// helpers of UtilsCUDA/knnCuda.h that the tracker calls) as C-ABI entry
// points with status codes. Each VI call runs its frame in a one-frame
// workspace kept per device between calls (grown when a frame does not fit),
// with the EM iteration replayed as a captured hipGraph and the convergence
// flags polled every four iterations; the results come back as one copy of
// the records formed on the device.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/tgmm_em.h"
#include "tgmm_runtime_internal.h"

using namespace tgmm;

#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      tgmm_set_error("%s failed: %s", #x, hipGetErrorString(e_));        \
      return e_ == hipErrorOutOfMemory ? TGMM_ERR_OOM : TGMM_ERR_DEVICE; \
    }                                                                    \
  } while (0)

namespace {
// Host statics of the reference: kNN scale (uploadScaleCUDA), regularisation
// scale (copyScaleToDvice) and regularizePrecisionMatrixConstants.
std::mutex g_mu;
float g_scale_knn[3] = {1.f, 1.f, 1.f};
float g_scale_reg[3] = {1.f, 1.f, 1.f};
double g_lmin = -1.0, g_lmax = -1.0, g_emax = -1.0;

// One cached one-frame workspace per device for the VI entry points (the
// calls on a device are serialised by g_run_mu).
std::mutex g_run_mu;
std::vector<tgmm_batch*> g_cache;

// Solve one frame in the device's cached workspace and write the records
// into gm (host), rnk/ind into the caller's device arrays (K rows of S,
// slot-major), and optionally the reference's debug dump.
int run_frame(int dev, const FrameSrc& f, const tgmm_em_opts& o, const float* sreg,
              GaussianMixtureModelCUDA* gm, float* rnkCUDA, int* indCUDA, int* indCUDAtr,
              long long S, const char* dump, int frame) {
  std::lock_guard<std::mutex> lk(g_run_mu);
  if ((int)g_cache.size() <= dev) g_cache.resize(dev + 1, nullptr);
  tgmm_batch* b = nullptr;
  int rc = batch_build(dev, std::vector<FrameSrc>{f}, o, sreg, nullptr, &b, g_cache[dev], 1.25);
  g_cache[dev] = rc ? nullptr : b;
  if (rc) return rc;
  rc = tgmm_batch_solve(b, TGMM_SOLVE_GRAPH);
  tgmm_frame_desc d;
  std::memset(&d, 0, sizeof(d));
  d.gm = gm;
  if (!rc) rc = tgmm_batch_fetch(b, &d, 1);
  if (rc) return rc;
  const long long Sc = b->ws.sh.Scap;
  const size_t w = (size_t)S * sizeof(float);
  const int K = o.K;
  if (S > 0) {
    hipError_t e = hipMemcpy2DAsync(rnkCUDA, w, b->ws.P.rnk, Sc * sizeof(float), w, K,
                                    hipMemcpyDeviceToDevice, b->st);
    if (e == hipSuccess)
      e = hipMemcpy2DAsync(indCUDA, w, b->ws.P.ind, Sc * sizeof(int), w, K,
                           hipMemcpyDeviceToDevice, b->st);
    if (e == hipSuccess && indCUDAtr)
      e = hipMemcpyAsync(indCUDAtr, indCUDA, K * w, hipMemcpyDeviceToDevice, b->st);
    if (e == hipSuccess) e = hipStreamSynchronize(b->st);
    if (e != hipSuccess) {
      tgmm_set_error("copy of rnk/ind to caller buffers failed: %s", hipGetErrorString(e));
      return TGMM_ERR_DEVICE;
    }
  }
  if (getenv("TGMM_VERBOSE"))
    printf("Frame=%d;iter=%d;Log-likelihood =%16.12f\n", frame, d.iterations + 1, d.ll);
  if (dump && dump[0]) {
    // {int S; int K; float rnk[S*K]; int ind[S*K]} (GMEM.cu:4364-4384)
    std::vector<float> r2((size_t)S * K);
    std::vector<int> i2((size_t)S * K);
    if (S > 0 && (hipMemcpy2D(r2.data(), w, b->ws.P.rnk, Sc * 4, w, K, hipMemcpyDeviceToHost) !=
                      hipSuccess ||
                  hipMemcpy2D(i2.data(), w, b->ws.P.ind, Sc * 4, w, K, hipMemcpyDeviceToHost) !=
                      hipSuccess)) {
      tgmm_set_error("download of rnk/ind for the debug dump failed");
      return TGMM_ERR_DEVICE;
    }
    FILE* fid = fopen(dump, "wb");
    if (!fid) {
      tgmm_set_error("opening file %s to save rnk values from variational inference", dump);
      return TGMM_ERR_IO;
    }
    const int Si = (int)S, Ki = K;
    fwrite(&Si, sizeof(int), 1, fid);
    fwrite(&Ki, sizeof(int), 1, fid);
    fwrite(r2.data(), sizeof(float), r2.size(), fid);
    fwrite(i2.data(), sizeof(int), i2.size(), fid);
    fclose(fid);
  }
  return TGMM_OK;
}
}  // namespace

extern "C" int tgmm_setRegularizePrecisionMatrixConstants(double lambdaMin, double lambdaMax,
                                                          double maxExcentricity) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_lmin = lambdaMin;
  g_lmax = lambdaMax;
  g_emax = maxExcentricity;
  return TGMM_OK;
}

extern "C" int tgmm_copyScaleToDvice(const float* scale) {
  if (!scale) {
    tgmm_set_error("copyScaleToDvice: scale is NULL");
    return TGMM_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  for (int a = 0; a < 3; a++) g_scale_reg[a] = scale[a];
  return TGMM_OK;
}

extern "C" int tgmm_GMEMinitializeMemoryWithSupervoxels(
    float** queryCUDA, const float* query, float** imgDataCUDA, const float* imgData,
    const float* scale, long long query_nb, float** rnkCUDA, int** indCUDA, int** indCUDAtr,
    unsigned short numLabels, float** centroidLabelPositionCUDA,
    const float* centroidLabelPosition, long long** labelListPtrCUDA,
    const long long* labelListPtrHOST) {
  if (!queryCUDA || !imgDataCUDA || !rnkCUDA || !indCUDA || !indCUDAtr ||
      !centroidLabelPositionCUDA || !labelListPtrCUDA || !scale || query_nb < 0 ||
      (query_nb > 0 && (!query || !imgData)) || !labelListPtrHOST ||
      (numLabels > 0 && !centroidLabelPosition)) {
    tgmm_set_error("GMEMinitializeMemoryWithSupervoxels: NULL argument or negative size");
    return TGMM_ERR_ARG;
  }
  const long long N = query_nb, S = numLabels, K = TGMM_LEGACY_K;
  auto M = [](void** p, size_t n) { return hipMalloc(p, n > 0 ? n : 4); };
  HIPCHK(M((void**)queryCUDA, 3 * N * sizeof(float)));
  HIPCHK(M((void**)imgDataCUDA, N * sizeof(float)));
  HIPCHK(M((void**)rnkCUDA, K * S * sizeof(float)));
  HIPCHK(M((void**)indCUDA, K * S * sizeof(int)));
  HIPCHK(M((void**)indCUDAtr, K * S * sizeof(int)));
  HIPCHK(M((void**)centroidLabelPositionCUDA, 3 * S * sizeof(float)));
  HIPCHK(M((void**)labelListPtrCUDA, (S + 1) * sizeof(long long)));
  if (N > 0) {
    HIPCHK(hipMemcpy(*queryCUDA, query, 3 * N * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(*imgDataCUDA, imgData, N * sizeof(float), hipMemcpyHostToDevice));
  }
  if (S > 0)
    HIPCHK(hipMemcpy(*centroidLabelPositionCUDA, centroidLabelPosition, 3 * S * sizeof(float),
                     hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(*labelListPtrCUDA, labelListPtrHOST, (S + 1) * sizeof(long long),
                   hipMemcpyHostToDevice));
  std::lock_guard<std::mutex> lk(g_mu);
  for (int a = 0; a < 3; a++) g_scale_knn[a] = scale[a];
  return TGMM_OK;
}

extern "C" int tgmm_GMEMinitializeMemoryFromRaw(float** queryCUDA, float** imgDataCUDA,
                                               const float* scale, float** rnkCUDA,
                                               int** indCUDA, int** indCUDAtr,
                                               float** centroidLabelPositionCUDA,
                                               long long** labelListPtrCUDA,
                                               const tgmm_raw_frame* frame) {
  if (!queryCUDA || !imgDataCUDA || !rnkCUDA || !indCUDA || !indCUDAtr ||
      !centroidLabelPositionCUDA || !labelListPtrCUDA || !scale || !frame || frame->N < 0 ||
      frame->S < 0) {
    tgmm_set_error("GMEMinitializeMemoryFromRaw: NULL argument or negative size");
    return TGMM_ERR_ARG;
  }
  if (frame->S > 65535) {
    tgmm_set_error("GMEMinitializeMemoryFromRaw: %lld supervoxels (the drop-in counts them in "
                   "an unsigned short)", frame->S);
    return TGMM_ERR_LIMIT;
  }
  const long long N = frame->N, S = frame->S, K = TGMM_LEGACY_K;
  float** outs_f[4] = {queryCUDA, imgDataCUDA, rnkCUDA, centroidLabelPositionCUDA};
  for (float** p : outs_f) *p = nullptr;
  *indCUDA = *indCUDAtr = nullptr;
  *labelListPtrCUDA = nullptr;
  auto M = [](void** p, size_t n) { return hipMalloc(p, n > 0 ? n : 4); };
  int rc = TGMM_OK;
  hipError_t e = hipSuccess;
  if ((e = M((void**)queryCUDA, 3 * N * sizeof(float))) != hipSuccess ||
      (e = M((void**)imgDataCUDA, N * sizeof(float))) != hipSuccess ||
      (e = M((void**)rnkCUDA, K * S * sizeof(float))) != hipSuccess ||
      (e = M((void**)indCUDA, K * S * sizeof(int))) != hipSuccess ||
      (e = M((void**)indCUDAtr, K * S * sizeof(int))) != hipSuccess ||
      (e = M((void**)centroidLabelPositionCUDA, 3 * S * sizeof(float))) != hipSuccess ||
      (e = M((void**)labelListPtrCUDA, (S + 1) * sizeof(long long))) != hipSuccess) {
    tgmm_set_error("GMEMinitializeMemoryFromRaw: hipMalloc failed: %s", hipGetErrorString(e));
    rc = e == hipErrorOutOfMemory ? TGMM_ERR_OOM : TGMM_ERR_DEVICE;
  }
  if (!rc)
    rc = tgmm_marshal_frame(frame, *queryCUDA, *imgDataCUDA, *centroidLabelPositionCUDA,
                            *labelListPtrCUDA);
  if (rc) {
    tgmm_GMEMreleaseMemoryWithSupervoxels(queryCUDA, imgDataCUDA, rnkCUDA, indCUDA, indCUDAtr,
                                          centroidLabelPositionCUDA, labelListPtrCUDA);
    return rc;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  for (int a = 0; a < 3; a++) g_scale_knn[a] = scale[a];
  return TGMM_OK;
}

// Voxel mode (GMEMvariationalInferenceCUDA, GMEMupdateCUDA.cu:3420-3973):
// the same plain per-frame solve over one-voxel supervoxels, with the
// reference's size limits (:3425-3440) and outputs of K x query_nb. The
// reference writes no debug dump in this mode (its dump code is commented
// out), so debugPath is ignored.
extern "C" int tgmm_GMEMvariationalInferenceCUDA(float* queryCUDA, float* imgDataCUDA,
                                                 float* rnkCUDA, float* rnkCUDAtr, int* indCUDA,
                                                 int* indCUDAtr,
                                                 GaussianMixtureModelCUDA* vecGMtemp,
                                                 long long query_nb, int ref_nb, int maxIterEM,
                                                 double tolLikelihood, int devCUDA, int frame,
                                                 int W4DOF, const char* debugPath) {
  (void)rnkCUDAtr;
  (void)debugPath;
  if (query_nb > 1024LL * 65535LL) {
    tgmm_set_error("Too many query points for the GPU resources (%lld)", query_nb);
    return TGMM_ERR_LIMIT;
  }
  if (ref_nb > 65535) {
    tgmm_set_error("Too many reference points for the GPU resources (%d)", ref_nb);
    return TGMM_ERR_LIMIT;
  }
  if (query_nb < 0 || ref_nb < 0 || (ref_nb > 0 && !vecGMtemp) ||
      (query_nb > 0 && (!queryCUDA || !imgDataCUDA || !rnkCUDA || !indCUDA))) {
    tgmm_set_error("GMEMvariationalInferenceCUDA: NULL argument");
    return TGMM_ERR_ARG;
  }
  tgmm_em_opts o;
  float sreg[3];
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_lmin < 0) {
      tgmm_set_error(
          "before GMEMregularizeWkKernel: regularizePrecisionMatrixConstants were not set");
      return TGMM_ERR_NOT_SET;
    }
    std::memset(&o, 0, sizeof(o));
    o.K = TGMM_LEGACY_K;
    o.max_iter = maxIterEM;
    o.tol = tolLikelihood;
    o.lambda_min = g_lmin;
    o.lambda_max = g_lmax;
    o.max_excentricity = g_emax;
    o.w4dof = W4DOF ? 1 : 0;
    for (int a = 0; a < 3; a++) {
      o.scale[a] = g_scale_knn[a];
      sreg[a] = g_scale_reg[a];
    }
    o.max_sv_per_gaussian = TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN;
    o.voxel_mode = 1;
  }
  HIPCHK(hipSetDevice(devCUDA));
  FrameSrc f;
  f.N = query_nb;
  f.G = ref_nb;
  f.query = queryCUDA;
  f.img = imgDataCUDA;
  f.gm = vecGMtemp;
  return run_frame(devCUDA, f, o, sreg, vecGMtemp, rnkCUDA, indCUDA, indCUDAtr, query_nb,
                   nullptr, frame);
}

extern "C" int tgmm_GMEMreleaseMemoryWithSupervoxels(float** queryCUDA, float** imgDataCUDA,
                                                     float** rnkCUDA, int** indCUDA,
                                                     int** indCUDAtr,
                                                     float** centroidLabelPositionCUDA,
                                                     long long** labelListPtrCUDA) {
  void** ps[7] = {(void**)queryCUDA, (void**)imgDataCUDA, (void**)rnkCUDA, (void**)indCUDA,
                  (void**)indCUDAtr, (void**)centroidLabelPositionCUDA, (void**)labelListPtrCUDA};
  int rc = TGMM_OK;
  for (void** p : ps) {
    if (!p) {
      rc = TGMM_ERR_ARG;
      continue;
    }
    if (*p) {
      hipError_t e = hipFree(*p);
      if (e != hipSuccess) {
        tgmm_set_error("hipFree failed: %s", hipGetErrorString(e));
        rc = TGMM_ERR_DEVICE;
      }
    }
    *p = nullptr;
  }
  return rc;
}

extern "C" int tgmm_GMEMvariationalInferenceCUDAWithSupervoxels(
    float* queryCUDA, float* imgDataCUDA, float* rnkCUDA, float* rnkCUDAtr, int* indCUDA,
    int* indCUDAtr, float* centroidLabelPositionCUDA, long long* labelListPtrCUDA,
    GaussianMixtureModelCUDA* vecGMtemp, long long query_nb, int ref_nb,
    unsigned short numLabels, int maxIterEM, double tolLikelihood, int devCUDA, int frame,
    int W4DOF, const char* debugPath) {
  (void)rnkCUDAtr;
  // limits of the reference (GMEM.cu:3992-4007)
  if (query_nb > 1024LL * 65535LL) {
    tgmm_set_error("Too many query points for the GPU resources (%lld)", query_nb);
    return TGMM_ERR_LIMIT;
  }
  if (ref_nb > 65535) {
    tgmm_set_error("Too many reference points for the GPU resources (%d)", ref_nb);
    return TGMM_ERR_LIMIT;
  }
  if (query_nb < 0 || ref_nb < 0 || !labelListPtrCUDA || (ref_nb > 0 && !vecGMtemp) ||
      (numLabels > 0 && (!rnkCUDA || !indCUDA || !centroidLabelPositionCUDA)) ||
      (query_nb > 0 && (!queryCUDA || !imgDataCUDA))) {
    tgmm_set_error("GMEMvariationalInferenceCUDAWithSupervoxels: NULL argument");
    return TGMM_ERR_ARG;
  }
  tgmm_em_opts o;
  float sreg[3];
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_lmin < 0) {
      tgmm_set_error(
          "before GMEMregularizeWkKernel: regularizePrecisionMatrixConstants were not set");
      return TGMM_ERR_NOT_SET;
    }
    std::memset(&o, 0, sizeof(o));
    o.K = TGMM_LEGACY_K;
    o.max_iter = maxIterEM;
    o.tol = tolLikelihood;
    o.lambda_min = g_lmin;
    o.lambda_max = g_lmax;
    o.max_excentricity = g_emax;
    o.w4dof = W4DOF ? 1 : 0;
    for (int a = 0; a < 3; a++) {
      o.scale[a] = g_scale_knn[a];
      sreg[a] = g_scale_reg[a];
    }
    o.max_sv_per_gaussian = TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN;
  }
  HIPCHK(hipSetDevice(devCUDA));
  const long long S = numLabels;
  std::vector<long long> ptr(S + 1);
  HIPCHK(hipMemcpy(ptr.data(), labelListPtrCUDA, (S + 1) * sizeof(long long),
                   hipMemcpyDeviceToHost));
  FrameSrc f;
  f.N = query_nb;
  f.S = S;
  f.G = ref_nb;
  f.query = queryCUDA;
  f.img = imgDataCUDA;
  f.cent = centroidLabelPositionCUDA;
  f.ptr = ptr.data();
  f.gm = vecGMtemp;
  return run_frame(devCUDA, f, o, sreg, vecGMtemp, rnkCUDA, indCUDA, indCUDAtr, S, debugPath,
                   frame);
}

// ------------------------------------------- UtilsCUDA/knnCuda.h helpers --
// selectCUDAdeviceMaxMemory (knnCuda.h:23, knnCuda.cu:27-47): the device with
// the most global memory, the first of equals; -1 when there is none.
extern "C" int tgmm_selectCUDAdeviceMaxMemory(int* device) {
  if (!device) {
    tgmm_set_error("selectCUDAdeviceMaxMemory: device is NULL");
    return TGMM_ERR_ARG;
  }
  *device = -1;
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  size_t best = 0;
  for (int i = 0; i < count; i++) {
    size_t total = 0;
    HIPCHK(hipDeviceTotalMem(&total, i));
    if (total > best) {
      best = total;
      *device = i;
    }
  }
  return TGMM_OK;
}

// setDeviceCUDA (knnCuda.h:20, knnCuda.cu:737-741)
extern "C" int tgmm_setDeviceCUDA(int devCUDA) {
  HIPCHK(hipSetDevice(devCUDA));
  return TGMM_OK;
}

// uploadScaleCUDA (knnCuda.h:21, knnCuda.cu:827-830): the kNN anisotropy the
// VI entry points use (the init entry points record it too).
extern "C" int tgmm_uploadScaleCUDA(const float* scale) {
  if (!scale) {
    tgmm_set_error("uploadScaleCUDA: scale is NULL");
    return TGMM_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  for (int a = 0; a < 3; a++) g_scale_knn[a] = scale[a];
  return TGMM_OK;
}

// GMEMinitializeMemory (GMEMupdateCUDA.h:129-131, GMEMupdateCUDA.cu:4538-4579):
// the voxel-mode device arrays query [3N] SoA, img [N], rnk, ind, indTr
// [5N] on the current device, the host arrays copied, scale recorded as the
// kNN anisotropy.
extern "C" int tgmm_GMEMinitializeMemory(float** queryCUDA, const float* query,
                                         float** imgDataCUDA, const float* imgData,
                                         const float* scale, long long query_nb, float** rnkCUDA,
                                         int** indCUDA, int** indCUDAtr) {
  if (!queryCUDA || !imgDataCUDA || !rnkCUDA || !indCUDA || !indCUDAtr || !scale ||
      query_nb < 0 || (query_nb > 0 && (!query || !imgData))) {
    tgmm_set_error("GMEMinitializeMemory: NULL argument or negative size");
    return TGMM_ERR_ARG;
  }
  const long long N = query_nb, K = TGMM_LEGACY_K;
  *queryCUDA = *imgDataCUDA = *rnkCUDA = nullptr;
  *indCUDA = *indCUDAtr = nullptr;
  auto M = [](void** p, size_t n) { return hipMalloc(p, n > 0 ? n : 4); };
  hipError_t e = hipSuccess;
  if ((e = M((void**)queryCUDA, 3 * N * sizeof(float))) != hipSuccess ||
      (e = M((void**)imgDataCUDA, N * sizeof(float))) != hipSuccess ||
      (e = M((void**)rnkCUDA, K * N * sizeof(float))) != hipSuccess ||
      (e = M((void**)indCUDA, K * N * sizeof(int))) != hipSuccess ||
      (e = M((void**)indCUDAtr, K * N * sizeof(int))) != hipSuccess ||
      (N > 0 && (e = hipMemcpy(*queryCUDA, query, 3 * N * sizeof(float),
                               hipMemcpyHostToDevice)) != hipSuccess) ||
      (N > 0 && (e = hipMemcpy(*imgDataCUDA, imgData, N * sizeof(float),
                               hipMemcpyHostToDevice)) != hipSuccess)) {
    tgmm_set_error("GMEMinitializeMemory: %s", hipGetErrorString(e));
    tgmm_GMEMreleaseMemory(queryCUDA, imgDataCUDA, rnkCUDA, indCUDA, indCUDAtr);
    return e == hipErrorOutOfMemory ? TGMM_ERR_OOM : TGMM_ERR_DEVICE;
  }
  return tgmm_uploadScaleCUDA(scale);
}

// GMEMreleaseMemory (GMEMupdateCUDA.h:138-139, GMEMupdateCUDA.cu:4637-4653):
// frees and nulls every pointer.
extern "C" int tgmm_GMEMreleaseMemory(float** queryCUDA, float** imgDataCUDA, float** rnkCUDA,
                                      int** indCUDA, int** indCUDAtr) {
  void** ps[5] = {(void**)queryCUDA, (void**)imgDataCUDA, (void**)rnkCUDA, (void**)indCUDA,
                  (void**)indCUDAtr};
  int rc = TGMM_OK;
  for (void** p : ps) {
    if (!p) {
      rc = TGMM_ERR_ARG;
      continue;
    }
    if (*p) {
      hipError_t e = hipFree(*p);
      if (e != hipSuccess) {
        tgmm_set_error("hipFree failed: %s", hipGetErrorString(e));
        rc = TGMM_ERR_DEVICE;
      }
    }
    *p = nullptr;
  }
  return rc;
}
