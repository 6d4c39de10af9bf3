// tgmm_local.hip -- local log-likelihood over lists of supervoxels
// (calculateLocalLikelihood, GMEMupdateCUDA.cu:2952-3050, with its kernel
// GMEMcomputeLocalLikelihoodKernelWithSupervoxels, :1989-2103, and the sum
// addLocalLikelihoodWithSupervoxels, :2876-2945): the before/after values of
// the tracker's cell-division likelihood-ratio test
// (mainTrackingGaussianMixtureModel.cpp:1336-1375).
//
// One call uploads the records and lists in one copy, evaluates every list
// in three launches on the null stream and downloads the list sums:
//   k_local_alpha  sum of alpha over all records, 256-record partials;
//   k_local_ll     one 16-lane group per (list, supervoxel) entry: the
//                  supervoxel's candidates (ind, K = 5 slots) staged in LDS,
//                  then each voxel's density evaluated in FP64 exactly as the
//                  reference writes it, img * log(density) summed per entry;
//   k_local_lists  one wave per list, the entries' sums in a fixed order.
// The result depends on neither the device nor the number of lists: the sum
// of alpha is the same fixed tree in every block, each entry's voxels go to
// lane (i mod 16) in increasing order with a fixed butterfly, and each list
// sums its entries by lane (p mod 64) then a fixed butterfly.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/tgmm_em.h"
#include "tgmm_device.h"
#include "tgmm_runtime_internal.h"

namespace {

constexpr int kLG = 16;           // lanes per listed supervoxel
constexpr int kLB = 256;          // threads per block
constexpr int kLPB = kLB / kLG;   // listed supervoxels per block
constexpr int kLK = TGMM_LEGACY_K;

struct LCand {
  double m[3];
  double W[6];
  double nu;
  double coef;   // sqrt(det W nu^3) / (2 pi)^1.5 (NaN for an indefinite W)
  double ratio;  // alpha / sum_alpha
  int ok;        // slot holds a Gaussian
  int pad;
};

// 256-record partial sums of alpha (GMEMtotalAlphaKernel, GMEM.cu:579-604).
__global__ __launch_bounds__(kLB) void k_local_alpha(const GaussianMixtureModelCUDA* rec, int G,
                                                     double* part) {
  __shared__ double t[kLB];
  const long long k = (long long)blockIdx.x * kLB + threadIdx.x;
  t[threadIdx.x] = k < G ? rec[k].alpha_k : 0.0;
  __syncthreads();
  for (int w = kLB / 2; w >= 1; w >>= 1) {
    if (threadIdx.x < w) t[threadIdx.x] += t[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = t[0];
}

// Entry p = blockIdx.x * 16 + group: supervoxel s = entries[p]. Per voxel,
// sum_j sqrt(det W_j nu_j^3) / (2 pi)^1.5 * exp(-nu_j d^T W_j d / 2) *
// alpha_j / sum_alpha over the occupied slots (GMEM.cu:2049-2058), and
// img * log of it where the sum is > 0 (:2067-2068; NaN and 0 add nothing).
__global__ __launch_bounds__(kLB) void k_local_ll(
    const float* __restrict__ query, const float* __restrict__ img, long long N,
    const long long* __restrict__ lptr, const int* __restrict__ ind, long long S,
    const GaussianMixtureModelCUDA* __restrict__ rec, int G, const double* __restrict__ part,
    int nparts, const int* __restrict__ entries, long long nentries,
    double* __restrict__ entry_ll) {
  __shared__ double red[kLB];
  __shared__ LCand cst[kLPB][kLK];
  // sum of alpha over every record: the same tree in every block
  double a = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kLB) a += part[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int w = kLB / 2; w >= 1; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double sum_alpha = red[0];
  const int lane = threadIdx.x % kLG, grp = threadIdx.x / kLG;
  const long long p = (long long)blockIdx.x * kLPB + grp;
  const bool live = p < nentries;
  const long long s = live ? entries[p] : 0;
  if (live && lane < kLK) {
    LCand& c = cst[grp][lane];
    const int k = ind[s + (long long)lane * S];
    c.ok = (k >= 0 && k < G) ? 1 : 0;
    if (c.ok) {
      const GaussianMixtureModelCUDA& g = rec[k];
      for (int i = 0; i < 3; i++) c.m[i] = g.m_k[i];
      for (int i = 0; i < 6; i++) c.W[i] = g.W_k[i];
      c.nu = g.nu_k;
      c.coef = sqrt(tgmm::det3(c.W) * pow(c.nu, 3.0)) / 15.749609945722419;  // (2 pi)^1.5
      c.ratio = g.alpha_k / sum_alpha;
    }
  }
  __syncthreads();
  if (!live) return;
  const long long beg = lptr[s] > 0 ? lptr[s] : 0;
  const long long end = lptr[s + 1] < N ? lptr[s + 1] : N;
  double acc = 0.0;
  for (long long v = beg + lane; v < end; v += kLG) {
    const float x0 = query[v], x1 = query[v + N], x2 = query[v + 2 * N];
    double like = 0.0;
#pragma unroll
    for (int j = 0; j < kLK; j++) {
      const LCand& c = cst[grp][j];
      if (!c.ok) continue;
      const double d0 = x0 - c.m[0], d1 = x1 - c.m[1], d2 = x2 - c.m[2];
      const double maha =
          c.nu * (c.W[0] * d0 * d0 + c.W[3] * d1 * d1 + c.W[5] * d2 * d2 + 2.0 * c.W[1] * d0 * d1 +
                  2.0 * c.W[2] * d0 * d2 + 2.0 * c.W[4] * d1 * d2);
      double aux = c.coef;
      aux *= exp(-0.5 * maha);
      aux *= c.ratio;
      like += aux;
    }
    if (like > 0.0) acc += (double)img[v] * log(like);
  }
#pragma unroll
  for (int m = kLG / 2; m >= 1; m >>= 1) acc += __shfl_xor(acc, m, kLG);
  if (lane == 0) entry_ll[p] = acc;
}

// List i = entries [list_ptr[i], list_ptr[i+1]); one wave per list.
__global__ __launch_bounds__(kLB) void k_local_lists(const double* __restrict__ entry_ll,
                                                     const long long* __restrict__ list_ptr,
                                                     int n_lists, double* __restrict__ out) {
  const int lane = threadIdx.x % 64;
  const long long i = (long long)blockIdx.x * (kLB / 64) + threadIdx.x / 64;
  if (i >= n_lists) return;  // whole waves
  double a = 0.0;
  for (long long q = list_ptr[i] + lane; q < list_ptr[i + 1]; q += 64) a += entry_ll[q];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) a += __shfl_xor(a, m, 64);
  if (lane == 0) out[i] = a;
}

inline size_t up256(size_t n) { return (n + 255) & ~size_t(255); }
inline unsigned nblocks(long long n, int per) { return (unsigned)((n + per - 1) / per); }

// Per-device workspace kept between calls (grow-only, held until the process
// ends): device memory and pinned host staging. Calls are serialised by one
// mutex, so a call's buffers are never in use by another.
struct LocalWs {
  void* dev = nullptr;
  size_t dev_cap = 0;
  void* host = nullptr;
  size_t host_cap = 0;
  hipError_t reserve(size_t dev_bytes, size_t host_bytes) {
    hipError_t e = hipSuccess;
    if (dev_bytes > dev_cap) {
      if (dev) (void)hipFree(dev);
      dev = nullptr;
      dev_cap = 0;
      if ((e = hipMalloc(&dev, dev_bytes)) != hipSuccess) return e;
      dev_cap = dev_bytes;
    }
    if (host_bytes > host_cap) {
      if (host) (void)hipHostFree(host);
      host = nullptr;
      host_cap = 0;
      if ((e = hipHostMalloc(&host, host_bytes, hipHostMallocDefault)) != hipSuccess) return e;
      host_cap = host_bytes;
    }
    return e;
  }
};
std::mutex g_ws_mu;
std::vector<LocalWs> g_ws;

LocalWs& workspace_for(int dev) {
  if ((int)g_ws.size() <= dev) g_ws.resize(dev + 1);
  return g_ws[dev];
}

int hip_fail(hipError_t e, const char* what) {
  tgmm_set_error("calculateLocalLikelihood: %s failed: %s", what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? TGMM_ERR_OOM : TGMM_ERR_DEVICE;
}

}  // namespace

extern "C" int tgmm_calculateLocalLikelihood(
    double* localLikelihood, const long long* list_ptr, const int* list_idx, int n_lists,
    const float* queryCUDA, const float* imgDataCUDA, const float* rnkCUDA, const int* indCUDA,
    const GaussianMixtureModelCUDA* vecGMtemp, const long long* labelListPtrCUDA,
    long long query_nb, int ref_nb, int numLabels) {
  (void)rnkCUDA;  // passed and unused, as in the reference
  if (n_lists < 0 || query_nb < 0 || ref_nb < 0 || numLabels < 0 ||
      (n_lists > 0 && (!localLikelihood || !list_ptr)) || (ref_nb > 0 && !vecGMtemp)) {
    tgmm_set_error("calculateLocalLikelihood: NULL argument or negative size");
    return TGMM_ERR_ARG;
  }
  if (n_lists == 0) return TGMM_OK;
  if (list_ptr[0] != 0) {
    tgmm_set_error("calculateLocalLikelihood: list_ptr[0] must be 0");
    return TGMM_ERR_ARG;
  }
  for (int i = 0; i < n_lists; i++)
    if (list_ptr[i + 1] < list_ptr[i]) {
      tgmm_set_error("calculateLocalLikelihood: list_ptr decreases at list %d", i);
      return TGMM_ERR_ARG;
    }
  const long long P = list_ptr[n_lists];
  if (P == 0) return TGMM_OK;  // every list empty: nothing written
  if (P > INT_MAX) {
    tgmm_set_error("calculateLocalLikelihood: %lld list entries (at most %d)", P, INT_MAX);
    return TGMM_ERR_LIMIT;
  }
  if (!list_idx) {
    tgmm_set_error("calculateLocalLikelihood: list_idx is NULL");
    return TGMM_ERR_ARG;
  }
  for (long long q = 0; q < P; q++)
    if (list_idx[q] < 0 || list_idx[q] >= numLabels) {
      tgmm_set_error("calculateLocalLikelihood: supervoxel index %d at entry %lld is outside "
                     "[0, %d)", list_idx[q], q, numLabels);
      return TGMM_ERR_ARG;
    }
  if (!indCUDA || !labelListPtrCUDA || (query_nb > 0 && (!queryCUDA || !imgDataCUDA))) {
    tgmm_set_error("calculateLocalLikelihood: NULL device argument");
    return TGMM_ERR_ARG;
  }
  // one device region: records, entries, list pointers (uploaded in one
  // copy from the same layout in pinned staging), alpha partials, entry
  // sums, list sums
  const int nparts = (ref_nb + kLB - 1) / kLB;
  const size_t o_rec = 0;
  const size_t o_ent = o_rec + up256((size_t)ref_nb * sizeof(GaussianMixtureModelCUDA));
  const size_t o_lp = o_ent + up256((size_t)P * sizeof(int));
  const size_t up_bytes = o_lp + (size_t)(n_lists + 1) * sizeof(long long);
  const size_t o_part = up256(up_bytes);
  const size_t o_ell = o_part + up256((size_t)nparts * sizeof(double));
  const size_t o_out = o_ell + up256((size_t)P * sizeof(double));
  const size_t total = o_out + up256((size_t)n_lists * sizeof(double));
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_ws_mu);
  LocalWs& ws = workspace_for(dev);
  if ((e = ws.reserve(total, up_bytes > (size_t)n_lists * sizeof(double)
                                 ? up_bytes : (size_t)n_lists * sizeof(double))) != hipSuccess)
    return hip_fail(e, "workspace allocation");
  char* base = (char*)ws.dev;
  char* stage = (char*)ws.host;
  if (ref_nb > 0)
    std::memcpy(stage + o_rec, vecGMtemp, (size_t)ref_nb * sizeof(GaussianMixtureModelCUDA));
  std::memcpy(stage + o_ent, list_idx, (size_t)P * sizeof(int));
  std::memcpy(stage + o_lp, list_ptr, (size_t)(n_lists + 1) * sizeof(long long));
  auto* rec = (GaussianMixtureModelCUDA*)(base + o_rec);
  auto* ent = (int*)(base + o_ent);
  auto* lp = (long long*)(base + o_lp);
  auto* part = (double*)(base + o_part);
  auto* ell = (double*)(base + o_ell);
  auto* out = (double*)(base + o_out);
  if ((e = hipMemcpyAsync(base, stage, up_bytes, hipMemcpyHostToDevice, 0)) != hipSuccess)
    return hip_fail(e, "upload of the records and lists");
  if (nparts > 0) k_local_alpha<<<nparts, kLB, 0, 0>>>(rec, ref_nb, part);
  k_local_ll<<<nblocks(P, kLPB), kLB, 0, 0>>>(queryCUDA, imgDataCUDA, query_nb, labelListPtrCUDA,
                                              indCUDA, numLabels, rec, ref_nb, part, nparts, ent,
                                              P, ell);
  k_local_lists<<<nblocks(n_lists, kLB / 64), kLB, 0, 0>>>(ell, lp, n_lists, out);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "kernel launch");
  if ((e = hipMemcpyAsync(stage, out, (size_t)n_lists * sizeof(double), hipMemcpyDeviceToHost,
                          0)) != hipSuccess ||
      (e = hipStreamSynchronize(0)) != hipSuccess)
    return hip_fail(e, "download of the results");
  const double* host = (const double*)stage;
  for (int i = 0; i < n_lists; i++)
    if (list_ptr[i + 1] > list_ptr[i]) localLikelihood[i] = host[i];
  return TGMM_OK;
}
