// tgmm_svknn.hip -- supervoxel nearest neighbours (knnCUDA_,
// temporalLogicalRules/knnCUDA/knnCuda.cu:142-232 with its kernel :56-140):
// for each query centroid, the K <= 11 nearest reference centroids under the
// scaled squared distance, in the reference's slot layout. The tracker calls
// it three times per frame (supervoxels against themselves, the previous and
// the next frame; mainTrackingGaussianMixtureModel.cpp:874-888,
// lineageHyperTree.cpp:2526-2775) and reads the slots in order.
//
// The slot layout is the outcome of a scan of the references in index
// order: a reference whose distance is not above the current threshold (the
// largest slot distance) replaces the first slot holding it. The kernel
// replays that scan exactly, one lane per query. Per block of 64 references
// each lane first marks, branch-free, the references not above its
// threshold at the block's start (wave-uniform loads), a superset of the
// ones the scan inserts since the threshold only falls; it then walks only
// its marks, in index order, with the exact test against the current
// threshold, so the wave pays the slot replacement for its busiest lane's
// candidates rather than for every reference. It skips whole blocks of 64 consecutive
// references that no lane of the wave can insert: a block's bounding box
// gives, per axis, a gap no larger than any of its references' |q - r|, and
// the bound is formed with the distance's own float operations in the same
// order, so by monotonicity of rounding it never exceeds a member's computed
// distance; a block is skipped only when the bound is above every lane's
// threshold, i.e. when the scan would insert none of its references. A block
// holding a NaN coordinate has an unbounded box and is never skipped, nor is
// any block for a query with a NaN coordinate. The exact test is the
// reference's, with its early exits after the x term and the x + y partial
// sum (a NaN full distance is taken only when those partial sums are not
// above the threshold).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/tgmm_em.h"
#include "tgmm_runtime_internal.h"

namespace {

constexpr int kKB = 64;  // references per bounding box
constexpr int kKM = TGMM_SV_KNN_MAX;
constexpr int kKT = 64;  // threads (queries) per block

// Bounding box of each block of 64 references: box[a][b] = min, box[3+a][b]
// = max of coordinate a (-inf, +inf when the block holds a NaN).
__global__ __launch_bounds__(kKB) void k_svknn_boxes(const float* __restrict__ ref, int R,
                                                     int Rp, int nblk, float* __restrict__ box) {
  const int b = blockIdx.x;
  const int r = b * kKB + threadIdx.x;
  float lo[3], hi[3];
  int nan = 0;
  for (int a = 0; a < 3; a++) {
    const float v = r < R ? ref[r + (long long)a * Rp] : NAN;
    const bool in = r < R;
    nan |= (in && isnan(v)) ? 1 : 0;
    lo[a] = in ? v : INFINITY;
    hi[a] = in ? v : -INFINITY;
  }
  for (int m = 32; m >= 1; m >>= 1)
    for (int a = 0; a < 3; a++) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], m, 64));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], m, 64));
    }
  const bool any_nan = __any(nan);
  if (threadIdx.x == 0)
    for (int a = 0; a < 3; a++) {
      box[(long long)a * nblk + b] = any_nan ? -INFINITY : lo[a];
      box[(long long)(3 + a) * nblk + b] = any_nan ? INFINITY : hi[a];
    }
}

// the reference's scaled squared distance, float, x then y then z, no contraction
__device__ inline float svdist(float q0, float q1, float q2, const float* __restrict__ ref,
                               int r, int R, float s0, float s1, float s2) {
#pragma clang fp contract(off)
  float a = (q0 - ref[r]) * s0;
  float d = a * a;
  a = (q1 - ref[r + R]) * s1;
  d += a * a;
  a = (q2 - ref[r + 2 * R]) * s2;
  d += a * a;
  return d;
}

// The reference's insertion test (knnCuda.cu:103-113): it leaves a reference
// as soon as the x term, or the x + y partial sum, is above the threshold,
// and takes it when the full sum is not. For finite coordinates the partial
// sums are non-decreasing and only the full sum decides; a NaN makes the sums
// after it NaN (never above the threshold), so the partial sums before it
// still decide. Returns true with d = the full sum when the reference is
// taken.
__device__ inline bool svtake(float q0, float q1, float q2, const float* __restrict__ ref,
                              int r, int R, float s0, float s1, float s2, float thr, float& d) {
#pragma clang fp contract(off)
  float a = (q0 - ref[r]) * s0;
  const float p1 = a * a;
  a = (q1 - ref[r + R]) * s1;
  const float p2 = p1 + a * a;
  a = (q2 - ref[r + 2 * R]) * s2;
  d = p2 + a * a;
  return !(p1 > thr) && !(p2 > thr) && !(d > thr);
}

// ref is SoA with axis stride Rp = nblk * 64 (padded with zeros past R)
__global__ __launch_bounds__(kKT) void k_svknn(const float* __restrict__ query, long long Q,
                                               const float* __restrict__ ref, int R, int Rp,
                                               int K,
                                               float s0, float s1, float s2,
                                               const float* __restrict__ box, int nblk,
                                               int* __restrict__ ind, float* __restrict__ dist) {
#pragma clang fp contract(off)
  const long long q = (long long)blockIdx.x * kKT + threadIdx.x;
  const bool live = q < Q;
  const float q0 = live ? query[q] : 0.0f, q1 = live ? query[q + Q] : 0.0f,
              q2 = live ? query[q + 2 * Q] : 0.0f;
  const float a0 = fabsf(s0), a1 = fabsf(s1), a2 = fabsf(s2);
  // a NaN query coordinate makes every distance NaN, which the scan inserts:
  // such a lane takes every block
  const bool qnan = live && (isnan(q0) || isnan(q1) || isnan(q2));
  float md[kKM];
  int mi[kKM];
#pragma unroll
  for (int j = 0; j < kKM; j++) {
    md[j] = 1e32f;
    mi[j] = -1;
  }
  float thr = live ? 1e32f : -INFINITY;  // a lane past the end inserts nothing
  int pos = 0;
  for (int b = 0; b < nblk; b++) {
    // the block's lower bound for this query, in the distance's operation order
    const float g0 = fmaxf(fmaxf(box[b] - q0, q0 - box[3 * nblk + b]), 0.0f) * a0;
    const float g1 = fmaxf(fmaxf(box[nblk + b] - q1, q1 - box[4 * nblk + b]), 0.0f) * a1;
    const float g2 = fmaxf(fmaxf(box[2 * nblk + b] - q2, q2 - box[5 * nblk + b]), 0.0f) * a2;
    float lb = g0 * g0;
    lb += g1 * g1;
    lb += g2 * g2;
    if (!__any(qnan || !(lb > thr))) continue;  // no lane would insert from this block
    const int rb = b * kKB;
    const int n = R - rb < kKB ? R - rb : kKB;
    // the references whose full distance is not above this lane's threshold
    // at the block's start: a superset of those the scan inserts (the
    // threshold only falls, and a NaN distance is marked)
    // (the reference arrays are padded to whole blocks, so the 64 loads of a
    // block are unconditional and batch; the padding's marks are cleared)
    unsigned long long cand = 0ull;
    const float t0 = thr;
#pragma unroll
    for (int u = 0; u < kKB; u++) {
      const float d = svdist(q0, q1, q2, ref, rb + u, Rp, s0, s1, s2);
      cand |= (d > t0) ? 0ull : (1ull << u);
    }
    const unsigned long long nmask = n < kKB ? (1ull << n) - 1ull : ~0ull;
    cand &= nmask;
    // the scan over the candidates, in index order, with the current threshold
    while (cand) {
      const int u = __builtin_ctzll(cand);
      cand &= cand - 1ull;
      float d;
      if (!svtake(q0, q1, q2, ref, rb + u, Rp, s0, s1, s2, thr, d)) continue;
#pragma unroll
      for (int j = 0; j < kKM; j++)
        if (j == pos) {
          md[j] = d;
          mi[j] = rb + u;
        }
      thr = md[0];
      pos = 0;
#pragma unroll
      for (int j = 1; j < kKM; j++)
        if (j < K && thr < md[j]) {
          thr = md[j];
          pos = j;
        }
      // a NaN in slot 0 makes the threshold NaN (findMaxPosition's `<` never
      // moves off it): the scan then takes the next reference, whatever its
      // distance, so the rest of the block is walked
      if (thr != thr && u + 1 < kKB) cand |= (~0ull << (u + 1)) & nmask;
    }
  }
  if (!live) return;
#pragma unroll
  for (int j = 0; j < kKM; j++)
    if (j < K) {
      ind[q + j * Q] = mi[j];
      if (dist) dist[q + j * Q] = md[j];
    }
}

inline size_t up256(size_t n) { return (n + 255) & ~size_t(255); }

struct KnnWs {
  void* dev = nullptr;
  size_t cap = 0;
};
std::mutex g_mu;
std::vector<KnnWs> g_ws;

int hip_fail(hipError_t e, const char* what) {
  tgmm_set_error("knnCUDA_: %s failed: %s", what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? TGMM_ERR_OOM : TGMM_ERR_DEVICE;
}

}  // namespace

extern "C" int tgmm_knnCUDA_(int* ind, float* dist, const float* query, const float* ref,
                             long long query_nb, int ref_nb, int KNN, const float* scale,
                             int devCUDA) {
  if (KNN < 1 || KNN > kKM) {
    tgmm_set_error("knnCUDA_: KNN = %d outside 1..%d (maxKNN)", KNN, kKM);
    return TGMM_ERR_LIMIT;
  }
  if (query_nb < 0 || !scale) {
    tgmm_set_error("knnCUDA_: negative query_nb or NULL scale");
    return TGMM_ERR_ARG;
  }
  if (ref_nb <= 0) {  // no possible assignments (knnCuda.cu:165-171)
    if (dist)
      for (long long i = 0; i < query_nb * KNN; i++) dist[i] = 1e32f;
    return TGMM_OK;
  }
  if (query_nb == 0) return TGMM_OK;
  if (!ind || !query || !ref) {
    tgmm_set_error("knnCUDA_: NULL argument");
    return TGMM_ERR_ARG;
  }
  hipError_t e = hipSetDevice(devCUDA);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  const long long Q = query_nb, R = ref_nb;
  const int nblk = (int)((R + kKB - 1) / kKB);
  const long long Rp = (long long)nblk * kKB;
  const size_t o_q = 0;
  const size_t o_r = o_q + up256((size_t)Q * 12);
  const size_t o_box = o_r + up256((size_t)Rp * 12);
  const size_t o_ind = o_box + up256((size_t)nblk * 24);
  const size_t o_dist = o_ind + up256((size_t)Q * KNN * 4);
  const size_t total = o_dist + up256((size_t)Q * KNN * 4);
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_ws.size() <= devCUDA) g_ws.resize(devCUDA + 1);
  KnnWs& ws = g_ws[devCUDA];
  if (total > ws.cap) {
    if (ws.dev) (void)hipFree(ws.dev);
    ws.dev = nullptr;
    ws.cap = 0;
    if ((e = hipMalloc(&ws.dev, total)) != hipSuccess) return hip_fail(e, "hipMalloc");
    ws.cap = total;
  }
  char* base = (char*)ws.dev;
  auto* dq = (float*)(base + o_q);
  auto* dr = (float*)(base + o_r);
  auto* dbox = (float*)(base + o_box);
  auto* dind = (int*)(base + o_ind);
  auto* ddist = (float*)(base + o_dist);
  if ((e = hipMemcpyAsync(dq, query, (size_t)Q * 12, hipMemcpyHostToDevice, 0)) != hipSuccess)
    return hip_fail(e, "upload of the queries");
  for (int a = 0; a < 3; a++)
    if ((e = hipMemcpyAsync(dr + a * Rp, ref + a * R, (size_t)R * 4, hipMemcpyHostToDevice, 0)) !=
            hipSuccess ||
        (Rp > R && (e = hipMemsetAsync(dr + a * Rp + R, 0, (size_t)(Rp - R) * 4, 0)) !=
                       hipSuccess))
      return hip_fail(e, "upload of the references");
  k_svknn_boxes<<<nblk, kKB, 0, 0>>>(dr, (int)R, (int)Rp, nblk, dbox);
  k_svknn<<<(unsigned)((Q + kKT - 1) / kKT), kKT, 0, 0>>>(dq, Q, dr, (int)R, (int)Rp, KNN,
                                                        scale[0], scale[1],
                                                      scale[2], dbox, nblk, dind,
                                                      dist ? ddist : nullptr);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "kernel launch");
  if ((e = hipMemcpy(ind, dind, (size_t)Q * KNN * 4, hipMemcpyDeviceToHost)) != hipSuccess ||
      (dist && (e = hipMemcpy(dist, ddist, (size_t)Q * KNN * 4, hipMemcpyDeviceToHost)) !=
                   hipSuccess))
    return hip_fail(e, "download of the neighbours");
  return TGMM_OK;
}
