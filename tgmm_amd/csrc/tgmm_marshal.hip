// tgmm_marshal.hip -- the EM's inputs built on the device from a frame's raw
// supervoxel voxels: the tracker's local background subtraction, conversion
// and scaling of the image (mainTrackingGaussianMixtureModel.cpp:598-669), its
// voxel SoA (:774-870) and the supervoxel centroids
// (supervoxel::weightedGaussianStatistics<uint16>, supervoxel.cpp:1410-1449).
//
// A frame arrives as the supervoxels' linear voxel indices (PixelIdxList,
// 8 B per voxel), their raw uint16 intensities (2 B per voxel, or gathered on
// the device from a resident volume), the supervoxel pointer and each
// supervoxel's local background threshold. Three kernels:
//   k_mar_sv     one 16-lane group per supervoxel: coordinates (x fastest,
//                mylib coord4idx, array.p:557-576), the subtracted intensity
//                (uint16 arithmetic, as the reference's unsigned short), the
//                centroid of the raw intensities accumulated in float in list
//                order (each lane forms its voxel's products; the group adds
//                them in order through shuffles, with no contraction), and
//                per block the range of the subtracted image;
//   k_mar_range  one block: the image range from the block partials;
//   k_mar_scale  lane per voxel: Scale_Array_To_Range(img, 0, 1) over the
//                whole volume (mylib array.p:2708-2820), where voxels outside
//                every supervoxel are 0 (main.cpp:644-649):
//                img = (float)((1.0 / (M - m)) * (img - m)) in double.
// Supervoxels are disjoint (a segmentation), so the volume has background
// voxels exactly when N < d0 d1 d2; then m = 0. Coordinates are split from
// the linear index by reciprocal multiplication with a remainder fix-up
// (exact below 2^53 voxels; 64-bit integer division is emulated and costly).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/tgmm_em.h"
#include "tgmm_runtime_internal.h"

namespace {

constexpr int kMG = 16;   // lanes per supervoxel
constexpr int kMB = 256;  // threads per block

// floor(n / d) for n < 2^53: the double product with a rounded reciprocal is
// within one of the quotient, and the remainder fixes it.
__device__ inline unsigned long long div_exact(unsigned long long n, unsigned long long d,
                                               double inv, unsigned long long& r) {
  unsigned long long q = (unsigned long long)((double)n * inv);
  long long rem = (long long)(n - q * d);
  if (rem < 0) {
    q -= 1;
    rem += (long long)d;
  } else if (rem >= (long long)d) {
    q += 1;
    rem -= (long long)d;
  }
  r = (unsigned long long)rem;
  return q;
}

__global__ void k_mar_init(unsigned* range) {
  if (threadIdx.x == 0) range[2] = 0u;  // set when an index lies outside the volume
}

// range[0], range[1] = max, min of the subtracted image over the listed
// voxels, from the per-block partials of k_mar_sv (one block).
__global__ __launch_bounds__(1024) void k_mar_range(const unsigned* __restrict__ part_hi,
                                                    const unsigned* __restrict__ part_lo,
                                                    int nparts, unsigned* __restrict__ range) {
  __shared__ unsigned sh[1024 / 64], sl[1024 / 64];
  unsigned hi = 0u, lo = 0xFFFFu;
  for (int i = threadIdx.x; i < nparts; i += 1024) {
    hi = part_hi[i] > hi ? part_hi[i] : hi;
    lo = part_lo[i] < lo ? part_lo[i] : lo;
  }
  for (int m = 32; m >= 1; m >>= 1) {
    const unsigned h2 = __shfl_xor(hi, m, 64), l2 = __shfl_xor(lo, m, 64);
    hi = h2 > hi ? h2 : hi;
    lo = l2 < lo ? l2 : lo;
  }
  if (threadIdx.x % 64 == 0) {
    sh[threadIdx.x / 64] = hi;
    sl[threadIdx.x / 64] = lo;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 1024 / 64; w++) {
      hi = sh[w] > hi ? sh[w] : hi;
      lo = sl[w] < lo ? sl[w] : lo;
    }
    range[0] = hi;
    range[1] = lo;
  }
}

__global__ __launch_bounds__(kMB) void k_mar_sv(
    const unsigned long long* __restrict__ idx, const unsigned short* __restrict__ raw,
    const unsigned short* __restrict__ volume, const long long* __restrict__ ptr,
    const unsigned short* __restrict__ thr, long long S, long long N, unsigned long long d0,
    unsigned long long d1, unsigned long long vol, double inv0, double inv1,
    float* __restrict__ query,
    float* __restrict__ img, float* __restrict__ cent, unsigned* __restrict__ part_hi,
    unsigned* __restrict__ part_lo, unsigned* __restrict__ range) {
  __shared__ unsigned red_hi[kMB / 64], red_lo[kMB / 64];
  const int lane = threadIdx.x % kMG;
  const long long s = (long long)blockIdx.x * (kMB / kMG) + threadIdx.x / kMG;
  unsigned hi = 0u, lo = 0xFFFFu, bad = 0u;
  if (s < S) {
    const long long beg = ptr[s], end = ptr[s + 1];
    const unsigned short th = thr[s];
    float nk = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (long long base = beg; base < end; base += kMG) {  // the same trips in a group
      const long long v = base + lane;
      float w = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
      if (v < end) {
        const unsigned long long id = idx[v];
        unsigned long long x = 0, y = 0, z = 0;
        unsigned short r = 0;
        if (id >= vol) {
          bad = 1u;
        } else {
          const unsigned long long q = div_exact(id, d0, inv0, x);
          z = div_exact(q, d1, inv1, y);
          r = raw ? raw[v] : volume[id];
        }
        query[v] = (float)x;
        query[v + N] = (float)y;
        query[v + 2 * N] = (float)z;
        const unsigned short sub = (unsigned short)(r - th);
        img[v] = (float)sub;
        hi = sub > hi ? sub : hi;
        lo = sub < lo ? sub : lo;
        w = (float)r;
        px = w * (float)x;
        py = w * (float)y;
        pz = w * (float)z;
      }
      // the chunk's voxels in list order; lanes past the end add +0.0,
      // which leaves the (non-negative) sums unchanged
#pragma unroll
      for (int j = 0; j < kMG; j++) {
#pragma clang fp contract(off)
        nk += __shfl(w, j, kMG);
        cx += __shfl(px, j, kMG);
        cy += __shfl(py, j, kMG);
        cz += __shfl(pz, j, kMG);
      }
    }
    if (lane == 0) {
      const bool zero = (double)nk < 1e-5;
      cent[s] = zero ? 0.0f : cx / nk;
      cent[s + S] = zero ? 0.0f : cy / nk;
      cent[s + 2 * S] = zero ? 0.0f : cz / nk;
    }
  }
  // block range, to this block's partial
  for (int m = 32; m >= 1; m >>= 1) {
    const unsigned h2 = __shfl_xor(hi, m, 64), l2 = __shfl_xor(lo, m, 64);
    hi = h2 > hi ? h2 : hi;
    lo = l2 < lo ? l2 : lo;
  }
  if (threadIdx.x % 64 == 0) {
    red_hi[threadIdx.x / 64] = hi;
    red_lo[threadIdx.x / 64] = lo;
  }
  if (bad) atomicOr(range + 2, 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kMB / 64; w++) {
      hi = red_hi[w] > hi ? red_hi[w] : hi;
      lo = red_lo[w] < lo ? red_lo[w] : lo;
    }
    part_hi[blockIdx.x] = hi;
    part_lo[blockIdx.x] = lo;
  }
}

__global__ __launch_bounds__(kMB) void k_mar_scale(float* __restrict__ img, long long N,
                                                   const unsigned* __restrict__ range,
                                                   int background) {
  const long long v = (long long)blockIdx.x * kMB + threadIdx.x;
  if (v >= N) return;
  const unsigned hi = range[0];
  const unsigned lo = background ? 0u : range[1];
  if (hi == lo) return;  // a constant image is left as it is
  const double f = 1.0 / ((double)hi - (double)lo);
  img[v] = (float)(f * ((double)img[v] - (double)lo));
}

inline size_t up256(size_t n) { return (n + 255) & ~size_t(255); }
inline unsigned nblocks(long long n, int per) { return (unsigned)((n + per - 1) / per); }

// Per-device staging of the inputs (grow-only, held until the process
// ends); calls are serialised by one mutex.
struct MarshalWs {
  void* dev = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (dev) (void)hipFree(dev);
    dev = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&dev, bytes);
    if (e == hipSuccess) cap = bytes;
    return e;
  }
};
std::mutex g_mu;
std::vector<MarshalWs> g_ws;

// The allocation-free variant's scratch: a file-scope device array (each
// device holds its own copy until the process ends), whose device address is
// looked up once per device with that device current. Calls are serialised
// by g_mu; every call overwrites what it reads (the supervoxel pointer, the
// block partials, the range; the error flag is cleared by a kernel).
constexpr size_t kScratchBytes = size_t(16) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
std::vector<void*> g_scratch;

int hip_fail(hipError_t e, const char* what) {
  tgmm_set_error("marshal_frame: %s failed: %s", what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? TGMM_ERR_OOM : TGMM_ERR_DEVICE;
}

int check_frame(const tgmm_raw_frame* f) {
  if (!f || f->N < 0 || f->S < 0 || !f->sv_ptr || (f->S > 0 && !f->thr) ||
      (f->N > 0 && (!f->pixel_idx || (!f->raw && !f->volume)))) {
    tgmm_set_error("marshal_frame: NULL argument or negative size");
    return TGMM_ERR_ARG;
  }
  for (int a = 0; a < 3; a++)
    if (f->dims[a] <= 0) {
      tgmm_set_error("marshal_frame: dims[%d] = %lld", a, f->dims[a]);
      return TGMM_ERR_ARG;
    }
  const double vol = (double)f->dims[0] * (double)f->dims[1] * (double)f->dims[2];
  if (vol >= 9007199254740992.0) {  // 2^53
    tgmm_set_error("marshal_frame: volume of %.0f voxels (at most 2^53)", vol);
    return TGMM_ERR_LIMIT;
  }
  if (f->sv_ptr[0] != 0 || f->sv_ptr[f->S] != f->N) {
    tgmm_set_error("marshal_frame: sv_ptr must run from 0 to N = %lld", f->N);
    return TGMM_ERR_ARG;
  }
  for (long long s = 0; s < f->S; s++)
    if (f->sv_ptr[s + 1] < f->sv_ptr[s]) {
      tgmm_set_error("marshal_frame: sv_ptr decreases at supervoxel %lld", s);
      return TGMM_ERR_ARG;
    }
  if ((double)f->N > vol) {
    tgmm_set_error("marshal_frame: %lld voxels listed in a volume of %.0f: supervoxels must be "
                   "disjoint", f->N, vol);
    return TGMM_ERR_ARG;
  }
  return TGMM_OK;
}

// The marshalling on the default stream, from device inputs (idx, raw,
// thr), with the device buffers it needs besides the outputs: the supervoxel
// pointer (uploaded here from f->sv_ptr), the range and error flag, and the
// block partials of the range.
int marshal_run(const tgmm_raw_frame* f, const unsigned long long* idx, const unsigned short* raw,
                const unsigned short* thr, long long* ptr, unsigned* range, unsigned* part_hi,
                unsigned* part_lo, int nsb, float* queryCUDA, float* imgDataCUDA,
                float* centroidLabelPositionCUDA, long long* labelListPtrCUDA) {
  const long long N = f->N, S = f->S;
  hipError_t e;
  if ((e = hipMemcpyAsync(ptr, f->sv_ptr, (size_t)(S + 1) * 8, hipMemcpyHostToDevice, 0)) !=
      hipSuccess)
    return hip_fail(e, "upload of the supervoxel pointer");
  if (labelListPtrCUDA &&
      (e = hipMemcpyAsync(labelListPtrCUDA, ptr, (size_t)(S + 1) * 8, hipMemcpyDeviceToDevice,
                          0)) != hipSuccess)
    return hip_fail(e, "copy of the supervoxel pointer");
  k_mar_init<<<1, 64, 0, 0>>>(range);
  if (S > 0)
    k_mar_sv<<<nsb, kMB, 0, 0>>>(
        idx, raw, f->volume, ptr, thr, S, N, (unsigned long long)f->dims[0],
        (unsigned long long)f->dims[1],
        (unsigned long long)f->dims[0] * (unsigned long long)f->dims[1] *
            (unsigned long long)f->dims[2],
        1.0 / (double)f->dims[0], 1.0 / (double)f->dims[1], queryCUDA, imgDataCUDA,
        centroidLabelPositionCUDA, part_hi, part_lo, range);
  if (N > 0) k_mar_range<<<1, 1024, 0, 0>>>(part_hi, part_lo, nsb, range);
  const double vol = (double)f->dims[0] * (double)f->dims[1] * (double)f->dims[2];
  if (N > 0)
    k_mar_scale<<<nblocks(N, kMB), kMB, 0, 0>>>(imgDataCUDA, N, range, (double)N < vol ? 1 : 0);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "kernel launch");
  unsigned hrange[3];
  if ((e = hipMemcpy(hrange, range, sizeof(hrange), hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_fail(e, "read-back of the range");
  if (hrange[2]) {
    tgmm_set_error("marshal_frame: a voxel index lies outside the %lld x %lld x %lld volume",
                   f->dims[0], f->dims[1], f->dims[2]);
    return TGMM_ERR_ARG;
  }
  return TGMM_OK;
}

}  // namespace

extern "C" int tgmm_marshal_frame(const tgmm_raw_frame* f, float* queryCUDA, float* imgDataCUDA,
                                  float* centroidLabelPositionCUDA, long long* labelListPtrCUDA) {
  int rc = check_frame(f);
  if (rc) return rc;
  const long long N = f->N, S = f->S;
  if ((N > 0 && (!queryCUDA || !imgDataCUDA)) || (S > 0 && !centroidLabelPositionCUDA)) {
    tgmm_set_error("marshal_frame: NULL output");
    return TGMM_ERR_ARG;
  }
  const bool dev_in = f->device_inputs != 0;
  const size_t o_idx = 0;
  const size_t o_raw = o_idx + up256(dev_in ? 0 : (size_t)N * 8);
  const size_t o_ptr = o_raw + up256(f->raw && !dev_in ? (size_t)N * 2 : 0);
  const size_t o_thr = o_ptr + up256((size_t)(S + 1) * 8);
  const size_t o_rng = o_thr + up256(dev_in ? 0 : (size_t)S * 2);
  const int nsb = S > 0 ? (int)nblocks(S, kMB / kMG) : 0;
  const size_t o_part = o_rng + 256;
  const size_t total = o_part + 2 * up256((size_t)nsb * sizeof(unsigned));
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_ws.size() <= dev) g_ws.resize(dev + 1);
  MarshalWs& ws = g_ws[dev];
  if ((e = ws.reserve(total)) != hipSuccess) return hip_fail(e, "workspace allocation");
  char* base = (char*)ws.dev;
  // device inputs are read in place; host inputs are uploaded
  const unsigned long long* idx = f->pixel_idx;
  const unsigned short* raw = f->raw;
  const unsigned short* thr = f->thr;
  auto* ptr = (long long*)(base + o_ptr);
  auto* range = (unsigned*)(base + o_rng);
  auto* part_hi = (unsigned*)(base + o_part);
  auto* part_lo = part_hi + up256((size_t)nsb * sizeof(unsigned)) / sizeof(unsigned);
  if (!dev_in) {
    auto* idx_d = (unsigned long long*)(base + o_idx);
    auto* raw_d = f->raw ? (unsigned short*)(base + o_raw) : nullptr;
    auto* thr_d = (unsigned short*)(base + o_thr);
    if (N > 0 && (e = hipMemcpyAsync(idx_d, f->pixel_idx, (size_t)N * 8, hipMemcpyHostToDevice,
                                     0)) != hipSuccess)
      return hip_fail(e, "upload of the voxel indices");
    if (raw_d && (e = hipMemcpyAsync(raw_d, f->raw, (size_t)N * 2, hipMemcpyHostToDevice, 0)) !=
                     hipSuccess)
      return hip_fail(e, "upload of the intensities");
    if (S > 0 && (e = hipMemcpyAsync(thr_d, f->thr, (size_t)S * 2, hipMemcpyHostToDevice, 0)) !=
                     hipSuccess)
      return hip_fail(e, "upload of the thresholds");
    idx = idx_d;
    raw = raw_d;
    thr = thr_d;
  }
  return marshal_run(f, idx, raw, thr, ptr, range, part_hi, part_lo, nsb, queryCUDA, imgDataCUDA,
                     centroidLabelPositionCUDA, labelListPtrCUDA);
}

// tgmm_marshal_frame with device inputs, allocating nothing: the supervoxel
// pointer, the range and the block partials live in the scratch array.
extern "C" int tgmm_marshal_frame_noalloc(const tgmm_raw_frame* f, float* queryCUDA,
                                          float* imgDataCUDA, float* centroidLabelPositionCUDA,
                                          long long* labelListPtrCUDA) {
  int rc = check_frame(f);
  if (rc) return rc;
  const long long N = f->N, S = f->S;
  if ((N > 0 && (!queryCUDA || !imgDataCUDA)) || (S > 0 && !centroidLabelPositionCUDA)) {
    tgmm_set_error("marshal_frame_noalloc: NULL output");
    return TGMM_ERR_ARG;
  }
  if (!f->device_inputs) {
    tgmm_set_error("marshal_frame_noalloc: the inputs must be device memory (device_inputs = 1); "
                   "host inputs need staging that this entry point does not allocate");
    return TGMM_ERR_ARG;
  }
  const int nsb = S > 0 ? (int)nblocks(S, kMB / kMG) : 0;
  const size_t o_ptr = 0;
  const size_t o_rng = o_ptr + up256((size_t)(S + 1) * 8);
  const size_t o_part = o_rng + 256;
  const size_t part_b = up256((size_t)nsb * sizeof(unsigned));
  const size_t total = o_part + 2 * part_b;
  if (total > kScratchBytes) {
    tgmm_set_error("marshal_frame_noalloc: %lld supervoxels need %zu bytes of scratch; the "
                   "allocation-free entry point holds %zu (about 1.97 million supervoxels)",
                   S, total, kScratchBytes);
    return TGMM_ERR_LIMIT;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1, nullptr);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess)
      return hip_fail(e, "hipGetSymbolAddress of the scratch array");
    g_scratch[dev] = p;
  }
  char* base = (char*)g_scratch[dev];
  auto* ptr = (long long*)(base + o_ptr);
  auto* range = (unsigned*)(base + o_rng);
  auto* part_hi = (unsigned*)(base + o_part);
  auto* part_lo = (unsigned*)(base + o_part + part_b);
  return marshal_run(f, f->pixel_idx, f->raw, f->thr, ptr, range, part_hi, part_lo, nsb,
                     queryCUDA, imgDataCUDA, centroidLabelPositionCUDA, labelListPtrCUDA);
}
