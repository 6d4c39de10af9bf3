// tgmm_kernels.hip -- gfx950 kernels of the supervoxel variational-inference
// EM (one EM iteration = launch_iteration). Wave64 throughout; no MFMA (the
// 3x3 forms are not dense contractions); no floating-point atomics, so every
// result is independent of scheduling and of which other frames share the
// launch.
//
// Once per frame (launch_prep): per-supervoxel FP64 moments, the frame's
// uniform kNN grid, and the supervoxel order sorted by grid cell.
// Per iteration, over all frames of a workspace:
//   gauss_prep     E[ln det Lambda], E[ln pi] (GMEM.cu:553-628), E-step
//                  constant, f32 means, kNN grid binning
//   scan/scatter   counting sort of the means into per-frame grid cells
//   knn            exact K nearest means per supervoxel centroid by ring
//                  search (knnCuda.cu:61-144 distance, (dist, idx) order)
//   estep          responsibilities from per-supervoxel moments
//                  (GMEM.cu:2244-2336), significant-pair counts
//   scan/scatter   pairs (s, j) with r > 1e-3 grouped per Gaussian
//   seg_sort       each Gaussian's pair list ordered by supervoxel
//   mstep_update   FP64 segmented sufficient statistics (GMEM.cu:1286-1455),
//                  posterior (1601-1677), regularisation (1693-1809),
//                  per-frame sum of alpha
//   death          dead-cell rule (1813-1843), log-likelihood parameters,
//                  per-frame sum of alpha
//   loglik         voxel stream: img * log sum_k pi_k N(x; m_k, (nu W)^-1)
//                  (1913-1978), per-frame total, convergence (4145-4148)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <climits>

#include "tgmm_device.h"
#include "tgmm_internal.h"

namespace tgmm {

// ---------------------------------------------------------------- helpers --
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Sum over a block of NW waves in a fixed order; result valid in thread 0.
template <int NW>
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  if (NW == 1) return v;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NW; i++) t += red[i];
  return t;
}

// Sum of a frame's per-block partials by one wave, in an order fixed by the
// partial index alone (lane l takes partials l, l+64, ... into four
// interleaved sums, then a fixed butterfly). Every lane returns the total.
// The partials were written by an earlier kernel, or by the calling block
// before a barrier.
__device__ inline double frame_sum_w64(const double* part, int nblk) {
  double a4[4] = {0.0, 0.0, 0.0, 0.0};
  const int l = threadIdx.x;
  int i = l;
  for (; i + 960 < nblk; i += 1024) {  // 16 loads in flight, the same order
    double v[16];
#pragma unroll
    for (int t = 0; t < 16; t++) v[t] = part[i + 64 * t];
#pragma unroll
    for (int t = 0; t < 16; t++) a4[t & 3] += v[t];
  }
  for (; i + 192 < nblk; i += 256) {
#pragma unroll
    for (int q = 0; q < 4; q++) a4[q] += part[i + 64 * q];
  }
  for (int q = 0; i < nblk; i += 64, q++) a4[q & 3] += part[i];
  return wave_sum((a4[0] + a4[1]) + (a4[2] + a4[3]));
}

__device__ inline double gsv(const DevPtrs& P, int f, long long k) { return P.gs[f * P.Gcap + k]; }

// reference loop condition (GMEM.cu:4145-4148)
__device__ inline int keep_going(double ll, double l1, double l2, double l3, double tol) {
  return (fabs(ll - l1) / fabs(l1) > tol) && (fabs(ll - l2) / fabs(l2) > tol) &&
         (fabs(ll - l3) / fabs(l3) > tol);
}

// FP32 kNN distance in the reference's order (knnCuda.cu:116-124), never
// contracted into FMAs so that host and device agree bit for bit.
__device__ inline float knn_dist(float q0, float q1, float q2, float m0, float m1, float m2,
                                 float s0, float s1, float s2) {
#pragma clang fp contract(off)
  const float a = (q0 - m0) * s0;
  float d = a * a;
  const float b = (q1 - m1) * s1;
  d = d + b * b;
  const float e = (q2 - m2) * s2;
  d = d + e * e;
  return d;
}

__device__ inline bool lex_less(float da, int ia, float db, int ib) {
  return da < db || (da == db && ia < ib);
}

// Grid cell (per axis) of scaled coordinate u; NaN and values below the
// origin go to 0, values beyond the last cell to the last cell.
__device__ inline int axis_cell(double u, float org, float h, int dim) {
  const double t = (u - (double)org) / (double)h;
  if (t >= (double)dim) return dim - 1;
  if (t > 0.0) return (int)floor(t);
  return 0;
}

// ------------------------------------------------------------ prep kernels --
// 16 lanes per supervoxel: unweighted moments (E-step) and intensity-weighted
// moments (M-step), FP64, centred about the supervoxel's own means (two
// passes over its voxels; sums over the group by a 4-step butterfly).
__device__ inline double group16_sum(double x) {
#pragma unroll
  for (int m = 8; m >= 1; m >>= 1) x += __shfl_xor(x, m, 16);
  return x;
}
__device__ inline float group16_min(float x) {
#pragma unroll
  for (int m = 8; m >= 1; m >>= 1) x = fminf(x, __shfl_xor(x, m, 16));
  return x;
}
__device__ inline float group16_max(float x) {
#pragma unroll
  for (int m = 8; m >= 1; m >>= 1) x = fmaxf(x, __shfl_xor(x, m, 16));
  return x;
}

__global__ __launch_bounds__(256) void k_moments(DevPtrs P) {
  // every exit below is uniform over a supervoxel's 16 lanes, which is all
  // the group sums need
  const long long s = P.s_lo + (long long)blockIdx.x * 16 + (threadIdx.x >> 4);
  const int lane = threadIdx.x & 15;
  if (s >= P.Scap) return;
  const int f = P.sv_frame[s];
  if (f < 0 || (P.only_frame >= 0 && f != P.only_frame)) return;
  const long long beg = P.sv_beg[s];
  const int n = P.sv_len[s];
  const long long S = P.Scap;
  if (n <= 0) {
    if (lane < 10) {
      P.mom[lane * S + s] = 0.0;
      P.wmom[s * 10 + lane] = 0.0;
    }
    if (lane < 3) P.anchor[lane * S + s] = 0.0f;
    if (lane == 0) P.sv_ball[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const float a0 = P.vx[beg], a1 = P.vy[beg], a2 = P.vz[beg];
  double sx = 0, sy = 0, sz = 0, sw = 0, swx = 0, swy = 0, swz = 0;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool fits = true;  // every offset from the anchor is an integer in [-127, 127]
  for (int i = lane; i < n; i += 16) {
    const float v[3] = {P.vx[beg + i], P.vy[beg + i], P.vz[beg + i]};
    for (int a = 0; a < 3; a++) {
      lo[a] = fminf(lo[a], v[a]);
      hi[a] = fmaxf(hi[a], v[a]);
    }
    {
#pragma clang fp contract(off)
      // the log-likelihood's f32 offsets, exactly
      const float d0 = v[0] - a0, d1 = v[1] - a1, d2 = v[2] - a2;
      const bool ok = fabsf(d0) <= 127.0f && fabsf(d1) <= 127.0f && fabsf(d2) <= 127.0f &&
                      d0 == truncf(d0) && d1 == truncf(d1) && d2 == truncf(d2);
      fits = fits && ok;
      // byte 3: the supervoxel's position in its 256-aligned group
      const unsigned pk = (ok ? ((unsigned)(int)d0 & 0xffu) | (((unsigned)(int)d1 & 0xffu) << 8) |
                                    (((unsigned)(int)d2 & 0xffu) << 16)
                              : 0u) |
                          ((unsigned)(s & 255) << 24);
      P.vpk[beg + i] = make_int2((int)pk, __float_as_int(P.vimg[beg + i]));
    }
    const double x = (double)v[0] - a0, y = (double)v[1] - a1, z = (double)v[2] - a2,
                 w = (double)P.vimg[beg + i];
    sx += x;
    sy += y;
    sz += z;
    sw += w;
    swx += w * x;
    swy += w * y;
    swz += w * z;
  }
  sx = group16_sum(sx);
  sy = group16_sum(sy);
  sz = group16_sum(sz);
  sw = group16_sum(sw);
  swx = group16_sum(swx);
  swy = group16_sum(swy);
  swz = group16_sum(swz);
  const double cx = sx / n, cy = sy / n, cz = sz / n;
  const double ux = sw > 0 ? swx / sw : cx, uy = sw > 0 ? swy / sw : cy,
               uz = sw > 0 ? swz / sw : cz;
  double q[6] = {0, 0, 0, 0, 0, 0}, c[6] = {0, 0, 0, 0, 0, 0};
  for (int i = lane; i < n; i += 16) {
    const double x = (double)P.vx[beg + i] - a0, y = (double)P.vy[beg + i] - a1,
                 z = (double)P.vz[beg + i] - a2, w = (double)P.vimg[beg + i];
    const double d0 = x - cx, d1 = y - cy, d2 = z - cz;
    q[0] += d0 * d0;
    q[1] += d0 * d1;
    q[2] += d0 * d2;
    q[3] += d1 * d1;
    q[4] += d1 * d2;
    q[5] += d2 * d2;
    const double e0 = x - ux, e1 = y - uy, e2 = z - uz;
    c[0] += w * e0 * e0;
    c[1] += w * e0 * e1;
    c[2] += w * e0 * e2;
    c[3] += w * e1 * e1;
    c[4] += w * e1 * e2;
    c[5] += w * e2 * e2;
  }
#pragma unroll
  for (int i = 0; i < 6; i++) {
    q[i] = group16_sum(q[i]);
    c[i] = group16_sum(c[i]);
  }
  if (lane == 0) {
    P.mom[0 * S + s] = (double)n;
    P.mom[1 * S + s] = (double)a0 + cx;
    P.mom[2 * S + s] = (double)a1 + cy;
    P.mom[3 * S + s] = (double)a2 + cz;
    for (int i = 0; i < 6; i++) P.mom[(4 + i) * S + s] = q[i];
    double* wm = P.wmom + s * 10;
    wm[0] = sw;
    wm[1] = (double)a0 + ux;
    wm[2] = (double)a1 + uy;
    wm[3] = (double)a2 + uz;
    for (int i = 0; i < 6; i++) wm[4 + i] = c[i];
    P.anchor[0 * S + s] = a0;
    P.anchor[1 * S + s] = a1;
    P.anchor[2 * S + s] = a2;
  }
  if (!fits) atomicAnd(&P.fr[f].packed, 0);
  // the ball around the voxels' box (NaN coordinates give a NaN ball, which
  // the log-likelihood's candidate bounds ignore)
  const float a[3] = {a0, a1, a2};
  double bc[3], r2 = 0.0;
  for (int i = 0; i < 3; i++) {
    const double l = (double)group16_min(lo[i]), h = (double)group16_max(hi[i]);
    bc[i] = 0.5 * (l + h) - (double)a[i];
    r2 += 0.25 * (h - l) * (h - l);
  }
  if (lane == 0)
    P.sv_ball[s] = make_float4((float)bc[0], (float)bc[1], (float)bc[2],
                               (float)(sqrt(r2) * (1.0 + 1e-6) + 1e-6 * (fabs(bc[0]) +
                                       fabs(bc[1]) + fabs(bc[2]))));
}

// Voxel mode: every supervoxel is one voxel, so its moments are the voxel
// itself (count 1, mean x, no scatter) and its M-step weight is the stored
// responsibility, which already carries the intensity (GMEM.cu:2216-2222).
__global__ __launch_bounds__(256) void k_moments_voxel(DevPtrs P) {
  const long long s = P.s_lo + (long long)blockIdx.x * 256 + threadIdx.x;
  if (s >= P.Scap) return;
  const int f = P.sv_frame[s];
  if (f < 0 || (P.only_frame >= 0 && f != P.only_frame)) return;
  const long long S = P.Scap, v = P.sv_beg[s];
  const float x[3] = {P.vx[v], P.vy[v], P.vz[v]};
  P.mom[s] = 1.0;
  P.wmom[s * 10] = 1.0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    P.mom[(1 + a) * S + s] = (double)x[a];
    P.wmom[s * 10 + 1 + a] = (double)x[a];
    P.anchor[a * S + s] = x[a];
  }
#pragma unroll
  for (int i = 4; i < 10; i++) {
    P.mom[i * S + s] = 0.0;
    P.wmom[s * 10 + i] = 0.0;
  }
  P.sv_ball[s] = make_float4(0.f, 0.f, 0.f, 0.f);
  // the voxel is its own anchor: offsets 0, unless a coordinate is not finite
  P.vpk[v] = make_int2((int)((unsigned)(s & 255) << 24), __float_as_int(P.vimg[v]));
  if (!(x[0] - x[0] == 0.0f && x[1] - x[1] == 0.0f && x[2] - x[2] == 0.0f))
    atomicAnd(&P.fr[f].packed, 0);
}

// One block per frame: bounding box of the scaled kNN centroids and the grid
// shape (about knn_cell means per cell, at most ncap cells).
__global__ __launch_bounds__(1024) void k_grid_setup(DevPtrs P, KOpts O) {
  const int f = (P.only_frame >= 0) ? P.only_frame : (int)blockIdx.x;
  if (f >= P.Fcap || (P.only_frame < 0 && f >= P.nframes)) return;
  FrameDev* F = P.fr + f;
  const long long s0 = F->s0, ns = F->ns;
  const long long S = P.Scap;
  __shared__ double red[6][1024];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (long long i = threadIdx.x; i < ns; i += 1024) {
    for (int a = 0; a < 3; a++) {
      const double u = (double)P.sv_cent[a * S + s0 + i] * (double)O.scale[a];
      lo[a] = fmin(lo[a], u);
      hi[a] = fmax(hi[a], u);
    }
  }
  for (int a = 0; a < 3; a++) {
    red[a][threadIdx.x] = lo[a];
    red[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int a = 0; a < 3; a++) {
        red[a][threadIdx.x] = fmin(red[a][threadIdx.x], red[a][threadIdx.x + st]);
        red[3 + a][threadIdx.x] = fmax(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double ext[3], org[3];
  for (int a = 0; a < 3; a++) {
    const bool ok = ns > 0 && isfinite(red[a][0]) && isfinite(red[3 + a][0]);
    org[a] = ok ? red[a][0] : 0.0;
    ext[a] = ok ? fmax(red[3 + a][0] - red[a][0], 0.0) : 0.0;
  }
  const double target = fmax(1.0, ceil((double)F->ng / (double)O.tiles.knn_cell));
  const double vol = fmax(ext[0], 1e-3) * fmax(ext[1], 1e-3) * fmax(ext[2], 1e-3);
  double h = cbrt(vol / target);
  if (!(h > 0.0) || !(h < 1e30)) h = 1.0;
  // the grid stores f32 origin and edge; cell of u is floor((u - org) / h)
  float orgf[3];
  for (int a = 0; a < 3; a++) {
    orgf[a] = (float)org[a];
    if ((double)orgf[a] > org[a]) orgf[a] = nextafterf(orgf[a], -INFINITY);
  }
  long long dims[3] = {1, 1, 1};
  float hf = (float)h;
  for (int guard = 0; guard < 400; guard++) {
    hf = (float)h;
    for (int a = 0; a < 3; a++) {
      const double top = org[a] + ext[a];
      long long d = (long long)floor((top - (double)orgf[a]) / (double)hf) + 1;
      if (d < 1) d = 1;
      // the last cell must reach past the largest coordinate
      while ((double)orgf[a] + (double)d * (double)hf <= top) d++;
      dims[a] = d;
    }
    if (dims[0] * dims[1] * dims[2] < F->ncap) break;
    h *= 1.25;
  }
  for (int a = 0; a < 3; a++) {
    F->org[a] = orgf[a];
    F->dims[a] = (int)dims[a];
  }
  F->h = hf;
  if (dims[0] * dims[1] * dims[2] >= F->ncap) {  // cannot happen for finite input
    F->dims[0] = F->dims[1] = F->dims[2] = 1;
    F->h = 3.0e38f;
  }
}

// Supervoxels sorted by the grid cell of their centroid (once per frame), so
// that neighbouring lanes of the kNN kernel walk the same cells.
__global__ __launch_bounds__(256) void k_qcell_count(DevPtrs P, KOpts O) {
  const long long sb = P.s_lo / 256 + blockIdx.x;
  const int f = P.sblk_frame[sb];
  if (f < 0 || (P.only_frame >= 0 && f != P.only_frame)) return;
  const long long s = sb * 256 + threadIdx.x;
  if (P.sv_frame[s] != f) return;
  const FrameDev* F = P.fr + f;
  const long long S = P.Scap;
  int c[3];
  for (int a = 0; a < 3; a++)
    c[a] = axis_cell((double)P.sv_cent[a * S + s] * (double)O.scale[a], F->org[a], F->h,
                     F->dims[a]);
  const long long cell = F->c0 + ((long long)c[2] * F->dims[1] + c[1]) * F->dims[0] + c[0];
  P.qcell[s] = (int)cell;
  atomicAdd(P.cell_cnt + cell, 1);
}

__global__ __launch_bounds__(256) void k_qcell_scatter(DevPtrs P) {
  const long long sb = P.s_lo / 256 + blockIdx.x;
  const int f = P.sblk_frame[sb];
  if (f < 0 || (P.only_frame >= 0 && f != P.only_frame)) return;
  const long long s = sb * 256 + threadIdx.x;
  if (P.sv_frame[s] != f) return;
  const FrameDev* F = P.fr + f;
  const int c = P.qcell[s];
  const long long pos = P.qcell_start[c] + atomicSub(P.cell_cnt + c, 1) - 1;
  const long long S = P.Scap;
  P.q4[pos] = make_float4(P.sv_cent[s], P.sv_cent[S + s], P.sv_cent[2 * S + s],
                          __int_as_float((int)s));
}

// Iteration state of every frame (reference: ll = -1, -10, -100, -1000 and
// numIterEM = 1 before the loop, GMEM.cu:4143-4144).
__global__ void k_iter_init(DevPtrs P) {
  const int f = (P.only_frame >= 0) ? (threadIdx.x == 0 && blockIdx.x == 0 ? P.only_frame : -1)
                                    : (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (f < 0 || f >= P.Fcap || (P.only_frame < 0 && f >= P.nframes)) return;
  FrameDev* F = P.fr + f;
  F->ll = -1.0;
  F->ll1 = -10.0;
  F->ll2 = -100.0;
  F->ll3 = -1000.0;
  F->iter = 1;
  F->active = (F->ns > 0 && F->ng > 0 && F->max_iter >= 1 &&
               keep_going(F->ll, F->ll1, F->ll2, F->ll3, F->tol))
                  ? 1
                  : 0;
  F->status = 0;
}

// Sum of alpha per frame before the first iteration (GMEM.cu:4167).
__global__ __launch_bounds__(kGB) void k_alpha_init(DevPtrs P) {
  __shared__ double red[1];
  const long long gb = P.g_lo / kGB + blockIdx.x;
  const int f = P.gblk_frame[gb];
  if (f < 0 || (P.only_frame >= 0 && f != P.only_frame)) return;
  FrameDev* F = P.fr + f;
  if (!F->active) return;
  const long long k = gb * kGB + threadIdx.x;
  const double a = (P.g_frame[k] == f) ? gsv(P, GS_ALPHA, k) : 0.0;
  const double t = block_sum<1>(a, red);
  if (threadIdx.x == 0) P.part_g[gb] = t;
}

// ---------------------------------------------------------------- scans --
constexpr int kSortShort = 128;  // longer pair lists are sorted by a block each
constexpr int kScanQ = 8;        // int4 loads per thread and pass, k_scan_frames (4 | kScanQ)
constexpr int kSortLds = 4096;   // longest list sorted in LDS

// Exclusive scan of each frame's counts, one 1024-thread block per frame:
// out[off + i] = base + sum(in[off .. off + i)) for i = 0..n, with
//   which 0: the frame's kNN grid cells (off c0, n cells, base g0: the
//            references of a frame's grid go to its Gaussian range),
//   which 1: the frame's Gaussians (off g0, n ng, base K*s0: a frame has at
//            most K*ns significant pairs),
//   which 2: the frame's grid cells, counts of kNN queries (base s0: the
//            positions of the frame's queries in cell order).
// Both ranges hold n+1 entries (the grid has fewer cells than its range, the
// Gaussian range at least one padding entry), so frames never share an entry.
// A pass covers kScanQ segments of 4096 entries; in segment q thread t loads
// and stores entries 4t .. 4t+3 as one int4 (consecutive threads, consecutive
// 16 bytes: a wave's load touches 8 cache lines, where a contiguous run per
// thread touched 64), and the ranges start on multiples of 4. The prefix of
// an int4 is the carry, the sums of the pass's earlier segments, and the sums
// of the segment's earlier threads: one wave scan per segment, the
// (segment, wave) totals scanned by wave 0.
__global__ __launch_bounds__(1024) void k_scan_frames(DevPtrs P, int which, int need_active,
                                                      int K) {
  constexpr int NW = 16;  // waves per block
  static_assert(kScanQ % 4 == 0, "whole values per lane in the totals scan");
  __shared__ int wsum[kScanQ * NW];
  __shared__ int total;
  // the grid scan opens an iteration: the kNN's tie list starts empty
  if (which == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    P.knn_ctl[0] = 0;
    P.knn_ctl[2] = 0;  // and the list of long pair lists
  }
  const int f = (P.only_frame >= 0) ? P.only_frame : (int)blockIdx.x;
  if (f >= P.Fcap || (P.only_frame < 0 && f >= P.nframes)) return;
  const FrameDev* F = P.fr + f;
  if (F->ncap == 0 || (need_active && !F->active)) return;  // empty slot / finished
  const int* in;
  int* out;
  long long n;
  int base;
  if (which == 0 || which == 2) {
    in = P.cell_cnt + F->c0;
    out = (which == 0 ? P.cell_start : P.qcell_start) + F->c0;
    n = (long long)F->dims[0] * F->dims[1] * F->dims[2];
    base = (int)(which == 0 ? F->g0 : F->s0);
  } else {
    in = P.seg_cnt + F->g0;
    out = P.seg_start + F->g0;
    n = F->ng;
    base = (int)(K * F->s0);
  }
  const long long n1 = n + 1;  // entry n carries the total
  // the pair scan lists the Gaussians whose pair list is long, for the
  // block-wide sort (k_seg_sort)
  const bool list_long = which == 1;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int carry = base;
  for (long long c0 = 0; c0 < n1; c0 += 4096LL * kScanQ) {
    int4 v[kScanQ];
    int x[kScanQ], sq[kScanQ];
#pragma unroll
    for (int q = 0; q < kScanQ; q++) {
      const long long e = c0 + 4096LL * q + 4LL * threadIdx.x;
      if (e + 4 <= n) {
        v[q] = *reinterpret_cast<const int4*>(in + e);
      } else {
        v[q].x = e < n ? in[e] : 0;
        v[q].y = e + 1 < n ? in[e + 1] : 0;
        v[q].z = e + 2 < n ? in[e + 2] : 0;
        v[q].w = e + 3 < n ? in[e + 3] : 0;
      }
    }
#pragma unroll
    for (int q = 0; q < kScanQ; q++) {
      sq[q] = v[q].x + v[q].y + v[q].z + v[q].w;
      x[q] = sq[q];
      if (list_long && (v[q].x > kSortShort || v[q].y > kSortShort || v[q].z > kSortShort ||
                        v[q].w > kSortShort)) {
        const long long e = c0 + 4096LL * q + 4LL * threadIdx.x;
        const int c[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
        for (int i = 0; i < 4; i++)
          if (c[i] > kSortShort) P.long_list[atomicAdd(P.knn_ctl + 2, 1)] = (int)(F->g0 + e + i);
      }
    }
    // inclusive wave scans, the segments side by side
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
      for (int q = 0; q < kScanQ; q++) {
        const int y = __shfl_up(x[q], o, 64);
        if (lane >= o) x[q] += y;
      }
    }
    if (lane == 63)
#pragma unroll
      for (int q = 0; q < kScanQ; q++) wsum[q * NW + w] = x[q];
    __syncthreads();
    if (w == 0) {
      // the kScanQ * NW totals in (segment, wave) order, PL per lane
      constexpr int PL = kScanQ * NW / 64;
      int a[PL], y = 0;
#pragma unroll
      for (int i = 0; i < PL; i++) {
        a[i] = wsum[PL * lane + i];
        y += a[i];
      }
      const int own = y;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(y, o, 64);
        if (lane >= o) y += t;
      }
      int ex = y - own;
#pragma unroll
      for (int i = 0; i < PL; i++) {
        wsum[PL * lane + i] = ex;
        ex += a[i];
      }
      if (lane == 63) total = y;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kScanQ; q++) {
      const long long e = c0 + 4096LL * q + 4LL * threadIdx.x;
      int4 r;
      r.x = carry + wsum[q * NW + w] + x[q] - sq[q];
      r.y = r.x + v[q].x;
      r.z = r.y + v[q].y;
      r.w = r.z + v[q].z;
      if (e + 4 <= n1) {
        *reinterpret_cast<int4*>(out + e) = r;
      } else {
        if (e < n1) out[e] = r.x;
        if (e + 1 < n1) out[e + 1] = r.y;
        if (e + 2 < n1) out[e + 2] = r.z;
      }
    }
    carry += total;
    __syncthreads();  // wsum and total are rewritten by the next pass
  }
}

// ------------------------------------------------------- iteration kernels --
// E-step constants and kNN grid binning of Gaussian k for the next iteration
// (GMEM.cu:553-628, 527-539), from its state (m, W, nu, beta, alpha). E[ln pi]
// needs the frame's sum of alpha after the death test, so the pack keeps
// digamma(alpha) and the E-step subtracts digamma(sum); E[ln det] and
// E[ln pi] reach the Gaussian's record when that iteration's M-step runs,
// as the reference computes them at the start of each iteration.
__device__ inline void gauss_prep_one(const DevPtrs& P, const KOpts& O, const FrameDev* F,
                                      long long k, bool alive, const double* m,
                                      const double* W, double nu, double beta, double alpha) {
  const long long G = P.Gcap;
  if (!alive) {
    P.eldn[k] = -1e32;
    P.gcell[k] = -1;
    P.m4[k] = make_float4(NAN, NAN, NAN, __int_as_float((int)k));  // no kNN distance
    return;
  }
  double eld = 3.0 * 0.69314718055994530942 + log(det3(W));
  for (int i = 0; i < 3; i++) eld += digamma(0.5 * (nu + 1.0 - i));
  P.eldn[k] = eld;
  EPack ep;
  for (int i = 0; i < 3; i++) ep.m[i] = m[i];
  for (int i = 0; i < 6; i++) ep.W[i] = W[i];
  ep.nu = nu;
  ep.kappa0 = 0.5 * eld - 1.5 / beta;
  ep.dpsi = digamma(alpha);
  P.epack[k] = ep;
  const float4 mf = make_float4((float)m[0], (float)m[1], (float)m[2], __int_as_float((int)k));
  P.m4[k] = mf;
  int c[3];
  c[0] = axis_cell((double)mf.x * (double)O.scale[0], F->org[0], F->h, F->dims[0]);
  c[1] = axis_cell((double)mf.y * (double)O.scale[1], F->org[1], F->h, F->dims[1]);
  c[2] = axis_cell((double)mf.z * (double)O.scale[2], F->org[2], F->h, F->dims[2]);
  const long long cell = F->c0 + ((long long)c[2] * F->dims[1] + c[1]) * F->dims[0] + c[0];
  P.gcell[k] = (int)cell;
  atomicAdd(P.cell_cnt + cell, 1);
}

// The first iteration's E-step constants and grid binning, from the initial
// state (later iterations get theirs from k_death).
__global__ __launch_bounds__(kGB) void k_gauss_prep(DevPtrs P, KOpts O) {
  const long long gb = P.g_lo / kGB + blockIdx.x;
  const int f = P.gblk_frame[gb];
  if (f < 0 || (P.only_frame >= 0 && f != P.only_frame)) return;
  const FrameDev* F = P.fr + f;
  if (!F->active) return;
  const long long k = gb * kGB + threadIdx.x;
  if (P.g_frame[k] != f) return;
  double m[3], W[6];
  for (int i = 0; i < 3; i++) m[i] = gsv(P, GS_M + i, k);
  for (int i = 0; i < 6; i++) W[i] = gsv(P, GS_W + i, k);
  gauss_prep_one(P, O, F, k, gsv(P, GS_M_O, k) > -1e31, m, W, gsv(P, GS_NU, k),
                 gsv(P, GS_BETA, k), gsv(P, GS_ALPHA, k));
}

// Counting-sort scatter of the means into their grid cells; every mean of an
// active frame (NaN when dead) is also kept in index order as the means of
// this iteration's kNN search (m4 is rewritten for the next iteration before
// the search's slot layout is replayed after the last one).
__global__ __launch_bounds__(kGB) void k_grid_scatter(DevPtrs P) {
  const int f = P.gblk_frame[blockIdx.x];
  if (f < 0 || !P.fr[f].active) return;
  const long long k = (long long)blockIdx.x * kGB + threadIdx.x;
  if (P.g_frame[k] != f) return;
  const float4 m = P.m4[k];
  P.mknn[k] = m;
  const int c = P.gcell[k];
  if (c < 0) return;
  const int pos = P.cell_start[c] + atomicSub(P.cell_cnt + c, 1) - 1;
  P.cg4[pos] = m;
}

// Inserts a candidate (distance d <= 1e32, index idx) into a list of the KM
// best kept sorted by (distance, index). `tie` records whether an element
// left out of the K best (the candidate itself, or the K-th it displaces) has
// the distance of the K-th after the insertion: then the reference's scan,
// whose slot replacement rule breaks such ties by position, may hold a
// different set than the lower-index rule, and the supervoxel is searched
// again in the reference's order. Every Gaussian at the final K-th distance
// is considered by the search (the termination and pruning bounds are
// strict), so a tie is never missed; an equality with a K-th distance that
// later falls flags a supervoxel needlessly, which changes nothing.
template <int KM>
__device__ inline void knn_try(float (&bd)[KM], int (&bi)[KM], float d, int idx, int K,
                               bool& tie, float& kd) {
  // kd is the list's K-th distance (infinite until K are held)
  if (!(d <= 1e32f)) return;
  if (!lex_less(d, idx, bd[KM - 1], bi[KM - 1])) {
    tie |= d == kd;
    return;
  }
  float cd = d;
  int ci = idx;
#pragma unroll
  for (int j = 0; j < KM; j++) {
    if (lex_less(cd, ci, bd[j], bi[j])) {
      const float td = bd[j];
      const int ti = bi[j];
      bd[j] = cd;
      bi[j] = ci;
      cd = td;
      ci = ti;
    }
    // cd is now the best element outside the K best
    if (j == K - 1) {
      tie |= (cd <= 1e32f) && cd == bd[j];
      kd = bd[j];
    }
  }
}

template <int KM, int QL>
__device__ inline void knn_merge(float (&bd)[KM], int (&bi)[KM], int K, bool& tie, float& kd) {
#pragma unroll
  for (int m = QL / 2; m >= 1; m >>= 1) {
    float od[KM];
    int oi[KM];
#pragma unroll
    for (int t = 0; t < KM; t++) {
      od[t] = __shfl_xor(bd[t], m, QL);
      oi[t] = __shfl_xor(bi[t], m, QL);
    }
#pragma unroll
    for (int t = 0; t < KM; t++) {
      bool dup = false;
#pragma unroll
      for (int j = 0; j < KM; j++) dup |= (bi[j] == oi[t]);
      if (!dup) knn_try<KM>(bd, bi, od[t], oi[t], K, tie, kd);
    }
  }
}

// Shells r >= 2 of the exact search around query cell c (scaled query u),
// until the K-th best is below the open-face bound of the visited block.
// consider(ref) may defer a reference; flush() (called by every lane of the
// wave after each shell) completes the deferred ones before the K-th best is
// read.
// Cells are pruned by the current K-th best: a row whose y/z boxes, or a cell
// whose box, lies farther than it (with the bound's relative margin) holds no
// reference that can enter the K best. QL lanes share a query (lane `sub`
// takes every QL-th reference of a range); their lists are merged after each
// shell.
template <int KM, int QL, typename Consider, typename Flush>
__device__ inline void knn_shells(const DevPtrs& P, const FrameDev* F, const int* c,
                                  const double* u, int K, int sub, float (&bd)[KM],
                                  int (&bi)[KM], bool& tie, float& kd, Consider consider,
                                  Flush flush) {
  const int nx = F->dims[0], ny = F->dims[1], nz = F->dims[2];
  const int dims[3] = {nx, ny, nz};
  const double h = F->h;
  const double org[3] = {F->org[0], F->org[1], F->org[2]};
  const long long c0 = F->c0;
  auto kth = [&]() -> double {
    float v = INFINITY;
#pragma unroll
    for (int t = 0; t < KM; t++)
      if (t == K - 1) v = bd[t];
    return (double)v;
  };
  auto lower_bound_sq = [&](int r) -> double {
    double lb = INFINITY;
    for (int a = 0; a < 3; a++) {
      if (c[a] - r > 0) lb = fmin(lb, u[a] - (org[a] + (double)(c[a] - r) * h));
      if (c[a] + r < dims[a] - 1) lb = fmin(lb, (org[a] + (double)(c[a] + r + 1) * h) - u[a]);
    }
    if (lb == INFINITY) return INFINITY;
    lb = fmax(lb, 0.0) * (1.0 - 1e-5);
    return lb * lb;
  };
  auto box_d2 = [&](int a, int ci) -> double {
    const double lo = org[a] + (double)ci * h, hi = org[a] + (double)(ci + 1) * h;
    const double d = fmax(fmax(lo - u[a], u[a] - hi), 0.0) * (1.0 - 1e-5);
    return d * d;
  };
  const int rmax = max(nx, max(ny, nz));
  for (int r = 1; r <= rmax; r++) {
    if (r > 1) {
      const double bound = kth();
      const int zlo = max(c[2] - r, 0), zhi = min(c[2] + r, nz - 1);
      const int ylo = max(c[1] - r, 0), yhi = min(c[1] + r, ny - 1);
      for (int z = zlo; z <= zhi; z++) {
        const double dz = box_d2(2, z);
        if (dz > bound) continue;
        const bool zface = (z == c[2] - r) || (z == c[2] + r);
        for (int y = ylo; y <= yhi; y++) {
          const double base = dz + box_d2(1, y);
          if (base > bound) continue;
          const bool face = zface || (y == c[1] - r) || (y == c[1] + r);
          const long long row = c0 + ((long long)z * ny + y) * nx;
#pragma unroll 1
          for (int g = 0; g < 2; g++) {
            int xa, xb;
            if (face) {
              if (g == 1) break;
              xa = max(c[0] - r, 0);
              xb = min(c[0] + r, nx - 1);
              while (xa < c[0] && base + box_d2(0, xa) > bound) xa++;
              while (xb > c[0] && base + box_d2(0, xb) > bound) xb--;
            } else {
              xa = xb = g == 0 ? c[0] - r : c[0] + r;
              if (xa < 0 || xa >= nx || base + box_d2(0, xa) > bound) continue;
            }
            const int pb = P.cell_start[row + xa], pe = P.cell_start[row + xb + 1];
            for (int p = pb + sub; p < pe; p += 4 * QL) {
              float4 v[4];
#pragma unroll
              for (int w = 0; w < 4; w++)
                v[w] = (p + w * QL < pe) ? P.cg4[p + w * QL] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
              for (int w = 0; w < 4; w++)
                if (p + w * QL < pe) consider(v[w]);
            }
          }
        }
      }
      flush();  // every lane of the wave takes part
      if (QL > 1) knn_merge<KM, QL>(bd, bi, K, tie, kd);
    }
    const double lb2 = lower_bound_sq(r);
    if (lb2 == INFINITY || kth() < lb2) break;  // every cell visited, or done
  }
}

// Exact K nearest means by expanding shells of grid cells. After shell r,
// every unvisited mean lies beyond an open face of the visited block, so its
// scaled distance is at least LB; the search stops once the K-th best
// distance is below LB^2 with a relative margin far above f32 rounding, so
// the result equals brute force (ties broken by the lower index). Queries are
// taken in grid-cell order (q4). QL lanes share a query: each scans every
// QL-th reference of the visited cells into its own sorted list, and the
// lists are merged by a butterfly after each shell (a reference seen by two
// lanes of a merged list is kept once), so every lane of the group holds the
// K nearest of all references seen.

// Candidates a lane queues before they are inserted, k_knn (dynamic LDS:
// kKnnQueue entries per thread).
constexpr int kKnnQueue = 20;

template <int KM, int KC, int QL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KM <= 8 ? 4 : 1))) void k_knn(DevPtrs P, KOpts O) {
  extern __shared__ float2 kq[];  // entry e of thread t at e * blockDim.x + t
  const int sub = threadIdx.x % QL;
  const long long qpb = blockDim.x / QL;  // queries per block
  const long long i = (long long)blockIdx.x * qpb + threadIdx.x / QL;
  const int f = P.sblk_frame[((long long)blockIdx.x * qpb) >> 8];
  if (f < 0 || !P.fr[f].active) return;
  const FrameDev* F = P.fr + f;
  if (i >= F->s0 + F->ns) return;  // the whole group of the query leaves
  const float4 qq = P.q4[i];
  const long long s = __float_as_int(qq.w);
  const long long S = P.Scap;
  const int K = O.K;
  const float s0 = O.scale[0], s1 = O.scale[1], s2 = O.scale[2];
  const float q0 = qq.x, q1 = qq.y, q2 = qq.z;
  const int nx = F->dims[0], ny = F->dims[1], nz = F->dims[2];
  const double u[3] = {(double)q0 * s0, (double)q1 * s1, (double)q2 * s2};
  const int dims[3] = {nx, ny, nz};
  int c[3];
  for (int a = 0; a < 3; a++) c[a] = axis_cell(u[a], F->org[a], F->h, dims[a]);
  float bd[KM];
  int bi[KM];
#pragma unroll
  for (int j = 0; j < KM; j++) {
    bd[j] = INFINITY;
    bi[j] = 0x7fffffff;
  }
  const long long c0 = F->c0;
  bool tie = false;
  float kd = INFINITY;
  // An upper bound of the K-th distance: the previous iteration's candidates
  // of this supervoxel, if all K are still living means of this frame (a
  // mean's distance is a real candidate distance; dead means are NaN).
  float kbound = INFINITY;
  {
    // the K indices, then the K means, each one round of loads (a slot past
    // K reads row 0, an index outside the frame reads mean 0; neither is
    // used), and the bound formed branch-free, so that no load waits for
    // the previous candidate's
    int kk[KM];
#pragma unroll
    for (int j = 0; j < KM; j++) kk[j] = P.ind[(long long)(j < K ? j : 0) * S + s];
    float4 mm[KM];
#pragma unroll
    for (int j = 0; j < KM; j++) {
      const bool in = kk[j] >= F->g0 && kk[j] < F->g0 + F->ng;
      mm[j] = P.m4[in ? kk[j] : 0];
    }
    float mx = -INFINITY;
    bool all = true;
#pragma unroll
    for (int j = 0; j < KM; j++) {
      const bool in = kk[j] >= F->g0 && kk[j] < F->g0 + F->ng;
      const float4 m = in ? mm[j] : make_float4(NAN, NAN, NAN, 0.f);
      const float d = knn_dist(q0, q1, q2, m.x, m.y, m.z, s0, s1, s2);
      all &= (j >= K) | (d <= 1e32f);
      mx = j < K ? fmaxf(mx, d) : mx;
    }
    if (all) kbound = mx;
  }
  // The 3x3x3 block's references are first filtered, each lane queueing in
  // LDS those within kbound and the list's K-th distance (a reference beyond
  // either cannot enter the K best nor tie the final K-th, which is at most
  // both), then inserted together: the wave then runs the insertion once
  // per queued entry of its longest queue rather than once per reference.
  float qthr = kbound;
  int nq = 0;
  auto enqueue = [&](const float4 cur) {
    const float d = knn_dist(q0, q1, q2, cur.x, cur.y, cur.z, s0, s1, s2);
    if (d <= qthr) {
      kq[nq * blockDim.x + threadIdx.x] = make_float2(d, cur.w);
      nq++;
    }
  };
  auto flush = [&]() {  // every lane of the wave takes part
    for (int e = 0; __any(e < nq); e++)
      if (e < nq) {
        const float2 cq = kq[e * blockDim.x + threadIdx.x];
        knn_try<KM>(bd, bi, cq.x, __float_as_int(cq.y), K, tie, kd);
      }
    nq = 0;
    qthr = fminf(kbound, kd);
  };
  // squared distance from the query to the box of cell index ci along axis a
  // (FP64 from the f32 grid values), shrunk by the same relative margin as
  // the shell bound; 0 inside. A mean binned into a cell (an edge cell holds
  // the means beyond the grid on its outer side) is at least that far.
  auto box_d2 = [&](int a, int ci) -> double {
    const double h = F->h, o = F->org[a];  // the frame's grid (scalar loads)
    const double lo = o + (double)ci * h, hi = o + (double)(ci + 1) * h;
    const double d = fmax(fmax(lo - u[a], u[a] - hi), 0.0) * (1.0 - 1e-5);
    return d * d;
  };
  // shells 0 and 1 together: the 9 rows of the 3x3x3 block, each trimmed to
  // the cells whose box is within the bound (a cell farther than the bound
  // holds no candidate that can beat the K-th). All row bounds are loaded
  // first.
  {
    double dx2[3], dy2[3], dz2[3];
#pragma unroll
    for (int o = 0; o < 3; o++) {
      dx2[o] = box_d2(0, c[0] + o - 1);
      dy2[o] = box_d2(1, c[1] + o - 1);
      dz2[o] = box_d2(2, c[2] + o - 1);
    }
    const double thr = (double)kbound;
    int pb[9], cum[10];
    cum[0] = 0;
#pragma unroll
    for (int t = 0; t < 9; t++) {
      const int z = c[2] + t / 3 - 1, y = c[1] + t % 3 - 1;
      const double base = dz2[t / 3] + dy2[t % 3];
      int xa = max(c[0] - 1, 0), xb = min(c[0] + 1, nx - 1);
      if (xa < c[0] && base + dx2[0] > thr) xa = c[0];
      if (xb > c[0] && base + dx2[2] > thr) xb = c[0];
      const bool ok = z >= 0 && z < nz && y >= 0 && y < ny && base + dx2[1] <= thr;
      const long long row = ok ? c0 + ((long long)z * ny + y) * nx : c0;
      const int b0 = ok ? P.cell_start[row + xa] : 0;
      const int e0 = ok ? P.cell_start[row + xb + 1] : 0;
      pb[t] = b0;
      cum[t + 1] = cum[t] + (e0 - b0);
    }
    // each z slab's 3 trimmed rows as one range of references: KC per lane
    // in flight across row boundaries, so a lane waits for memory once per
    // KC references rather than once per row, and finds a reference's row
    // among 3 (the K best do not depend on the order references are
    // considered in)
#pragma unroll
    for (int zs = 0; zs < 3; zs++) {
      const int t0 = 3 * zs;
      const int c1 = cum[t0 + 1] - cum[t0], c2 = cum[t0 + 2] - cum[t0];
      const int tot = cum[t0 + 3] - cum[t0];
      const int o0 = pb[t0], o1 = pb[t0 + 1] - c1, o2 = pb[t0 + 2] - c2;
      for (int p0 = sub; __any(p0 < tot); p0 += KC * QL) {
        float4 v[KC];
#pragma unroll
        for (int w = 0; w < KC; w++) {
          const int p = p0 + w * QL;
          v[w] = make_float4(0.f, 0.f, 0.f, 0.f);
          if (p < tot) v[w] = P.cg4[(p >= c2 ? o2 : (p >= c1 ? o1 : o0)) + p];
        }
#pragma unroll
        for (int w = 0; w < KC; w++)
          if (p0 + w * QL < tot) enqueue(v[w]);
        if (__any(nq > kKnnQueue - KC)) flush();
      }
    }
    flush();
  }
  if (QL > 1) knn_merge<KM, QL>(bd, bi, K, tie, kd);
  // the further shells queue as well; a full queue inserts directly (the
  // shells' loops differ between lanes, so the wave flushes only after each)
  auto defer = [&](const float4 cur) {
    const float d = knn_dist(q0, q1, q2, cur.x, cur.y, cur.z, s0, s1, s2);
    if (d <= qthr) {
      if (nq < kKnnQueue) {
        kq[nq * blockDim.x + threadIdx.x] = make_float2(d, cur.w);
        nq++;
      } else {
        knn_try<KM>(bd, bi, d, __float_as_int(cur.w), K, tie, kd);
      }
    }
  };
  knn_shells<KM, QL>(P, F, c, u, K, sub, bd, bi, tie, kd, defer, flush);
#pragma unroll
  for (int m = QL / 2; m >= 1; m >>= 1) tie |= __shfl_xor((int)tie, m, QL) != 0;
  if (sub == 0) {
#pragma unroll
    for (int j = 0; j < KM; j++)
      if (j < K) P.ind[(long long)j * S + s] = (bd[j] <= 1e32f) ? bi[j] : -1;
    if (tie) P.tie_list[atomicAdd(P.knn_ctl, 1)] = (int)s;
  }
}

// Exact K nearest means, one block per grid cell (the same search as k_knn,
// organised by cell): every query of the cell has the same 3x3x3 block, so
// the block stages the block's references in LDS once (kKnnCap at a time).
// Lane l of each of the NW waves takes query l of the cell; wave w scans the
// references w, w+NW, ... into its own sorted list (LDS reads are
// broadcasts), the NW lists of a query are merged by wave 0, and a query
// whose K-th best is not below the block's open-face bound continues with the
// further shells on its own, from HBM, as in k_knn.
constexpr int kKnnCap = 512;

template <int KM, int NW>
__global__ __launch_bounds__(64 * NW) void k_knn_cell(DevPtrs P, KOpts O) {
  __shared__ float4 cand[kKnnCap];
  __shared__ float ld[NW > 1 ? NW - 1 : 1][KM][64];
  __shared__ int li[NW > 1 ? NW - 1 : 1][KM][64];
  __shared__ int lt[NW > 1 ? NW - 1 : 1][64];
  const long long cc = blockIdx.x;  // global cell index
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int f = P.cblk_frame[cc / kCB];
  if (f < 0) return;
  const FrameDev* F = P.fr + f;
  if (!F->active) return;
  const int nx = F->dims[0], ny = F->dims[1], nz = F->dims[2];
  const long long lc = cc - F->c0;
  if (lc >= (long long)nx * ny * nz) return;
  const int qs = P.qcell_start[cc], qe = P.qcell_start[cc + 1];
  if (qs >= qe) return;
  const long long c0 = F->c0;
  const int c[3] = {(int)(lc % nx), (int)((lc / nx) % ny), (int)(lc / ((long long)nx * ny))};
  // the 9 rows of the block: lanes 0..8 read the row starts, 9..17 the ends
  const int xlo = max(c[0] - 1, 0), xhi = min(c[0] + 1, nx - 1);
  int rv = 0;
  if (lane < 18) {
    const int t = lane % 9;
    const int z = c[2] + t / 3 - 1, y = c[1] + t % 3 - 1;
    if (z >= 0 && z < nz && y >= 0 && y < ny)
      rv = P.cell_start[c0 + ((long long)z * ny + y) * nx + (lane < 9 ? xlo : xhi + 1)];
  }
  int pb[9], cum[10];
  cum[0] = 0;
#pragma unroll
  for (int t = 0; t < 9; t++) {
    const int b0 = __builtin_amdgcn_readlane(rv, t), e0 = __builtin_amdgcn_readlane(rv, t + 9);
    pb[t] = b0;
    cum[t + 1] = cum[t] + (e0 - b0);
  }
  const int T = cum[9];
  const long long S = P.Scap;
  const int K = O.K;
  const float s0 = O.scale[0], s1 = O.scale[1], s2 = O.scale[2];
  const double h = F->h;
  const double org[3] = {F->org[0], F->org[1], F->org[2]};
  const int dims[3] = {nx, ny, nz};
  for (int qb = qs; qb < qe; qb += 64) {
    const int qi = qb + lane;
    const bool act = qi < qe;
    const float4 qq = act ? P.q4[qi] : make_float4(0.f, 0.f, 0.f, 0.f);
    const long long s = __float_as_int(qq.w);
    const float q0 = qq.x, q1 = qq.y, q2 = qq.z;
    float bd[KM];
    int bi[KM];
#pragma unroll
    for (int j = 0; j < KM; j++) {
      bd[j] = INFINITY;
      bi[j] = 0x7fffffff;
    }
    bool tie = false;
    float kd = INFINITY;
    auto consider = [&](const float4 cur) {
      const float d = knn_dist(q0, q1, q2, cur.x, cur.y, cur.z, s0, s1, s2);
      knn_try<KM>(bd, bi, d, __float_as_int(cur.w), K, tie, kd);
    };
    for (int cb = 0; cb < T; cb += kKnnCap) {
      const int n = min(kKnnCap, T - cb);
      __syncthreads();  // the previous chunk has been read
      for (int x = threadIdx.x; x < n; x += 64 * NW) {
        const int y = cb + x;
        int o = pb[0];  // row of y: the last t with cum[t] <= y
#pragma unroll
        for (int t = 1; t < 9; t++) o = y >= cum[t] ? pb[t] - cum[t] : o;
        cand[x] = P.cg4[o + y];
      }
      __syncthreads();
      if (act) {
        int x = w;
        for (; x + 3 * NW < n; x += 4 * NW) {
          const float4 a0 = cand[x], a1 = cand[x + NW], a2 = cand[x + 2 * NW],
                       a3 = cand[x + 3 * NW];
          consider(a0);
          consider(a1);
          consider(a2);
          consider(a3);
        }
        for (; x < n; x += NW) consider(cand[x]);
      }
    }
    if (NW > 1) {
      __syncthreads();
      if (w > 0) {
#pragma unroll
        for (int j = 0; j < KM; j++) {
          ld[w - 1][j][lane] = bd[j];
          li[w - 1][j][lane] = bi[j];
        }
        lt[w - 1][lane] = tie ? 1 : 0;
      }
      __syncthreads();
      if (w > 0) continue;  // wave 0 finishes the cell's queries
#pragma unroll
      for (int v = 0; v < NW - 1; v++) {
#pragma unroll
        for (int j = 0; j < KM; j++)
          knn_try<KM>(bd, bi, ld[v][j][lane], li[v][j][lane], K, tie, kd);
        tie |= lt[v][lane] != 0;
      }
    }
    if (!act) continue;
    const double u[3] = {(double)q0 * s0, (double)q1 * s1, (double)q2 * s2};
    knn_shells<KM, 1>(P, F, c, u, K, 0, bd, bi, tie, kd, consider, [] {});
#pragma unroll
    for (int j = 0; j < KM; j++)
      if (j < K) P.ind[(long long)j * S + s] = (bd[j] <= 1e32f) ? bi[j] : -1;
    if (tie) P.tie_list[atomicAdd(P.knn_ctl, 1)] = (int)s;
  }
}

// ------------------------------------------- the reference's scan order --
// The reference's kNN (knnKernel, knnCuda.cu:61-144) scans the Gaussians in
// index order: K slots of (distance, index) start at 1e32 with threshold
// 1e32; it leaves a Gaussian as soon as the x term, the x + y partial sum or
// the full sum is above the threshold (:116-124), otherwise the Gaussian
// replaces the slot at minPos, and minPos becomes the first slot holding the
// largest distance (findMaxPosition, :49-59). The slot layout, and under a
// tie at the K-th distance the set itself, are the outcome of that scan.
// A dead Gaussian (its mean -1e32, an infinite distance) is never taken; here
// its kNN mean is NaN and it is skipped.
template <int KM>
struct ScanSlots {
  float md[KM];
  int mi[KM];
  float thr;
  int pos;
  // a lane without a query takes nothing and keeps its slots (the lanes of a
  // wave may be replayed in turns)
  __device__ inline void init(bool live) {
#pragma unroll
    for (int j = 0; j < KM; j++) {
      md[j] = live ? 1e32f : md[j];
      mi[j] = live ? -1 : mi[j];
    }
    thr = live ? 1e32f : -INFINITY;
    pos = 0;
  }
  // distance terms of one Gaussian in the reference's order
  __device__ static inline void terms(float q0, float q1, float q2, float4 r, float s0, float s1,
                                      float s2, float& p1, float& p2, float& d) {
#pragma clang fp contract(off)
    float a = (q0 - r.x) * s0;
    p1 = a * a;
    a = (q1 - r.y) * s1;
    p2 = p1 + a * a;
    a = (q2 - r.z) * s2;
    d = p2 + a * a;
  }
  __device__ inline float threshold() const { return thr; }
  __device__ inline float dist(int j) const { return md[j]; }
  __device__ inline bool passes(float p1, float p2, float d) const {
    return !(p1 > thr) && !(p2 > thr) && !(d > thr);
  }
  __device__ inline void insert(float d, int k, int K) {
#pragma unroll
    for (int j = 0; j < KM; j++) {
      const bool at = j == pos;
      md[j] = at ? d : md[j];
      mi[j] = at ? k : mi[j];
    }
    float t = md[0];
    int p = 0;
#pragma unroll
    for (int j = 1; j < KM; j++) {
      const bool c = j < K && t < md[j];
      t = c ? md[j] : t;
      p = c ? j : p;
    }
    thr = t;
    pos = p;
  }
  __device__ inline void take(float q0, float q1, float q2, float4 r, int k, float s0, float s1,
                              float s2, int K) {
    if (r.x != r.x) return;  // dead
    float p1, p2, d;
    terms(q0, q1, q2, r, s0, s1, s2, p1, p2, d);
    if (passes(p1, p2, d)) insert(d, k, K);
  }
};

// The same slots for a wave that scans one query (every lane holds the same
// state). Insertion and the search for the largest slot are written as
// selects, so that they run without branches: the slot index is wave-uniform,
// and a branch per slot costs more than the selects themselves.
template <int KM>
struct ScanSlotsOne {
  float md[KM];
  int mi[KM];
  float thr;
  int pos;
  __device__ inline void init(bool) {
#pragma unroll
    for (int j = 0; j < KM; j++) {
      md[j] = 1e32f;
      mi[j] = -1;
    }
    thr = 1e32f;
    pos = 0;
  }
  __device__ inline float threshold() const { return thr; }
  __device__ inline float dist(int j) const { return md[j]; }
  __device__ inline bool passes(float p1, float p2, float d) const {
    return !(p1 > thr) && !(p2 > thr) && !(d > thr);
  }
  __device__ inline void insert(float d, int k, int K) {
#pragma unroll
    for (int j = 0; j < KM; j++) {
      const bool at = j == pos;
      md[j] = at ? d : md[j];
      mi[j] = at ? k : mi[j];
    }
    float t = md[0];
    int p = 0;
#pragma unroll
    for (int j = 1; j < KM; j++) {
      const bool c = j < K && t < md[j];
      t = c ? md[j] : t;
      p = c ? j : p;
    }
    thr = t;
    pos = p;
  }
};

template <int KM, bool ONE>
using SlotsOf = typename std::conditional<ONE, ScanSlotsOne<KM>, ScanSlots<KM>>::type;

// The scan replayed exactly against the frame's Gaussians of its last kNN
// search (mknn, in index order, and its grid), in one of two shapes:
//   ONE = false: one query per lane (dead lanes take nothing); the Gaussians
//     are loaded 64 at a time (one a lane), and those inside the box of the
//     wave's balls (below) are broadcast to the lanes in index order;
//   ONE = true: the wave's single query (every lane holds it); 64 Gaussians
//     are tested at once (one a lane) and the sequential rule is applied to
//     the lanes that pass the current threshold, lowest index first, the
//     test repeated after every insertion (a Gaussian skipped at the current
//     threshold is skipped by the scan: the threshold changes only when one
//     is taken). The slots are the same in every lane.
// The box of the wave's balls: the box spanned by the wave's queries
// (scaled), widened by sqrt(max threshold) with a relative margin far above
// float rounding. A Gaussian outside it is farther than every query's
// threshold on one axis, so the partial sum that adds that axis is already
// above the threshold and the scan leaves it. Thresholds only fall, so a box
// stays valid while the scan goes on; it is formed again every round of
// Gaussians. There is none (everything is scanned) while a query has fewer
// than K taken or a NaN threshold.
// The scan runs in two phases:
//  1. The first m Gaussians in index order are scanned. m grows in steps of
//     256 (512 for ONE) until the box holds no more Gaussians (on average)
//     than have been scanned.
//  2. The indices of the grid cells' Gaussians in the box are marked in an
//     LDS bitmap, compacted in index order into an LDS list, and scanned.
//     When the box or the bitmap is too large (no box yet, or more Gaussians
//     than the bitmap holds), every later Gaussian is scanned.
// LDS: words_cap bitmap words, then kListCap list entries.
constexpr int kLayoutCells = 8192;  // most grid cells marked per wave
constexpr int kLayoutWords = 4096;  // largest bitmap (Gaussians / 64) in LDS: 32 KiB
constexpr int kListCap = 1024;      // marked Gaussians listed at a time
constexpr int kMarkU = 8;           // grid entries a lane marks per round

struct ScanBox {
  double qlo[3], qhi[3];  // the wave's queries (scaled)
  double lo[3], hi[3];    // widened by the ball radius
  bool wide;              // no box: a query without K taken, or a NaN threshold
};

// Refreshes the box from the current thresholds (wave-uniform).
template <int KM, bool ONE>
__device__ inline void scan_box(ScanBox& b, const SlotsOf<KM, ONE>& sl, bool live) {
  float R2 = live ? sl.threshold() : 0.0f;
  bool wide = !(R2 < 1e32f);
  if constexpr (!ONE) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) R2 = fmaxf(R2, __shfl_xor(R2, o, 64));
    wide = __any(wide);
  }
  b.wide = wide;
  const double R = wide ? 0.0 : sqrt((double)R2) * (1.0 + 1e-5) + 1e-30;
  for (int a = 0; a < 3; a++) {
    b.lo[a] = b.qlo[a] - R;
    b.hi[a] = b.qhi[a] + R;
  }
}

// Up to 64 Gaussians, one a lane (index k relative to the frame, -1 past the
// end, in increasing order over the lanes), through the scan.
template <int KM, bool ONE>
__device__ inline void scan_take64(SlotsOf<KM, ONE>& sl, const ScanBox& box, float q0, float q1,
                                   float q2, float4 r, int k, float s0, float s1, float s2,
                                   int K) {
  const int lane = threadIdx.x & 63;
  const bool ok = k >= 0 && !(r.x != r.x);  // in range and alive
  if constexpr (ONE) {
    float p1, p2, d;
    ScanSlots<KM>::terms(q0, q1, q2, r, s0, s1, s2, p1, p2, d);
    unsigned long long mask = __ballot(ok && sl.passes(p1, p2, d));
    while (mask) {
      const int u = __builtin_ctzll(mask);
      sl.insert(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(d), u)),
                __builtin_amdgcn_readlane(k, u), K);
      mask = __ballot(ok && lane > u && sl.passes(p1, p2, d));
    }
  } else {
    bool in = ok;
    if (!box.wide) {  // comparisons with NaN keep the Gaussian
      const double u0 = (double)r.x * s0, u1 = (double)r.y * s1, u2 = (double)r.z * s2;
      in = in && !(u0 < box.lo[0] || u0 > box.hi[0] || u1 < box.lo[1] || u1 > box.hi[1] ||
                   u2 < box.lo[2] || u2 > box.hi[2]);
    }
    unsigned long long mask = __ballot(in);
    while (mask) {
      const int u = __builtin_ctzll(mask);
      mask &= mask - 1ull;
      float4 ru;
      ru.x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.x), u));
      ru.y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.y), u));
      ru.z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.z), u));
      ru.w = 0.0f;
      sl.take(q0, q1, q2, ru, __builtin_amdgcn_readlane(k, u), s0, s1, s2, K);
    }
  }
}

// n Gaussians through the scan, 64 * U per round (U loads a lane issued
// together), the box refreshed every round. idx(i) is the i-th Gaussian's
// frame-relative index (increasing in i).
template <int KM, bool ONE, class Idx>
__device__ inline void scan_list(const float4* mk, int n, Idx idx, SlotsOf<KM, ONE>& sl, ScanBox& box,
                                 bool live, float q0, float q1, float q2, float s0, float s1,
                                 float s2, int K) {
  constexpr int U = ONE ? 8 : 4;  // loads a lane in flight
  const int lane = threadIdx.x & 63;
  for (int b = 0; b < n; b += 64 * U) {
    float4 r[U];
    int kk[U];
#pragma unroll
    for (int c = 0; c < U; c++) {
      const int i = b + 64 * c + lane;
      kk[c] = i < n ? idx(i) : -1;
      r[c] = kk[c] >= 0 ? mk[kk[c]] : make_float4(NAN, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int c = 0; c < U; c++)
      if (b + 64 * c < n) scan_take64<KM, ONE>(sl, box, q0, q1, q2, r[c], kk[c], s0, s1, s2, K);
    scan_box<KM, ONE>(box, sl, live);
  }
}

// dynamic LDS of a kernel calling scan_replay with words_cap = words
inline size_t scan_lds_bytes(int words) {
  return (size_t)std::max(words, 1) * 8 + (size_t)kListCap * sizeof(int);
}

template <int KM, bool ONE>
__device__ inline void scan_replay(const DevPtrs& P, const KOpts& O, const FrameDev* F, bool live,
                                   float q0, float q1, float q2, int words_cap,
                                   unsigned long long* bits, SlotsOf<KM, ONE>& sl) {
  const int lane = threadIdx.x & 63;
  const long long g0 = F->g0;
  const int ng = (int)F->ng;
  const int K = O.K;
  const float s0 = O.scale[0], s1 = O.scale[1], s2 = O.scale[2];
  const float4* mk = P.mknn + g0;
  int* list = reinterpret_cast<int*>(bits + words_cap);
  sl.init(live);
  ScanBox box;
  for (int a = 0; a < 3; a++) {
    box.qlo[a] = 1e300;
    box.qhi[a] = -1e300;
  }
  if (live) {
    const double u[3] = {(double)q0 * s0, (double)q1 * s1, (double)q2 * s2};
    for (int a = 0; a < 3; a++) {
      box.qlo[a] = u[a];
      box.qhi[a] = u[a];
    }
  }
  if constexpr (!ONE) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
      for (int a = 0; a < 3; a++) {
        box.qlo[a] = fmin(box.qlo[a], __shfl_xor(box.qlo[a], o, 64));
        box.qhi[a] = fmax(box.qhi[a], __shfl_xor(box.qhi[a], o, 64));
      }
  }
  box.wide = true;
  const double avg_per_cell =
      (double)ng / ((double)F->dims[0] * (double)F->dims[1] * (double)F->dims[2]);
  int m = 0;
  int cl[3] = {0, 0, 0}, ch[3] = {0, 0, 0};
  long long ncell = 0;
  for (;;) {
    const int m2 = min(ng, m + (ONE ? 512 : 256));
    const int mb = m;
    scan_list<KM, ONE>(mk, m2 - mb, [mb](int i) { return mb + i; }, sl, box, live, q0, q1, q2,
                       s0, s1, s2, K);
    m = m2;
    if (m >= ng) break;
    if (!box.wide) {
      ncell = 1;
      for (int a = 0; a < 3; a++) {
        cl[a] = axis_cell(box.lo[a], F->org[a], F->h, F->dims[a]);
        ch[a] = axis_cell(box.hi[a], F->org[a], F->h, F->dims[a]);
        ncell *= (long long)(ch[a] - cl[a] + 1);
      }
      // stop when the box costs no more than the Gaussians scanned so far
      if (ncell <= kLayoutCells && (double)ncell * avg_per_cell <= fmax(1024.0, (double)m)) break;
    }
    if (m >= 16384) break;  // the box or the rest of the scan, whichever applies
  }
  if (m >= ng) return;
  const int words = (ng - m + 63) / 64;
  if (box.wide || words > words_cap || ncell > kLayoutCells) {
    const int mb = m;
    scan_list<KM, ONE>(mk, ng - mb, [mb](int i) { return mb + i; }, sl, box, live, q0, q1, q2,
                       s0, s1, s2, K);
    return;
  }
  for (int w = lane; w < words; w += 64) bits[w] = 0ull;
  // The box's rows of cells along x are contiguous runs of cg4: each lane
  // takes a row (its run's ends), the runs are laid end to end by a prefix
  // over the lanes (kept in the list area), and their entries are marked
  // 64 * kMarkU at a time, kMarkU loads a lane in flight.
  const int ny = ch[1] - cl[1] + 1;
  const int nrow = ny * (ch[2] - cl[2] + 1);
  int* rpre = list;       // 64 prefixes
  int* rbeg = list + 64;  // 64 run starts
  for (int rb = 0; rb < nrow; rb += 64) {
    const int r = rb + lane;
    int b = 0, len = 0;
    if (r < nrow) {
      const long long row =
          F->c0 + ((long long)(cl[2] + r / ny) * F->dims[1] + (cl[1] + r % ny)) * F->dims[0];
      b = P.cell_start[row + cl[0]];
      len = P.cell_start[row + ch[0] + 1] - b;
    }
    int x = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    const int tot = __shfl(x, 63, 64);
    __syncthreads();  // the previous rows' prefixes are read
    rpre[lane] = x - len;
    rbeg[lane] = b;
    __syncthreads();
    for (int j0 = 0; j0 < tot; j0 += 64 * kMarkU) {
      int pk[kMarkU];
#pragma unroll
      for (int t = 0; t < kMarkU; t++) {
        const int j = j0 + 64 * t + lane;
        pk[t] = -1;
        if (j < tot) {
          int lo = 0;  // the last row whose run starts at or before j
#pragma unroll
          for (int st = 32; st >= 1; st >>= 1)
            if (lo + st < 64 && rpre[lo + st] <= j) lo += st;
          pk[t] = rbeg[lo] + (j - rpre[lo]);
        }
      }
      int kk[kMarkU];
#pragma unroll
      for (int t = 0; t < kMarkU; t++) kk[t] = pk[t] >= 0 ? __float_as_int(P.cg4[pk[t]].w) : -1;
#pragma unroll
      for (int t = 0; t < kMarkU; t++) {
        const int k = pk[t] >= 0 ? (int)(kk[t] - g0) - m : -1;
        if (k >= 0) atomicOr(bits + (k >> 6), 1ull << (k & 63));
      }
    }
  }
  __syncthreads();
  // the marked Gaussians in index order: lane l counts the marks of its run
  // of words; list positions follow from the lanes' prefix; kListCap at a time
  const int wpl = (words + 63) / 64;
  const int w0 = min(words, lane * wpl), w1 = min(words, w0 + wpl);
  int c = 0;
  for (int w = w0; w < w1; w++) c += __popcll(bits[w]);
  int x = c;  // inclusive scan of the counts over the lanes
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  const int tot = __shfl(x, 63, 64);
  const int pre = x - c;
  for (int base = 0; base < tot; base += kListCap) {
    if (pre < base + kListCap && pre + c > base) {
      int o = pre;
      for (int w = w0; w < w1 && o < base + kListCap; w++) {
        unsigned long long word = bits[w];
        while (word) {
          if (o >= base && o < base + kListCap)
            list[o - base] = m + 64 * w + __builtin_ctzll(word);
          o++;
          word &= word - 1ull;
        }
      }
    }
    __syncthreads();
    scan_list<KM, ONE>(mk, min(kListCap, tot - base), [list](int i) { return list[i]; }, sl, box,
                       live, q0, q1, q2, s0, s1, s2, K);
    __syncthreads();  // the list is rewritten next
  }
}

// Supervoxels whose K-th distance was tied (listed by the kNN search): their
// set as the reference's scan keeps it, one wave per listed supervoxel,
// written in the search's (distance, index) order, as every other
// supervoxel's: the E-step's sums then do not depend on which supervoxels
// were searched again (the reference's slot layout is formed once, after the
// last iteration).
template <int KM>
__global__ __launch_bounds__(64) void k_knn_exact(DevPtrs P, KOpts O, int words_cap) {
  extern __shared__ unsigned long long bits[];
  const int cnt = P.knn_ctl[0];
  const long long S = P.Scap;
  const int K = O.K;
  for (int t = blockIdx.x; t < cnt; t += gridDim.x) {
    const long long s = P.tie_list[t];
    const FrameDev* F = P.fr + P.sv_frame[s];
    const float q0 = P.sv_cent[s], q1 = P.sv_cent[S + s], q2 = P.sv_cent[2 * S + s];
    ScanSlotsOne<KM> sl;
    scan_replay<KM, true>(P, O, F, true, q0, q1, q2, words_cap, bits, sl);
    if (threadIdx.x == 0) {
      float od[KM];
      int oi[KM];
#pragma unroll
      for (int j = 0; j < KM; j++) {
        od[j] = INFINITY;
        oi[j] = 0x7fffffff;
      }
      bool unused = false;
      float kd = INFINITY;
#pragma unroll
      for (int j = 0; j < KM; j++)
        if (j < K && sl.mi[j] >= 0) knn_try<KM>(od, oi, sl.dist(j), sl.mi[j], K, unused, kd);
#pragma unroll
      for (int j = 0; j < KM; j++)
        if (j < K) P.ind[(long long)j * S + s] = od[j] <= 1e32f ? (int)(F->g0 + oi[j]) : -1;
    }
    __syncthreads();  // the bitmap is free for the next listed supervoxel
  }
}

// The slot layout of every supervoxel after a solve: the reference's scan of
// its last kNN search replayed exactly (scan_replay), 64 supervoxels of
// one frame in grid-cell order per wave. The wave's supervoxels are replayed
// one row of grid cells (z, y) at a time, so that the box of the queries
// replayed together stays compact where the wave's cells wrap from the end
// of one row to the start of the next. The candidates found by the
// iterations' search (ind, rnk) are then put in these slots,
// responsibilities moving with their Gaussian; a supervoxel whose set
// differs counts in knn_ctl[1] (never, by construction). win[s] receives the
// (Gaussian, slot) whose responsibility is above 0.5, if any.
template <int KM>
__global__ __launch_bounds__(64) void k_knn_layout(DevPtrs P, KOpts O, int only_frame,
                                                   int words_cap) {
  extern __shared__ unsigned long long bits[];
  const int lane = threadIdx.x;
  const long long qb = P.s_lo + (long long)blockIdx.x * 64;
  if (qb >= P.Scap) return;
  const int f = P.sblk_frame[qb >> 8];
  if (f < 0 || (only_frame >= 0 && f != only_frame)) return;
  const FrameDev* F = P.fr + f;
  if (F->iter <= 1 || F->ng <= 0 || qb >= F->s0 + F->ns) return;  // no kNN ran
  const long long S = P.Scap;
  const long long g0 = F->g0;
  const int K = O.K;
  const long long i = qb + lane;
  const bool in = i < F->s0 + F->ns;
  const float4 qq = in ? P.q4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const long long s = in ? __float_as_int(qq.w) : 0;
  // the row of grid cells of the lane's query
  const int row = in ? (int)((P.qcell[s] - F->c0) / F->dims[0]) : 0x7fffffff;
  ScanSlots<KM> sl;
#pragma unroll
  for (int j = 0; j < KM; j++) {
    sl.md[j] = 1e32f;
    sl.mi[j] = -1;
  }
  bool todo = in;
  while (__any(todo)) {
    int r0 = todo ? row : 0x7fffffff;  // the lowest row left
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) r0 = min(r0, __shfl_xor(r0, o, 64));
    const bool live = todo && row == r0;
    scan_replay<KM, false>(P, O, F, live, qq.x, qq.y, qq.z, words_cap, bits, sl);
    __syncthreads();  // the bitmap and list are free for the next row
    todo = todo && !live;
  }
  if (in) {
    // the search's candidates into the reference's slots
    int oi[KM];
    float orr[KM];
    // one round of loads (a slot past K reads row 0, not used): no load
    // waits for the previous slot's
#pragma unroll
    for (int j = 0; j < KM; j++) {
      const long long row = j < K ? j : 0;
      const int ij = P.ind[row * S + s];
      const float rj = P.rnk[row * S + s];
      oi[j] = j < K ? ij : -1;
      orr[j] = j < K ? rj : 0.0f;
    }
    int n_old = 0, n_new = 0;
#pragma unroll
    for (int j = 0; j < KM; j++) {
      n_old += (j < K && oi[j] >= 0) ? 1 : 0;
      n_new += (j < K && sl.mi[j] >= 0) ? 1 : 0;
    }
    bool same = n_old == n_new;
    int2 w = make_int2(-1, -1);
#pragma unroll
    for (int j = 0; j < KM; j++) {
      if (j >= K) continue;
      const int k = sl.mi[j] >= 0 ? (int)(g0 + sl.mi[j]) : -1;
      float r = 0.0f;
      bool found = k < 0;
#pragma unroll
      for (int t = 0; t < KM; t++)
        if (t < K && k >= 0 && oi[t] == k) {
          r = orr[t];
          found = true;
        }
      same &= found;
      P.ind[(long long)j * S + s] = k;
      P.rnk[(long long)j * S + s] = r;
      if (k >= 0 && r > 0.5f) w = make_int2(k, j);
    }
    P.win[s] = w;
    if (!same) atomicAdd(P.knn_ctl + 1, 1);
  }
}

// The returned records (GaussianMixtureModelCUDA, 288 bytes), one lane per
// Gaussian: the 27 state doubles, the fixed flag, and the supervoxels
// assigned to the Gaussian: for slots j = 0, 1, ... in turn, the supervoxels
// s in increasing order whose candidate in slot j is this Gaussian with
// r > 0.5, at most `cap` (the reference's slot-major scan, GMEM.cu:4328-4350).
// They are found among the Gaussian's M-step pairs (r > 1e-3), ordered by
// supervoxel. Entries past the count are as given; voxel mode keeps the
// given lists (no assignment, GMEM.cu:3420-3973).
__global__ __launch_bounds__(64) void k_finalize(DevPtrs P, KOpts O, int cap, int only_frame,
                                                 int writeback) {
  const long long k = P.g_lo + (long long)blockIdx.x * 64 + threadIdx.x;
  if (k >= P.Gcap) return;
  const int f = P.g_frame[k];
  if (f < 0 || (only_frame >= 0 && f != only_frame)) return;
  const FrameDev* F = P.fr + f;
  const long long G = P.Gcap;
  GaussianMixtureModelCUDA* R = P.rec + k;
  double* d = (double*)R;
#pragma unroll
  for (int i = 0; i < GS_NFIELDS; i++) d[i] = P.gs[(long long)i * G + k];
  int* kp = P.svkeep + 17 * k;
  int idx[TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN];
#pragma unroll
  for (int i = 0; i < TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN; i++) idx[i] = kp[i];
  int n = kp[16];
  if (!O.voxel) {
    n = 0;
    if (F->iter > 1 && cap > 0) {
      const int st = P.seg_start[k], en = P.seg_start[k + 1];
      for (int j = 0; j < O.K && n < cap; j++)
        for (int e = st; e < en && n < cap; e++) {
          const long long sv = P.pairs_sorted[e] >> kPairShift;
          const int2 w = P.win[sv];
          if (w.x == (int)k && w.y == j) {
#pragma unroll
            for (int t = 0; t < TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN; t++)
              if (t == n) idx[t] = (int)(sv - F->s0);
            n++;
          }
        }
    }
  }
  unsigned int* u = (unsigned int*)((char*)R + 216);  // fixed and its padding
  *u = (unsigned int)P.fixed[k];
#pragma unroll
  for (int i = 0; i < TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN; i++) R->supervoxelIdx[i] = idx[i];
  R->supervoxelNum = n;
  if (writeback && !O.voxel) {
#pragma unroll
    for (int i = 0; i < TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN; i++) kp[i] = idx[i];
    kp[16] = n;
  }
}

// Responsibilities of the K candidates of each supervoxel. The reference's
// per-voxel sum of -0.5(3/beta + nu d^T W d) + E[ln pi] + 0.5 E[ln det]
// (GMEM.cu:2281-2298) equals, exactly,
//   n kappa - 0.5 nu (tr(W Q) + n (c - m)^T W (c - m))
// with n, c, Q the voxel count, mean and scatter of the supervoxel, so each
// (s, k) score is one FP64 evaluation; normalisation is max-subtracted
// log-sum-exp with FP32 exponentials.
// Runs of equal keys in consecutive lanes of a wave (key < 0: no entry).
// Neighbouring supervoxels, and in voxel mode neighbouring voxels, mostly
// share their j-th candidate, so one atomic per run replaces one per lane.
// Every lane of the wave must call these together.
struct WaveRun {
  int head;  // lane that starts this lane's run
  int len;   // run length (valid on the head lane)
};

__device__ inline WaveRun wave_run(int key) {
  const int lane = __lane_id();
  const int prev = __shfl_up(key, 1, 64);
  const bool cont = key >= 0 && lane > 0 && prev == key;
  const unsigned long long m = __ballot(cont);
  const unsigned long long heads = __ballot(key >= 0 && !cont);
  WaveRun r;
  const unsigned long long below = heads & (lane == 63 ? ~0ULL : ((2ULL << lane) - 1));
  r.head = below ? 63 - __builtin_clzll(below) : lane;
  const unsigned long long rest = lane == 63 ? ~0ULL : ~(m >> (lane + 1));
  r.len = 1 + __builtin_ctzll(rest);
  return r;
}

// cnt[key] += 1 for every lane with key >= 0.
__device__ inline void wave_count(int* cnt, int key) {
  const WaveRun r = wave_run(key);
  if (key >= 0 && r.head == __lane_id()) atomicAdd(cnt + key, r.len);
}

// The E-step of supervoxel s (ok: a real supervoxel of an active frame):
// its responsibilities from the candidates' log scores, and the pair counts
// of its significant candidates. Every lane of the wave calls it together
// (the pair counting is wave-wide).
template <int KM>
__device__ inline void estep_one(const DevPtrs& P, const KOpts& O, long long s, bool ok) {
  const long long S = P.Scap;
  const int K = O.K;
  // The candidates' indices and parameters and the moments depend on s
  // alone (s < Scap; an index is -1 or in range): they are loaded whether or
  // not the supervoxel is live, without waiting for the frame table, and the
  // results of a supervoxel that is not are dropped.
  int kraw[KM];
#pragma unroll
  for (int j = 0; j < KM; j++) kraw[j] = j < K ? P.ind[(long long)j * S + s] : -1;
  const int f = ok ? P.sv_frame[s] : 0;
  const double psi_sum = ok ? P.fr[f].psi_sum_alpha : 0.0;
  double n = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0, Q[6] = {0, 0, 0, 0, 0, 0};
  if (!O.voxel) {
    n = P.mom[s];
    c0 = P.mom[S + s];
    c1 = P.mom[2 * S + s];
    c2 = P.mom[3 * S + s];
#pragma unroll
    for (int i = 0; i < 6; i++) Q[i] = P.mom[(4 + i) * S + s];
  } else if (ok) {  // the voxel itself: count 1, no scatter (the same values as mom)
    const long long v = P.sv_beg[s];
    n = 1.0;
    c0 = (double)P.vx[v];
    c1 = (double)P.vy[v];
    c2 = (double)P.vz[v];
#pragma unroll
    for (int i = 0; i < 6; i++) Q[i] = 0.0;
  }
  double L[KM];
  int kk[KM];
  double mx = -INFINITY;
#pragma unroll
  for (int j = 0; j < KM; j++) {
    L[j] = -INFINITY;
    kk[j] = -1;
    if (j < K) {
      const int k = kraw[j];
      kk[j] = ok ? k : -1;
      if (k >= 0) {
        const EPack ep = P.epack[k];
        const double d0 = c0 - ep.m[0], d1 = c1 - ep.m[1], d2 = c2 - ep.m[2];
        const double* W = ep.W;
        const double qf = W[0] * d0 * d0 + W[3] * d1 * d1 + W[5] * d2 * d2 +
                          2.0 * (W[1] * d0 * d1 + W[2] * d0 * d2 + W[4] * d1 * d2);
        const double tq = W[0] * Q[0] + W[3] * Q[3] + W[5] * Q[5] +
                          2.0 * (W[1] * Q[1] + W[2] * Q[2] + W[4] * Q[4]);
        const double kappa = (ep.dpsi - psi_sum) + ep.kappa0;
        L[j] = n * kappa - 0.5 * ep.nu * (tq + n * qf);
        // voxel mode: a score <= -100 counts as zero (GMEM.cu:2208)
        if (O.voxel && !(L[j] > -100.0)) kk[j] = -2;
        if (kk[j] >= 0) mx = fmax(mx, L[j]);
      }
    }
  }
  float e[KM];
  float den = 0.0f;
#pragma unroll
  for (int j = 0; j < KM; j++) {
    e[j] = 0.0f;
    if (j < K && kk[j] >= 0 && mx > -INFINITY) {
      e[j] = expf((float)(L[j] - mx));
      den += e[j];
    }
  }
  // voxel mode stores r times the voxel's intensity (GMEM.cu:2216-2222)
  const float wv = (ok && O.voxel) ? P.vimg[P.sv_beg[s]] : 1.0f;
#pragma unroll
  for (int j = 0; j < KM; j++) {
    if (j < K) {  // K is uniform: the whole wave runs this
      const float r = ((den > 0.0f || den != den) ? e[j] / den : 0.0f) * wv;
      if (ok) P.rnk[(long long)j * S + s] = r;
      wave_count(P.seg_cnt, (ok && kk[j] >= 0 && pair_significant(r)) ? kk[j] : -1);
    }
  }
}

// The E-step of every supervoxel of the active frames.
template <int KM>
__global__ __launch_bounds__(256) void k_estep(DevPtrs P, KOpts O) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = P.sblk_frame[s >> 8];
  // no early exit: every lane takes part in the wave-wide pair counting
  const bool ok = f >= 0 && P.fr[f].active && P.sv_frame[s] == f;
  estep_one<KM>(P, O, s, ok);
}

// Each significant pair's code at its Gaussian's list, in a slot taken from
// the count (the lists are ordered by k_seg_sort). The candidates' indices
// and responsibilities depend on s alone (s < Scap) and are loaded before the
// frame checks; the slot reservations of all candidates are issued back to
// back and their results used after, so that a wave waits for one round of
// atomics, not K. Branch-free over the candidates (slots j >= K read row 0
// and hold no pair), so that no wait is forced at branch joins.
template <int KM>
__global__ __launch_bounds__(256) void k_pair_scatter(DevPtrs P, KOpts O) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long S = P.Scap;
  const int K = O.K;
  const int lane = __lane_id();
  int kr[KM];
  float rr[KM];
#pragma unroll
  for (int j = 0; j < KM; j++) {
    const long long row = j < K ? j : 0;
    kr[j] = P.ind[row * S + s];
    rr[j] = P.rnk[row * S + s];
  }
  const int f = P.sblk_frame[s >> 8];
  // no early exit: every lane takes part in the wave-wide slot reservation
  const bool ok = f >= 0 && P.fr[f].active && P.sv_frame[s] == f;
  int key[KM], head[KM], old[KM], st[KM];
#pragma unroll
  for (int j = 0; j < KM; j++) {
    key[j] = (j < K && ok && kr[j] >= 0 && pair_significant(rr[j])) ? kr[j] : -1;
    const WaveRun r = wave_run(key[j]);
    head[j] = r.head;
    int o = 0;
    if (key[j] >= 0 && r.head == lane) o = atomicSub(P.seg_cnt + key[j], r.len);
    old[j] = o;
    st[j] = P.seg_start[key[j] >= 0 ? key[j] : 0];
  }
#pragma unroll
  for (int j = 0; j < KM; j++) {
    // a run of lanes with one key takes the slots below the value its head
    // took, counting down in lane order
    const int o = __shfl(old[j], head[j], 64);
    if (key[j] >= 0) P.pairs[st[j] + o - 1 - (lane - head[j])] = (int)((s << kPairShift) | j);
  }
}

// Orders of the LDS accesses of one wave's lanes (no block barrier).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Pair lists ordered by code (= by supervoxel), so that the FP64
// accumulation order is fixed; codes are distinct (a supervoxel lists a
// Gaussian at most once). The first nshort blocks take 16 Gaussians each, 16
// lanes per Gaussian: a bitonic sort in LDS over the next power of two, for
// lists up to kSortShort. The kSortLongBlocks blocks after them take the
// Gaussians with longer lists (listed by the pair scan), one at a time: a
// bitonic sort in LDS, or a block-wide rank sort beyond kSortLds entries.
constexpr int kSortLongBlocks = 64;

__device__ inline void seg_sort_short(const DevPtrs& P, int* sh) {
  const int grp = threadIdx.x >> 4, lane = threadIdx.x & 15;
  const long long k = (long long)blockIdx.x * 16 + grp;
  // the list's bounds and the Gaussian's frame are in range for every k <
  // Gext; they are loaded before the frame checks, used only after them
  const int gf = P.g_frame[k];
  const int st = P.seg_start[k], en = P.seg_start[k + 1];
  const int f = P.gblk_frame[k / kGB];
  if (f < 0 || !P.fr[f].active) return;
  if (gf != f) return;
  const int L = en - st;
  if (L <= 1 || L > kSortShort) {
    if (L == 1 && lane == 0) P.pairs_sorted[st] = P.pairs[st];
    return;
  }
  int* a = sh + grp * kSortShort;
  int n = 2;
  while (n < L) n <<= 1;
  for (int i = lane; i < n; i += 16) a[i] = i < L ? P.pairs[st + i] : INT_MAX;
  wave_lds_sync();
  for (int w = 2; w <= n; w <<= 1)
    for (int j = w >> 1; j > 0; j >>= 1) {
      // pairs (lo, lo + j) with bit j of lo clear; ascending where bit w is
      for (int i = lane; i < (n >> 1); i += 16) {
        const int lo = 2 * i - (i & (j - 1)), hi = lo + j;
        const int x = a[lo], y = a[hi];
        if ((x > y) == ((lo & w) == 0)) {
          a[lo] = y;
          a[hi] = x;
        }
      }
      wave_lds_sync();
    }
  for (int i = lane; i < L; i += 16) P.pairs_sorted[st + i] = a[i];
}

__device__ inline void seg_sort_long(const DevPtrs& P, int* sh, int first) {
  const int cnt = P.knn_ctl[2];
  for (int li = first; li < cnt; li += kSortLongBlocks) {
    const long long k = P.long_list[li];
    const int st = P.seg_start[k], L = P.seg_start[k + 1] - st;
    const int t = threadIdx.x;
    if (L > kSortLds) {
      for (int i = t; i < L; i += 256) {
        const int e = P.pairs[st + i];
        int rank = 0;
        for (int j = 0; j < L; j++) rank += (P.pairs[st + j] < e) ? 1 : 0;
        P.pairs_sorted[st + rank] = e;
      }
      continue;
    }
    int n = 1;
    while (n < L) n <<= 1;
    for (int i = t; i < n; i += 256) sh[i] = i < L ? P.pairs[st + i] : INT_MAX;
    __syncthreads();
    for (int w = 2; w <= n; w <<= 1)
      for (int j = w >> 1; j > 0; j >>= 1) {
        for (int i = t; i < n; i += 256) {
          const int p = i ^ j;
          if (p > i) {
            const int a = sh[i], b = sh[p];
            if ((a > b) == ((i & w) == 0)) {
              sh[i] = b;
              sh[p] = a;
            }
          }
        }
        __syncthreads();
      }
    for (int i = t; i < L; i += 256) P.pairs_sorted[st + i] = sh[i];
    __syncthreads();  // sh is free for the next list
  }
}

__global__ __launch_bounds__(256) void k_seg_sort(DevPtrs P, int nshort) {
  __shared__ int sh[kSortLds > 16 * kSortShort ? kSortLds : 16 * kSortShort];
  if ((int)blockIdx.x < nshort)
    seg_sort_short(P, sh);
  else
    seg_sort_long(P, sh, (int)blockIdx.x - nshort);
}

// Posterior (GMEM.cu:1601-1677) and regularisation (GMEM.cu:1693-1809) of
// Gaussian k from its M-step sums st = (N, S1[3], S2[6]), for an alive,
// non-fixed Gaussian; commits this iteration's E[ln det] and E[ln pi].
// Returns alpha_k after the update. Not inlined: one compiled body for every
// M-step launch shape, so that every shape rounds identically (contraction
// into fused multiply-adds is decided per compiled body).
__device__ __noinline__ double mstep_update_one(double* __restrict__ gs,
                                                const double* __restrict__ eldn,
                                                const EPack* __restrict__ epack,
                                                const unsigned char* __restrict__ fixed_flags,
                                                long long G, long long k, double psi_sum,
                                                const double* st, int w4dof, float sr0,
                                                float sr1, float sr2, double lmin, double lmax,
                                                double emax) {
#define GSV(fidx, kk) gs[(long long)(fidx) * G + (kk)]
  const bool alive = GSV(GS_M_O, k) > -1e31;
  const bool fixed = fixed_flags[k] != 0;
  // this iteration's expectations, as used by its E-step
  gs[GS_ELD * G + k] = eldn[k];
  gs[GS_ELP * G + k] = alive ? epack[k].dpsi - psi_sum : -1e32;
  if (alive && !fixed) {
    const double r0 = GSV(GS_M, k), r1 = GSV(GS_M + 1, k), r2 = GSV(GS_M + 2, k);
    const double N = st[0];
    double S1[3], S2[6];
    for (int i = 0; i < 3; i++) S1[i] = st[1 + i];
    for (int i = 0; i < 6; i++) S2[i] = st[4 + i];
    double xk[3] = {0, 0, 0}, Ck[6] = {0, 0, 0, 0, 0, 0};
    if (N > 0.0) {
      const double m0 = S1[0] / N, m1 = S1[1] / N, m2 = S1[2] / N;
      xk[0] = r0 + m0;
      xk[1] = r1 + m1;
      xk[2] = r2 + m2;
      Ck[0] = S2[0] - S1[0] * m0;
      Ck[1] = S2[1] - S1[0] * m1;
      Ck[2] = S2[2] - S1[0] * m2;
      Ck[3] = S2[3] - S1[1] * m1;
      Ck[4] = S2[4] - S1[1] * m2;
      Ck[5] = S2[5] - S1[2] * m2;
    }
    // posterior (FP64 throughout)
    const double beta_o = GSV(GS_BETA_O, k), nu_o = GSV(GS_NU_O, k),
                 alpha_o = GSV(GS_ALPHA_O, k);
    double mo[3], Wo[6];
    for (int i = 0; i < 3; i++) mo[i] = GSV(GS_M_O + i, k);
    for (int i = 0; i < 6; i++) Wo[i] = GSV(GS_W_O + i, k);
    const double beta = beta_o + N, nu = nu_o + N, alpha = fmax(alpha_o + N, 0.0);
    double m[3];
    for (int i = 0; i < 3; i++) m[i] = (beta_o * mo[i] + N * xk[i]) / beta;
    if (Wo[5] < 1e-8) {
      Wo[5] = 0.5 * (Wo[0] + Wo[3]);
      gs[(GS_W_O + 5) * G + k] = Wo[5];
    }
    double A[6], W[6];
    inv3(Wo, A);
    const double dx[3] = {xk[0] - mo[0], xk[1] - mo[1], xk[2] - mo[2]};
    const double fct = beta_o / (beta_o + N);
    A[0] += Ck[0] + N * fct * dx[0] * dx[0];
    A[1] += Ck[1] + N * fct * dx[0] * dx[1];
    A[2] += Ck[2] + N * fct * dx[0] * dx[2];
    A[3] += Ck[3] + N * fct * dx[1] * dx[1];
    A[4] += Ck[4] + N * fct * dx[1] * dx[2];
    A[5] += Ck[5] + N * fct * dx[2] * dx[2];
    inv3(A, W);
    const float sreg[3] = {sr0, sr1, sr2};
    regularize_W(W, nu, w4dof, sreg, lmin, lmax, emax);
    gs[GS_BETA * G + k] = beta;
    gs[GS_NU * G + k] = nu;
    gs[GS_ALPHA * G + k] = alpha;
    for (int i = 0; i < 3; i++) gs[(GS_M + i) * G + k] = m[i];
    for (int i = 0; i < 6; i++) gs[(GS_W + i) * G + k] = W[i];
    return alpha;
  } else {
    return GSV(GS_ALPHA, k);
  }
#undef GSV
}

// One pair's terms added to the M-step sums c (N, S1[3], S2[6]) about the
// mean (r0, r1, r2): r its responsibility, m01..m89 its supervoxel's weighted
// moments (W_s first). The fused multiply-adds are written out and nothing
// else is contracted, so that every launch shape rounds the sums alike; an
// entry with use false adds exact zeros (w = r = 0, finite moments).
__device__ inline void mstep_add(double* c, bool use, double rr, double2 m01, double2 m23,
                                 double2 m45, double2 m67, double2 m89, double r0, double r1,
                                 double r2) {
#pragma clang fp contract(off)
  const double r = use ? rr : 0.0;
  const double w = use ? rr * m01.x : 0.0;
  const double d0 = m01.y - r0, d1 = m23.x - r1, d2 = m23.y - r2;
  const double wd0 = w * d0, wd1 = w * d1, wd2 = w * d2;
  c[0] = c[0] + w;
  c[1] = c[1] + wd0;
  c[2] = c[2] + wd1;
  c[3] = c[3] + wd2;
  c[4] = c[4] + fma(r, m45.x, wd0 * d0);
  c[5] = c[5] + fma(r, m45.y, wd0 * d1);
  c[6] = c[6] + fma(r, m67.x, wd0 * d2);
  c[7] = c[7] + fma(r, m67.y, wd1 * d1);
  c[8] = c[8] + fma(r, m89.x, wd1 * d2);
  c[9] = c[9] + fma(r, m89.y, wd2 * d2);
}

// M-step: segmented FP64 sufficient statistics over each Gaussian's
// significant pairs, accumulated about the current mean m_k,
//   N   = sum r W_s
//   S1  = sum r W_s (mu_s - m_k)
//   S2  = sum r [C_s + W_s (mu_s - m_k)(mu_s - m_k)^T]
// so xbar = m_k + S1/N and sum w (x - xbar)(x - xbar)^T = S2 - S1 S1^T / N,
// identical to the reference's per-voxel sums over r_sk * img_v
// (GMEM.cu:1338-1453); then, in the same block, the posterior
// (GMEM.cu:1601-1677) and regularisation (GMEM.cu:1693-1809) of the alive,
// non-fixed Gaussians, and the per-64-Gaussian sums of alpha for the death
// test. A block covers 64 Gaussians with ML lanes each; their sums go
// through LDS to wave 0, one lane per Gaussian for the update.
// Summation order, identical for every ML: sorted pair e of a Gaussian is
// added, in increasing e, to class sum e mod 16, and the class sums are
// combined as ((c + c^8) + ...^4) ...^1.
template <int ML>
__global__ __launch_bounds__(64 * ML) void k_mstep(DevPtrs P, KOpts O) {
  constexpr int NC = 16 / ML;
  __shared__ double sums[10][64];
  __shared__ double red[1];
  const int f = P.gblk_frame[blockIdx.x];
  if (f < 0) return;  // uniform over the block
  FrameDev* F = P.fr + f;
  if (!F->active) return;
  const long long G = P.Gcap, S = P.Scap;
  {
    const int g = threadIdx.x / ML;
    const long long k = (long long)blockIdx.x * 64 + g;
    const int lane = threadIdx.x % ML;
    const bool mine = P.g_frame[k] == f;
    double a[NC][10];
#pragma unroll
    for (int q = 0; q < NC; q++)
#pragma unroll
      for (int i = 0; i < 10; i++) a[q][i] = 0.0;
    if (mine) {
      const double r0 = gsv(P, GS_M, k), r1 = gsv(P, GS_M + 1, k), r2 = gsv(P, GS_M + 2, k);
      const int st = P.seg_start[k], en = P.seg_start[k + 1];
      for (int e0 = st; e0 < en; e0 += 16) {
        if constexpr (NC == 4) {
          // four entries per lane: one at a time, an entry past the list or
          // of zero weight skipped (its loads are not issued)
#pragma unroll
          for (int q = 0; q < NC; q++) {
            const int e = e0 + lane + ML * q;
            if (e >= en) continue;
            const int code = P.pairs_sorted[e];
            const long long s = code >> kPairShift;
            const int j = code & ((1 << kPairShift) - 1);
            const double r = (double)P.rnk[(long long)j * S + s];
            const double2* wm = reinterpret_cast<const double2*>(P.wmom + s * 10);
            const double2 m01 = wm[0], m23 = wm[1], m45 = wm[2], m67 = wm[3], m89 = wm[4];
            if (!(m01.x != 0.0)) continue;
            mstep_add(a[q], true, r, m01, m23, m45, m67, m89, r0, r1, r2);
          }
        } else {
          // one or two entries per lane: the entries' loads first (an entry
          // past the list reads the list's last pair, in range since
          // e0 < en), then their sums, where an entry past the list or of
          // zero weight adds exact zeros (its w and r are 0 and its moments
          // finite): no branch, so no load waits for the previous entry's
          int code[NC];
#pragma unroll
          for (int q = 0; q < NC; q++) {
            const int e = e0 + lane + ML * q;
            code[q] = P.pairs_sorted[e < en ? e : en - 1];
          }
          float rq[NC];
          double2 m01[NC], m23[NC], m45[NC], m67[NC], m89[NC];
#pragma unroll
          for (int q = 0; q < NC; q++) {
            const long long s = code[q] >> kPairShift;
            const int j = code[q] & ((1 << kPairShift) - 1);
            rq[q] = P.rnk[(long long)j * S + s];
            // the supervoxel's 10 weighted moments, contiguous (80 bytes)
            const double2* wm = reinterpret_cast<const double2*>(P.wmom + s * 10);
            m01[q] = wm[0];
            m23[q] = wm[1];
            m45[q] = wm[2];
            m67[q] = wm[3];
            m89[q] = wm[4];
          }
#pragma unroll
          for (int q = 0; q < NC; q++) {
            const int e = e0 + lane + ML * q;
            const bool use = e < en && m01[q].x != 0.0;
            mstep_add(a[q], use, (double)rq[q], m01[q], m23[q], m45[q], m67[q], m89[q], r0, r1,
                      r2);
          }
        }
      }
    }
#pragma unroll
    for (int m = 8; m >= ML; m >>= 1) {
      const int qm = m / ML;
#pragma unroll
      for (int q = 0; q < qm; q++)
#pragma unroll
        for (int i = 0; i < 10; i++) a[q][i] += a[q + qm][i];
    }
#pragma unroll
    for (int i = 0; i < 10; i++) {
#pragma unroll
      for (int m = ML / 2; m >= 1; m >>= 1) a[0][i] += __shfl_xor(a[0][i], m, ML);
    }
    if (lane == 0)
#pragma unroll
      for (int i = 0; i < 10; i++) sums[i][g] = a[0][i];
  }
  __syncthreads();
  if (threadIdx.x >= 64) return;
  const long long k = (long long)blockIdx.x * 64 + threadIdx.x;
  double st[10];
#pragma unroll
  for (int i = 0; i < 10; i++) st[i] = sums[i][threadIdx.x];
  double alpha_out = 0.0;
  if (P.g_frame[k] == f)
    alpha_out = mstep_update_one(P.gs, P.eldn, P.epack, P.fixed, G, k, F->psi_sum_alpha, st,
                                 O.w4dof, O.scale_reg[0], O.scale_reg[1], O.scale_reg[2], O.lmin,
                                 O.lmax, O.emax);
  const double t = block_sum<1>(alpha_out, red);
  if (threadIdx.x == 0) P.part_g2[blockIdx.x] = t;
}

// Dead-cell rule (GMEM.cu:1813-1843), then the log-likelihood parameters and
// the per-frame sum of alpha used by the log-likelihood and by the next
// iteration's E[ln pi]. The log-likelihood parameters are stored in base 2:
// M = nu W log2(e)/2 and lg2 = (log alpha + log det(nu W)/2 - 1.5 log 2pi) log2(e).
__global__ __launch_bounds__(kGB) void k_death(DevPtrs P, KOpts O) {
  __shared__ double red[1];
  const int f = P.gblk_frame[blockIdx.x];
  if (f < 0) return;
  FrameDev* F = P.fr + f;
  if (!F->active) return;
  // the frame's sum of alpha after the posterior update, from the blocks'
  // partials in index order (every block of the frame forms the same value)
  const double sum_alpha2 = frame_sum_w64(P.part_g2 + F->g0 / kGB, F->nbg);
  const long long k = (long long)blockIdx.x * kGB + threadIdx.x;
  const long long G = P.Gcap;
  double alpha_out = 0.0;
  if (P.g_frame[k] == f) {
    double W[6];
    bool nanW = isnan(gsv(P, GS_ELD, k));
    for (int i = 0; i < 6; i++) {
      W[i] = gsv(P, GS_W + i, k);
      nanW |= isnan(W[i]);
    }
    double alpha = gsv(P, GS_ALPHA, k);
    bool alive = gsv(P, GS_M_O, k) > -1e31;
    const bool fixed = P.fixed[k] != 0;
    if (nanW || (alpha / sum_alpha2 < 1e-5 && alive && !fixed)) {
      alpha = 0.0;
      P.gs[GS_ALPHA * G + k] = 0.0;
      for (int i = 0; i < 3; i++) {
        P.gs[(GS_M_O + i) * G + k] = -1e32;
        P.gs[(GS_M + i) * G + k] = -1e32;
      }
      const double I6[6] = {1, 0, 0, 1, 0, 1};
      for (int i = 0; i < 6; i++) {
        W[i] = I6[i];
        P.gs[(GS_W + i) * G + k] = I6[i];
      }
      alive = false;
    }
    const double LOG2E = 1.4426950408889634074;
    const double nu = gsv(P, GS_NU, k);
    double m[3];
    for (int i = 0; i < 3; i++) m[i] = gsv(P, GS_M + i, k);
    gauss_prep_one(P, O, F, k, alive, m, W, nu, gsv(P, GS_BETA, k), alpha);
    LLPack lp;
    for (int i = 0; i < 3; i++) lp.m[i] = m[i];
    // Cholesky factor (FP64) of M = nu W log2(e)/2; a precision matrix that
    // is not positive definite has no density, its terms are dropped
    double M[6];
    for (int i = 0; i < 6; i++) M[i] = nu * W[i] * (0.5 * LOG2E);
    const double u00 = sqrt(M[0]);
    const double u01 = M[1] / u00, u02 = M[2] / u00;
    const double u11 = sqrt(M[3] - u01 * u01);
    const double u12 = (M[4] - u01 * u02) / u11;
    const double u22 = sqrt(M[5] - u02 * u02 - u12 * u12);
    const bool pd = u00 > 0.0 && u11 > 0.0 && u22 > 0.0;  // false for NaN too
    lp.U[0] = pd ? (float)u00 : 0.0f;
    lp.U[1] = pd ? (float)u01 : 0.0f;
    lp.U[2] = pd ? (float)u02 : 0.0f;
    lp.U[3] = pd ? (float)u11 : 0.0f;
    lp.U[4] = pd ? (float)u12 : 0.0f;
    lp.U[5] = pd ? (float)u22 : 0.0f;
    {
      double f2 = 0.0;
      for (int i = 0; i < 6; i++) f2 += (double)lp.U[i] * (double)lp.U[i];
      lp.nf = (float)(sqrt(f2) * (1.0 + 1e-6));
    }
    lp.pad = 0.0f;
    const double dt = det3(W) * nu * nu * nu;
    lp.lg2 = (alive && alpha > 0.0 && dt > 0.0 && pd)
                 ? (log(alpha) + 0.5 * log(dt) - 1.5 * 1.8378770664093454836) * LOG2E
                 : -INFINITY;
    P.llpack[k] = lp;
    alpha_out = alpha;
  }
  const double t = block_sum<1>(alpha_out, red);
  if (threadIdx.x == 0) P.part_g[blockIdx.x] = t;
}

// A frame's sum of alpha from the Gaussian-block partials of the kernel
// before, one wave per frame, partials in index order (lane l takes l, l+64,
// ... into four interleaved sums, then a fixed butterfly): the sum used by
// E[ln pi] and the log-likelihood, with its digamma and log.
__global__ __launch_bounds__(64) void k_alpha_sum(DevPtrs P, int need_active) {
  const int f = (P.only_frame >= 0) ? P.only_frame : (int)blockIdx.x;
  if (f >= P.Fcap || (P.only_frame < 0 && f >= P.nframes)) return;
  FrameDev* F = P.fr + f;
  if (F->ncap == 0 || (need_active && !F->active)) return;
  const double tot = frame_sum_w64(P.part_g + F->g0 / kGB, F->nbg);
  if (threadIdx.x == 0) {
    F->sum_alpha = tot;
    F->psi_sum_alpha = digamma(tot);
    F->log_sum_alpha = log(tot);
  }
}


// Log-likelihood over the voxel stream (GMEM.cu:1913-1978). A block of 256
// threads takes B consecutive supervoxels of one frame (B divides 256):
//  1. Candidate constants, LPS lanes per supervoxel (lane j < K: candidate
//     j), kept in LDS: the Cholesky factor U and e = U (a_s - m_k), with a_s
//     an anchor voxel of the supervoxel, formed in FP64 and rounded once, and
//     the log2 weight g. Per voxel and candidate the scaled quadratic form is
//     |y|^2 with y = U (x - a_s) + e (x - a_s is exact in f32), so the f32
//     form keeps full relative precision far from the origin (9 FMA-class
//     operations per candidate).
//  2. The supervoxels' voxels, which are contiguous in the stream, in chunks
//     of 256*UV: thread t takes voxels t, t + 256, ... (coalesced loads, UV
//     a thread issued together), finds its supervoxel by a binary search of
//     the block's voxel starts, and forms the voxel's value into LDS.
//  3. The chunk's values are added to the supervoxels' class sums.
// Terms are evaluated in base 2 (hardware exp2/log2), log-sum-exp with the
// largest term factored out; a voxel's value img * term is rounded to FP64
// and accumulated in FP64. A candidate whose term underflows in the
// reference's double evaluation (its exponential or the whole term below
// 2^-1075) adds nothing, and a voxel with no term left contributes nothing
// (GMEM.cu:1958-1959).
//
// Candidates that are negligible at every voxel of the supervoxel are left
// out. With c the centre and r the radius of the ball around the
// supervoxel's voxels (sv_ball) and nf >= |U|_2, every voxel x has
// |U (x - m)| within sqrt(q_c) -+ nf r, where q_c is the form at c (FP64),
// so t(x) = g - q(x) lies in [g - hi, g - lo]; both bounds are widened by far
// more than the f32 evaluation's rounding. A candidate with g - hi >= -1000
// and hi <= 1000 (never cut as underflowing) bounds every voxel's largest
// term from below by L = max(g - hi); a candidate with g - lo < L - 40 then
// adds less than 2^-39 of the largest term at every voxel (the bounds are
// formed in f32), and is dropped.
// The others keep their slot order. The decision depends only on the
// supervoxel and its candidates, not on the launch shape. In voxel mode
// every candidate is kept, as k_loglik_voxel keeps them.
//
// Summation order, identical for every (B, UV): voxel i of a supervoxel is
// added, in increasing i, to class sum i mod 16 (starting from 0); the 16
// class sums are added as ((c + c^8) + ...^4) ...^1 (the butterfly of a
// 16-lane group, read at lane 0); the sums of each 256 consecutive
// supervoxels by one pairwise tree (a block forms the subtree over its B
// leaves, k_ll_final the levels above); the frame's 256-supervoxel partials
// by frame_sum_w64. No block waits for another: the frame total is formed by
// k_ll_final after this kernel.
template <int N>
__device__ inline double pairwise_tree(double* v) {  // in place, v[0] = result
#pragma unroll
  for (int w = N / 2; w >= 1; w >>= 1)
#pragma unroll
    for (int i = 0; i < w; i++) v[i] = v[2 * i] + v[2 * i + 1];
  return v[0];
}

// The butterfly sum of 16 class sums as lane 0 of a 16-lane group forms it
// with tot += shfl_xor(tot, m) for m = 8, 4, 2, 1.
__device__ inline double class_butterfly(const double* c) {
  double a[8], b[4], d[2];
#pragma unroll
  for (int l = 0; l < 8; l++) a[l] = c[l] + c[l + 8];
#pragma unroll
  for (int l = 0; l < 4; l++) b[l] = a[l] + a[l + 4];
#pragma unroll
  for (int l = 0; l < 2; l++) d[l] = b[l] + b[l + 2];
  return d[0] + d[1];
}

template <int KM, int B, int UV, int NT>
__global__ __launch_bounds__(NT) void k_loglik(DevPtrs P, KOpts O) {
  constexpr int LPS = KM <= 8 ? 8 : 16;  // lanes per supervoxel, constants
  constexpr int SPP = NT / LPS;          // supervoxels per constants pass
  constexpr int CH = NT * UV;            // voxels per chunk
  static_assert(B % SPP == 0 && 256 % B == 0 && B >= 4, "block shape");
  __shared__ float4 cst[B * KM * 3];
  __shared__ float4 anc[B];   // anchor, and the number of candidates kept (int bits)
  __shared__ int vst[B + 1];  // first voxel of each supervoxel, relative to vbase
  __shared__ double vals[CH];
  __shared__ double cls[B * 16];
  __shared__ int vend_s;
  const long long sb = (long long)blockIdx.x * B;
  const long long S = P.Scap;
  const int K = O.K;
  // A lane's loads for its (supervoxel, slot) of a constants pass: they
  // depend on the block's supervoxels alone, so the first pass's are issued
  // before the frame's, together, and the candidate's parameters right after
  // its index (in range whenever it is not -1).
  struct Fetch {
    int frame, len, kraw;
    long long beg;
    float a0, a1, a2;
    float4 ball;
    LLPack lp;
  };
  auto fetch = [&](int p) {
    Fetch r;
    const long long s = sb + p + threadIdx.x / LPS;
    const int lane = threadIdx.x % LPS;
    r.frame = P.sv_frame[s];
    r.len = P.sv_len[s];
    r.beg = P.sv_beg[s];
    r.a0 = P.anchor[s];
    r.a1 = P.anchor[S + s];
    r.a2 = P.anchor[2 * S + s];
    r.ball = P.sv_ball[s];
    r.kraw = lane < K ? P.ind[(long long)lane * S + s] : -1;
    if (r.kraw >= 0) r.lp = P.llpack[r.kraw];
    return r;
  };
  const int fr0 = P.sv_frame[sb];
  const long long beg0 = P.sv_beg[sb];
  const Fetch first = fetch(0);
  const int f = P.sblk_frame[sb >> 8];
  if (f < 0) return;
  const FrameDev* F = P.fr + f;
  if (!F->active) return;
  const double lsa2 = F->log_sum_alpha * 1.4426950408889634074;
  // the block's first supervoxel is real unless the whole block is padding
  const long long vbase = fr0 == f ? beg0 : 0;
  // the voxels of a chunk, raw (packed: .x .y of vpk; else x, y, z, img
  // bits), loaded within the frame's voxel range; the first chunk's loads
  // are issued before the constants are formed, each next chunk's before
  // the current one is evaluated
  const bool packed = F->packed != 0;
  const long long vlim = F->v0 + F->nv;
  auto load_chunk = [&](int c0, int4 (&raw)[UV]) {
#pragma unroll
    for (int u = 0; u < UV; u++) {
      const long long g = vbase + c0 + u * NT + threadIdx.x;
      raw[u] = make_int4(0, 0, 0, 0);
      if (g < vlim) {  // streamed once: non-temporal loads
        if (packed) {
          const unsigned long long pk =
              __builtin_nontemporal_load(reinterpret_cast<const unsigned long long*>(P.vpk) + g);
          raw[u].x = (int)(unsigned)pk;
          raw[u].y = (int)(unsigned)(pk >> 32);
        } else {
          raw[u] = make_int4(__float_as_int(__builtin_nontemporal_load(P.vx + g)),
                             __float_as_int(__builtin_nontemporal_load(P.vy + g)),
                             __float_as_int(__builtin_nontemporal_load(P.vz + g)),
                             __float_as_int(__builtin_nontemporal_load(P.vimg + g)));
        }
      }
    }
  };
  int4 cur[UV];
  load_chunk(0, cur);
  if (threadIdx.x == 0) vend_s = 0;
  for (int x = threadIdx.x; x < B * 16; x += NT) cls[x] = 0.0;
  __syncthreads();
  // 1. constants
#pragma unroll
  for (int p = 0; p < B; p += SPP) {
    const int i = p + threadIdx.x / LPS;
    const int lane = threadIdx.x % LPS;
    const Fetch r = p == 0 ? first : fetch(p);
    const bool real = r.frame == f;
    const int n = real ? r.len : 0;
    const float a0 = r.a0, a1 = r.a1, a2 = r.a2;
    const float4 ball = r.ball;
    const int kk = n > 0 ? r.kraw : -1;
    float pe0 = 0.0f, pe1 = 0.0f, pe2 = 0.0f, pu[6] = {0, 0, 0, 0, 0, 0}, pg = -INFINITY;
    double tlo = -INFINITY, thi = INFINITY;  // bounds of the candidate's terms
    if (kk >= 0) {
      const LLPack& lp = r.lp;
      const double d0 = (double)a0 - lp.m[0], d1 = (double)a1 - lp.m[1], d2 = (double)a2 - lp.m[2];
#pragma unroll
      for (int q = 0; q < 6; q++) pu[q] = lp.U[q];
      pe0 = (float)((double)pu[0] * d0 + (double)pu[1] * d1 + (double)pu[2] * d2);
      pe1 = (float)((double)pu[3] * d1 + (double)pu[4] * d2);
      pe2 = (float)((double)pu[5] * d2);
      pg = (float)(lp.lg2 - lsa2);
      {
        // in f32: the rounding of y = U (c - m) is below 1e-6 nf |c - m|,
        // added to the ball's reach, and a relative error of sq below 1e-6
        // is far inside err's margin
#pragma clang fp contract(off)
        const float b0 = (float)(d0 + (double)ball.x), b1 = (float)(d1 + (double)ball.y),
                    b2 = (float)(d2 + (double)ball.z);  // centre - mean
        const float y0 = pu[0] * b0 + pu[1] * b1 + pu[2] * b2;
        const float y1 = pu[3] * b1 + pu[4] * b2;
        const float y2 = pu[5] * b2;
        const float sq = sqrtf(y0 * y0 + y1 * y1 + y2 * y2);
        const float bn = sqrtf(b0 * b0 + b1 * b1 + b2 * b2);
        const float rr = lp.nf * (ball.w + 1e-6f * bn);
        const float w = sq + 3.0f * rr;
        const float err = 1e-5f * (w * w + fabsf(pg) + 1.0f);
        const float lo = fmaxf(sq - rr, 0.0f);
        const float hi = (sq + rr) * (sq + rr) + err;
        if (pg > -INFINITY) {
          thi = (double)(pg - (lo * lo - err));
          if (hi <= 1000.0f) tlo = (double)(pg - hi);
        }
      }
    }
    // the candidates kept, compacted in slot order
    double L = tlo;
#pragma unroll
    for (int m = LPS / 2; m >= 1; m >>= 1) L = fmax(L, __shfl_xor(L, m, LPS));
    const bool live =
        kk >= 0 && pg > -INFINITY && !(!O.voxel && L >= -1000.0 && thi < L - 40.0);
    const unsigned long long gm =
        (__ballot(live) >> ((threadIdx.x & 63) & ~(LPS - 1))) & ((1ull << LPS) - 1ull);
    const int slot = __popcll(gm & ((1ull << lane) - 1ull));
    if (live) {
      // .z of the third: the largest q whose term the reference does not
      // flush to zero (2^-q and 2^(pg - q) both at least 2^-1075)
      float4* d = cst + (i * KM + slot) * 3;
      d[0] = make_float4(pu[0], pu[1], pu[2], pe0);
      d[1] = make_float4(pu[3], pu[4], pe1, pu[5]);
      d[2] = make_float4(pe2, pg, fminf(1075.0f, pg + 1075.0f), 0.0f);
    }
    if (lane == 0) {
      anc[i] = make_float4(a0, a1, a2, __int_as_float(__popcll(gm)));
      vst[i] = real ? (int)(r.beg - vbase) : -1;
      if (real) atomicMax(&vend_s, (int)(r.beg + n - vbase));
    }
  }
  __syncthreads();
  const int vend = vend_s;
  // padding supervoxels (after the frame's last) start where the real ones end
  for (int i = threadIdx.x; i <= B; i += NT)
    if (i == B || vst[i] < 0) vst[i] = vend;
  __syncthreads();
  // 2. and 3. the voxels, chunk by chunk
  for (int c0 = 0; c0 < vend; c0 += CH) {
    int4 nxt[UV];
#pragma unroll
    for (int u = 0; u < UV; u++) nxt[u] = make_int4(0, 0, 0, 0);
    if (c0 + CH < vend) load_chunk(c0 + CH, nxt);
    // coordinates (or offsets from the anchor, packed) and intensity
    float4 vv[UV];
    int svi[UV];  // the voxels' supervoxels in the block (packed form)
#pragma unroll
    for (int u = 0; u < UV; u++) {
      if (packed) {
        vv[u] = make_float4((float)(signed char)(cur[u].x & 0xff),
                            (float)(signed char)((cur[u].x >> 8) & 0xff),
                            (float)(signed char)((cur[u].x >> 16) & 0xff), __int_as_float(cur[u].y));
        svi[u] = (int)(((unsigned)cur[u].x >> 24) - (unsigned)(sb & 255));
      } else {
        vv[u] = make_float4(__int_as_float(cur[u].x), __int_as_float(cur[u].y),
                            __int_as_float(cur[u].z), __int_as_float(cur[u].w));
        svi[u] = 0;
      }
    }
#pragma unroll
    for (int u = 0; u < UV; u++) {
      const int v = c0 + u * NT + threadIdx.x;
      if (v >= vend) continue;
      int i = 0;  // the last supervoxel starting at or before v
      if (packed) {
        i = svi[u];
      } else {
#pragma unroll
        for (int st = B / 2; st >= 1; st >>= 1)
          if (vst[i + st] <= v) i += st;
      }
      const float4 an = anc[i];
      const int nlive = __float_as_int(an.w);
      float tj[KM];
      int nm = 0;          // candidate slots any lane of the wave uses
      float mx = -INFINITY;  // the largest term
      {
#pragma clang fp contract(off)
        const float x0 = packed ? vv[u].x : vv[u].x - an.x;
        const float x1 = packed ? vv[u].y : vv[u].y - an.y;
        const float x2 = packed ? vv[u].z : vv[u].z - an.z;
        // the supervoxel's constants; slots at or past nlive are evaluated
        // too (within the block's constants) and their terms dropped, so
        // that the wave runs the slots without branching
        const float4* c4 = cst + i * (KM * 3);
#pragma unroll
        for (int j = 0; j < KM; j++) {
          tj[j] = -INFINITY;
          if (!__any(j < nlive)) break;
          nm = j + 1;
          const float4 A = c4[3 * j], Bq = c4[3 * j + 1], C = c4[3 * j + 2];
          const float y0 = __builtin_fmaf(A.x, x0, __builtin_fmaf(A.y, x1,
                                                                  __builtin_fmaf(A.z, x2, A.w)));
          const float y1 = __builtin_fmaf(Bq.x, x1, __builtin_fmaf(Bq.y, x2, Bq.z));
          const float y2 = __builtin_fmaf(Bq.w, x2, C.x);
          const float q = __builtin_fmaf(y0, y0, __builtin_fmaf(y1, y1, y2 * y2));
          // the reference's double term underflows to 0 when its exponential
          // (2^-q) or the whole term (2^(pg - q)) is below 2^-1075
          const float t = C.y - q;
          tj[j] = ((j < nlive) & (q <= C.z)) ? t : -INFINITY;
          mx = __builtin_amdgcn_fmed3f(tj[j], mx, INFINITY);  // max; terms are never NaN
        }
      }
      nm = __builtin_amdgcn_readfirstlane(nm);  // the same in every lane
      double val = 0.0;
      {
#pragma clang fp contract(off)
        if (nlive == 1) {
          // a single term t: the value is img * t (k_loglik_voxel does the same)
          if (mx > -INFINITY) val = (double)vv[u].w * (double)mx;
        } else if (mx > -INFINITY) {
          // slots at or past nlive hold -inf and add exp2(-inf) = +0
          float sum = 0.0f;
#pragma unroll
          for (int j = 0; j < KM; j++) {
            if (j >= nm) break;
            sum += __builtin_amdgcn_exp2f(tj[j] - mx);
          }
          val = (double)vv[u].w * (double)(mx + __builtin_amdgcn_logf(sum));
        }
      }
      vals[u * NT + threadIdx.x] = val;
    }
    __syncthreads();
    // class (i, c): the chunk's voxels of supervoxel i at positions = c mod 16,
    // for the supervoxels the chunk overlaps
    int ia = 0, ib = 0;
    {
      const int vl = min(c0 + CH, vend) - 1;
#pragma unroll
      for (int st = B / 2; st >= 1; st >>= 1) {
        if (vst[ia + st] <= c0) ia += st;
        if (vst[ib + st] <= vl) ib += st;
      }
    }
    for (int task = ia * 16 + threadIdx.x; task < (ib + 1) * 16; task += NT) {
      const int i = task >> 4, c = task & 15;
      const int s0 = vst[i], lo = max(s0, c0), hi = min(vst[i + 1], c0 + CH);
      if (lo >= hi) continue;
      int v = s0 + c;
      if (v < lo) v += ((lo - v + 15) >> 4) << 4;
      double a = cls[task];
      for (; v < hi; v += 16) a += vals[v - c0];
      cls[task] = a;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < UV; u++) cur[u] = nxt[u];
  }
  // the supervoxel sums, then pairwise_tree<B>'s levels in LDS
  if (threadIdx.x < B) vals[threadIdx.x] = class_butterfly(cls + 16 * threadIdx.x) * 0.69314718055994530942;
  __syncthreads();
  for (int w = B / 2; w >= 1; w >>= 1) {
    double a = 0.0;
    if ((int)threadIdx.x < w) a = vals[2 * threadIdx.x] + vals[2 * threadIdx.x + 1];
    __syncthreads();
    if ((int)threadIdx.x < w) vals[threadIdx.x] = a;
    __syncthreads();
  }
  if (threadIdx.x == 0) P.part_s[blockIdx.x] = vals[0];
}

// Log-likelihood in voxel mode, where every supervoxel is one voxel (or
// none, as padding): one lane per supervoxel, 256 per block. A lane forms its
// voxel's term with k_loglik's operations (the anchor is the voxel itself, so
// y = U (x - a) + e = e) over all K candidates (bounding a one-voxel
// supervoxel's terms costs as much as evaluating them), its supervoxel sum
// as k_loglik's class sums leave it (0 + term), and the block the same
// pairwise tree over its 256 supervoxels, without 15 idle lanes of 16.
template <int KM>
__global__ __launch_bounds__(256) void k_loglik_voxel(DevPtrs P, KOpts O) {
  __shared__ double tree[256];
  const int f = P.sblk_frame[blockIdx.x];
  if (f < 0) return;
  const FrameDev* F = P.fr + f;
  if (!F->active) return;
  const long long S = P.Scap;
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
  const int K = O.K;
  const double lsa2 = F->log_sum_alpha * 1.4426950408889634074;
  const int n = (P.sv_frame[s] == f) ? P.sv_len[s] : 0;
  double t = 0.0;
  if (n > 0) {
    const long long g = P.sv_beg[s];
    const float4 vv = make_float4(P.vx[g], P.vy[g], P.vz[g], P.vimg[g]);
    const float ca0 = P.anchor[s], ca1 = P.anchor[S + s], ca2 = P.anchor[2 * S + s];
    float tj[KM];
    int nterm = 0;  // candidates with a density (kept by k_loglik's rule in voxel mode)
#pragma unroll
    for (int j = 0; j < KM; j++) {
      const int kk = (j < K) ? P.ind[(long long)j * S + s] : -1;
      float pe0 = 0.0f, pe1 = 0.0f, pe2 = 0.0f, pu[6] = {0, 0, 0, 0, 0, 0}, pg = -INFINITY;
      if (kk >= 0) {
        const LLPack lp = P.llpack[kk];
        const double d0 = (double)ca0 - lp.m[0], d1 = (double)ca1 - lp.m[1],
                     d2 = (double)ca2 - lp.m[2];
#pragma unroll
        for (int i = 0; i < 6; i++) pu[i] = lp.U[i];
        pe0 = (float)((double)pu[0] * d0 + (double)pu[1] * d1 + (double)pu[2] * d2);
        pe1 = (float)((double)pu[3] * d1 + (double)pu[4] * d2);
        pe2 = (float)((double)pu[5] * d2);
        pg = (float)(lp.lg2 - lsa2);
      }
      nterm += (kk >= 0 && pg > -INFINITY) ? 1 : 0;
      {
#pragma clang fp contract(off)
        const float x0 = vv.x - ca0, x1 = vv.y - ca1, x2 = vv.z - ca2;
        const float y0 = __builtin_fmaf(pu[0], x0, __builtin_fmaf(pu[1], x1,
                                                                  __builtin_fmaf(pu[2], x2, pe0)));
        const float y1 = __builtin_fmaf(pu[3], x1, __builtin_fmaf(pu[4], x2, pe1));
        const float y2 = __builtin_fmaf(pu[5], x2, pe2);
        const float q = __builtin_fmaf(y0, y0, __builtin_fmaf(y1, y1, y2 * y2));
        const float tt = pg - q;
        tj[j] = (q <= 1075.0f && tt >= -1075.0f) ? tt : -INFINITY;
      }
    }
    {
#pragma clang fp contract(off)
      float mx = -INFINITY;
#pragma unroll
      for (int j = 0; j < KM; j++) mx = fmaxf(mx, tj[j]);
      if (nterm == 1) {
        if (mx > -INFINITY) t = (double)vv.w * (double)mx;
      } else if (mx > -INFINITY) {
        float sum = 0.0f;
#pragma unroll
        for (int j = 0; j < KM; j++) sum += __builtin_amdgcn_exp2f(tj[j] - mx);
        t = (double)vv.w * (double)(mx + __builtin_amdgcn_logf(sum));
      }
    }
  }
  double cls = 0.0;
  cls += t;
  tree[threadIdx.x] = cls * 0.69314718055994530942;
  __syncthreads();
  // the pairwise tree of pairwise_tree<256>, one level per step
  for (int w = 128; w >= 1; w >>= 1) {
    double a = 0.0;
    if (threadIdx.x < w) a = tree[2 * threadIdx.x] + tree[2 * threadIdx.x + 1];
    __syncthreads();
    if (threadIdx.x < w) tree[threadIdx.x] = a;
    __syncthreads();
  }
  if (threadIdx.x == 0) P.part_s[blockIdx.x] = tree[0];
}

// The upper levels of each 256-supervoxel group's tree from the block
// partials of k_loglik (BPG = 256 / B blocks per group), one thread per
// group of an active frame, all frames together.
template <int BPG>
__global__ __launch_bounds__(256) void k_ll_groups(DevPtrs P) {
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  if (g >= (P.Sext + 255) / 256) return;
  const int f = P.sblk_frame[g];
  if (f < 0 || !P.fr[f].active) return;
  double q[BPG];
#pragma unroll
  for (int i = 0; i < BPG; i++) q[i] = P.part_s[g * BPG + i];
  P.part_sg[g] = pairwise_tree<BPG>(q);
}

// The frame's total (frame_sum_w64 over its group sums, in lanes 0..63 of
// the calling block), the convergence test (GMEM.cu:4145-4148) and the
// iteration count.
__device__ inline void ll_frame_total(FrameDev* F, const double* groups) {
  const double ll = frame_sum_w64(groups + (F->s0 >> 8), F->nbl);
  if (threadIdx.x == 0) {
    F->ll3 = F->ll2;
    F->ll2 = F->ll1;
    F->ll1 = F->ll;
    F->ll = ll;
    F->iter += 1;
    F->active = (keep_going(F->ll, F->ll1, F->ll2, F->ll3, F->tol) && F->iter <= F->max_iter) ? 1 : 0;
  }
}

// Frame totals of the log-likelihood, one wave per frame, from the group
// sums (the block partials themselves when BPG = 1).
template <int BPG>
__global__ __launch_bounds__(64) void k_ll_final(DevPtrs P) {
  FrameDev* F = P.fr + blockIdx.x;
  if (!F->active) return;
  ll_frame_total(F, BPG == 1 ? P.part_s : P.part_sg);
}

// Both steps in one block per frame, for batches with few groups (the
// groups' trees a thread at a time, then the total): one launch fewer.
constexpr int kLLFrameGroups = 2048;  // most groups of all frames for this form
template <int BPG>
__global__ __launch_bounds__(1024) void k_ll_frame(DevPtrs P) {
  FrameDev* F = P.fr + blockIdx.x;
  if (!F->active) return;
  const long long g0 = F->s0 >> 8;
  for (int g = threadIdx.x; g < F->nbl; g += 1024) {
    double q[BPG];
#pragma unroll
    for (int i = 0; i < BPG; i++) q[i] = P.part_s[(g0 + g) * BPG + i];
    P.part_sg[g0 + g] = pairwise_tree<BPG>(q);
  }
  __syncthreads();
  if (threadIdx.x >= 64) return;
  ll_frame_total(F, P.part_sg);
}

// ---------------------------------------------------------------- launchers --
int kmax_for(int K) { return K <= 5 ? 5 : (K <= 8 ? 8 : 16); }

static inline unsigned nb(long long n, int per) { return (unsigned)((n + per - 1) / per); }


hipError_t launch_prep(const DevPtrs& P, const KOpts& O, hipStream_t st) {
  if (O.voxel)
    k_moments_voxel<<<nb(P.Sext, 256), 256, 0, st>>>(P);
  else
    k_moments<<<nb(P.Sext, 16), 256, 0, st>>>(P);
  k_grid_setup<<<P.only_frame >= 0 ? 1 : P.nframes, 1024, 0, st>>>(P, O);
  k_qcell_count<<<nb(P.Sext, 256), 256, 0, st>>>(P, O);
  k_scan_frames<<<P.only_frame >= 0 ? 1 : P.nframes, 1024, 0, st>>>(P, 2, 0, O.K);
  k_qcell_scatter<<<nb(P.Sext, 256), 256, 0, st>>>(P);
  return hipGetLastError();
}

hipError_t launch_reset(const DevPtrs& P, const KOpts& O, hipStream_t st) {
  k_iter_init<<<nb(P.nframes, 64), 64, 0, st>>>(P);
  k_alpha_init<<<nb(P.Gext, kGB), kGB, 0, st>>>(P);
  k_alpha_sum<<<P.only_frame >= 0 ? 1 : P.nframes, 64, 0, st>>>(P, 1);
  k_gauss_prep<<<nb(P.Gext, kGB), kGB, 0, st>>>(P, O);
  return hipGetLastError();
}

// GaussianMixtureModel::updatePriors (GaussianMixtureModel.cpp:892-910) with
// N_k = nu_k - nu_o (GaussianMixtureModel.h:169), then the tracker's
// nu_k *= 2, W_k /= 2 (main.cpp:961-969).
__global__ void k_next_priors(DevPtrs P, long long g0, long long ng, double bpct, double npct,
                              double apct, double alpha_total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ng) return;
  const long long k = g0 + i;
  const long long G = P.Gcap;
  double* gs = P.gs;
  const double nu_k = gs[GS_NU * G + k], nu_o = gs[GS_NU_O * G + k];
  const double Nk = nu_k - nu_o;
  for (int a = 0; a < 3; a++) gs[(GS_M_O + a) * G + k] = gs[(GS_M + a) * G + k];
  gs[GS_BETA_O * G + k] = fmax(bpct * Nk, 1e-5);
  const double nu_o_new = fmax(npct * Nk, 3.0);
  gs[GS_NU_O * G + k] = nu_o_new;
  for (int a = 0; a < 6; a++) gs[(GS_W_O + a) * G + k] = gs[(GS_W + a) * G + k] * nu_k / nu_o_new;
  if (alpha_total >= 0.0) gs[GS_ALPHA_O * G + k] = alpha_total * 1e-5 * apct;
  gs[GS_NU * G + k] = nu_k * 2.0;
  for (int a = 0; a < 6; a++) gs[(GS_W + a) * G + k] /= 2.0;
}

hipError_t launch_next_priors(const DevPtrs& P, long long g0, long long ng, double beta_pct,
                              double nu_pct, double alpha_pct, double alpha_total,
                              hipStream_t st) {
  if (ng <= 0) return hipSuccess;
  k_next_priors<<<nb(ng, 256), 256, 0, st>>>(P, g0, ng, beta_pct, nu_pct, alpha_pct,
                                             alpha_total);
  return hipGetLastError();
}

// ev_after (optional) is recorded between the log-likelihood kernel and the
// per-frame totals
template <int KM>
static void launch_loglik(const DevPtrs& P, const KOpts& O, hipStream_t st,
                          hipEvent_t ev_after) {
  const int gw = O.tiles.ll_gw, spg = O.tiles.ll_u;
  // timed runs: kTimedLaunches back-to-back launches (the kernel only writes
  // its block partials, so every launch gives the same values)
  const int reps = ev_after ? kTimedLaunches : 1;
  if (gw == 1) {  // one lane per one-voxel supervoxel (voxel mode)
    for (int r = 0; r < reps; r++) k_loglik_voxel<KM><<<nb(P.Sext, 256), 256, 0, st>>>(P, O);
    if (ev_after) (void)hipEventRecord(ev_after, st);
    k_ll_final<1><<<P.nframes, 64, 0, st>>>(P);
    return;
  }
  // ll_group 16 and 32: 2 * ll_group supervoxels per block of 256 threads
  // (32 for K > 8, whose constants take twice the LDS); 64: 16 supervoxels
  // per block of 128 threads (more blocks in flight). ll_unroll voxels per
  // thread in flight.
  const int B = gw == 64 ? 16 : std::min(2 * gw, KM > 8 ? 32 : 64);
  const unsigned g = nb(P.Sext, B);
#define TGMM_LL(B_, UV_, NT_)                                                   \
  if (B == B_ && spg == UV_ && (NT_ == 128) == (gw == 64)) {                    \
    for (int r = 0; r < reps; r++) k_loglik<KM, B_, UV_, NT_><<<g, NT_, 0, st>>>(P, O); \
    if (ev_after) (void)hipEventRecord(ev_after, st);                           \
    if (nb(P.Sext, 256) <= (unsigned)kLLFrameGroups) {                         \
      k_ll_frame<256 / B_><<<P.nframes, 1024, 0, st>>>(P);                      \
    } else {                                                                    \
      k_ll_groups<256 / B_><<<nb(nb(P.Sext, 256), 256), 256, 0, st>>>(P);      \
      k_ll_final<256 / B_><<<P.nframes, 64, 0, st>>>(P);                        \
    }                                                                           \
    return;                                                                     \
  }
  TGMM_LL(32, 1, 256)
  TGMM_LL(32, 2, 256)
  TGMM_LL(32, 4, 256)
  TGMM_LL(32, 8, 256)
  if constexpr (KM <= 8) {
    TGMM_LL(64, 1, 256)
    TGMM_LL(64, 2, 256)
    TGMM_LL(64, 4, 256)
    TGMM_LL(16, 1, 128)
    TGMM_LL(16, 2, 128)
    TGMM_LL(16, 4, 128)
  } else {
    TGMM_LL(16, 1, 128)
    TGMM_LL(16, 2, 128)
    TGMM_LL(16, 4, 128)
  }
#undef TGMM_LL
}

static void launch_mstep(const DevPtrs& P, const KOpts& O, hipStream_t st) {
  const unsigned g = nb(P.Gext, 64);
  if (O.tiles.ms_lanes == 4)
    k_mstep<4><<<g, 256, 0, st>>>(P, O);
  else if (O.tiles.ms_lanes == 8)
    k_mstep<8><<<g, 512, 0, st>>>(P, O);
  else
    k_mstep<16><<<g, 1024, 0, st>>>(P, O);
}

template <int KM>
static void launch_knn(const DevPtrs& P, const KOpts& O, hipStream_t st) {
  if (O.tiles.knn_lanes == 0) {  // one block of knn_block/64 waves per grid cell
    const unsigned g = (unsigned)P.Cext;
    if (O.tiles.knn_block == 64)
      k_knn_cell<KM, 1><<<g, 64, 0, st>>>(P, O);
    else if (O.tiles.knn_block == 128)
      k_knn_cell<KM, 2><<<g, 128, 0, st>>>(P, O);
    else
      k_knn_cell<KM, 4><<<g, 256, 0, st>>>(P, O);
    return;
  }
  const int bk = O.tiles.knn_block, ql = O.tiles.knn_lanes;
  const unsigned g = nb(P.Sext, bk / ql);
  const size_t lds = (size_t)bk * kKnnQueue * sizeof(float2);
  if (O.tiles.knn_chunk == 4) {
    if (ql == 1) k_knn<KM, 4, 1><<<g, bk, lds, st>>>(P, O);
    else if (ql == 2) k_knn<KM, 4, 2><<<g, bk, lds, st>>>(P, O);
    else k_knn<KM, 4, 4><<<g, bk, lds, st>>>(P, O);
  } else {
    if (ql == 1) k_knn<KM, 8, 1><<<g, bk, lds, st>>>(P, O);
    else if (ql == 2) k_knn<KM, 8, 2><<<g, bk, lds, st>>>(P, O);
    else k_knn<KM, 8, 4><<<g, bk, lds, st>>>(P, O);
  }
}

template <int KM>
static void launch_knn_layout_t(const DevPtrs& P, const KOpts& O, int only_frame, int max_ng,
                                hipStream_t st) {
  const int words = (int)std::min<long long>(kLayoutWords, ((long long)max_ng + 63) / 64);
  k_knn_layout<KM><<<nb(P.Sext, 64), 64, scan_lds_bytes(words), st>>>(P, O, only_frame,
                                                                               words);
}

hipError_t launch_knn_layout(const DevPtrs& P, const KOpts& O, int only_frame, int max_ng,
                             hipStream_t st) {
  switch (kmax_for(O.K)) {
    case 5:
      launch_knn_layout_t<5>(P, O, only_frame, max_ng, st);
      break;
    case 8:
      launch_knn_layout_t<8>(P, O, only_frame, max_ng, st);
      break;
    default:
      launch_knn_layout_t<16>(P, O, only_frame, max_ng, st);
  }
  return hipGetLastError();
}

hipError_t launch_finalize(const DevPtrs& P, const KOpts& O, int cap, int only_frame,
                           int writeback, hipStream_t st) {
  k_finalize<<<nb(P.Gext, 64), 64, 0, st>>>(P, O, cap, only_frame, writeback);
  return hipGetLastError();
}

template <int KM>
static hipError_t iteration_t(const DevPtrs& P, const KOpts& O, hipStream_t st, hipEvent_t e0,
                              hipEvent_t e1, int tk) {
  const unsigned gb = nb(P.Gext, kGB);
  const int bs = O.tiles.sv_block;
  k_scan_frames<<<P.nframes, 1024, 0, st>>>(P, 0, 1, O.K);
  k_grid_scatter<<<gb, kGB, 0, st>>>(P);
  if (tk == TK_KNN && e0) (void)hipEventRecord(e0, st);
  launch_knn<KM>(P, O, st);
  if (tk == TK_KNN && e1) (void)hipEventRecord(e1, st);
  {
    const int words = (int)std::min<long long>(kLayoutWords, (P.Gext + 63) / 64);
    k_knn_exact<KM><<<128, 64, scan_lds_bytes(words), st>>>(P, O, words);
  }
  if (tk == TK_ESTEP && e0) (void)hipEventRecord(e0, st);
  k_estep<KM><<<nb(P.Sext, bs), bs, 0, st>>>(P, O);
  if (tk == TK_ESTEP && e1) (void)hipEventRecord(e1, st);
  k_scan_frames<<<P.nframes, 1024, 0, st>>>(P, 1, 1, O.K);
  k_pair_scatter<KM><<<nb(P.Sext, bs), bs, 0, st>>>(P, O);
  k_seg_sort<<<nb(P.Gext, 16) + kSortLongBlocks, 256, 0, st>>>(P, (int)nb(P.Gext, 16));
  if (tk == TK_MSTEP && e0) (void)hipEventRecord(e0, st);
  launch_mstep(P, O, st);
  if (tk == TK_MSTEP && e1) (void)hipEventRecord(e1, st);
  k_death<<<gb, kGB, 0, st>>>(P, O);
  k_alpha_sum<<<P.nframes, 64, 0, st>>>(P, 1);
  if (tk == TK_LOGLIK && e0) (void)hipEventRecord(e0, st);
  launch_loglik<KM>(P, O, st, tk == TK_LOGLIK ? e1 : nullptr);
  return hipGetLastError();
}

hipError_t launch_iteration(const DevPtrs& P, const KOpts& O, hipStream_t st, hipEvent_t e0,
                            hipEvent_t e1, int tk) {
  if (!tiles_valid(O.tiles) || (O.tiles.ll_gw == 1 && !O.voxel)) return hipErrorInvalidValue;
  switch (kmax_for(O.K)) {
    case 5:
      return iteration_t<5>(P, O, st, e0, e1, tk);
    case 8:
      return iteration_t<8>(P, O, st, e0, e1, tk);
    default:
      return iteration_t<16>(P, O, st, e0, e1, tk);
  }
}

}  // namespace tgmm
