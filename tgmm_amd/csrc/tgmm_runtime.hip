// tgmm_runtime.hip -- host runtime of libtgmm_em.so: device workspaces for
// batches of frames, hipGraph-replayed EM iterations, the drop-in entry
// points of the reference host API, and the C-ABI error convention.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tgmm_em.h"
#include "tgmm_internal.h"
#include "tgmm_runtime_internal.h"

using namespace tgmm;

// ------------------------------------------------------------ error state --
static thread_local std::string g_err = "";

void tgmm_set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  if (getenv("TGMM_VERBOSE")) fprintf(stderr, "tgmm: %s\n", buf);
}

extern "C" const char* tgmm_last_error(void) { return g_err.c_str(); }

#define HIPCHK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      tgmm_set_error("%s failed: %s", #x, hipGetErrorString(e_));            \
      return e_ == hipErrorOutOfMemory ? TGMM_ERR_OOM : TGMM_ERR_DEVICE;     \
    }                                                                        \
  } while (0)

extern "C" int tgmm_device_count(int* count) {
  if (!count) return TGMM_ERR_ARG;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = (e == hipSuccess) ? n : 0;
  return TGMM_OK;
}

// -------------------------------------------------------------- workspace --
static long long round_up(long long x, long long a) { return ((x + a - 1) / a) * a; }

int ws_alloc(Workspace* w, const WsShape& sh, int K) {
  w->sh = sh;
  w->K = K;
  DevPtrs& P = w->P;
  std::memset(&P, 0, sizeof(P));
  P.Scap = sh.Scap;
  P.Gcap = sh.Gcap;
  P.Ccap = sh.Ccap;
  P.Sext = sh.Scap;
  P.Gext = sh.Gcap;
  P.Cext = sh.Ccap;
  P.Vcap = sh.Vcap;
  P.Fcap = sh.Fcap;
  auto A = [&](void** p, size_t bytes) -> int {
    bytes = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
      tgmm_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
      return TGMM_ERR_OOM;
    }
    w->allocs.push_back(*p);
    w->bytes += bytes;
    return TGMM_OK;
  };
#define AL(ptr, n) \
  if (int rc_ = A((void**)&(ptr), (size_t)(n) * sizeof(*(ptr)))) return rc_
  const long long S = sh.Scap, G = sh.Gcap, C = sh.Ccap, V = sh.Vcap, F = sh.Fcap;
  AL(P.vx, V);
  AL(P.vy, V);
  AL(P.vz, V);
  AL(P.vimg, V);
  AL(P.vpk, V);
  AL(P.sv_frame, S);
  AL(P.sv_beg, S);
  AL(P.sv_len, S);
  AL(P.sv_cent, 3 * S);
  AL(P.mom, 10 * S);
  AL(P.wmom, 10 * S);
  AL(P.anchor, 3 * S);
  AL(P.sv_ball, S);
  AL(P.ind, (long long)K * S);
  AL(P.rnk, (long long)K * S);
  AL(P.sblk_frame, S / kAlign);
  AL(P.q4, S);
  AL(P.qcell, S);
  AL(P.g_frame, G);
  AL(P.gs, GS_NFIELDS * G);
  AL(P.fixed, G);
  AL(P.epack, G);
  AL(P.eldn, G);
  AL(P.llpack, G);
  AL(P.m4, G);
  AL(P.gcell, G);
  AL(P.gblk_frame, G / kGB);
  AL(P.seg_cnt, G + 1);
  AL(P.seg_start, G + 1);
  AL(P.pairs, (long long)K * S);
  AL(P.pairs_sorted, (long long)K * S);
  AL(P.cell_cnt, C + 1);
  AL(P.cell_start, C + 1);
  AL(P.qcell_start, C + 1);
  AL(P.cblk_frame, C / kCB + 1);
  AL(P.cg4, G);
  AL(P.mknn, G);
  AL(P.tie_list, S);
  AL(P.knn_ctl, 64);
  AL(P.long_list, G);
  AL(P.win, S);
  AL(P.rec, G);
  AL(P.svkeep, 17 * G);
  AL(P.part_g, G / kGB);
  AL(P.part_g2, G / kGB);
  AL(P.part_s, S / kLLminB);
  AL(P.part_sg, S / 256);
  AL(P.fr, F);
  AL(w->gs_init, GS_NFIELDS * G);
#undef AL
  return TGMM_OK;
}

void ws_free(Workspace* w) {
  for (void* p : w->allocs) (void)hipFree(p);
  w->allocs.clear();
  w->bytes = 0;
}

long long ws_bytes_estimate(const WsShape& sh, int K) {
  const long long S = sh.Scap, G = sh.Gcap, C = sh.Ccap, V = sh.Vcap;
  return 24 * V + S * (16 + 8 + 4 + 12 + 80 + 80 + 12 + 16 + 16LL * K + 1 + 4 + 8) +
         G * (4 + 8 * GS_NFIELDS * 2 + 1 + 96 + 8 + 64 + 12 + 4 + 8 + 12 + 4 + 1 + 16 + 288 +
              68 + 4) +
         8 * C + 64 * sh.Fcap;
}

// ------------------------------------------------------------------ batch --
static int validate_opts(const tgmm_em_opts* o) {
  if (!o) {
    tgmm_set_error("opts is NULL");
    return TGMM_ERR_ARG;
  }
  if (o->K < 1 || o->K > TGMM_KMAX) {
    tgmm_set_error("K=%d outside [1, %d]", o->K, TGMM_KMAX);
    return TGMM_ERR_ARG;
  }
  if (o->max_iter < 0) {
    tgmm_set_error("max_iter < 0");
    return TGMM_ERR_ARG;
  }
  if (o->lambda_min < 0 || o->lambda_max < 0 || o->max_excentricity < 0) {
    tgmm_set_error("regularisation constants were not set (negative)");
    return TGMM_ERR_NOT_SET;
  }
  for (int a = 0; a < 3; a++)
    if (!(o->scale[a] > 1e-3f)) {
      tgmm_set_error("scale[%d]=%g must be > 1e-3", a, (double)o->scale[a]);
      return TGMM_ERR_ARG;
    }
  if (o->max_sv_per_gaussian < 0 || o->max_sv_per_gaussian > TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN) {
    tgmm_set_error("max_sv_per_gaussian outside [0, 16]");
    return TGMM_ERR_ARG;
  }
  if (o->voxel_mode != 0 && o->voxel_mode != 1) {
    tgmm_set_error("voxel_mode must be 0 or 1");
    return TGMM_ERR_ARG;
  }
  return TGMM_OK;
}

KOpts make_kopts(const tgmm_em_opts& o, const float* scale_reg) {
  KOpts k;
  std::memset(&k, 0, sizeof(k));
  k.K = o.K;
  k.w4dof = o.w4dof ? 1 : 0;
  for (int a = 0; a < 3; a++) {
    k.scale[a] = o.scale[a];
    k.scale_reg[a] = scale_reg ? scale_reg[a] : o.scale[a];
  }
  k.lmin = o.lambda_min;
  k.lmax = o.lambda_max;
  k.emax = o.max_excentricity;
  k.tiles = kDefaultTiles;
  k.voxel = o.voxel_mode ? 1 : 0;
  if (k.voxel) {  // the log-likelihood's lane-per-voxel form
    k.tiles.ll_gw = 1;
    k.tiles.ll_u = 1;
  }
  return k;
}

bool tgmm::tiles_valid(const Tiles& t) {
  auto in = [](int v, int a, int b, int c) { return v == a || v == b || v == c; };
  return ((in(t.ll_gw, 16, 32, 64) && (in(t.ll_u, 1, 2, 4) || (t.ll_u == 8 && t.ll_gw == 16))) ||
          (t.ll_gw == 1 && t.ll_u == 1)) &&
         in(t.ms_lanes, 4, 8, 16) &&
         in(t.sv_block, 64, 128, 256) && in(t.knn_block, 64, 128, 256) &&
         (t.knn_chunk == 4 || t.knn_chunk == 8) && in(t.knn_cell, 2, 4, 6) &&
         (t.knn_lanes == 0 || in(t.knn_lanes, 1, 2, 4));
}

// records (AoS, host) -> 27 x G SoA columns at [g0, g0+G) of a Gcap table,
// the fixed flags, and (if svkeep) each record's supervoxelIdx[16] and
// supervoxelNum as 17 ints at 17 * (g0 + k)
void records_to_soa(const GaussianMixtureModelCUDA* gm, long long G, long long g0, long long Gcap,
                    double* gs, unsigned char* fixed, int* svkeep) {
  for (long long k = 0; k < G; k++) {
    const double* d = (const double*)&gm[k];
    for (int fidx = 0; fidx < GS_NFIELDS; fidx++) gs[fidx * Gcap + g0 + k] = d[fidx];
    fixed[g0 + k] = gm[k].fixed ? 1 : 0;
    if (svkeep) {
      int* o = svkeep + 17 * (g0 + k);
      std::memcpy(o, gm[k].supervoxelIdx, sizeof(int) * TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN);
      o[16] = gm[k].supervoxelNum;
    }
  }
}

void soa_to_records(const double* gs, long long G, long long g0, long long Gcap,
                    GaussianMixtureModelCUDA* gm) {
  for (long long k = 0; k < G; k++) {
    double* d = (double*)&gm[k];
    for (int fidx = 0; fidx < GS_NFIELDS; fidx++) d[fidx] = gs[fidx * Gcap + g0 + k];
  }
}

int batch_build(int device, const std::vector<FrameSrc>& fs_in, const tgmm_em_opts& o,
                const float* scale_reg, hipStream_t stream, tgmm_batch** out,
                tgmm_batch* reuse, double headroom) {
  if (int rc = validate_opts(&o)) return rc;
  HIPCHK(hipSetDevice(device));
  // voxel mode (GMEMupdateCUDA.cu:3420-3973): every voxel is its own kNN
  // query and a one-voxel supervoxel whose centroid is the voxel itself
  std::vector<FrameSrc> vfs;
  std::vector<std::vector<long long>> vptr;
  if (o.voxel_mode) {
    vfs = fs_in;
    vptr.resize(fs_in.size());
    for (size_t i = 0; i < fs_in.size(); i++) {
      if (fs_in[i].N < 0) {
        tgmm_set_error("frame %zu: negative size", i);
        return TGMM_ERR_ARG;
      }
      vptr[i].resize(fs_in[i].N + 1);
      for (long long v = 0; v <= fs_in[i].N; v++) vptr[i][v] = v;
      vfs[i].S = fs_in[i].N;
      vfs[i].cent = fs_in[i].query;
      vfs[i].ptr = vptr[i].data();
    }
  }
  const std::vector<FrameSrc>& fs = o.voxel_mode ? vfs : fs_in;
  // placement
  std::vector<FrameDev> fh(fs.size());
  WsShape sh;
  long long v = 0, s = 0, g = 0, c = 0;
  for (size_t i = 0; i < fs.size(); i++) {
    const FrameSrc& f = fs[i];
    if (f.N < 0 || f.S < 0 || f.G < 0) {
      tgmm_set_error("frame %zu: negative size", i);
      return TGMM_ERR_ARG;
    }
    FrameDev& F = fh[i];
    std::memset(&F, 0, sizeof(F));
    F.v0 = v;
    F.nv = f.N;
    F.s0 = s;
    F.ns = f.S;
    F.g0 = g;
    F.ng = f.G;
    F.c0 = c;
    // grid-cell range: room for the densest grid (two means per cell) plus
    // one entry past the last cell (the per-frame scan writes it), a multiple
    // of 64 (cell block -> frame table)
    F.ncap = round_up(f.G / 2 + kCB, kCB);
    // every block mapped to the frame's 256-aligned ranges takes part in the
    // per-frame reductions, including trailing blocks with no valid element;
    // the Gaussian range has at least one padding entry (segment end)
    F.nbg = (int)(round_up(f.G + 1, kAlign) / kGB);
    F.nbl = (int)(round_up(std::max<long long>(f.S, 1), kAlign) / 256);
    F.max_iter = o.max_iter;
    F.tol = o.tol;
    F.packed = 1;
    v += f.N;
    s += round_up(std::max<long long>(f.S, 1), kAlign);
    g += round_up(f.G + 1, kAlign);
    c += F.ncap;
  }
  sh.Vcap = std::max<long long>(v, 1);
  sh.Scap = std::max<long long>(s, kAlign);
  sh.Gcap = std::max<long long>(g, kAlign);
  sh.Ccap = std::max<long long>(c, kCB);
  sh.Fcap = std::max<int>((int)fs.size(), 1);
  const WsShape need = sh;  // the extents the kernels cover
  if ((sh.Scap << kPairShift) >= (1LL << 31) || (long long)o.K * sh.Scap >= (1LL << 31) ||
      sh.Gcap >= (1LL << 31) || sh.Ccap >= (1LL << 31)) {
    tgmm_set_error("workspace exceeds 32-bit index limits (S=%lld, G=%lld)", sh.Scap, sh.Gcap);
    return TGMM_ERR_LIMIT;
  }
  tgmm_batch* b = nullptr;
  if (reuse) {
    const WsShape& r = reuse->ws.sh;
    if (reuse->device == device && reuse->ws.K == o.K && r.Vcap >= sh.Vcap &&
        r.Scap >= sh.Scap && r.Gcap >= sh.Gcap && r.Ccap >= sh.Ccap && r.Fcap >= sh.Fcap) {
      b = reuse;  // the workspace fits: keep its allocations, stream and graph
    } else {
      tgmm_batch_destroy(reuse);
    }
  }
  if (!b) {
    b = new tgmm_batch();
    b->device = device;
    if (headroom > 1.0) {  // grow-only caches leave room for larger frames
      auto up = [&](long long x, long long a) {
        return round_up((long long)std::ceil((double)x * headroom), a);
      };
      sh.Vcap = up(sh.Vcap, 1);
      sh.Scap = up(sh.Scap, kAlign);
      sh.Gcap = up(sh.Gcap, kAlign);
      sh.Ccap = up(sh.Ccap, kCB);
      if ((sh.Scap << kPairShift) >= (1LL << 31) || (long long)o.K * sh.Scap >= (1LL << 31) ||
          sh.Gcap >= (1LL << 31) || sh.Ccap >= (1LL << 31))
        sh = need;
    }
    if (int rc = ws_alloc(&b->ws, sh, o.K)) {
      ws_free(&b->ws);
      delete b;
      return rc;
    }
    if (stream) {
      b->st = stream;
      b->own_stream = false;
    } else {
      if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
        ws_free(&b->ws);
        delete b;
        tgmm_set_error("hipStreamCreate failed");
        return TGMM_ERR_DEVICE;
      }
      b->own_stream = true;
    }
  }
  b->opts = o;
  b->ko = make_kopts(o, scale_reg);
  b->nframes = (int)fs.size();
  b->fh = fh;
  b->fh.resize(b->ws.sh.Fcap);  // unused slots stay zero: no supervoxels, no Gaussians
  for (size_t i = fs.size(); i < b->fh.size(); i++) std::memset(&b->fh[i], 0, sizeof(FrameDev));
  DevPtrs& P = b->ws.P;
  P.nframes = (int)fs.size();
  P.only_frame = -1;
  P.s_lo = P.g_lo = 0;
  P.Sext = need.Scap;
  P.Gext = need.Gcap;
  P.Cext = need.Ccap;
  int rc = batch_upload(b, fs);
  if (rc) {
    tgmm_batch_destroy(b);
    return rc;
  }
  *out = b;
  return TGMM_OK;
}

int batch_upload(tgmm_batch* b, const std::vector<FrameSrc>& fs) {
  DevPtrs& P = b->ws.P;
  const WsShape& sh = b->ws.sh;
  hipStream_t st = b->st;
  const long long S = sh.Scap, G = sh.Gcap;
  std::vector<int> sv_frame(S, -1), sv_len(S, 0), g_frame(G, -1);
  std::vector<long long> sv_beg(S, 0);
  std::vector<int> sblk(S / kAlign, -1), gblk(G / kGB, -1), cblk(b->ws.sh.Ccap / kCB + 1, -1);
  std::vector<double> gs((size_t)GS_NFIELDS * G, 0.0);
  std::vector<unsigned char> fixed(G, 0);
  std::vector<int> svkeep((size_t)17 * G, 0);
  for (size_t i = 0; i < fs.size(); i++) {
    const FrameSrc& f = fs[i];
    const FrameDev& F = b->fh[i];
    for (long long j = 0; j < f.S; j++) {
      const long long a = f.ptr[j], e = f.ptr[j + 1];
      if (a < 0 || e < a || e > f.N) {
        tgmm_set_error("frame %zu: supervoxel pointer array is not monotone in [0, N]", i);
        return TGMM_ERR_ARG;
      }
      if (e - a > 0x7fffffffLL) {
        tgmm_set_error("frame %zu: supervoxel with more than 2^31 voxels", i);
        return TGMM_ERR_LIMIT;
      }
      sv_frame[F.s0 + j] = (int)i;
      sv_beg[F.s0 + j] = F.v0 + a;
      sv_len[F.s0 + j] = (int)(e - a);
    }
    for (long long j = 0; j < round_up(std::max<long long>(f.S, 1), kAlign) / kAlign; j++)
      sblk[F.s0 / kAlign + j] = (int)i;
    for (long long j = 0; j < F.nbg; j++) gblk[F.g0 / kGB + j] = (int)i;
    for (long long j = 0; j < F.ncap / kCB; j++) cblk[F.c0 / kCB + j] = (int)i;
    for (long long k = 0; k < f.G; k++) g_frame[F.g0 + k] = (int)i;
    if (f.G > 0) records_to_soa(f.gm, f.G, F.g0, G, gs.data(), fixed.data(), svkeep.data());
  }
  const hipMemcpyKind kv = hipMemcpyDefault;
  for (size_t i = 0; i < fs.size(); i++) {
    const FrameSrc& f = fs[i];
    const FrameDev& F = b->fh[i];
    if (f.N > 0) {
      HIPCHK(hipMemcpyAsync(P.vx + F.v0, f.query, f.N * 4, kv, st));
      HIPCHK(hipMemcpyAsync(P.vy + F.v0, f.query + f.N, f.N * 4, kv, st));
      HIPCHK(hipMemcpyAsync(P.vz + F.v0, f.query + 2 * f.N, f.N * 4, kv, st));
      HIPCHK(hipMemcpyAsync(P.vimg + F.v0, f.img, f.N * 4, kv, st));
    }
    if (f.S > 0)
      for (int a = 0; a < 3; a++)
        HIPCHK(hipMemcpyAsync(P.sv_cent + a * S + F.s0, f.cent + a * f.S, f.S * 4, kv, st));
  }
  HIPCHK(hipMemcpyAsync(P.sv_frame, sv_frame.data(), S * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.sv_beg, sv_beg.data(), S * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.sv_len, sv_len.data(), S * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.sblk_frame, sblk.data(), sblk.size() * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.cblk_frame, cblk.data(), cblk.size() * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.g_frame, g_frame.data(), G * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.gblk_frame, gblk.data(), gblk.size() * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.gs, gs.data(), gs.size() * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(b->ws.gs_init, gs.data(), gs.size() * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.fixed, fixed.data(), G, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.svkeep, svkeep.data(), svkeep.size() * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(P.knn_ctl, 0, 64 * 4, st));
  HIPCHK(hipMemcpyAsync(P.fr, b->fh.data(), b->fh.size() * sizeof(FrameDev), hipMemcpyHostToDevice,
                        st));
  HIPCHK(hipMemsetAsync(P.ind, 0xFF, (size_t)b->ws.K * S * 4, st));
  HIPCHK(hipMemsetAsync(P.rnk, 0, (size_t)b->ws.K * S * 4, st));
  HIPCHK(hipMemsetAsync(P.seg_cnt, 0, (G + 1) * 4, st));
  HIPCHK(hipMemsetAsync(P.cell_cnt, 0, (sh.Ccap + 1) * 4, st));
  HIPCHK(launch_prep(P, b->ko, st));
  b->grid_cell = b->ko.tiles.knn_cell;
  HIPCHK(launch_reset(P, b->ko, st));
  HIPCHK(hipStreamSynchronize(st));
  b->sizes.clear();
  b->max_ng = 0;
  for (const FrameSrc& f : fs) {
    b->sizes.push_back({f.N, f.S, (long long)f.G});
    b->max_ng = std::max<long long>(b->max_ng, f.G);
  }
  return TGMM_OK;
}

extern "C" int tgmm_batch_create(int device, const tgmm_frame_desc* frames, int nframes,
                                 const tgmm_em_opts* opts, tgmm_batch** out) {
  if (!frames || nframes < 1 || !out) {
    tgmm_set_error("tgmm_batch_create: bad arguments");
    return TGMM_ERR_ARG;
  }
  if (int rc = validate_opts(opts)) return rc;
  const bool vox = opts->voxel_mode != 0;  // supervoxel arrays are not used
  std::vector<FrameSrc> fs(nframes);
  for (int i = 0; i < nframes; i++) {
    const tgmm_frame_desc& d = frames[i];
    if (d.n_voxels < 0 || d.n_supervoxels < 0 || d.n_gaussians < 0 ||
        (d.n_voxels > 0 && (!d.query || !d.img)) ||
        (!vox && ((d.n_supervoxels > 0 && !d.centroids) || !d.sv_ptr)) ||
        (d.n_gaussians > 0 && !d.gm)) {
      tgmm_set_error("frame %d: NULL array or negative size", i);
      return TGMM_ERR_ARG;
    }
    FrameSrc& f = fs[i];
    f.N = d.n_voxels;
    f.S = d.n_supervoxels;
    f.G = d.n_gaussians;
    f.query = d.query;
    f.img = d.img;
    f.cent = d.centroids;
    f.ptr = d.sv_ptr;
    f.gm = d.gm;
  }
  return batch_build(device, fs, *opts, nullptr, nullptr, out);
}

extern "C" int tgmm_batch_reset(tgmm_batch* b) {
  if (!b) return TGMM_ERR_ARG;
  HIPCHK(hipSetDevice(b->device));
  DevPtrs& P = b->ws.P;
  HIPCHK(hipMemcpyAsync(P.gs, b->ws.gs_init, (size_t)GS_NFIELDS * b->ws.sh.Gcap * 8,
                        hipMemcpyDeviceToDevice, b->st));
  // grid counts left by frames that stopped after their last binning
  HIPCHK(hipMemsetAsync(P.cell_cnt, 0, (b->ws.sh.Ccap + 1) * 4, b->st));
  HIPCHK(launch_reset(P, b->ko, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  return TGMM_OK;
}

static int any_active(tgmm_batch* b, bool* act) {
  DevPtrs& P = b->ws.P;
  HIPCHK(hipMemcpyAsync(b->fh.data(), P.fr, b->fh.size() * sizeof(FrameDev),
                        hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  *act = false;
  for (const FrameDev& F : b->fh) *act |= F.active != 0;
  return TGMM_OK;
}

static void destroy_graphs(tgmm_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->gexec_n) (void)hipGraphExecDestroy(b->gexec_n);
  if (b->graph_n) (void)hipGraphDestroy(b->graph_n);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->gexec_n = nullptr;
  b->graph_n = nullptr;
  b->cap_n = 0;
}

// One capture of n iterations on the batch's stream.
static int capture_iterations(tgmm_batch* b, int n, hipGraph_t* graph, hipGraphExec_t* exec) {
  HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int i = 0; i < n && le == hipSuccess; i++) le = launch_iteration(b->ws.P, b->ko, b->st);
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(b->st, &g);
  if (le != hipSuccess || ee != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    tgmm_set_error("graph capture failed: %s / %s", hipGetErrorString(le), hipGetErrorString(ee));
    return TGMM_ERR_DEVICE;
  }
  *graph = g;
  HIPCHK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return TGMM_OK;
}

// The solve's graphs: one iteration, and n iterations back to back (n >= 2),
// so that a solve launches few graphs (each launch boundary idles the GPU
// for several microseconds).
int batch_capture(tgmm_batch* b, int n) {
  // a captured iteration holds its kernel arguments: capture again when they
  // changed (a cached workspace reused for frames of another size)
  if (b->gexec && (std::memcmp(&b->cap_P, &b->ws.P, sizeof(DevPtrs)) != 0 ||
                   std::memcmp(&b->cap_ko, &b->ko, sizeof(KOpts)) != 0 ||
                   b->cap_n != (n >= 2 ? n : 0)))
    destroy_graphs(b);
  if (b->gexec) return TGMM_OK;
  std::memcpy(&b->cap_P, &b->ws.P, sizeof(DevPtrs));  // padding included, as compared
  std::memcpy(&b->cap_ko, &b->ko, sizeof(KOpts));
  if (int rc = capture_iterations(b, 1, &b->graph, &b->gexec)) {
    destroy_graphs(b);
    return rc;
  }
  if (n >= 2) {
    if (int rc = capture_iterations(b, n, &b->graph_n, &b->gexec_n)) {
      destroy_graphs(b);
      return rc;
    }
    b->cap_n = n;
  }
  return TGMM_OK;
}

extern "C" int tgmm_batch_solve(tgmm_batch* b, int mode) {
  if (!b) return TGMM_ERR_ARG;
  HIPCHK(hipSetDevice(b->device));
  const int iters = b->opts.max_iter;
  const bool poll = b->opts.tol >= 0.0;
  // with a tolerance, the host looks at the frames' states every kPollEvery
  // iterations (frames that converged have stopped on the device already)
  if (mode == TGMM_SOLVE_GRAPH) {
    if (int rc = batch_capture(b, poll ? kPollEvery : std::min(iters, kGraphIters))) return rc;
  }
  int it = 0;  // iterations launched
  while (it < iters) {
    if (mode == TGMM_SOLVE_GRAPH) {
      if (b->gexec_n && iters - it >= b->cap_n) {
        HIPCHK(hipGraphLaunch(b->gexec_n, b->st));
        it += b->cap_n;
      } else {
        HIPCHK(hipGraphLaunch(b->gexec, b->st));
        it++;
      }
    } else {
      HIPCHK(launch_iteration(b->ws.P, b->ko, b->st));
      it++;
    }
    const bool check = (mode == TGMM_SOLVE_HOSTSYNC) || (poll && it % kPollEvery == 0);
    if (check && it < iters) {
      bool act = true;
      if (int rc = any_active(b, &act)) return rc;
      if (!act) break;
    }
  }
  // the last search's candidates in the reference's slot layout
  if (iters > 0) HIPCHK(launch_knn_layout(b->ws.P, b->ko, -1, (int)b->max_ng, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  return TGMM_OK;
}

// Results of every frame: the records are formed on the device (state, fixed
// flag, r > 0.5 supervoxel assignment) and copied into each frame's gm in one
// transfer; rnk/ind (optional, host) are copied slot by slot, ind made
// frame-local; ll and the iteration count come from the frame table.
extern "C" int tgmm_batch_fetch(tgmm_batch* b, tgmm_frame_desc* frames, int nframes) {
  if (!b || !frames || nframes != b->nframes) {
    tgmm_set_error("tgmm_batch_fetch: bad arguments");
    return TGMM_ERR_ARG;
  }
  HIPCHK(hipSetDevice(b->device));
  DevPtrs& P = b->ws.P;
  const long long S = b->ws.sh.Scap;
  const int K = b->ws.K;
  HIPCHK(launch_finalize(P, b->ko, b->opts.max_sv_per_gaussian, -1, 0, b->st));
  for (int i = 0; i < nframes; i++) {
    const FrameDev& F = b->fh[i];
    tgmm_frame_desc& d = frames[i];
    if (d.gm && F.ng > 0)
      HIPCHK(hipMemcpyAsync(d.gm, P.rec + F.g0, F.ng * sizeof(GaussianMixtureModelCUDA),
                            hipMemcpyDeviceToHost, b->st));
    const size_t w = (size_t)F.ns * 4;
    if (d.rnk && F.ns > 0)
      HIPCHK(hipMemcpy2DAsync(d.rnk, w, P.rnk + F.s0, S * 4, w, K, hipMemcpyDeviceToHost, b->st));
    if (d.ind && F.ns > 0)
      HIPCHK(hipMemcpy2DAsync(d.ind, w, P.ind + F.s0, S * 4, w, K, hipMemcpyDeviceToHost, b->st));
  }
  HIPCHK(hipMemcpyAsync(b->fh.data(), P.fr, b->fh.size() * sizeof(FrameDev),
                        hipMemcpyDeviceToHost, b->st));
  int ctl[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(ctl, P.knn_ctl, sizeof(ctl), hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  for (int i = 0; i < nframes; i++) {
    const FrameDev& F = b->fh[i];
    tgmm_frame_desc& d = frames[i];
    if (d.ind && F.g0 > 0)
      for (long long e = 0; e < (long long)K * F.ns; e++)
        if (d.ind[e] >= 0) d.ind[e] -= (int)F.g0;
    d.ll = F.ll;
    d.iterations = F.iter - 1;
    d.status = TGMM_OK;
  }
  b->layout_mismatch = ctl[1];
  return TGMM_OK;
}

extern "C" int tgmm_batch_stats(tgmm_batch* b, tgmm_batch_stats_t* st) {
  if (!b || !st) return TGMM_ERR_ARG;
  HIPCHK(hipSetDevice(b->device));
  int ctl[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(ctl, b->ws.P.knn_ctl, sizeof(ctl), hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  st->knn_ties_last_iteration = ctl[0];
  st->layout_mismatches = ctl[1];
  st->workspace_bytes = b->ws.bytes;
  return TGMM_OK;
}

// Algorithmic bytes of one launch of a kernel over all frames of the batch
// (compulsory traffic; Gaussian-side gathers counted once per Gaussian).
static double kernel_bytes(tgmm_batch* b, int tk) {
  const int K = b->ws.K;
  double t = 0.0;
  // whether each frame's voxels are read packed (decided on the device)
  std::vector<FrameDev> fd(b->fh.size());
  if (!fd.empty() &&
      hipMemcpy(fd.data(), b->ws.P.fr, fd.size() * sizeof(FrameDev), hipMemcpyDeviceToHost) !=
          hipSuccess)
    fd.clear();
  for (size_t i = 0; i < b->sizes.size(); i++) {
    const auto& z = b->sizes[i];
    const double N = (double)z.N, S = (double)z.S, G = (double)z.G;
    const bool packed = i < fd.size() && fd[i].packed != 0 && !b->ko.voxel;
    switch (tk) {
      case TK_LOGLIK:
        // voxels (8 B packed, else 16 B); per supervoxel its frame, length,
        // first voxel, anchor, ball and K candidates; a 64-byte pack per
        // Gaussian
        t += (packed ? 8.0 : 16.0) * N + S * (4 + 4 + 8 + 12 + 16 + 4.0 * K) + G * 64;
        break;
      case TK_ESTEP:
        t += S * (4 + 80 + 4.0 * K + 4.0 * K) + G * (8 * 11);
        break;
      case TK_KNN:
        // per query its centroid and index (16 B), the previous candidates
        // (the bound) and the new ones (4 B each); per Gaussian its binned
        // reference and its mean for the bound (16 B each)
        t += S * (16 + 8.0 * K) + G * 32;
        break;
      case TK_MSTEP:
        t += G * (27 * 8 * 2 + 8) + S * (4 + 4 + 80);
        break;
    }
  }
  return t;
}

extern "C" int tgmm_batch_time_kernel(tgmm_batch* b, const char* name, int reps, double* avg_ms,
                                      double* bytes_per_launch) {
  if (!b || !name || reps < 1 || !avg_ms) return TGMM_ERR_ARG;
  int tk = -1;
  if (!strcmp(name, "loglik")) tk = TK_LOGLIK;
  else if (!strcmp(name, "estep")) tk = TK_ESTEP;
  else if (!strcmp(name, "knn")) tk = TK_KNN;
  else if (!strcmp(name, "mstep")) tk = TK_MSTEP;
  if (tk < 0) {
    tgmm_set_error("unknown kernel name '%s'", name);
    return TGMM_ERR_ARG;
  }
  if (b->opts.max_iter < 1) return TGMM_ERR_ARG;
  HIPCHK(hipSetDevice(b->device));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // eager iterations from the initial state; `reps` launches of the kernel
  double tot = 0.0;
  int it = b->opts.max_iter;
  for (int r = 0; r < reps; r++) {
    if (it == b->opts.max_iter) {
      if (int rc = tgmm_batch_reset(b)) return rc;
      it = 0;
    }
    HIPCHK(launch_iteration(b->ws.P, b->ko, b->st, e0, e1, tk));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    tot += ms;
    it++;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = tot / ((double)reps * timed_launches(tk));
  if (bytes_per_launch) *bytes_per_launch = kernel_bytes(b, tk);
  return TGMM_OK;
}

// Rebuild every frame's kNN grid for ko.tiles.knn_cell. The grid counts
// are cleared first (a frame's last binning may have left counts); the
// iteration state must be reset afterwards.
static int regrid(tgmm_batch* b, const KOpts& ko) {
  DevPtrs& P = b->ws.P;
  HIPCHK(hipMemsetAsync(P.cell_cnt, 0, (b->ws.sh.Ccap + 1) * 4, b->st));
  HIPCHK(launch_prep(P, ko, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  b->grid_cell = ko.tiles.knn_cell;
  return TGMM_OK;
}

static Tiles from_public(const tgmm_tiles& t) {
  return Tiles{t.ll_group,  t.ll_unroll, t.mstep_lanes, t.sv_block,
               t.knn_block, t.knn_chunk, t.knn_cell,    t.knn_lanes};
}
static tgmm_tiles to_public(const Tiles& t) {
  return tgmm_tiles{t.ll_gw,     t.ll_u,      t.ms_lanes, t.sv_block,
                    t.knn_block, t.knn_chunk, t.knn_cell, t.knn_lanes};
}

extern "C" int tgmm_batch_set_tiles(tgmm_batch* b, const tgmm_tiles* t) {
  if (!b || !t) return TGMM_ERR_ARG;
  const Tiles x = from_public(*t);
  if (!tiles_valid(x)) {
    tgmm_set_error("tgmm_batch_set_tiles: invalid tile shape {%d, %d, %d, %d, %d}", x.ll_gw,
                   x.ll_u, x.ms_lanes, x.sv_block, x.knn_block);
    return TGMM_ERR_ARG;
  }
  if (x.ll_gw == 1 && !b->ko.voxel) {
    tgmm_set_error("tgmm_batch_set_tiles: ll_group 1 (a lane per voxel) needs voxel mode");
    return TGMM_ERR_ARG;
  }
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->st));
  b->ko.tiles = x;
  destroy_graphs(b);  // recaptured by the next graph solve
  if (x.knn_cell != b->grid_cell) {  // the kNN grids are sized by knn_cell
    if (int rc = regrid(b, b->ko)) return rc;
    return tgmm_batch_reset(b);
  }
  return TGMM_OK;
}

extern "C" int tgmm_batch_get_tiles(tgmm_batch* b, tgmm_tiles* t) {
  if (!b || !t) return TGMM_ERR_ARG;
  *t = to_public(b->ko.tiles);
  return TGMM_OK;
}

// Device time of one launch of kernel `tk`, averaged over `iters` eager
// iterations from the initial state, with launch shapes `t`.
static int time_shape(tgmm_batch* b, const Tiles& t, int tk, int iters, hipEvent_t e0,
                      hipEvent_t e1, double* ms_out) {
  KOpts ko = b->ko;
  ko.tiles = t;
  if (t.knn_cell != b->grid_cell)  // rebuild the kNN grids for this density
    if (int rc = regrid(b, ko)) return rc;
  if (int rc = tgmm_batch_reset(b)) return rc;
  double tot = 0.0;
  for (int it = 0; it < iters; it++) {
    HIPCHK(launch_iteration(b->ws.P, ko, b->st, e0, e1, tk));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    tot += ms;
  }
  *ms_out = tot / ((double)iters * timed_launches(tk));
  return TGMM_OK;
}

extern "C" int tgmm_batch_autotune(tgmm_batch* b, int reps, tgmm_tiles* chosen, char* report,
                                   int report_cap) {
  if (!b || reps < 1) return TGMM_ERR_ARG;
  if (b->opts.max_iter < 1) {
    tgmm_set_error("tgmm_batch_autotune: the batch runs no iterations");
    return TGMM_ERR_ARG;
  }
  HIPCHK(hipSetDevice(b->device));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int iters = std::min(reps, b->opts.max_iter);
  Tiles best = b->ko.tiles;
  std::string rep;
  char line[160];
  struct Knob {
    const char* name;
    int tk;
    int n;
    int v[15][3];
  };
  const Knob knobs[5] = {
      {"loglik", TK_LOGLIK, 10, {{16, 1}, {16, 2}, {16, 4}, {16, 8}, {32, 1}, {32, 2}, {32, 4},
                                 {64, 1}, {64, 2}, {64, 4}}},
      {"mstep", TK_MSTEP, 3, {{4, 0}, {8, 0}, {16, 0}}},
      {"estep", TK_ESTEP, 3, {{64, 0}, {128, 0}, {256, 0}}},
      {"knn_cell", TK_KNN, 3, {{2, 0}, {4, 0}, {6, 0}}},
      {"knn", TK_KNN, 15, {{64, 8, 0},  {128, 8, 0}, {256, 8, 0}, {64, 4, 1},  {256, 4, 1},
                           {64, 8, 1},  {256, 8, 1}, {64, 4, 2},  {256, 4, 2}, {64, 8, 2},
                           {256, 8, 2}, {64, 4, 4},  {256, 4, 4}, {64, 8, 4},  {256, 8, 4}}},
  };
  int rc = TGMM_OK;
  for (int q = 0; q < 5 && rc == TGMM_OK; q++) {
    const Knob& kb = knobs[q];
    double tbest = 1e300;
    Tiles pick = best;
    const int c0 = (q == 0 && b->ko.voxel) ? -1 : 0;  // voxel mode: the lane-per-voxel form too
    for (int c = c0; c < kb.n; c++) {
      Tiles t = best;
      if (c < 0) {
        t.ll_gw = 1;
        t.ll_u = 1;
        double ms = 0.0;
        if ((rc = time_shape(b, t, kb.tk, iters, e0, e1, &ms)) != TGMM_OK) break;
        snprintf(line, sizeof(line), "%s 1x1 %.2f\n", kb.name, 1000.0 * ms);
        rep += line;
        if (ms < tbest) {
          tbest = ms;
          pick = t;
        }
        continue;
      }
      switch (q) {
        case 0: t.ll_gw = kb.v[c][0]; t.ll_u = kb.v[c][1]; break;
        case 1: t.ms_lanes = kb.v[c][0]; break;
        case 2: t.sv_block = kb.v[c][0]; break;
        case 3: t.knn_cell = kb.v[c][0]; break;
        default:
          t.knn_block = kb.v[c][0];
          t.knn_chunk = kb.v[c][1];
          t.knn_lanes = kb.v[c][2];
          break;
      }
      double ms = 0.0;
      if ((rc = time_shape(b, t, kb.tk, iters, e0, e1, &ms)) != TGMM_OK) break;
      if (q == 4 && kb.v[c][2] == 0)
        snprintf(line, sizeof(line), "%s cell%d %.2f\n", kb.name, kb.v[c][0] / 64, 1000.0 * ms);
      else if (kb.v[c][2])
        snprintf(line, sizeof(line), "%s %dx%dx%d %.2f\n", kb.name, kb.v[c][0], kb.v[c][1],
                 kb.v[c][2], 1000.0 * ms);
      else if (kb.v[c][1])
        snprintf(line, sizeof(line), "%s %dx%d %.2f\n", kb.name, kb.v[c][0], kb.v[c][1],
                 1000.0 * ms);
      else
        snprintf(line, sizeof(line), "%s %d %.2f\n", kb.name, kb.v[c][0], 1000.0 * ms);
      rep += line;
      if (ms < tbest) {
        tbest = ms;
        pick = t;
      }
    }
    best = pick;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != TGMM_OK) return rc;
  b->ko.tiles = best;
  destroy_graphs(b);
  if (best.knn_cell != b->grid_cell)
    if (int r2 = regrid(b, b->ko)) return r2;
  if (int r2 = tgmm_batch_reset(b)) return r2;
  if (chosen) *chosen = to_public(best);
  if (report && report_cap > 0) {
    const size_t n = std::min((size_t)report_cap - 1, rep.size());
    std::memcpy(report, rep.data(), n);
    report[n] = 0;
  }
  return TGMM_OK;
}

extern "C" int tgmm_batch_destroy(tgmm_batch* b) {
  if (!b) return TGMM_ERR_ARG;
  (void)hipSetDevice(b->device);
  destroy_graphs(b);
  if (b->st) (void)hipStreamSynchronize(b->st);
  ws_free(&b->ws);
  if (b->own_stream && b->st) (void)hipStreamDestroy(b->st);
  delete b;
  return TGMM_OK;
}
