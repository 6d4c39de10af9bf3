// tgmm_server.hip -- continuous-batching server for concurrent time-lapse
// streams (no reference counterpart: the reference solves one frame at a
// time, main.cpp:387, 1100).
//
// One device workspace of fixed capacity, sized from an HBM budget, holds
// every in-flight frame. Each scheduling step
//   1. retires frames whose EM has stopped: a kernel gathers each one's
//      results into a staging block that a third stream copies back to
//      page-locked host memory in one transfer while the iterations go on;
//      the ranges are released at once (later uploads wait for the gathers)
//      and the ticket is done when its copy has landed,
//   2. admits staged frames, one frame per stream at a time (the compute
//      stream waits for the frame's upload on the device, not on the host),
//   3. replays one captured EM iteration over every in-flight frame, without
//      waiting for it (the next step does),
//   4. stages queued frames whose voxel, supervoxel, Gaussian and grid-cell
//      ranges fit (first fit, in arrival order) while that iteration runs:
//      their inputs are uploaded on copy streams, so the upload of a
//      stream's next frame overlaps the iterations of the frames in flight.
//      With nothing in flight, frames are staged and admitted at once.
// Frames join and leave at iteration boundaries, so requests that arrive
// while others are on the GPU start as soon as capacity frees up and their
// iterations share launches with the frames already running. A stream's
// posterior Gaussians stay resident in the workspace and become, in place,
// the priors of the stream's next frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <set>
#include <vector>

#include "../../include/tgmm_em.h"
#include "tgmm_runtime_internal.h"
#include "tgmm_staging.h"

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

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

long long rup(long long x, long long a) { return ((x + a - 1) / a) * a; }
size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

// first-fit allocator of aligned ranges in [0, cap)
class RangeAlloc {
 public:
  void init(long long cap) {
    free_.clear();
    hwm_ = 0;
    if (cap > 0) free_[0] = cap;
  }
  long long high_water() const { return hwm_; }  // largest end ever allocated
  bool alloc(long long n, long long align, long long* out) {
    if (n <= 0) {
      *out = 0;
      return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      const long long st = rup(it->first, align);
      const long long end = it->first + it->second;
      if (st + n <= end) {
        const long long a = it->first, b = end;
        free_.erase(it);
        if (st > a) free_[a] = st - a;
        if (st + n < b) free_[st + n] = b - (st + n);
        *out = st;
        hwm_ = std::max(hwm_, st + n);
        return true;
      }
    }
    return false;
  }
  void release(long long st, long long n) {
    if (n <= 0) return;
    auto it = free_.emplace(st, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<long long, long long> free_;
  long long hwm_ = 0;
};

// Recycled result buffers: a device staging area and its page-locked host
// mirror, of the same power-of-two capacity.
struct ResBuf {
  void* h = nullptr;
  void* d = nullptr;
  size_t cap = 0;
};

class ResPool {
 public:
  ~ResPool() {
    for (auto& kv : free_) {
      (void)hipHostFree(kv.second.h);
      (void)hipFree(kv.second.d);
    }
  }
  bool get(size_t bytes, ResBuf* r) {
    size_t c = 1 << 16;
    while (c < bytes) c <<= 1;
    auto it = free_.find(c);
    if (it != free_.end()) {
      *r = it->second;
      free_.erase(it);
      return true;
    }
    r->cap = c;
    if (hipHostMalloc(&r->h, c, 0) != hipSuccess) return false;
    if (hipMalloc(&r->d, c) != hipSuccess) {
      (void)hipHostFree(r->h);
      return false;
    }
    return true;
  }
  void put(ResBuf& r) {
    if (r.h) free_.emplace(r.cap, r);
    r = ResBuf{};
  }

 private:
  std::multimap<size_t, ResBuf> free_;
};

struct Ticket {
  long long id = 0;
  int stream = -1;
  long long N = 0, S = 0, G = 0;
  // inputs: the caller's arrays (borrowed) or a page-locked copy (in_blk)
  const float *q = nullptr, *im = nullptr, *ce = nullptr;
  const long long* pt = nullptr;
  HostBlock in_blk;
  std::vector<GaussianMixtureModelCUDA> gm_in;  // empty: priors from the stream
  int state = 0;  // 0 queued, 3 staged (ranges held, upload issued), 1 running,
                  // 4 retiring (results being copied back), 2 done
  double t_submit = 0, t_admit = 0, t_done = 0;
  int slot = -1;
  long long v0 = 0, s0 = 0, sa = 0, g0 = 0, ga = 0, c0 = 0, ncap = 0;
  bool own_g = false;         // the Gaussian range was allocated for this ticket
  hipEvent_t up = nullptr;    // the upload has completed on the copy streams
  // results: gathered on the device at retirement and copied into a
  // page-locked buffer (state 4 until `down` completes): the records, then
  // rnk and frame-local ind
  ResBuf res;
  size_t res_bytes = 0;
  hipEvent_t down = nullptr;
  double ll = 0;
  int iters = 0;
  int status = TGMM_OK;
};

struct Stream {
  long long g0 = -1, ga = 0, G = 0;  // resident Gaussian range
  long long G_next = -1;             // Gaussians of the last submitted frame
  double alpha_total = -1.0;
  bool busy = false;                 // a frame of this stream is in flight
};

}  // namespace

struct tgmm_server {
  int device = 0;
  hipStream_t st = nullptr;   // compute
  // uploads: a frame's voxel arrays go out on several copy streams at once
  // (each stream's copies run on one DMA engine, which alone does not fill
  // the host link); the first stream joins the others before the frame's
  // upload event
  static constexpr int kUp = 4;
  hipStream_t cst[kUp] = {};
  hipEvent_t up_join[kUp] = {};
  hipStream_t dst = nullptr;  // result downloads
  hipEvent_t gathered = nullptr;        // the step's result gathers are done
  std::vector<long long> to_download;   // retired this step, copies not yet queued
  tgmm_server_opts o;
  KOpts ko;
  Workspace ws;
  RangeAlloc av, as, ag, ac;
  std::vector<int> slot_ticket;
  std::vector<FrameDev> fh;
  std::map<long long, Ticket> tickets;
  std::deque<long long> queue;
  std::map<int, Stream> streams;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long long next_id = 0;
  int in_flight = 0;
  long long anon_queued = 0;  // queued frames with no stream
  ResPool pool;
  // page-locked host sources of the uploads, held until their copies are done
  static bool host_alloc(void** p, size_t n) { return hipHostMalloc(p, n, 0) == hipSuccess; }
  static void host_free(void* p) { (void)hipHostFree(p); }
  HostPool<bool (*)(void**, size_t), void (*)(void*)> hpool{&host_alloc, &host_free};
  InFlight<hipEvent_t> uploads;
  std::vector<long long> retiring;
  hipEvent_t released = nullptr;  // recorded after the last retirement's copies
  bool release_pending = false;   // uploads must wait for `released`
  // cumulative host seconds: submit, admit, iterate (launch + wait), retire
  double t_phase[4] = {0, 0, 0, 0};
  long long n_steps = 0, n_iter_frames = 0;
  hipEvent_t ev_it[2] = {nullptr, nullptr};  // around each graph replay
  hipEvent_t ev_fh = nullptr;                // frame states copied back
  FrameDev* fh_pin = nullptr;                // page-locked landing area for them
  bool launched = false;                     // an iteration is in flight
  double t_gpu = 0;
};

static int server_capture(tgmm_server* s) {
  if (s->gexec) return TGMM_OK;
  HIPCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
  hipError_t le = launch_iteration(s->ws.P, s->ko, s->st);
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(s->st, &g);
  if (le != hipSuccess || ee != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    tgmm_set_error("server graph capture failed");
    return TGMM_ERR_DEVICE;
  }
  s->graph = g;
  HIPCHK(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
  return TGMM_OK;
}

extern "C" int tgmm_server_create(int device, const tgmm_server_opts* opts, tgmm_server** out) {
  if (!opts || !out || opts->max_frames_in_flight < 1 || opts->hbm_budget_bytes <= 0) {
    tgmm_set_error("tgmm_server_create: bad arguments");
    return TGMM_ERR_ARG;
  }
  const tgmm_em_opts& e = opts->em;
  if (e.K < 1 || e.K > TGMM_KMAX || e.max_iter < 0 || e.lambda_min < 0) {
    tgmm_set_error("tgmm_server_create: bad EM options");
    return e.lambda_min < 0 ? TGMM_ERR_NOT_SET : TGMM_ERR_ARG;
  }
  if (e.voxel_mode) {
    tgmm_set_error("tgmm_server_create: voxel mode is served by batches, not by the server");
    return TGMM_ERR_ARG;
  }
  HIPCHK(hipSetDevice(device));
  tgmm_server* s = new tgmm_server();
  s->device = device;
  s->o = *opts;
  s->ko = make_kopts(e, nullptr);
  // capacity from the budget: supervoxels >= 8 voxels and Gaussians >= 32
  // voxels on average, two grid cells per Gaussian plus slack per slot
  WsShape sh;
  sh.Fcap = opts->max_frames_in_flight;
  long long V = 1 << 20;
  for (;;) {
    WsShape t;
    t.Fcap = sh.Fcap;
    const long long V2 = V * 2;
    t.Vcap = V2;
    t.Scap = rup(V2 / 8 + 256LL * t.Fcap, kAlign);
    t.Gcap = rup(V2 / 32 + 256LL * t.Fcap, kAlign);
    t.Ccap = rup(t.Gcap / 2 + 2LL * kCB * t.Fcap, kCB);
    if (ws_bytes_estimate(t, e.K) > opts->hbm_budget_bytes || V2 > (1LL << 31)) break;
    V = V2;
  }
  sh.Vcap = V;
  sh.Scap = rup(V / 8 + 256LL * sh.Fcap, kAlign);
  sh.Gcap = rup(V / 32 + 256LL * sh.Fcap, kAlign);
  sh.Ccap = rup(sh.Gcap / 2 + 2LL * kCB * sh.Fcap, kCB);
  if (int rc = ws_alloc(&s->ws, sh, e.K)) {
    ws_free(&s->ws);
    delete s;
    return rc;
  }
  DevPtrs& P = s->ws.P;
  P.nframes = sh.Fcap;
  P.only_frame = -1;
  P.Sext = kAlign;  // grown to the allocated ranges as frames are admitted
  P.Gext = kAlign;
  P.Cext = kCB;
  bool streams_ok = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) == hipSuccess &&
                    hipStreamCreateWithFlags(&s->dst, hipStreamNonBlocking) == hipSuccess;
  for (int i = 0; i < tgmm_server::kUp && streams_ok; i++)
    streams_ok = hipStreamCreateWithFlags(&s->cst[i], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&s->up_join[i], hipEventDisableTiming) == hipSuccess;
  if (!streams_ok) {
    if (s->st) (void)hipStreamDestroy(s->st);
    if (s->dst) (void)hipStreamDestroy(s->dst);
    for (int i = 0; i < tgmm_server::kUp; i++) {
      if (s->cst[i]) (void)hipStreamDestroy(s->cst[i]);
      if (s->up_join[i]) (void)hipEventDestroy(s->up_join[i]);
    }
    ws_free(&s->ws);
    delete s;
    tgmm_set_error("hipStreamCreate failed");
    return TGMM_ERR_DEVICE;
  }
  s->av.init(sh.Vcap);
  s->as.init(sh.Scap);
  s->ag.init(sh.Gcap);
  s->ac.init(sh.Ccap);
  s->slot_ticket.assign(sh.Fcap, -1);
  s->fh.assign(sh.Fcap, FrameDev());
  std::memset(s->fh.data(), 0, sizeof(FrameDev) * sh.Fcap);
  if (hipHostMalloc((void**)&s->fh_pin, sizeof(FrameDev) * sh.Fcap, 0) != hipSuccess ||
      hipEventCreate(&s->ev_it[0]) != hipSuccess || hipEventCreate(&s->ev_it[1]) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev_fh, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->gathered, hipEventDisableTiming) != hipSuccess) {
    tgmm_set_error("server: page-locked frame table or events not available");
    tgmm_server_destroy(s);
    return TGMM_ERR_OOM;
  }
  const long long S = sh.Scap, G = sh.Gcap;
  auto fail = [&](hipError_t e2) {
    tgmm_set_error("server init: %s", hipGetErrorString(e2));
    tgmm_server_destroy(s);
    return TGMM_ERR_DEVICE;
  };
  hipError_t er;
  if ((er = hipMemsetAsync(P.sv_frame, 0xFF, S * 4, s->st)) ||
      (er = hipMemsetAsync(P.sblk_frame, 0xFF, (S / kAlign) * 4, s->st)) ||
      (er = hipMemsetAsync(P.cblk_frame, 0xFF, (sh.Ccap / kCB + 1) * 4, s->st)) ||
      (er = hipMemsetAsync(P.g_frame, 0xFF, G * 4, s->st)) ||
      (er = hipMemsetAsync(P.gblk_frame, 0xFF, (G / kGB) * 4, s->st)) ||
      (er = hipMemsetAsync(P.ind, 0xFF, (size_t)e.K * S * 4, s->st)) ||
      (er = hipMemsetAsync(P.rnk, 0, (size_t)e.K * S * 4, s->st)) ||
      (er = hipMemsetAsync(P.seg_cnt, 0, (G + 1) * 4, s->st)) ||
      (er = hipMemsetAsync(P.cell_cnt, 0, (sh.Ccap + 1) * 4, s->st)) ||
      (er = hipMemsetAsync(P.knn_ctl, 0, 64 * 4, s->st)) ||
      (er = hipMemsetAsync(P.svkeep, 0, (size_t)17 * G * 4, s->st)) ||
      (er = hipMemcpyAsync(P.fr, s->fh.data(), sh.Fcap * sizeof(FrameDev),
                           hipMemcpyHostToDevice, s->st)) ||
      (er = hipStreamSynchronize(s->st)))
    return fail(er);
  *out = s;
  return TGMM_OK;
}

extern "C" int tgmm_server_submit(tgmm_server* s, int stream_id, const tgmm_frame_desc* d,
                                  long long* ticket) {
  if (!s || !d || !ticket || d->n_voxels < 0 || d->n_supervoxels < 0 || d->n_gaussians < 0 ||
      !d->sv_ptr || (d->n_voxels > 0 && (!d->query || !d->img)) ||
      (d->n_supervoxels > 0 && !d->centroids)) {
    tgmm_set_error("tgmm_server_submit: bad arguments");
    return TGMM_ERR_ARG;
  }
  const double t_enter = now_s();
  Ticket t;
  t.id = s->next_id++;
  t.stream = stream_id;
  t.N = d->n_voxels;
  t.S = d->n_supervoxels;
  for (long long j = 0; j < t.S; j++)
    if (d->sv_ptr[j] < 0 || d->sv_ptr[j + 1] < d->sv_ptr[j] || d->sv_ptr[j + 1] > t.N) {
      tgmm_set_error("tgmm_server_submit: supervoxel pointer array not monotone in [0, N]");
      return TGMM_ERR_ARG;
    }
  if (s->o.borrow_inputs) {
    t.q = d->query;
    t.im = d->img;
    t.ce = d->centroids;
    t.pt = d->sv_ptr;
  } else {  // a page-locked copy, uploaded from and returned after its copies
    const size_t oq = 0, oi = up256(12 * (size_t)t.N), oc = oi + up256(4 * (size_t)t.N),
                 op = oc + up256(12 * (size_t)t.S), total = op + 8 * (size_t)(t.S + 1);
    if (!s->hpool.get(total, &t.in_blk)) {
      tgmm_set_error("tgmm_server_submit: %zu bytes of page-locked memory not available", total);
      return TGMM_ERR_OOM;
    }
    char* b = (char*)t.in_blk.p;
    std::memcpy(b + oq, d->query, 12 * (size_t)t.N);
    std::memcpy(b + oi, d->img, 4 * (size_t)t.N);
    std::memcpy(b + oc, d->centroids, 12 * (size_t)t.S);
    std::memcpy(b + op, d->sv_ptr, 8 * (size_t)(t.S + 1));
    t.q = (const float*)(b + oq);
    t.im = (const float*)(b + oi);
    t.ce = (const float*)(b + oc);
    t.pt = (const long long*)(b + op);
  }
  if (d->gm) {
    t.G = d->n_gaussians;
    t.gm_in.assign(d->gm, d->gm + t.G);
  } else {
    auto it = s->streams.find(stream_id);
    if (stream_id < 0 || it == s->streams.end() || it->second.G_next < 0) {
      tgmm_set_error("tgmm_server_submit: no Gaussians given and no earlier frame of the stream");
      s->hpool.put(t.in_blk);
      return TGMM_ERR_ARG;
    }
    t.G = it->second.G_next;  // priors come from the stream's previous posterior
  }
  if (stream_id >= 0) s->streams[stream_id].G_next = t.G;
  const WsShape& sh = s->ws.sh;
  if (t.N > sh.Vcap || rup(std::max<long long>(t.S, 1), kAlign) > sh.Scap ||
      rup(t.G + 1, kAlign) > sh.Gcap) {
    tgmm_set_error("frame exceeds the server's capacity (N=%lld S=%lld G=%lld)", t.N, t.S, t.G);
    s->hpool.put(t.in_blk);
    return TGMM_ERR_LIMIT;
  }
  t.t_submit = now_s();
  s->t_phase[0] += t.t_submit - t_enter;
  if (stream_id < 0) s->anon_queued++;
  s->queue.push_back(t.id);
  s->tickets.emplace(t.id, std::move(t));
  *ticket = t.id;
  return TGMM_OK;
}

// Kernel grids cover the ranges allocated so far (high-water marks), not the
// whole budget-sized workspace; the captured iteration is re-captured when
// they grow.
static int update_extents(tgmm_server* s) {
  DevPtrs& P = s->ws.P;
  const long long se = std::max<long long>(rup(s->as.high_water(), kAlign), kAlign);
  const long long ge = std::max<long long>(rup(s->ag.high_water(), kAlign), kAlign);
  const long long ce = std::max<long long>(rup(s->ac.high_water(), kCB), kCB);
  if (se == P.Sext && ge == P.Gext && ce == P.Cext) return TGMM_OK;
  P.Sext = std::min(se, P.Scap);
  P.Gext = std::min(ge, P.Gcap);
  P.Cext = std::min(ce, P.Ccap);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
  return TGMM_OK;
}

// The compute stream issues no copies of its own while serving: the copy
// engines are kept busy by the uploads and downloads, and a small copy queued
// behind them would hold up the iterations. Frame-table entries travel as
// kernel arguments and the table comes back through page-locked host memory
// written by a kernel.
__global__ void k_set_frame(FrameDev* fr, FrameDev F) {
  if (threadIdx.x == 0) *fr = F;
}

__global__ void __launch_bounds__(256) k_frames_to_host(int* __restrict__ out,
                                                        const int* __restrict__ fr, int words) {
  for (int i = threadIdx.x; i < words; i += blockDim.x) out[i] = fr[i];
}

__global__ void __launch_bounds__(256) k_fill_i32(int* p, long long n, int v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// One per-element frame table range: n entries from p, the first `valid` of
// them set to the slot and the rest to -1.
struct MapSeg {
  int* p;
  long long n, valid;
};

__global__ void __launch_bounds__(256) k_map_ranges(MapSeg s0, MapSeg s1, MapSeg s2, MapSeg s3,
                                                    MapSeg s4, int f) {
  const MapSeg m = blockIdx.y == 0 ? s0 : blockIdx.y == 1 ? s1 : blockIdx.y == 2 ? s2
                   : blockIdx.y == 3 ? s3 : s4;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m.n;
       i += (long long)gridDim.x * blockDim.x)
    m.p[i] = i < m.valid ? f : -1;
}

// Map the frame's ranges to slot f (or -1) in the per-element frame tables
// (on the compute stream, so ordered with every kernel that reads them).
static int map_ranges(tgmm_server* s, Ticket& t, int f) {
  DevPtrs& P = s->ws.P;
  const long long vs = f >= 0 ? t.S : 0, vg = f >= 0 ? t.G : 0;
  MapSeg a{P.sv_frame + t.s0, t.sa, vs};
  MapSeg b{P.sblk_frame + t.s0 / kAlign, t.sa / kAlign, f >= 0 ? t.sa / kAlign : 0};
  MapSeg c{P.g_frame + t.g0, t.ga, vg};
  MapSeg d{P.gblk_frame + t.g0 / kGB, t.ga / kGB, f >= 0 ? t.ga / kGB : 0};
  MapSeg e{P.cblk_frame + t.c0 / kCB, t.ncap / kCB, f >= 0 ? t.ncap / kCB : 0};
  const long long mx = std::max(std::max(t.sa, t.ga), t.ncap / kCB);
  const unsigned gx = (unsigned)std::max(1LL, std::min(1024LL, (mx + 255) / 256));
  k_map_ranges<<<dim3(gx, 5), 256, 0, s->st>>>(a, b, c, d, e, f);
  HIPCHK(hipGetLastError());
  return TGMM_OK;
}

static void release_ranges(tgmm_server* s, Ticket& t) {
  s->av.release(t.v0, t.N);
  s->as.release(t.s0, t.sa);
  s->ac.release(t.c0, t.ncap);
  if (t.own_g) s->ag.release(t.g0, t.ga);
  t.own_g = false;
}

// Allocate the ticket's ranges and issue its uploads on the copy streams.
// Returns TGMM_OK with t.state = 3, or TGMM_ERR_BUSY if a range does not fit.
static int stage(tgmm_server* s, Ticket& t) {
  DevPtrs& P = s->ws.P;
  hipStream_t* cs = s->cst;
  constexpr int nu = tgmm_server::kUp;
  const long long S = s->ws.sh.Scap, G = s->ws.sh.Gcap;
  t.sa = rup(std::max<long long>(t.S, 1), kAlign);
  t.ga = rup(t.G + 1, kAlign);  // at least one padding entry (segment end)
  t.ncap = rup(t.G / 2 + kCB, kCB);
  t.own_g = false;
  if (!s->av.alloc(t.N, 1, &t.v0)) return TGMM_ERR_BUSY;
  if (!s->as.alloc(t.sa, kAlign, &t.s0)) {
    s->av.release(t.v0, t.N);
    return TGMM_ERR_BUSY;
  }
  if (!s->ac.alloc(t.ncap, kCB, &t.c0)) {
    s->av.release(t.v0, t.N);
    s->as.release(t.s0, t.sa);
    return TGMM_ERR_BUSY;
  }
  if (!t.gm_in.empty()) {  // records given: a range of its own
    if (!s->ag.alloc(t.ga, kAlign, &t.g0)) {
      release_ranges(s, t);
      return TGMM_ERR_BUSY;
    }
    t.own_g = true;
  }
  if (s->release_pending) {  // ranges released by retirements are read back first
    for (int i = 0; i < nu; i++) HIPCHK(hipStreamWaitEvent(cs[i], s->released, 0));
    s->release_pending = false;
  }
  if (t.N > 0) {
    HIPCHK(hipMemcpyAsync(P.vx + t.v0, t.q, t.N * 4, hipMemcpyHostToDevice, cs[0]));
    HIPCHK(hipMemcpyAsync(P.vy + t.v0, t.q + t.N, t.N * 4, hipMemcpyHostToDevice, cs[1 % nu]));
    HIPCHK(hipMemcpyAsync(P.vz + t.v0, t.q + 2 * t.N, t.N * 4, hipMemcpyHostToDevice,
                          cs[2 % nu]));
    HIPCHK(hipMemcpyAsync(P.vimg + t.v0, t.im, t.N * 4, hipMemcpyHostToDevice, cs[3 % nu]));
  }
  // the supervoxel tables and, for given records, the Gaussian state, fixed
  // flags and supervoxel lists, packed into one page-locked block
  const size_t ob = 0, ol = up256(8 * (size_t)t.S), og = ol + up256(4 * (size_t)t.S);
  const size_t nG = t.own_g ? (size_t)t.G : 0;
  const size_t of = og + up256(8 * GS_NFIELDS * nG), ok = of + up256(nG);
  const size_t total = ok + 68 * nG;
  HostBlock tab;
  if (!s->hpool.get(total, &tab)) {
    tgmm_set_error("server: %zu bytes of page-locked memory not available", total);
    release_ranges(s, t);
    return TGMM_ERR_OOM;
  }
  char* tb = (char*)tab.p;
  long long* beg = (long long*)(tb + ob);
  int* len = (int*)(tb + ol);
  for (long long j = 0; j < t.S; j++) {
    beg[j] = t.v0 + t.pt[j];
    len[j] = (int)(t.pt[j + 1] - t.pt[j]);
  }
  if (t.S > 0) {
    for (int a = 0; a < 3; a++)
      HIPCHK(hipMemcpyAsync(P.sv_cent + a * S + t.s0, t.ce + a * t.S, t.S * 4,
                            hipMemcpyHostToDevice, cs[0]));
    HIPCHK(hipMemcpyAsync(P.sv_beg + t.s0, beg, t.S * 8, hipMemcpyHostToDevice, cs[0]));
    HIPCHK(hipMemcpyAsync(P.sv_len + t.s0, len, t.S * 4, hipMemcpyHostToDevice, cs[0]));
  }
  if (t.own_g && t.G > 0) {
    double* gs = (double*)(tb + og);
    unsigned char* fx = (unsigned char*)(tb + of);
    int* kp = (int*)(tb + ok);
    records_to_soa(t.gm_in.data(), t.G, 0, t.G, gs, fx, kp);
    for (int fi = 0; fi < GS_NFIELDS; fi++)
      HIPCHK(hipMemcpyAsync(P.gs + fi * G + t.g0, gs + fi * t.G, t.G * 8, hipMemcpyHostToDevice,
                            cs[0]));
    HIPCHK(hipMemcpyAsync(P.fixed + t.g0, fx, t.G, hipMemcpyHostToDevice, cs[0]));
    HIPCHK(hipMemcpyAsync(P.svkeep + 17 * t.g0, kp, 68 * t.G, hipMemcpyHostToDevice, cs[0]));
  }
  if (!t.up) HIPCHK(hipEventCreateWithFlags(&t.up, hipEventDisableTiming));
  for (int i = 1; i < nu; i++) {
    HIPCHK(hipEventRecord(s->up_join[i], cs[i]));
    HIPCHK(hipStreamWaitEvent(cs[0], s->up_join[i], 0));
  }
  HIPCHK(hipEventRecord(t.up, cs[0]));
  // the blocks read by these copies return to the pool once t.up completes
  s->uploads.hold(t.id, tab, t.up);
  if (t.in_blk.p) {
    s->uploads.hold(t.id, t.in_blk, t.up);
    t.in_blk = HostBlock{};
  }
  t.state = 3;
  return TGMM_OK;
}

// Put a staged ticket into slot f: the compute stream waits for its upload,
// the stream's resident posterior becomes its priors (or its own records
// replace the resident state), then the frame's set-up kernels.
static int activate(tgmm_server* s, Ticket& t, int f) {
  DevPtrs& P = s->ws.P;
  hipStream_t st = s->st;
  Stream& sr = s->streams[t.stream];
  HIPCHK(hipStreamWaitEvent(st, t.up, 0));
  if (t.own_g) {
    if (t.stream >= 0 && sr.g0 >= 0) {  // new records replace the resident state
      s->ag.release(sr.g0, sr.ga);
      sr.g0 = -1;
    }
  } else {
    t.g0 = sr.g0;
    HIPCHK(launch_next_priors(P, t.g0, t.G, s->o.beta_percentage_of_nk,
                              s->o.nu_percentage_of_nk, s->o.alpha_percentage, sr.alpha_total,
                              st));
  }
  FrameDev& F = s->fh[f];
  std::memset(&F, 0, sizeof(F));
  F.v0 = t.v0;
  F.nv = t.N;
  F.s0 = t.s0;
  F.ns = t.S;
  F.g0 = t.g0;
  F.ng = t.G;
  F.c0 = t.c0;
  F.ncap = t.ncap;
  F.nbg = (int)(t.ga / kGB);
  F.nbl = (int)(t.sa / 256);
  F.max_iter = s->o.em.max_iter;
  F.tol = s->o.em.tol;
  F.packed = 1;
  k_set_frame<<<1, 64, 0, st>>>(P.fr + f, F);
  HIPCHK(hipGetLastError());
  if (int rc = map_ranges(s, t, f)) return rc;
  if (int rc = update_extents(s)) return rc;  // grids must cover the new ranges
  DevPtrs Q = P;  // set-up kernels over this frame's ranges only
  Q.only_frame = f;
  Q.s_lo = t.s0;
  Q.Sext = t.sa;
  Q.g_lo = t.g0;
  Q.Gext = t.ga;
  // the cell range may hold counts left by an earlier frame's last binning
  k_fill_i32<<<(unsigned)std::min<long long>(1024, (t.ncap + 255) / 256), 256, 0, st>>>(
      P.cell_cnt + t.c0, t.ncap, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(launch_prep(Q, s->ko, st));
  HIPCHK(launch_reset(Q, s->ko, st));
  t.slot = f;
  t.state = 1;
  t.t_admit = now_s();
  s->slot_ticket[f] = (int)t.id;
  s->in_flight++;
  sr.busy = true;
  return TGMM_OK;
}

// Stage queued frames (a stream's frames strictly in order, at most its next
// one staged ahead).
static int stage_queued(tgmm_server* s) {
  std::set<int> ahead;  // streams whose earliest unfinished frame was passed
  int staged = 0;
  for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
    Ticket& t = s->tickets[*it];
    if (t.state == 3) {
      staged++;
      if (t.stream >= 0) ahead.insert(t.stream);
      continue;
    }
    if (t.stream >= 0 && ahead.count(t.stream)) continue;  // not this stream's next
    if (t.stream >= 0) ahead.insert(t.stream);
    if (staged >= s->ws.sh.Fcap) continue;
    const int rc = stage(s, t);
    if (rc == TGMM_OK) {
      staged++;
    } else if (rc != TGMM_ERR_BUSY) {
      return rc;
    }
  }
  return TGMM_OK;
}

// Admit staged frames whose stream is free, in arrival order.
static int admit_staged(tgmm_server* s) {
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    if (s->in_flight >= s->ws.sh.Fcap) break;
    Ticket& t = s->tickets[*it];
    if (t.state != 3 || (t.stream >= 0 && s->streams[t.stream].busy)) {
      ++it;
      continue;
    }
    if (!t.own_g && s->streams[t.stream].g0 < 0) {  // no resident state (cannot happen)
      ++it;
      continue;
    }
    int f = -1;
    for (int i = 0; i < (int)s->slot_ticket.size(); i++)
      if (s->slot_ticket[i] < 0) {
        f = i;
        break;
      }
    if (f < 0) break;
    if (int rc = activate(s, t, f)) return rc;
    if (t.stream < 0) s->anon_queued--;
    it = s->queue.erase(it);
  }
  return TGMM_OK;
}

// A retired frame's results gathered into one contiguous block: its records
// [G] (288 bytes each, formed by k_finalize), the responsibilities [K][S]
// (f32) and the candidate indices [K][S] (i32, made frame-local; -1 stays
// -1).
__global__ void __launch_bounds__(256) k_gather_results(const GaussianMixtureModelCUDA* rec,
    const float* __restrict__ rnk, const int* __restrict__ ind, long long Scap, long long g0,
    long long G, long long s0, long long S, int K, char* __restrict__ out) {
  const long long recb = G * (long long)sizeof(GaussianMixtureModelCUDA), rb = K * S * 4;
  const long long step = (long long)gridDim.x * blockDim.x;
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  switch (blockIdx.y) {
    case 0: {
      const double* src = (const double*)(rec + g0);
      for (long long i = i0; i < recb / 8; i += step) ((double*)out)[i] = src[i];
      break;
    }
    case 1:
      for (long long i = i0; i < K * S; i += step) {
        const long long j = i / S, v = i - j * S;
        ((float*)(out + recb))[i] = rnk[j * Scap + s0 + v];
      }
      break;
    default:
      for (long long i = i0; i < K * S; i += step) {
        const long long j = i / S, v = i - j * S;
        const int k = ind[j * Scap + s0 + v];
        ((int*)(out + recb + rb))[i] = k >= 0 ? k - (int)g0 : k;
      }
  }
}

// Retire the frame of slot f: its results are gathered into a staging block
// on the compute stream (ordered before any later use of its ranges there;
// uploads on the copy streams wait for `released`), to be copied back on the
// download stream (queue_downloads), its ranges are unmapped and released,
// and the stream keeps its Gaussian range for the next frame. The ticket is
// done once its copy has landed.
static int retire(tgmm_server* s, int f) {
  const long long id = s->slot_ticket[f];
  Ticket& t = s->tickets[id];
  DevPtrs& P = s->ws.P;
  hipStream_t st = s->st;
  const long long S = s->ws.sh.Scap;
  const int K = s->ws.K;
  const size_t recb = (size_t)t.G * sizeof(GaussianMixtureModelCUDA), rb = (size_t)K * t.S * 4;
  const size_t bytes = recb + 2 * rb;
  if (!s->pool.get(bytes + 64, &t.res)) {
    tgmm_set_error("result buffers of %zu bytes not available", bytes);
    return TGMM_ERR_OOM;
  }
  // the last search's candidates in the reference's slot layout, then the
  // records with the r > 0.5 supervoxel assignment (a stream's next frame
  // starts from these lists)
  DevPtrs Q = P;
  Q.s_lo = t.s0;
  Q.Sext = t.sa;
  Q.g_lo = t.g0;
  Q.Gext = t.ga;
  HIPCHK(launch_knn_layout(Q, s->ko, f, (int)t.G, st));
  HIPCHK(launch_finalize(Q, s->ko, s->o.em.max_sv_per_gaussian, f, t.stream >= 0 ? 1 : 0, st));
  const long long mx = std::max<long long>((long long)(recb / 8), (long long)K * t.S);
  const unsigned gx = (unsigned)std::max(1LL, std::min(2048LL, (mx + 255) / 256));
  k_gather_results<<<dim3(gx, 3), 256, 0, st>>>(P.rec, P.rnk, P.ind, S, t.g0, t.G, t.s0, t.S, K,
                                                 (char*)t.res.d);
  HIPCHK(hipGetLastError());
  t.res_bytes = bytes;
  s->to_download.push_back(t.id);
  const FrameDev F = s->fh[f];  // a copy: the slot's entry is cleared below
  t.ll = F.ll;
  t.iters = F.iter - 1;
  t.state = 4;
  s->retiring.push_back(t.id);
  // unmap and release
  if (int rc = map_ranges(s, t, -1)) return rc;
  FrameDev z;
  std::memset(&z, 0, sizeof(z));
  s->fh[f] = z;
  k_set_frame<<<1, 64, 0, st>>>(P.fr + f, z);
  HIPCHK(hipGetLastError());
  s->av.release(t.v0, t.N);
  s->as.release(t.s0, t.sa);
  s->ac.release(t.c0, t.ncap);
  Stream& sr = s->streams[t.stream];
  if (t.stream >= 0) {
    sr.g0 = t.g0;
    sr.ga = t.ga;
    sr.G = t.G;
    sr.alpha_total = F.sum_alpha;
  } else {
    s->ag.release(t.g0, t.ga);
  }
  sr.busy = false;
  s->slot_ticket[f] = -1;
  s->in_flight--;
  s->release_pending = true;
  return TGMM_OK;
}

// Copy the staging blocks gathered in this step back to page-locked host
// memory on the download stream, after the gathers, overlapping the
// iterations that follow on the compute stream.
static int queue_downloads(tgmm_server* s) {
  if (s->to_download.empty()) return TGMM_OK;
  HIPCHK(hipEventRecord(s->gathered, s->st));
  HIPCHK(hipStreamWaitEvent(s->dst, s->gathered, 0));
  for (long long id : s->to_download) {
    Ticket& t = s->tickets[id];
    HIPCHK(hipMemcpyAsync(t.res.h, t.res.d, t.res_bytes, hipMemcpyDeviceToHost, s->dst));
    if (!t.down) HIPCHK(hipEventCreateWithFlags(&t.down, hipEventDisableTiming));
    HIPCHK(hipEventRecord(t.down, s->dst));
  }
  s->to_download.clear();
  return TGMM_OK;
}

// Upload blocks whose copies have completed go back to the pool.
static int reclaim_uploads(tgmm_server* s) {
  const int rc = s->uploads.reclaim(s->hpool, [](hipEvent_t e) {
    const hipError_t q = hipEventQuery(e);
    return q == hipSuccess ? 1 : q == hipErrorNotReady ? 0 : -1;
  });
  if (rc < 0) {
    tgmm_set_error("server: an upload failed");
    return TGMM_ERR_DEVICE;
  }
  return TGMM_OK;
}

// Retiring tickets whose results have landed become done.
static int collect(tgmm_server* s, bool block) {
  for (size_t i = 0; i < s->retiring.size();) {
    Ticket& t = s->tickets[s->retiring[i]];
    hipError_t q = block ? hipEventSynchronize(t.down) : hipEventQuery(t.down);
    if (q == hipErrorNotReady) {
      ++i;
      continue;
    }
    HIPCHK(q);
    t.state = 2;
    t.t_done = now_s();
    std::vector<GaussianMixtureModelCUDA>().swap(t.gm_in);
    s->retiring.erase(s->retiring.begin() + i);
  }
  return reclaim_uploads(s);
}

// Wait for the iteration launched by the previous step and take the frame
// states it copied back.
static int drain(tgmm_server* s) {
  if (!s->launched) return TGMM_OK;
  const double t = now_s();
  HIPCHK(hipEventSynchronize(s->ev_fh));
  std::memcpy(s->fh.data(), s->fh_pin, s->fh.size() * sizeof(FrameDev));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, s->ev_it[0], s->ev_it[1]));
  s->t_gpu += 1e-3 * ms;
  s->launched = false;
  s->t_phase[2] += now_s() - t;
  return TGMM_OK;
}

extern "C" int tgmm_server_step(tgmm_server* s, int* pending) {
  if (!s) return TGMM_ERR_ARG;
  HIPCHK(hipSetDevice(s->device));
  if (int rc = drain(s)) return rc;
  // retire frames that stopped in the previous iteration
  const double t0 = now_s();
  for (int f = 0; f < (int)s->slot_ticket.size(); f++)
    if (s->slot_ticket[f] >= 0 && !s->fh[f].active)
      if (int rc = retire(s, f)) return rc;
  if (int rc = queue_downloads(s)) return rc;
  if (s->release_pending) {
    if (!s->released) HIPCHK(hipEventCreateWithFlags(&s->released, hipEventDisableTiming));
    HIPCHK(hipEventRecord(s->released, s->st));
  }
  const double t1 = now_s();
  if (int rc = admit_staged(s)) return rc;
  if (s->in_flight == 0) {  // idle: stage and admit at once
    if (int rc = stage_queued(s)) return rc;
    if (int rc = admit_staged(s)) return rc;
  }
  if (int rc = update_extents(s)) return rc;
  const double t2 = now_s();
  if (s->in_flight > 0) {
    if (int rc = server_capture(s)) return rc;
    HIPCHK(hipEventRecord(s->ev_it[0], s->st));
    HIPCHK(hipGraphLaunch(s->gexec, s->st));
    HIPCHK(hipEventRecord(s->ev_it[1], s->st));
    static_assert(sizeof(FrameDev) % 4 == 0, "frame table copied in words");
    k_frames_to_host<<<1, 256, 0, s->st>>>((int*)s->fh_pin, (const int*)s->ws.P.fr,
                                           (int)(s->fh.size() * sizeof(FrameDev) / 4));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(s->ev_fh, s->st));
    s->launched = true;  // the next step (or drain) waits for it
    s->n_steps++;
    s->n_iter_frames += s->in_flight;
  }
  const double t3 = now_s();
  // staging (host tables, uploads) overlaps the iteration just launched;
  // these frames are admitted by the next step
  if (int rc = stage_queued(s)) return rc;
  const double t4 = now_s();
  s->t_phase[3] += t1 - t0;
  s->t_phase[1] += (t2 - t1) + (t4 - t3);
  s->t_phase[2] += t3 - t2;
  // results that have landed; with nothing left to run, wait for them
  if (int rc = collect(s, !s->launched)) return rc;
  if (pending) *pending = s->in_flight + (int)s->queue.size() + (int)s->retiring.size();
  return TGMM_OK;
}

extern "C" int tgmm_server_wait(tgmm_server* s, long long ticket) {
  if (!s) return TGMM_ERR_ARG;
  if (ticket >= 0 && !s->tickets.count(ticket)) {
    tgmm_set_error("unknown ticket %lld", ticket);
    return TGMM_ERR_ARG;
  }
  for (;;) {
    int pending = 0;
    if (ticket >= 0 && s->tickets[ticket].state == 2) return TGMM_OK;
    if (int rc = tgmm_server_step(s, &pending)) return rc;
    if (ticket < 0 && pending == 0) return TGMM_OK;
    const int ts = ticket >= 0 ? s->tickets[ticket].state : 0;
    if (ticket >= 0 && (ts == 0 || ts == 3) && s->in_flight == 0) {
      // nothing in flight and the ticket was not admitted: it never fits
      int again = 0;
      if (int rc = tgmm_server_step(s, &again)) return rc;
      const int t2 = s->tickets[ticket].state;
      if ((t2 == 0 || t2 == 3) && s->in_flight == 0) {
        tgmm_set_error("ticket %lld cannot be admitted (capacity)", ticket);
        return TGMM_ERR_BUSY;
      }
    }
  }
}

extern "C" int tgmm_server_query(tgmm_server* s, long long ticket, int* state) {
  if (!s || !state || !s->tickets.count(ticket)) {
    tgmm_set_error("tgmm_server_query: bad arguments");
    return TGMM_ERR_ARG;
  }
  const int st = s->tickets[ticket].state;
  *state = (st == 3) ? 0 : (st == 4) ? 1 : st;  // staged: queued; retiring: running
  return TGMM_OK;
}

extern "C" int tgmm_server_fetch(tgmm_server* s, long long ticket, tgmm_frame_desc* d,
                                 double* t_submit, double* t_admit, double* t_done) {
  if (!s || !d || !s->tickets.count(ticket)) {
    tgmm_set_error("tgmm_server_fetch: bad arguments");
    return TGMM_ERR_ARG;
  }
  Ticket& t = s->tickets[ticket];
  if (t.state != 2) {
    tgmm_set_error("ticket %lld not finished", ticket);
    return TGMM_ERR_BUSY;
  }
  const int K = s->ws.K;
  const size_t recb = (size_t)t.G * sizeof(GaussianMixtureModelCUDA), rb = (size_t)K * t.S * 4;
  const char* base = (const char*)t.res.h;
  if (d->gm) std::memcpy(d->gm, base, recb);
  if (d->rnk) std::memcpy(d->rnk, base + recb, rb);
  if (d->ind) std::memcpy(d->ind, base + recb + rb, rb);
  // the ticket's upload blocks are back in the pool before its event goes
  if (s->uploads.held_by(ticket)) {
    HIPCHK(hipEventSynchronize(t.up));
    if (int rc = reclaim_uploads(s)) return rc;
  }
  s->pool.put(t.res);
  if (t.down) (void)hipEventDestroy(t.down);
  d->ll = t.ll;
  d->iterations = t.iters;
  d->status = t.status;
  d->n_gaussians = (int)t.G;
  if (t_submit) *t_submit = t.t_submit;
  if (t_admit) *t_admit = t.t_admit;
  if (t_done) *t_done = t.t_done;
  if (t.up) (void)hipEventDestroy(t.up);
  s->tickets.erase(ticket);
  return TGMM_OK;
}

extern "C" int tgmm_server_stats(tgmm_server* s, tgmm_server_stats_t* st) {
  if (!s || !st) return TGMM_ERR_ARG;
  st->submit_s = s->t_phase[0];
  st->admit_s = s->t_phase[1];
  st->iterate_s = s->t_phase[2];
  st->retire_s = s->t_phase[3];
  st->iterate_gpu_s = s->t_gpu;
  st->steps = s->n_steps;
  st->frame_iterations = s->n_iter_frames;
  st->in_flight = s->in_flight;
  st->queued = (long long)s->queue.size();
  return TGMM_OK;
}

extern "C" int tgmm_server_set_tiles(tgmm_server* s, const tgmm_tiles* t) {
  if (!s || !t) return TGMM_ERR_ARG;
  const Tiles x{t->ll_group,  t->ll_unroll, t->mstep_lanes, t->sv_block,
                t->knn_block, t->knn_chunk, t->knn_cell,    t->knn_lanes};
  if (!tiles_valid(x) || x.ll_gw == 1) {  // the server runs supervoxel mode only
    tgmm_set_error("tgmm_server_set_tiles: invalid tile shape");
    return TGMM_ERR_ARG;
  }
  if (x.knn_cell != s->ko.tiles.knn_cell && s->in_flight > 0) {
    // grids are built when a frame is admitted
    tgmm_set_error("tgmm_server_set_tiles: knn_cell cannot change while frames are in flight");
    return TGMM_ERR_BUSY;
  }
  HIPCHK(hipSetDevice(s->device));
  if (int rc = drain(s)) return rc;
  HIPCHK(hipStreamSynchronize(s->st));
  s->ko.tiles = x;
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
  return TGMM_OK;
}

extern "C" int tgmm_server_destroy(tgmm_server* s) {
  if (!s) return TGMM_ERR_ARG;
  (void)hipSetDevice(s->device);
  for (hipStream_t c : s->cst)
    if (c) (void)hipStreamSynchronize(c);
  if (s->st) (void)hipStreamSynchronize(s->st);
  if (s->dst) (void)hipStreamSynchronize(s->dst);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->uploads.drain(s->hpool, [](hipEvent_t e) { (void)hipEventSynchronize(e); });
  for (auto& kv : s->tickets) {
    if (kv.second.up) (void)hipEventDestroy(kv.second.up);
    if (kv.second.down) (void)hipEventDestroy(kv.second.down);
    s->pool.put(kv.second.res);
    s->hpool.put(kv.second.in_blk);
  }
  s->hpool.clear();
  if (s->released) (void)hipEventDestroy(s->released);
  for (hipEvent_t e : s->ev_it)
    if (e) (void)hipEventDestroy(e);
  if (s->ev_fh) (void)hipEventDestroy(s->ev_fh);
  if (s->fh_pin) (void)hipHostFree(s->fh_pin);
  ws_free(&s->ws);
  if (s->st) (void)hipStreamDestroy(s->st);
  for (int i = 0; i < tgmm_server::kUp; i++) {
    if (s->cst[i]) (void)hipStreamDestroy(s->cst[i]);
    if (s->up_join[i]) (void)hipEventDestroy(s->up_join[i]);
  }
  if (s->dst) (void)hipStreamDestroy(s->dst);
  if (s->gathered) (void)hipEventDestroy(s->gathered);
  delete s;
  return TGMM_OK;
}
