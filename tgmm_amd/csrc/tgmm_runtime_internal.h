// tgmm_runtime_internal.h -- host-side structures shared by the runtime
// translation units (batch, drop-in entry points, server).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/tgmm_em.h"
#include "tgmm_internal.h"

struct WsShape {
  long long Vcap = 0, Scap = 0, Gcap = 0, Ccap = 0;
  int Fcap = 0;
};

struct Workspace {
  WsShape sh;
  int K = 0;
  tgmm::DevPtrs P;
  double* gs_init = nullptr;  // initial Gaussian state, for reset
  std::vector<void*> allocs;
  long long bytes = 0;
};

// One frame's inputs: pointers may be host or device memory (copied with
// hipMemcpyDefault); ptr and gm are host memory.
struct FrameSrc {
  long long N = 0, S = 0;
  int G = 0;
  const float* query = nullptr;
  const float* img = nullptr;
  const float* cent = nullptr;
  const long long* ptr = nullptr;
  const GaussianMixtureModelCUDA* gm = nullptr;
};

struct FrameSize {
  long long N, S, G;
};

struct tgmm_batch {
  int device = 0;
  hipStream_t st = nullptr;
  bool own_stream = false;
  tgmm_em_opts opts;
  tgmm::KOpts ko;
  int nframes = 0;
  std::vector<tgmm::FrameDev> fh;  // host mirror of the frame table
  std::vector<FrameSize> sizes;
  Workspace ws;
  hipGraph_t graph = nullptr;  // one iteration
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph_n = nullptr;  // cap_n iterations back to back (none when cap_n < 2)
  hipGraphExec_t gexec_n = nullptr;
  int cap_n = 0;
  tgmm::DevPtrs cap_P;  // launch arguments the graphs were captured with
  tgmm::KOpts cap_ko;
  int grid_cell = 0;  // knn_cell the kNN grids were last built with
  long long max_ng = 0;      // most Gaussians of one frame
  int layout_mismatch = 0;   // supervoxels whose final slot layout found another set
};

void tgmm_set_error(const char* fmt, ...);
int ws_alloc(Workspace* w, const WsShape& sh, int K);
void ws_free(Workspace* w);
long long ws_bytes_estimate(const WsShape& sh, int K);
tgmm::KOpts make_kopts(const tgmm_em_opts& o, const float* scale_reg);
void records_to_soa(const GaussianMixtureModelCUDA* gm, long long G, long long g0, long long Gcap,
                    double* gs, unsigned char* fixed, int* svkeep);
void soa_to_records(const double* gs, long long G, long long g0, long long Gcap,
                    GaussianMixtureModelCUDA* gm);
// Places the frames in a workspace and uploads them. With `reuse`, that
// batch's workspace, stream and graph are kept if the frames fit (else it is
// destroyed); a new workspace is `headroom` times the size needed (>= 1).
int batch_build(int device, const std::vector<FrameSrc>& fs, const tgmm_em_opts& o,
                const float* scale_reg, hipStream_t stream, tgmm_batch** out,
                tgmm_batch* reuse = nullptr, double headroom = 1.0);
int batch_upload(tgmm_batch* b, const std::vector<FrameSrc>& fs);
// Iterations between the host's looks at the frames' states when the solve
// has a tolerance, and the most iterations captured into one graph without one.
constexpr int kPollEvery = 4;
constexpr int kGraphIters = 32;
int batch_capture(tgmm_batch* b, int n);

