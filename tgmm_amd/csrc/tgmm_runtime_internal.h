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
  std::vector<unsigned char> fixed_host;
  Workspace ws;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int grid_cell = 0;  // knn_cell the kNN grids were last built with
  // voxel mode: each frame's input supervoxelIdx[16] + supervoxelNum per
  // Gaussian, returned unchanged by the fetch
  std::vector<std::vector<int>> sv_keep;
};

void tgmm_set_error(const char* fmt, ...);
int ws_alloc(Workspace* w, const WsShape& sh, int K);
void ws_free(Workspace* w);
long long ws_bytes_estimate(const WsShape& sh, int K);
tgmm::KOpts make_kopts(const tgmm_em_opts& o, const float* scale_reg);
void records_to_soa(const GaussianMixtureModelCUDA* gm, long long G, long long g0, long long Gcap,
                    double* gs, unsigned char* fixed);
void soa_to_records(const double* gs, long long G, long long g0, long long Gcap,
                    GaussianMixtureModelCUDA* gm);
int batch_build(int device, const std::vector<FrameSrc>& fs, const tgmm_em_opts& o,
                const float* scale_reg, hipStream_t stream, tgmm_batch** out);
int batch_upload(tgmm_batch* b, const std::vector<FrameSrc>& fs);
int batch_capture(tgmm_batch* b);
int batch_fetch_host(tgmm_batch* b, std::vector<double>& gs, std::vector<float>& rnk,
                     std::vector<int>& ind);
void frame_results(tgmm_batch* b, int i, const std::vector<double>& gs,
                   const std::vector<float>& rnk, const std::vector<int>& ind,
                   GaussianMixtureModelCUDA* gm, float* rnk_out, int* ind_out, double* ll,
                   int* iters);
