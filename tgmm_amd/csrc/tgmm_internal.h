// tgmm_internal.h -- device workspace layout shared by the kernels
// (tgmm_kernels.hip) and the host runtime (tgmm_runtime.hip).
//
// One workspace holds any number of frames. Every per-element array is a
// flat, capacity-sized structure of arrays:
//   voxels       [Vcap]   x, y, z, img (f32)
//   supervoxels  [Scap]   frame id, voxel range, kNN centroid, moments,
//                          K candidate slots (slot-major: j*Scap + s)
//   Gaussians    [Gcap]   27-double state (record field order), E-step and
//                          log-likelihood parameters, pair segments
//   grid cells   [Ccap]   per-frame uniform kNN grids
// A frame owns a contiguous range in each space; supervoxel and Gaussian
// ranges start on multiples of 256 so that every 256-wide block (and every
// 64-wide log-likelihood block) belongs to exactly one frame, which makes the
// per-frame reductions deterministic without atomics on floating point.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tgmm_em.h"

namespace tgmm {

constexpr int kAlign = 256;      // supervoxel / Gaussian range alignment
constexpr int kGB = 64;          // threads per block, Gaussian-lane kernels
constexpr int kCB = 64;          // grid-cell range alignment (cell block -> frame table)
constexpr int kSB = 256;         // threads per block, supervoxel-lane kernels
constexpr int kLLminB = 4;       // fewest supervoxels per log-likelihood block
constexpr int kPairShift = 4;    // pair code = (s << 4) | j, j < 16

// state field indices (record order, GMEMupdateCUDA.h:22-34)
enum {
  GS_BETA = 0, GS_NU = 1, GS_ALPHA = 2, GS_M = 3, GS_W = 6, GS_ELD = 12, GS_ELP = 13,
  GS_SPLIT = 14, GS_BETA_O = 15, GS_NU_O = 16, GS_ALPHA_O = 17, GS_M_O = 18, GS_W_O = 21,
  GS_NFIELDS = 27
};

struct FrameDev {
  long long v0, nv;   // voxel range
  long long s0, ns;   // supervoxel range (s0 % 256 == 0)
  long long g0, ng;   // Gaussian range (g0 % 256 == 0)
  long long c0, ncap; // grid-cell range
  float org[3];       // grid origin in scaled coordinates
  float h;            // cell edge
  int dims[3];
  int nbg;            // kGB-blocks over the 256-aligned Gaussian range
  int nbl;            // 256-supervoxel groups over the aligned range (log-likelihood)
  int max_iter;
  double tol;
  double sum_alpha, psi_sum_alpha, log_sum_alpha;
  double ll, ll1, ll2, ll3;
  int iter;           // reference numIterEM (starts at 1)
  int active;         // keeps iterating
  int status;
  int packed;         // every voxel's offset from its supervoxel's anchor is a
                      // small integer: the log-likelihood reads vpk (set to 1
                      // on the host, cleared by the moments kernel)
};

// Per-Gaussian parameter packs gathered by the supervoxel kernels: one
// 64-byte (log-likelihood) or 96-byte (E-step) record per Gaussian instead of
// one cache line per field.
struct alignas(64) LLPack {
  double m[3];  // mean
  double lg2;   // (log alpha + log det(nu W)/2 - 1.5 log 2pi) * log2(e)
  float U[6];   // upper Cholesky factor of nu W * log2(e) / 2:
                // u00 u01 u02 u11 u12 u22, so d^T (nu W log2(e)/2) d = |U d|^2
  float nf;     // Frobenius norm of U (rounded up): |U d| <= nf |d|
  float pad;
};
struct alignas(32) EPack {
  double m[3];
  double W[6];
  double nu;
  double kappa0;  // E[ln det]/2 - 1.5/beta
  double dpsi;    // digamma(alpha); E[ln pi] = dpsi - digamma(sum of alpha)
};

// Launch shapes of the tunable kernels (chosen per workload by the
// autotuner, or the defaults below). Every choice gives bitwise the same
// results: the kernels fix their floating-point summation order
// independently of the shape.
struct Tiles {
  int ll_gw;      // log-likelihood: 16, 32 for 2*ll_gw supervoxels per 256-thread block
                  // (32 for K > 8), 64 for 16 per 128-thread block; 1 = a lane per
                  // voxel (voxel mode)
  int ll_u;       // log-likelihood: voxels a thread loads at once, 1, 2, 4 (8 with ll_gw 16)
  int ms_lanes;   // lanes per Gaussian, M-step statistics: 4, 8, 16
  int sv_block;   // threads per block, E-step and pair scatter: 64, 128, 256
  int knn_block;  // threads per block, kNN search: 64, 128, 256
  int knn_chunk;  // candidate references loaded together, kNN search: 4, 8
  int knn_cell;   // mean Gaussians per kNN grid cell: 2, 4, 6 (set per frame
                  // when its grid is built)
  int knn_lanes;  // kNN search: 0 = one block of knn_block/64 waves per grid
                  // cell (its 3x3x3 block staged in LDS for all of the cell's
                  // queries); 1, 2, 4 = lanes per query
};
constexpr Tiles kDefaultTiles = {16, 4, 16, 256, 256, 8, 6, 2};
bool tiles_valid(const Tiles& t);

struct KOpts {
  int K;
  int w4dof;
  float scale[3];      // kNN anisotropy (uploadScaleCUDA, knnCuda.cu:827-830)
  float scale_reg[3];  // regularisation anisotropy (copyScaleToDvice)
  double lmin, lmax, emax;
  Tiles tiles;
  int voxel;  // per-voxel EM (one-voxel supervoxels, cut and intensity in the E-step)
};

struct DevPtrs {
  // voxels
  float *vx, *vy, *vz, *vimg;
  int2* vpk;        // Vcap: packed voxel: .x bytes 0..2 the offsets from the
                    // supervoxel's anchor (int8), byte 3 the supervoxel's index
                    // mod 256; .y the intensity bits
  // supervoxels
  int* sv_frame;
  long long* sv_beg;
  int* sv_len;
  float* sv_cent;   // 3*Scap SoA
  double* mom;      // 10*Scap: n, c[3], Q[6]  (unweighted, FP64)
  double* wmom;     // 10*Scap: W, mu[3], C[6] (intensity-weighted, FP64), the 10 values
                    // of a supervoxel contiguous (wmom[10 s + i]: one M-step pair reads
                    // 80 bytes in a row)
  float* anchor;    // 3*Scap: first voxel of the supervoxel
  float4* sv_ball;  // Scap: ball around the supervoxel's voxels: centre
                    // relative to the anchor, radius (rounded up)
  int* ind;         // K*Scap
  float* rnk;       // K*Scap
  int* sblk_frame;  // Scap/256
  float4* q4;       // Scap: kNN queries of each frame in grid-cell order:
                    // centroid x, y, z and the supervoxel index (int bits)
  int* qcell;       // Scap: grid cell of each supervoxel centroid
  // Gaussians
  int* g_frame;
  double* gs;       // 27*Gcap
  unsigned char* fixed;
  EPack* epack;     // Gcap: E-step parameters of the next iteration
  double* eldn;     // Gcap: E[ln det] of the next iteration (committed to the state
                    // when that iteration's M-step runs)
  LLPack* llpack;   // Gcap: log-likelihood parameters (written by death)
  float4* m4;       // Gcap: f32 mean and index bits (kNN references; NaN: dead)
  float4* mknn;     // Gcap: the means of the last kNN search, in index order
                    // (written when they are binned; NaN: dead)
  int* gcell;
  int* gblk_frame;  // Gcap/kGB
  // M-step segments
  int* seg_cnt;     // Gcap+1
  int* seg_start;   // Gcap+1
  int* pairs;       // K*Scap
  int* pairs_sorted;
  // grid
  int* cell_cnt;    // Ccap+1
  int* cell_start;  // Ccap+1
  int* qcell_start; // Ccap+1: first kNN query of each cell (positions in q4)
  int* cblk_frame;  // Ccap/kCB: frame of each 64-cell block, -1 if none
  float4* cg4;      // Gcap: cell-sorted references (x, y, z, index bits)
  // kNN: supervoxels whose K-th distance is tied with a Gaussian outside
  // their K best (their candidate set follows the reference's scan order)
  int* tie_list;    // Scap
  int* knn_ctl;     // [0] ties listed this iteration, [1] supervoxels whose
                    //     final slot layout found a different set (stays 0),
                    //     [2] Gaussians in long_list this iteration
  int* long_list;   // Gcap: Gaussians whose pair list is long (sorted by a block each)
  int2* win;        // Scap: (Gaussian, slot) of the candidate with r > 0.5, or -1
  // results
  GaussianMixtureModelCUDA* rec;  // Gcap: records as returned (288 B each)
  int* svkeep;      // 17*Gcap: supervoxelIdx[16] and supervoxelNum as given
  // reductions
  double* part_g;   // Gcap/kGB
  double* part_g2;  // Gcap/kGB: sums of alpha after the posterior update (death test)
  double* part_s;   // Scap/kLLminB: log-likelihood block partials
  double* part_sg;  // Scap/256: partials of 256-supervoxel groups
  FrameDev* fr;
  long long Scap, Gcap, Ccap, Vcap;
  long long Sext, Gext, Cext;  // launch extents (<= capacities): grids cover
                               // only the ranges in use
  int Fcap;
  int nframes;
  int only_frame;   // prep/reset kernels: -1 = every frame, else this frame
  // prep/reset kernels for one frame: its supervoxel and Gaussian ranges
  // start here (256-aligned) and Sext/Gext count from these bases
  long long s_lo = 0, g_lo = 0;
};

// kernel launchers (tgmm_kernels.hip); all enqueue on `st`
// once per frame: supervoxel moments, kNN grid shape, queries sorted by cell
hipError_t launch_prep(const DevPtrs& P, const KOpts& O, hipStream_t st);
// iteration state, initial sum of alpha and first E-step constants per frame;
// the frames' kNN grid counts must be zero
hipError_t launch_reset(const DevPtrs& P, const KOpts& O, hipStream_t st);
// one EM iteration of every active frame; optional events around one kernel
hipError_t launch_iteration(const DevPtrs& P, const KOpts& O, hipStream_t st,
                            hipEvent_t ev_before = nullptr, hipEvent_t ev_after = nullptr,
                            int timed_kernel = -1);
int kmax_for(int K);

// Priors of a stream's next frame from its previous posterior, in place on
// the Gaussian range [g0, g0+ng) (GaussianMixtureModel::updatePriors,
// GaussianMixtureModel.cpp:892-910, then nu_k *= 2, W_k /= 2 as the tracker
// does, main.cpp:961-969).
hipError_t launch_next_priors(const DevPtrs& P, long long g0, long long ng, double beta_pct,
                              double nu_pct, double alpha_pct, double alpha_total,
                              hipStream_t st);

// After a solve: each supervoxel's K slots in the reference's layout (its
// scan of the Gaussians in index order, knnCuda.cu:61-144) with the
// responsibilities moved along, and the (Gaussian, slot) of each supervoxel's
// candidate with r > 0.5. Frames that ran no iteration are left as they are.
// only_frame >= 0: that frame only (its supervoxel range from P.s_lo, P.Sext).
hipError_t launch_knn_layout(const DevPtrs& P, const KOpts& O, int only_frame, int max_ng,
                             hipStream_t st);
// The 288-byte records of every Gaussian (of only_frame, from P.g_lo and
// P.Gext, when >= 0): state, fixed flag, and supervoxelIdx/supervoxelNum
// from the r > 0.5 assignment (GMEM.cu:4328-4350; entries past the count as
// given) or, in voxel mode, as given. writeback: the returned lists become
// the given ones (a stream's next frame starts from its posterior records).
hipError_t launch_finalize(const DevPtrs& P, const KOpts& O, int cap, int only_frame,
                           int writeback, hipStream_t st);

// timed_kernel ids for launch_iteration
enum { TK_KNN = 0, TK_ESTEP = 1, TK_MSTEP = 2, TK_LOGLIK = 3 };
// launches between the timing events: the log-likelihood, which writes only
// its block partials, runs kTimedLaunches times back to back, so the event
// pair's own dispatch gap is spread over them; every other kernel updates
// state that its next launch would read (the E-step adds to the pair
// counts) and runs once
constexpr int kTimedLaunches = 8;
inline int timed_launches(int tk) { return tk == TK_LOGLIK ? kTimedLaunches : 1; }

}  // namespace tgmm
