/*
 * tgmm_em.h -- C-ABI of the MI355X (gfx950) variational-inference EM for
 * supervoxel-based Gaussian mixtures (the per-frame fit of the TGMM tracker).
 *
 * Two layers are exported from libtgmm_em.so:
 *
 *  1. Drop-in entry points with the argument lists of the reference host API
 *     (nmeth.3036-S2/src/UtilsCUDA/GMEMupdateCUDA.h). They are exported with a
 *     "tgmm_" prefix and return a status code instead of calling exit(); the
 *     C++ header tgmm_em_compat.hpp wraps them in inline functions carrying the
 *     reference's exact names, parameter lists and exit-on-error behaviour.
 *
 *  2. A widened, handle-based API for batches of frames of uneven size
 *     (64-bit voxel/supervoxel counts, runtime K <= 16, more than 65535
 *     Gaussians), hipGraph-replayed iterations, and a continuous-batching
 *     server for concurrent time-lapse streams.
 *
 * All pointers are plain host or device pointers; no torch types appear here.
 * Every function returns TGMM_OK (0) or a negative TGMM_ERR_* code; the last
 * error message is available from tgmm_last_error().
 */
#ifndef TGMM_EM_H_
#define TGMM_EM_H_

#ifdef __cplusplus
extern "C" {
#endif

#define TGMM_DIMS 3
#define TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN 16 /* GMEMupdateCUDA.h:20 */
#define TGMM_KMAX 16                         /* widened API: K in [1, 16] */
#define TGMM_LEGACY_K 5                      /* constants.h:15 maxGaussiansPerVoxel */
#define TGMM_SV_KNN_MAX 11 /* temporalLogicalRules/knnCUDA/knnCuda.h:28 maxKNN */

/* status codes */
#define TGMM_OK 0
#define TGMM_ERR_ARG (-1)     /* invalid argument (NULL, negative size, bad K) */
#define TGMM_ERR_LIMIT (-2)   /* size beyond a compiled or configured limit */
#define TGMM_ERR_OOM (-3)     /* device or host allocation failed / budget */
#define TGMM_ERR_DEVICE (-4)  /* HIP runtime error */
#define TGMM_ERR_NOT_SET (-5) /* regularisation constants were never set */
#define TGMM_ERR_IO (-6)      /* debug dump could not be written */
#define TGMM_ERR_BUSY (-7)    /* server: no free slot / capacity right now */

/*
 * Per-Gaussian record, 288 bytes, byte-for-byte the layout of
 * GaussianMixtureModelCUDA (GMEMupdateCUDA.h:22-59): 27 doubles, the `fixed`
 * flag (one byte, C++ bool), 16 supervoxel indices and their count.
 * Offsets: beta_k 0, nu_k 8, alpha_k 16, m_k 24, W_k 48, E[ln det] 96,
 * E[ln pi] 104, splitScore 112, beta_o 120, nu_o 128, alpha_o 136, m_o 144,
 * W_o 168, fixed 216, supervoxelIdx 220, supervoxelNum 284.
 * W_k and W_o are packed symmetric [xx, xy, xz, yy, yz, zz].
 * A Gaussian is dead when m_o[0] < -1e31.
 */
typedef struct GaussianMixtureModelCUDA {
  double beta_k, nu_k, alpha_k;
  double m_k[TGMM_DIMS];
  double W_k[6];
  double expectedLogDetCovarianceCUDA;
  double expectedLogResponsivityCUDA;
  double splitScore;
  double beta_o, nu_o, alpha_o;
  double m_o[TGMM_DIMS];
  double W_o[6];
  unsigned char fixed;
  int supervoxelIdx[TGMM_MAX_SUPERVOXELS_PER_GAUSSIAN];
  int supervoxelNum;
} GaussianMixtureModelCUDA;

/* Last error message of the calling thread (never NULL). */
const char* tgmm_last_error(void);
/* Number of HIP devices visible (0 when none); never initialises streams. */
int tgmm_device_count(int* count);

/* ---------------------------------------------------------------------------
 * 1. Drop-in entry points (K = 5, unsigned short supervoxel count, G <= 65535)
 * ------------------------------------------------------------------------ */

/* Replaces GMEMinitializeMemoryWithSupervoxels (GMEMupdateCUDA.h:132-137,
 * GMEMupdateCUDA.cu:4580-4636). Allocates query[3N] f32 SoA, img[N],
 * rnk[5S], ind[5S], indTr[5S], centroids[3S] SoA, ptr[S+1] on the current
 * device, copies the host arrays, and records `scale` as the kNN anisotropy
 * (uploadScaleCUDA, knnCuda.cu:827-830). */
int tgmm_GMEMinitializeMemoryWithSupervoxels(
    float** queryCUDA, const float* query, float** imgDataCUDA,
    const float* imgData, const float* scale, long long query_nb,
    float** rnkCUDA, int** indCUDA, int** indCUDAtr, unsigned short numLabels,
    float** centroidLabelPositionCUDA, const float* centroidLabelPosition,
    long long** labelListPtrCUDA, const long long* labelListPtrHOST);

/* Replaces copyScaleToDvice (GMEMupdateCUDA.h:82, GMEMupdateCUDA.cu:4760):
 * anisotropy used by the precision-matrix regularisation. */
int tgmm_copyScaleToDvice(const float* scale);

/* Replaces the host statics regularizePrecisionMatrixConstants::setConstants
 * (constants.h:38-44) that GMEMvariationalInferenceCUDAWithSupervoxels reads
 * (GMEMupdateCUDA.cu:4217-4226). Until called, the EM returns
 * TGMM_ERR_NOT_SET (the reference exits with status 3). */
int tgmm_setRegularizePrecisionMatrixConstants(double lambdaMin,
                                               double lambdaMax,
                                               double maxExcentricity);

/* Replaces GMEMvariationalInferenceCUDAWithSupervoxels (GMEMupdateCUDA.h:92-98,
 * GMEMupdateCUDA.cu:3985-4421). Device inputs from the init call; vecGMtemp is
 * host in/out [ref_nb]. On return: posterior and expectation fields updated,
 * supervoxelIdx/supervoxelNum rebuilt from r > 0.5 (<= 16, slot-major scan;
 * entries past the count as given), rnkCUDA/indCUDA (device, slot-major
 * s + j*S) hold the last iteration in the reference's slot layout (the
 * outcome of its scan of the Gaussians in index order, knnCuda.cu:61-144)
 * and unused slots ind = -1, r = 0. rnkCUDAtr may be NULL (unused, as in the
 * reference); indCUDAtr receives a copy of indCUDA. debugPath NULL or ""
 * disables the {int S; int K; float rnk[]; int ind[]} dump. The frame runs in
 * a one-frame workspace kept per device between calls (grown when a frame
 * does not fit, held until the process ends), its iterations replayed as
 * captured hipGraphs, convergence polled every four iterations. */
int tgmm_GMEMvariationalInferenceCUDAWithSupervoxels(
    float* queryCUDA, float* imgDataCUDA, float* rnkCUDA, float* rnkCUDAtr,
    int* indCUDA, int* indCUDAtr, float* centroidLabelPositionCUDA,
    long long* labelListPtrCUDA, GaussianMixtureModelCUDA* vecGMtemp,
    long long query_nb, int ref_nb, unsigned short numLabels, int maxIterEM,
    double tolLikelihood, int devCUDA, int frame, int W4DOF,
    const char* debugPath);

/* Replaces GMEMvariationalInferenceCUDA (GMEMupdateCUDA.h:87-91,
 * GMEMupdateCUDA.cu:3420-3973), the per-voxel form of the EM: device inputs
 * queryCUDA [3*query_nb] SoA and imgDataCUDA [query_nb]; vecGMtemp is host
 * in/out [ref_nb]. Every voxel takes its K = 5 nearest Gaussians; its
 * responsibilities are cut to zero where the log score is <= -100 and are
 * multiplied by the voxel's intensity (GMEMupdateCUDA.cu:2177-2242); the
 * M-step keeps pairs whose stored value exceeds 1e-3. On return: posterior
 * and expectation fields updated, supervoxelIdx/supervoxelNum as given,
 * rnkCUDA/indCUDA (device, slot-major v + j*query_nb) hold the last
 * iteration. rnkCUDAtr may be NULL; indCUDAtr (may be NULL) receives a copy
 * of indCUDA. debugPath is accepted and unused, as in the reference. */
int tgmm_GMEMvariationalInferenceCUDA(float* queryCUDA, float* imgDataCUDA, float* rnkCUDA,
                                      float* rnkCUDAtr, int* indCUDA, int* indCUDAtr,
                                      GaussianMixtureModelCUDA* vecGMtemp, long long query_nb,
                                      int ref_nb, int maxIterEM, double tolLikelihood,
                                      int devCUDA, int frame, int W4DOF, const char* debugPath);

/* Replaces calculateLocalLikelihood (GMEMupdateCUDA.h:116-127,
 * GMEMupdateCUDA.cu:2952-3050): for each list of supervoxel indices, the sum
 * over its supervoxels s and their voxels v of
 *   img_v * log( sum_j sqrt(det W_j nu_j^3) / (2 pi)^1.5
 *                      * exp(-nu_j (x_v - m_j)^T W_j (x_v - m_j) / 2)
 *                      * alpha_j / sum_alpha )
 * over the K = 5 candidates j = indCUDA[s + j*numLabels] (slots < 0 skipped),
 * sum_alpha over all ref_nb records of vecGMtemp (host, read only); a voxel
 * whose density is not > 0 adds nothing; a supervoxel listed twice counts
 * twice. The reference's vector<vector<int>> is passed in CSR form: list i is
 * list_idx[list_ptr[i] .. list_ptr[i+1]), list_ptr[0] = 0, n_lists + 1
 * entries. localLikelihood[i] (host) is written for every non-empty list and
 * left as given for an empty one (the reference skips it). Device inputs
 * queryCUDA, imgDataCUDA, indCUDA, labelListPtrCUDA as after
 * tgmm_GMEMvariationalInferenceCUDAWithSupervoxels; rnkCUDA is unused, as in
 * the reference. Indices outside [0, numLabels) are rejected (TGMM_ERR_ARG).
 * Runs on the current device. */
int tgmm_calculateLocalLikelihood(double* localLikelihood, const long long* list_ptr,
                                  const int* list_idx, int n_lists, const float* queryCUDA,
                                  const float* imgDataCUDA, const float* rnkCUDA,
                                  const int* indCUDA, const GaussianMixtureModelCUDA* vecGMtemp,
                                  const long long* labelListPtrCUDA, long long query_nb,
                                  int ref_nb, int numLabels);

/* Replaces GMEMreleaseMemoryWithSupervoxels (GMEMupdateCUDA.h:140-144,
 * GMEMupdateCUDA.cu:4654-4676): frees and nulls every pointer. */
int tgmm_GMEMreleaseMemoryWithSupervoxels(float** queryCUDA,
                                          float** imgDataCUDA, float** rnkCUDA,
                                          int** indCUDA, int** indCUDAtr,
                                          float** centroidLabelPositionCUDA,
                                          long long** labelListPtrCUDA);

/* Replaces GMEMinitializeMemory (GMEMupdateCUDA.h:129-131,
 * GMEMupdateCUDA.cu:4538-4579), the voxel-mode initialiser: allocates
 * query[3N] f32 SoA, img[N], rnk[5N], ind[5N], indTr[5N] on the current
 * device, copies the host arrays, and records `scale` as the kNN anisotropy
 * (uploadScaleCUDA). */
int tgmm_GMEMinitializeMemory(float** queryCUDA, const float* query, float** imgDataCUDA,
                              const float* imgData, const float* scale, long long query_nb,
                              float** rnkCUDA, int** indCUDA, int** indCUDAtr);

/* Replaces GMEMreleaseMemory (GMEMupdateCUDA.h:138-139,
 * GMEMupdateCUDA.cu:4637-4653): frees and nulls every pointer. */
int tgmm_GMEMreleaseMemory(float** queryCUDA, float** imgDataCUDA, float** rnkCUDA,
                           int** indCUDA, int** indCUDAtr);

/* Replaces selectCUDAdeviceMaxMemory (UtilsCUDA/knnCuda.h:23,
 * knnCuda.cu:27-47), called by the tracker at start-up (main.cpp:207): the
 * visible device with the most global memory (the first of equals), or -1
 * when there is none. */
int tgmm_selectCUDAdeviceMaxMemory(int* device);

/* Replaces setDeviceCUDA (UtilsCUDA/knnCuda.h:20, knnCuda.cu:737-741), called
 * by the tracker before each frame (main.cpp:778): makes devCUDA current. */
int tgmm_setDeviceCUDA(int devCUDA);

/* Replaces uploadScaleCUDA (UtilsCUDA/knnCuda.h:21, knnCuda.cu:827-830): the
 * kNN anisotropy used by the VI entry points (the init entry points also set
 * it, as the reference's do). */
int tgmm_uploadScaleCUDA(const float* scale);

/* ---------------------------------------------------------------------------
 * 2. Widened API: batches of frames, graphs, serving
 * ------------------------------------------------------------------------ */

typedef struct tgmm_em_opts {
  int K;                   /* candidates per supervoxel, 1..TGMM_KMAX */
  int max_iter;            /* maxIterEM */
  double tol;              /* tolLikelihood; < 0 runs exactly max_iter */
  double lambda_min;       /* regularisation constants (constants.h:19-45) */
  double lambda_max;
  double max_excentricity;
  int w4dof;               /* zero xz/yz of W in regularisation */
  float scale[TGMM_DIMS];  /* anisotropy for kNN and regularisation */
  int max_sv_per_gaussian; /* cap of supervoxelIdx, 0..16 (16 = reference) */
  int voxel_mode;          /* 1: per-voxel EM (GMEMvariationalInferenceCUDA,
                              GMEMupdateCUDA.cu:3420-3973): every voxel is its
                              own query with K candidates; n_supervoxels,
                              sv_ptr and centroids are ignored, rnk/ind have
                              K*n_voxels entries and rnk is multiplied by the
                              voxel's intensity; supervoxelIdx/Num are left as
                              given */
} tgmm_em_opts;

/* One frame. Input arrays are HOST memory with the reference's boundary
 * layout (query SoA x..,y..,z.., voxels of supervoxel s in
 * [sv_ptr[s], sv_ptr[s+1]), centroids SoA). gm is host in/out; rnk/ind are
 * optional host outputs of K*n_supervoxels entries, slot-major (s + j*S). */
typedef struct tgmm_frame_desc {
  long long n_voxels;
  long long n_supervoxels;
  int n_gaussians;
  const float* query;
  const float* img;
  const float* centroids;
  const long long* sv_ptr;
  GaussianMixtureModelCUDA* gm;
  float* rnk;
  int* ind;
  double ll;      /* out: final log-likelihood (GMEM.cu:4254-4262) */
  int iterations; /* out: EM iterations run */
  int status;     /* out: TGMM_OK or error code for this frame */
} tgmm_frame_desc;

typedef struct tgmm_batch tgmm_batch;

/* Solve modes */
#define TGMM_SOLVE_EAGER 0     /* kernels launched per iteration, no host sync */
#define TGMM_SOLVE_GRAPH 1     /* iterations captured in hipGraphs, replayed */
#define TGMM_SOLVE_HOSTSYNC 2  /* plain form: host reads ll/convergence each iteration */

/* Uploads all frames into one device workspace (one allocation set), builds
 * per-supervoxel moments and per-frame kNN grids. Host arrays may be freed
 * after this returns. */
int tgmm_batch_create(int device, const tgmm_frame_desc* frames, int nframes,
                      const tgmm_em_opts* opts, tgmm_batch** out);
/* Re-loads every frame's initial Gaussian records (as given at create) into
 * the device state, device-to-device. */
int tgmm_batch_reset(tgmm_batch* b);
/* Runs EM on all frames; stream-ordered, returns after the last iteration has
 * been enqueued and completed (synchronises the batch stream). */
int tgmm_batch_solve(tgmm_batch* b, int mode);
/* Copies results into frames[i].gm/rnk/ind/ll/iterations (host); gm, rnk
 * and ind may each be NULL. The records are formed on the device (state,
 * fixed flag, the r > 0.5 supervoxel assignment; supervoxelIdx entries past
 * supervoxelNum as given) and copied in one transfer per frame; rnk/ind are
 * in the reference's slot layout with frame-local indices. */
int tgmm_batch_fetch(tgmm_batch* b, tgmm_frame_desc* frames, int nframes);

/* Diagnostics of the last solve: supervoxels whose K-th kNN distance was
 * tied in the last iteration (searched again in the reference's scan order),
 * supervoxels whose final slot layout found a different candidate set than
 * the iterations' search (0 unless something is wrong), workspace bytes. */
typedef struct tgmm_batch_stats_t {
  int knn_ties_last_iteration;
  int layout_mismatches;
  long long workspace_bytes;
} tgmm_batch_stats_t;
int tgmm_batch_stats(tgmm_batch* b, tgmm_batch_stats_t* stats);
/* Launch shapes of the tunable kernels. Every valid choice gives bitwise the
 * same results (the kernels fix their summation order independently of the
 * shape); only the speed changes. Defaults: {16, 4, 16, 256, 256, 8, 6, 2}. */
typedef struct tgmm_tiles {
  int ll_group;    /* log-likelihood: 16 or 32 for 2 * ll_group supervoxels
                      per block of 256 threads (32 for K > 8), 64 for 16
                      per block of 128;
                      1 (with ll_unroll 1) = one lane per one-voxel
                      supervoxel, voxel mode only, its default there */
  int ll_unroll;   /* log-likelihood: voxels a thread loads at once, 1, 2
                      or 4 (8 with ll_group 16) */
  int mstep_lanes; /* lanes per Gaussian, M-step statistics: 4, 8 or 16 */
  int sv_block;    /* threads per block, E-step and pair scatter: 64, 128, 256 */
  int knn_block;   /* threads per block, kNN search: 64, 128 or 256 */
  int knn_chunk;   /* candidate references loaded together, kNN search: 4 or 8 */
  int knn_cell;    /* mean Gaussians per kNN grid cell: 2, 4 or 6 */
  int knn_lanes;   /* kNN search: 0 = one block of knn_block/64 waves per grid
                      cell, its 3x3x3 block staged in LDS for all of the
                      cell's queries; 1, 2 or 4 = lanes per query */
} tgmm_tiles;

/* A change of knn_cell rebuilds the kNN grids and resets the batch to its
 * initial state (as tgmm_batch_reset). */
int tgmm_batch_set_tiles(tgmm_batch* b, const tgmm_tiles* tiles);
int tgmm_batch_get_tiles(tgmm_batch* b, tgmm_tiles* tiles);
/* Autotuner. For each knob in turn (log-likelihood shape, M-step lanes,
 * E-step block, kNN block) it times every candidate, the other knobs held at
 * the best so far, over `reps` eager iterations of the batch's own frames
 * from their initial state (hipEvents around the kernel on the batch
 * stream), and keeps the fastest. Leaves the batch in its initial state with
 * the chosen tiles. `chosen` and `report` (one line per candidate: knob,
 * shape, microseconds per launch) are optional. */
int tgmm_batch_autotune(tgmm_batch* b, int reps, tgmm_tiles* chosen, char* report,
                        int report_cap);

/* Average device time (ms) of one launch of the named kernel over `reps`
 * eager iterations, measured with hipEvents on the batch stream. In each
 * iteration the events bracket 8 back-to-back launches of "loglik" (it
 * writes only its block partials) and one launch of "estep", "knn" and
 * "mstep" (they update state that a repeated launch would read). Leaves the
 * state advanced. */
int tgmm_batch_time_kernel(tgmm_batch* b, const char* name, int reps,
                           double* avg_ms, double* bytes_per_launch);
int tgmm_batch_destroy(tgmm_batch* b);

/* Continuous-batching server for concurrent time-lapse streams. Frames are
 * admitted into a fixed-capacity device workspace sized by an HBM budget;
 * every hipGraph replay advances every in-flight frame by one EM iteration;
 * frames join as capacity frees up and leave when done; each stream's
 * posterior Gaussians stay on the device and become the next frame's priors
 * (GaussianMixtureModel::updatePriors, GaussianMixtureModel.cpp:892-910,
 * plus the tracker's nu_k*2, W_k/2 step, main.cpp:961-969). */
typedef struct tgmm_server tgmm_server;

typedef struct tgmm_server_opts {
  tgmm_em_opts em;
  long long hbm_budget_bytes; /* device workspace budget */
  int max_frames_in_flight;
  double beta_percentage_of_nk;  /* updatePriors arguments */
  double nu_percentage_of_nk;
  double alpha_percentage;
  int borrow_inputs; /* nonzero: submit keeps pointers to the caller's query,
                        img, centroids and sv_ptr arrays (which must stay valid
                        until the ticket is fetched) instead of copying them;
                        page-locked arrays then upload without blocking */
} tgmm_server_opts;

int tgmm_server_create(int device, const tgmm_server_opts* opts,
                       tgmm_server** out);
/* Queue a frame. stream_id >= 0 names a time-lapse; when the stream already
 * has device-resident Gaussians and frame->gm is NULL, the priors come from
 * the stream's previous posterior. Returns a ticket >= 0 or an error. Input
 * host arrays are copied before return unless borrow_inputs is set; records
 * (gm) are always copied. */
int tgmm_server_submit(tgmm_server* s, int stream_id,
                       const tgmm_frame_desc* frame, long long* ticket);
/* Runs scheduling steps until the ticket completes (or, with ticket < 0,
 * until every queued frame completes). */
int tgmm_server_wait(tgmm_server* s, long long ticket);
/* Advances the server by one scheduling step (admit, one iteration, retire);
 * returns the number of frames still in flight or queued via *pending. */
int tgmm_server_step(tgmm_server* s, int* pending);
/* State of a ticket that was not fetched yet: 0 queued, 1 in flight,
 * 2 done (results on the host, ready for tgmm_server_fetch). */
int tgmm_server_query(tgmm_server* s, long long ticket, int* state);
/* Fetch a completed ticket's outputs into frame (gm must have room for the
 * frame's Gaussians; rnk/ind optional). Also returns admission and
 * completion wall times (seconds, monotonic clock). */
int tgmm_server_fetch(tgmm_server* s, long long ticket, tgmm_frame_desc* frame,
                      double* t_submit, double* t_admit, double* t_done);
/* Cumulative accounting of a server: host seconds spent in submit (input
 * copies), admission (uploads, prior update, per-frame set-up), iteration
 * (graph launch and wait) and retirement (results gathered and queued for
 * copy-back); scheduling steps that ran an iteration, frame-iterations run,
 * frames in flight and queued now; device seconds of the iteration graph
 * replays (HIP events around each replay). */
typedef struct tgmm_server_stats_t {
  double submit_s, admit_s, iterate_s, retire_s;
  long long steps, frame_iterations;
  long long in_flight, queued;
  double iterate_gpu_s;
} tgmm_server_stats_t;
int tgmm_server_stats(tgmm_server* s, tgmm_server_stats_t* stats);

/* Launch shapes for the server's iterations (for example those the
 * autotuner chose on a representative batch). knn_cell can change only while
 * no frame is in flight (TGMM_ERR_BUSY otherwise). */
int tgmm_server_set_tiles(tgmm_server* s, const tgmm_tiles* tiles);
int tgmm_server_destroy(tgmm_server* s);

/* ---------------------------------------------------------------------------
 * 3. Input marshalling on the device
 * ------------------------------------------------------------------------ */

/* One frame as raw supervoxel voxels: what the tracker holds after trimming
 * (mainTrackingGaussianMixtureModel.cpp:556-649). Supervoxel s is
 * pixel_idx[sv_ptr[s] .. sv_ptr[s+1]) (its PixelIdxList, linear indices
 * x + d0 (y + d1 z)); supervoxels are disjoint. Intensities are the raw
 * uint16 values of the listed voxels (raw), or, when raw is NULL, read from
 * volume, the whole raw image in device memory. thr[s] is the supervoxel's
 * local background (the trimSupervoxel threshold, :558-559). pixel_idx, raw
 * and thr are host memory, or device memory read in place when
 * device_inputs is 1; sv_ptr is host memory. Volumes up to 2^53 voxels. */
typedef struct tgmm_raw_frame {
  long long N;                         /* voxels listed */
  long long S;                         /* supervoxels */
  long long dims[TGMM_DIMS];           /* image size, x fastest */
  const unsigned long long* pixel_idx; /* [N] */
  const unsigned short* raw;           /* [N] or NULL */
  const unsigned short* volume;        /* [d0 d1 d2] device, read where raw is NULL */
  const long long* sv_ptr;             /* [S+1], sv_ptr[0] = 0, sv_ptr[S] = N */
  const unsigned short* thr;           /* [S] */
  int device_inputs;                   /* 1: pixel_idx, raw, thr are device memory */
} tgmm_raw_frame;

/* No single reference function: replaces the tracker's host construction of
 * the EM inputs, mainTrackingGaussianMixtureModel.cpp:598-669 (local
 * background subtraction in uint16, zeroing outside the supervoxels,
 * Convert_Array_Inplace to FLOAT32 and Scale_Array_To_Range(img, 0, 1), mylib
 * array.p:2708-2820) and :774-870 (the voxel SoA from Idx2CoordA), with the
 * centroids of supervoxel::weightedGaussianStatistics<uint16>
 * (supervoxel.cpp:1410-1449; raw intensities, float sums in list order).
 * Writes, on the current device: queryCUDA [3N] SoA coordinates,
 * imgDataCUDA [N] scaled intensities, centroidLabelPositionCUDA [3S] SoA and,
 * if not NULL, labelListPtrCUDA [S+1] (a copy of sv_ptr). Returns
 * TGMM_ERR_ARG for an index outside the volume (mylib exits there) and for
 * more voxels than the volume holds (the lists overlap). */
int tgmm_marshal_frame(const tgmm_raw_frame* frame, float* queryCUDA, float* imgDataCUDA,
                       float* centroidLabelPositionCUDA, long long* labelListPtrCUDA);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). tgmm_marshal_frame with the same outputs, bitwise,
 * that allocates nothing: its working buffers (the uploaded supervoxel
 * pointer, the intensity range, the block partials) live in a 16 MiB
 * device array of the library, one per device, kept until the process
 * ends. The inputs must be device memory (device_inputs = 1). Frames whose
 * buffers exceed the array (about 1.97 million supervoxels) return
 * TGMM_ERR_LIMIT. Every call on a device shares that device's array and
 * leaves its contents behind; two calls overlapping on one device would race
 * on it, so the library serialises the calls of a process and each call
 * returns only after its kernels have finished (each process has its own
 * array). */
int tgmm_marshal_frame_noalloc(const tgmm_raw_frame* frame, float* queryCUDA, float* imgDataCUDA,
                               float* centroidLabelPositionCUDA, long long* labelListPtrCUDA);

/* tgmm_GMEMinitializeMemoryWithSupervoxels (GMEMupdateCUDA.h:132-137) fed
 * from a raw frame instead of the host float arrays: allocates the same
 * seven device arrays (query_nb = frame->N, numLabels = frame->S <= 65535),
 * fills them with tgmm_marshal_frame and records scale as the kNN
 * anisotropy. Release with tgmm_GMEMreleaseMemoryWithSupervoxels. */
int tgmm_GMEMinitializeMemoryFromRaw(float** queryCUDA, float** imgDataCUDA, const float* scale,
                                     float** rnkCUDA, int** indCUDA, int** indCUDAtr,
                                     float** centroidLabelPositionCUDA,
                                     long long** labelListPtrCUDA, const tgmm_raw_frame* frame);

/* ---------------------------------------------------------------------------
 * 4. Supervoxel nearest neighbours
 * ------------------------------------------------------------------------ */

/* Replaces knnCUDA_ (temporalLogicalRules/knnCUDA/knnCuda.h:56-58,
 * knnCuda.cu:142-232), the tracker's supervoxel neighbour search in space and
 * time (lineageHyperTree.cpp:2526-2775, supervoxel.cpp:700-1290). Host
 * inputs query [3*query_nb] and ref [3*ref_nb] (SoA centroids), scale[3];
 * host outputs ind and dist (may be NULL) [KNN*query_nb], slot-major
 * (q + j*query_nb): for each query the references are scanned in index
 * order and a reference whose scaled squared distance
 * ((q-r)_x s_x)^2 + ((q-r)_y s_y)^2 + ((q-r)_z s_z)^2 (float, in that order)
 * is not above the largest slot distance replaces the first slot holding it
 * (slots start at 1e32), the reference's exact slot layout; slots never
 * filled have ind = -1 (uninitialised in the reference) and dist = 1e32.
 * ref_nb <= 0 sets dist to 1e32 and leaves ind as given. KNN outside
 * 1..TGMM_SV_KNN_MAX returns TGMM_ERR_LIMIT (the reference's kernel writes
 * nothing then). Runs on device devCUDA. */
int tgmm_knnCUDA_(int* ind, float* dist, const float* query, const float* ref,
                  long long query_nb, int ref_nb, int KNN, const float* scale, int devCUDA);

#ifdef __cplusplus
}
#endif

#endif /* TGMM_EM_H_ */
