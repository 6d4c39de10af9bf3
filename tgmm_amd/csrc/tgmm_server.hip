// tgmm_server.hip -- continuous-batching server (placeholder until built).
#include "../../include/tgmm_em.h"
#include "tgmm_runtime_internal.h"

extern "C" int tgmm_server_create(int, const tgmm_server_opts*, tgmm_server**) {
  tgmm_set_error("server not built yet");
  return TGMM_ERR_ARG;
}
extern "C" int tgmm_server_submit(tgmm_server*, int, const tgmm_frame_desc*, long long*) {
  return TGMM_ERR_ARG;
}
extern "C" int tgmm_server_wait(tgmm_server*, long long) { return TGMM_ERR_ARG; }
extern "C" int tgmm_server_step(tgmm_server*, int*) { return TGMM_ERR_ARG; }
extern "C" int tgmm_server_fetch(tgmm_server*, long long, tgmm_frame_desc*, double*, double*,
                                 double*) {
  return TGMM_ERR_ARG;
}
extern "C" int tgmm_server_destroy(tgmm_server*) { return TGMM_ERR_ARG; }
