// mcw_serve.cpp -- continuous-batching request server (placeholder until the
// serving loop lands; every entry point reports that it is unavailable).
#include "../../include/macaw_amd.h"

struct mcw_server {
  int unused;
};

extern "C" {
int mcw_server_create(mcw_server **srv, mcw_ctx *const *, int32_t, int32_t, int64_t)
{
  if (srv) *srv = nullptr;
  return 1;
}
int mcw_server_destroy(mcw_server *) { return 0; }
const char *mcw_server_last_error(const mcw_server *) { return "serving is not available in this build"; }
int mcw_submit(mcw_server *, const mcw_request *) { return 1; }
int mcw_poll(mcw_server *, int32_t *) { return 1; }
int mcw_result(mcw_server *, mcw_request_status *, double *, size_t) { return 1; }
}
