// bpmax_server.cpp -- continuous-batching request server over the engine (placeholder
// until the serving loop lands).
#include "bpmax_engine_impl.h"

void bpx_server_shutdown(bpmax_ctx *ctx) { (void)ctx; }
