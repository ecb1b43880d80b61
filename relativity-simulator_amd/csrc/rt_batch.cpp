// rt_batch.cpp — batched frames and the serving loop (rt_trace_batch, rt_server_*).
#include <hip/hip_runtime.h>

#include "../../include/rt_tracer.h"
#include "rt_host.h"

using namespace rt;

extern "C" {

rt_status rt_trace_batch(rt_context*, const rt_request*, size_t, rt_results*) {
    return fail(RT_ERR_INVALID_ARG, "rt_trace_batch: not available in this build");
}
rt_status rt_server_create(int, const rt_server_config*, rt_server** out) {
    if (out) *out = nullptr;
    return fail(RT_ERR_INVALID_ARG, "rt_server: not available in this build");
}
rt_status rt_server_destroy(rt_server*) { return RT_OK; }
rt_status rt_server_submit(rt_server*, const rt_request*, uint64_t*) { return fail(RT_ERR_INVALID_ARG, "no server"); }
rt_status rt_server_poll(rt_server*, uint64_t, int*) { return fail(RT_ERR_INVALID_ARG, "no server"); }
rt_status rt_server_wait(rt_server*, uint64_t, int64_t) { return fail(RT_ERR_INVALID_ARG, "no server"); }
rt_status rt_server_take(rt_server*, uint64_t, rt_results*) { return fail(RT_ERR_INVALID_ARG, "no server"); }
rt_status rt_server_cancel(rt_server*, uint64_t) { return fail(RT_ERR_INVALID_ARG, "no server"); }

}  // extern "C"
