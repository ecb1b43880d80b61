// rt_host.h — host-side helpers shared by the context, batch and server code.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/rt_tracer.h"
#include "rt_params.h"

namespace rt {

rt_status fail(rt_status st, const std::string& msg);
rt_status hip_fail(hipError_t e, const char* what);

bool valid_request(const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm, std::string& why);
bool valid_camera(const rt_camera* c, std::string& why);
void fill_params(const rt_metric* mt, const rt_integrator* ig, const rt_termination* tm, TraceParams& P);
void fill_camera(const rt_camera* c, TraceParams& P);
double kerr_isco_host(double m, double a_spin);

// Grow-only device result arrays of one context (SoA, one allocation).
struct DeviceResults {
    void* base = nullptr;
    size_t capacity = 0;
    OutPtrs ptrs{};
    ~DeviceResults();
    rt_status ensure(size_t n);
    void release();
    static size_t bytes_per_ray();
};

rt_status copy_results_to_host(const OutPtrs& d, size_t n, const rt_results* out, hipStream_t s);

}  // namespace rt
