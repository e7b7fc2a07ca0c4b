// oc_autotune.cpp -- launch-shape autotuner of a Mooney-Rivlin equations set.
// Times, on the set's own mesh(es) and current DOFs, every candidate of
//   * the element-matrix kernel's persistent grid (blocks per CU),
//   * the CSR row gather's prefetch depth (element rows per chunk),
//   * the SpMV's lanes per CSR row,
// keeps the fastest of each in the set (later launches use them) and
// reports the timings as JSON.
#include <cstdio>
#include <string>
#include <vector>

#include "oc_internal.h"

namespace oc {
namespace {

template <typename F>
int best_of(F launch, int reps, float *ms_out)
{
    Runtime &rt = runtime();
    hipEvent_t a = rt.events[62], b = rt.events[63];
    float best = 1e30f;
    OC_CHECK(launch());  // warm-up
    for (int r = 0; r < reps; ++r) {
        OC_HIP_CHECK(hipEventRecord(a, rt.stream));
        OC_CHECK(launch());
        OC_HIP_CHECK(hipEventRecord(b, rt.stream));
        OC_HIP_CHECK(hipEventSynchronize(b));
        float ms = 0.0f;
        OC_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
        if (ms < best) best = ms;
    }
    *ms_out = best;
    return OC_NO_ERROR;
}

}  // namespace
}  // namespace oc

using namespace oc;

extern "C" int oc_autotune(oc_equations_set h, int precision, char *report, int report_size)
{
    OC_CHECK(require_runtime());
    if (!h) return set_error(OC_ERR_INVALID_HANDLE, "equations set handle is NULL");
    EquationsSet &s = h->s;
    if (s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN)
        return set_error(OC_ERR_INVALID_ARGUMENT, "the autotuner covers Mooney-Rivlin sets");
    if (precision != OC_FP64 && precision != OC_FP32)
        return set_error(OC_ERR_INVALID_ARGUMENT, "precision must be OC_FP64 or OC_FP32");
    OC_CHECK(mr_jacobian_buffers(s, precision));
    OC_CHECK(launch_mr_gauss(s));
    std::string js = "{\"precision\": \"";
    js += precision == OC_FP64 ? "fp64" : "fp32";
    js += "\"";
    char buf[256];
    const int reps = 3;

    js += ", \"element_matrix_ms\": {";
    int best_bpc = 4;
    float best_ms = 1e30f;
    const int bpcs[] = {1, 2, 3, 4, 6, 8};
    for (int i = 0; i < 6; ++i) {
        float ms = 0.0f;
        OC_CHECK(best_of([&] { return launch_mr_element_matrices(s, precision, bpcs[i]); }, reps, &ms));
        std::snprintf(buf, sizeof buf, "%s\"%d\": %.4f", i ? ", " : "", bpcs[i], ms);
        js += buf;
        if (ms < best_ms) {
            best_ms = ms;
            best_bpc = bpcs[i];
        }
    }
    js += "}";

    js += ", \"row_gather_ms\": {";
    int best_chunk = 8;
    best_ms = 1e30f;
    const int chunks[] = {4, 8, 16};
    for (int i = 0; i < 3; ++i) {
        float ms = 0.0f;
        OC_CHECK(best_of([&] { return launch_mr_row_gather(s, precision, chunks[i]); }, reps, &ms));
        std::snprintf(buf, sizeof buf, "%s\"%d\": %.4f", i ? ", " : "", chunks[i], ms);
        js += buf;
        if (ms < best_ms) {
            best_ms = ms;
            best_chunk = chunks[i];
        }
    }
    js += "}";

    // SpMV: lanes per row (per node for node rows) x neighbours in flight per lane
    OC_CHECK(prepare_node_rows(s));
    js += ", \"spmv_ms\": {";
    DevBuf<double> y;
    OC_CHECK(y.alloc(s.n_rows));
    int best_lanes = 32, best_unroll = 1;
    best_ms = 1e30f;
    const int lanes[] = {8, 16, 32, 64};
    bool first = true;
    for (int i = 0; i < 4; ++i)
        for (int u = 1; u <= 2; ++u) {
            if (u == 2 && (s.node_rows != 1 || lanes[i] < 16)) continue;  // unrolled form: node rows only
            float ms = 0.0f;
            OC_CHECK(best_of(
                [&] { return launch_spmv_lanes(s, precision == OC_FP32, lanes[i], s.dofs.ptr, y.ptr, u); }, reps,
                &ms));
            std::snprintf(buf, sizeof buf, "%s\"%dx%d\": %.4f", first ? "" : ", ", lanes[i], u, ms);
            first = false;
            js += buf;
            if (ms < best_ms) {
                best_ms = ms;
                best_lanes = lanes[i];
                best_unroll = u;
            }
        }
    js += "}";
    // level-0 multigrid node kernel (residual form), when the rows have the node layout
    int best_smoother = -1;
    if (s.node_rows == 1 && !s.free_mask.ptr) {  // no boundary conditions yet: every DOF free
        OC_CHECK(s.free_mask.alloc(s.n_rows));
        OC_HIP_CHECK(hipMemset(s.free_mask.ptr, 1, s.n_rows));
    }
    if (s.node_rows == 1) {
        js += ", \"smoother_residual_ms\": {";
        DevBuf<double> r;
        OC_CHECK(r.alloc(static_cast<size_t>(3) * s.n_nodes));
        best_ms = 1e30f;
        const char *names[] = {"32x1", "32x2", "16x2", "16x4", "64x1"};
        for (int v = 0; v < 5; ++v) {
            float ms = 0.0f;
            OC_CHECK(best_of([&] { return launch_smoother_residual(s, precision, v, s.dofs.ptr, s.dofs.ptr, r.ptr); },
                             reps, &ms));
            std::snprintf(buf, sizeof buf, "%s\"%s\": %.4f", v ? ", " : "", names[v], ms);
            js += buf;
            if (ms < best_ms) {
                best_ms = ms;
                best_smoother = v;
            }
        }
        js += "}";
    }
    OC_HIP_CHECK(hipStreamSynchronize(runtime().stream));
    s.tuned_element_blocks_per_cu = best_bpc;
    s.tuned_gather_chunk = best_chunk;
    s.tuned_spmv_lanes = best_lanes;
    s.tuned_spmv_unroll = best_unroll;
    if (best_smoother >= 0) s.tuned_smoother = best_smoother;
    std::snprintf(buf, sizeof buf, ", \"chosen\": {\"element_blocks_per_cu\": %d, \"gather_chunk\": %d, "
                                   "\"spmv_lanes\": %d, \"spmv_unroll\": %d, \"smoother_variant\": %d}}",
                  best_bpc, best_chunk, best_lanes, best_unroll, best_smoother);
    js += buf;
    if (report && report_size > 0) std::snprintf(report, static_cast<size_t>(report_size), "%s", js.c_str());
    return OC_NO_ERROR;
}
