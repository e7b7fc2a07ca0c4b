// oc_autotune.cpp -- launch-shape autotuner of a Mooney-Rivlin equations set.
// Times, on the set's own mesh(es) and current DOFs, every candidate of
//   * the element-matrix kernel's persistent grid (blocks per CU),
//   * the CSR row gather's prefetch depth (element rows per chunk),
//   * the SpMV's lanes per CSR row and neighbours in flight per lane,
//   * the level-0 multigrid node kernel's shape,
// keeps in the set (for later launches) the fastest of each when it beats the
// default by more than the timing noise, and reports the timings as JSON.
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

// A candidate replaces the default shape only when it is faster by more than
// kMargin: the timings repeat to about 1-2%, and a choice inside that band is
// noise.
constexpr float kMargin = 0.03f;

struct Choice {
    int index;       // chosen candidate
    float ms;        // its best time
    float default_ms;
};

// Times candidates 0..n-1 (best of `reps` each), appends "name": {label: ms,
// ...} to js and picks the fastest unless it beats candidate `def` by less
// than kMargin.
template <typename Launch, typename Label>
int tune(std::string &js, const char *name, int n, int def, int reps, Launch launch, Label label, Choice *out)
{
    char buf[128];
    std::vector<float> ms(static_cast<size_t>(n));
    js += ", \"";
    js += name;
    js += "\": {";
    for (int i = 0; i < n; ++i) {
        OC_CHECK(best_of([&] { return launch(i); }, reps, &ms[static_cast<size_t>(i)]));
        std::snprintf(buf, sizeof buf, "%s\"%s\": %.4f", i ? ", " : "", label(i).c_str(), ms[static_cast<size_t>(i)]);
        js += buf;
    }
    js += "}";
    int best = def;
    for (int i = 0; i < n; ++i)
        if (ms[static_cast<size_t>(i)] < ms[static_cast<size_t>(best)]) best = i;
    if (ms[static_cast<size_t>(best)] > (1.0f - kMargin) * ms[static_cast<size_t>(def)]) best = def;
    *out = {best, ms[static_cast<size_t>(best)], ms[static_cast<size_t>(def)]};
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
    const bool fp32 = precision == OC_FP32;
    std::string js = "{\"precision\": \"";
    js += fp32 ? "fp32" : "fp64";
    js += "\"";
    const int reps = 3;
    auto num = [](int v) { return std::to_string(v); };

    // element-matrix grid (blocks per CU; default 1)
    const int bpcs[] = {1, 2, 3, 4, 6, 8};
    Choice c_elem{};
    OC_CHECK(tune(js, "element_matrix_ms", 6, 0, reps,
                  [&](int i) { return launch_mr_element_matrices(s, precision, bpcs[i]); },
                  [&](int i) { return num(bpcs[i]); }, &c_elem));
    // CSR row gather prefetch depth (element rows per chunk; default 8)
    const int chunks[] = {4, 8, 16};
    Choice c_gather{};
    OC_CHECK(tune(js, "row_gather_ms", 3, 1, reps, [&](int i) { return launch_mr_row_gather(s, precision, chunks[i]); },
                  [&](int i) { return num(chunks[i]); }, &c_gather));

    // SpMV: lanes per row (per node for node rows) x neighbours in flight per
    // lane; default 32 x 1 (FP32 values) or 16 x 2 (FP64) on node rows
    OC_CHECK(prepare_node_rows(s));
    DevBuf<double> y;
    OC_CHECK(y.alloc(s.n_rows));
    std::vector<std::pair<int, int>> spmv;
    for (int l : {8, 16, 32, 64})
        for (int u = 1; u <= 2; ++u)
            if (u == 1 || (s.node_rows == 1 && l >= 16)) spmv.push_back({l, u});  // unrolled form: node rows only
    const std::pair<int, int> spmv_def = s.node_rows == 1 && !fp32 ? std::make_pair(16, 2) : std::make_pair(32, 1);
    int spmv_def_i = 0;
    for (size_t i = 0; i < spmv.size(); ++i)
        if (spmv[i] == spmv_def) spmv_def_i = static_cast<int>(i);
    Choice c_spmv{};
    OC_CHECK(tune(js, "spmv_ms", static_cast<int>(spmv.size()), spmv_def_i, reps,
                  [&](int i) {
                      return launch_spmv_lanes(s, fp32, spmv[static_cast<size_t>(i)].first, s.dofs.ptr, y.ptr,
                                               spmv[static_cast<size_t>(i)].second);
                  },
                  [&](int i) {
                      return num(spmv[static_cast<size_t>(i)].first) + "x" + num(spmv[static_cast<size_t>(i)].second);
                  },
                  &c_spmv));

    // level-0 multigrid node kernel (residual form), when the rows have the
    // node layout: FP64 values (default 16 x 2), or for an FP32 Jacobian the
    // packed bf16 copy its smoother reads (default 16 lanes x 1 quad)
    int best_smoother = -1;
    if (s.node_rows == 1 && !s.free_mask.ptr) {  // no boundary conditions yet: every DOF free
        OC_CHECK(s.free_mask.alloc(s.n_rows));
        OC_HIP_CHECK(hipMemset(s.free_mask.ptr, 1, s.n_rows));
    }
    if (s.node_rows == 1) {
        DevBuf<double> r;
        OC_CHECK(r.alloc(static_cast<size_t>(3) * s.n_nodes));
        const char *names64[] = {"32x1", "32x2", "16x2", "16x4", "64x1"};
        const char *names_bf[] = {"16x1q", "8x1q", "32x1q", "8x2q", "16x2q"};
        const char **names = fp32 ? names_bf : names64;
        if (fp32) {
            OC_CHECK(prepare_uu_bf16(s));
            OC_CHECK(launch_uu_pack(s));
        }
        Choice c_sm{};
        OC_CHECK(tune(js, "smoother_residual_ms", 5, fp32 ? 0 : 2, reps,
                      [&](int v) { return launch_smoother_residual(s, precision, v, s.dofs.ptr, s.dofs.ptr, r.ptr); },
                      [&](int v) { return std::string(names[v]); }, &c_sm));
        best_smoother = c_sm.index;
    }
    OC_HIP_CHECK(hipStreamSynchronize(runtime().stream));
    const int best_bpc = bpcs[c_elem.index], best_chunk = chunks[c_gather.index];
    const int best_lanes = spmv[static_cast<size_t>(c_spmv.index)].first;
    const int best_unroll = spmv[static_cast<size_t>(c_spmv.index)].second;
    s.tuned_element_blocks_per_cu = best_bpc;
    s.tuned_gather_chunk = best_chunk;
    s.tuned_spmv_lanes = best_lanes;
    s.tuned_spmv_unroll = best_unroll;
    if (best_smoother >= 0) s.tuned_smoother = best_smoother;
    char buf[256];
    std::snprintf(buf, sizeof buf, ", \"margin\": %.2f, \"chosen\": {\"element_blocks_per_cu\": %d, \"gather_chunk\": %d, "
                                   "\"spmv_lanes\": %d, \"spmv_unroll\": %d, \"smoother_variant\": %d}}",
                  kMargin, best_bpc, best_chunk, best_lanes, best_unroll, best_smoother);
    js += buf;
    if (report && report_size > 0) std::snprintf(report, static_cast<size_t>(report_size), "%s", js.c_str());
    return OC_NO_ERROR;
}
