// oc_api.cpp -- the C ABI of liboc_asm (declared in include/oc_asm.h):
// runtime, error reporting, equations-set lifetime, data movement and the
// assembly entry points that dispatch to the HIP kernels.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "oc_internal.h"

namespace oc {

namespace {
thread_local std::string g_last_error;
Runtime g_runtime;
std::mutex g_runtime_mutex;
}  // namespace

int set_error(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

Runtime &runtime() { return g_runtime; }

// ---- kernel timing -------------------------------------------------------------
namespace {
struct PendingTiming {
    std::string name;
    hipEvent_t start, stop;
};
struct TimingState {
    bool enabled = false;
    std::string only;  // when set, only this kernel is timed
    std::vector<PendingTiming> pending;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<std::string, std::pair<double, long long>>> totals;
} g_timing;

hipEvent_t timing_event()
{
    if (!g_timing.pool.empty()) {
        hipEvent_t e = g_timing.pool.back();
        g_timing.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int timing_resolve()
{
    if (g_timing.pending.empty()) return OC_NO_ERROR;
    OC_HIP_CHECK(hipEventSynchronize(g_timing.pending.back().stop));
    for (auto &p : g_timing.pending) {
        float ms = 0.0f;
        OC_HIP_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
        bool found = false;
        for (auto &t : g_timing.totals)
            if (t.first == p.name) {
                t.second.first += ms;
                t.second.second += 1;
                found = true;
                break;
            }
        if (!found) g_timing.totals.push_back({p.name, {static_cast<double>(ms), 1}});
        g_timing.pool.push_back(p.start);
        g_timing.pool.push_back(p.stop);
    }
    g_timing.pending.clear();
    return OC_NO_ERROR;
}
}  // namespace

namespace {
bool g_capturing = false;
}
void set_capturing(bool on) { g_capturing = on; }
bool capturing() { return g_capturing; }

KernelTimer::KernelTimer(const char *name) : name_(name)
{
    if (!g_timing.enabled || g_capturing) return;
    if (!g_timing.only.empty() && g_timing.only != name) return;
    start_ = timing_event();
    if (start_) (void)hipEventRecord(start_, g_runtime.stream);
}

KernelTimer::~KernelTimer()
{
    if (!start_) return;
    hipEvent_t stop = timing_event();
    if (!stop) return;
    (void)hipEventRecord(stop, g_runtime.stream);
    g_timing.pending.push_back({name_, start_, stop});
    if (g_timing.pending.size() > 4096) (void)timing_resolve();
}

int require_runtime()
{
    if (g_runtime.stream == nullptr)
        return set_error(OC_ERR_NOT_INITIALISED, "oc_initialise has not been called");
    return OC_NO_ERROR;
}

static int to_zero_based(const int *in, size_t n, int base, int limit, std::vector<int32_t> &out, const char *what)
{
    if (base != 0 && base != 1) return set_error(OC_ERR_INVALID_ARGUMENT, "index_base must be 0 or 1");
    out.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const int v = in[i] - base;
        if (v < 0 || v >= limit)
            return set_error(OC_ERR_INVALID_ARGUMENT, "%s[%zu] = %d out of range [%d, %d)", what, i, in[i], base,
                             limit + base);
        out[i] = v;
    }
    return OC_NO_ERROR;
}

static int finish_structure(EquationsSet &s, int ncomp, const int32_t *nn, const int32_t *nen,
                            const int32_t *const *conn)
{
    OC_CHECK(build_pattern(ncomp, s.n_elements, nn, nen, conn, s.h_row_ptr, s.h_col_idx));
    s.nnz = static_cast<int64_t>(s.h_col_idx.size());
    OC_CHECK(s.row_ptr.upload(s.h_row_ptr.data(), s.h_row_ptr.size()));
    OC_CHECK(s.col_idx.upload(s.h_col_idx.data(), s.h_col_idx.size()));
    OC_CHECK(s.dofs.alloc(s.n_rows));
    OC_CHECK(s.residual.alloc(s.n_rows));
    OC_CHECK(s.rhs.alloc(s.n_rows));
    OC_HIP_CHECK(hipMemset(s.dofs.ptr, 0, sizeof(double) * s.n_rows));
    OC_HIP_CHECK(hipMemset(s.residual.ptr, 0, sizeof(double) * s.n_rows));
    OC_HIP_CHECK(hipMemset(s.rhs.ptr, 0, sizeof(double) * s.n_rows));
    return OC_NO_ERROR;
}

static int check_handle(oc_equations_set h)
{
    if (h == nullptr) return set_error(OC_ERR_INVALID_HANDLE, "equations set handle is NULL");
    return require_runtime();
}

}  // namespace oc

using namespace oc;

extern "C" {

int oc_initialise(int device)
{
    std::lock_guard<std::mutex> lock(g_runtime_mutex);
    Runtime &rt = g_runtime;
    if (rt.stream != nullptr) {
        if (rt.device == device) return OC_NO_ERROR;
        return set_error(OC_ERR_INVALID_ARGUMENT, "already initialised on device %d", rt.device);
    }
    int count = 0;
    OC_HIP_CHECK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return set_error(OC_ERR_INVALID_ARGUMENT, "device %d not in [0, %d)", device, count);
    OC_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    OC_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(OC_ERR_UNSUPPORTED, "device %d is %s; this library is built for gfx950 only", device,
                         prop.gcnArchName);
    rt.num_cus = prop.multiProcessorCount;
    OC_HIP_CHECK(hipStreamCreateWithFlags(&rt.stream, hipStreamNonBlocking));
    for (auto &e : rt.events) OC_HIP_CHECK(hipEventCreate(&e));
    rt.device = device;
    // device-side set-up that later calls must not allocate: the basis tables
    // and the address of this device's copy of the allocation-free scratch
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    OC_CHECK(noalloc_scratch_init(device));
    return OC_NO_ERROR;
}

int oc_finalise(void)
{
    std::lock_guard<std::mutex> lock(g_runtime_mutex);
    Runtime &rt = g_runtime;
    if (rt.stream == nullptr) return OC_NO_ERROR;
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    for (auto &e : rt.events) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    (void)hipStreamDestroy(rt.stream);
    rt.stream = nullptr;
    rt.device = -1;
    return OC_NO_ERROR;
}

int oc_last_error(char *buffer, int buffer_size)
{
    if (buffer == nullptr || buffer_size <= 0) return OC_ERR_INVALID_ARGUMENT;
    std::snprintf(buffer, static_cast<size_t>(buffer_size), "%s", g_last_error.c_str());
    return OC_NO_ERROR;
}

int oc_device_get(int *device)
{
    if (!device) return set_error(OC_ERR_INVALID_ARGUMENT, "device is NULL");
    *device = g_runtime.device;
    return OC_NO_ERROR;
}

int oc_laplace_create(int n_nodes, const double *coords, int n_elements, const int *elem_nodes, int index_base,
                      oc_equations_set *equations_set)
{
    OC_CHECK(require_runtime());
    if (!coords || !elem_nodes || !equations_set || n_nodes <= 0 || n_elements <= 0)
        return set_error(OC_ERR_INVALID_ARGUMENT, "oc_laplace_create: invalid arguments");
    auto *h = new (std::nothrow) oc_equations_set_s();
    if (!h) return set_error(OC_ERR_ALLOCATION, "oc_laplace_create: out of host memory");
    EquationsSet &s = h->s;
    s.cls = OC_EQUATIONS_SET_LAPLACE;
    s.n_nodes = n_nodes;
    s.n_rows = n_nodes;
    s.n_elements = n_elements;
    s.nen = 8;
    std::vector<int32_t> conn;
    int rc = to_zero_based(elem_nodes, static_cast<size_t>(n_elements) * 8, index_base, n_nodes, conn, "elem_nodes");
    const int32_t nn[1] = {n_nodes}, nen[1] = {8};
    const int32_t *cp[1] = {conn.data()};
    if (rc == OC_NO_ERROR) rc = finish_structure(s, 1, nn, nen, cp);
    if (rc == OC_NO_ERROR) rc = s.conn.upload(conn.data(), conn.size());
    if (rc == OC_NO_ERROR) rc = s.coords.upload(coords, static_cast<size_t>(3) * n_nodes);
    if (rc == OC_NO_ERROR) rc = s.values.alloc(s.nnz);
    if (rc != OC_NO_ERROR) {
        delete h;
        return rc;
    }
    laplace_detect_grid(s, conn);
    *equations_set = h;
    return OC_NO_ERROR;
}

int oc_mooney_rivlin_create(int n_nodes_q, const double *coords_q, int n_nodes_l, int n_elements,
                            const int *elem_nodes_q, const int *elem_nodes_l, int index_base, double c10, double c01,
                            oc_equations_set *equations_set)
{
    return oc_mooney_rivlin_batch_create(1, &n_nodes_q, &n_nodes_l, &n_elements, coords_q, elem_nodes_q,
                                         elem_nodes_l, &c10, &c01, index_base, equations_set);
}

int oc_mooney_rivlin_batch_create(int n_meshes, const int *n_nodes_q, const int *n_nodes_l, const int *n_elements,
                                  const double *coords_q, const int *elem_nodes_q, const int *elem_nodes_l,
                                  const double *c10, const double *c01, int index_base,
                                  oc_equations_set *equations_set)
{
    OC_CHECK(require_runtime());
    if (n_meshes < 1 || !n_nodes_q || !n_nodes_l || !n_elements || !coords_q || !elem_nodes_q || !elem_nodes_l ||
        !c10 || !c01 || !equations_set)
        return set_error(OC_ERR_INVALID_ARGUMENT, "oc_mooney_rivlin_batch_create: invalid arguments");
    std::vector<int32_t> qo(n_meshes + 1, 0), lo(n_meshes + 1, 0), eo(n_meshes + 1, 0);
    for (int m = 0; m < n_meshes; ++m) {
        if (n_nodes_q[m] <= 0 || n_nodes_l[m] <= 0 || n_elements[m] <= 0)
            return set_error(OC_ERR_INVALID_ARGUMENT, "mesh %d: node and element counts must be positive", m);
        const int64_t q = static_cast<int64_t>(qo[m]) + n_nodes_q[m], l = static_cast<int64_t>(lo[m]) + n_nodes_l[m],
                      e = static_cast<int64_t>(eo[m]) + n_elements[m];
        if (3 * q + l >= INT32_MAX || e * 27 >= INT32_MAX)
            return set_error(OC_ERR_UNSUPPORTED, "batch exceeds int32 DOF or connectivity indices at mesh %d", m);
        qo[m + 1] = static_cast<int32_t>(q);
        lo[m + 1] = static_cast<int32_t>(l);
        eo[m + 1] = static_cast<int32_t>(e);
    }
    const int NQ = qo[n_meshes], NL = lo[n_meshes], NE = eo[n_meshes];
    auto *h = new (std::nothrow) oc_equations_set_s();
    if (!h) return set_error(OC_ERR_ALLOCATION, "oc_mooney_rivlin_batch_create: out of host memory");
    EquationsSet &s = h->s;
    s.cls = OC_EQUATIONS_SET_MOONEY_RIVLIN;
    s.n_nodes = NQ;
    s.n_nodes_l = NL;
    s.n_rows = 3 * NQ + NL;
    s.n_elements = NE;
    s.nen = 27;
    s.n_meshes = n_meshes;
    s.c10 = c10[0];
    s.c01 = c01[0];
    s.h_q_off = qo;
    s.h_l_off = lo;
    s.h_e_off = eo;
    // union numbering: mesh-local nodes shifted by the mesh's node offset;
    // coordinates re-laid separated over the union
    std::vector<int32_t> cq(static_cast<size_t>(NE) * 27), cl(static_cast<size_t>(NE) * 8);
    std::vector<double> X(static_cast<size_t>(3) * NQ), mat(static_cast<size_t>(2) * NE);
    int rc = OC_NO_ERROR;
    int64_t cpos = 0;
    for (int m = 0; m < n_meshes && rc == OC_NO_ERROR; ++m) {
        const int nq = n_nodes_q[m], nl = n_nodes_l[m], ne = n_elements[m];
        std::vector<int32_t> tq, tl;
        rc = to_zero_based(elem_nodes_q + static_cast<int64_t>(eo[m]) * 27, static_cast<size_t>(ne) * 27,
                           index_base, nq, tq, "elem_nodes_q");
        if (rc == OC_NO_ERROR)
            rc = to_zero_based(elem_nodes_l + static_cast<int64_t>(eo[m]) * 8, static_cast<size_t>(ne) * 8,
                               index_base, nl, tl, "elem_nodes_l");
        if (rc != OC_NO_ERROR) break;
        for (size_t k = 0; k < tq.size(); ++k) cq[static_cast<size_t>(eo[m]) * 27 + k] = tq[k] + qo[m];
        for (size_t k = 0; k < tl.size(); ++k) cl[static_cast<size_t>(eo[m]) * 8 + k] = tl[k] + lo[m];
        for (int c = 0; c < 3; ++c)
            for (int i = 0; i < nq; ++i) X[static_cast<size_t>(c) * NQ + qo[m] + i] = coords_q[cpos + c * nq + i];
        cpos += 3 * static_cast<int64_t>(nq);
        for (int e = 0; e < ne; ++e) {
            mat[2 * (static_cast<size_t>(eo[m]) + e)] = c10[m];
            mat[2 * (static_cast<size_t>(eo[m]) + e) + 1] = c01[m];
        }
    }
    // four dependent components: z1, z2, z3 on the quadratic mesh component,
    // p on the linear one (separated-component DOF order)
    const int32_t nn[4] = {NQ, NQ, NQ, NL}, nen[4] = {27, 27, 27, 8};
    const int32_t *cp[4] = {cq.data(), cq.data(), cq.data(), cl.data()};
    if (rc == OC_NO_ERROR) rc = finish_structure(s, 4, nn, nen, cp);
    if (rc == OC_NO_ERROR) rc = s.conn.upload(cq.data(), cq.size());
    if (rc == OC_NO_ERROR) rc = s.conn_l.upload(cl.data(), cl.size());
    if (rc == OC_NO_ERROR) rc = s.coords.upload(X.data(), X.size());
    if (rc == OC_NO_ERROR) rc = s.mat.upload(mat.data(), mat.size());
    if (rc == OC_NO_ERROR) rc = s.q_off.upload(qo.data(), qo.size());
    if (rc == OC_NO_ERROR) rc = s.l_off.upload(lo.data(), lo.size());
    if (rc == OC_NO_ERROR) rc = mr_prepare(s);  // scatter map and gather lists
    if (rc != OC_NO_ERROR) {
        delete h;
        return rc;
    }
    *equations_set = h;
    return OC_NO_ERROR;
}

int oc_equations_set_mesh_get(oc_equations_set h, int mesh, int *node_offset_q, int *n_nodes_q, int *node_offset_l,
                              int *n_nodes_l, int *element_offset, int *n_elements)
{
    OC_CHECK(check_handle(h));
    const EquationsSet &s = h->s;
    if (s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN || mesh < 0 || mesh >= s.n_meshes)
        return set_error(OC_ERR_INVALID_ARGUMENT, "mesh %d of a set with %d Mooney-Rivlin meshes", mesh,
                         s.cls == OC_EQUATIONS_SET_MOONEY_RIVLIN ? s.n_meshes : 0);
    if (node_offset_q) *node_offset_q = s.h_q_off[mesh];
    if (n_nodes_q) *n_nodes_q = s.h_q_off[mesh + 1] - s.h_q_off[mesh];
    if (node_offset_l) *node_offset_l = s.h_l_off[mesh];
    if (n_nodes_l) *n_nodes_l = s.h_l_off[mesh + 1] - s.h_l_off[mesh];
    if (element_offset) *element_offset = s.h_e_off[mesh];
    if (n_elements) *n_elements = s.h_e_off[mesh + 1] - s.h_e_off[mesh];
    return OC_NO_ERROR;
}

int oc_mesh_matrix_get(oc_equations_set h, int mesh, int precision, int index_base, int *n_rows, long long *nnz,
                       int *row_ptr, int *col_idx, void *values)
{
    OC_CHECK(check_handle(h));
    const EquationsSet &s = h->s;
    if (s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN || mesh < 0 || mesh >= s.n_meshes)
        return set_error(OC_ERR_INVALID_ARGUMENT, "mesh %d of a set with %d Mooney-Rivlin meshes", mesh,
                         s.cls == OC_EQUATIONS_SET_MOONEY_RIVLIN ? s.n_meshes : 0);
    if ((precision != OC_FP64 && precision != OC_FP32) || (index_base != 0 && index_base != 1))
        return set_error(OC_ERR_INVALID_ARGUMENT, "oc_mesh_matrix_get: precision or index_base");
    const int NQ = s.n_nodes;
    const int q0 = s.h_q_off[mesh], nq = s.h_q_off[mesh + 1] - q0;
    const int l0 = s.h_l_off[mesh], nl = s.h_l_off[mesh + 1] - l0;
    // the mesh's rows: four contiguous ranges of union rows (z1, z2, z3, p)
    const int first[4] = {q0, NQ + q0, 2 * NQ + q0, 3 * NQ + l0}, count[4] = {nq, nq, nq, nl};
    int64_t total = 0;
    for (int c = 0; c < 4; ++c) total += s.h_row_ptr[first[c] + count[c]] - s.h_row_ptr[first[c]];
    if (n_rows) *n_rows = 3 * nq + nl;
    if (nnz) *nnz = total;
    if ((row_ptr || col_idx) && total + index_base > INT32_MAX)
        return set_error(OC_ERR_UNSUPPORTED, "mesh %d has %lld nonzeros, beyond the 4-byte indices", mesh,
                         static_cast<long long>(total));
    const void *dev = precision == OC_FP64 ? static_cast<const void *>(s.values.ptr)
                                           : static_cast<const void *>(s.values_sp.ptr);
    if (values && !dev) return set_error(OC_ERR_INVALID_ARGUMENT, "no Jacobian of that precision assembled");
    const size_t vsize = precision == OC_FP64 ? sizeof(double) : sizeof(float);
    if (values) OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
    int64_t pos = 0;
    int row = 0;
    for (int c = 0; c < 4; ++c) {
        const int64_t k0 = s.h_row_ptr[first[c]], k1 = s.h_row_ptr[first[c] + count[c]];
        if (row_ptr)
            for (int i = 0; i < count[c]; ++i)
                row_ptr[row + i] = static_cast<int>(pos + s.h_row_ptr[first[c] + i] - k0 + index_base);
        if (col_idx)  // union column -> mesh-local column (the block is diagonal: every column is the mesh's own)
            for (int64_t k = k0; k < k1; ++k) {
                const int col = s.h_col_idx[k];
                const int local = col < 3 * NQ ? (col / NQ) * nq + (col % NQ - q0) : 3 * nq + (col - 3 * NQ - l0);
                col_idx[pos + k - k0] = local + index_base;
            }
        if (values && k1 > k0)
            OC_HIP_CHECK(hipMemcpy(static_cast<char *>(values) + pos * vsize,
                                   static_cast<const char *>(dev) + k0 * vsize, (k1 - k0) * vsize,
                                   hipMemcpyDeviceToHost));
        pos += k1 - k0;
        row += count[c];
    }
    if (row_ptr) row_ptr[row] = static_cast<int>(pos + index_base);
    return OC_NO_ERROR;
}

int oc_equations_set_subtype_set(oc_equations_set h, int subtype)
{
    OC_CHECK(check_handle(h));
    if (h->s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN)
        return set_error(OC_ERR_INVALID_ARGUMENT, "subtypes are defined for Mooney-Rivlin sets");
    if (subtype != OC_MOONEY_RIVLIN_SUBTYPE && subtype != OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE)
        return set_error(OC_ERR_INVALID_ARGUMENT, "unknown Mooney-Rivlin subtype %d", subtype);
    EquationsSet &s = h->s;
    if (s.subtype != subtype && s.step_graph) {  // the Newton-step graph holds the other subtype's kernels
        OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
        (void)hipGraphExecDestroy(s.step_graph);
        s.step_graph = nullptr;
    }
    s.subtype = subtype;
    return OC_NO_ERROR;
}

int oc_laplace_fused_set(oc_equations_set h, int fused, int *active)
{
    OC_CHECK(check_handle(h));
    if (h->s.cls != OC_EQUATIONS_SET_LAPLACE)
        return set_error(OC_ERR_INVALID_ARGUMENT, "the fused assembly form is defined for Laplace sets");
    if (fused < 0 || fused > 2) return set_error(OC_ERR_INVALID_ARGUMENT, "fused form %d not in [0, 2]", fused);
    h->s.lap_fused = fused;
    if (active) *active = laplace_fused_serves(h->s) ? 1 : 0;
    return OC_NO_ERROR;
}


int oc_mesh_count(oc_equations_set h, int *n_meshes)
{
    OC_CHECK(check_handle(h));
    if (!n_meshes) return set_error(OC_ERR_INVALID_ARGUMENT, "n_meshes is NULL");
    *n_meshes = h->s.n_meshes;
    return OC_NO_ERROR;
}

int oc_equations_set_destroy(oc_equations_set h)
{
    if (!h) return set_error(OC_ERR_INVALID_HANDLE, "equations set handle is NULL");
    if (g_runtime.stream) (void)hipStreamSynchronize(g_runtime.stream);
    delete h;
    return OC_NO_ERROR;
}

int oc_equations_set_info(oc_equations_set h, int *set_class, int *n_rows, long long *nnz, int *n_elements,
                          int *n_colours)
{
    OC_CHECK(check_handle(h));
    if (set_class) *set_class = h->s.cls;
    if (n_rows) *n_rows = h->s.n_rows;
    if (nnz) *nnz = h->s.nnz;
    if (n_elements) *n_elements = h->s.n_elements;
    if (n_colours) *n_colours = 0;  // assembly needs no element colouring
    return OC_NO_ERROR;
}

int oc_materials_set(oc_equations_set h, double c10, double c01)
{
    OC_CHECK(check_handle(h));
    if (h->s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN)
        return set_error(OC_ERR_INVALID_ARGUMENT, "materials are defined for Mooney-Rivlin sets only");
    std::vector<double> c10s(h->s.n_meshes, c10), c01s(h->s.n_meshes, c01);
    return oc_mesh_materials_set(h, c10s.data(), c01s.data());
}

int oc_mesh_materials_set(oc_equations_set h, const double *c10, const double *c01)
{
    OC_CHECK(check_handle(h));
    EquationsSet &s = h->s;
    if (s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN || !c10 || !c01)
        return set_error(OC_ERR_INVALID_ARGUMENT, "per-mesh materials of a Mooney-Rivlin set");
    std::vector<double> mat(static_cast<size_t>(2) * s.n_elements);
    for (int m = 0; m < s.n_meshes; ++m)
        for (int e = s.h_e_off[m]; e < s.h_e_off[m + 1]; ++e) {
            mat[2 * static_cast<size_t>(e)] = c10[m];
            mat[2 * static_cast<size_t>(e) + 1] = c01[m];
        }
    OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));  // no launch may still read the old values
    OC_HIP_CHECK(hipMemcpy(s.mat.ptr, mat.data(), mat.size() * sizeof(double), hipMemcpyHostToDevice));
    s.node_mat.release();
    s.c10 = c10[0];
    s.c01 = c01[0];
    return OC_NO_ERROR;
}

int oc_matrix_structure_get(oc_equations_set h, int index_base, int *row_ptr, int *col_idx)
{
    OC_CHECK(check_handle(h));
    if (index_base != 0 && index_base != 1) return set_error(OC_ERR_INVALID_ARGUMENT, "index_base must be 0 or 1");
    const EquationsSet &s = h->s;
    if (s.nnz + index_base > INT32_MAX)
        return set_error(OC_ERR_UNSUPPORTED, "%lld nonzeros exceed the 4-byte integer structure arrays",
                         static_cast<long long>(s.nnz));
    if (row_ptr)
        for (size_t i = 0; i < s.h_row_ptr.size(); ++i) row_ptr[i] = static_cast<int>(s.h_row_ptr[i] + index_base);
    if (col_idx)
        for (size_t i = 0; i < s.h_col_idx.size(); ++i) col_idx[i] = s.h_col_idx[i] + index_base;
    return OC_NO_ERROR;
}

int oc_dofs_set(oc_equations_set h, const double *dofs)
{
    OC_CHECK(check_handle(h));
    if (!dofs) return set_error(OC_ERR_INVALID_ARGUMENT, "dofs is NULL");
    h->s.plain_start_valid = false;
    OC_HIP_CHECK(hipMemcpyAsync(h->s.dofs.ptr, dofs, sizeof(double) * h->s.n_rows, hipMemcpyHostToDevice,
                                g_runtime.stream));
    OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
    return OC_NO_ERROR;
}

int oc_dofs_set_device(oc_equations_set h, const double *dofs_device)
{
    OC_CHECK(check_handle(h));
    if (!dofs_device) return set_error(OC_ERR_INVALID_ARGUMENT, "dofs_device is NULL");
    h->s.plain_start_valid = false;
    OC_HIP_CHECK(hipMemcpyAsync(h->s.dofs.ptr, dofs_device, sizeof(double) * h->s.n_rows, hipMemcpyDeviceToDevice,
                                g_runtime.stream));
    return OC_NO_ERROR;
}

int oc_dofs_get(oc_equations_set h, double *dofs)
{
    OC_CHECK(check_handle(h));
    if (!dofs) return set_error(OC_ERR_INVALID_ARGUMENT, "dofs is NULL");
    OC_HIP_CHECK(hipMemcpyAsync(dofs, h->s.dofs.ptr, sizeof(double) * h->s.n_rows, hipMemcpyDeviceToHost,
                                g_runtime.stream));
    OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
    return OC_NO_ERROR;
}

int oc_equations_set_assemble(oc_equations_set h)
{
    OC_CHECK(check_handle(h));
    if (h->s.cls != OC_EQUATIONS_SET_LAPLACE)
        return set_error(OC_ERR_INVALID_ARGUMENT, "linear assembly is defined for Laplace sets only");
    return launch_laplace_assemble(h->s);
}

int oc_equations_set_residual_evaluate(oc_equations_set h)
{
    OC_CHECK(check_handle(h));
    if (h->s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN)
        return set_error(OC_ERR_INVALID_ARGUMENT, "residual evaluation is defined for Mooney-Rivlin sets only");
    return launch_mr_residual(h->s);
}

int oc_equations_set_jacobian_evaluate(oc_equations_set h, int precision)
{
    OC_CHECK(check_handle(h));
    if (h->s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN)
        return set_error(OC_ERR_INVALID_ARGUMENT, "Jacobian evaluation is defined for Mooney-Rivlin sets only");
    if (precision != OC_FP64 && precision != OC_FP32)
        return set_error(OC_ERR_INVALID_ARGUMENT, "precision must be OC_FP64 or OC_FP32");
    EquationsSet &s = h->s;
    if (precision == OC_FP64 && s.values.ptr == nullptr) OC_CHECK(s.values.alloc(s.nnz));
    if (precision == OC_FP32 && s.values_sp.ptr == nullptr) OC_CHECK(s.values_sp.alloc(s.nnz));
    return launch_mr_jacobian(s, precision);
}

int oc_equations_set_jacobian_evaluate_fd(oc_equations_set h)
{
    OC_CHECK(check_handle(h));
    if (h->s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN)
        return set_error(OC_ERR_INVALID_ARGUMENT, "Jacobian evaluation is defined for Mooney-Rivlin sets only");
    return launch_mr_jacobian_fd(h->s);
}

int oc_matrix_values_get(oc_equations_set h, double *values)
{
    OC_CHECK(check_handle(h));
    if (!values) return set_error(OC_ERR_INVALID_ARGUMENT, "values is NULL");
    if (!h->s.values.ptr) return set_error(OC_ERR_INVALID_ARGUMENT, "no FP64 matrix has been assembled");
    OC_HIP_CHECK(hipMemcpyAsync(values, h->s.values.ptr, sizeof(double) * h->s.nnz, hipMemcpyDeviceToHost,
                                g_runtime.stream));
    OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
    return OC_NO_ERROR;
}

int oc_matrix_values_get_sp(oc_equations_set h, float *values)
{
    OC_CHECK(check_handle(h));
    if (!values) return set_error(OC_ERR_INVALID_ARGUMENT, "values is NULL");
    if (!h->s.values_sp.ptr) return set_error(OC_ERR_INVALID_ARGUMENT, "no FP32 matrix has been assembled");
    OC_HIP_CHECK(hipMemcpyAsync(values, h->s.values_sp.ptr, sizeof(float) * h->s.nnz, hipMemcpyDeviceToHost,
                                g_runtime.stream));
    OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
    return OC_NO_ERROR;
}

static int vector_get(const DevBuf<double> &b, int n, double *out)
{
    if (!out) return set_error(OC_ERR_INVALID_ARGUMENT, "output is NULL");
    OC_HIP_CHECK(hipMemcpyAsync(out, b.ptr, sizeof(double) * n, hipMemcpyDeviceToHost, g_runtime.stream));
    OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
    return OC_NO_ERROR;
}

int oc_residual_get(oc_equations_set h, double *residual)
{
    OC_CHECK(check_handle(h));
    return vector_get(h->s.residual, h->s.n_rows, residual);
}

int oc_rhs_get(oc_equations_set h, double *rhs)
{
    OC_CHECK(check_handle(h));
    return vector_get(h->s.rhs, h->s.n_rows, rhs);
}

int oc_device_pointer_get(oc_equations_set h, int buffer, void **device_pointer)
{
    OC_CHECK(check_handle(h));
    if (!device_pointer) return set_error(OC_ERR_INVALID_ARGUMENT, "device_pointer is NULL");
    EquationsSet &s = h->s;
    void *p = nullptr;
    switch (buffer) {
    case OC_BUFFER_MATRIX_VALUES: p = s.values.ptr; break;
    case OC_BUFFER_MATRIX_VALUES_SP: p = s.values_sp.ptr; break;
    case OC_BUFFER_ROW_PTR: p = s.row_ptr.ptr; break;
    case OC_BUFFER_COL_IDX: p = s.col_idx.ptr; break;
    case OC_BUFFER_DOFS: p = s.dofs.ptr; break;
    case OC_BUFFER_RESIDUAL: p = s.residual.ptr; break;
    case OC_BUFFER_RHS: p = s.rhs.ptr; break;
    case OC_BUFFER_SOLVER_VALUES: p = s.solver.values.ptr; break;
    case OC_BUFFER_SOLVER_VALUES_SP: p = s.solver.values_sp.ptr; break;
    case OC_BUFFER_SOLVER_ROW_PTR: p = s.solver.row_ptr.ptr; break;
    case OC_BUFFER_SOLVER_COL_IDX: p = s.solver.col_idx.ptr; break;
    case OC_BUFFER_SOLVER_VECTOR: p = s.solver.vector.ptr; break;
    default: return set_error(OC_ERR_INVALID_ARGUMENT, "unknown buffer %d", buffer);
    }
    if (!p) return set_error(OC_ERR_INVALID_ARGUMENT, "buffer %d has not been allocated yet", buffer);
    *device_pointer = p;
    return OC_NO_ERROR;
}

int oc_solver_matrices_assemble(oc_equations_set h, int precision)
{
    OC_CHECK(check_handle(h));
    return solver_matrices_assemble(h->s, precision);
}

static int solver_ready(oc_equations_set h)
{
    OC_CHECK(check_handle(h));
    if (!h->s.solver.built || h->s.solver.bc_version != h->s.bc_version)
        return set_error(OC_ERR_INVALID_ARGUMENT, "the solver matrices have not been assembled for the current "
                                                  "boundary conditions (oc_solver_matrices_assemble)");
    return OC_NO_ERROR;
}

int oc_solver_matrix_info(oc_equations_set h, int *n_rows, long long *nnz)
{
    OC_CHECK(solver_ready(h));
    if (n_rows) *n_rows = h->s.solver.n;
    if (nnz) *nnz = static_cast<long long>(h->s.solver.nnz);
    return OC_NO_ERROR;
}

int oc_solver_matrix_structure_get(oc_equations_set h, int index_base, int *row_ptr, int *col_idx)
{
    OC_CHECK(solver_ready(h));
    if (index_base != 0 && index_base != 1) return set_error(OC_ERR_INVALID_ARGUMENT, "index_base must be 0 or 1");
    const SolverSystem &S = h->s.solver;
    if (S.nnz + index_base > INT32_MAX)
        return set_error(OC_ERR_UNSUPPORTED, "%lld solver nonzeros exceed the 4-byte integer structure arrays",
                         static_cast<long long>(S.nnz));
    if (row_ptr) {
        std::vector<int64_t> p(static_cast<size_t>(S.n) + 1);
        OC_HIP_CHECK(hipMemcpyAsync(p.data(), S.row_ptr.ptr, p.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                                    g_runtime.stream));
        OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
        for (size_t i = 0; i < p.size(); ++i) row_ptr[i] = static_cast<int>(p[i] + index_base);
    }
    if (col_idx && S.nnz > 0) {
        OC_HIP_CHECK(hipMemcpyAsync(col_idx, S.col_idx.ptr, static_cast<size_t>(S.nnz) * sizeof(int),
                                    hipMemcpyDeviceToHost, g_runtime.stream));
        OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
        if (index_base)
            for (int64_t k = 0; k < S.nnz; ++k) col_idx[k] += index_base;
    }
    return OC_NO_ERROR;
}

int oc_solver_matrix_values_get(oc_equations_set h, double *values)
{
    OC_CHECK(solver_ready(h));
    if (!values) return set_error(OC_ERR_INVALID_ARGUMENT, "values is NULL");
    const SolverSystem &S = h->s.solver;
    if (!S.values.ptr && S.nnz) return set_error(OC_ERR_INVALID_ARGUMENT, "no FP64 solver matrix has been assembled");
    if (S.nnz)
        OC_HIP_CHECK(hipMemcpyAsync(values, S.values.ptr, static_cast<size_t>(S.nnz) * sizeof(double),
                                    hipMemcpyDeviceToHost, g_runtime.stream));
    OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
    return OC_NO_ERROR;
}

int oc_solver_matrix_values_get_sp(oc_equations_set h, float *values)
{
    OC_CHECK(solver_ready(h));
    if (!values) return set_error(OC_ERR_INVALID_ARGUMENT, "values is NULL");
    const SolverSystem &S = h->s.solver;
    if (!S.values_sp.ptr && S.nnz) return set_error(OC_ERR_INVALID_ARGUMENT, "no FP32 solver matrix has been assembled");
    if (S.nnz)
        OC_HIP_CHECK(hipMemcpyAsync(values, S.values_sp.ptr, static_cast<size_t>(S.nnz) * sizeof(float),
                                    hipMemcpyDeviceToHost, g_runtime.stream));
    OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
    return OC_NO_ERROR;
}

int oc_solver_vector_get(oc_equations_set h, double *vector)
{
    OC_CHECK(solver_ready(h));
    if (!vector) return set_error(OC_ERR_INVALID_ARGUMENT, "vector is NULL");
    const SolverSystem &S = h->s.solver;
    if (S.n)
        OC_HIP_CHECK(hipMemcpyAsync(vector, S.vector.ptr, static_cast<size_t>(S.n) * sizeof(double),
                                    hipMemcpyDeviceToHost, g_runtime.stream));
    OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
    return OC_NO_ERROR;
}

int oc_event_record(int slot)
{
    OC_CHECK(require_runtime());
    if (slot < 0 || slot >= 64) return set_error(OC_ERR_INVALID_ARGUMENT, "event slot %d not in [0,64)", slot);
    OC_HIP_CHECK(hipEventRecord(g_runtime.events[slot], g_runtime.stream));
    return OC_NO_ERROR;
}

int oc_event_elapsed(int start, int end, float *milliseconds)
{
    OC_CHECK(require_runtime());
    if (start < 0 || start >= 64 || end < 0 || end >= 64 || !milliseconds)
        return set_error(OC_ERR_INVALID_ARGUMENT, "invalid event slots");
    OC_HIP_CHECK(hipEventSynchronize(g_runtime.events[end]));
    OC_HIP_CHECK(hipEventElapsedTime(milliseconds, g_runtime.events[start], g_runtime.events[end]));
    return OC_NO_ERROR;
}

int oc_synchronise(void)
{
    OC_CHECK(require_runtime());
    OC_HIP_CHECK(hipStreamSynchronize(g_runtime.stream));
    return OC_NO_ERROR;
}

int oc_kernel_timing_enable(int enable)
{
    OC_CHECK(require_runtime());
    OC_CHECK(timing_resolve());
    g_timing.totals.clear();
    g_timing.enabled = enable != 0;
    return OC_NO_ERROR;
}

int oc_kernel_timing_select(const char *kernel_name)
{
    OC_CHECK(require_runtime());
    OC_CHECK(timing_resolve());
    g_timing.only = kernel_name ? kernel_name : "";
    return OC_NO_ERROR;
}

int oc_kernel_timing_get(const char *kernel_name, double *total_ms, long long *launches)
{
    OC_CHECK(require_runtime());
    if (!kernel_name || !total_ms || !launches) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_CHECK(timing_resolve());
    *total_ms = 0.0;
    *launches = 0;
    for (auto &t : g_timing.totals)
        if (t.first == kernel_name) {
            *total_ms = t.second.first;
            *launches = t.second.second;
        }
    return OC_NO_ERROR;
}

}  // extern "C"
