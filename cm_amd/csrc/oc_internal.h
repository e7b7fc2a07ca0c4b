// oc_internal.h -- internal declarations shared by the liboc_asm sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/oc_asm.h"

namespace oc {

// ---- error handling ---------------------------------------------------------
int set_error(int code, const char *fmt, ...);
#define OC_HIP_CHECK(call)                                                                              \
    do {                                                                                                \
        hipError_t _e = (call);                                                                         \
        if (_e != hipSuccess)                                                                           \
            return ::oc::set_error(OC_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                                   __FILE__, __LINE__);                                                 \
    } while (0)
#define OC_CHECK(call)              \
    do {                            \
        int _rc = (call);           \
        if (_rc != OC_NO_ERROR)     \
            return _rc;             \
    } while (0)

struct Runtime {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t events[64] = {};
    int num_cus = 0;
};
constexpr int kArnoldiHostSlots = 2 * 1024;
Runtime &runtime();
int require_runtime();

// Brackets one kernel launch with events on the library stream when kernel
// timing is enabled (oc_kernel_timing_enable); no-op otherwise.
class KernelTimer {
public:
    explicit KernelTimer(const char *name);
    ~KernelTimer();
    KernelTimer(const KernelTimer &) = delete;
    KernelTimer &operator=(const KernelTimer &) = delete;

private:
    const char *name_;
    hipEvent_t start_ = nullptr;
};

// Device buffer owned by the library.  Move-only: a copy would free the
// allocation twice (containers that grow move their elements).
template <typename T>
struct DevBuf {
    T *ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), n(o.n)
    {
        o.ptr = nullptr;
        o.n = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            ptr = o.ptr;
            n = o.n;
            o.ptr = nullptr;
            o.n = 0;
        }
        return *this;
    }
    int alloc(size_t count)
    {
        release();
        if (count == 0) return OC_NO_ERROR;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(T));
        if (e != hipSuccess) {
            ptr = nullptr;
            return set_error(OC_ERR_ALLOCATION, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                             hipGetErrorString(e));
        }
        n = count;
        return OC_NO_ERROR;
    }
    int upload(const T *host, size_t count)
    {
        OC_CHECK(alloc(count));
        if (count) OC_HIP_CHECK(hipMemcpy(ptr, host, count * sizeof(T), hipMemcpyHostToDevice));
        return OC_NO_ERROR;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

// Geometric multigrid hierarchy of the displacement block (oc_multigrid.hip).
// Level 0 is the triquadratic CSR operator itself; level 1 the trilinear
// space on the same elements (its nodes are the element corners, numbered as
// the pressure nodes); level l+1 coarsens level l by 2 per direction.  Levels
// >= 1 store a 27-point 3x3-block stencil per node (node-major vectors).
struct MgLevel {
    int nn = 0;                           // nodes over all meshes
    std::vector<int32_t> h_dim, h_off;    // nodes per direction / first node, per mesh (h_off: M+1)
    DevBuf<int32_t> dim, off;
    DevBuf<double> A;                     // nn*27*9: A[(node*27 + d)*9 + 3c + c']
    DevBuf<double> Dinv;                  // nn*9 inverse diagonal blocks
    DevBuf<double> As;                    // A(node, d) Dinv(neighbour d), levels with a coarser level
    DevBuf<uint8_t> freed;                // nn*3: DOF not fixed
    DevBuf<double> x, b, r;               // nn*3 each
};
struct Multigrid {
    std::vector<int32_t> h_n;             // elements per direction of each mesh
    DevBuf<int32_t> n, e_off;
    std::vector<MgLevel> levels;          // levels[0] holds only the per-mesh tables
    DevBuf<double> ke1;                   // Galerkin trilinear element matrices, E*24*24
    DevBuf<double> t0, r0, x0;            // level-0 scratch (3*n_nodes each)
    DevBuf<double> coarse_inv;            // per mesh dense inverse of the coarsest level
    int coarse_stride = 0;                // (max DOFs per mesh at the coarsest level)^2
    int pre = 1, post = 2;                // level-0 sweeps (oc_multigrid_options_set)
    // the trilinear and coarser stencil levels: one sweep each way (more
    // sweeps there cost Krylov iterations rather than saving them)
    int stencil_pre = 1, stencil_post = 1;
    double omega = 0.8;
    int coarse_sweeps = 0;                // 0: dense solve of the coarsest level
};

// Solver system (oc_solver_matrices.hip): the equations matrix or Jacobian
// restricted to the free DOFs.  index[d] = solver number of DOF d (-1 when
// fixed), rows[i] = equations row of solver row i; CSR with 64-bit offsets.
struct SolverSystem {
    bool built = false;
    bool columns_written = false;   // col_idx is written by the first values pass
    uint64_t bc_version = 0;        // boundary conditions the structure was built for
    int n = 0;
    int64_t nnz = 0;
    int precision = -1;             // precision of the last assembly
    DevBuf<int32_t> index, rows;
    DevBuf<int64_t> row_ptr;
    DevBuf<int32_t> col_idx;
    DevBuf<double> values;
    DevBuf<float> values_sp;
    DevBuf<double> vector;          // RHS with the fixed columns moved over (linear) or residual
};

// Per-equations-set state.  Element arrays are element-major; node and DOF
// indices are 0-based; CSR col_idx 0-based int32, row offsets int64 on the
// device (a batch of meshes can exceed 2^31 nonzeros).
//
// A Mooney-Rivlin set holds one mesh or a batch of independent meshes as the
// disjoint union: mesh m owns quadratic nodes [q_off[m], q_off[m+1]), pressure
// nodes [l_off[m], l_off[m+1]) and elements [e_off[m], e_off[m+1]); DOFs are
// component-major over the union (z1 of every node, z2, z3, then p), so the
// batch is one block-diagonal system and every kernel runs over all meshes
// in one launch.
struct EquationsSet {
    int cls = 0;
    int n_rows = 0;          // DOFs
    int64_t nnz = 0;
    int n_elements = 0;
    int nen = 0;             // element nodes of the main component (8 or 27)
    int n_nodes = 0;         // nodes of the main component
    int n_nodes_l = 0;       // pressure component nodes (MR)
    double c10 = 0.0, c01 = 0.0;  // materials of a single-mesh set
    int subtype = OC_MOONEY_RIVLIN_SUBTYPE;  // Mooney-Rivlin sets: constitutive subtype
    int n_meshes = 1;
    std::vector<int32_t> h_q_off, h_l_off, h_e_off;  // n_meshes + 1 each
    DevBuf<int32_t> q_off, l_off;
    DevBuf<double> mat;       // (c10, c01) per element
    DevBuf<double> node_mat;  // (c10, c01) per quadratic node (warm start)

    // host copies of the structure
    std::vector<int64_t> h_row_ptr;
    std::vector<int32_t> h_col_idx;

    DevBuf<double> coords;   // 3 * n_nodes, separated
    DevBuf<int32_t> conn;    // nen * E
    DevBuf<int32_t> conn_l;  // 8 * E (MR pressure component)
    DevBuf<int32_t> lin_to_quad;  // pressure node -> coincident quadratic node (warm start)
    DevBuf<int64_t> row_ptr;
    DevBuf<int32_t> col_idx;
    DevBuf<double> values;
    DevBuf<float> values_sp;
    // bf16 copy of the displacement block K_uu of an FP32 Jacobian, packed for
    // the level-0 smoother (node rows only): node i's neighbours, padded to a
    // multiple of four (padding: the node itself, value 0), start at
    // uu_off[i]; columns uu_col[uu_off[i] + p], values
    // uu_bf[9 uu_off[i] + (3 r + c) degp + p] for row r and component c
    DevBuf<int64_t> uu_off;
    DevBuf<int32_t> uu_col;
    DevBuf<uint16_t> uu_bf;
    DevBuf<float> uu_x4;  // the iterate the smoother gathers, per node (x, y, z, -) in FP32; two buffers
    DevBuf<double> dofs, residual, rhs;

    // Mooney-Rivlin: per-Gauss-point records (field-major, see
    // oc_elasticity.hip) and the CSR scatter map: element column j of row
    // dof_i of element e sits at row_ptr[dof_i] + (j / 27) * row_nb[dof_i] +
    // col_rank[(e*35 + k)*35 + j % 27] for j < 81 (3 * row_nb + rank of
    // pressure node j - 81 otherwise), k the element node of row i.
    DevBuf<double> gp_records;
    DevBuf<uint8_t> col_rank;
    DevBuf<uint16_t> row_nb;
    DevBuf<double> kbuf64;      // element matrices, E*89*89, row-major
    DevBuf<float> kbuf32;
    int max_row_len = 0;
    // element residuals (E*89) and the DOF -> element-entry gather lists
    DevBuf<double> relem;
    DevBuf<int32_t> gather_ptr, gather_idx;
    // Laplace: 8-bit offsets (in the row-gather order: entry p*8+j belongs to
    // the p-th element row of the gather lists); gather_idx holds the gather
    // position of element row e*8+i; element matrices in kbuf64 (E*64, rows
    // in gather order)
    DevBuf<uint8_t> slot_offsets8;
    // Laplace sets on a structured grid: elements per direction (zero when
    // the mesh is not one), and when the fused assembly serves it: 0 never,
    // 1 when the grid fills the GPU, 2 always (oc_laplace_fused_set)
    int lap_grid[3] = {0, 0, 0};
    int lap_fused = 1;

    // Dirichlet conditions (0-based DOFs, full values) and the free-DOF mask;
    // bc_version counts oc_boundary_conditions_set calls
    int n_bc = 0;
    uint64_t bc_version = 0;
    DevBuf<int32_t> bc_idx;
    DevBuf<double> bc_full;
    DevBuf<uint8_t> free_mask;
    // solver weights: wgt = 1/||F_ref|| of the row's mesh on free rows, 0 on
    // fixed rows; iwgt = 1/wgt on free rows, 0 on fixed rows
    DevBuf<double> wgt, iwgt;
    // solver workspace (oc_solver.hip)
    DevBuf<double> krylov;      // (restart+1) * n_rows
    DevBuf<double> work;        // vectors
    DevBuf<double> precond;     // preconditioner storage
    DevBuf<double> partials;    // reduction partial sums
    DevBuf<double> arnoldi;     // Hessenberg factor, rotations and coefficients of the Arnoldi process
    // Krylov progress read back without a stream synchronisation: pinned,
    // mapped, coherent host memory the Arnoldi kernels write (kArnoldiHostSlots
    // doubles, host and device addresses), and two timing-free events marking
    // alternate iterations; made at the first solve, per set
    double *arnoldi_host = nullptr;
    double *arnoldi_host_dev = nullptr;
    hipEvent_t arnoldi_events[2] = {};
    // Mooney-Rivlin sets: 1 when the displacement rows c*nq + i (c < 3) of
    // every node i hold three blocks of the same sorted neighbour nodes
    // (components 0, 1, 2) and then one common list of pressure columns, so
    // kernels can read one column list per node; 0 when not; -1 unchecked
    int node_rows = -1;
    DevBuf<int64_t> pstart;     // first pressure column of each displacement row
    int krylov_capacity = 0;

    // launch shapes (oc_autotune); 0 = default
    int tuned_element_blocks_per_cu = 0;
    int tuned_spmv_lanes = 0;
    int tuned_gather_chunk = 0;
    int tuned_smoother = -1;     // level-0 node-kernel shape (oc_multigrid.hip); -1 = default
    int tuned_spmv_unroll = 0;   // neighbours per lane in flight in the node-wise SpMV

    // Newton-step graph: Jacobian assembly + preconditioner setup, recorded for
    // {precision, preconditioner, element blocks per CU, gather chunk,
    // boundary-condition version}
    hipGraphExec_t step_graph = nullptr;
    long long step_graph_key[5] = {-1, -1, -1, -1, -1};

    // The plain start of a warm-started load step: the DOFs before
    // oc_mlp_warm_start, advanced by the same load increments; the Newton
    // fallback restarts from it (valid until the next solve or DOF upload)
    DevBuf<double> plain_start;
    bool plain_start_valid = false;

    SolverSystem solver;      // oc_solver_matrices_assemble
    Multigrid *mg = nullptr;  // set by oc_multigrid_structure_set
    ~EquationsSet()
    {
        if (step_graph) (void)hipGraphExecDestroy(step_graph);
        for (hipEvent_t e : arnoldi_events)
            if (e) (void)hipEventDestroy(e);
        if (arnoldi_host) (void)hipHostFree(arnoldi_host);
        delete mg;
    }
};

// ---- shared device resources (oc_assembly.hip) --------------------------------
struct BasisTables;
int ensure_tables(const BasisTables **out);
const BasisTables *device_basis_tables();
// address of this device's allocation-free scratch (oc_elasticity.hip), looked
// up once per device with the device current
int noalloc_scratch_init(int device);
template <typename T>
int fill(T *p, int64_t n, T v);

// ---- host-side setup helpers (oc_setup.cpp) ---------------------------------
int build_pattern(int ncomp, int n_elements, const int32_t *nn, const int32_t *nen, const int32_t *const *conn,
                  std::vector<int64_t> &row_ptr, std::vector<int32_t> &col_idx);

// ---- kernel launchers --------------------------------------------------------
int launch_laplace_assemble(EquationsSet &s);
void laplace_detect_grid(EquationsSet &s, const std::vector<int32_t> &conn);
bool laplace_fused_serves(const EquationsSet &s);  // the fused form assembles this set
int mr_prepare(EquationsSet &s);  // builds the scatter map once
int launch_mr_residual(EquationsSet &s);
int launch_mr_jacobian(EquationsSet &s, int precision);
int launch_mr_jacobian_fd(EquationsSet &s);  // forward differences of the element residuals
int mr_node_materials(EquationsSet &s);  // fills node_mat from mat
int mr_jacobian_buffers(EquationsSet &s, int precision);  // element-matrix and CSR value buffers
int launch_mr_element_matrices(EquationsSet &s, int precision, int blocks_per_cu);
int launch_mr_row_gather(EquationsSet &s, int precision, int chunk);
int launch_mr_gauss(EquationsSet &s);
// pressure-block Schur estimate from the pressure mass (analytic MR Jacobians:
// it reads the Gauss records of the last Jacobian assembly)
int launch_schur_mass(EquationsSet &s, double *inv_p);
// multigrid (oc_multigrid.hip): setup after a Jacobian assembly in
// `precision`; apply z_u = V-cycle(t) for the displacement block
int mg_setup(EquationsSet &s, int precision);
int mg_apply(EquationsSet &s, int precision, const double *t, double *z);
// pressure-column starts and the node-row layout check (oc_solver.hip)
int prepare_node_rows(EquationsSet &s);
// the packed bf16 displacement block's structure (uu_off, uu_col) and value
// buffer, once per set (oc_multigrid.hip)
int prepare_uu_bf16(EquationsSet &s);
int launch_uu_pack(EquationsSet &s);  // the FP32 Jacobian's values into it
// x4[j] = (x[j], x[nq + j], x[2 nq + j], 0) in FP32, on the runtime stream
void pack_x4(int nq, const double *x, float4 *x4);
// r = b - K_uu x with the level-0 node kernel of shape `variant` (autotuner)
int launch_smoother_residual(EquationsSet &s, int precision, int variant, const double *b, const double *x,
                             double *r);
// solver matrix and vector from the last assembled matrix (oc_solver_matrices.hip)
int solver_matrices_assemble(EquationsSet &s, int precision);
// y = A x with `lanes` lanes per CSR row (A in FP64 or FP32 values)
int launch_spmv_lanes(EquationsSet &s, bool fp32, int lanes, const double *x, double *y, int unroll = 1);

// Kernel timers are suspended while a stream capture records a graph.
void set_capturing(bool on);
bool capturing();

}  // namespace oc

// The opaque handle of the C ABI.
struct oc_equations_set_s {
    oc::EquationsSet s;
};
