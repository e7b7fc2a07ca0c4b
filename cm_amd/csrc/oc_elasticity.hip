// oc_elasticity.hip -- gfx950 Mooney-Rivlin finite elasticity assembly:
// residual and analytic Jacobian (finite_elasticity_routines.f90:842-1175,
// 1260-1640) scattered deterministically into the CSR equations matrix.
//
// Jacobian: mr_gauss_kernel (one thread per element Gauss point, all
// elements) evaluates the FP64 kinematics and tangent into Gauss-point
// records (field-major, see GpField); mr_element_matrix_f64_kernel (FP64) or
// mr_element_matrix_kernel<float> (FP32) contracts the records of every
// element into its 89x89 element matrix on MFMA; and csr_row_gather_kernel
// assembles each CSR row from the element rows that touch it in ascending
// element order (deterministic, write-once), placing them through the
// neighbour-rank scatter map.
// Residual: mr_residual_element_kernel evaluates element residuals of all
// elements into a buffer and dof_gather_kernel sums them per DOF in
// ascending element order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "oc_fe_device.h"
#include "oc_internal.h"

namespace oc {

// Gauss-point record fields.  Record of Gauss point q = e*27 + g stores
// field f at rec[f * n_gp + q] (coalesced across Gauss points).
enum GpField : int {
    GP_DXIDZ = 0,   // 9: DXI_DX of the deformed metrics, column-major
    GP_JGW = 9,     // J_z * w_g
    GP_C = 10,      // 36: spatial elasticity tensor (Voigt, column-major)
    GP_SIGJ = 46,   // 6: Cauchy stress from the elasticity routine
    GP_NFIELDS = 52
};

constexpr int kGaussElemsPerBlock = 9;  // 9 elements x 27 Gauss points = 243 threads

// Restates, per Gauss point, the interpolation/metrics/F/J block of the
// Jacobian routine (:954-989) and FINITE_ELASTICITY_GAUSS_ELASTICITY_TENSOR
// (:609-740).
__global__ void __launch_bounds__(256)
mr_gauss_kernel(int n_elements, const int *__restrict__ conn_q, const int *__restrict__ conn_l,
                const double *__restrict__ coords, int nq, const double *__restrict__ dofs,
                const double *__restrict__ mat, const BasisTables *__restrict__ T, double *__restrict__ rec)
{
    __shared__ double s_tab[kTabSize];
    __shared__ double s_X[kGaussElemsPerBlock][81];
    __shared__ double s_z[kGaussElemsPerBlock][81];
    __shared__ double s_p[kGaussElemsPerBlock][8];
    constexpr int kTabPer = (27 * 27 * 3 + 255) / 256;
    constexpr int kNodePer = (kGaussElemsPerBlock * 81 + 255) / 256;
    const int t = threadIdx.x;
    const int e0 = xcd_block(blockIdx.x, gridDim.x) * kGaussElemsPerBlock;
    // all global loads are issued before the LDS writes: the table and the
    // connectivity, then the nodal values they index
    double tab[kTabPer];
#pragma unroll
    for (int i = 0; i < kTabPer; ++i) {  // [ni][n][g] <- quad27[(g*27 + n)*4 + 1 + ni]
        const int k = t + 256 * i;
        const int ni = k / 729, n = (k / 27) % 27, g = k % 27;
        tab[i] = k < 27 * 27 * 3 ? T->quad27[(g * 27 + n) * 4 + 1 + ni] : 0.0;
    }
    int node[kNodePer];
#pragma unroll
    for (int i = 0; i < kNodePer; ++i) {
        const int k = t + 256 * i;
        const int e = e0 + k / 81;
        node[i] = (k < kGaussElemsPerBlock * 81 && e < n_elements) ? conn_q[e * 27 + k % 27] : -1;
    }
    const int pe = e0 + t / 8;
    const int pnode = (t < kGaussElemsPerBlock * 8 && pe < n_elements) ? conn_l[pe * 8 + t % 8] : -1;
    double xv[kNodePer], zv[kNodePer];
#pragma unroll
    for (int i = 0; i < kNodePer; ++i) {
        const int k = t + 256 * i, c = (k % 81) / 27;
        xv[i] = node[i] >= 0 ? coords[static_cast<int64_t>(c) * nq + node[i]] : 0.0;
        zv[i] = node[i] >= 0 ? dofs[static_cast<int64_t>(c) * nq + node[i]] : 0.0;
    }
    const double pval = pnode >= 0 ? dofs[static_cast<int64_t>(3) * nq + pnode] : 0.0;
#pragma unroll
    for (int i = 0; i < kTabPer; ++i) {
        const int k = t + 256 * i;
        const int ni = k / 729, n = (k / 27) % 27, g = k % 27;
        if (k < 27 * 27 * 3) s_tab[(ni * 27 + n) * kTabStride + g] = tab[i];
    }
#pragma unroll
    for (int i = 0; i < kNodePer; ++i) {
        const int k = t + 256 * i;
        if (node[i] >= 0) {
            s_X[k / 81][k % 81] = xv[i];
            s_z[k / 81][k % 81] = zv[i];
        }
    }
    if (pnode >= 0) s_p[t / 8][t % 8] = pval;
    __syncthreads();
    const int le = t / 27, g = t % 27;
    const int e = e0 + le;
    if (t >= kGaussElemsPerBlock * 27 || e >= n_elements) return;
    GaussKinematics k;
    gauss_kinematics(s_tab, g, s_z[le], s_X[le], k);
    double P = 0.0;
    for (int n = 0; n < 8; ++n) P += T->lin27[(g * 8 + n) * 4] * s_p[le][n];
    const int64_t n_gp = static_cast<int64_t>(n_elements) * 27;
    const int64_t q = static_cast<int64_t>(e) * 27 + g;
#pragma unroll
    for (int i = 0; i < 9; ++i) rec[(GP_DXIDZ + i) * n_gp + q] = k.dxidz[i];
    rec[GP_JGW * n_gp + q] = k.jz * T->w27[g];
    double c[36], sigma[6];
    mr_elasticity_tensor(k.F, k.Jznu, P, mat[2 * e], mat[2 * e + 1], c, sigma);
#pragma unroll
    for (int i = 0; i < 36; ++i) rec[(GP_C + i) * n_gp + q] = c[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) rec[(GP_SIGJ + i) * n_gp + q] = sigma[i];
}

// Scatter map.  A CSR row of node n (any of its three displacement rows, or
// a pressure row) holds, in ascending DOF order, the component-0 block of its
// nb displacement-node neighbours, the component-1 and component-2 blocks of
// the same nodes, then its pressure-node neighbours.  Element column j of
// row dof_i therefore sits at
//   (j / 27) * nb + rank[e][k][j % 27]           (j < 81)
//   3 * nb + rank[e][k][27 + j - 81]             (j >= 81),
// where k is the element node of row i (0..26 displacement, 27..34 pressure)
// and rank[e][k][m] is the position of element node m among the neighbours of
// node k: 35 x 35 bytes per element instead of 89 x 89 offsets, shared by the
// three displacement rows of a node.  nb is stored per row.
constexpr int kRankNodes = 35;

// nb of every row: the number of its columns below nq (component-0 block)
__global__ void __launch_bounds__(256)
row_neighbours_kernel(int n_rows, int nq, const int64_t *__restrict__ row_ptr, const int *__restrict__ col_idx,
                      uint16_t *__restrict__ nb)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    int64_t lo = row_ptr[r], hi = row_ptr[r + 1];  // first column >= nq
    const int64_t start = lo;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (col_idx[mid] < nq) lo = mid + 1;
        else hi = mid;
    }
    nb[r] = static_cast<uint16_t>(lo - start);
}

// rank[e][k][m], checked by bisection against every component pair it stands for
__global__ void __launch_bounds__(256)
column_rank_kernel(int n_elements, const int *__restrict__ conn_q, const int *__restrict__ conn_l, int nq,
                   const int64_t *__restrict__ row_ptr, const int *__restrict__ col_idx,
                   const uint16_t *__restrict__ nb, uint8_t *__restrict__ rank, int *__restrict__ n_bad)
{
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<int64_t>(n_elements) * kRankNodes * kRankNodes) return;
    const int e = static_cast<int>(gid / (kRankNodes * kRankNodes));
    const int k = static_cast<int>((gid / kRankNodes) % kRankNodes), m = static_cast<int>(gid % kRankNodes);
    const bool qrow = k < 27, qcol = m < 27;
    const int rnode = qrow ? conn_q[e * 27 + k] : conn_l[e * 8 + k - 27];
    const int cnode = qcol ? conn_q[e * 27 + m] : conn_l[e * 8 + m - 27];
    const int row0 = qrow ? rnode : 3 * nq + rnode;
    const int col0 = qcol ? cnode : 3 * nq + cnode;
    const int64_t slot0 = csr_slot(row_ptr, col_idx, row0, col0);
    const int64_t rk = slot0 - row_ptr[row0] - (qcol ? 0 : 3 * static_cast<int64_t>(nb[row0]));
    bool ok = slot0 >= 0 && rk >= 0 && rk < 256;
    for (int c = 0; c < (qrow ? 3 : 1) && ok; ++c) {
        const int row = qrow ? c * nq + rnode : row0;
        for (int cc = 0; cc < (qcol ? 3 : 1) && ok; ++cc) {
            const int col = qcol ? cc * nq + cnode : col0;
            const int64_t want = row_ptr[row] + (qcol ? cc : 3) * static_cast<int64_t>(nb[row]) + rk;
            ok = csr_slot(row_ptr, col_idx, row, col) == want;
        }
    }
    if (!ok) atomicAdd(n_bad, 1);
    rank[gid] = ok ? static_cast<uint8_t>(rk) : 0;
}

int mr_prepare(EquationsSet &s)
{
    if (s.col_rank.ptr) return OC_NO_ERROR;
    int64_t max_row = 0;
    for (int r = 0; r < s.n_rows; ++r) max_row = std::max(max_row, s.h_row_ptr[r + 1] - s.h_row_ptr[r]);
    if (max_row > 8192)
        return set_error(OC_ERR_UNSUPPORTED, "CSR row of %lld entries exceeds the row-gather capacity",
                         static_cast<long long>(max_row));
    if (static_cast<int64_t>(s.n_elements) * 89 >= INT32_MAX)
        return set_error(OC_ERR_UNSUPPORTED, "%d elements exceed the int32 element-entry gather lists", s.n_elements);
    s.max_row_len = static_cast<int>(max_row);
    Runtime &rt = runtime();
    OC_CHECK(s.row_nb.alloc(static_cast<size_t>(s.n_rows)));
    row_neighbours_kernel<<<static_cast<unsigned>((s.n_rows + 255) / 256), 256, 0, rt.stream>>>(
        s.n_rows, s.n_nodes, s.row_ptr.ptr, s.col_idx.ptr, s.row_nb.ptr);
    OC_HIP_CHECK(hipGetLastError());
    const int64_t nranks = static_cast<int64_t>(s.n_elements) * kRankNodes * kRankNodes;
    OC_CHECK(s.col_rank.alloc(static_cast<size_t>(nranks)));
    DevBuf<int> bad;
    OC_CHECK(bad.alloc(1));
    OC_HIP_CHECK(hipMemsetAsync(bad.ptr, 0, sizeof(int), rt.stream));
    column_rank_kernel<<<static_cast<unsigned>((nranks + 255) / 256), 256, 0, rt.stream>>>(
        s.n_elements, s.conn.ptr, s.conn_l.ptr, s.n_nodes, s.row_ptr.ptr, s.col_idx.ptr, s.row_nb.ptr,
        s.col_rank.ptr, bad.ptr);
    OC_HIP_CHECK(hipGetLastError());
    int h_bad = 0;
    OC_HIP_CHECK(hipMemcpyAsync(&h_bad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, rt.stream));
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    if (h_bad) {
        s.col_rank.release();
        return set_error(OC_ERR_PATTERN,
                         "%d element entries missing from the CSR structure, not in its component-block layout or past 255 neighbour nodes",
                         h_bad);
    }
    OC_CHECK(s.gp_records.alloc(static_cast<size_t>(s.n_elements) * 27 * GP_NFIELDS));
    OC_CHECK(s.relem.alloc(static_cast<size_t>(s.n_elements) * 89));
    // DOF -> element entries, ascending element order
    std::vector<int32_t> hconn(static_cast<size_t>(s.n_elements) * 27), hconn_l(static_cast<size_t>(s.n_elements) * 8);
    OC_HIP_CHECK(hipMemcpy(hconn.data(), s.conn.ptr, hconn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    OC_HIP_CHECK(hipMemcpy(hconn_l.data(), s.conn_l.ptr, hconn_l.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> ptr(static_cast<size_t>(s.n_rows) + 1, 0), idx(static_cast<size_t>(s.n_elements) * 89);
    // element DOF i of element e (component -> node order of the reference
    // element matrix, equations_matrices_routines.f90:684-941): 0..80
    // displacement, 81..88 pressure
    auto dof_of = [&](int64_t e, int i) {
        return i < 81 ? (i / 27) * s.n_nodes + hconn[e * 27 + i % 27] : 3 * s.n_nodes + hconn_l[e * 8 + i - 81];
    };
    for (int64_t e = 0; e < s.n_elements; ++e)
        for (int i = 0; i < 89; ++i) ptr[dof_of(e, i) + 1]++;
    for (int d = 0; d < s.n_rows; ++d) ptr[d + 1] += ptr[d];
    std::vector<int32_t> fillp(ptr.begin(), ptr.end() - 1);
    for (int64_t e = 0; e < s.n_elements; ++e)
        for (int i = 0; i < 89; ++i) idx[fillp[dof_of(e, i)]++] = static_cast<int32_t>(e * 89 + i);
    OC_CHECK(s.gather_ptr.upload(ptr.data(), ptr.size()));
    OC_CHECK(s.gather_idx.upload(idx.data(), idx.size()));
    return OC_NO_ERROR;
}

int launch_mr_gauss(EquationsSet &s)
{
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    const unsigned blocks = static_cast<unsigned>((s.n_elements + kGaussElemsPerBlock - 1) / kGaussElemsPerBlock);
    KernelTimer timer("mr_gauss_jacobian");
    mr_gauss_kernel<<<blocks, 256, 0, rt.stream>>>(s.n_elements, s.conn.ptr, s.conn_l.ptr, s.coords.ptr, s.n_nodes,
                                                   s.dofs.ptr, s.mat.ptr, T, s.gp_records.ptr);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

// ---------------------------------------------------------------------------
// Residual (:1491-1574), fused with the Gauss-point evaluation: one block of
// 9 elements x 27 Gauss points evaluates kinematics and the stress routine,
// then sums each element residual entry
//   R_(a,m) += sum_i DPHIDZ(i,m) JGW sigma(i,a),  R_(p,m) += J_X w (J-1) psi_m
// over the Gauss points in order into a per-element buffer.  With
// DPHIDZ(i,m) = sum_j dphi_m/dxi_j DXI_DX(j,i), each Gauss point keeps in LDS
// only T(j,a) = sum_i DXI_DX(j,i) JGW sigma(i,a), and the entry sums
// dphi_m/dxi_j T(j,a) over j and the Gauss points.  A gather kernel
// then adds the element contributions of every DOF in ascending element order
// (the reference's element-loop order, equations_set_routines.f90:5535-5568).
//
// With kFD, "element" v of the launch is the pair (e = v / 89, column col =
// v % 89): element e's residual with its DOF col raised by delta, written as
// column col of e's element matrix (R(x + delta e_col) - R(x)) / delta, R(x)
// from relem (EquationsSet_FiniteElementJacobianEvaluateFD,
// equations_set_routines.f90:3032-3150).
// 320 threads: 243 evaluate the Gauss points, then each of the 9 x 35
// residual work items has a thread of its own.
constexpr int kResThreads = 320;

template <bool kFD, int kSub>
__global__ void __launch_bounds__(kResThreads)
mr_residual_element_kernel(int n_elements, const int *__restrict__ conn_q, const int *__restrict__ conn_l,
                           const double *__restrict__ coords, int nq, const double *__restrict__ dofs,
                           const double *__restrict__ mat, const BasisTables *__restrict__ T,
                           double *__restrict__ relem, double delta, double *__restrict__ kfd)
{
    // The nodal values (X, z) are only read by the Gauss-point phase and the
    // Gauss-point results (DXI_DX, JGW sigma, the pressure weight) only
    // written after it, so they share LDS (the block stays under 64 KB).
    constexpr int kNodal = 2 * kGaussElemsPerBlock * 81;
    constexpr int kResults = 2 * kGaussElemsPerBlock * 9 * 27 + kGaussElemsPerBlock * 27;
    __shared__ double s_tab[kTabSize];
    __shared__ double s_p[kGaussElemsPerBlock][8];
    __shared__ double s_u[kNodal > kResults ? kNodal : kResults];
    double(*s_X)[81] = reinterpret_cast<double(*)[81]>(s_u);
    double(*s_z)[81] = reinterpret_cast<double(*)[81]>(s_u + kGaussElemsPerBlock * 81);
    double(*s_T)[9][27] = reinterpret_cast<double(*)[9][27]>(s_u);  // T(j, a): [3a + j][g]
    double(*s_rp)[27] = reinterpret_cast<double(*)[27]>(s_u + kGaussElemsPerBlock * 9 * 27);
    constexpr int kTabPer = (27 * 27 * 3 + kResThreads - 1) / kResThreads;
    constexpr int kNodePer = (kGaussElemsPerBlock * 81 + kResThreads - 1) / kResThreads;
    const int t = threadIdx.x;
    const int e0 = xcd_block(blockIdx.x, gridDim.x) * kGaussElemsPerBlock;
    // all global loads are issued before the LDS writes: the table and the
    // connectivity, then the nodal values they index
    double tab[kTabPer];
#pragma unroll
    for (int i = 0; i < kTabPer; ++i) {  // [ni][n][g] <- quad27[(g*27 + n)*4 + 1 + ni]
        const int k = t + kResThreads * i;
        const int ni = k / 729, n = (k / 27) % 27, g = k % 27;
        tab[i] = k < 27 * 27 * 3 ? T->quad27[(g * 27 + n) * 4 + 1 + ni] : 0.0;
    }
    // n_elements counts launch elements (element-column pairs with kFD)
    int node[kNodePer];
#pragma unroll
    for (int i = 0; i < kNodePer; ++i) {
        const int k = t + kResThreads * i;
        const int le = k / 81, n = k % 27, v = e0 + le;
        node[i] = (k < kGaussElemsPerBlock * 81 && v < n_elements) ? conn_q[(kFD ? v / 89 : v) * 27 + n] : -1;
    }
    const int pv = e0 + t / 8;
    const int pnode = (t < kGaussElemsPerBlock * 8 && pv < n_elements) ? conn_l[(kFD ? pv / 89 : pv) * 8 + t % 8] : -1;
    double xv[kNodePer], zv[kNodePer];
#pragma unroll
    for (int i = 0; i < kNodePer; ++i) {
        const int k = t + kResThreads * i, c = (k % 81) / 27;
        xv[i] = node[i] >= 0 ? coords[static_cast<int64_t>(c) * nq + node[i]] : 0.0;
        zv[i] = node[i] >= 0 ? dofs[static_cast<int64_t>(c) * nq + node[i]] : 0.0;
    }
    const double pval = pnode >= 0 ? dofs[static_cast<int64_t>(3) * nq + pnode] : 0.0;
#pragma unroll
    for (int i = 0; i < kTabPer; ++i) {
        const int k = t + kResThreads * i;
        const int ni = k / 729, n = (k / 27) % 27, g = k % 27;
        if (k < 27 * 27 * 3) s_tab[(ni * 27 + n) * kTabStride + g] = tab[i];
    }
#pragma unroll
    for (int i = 0; i < kNodePer; ++i) {
        const int k = t + kResThreads * i;
        const int le = k / 81, r = k % 81;
        if (node[i] >= 0) {
            s_X[le][r] = xv[i];
            s_z[le][r] = zv[i] + ((kFD && (e0 + le) % 89 == r) ? delta : 0.0);
        }
    }
    if (pnode >= 0) s_p[t / 8][t % 8] = pval + ((kFD && pv % 89 == 81 + t % 8) ? delta : 0.0);
    __syncthreads();
    {
        const int le = t / 27, g = t % 27;
        const int v = e0 + le, e = kFD ? v / 89 : v;
        const bool active = t < kGaussElemsPerBlock * 27 && v < n_elements;
        double rT[9], rrp = 0.0;
        if (active) {
            GaussKinematics k;
            gauss_kinematics(s_tab, g, s_z[le], s_X[le], k);
            double P = 0.0;
            for (int n = 0; n < 8; ++n) P += T->lin27[(g * 8 + n) * 4] * s_p[le][n];
            double sigma[6], jgw, J;
            if (kSub == OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE) {
                // the 4-component residual loop of the subtype (:1654-1761):
                // weight w Jxxi Jznu, Jznu from the stress routine
                J = imr_cauchy_stress(k.F, P, mat[2 * e], mat[2 * e + 1], sigma);
                jgw = T->w27[g] * k.jx * J;
            } else {
                mr_cauchy_stress(k.F, k.Jznu, P, mat[2 * e], mat[2 * e + 1], sigma);
                jgw = k.jz * T->w27[g];
                J = k.Jznu;
            }
            // dxidz[j + 3 i] = DXI_DX(j, i)
            for (int a = 0; a < 3; ++a)
                for (int j = 0; j < 3; ++j) {
                    double acc = 0.0;
                    for (int i = 0; i < 3; ++i) acc = fma(k.dxidz[j + 3 * i], jgw * sigma[t2v(i, a)], acc);
                    rT[3 * a + j] = acc;
                }
            rrp = k.jx * T->w27[g] * (J - 1.0);
        }
        __syncthreads();  // X and z consumed: their LDS takes the results
        if (active) {
            for (int i = 0; i < 9; ++i) s_T[le][i][g] = rT[i];
            s_rp[le][g] = rrp;
        }
    }
    __syncthreads();
    // one work item per (element, displacement node m), for its three
    // component rows; then one item per pressure row
    for (int w = t; w < kGaussElemsPerBlock * 35; w += kResThreads) {
        const int le = w / 35, q = w % 35;
        const int v = e0 + le;
        if (v >= n_elements) continue;
        double acc[3] = {0.0, 0.0, 0.0};
        if (q < 27) {
            const int m = q;
#pragma unroll 3
            for (int g = 0; g < 27; ++g) {
                const double f[3] = {s_tab[(0 * 27 + m) * kTabStride + g], s_tab[(1 * 27 + m) * kTabStride + g],
                                     s_tab[(2 * 27 + m) * kTabStride + g]};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double dot = fma(f[2], s_T[le][3 * a + 2][g], fma(f[1], s_T[le][3 * a + 1][g], f[0] * s_T[le][3 * a][g]));
                    acc[a] = acc[a] + dot;
                }
            }
        } else {
            const int m = q - 27;
#pragma unroll 1
            for (int g = 0; g < 27; ++g) acc[0] = acc[0] + s_rp[le][g] * T->lin27[(g * 8 + m) * 4];
        }
        const int nr = q < 27 ? 3 : 1;
        for (int a = 0; a < nr; ++a) {
            const int r = q < 27 ? a * 27 + q : 81 + (q - 27);
            if (kFD) {
                const int64_t e = v / 89, col = v % 89;
                kfd[(e * 89 + r) * 89 + col] = (acc[a] - relem[e * 89 + r]) / delta;
            } else {
                relem[static_cast<int64_t>(v) * 89 + r] = acc[a];
            }
        }
    }
}

// Schur-complement estimate of the pressure block, S = B K^-1 B^T ~ M_p / mu
// (the pressure mass matrix over the shear modulus, as for Stokes flow):
// lumped, in the current configuration, each element's part divided by its
// linearised shear modulus mu = 2 (c10 + c01).  One thread per pressure node
// sums its elements' integrals of psi_m J_z w (its gather list: entries
// e*89 + 81 + m, ascending elements); inv_p = -1 / S.
__global__ void __launch_bounds__(256)
schur_mass_kernel(int nl, const int *__restrict__ gptr, const int *__restrict__ gidx, int n_elements,
                  const double *__restrict__ rec, const BasisTables *__restrict__ T, const double *__restrict__ mat,
                  double *__restrict__ inv_p)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nl) return;
    const int64_t n_gp = static_cast<int64_t>(n_elements) * 27;
    double sum = 0.0;
    for (int k = gptr[p]; k < gptr[p + 1]; ++k) {
        const int64_t e = gidx[k] / 89;
        const int m = gidx[k] % 89 - 81;
        double acc = 0.0;
        for (int g = 0; g < 27; ++g) acc += T->lin27[(g * 8 + m) * 4] * rec[GP_JGW * n_gp + e * 27 + g];
        sum += acc / (2.0 * (mat[2 * e] + mat[2 * e + 1]));
    }
    inv_p[p] = sum > 0.0 ? -1.0 / sum : -1.0;
}

// residual[d] = sum over the element entries of DOF d, ascending elements.
__global__ void __launch_bounds__(256)
dof_gather_kernel(int n_rows, const int *__restrict__ ptr, const int *__restrict__ idx,
                  const double *__restrict__ relem, double *__restrict__ out)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n_rows) return;
    double acc = 0.0;
    for (int k = ptr[d]; k < ptr[d + 1]; ++k) acc = acc + relem[idx[k]];
    out[d] = acc;
}

// element residuals of the set's subtype into relem (kFD: the forward-
// difference columns, `blocks` covering every (element, column) pair)
template <bool kFD>
static int launch_residual_elements(const EquationsSet &s, unsigned blocks, const double *dofs,
                                    const BasisTables *T, double *relem, double delta, double *kfd)
{
    Runtime &rt = runtime();
    const int nv = kFD ? s.n_elements * 89 : s.n_elements;
    if (s.subtype == OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE)
        mr_residual_element_kernel<kFD, OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE><<<blocks, kResThreads, 0, rt.stream>>>(
            nv, s.conn.ptr, s.conn_l.ptr, s.coords.ptr, s.n_nodes, dofs, s.mat.ptr, T, relem, delta, kfd);
    else
        mr_residual_element_kernel<kFD, OC_MOONEY_RIVLIN_SUBTYPE><<<blocks, kResThreads, 0, rt.stream>>>(
            nv, s.conn.ptr, s.conn_l.ptr, s.coords.ptr, s.n_nodes, dofs, s.mat.ptr, T, relem, delta, kfd);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

int launch_mr_residual(EquationsSet &s)
{
    OC_CHECK(mr_prepare(s));
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    const unsigned blocks = static_cast<unsigned>((s.n_elements + kGaussElemsPerBlock - 1) / kGaussElemsPerBlock);
    {
        KernelTimer timer("mr_residual");
        OC_CHECK(launch_residual_elements<false>(s, blocks, s.dofs.ptr, T, s.relem.ptr, 0.0, nullptr));
    }
    {
        KernelTimer timer("dof_gather");
        dof_gather_kernel<<<static_cast<unsigned>((s.n_rows + 255) / 256), 256, 0, rt.stream>>>(
            s.n_rows, s.gather_ptr.ptr, s.gather_idx.ptr, s.relem.ptr, s.residual.ptr);
        OC_HIP_CHECK(hipGetLastError());
    }
    return OC_NO_ERROR;
}

// ---------------------------------------------------------------------------
// Analytic Jacobian (FiniteElasticity_FiniteElementJacobianEvaluate,
// finite_elasticity_routines.f90:842-1175) in two phases.
//
// Phase A (mr_element_matrix_f64_kernel for FP64, mr_element_matrix_kernel<float>
// for FP32; all elements, no colours): for each of the six lower displacement
// blocks (a, b), b <= a,
//   K_(a,m),(b,n) = sum_(g,j) U_g^{ab}[m][j] D_g[j][n],
//   U_g^{ab}[m][j] = sum_i D_g[i][m] W_g[(a,i),(b,j)],
//   W_g[(a,i),(b,j)] = J_z w_g (c(V(i,a),V(j,b)) + delta_ab sigma_ij),
// i.e. the reference's TEMPVEC / DOT_PRODUCT contraction (:1001-1037)
// re-associated as a 27x81 by 81x27 GEMM on MFMA; pressure rows
// K_(p,m),(b,n) = sum_g (J_z w_g psi_m) D_g[b][n] (:1039-1054) as an 8x27 by
// 27x81 GEMM.  Both kernels run a producer/consumer pipeline per block:
// producer waves build D, W and the pressure operand of the next element in
// double-buffered LDS while consumer waves contract the current one, with D in
// registers and U formed in registers from D and W.  Lower triangle kept,
// mirrored (:1149-1153); the element matrix goes to an HBM buffer, row-major.
// Phase B (csr_row_gather_kernel<Real>): one wavefront per CSR row adds the
// element rows that touch it, in ascending element order, into an LDS copy of
// the row and writes the row once (no zero fill, no read-modify-write).
typedef float mfma_f32x4 __attribute__((ext_vector_type(4)));
typedef double mfma_f64x4 __attribute__((ext_vector_type(4)));

template <typename Real>
struct Mfma;

template <>
struct Mfma<float> {
    using acc_t = mfma_f32x4;
    __device__ static acc_t step(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    // output row of accumulator register r for this lane (column = lane % 16)
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static constexpr int kRowStep = 1;
};

template <>
struct Mfma<double> {
    using acc_t = mfma_f64x4;
    __device__ static acc_t step(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
    static constexpr int kRowStep = 4;
};

// Per-lane operand slices of D_g[j][n] = sum_mi DXI_DX(mi, j) dphi_n/dxi_mi
// (:974-986).  The contraction index is k = (j, g) with g padded to 28, so
// MFMA k-step ks covers j = ks / 7 and g = 4 (ks % 7) + kk (kk = lane / 16).
// Lane (kk, l) keeps D_g[j][16 t + l] for its seven Gauss points g = 4 g' + kk
// and j = 0..2 in D[t][g'][j] (t = 0, 1: the two 16-wide tiles of the 27 basis
// functions).  These registers are the B operand of every displacement block
// and, contracted with W, form its A operand
//   U^{ab}_g[m][j] = sum_i D_g[i][m] W_g[(a,i),(b,j)]
// on the fly, so U never goes through LDS.
// Element-matrix stores go through a buffer resource over the 89x89 block of
// one element: the lane part of each offset is one register and the tile part
// a constant, so a store costs no address arithmetic.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
template <typename Real>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t element_rsrc(Real *Ke)
{
    return __builtin_amdgcn_make_buffer_rsrc(Ke, 0, 89 * 89 * static_cast<int>(sizeof(Real)), 0x00020000);
}
__device__ __forceinline__ void element_store(__amdgpu_buffer_rsrc_t rs, double v, int voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rs, voff, soff, 0);
}
__device__ __forceinline__ void element_store(__amdgpu_buffer_rsrc_t rs, float v, int voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, voff, soff, 0);
}

template <typename Real>
struct DSlice {
    Real v[2][7][3];
};

// One displacement block (a, b), row tile MT, column tiles 0 (and 1 if TWO):
// 21 k-steps of mfma 16x16x4, then the tiles go to the element buffer, the
// (b, a) block as the mirror image.  A diagonal block keeps its lower triangle
// (:1149-1153).
template <typename Real, int MT, bool TWO>
__device__ __forceinline__ void mr_uu_block(int a, int b, int lane, const DSlice<Real> &D, const Real (*s_W)[81],
                                            Real *__restrict__ Ke)
{
    using M = Mfma<Real>;
    const int kk = lane >> 4, l = lane & 15;
    typename M::acc_t acc0 = {Real(0), Real(0), Real(0), Real(0)};
    typename M::acc_t acc1 = {Real(0), Real(0), Real(0), Real(0)};
    // k-step s = 7 j + gp; software pipeline: the W loads of step s + 2 and
    // the A operand of step s + 1 are issued behind the MFMAs of step s, and a
    // scheduling barrier per step keeps that order (and the register count)
    const Real *wbase = &s_W[kk][a * 27 + b * 3];
    auto load_w = [&](int st, Real *w) {
        const Real *p = wbase + (4 * (st % 7)) * 81 + st / 7;
        w[0] = p[0];
        w[1] = p[9];
        w[2] = p[18];
    };
    auto a_operand = [&](int st, const Real *w) {
        const int gp = st % 7;
        return D.v[MT][gp][0] * w[0] + D.v[MT][gp][1] * w[1] + D.v[MT][gp][2] * w[2];
    };
    Real wq[2][3];
    load_w(0, wq[0]);
    load_w(1, wq[1]);
    Real u = a_operand(0, wq[0]);
#pragma unroll
    for (int st = 0; st < 21; ++st) {
        const int j = st / 7, gp = st % 7;
        if (st + 2 < 21) load_w(st + 2, wq[st & 1]);
        acc0 = M::step(u, D.v[0][gp][j], acc0);
        if (TWO) acc1 = M::step(u, D.v[1][gp][j], acc1);
        if (st + 1 < 21) u = a_operand(st + 1, wq[(st + 1) & 1]);
        __builtin_amdgcn_sched_barrier(0);
    }
    // entry (m, n) of the tile at (a*27 + m)*89 + b*27 + n and its mirror at
    // (b*27 + n)*89 + a*27 + m, with m = MT*16 + row0 + kRowStep*r and
    // n = nt*16 + l: lane part row0*89 + l (mirror l*89 + row0), the rest
    // constant
    const __amdgpu_buffer_rsrc_t rs = element_rsrc(Ke);
    const int row0 = M::row(lane, 0);
    const int vdir = (row0 * 89 + l) * static_cast<int>(sizeof(Real));
    const int vmir = (l * 89 + row0) * static_cast<int>(sizeof(Real));
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int m = MT * 16 + row0 + M::kRowStep * r;
        if (m >= 27) continue;
#pragma unroll
        for (int nt = 0; nt < (TWO ? 2 : 1); ++nt) {
            const int n = nt * 16 + l;
            if (n >= 27 || (a == b && m < n)) continue;
            const Real v = nt ? acc1[r] : acc0[r];
            const int cdir = ((a * 27 + MT * 16 + M::kRowStep * r) * 89 + b * 27 + nt * 16) * static_cast<int>(sizeof(Real));
            const int cmir = ((b * 27 + nt * 16) * 89 + a * 27 + MT * 16 + M::kRowStep * r) * static_cast<int>(sizeof(Real));
            element_store(rs, v, vdir, cdir);
            element_store(rs, v, vmir, cmir);
        }
    }
}

// Pressure rows K_(p,m),(b,n) = sum_g (J_z w_g psi_m(g)) D_g[b][n]
// (:1039-1054): k = g only, so k-step ks is g = 4 ks + kk and the B operand
// is D[t][ks][b]; six 16x16 tiles (b, t), rows 8..15 padding.  Also writes
// the zero p-p block.
template <typename Real>
__device__ __forceinline__ void mr_pressure_rows(int lane, const DSlice<Real> &D, const Real *s_Pa,
                                                 Real *__restrict__ Ke)
{
    using M = Mfma<Real>;
    const int kk = lane >> 4, l = lane & 15;
    Real pa[7];
#pragma unroll
    for (int ks = 0; ks < 7; ++ks) pa[ks] = s_Pa[(4 * ks + kk) * 16 + l];
    const __amdgpu_buffer_rsrc_t rs = element_rsrc(Ke);
    const int row0 = M::row(lane, 0);
    const int vdir = (row0 * 89 + l) * static_cast<int>(sizeof(Real));
    const int vmir = (l * 89 + row0) * static_cast<int>(sizeof(Real));
#pragma unroll
    for (int b = 0; b < 3; ++b) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            typename M::acc_t acc = {Real(0), Real(0), Real(0), Real(0)};
#pragma unroll
            for (int ks = 0; ks < 7; ++ks) acc = M::step(pa[ks], D.v[t][ks][b], acc);
            const int n = t * 16 + l;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = row0 + M::kRowStep * r;
                if (m >= 8 || n >= 27) continue;
                const int cdir = ((81 + M::kRowStep * r) * 89 + b * 27 + t * 16) * static_cast<int>(sizeof(Real));
                const int cmir = ((b * 27 + t * 16) * 89 + 81 + M::kRowStep * r) * static_cast<int>(sizeof(Real));
                element_store(rs, acc[r], vdir, cdir);
                element_store(rs, acc[r], vmir, cmir);
            }
        }
    }
    Ke[(81 + lane / 8) * 89 + 81 + lane % 8] = Real(0);
}

// Producer/consumer element-matrix kernel: one block of eight waves walks its
// elements in a two-stage pipeline.  Waves 4..7 (producers) build, for
// element i + 1, the operands D, W and the pressure A operand in one half of
// double-buffered LDS and load the Gauss-point records of element i + 2;
// waves 0..3 (consumers) meanwhile load their D slices of element i from the
// other half and run its MFMA contraction.  One block barrier per element.
template <typename Real>
__global__ void __launch_bounds__(512)
mr_element_matrix_kernel(int n_elements, const double *__restrict__ rec, const BasisTables *__restrict__ T,
                         Real *__restrict__ kbuf)
{
    __shared__ double s_tab[27][3][27];  // dphi_n/dxi_mi at g: [g][mi][n]
    __shared__ double s_psi[27][8];      // trilinear psi_m(g)
    __shared__ double s_rec[2][52][27];  // Gauss-point records
    __shared__ Real s_D[2][28][3][32];   // D_g[j][n]; g = 27 and n >= 27 zero
    __shared__ Real s_W[2][28][81];      // W_g[(a,i),(b,j)] at [g][(3a+i)*9 + 3b+j]; g = 27 zero
    __shared__ Real s_Pa[2][28 * 16];    // J_z w_g psi_m(g) at [g*16 + m]; m >= 8 and g = 27 zero
    constexpr int kRec = 52 * 27, kRecPerThread = (kRec + 255) / 256;
    const int t = threadIdx.x, lane = t & 63;
    const bool producer = t >= 256;
    const int pt = t - 256;
    // producer thread pt < 243 owns column rq = pt % 81 of W_g (row (a,i),
    // column (b,j)) and of D_g (j = rq / 27, n = rq % 27) for g = pg (mod 3)
    const int rq = (pt >= 0 ? pt : 0) % 81, pg = (pt >= 0 ? pt : 0) / 81;
    const int wa = rq / 27, wi = (rq / 9) % 3, wb = (rq % 9) / 3, wj = rq % 3;
    const int wf1 = GP_C + t2v(wi, wa) + 6 * t2v(wj, wb), wf2 = GP_SIGJ + t2v(wi, wj);
    const bool wdiag = wa == wb;
    const int dj = rq / 27, dn = rq % 27;
    const int64_t n_gp = static_cast<int64_t>(n_elements) * 27;
    const int n_mine = static_cast<int>(blockIdx.x) < n_elements
                           ? (n_elements - 1 - static_cast<int>(blockIdx.x)) / static_cast<int>(gridDim.x) + 1
                           : 0;
    for (int k = t; k < 27 * 27 * 3; k += 512) {
        const int g = k / 81, mi = (k / 27) % 3, n = k % 27;
        s_tab[g][mi][n] = T->quad27[(g * 27 + n) * 4 + 1 + mi];
    }
    for (int k = t; k < 27 * 8; k += 512) s_psi[k / 8][k % 8] = T->lin27[k * 4];
    for (int k = t; k < 2 * 28 * 3 * 32; k += 512) (&s_D[0][0][0][0])[k] = Real(0);
    for (int k = t; k < 2 * 28 * 81; k += 512) (&s_W[0][0][0])[k] = Real(0);
    for (int k = t; k < 2 * 28 * 16; k += 512) (&s_Pa[0][0])[k] = Real(0);
    if (producer && n_mine > 0) {
        const int64_t q0 = static_cast<int64_t>(blockIdx.x) * 27;
        for (int k = pt; k < kRec; k += 256) s_rec[0][k / 27][k % 27] = rec[(k / 27) * n_gp + q0 + (k % 27)];
    }
    __syncthreads();

    for (int it = 0; it <= n_mine; ++it) {
        if (producer) {
            if (it < n_mine) {
                const int p = it & 1;
                // records of the element after this one: issued first, stored last
                const int e_next = static_cast<int>(blockIdx.x) + (it + 1) * static_cast<int>(gridDim.x);
                double pre[kRecPerThread];
#pragma unroll
                for (int i = 0; i < kRecPerThread; ++i) {
                    const int k = pt + 256 * i;
                    pre[i] = (it + 1 < n_mine && k < kRec)
                                 ? rec[(k / 27) * n_gp + static_cast<int64_t>(e_next) * 27 + (k % 27)]
                                 : 0.0;
                }
                const double(*R)[27] = s_rec[p];
                // all LDS reads of the phase before its writes: one round trip
                if (pt < 243) {  // column rq of W_g and of D_g for g = pg + 3k
                    Real wv[9], dv[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        const int g = pg + 3 * k;
                        double v = R[wf1][g];
                        if (wdiag) v = v + R[wf2][g];
                        wv[k] = static_cast<Real>(R[GP_JGW][g] * v);
                        dv[k] = static_cast<Real>(R[GP_DXIDZ + 0 + 3 * dj][g] * s_tab[g][0][dn] +
                                                  R[GP_DXIDZ + 1 + 3 * dj][g] * s_tab[g][1][dn] +
                                                  R[GP_DXIDZ + 2 + 3 * dj][g] * s_tab[g][2][dn]);
                    }
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        s_W[p][pg + 3 * k][rq] = wv[k];
                        s_D[p][pg + 3 * k][dj][dn] = dv[k];
                    }
                }
                if (pt < 27 * 8) s_Pa[p][pt / 8 * 16 + pt % 8] = static_cast<Real>(R[GP_JGW][pt / 8] * s_psi[pt / 8][pt % 8]);
                if (it + 1 < n_mine) {
#pragma unroll
                    for (int i = 0; i < kRecPerThread; ++i) {
                        const int k = pt + 256 * i;
                        if (k < kRec) s_rec[p ^ 1][k / 27][k % 27] = pre[i];
                    }
                }
            }
        } else if (it >= 1) {
            const int p = (it - 1) & 1;
            const int e = static_cast<int>(blockIdx.x) + (it - 1) * static_cast<int>(gridDim.x);
            Real *Ke = kbuf + static_cast<int64_t>(e) * 89 * 89;
            // the lane index is made opaque per element so that the W and store
            // offsets derived from it are recomputed here rather than hoisted
            // out of the element loop and held in registers across it
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const int kk = ln >> 4, l = ln & 15, wave = t >> 6;
            DSlice<Real> D;
#pragma unroll
            for (int tt = 0; tt < 2; ++tt)
#pragma unroll
                for (int gp = 0; gp < 7; ++gp)
#pragma unroll
                    for (int j = 0; j < 3; ++j) D.v[tt][gp][j] = s_D[p][4 * gp + kk][j][16 * tt + l];
            const Real(*W)[81] = s_W[p];
            // 483 MFMA k-steps per element over four waves (126/126/126/105)
            if (wave == 0) {
                mr_uu_block<Real, 0, true>(1, 0, ln, D, W, Ke);
                mr_uu_block<Real, 1, true>(1, 0, ln, D, W, Ke);
                mr_uu_block<Real, 1, true>(0, 0, ln, D, W, Ke);
            } else if (wave == 1) {
                mr_uu_block<Real, 0, true>(2, 0, ln, D, W, Ke);
                mr_uu_block<Real, 1, true>(2, 0, ln, D, W, Ke);
                mr_uu_block<Real, 1, true>(1, 1, ln, D, W, Ke);
            } else if (wave == 2) {
                mr_uu_block<Real, 0, true>(2, 1, ln, D, W, Ke);
                mr_uu_block<Real, 1, true>(2, 1, ln, D, W, Ke);
                mr_uu_block<Real, 1, true>(2, 2, ln, D, W, Ke);
            } else {
                mr_pressure_rows<Real>(ln, D, s_Pa[p], Ke);
                mr_uu_block<Real, 0, false>(0, 0, ln, D, W, Ke);
                mr_uu_block<Real, 0, false>(1, 1, ln, D, W, Ke);
                mr_uu_block<Real, 0, false>(2, 2, ln, D, W, Ke);
            }
        }
        __syncthreads();
    }
}

// FP64 element matrices on v_mfma_f64_4x4x4f64 (four 4x4x4 blocks per
// instruction; lane = c + 4 blk + 16 h holds A_blk[c][h], B_blk[h][c] and the
// result D_blk[h][c]).  The FP64 16x16x4 instruction runs at about 60% of the
// 4x4x4 rate here, and the 4-wide tiles also pad 27 to 28 instead of 32.
// Same producer/consumer pipeline as mr_element_matrix_kernel.  A consumer
// instruction covers the row tiles I = 4t + blk of one displacement block
// (rows m = 16t + 4blk + i) against one column tile J (n = 4J + j):
//   A = U^{ab}_g[16t + l][j_dim] from the lane's D slice and W (l = c + 4 blk),
//   B = D_g[j_dim][4J + c], shared by the four row tiles, read from LDS;
// k-step s covers j_dim = s / 7 and g = 4 (s % 7) + h.  Lower block (a, b)
// is numbered q = a (a + 1) / 2 + b (F64Unit).

__device__ __forceinline__ double2 lds_pair(const double *p) { return *reinterpret_cast<const double2 *>(p); }

template <int Q, int T, int NJ>
struct F64Unit {
    static constexpr int q = Q, t = T, nj = NJ;
    static constexpr int a = Q == 0 ? 0 : (Q <= 2 ? 1 : 2);
    static constexpr int b = Q - (a * (a + 1)) / 2;
};

// stores the tiles of one unit: acc[J] at m = 16T + 4blk + h, n = 4J + c
template <typename U>
__device__ __forceinline__ void f64_unit_store(const double *acc, int blk, int h, int c, __amdgpu_buffer_rsrc_t rs)
{
    const int m = 16 * U::t + 4 * blk + h;
    const int vdir = ((4 * blk + h) * 89 + c) * 8, vmir = (c * 89 + 4 * blk + h) * 8;
#pragma unroll
    for (int J = 0; J < U::nj; ++J) {
        const int n = 4 * J + c;
        if (m >= 27 || n >= 27 || (U::a == U::b && m < n)) continue;
        element_store(rs, acc[J], vdir, ((U::a * 27 + 16 * U::t) * 89 + U::b * 27 + 4 * J) * 8);
        element_store(rs, acc[J], vmir, ((U::b * 27 + 4 * J) * 89 + U::a * 27 + 16 * U::t) * 8);
    }
}

// three units of one wave, all on row half T, sharing the B reads of every
// k-step; the B and W operands of step s + 1 are loaded while the MFMAs of
// step s run
template <int T, typename U0, typename U1, typename U2>
__device__ __forceinline__ void f64_units(int lane, const double (&DT)[7][3], const double *sD, const double *sW,
                                          __amdgpu_buffer_rsrc_t rs)
{
    const int c = lane & 3, blk = (lane >> 2) & 3, h = lane >> 4;
    double acc0[7], acc1[7], acc2[7];
#pragma unroll
    for (int J = 0; J < 7; ++J) acc0[J] = acc1[J] = acc2[J] = 0.0;
    double B[2][7], W[2][3][3];
    auto load = [&](int st, double (&b)[7], double (&w)[3][3]) {
        const int jd = st / 7, g = 4 * (st % 7) + h;
        const double *bp = sD + ((g * 3 + jd) * 4 + c) * 8;  // D_g[jd][4J + c], J = 0..6
        const double2 b01 = lds_pair(bp), b23 = lds_pair(bp + 2), b45 = lds_pair(bp + 4);
        b[0] = b01.x; b[1] = b01.y; b[2] = b23.x; b[3] = b23.y; b[4] = b45.x; b[5] = b45.y; b[6] = bp[6];
        const int qs[3] = {U0::q, U1::q, U2::q};
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double *wp = sW + ((g * 6 + qs[x]) * 3 + jd) * 4;  // W_g[(a,i),(b,jd)], i = 0..2
            const double2 w01 = lds_pair(wp);
            w[x][0] = w01.x; w[x][1] = w01.y; w[x][2] = wp[2];
        }
    };
    load(0, B[0], W[0]);
#pragma unroll
    for (int st = 0; st < 21; ++st) {
        const int cur = st & 1, gp = st % 7;
        if (st + 1 < 21) load(st + 1, B[cur ^ 1], W[cur ^ 1]);
        double u[3];
#pragma unroll
        for (int x = 0; x < 3; ++x)
            u[x] = DT[gp][0] * W[cur][x][0] + DT[gp][1] * W[cur][x][1] + DT[gp][2] * W[cur][x][2];
#pragma unroll
        for (int J = 0; J < 7; ++J) {
            if (J < U0::nj) acc0[J] = __builtin_amdgcn_mfma_f64_4x4x4f64(u[0], B[cur][J], acc0[J], 0, 0, 0);
            if (J < U1::nj) acc1[J] = __builtin_amdgcn_mfma_f64_4x4x4f64(u[1], B[cur][J], acc1[J], 0, 0, 0);
            if (J < U2::nj) acc2[J] = __builtin_amdgcn_mfma_f64_4x4x4f64(u[2], B[cur][J], acc2[J], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    f64_unit_store<U0>(acc0, blk, h, c, rs);
    f64_unit_store<U1>(acc1, blk, h, c, rs);
    f64_unit_store<U2>(acc2, blk, h, c, rs);
}

// pressure rows K_(p,m),(b,n), m = 4I + h, n = 16t + 4blk + c, for all b and
// t; the B operand D_g[b][n] comes from LDS
__device__ __forceinline__ void f64_pressure(int lane, const double *sD, const double *sPa, __amdgpu_buffer_rsrc_t rs)
{
    const int c = lane & 3, blk = (lane >> 2) & 3, h = lane >> 4;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
        for (int gp = 0; gp < 7; ++gp) {
            const int g = 4 * gp + h;
            const double *pp = sPa + g * 8 + c;
            const double a0 = pp[0], a1 = pp[4];
            const double *bp = sD + ((g * 3 + b) * 4 + c) * 8 + blk;  // D_g[b][16t + 4blk + c] at +4t
            const double b0 = bp[0], b1 = bp[4];
            acc[0][0] = __builtin_amdgcn_mfma_f64_4x4x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_4x4x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_4x4x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_4x4x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        const int vdir = (h * 89 + 4 * blk + c) * 8, vmir = ((4 * blk + c) * 89 + h) * 8;
#pragma unroll
        for (int I = 0; I < 2; ++I)
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                if (16 * t + 4 * blk + c >= 27) continue;
                element_store(rs, acc[I][t], vdir, ((81 + 4 * I) * 89 + b * 27 + 16 * t) * 8);
                element_store(rs, acc[I][t], vmir, ((b * 27 + 16 * t) * 89 + 81 + 4 * I) * 8);
            }
    }
}

__global__ void __launch_bounds__(512)
mr_element_matrix_f64_kernel(int n_elements, const double *__restrict__ rec, const BasisTables *__restrict__ T,
                             double *__restrict__ kbuf)
{
    __shared__ double s_tab[27][3][27];          // dphi_n/dxi_mi at g: [g][mi][n]
    __shared__ double s_psi[27][8];              // trilinear psi_m(g)
    __shared__ double s_rec[2][52][27];          // Gauss-point records
    __shared__ __align__(16) double s_D[2][28][3][4][8];  // D_g[j][n] at [g][j][n % 4][n / 4]; padding zero
    __shared__ __align__(16) double s_W[2][28][6][3][4];  // W_g[(a,i),(b,j)], lower block q, at [g][q][j][i]; padding zero
    __shared__ double s_Pa[2][28][8];            // J_z w_g psi_m(g); g = 27 zero
    constexpr int kRec = 52 * 27, kRecPerThread = (kRec + 255) / 256;
    const int t = threadIdx.x, lane = t & 63;
    const bool producer = t >= 256;
    const int pt = producer ? t - 256 : 0;
    // producer thread pt < 243 owns (j, n) = (rq / 27, rq % 27) of D_g for
    // g = pg (mod 3); pt < 216 owns (q, j, i) = wq of W_g for g = pw (mod 4)
    const int rq = pt % 81, pg = pt / 81, dj = rq / 27, dn = rq % 27;
    const int wq = pt % 54, pw = pt / 54;
    const int wqq = wq / 9, wj = (wq / 3) % 3, wi = wq % 3;
    const int wa = wqq == 0 ? 0 : (wqq <= 2 ? 1 : 2), wb = wqq - (wa * (wa + 1)) / 2;
    const int wf1 = GP_C + t2v(wi, wa) + 6 * t2v(wj, wb), wf2 = GP_SIGJ + t2v(wi, wj);
    const bool wdiag = wa == wb;
    const int64_t n_gp = static_cast<int64_t>(n_elements) * 27;
    const int n_mine = static_cast<int>(blockIdx.x) < n_elements
                           ? (n_elements - 1 - static_cast<int>(blockIdx.x)) / static_cast<int>(gridDim.x) + 1
                           : 0;
    for (int k = t; k < 27 * 27 * 3; k += 512) {
        const int g = k / 81, mi = (k / 27) % 3, n = k % 27;
        s_tab[g][mi][n] = T->quad27[(g * 27 + n) * 4 + 1 + mi];
    }
    for (int k = t; k < 27 * 8; k += 512) s_psi[k / 8][k % 8] = T->lin27[k * 4];
    for (int k = t; k < 2 * 28 * 3 * 4 * 8; k += 512) (&s_D[0][0][0][0][0])[k] = 0.0;
    for (int k = t; k < 2 * 28 * 6 * 3 * 4; k += 512) (&s_W[0][0][0][0][0])[k] = 0.0;
    for (int k = t; k < 2 * 28 * 8; k += 512) (&s_Pa[0][0][0])[k] = 0.0;
    if (producer && n_mine > 0) {
        const int64_t q0 = static_cast<int64_t>(blockIdx.x) * 27;
        for (int k = pt; k < kRec; k += 256) s_rec[0][k / 27][k % 27] = rec[(k / 27) * n_gp + q0 + (k % 27)];
    }
    __syncthreads();

    for (int it = 0; it <= n_mine; ++it) {
        if (producer) {
            if (it < n_mine) {
                const int p = it & 1;
                const int e_next = static_cast<int>(blockIdx.x) + (it + 1) * static_cast<int>(gridDim.x);
                double pre[kRecPerThread];
#pragma unroll
                for (int i = 0; i < kRecPerThread; ++i) {
                    const int k = pt + 256 * i;
                    pre[i] = (it + 1 < n_mine && k < kRec)
                                 ? rec[(k / 27) * n_gp + static_cast<int64_t>(e_next) * 27 + (k % 27)]
                                 : 0.0;
                }
                const double(*R)[27] = s_rec[p];
                // every LDS read of a phase is issued before its writes (which
                // the compiler may not move reads across), so a phase waits
                // for one LDS round trip rather than one per Gauss point
                if (pt < 243) {  // D_g[j][n] = sum_mi DXI_DX(mi, j) dphi_n/dxi_mi (:974-986), g = pg + 3k
                    double dv[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        const int g = pg + 3 * k;
                        dv[k] = R[GP_DXIDZ + 0 + 3 * dj][g] * s_tab[g][0][dn] +
                                R[GP_DXIDZ + 1 + 3 * dj][g] * s_tab[g][1][dn] +
                                R[GP_DXIDZ + 2 + 3 * dj][g] * s_tab[g][2][dn];
                    }
#pragma unroll
                    for (int k = 0; k < 9; ++k) s_D[p][pg + 3 * k][dj][dn % 4][dn / 4] = dv[k];
                }
                if (pt < 216) {  // W_g[(a,i),(b,j)] = J_z w_g (c + delta_ab sigma_ij), g = pw + 4k < 27
                    double wv[7];
#pragma unroll
                    for (int k = 0; k < 7; ++k) {
                        const int g = min(pw + 4 * k, 26);
                        double v = R[wf1][g];
                        if (wdiag) v = v + R[wf2][g];
                        wv[k] = R[GP_JGW][g] * v;
                    }
                    const double pa = R[GP_JGW][pt / 8] * s_psi[pt / 8][pt % 8];
#pragma unroll
                    for (int k = 0; k < 7; ++k)
                        if (pw + 4 * k < 27) s_W[p][pw + 4 * k][wqq][wj][wi] = wv[k];
                    s_Pa[p][pt / 8][pt % 8] = pa;
                }
                if (it + 1 < n_mine) {
#pragma unroll
                    for (int i = 0; i < kRecPerThread; ++i) {
                        const int k = pt + 256 * i;
                        if (k < kRec) s_rec[p ^ 1][k / 27][k % 27] = pre[i];
                    }
                }
            }
        } else if (it >= 1) {
            const int p = (it - 1) & 1;
            const int e = static_cast<int>(blockIdx.x) + (it - 1) * static_cast<int>(gridDim.x);
            // the lane index is made opaque per element so that offsets derived
            // from it are recomputed here rather than held across the loop
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const int c = ln & 3, blk = (ln >> 2) & 3, h = ln >> 4, wave = t >> 6;
            // waves 0, 1: row half t = 0 (and the pressure rows); waves 2, 3: t = 1.
            // The lane keeps D_g[i][16t + 4blk + c] for its Gauss points g = 4gp + h.
            const int tw = wave >> 1;
            double DT[7][3];
#pragma unroll
            for (int gp = 0; gp < 7; ++gp)
#pragma unroll
                for (int j = 0; j < 3; ++j) DT[gp][j] = s_D[p][4 * gp + h][j][c][4 * tw + blk];
            const double *sD = &s_D[p][0][0][0][0], *sW = &s_W[p][0][0][0][0];
            const __amdgpu_buffer_rsrc_t rs = element_rsrc(kbuf + static_cast<int64_t>(e) * 89 * 89);
            // 1659 k-steps per element: 378 / 399 / 441 / 441
            if (wave == 0) {
                f64_units<0, F64Unit<1, 0, 7>, F64Unit<3, 0, 7>, F64Unit<0, 0, 4>>(ln, DT, sD, sW, rs);
            } else if (wave == 1) {
                f64_units<0, F64Unit<4, 0, 7>, F64Unit<2, 0, 4>, F64Unit<5, 0, 4>>(ln, DT, sD, sW, rs);
                f64_pressure(ln, sD, &s_Pa[p][0][0], rs);
            } else if (wave == 2) {
                f64_units<1, F64Unit<1, 1, 7>, F64Unit<3, 1, 7>, F64Unit<0, 1, 7>>(ln, DT, sD, sW, rs);
            } else {
                f64_units<1, F64Unit<4, 1, 7>, F64Unit<2, 1, 7>, F64Unit<5, 1, 7>>(ln, DT, sD, sW, rs);
                // zero p-p block
                double *Ke = kbuf + static_cast<int64_t>(e) * 89 * 89;
                Ke[(81 + ln / 8) * 89 + 81 + ln % 8] = 0.0;
            }
        }
        __syncthreads();
    }
}

template <typename Real, int kChunk>
__global__ void __launch_bounds__(64)
csr_row_gather_kernel(int n_rows, const int64_t *__restrict__ row_ptr, const int *__restrict__ gptr,
                      const int *__restrict__ gidx, const Real *__restrict__ kbuf,
                      const uint8_t *__restrict__ rank, const uint16_t *__restrict__ row_nb,
                      Real *__restrict__ values)
{
    extern __shared__ unsigned char s_raw[];
    Real *s_row = reinterpret_cast<Real *>(s_raw);
    const int r = blockIdx.x, lane = threadIdx.x;
    if (r >= n_rows) return;
    const int64_t r0 = row_ptr[r];
    const int len = static_cast<int>(row_ptr[r + 1] - r0);
    const int nb = row_nb[r];
    for (int k = lane; k < len; k += 64) s_row[k] = Real(0);
    __syncthreads();
    // lane j holds element columns j and j + 64: component block and rank slot
    const int blk0 = (lane / 27) * nb, m0 = lane % 27;
    const bool hi = lane + 64 < 89;
    const int j1 = lane + 64;
    const int blk1 = j1 < 81 ? (j1 / 27) * nb : 3 * nb, m1 = j1 < 81 ? j1 % 27 : 27 + j1 - 81;
    // contributions in chunks of kChunk element rows: all global loads of a
    // chunk are issued before its LDS accumulation
    for (int c0 = gptr[r]; c0 < gptr[r + 1]; c0 += kChunk) {
        const int nc = min(kChunk, gptr[r + 1] - c0);
        Real v0[kChunk], v1[kChunk];
        int o0[kChunk], o1[kChunk];
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
            if (c < nc) {
                const int ei = gidx[c0 + c];  // e*89 + i
                const int e = ei / 89, i = ei % 89;
                const int64_t base = static_cast<int64_t>(ei) * 89;
                const uint8_t *rk = rank + (static_cast<int64_t>(e) * kRankNodes + (i < 81 ? i % 27 : i - 54)) *
                                               kRankNodes;
                v0[c] = kbuf[base + lane];
                o0[c] = blk0 + rk[m0];
                v1[c] = hi ? kbuf[base + j1] : Real(0);
                o1[c] = hi ? blk1 + rk[m1] : -1;
            }
        }
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
            if (c < nc) {
                s_row[o0[c]] += v0[c];
                if (o1[c] >= 0) s_row[o1[c]] += v1[c];
            }
            __syncthreads();
        }
    }
    for (int k = lane; k < len; k += 64) values[r0 + k] = s_row[k];
}

template <typename Real>
static int launch_row_gather(EquationsSet &s, const Real *kbuf, Real *values, int chunk)
{
    Runtime &rt = runtime();
    const size_t row_lds = static_cast<size_t>(s.max_row_len) * sizeof(Real);
    const unsigned grid = static_cast<unsigned>(s.n_rows);
    switch (chunk) {
    case 4:
        csr_row_gather_kernel<Real, 4><<<grid, 64, row_lds, rt.stream>>>(s.n_rows, s.row_ptr.ptr, s.gather_ptr.ptr,
                                                                        s.gather_idx.ptr, kbuf, s.col_rank.ptr, s.row_nb.ptr,
                                                                        values);
        break;
    case 16:
        csr_row_gather_kernel<Real, 16><<<grid, 64, row_lds, rt.stream>>>(s.n_rows, s.row_ptr.ptr, s.gather_ptr.ptr,
                                                                         s.gather_idx.ptr, kbuf, s.col_rank.ptr, s.row_nb.ptr,
                                                                         values);
        break;
    default:
        csr_row_gather_kernel<Real, 8><<<grid, 64, row_lds, rt.stream>>>(s.n_rows, s.row_ptr.ptr, s.gather_ptr.ptr,
                                                                        s.gather_idx.ptr, kbuf, s.col_rank.ptr, s.row_nb.ptr,
                                                                        values);
        break;
    }
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

int launch_schur_mass(EquationsSet &s, double *inv_p)
{
    if (!s.gp_records.ptr || s.subtype != OC_MOONEY_RIVLIN_SUBTYPE)
        return set_error(OC_ERR_UNSUPPORTED, "the pressure-mass Schur estimate needs the analytic Jacobian's Gauss records");
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    const int nl = s.n_nodes_l;
    schur_mass_kernel<<<static_cast<unsigned>((nl + 255) / 256), 256, 0, runtime().stream>>>(
        nl, s.gather_ptr.ptr + 3 * static_cast<int64_t>(s.n_nodes), s.gather_idx.ptr, s.n_elements, s.gp_records.ptr, T,
        s.mat.ptr, inv_p);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

int mr_jacobian_buffers(EquationsSet &s, int precision)
{
    OC_CHECK(mr_prepare(s));
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    const size_t kn = static_cast<size_t>(s.n_elements) * 89 * 89;
    if (precision == OC_FP64) {
        if (!s.kbuf64.ptr) OC_CHECK(s.kbuf64.alloc(kn));
        if (!s.values.ptr) OC_CHECK(s.values.alloc(s.nnz));
    } else {
        if (!s.kbuf32.ptr) OC_CHECK(s.kbuf32.alloc(kn));
        if (!s.values_sp.ptr) OC_CHECK(s.values_sp.alloc(s.nnz));
    }
    return OC_NO_ERROR;
}

// Element-matrix phase alone (the autotuner times it per launch shape).
int launch_mr_element_matrices(EquationsSet &s, int precision, int blocks_per_cu)
{
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    // default: one resident block of eight waves per CU
    const int bpc = blocks_per_cu > 0 ? blocks_per_cu : 1;
    const unsigned blocks = static_cast<unsigned>(std::min(s.n_elements, rt.num_cus * bpc));
    if (precision == OC_FP64) {
        KernelTimer timer("mr_element_matrix_fp64");
        mr_element_matrix_f64_kernel<<<blocks, 512, 0, rt.stream>>>(s.n_elements, s.gp_records.ptr, T, s.kbuf64.ptr);
    } else {
        KernelTimer timer("mr_element_matrix_fp32");
        mr_element_matrix_kernel<float><<<blocks, 512, 0, rt.stream>>>(s.n_elements, s.gp_records.ptr, T,
                                                                      s.kbuf32.ptr);
    }
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

int launch_mr_row_gather(EquationsSet &s, int precision, int chunk)
{
    if (precision == OC_FP64) {
        KernelTimer timer("csr_row_gather_fp64");
        return launch_row_gather<double>(s, s.kbuf64.ptr, s.values.ptr, chunk);
    }
    KernelTimer timer("csr_row_gather_fp32");
    return launch_row_gather<float>(s, s.kbuf32.ptr, s.values_sp.ptr, chunk);
}

// Forward-difference Jacobian: element residuals at x, then one launch over
// all (element, column) pairs, then the same ordered CSR row gather.
int launch_mr_jacobian_fd(EquationsSet &s)
{
    OC_CHECK(mr_jacobian_buffers(s, OC_FP64));
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    // delta = (1 + ||x||_2) 1e-7 over the whole dependent vector (:3086-3087)
    std::vector<double> x(s.n_rows);
    OC_HIP_CHECK(hipMemcpyAsync(x.data(), s.dofs.ptr, sizeof(double) * x.size(), hipMemcpyDeviceToHost, rt.stream));
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    double ss = 0.0;
    for (double v : x) ss += v * v;
    const double delta = (1.0 + std::sqrt(ss)) * 1e-7;
    const unsigned blocks = static_cast<unsigned>((s.n_elements + kGaussElemsPerBlock - 1) / kGaussElemsPerBlock);
    OC_CHECK(launch_residual_elements<false>(s, blocks, s.dofs.ptr, T, s.relem.ptr, 0.0, nullptr));
    const int64_t nv = static_cast<int64_t>(s.n_elements) * 89;
    if (nv >= INT32_MAX) return set_error(OC_ERR_UNSUPPORTED, "too many elements for the finite-difference launch");
    {
        KernelTimer timer("mr_fd_columns");
        const unsigned fb = static_cast<unsigned>((nv + kGaussElemsPerBlock - 1) / kGaussElemsPerBlock);
        OC_CHECK(launch_residual_elements<true>(s, fb, s.dofs.ptr, T, s.relem.ptr, delta, s.kbuf64.ptr));
    }
    return launch_mr_row_gather(s, OC_FP64, s.tuned_gather_chunk);
}

int launch_mr_jacobian(EquationsSet &s, int precision)
{
    if (s.subtype == OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE) {
        // the reference evaluates this subtype's Jacobian by forward differences
        if (precision != OC_FP64)
            return set_error(OC_ERR_UNSUPPORTED, "the incompressible Mooney-Rivlin subtype has FP64 "
                                                 "forward-difference Jacobians only");
        return launch_mr_jacobian_fd(s);
    }
    OC_CHECK(mr_jacobian_buffers(s, precision));
    OC_CHECK(launch_mr_gauss(s));
    OC_CHECK(launch_mr_element_matrices(s, precision, s.tuned_element_blocks_per_cu));
    return launch_mr_row_gather(s, precision, s.tuned_gather_chunk);
}

// ---------------------------------------------------------------------------
// Allocation-free residual evaluation (oc_residual_evaluate_noalloc): the
// element residuals go through `scratch`, a fixed array in device memory (one
// copy per GPU, held until the process ends) instead of the set's buffer.
// Host code never passes the symbol itself to a kernel (on the host it only
// stands for the device variable): it uses the address hipGetSymbolAddress
// returned for the device, kept in g_scratch_ptr.  Every element residual
// entry is written before the DOF gather reads it, so nothing is cleared
// between calls; calls that overlap on one GPU share the array and race.
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[static_cast<size_t>(OC_NOALLOC_MAX_ELEMENTS) * 89];
static double *g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return set_error(OC_ERR_UNSUPPORTED, "device %d beyond the %d devices of the scratch table", device, kMaxDevices);
    if (g_scratch_ptr[device]) return OC_NO_ERROR;
    void *p = nullptr;
    OC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(OC_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = static_cast<double *>(p);
    return OC_NO_ERROR;
}

// (c10, c01) of every quadratic node from its elements (all elements around
// a node belong to one mesh, so every writer of an entry writes the same pair)
__global__ void __launch_bounds__(256) node_materials_kernel(int n_elements, const int *__restrict__ conn_q,
                                                            const double *__restrict__ mat,
                                                            double *__restrict__ node_mat)
{
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n_elements * 27; t += gridDim.x * blockDim.x) {
        const int e = t / 27, node = conn_q[t];
        node_mat[2 * node] = mat[2 * e];
        node_mat[2 * node + 1] = mat[2 * e + 1];
    }
}

int mr_node_materials(EquationsSet &s)
{
    if (!s.node_mat.ptr) OC_CHECK(s.node_mat.alloc(static_cast<size_t>(2) * s.n_nodes));
    const int64_t n = static_cast<int64_t>(s.n_elements) * 27;
    const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 8192));
    node_materials_kernel<<<blocks, 256, 0, runtime().stream>>>(s.n_elements, s.conn.ptr, s.mat.ptr, s.node_mat.ptr);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

}  // namespace oc

using namespace oc;

extern "C" int oc_residual_evaluate_noalloc(oc_equations_set h, const double *dofs_device, double *residual_device)
{
    OC_CHECK(require_runtime());
    if (!h) return set_error(OC_ERR_INVALID_HANDLE, "equations set handle is NULL");
    if (!dofs_device || !residual_device)
        return set_error(OC_ERR_INVALID_ARGUMENT, "oc_residual_evaluate_noalloc: NULL device array");
    EquationsSet &s = h->s;
    if (s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN)
        return set_error(OC_ERR_INVALID_ARGUMENT, "oc_residual_evaluate_noalloc is defined for Mooney-Rivlin sets");
    if (s.n_elements > OC_NOALLOC_MAX_ELEMENTS)
        return set_error(OC_ERR_UNSUPPORTED,
                         "%d elements exceed the %d elements of the allocation-free residual's scratch; use "
                         "oc_equations_set_residual_evaluate",
                         s.n_elements, OC_NOALLOC_MAX_ELEMENTS);
    Runtime &rt = runtime();
    const BasisTables *T = device_basis_tables();
    double *scr = (rt.device >= 0 && rt.device < kMaxDevices) ? g_scratch_ptr[rt.device] : nullptr;
    if (!T || !scr || !s.gather_ptr.ptr)
        return set_error(OC_ERR_NOT_INITIALISED, "device tables, scratch or the set's gather lists are not set up");
    const unsigned blocks = static_cast<unsigned>((s.n_elements + kGaussElemsPerBlock - 1) / kGaussElemsPerBlock);
    OC_CHECK(launch_residual_elements<false>(s, blocks, dofs_device, T, scr, 0.0, nullptr));
    dof_gather_kernel<<<static_cast<unsigned>((s.n_rows + 255) / 256), 256, 0, rt.stream>>>(
        s.n_rows, s.gather_ptr.ptr, s.gather_idx.ptr, scr, residual_device);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}
