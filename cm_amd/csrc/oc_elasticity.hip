// oc_elasticity.hip -- gfx950 Mooney-Rivlin finite elasticity assembly:
// residual and analytic Jacobian (finite_elasticity_routines.f90:842-1175,
// 1260-1640) scattered deterministically into the CSR equations matrix.
//
// Jacobian: mr_gauss_kernel (one thread per element Gauss point, all
// elements) evaluates the FP64 kinematics and tangent into Gauss-point
// records (field-major, see GpField); mr_jacobian_element_kernel contracts
// the records of each element of one colour into the 89x89 element Jacobian
// and scatters it through the precomputed slot-offset map.  Elements of one
// colour share no DOF, so each launch writes every CSR entry at most once and
// the colour order fixes the summation order.
// Residual: mr_residual_element_kernel evaluates element residuals of all
// elements into a buffer and dof_gather_kernel sums them per DOF in
// ascending element order.
#include <algorithm>
#include <cstdint>

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
                const double *__restrict__ coords, int nq, const double *__restrict__ dofs, double c10, double c01,
                const BasisTables *__restrict__ T, double *__restrict__ rec)
{
    __shared__ double s_fq[27 * 27 * 4];
    __shared__ double s_X[kGaussElemsPerBlock][81];
    __shared__ double s_z[kGaussElemsPerBlock][81];
    __shared__ double s_p[kGaussElemsPerBlock][8];
    const int t = threadIdx.x;
    const int e0 = blockIdx.x * kGaussElemsPerBlock;
    for (int k = t; k < 27 * 27 * 4; k += 256) s_fq[k] = T->quad27[k];
    for (int k = t; k < kGaussElemsPerBlock * 81; k += 256) {
        const int le = k / 81, r = k % 81, c = r / 27, n = r % 27;
        const int e = e0 + le;
        if (e < n_elements) {
            const int node = conn_q[e * 27 + n];
            s_X[le][r] = coords[static_cast<int64_t>(c) * nq + node];
            s_z[le][r] = dofs[static_cast<int64_t>(c) * nq + node];
        }
    }
    for (int k = t; k < kGaussElemsPerBlock * 8; k += 256) {
        const int le = k / 8, n = k % 8;
        const int e = e0 + le;
        if (e < n_elements) s_p[le][n] = dofs[static_cast<int64_t>(3) * nq + conn_l[e * 8 + n]];
    }
    __syncthreads();
    const int le = t / 27, g = t % 27;
    const int e = e0 + le;
    if (t >= kGaussElemsPerBlock * 27 || e >= n_elements) return;
    GaussKinematics k;
    gauss_kinematics(s_fq + g * 27 * 4, s_z[le], s_X[le], k);
    double P = 0.0;
    for (int n = 0; n < 8; ++n) P += T->lin27[(g * 8 + n) * 4] * s_p[le][n];
    const int64_t n_gp = static_cast<int64_t>(n_elements) * 27;
    const int64_t q = static_cast<int64_t>(e) * 27 + g;
#pragma unroll
    for (int i = 0; i < 9; ++i) rec[(GP_DXIDZ + i) * n_gp + q] = k.dxidz[i];
    rec[GP_JGW * n_gp + q] = k.jz * T->w27[g];
    double c[36], sigma[6];
    mr_elasticity_tensor(k.F, k.Jznu, P, c10, c01, c, sigma);
#pragma unroll
    for (int i = 0; i < 36; ++i) rec[(GP_C + i) * n_gp + q] = c[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) rec[(GP_SIGJ + i) * n_gp + q] = sigma[i];
}

// Element DOFs (component -> node order of the reference element matrix,
// equations_matrices_routines.f90:684-941): 0..80 displacement, 81..88 p.
__device__ __forceinline__ int element_dof(const int *__restrict__ conn_q, const int *__restrict__ conn_l, int nq,
                                           int e, int i)
{
    if (i < 81) return (i / 27) * nq + conn_q[e * 27 + (i % 27)];
    return 3 * nq + conn_l[e * 8 + (i - 81)];
}

// Scatter map: offset of element column j inside CSR row dof_i.
__global__ void __launch_bounds__(128)
build_slot_offsets_kernel(int n_elements, const int *__restrict__ conn_q, const int *__restrict__ conn_l, int nq,
                          const int *__restrict__ row_ptr, const int *__restrict__ col_idx,
                          uint16_t *__restrict__ offs, int *__restrict__ n_missing)
{
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<int64_t>(n_elements) * 89) return;
    const int e = static_cast<int>(gid / 89), i = static_cast<int>(gid % 89);
    const int row = element_dof(conn_q, conn_l, nq, e, i);
    const int base = row_ptr[row];
    for (int j = 0; j < 89; ++j) {
        const int slot = csr_slot(row_ptr, col_idx, row, element_dof(conn_q, conn_l, nq, e, j));
        if (slot < 0) {
            atomicAdd(n_missing, 1);
            offs[gid * 89 + j] = 0;
        } else {
            offs[gid * 89 + j] = static_cast<uint16_t>(slot - base);
        }
    }
}

int mr_prepare(EquationsSet &s)
{
    if (s.slot_offsets.ptr) return OC_NO_ERROR;
    int max_row = 0;
    for (int r = 0; r < s.n_rows; ++r) max_row = std::max(max_row, s.h_row_ptr[r + 1] - s.h_row_ptr[r]);
    if (max_row > 65535)
        return set_error(OC_ERR_UNSUPPORTED, "CSR row of %d entries exceeds the 16-bit scatter map", max_row);
    OC_CHECK(s.slot_offsets.alloc(static_cast<size_t>(s.n_elements) * 89 * 89));
    DevBuf<int> missing;
    OC_CHECK(missing.alloc(1));
    Runtime &rt = runtime();
    OC_HIP_CHECK(hipMemsetAsync(missing.ptr, 0, sizeof(int), rt.stream));
    const int64_t nthreads = static_cast<int64_t>(s.n_elements) * 89;
    build_slot_offsets_kernel<<<static_cast<unsigned>((nthreads + 127) / 128), 128, 0, rt.stream>>>(
        s.n_elements, s.conn.ptr, s.conn_l.ptr, s.n_nodes, s.row_ptr.ptr, s.col_idx.ptr, s.slot_offsets.ptr,
        missing.ptr);
    OC_HIP_CHECK(hipGetLastError());
    int h_missing = 0;
    OC_HIP_CHECK(hipMemcpyAsync(&h_missing, missing.ptr, sizeof(int), hipMemcpyDeviceToHost, rt.stream));
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    if (h_missing) return set_error(OC_ERR_PATTERN, "%d element entries missing from the CSR structure", h_missing);
    OC_CHECK(s.gp_records.alloc(static_cast<size_t>(s.n_elements) * 27 * GP_NFIELDS));
    OC_CHECK(s.relem.alloc(static_cast<size_t>(s.n_elements) * 89));
    // DOF -> element entries, ascending element order
    std::vector<int32_t> hconn(static_cast<size_t>(s.n_elements) * 27), hconn_l(static_cast<size_t>(s.n_elements) * 8);
    OC_HIP_CHECK(hipMemcpy(hconn.data(), s.conn.ptr, hconn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    OC_HIP_CHECK(hipMemcpy(hconn_l.data(), s.conn_l.ptr, hconn_l.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> ptr(static_cast<size_t>(s.n_rows) + 1, 0), idx(static_cast<size_t>(s.n_elements) * 89);
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

static int launch_gauss(EquationsSet &s)
{
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    const unsigned blocks = static_cast<unsigned>((s.n_elements + kGaussElemsPerBlock - 1) / kGaussElemsPerBlock);
    KernelTimer timer("mr_gauss_jacobian");
    mr_gauss_kernel<<<blocks, 256, 0, rt.stream>>>(s.n_elements, s.conn.ptr, s.conn_l.ptr, s.coords.ptr, s.n_nodes,
                                                   s.dofs.ptr, s.c10, s.c01, T, s.gp_records.ptr);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

// ---------------------------------------------------------------------------
// Residual (:1491-1574), fused with the Gauss-point evaluation: one block of
// 9 elements x 27 Gauss points evaluates kinematics and the stress routine,
// keeps JGW*sigma and DXI_DX in LDS, then sums each element residual entry
//   R_(a,m) += sum_i DPHIDZ(i,m) JGW sigma(i,a),  R_(p,m) += J_X w (J-1) psi_m
// over the Gauss points in order into a per-element buffer.  A gather kernel
// then adds the element contributions of every DOF in ascending element order
// (the reference's element-loop order, equations_set_routines.f90:5535-5568).
__global__ void __launch_bounds__(256)
mr_residual_element_kernel(int n_elements, const int *__restrict__ conn_q, const int *__restrict__ conn_l,
                           const double *__restrict__ coords, int nq, const double *__restrict__ dofs, double c10,
                           double c01, const BasisTables *__restrict__ T, double *__restrict__ relem)
{
    __shared__ double s_fq[27 * 27 * 4];
    __shared__ double s_X[kGaussElemsPerBlock][81];
    __shared__ double s_z[kGaussElemsPerBlock][81];
    __shared__ double s_p[kGaussElemsPerBlock][8];
    __shared__ double s_dx[kGaussElemsPerBlock][9][27];
    __shared__ double s_S[kGaussElemsPerBlock][9][27];  // JGW * sigma(i, a): [i + 3a][g]
    __shared__ double s_rp[kGaussElemsPerBlock][27];
    const int t = threadIdx.x;
    const int e0 = blockIdx.x * kGaussElemsPerBlock;
    for (int k = t; k < 27 * 27 * 4; k += 256) s_fq[k] = T->quad27[k];
    for (int k = t; k < kGaussElemsPerBlock * 81; k += 256) {
        const int le = k / 81, r = k % 81, c = r / 27, n = r % 27;
        const int e = e0 + le;
        if (e < n_elements) {
            const int node = conn_q[e * 27 + n];
            s_X[le][r] = coords[static_cast<int64_t>(c) * nq + node];
            s_z[le][r] = dofs[static_cast<int64_t>(c) * nq + node];
        }
    }
    for (int k = t; k < kGaussElemsPerBlock * 8; k += 256) {
        const int le = k / 8, n = k % 8;
        const int e = e0 + le;
        if (e < n_elements) s_p[le][n] = dofs[static_cast<int64_t>(3) * nq + conn_l[e * 8 + n]];
    }
    __syncthreads();
    {
        const int le = t / 27, g = t % 27;
        const int e = e0 + le;
        if (t < kGaussElemsPerBlock * 27 && e < n_elements) {
            GaussKinematics k;
            gauss_kinematics(s_fq + g * 27 * 4, s_z[le], s_X[le], k);
            double P = 0.0;
            for (int n = 0; n < 8; ++n) P += T->lin27[(g * 8 + n) * 4] * s_p[le][n];
            double sigma[6];
            mr_cauchy_stress(k.F, k.Jznu, P, c10, c01, sigma);
            const double jgw = k.jz * T->w27[g];
            for (int i = 0; i < 9; ++i) s_dx[le][i][g] = k.dxidz[i];
            for (int a = 0; a < 3; ++a)
                for (int i = 0; i < 3; ++i) s_S[le][i + 3 * a][g] = jgw * sigma[t2v(i, a)];
            s_rp[le][g] = k.jx * T->w27[g] * (k.Jznu - 1.0);
        }
    }
    __syncthreads();
    for (int w = t; w < kGaussElemsPerBlock * 89; w += 256) {
        const int le = w / 89, r = w % 89;
        const int e = e0 + le;
        if (e >= n_elements) continue;
        double acc = 0.0;
        if (r < 81) {
            const int a = r / 27, m = r % 27;
#pragma unroll 1
            for (int g = 0; g < 27; ++g) {
                const double *f = s_fq + (g * 27 + m) * 4;
                double d[3];
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    d[i] = s_dx[le][0 + 3 * i][g] * f[1] + s_dx[le][1 + 3 * i][g] * f[2] +
                           s_dx[le][2 + 3 * i][g] * f[3];
                const double dot = d[0] * s_S[le][0 + 3 * a][g] + d[1] * s_S[le][1 + 3 * a][g] +
                                   d[2] * s_S[le][2 + 3 * a][g];
                acc = acc + dot;
            }
        } else {
            const int m = r - 81;
#pragma unroll 1
            for (int g = 0; g < 27; ++g) acc = acc + s_rp[le][g] * T->lin27[(g * 8 + m) * 4];
        }
        relem[static_cast<int64_t>(e) * 89 + r] = acc;
    }
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

int launch_mr_residual(EquationsSet &s)
{
    OC_CHECK(mr_prepare(s));
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    const unsigned blocks = static_cast<unsigned>((s.n_elements + kGaussElemsPerBlock - 1) / kGaussElemsPerBlock);
    {
        KernelTimer timer("mr_residual");
        mr_residual_element_kernel<<<blocks, 256, 0, rt.stream>>>(s.n_elements, s.conn.ptr, s.conn_l.ptr,
                                                                  s.coords.ptr, s.n_nodes, s.dofs.ptr, s.c10, s.c01,
                                                                  T, s.relem.ptr);
        OC_HIP_CHECK(hipGetLastError());
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
// Stage 2: analytic Jacobian as
//   K_(a,m),(b,n) = sum_g sum_i D_g[i][m] T_g^{ab}[i][n],
//   T_g^{ab}[i][n] = sum_j W_g[(a,i),(b,j)] D_g[j][n],
//   W_g[(a,i),(b,j)] = J_z w_g (c(V(i,a),V(j,b)) + delta_ab sigma_ij),
// the reference's TEMPVEC / DOT_PRODUCT contraction (:1001-1037); pressure
// rows K_(p,m),(b,n) = sum_g J_z w_g D_g[b][n] psi_m (:1039-1054); the
// off-diagonal and pressure blocks are mirrored (:1149-1153).  The
// contraction runs in Real (double, or float for FP32 Jacobians).
// 162 threads own 3x9 tiles of the six lower displacement blocks, 81 own a
// pressure column.  Persistent blocks keep the basis table in LDS.  Every
// entry of one element has its own CSR slot, so the scatter issues all loads
// of a batch before its stores.
template <typename Real>
__global__ void __launch_bounds__(256)
mr_jacobian_element_kernel(const int *__restrict__ colour_elems, int n_in_colour, int n_elements,
                           const double *__restrict__ rec, const int *__restrict__ conn_q,
                           const int *__restrict__ conn_l, int nq, const int *__restrict__ row_ptr,
                           const uint16_t *__restrict__ offs, const BasisTables *__restrict__ T,
                           Real *__restrict__ values)
{
    __shared__ double s_df[27 * 27 * 3];
    __shared__ Real s_psi[27 * 8];
    __shared__ double s_dx[9][27];
    __shared__ double s_c[36][27];
    __shared__ double s_sig[6][27];
    __shared__ double s_jgw[27];
    __shared__ Real s_D[27][81];
    __shared__ Real s_W[27][81];
    __shared__ Real s_T[6][81];
    __shared__ int s_base[89];
    const int blk_a[6] = {0, 1, 1, 2, 2, 2}, blk_b[6] = {0, 0, 1, 0, 1, 2};
    const int t = threadIdx.x;
    for (int k = t; k < 27 * 27 * 3; k += 256) s_df[k] = T->quad27[(k / 3) * 4 + 1 + k % 3];
    for (int k = t; k < 27 * 8; k += 256) s_psi[k] = static_cast<Real>(T->lin27[k * 4]);
    const int64_t n_gp = static_cast<int64_t>(n_elements) * 27;

    const bool is_uu = t < 162, is_p = t >= 162 && t < 243;
    int blk = 0, m0 = 0, n0 = 0, pb = 0, pn = 0;
    if (is_uu) {
        blk = t / 27;
        const int r = t % 27;
        m0 = (r / 3) * 3;
        n0 = (r % 3) * 9;
    } else if (is_p) {
        pb = (t - 162) / 27;
        pn = (t - 162) % 27;
    }

    for (int idx = blockIdx.x; idx < n_in_colour; idx += gridDim.x) {
        const int e = colour_elems[idx];
        const int64_t q0 = static_cast<int64_t>(e) * 27;
        __syncthreads();
        for (int k = t; k < 52 * 27; k += 256) {
            const int f = k / 27, g = k % 27;
            const double v = rec[f * n_gp + q0 + g];
            if (f < 9) s_dx[f][g] = v;
            else if (f == GP_JGW) s_jgw[g] = v;
            else if (f < GP_SIGJ) s_c[f - GP_C][g] = v;
            else s_sig[f - GP_SIGJ][g] = v;
        }
        if (t < 89) s_base[t] = row_ptr[element_dof(conn_q, conn_l, nq, e, t)];
        __syncthreads();
        for (int k = t; k < 27 * 81; k += 256) {
            const int g = k / 81, r = k % 81;
            {  // D_g[i][n] = sum_mi DXI_DX(mi, i) dphi_n/dxi_mi (:974-986)
                const int i = r / 27, n = r % 27;
                const double *f = s_df + (g * 27 + n) * 3;
                s_D[g][r] = static_cast<Real>(s_dx[0 + 3 * i][g] * f[0] + s_dx[1 + 3 * i][g] * f[1] +
                                              s_dx[2 + 3 * i][g] * f[2]);
            }
            {  // W_g[(a,i),(b,j)]
                const int row = r / 9, col = r % 9;
                const int a = row / 3, i = row % 3, b = col / 3, j = col % 3;
                double v = s_c[t2v(i, a) + 6 * t2v(j, b)][g];
                if (a == b) v = v + s_sig[t2v(i, j)][g];
                s_W[g][r] = static_cast<Real>(s_jgw[g] * v);
            }
        }
        __syncthreads();

        Real acc[27];
#pragma unroll
        for (int i = 0; i < 27; ++i) acc[i] = Real(0);
        for (int g = 0; g < 27; ++g) {
            for (int q = t; q < 486; q += 256) {
                const int bq = q / 81, rem = q % 81, i = rem / 27, n = rem % 27;
                const int a = blk_a[bq], b = blk_b[bq];
                const Real *w = &s_W[g][(a * 3 + i) * 9 + b * 3];
                s_T[bq][i * 27 + n] = w[0] * s_D[g][n] + w[1] * s_D[g][27 + n] + w[2] * s_D[g][54 + n];
            }
            __syncthreads();
            if (is_uu) {
#pragma unroll
                for (int mm = 0; mm < 3; ++mm) {
                    const Real d0 = s_D[g][m0 + mm], d1 = s_D[g][27 + m0 + mm], d2 = s_D[g][54 + m0 + mm];
#pragma unroll
                    for (int nn = 0; nn < 9; ++nn) {
                        const Real dot = d0 * s_T[blk][n0 + nn] + d1 * s_T[blk][27 + n0 + nn] +
                                         d2 * s_T[blk][54 + n0 + nn];
                        acc[mm * 9 + nn] = acc[mm * 9 + nn] + dot;
                    }
                }
            } else if (is_p) {
                const Real jd = static_cast<Real>(s_jgw[g]) * s_D[g][pb * 27 + pn];
#pragma unroll
                for (int m = 0; m < 8; ++m) acc[m] = acc[m] + jd * s_psi[g * 8 + m];
            }
            __syncthreads();
        }

        const uint16_t *eo = offs + static_cast<int64_t>(e) * 89 * 89;
        if (is_uu) {
            const int a = blk_a[blk], b = blk_b[blk];
            int slot[27];
            Real old[27];
#pragma unroll
            for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                for (int nn = 0; nn < 9; ++nn) {
                    const int i = a * 27 + m0 + mm, j = b * 27 + n0 + nn;
                    slot[mm * 9 + nn] = s_base[i] + eo[i * 89 + j];
                }
#pragma unroll
            for (int k = 0; k < 27; ++k) old[k] = values[slot[k]];
#pragma unroll
            for (int k = 0; k < 27; ++k) values[slot[k]] = old[k] + acc[k];
            if (a != b) {
#pragma unroll
                for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                    for (int nn = 0; nn < 9; ++nn) {
                        const int i = a * 27 + m0 + mm, j = b * 27 + n0 + nn;
                        slot[mm * 9 + nn] = s_base[j] + eo[j * 89 + i];
                    }
#pragma unroll
                for (int k = 0; k < 27; ++k) old[k] = values[slot[k]];
#pragma unroll
                for (int k = 0; k < 27; ++k) values[slot[k]] = old[k] + acc[k];
            }
        } else if (is_p) {
            const int j = pb * 27 + pn;
            int slot[16];
            Real old[16];
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int i = 81 + m;
                slot[m] = s_base[i] + eo[i * 89 + j];
                slot[8 + m] = s_base[j] + eo[j * 89 + i];
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) old[k] = values[slot[k]];
#pragma unroll
            for (int k = 0; k < 16; ++k) values[slot[k]] = old[k] + acc[k & 7];
        }
    }
}

int launch_mr_jacobian(EquationsSet &s, int precision)
{
    OC_CHECK(mr_prepare(s));
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    OC_CHECK(launch_gauss(s));
    if (precision == OC_FP64) OC_CHECK(fill<double>(s.values.ptr, s.nnz, 0.0));
    else OC_CHECK(fill<float>(s.values_sp.ptr, s.nnz, 0.0f));
    const int ncol = static_cast<int>(s.colour_ptr.size()) - 1;
    for (int c = 0; c < ncol; ++c) {
        const int n = s.colour_ptr[c + 1] - s.colour_ptr[c];
        if (n == 0) continue;
        const int *ce = s.colour_elems.ptr + s.colour_ptr[c];
        const unsigned blocks = static_cast<unsigned>(std::min(n, rt.num_cus * 2));
        KernelTimer timer(precision == OC_FP64 ? "mr_jacobian_fp64" : "mr_jacobian_fp32");
        if (precision == OC_FP64)
            mr_jacobian_element_kernel<double><<<blocks, 256, 0, rt.stream>>>(
                ce, n, s.n_elements, s.gp_records.ptr, s.conn.ptr, s.conn_l.ptr, s.n_nodes, s.row_ptr.ptr,
                s.slot_offsets.ptr, T, s.values.ptr);
        else
            mr_jacobian_element_kernel<float><<<blocks, 256, 0, rt.stream>>>(
                ce, n, s.n_elements, s.gp_records.ptr, s.conn.ptr, s.conn_l.ptr, s.n_nodes, s.row_ptr.ptr,
                s.slot_offsets.ptr, T, s.values_sp.ptr);
        OC_HIP_CHECK(hipGetLastError());
    }
    return OC_NO_ERROR;
}

}  // namespace oc
