// oc_assembly.hip -- gfx950 assembly kernels: Laplace linear matrix, Mooney-
// Rivlin residual and analytic Jacobian, with deterministic CSR scatter.
//
// Determinism: elements are greedily coloured on the host so that no two
// elements of one colour share a DOF.  Each colour is one launch; inside a
// launch every CSR entry / residual entry receives at most one contribution,
// and the colour order fixes the summation order of every entry.
#include <cmath>
#include <cstring>

#include "oc_fe_device.h"
#include "oc_internal.h"

namespace oc {

// ---------------------------------------------------------------------------
// Basis tables (host).  GAUSS_LEGENDRE closed forms on [0,1]
// (basis_routines.f90:5826-5875); Lagrange 1-D polynomials (:7015-7137);
// tensor products with xi1 fastest for both Gauss points (:2863-2890) and
// basis nodes (:1594-1630).
static double lagrange_1d(int order, int node, bool deriv, double xi)
{
    if (order == 1) {
        if (!deriv) return node == 0 ? 1.0 - xi : xi;
        return node == 0 ? -1.0 : 1.0;
    }
    if (!deriv) {
        if (node == 0) return 1.0 - 3.0 * xi + 2.0 * xi * xi;
        if (node == 1) return 4.0 * xi * (1.0 - xi);
        return xi * (xi + xi - 1.0);
    }
    if (node == 0) return 4.0 * xi - 3.0;
    if (node == 1) return 4.0 - 8.0 * xi;
    return 4.0 * xi - 1.0;
}

static void gauss_1d(int n, double *x, double *w)
{
    if (n == 2) {
        const double s = 1.0 / std::sqrt(3.0);
        x[0] = (-s + 1.0) / 2.0;
        x[1] = (s + 1.0) / 2.0;
        w[0] = w[1] = 0.5;
    } else {  // n == 3
        const double s = std::sqrt(0.6);
        x[0] = (-s + 1.0) / 2.0;
        x[1] = 0.5;
        x[2] = (s + 1.0) / 2.0;
        w[0] = 2.5 / 9.0;
        w[1] = 4.0 / 9.0;
        w[2] = 2.5 / 9.0;
    }
}

static void tabulate(int order, int ng1, double *w, double *fns)
{
    double gx[3], gw[3];
    gauss_1d(ng1, gx, gw);
    const int nx = order + 1, ns = nx * nx * nx;
    int g = 0;
    for (int k = 0; k < ng1; ++k)
        for (int j = 0; j < ng1; ++j)
            for (int i = 0; i < ng1; ++i, ++g) {
                const double xi[3] = {gx[i], gx[j], gx[k]};
                w[g] = gw[i] * gw[j] * gw[k];
                int s = 0;
                for (int n3 = 0; n3 < nx; ++n3)
                    for (int n2 = 0; n2 < nx; ++n2)
                        for (int n1 = 0; n1 < nx; ++n1, ++s) {
                            const int pos[3] = {n1, n2, n3};
                            for (int nu = 0; nu < 4; ++nu) {
                                double v = 1.0;
                                for (int d = 0; d < 3; ++d) v *= lagrange_1d(order, pos[d], nu == d + 1, xi[d]);
                                fns[(g * ns + s) * 4 + nu] = v;
                            }
                        }
            }
}

void build_basis_tables(BasisTables &t)
{
    double wtmp[27];
    tabulate(2, 3, t.w27, t.quad27);
    tabulate(1, 3, wtmp, t.lin27);
    tabulate(1, 2, t.w8, t.lin8);
}

static BasisTables *g_device_tables = nullptr;
static int g_tables_device = -1;

static int ensure_tables(const BasisTables **out)
{
    Runtime &rt = runtime();
    if (g_device_tables == nullptr || g_tables_device != rt.device) {
        static BasisTables host;
        build_basis_tables(host);
        void *p = nullptr;
        OC_HIP_CHECK(hipMalloc(&p, sizeof(BasisTables)));
        OC_HIP_CHECK(hipMemcpy(p, &host, sizeof(BasisTables), hipMemcpyHostToDevice));
        g_device_tables = static_cast<BasisTables *>(p);
        g_tables_device = rt.device;
    }
    *out = g_device_tables;
    return OC_NO_ERROR;
}

const BasisTables *device_basis_tables() { return g_device_tables; }

// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) fill_kernel(T *__restrict__ p, int64_t n, T v)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <typename T>
static int fill(T *p, int64_t n, T v)
{
    if (n <= 0) return OC_NO_ERROR;
    Runtime &rt = runtime();
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, static_cast<int64_t>(rt.num_cus) * 16);
    fill_kernel<T><<<static_cast<unsigned>(blocks), 256, 0, rt.stream>>>(p, n, v);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

// ---------------------------------------------------------------------------
// Laplace, standard subtype (Laplace_equations_routines.f90:474-571):
// K_e(ms,ns) += sum_{mi,ni} dphi_ms/dxi_mi GU(mi,ni) dphi_ns/dxi_ni * J w.
// Eight lanes per element: lane g first evaluates the metric of Gauss point
// g, then lane ms owns element row ms and scatters it into CSR row node(ms).
constexpr int kLapElemsPerBlock = 32;

__global__ void __launch_bounds__(256)
laplace_assemble_kernel(const int *__restrict__ colour_elems, int n_in_colour, const int *__restrict__ conn,
                        const double *__restrict__ coords, int n_nodes, const int *__restrict__ row_ptr,
                        const int *__restrict__ col_idx, double *__restrict__ values,
                        const BasisTables *__restrict__ T)
{
    __shared__ double s_x[kLapElemsPerBlock][3][8];
    __shared__ double s_gu[kLapElemsPerBlock][8][10];
    __shared__ int s_node[kLapElemsPerBlock][8];
    const int local = threadIdx.x >> 3, lane = threadIdx.x & 7;
    const int idx = blockIdx.x * kLapElemsPerBlock + local;
    const bool active = idx < n_in_colour;
    if (active) {
        const int e = colour_elems[idx];
        const int node = conn[e * 8 + lane];
        s_node[local][lane] = node;
        for (int c = 0; c < 3; ++c) s_x[local][c][lane] = coords[static_cast<int64_t>(c) * n_nodes + node];
    }
    __syncthreads();
    if (active) {
        const int g = lane;
        const double *f = T->lin8 + g * 8 * 4;
        double dx[9], gu[9];
        for (int x = 0; x < 3; ++x)
            for (int ni = 0; ni < 3; ++ni) {
                double s = 0.0;
                for (int n = 0; n < 8; ++n) s += f[n * 4 + 1 + ni] * s_x[local][x][n];
                dx[x + 3 * ni] = s;
            }
        const double jac = metric_jacobian(dx, gu);
        for (int i = 0; i < 9; ++i) s_gu[local][g][i] = gu[i];
        s_gu[local][g][9] = jac * T->w8[g];
    }
    __syncthreads();
    if (!active) return;
    const int ms = lane;
    double k[8];
#pragma unroll
    for (int ns = 0; ns < 8; ++ns) k[ns] = 0.0;
    for (int g = 0; g < 8; ++g) {
        const double *f = T->lin8 + g * 8 * 4;
        const double *gu = s_gu[local][g];
        double v[3];
#pragma unroll
        for (int ni = 0; ni < 3; ++ni)
            v[ni] = f[ms * 4 + 1] * gu[0 + 3 * ni] + f[ms * 4 + 2] * gu[1 + 3 * ni] + f[ms * 4 + 3] * gu[2 + 3 * ni];
        const double rwg = gu[9];
#pragma unroll
        for (int ns = 0; ns < 8; ++ns) {
            const double sum = v[0] * f[ns * 4 + 1] + v[1] * f[ns * 4 + 2] + v[2] * f[ns * 4 + 3];
            k[ns] = k[ns] + sum * rwg;
        }
    }
    const int row = s_node[local][ms];
#pragma unroll
    for (int ns = 0; ns < 8; ++ns) {
        const int slot = csr_slot(row_ptr, col_idx, row, s_node[local][ns]);
        if (slot >= 0) values[slot] += k[ns];
    }
}

int launch_laplace_assemble(EquationsSet &s)
{
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    OC_CHECK(fill<double>(s.values.ptr, s.nnz, 0.0));
    OC_CHECK(fill<double>(s.rhs.ptr, s.n_rows, 0.0));  // RHS element vector is 0 (:542)
    const int ncol = static_cast<int>(s.colour_ptr.size()) - 1;
    for (int c = 0; c < ncol; ++c) {
        const int n = s.colour_ptr[c + 1] - s.colour_ptr[c];
        if (n == 0) continue;
        const unsigned blocks = static_cast<unsigned>((n + kLapElemsPerBlock - 1) / kLapElemsPerBlock);
        laplace_assemble_kernel<<<blocks, 256, 0, rt.stream>>>(s.colour_elems.ptr + s.colour_ptr[c], n, s.conn.ptr,
                                                              s.coords.ptr, s.n_nodes, s.row_ptr.ptr, s.col_idx.ptr,
                                                              s.values.ptr, T);
        OC_HIP_CHECK(hipGetLastError());
    }
    return OC_NO_ERROR;
}

// ---------------------------------------------------------------------------
// Mooney-Rivlin residual (FiniteElasticity_FiniteElementResidualEvaluate,
// MR branch, finite_elasticity_routines.f90:1488-1575).  One wavefront per
// element: lanes 0..26 evaluate one Gauss point each (kinematics, Cauchy
// stress, dphi/dz), then lanes own element residual entries and sum the
// Gauss points in order.
__global__ void __launch_bounds__(64)
mr_residual_kernel(const int *__restrict__ colour_elems, const int *__restrict__ conn_q,
                   const int *__restrict__ conn_l, const double *__restrict__ coords, int nq,
                   const double *__restrict__ dofs, double c10, double c01, double *__restrict__ residual,
                   const BasisTables *__restrict__ T)
{
    __shared__ double s_X[81], s_z[81], s_p[8];
    __shared__ double s_D[27][81];  // dphi_n/dz_i at g: [i*27 + n]
    __shared__ double s_S[27][9];   // JGW * sigma(i, a): [i + 3a]
    __shared__ double s_t[27];      // JGW_X * (J - 1)
    __shared__ int s_dof[89];
    const int e = colour_elems[blockIdx.x];
    const int t = threadIdx.x;
    for (int k = t; k < 81; k += 64) {
        const int c = k / 27, n = k % 27;
        const int node = conn_q[e * 27 + n];
        s_X[k] = coords[static_cast<int64_t>(c) * nq + node];
        s_z[k] = dofs[static_cast<int64_t>(c) * nq + node];
        s_dof[k] = c * nq + node;
    }
    if (t < 8) {
        const int pn = conn_l[e * 8 + t];
        s_p[t] = dofs[static_cast<int64_t>(3) * nq + pn];
        s_dof[81 + t] = 3 * nq + pn;
    }
    __syncthreads();
    if (t < 27) {
        const int g = t;
        const double *fq = T->quad27 + g * 27 * 4;
        GaussKinematics k;
        gauss_kinematics(fq, s_z, s_X, k);
        double P = 0.0;
        for (int n = 0; n < 8; ++n) P += T->lin27[(g * 8 + n) * 4] * s_p[n];
        double sigma[6];
        mr_cauchy_stress(k.F, k.Jznu, P, c10, c01, sigma);
        const double jgw = k.jz * T->w27[g];
        for (int a = 0; a < 3; ++a)
            for (int i = 0; i < 3; ++i) s_S[g][i + 3 * a] = jgw * sigma[t2v(i, a)];
        for (int n = 0; n < 27; ++n)
            for (int mh = 0; mh < 3; ++mh)
                s_D[g][mh * 27 + n] = k.dxidz[0 + 3 * mh] * fq[n * 4 + 1] + k.dxidz[1 + 3 * mh] * fq[n * 4 + 2] +
                                      k.dxidz[2 + 3 * mh] * fq[n * 4 + 3];
        s_t[g] = k.jx * T->w27[g] * (k.Jznu - 1.0);
    }
    __syncthreads();
    for (int r = t; r < 89; r += 64) {
        double acc = 0.0;
        if (r < 81) {
            const int a = r / 27, m = r % 27;
            for (int g = 0; g < 27; ++g) {
                const double dot = s_D[g][m] * s_S[g][0 + 3 * a] + s_D[g][27 + m] * s_S[g][1 + 3 * a] +
                                   s_D[g][54 + m] * s_S[g][2 + 3 * a];
                acc = acc + dot;
            }
        } else {
            const int m = r - 81;
            for (int g = 0; g < 27; ++g) acc = acc + s_t[g] * T->lin27[(g * 8 + m) * 4];
        }
        residual[s_dof[r]] += acc;
    }
}

int launch_mr_residual(EquationsSet &s)
{
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    OC_CHECK(fill<double>(s.residual.ptr, s.n_rows, 0.0));
    const int ncol = static_cast<int>(s.colour_ptr.size()) - 1;
    for (int c = 0; c < ncol; ++c) {
        const int n = s.colour_ptr[c + 1] - s.colour_ptr[c];
        if (n == 0) continue;
        mr_residual_kernel<<<static_cast<unsigned>(n), 64, 0, rt.stream>>>(
            s.colour_elems.ptr + s.colour_ptr[c], s.conn.ptr, s.conn_l.ptr, s.coords.ptr, s.n_nodes, s.dofs.ptr,
            s.c10, s.c01, s.residual.ptr, T);
        OC_HIP_CHECK(hipGetLastError());
    }
    return OC_NO_ERROR;
}

// ---------------------------------------------------------------------------
// Mooney-Rivlin analytic Jacobian (FiniteElasticity_FiniteElementJacobian
// Evaluate, finite_elasticity_routines.f90:842-1175) as
//   K_(a,m),(b,n) = sum_g sum_i D_g[i][m] T_g^{ab}[i][n],
//   T_g^{ab}[i][n] = sum_j W_g[(a,i),(b,j)] D_g[j][n],
//   W_g[(a,i),(b,j)] = J_z w_g (c(V(i,a),V(j,b)) + delta_ab sigma_ij),
// which is the reference's TEMPVEC/DOT_PRODUCT contraction (:1001-1037);
// pressure rows K_(p,m),(b,n) = sum_g J_z w_g D_g[b][n] psi_m (:1039-1054)
// and the symmetric mirror (:1149-1153).  Gauss-point state is FP64; the
// contraction runs in Real (double, or float for FP32 Jacobians).
// 256 threads per element: 162 threads own 3x9 tiles of the six lower
// displacement blocks, 81 own a pressure column.
template <typename Real>
__global__ void __launch_bounds__(256)
mr_jacobian_kernel(const int *__restrict__ colour_elems, const int *__restrict__ conn_q,
                   const int *__restrict__ conn_l, const double *__restrict__ coords, int nq,
                   const double *__restrict__ dofs, double c10, double c01, const int *__restrict__ row_ptr,
                   const int *__restrict__ col_idx, Real *__restrict__ values, const BasisTables *__restrict__ T)
{
    __shared__ double s_X[81], s_z[81], s_p[8];
    __shared__ Real s_D[27][81];
    __shared__ Real s_W[27][81];
    __shared__ Real s_jgw[27];
    __shared__ Real s_T[6][81];
    __shared__ int s_dof[89];
    const int blk_a[6] = {0, 1, 1, 2, 2, 2}, blk_b[6] = {0, 0, 1, 0, 1, 2};
    const int e = colour_elems[blockIdx.x];
    const int t = threadIdx.x;
    for (int k = t; k < 81; k += 256) {
        const int c = k / 27, n = k % 27;
        const int node = conn_q[e * 27 + n];
        s_X[k] = coords[static_cast<int64_t>(c) * nq + node];
        s_z[k] = dofs[static_cast<int64_t>(c) * nq + node];
        s_dof[k] = c * nq + node;
    }
    if (t < 8) {
        const int pn = conn_l[e * 8 + t];
        s_p[t] = dofs[static_cast<int64_t>(3) * nq + pn];
        s_dof[81 + t] = 3 * nq + pn;
    }
    __syncthreads();
    if (t < 27) {
        const int g = t;
        const double *fq = T->quad27 + g * 27 * 4;
        GaussKinematics k;
        gauss_kinematics(fq, s_z, s_X, k);
        double P = 0.0;
        for (int n = 0; n < 8; ++n) P += T->lin27[(g * 8 + n) * 4] * s_p[n];
        double c[36], sigma[6];
        mr_elasticity_tensor(k.F, k.Jznu, P, c10, c01, c, sigma);
        const double jgw = k.jz * T->w27[g];
        for (int a = 0; a < 3; ++a)
            for (int i = 0; i < 3; ++i)
                for (int b = 0; b < 3; ++b)
                    for (int j = 0; j < 3; ++j) {
                        double v = c[t2v(i, a) + 6 * t2v(j, b)];
                        if (a == b) v = v + sigma[t2v(i, j)];
                        s_W[g][(a * 3 + i) * 9 + b * 3 + j] = static_cast<Real>(jgw * v);
                    }
        for (int n = 0; n < 27; ++n)
            for (int mh = 0; mh < 3; ++mh)
                s_D[g][mh * 27 + n] = static_cast<Real>(k.dxidz[0 + 3 * mh] * fq[n * 4 + 1] +
                                                        k.dxidz[1 + 3 * mh] * fq[n * 4 + 2] +
                                                        k.dxidz[2 + 3 * mh] * fq[n * 4 + 3]);
        s_jgw[g] = static_cast<Real>(jgw);
    }
    __syncthreads();

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
            const Real jd = s_jgw[g] * s_D[g][pb * 27 + pn];
#pragma unroll
            for (int m = 0; m < 8; ++m) acc[m] = acc[m] + jd * static_cast<Real>(T->lin27[(g * 8 + m) * 4]);
        }
        __syncthreads();
    }
    if (is_uu) {
        const int a = blk_a[blk], b = blk_b[blk];
        for (int mm = 0; mm < 3; ++mm)
            for (int nn = 0; nn < 9; ++nn) {
                const int r = s_dof[a * 27 + m0 + mm], c = s_dof[b * 27 + n0 + nn];
                const Real v = acc[mm * 9 + nn];
                int slot = csr_slot(row_ptr, col_idx, r, c);
                if (slot >= 0) values[slot] += v;
                if (a != b) {
                    slot = csr_slot(row_ptr, col_idx, c, r);
                    if (slot >= 0) values[slot] += v;
                }
            }
    } else if (is_p) {
        const int c = s_dof[pb * 27 + pn];
        for (int m = 0; m < 8; ++m) {
            const int r = s_dof[81 + m];
            int slot = csr_slot(row_ptr, col_idx, r, c);
            if (slot >= 0) values[slot] += acc[m];
            slot = csr_slot(row_ptr, col_idx, c, r);
            if (slot >= 0) values[slot] += acc[m];
        }
    }
}

int launch_mr_jacobian(EquationsSet &s, int precision)
{
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    if (precision == OC_FP64) OC_CHECK(fill<double>(s.values.ptr, s.nnz, 0.0));
    else OC_CHECK(fill<float>(s.values_sp.ptr, s.nnz, 0.0f));
    const int ncol = static_cast<int>(s.colour_ptr.size()) - 1;
    for (int c = 0; c < ncol; ++c) {
        const int n = s.colour_ptr[c + 1] - s.colour_ptr[c];
        if (n == 0) continue;
        const int *ce = s.colour_elems.ptr + s.colour_ptr[c];
        if (precision == OC_FP64)
            mr_jacobian_kernel<double><<<static_cast<unsigned>(n), 256, 0, rt.stream>>>(
                ce, s.conn.ptr, s.conn_l.ptr, s.coords.ptr, s.n_nodes, s.dofs.ptr, s.c10, s.c01, s.row_ptr.ptr,
                s.col_idx.ptr, s.values.ptr, T);
        else
            mr_jacobian_kernel<float><<<static_cast<unsigned>(n), 256, 0, rt.stream>>>(
                ce, s.conn.ptr, s.conn_l.ptr, s.coords.ptr, s.n_nodes, s.dofs.ptr, s.c10, s.c01, s.row_ptr.ptr,
                s.col_idx.ptr, s.values_sp.ptr, T);
        OC_HIP_CHECK(hipGetLastError());
    }
    return OC_NO_ERROR;
}

}  // namespace oc
