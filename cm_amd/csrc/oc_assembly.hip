// oc_assembly.hip -- basis tables, fill kernel and the gfx950 Laplace
// linear-matrix assembly: element matrices, then a deterministic write-once
// CSR row gather in ascending element order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

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

int ensure_tables(const BasisTables **out)
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
int fill(T *p, int64_t n, T v)
{
    if (n <= 0) return OC_NO_ERROR;
    Runtime &rt = runtime();
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, static_cast<int64_t>(rt.num_cus) * 16);
    KernelTimer timer("fill");
    fill_kernel<T><<<static_cast<unsigned>(blocks), 256, 0, rt.stream>>>(p, n, v);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}
template int fill<double>(double *, int64_t, double);
template int fill<float>(float *, int64_t, float);

// ---------------------------------------------------------------------------
// Laplace, standard subtype (Laplace_equations_routines.f90:474-571):
// K_e(ms,ns) += sum_{mi,ni} dphi_ms/dxi_mi GU(mi,ni) dphi_ns/dxi_ni * J w.
// Two phases, as the Jacobian: laplace_element_kernel writes every element
// matrix to a buffer (eight lanes per element: lane g evaluates the metric of
// Gauss point g, then lane ms owns element row ms and writes it as 64
// contiguous bytes); laplace_row_gather_kernel gives each CSR row one thread,
// which adds the element rows touching it in ascending element order (the
// reference's element-loop order) into an LDS copy of its block's rows, and
// the block writes its contiguous CSR segment once: no zero fill, no
// read-modify-write, no colours.
constexpr int kLapElemsPerBlock = 32;

__global__ void __launch_bounds__(256)
laplace_slot_offsets_kernel(int n_elements, const int *__restrict__ conn, const int64_t *__restrict__ row_ptr,
                            const int *__restrict__ col_idx, uint8_t *__restrict__ offs, int *__restrict__ n_bad)
{
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<int64_t>(n_elements) * 8) return;
    const int64_t e = gid / 8;
    const int row = conn[gid];
    const int64_t base = row_ptr[row];
    for (int j = 0; j < 8; ++j) {
        const int64_t slot = csr_slot(row_ptr, col_idx, row, conn[e * 8 + j]);
        if (slot < 0 || slot - base > 255) {
            atomicAdd(n_bad, 1);
            offs[gid * 8 + j] = 0;
        } else {
            offs[gid * 8 + j] = static_cast<uint8_t>(slot - base);
        }
    }
}

__global__ void __launch_bounds__(256)
laplace_element_kernel(int n_elements, const int *__restrict__ conn, const double *__restrict__ coords, int n_nodes,
                       const BasisTables *__restrict__ T, double *__restrict__ kbuf)
{
    __shared__ double s_x[kLapElemsPerBlock][3][8];
    __shared__ double s_gu[kLapElemsPerBlock][8][10];
    __shared__ double s_f[8 * 8 * 4];  // trilinear basis at the 2x2x2 Gauss points
    const int local = threadIdx.x >> 3, lane = threadIdx.x & 7;
    const int e = blockIdx.x * kLapElemsPerBlock + local;
    const bool active = e < n_elements;
    s_f[threadIdx.x] = T->lin8[threadIdx.x];  // 256 entries, one per thread
    if (active) {
        const int node = conn[static_cast<int64_t>(e) * 8 + lane];
        for (int c = 0; c < 3; ++c) s_x[local][c][lane] = coords[static_cast<int64_t>(c) * n_nodes + node];
    }
    __syncthreads();
    if (active) {
        const int g = lane;
        const double *f = s_f + g * 8 * 4;
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
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        const double *f = s_f + g * 8 * 4;
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
    double2 *out = reinterpret_cast<double2 *>(kbuf + (static_cast<int64_t>(e) * 8 + ms) * 8);
#pragma unroll
    for (int q = 0; q < 4; ++q) out[q] = make_double2(k[2 * q], k[2 * q + 1]);
}

template <int kRows>
__global__ void __launch_bounds__(kRows)
laplace_row_gather_kernel(int n_rows, const int64_t *__restrict__ row_ptr, const int *__restrict__ gptr,
                          const int *__restrict__ gidx, const double *__restrict__ kbuf,
                          const uint8_t *__restrict__ offs, double *__restrict__ values)
{
    extern __shared__ unsigned char s_raw[];
    double *s_seg = reinterpret_cast<double *>(s_raw);
    const int r0 = blockIdx.x * kRows, t = threadIdx.x, r = r0 + t;
    const int r_end = min(r0 + kRows, n_rows);
    const int64_t base = row_ptr[r0];
    const int seg = static_cast<int>(row_ptr[r_end] - base);
    for (int k = t; k < seg; k += kRows) s_seg[k] = 0.0;
    __syncthreads();
    if (r < n_rows) {
        double *row = s_seg + (row_ptr[r] - base);
        for (int c = gptr[r]; c < gptr[r + 1]; ++c) {
            const int64_t ei = gidx[c];  // element * 8 + element row
            const uint64_t packed = *reinterpret_cast<const uint64_t *>(offs + ei * 8);
            const double2 *kr = reinterpret_cast<const double2 *>(kbuf + ei * 8);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double2 v = kr[q];
                row[(packed >> (16 * q)) & 0xff] += v.x;
                row[(packed >> (16 * q + 8)) & 0xff] += v.y;
            }
        }
    }
    __syncthreads();
    for (int k = t; k < seg; k += kRows) values[base + k] = s_seg[k];
}

static int laplace_prepare(EquationsSet &s)
{
    if (s.slot_offsets8.ptr) return OC_NO_ERROR;
    OC_CHECK(s.slot_offsets8.alloc(static_cast<size_t>(s.n_elements) * 64));
    DevBuf<int> bad;
    OC_CHECK(bad.alloc(1));
    Runtime &rt = runtime();
    OC_HIP_CHECK(hipMemsetAsync(bad.ptr, 0, sizeof(int), rt.stream));
    const int64_t nthreads = static_cast<int64_t>(s.n_elements) * 8;
    laplace_slot_offsets_kernel<<<static_cast<unsigned>((nthreads + 255) / 256), 256, 0, rt.stream>>>(
        s.n_elements, s.conn.ptr, s.row_ptr.ptr, s.col_idx.ptr, s.slot_offsets8.ptr, bad.ptr);
    OC_HIP_CHECK(hipGetLastError());
    int h_bad = 0;
    OC_HIP_CHECK(hipMemcpyAsync(&h_bad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, rt.stream));
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    if (h_bad) {
        s.slot_offsets8.release();
        return set_error(OC_ERR_PATTERN, "%d Laplace element entries missing from the CSR structure or beyond "
                                         "the 8-bit scatter map", h_bad);
    }
    if (static_cast<int64_t>(s.n_elements) * 8 >= INT32_MAX)
        return set_error(OC_ERR_UNSUPPORTED, "%d elements exceed the int32 element-row gather lists", s.n_elements);
    // rows per gather block: the block's CSR segment must fit the LDS
    int64_t max_seg = 0;
    for (int r0 = 0; r0 < s.n_rows; r0 += 256)
        max_seg = std::max(max_seg, s.h_row_ptr[std::min(r0 + 256, s.n_rows)] - s.h_row_ptr[r0]);
    s.lap_rows_per_block = 256;
    if (max_seg > 8192) {
        max_seg = 0;
        for (int r0 = 0; r0 < s.n_rows; r0 += 64)
            max_seg = std::max(max_seg, s.h_row_ptr[std::min(r0 + 64, s.n_rows)] - s.h_row_ptr[r0]);
        s.lap_rows_per_block = 64;
        if (max_seg > 8192)
            return set_error(OC_ERR_UNSUPPORTED, "Laplace CSR rows too long for the row gather (%lld entries in 64 rows)",
                             static_cast<long long>(max_seg));
    }
    s.max_row_len = static_cast<int>(max_seg);
    // DOF -> element rows, ascending element order
    std::vector<int32_t> hconn(static_cast<size_t>(s.n_elements) * 8);
    OC_HIP_CHECK(hipMemcpy(hconn.data(), s.conn.ptr, hconn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> ptr(static_cast<size_t>(s.n_rows) + 1, 0), idx(hconn.size());
    for (size_t k = 0; k < hconn.size(); ++k) ptr[hconn[k] + 1]++;
    for (int d = 0; d < s.n_rows; ++d) ptr[d + 1] += ptr[d];
    std::vector<int32_t> fillp(ptr.begin(), ptr.end() - 1);
    for (size_t k = 0; k < hconn.size(); ++k) idx[fillp[hconn[k]]++] = static_cast<int32_t>(k);
    OC_CHECK(s.gather_ptr.upload(ptr.data(), ptr.size()));
    OC_CHECK(s.gather_idx.upload(idx.data(), idx.size()));
    OC_CHECK(s.kbuf64.alloc(static_cast<size_t>(s.n_elements) * 64));
    return OC_NO_ERROR;
}

int launch_laplace_assemble(EquationsSet &s)
{
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    OC_CHECK(laplace_prepare(s));
    Runtime &rt = runtime();
    OC_CHECK(fill<double>(s.rhs.ptr, s.n_rows, 0.0));  // RHS element vector is 0 (:542)
    {
        KernelTimer timer("laplace_element");
        const unsigned blocks = static_cast<unsigned>((s.n_elements + kLapElemsPerBlock - 1) / kLapElemsPerBlock);
        laplace_element_kernel<<<blocks, 256, 0, rt.stream>>>(s.n_elements, s.conn.ptr, s.coords.ptr, s.n_nodes, T,
                                                             s.kbuf64.ptr);
        OC_HIP_CHECK(hipGetLastError());
    }
    KernelTimer timer("laplace_row_gather");
    const size_t lds = static_cast<size_t>(s.max_row_len) * sizeof(double);
    if (s.lap_rows_per_block == 256)
        laplace_row_gather_kernel<256><<<static_cast<unsigned>((s.n_rows + 255) / 256), 256, lds, rt.stream>>>(
            s.n_rows, s.row_ptr.ptr, s.gather_ptr.ptr, s.gather_idx.ptr, s.kbuf64.ptr, s.slot_offsets8.ptr,
            s.values.ptr);
    else
        laplace_row_gather_kernel<64><<<static_cast<unsigned>((s.n_rows + 63) / 64), 64, lds, rt.stream>>>(
            s.n_rows, s.row_ptr.ptr, s.gather_ptr.ptr, s.gather_idx.ptr, s.kbuf64.ptr, s.slot_offsets8.ptr,
            s.values.ptr);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

}  // namespace oc
