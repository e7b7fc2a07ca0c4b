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
// matrix to a buffer (one thread per element); laplace_row_gather_kernel
// gives each CSR row eight lanes, which add the element rows touching it in
// ascending element order (the reference's element-loop order) into an LDS
// copy of its block's rows, and the block writes its contiguous CSR segment
// once: no zero fill, no read-modify-write, no colours.
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

// The 36 upper-triangle entries K[ms*(15-ms)/2 + ns] (ms <= ns) of one
// trilinear element matrix from its node coordinates x[c][n]
// (Laplace_equations_routines.f90:474-571).  Restated for the tensor-product
// structure of the trilinear basis at the 2x2x2 Gauss points:
//   * dX/dxi by sum factorisation: along xi_d a trilinear map's derivative is
//     the bilinear interpolant, in the other two coordinates, of its four edge
//     differences along d, so 36 values serve all eight Gauss points;
//   * the basis derivatives are products of the 1-D factors 1 - xi, xi at
//     the Gauss abscissae, compile-time constants;
//   * GU J w = adj(GL) w / sqrt(det GL), since J = sqrt(det GL) and
//     GU = adj(GL) / det GL (COORDINATE_METRICS_CALCULATE and
//     INVERT_FULL_DP, whose zero-tolerance rule it keeps);
//   * the basis derivatives sum to zero at every point, so each row of K_e
//     sums to zero: the loop accumulates the 28 entries among local nodes
//     0-6 and the entries of node 7 follow as negated row sums.
// The Gauss loop unrolls, so every index is a constant.
__device__ __forceinline__ int lap_upper(int ms, int ns) { return ms * (15 - ms) / 2 + ns; }

__device__ __forceinline__ double lerp(double a, double b, double q) { return fma(q, b - a, a); }

// the 2-point Gauss abscissae on [0, 1], (1 -+ 1/sqrt 3) / 2 rounded as the
// host basis tables compute them, and the linear factors L[g][n] of local
// coordinate n (0 or 1) at abscissa g: 1 - xi, xi
constexpr double kLapXi[2] = {0x1.b0cb174df99c6p-3, 0x1.93cd3a2c8198ep-1};
__device__ constexpr double lap_factor(int g, int n) { return n ? kLapXi[g] : 1.0 - kLapXi[g]; }

__device__ __forceinline__ void laplace_element_upper(const double (&x)[3][8], double (&K)[36])
{
    // every multiply-add is an explicit fma (contraction is off here), so the
    // rounding is the same wherever the function is compiled: the fused and
    // the two-pass assembly write bit-identical values
#pragma clang fp contract(off)
    // D[c][d][2 gb + ga]: dX_c/dxi_d at Gauss indices (ga, gb) of the other
    // two directions (ascending)
    double D[3][3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int sd = 1 << d, sa = d == 0 ? 2 : 1, sb = d == 2 ? 2 : 4;
            double e[2][2];
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                for (int na = 0; na < 2; ++na) e[nb][na] = x[c][sd + sa * na + sb * nb] - x[c][sa * na + sb * nb];
#pragma unroll
            for (int gb = 0; gb < 2; ++gb)
#pragma unroll
                for (int ga = 0; ga < 2; ++ga)
                    D[c][d][2 * gb + ga] = lerp(lerp(e[0][0], e[0][1], kLapXi[ga]), lerp(e[1][0], e[1][1], kLapXi[ga]),
                                                kLapXi[gb]);
        }
    double Kr[28];  // entries (ms, ns), ms <= ns <= 6, at lap_upper(ms, ns) - ms
#pragma unroll
    for (int i = 0; i < 28; ++i) Kr[i] = 0.0;
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        const int g1 = g & 1, g2 = (g >> 1) & 1, g3 = g >> 2;
        double dx[9];  // dx[c + 3 mi] = dX_c/dxi_mi
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dx[c] = D[c][0][2 * g3 + g2];
            dx[c + 3] = D[c][1][2 * g3 + g1];
            dx[c + 6] = D[c][2][2 * g2 + g1];
        }
        double gl[6];  // GL = dX^T dX: 00 11 22 01 02 12
        auto dot3 = [](double a0, double b0, double a1, double b1, double a2, double b2) {
            return fma(a2, b2, fma(a1, b1, a0 * b0));
        };
        gl[0] = dot3(dx[0], dx[0], dx[1], dx[1], dx[2], dx[2]);
        gl[1] = dot3(dx[3], dx[3], dx[4], dx[4], dx[5], dx[5]);
        gl[2] = dot3(dx[6], dx[6], dx[7], dx[7], dx[8], dx[8]);
        gl[3] = dot3(dx[0], dx[3], dx[1], dx[4], dx[2], dx[5]);
        gl[4] = dot3(dx[0], dx[6], dx[1], dx[7], dx[2], dx[8]);
        gl[5] = dot3(dx[3], dx[6], dx[4], dx[7], dx[5], dx[8]);
        auto diff2 = [](double a, double b, double c, double d) { return fma(a, b, -(c * d)); };  // ab - cd
        const double c00 = diff2(gl[1], gl[2], gl[5], gl[5]), c11 = diff2(gl[0], gl[2], gl[4], gl[4]);
        const double c22 = diff2(gl[0], gl[1], gl[3], gl[3]), c01 = diff2(gl[4], gl[5], gl[3], gl[2]);
        const double c02 = diff2(gl[3], gl[5], gl[4], gl[1]), c12 = diff2(gl[3], gl[4], gl[0], gl[5]);
        const double det = dot3(gl[0], c00, gl[3], c01, gl[4], c02);  // det GL = det(dX)^2
        // weight 1/8 of the 2x2x2 rule on [0, 1]^3
        const double sc = fabs(det) > 5.0 * 2.220446049250313e-16 ? copysign(0.125 * rsqrt(fabs(det)), det) : 0.0;
        const double A[9] = {c00 * sc, c01 * sc, c02 * sc, c01 * sc, c11 * sc, c12 * sc, c02 * sc, c12 * sc, c22 * sc};
        // dphi_n/dxi at g: (+-1) times the other two directions' factors
        double f[8][3];
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const int n1 = n & 1, n2 = (n >> 1) & 1, n3 = n >> 2;
            f[n][0] = (n1 ? 1.0 : -1.0) * (lap_factor(g2, n2) * lap_factor(g3, n3));
            f[n][1] = (n2 ? 1.0 : -1.0) * (lap_factor(g1, n1) * lap_factor(g3, n3));
            f[n][2] = (n3 ? 1.0 : -1.0) * (lap_factor(g1, n1) * lap_factor(g2, n2));
        }
        int idx = 0;
#pragma unroll
        for (int ms = 0; ms < 7; ++ms) {
            double v[3];
#pragma unroll
            for (int ni = 0; ni < 3; ++ni)
                v[ni] = dot3(f[ms][0], A[0 + 3 * ni], f[ms][1], A[1 + 3 * ni], f[ms][2], A[2 + 3 * ni]);
#pragma unroll
            for (int ns = ms; ns < 7; ++ns, ++idx)
                Kr[idx] = fma(v[2], f[ns][2], fma(v[1], f[ns][1], fma(v[0], f[ns][0], Kr[idx])));
        }
    }
    int idx = 0;
#pragma unroll
    for (int ms = 0; ms < 7; ++ms)
#pragma unroll
        for (int ns = ms; ns < 7; ++ns, ++idx) K[lap_upper(ms, ns)] = Kr[idx];
    double k77 = 0.0;
#pragma unroll
    for (int ms = 0; ms < 7; ++ms) {
        double r = 0.0;
#pragma unroll
        for (int ns = 0; ns < 7; ++ns) r += K[ms <= ns ? lap_upper(ms, ns) : lap_upper(ns, ms)];
        K[lap_upper(ms, 7)] = -r;
        k77 += r;
    }
    K[lap_upper(7, 7)] = k77;
}

// One thread per element; each row of K_e goes to its place in the
// row-gather order, so the gather reads every CSR row's contributions
// contiguously.
constexpr int kLapThreads = 64;

__global__ void __launch_bounds__(kLapThreads)
laplace_element_kernel(int n_elements, const int *__restrict__ conn, const double *__restrict__ coords, int n_nodes,
                       const int *__restrict__ gpos, double *__restrict__ kbuf)
{
    const int e = blockIdx.x * kLapThreads + threadIdx.x;
    if (e >= n_elements) return;
    double x[3][8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int node = conn[static_cast<int64_t>(e) * 8 + n];
#pragma unroll
        for (int c = 0; c < 3; ++c) x[c][n] = coords[static_cast<int64_t>(c) * n_nodes + node];
    }
    double K[36];
    laplace_element_upper(x, K);
    // row ms of this element goes to its position in the row-gather order
    double row[64];
    int idx = 0;
#pragma unroll
    for (int ms = 0; ms < 8; ++ms)
#pragma unroll
        for (int ns = ms; ns < 8; ++ns, ++idx) {
            row[ms * 8 + ns] = K[idx];
            row[ns * 8 + ms] = K[idx];
        }
#pragma unroll
    for (int ms = 0; ms < 8; ++ms) {
        double2 *out = reinterpret_cast<double2 *>(kbuf + static_cast<int64_t>(gpos[static_cast<int64_t>(e) * 8 + ms]) * 8);
#pragma unroll
        for (int q = 0; q < 4; ++q) out[q] = make_double2(row[ms * 8 + 2 * q], row[ms * 8 + 2 * q + 1]);
    }
}

// slot offsets permuted into the row-gather order
__global__ void __launch_bounds__(256) permute_offsets_kernel(int64_t n, const int *__restrict__ idx,
                                                             const uint64_t *__restrict__ offs,
                                                             uint64_t *__restrict__ out)
{
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < n;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x)
        out[p] = offs[idx[p]];
}

// Eight lanes per CSR row (32 rows per block): lane j adds column j of each
// of the row's element rows, in ascending element order, into the block's
// LDS copy of its CSR segment; a chunk's global loads (eight element rows)
// are issued before its LDS adds.  The block then writes the segment once.
constexpr int kLapGatherRows = 32;

__global__ void __launch_bounds__(8 * kLapGatherRows)
laplace_row_gather_kernel(int n_rows, const int64_t *__restrict__ row_ptr, const int *__restrict__ gptr,
                          const double *__restrict__ kbuf, const uint8_t *__restrict__ offs,
                          double *__restrict__ values)
{
    extern __shared__ unsigned char s_raw[];
    double *s_seg = reinterpret_cast<double *>(s_raw);
    const int t = threadIdx.x, j = t & 7;
    const int r0 = blockIdx.x * kLapGatherRows, r = r0 + (t >> 3);
    const int r_end = min(r0 + kLapGatherRows, n_rows);
    const int64_t base = row_ptr[r0];
    const int seg = static_cast<int>(row_ptr[r_end] - base);
    for (int k = t; k < seg; k += 8 * kLapGatherRows) s_seg[k] = 0.0;
    __syncthreads();
    if (r < n_rows) {
        double *row = s_seg + (row_ptr[r] - base);
        const int c_end = gptr[r + 1];
        for (int c0 = gptr[r]; c0 < c_end; c0 += 8) {
            double v[8];
            int o[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                if (c0 + c < c_end) {
                    const int64_t p = static_cast<int64_t>(c0 + c) * 8 + j;
                    v[c] = kbuf[p];
                    o[c] = offs[p];
                }
            }
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (c0 + c < c_end) row[o[c]] += v[c];
        }
    }
    __syncthreads();
    for (int k = t; k < seg; k += 8 * kLapGatherRows) values[base + k] = s_seg[k];
}

// ---------------------------------------------------------------------------
// Fused Laplace assembly of a structured hexahedral grid (nx x ny x nz
// elements, nodes and elements numbered x fastest, local nodes xi1 fastest):
// no element-matrix buffer.  A block owns the CSR rows of a kLapTX x kLapTY
// column of nodes over `seg` element layers and walks up the layers.  Per
// layer, each thread computes one element matrix of the column's element
// footprint (one element wider on the low x and y sides) into LDS; each node
// thread then
//   * finishes the row of its node in the layer's bottom plane: the partial
//     sums from the layer below (kept in its registers) plus the four
//     elements above the plane, and writes it to its CSR segment once;
//   * starts the row of its node in the top plane with the layer's four
//     elements below that plane.
// A row's entries thus add the element matrices that hold them in ascending
// element order from 0, as the two-pass row gather does, so values are
// bit-identical to it.  Blocks share only recomputed footprint elements
// (x, y halo and the layer below a segment), never output.
constexpr int kLapTX = 15, kLapTY = 7, kLapSeg = 8;
constexpr int kLapFX = kLapTX + 1, kLapFusedThreads = (kLapTX + 1) * (kLapTY + 1);  // 128

__global__ void __launch_bounds__(kLapFusedThreads)
laplace_fused_kernel(int nx, int ny, int nz, int seg, const double *__restrict__ coords, int n_nodes,
                     const int64_t *__restrict__ row_ptr, double *__restrict__ values)
{
    __shared__ double sK[36][kLapFusedThreads];  // element matrices of the layer (upper triangles)
    const int t = threadIdx.x;
    const int tiles_x = (nx + kLapTX) / kLapTX, tiles_y = (ny + kLapTY) / kLapTY;  // nx + 1 nodes
    int b = static_cast<int>(blockIdx.x);
    const int i0 = (b % tiles_x) * kLapTX;
    b /= tiles_x;
    const int j0 = (b % tiles_y) * kLapTY;
    const int k0 = (b / tiles_y) * seg, k1 = min(nz, k0 + seg);
    const int NX = nx + 1;
    const int64_t NXY = static_cast<int64_t>(nx + 1) * (ny + 1);
    // this thread's footprint element and node column
    const int ie = i0 - 1 + t % kLapFX, je = j0 - 1 + t / kLapFX;
    const bool has_elem = ie >= 0 && ie < nx && je >= 0 && je < ny;
    const int ni = i0 + t % kLapTX, nj = j0 + t / kLapTX;
    const bool has_node = t < kLapTX * kLapTY && ni <= nx && nj <= ny;
    const int dxlo = ni > 0 ? -1 : 0, cx = (ni < nx ? 2 : 1) - dxlo;
    const int dylo = nj > 0 ? -1 : 0, cy = (nj < ny ? 2 : 1) - dylo;
    // partial row of the column's node in the next plane (slots dz = -1, 0):
    // the thread keeps its column over the layers, so it stays in registers
    double P[18];
#pragma unroll
    for (int q = 0; q < 18; ++q) P[q] = 0.0;
    // CSR row of node (ni, nj, k): entry (dx, dy, dz) of its stencil is at
    // ((dz - dzlo) cy + dy - dylo) cx + dx - dxlo (columns ascend); the loops
    // unroll, so R stays in registers
    auto write_row = [&](int k, const double(&R)[27]) {
        const int64_t base = row_ptr[ni + NX * nj + NXY * k];
        const int dzlo = k > 0 ? -1 : 0, dzhi = k < nz ? 1 : 0;
#pragma unroll
        for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    if (dz < dzlo || dz > dzhi || dy < dylo || dy >= dylo + cy || dx < dxlo || dx >= dxlo + cx)
                        continue;
                    values[base + ((dz - dzlo) * cy + dy - dylo) * cx + dx - dxlo] =
                        R[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1];
                }
    };
    for (int k = max(k0 - 1, 0); k < k1; ++k) {
        double K[36];
        if (has_elem) {
            double x[3][8];
            const int64_t base = ie + static_cast<int64_t>(NX) * je + NXY * k;
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                const int64_t node = base + (n & 1) + NX * ((n >> 1) & 1) + NXY * (n >> 2);
#pragma unroll
                for (int c = 0; c < 3; ++c) x[c][n] = coords[static_cast<int64_t>(c) * n_nodes + node];
            }
            laplace_element_upper(x, K);
        } else {
#pragma unroll
            for (int q = 0; q < 36; ++q) K[q] = 0.0;
        }
        __syncthreads();  // the previous layer's rows are read
#pragma unroll
        for (int q = 0; q < 36; ++q) sK[q][t] = K[q];
        __syncthreads();
        if (!has_node) continue;
        // the elements around the column in this layer, ascending:
        // (ni-1, nj-1), (ni, nj-1), (ni-1, nj), (ni, nj); local node a of the
        // column's node is (a1, a2) = (ni - ex, nj - ey) in the plane
        if (k >= k0) {  // plane k is this block's: finish and write its rows
            double R[27];
#pragma unroll
            for (int q = 0; q < 18; ++q) R[q] = P[q];
#pragma unroll
            for (int q = 18; q < 27; ++q) R[q] = 0.0;
#pragma unroll
            for (int a2 = 1; a2 >= 0; --a2)
#pragma unroll
                for (int a1 = 1; a1 >= 0; --a1) {
                    const int ex = ni - a1, ey = nj - a2;
                    if (ex < 0 || ex >= nx || ey < 0 || ey >= ny) continue;
                    const int te = ex - i0 + 1 + kLapFX * (ey - j0 + 1);
                    const int ms = a1 + 2 * a2;  // bottom face of the element
#pragma unroll
                    for (int ns = 0; ns < 8; ++ns) {
                        const int b1 = ns & 1, b2 = (ns >> 1) & 1, b3 = ns >> 2;
                        R[(b3 + 1) * 9 + (b2 - a2 + 1) * 3 + b1 - a1 + 1] +=
                            sK[ms <= ns ? lap_upper(ms, ns) : lap_upper(ns, ms)][te];
                    }
                }
            write_row(k, R);
        }
#pragma unroll
        for (int q = 0; q < 18; ++q) P[q] = 0.0;
#pragma unroll
        for (int a2 = 1; a2 >= 0; --a2)
#pragma unroll
            for (int a1 = 1; a1 >= 0; --a1) {
                const int ex = ni - a1, ey = nj - a2;
                if (ex < 0 || ex >= nx || ey < 0 || ey >= ny) continue;
                const int te = ex - i0 + 1 + kLapFX * (ey - j0 + 1);
                const int ms = a1 + 2 * a2 + 4;  // top face of the element
#pragma unroll
                for (int ns = 0; ns < 8; ++ns) {
                    const int b1 = ns & 1, b2 = (ns >> 1) & 1, b3 = ns >> 2;
                    P[b3 * 9 + (b2 - a2 + 1) * 3 + b1 - a1 + 1] +=
                        sK[ms <= ns ? lap_upper(ms, ns) : lap_upper(ns, ms)][te];
                }
            }
    }
    if (k1 == nz && has_node) {  // the top plane has no elements above it
        double R[27];
#pragma unroll
        for (int q = 0; q < 18; ++q) R[q] = P[q];
#pragma unroll
        for (int q = 18; q < 27; ++q) R[q] = 0.0;
        write_row(nz, R);
    }
}

// Structured-grid check of a Laplace set: nodes and elements numbered x
// fastest on an nx x ny x nz grid with xi1-fastest local nodes, and the CSR
// rows the 27-point stencils with ascending columns.  Sets s.lap_grid, or
// leaves it zero (the two-pass assembly then serves the mesh).
void laplace_detect_grid(EquationsSet &s, const std::vector<int32_t> &conn)
{
    s.lap_grid[0] = s.lap_grid[1] = s.lap_grid[2] = 0;
    if (s.n_elements < 1 || conn.size() < 8 || conn[0] != 0 || conn[1] != 1) return;
    const int64_t NX = conn[2];
    if (NX < 2 || conn[4] % NX != 0) return;
    const int64_t NY = conn[4] / NX;
    if (NY < 2 || s.n_nodes % (NX * NY) != 0) return;
    const int64_t NZ = s.n_nodes / (NX * NY);
    const int64_t nx = NX - 1, ny = NY - 1, nz = NZ - 1;
    if (nz < 1 || nx * ny * nz != s.n_elements) return;
    for (int64_t e = 0; e < s.n_elements; ++e) {
        const int64_t ie = e % nx, je = (e / nx) % ny, ke = e / (nx * ny);
        const int64_t base = ie + NX * (je + NY * ke);
        for (int n = 0; n < 8; ++n)
            if (conn[e * 8 + n] != base + (n & 1) + NX * ((n >> 1) & 1) + NX * NY * (n >> 2)) return;
    }
    for (int64_t node = 0; node < s.n_nodes; ++node) {
        const int64_t i = node % NX, j = (node / NX) % NY, k = node / (NX * NY);
        int64_t p = s.h_row_ptr[node];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (i + dx < 0 || i + dx > nx || j + dy < 0 || j + dy > ny || k + dz < 0 || k + dz > nz) continue;
                    if (p >= s.h_row_ptr[node + 1] || s.h_col_idx[p] != node + dx + NX * dy + NX * NY * dz) return;
                    ++p;
                }
        if (p != s.h_row_ptr[node + 1]) return;
    }
    s.lap_grid[0] = static_cast<int>(nx);
    s.lap_grid[1] = static_cast<int>(ny);
    s.lap_grid[2] = static_cast<int>(nz);
}

// Blocks of the fused launch with kLapSeg-layer segments; the fused form
// serves a grid only when they give every CU a block (a block walks its
// layers in sequence, so a small grid runs faster as the two passes: at 64^3,
// 1.6 blocks per CU, the fused form is 4% faster; at 8^3 2.7x slower).
static int64_t laplace_fused_blocks(const EquationsSet &s)
{
    const int nx = s.lap_grid[0], ny = s.lap_grid[1], nz = s.lap_grid[2];
    return static_cast<int64_t>((nx + kLapTX) / kLapTX) * ((ny + kLapTY) / kLapTY) * ((nz + kLapSeg - 1) / kLapSeg);
}

bool laplace_fused_serves(const EquationsSet &s)
{
    if (s.lap_grid[0] <= 0 || s.lap_fused == 0) return false;
    return s.lap_fused == 2 || laplace_fused_blocks(s) >= static_cast<int64_t>(runtime().num_cus);
}

int launch_laplace_fused(EquationsSet &s)
{
    Runtime &rt = runtime();
    const int nx = s.lap_grid[0], ny = s.lap_grid[1], nz = s.lap_grid[2];
    if (nx < 1 || ny < 1 || nz < 1) return set_error(OC_ERR_UNSUPPORTED, "the Laplace set is not a structured grid");
    OC_CHECK(fill<double>(s.rhs.ptr, s.n_rows, 0.0));  // RHS element vector is 0 (:542)
    KernelTimer timer("laplace_fused");
    laplace_fused_kernel<<<static_cast<unsigned>(laplace_fused_blocks(s)), kLapFusedThreads, 0, rt.stream>>>(
        nx, ny, nz, kLapSeg, s.coords.ptr, s.n_nodes, s.row_ptr.ptr, s.values.ptr);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
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
    // the CSR segment of a row-gather block must fit the LDS
    int64_t max_seg = 0;
    for (int r0 = 0; r0 < s.n_rows; r0 += kLapGatherRows)
        max_seg = std::max(max_seg, s.h_row_ptr[std::min(r0 + kLapGatherRows, s.n_rows)] - s.h_row_ptr[r0]);
    if (max_seg > 8192)
        return set_error(OC_ERR_UNSUPPORTED, "Laplace CSR rows too long for the row gather (%lld entries in %d rows)",
                         static_cast<long long>(max_seg), kLapGatherRows);
    s.max_row_len = static_cast<int>(max_seg);
    // DOF -> element rows, ascending element order
    std::vector<int32_t> hconn(static_cast<size_t>(s.n_elements) * 8);
    OC_HIP_CHECK(hipMemcpy(hconn.data(), s.conn.ptr, hconn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> ptr(static_cast<size_t>(s.n_rows) + 1, 0), idx(hconn.size());
    for (size_t k = 0; k < hconn.size(); ++k) ptr[hconn[k] + 1]++;
    for (int d = 0; d < s.n_rows; ++d) ptr[d + 1] += ptr[d];
    std::vector<int32_t> fillp(ptr.begin(), ptr.end() - 1);
    for (size_t k = 0; k < hconn.size(); ++k) idx[fillp[hconn[k]]++] = static_cast<int32_t>(k);
    std::vector<int32_t> gpos(idx.size());
    for (size_t p = 0; p < idx.size(); ++p) gpos[idx[p]] = static_cast<int32_t>(p);
    OC_CHECK(s.gather_ptr.upload(ptr.data(), ptr.size()));
    OC_CHECK(s.gather_idx.upload(gpos.data(), gpos.size()));  // element row -> gather position
    DevBuf<int32_t> didx;
    OC_CHECK(didx.upload(idx.data(), idx.size()));
    DevBuf<uint8_t> permuted;
    OC_CHECK(permuted.alloc(s.slot_offsets8.n));
    const int64_t n = static_cast<int64_t>(idx.size());
    permute_offsets_kernel<<<static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 65535)), 256, 0, rt.stream>>>(
        n, didx.ptr, reinterpret_cast<const uint64_t *>(s.slot_offsets8.ptr), reinterpret_cast<uint64_t *>(permuted.ptr));
    OC_HIP_CHECK(hipGetLastError());
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    s.slot_offsets8 = std::move(permuted);
    OC_CHECK(s.kbuf64.alloc(static_cast<size_t>(s.n_elements) * 64));
    return OC_NO_ERROR;
}

int launch_laplace_assemble(EquationsSet &s)
{
    if (laplace_fused_serves(s)) return launch_laplace_fused(s);
    OC_CHECK(laplace_prepare(s));
    Runtime &rt = runtime();
    OC_CHECK(fill<double>(s.rhs.ptr, s.n_rows, 0.0));  // RHS element vector is 0 (:542)
    {
        KernelTimer timer("laplace_element");
        const unsigned blocks = static_cast<unsigned>((s.n_elements + kLapThreads - 1) / kLapThreads);
        laplace_element_kernel<<<blocks, kLapThreads, 0, rt.stream>>>(s.n_elements, s.conn.ptr, s.coords.ptr,
                                                                     s.n_nodes, s.gather_idx.ptr, s.kbuf64.ptr);
        OC_HIP_CHECK(hipGetLastError());
    }
    KernelTimer timer("laplace_row_gather");
    const size_t lds = static_cast<size_t>(s.max_row_len) * sizeof(double);
    laplace_row_gather_kernel<<<static_cast<unsigned>((s.n_rows + kLapGatherRows - 1) / kLapGatherRows),
                                8 * kLapGatherRows, lds, rt.stream>>>(s.n_rows, s.row_ptr.ptr, s.gather_ptr.ptr,
                                                                       s.kbuf64.ptr, s.slot_offsets8.ptr, s.values.ptr);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

}  // namespace oc
