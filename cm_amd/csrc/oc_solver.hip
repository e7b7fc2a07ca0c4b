// oc_solver.hip -- the GPU solver that checks the assembly path end to end: Dirichlet
// elimination by masking, CSR SpMV, block preconditioners, right-
// preconditioned GMRES and Newton with backtracking line search, plus the
// load-increment rule for fixed-incremented boundary conditions.
//
// The reference solves with PETSc SNES/KSP (solver_routines.f90:18306-19028,
// 10957-11006) on a solver matrix from which Dirichlet rows and columns are
// removed (solver_matrices_routines.f90:1258-1483).  Here the equations CSR is
// kept and the operator is applied as  y_f = J_ff x_f, y_d = x_d  with x_d = 0
// in every Krylov vector, which is the same eliminated system.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <vector>

#include "oc_fe_device.h"
#include "oc_internal.h"

namespace oc {

namespace {

constexpr int kReduceBlocks = 256;

// ---- vector kernels -----------------------------------------------------------
__global__ void __launch_bounds__(256) axpby_kernel(int n, double a, const double *__restrict__ x, double b,
                                                   double *__restrict__ y)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}

__global__ void __launch_bounds__(256) copy_scale_kernel(int n, double a, const double *__restrict__ x,
                                                        double *__restrict__ y)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = a * x[i];
}

__global__ void __launch_bounds__(256) scale_kernel(int n, const double *__restrict__ w, double *__restrict__ y)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = w[i] * y[i];
}

// y = a * w .* x
__global__ void __launch_bounds__(256) weighted_copy_kernel(int n, double a, const double *__restrict__ w,
                                                           const double *__restrict__ x, double *__restrict__ y)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = a * w[i] * x[i];
}

// Solver weights of a Mooney-Rivlin union: mesh m = blockIdx.y owns rows
// c*NQ + [q_off[m], q_off[m+1]) (c < 3) and 3NQ + [l_off[m], l_off[m+1]).
__device__ __forceinline__ int64_t mesh_row(int m, int64_t i, const int *__restrict__ q_off,
                                            const int *__restrict__ l_off, int nq_total, bool *valid)
{
    const int64_t nq = q_off[m + 1] - q_off[m], nl = l_off[m + 1] - l_off[m];
    *valid = i < 3 * nq + nl;
    if (i < 3 * nq) return (i / nq) * nq_total + q_off[m] + i % nq;
    return 3 * static_cast<int64_t>(nq_total) + l_off[m] + (i - 3 * nq);
}

__global__ void __launch_bounds__(256) mesh_weights_kernel(const int *__restrict__ q_off,
                                                          const int *__restrict__ l_off, int nq_total,
                                                          const double *__restrict__ w_mesh,
                                                          const uint8_t *__restrict__ free_mask,
                                                          double *__restrict__ wgt, double *__restrict__ iwgt)
{
    const int m = blockIdx.y;
    const double w = w_mesh[m];
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;; i += gridDim.x * blockDim.x) {
        bool valid;
        const int64_t r = mesh_row(m, i, q_off, l_off, nq_total, &valid);
        if (!valid) break;
        const bool f = free_mask[r] != 0;
        wgt[r] = f ? w : 0.0;
        iwgt[r] = f ? 1.0 / w : 0.0;
    }
}

// partial[m * gridDim.x + blockIdx.x] = sum of x_r y_r over the free rows of mesh m
__global__ void __launch_bounds__(256) mesh_dots_kernel(const int *__restrict__ q_off, const int *__restrict__ l_off,
                                                       int nq_total, const uint8_t *__restrict__ free_mask,
                                                       const double *__restrict__ x, const double *__restrict__ y,
                                                       double *__restrict__ partial)
{
    __shared__ double sh[256];
    const int m = blockIdx.y;
    double acc = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;; i += gridDim.x * blockDim.x) {
        bool valid;
        const int64_t r = mesh_row(m, i, q_off, l_off, nq_total, &valid);
        if (!valid) break;
        if (free_mask[r]) acc += x[r] * y[r];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[m * gridDim.x + blockIdx.x] = sh[0];
}

// x = x0 + lambda[m] dx on every row of mesh m (the per-mesh line search)
__global__ void __launch_bounds__(256) mesh_axpy_kernel(const int *__restrict__ q_off, const int *__restrict__ l_off,
                                                       int nq_total, const double *__restrict__ lambda,
                                                       const double *__restrict__ dx, const double *__restrict__ x0,
                                                       double *__restrict__ x)
{
    const int m = blockIdx.y;
    const double a = lambda[m];
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;; i += gridDim.x * blockDim.x) {
        bool valid;
        const int64_t r = mesh_row(m, i, q_off, l_off, nq_total, &valid);
        if (!valid) break;
        x[r] = a * dx[r] + x0[r];
    }
}

// partial[j * gridDim.x + blockIdx.x] = sum over this block's slice of V_j . w
// for j = 0..k-1 (blockIdx.y = j).  Deterministic: fixed grid, fixed tree.
__global__ void __launch_bounds__(256) multi_dot_kernel(int n, const double *__restrict__ V, int64_t ldv,
                                                       const double *__restrict__ w, double *__restrict__ partial,
                                                       const double *__restrict__ gate)
{
    if (gate && *gate == 0.0) return;  // a pass the Arnoldi process did not ask for
    __shared__ double s[256];
    const double *v = V + static_cast<int64_t>(blockIdx.y) * ldv;
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += v[i] * w[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = s[0];
}

// w += sign * sum_j c_j V_j  (j < k <= kCoefChunk); coefficients by value
constexpr int kCoefChunk = 64;
struct Coefficients {
    double c[kCoefChunk];
};

__global__ void __launch_bounds__(256) combine_kernel(int n, int k, const double *__restrict__ V, int64_t ldv,
                                                     Coefficients coef, double sign, double *__restrict__ w)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < k; ++j) acc += coef.c[j] * V[static_cast<int64_t>(j) * ldv + i];
        w[i] += sign * acc;
    }
}

// ---- Arnoldi process on the device ---------------------------------------------
// Classical Gram-Schmidt with one reorthogonalisation pass when the projection
// removes more than 3/4 of |w|^2 (CGS2), and Givens rotations of the
// Hessenberg column, all without a host round trip.  Layout of the state for
// restart m: R (m x m, row-major, column j written by step j), h and h2 (m+2),
// cs, sn (m), g (m+1), then flag, hn and 1/hn.
struct Arnoldi {
    double *R, *h, *h2, *cs, *sn, *g, *flag, *hn, *inv;
    static size_t size(int m) { return static_cast<size_t>(m) * m + 2 * (m + 2) + 2 * m + (m + 1) + 3; }
    static Arnoldi at(double *base, int m)
    {
        Arnoldi a;
        a.R = base;
        a.h = a.R + static_cast<size_t>(m) * m;
        a.h2 = a.h + m + 2;
        a.cs = a.h2 + m + 2;
        a.sn = a.cs + m;
        a.g = a.sn + m;
        a.flag = a.g + m + 1;
        a.hn = a.flag + 1;
        a.inv = a.hn + 1;
        return a;
    }
};

// sum of the kReduceBlocks partials of dot product i, in block order
__device__ __forceinline__ double partial_sum(const double *__restrict__ partial, int i)
{
    double acc = 0.0;
    for (int b = 0; b < kReduceBlocks; ++b) acc += partial[static_cast<size_t>(i) * kReduceBlocks + b];
    return acc;
}

// g = (beta, 0, ..., 0)
__global__ void __launch_bounds__(256) arnoldi_start_kernel(int m, double beta, Arnoldi a)
{
    for (int i = threadIdx.x; i <= m; i += blockDim.x) a.g[i] = i == 0 ? beta : 0.0;
}

// step j, first pass: h_i = V_i . w (i <= j), |w|^2 = h_{j+1}; asks for the
// second pass or takes |w - V h| by Pythagoras
__global__ void __launch_bounds__(256) arnoldi_project_kernel(int j, const double *__restrict__ partial, Arnoldi a)
{
    for (int i = threadIdx.x; i < j + 2; i += blockDim.x) a.h[i] = partial_sum(partial, i);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double wn2 = a.h[j + 1];
        double proj2 = 0.0;
        for (int i = 0; i <= j; ++i) proj2 += a.h[i] * a.h[i];
        const bool again = wn2 - proj2 < 0.25 * wn2;
        *a.flag = again ? 1.0 : 0.0;
        *a.hn = again ? 0.0 : sqrt(wn2 - proj2);
    }
}

// step j, second pass (only with the flag set): h2_i = V_i . w, h_i += h2_i
__global__ void __launch_bounds__(256) arnoldi_correct_kernel(int j, const double *__restrict__ partial, Arnoldi a)
{
    if (*a.flag == 0.0) return;
    for (int i = threadIdx.x; i <= j; i += blockDim.x) {
        const double c = partial_sum(partial, i);
        a.h2[i] = c;
        a.h[i] += c;
    }
}

// w += sign * sum_{j<k} coef_j V_j, coefficients in device memory
__global__ void __launch_bounds__(256) combine_device_kernel(int n, int k, const double *__restrict__ V, int64_t ldv,
                                                            const double *__restrict__ coef, double sign,
                                                            double *__restrict__ w, const double *__restrict__ gate)
{
    if (gate && *gate == 0.0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < k; ++j) acc += coef[j] * V[static_cast<int64_t>(j) * ldv + i];
        w[i] += sign * acc;
    }
}

// step j, last: the norm of the orthogonalised vector (from the partials of
// w . w after a second pass), the rotations, R's column j, g, and the
// residual estimate |g_{j+1}| and hn into the host-visible slots 2j, 2j+1
__global__ void __launch_bounds__(64) arnoldi_finalize_kernel(int j, int m, const double *__restrict__ partial,
                                                            Arnoldi a, double *__restrict__ host)
{
    if (threadIdx.x != 0) return;
    double hn = *a.hn;
    if (*a.flag != 0.0) hn = sqrt(fmax(partial_sum(partial, 0), 0.0));
    *a.inv = hn > 0.0 ? 1.0 / hn : 0.0;
    double hi = a.h[0];
    for (int i = 0; i < j; ++i) {
        const double hi1 = a.h[i + 1];
        a.R[static_cast<size_t>(i) * m + j] = a.cs[i] * hi + a.sn[i] * hi1;
        hi = -a.sn[i] * hi + a.cs[i] * hi1;
    }
    // hi is the rotated h_j; h_{j+1} = hn
    const double den = hypot(hi, hn);
    const double c = den > 0.0 ? hi / den : 1.0, sgn = den > 0.0 ? hn / den : 0.0;
    a.cs[j] = c;
    a.sn[j] = sgn;
    a.R[static_cast<size_t>(j) * m + j] = den;
    const double gj = a.g[j];
    a.g[j + 1] = -sgn * gj;
    a.g[j] = c * gj;
    host[2 * j] = fabs(a.g[j + 1]);
    host[2 * j + 1] = hn;
}

// v *= 1/hn unless hn is zero
__global__ void __launch_bounds__(256) arnoldi_scale_kernel(int n, const double *__restrict__ inv, double *__restrict__ v)
{
    const double a = *inv;
    if (a == 0.0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = a * v[i];
}

// ---- CSR SpMV of the weighted Dirichlet operator: y_r = w_r sum_k v_k x_c
// (w_r = 0 on fixed rows, whose sum is skipped), or y = A x when w is NULL.
// `lanes` lanes per row, shuffle reduction.
template <typename VT, int kLanes>
__global__ void __launch_bounds__(256) spmv_kernel(int n, const int64_t *__restrict__ row_ptr,
                                                  const int *__restrict__ col_idx, const VT *__restrict__ vals,
                                                  const double *__restrict__ w, const double *__restrict__ x,
                                                  double *__restrict__ y)
{
    const int lane = threadIdx.x % kLanes;
    const int64_t xb = xcd_block(blockIdx.x, gridDim.x);
    const int row = static_cast<int>((xb * blockDim.x + threadIdx.x) / kLanes);
    if (row >= n) return;
    double acc = 0.0;
    const double wr = w == nullptr ? 1.0 : w[row];
    if (wr != 0.0) acc = csr_row_dot<VT, kLanes>(row_ptr[row], row_ptr[row + 1], lane, col_idx, vals, x);
#pragma unroll
    for (int o = kLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kLanes);
    if (lane == 0) y[row] = wr * acc;
}

// ---- Mooney-Rivlin rows by nodes ------------------------------------------------------
// Node i's displacement rows c*nq + i (c < 3): [row_ptr, pstart) holds three
// blocks of deg columns, block c' being c'*nq + j over the same sorted
// neighbour nodes j, and [pstart, row_ptr + 1) the same pressure columns in
// all three rows.  bad = 1 when any node breaks that layout.
__global__ void __launch_bounds__(256) node_pattern_check_kernel(int nq, const int64_t *__restrict__ row_ptr,
                                                                const int64_t *__restrict__ pstart,
                                                                const int *__restrict__ col_idx, int *__restrict__ bad)
{
    const int i = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 5);
    const int lane = threadIdx.x & 31;
    if (i >= nq) return;
    const int64_t d0 = pstart[i] - row_ptr[i], dp = row_ptr[i + 1] - pstart[i];
    bool ok = d0 % 3 == 0;
    for (int c = 1; c < 3; ++c) {
        const int64_t r = static_cast<int64_t>(c) * nq + i;
        ok = ok && pstart[r] - row_ptr[r] == d0 && row_ptr[r + 1] - pstart[r] == dp;
    }
    if (ok) {
        const int64_t deg = d0 / 3;
        for (int64_t p = lane; p < deg; p += 32) {
            const int j = col_idx[row_ptr[i] + p];
            ok = ok && j >= 0 && j < nq;
            for (int c = 0; c < 3; ++c)
                for (int cc = 0; cc < 3; ++cc)
                    ok = ok && col_idx[row_ptr[static_cast<int64_t>(c) * nq + i] + cc * deg + p] ==
                                   static_cast<int64_t>(cc) * nq + j;
        }
        for (int64_t p = lane; p < dp; p += 32)
            for (int c = 1; c < 3; ++c)
                ok = ok && col_idx[pstart[static_cast<int64_t>(c) * nq + i] + p] == col_idx[pstart[i] + p];
    }
    if (!ok) *bad = 1;
}

// y = W (A x) (or A x when w is NULL) of a Mooney-Rivlin set with node rows:
// groups g < nq take node g's three displacement rows (one read of the
// neighbour and pressure column lists), groups nq + l take pressure row l
template <typename VT, int kLanes, int kU>
__global__ void __launch_bounds__(256) spmv_node_kernel(int nq, int nl, const int64_t *__restrict__ row_ptr,
                                                       const int64_t *__restrict__ pstart,
                                                       const int *__restrict__ col_idx, const VT *__restrict__ vals,
                                                       const double *__restrict__ w, const double *__restrict__ x,
                                                       double *__restrict__ y)
{
    const int lane = threadIdx.x % kLanes;
    const int64_t xb = xcd_block(blockIdx.x, gridDim.x);
    const int g = static_cast<int>((xb * blockDim.x + threadIdx.x) / kLanes);
    const int64_t nq64 = nq;
    if (g >= nq + nl) return;  // whole groups
    if (g >= nq) {
        const int64_t row = 3 * nq64 + (g - nq);
        const double wr = w == nullptr ? 1.0 : w[row];
        double acc = 0.0;
        if (wr != 0.0) acc = csr_row_dot<VT, kLanes>(row_ptr[row], row_ptr[row + 1], lane, col_idx, vals, x);
#pragma unroll
        for (int o = kLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kLanes);
        if (lane == 0) y[row] = wr * acc;
        return;
    }
    const int i = g;
    const int64_t sr[3] = {row_ptr[i], row_ptr[nq64 + i], row_ptr[2 * nq64 + i]};
    const int64_t p0 = pstart[i], p1 = pstart[nq64 + i], p2 = pstart[2 * nq64 + i];
    const int deg = static_cast<int>((p0 - sr[0]) / 3), dp = static_cast<int>(row_ptr[i + 1] - p0);
    const double w0 = w == nullptr ? 1.0 : w[i], w1 = w == nullptr ? 1.0 : w[nq64 + i],
                 w2 = w == nullptr ? 1.0 : w[2 * nq64 + i];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (w0 != 0.0 || w1 != 0.0 || w2 != 0.0) {
        double a[3] = {0.0, 0.0, 0.0};
        node_rows_uu<VT, kLanes, kU>(sr, deg, lane, nq64, col_idx, vals, x, a);
        a0 = a[0];
        a1 = a[1];
        a2 = a[2];
        for (int p = lane; p < dp; p += kLanes) {
            const double xp = x[col_idx[p0 + p]];
            a0 += static_cast<double>(vals[p0 + p]) * xp;
            a1 += static_cast<double>(vals[p1 + p]) * xp;
            a2 += static_cast<double>(vals[p2 + p]) * xp;
        }
    }
#pragma unroll
    for (int o = kLanes / 2; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o, kLanes);
        a1 += __shfl_xor(a1, o, kLanes);
        a2 += __shfl_xor(a2, o, kLanes);
    }
    if (lane < 3) {
        const int64_t row = lane * nq64 + i;
        y[row] = lane == 0 ? w0 * a0 : (lane == 1 ? w1 * a1 : w2 * a2);
    }
}

// Pressure Jacobi z_p = inv_p .* r_p (written by the groups i < nl) and
// t = r_u - B^T z_p on free displacement rows, 0 on fixed ones, by nodes; z_p
// is formed again where B^T reads it, so one launch does both (nl <= nq)
template <typename VT, int kLanes>
__global__ void __launch_bounds__(256) schur_rhs_node_kernel(int nq, int nl, const int64_t *__restrict__ row_ptr,
                                                            const int64_t *__restrict__ pstart,
                                                            const int *__restrict__ col_idx,
                                                            const VT *__restrict__ vals,
                                                            const uint8_t *__restrict__ free_mask,
                                                            const double *__restrict__ inv_p,
                                                            const double *__restrict__ r, double *__restrict__ z,
                                                            double *__restrict__ t)
{
    const int lane = threadIdx.x % kLanes;
    const int64_t xb = xcd_block(blockIdx.x, gridDim.x);
    const int i = static_cast<int>((xb * blockDim.x + threadIdx.x) / kLanes);
    if (i >= nq) return;  // whole groups
    const int64_t nq64 = nq;
    if (i < nl && lane == 0) z[3 * nq64 + i] = inv_p[i] * r[3 * nq64 + i];
    const int64_t p0 = pstart[i], p1 = pstart[nq64 + i], p2 = pstart[2 * nq64 + i];
    const int dp = static_cast<int>(row_ptr[i + 1] - p0);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int p = lane; p < dp; p += kLanes) {
        const int c = col_idx[p0 + p];
        const double zp = inv_p[c - 3 * nq64] * r[c];
        a0 += static_cast<double>(vals[p0 + p]) * zp;
        a1 += static_cast<double>(vals[p1 + p]) * zp;
        a2 += static_cast<double>(vals[p2 + p]) * zp;
    }
#pragma unroll
    for (int o = kLanes / 2; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o, kLanes);
        a1 += __shfl_xor(a1, o, kLanes);
        a2 += __shfl_xor(a2, o, kLanes);
    }
    if (lane < 3) {
        const int64_t row = lane * nq64 + i;
        const double a = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
        t[row] = free_mask[row] ? r[row] - a : 0.0;
    }
}

// ---- preconditioners ----------------------------------------------------------------
template <typename VT>
__device__ __forceinline__ double csr_entry_t(const int64_t *__restrict__ row_ptr, const int *__restrict__ col_idx,
                                              const VT *__restrict__ vals, int r, int c)
{
    int64_t lo = row_ptr[r], hi = row_ptr[r + 1] - 1;
    while (lo <= hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int v = col_idx[mid];
        if (v == c) return static_cast<double>(vals[mid]);
        if (v < c) lo = mid + 1;
        else hi = mid - 1;
    }
    return 0.0;
}

// point Jacobi: inv[d] = 1/a_dd (1 on zero diagonals and fixed rows, as PCJACOBI)
template <typename VT>
__global__ void __launch_bounds__(256) jacobi_setup_kernel(int n, const int64_t *__restrict__ row_ptr,
                                                          const int *__restrict__ col_idx,
                                                          const VT *__restrict__ vals,
                                                          const uint8_t *__restrict__ free_mask,
                                                          double *__restrict__ inv)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const double a = csr_entry_t(row_ptr, col_idx, vals, d, d);
    inv[d] = (!free_mask[d] || a == 0.0) ? 1.0 : 1.0 / a;
}

// nodal 3x3 blocks of the displacement DOFs (d, d+nq, d+2nq) with fixed
// rows/columns replaced by the identity, inverted; layout inv[node*9 + a*3+b].
template <typename VT>
__global__ void __launch_bounds__(256) block_setup_kernel(int nq, const int64_t *__restrict__ row_ptr,
                                                         const int *__restrict__ col_idx,
                                                         const VT *__restrict__ vals,
                                                         const uint8_t *__restrict__ free_mask,
                                                         double *__restrict__ inv, double *__restrict__ diag_u)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nq) return;
    double A[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const int ra = a * nq + node, cb = b * nq + node;
            double v;
            if (free_mask[ra] && free_mask[cb]) v = csr_entry_t(row_ptr, col_idx, vals, ra, cb);
            else v = (a == b) ? 1.0 : 0.0;
            A[a * 3 + b] = v;
        }
    for (int a = 0; a < 3; ++a) diag_u[a * nq + node] = A[a * 3 + a];
    const double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8], c02 = A[3] * A[7] - A[4] * A[6];
    const double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
    double *o = inv + static_cast<int64_t>(node) * 9;
    if (fabs(det) < 1e-300) {
        for (int i = 0; i < 9; ++i) o[i] = (i % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    const double id = 1.0 / det;
    o[0] = c00 * id;
    o[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    o[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    o[3] = c01 * id;
    o[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    o[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    o[6] = c02 * id;
    o[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    o[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

// pressure rows: S_p = sum_j B_pj^2 / K_jj over free displacement columns
// (diagonal estimate of the Schur complement B diag(K)^-1 B^T); inv = -1/S.
// The analytic Mooney-Rivlin subtype uses the pressure-mass estimate
// (launch_schur_mass) instead: at 32^3 it needs 0.6x the Krylov iterations of
// plain Newton and 0.85x of the learned start's; the forward-difference
// subtype, which has no Gauss records, keeps this one.
template <typename VT>
__global__ void __launch_bounds__(256) schur_setup_kernel(int nq, int nl, const int64_t *__restrict__ row_ptr,
                                                         const int *__restrict__ col_idx,
                                                         const VT *__restrict__ vals,
                                                         const uint8_t *__restrict__ free_mask,
                                                         const double *__restrict__ diag_u, double *__restrict__ inv)
{
    // one wavefront per pressure row: lanes over the row's displacement
    // columns, a fixed butterfly for the sum
    const int p = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= nl) return;  // whole wavefronts
    const int r = 3 * nq + p;
    double s = 0.0;
    for (int64_t k = row_ptr[r] + lane; k < row_ptr[r + 1]; k += 64) {
        const int c = col_idx[k];
        if (c >= 3 * nq || !free_mask[c]) continue;
        const double b = static_cast<double>(vals[k]);
        const double kd = diag_u[c];
        if (kd != 0.0) s += b * b / kd;
    }
#pragma unroll
    for (int sh = 32; sh > 0; sh >>= 1) s += __shfl_xor(s, sh, 64);
    if (lane == 0) inv[p] = s > 0.0 ? -1.0 / s : -1.0;
}

__global__ void __launch_bounds__(256) apply_jacobi_kernel(int n, const double *__restrict__ inv,
                                                          const double *__restrict__ r, double *__restrict__ z)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) z[i] = inv[i] * r[i];
}

// Block upper-triangular application, displacement part, one thread per node:
// t = r_u - B^T z_p over the pressure columns of the node's three rows
// (they start at pstart[row], columns being sorted), then z_u = Kblock^-1 t.
// Fixed rows contribute 0.
template <typename VT>
__global__ void __launch_bounds__(256) triangular_u_kernel(int nq, const int64_t *__restrict__ row_ptr,
                                                          const int64_t *__restrict__ pstart,
                                                          const int *__restrict__ col_idx,
                                                          const VT *__restrict__ vals,
                                                          const uint8_t *__restrict__ free_mask,
                                                          const double *__restrict__ inv_u,
                                                          const double *__restrict__ r, double *__restrict__ z)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nq) return;
    double t[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int row = a * nq + node;
        double acc = 0.0;
        if (free_mask[row]) {
            acc = r[row];
            for (int64_t k = pstart[row]; k < row_ptr[row + 1]; ++k)
                acc -= static_cast<double>(vals[k]) * z[col_idx[k]];
        }
        t[a] = acc;
    }
    const double *o = inv_u + static_cast<int64_t>(node) * 9;
    z[node] = o[0] * t[0] + o[1] * t[1] + o[2] * t[2];
    z[nq + node] = o[3] * t[0] + o[4] * t[1] + o[5] * t[2];
    z[2 * nq + node] = o[6] * t[0] + o[7] * t[1] + o[8] * t[2];
}

// t = r_u - B^T z_p on the free displacement rows (pressure columns start at
// pstart[row]); 8 lanes per row, shuffle reduction in a fixed order
template <typename VT>
__global__ void __launch_bounds__(256) schur_rhs_kernel(int nu, const int64_t *__restrict__ row_ptr,
                                                       const int64_t *__restrict__ pstart,
                                                       const int *__restrict__ col_idx, const VT *__restrict__ vals,
                                                       const uint8_t *__restrict__ free_mask,
                                                       const double *__restrict__ r, const double *__restrict__ z,
                                                       double *__restrict__ t)
{
    constexpr int kLanes = 8;
    const int lane = threadIdx.x % kLanes;
    const int row = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kLanes);
    if (row >= nu) return;
    const bool f = free_mask[row] != 0;
    double acc = 0.0;
    if (f)
        for (int64_t k = pstart[row] + lane; k < row_ptr[row + 1]; k += kLanes)
            acc += static_cast<double>(vals[k]) * z[col_idx[k]];
#pragma unroll
    for (int o = kLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kLanes);
    if (lane == 0) t[row] = f ? r[row] - acc : 0.0;
}

__global__ void __launch_bounds__(256) pressure_start_kernel(int n_u_rows, int first_p_col,
                                                            const int64_t *__restrict__ row_ptr,
                                                            const int *__restrict__ col_idx,
                                                            int64_t *__restrict__ pstart)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_u_rows) return;
    int64_t lo = row_ptr[r], hi = row_ptr[r + 1];
    while (lo < hi) {  // first k with col_idx[k] >= first_p_col
        const int64_t mid = (lo + hi) >> 1;
        if (col_idx[mid] < first_p_col) lo = mid + 1;
        else hi = mid;
    }
    pstart[r] = lo;
}

__global__ void __launch_bounds__(256) apply_block_kernel(int nq, int nl, const double *__restrict__ inv_u,
                                                         const double *__restrict__ inv_p,
                                                         const double *__restrict__ r, double *__restrict__ z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) {
        const double *o = inv_u + static_cast<int64_t>(i) * 9;
        const double r0 = r[i], r1 = r[nq + i], r2 = r[2 * nq + i];
        z[i] = o[0] * r0 + o[1] * r1 + o[2] * r2;
        z[nq + i] = o[3] * r0 + o[4] * r1 + o[5] * r2;
        z[2 * nq + i] = o[6] * r0 + o[7] * r1 + o[8] * r2;
    }
    if (i < nl) z[3 * nq + i] = inv_p[i] * r[3 * nq + i];
}

// ---- boundary conditions ----------------------------------------------------------
__global__ void __launch_bounds__(256) bc_increment_kernel(int n, const int *__restrict__ idx,
                                                          const double *__restrict__ full, double *__restrict__ dofs,
                                                          int iteration, int max_iterations)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = idx[i];
    if (iteration >= max_iterations) dofs[d] = full[i];
    else dofs[d] = dofs[d] + (full[i] - dofs[d]) / static_cast<double>(max_iterations - iteration + 1);
}

__global__ void __launch_bounds__(256) unit_weights_kernel(int n, const uint8_t *__restrict__ free_mask,
                                                          double *__restrict__ wgt, double *__restrict__ iwgt)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = free_mask[i] ? 1.0 : 0.0;
        wgt[i] = v;
        iwgt[i] = v;
    }
}

unsigned grid_for(int64_t n, int per_block = 256)
{
    const int64_t b = (n + per_block - 1) / per_block;
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(b, runtime().num_cus * 8)));
}

// ---- host-side solver ------------------------------------------------------------------
// GMRES works on the weighted system W J dx = -W F, W = diag(wgt) constant on
// each mesh of a batch (1/||F_ref|| of that mesh), with the right
// preconditioner W^-1 P^-1: per mesh the operator is J_m P_m^-1, unscaled,
// while the Krylov norms weigh every mesh's residual relative to its own
// reference.  A single mesh (or the Laplace set, with unit weights) is the
// one-block case.
struct Krylov {
    EquationsSet &s;
    int n;
    int m;                   // restart
    double *V;               // (m+1) x n
    double *r, *z, *w, *u;   // work vectors
    double *inv_u, *inv_p, *diag_u, *inv_pt;  // preconditioner
    const int64_t *pstart = nullptr;
    int precond = OC_PRECONDITIONER_BLOCK_DIAGONAL;
    bool fp32 = false;
    Arnoldi ar{};
    std::vector<double> hpart;

    // y = W (A x)
    int spmv(const double *x, double *y)
    {
        KernelTimer timer("spmv");
        const bool lap = s.cls == OC_EQUATIONS_SET_LAPLACE;
        // defaults: 8 lanes per Laplace row; per Mooney-Rivlin node 32 lanes
        // (FP32) or 16 lanes with two neighbours in flight each (FP64)
        const int lanes = s.tuned_spmv_lanes > 0 ? s.tuned_spmv_lanes : (lap ? 8 : (fp32 ? 32 : 16));
        if (lap) return launch_spmv<double>(s.values.ptr, lanes, s.wgt.ptr, x, y);
        if (fp32) return launch_spmv<float>(s.values_sp.ptr, lanes, s.wgt.ptr, x, y);
        return launch_spmv<double>(s.values.ptr, lanes, s.wgt.ptr, x, y);
    }
    template <typename VT>
    int launch_spmv(const VT *vals, int lanes, const double *wv, const double *x, double *y, int unroll = -1)
    {
        if (unroll < 0) unroll = s.tuned_spmv_unroll > 0 ? s.tuned_spmv_unroll : (fp32 ? 1 : 2);
        Runtime &rt = runtime();
        if (s.cls == OC_EQUATIONS_SET_MOONEY_RIVLIN && s.node_rows == 1 && lanes >= 16) {
            const int nq = s.n_nodes, nl = s.n_nodes_l;
            const unsigned gn = static_cast<unsigned>((static_cast<int64_t>(nq + nl) * lanes + 255) / 256);
            const bool u2 = unroll == 2;
            // the Krylov operator gathers the iterate in FP64: FP32 node
            // quadruples, as the bf16 smoother reads, made the product about
            // 10% faster but perturbed it enough to cost Newton iterations
#define OC_SPMV_NODE(L, U) \
    spmv_node_kernel<VT, L, U><<<gn, 256, 0, rt.stream>>>(nq, nl, s.row_ptr.ptr, s.pstart.ptr, s.col_idx.ptr, vals, wv, x, y)
            if (lanes == 16) {
                if (u2) OC_SPMV_NODE(16, 2);
                else OC_SPMV_NODE(16, 1);
            } else if (lanes == 64) {
                if (u2) OC_SPMV_NODE(64, 2);
                else OC_SPMV_NODE(64, 1);
            } else {
                if (u2) OC_SPMV_NODE(32, 2);
                else OC_SPMV_NODE(32, 1);
            }
#undef OC_SPMV_NODE
            OC_HIP_CHECK(hipGetLastError());
            return OC_NO_ERROR;
        }
        const unsigned g = static_cast<unsigned>((static_cast<int64_t>(n) * lanes + 255) / 256);
        switch (lanes) {
        case 8: spmv_kernel<VT, 8><<<g, 256, 0, rt.stream>>>(n, s.row_ptr.ptr, s.col_idx.ptr, vals, wv, x, y); break;
        case 16: spmv_kernel<VT, 16><<<g, 256, 0, rt.stream>>>(n, s.row_ptr.ptr, s.col_idx.ptr, vals, wv, x, y); break;
        case 64: spmv_kernel<VT, 64><<<g, 256, 0, rt.stream>>>(n, s.row_ptr.ptr, s.col_idx.ptr, vals, wv, x, y); break;
        default: spmv_kernel<VT, 32><<<g, 256, 0, rt.stream>>>(n, s.row_ptr.ptr, s.col_idx.ptr, vals, wv, x, y); break;
        }
        OC_HIP_CHECK(hipGetLastError());
        return OC_NO_ERROR;
    }

    int setup_preconditioner()
    {
        Runtime &rt = runtime();
        KernelTimer timer("preconditioner_setup");
        if (s.cls == OC_EQUATIONS_SET_LAPLACE || precond == OC_PRECONDITIONER_JACOBI) {
            if (fp32)
                jacobi_setup_kernel<float><<<grid_rows(n), 256, 0, rt.stream>>>(n, s.row_ptr.ptr, s.col_idx.ptr,
                                                                                s.values_sp.ptr, s.free_mask.ptr,
                                                                                inv_u);
            else
                jacobi_setup_kernel<double><<<grid_rows(n), 256, 0, rt.stream>>>(n, s.row_ptr.ptr, s.col_idx.ptr,
                                                                                 s.values.ptr, s.free_mask.ptr,
                                                                                 inv_u);
            OC_HIP_CHECK(hipGetLastError());
            return OC_NO_ERROR;
        }
        const int nq = s.n_nodes, nl = s.n_nodes_l;
        if (precond == OC_PRECONDITIONER_MULTIGRID) OC_CHECK(mg_setup(s, fp32 ? OC_FP32 : OC_FP64));
        const bool mass_schur = s.subtype == OC_MOONEY_RIVLIN_SUBTYPE && s.gp_records.ptr;
        if (fp32) {
            block_setup_kernel<float><<<grid_rows(nq), 256, 0, rt.stream>>>(nq, s.row_ptr.ptr, s.col_idx.ptr,
                                                                            s.values_sp.ptr, s.free_mask.ptr, inv_u,
                                                                            diag_u);
            if (mass_schur)
                OC_CHECK(launch_schur_mass(s, inv_p));
            else
                schur_setup_kernel<float><<<grid_rows(static_cast<int64_t>(nl) * 64), 256, 0, rt.stream>>>(nq, nl, s.row_ptr.ptr, s.col_idx.ptr,
                                                                                s.values_sp.ptr, s.free_mask.ptr, diag_u,
                                                                                inv_p);
        } else {
            block_setup_kernel<double><<<grid_rows(nq), 256, 0, rt.stream>>>(nq, s.row_ptr.ptr, s.col_idx.ptr,
                                                                             s.values.ptr, s.free_mask.ptr, inv_u,
                                                                             diag_u);
            if (mass_schur)
                OC_CHECK(launch_schur_mass(s, inv_p));
            else
                schur_setup_kernel<double><<<grid_rows(static_cast<int64_t>(nl) * 64), 256, 0, rt.stream>>>(nq, nl, s.row_ptr.ptr, s.col_idx.ptr,
                                                                                 s.values.ptr, s.free_mask.ptr, diag_u,
                                                                                 inv_p);
        }
        OC_HIP_CHECK(hipGetLastError());
        return OC_NO_ERROR;
    }
    static unsigned grid_rows(int64_t count) { return static_cast<unsigned>(std::max<int64_t>(1, (count + 255) / 256)); }

    // z = W^-1 P^-1 r, with P block upper-triangular (z_p = S^-1 r_p, z_u =
    // K^-1 (r_u - B^T z_p)), block diagonal or point Jacobi
    int apply_preconditioner(const double *rin, double *zout)
    {
        Runtime &rt = runtime();
        if (s.cls == OC_EQUATIONS_SET_LAPLACE || precond == OC_PRECONDITIONER_JACOBI) {
            apply_jacobi_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, inv_u, rin, zout);
        } else {
            const int nq = s.n_nodes, nl = s.n_nodes_l;
            if (precond == OC_PRECONDITIONER_MULTIGRID) {
                const bool nodes = s.node_rows == 1 && nl <= nq;
                if (!nodes)
                    apply_jacobi_kernel<<<grid_for(nl), 256, 0, rt.stream>>>(
                        nl, inv_p, rin + 3 * static_cast<int64_t>(nq), zout + 3 * static_cast<int64_t>(nq));
                const int nu = 3 * nq;
                double *t = s.mg->t0.ptr;
                const unsigned gs = grid_rows(static_cast<int64_t>(nu) * 8);
                const unsigned gsn = grid_rows(static_cast<int64_t>(nq) * 16);
                if (nodes && fp32)
                    schur_rhs_node_kernel<float, 16><<<gsn, 256, 0, rt.stream>>>(
                        nq, nl, s.row_ptr.ptr, pstart, s.col_idx.ptr, s.values_sp.ptr, s.free_mask.ptr, inv_p, rin,
                        zout, t);
                else if (nodes)
                    schur_rhs_node_kernel<double, 16><<<gsn, 256, 0, rt.stream>>>(
                        nq, nl, s.row_ptr.ptr, pstart, s.col_idx.ptr, s.values.ptr, s.free_mask.ptr, inv_p, rin,
                        zout, t);
                else if (fp32)
                    schur_rhs_kernel<float><<<gs, 256, 0, rt.stream>>>(nu, s.row_ptr.ptr, pstart, s.col_idx.ptr,
                                                                        s.values_sp.ptr, s.free_mask.ptr, rin, zout, t);
                else
                    schur_rhs_kernel<double><<<gs, 256, 0, rt.stream>>>(nu, s.row_ptr.ptr, pstart, s.col_idx.ptr,
                                                                         s.values.ptr, s.free_mask.ptr, rin, zout, t);
                OC_HIP_CHECK(hipGetLastError());
                OC_CHECK(mg_apply(s, fp32 ? OC_FP32 : OC_FP64, t, zout));
            } else if (precond == OC_PRECONDITIONER_BLOCK_TRIANGULAR) {
                apply_jacobi_kernel<<<grid_for(nl), 256, 0, rt.stream>>>(nl, inv_p, rin + 3 * static_cast<int64_t>(nq),
                                                                         zout + 3 * static_cast<int64_t>(nq));
                if (fp32)
                    triangular_u_kernel<float><<<grid_rows(nq), 256, 0, rt.stream>>>(
                        nq, s.row_ptr.ptr, pstart, s.col_idx.ptr, s.values_sp.ptr, s.free_mask.ptr, inv_u, rin, zout);
                else
                    triangular_u_kernel<double><<<grid_rows(nq), 256, 0, rt.stream>>>(
                        nq, s.row_ptr.ptr, pstart, s.col_idx.ptr, s.values.ptr, s.free_mask.ptr, inv_u, rin, zout);
            } else {
                apply_block_kernel<<<grid_rows(std::max(nq, nl)), 256, 0, rt.stream>>>(nq, nl, inv_u, inv_p, rin,
                                                                                       zout);
            }
        }
        OC_HIP_CHECK(hipGetLastError());
        scale_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, s.iwgt.ptr, zout);
        OC_HIP_CHECK(hipGetLastError());
        return OC_NO_ERROR;
    }

    // h[j] = V_j . w for j < k (one device->host copy)
    int multi_dot(const double *Vb, int k, const double *wv, double *h)
    {
        Runtime &rt = runtime();
        dim3 grid(kReduceBlocks, k);
        multi_dot_kernel<<<grid, 256, 0, rt.stream>>>(n, Vb, n, wv, s.partials.ptr, nullptr);
        OC_HIP_CHECK(hipGetLastError());
        hpart.resize(static_cast<size_t>(k) * kReduceBlocks);
        OC_HIP_CHECK(hipMemcpyAsync(hpart.data(), s.partials.ptr, sizeof(double) * hpart.size(),
                                    hipMemcpyDeviceToHost, rt.stream));
        OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
        for (int j = 0; j < k; ++j) {
            double acc = 0.0;
            for (int b = 0; b < kReduceBlocks; ++b) acc += hpart[static_cast<size_t>(j) * kReduceBlocks + b];
            h[j] = acc;
        }
        return OC_NO_ERROR;
    }

    int norm(const double *x, double *out)
    {
        double h = 0.0;
        OC_CHECK(multi_dot(x, 1, x, &h));
        *out = std::sqrt(std::max(h, 0.0));
        return OC_NO_ERROR;
    }

    int combine(const double *Vb, int k, const std::vector<double> &coef, double sign, double *out)
    {
        Runtime &rt = runtime();
        for (int j0 = 0; j0 < k; j0 += kCoefChunk) {
            const int kk = std::min(kCoefChunk, k - j0);
            Coefficients c{};
            for (int j = 0; j < kk; ++j) c.c[j] = coef[j0 + j];
            combine_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, kk, Vb + static_cast<int64_t>(j0) * n, n, c, sign,
                                                               out);
            OC_HIP_CHECK(hipGetLastError());
        }
        return OC_NO_ERROR;
    }

    // One Arnoldi step j on the device: V_{j+1} = orthonormalised W A P^-1 W^-1 V_j,
    // R's column j and g rotated, |g_{j+1}| and hn into the host slots 2j, 2j+1.
    int arnoldi_step(int j)
    {
        Runtime &rt = runtime();
        double *vj = V + static_cast<int64_t>(j) * n, *vn = V + static_cast<int64_t>(j + 1) * n;
        OC_CHECK(apply_preconditioner(vj, z));
        OC_CHECK(spmv(z, vn));
        // classical Gram-Schmidt: one fused reduction gives V_0..V_j . w and
        // w . w (V_{j+1} is w itself); the norm after projection follows by
        // Pythagoras unless cancellation asks for a second pass and an
        // explicit norm, whose kernels then run (they return at once otherwise)
        multi_dot_kernel<<<dim3(kReduceBlocks, j + 2), 256, 0, rt.stream>>>(n, V, n, vn, s.partials.ptr, nullptr);
        arnoldi_project_kernel<<<1, 256, 0, rt.stream>>>(j, s.partials.ptr, ar);
        combine_device_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, j + 1, V, n, ar.h, -1.0, vn, nullptr);
        multi_dot_kernel<<<dim3(kReduceBlocks, j + 1), 256, 0, rt.stream>>>(n, V, n, vn, s.partials.ptr, ar.flag);
        arnoldi_correct_kernel<<<1, 256, 0, rt.stream>>>(j, s.partials.ptr, ar);
        combine_device_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, j + 1, V, n, ar.h2, -1.0, vn, ar.flag);
        multi_dot_kernel<<<dim3(kReduceBlocks, 1), 256, 0, rt.stream>>>(n, vn, n, vn, s.partials.ptr, ar.flag);
        arnoldi_finalize_kernel<<<1, 64, 0, rt.stream>>>(j, m, s.partials.ptr, ar, s.arnoldi_host_dev);
        arnoldi_scale_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, ar.inv, vn);
        OC_HIP_CHECK(hipGetLastError());
        return OC_NO_ERROR;
    }

    // Solves W A x = b (x starts at 0) to ||b - W A x|| <= tol_abs; b is zero
    // on fixed DOFs.  Right preconditioning: x = W^-1 P^-1 V y.  The host
    // stays one Arnoldi step ahead of the device: it queues step j, then
    // reads step j-1's residual estimate; a converged j-1 leaves step j's
    // work unused (R's column j and g_j, g_{j+1} are not read).
    int gmres(const double *b, double *x, double tol_abs, int max_iter, int *iters, double *res_out)
    {
        Runtime &rt = runtime();
        if (2 * m > kArnoldiHostSlots) return set_error(OC_ERR_INVALID_ARGUMENT, "krylov restart %d too large", m);
        OC_HIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * n, rt.stream));
        int total = 0;
        double beta = 0.0;
        OC_CHECK(norm(b, &beta));
        double res = beta;
        std::vector<double> R(static_cast<size_t>(m) * m), g(m + 1), y(m);
        const volatile double *slots = s.arnoldi_host;
        bool first = true;
        while (total < max_iter && res > tol_abs) {
            // r = b - W A x
            if (first) {
                OC_HIP_CHECK(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, rt.stream));
                first = false;
            } else {
                OC_CHECK(spmv(x, w));
                OC_HIP_CHECK(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, rt.stream));
                axpby_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, -1.0, w, 1.0, r);
                OC_CHECK(norm(r, &beta));
                res = beta;
                if (res <= tol_abs) break;
            }
            copy_scale_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, 1.0 / beta, r, V);
            arnoldi_start_kernel<<<1, 256, 0, rt.stream>>>(m, beta, ar);
            const int steps = std::min(m, max_iter - total);
            int cols = 0;
            for (int j = 0; j < steps && cols == 0; ++j) {
                OC_CHECK(arnoldi_step(j));
                OC_HIP_CHECK(hipEventRecord(s.arnoldi_events[j & 1], rt.stream));
                // read step j-1; the cycle's last step is read as soon as it is queued
                for (int k = j - 1; k <= j && cols == 0; ++k) {
                    if (k < 0 || (k == j && j != steps - 1)) continue;
                    OC_HIP_CHECK(hipEventSynchronize(s.arnoldi_events[k & 1]));
                    const double rk = slots[2 * k], hk = slots[2 * k + 1];
                    res = rk;
                    if (rk <= tol_abs || hk == 0.0 || k == steps - 1) cols = k + 1;
                }
            }
            total += cols;
            // y = R^-1 g over the first cols columns; x += W^-1 P^-1 (V y)
            OC_HIP_CHECK(hipMemcpyAsync(R.data(), ar.R, sizeof(double) * static_cast<size_t>(cols) * m,
                                        hipMemcpyDeviceToHost, rt.stream));
            OC_HIP_CHECK(hipMemcpyAsync(g.data(), ar.g, sizeof(double) * cols, hipMemcpyDeviceToHost, rt.stream));
            OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
            for (int i = cols - 1; i >= 0; --i) {
                double t = g[i];
                for (int k = i + 1; k < cols; ++k) t -= R[static_cast<size_t>(i) * m + k] * y[k];
                y[i] = t / R[static_cast<size_t>(i) * m + i];
            }
            OC_CHECK(fill<double>(w, n, 0.0));
            std::vector<double> yc(y.begin(), y.begin() + cols);
            OC_CHECK(combine(V, cols, yc, 1.0, w));
            OC_CHECK(apply_preconditioner(w, z));
            axpby_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, 1.0, z, 1.0, x);
        }
        *iters = total;
        *res_out = res;
        return OC_NO_ERROR;
    }
};

}  // namespace

// Work vectors: x_prev, dx, rhs, r, z, w, u, tmp -> 8 * n; preconditioner:
// 9*nq + nl + n + n (inv_u, inv_p, diag_u, triangular scratch)
static int solver_prepare(EquationsSet &s, int restart, Krylov &K)
{
    const int n = s.n_rows;
    if (restart < 1 || restart + 1 > 1024) return set_error(OC_ERR_INVALID_ARGUMENT, "krylov_restart must be in [1, 1023]");
    if (K.precond < OC_PRECONDITIONER_JACOBI || K.precond > OC_PRECONDITIONER_MULTIGRID)
        return set_error(OC_ERR_INVALID_ARGUMENT, "unknown preconditioner %d", K.precond);
    if (K.precond == OC_PRECONDITIONER_MULTIGRID && (s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN || !s.mg))
        return set_error(OC_ERR_INVALID_ARGUMENT,
                         "the multigrid preconditioner needs oc_multigrid_structure_set on a Mooney-Rivlin set");
    if (!s.free_mask.ptr) {
        OC_CHECK(s.free_mask.alloc(n));
        OC_HIP_CHECK(hipMemset(s.free_mask.ptr, 1, n));
    }
    if (!s.wgt.ptr) OC_CHECK(s.wgt.alloc(n));
    if (!s.iwgt.ptr) OC_CHECK(s.iwgt.alloc(n));
    if (s.krylov_capacity < restart) {
        OC_CHECK(s.krylov.alloc(static_cast<size_t>(restart + 1) * n));
        s.krylov_capacity = restart;
    }
    if (s.arnoldi.n < Arnoldi::size(restart)) OC_CHECK(s.arnoldi.alloc(Arnoldi::size(restart)));
    if (!s.arnoldi_host) {
        for (auto &e : s.arnoldi_events)
            if (!e) OC_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        OC_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&s.arnoldi_host), sizeof(double) * kArnoldiHostSlots,
                                   hipHostMallocMapped | hipHostMallocCoherent));
        const hipError_t e = hipHostGetDevicePointer(reinterpret_cast<void **>(&s.arnoldi_host_dev), s.arnoldi_host, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(s.arnoldi_host);
            s.arnoldi_host = nullptr;
            return set_error(OC_ERR_HIP, "hipHostGetDevicePointer failed: %s", hipGetErrorString(e));
        }
    }
    if (!s.work.ptr) OC_CHECK(s.work.alloc(static_cast<size_t>(8) * n));
    if (!s.precond.ptr) OC_CHECK(s.precond.alloc(static_cast<size_t>(9) * s.n_nodes + 3 * static_cast<size_t>(n)));
    const size_t npart = std::max<size_t>(static_cast<size_t>(kReduceBlocks) * (1024 + 1),
                                          static_cast<size_t>(s.n_meshes) * 256);
    if (!s.partials.ptr || s.partials.n < npart) OC_CHECK(s.partials.alloc(npart));
    K.n = n;
    K.m = restart;
    K.V = s.krylov.ptr;
    K.ar = Arnoldi::at(s.arnoldi.ptr, restart);
    K.r = s.work.ptr + 3 * static_cast<int64_t>(n);
    K.z = s.work.ptr + 4 * static_cast<int64_t>(n);
    K.w = s.work.ptr + 5 * static_cast<int64_t>(n);
    K.u = s.work.ptr + 6 * static_cast<int64_t>(n);
    double *pc = s.precond.ptr;
    if (s.cls == OC_EQUATIONS_SET_LAPLACE) {
        K.inv_u = pc;
        K.inv_p = nullptr;
        K.diag_u = nullptr;
        K.inv_pt = nullptr;
    } else {
        K.inv_u = pc;
        K.inv_p = pc + static_cast<int64_t>(9) * s.n_nodes;
        K.diag_u = K.inv_p + n;
        K.inv_pt = K.diag_u + n;
        if (K.precond == OC_PRECONDITIONER_JACOBI) K.inv_u = K.inv_pt;  // n-sized point Jacobi
        OC_CHECK(prepare_node_rows(s));
        K.pstart = s.pstart.ptr;
    }
    return OC_NO_ERROR;
}

// First pressure column of every displacement row, and whether the set's rows
// have the node layout the node-wise kernels read (once per set)
int prepare_node_rows(EquationsSet &s)
{
    if (s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN) return OC_NO_ERROR;
    Runtime &rt = runtime();
    if (!s.pstart.ptr) {
        const int nu = 3 * s.n_nodes;
        OC_CHECK(s.pstart.alloc(nu));
        pressure_start_kernel<<<static_cast<unsigned>((nu + 255) / 256), 256, 0, rt.stream>>>(
            nu, nu, s.row_ptr.ptr, s.col_idx.ptr, s.pstart.ptr);
        OC_HIP_CHECK(hipGetLastError());
    }
    if (s.node_rows < 0) {
        DevBuf<int32_t> bad_d;
        OC_CHECK(bad_d.alloc(1));
        OC_HIP_CHECK(hipMemsetAsync(bad_d.ptr, 0, sizeof(int32_t), rt.stream));
        node_pattern_check_kernel<<<static_cast<unsigned>((static_cast<int64_t>(s.n_nodes) * 32 + 255) / 256), 256,
                                    0, rt.stream>>>(s.n_nodes, s.row_ptr.ptr, s.pstart.ptr, s.col_idx.ptr,
                                                    bad_d.ptr);
        OC_HIP_CHECK(hipGetLastError());
        int32_t bad = 1;
        OC_HIP_CHECK(hipMemcpyAsync(&bad, bad_d.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, rt.stream));
        OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
        s.node_rows = bad ? 0 : 1;
    }
    return OC_NO_ERROR;
}

int launch_spmv_lanes(EquationsSet &s, bool fp32, int lanes, const double *x, double *y, int unroll)
{
    Krylov K{s};
    K.n = s.n_rows;
    if (fp32) return K.launch_spmv<float>(s.values_sp.ptr, lanes, nullptr, x, y, unroll);
    return K.launch_spmv<double>(s.values.ptr, lanes, nullptr, x, y, unroll);
}

// Per-mesh dot products x . y over free rows (Mooney-Rivlin sets), one
// device->host copy.
static int mesh_dots(EquationsSet &s, const double *x, const double *y, std::vector<double> &out)
{
    Runtime &rt = runtime();
    const int M = s.n_meshes;
    const int chunks = M == 1 ? 256 : 32;
    dim3 grid(chunks, M);
    mesh_dots_kernel<<<grid, 256, 0, rt.stream>>>(s.q_off.ptr, s.l_off.ptr, s.n_nodes, s.free_mask.ptr, x, y,
                                                  s.partials.ptr);
    OC_HIP_CHECK(hipGetLastError());
    std::vector<double> h(static_cast<size_t>(M) * chunks);
    OC_HIP_CHECK(hipMemcpyAsync(h.data(), s.partials.ptr, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                                rt.stream));
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    out.assign(M, 0.0);
    for (int m = 0; m < M; ++m) {
        double acc = 0.0;
        for (int c = 0; c < chunks; ++c) acc += h[static_cast<size_t>(m) * chunks + c];
        out[m] = acc;
    }
    return OC_NO_ERROR;
}

// Per-mesh norms of x over free rows.
static int mesh_norms(EquationsSet &s, const double *x, std::vector<double> &out)
{
    OC_CHECK(mesh_dots(s, x, x, out));
    for (double &v : out) v = std::sqrt(v);
    return OC_NO_ERROR;
}

static int residual_mesh_norms(EquationsSet &s, std::vector<double> &out)
{
    OC_CHECK(launch_mr_residual(s));
    return mesh_norms(s, s.residual.ptr, out);
}

// Jacobian assembly (and, with K, the preconditioner setup) of one Newton
// step, recorded once as a graph of kernel nodes and replayed on later steps.
static int step_assembly(EquationsSet &s, Krylov *K, int precision)
{
    Runtime &rt = runtime();
    if (s.subtype == OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE) {
        // forward-difference Jacobians take their step size from the host:
        // launched eagerly, not recorded
        OC_CHECK(launch_mr_jacobian(s, precision));
        return K ? K->setup_preconditioner() : OC_NO_ERROR;
    }
    // the preconditioner kernels read the free-DOF mask, so a graph belongs to
    // one set of boundary conditions
    const long long key[5] = {precision, K ? K->precond : -1, s.tuned_element_blocks_per_cu, s.tuned_gather_chunk,
                              static_cast<long long>(s.bc_version)};
    if (!s.step_graph || !std::equal(key, key + 5, s.step_graph_key)) {
        if (s.step_graph) {
            (void)hipGraphExecDestroy(s.step_graph);
            s.step_graph = nullptr;
        }
        hipGraph_t graph = nullptr;
        OC_HIP_CHECK(hipStreamBeginCapture(rt.stream, hipStreamCaptureModeThreadLocal));
        set_capturing(true);
        int rc = launch_mr_jacobian(s, precision);
        if (rc == OC_NO_ERROR && K) rc = K->setup_preconditioner();
        set_capturing(false);
        const hipError_t ec = hipStreamEndCapture(rt.stream, &graph);
        OC_CHECK(rc);
        if (ec != hipSuccess) return set_error(OC_ERR_HIP, "stream capture failed: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&s.step_graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            s.step_graph = nullptr;
            return set_error(OC_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
        }
        std::copy(key, key + 5, s.step_graph_key);
    }
    KernelTimer timer("step_graph");
    OC_HIP_CHECK(hipGraphLaunch(s.step_graph, rt.stream));
    return OC_NO_ERROR;
}

static double weighted_norm(const std::vector<double> &nrm, const std::vector<double> &w)
{
    double acc = 0.0;
    for (size_t m = 0; m < nrm.size(); ++m) acc += (w[m] * nrm[m]) * (w[m] * nrm[m]);
    return std::sqrt(acc);
}

static double plain_norm(const std::vector<double> &nrm)
{
    double acc = 0.0;
    for (double v : nrm) acc += v * v;
    return std::sqrt(acc);
}

// The backtracking line search of a batch, one per mesh: mesh m steps to
// x_prev + lam_m dx and backtracks (quadratic, then cubic interpolation,
// safeguarded to [0.1, 0.5] lam) until f_m = 1/2 (w_m ||F_m||)^2 meets
// f_m(0) + alpha lam_m slope_m, with slope_m = (W F)_m . (W J dx)_m; a mesh
// that has met it keeps its step while the others go on.  A mesh already
// within its tolerance stays where it is (lam_m = 0), as its own solve would
// have stopped.  Every trial is one residual launch over all meshes.  On return lam holds the steps taken, trial
// the per-mesh residual norms there, *evals the trials run and *decreased
// whether every mesh met its condition within 20 trials (the last trial is
// kept otherwise, as for one mesh).
static int batch_line_search(EquationsSet &s, const oc_newton_options *opt, const std::vector<double> &w,
                             const std::vector<double> &target, const std::vector<double> &nrm, const double *rhs,
                             const double *wjdx, const double *dx,
                             const double *x_prev, std::vector<double> &lam, std::vector<double> &trial, int *evals,
                             bool *decreased)
{
    Runtime &rt = runtime();
    const int M = s.n_meshes;
    std::vector<double> fjdx;
    OC_CHECK(mesh_dots(s, rhs, wjdx, fjdx));  // rhs = -W F
    std::vector<double> slope(M), f0(M), lam_prev(M, 0.0), f_prev(M, 0.0);
    std::vector<int> tries(M, 0);
    std::vector<char> done(M, 0);
    for (int m = 0; m < M; ++m) {
        slope[m] = -fjdx[m];
        if (slope[m] > 0.0) slope[m] = -slope[m];
        f0[m] = 0.5 * (w[m] * nrm[m]) * (w[m] * nrm[m]);
        if (slope[m] == 0.0) slope[m] = -1.0;
    }
    lam.assign(M, 1.0);
    for (int m = 0; m < M; ++m)
        if (nrm[m] <= target[m]) {
            lam[m] = 0.0;
            done[m] = 1;
        }
    DevBuf<double> lam_dev;
    OC_CHECK(lam_dev.alloc(M));
    *decreased = false;
    int ls = 0;
    for (; ls < 20; ++ls) {
        OC_HIP_CHECK(hipMemcpyAsync(lam_dev.ptr, lam.data(), sizeof(double) * M, hipMemcpyHostToDevice, rt.stream));
        mesh_axpy_kernel<<<dim3(32, M), 256, 0, rt.stream>>>(s.q_off.ptr, s.l_off.ptr, s.n_nodes, lam_dev.ptr, dx,
                                                              x_prev, s.dofs.ptr);
        OC_HIP_CHECK(hipGetLastError());
        OC_CHECK(residual_mesh_norms(s, trial));
        if (!opt->line_search) {
            *decreased = true;
            ++ls;
            break;
        }
        bool all = true;
        for (int m = 0; m < M; ++m) {
            if (done[m]) continue;
            const double f1 = 0.5 * (w[m] * trial[m]) * (w[m] * trial[m]);
            if (std::isfinite(f1) && f1 <= f0[m] + opt->line_search_alpha * lam[m] * slope[m]) {
                done[m] = 1;
                continue;
            }
            all = false;
            const double l = lam[m], sl = slope[m];
            double lnew;
            if (tries[m] == 0 || !std::isfinite(f1)) {
                lnew = std::isfinite(f1) ? -sl / (2.0 * (f1 - f0[m] - sl)) : 0.1;
            } else {
                const double lp = lam_prev[m];
                const double r1 = f1 - f0[m] - l * sl, r2 = f_prev[m] - f0[m] - lp * sl;
                const double a = (r1 / (l * l) - r2 / (lp * lp)) / (l - lp);
                const double b = (-lp * r1 / (l * l) + l * r2 / (lp * lp)) / (l - lp);
                if (a == 0.0) lnew = -sl / (2.0 * b);
                else {
                    const double disc = b * b - 3.0 * a * sl;
                    lnew = disc < 0.0 ? 0.5 * l : (-b + std::sqrt(disc)) / (3.0 * a);
                }
            }
            lam_prev[m] = l;
            f_prev[m] = f1;
            ++tries[m];
            lam[m] = std::min(0.5 * l, std::max(0.1 * l, lnew));
        }
        if (all) {
            *decreased = true;
            ++ls;
            break;
        }
    }
    if (!*decreased) {  // the meshes still backtracking keep their last trial
        for (int m = 0; m < M; ++m)
            if (!done[m]) lam[m] = lam_prev[m];
    }
    *evals = ls;
    return OC_NO_ERROR;
}

// Inexact Newton with backtracking line search on every mesh of the set at
// once; mesh m has converged when ||F_m|| <= max(atol, rtol * ref_m).
static int newton_core(EquationsSet &s, const oc_newton_options *opt, const double *refs_in,
                       std::vector<double> &refs_out, std::vector<double> &finals_out, oc_newton_stats *stats)
{
    if (opt->krylov_restart < 1 || opt->maximum_iterations < 0)
        return set_error(OC_ERR_INVALID_ARGUMENT, "invalid Newton options");
    if (opt->jacobian_precision != OC_FP64 && opt->jacobian_precision != OC_FP32)
        return set_error(OC_ERR_INVALID_ARGUMENT, "jacobian_precision must be OC_FP64 or OC_FP32");
    const auto t0 = std::chrono::steady_clock::now();
    *stats = oc_newton_stats{};
    OC_CHECK(mr_prepare(s));
    Krylov K{s};
    K.precond = opt->preconditioner;
    K.fp32 = opt->jacobian_precision == OC_FP32;
    OC_CHECK(solver_prepare(s, opt->krylov_restart, K));
    // an FP32 Jacobian's multigrid smoother reads a packed bf16 copy of K_uu
    // (filled in the preconditioner setup) and gathers the iterate as FP32
    // node quadruples: allocated here, outside the step graph, and a graph
    // recorded before the copy existed is recorded again
    if (K.fp32 && K.precond == OC_PRECONDITIONER_MULTIGRID && s.node_rows == 1 && !s.uu_bf.ptr) {
        if (!s.values_sp.ptr) OC_CHECK(s.values_sp.alloc(s.nnz));
        OC_CHECK(prepare_uu_bf16(s));
        if (s.step_graph) {
            (void)hipGraphExecDestroy(s.step_graph);
            s.step_graph = nullptr;
        }
    }
    OC_CHECK(mr_jacobian_buffers(s, opt->jacobian_precision));
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    Runtime &rt = runtime();
    const int n = s.n_rows, M = s.n_meshes;
    double *x_prev = s.work.ptr, *dx = s.work.ptr + n, *rhs = s.work.ptr + 2 * static_cast<int64_t>(n);

    std::vector<double> nrm, ref(M), w(M), target(M);
    OC_CHECK(residual_mesh_norms(s, nrm));
    stats->residual_evaluations++;
    double min_wt = HUGE_VAL;
    for (int m = 0; m < M; ++m) {
        ref[m] = refs_in ? refs_in[m] : nrm[m];
        w[m] = ref[m] > 0.0 ? 1.0 / ref[m] : 1.0;
        target[m] = std::max(opt->absolute_tolerance, opt->relative_tolerance * ref[m]);
        min_wt = std::min(min_wt, w[m] * target[m]);
    }
    {
        DevBuf<double> wd;
        OC_CHECK(wd.upload(w.data(), w.size()));
        dim3 grid(64, M);
        mesh_weights_kernel<<<grid, 256, 0, rt.stream>>>(s.q_off.ptr, s.l_off.ptr, s.n_nodes, wd.ptr,
                                                         s.free_mask.ptr, s.wgt.ptr, s.iwgt.ptr);
        OC_HIP_CHECK(hipGetLastError());
        OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    }
    auto converged = [&](const std::vector<double> &v) {
        for (int m = 0; m < M; ++m)
            if (!(v[m] <= target[m])) return false;
        return true;
    };
    double fnorm = weighted_norm(nrm, w);
    const double ref_global = weighted_norm(ref, w);
    stats->initial_residual_norm = plain_norm(nrm);
    double fnorm_prev = fnorm, eta_prev = 0.5;
    int it = 0;
    // PETSc's default convergence test (SNESConvergedDefault, the test the
    // reference selects): a non-finite residual diverges; the residual test;
    // the function-evaluation limit; after a step, ||step|| < stol ||x||.
    // Deviation: the residual test (per mesh, abs or rel against ref_m) may
    // also hold before the first step, so that a start already inside the
    // tolerance (a warm start judged against the plain start) costs no step.
    int reason = 0;
    auto convergence_test = [&](double snorm, double xnorm) {
        bool finite = true;
        for (double v : nrm) finite = finite && std::isfinite(v);
        if (!finite) {
            reason = OC_DIVERGED_FNORM_NAN;
        } else if (converged(nrm)) {
            bool abs_all = true;
            for (int m = 0; m < M; ++m) abs_all = abs_all && nrm[m] <= opt->absolute_tolerance;
            reason = abs_all ? OC_CONVERGED_FNORM_ABS : OC_CONVERGED_FNORM_RELATIVE;
        } else if (opt->maximum_function_evaluations >= 0 &&
                   stats->residual_evaluations >= opt->maximum_function_evaluations) {
            reason = OC_DIVERGED_FUNCTION_COUNT;
        } else if (it > 0 && opt->solution_tolerance > 0.0 && snorm < opt->solution_tolerance * xnorm) {
            reason = OC_CONVERGED_SNORM_RELATIVE;
        }
    };
    convergence_test(0.0, 0.0);
    while (reason == 0 && it < opt->maximum_iterations) {
        OC_CHECK(step_assembly(s, &K, opt->jacobian_precision));
        stats->jacobian_evaluations++;
        // forcing term: fixed, or Eisenstat-Walker (choice 2, gamma 0.9, alpha
        // (1+sqrt5)/2) capped by ||F|| / ||F_ref||, the Dembo-Eisenstat-Steihaug
        // rule that keeps Newton quadratic once the iterate is close: a start
        // already near the solution is solved accurately from its first step
        double eta = opt->krylov_relative_tolerance;
        if (eta <= 0.0) {
            if (it == 0) eta = 0.5;
            else {
                eta = 0.9 * std::pow(fnorm / fnorm_prev, 1.618);
                const double safe = 0.9 * std::pow(eta_prev, 1.618);
                if (safe > 0.1) eta = std::max(eta, safe);
                eta = std::min(eta, 0.5);
            }
            if (ref_global > 0.0) eta = std::min(eta, fnorm / ref_global);
            eta = std::max(eta, 0.5 * min_wt / fnorm);  // do not over-solve the last step
            eta = std::max(eta, 1e-14);
            eta_prev = eta;
        }
        // rhs = -W F (zero on fixed DOFs)
        weighted_copy_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, -1.0, s.wgt.ptr, s.residual.ptr, rhs);
        OC_HIP_CHECK(hipGetLastError());
        int kit = 0;
        double kres = 0.0;
        // the linear solve also stops once its residual is below half the
        // Newton target, which the step then meets: an absolute floor as the
        // KSP's ABSOLUTE_TOLERANCE (solver_routines.f90:10983-10984), tied to
        // the Newton test so that the last step of a load step is not solved
        // further than that test asks
        const double ktol = std::max(eta * fnorm, 0.5 * min_wt);
        OC_CHECK(K.gmres(rhs, dx, ktol, opt->krylov_maximum_iterations, &kit, &kres));
        stats->krylov_iterations += kit;
        // a linear solve that ended above its tolerance (KSP_DIVERGED_ITS):
        // PETSc's SNES stops with SNES_DIVERGED_LINEAR_SOLVE once
        // maxLinearSolveFailures of them have happened, before the line search
        if (!(kres <= ktol)) {
            ++stats->linear_solve_failures;
            if (stats->linear_solve_failures >= std::max(1, opt->maximum_linear_solve_failures)) {
                reason = OC_DIVERGED_LINEAR_SOLVE;
                if (opt->output_type >= OC_SOLVER_PROGRESS_OUTPUT)
                    std::fprintf(stderr, "  newton %3d  ||WF|| %.6e  gmres %d its ended at %.3e > %.3e: diverged linear solve\n",
                                 it, fnorm, kit, kres, ktol);
                break;
            }
        }
        // line search on f = 1/2 ||W F||^2; initial slope (W F) . (W J dx), as
        // PETSc's bt line search takes it (made negative if it is not)
        OC_CHECK(K.spmv(dx, K.u));
        OC_HIP_CHECK(hipMemcpyAsync(x_prev, s.dofs.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice, rt.stream));
        double lambda = 1.0, lambda_taken = 1.0, fnew = 0.0;
        const double f0 = 0.5 * fnorm * fnorm;
        std::vector<double> trial, lam_mesh;  // lam_mesh: the step length of each mesh of a batch
        bool decreased = !opt->line_search;
        int ls = 0;
        double slope = 0.0;
        if (M > 1) {
            // a batch backtracks every mesh on its own f_m = 1/2 ||W F_m||^2,
            // as independent solves would, in one residual launch per trial
            OC_CHECK(batch_line_search(s, opt, w, target, nrm, rhs, K.u, dx, x_prev, lam_mesh, trial, &ls, &decreased));
            stats->residual_evaluations += ls;
            fnew = weighted_norm(trial, w);
            lambda_taken = 0.0;  // the shortest step of the meshes that moved (monitor output)
            for (double l : lam_mesh)
                if (l > 0.0) lambda_taken = lambda_taken > 0.0 ? std::min(lambda_taken, l) : l;
            if (!decreased) ls = 20;
            else --ls;
        } else {
        double fjdx = 0.0;
        OC_CHECK(K.multi_dot(rhs, 1, K.u, &fjdx));  // rhs = -W F
        slope = -fjdx;
        if (slope > 0.0) slope = -slope;
        if (slope == 0.0) slope = -1.0;
        double lambda_prev = 0.0, f_prev_trial = 0.0;
        for (; ls < 20; ++ls) {
            OC_HIP_CHECK(hipMemcpyAsync(s.dofs.ptr, x_prev, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                        rt.stream));
            axpby_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, lambda, dx, 1.0, s.dofs.ptr);
            lambda_taken = lambda;
            OC_CHECK(residual_mesh_norms(s, trial));
            fnew = weighted_norm(trial, w);
            stats->residual_evaluations++;
            const double f1 = 0.5 * fnew * fnew;
            if (!opt->line_search) break;
            if (std::isfinite(f1) && f1 <= f0 + opt->line_search_alpha * lambda * slope) {
                decreased = true;
                break;
            }
            // quadratic (first backtrack) or cubic interpolation, safeguarded to [0.1, 0.5] lambda
            double lnew;
            if (ls == 0 || !std::isfinite(f1)) {
                lnew = std::isfinite(f1) ? -slope / (2.0 * (f1 - f0 - slope)) : 0.1;
            } else {
                const double r1 = f1 - f0 - lambda * slope, r2 = f_prev_trial - f0 - lambda_prev * slope;
                const double a = (r1 / (lambda * lambda) - r2 / (lambda_prev * lambda_prev)) / (lambda - lambda_prev);
                const double b = (-lambda_prev * r1 / (lambda * lambda) + lambda * r2 / (lambda_prev * lambda_prev)) /
                                 (lambda - lambda_prev);
                if (a == 0.0) lnew = -slope / (2.0 * b);
                else {
                    const double disc = b * b - 3.0 * a * slope;
                    lnew = disc < 0.0 ? 0.5 * lambda : (-b + std::sqrt(disc)) / (3.0 * a);
                }
            }
            lambda_prev = lambda;
            f_prev_trial = f1;
            lambda = std::min(0.5 * lambda, std::max(0.1 * lambda, lnew));
        }
        }
        nrm = trial;
        if (opt->output_type >= OC_SOLVER_PROGRESS_OUTPUT)
            std::fprintf(stderr, "  newton %3d  ||WF|| %.6e -> %.6e  eta %.2e  gmres %4d its %.3e/%.3e  slope/f0 %.3f  lambda %.3g (%d trials)\n",
                         it, fnorm, fnew, eta, kit, kres, ktol, slope / (2.0 * f0), lambda_taken, ls + (decreased ? 1 : 0));
        fnorm_prev = fnorm;
        fnorm = fnew;
        ++it;
        if (!decreased) {  // backtracking ended without sufficient decrease; the last trial is kept
            ++stats->step_failures;
            if (stats->step_failures >= std::max(1, opt->maximum_step_failures)) {
                reason = OC_DIVERGED_LINE_SEARCH;
                break;
            }
        }
        double snorm = 0.0, xnorm = 0.0;
        if (opt->solution_tolerance > 0.0) {  // the step taken and the new iterate, free DOFs
            std::vector<double> part;
            OC_CHECK(mesh_norms(s, dx, part));
            if (M > 1) {
                for (int m = 0; m < M; ++m) part[m] *= lam_mesh[m];
                snorm = plain_norm(part);
            } else {
                snorm = lambda_taken * plain_norm(part);
            }
            OC_CHECK(mesh_norms(s, s.dofs.ptr, part));
            xnorm = plain_norm(part);
        }
        convergence_test(snorm, xnorm);
    }
    if (reason == 0) reason = OC_DIVERGED_MAX_IT;
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    stats->newton_iterations = it;
    stats->final_residual_norm = plain_norm(nrm);
    stats->reason = reason;
    stats->converged = reason > 0 ? 1 : 0;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    refs_out = ref;
    finals_out = nrm;
    if (opt->output_type >= OC_SOLVER_PROGRESS_OUTPUT)
        std::fprintf(stderr, "  newton: reason %d after %d iterations, %d Krylov, ||F|| %.6e\n", reason, it,
                     stats->krylov_iterations, stats->final_residual_norm);
    if (!stats->converged) {
        int worst = 0;
        double worst_ratio = -1.0;
        for (int m = 0; m < M; ++m) {
            const double ratio = std::isfinite(nrm[m]) ? nrm[m] / std::max(target[m], 1e-300) : HUGE_VAL;
            if (ratio > worst_ratio) {
                worst_ratio = ratio;
                worst = m;
            }
        }
        return set_error(OC_ERR_NOT_CONVERGED,
                         "Newton did not converge after %d iterations (reason %d): mesh %d ||F|| = %.3e > %.3e", it,
                         reason, worst, nrm[worst], target[worst]);
    }
    return OC_NO_ERROR;
}

// dst = src on the DOFs of the meshes with redo[m] set (quadratic nodes of
// mesh m are [q_off[m], q_off[m+1]) in each of the three components, its
// pressure nodes [l_off[m], l_off[m+1]))
__global__ void __launch_bounds__(256) restore_meshes_kernel(int n_rows, int nq, int n_meshes,
                                                            const int *__restrict__ q_off,
                                                            const int *__restrict__ l_off,
                                                            const uint8_t *__restrict__ redo,
                                                            const double *__restrict__ src, double *__restrict__ dst)
{
    for (int d = blockIdx.x * blockDim.x + threadIdx.x; d < n_rows; d += gridDim.x * blockDim.x) {
        const bool disp = d < 3 * nq;
        const int node = disp ? d % nq : d - 3 * nq;
        const int *off = disp ? q_off : l_off;
        int lo = 0, hi = n_meshes - 1;  // mesh with off[m] <= node < off[m + 1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= node) lo = mid;
            else hi = mid - 1;
        }
        if (redo[lo]) dst[d] = src[d];
    }
}

// Newton solve with the plain-start fallback: when the attempt fails and the
// load step was warm-started, the meshes that did not converge restart from
// the plain start with an FP64 Jacobian, judged against the same targets.
static int newton_solve_set(EquationsSet &s, const oc_newton_options *opt, const double *refs_in, double *finals_out,
                            oc_newton_stats *stats)
{
    const bool can_fall_back = opt->fallback != 0 && s.plain_start_valid;
    std::vector<double> refs, finals;
    int rc = newton_core(s, opt, refs_in, refs, finals, stats);
    if (rc == OC_ERR_NOT_CONVERGED && can_fall_back) {
        const int M = s.n_meshes;
        std::vector<uint8_t> redo(M, 0);
        for (int m = 0; m < M; ++m) {
            const double target = std::max(opt->absolute_tolerance, opt->relative_tolerance * refs[m]);
            redo[m] = (finals[m] <= target) ? 0 : 1;
        }
        DevBuf<uint8_t> redo_d;
        OC_CHECK(redo_d.upload(redo.data(), redo.size()));
        Runtime &rt = runtime();
        restore_meshes_kernel<<<grid_for(s.n_rows), 256, 0, rt.stream>>>(
            s.n_rows, s.n_nodes, M, s.q_off.ptr, s.l_off.ptr, redo_d.ptr, s.plain_start.ptr, s.dofs.ptr);
        OC_HIP_CHECK(hipGetLastError());
        OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
        oc_newton_options plain = *opt;
        plain.jacobian_precision = OC_FP64;
        plain.fallback = 0;
        if (opt->fallback_maximum_iterations > 0) plain.maximum_iterations = opt->fallback_maximum_iterations;
        if (opt->output_type >= OC_SOLVER_PROGRESS_OUTPUT)
            std::fprintf(stderr, "  newton: warm-started attempt failed (reason %d); plain FP64 restart\n", stats->reason);
        oc_newton_stats first = *stats, second{};
        rc = newton_core(s, &plain, refs.data(), refs, finals, &second);
        *stats = second;
        stats->newton_iterations += first.newton_iterations;
        stats->krylov_iterations += first.krylov_iterations;
        stats->residual_evaluations += first.residual_evaluations;
        stats->jacobian_evaluations += first.jacobian_evaluations;
        stats->linear_solve_failures += first.linear_solve_failures;
        stats->step_failures += first.step_failures;
        stats->initial_residual_norm = first.initial_residual_norm;
        stats->seconds += first.seconds;
        stats->fallback_used = 1;
        stats->warm_start_reason = first.reason;
    }
    s.plain_start_valid = false;
    if (finals_out && finals.size() == static_cast<size_t>(s.n_meshes))
        std::copy(finals.begin(), finals.end(), finals_out);
    return rc;
}

}  // namespace oc

using namespace oc;

extern "C" {

int oc_boundary_conditions_set(oc_equations_set h, int n, const int *dof_indices, const double *full_values,
                               int index_base)
{
    OC_CHECK(require_runtime());
    if (!h) return set_error(OC_ERR_INVALID_HANDLE, "equations set handle is NULL");
    if (n < 0 || (n > 0 && (!dof_indices || !full_values)) || (index_base != 0 && index_base != 1))
        return set_error(OC_ERR_INVALID_ARGUMENT, "oc_boundary_conditions_set: invalid arguments");
    EquationsSet &s = h->s;
    std::vector<int32_t> idx(n);
    std::vector<uint8_t> mask(s.n_rows, 1);
    for (int i = 0; i < n; ++i) {
        const int d = dof_indices[i] - index_base;
        if (d < 0 || d >= s.n_rows)
            return set_error(OC_ERR_INVALID_ARGUMENT, "Dirichlet DOF %d out of range", dof_indices[i]);
        if (!mask[d]) return set_error(OC_ERR_INVALID_ARGUMENT, "Dirichlet DOF %d given twice", dof_indices[i]);
        mask[d] = 0;
        idx[i] = d;
    }
    // the free-DOF mask keeps its device address (a recorded Newton-step graph
    // holds it; that graph is also re-recorded, its key has the BC version)
    if (s.bc_idx.n != static_cast<size_t>(n)) {
        OC_CHECK(s.bc_idx.alloc(static_cast<size_t>(n)));
        OC_CHECK(s.bc_full.alloc(static_cast<size_t>(n)));
    }
    if (!s.free_mask.ptr) OC_CHECK(s.free_mask.alloc(mask.size()));
    Runtime &rt = runtime();
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));  // kernels in flight may read the old values
    if (n) {
        OC_HIP_CHECK(hipMemcpy(s.bc_idx.ptr, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        OC_HIP_CHECK(hipMemcpy(s.bc_full.ptr, full_values, sizeof(double) * n, hipMemcpyHostToDevice));
    }
    OC_HIP_CHECK(hipMemcpy(s.free_mask.ptr, mask.data(), mask.size(), hipMemcpyHostToDevice));
    s.n_bc = n;
    ++s.bc_version;
    return OC_NO_ERROR;
}

int oc_load_increment_apply(oc_equations_set h, int iteration, int max_iterations)
{
    OC_CHECK(require_runtime());
    if (!h) return set_error(OC_ERR_INVALID_HANDLE, "equations set handle is NULL");
    if (max_iterations < 1 || iteration < 1 || iteration > max_iterations)
        return set_error(OC_ERR_INVALID_ARGUMENT, "load step %d of %d", iteration, max_iterations);
    EquationsSet &s = h->s;
    if (s.n_bc == 0) return OC_NO_ERROR;
    bc_increment_kernel<<<static_cast<unsigned>((s.n_bc + 255) / 256), 256, 0, runtime().stream>>>(
        s.n_bc, s.bc_idx.ptr, s.bc_full.ptr, s.dofs.ptr, iteration, max_iterations);
    OC_HIP_CHECK(hipGetLastError());
    if (s.plain_start_valid) {  // the plain start of a warm-started step takes the same increment
        bc_increment_kernel<<<static_cast<unsigned>((s.n_bc + 255) / 256), 256, 0, runtime().stream>>>(
            s.n_bc, s.bc_idx.ptr, s.bc_full.ptr, s.plain_start.ptr, iteration, max_iterations);
        OC_HIP_CHECK(hipGetLastError());
    }
    return OC_NO_ERROR;
}

int oc_newton_options_default(oc_newton_options *o)
{
    if (!o) return set_error(OC_ERR_INVALID_ARGUMENT, "options is NULL");
    o->absolute_tolerance = 1e-10;
    o->relative_tolerance = 1e-5;
    o->reference_norm = 0.0;
    o->maximum_iterations = 50;
    o->jacobian_precision = OC_FP64;
    o->krylov_restart = 30;
    o->krylov_maximum_iterations = 100000;
    o->krylov_relative_tolerance = 1e-5;
    o->line_search = 1;
    o->line_search_alpha = 1e-4;
    o->preconditioner = OC_PRECONDITIONER_JACOBI;
    o->solution_tolerance = 1e-5;
    o->maximum_function_evaluations = 1000;
    o->maximum_linear_solve_failures = 1;
    o->maximum_step_failures = 1;
    o->output_type = OC_SOLVER_NO_OUTPUT;
    o->fallback = 0;
    o->fallback_maximum_iterations = 0;
    return OC_NO_ERROR;
}

static int mr_solver_ready(oc_equations_set h)
{
    OC_CHECK(require_runtime());
    if (!h) return set_error(OC_ERR_INVALID_HANDLE, "equations set handle is NULL");
    if (h->s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN)
        return set_error(OC_ERR_INVALID_ARGUMENT, "defined for Mooney-Rivlin sets");
    OC_CHECK(mr_prepare(h->s));
    Krylov K{h->s};
    return solver_prepare(h->s, 1, K);
}

int oc_residual_norm(oc_equations_set h, double *norm)
{
    if (!norm) return set_error(OC_ERR_INVALID_ARGUMENT, "norm is NULL");
    OC_CHECK(mr_solver_ready(h));
    std::vector<double> v;
    OC_CHECK(residual_mesh_norms(h->s, v));
    *norm = plain_norm(v);
    return OC_NO_ERROR;
}

int oc_mesh_residual_norms(oc_equations_set h, double *norms)
{
    if (!norms) return set_error(OC_ERR_INVALID_ARGUMENT, "norms is NULL");
    OC_CHECK(mr_solver_ready(h));
    std::vector<double> v;
    OC_CHECK(residual_mesh_norms(h->s, v));
    std::copy(v.begin(), v.end(), norms);
    return OC_NO_ERROR;
}

int oc_newton_solve(oc_equations_set h, const oc_newton_options *opt, oc_newton_stats *stats)
{
    if (!opt || !stats) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_CHECK(mr_solver_ready(h));
    if (h->s.n_meshes != 1 && opt->reference_norm > 0.0)
        return set_error(OC_ERR_INVALID_ARGUMENT, "a batch takes per-mesh reference norms (oc_newton_solve_batch)");
    std::vector<double> refs(h->s.n_meshes, opt->reference_norm);
    return newton_solve_set(h->s, opt, opt->reference_norm > 0.0 ? refs.data() : nullptr, nullptr, stats);
}

int oc_newton_solve_batch(oc_equations_set h, const oc_newton_options *opt, const double *reference_norms,
                          double *final_norms, oc_newton_stats *stats)
{
    if (!opt || !stats) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_CHECK(mr_solver_ready(h));
    return newton_solve_set(h->s, opt, reference_norms, final_norms, stats);
}

int oc_jacobian_evaluate_graph(oc_equations_set h, int precision)
{
    OC_CHECK(require_runtime());
    if (!h) return set_error(OC_ERR_INVALID_HANDLE, "equations set handle is NULL");
    if (h->s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN)
        return set_error(OC_ERR_INVALID_ARGUMENT, "Jacobian evaluation is defined for Mooney-Rivlin sets only");
    if (precision != OC_FP64 && precision != OC_FP32)
        return set_error(OC_ERR_INVALID_ARGUMENT, "precision must be OC_FP64 or OC_FP32");
    OC_CHECK(mr_jacobian_buffers(h->s, precision));
    const BasisTables *T = nullptr;
    OC_CHECK(ensure_tables(&T));
    return step_assembly(h->s, nullptr, precision);
}

int oc_linear_solve(oc_equations_set h, const oc_newton_options *opt, oc_newton_stats *stats)
{
    OC_CHECK(require_runtime());
    if (!h || !opt || !stats) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    EquationsSet &s = h->s;
    if (s.cls != OC_EQUATIONS_SET_LAPLACE)
        return set_error(OC_ERR_INVALID_ARGUMENT, "linear solve is defined for Laplace sets");
    const auto t0 = std::chrono::steady_clock::now();
    *stats = oc_newton_stats{};
    Krylov K{s};
    K.precond = OC_PRECONDITIONER_JACOBI;
    OC_CHECK(solver_prepare(s, opt->krylov_restart, K));
    Runtime &rt = runtime();
    const int n = s.n_rows;
    double *rhs = s.work.ptr + 2 * static_cast<int64_t>(n), *dx = s.work.ptr + n, *tmp = s.work.ptr + 7 * static_cast<int64_t>(n);
    unit_weights_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, s.free_mask.ptr, s.wgt.ptr, s.iwgt.ptr);
    OC_HIP_CHECK(hipGetLastError());
    OC_CHECK(launch_laplace_assemble(s));
    // u = Dirichlet values on fixed DOFs, 0 elsewhere; rhs = RHS - K u on free rows
    OC_CHECK(fill<double>(s.dofs.ptr, n, 0.0));
    if (s.n_bc) OC_CHECK(oc_load_increment_apply(h, 1, 1));
    OC_CHECK(K.launch_spmv<double>(s.values.ptr, 8, nullptr, s.dofs.ptr, tmp));
    OC_HIP_CHECK(hipMemcpyAsync(rhs, s.rhs.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice, rt.stream));
    axpby_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, -1.0, tmp, 1.0, rhs);
    scale_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, s.wgt.ptr, rhs);
    OC_CHECK(K.setup_preconditioner());
    double bnorm = 0.0;
    OC_CHECK(K.norm(rhs, &bnorm));
    stats->initial_residual_norm = bnorm;
    const double target = std::max(opt->absolute_tolerance, opt->krylov_relative_tolerance * bnorm);
    int kit = 0;
    double kres = 0.0;
    OC_CHECK(K.gmres(rhs, dx, target, opt->krylov_maximum_iterations, &kit, &kres));
    axpby_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, 1.0, dx, 1.0, s.dofs.ptr);
    // true residual of the free rows
    OC_CHECK(K.launch_spmv<double>(s.values.ptr, 8, nullptr, s.dofs.ptr, tmp));
    axpby_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, 1.0, s.rhs.ptr, -1.0, tmp);
    scale_kernel<<<grid_for(n), 256, 0, rt.stream>>>(n, s.wgt.ptr, tmp);
    double fres = 0.0;
    OC_CHECK(K.norm(tmp, &fres));
    stats->krylov_iterations = kit;
    stats->newton_iterations = 1;
    stats->final_residual_norm = fres;
    stats->converged = fres <= std::max(target * 1.0001, opt->absolute_tolerance) ? 1 : 0;
    stats->reason = stats->converged ? (fres <= opt->absolute_tolerance ? OC_CONVERGED_FNORM_ABS : OC_CONVERGED_FNORM_RELATIVE)
                                     : OC_DIVERGED_MAX_IT;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!stats->converged)
        return set_error(OC_ERR_NOT_CONVERGED, "linear solve did not converge: ||r|| = %.3e > %.3e", fres, target);
    return OC_NO_ERROR;
}

}  // extern "C"
