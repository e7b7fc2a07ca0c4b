// oc_multigrid.hip -- geometric multigrid V-cycle for the displacement block
// K_uu of the Mooney-Rivlin Jacobian on regular generated meshes (no
// counterpart in the reference, whose PETSc KSP uses the preconditioner the
// user selects; solver_routines.f90:10957-11006).
//
// Hierarchy, per mesh of n^3 elements:
//   level 0  triquadratic nodes (2n+1)^3: the assembled CSR itself (u-u part,
//            columns left of the first pressure column of each row);
//   level 1  trilinear nodes (n+1)^3 on the same elements (numbered like the
//            pressure nodes); operator P0^T K P0 formed element by element
//            from the element matrices of the last Jacobian assembly,
//            K1_e = W^T M_e K_e M_e W (W: trilinear values at the 27 element
//            nodes, M_e: free-DOF mask), assembled into a 27-point stencil;
//   level l+1 (ceil(n_l/2)+1)^3 nodes, linear interpolation in each lattice
//            direction (even fine index -> coarse i/2, odd -> mean of its two
//            coarse neighbours), operator P^T A_l P on the stencils.
// Dirichlet DOFs: interpolation rows of fixed fine DOFs are dropped, a coarse
// DOF coinciding with a fixed fine DOF is fixed (its column dropped), and a
// coarse DOF left with an empty row is fixed too; fixed rows and columns are
// the identity.
// Smoother: damped nodal 3x3 block Jacobi; coarsest level (<= 3 nodes per
// direction): dense inverse per mesh.  Every reduction order is fixed, so the
// V-cycle is a fixed linear operator (usable inside GMRES).
#include <algorithm>
#include <cstdint>
#include <memory>
#include <utility>

#include "oc_fe_device.h"
#include "oc_internal.h"

namespace oc {
namespace {

__device__ __forceinline__ int find_mesh(const int *__restrict__ off, int M, int node)
{
    int lo = 0, hi = M - 1;  // off[lo] <= node < off[lo + 1]
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= node) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ double lin_weight(int delta) { return delta == 0 ? 1.0 : 0.5; }

// ---- setup -----------------------------------------------------------------------------
template <typename Real>
__global__ void __launch_bounds__(256)
q1_galerkin_element_kernel(int n_elements, const Real *__restrict__ kbuf, const int *__restrict__ conn_q, int nq,
                           const uint8_t *__restrict__ free_mask, double *__restrict__ ke1)
{
    __shared__ Real sK[81][82];  // K_e in its own precision: an FP32 Jacobian halves the block's LDS
    __shared__ double sT[81][25];
    // W[a][c], the trilinear corner function c at quadratic node a, is nonzero
    // at exactly the 2 x 2 x 2 nodes a on c's side of the element: both
    // products below run over those eight (ascending a), which leaves out only
    // exact zero terms and so gives the same sums as the full 27-term loops
    __shared__ double sWn[8][8];
    __shared__ int sA[8][8];
    __shared__ uint8_t sf[81];
    const int t = threadIdx.x;
    if (t < 8) {
        const int c = t;
        int q = 0;
        for (int a = 0; a < 27; ++a) {
            double w = 1.0;
            const int ad[3] = {a % 3, (a / 3) % 3, a / 9};
            for (int d = 0; d < 3; ++d) {
                const double x = 0.5 * ad[d];
                w *= ((c >> d) & 1) ? x : 1.0 - x;
            }
            if (w != 0.0 && q < 8) {
                sA[c][q] = a;
                sWn[c][q] = w;
                ++q;
            }
        }
    }
    for (int e = blockIdx.x; e < n_elements; e += gridDim.x) {
        __syncthreads();
        if (t < 81) sf[t] = free_mask[static_cast<int64_t>(t / 27) * nq + conn_q[static_cast<int64_t>(e) * 27 + t % 27]];
        __syncthreads();
        const Real *Ke = kbuf + static_cast<int64_t>(e) * 89 * 89;
        for (int k = t; k < 81 * 81; k += 256) {
            const int i = k / 81, j = k % 81;
            sK[i][j] = (sf[i] && sf[j]) ? Ke[i * 89 + j] : static_cast<Real>(0);
        }
        __syncthreads();
        for (int k = t; k < 81 * 24; k += 256) {
            const int i = k / 24, col = k % 24, cp = col / 8, kk = col % 8;
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) acc += static_cast<double>(sK[i][cp * 27 + sA[kk][q]]) * sWn[kk][q];
            sT[i][col] = acc;
        }
        __syncthreads();
        for (int k = t; k < 24 * 24; k += 256) {
            const int row = k / 24, col = k % 24, c = row / 8, kr = row % 8;
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) acc += sWn[kr][q] * sT[c * 27 + sA[kr][q]][col];
            ke1[static_cast<int64_t>(e) * 576 + k] = acc;
        }
    }
}

// level-1 stencil: block (I, I+d) = sum over the elements holding both
// corners (ascending element order) of K1_e[(c, corner I), (c', corner J)]
__global__ void __launch_bounds__(256)
q1_stencil_kernel(int M, const int *__restrict__ n_mesh, const int *__restrict__ e_off, const int *__restrict__ off1,
                  int nn, const double *__restrict__ ke1, double *__restrict__ A)
{
    const int64_t total = static_cast<int64_t>(nn) * 27;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int node = static_cast<int>(t / 27), d = static_cast<int>(t % 27);
        const int m = find_mesh(off1, M, node);
        const int n = n_mesh[m], N = n + 1, local = node - off1[m];
        const int I[3] = {local % N, (local / N) % N, local / (N * N)};
        const int J[3] = {I[0] + d % 3 - 1, I[1] + (d / 3) % 3 - 1, I[2] + d / 9 - 1};
        double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (J[0] >= 0 && J[0] < N && J[1] >= 0 && J[1] < N && J[2] >= 0 && J[2] < N) {
            int lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                lo[k] = max(max(I[k], J[k]) - 1, 0);
                hi[k] = min(min(I[k], J[k]), n - 1);
            }
            for (int ez = lo[2]; ez <= hi[2]; ++ez)
                for (int ey = lo[1]; ey <= hi[1]; ++ey)
                    for (int ex = lo[0]; ex <= hi[0]; ++ex) {
                        const int64_t e = e_off[m] + ex + static_cast<int64_t>(n) * (ey + static_cast<int64_t>(n) * ez);
                        const int kI = (I[0] - ex) + 2 * (I[1] - ey) + 4 * (I[2] - ez);
                        const int kJ = (J[0] - ex) + 2 * (J[1] - ey) + 4 * (J[2] - ez);
                        const double *k1 = ke1 + e * 576;
                        for (int c = 0; c < 3; ++c)
                            for (int cp = 0; cp < 3; ++cp) blk[3 * c + cp] += k1[(c * 8 + kI) * 24 + cp * 8 + kJ];
                    }
        }
        for (int k = 0; k < 9; ++k) A[t * 9 + k] = blk[k];
    }
}

// Galerkin coarsening of a stencil level (masked by the fine level's free
// DOFs): coarse block (I, J) = sum over fine nodes i near I and j near J of
// w_i w_j A_f(i, j).  A block of 243 threads takes nine (coarse node,
// direction) pairs; thread (pair, s) sums fine node s of I's neighbourhood
// against every j, and the 27 partial blocks are added in s order, so the
// sum is deterministic and one pair's work is spread over 27 threads (the
// coarse levels have few nodes).
constexpr int kGalPairs = 9;

__global__ void __launch_bounds__(27 * kGalPairs)
galerkin_coarsen_kernel(int M, const int *__restrict__ dim_f, const int *__restrict__ off_f,
                        const double *__restrict__ Af, const uint8_t *__restrict__ freed_f,
                        const int *__restrict__ dim_c, const int *__restrict__ off_c, int nn_c,
                        double *__restrict__ Ac)
{
    __shared__ double s_part[kGalPairs][27][9];
    const int p = threadIdx.x / 27, s = threadIdx.x % 27;
    const int64_t total = static_cast<int64_t>(nn_c) * 27;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kGalPairs + p;
    double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (t < total) {
        const int node = static_cast<int>(t / 27), d = static_cast<int>(t % 27);
        const int m = find_mesh(off_c, M, node);
        const int Nc = dim_c[m], Nf = dim_f[m], local = node - off_c[m];
        const int I[3] = {local % Nc, (local / Nc) % Nc, local / (Nc * Nc)};
        const int J[3] = {I[0] + d % 3 - 1, I[1] + (d / 3) % 3 - 1, I[2] + d / 9 - 1};
        const int i[3] = {2 * I[0] + s % 3 - 1, 2 * I[1] + (s / 3) % 3 - 1, 2 * I[2] + s / 9 - 1};
        if (J[0] >= 0 && J[0] < Nc && J[1] >= 0 && J[1] < Nc && J[2] >= 0 && J[2] < Nc && i[0] >= 0 && i[0] < Nf &&
            i[1] >= 0 && i[1] < Nf && i[2] >= 0 && i[2] < Nf) {
            const double wi = lin_weight(s % 3 - 1) * lin_weight((s / 3) % 3 - 1) * lin_weight(s / 9 - 1);
            const int64_t fi = off_f[m] + i[0] + static_cast<int64_t>(Nf) * (i[1] + static_cast<int64_t>(Nf) * i[2]);
            // fine neighbours j = 2J + q - 1 of J inside the grid and within
            // one node of i, per direction (the other q add nothing), in
            // ascending q = q0 + 3 q1 + 9 q2
            int qlo[3], qhi[3];
            for (int k = 0; k < 3; ++k) {
                qlo[k] = max(max(0, 1 - 2 * J[k]), i[k] - 2 * J[k]);
                qhi[k] = min(min(2, Nf - 2 * J[k]), i[k] - 2 * J[k] + 2);
            }
            for (int q2 = qlo[2]; q2 <= qhi[2]; ++q2)
                for (int q1 = qlo[1]; q1 <= qhi[1]; ++q1)
                    for (int q0 = qlo[0]; q0 <= qhi[0]; ++q0) {
                        const int j[3] = {2 * J[0] + q0 - 1, 2 * J[1] + q1 - 1, 2 * J[2] + q2 - 1};
                        const int e0 = j[0] - i[0], e1 = j[1] - i[1], e2 = j[2] - i[2];
                        const double w = wi * lin_weight(q0 - 1) * lin_weight(q1 - 1) * lin_weight(q2 - 1);
                        const int64_t fj =
                            off_f[m] + j[0] + static_cast<int64_t>(Nf) * (j[1] + static_cast<int64_t>(Nf) * j[2]);
                        const double *a = Af + (fi * 27 + (e0 + 1) + 3 * (e1 + 1) + 9 * (e2 + 1)) * 9;
                        for (int c = 0; c < 3; ++c) {
                            if (!freed_f[fi * 3 + c]) continue;
                            for (int cp = 0; cp < 3; ++cp)
                                if (freed_f[fj * 3 + cp]) blk[3 * c + cp] += w * a[3 * c + cp];
                        }
                    }
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) s_part[p][s][k] = blk[k];
    __syncthreads();
    if (threadIdx.x < kGalPairs * 9) {
        const int pp = threadIdx.x / 9, k = threadIdx.x % 9;
        const int64_t tt = static_cast<int64_t>(blockIdx.x) * kGalPairs + pp;
        if (tt < total) {
            double acc = 0.0;
            for (int ss = 0; ss < 27; ++ss) acc += s_part[pp][ss][k];
            Ac[tt * 9 + k] = acc;
        }
    }
}

// Coarse DOFs that coincide with a fixed fine DOF are fixed (a Dirichlet
// face stays Dirichlet on every level).  Level 1 from the triquadratic mask:
// coarse node I coincides with fine lattice node 2I.
__global__ void __launch_bounds__(256) coincident_q2_kernel(int M, const int *__restrict__ n_mesh,
                                                           const int *__restrict__ q_off, int nq_total,
                                                           const int *__restrict__ off1, int nn1,
                                                           const uint8_t *__restrict__ free_mask,
                                                           uint8_t *__restrict__ freed1)
{
    for (int node = blockIdx.x * blockDim.x + threadIdx.x; node < nn1; node += gridDim.x * blockDim.x) {
        const int m = find_mesh(off1, M, node);
        const int n = n_mesh[m], N = n + 1, T = 2 * n + 1, local = node - off1[m];
        const int64_t q = q_off[m] + 2 * (local % N) +
                          static_cast<int64_t>(T) * (2 * ((local / N) % N) + static_cast<int64_t>(T) * 2 * (local / (N * N)));
        for (int c = 0; c < 3; ++c) freed1[static_cast<int64_t>(node) * 3 + c] = free_mask[static_cast<int64_t>(c) * nq_total + q];
    }
}

// level l -> l+1: coarse I coincides with fine 2I; the extra coarse node of
// an odd fine grid (2I beyond the last fine index) takes the flag of the
// boundary node it extends, so every Dirichlet face stays one on the coarse
// grid and the masked interpolation keeps full column rank
__global__ void __launch_bounds__(256) coincident_kernel(int M, const int *__restrict__ dim_f,
                                                        const int *__restrict__ off_f,
                                                        const uint8_t *__restrict__ freed_f,
                                                        const int *__restrict__ dim_c, const int *__restrict__ off_c,
                                                        int nn_c, uint8_t *__restrict__ freed_c)
{
    for (int node = blockIdx.x * blockDim.x + threadIdx.x; node < nn_c; node += gridDim.x * blockDim.x) {
        const int m = find_mesh(off_c, M, node);
        const int Nc = dim_c[m], Nf = dim_f[m], local = node - off_c[m];
        const int i0 = min(2 * (local % Nc), Nf - 1), i1 = min(2 * ((local / Nc) % Nc), Nf - 1),
                  i2 = min(2 * (local / (Nc * Nc)), Nf - 1);
        const int64_t f = off_f[m] + i0 + static_cast<int64_t>(Nf) * (i1 + static_cast<int64_t>(Nf) * i2);
        for (int c = 0; c < 3; ++c) freed_c[static_cast<int64_t>(node) * 3 + c] = freed_f[f * 3 + c];
    }
}

// freed &= (diagonal != 0): a DOF whose Galerkin row is empty is fixed too
__global__ void __launch_bounds__(256) level_freed_kernel(int nn, const double *__restrict__ A,
                                                         uint8_t *__restrict__ freed)
{
    for (int node = blockIdx.x * blockDim.x + threadIdx.x; node < nn; node += gridDim.x * blockDim.x) {
        const double *D = A + (static_cast<int64_t>(node) * 27 + 13) * 9;
        for (int c = 0; c < 3; ++c) {
            const int64_t k = static_cast<int64_t>(node) * 3 + c;
            freed[k] = (freed[k] && D[4 * c] != 0.0) ? 1 : 0;
        }
    }
}

// Fixed rows and columns replaced by the identity; inverse diagonal blocks.
__global__ void __launch_bounds__(256) level_mask_kernel(int M, const int *__restrict__ dim,
                                                        const int *__restrict__ off, int nn, double *__restrict__ A,
                                                        const uint8_t *__restrict__ freed, double *__restrict__ Dinv)
{
    const int64_t total = static_cast<int64_t>(nn) * 27;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int node = static_cast<int>(t / 27), d = static_cast<int>(t % 27);
        const int m = find_mesh(off, M, node);
        const int N = dim[m], local = node - off[m];
        const int J[3] = {local % N + d % 3 - 1, (local / N) % N + (d / 3) % 3 - 1, local / (N * N) + d / 9 - 1};
        double *a = A + t * 9;
        if (J[0] < 0 || J[0] >= N || J[1] < 0 || J[1] >= N || J[2] < 0 || J[2] >= N) {
            for (int k = 0; k < 9; ++k) a[k] = 0.0;
            continue;
        }
        const int64_t j = off[m] + J[0] + static_cast<int64_t>(N) * (J[1] + static_cast<int64_t>(N) * J[2]);
        for (int c = 0; c < 3; ++c)
            for (int cp = 0; cp < 3; ++cp) {
                const bool keep = freed[static_cast<int64_t>(node) * 3 + c] && freed[j * 3 + cp];
                if (!keep) a[3 * c + cp] = (d == 13 && c == cp) ? 1.0 : 0.0;
            }
        if (d != 13) continue;
        const double *D = a;
        const double c00 = D[4] * D[8] - D[5] * D[7], c01 = D[5] * D[6] - D[3] * D[8], c02 = D[3] * D[7] - D[4] * D[6];
        const double det = D[0] * c00 + D[1] * c01 + D[2] * c02;
        double *o = Dinv + static_cast<int64_t>(node) * 9;
        if (fabs(det) < 1e-300) {
            for (int i = 0; i < 9; ++i) o[i] = (i % 4 == 0) ? 1.0 : 0.0;
            continue;
        }
        const double id = 1.0 / det;
        o[0] = c00 * id;
        o[1] = (D[2] * D[7] - D[1] * D[8]) * id;
        o[2] = (D[1] * D[5] - D[2] * D[4]) * id;
        o[3] = c01 * id;
        o[4] = (D[0] * D[8] - D[2] * D[6]) * id;
        o[5] = (D[2] * D[3] - D[0] * D[5]) * id;
        o[6] = c02 * id;
        o[7] = (D[1] * D[6] - D[0] * D[7]) * id;
        o[8] = (D[0] * D[4] - D[1] * D[3]) * id;
    }
}

constexpr int kCoarseMax = 81;  // 27 nodes x 3 components

// dense inverse of each mesh's coarsest level (in-place Gauss-Jordan, one
// block per mesh; the masked operator is symmetric positive definite)
__global__ void __launch_bounds__(256) coarse_inverse_kernel(int M, const int *__restrict__ dim,
                                                            const int *__restrict__ off,
                                                            const double *__restrict__ A, int stride,
                                                            double *__restrict__ inv)
{
    __shared__ double a[kCoarseMax][kCoarseMax + 1];
    __shared__ double fcol[kCoarseMax];
    __shared__ double piv;
    const int m = blockIdx.x, t = threadIdx.x;
    if (m >= M) return;
    const int N = dim[m], nodes = N * N * N, nd = 3 * nodes;
    for (int k = t; k < nd * nd; k += 256) {
        const int r = k / nd, col = k % nd;
        const int nr = r / 3, c = r % 3, nc = col / 3, cp = col % 3;
        const int e0 = nc % N - nr % N, e1 = (nc / N) % N - (nr / N) % N, e2 = nc / (N * N) - nr / (N * N);
        double v = 0.0;
        if (e0 >= -1 && e0 <= 1 && e1 >= -1 && e1 <= 1 && e2 >= -1 && e2 <= 1)
            v = A[((static_cast<int64_t>(off[m]) + nr) * 27 + (e0 + 1) + 3 * (e1 + 1) + 9 * (e2 + 1)) * 9 + 3 * c + cp];
        a[r][col] = v;
    }
    __syncthreads();
    for (int k = 0; k < nd; ++k) {
        if (t == 0) piv = (fabs(a[k][k]) > 1e-300) ? a[k][k] : 1.0;  // fixed DOFs are identity rows
        __syncthreads();
        const double p = piv;
        for (int j = t; j < nd; j += 256) a[k][j] = (j == k ? 1.0 : a[k][j]) / p;
        for (int i = t; i < nd; i += 256) fcol[i] = (i == k) ? 0.0 : a[i][k];
        __syncthreads();
        for (int i = t >> 5; i < nd; i += 8) {  // 8 x 32 threads over rows and columns
            if (i == k) continue;
            const double f = fcol[i];
            for (int j = t & 31; j < nd; j += 32) a[i][j] = (j == k ? 0.0 : a[i][j]) - f * a[k][j];
        }
        __syncthreads();
    }
    for (int k = t; k < nd * nd; k += 256) inv[static_cast<int64_t>(m) * stride + k] = a[k / nd][k % nd];
}

// ---- V-cycle kernels -------------------------------------------------------------------
// level 0: r = b - K_uu x on free rows (columns left of the first pressure column)
template <typename VT, int kLanes>
__global__ void __launch_bounds__(256)
uu_residual_kernel(int nu, const int64_t *__restrict__ row_ptr, const int64_t *__restrict__ pstart,
                   const int *__restrict__ col_idx, const VT *__restrict__ vals, const uint8_t *__restrict__ free_mask,
                   const double *__restrict__ b, const double *__restrict__ x, double *__restrict__ r)
{
    const int lane = threadIdx.x % kLanes;
    const int row = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kLanes);
    if (row >= nu) return;
    const bool f = free_mask[row] != 0;
    double acc = 0.0;
    if (f) acc = csr_row_dot<VT, kLanes>(row_ptr[row], pstart[row], lane, col_idx, vals, x);
#pragma unroll
    for (int o = kLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kLanes);
    if (lane == 0) r[row] = f ? b[row] - acc : 0.0;
}

// bf16 values: the upper 16 bits of an FP32 value, rounded to nearest even
// (NaN kept quiet)
__device__ __forceinline__ uint16_t bf16_round(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

// The packed bf16 copy of K_uu (EquationsSet::uu_bf): padded neighbour
// counts, then the padded column lists, once per set
__global__ void __launch_bounds__(256) uu_degp_kernel(int nq, const int64_t *__restrict__ row_ptr,
                                                     const int64_t *__restrict__ pstart, int64_t *__restrict__ degp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const int64_t deg = (pstart[i] - row_ptr[i]) / 3;
    degp[i] = (deg + 3) & ~static_cast<int64_t>(3);
}

constexpr int kPackLanes = 32;  // lanes per node in the packing kernels

__global__ void __launch_bounds__(256) uu_cols_kernel(int nq, const int64_t *__restrict__ row_ptr,
                                                     const int64_t *__restrict__ pstart, const int *__restrict__ col_idx,
                                                     const int64_t *__restrict__ uu_off, int *__restrict__ uu_col)
{
    const int lane = threadIdx.x % kPackLanes;
    const int i = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kPackLanes);
    if (i >= nq) return;
    const int64_t s0 = row_ptr[i], off = uu_off[i];
    const int deg = static_cast<int>((pstart[i] - s0) / 3), degp = static_cast<int>(uu_off[i + 1] - off);
    for (int p = lane; p < degp; p += kPackLanes) uu_col[off + p] = p < deg ? col_idx[s0 + p] : i;
}

// values of one FP32 Jacobian into the packed bf16 copy (each Newton step)
__global__ void __launch_bounds__(256) uu_pack_kernel(int nq, const int64_t *__restrict__ row_ptr,
                                                     const int64_t *__restrict__ pstart, const float *__restrict__ v,
                                                     const int64_t *__restrict__ uu_off, uint16_t *__restrict__ uu_bf)
{
    const int lane = threadIdx.x % kPackLanes;
    const int i = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kPackLanes);
    if (i >= nq) return;
    const int64_t nq64 = nq, off = uu_off[i];
    const int deg = static_cast<int>((pstart[i] - row_ptr[i]) / 3), degp = static_cast<int>(uu_off[i + 1] - off);
    const int64_t rows[3] = {row_ptr[i], row_ptr[nq64 + i], row_ptr[2 * nq64 + i]};
    // run (r, c) starts at 9 off + (3 r + c) degp, a multiple of four (degp
    // and every offset are), so a lane writes two neighbours as one 4-byte
    // word: the lower address holds the even neighbour
    uint32_t *o = reinterpret_cast<uint32_t *>(uu_bf + 9 * off);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float *src = v + rows[r] + static_cast<int64_t>(c) * deg;
            uint32_t *dst = o + (3 * r + c) * (degp / 2);
            for (int p = 2 * lane; p < degp; p += 2 * kPackLanes) {
                const uint32_t lo = p < deg ? bf16_round(src[p]) : 0u;
                const uint32_t hi = p + 1 < deg ? bf16_round(src[p + 1]) : 0u;
                dst[p / 2] = lo | (hi << 16);
            }
        }
}

// the iterate as FP32 node quadruples (x, y, z, 0) for the bf16 kernels'
// gathers: one 16-byte load per neighbour, and no accuracy lost against the
// bf16 values' 8-bit significands
__global__ void __launch_bounds__(256) pack_x4_kernel(int nq, const double *__restrict__ x, float4 *__restrict__ x4)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nq) return;
    const int64_t nq64 = nq;
    x4[j] = make_float4(static_cast<float>(x[j]), static_cast<float>(x[nq64 + j]), static_cast<float>(x[2 * nq64 + j]),
                        0.0f);
}

// level 0 residual by nodes from the packed bf16 K_uu: one group of kLanes
// lanes per node; a lane takes four neighbours at a time (kQ quads in
// flight): one 16-byte column load, one 8-byte load per (row, component)
// run and one 16-byte iterate load per neighbour (x4: pack_x4_kernel of x).
// r = b - K_uu x on free rows, 0 on fixed ones; kSweep 1 writes the block
// Jacobi update x + omega Kblock^-1 r instead (as uu_node_residual_kernel),
// and with out4 also its FP32 node quadruples for the next pass.
template <int kLanes, int kSweep, int kQ>
__global__ void __launch_bounds__(256)
uu_node_bf16_kernel(int nq, const int64_t *__restrict__ uu_off, const int *__restrict__ uu_col,
                    const uint16_t *__restrict__ uu_bf, const uint8_t *__restrict__ free_mask,
                    const double *__restrict__ b, const float4 *__restrict__ x4, const double *__restrict__ x,
                    const double *__restrict__ inv_u, double omega, double *__restrict__ out,
                    float *__restrict__ out4)
{
    const int lane = threadIdx.x % kLanes;
    const int64_t xb = xcd_block(blockIdx.x, gridDim.x);
    const int i = static_cast<int>((xb * blockDim.x + threadIdx.x) / kLanes);
    if (i >= nq) return;  // whole groups
    const int64_t nq64 = nq, off = uu_off[i];
    const int n4 = static_cast<int>((uu_off[i + 1] - off) >> 2);
    const int4 *cols = reinterpret_cast<const int4 *>(uu_col + off);
    const uint2 *vals = reinterpret_cast<const uint2 *>(uu_bf + 9 * off);  // run rc: vals[rc n4 + q]
    double a[3] = {0.0, 0.0, 0.0};
    for (int q0 = lane; q0 < n4; q0 += kLanes * kQ) {
        int4 cj[kQ];
        uint2 v[kQ][9];
#pragma unroll
        for (int u = 0; u < kQ; ++u) {
            const int q = q0 + u * kLanes;
            const bool ok = q < n4;
            cj[u] = ok ? cols[q] : make_int4(i, i, i, i);
#pragma unroll
            for (int rc = 0; rc < 9; ++rc) v[u][rc] = ok ? vals[rc * n4 + q] : make_uint2(0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < kQ; ++u) {
            const int j[4] = {cj[u].x, cj[u].y, cj[u].z, cj[u].w};
            double xv[4][3];
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const float4 q4 = x4[j[n]];
                xv[n][0] = q4.x;
                xv[n][1] = q4.y;
                xv[n][2] = q4.z;
            }
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const uint2 w = v[u][3 * r + c];
                    const float f[4] = {__uint_as_float(w.x << 16), __uint_as_float(w.x & 0xffff0000u),
                                        __uint_as_float(w.y << 16), __uint_as_float(w.y & 0xffff0000u)};
#pragma unroll
                    for (int n = 0; n < 4; ++n) a[r] = fma(static_cast<double>(f[n]), xv[n][c], a[r]);
                }
        }
    }
    double a0 = a[0], a1 = a[1], a2 = a[2];
#pragma unroll
    for (int o = kLanes / 2; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o, kLanes);
        a1 += __shfl_xor(a1, o, kLanes);
        a2 += __shfl_xor(a2, o, kLanes);
    }
    if (kSweep == 0) {
        if (lane < 3) {
            const int64_t row = lane * nq64 + i;
            const double a = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
            out[row] = free_mask[row] ? b[row] - a : 0.0;
        }
    } else if (lane < 3) {
        const double r0 = free_mask[i] ? b[i] - a0 : 0.0;
        const double r1 = free_mask[nq64 + i] ? b[nq64 + i] - a1 : 0.0;
        const double r2 = free_mask[2 * nq64 + i] ? b[2 * nq64 + i] - a2 : 0.0;
        const double *o = inv_u + static_cast<int64_t>(i) * 9 + 3 * lane;
        const int64_t row = lane * nq64 + i;
        const double xn = x[row] + omega * (o[0] * r0 + o[1] * r1 + o[2] * r2);
        out[row] = xn;
        if (out4) out4[4 * static_cast<int64_t>(i) + lane] = static_cast<float>(xn);
    }
}

// level 0 residual by nodes (EquationsSet::node_rows): one group of kLanes
// lanes per node reads the neighbour list once and the nine value blocks of
// its three rows; r = b - K_uu x on free rows, 0 on fixed ones.  kSweep 0
// writes r to out; kSweep 1 writes the block Jacobi update
// x + omega Kblock^-1 r to out (not x: neighbours still read the old x).
template <typename VT, int kLanes, int kSweep, int kU>
__global__ void __launch_bounds__(256)
uu_node_residual_kernel(int nq, const int64_t *__restrict__ row_ptr, const int64_t *__restrict__ pstart,
                        const int *__restrict__ col_idx, const VT *__restrict__ vals,
                        const uint8_t *__restrict__ free_mask, const double *__restrict__ b,
                        const double *__restrict__ x, const double *__restrict__ inv_u, double omega,
                        double *__restrict__ out)
{
    const int lane = threadIdx.x % kLanes;
    const int64_t xb = xcd_block(blockIdx.x, gridDim.x);
    const int i = static_cast<int>((xb * blockDim.x + threadIdx.x) / kLanes);
    if (i >= nq) return;  // whole groups
    const int64_t nq64 = nq;
    const int64_t sr[3] = {row_ptr[i], row_ptr[nq64 + i], row_ptr[2 * nq64 + i]};
    const int deg = static_cast<int>((pstart[i] - sr[0]) / 3);
    double a[3] = {0.0, 0.0, 0.0};
    node_rows_uu<VT, kLanes, kU>(sr, deg, lane, nq64, col_idx, vals, x, a);
    double a0 = a[0], a1 = a[1], a2 = a[2];
#pragma unroll
    for (int o = kLanes / 2; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o, kLanes);
        a1 += __shfl_xor(a1, o, kLanes);
        a2 += __shfl_xor(a2, o, kLanes);
    }
    if (kSweep == 0) {
        if (lane < 3) {
            const int64_t row = lane * nq64 + i;
            const double a = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
            out[row] = free_mask[row] ? b[row] - a : 0.0;
        }
    } else if (lane < 3) {
        const double r0 = free_mask[i] ? b[i] - a0 : 0.0;
        const double r1 = free_mask[nq64 + i] ? b[nq64 + i] - a1 : 0.0;
        const double r2 = free_mask[2 * nq64 + i] ? b[2 * nq64 + i] - a2 : 0.0;
        const double *o = inv_u + static_cast<int64_t>(i) * 9 + 3 * lane;
        const int64_t row = lane * nq64 + i;
        out[row] = x[row] + omega * (o[0] * r0 + o[1] * r1 + o[2] * r2);
    }
}

// level 0: x = (first ? 0 : x) + omega Kblock^-1 r, component-major (and,
// with x4, the result's FP32 node quadruples for the bf16 passes)
__global__ void __launch_bounds__(256) q2_jacobi_kernel(int nq, const double *__restrict__ inv_u,
                                                       const double *__restrict__ r, double *__restrict__ x,
                                                       double omega, int first, float *__restrict__ x4)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
        const double *o = inv_u + static_cast<int64_t>(i) * 9;
        const double r0 = r[i], r1 = r[nq + i], r2 = r[2 * static_cast<int64_t>(nq) + i];
        const double d[3] = {o[0] * r0 + o[1] * r1 + o[2] * r2, o[3] * r0 + o[4] * r1 + o[5] * r2,
                             o[6] * r0 + o[7] * r1 + o[8] * r2};
        for (int c = 0; c < 3; ++c) {
            const int64_t k = static_cast<int64_t>(c) * nq + i;
            const double xn = (first ? 0.0 : x[k]) + omega * d[c];
            x[k] = xn;
            if (x4) x4[4 * static_cast<int64_t>(i) + c] = static_cast<float>(xn);
        }
    }
}

// level 0 -> 1 restriction b1 = P0^T r0 (r0 is zero on fixed DOFs)
__global__ void __launch_bounds__(256) restrict_q2_kernel(int M, const int *__restrict__ n_mesh,
                                                         const int *__restrict__ q_off, int nq_total,
                                                         const int *__restrict__ off1, int nn1,
                                                         const uint8_t *__restrict__ freed1,
                                                         const double *__restrict__ r0, double *__restrict__ b1)
{
    // one 32-lane group per coarse node: lane s < 27 takes fine node s of its
    // neighbourhood, a fixed butterfly adds them
    const int node = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 5);
    const int s = threadIdx.x & 31;
    if (node >= nn1) return;  // whole 32-lane groups
    const int m = find_mesh(off1, M, node);
    const int n = n_mesh[m], N = n + 1, T = 2 * n + 1, local = node - off1[m];
    const int I[3] = {local % N, (local / N) % N, local / (N * N)};
    double acc[3] = {0.0, 0.0, 0.0};
    if (s < 27) {
        const int i[3] = {2 * I[0] + s % 3 - 1, 2 * I[1] + (s / 3) % 3 - 1, 2 * I[2] + s / 9 - 1};
        if (i[0] >= 0 && i[0] < T && i[1] >= 0 && i[1] < T && i[2] >= 0 && i[2] < T) {
            const double w = lin_weight(s % 3 - 1) * lin_weight((s / 3) % 3 - 1) * lin_weight(s / 9 - 1);
            const int64_t q = q_off[m] + i[0] + static_cast<int64_t>(T) * (i[1] + static_cast<int64_t>(T) * i[2]);
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] = w * r0[static_cast<int64_t>(c) * nq_total + q];
        }
    }
#pragma unroll
    for (int sh = 16; sh > 0; sh >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += __shfl_xor(acc[c], sh, 32);
    if (s != 0) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int64_t k = static_cast<int64_t>(node) * 3 + c;
        b1[k] = freed1[k] ? acc[c] : 0.0;
    }
}

// level 1 -> 0 prolongation x0 += P0 x1 on free DOFs (and, with x4, the
// result's FP32 node quadruples for the bf16 passes)
__global__ void __launch_bounds__(256) prolong_q2_kernel(int M, const int *__restrict__ n_mesh,
                                                        const int *__restrict__ q_off, int nq_total,
                                                        const int *__restrict__ off1,
                                                        const uint8_t *__restrict__ free_mask,
                                                        const double *__restrict__ x1, double *__restrict__ x0,
                                                        float *__restrict__ x4)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq_total; q += gridDim.x * blockDim.x) {
        const int m = find_mesh(q_off, M, q);
        const int n = n_mesh[m], N = n + 1, T = 2 * n + 1, local = q - q_off[m];
        const int i[3] = {local % T, (local / T) % T, local / (T * T)};
        int p0[3], p1[3];
        double w0[3], w1[3];
        for (int k = 0; k < 3; ++k) {
            p0[k] = i[k] >> 1;
            p1[k] = (i[k] + 1) >> 1;
            w0[k] = (i[k] & 1) ? 0.5 : 1.0;
            w1[k] = (i[k] & 1) ? 0.5 : 0.0;
        }
        double v[3] = {0.0, 0.0, 0.0};
        for (int s = 0; s < 8; ++s) {
            const int b0 = s & 1, b1 = (s >> 1) & 1, b2 = s >> 2;
            const double w = (b0 ? w1[0] : w0[0]) * (b1 ? w1[1] : w0[1]) * (b2 ? w1[2] : w0[2]);
            if (w == 0.0) continue;
            const int64_t P = off1[m] + (b0 ? p1[0] : p0[0]) +
                              static_cast<int64_t>(N) * ((b1 ? p1[1] : p0[1]) + static_cast<int64_t>(N) * (b2 ? p1[2] : p0[2]));
            for (int c = 0; c < 3; ++c) v[c] += w * x1[P * 3 + c];
        }
        for (int c = 0; c < 3; ++c) {
            const int64_t k = static_cast<int64_t>(c) * nq_total + q;
            if (free_mask[k]) x0[k] += v[c];
            if (x4) x4[4 * static_cast<int64_t>(q) + c] = static_cast<float>(x0[k]);
        }
    }
}

// As(node, d) = A(node, d) Dinv(neighbour d): the first Jacobi step from
// x = 0 (x1 = omega Dinv b) folded into the stencil, so r = b - A x1 =
// b - omega sum_d As(node, d) b(neighbour d) needs no x1 from other nodes
__global__ void __launch_bounds__(256) scaled_stencil_kernel(int M, const int *__restrict__ dim,
                                                            const int *__restrict__ off, int nn,
                                                            const double *__restrict__ A,
                                                            const double *__restrict__ Dinv, double *__restrict__ As)
{
    const int64_t total = static_cast<int64_t>(nn) * 27;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int node = static_cast<int>(t / 27), d = static_cast<int>(t % 27);
        const int m = find_mesh(off, M, node);
        const int N = dim[m], local = node - off[m];
        const int J[3] = {local % N + d % 3 - 1, (local / N) % N + (d / 3) % 3 - 1, local / (N * N) + d / 9 - 1};
        double *o = As + t * 9;
        if (J[0] < 0 || J[0] >= N || J[1] < 0 || J[1] >= N || J[2] < 0 || J[2] >= N) {
            for (int k = 0; k < 9; ++k) o[k] = 0.0;
            continue;
        }
        const int64_t j = off[m] + J[0] + static_cast<int64_t>(N) * (J[1] + static_cast<int64_t>(N) * J[2]);
        const double *a = A + t * 9, *q = Dinv + j * 9;
        for (int c = 0; c < 3; ++c)
            for (int cp = 0; cp < 3; ++cp) o[3 * c + cp] = a[3 * c] * q[cp] + a[3 * c + 1] * q[3 + cp] + a[3 * c + 2] * q[6 + cp];
    }
}

// First smoothing step of a stencil level and its residual in one launch:
// x = omega Dinv b, r = b - omega sum_d As(node, d) b(neighbour d) on free
// DOFs (0 on fixed ones); one 32-lane group per node as in stencil_apply_kernel
__global__ void __launch_bounds__(256) stencil_first_residual_kernel(int M, const int *__restrict__ dim,
                                                                    const int *__restrict__ off, int nn,
                                                                    const double *__restrict__ As,
                                                                    const uint8_t *__restrict__ freed,
                                                                    const double *__restrict__ b,
                                                                    const double *__restrict__ Dinv, double omega,
                                                                    double *__restrict__ x, double *__restrict__ r)
{
    const int node = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 5);
    const int d = threadIdx.x & 31;
    if (node >= nn) return;  // whole 32-lane groups
    const int m = find_mesh(off, M, node);
    const int N = dim[m], local = node - off[m];
    double p[3] = {0.0, 0.0, 0.0};
    if (d < 27) {
        const int J[3] = {local % N + d % 3 - 1, (local / N) % N + (d / 3) % 3 - 1, local / (N * N) + d / 9 - 1};
        if (J[0] >= 0 && J[0] < N && J[1] >= 0 && J[1] < N && J[2] >= 0 && J[2] < N) {
            const int64_t j = off[m] + J[0] + static_cast<int64_t>(N) * (J[1] + static_cast<int64_t>(N) * J[2]);
            const double *a = As + (static_cast<int64_t>(node) * 27 + d) * 9;
            const double b0 = b[j * 3], b1 = b[j * 3 + 1], b2 = b[j * 3 + 2];
#pragma unroll
            for (int c = 0; c < 3; ++c) p[c] = a[3 * c] * b0 + a[3 * c + 1] * b1 + a[3 * c + 2] * b2;
        }
    }
#pragma unroll
    for (int sh = 16; sh > 0; sh >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] += __shfl_xor(p[c], sh, 32);
    if (d != 0) return;
    const double *o = Dinv + static_cast<int64_t>(node) * 9;
    const int64_t k0 = static_cast<int64_t>(node) * 3;
    const double bn[3] = {b[k0], b[k0 + 1], b[k0 + 2]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r[k0 + c] = freed[k0 + c] ? bn[c] - omega * p[c] : 0.0;
        x[k0 + c] = omega * (o[3 * c] * bn[0] + o[3 * c + 1] * bn[1] + o[3 * c + 2] * bn[2]);
    }
}

__global__ void __launch_bounds__(256) stencil_jacobi_kernel(int nn, const double *__restrict__ Dinv,
                                                            const double *__restrict__ r, double *__restrict__ x,
                                                            double omega, int first)
{
    for (int node = blockIdx.x * blockDim.x + threadIdx.x; node < nn; node += gridDim.x * blockDim.x) {
        const double *o = Dinv + static_cast<int64_t>(node) * 9;
        const double *rr = r + static_cast<int64_t>(node) * 3;
        for (int c = 0; c < 3; ++c) {
            const int64_t k = static_cast<int64_t>(node) * 3 + c;
            const double d = o[3 * c] * rr[0] + o[3 * c + 1] * rr[1] + o[3 * c + 2] * rr[2];
            x[k] = (first ? 0.0 : x[k]) + omega * d;
        }
    }
}

// levels >= 1, one 32-lane group per node (the coarse levels have few
// nodes, so a node's 27 neighbour products run in parallel): lane d < 27
// forms A(node, d) x(neighbour d), and a fixed butterfly adds the 27
// products.  kSweep 0: out = r = b - A x on free DOFs.  kSweep 1: one Jacobi
// sweep, out = x + omega Dinv r (out and x distinct buffers).
template <int kSweep>
__global__ void __launch_bounds__(256) stencil_apply_kernel(int M, const int *__restrict__ dim,
                                                           const int *__restrict__ off, int nn,
                                                           const double *__restrict__ A,
                                                           const uint8_t *__restrict__ freed,
                                                           const double *__restrict__ b,
                                                           const double *__restrict__ x,
                                                           const double *__restrict__ Dinv, double omega,
                                                           double *__restrict__ out)
{
    const int node = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 5);
    const int d = threadIdx.x & 31;
    if (node >= nn) return;  // whole 32-lane groups
    const int m = find_mesh(off, M, node);
    const int N = dim[m], local = node - off[m];
    double p[3] = {0.0, 0.0, 0.0};
    if (d < 27) {
        const int J[3] = {local % N + d % 3 - 1, (local / N) % N + (d / 3) % 3 - 1, local / (N * N) + d / 9 - 1};
        if (J[0] >= 0 && J[0] < N && J[1] >= 0 && J[1] < N && J[2] >= 0 && J[2] < N) {
            const int64_t j = off[m] + J[0] + static_cast<int64_t>(N) * (J[1] + static_cast<int64_t>(N) * J[2]);
            const double *a = A + (static_cast<int64_t>(node) * 27 + d) * 9;
            const double x0 = x[j * 3], x1 = x[j * 3 + 1], x2 = x[j * 3 + 2];
#pragma unroll
            for (int c = 0; c < 3; ++c) p[c] = a[3 * c] * x0 + a[3 * c + 1] * x1 + a[3 * c + 2] * x2;
        }
    }
#pragma unroll
    for (int sh = 16; sh > 0; sh >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] += __shfl_xor(p[c], sh, 32);
    if (d != 0) return;
    double r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int64_t k = static_cast<int64_t>(node) * 3 + c;
        r[c] = freed[k] ? b[k] - p[c] : 0.0;
    }
    if (kSweep == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) out[static_cast<int64_t>(node) * 3 + c] = r[c];
    } else {
        const double *o = Dinv + static_cast<int64_t>(node) * 9;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t k = static_cast<int64_t>(node) * 3 + c;
            out[k] = x[k] + omega * (o[3 * c] * r[0] + o[3 * c + 1] * r[1] + o[3 * c + 2] * r[2]);
        }
    }
}

__global__ void __launch_bounds__(256) restrict_stencil_kernel(int M, const int *__restrict__ dim_f,
                                                              const int *__restrict__ off_f,
                                                              const int *__restrict__ dim_c,
                                                              const int *__restrict__ off_c, int nn_c,
                                                              const uint8_t *__restrict__ freed_c,
                                                              const double *__restrict__ rf, double *__restrict__ bc)
{
    // one 32-lane group per coarse node, as restrict_q2_kernel
    const int node = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 5);
    const int s = threadIdx.x & 31;
    if (node >= nn_c) return;  // whole 32-lane groups
    const int m = find_mesh(off_c, M, node);
    const int Nc = dim_c[m], Nf = dim_f[m], local = node - off_c[m];
    const int I[3] = {local % Nc, (local / Nc) % Nc, local / (Nc * Nc)};
    double acc[3] = {0.0, 0.0, 0.0};
    if (s < 27) {
        const int i[3] = {2 * I[0] + s % 3 - 1, 2 * I[1] + (s / 3) % 3 - 1, 2 * I[2] + s / 9 - 1};
        if (i[0] >= 0 && i[0] < Nf && i[1] >= 0 && i[1] < Nf && i[2] >= 0 && i[2] < Nf) {
            const double w = lin_weight(s % 3 - 1) * lin_weight((s / 3) % 3 - 1) * lin_weight(s / 9 - 1);
            const int64_t f = off_f[m] + i[0] + static_cast<int64_t>(Nf) * (i[1] + static_cast<int64_t>(Nf) * i[2]);
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] = w * rf[f * 3 + c];
        }
    }
#pragma unroll
    for (int sh = 16; sh > 0; sh >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += __shfl_xor(acc[c], sh, 32);
    if (s != 0) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int64_t k = static_cast<int64_t>(node) * 3 + c;
        bc[k] = freed_c[k] ? acc[c] : 0.0;
    }
}

__global__ void __launch_bounds__(256) prolong_stencil_kernel(int M, const int *__restrict__ dim_f,
                                                             const int *__restrict__ off_f, int nn_f,
                                                             const int *__restrict__ dim_c,
                                                             const int *__restrict__ off_c,
                                                             const uint8_t *__restrict__ freed_f,
                                                             const double *__restrict__ xc, double *__restrict__ xf)
{
    for (int node = blockIdx.x * blockDim.x + threadIdx.x; node < nn_f; node += gridDim.x * blockDim.x) {
        const int m = find_mesh(off_f, M, node);
        const int Nf = dim_f[m], Nc = dim_c[m], local = node - off_f[m];
        const int i[3] = {local % Nf, (local / Nf) % Nf, local / (Nf * Nf)};
        int p0[3], p1[3];
        double w0[3], w1[3];
        for (int k = 0; k < 3; ++k) {
            p0[k] = i[k] >> 1;
            p1[k] = (i[k] + 1) >> 1;
            w0[k] = (i[k] & 1) ? 0.5 : 1.0;
            w1[k] = (i[k] & 1) ? 0.5 : 0.0;
        }
        double v[3] = {0.0, 0.0, 0.0};
        for (int s = 0; s < 8; ++s) {
            const int b0 = s & 1, b1 = (s >> 1) & 1, b2 = s >> 2;
            const double w = (b0 ? w1[0] : w0[0]) * (b1 ? w1[1] : w0[1]) * (b2 ? w1[2] : w0[2]);
            if (w == 0.0) continue;
            const int64_t P = off_c[m] + (b0 ? p1[0] : p0[0]) +
                              static_cast<int64_t>(Nc) * ((b1 ? p1[1] : p0[1]) + static_cast<int64_t>(Nc) * (b2 ? p1[2] : p0[2]));
            for (int c = 0; c < 3; ++c) v[c] += w * xc[P * 3 + c];
        }
        for (int c = 0; c < 3; ++c) {
            const int64_t k = static_cast<int64_t>(node) * 3 + c;
            if (freed_f[k]) xf[k] += v[c];
        }
    }
}

__global__ void __launch_bounds__(256) coarse_apply_kernel(int M, const int *__restrict__ dim,
                                                          const int *__restrict__ off, int nn,
                                                          const double *__restrict__ inv, int stride,
                                                          const double *__restrict__ b, double *__restrict__ x)
{
    // one wavefront per coarse DOF: its row of the dense inverse against b,
    // lanes over the row, a fixed butterfly for the sum
    const int k = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= 3 * nn) return;  // whole wavefronts
    const int node = k / 3;
    const int m = find_mesh(off, M, node);
    const int N = dim[m], nd = 3 * N * N * N;
    const int r = k - 3 * off[m];
    const double *row = inv + static_cast<int64_t>(m) * stride + static_cast<int64_t>(r) * nd;
    const double *bm = b + static_cast<int64_t>(off[m]) * 3;
    double acc = 0.0;
    for (int j = lane; j < nd; j += 64) acc += row[j] * bm[j];
#pragma unroll
    for (int sh = 32; sh > 0; sh >>= 1) acc += __shfl_xor(acc, sh, 64);
    if (lane == 0) x[k] = acc;
}

unsigned grid1(int64_t n) { return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384))); }

}  // namespace

int mg_setup(EquationsSet &s, int precision)
{
    Multigrid &g = *s.mg;
    Runtime &rt = runtime();
    const int M = s.n_meshes;
    const unsigned eb = static_cast<unsigned>(std::min(s.n_elements, rt.num_cus * (precision == OC_FP32 ? 3 : 2)));
    if (precision == OC_FP64)
        q1_galerkin_element_kernel<double><<<eb, 256, 0, rt.stream>>>(s.n_elements, s.kbuf64.ptr, s.conn.ptr,
                                                                     s.n_nodes, s.free_mask.ptr, g.ke1.ptr);
    else
        q1_galerkin_element_kernel<float><<<eb, 256, 0, rt.stream>>>(s.n_elements, s.kbuf32.ptr, s.conn.ptr,
                                                                    s.n_nodes, s.free_mask.ptr, g.ke1.ptr);
    OC_HIP_CHECK(hipGetLastError());
    if (precision == OC_FP32 && s.uu_bf.ptr) OC_CHECK(launch_uu_pack(s));
    MgLevel &L1 = g.levels[1];
    q1_stencil_kernel<<<grid1(static_cast<int64_t>(L1.nn) * 27), 256, 0, rt.stream>>>(
        M, g.n.ptr, g.e_off.ptr, L1.off.ptr, L1.nn, g.ke1.ptr, L1.A.ptr);
    coincident_q2_kernel<<<grid1(L1.nn), 256, 0, rt.stream>>>(M, g.n.ptr, s.q_off.ptr, s.n_nodes, L1.off.ptr, L1.nn,
                                                              s.free_mask.ptr, L1.freed.ptr);
    level_freed_kernel<<<grid1(L1.nn), 256, 0, rt.stream>>>(L1.nn, L1.A.ptr, L1.freed.ptr);
    level_mask_kernel<<<grid1(static_cast<int64_t>(L1.nn) * 27), 256, 0, rt.stream>>>(
        M, L1.dim.ptr, L1.off.ptr, L1.nn, L1.A.ptr, L1.freed.ptr, L1.Dinv.ptr);
    OC_HIP_CHECK(hipGetLastError());
    if (L1.As.ptr)
        scaled_stencil_kernel<<<grid1(static_cast<int64_t>(L1.nn) * 27), 256, 0, rt.stream>>>(
            M, L1.dim.ptr, L1.off.ptr, L1.nn, L1.A.ptr, L1.Dinv.ptr, L1.As.ptr);
    const int nl = static_cast<int>(g.levels.size());
    for (int l = 1; l + 1 < nl; ++l) {
        MgLevel &F = g.levels[l], &C = g.levels[l + 1];
        galerkin_coarsen_kernel<<<static_cast<unsigned>((static_cast<int64_t>(C.nn) * 27 + kGalPairs - 1) / kGalPairs),
                                  27 * kGalPairs, 0, rt.stream>>>(
            M, F.dim.ptr, F.off.ptr, F.A.ptr, F.freed.ptr, C.dim.ptr, C.off.ptr, C.nn, C.A.ptr);
        coincident_kernel<<<grid1(C.nn), 256, 0, rt.stream>>>(M, F.dim.ptr, F.off.ptr, F.freed.ptr, C.dim.ptr,
                                                             C.off.ptr, C.nn, C.freed.ptr);
        level_freed_kernel<<<grid1(C.nn), 256, 0, rt.stream>>>(C.nn, C.A.ptr, C.freed.ptr);
        level_mask_kernel<<<grid1(static_cast<int64_t>(C.nn) * 27), 256, 0, rt.stream>>>(
            M, C.dim.ptr, C.off.ptr, C.nn, C.A.ptr, C.freed.ptr, C.Dinv.ptr);
        if (C.As.ptr)
            scaled_stencil_kernel<<<grid1(static_cast<int64_t>(C.nn) * 27), 256, 0, rt.stream>>>(
                M, C.dim.ptr, C.off.ptr, C.nn, C.A.ptr, C.Dinv.ptr, C.As.ptr);
        OC_HIP_CHECK(hipGetLastError());
    }
    MgLevel &Lc = g.levels[nl - 1];
    coarse_inverse_kernel<<<static_cast<unsigned>(M), 256, 0, rt.stream>>>(M, Lc.dim.ptr, Lc.off.ptr, Lc.A.ptr,
                                                                         g.coarse_stride, g.coarse_inv.ptr);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}


static unsigned group_grid(int nn) { return static_cast<unsigned>((static_cast<int64_t>(nn) * 32 + 255) / 256); }

// one Jacobi sweep of a stencil level: x <- x + omega Dinv (b - A x), through
// the level's r buffer (the two buffers trade places)
static int stencil_sweep(int M, MgLevel &F, double omega)
{
    stencil_apply_kernel<1><<<group_grid(F.nn), 256, 0, runtime().stream>>>(
        M, F.dim.ptr, F.off.ptr, F.nn, F.A.ptr, F.freed.ptr, F.b.ptr, F.x.ptr, F.Dinv.ptr, omega, F.r.ptr);
    OC_HIP_CHECK(hipGetLastError());
    std::swap(F.x, F.r);
    return OC_NO_ERROR;
}

// Launch shapes of the level-0 node kernels: lanes per node (below) x
// neighbours per lane in flight (1, 2, 2, 4, 1; the cases of launch_uu_node);
// the autotuner picks one (EquationsSet::tuned_smoother).
constexpr int kNodeVariants = 5;
constexpr int kNodeVariantLanes[kNodeVariants] = {32, 32, 16, 16, 64};

template <typename VT, int kSweep>
static int launch_uu_node(int variant, int nq, const EquationsSet &s, const VT *vals, const double *b,
                          const double *x, const double *inv_u, double omega, double *out)
{
    Runtime &rt = runtime();
    if (variant < 0 || variant >= kNodeVariants) variant = 0;
    const unsigned grid =
        static_cast<unsigned>((static_cast<int64_t>(nq) * kNodeVariantLanes[variant] + 255) / 256);
#define OC_UU_NODE(L, U)                                                                                       \
    uu_node_residual_kernel<VT, L, kSweep, U><<<grid, 256, 0, rt.stream>>>(nq, s.row_ptr.ptr, s.pstart.ptr,   \
                                                                         s.col_idx.ptr, vals, s.free_mask.ptr, \
                                                                         b, x, inv_u, omega, out)
    switch (variant) {
    case 1: OC_UU_NODE(32, 2); break;
    case 2: OC_UU_NODE(16, 2); break;
    case 3: OC_UU_NODE(16, 4); break;
    case 4: OC_UU_NODE(64, 1); break;
    default: OC_UU_NODE(32, 1); break;
    }
#undef OC_UU_NODE
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

// The packed bf16 level-0 kernels' shapes: lanes per node x quads of
// neighbours per lane in flight (variants as kNodeVariants).
constexpr int kBf16VariantLanes[kNodeVariants] = {16, 8, 32, 8, 16};

template <int kSweep>
// x4: the FP32 node quadruples of x when the caller has them (nullptr: packed
// here, into the first of the set's two quadruple buffers); out4: where a
// sweep also writes its result's quadruples (nullptr: not written)
static int launch_uu_bf16(int variant, const EquationsSet &s, const double *b, const double *x, const double *inv_u,
                          double omega, double *out, const float *x4_in = nullptr, float *out4 = nullptr)
{
    Runtime &rt = runtime();
    if (variant < 0 || variant >= kNodeVariants) variant = 0;
    const int nq = s.n_nodes;
    const unsigned grid = static_cast<unsigned>((static_cast<int64_t>(nq) * kBf16VariantLanes[variant] + 255) / 256);
    const float4 *x4 = reinterpret_cast<const float4 *>(x4_in);
    if (!x4) {
        pack_x4(nq, x, reinterpret_cast<float4 *>(s.uu_x4.ptr));
        x4 = reinterpret_cast<const float4 *>(s.uu_x4.ptr);
    }
#define OC_UU_BF16(L, Q)                                                                                          \
    uu_node_bf16_kernel<L, kSweep, Q><<<grid, 256, 0, rt.stream>>>(nq, s.uu_off.ptr, s.uu_col.ptr, s.uu_bf.ptr, \
                                                                 s.free_mask.ptr, b, x4, x, inv_u, omega, out, out4)
    switch (variant) {
    case 1: OC_UU_BF16(8, 1); break;
    case 2: OC_UU_BF16(32, 1); break;
    case 3: OC_UU_BF16(8, 2); break;
    case 4: OC_UU_BF16(16, 2); break;
    default: OC_UU_BF16(16, 1); break;
    }
#undef OC_UU_BF16
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

void pack_x4(int nq, const double *x, float4 *x4)
{
    pack_x4_kernel<<<static_cast<unsigned>((nq + 255) / 256), 256, 0, runtime().stream>>>(nq, x, x4);
}

int launch_uu_pack(EquationsSet &s)
{
    uu_pack_kernel<<<static_cast<unsigned>((static_cast<int64_t>(s.n_nodes) * kPackLanes + 255) / 256), 256, 0,
                     runtime().stream>>>(s.n_nodes, s.row_ptr.ptr, s.pstart.ptr, s.values_sp.ptr, s.uu_off.ptr,
                                         s.uu_bf.ptr);
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

int prepare_uu_bf16(EquationsSet &s)
{
    if (s.uu_bf.ptr) return OC_NO_ERROR;
    if (s.node_rows != 1 || !s.pstart.ptr) return set_error(OC_ERR_UNSUPPORTED, "the packed bf16 block needs node rows");
    Runtime &rt = runtime();
    const int nq = s.n_nodes;
    DevBuf<int64_t> degp;
    OC_CHECK(degp.alloc(nq));
    uu_degp_kernel<<<static_cast<unsigned>((nq + 255) / 256), 256, 0, rt.stream>>>(nq, s.row_ptr.ptr, s.pstart.ptr,
                                                                                   degp.ptr);
    OC_HIP_CHECK(hipGetLastError());
    std::vector<int64_t> off(static_cast<size_t>(nq) + 1, 0);
    OC_HIP_CHECK(hipMemcpyAsync(off.data() + 1, degp.ptr, sizeof(int64_t) * nq, hipMemcpyDeviceToHost, rt.stream));
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    for (int i = 0; i < nq; ++i) off[i + 1] += off[i];
    const int64_t total = off[nq];
    OC_CHECK(s.uu_off.upload(off.data(), off.size()));
    OC_CHECK(s.uu_col.alloc(static_cast<size_t>(total)));
    OC_CHECK(s.uu_bf.alloc(static_cast<size_t>(total) * 9));
    OC_CHECK(s.uu_x4.alloc(static_cast<size_t>(nq) * 8));  // two quadruple buffers
    uu_cols_kernel<<<static_cast<unsigned>((static_cast<int64_t>(nq) * kPackLanes + 255) / 256), 256, 0, rt.stream>>>(
        nq, s.row_ptr.ptr, s.pstart.ptr, s.col_idx.ptr, s.uu_off.ptr, s.uu_col.ptr);
    OC_HIP_CHECK(hipGetLastError());
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    return OC_NO_ERROR;
}

int launch_smoother_residual(EquationsSet &s, int precision, int variant, const double *b, const double *x,
                             double *r)
{
    if (s.node_rows != 1) return set_error(OC_ERR_UNSUPPORTED, "the node-wise smoother needs node rows");
    // every array the kernel reads must exist: the free-DOF mask only does once
    // boundary conditions are set or a solve has prepared the set
    const void *vals = precision == OC_FP32 ? (s.uu_bf.ptr ? static_cast<const void *>(s.uu_bf.ptr)
                                                           : static_cast<const void *>(s.values_sp.ptr))
                                            : static_cast<const void *>(s.values.ptr);
    if (!s.free_mask.ptr || !s.pstart.ptr || !vals || !b || !x || !r)
        return set_error(OC_ERR_INVALID_ARGUMENT, "smoother residual: free-DOF mask, pressure starts or values missing");
    if (precision == OC_FP32 && s.uu_bf.ptr) return launch_uu_bf16<0>(variant, s, b, x, nullptr, 0.0, r);
    if (precision == OC_FP32) return launch_uu_node<float, 0>(variant, s.n_nodes, s, s.values_sp.ptr, b, x, nullptr, 0.0, r);
    return launch_uu_node<double, 0>(variant, s.n_nodes, s, s.values.ptr, b, x, nullptr, 0.0, r);
}

// z (displacement part, component-major) = V-cycle applied to t
int mg_apply(EquationsSet &s, int precision, const double *t, double *z)
{
    Multigrid &g = *s.mg;
    Runtime &rt = runtime();
    const int M = s.n_meshes, nq = s.n_nodes, nu = 3 * nq, nl = static_cast<int>(g.levels.size());
    const double *inv_u = s.precond.ptr;  // nodal block inverses of the set's block preconditioner
    const unsigned gu = static_cast<unsigned>((static_cast<int64_t>(nu) * 32 + 255) / 256);
    const bool bf = precision == OC_FP32 && s.node_rows == 1 && s.uu_bf.ptr;
    // default shape: 16 lanes x 2 neighbours in flight for FP64 values (6%
    // faster than 32 x 1 at 32^3), 32 x 1 for bf16 and FP32
    const int variant = s.tuned_smoother >= 0 ? s.tuned_smoother : (precision == OC_FP64 ? 2 : 0);
    // with the packed bf16 values the kernels that produce the iterate also
    // write its FP32 node quadruples, into the two halves of uu_x4 in turn
    // (a sweep gathers one half and writes the other); q4 is the current one
    float *q4 = bf ? s.uu_x4.ptr : nullptr, *q4_other = bf ? s.uu_x4.ptr + 4 * static_cast<int64_t>(nq) : nullptr;
    auto residual0 = [&](const double *b, const double *x, double *r) -> int {
        if (bf)
            OC_CHECK(launch_uu_bf16<0>(variant, s, b, x, nullptr, 0.0, r, q4));
        else if (s.node_rows == 1 && precision == OC_FP32)
            OC_CHECK((launch_uu_node<float, 0>(variant, nq, s, s.values_sp.ptr, b, x, nullptr, 0.0, r)));
        else if (s.node_rows == 1)
            OC_CHECK((launch_uu_node<double, 0>(variant, nq, s, s.values.ptr, b, x, nullptr, 0.0, r)));
        else if (precision == OC_FP32)
            uu_residual_kernel<float, 32><<<gu, 256, 0, rt.stream>>>(nu, s.row_ptr.ptr, s.pstart.ptr, s.col_idx.ptr,
                                                                    s.values_sp.ptr, s.free_mask.ptr, b, x, r);
        else
            uu_residual_kernel<double, 32><<<gu, 256, 0, rt.stream>>>(nu, s.row_ptr.ptr, s.pstart.ptr, s.col_idx.ptr,
                                                                      s.values.ptr, s.free_mask.ptr, b, x, r);
        OC_HIP_CHECK(hipGetLastError());
        return OC_NO_ERROR;
    };
    // level 0 pre-smoothing from x = 0.  With node rows a sweep is one fused
    // kernel writing the other of z and x0; the first Jacobi step picks the
    // buffer that leaves the last post-smoothing sweep in z.
    const bool fused = s.node_rows == 1;
    const int post = g.post;
    double *cur = z, *other = g.x0.ptr;
    if (fused && ((std::max(g.pre - 1, 0) + post) & 1)) std::swap(cur, other);
    auto sweep0 = [&]() -> int {
        if (bf)
            OC_CHECK(launch_uu_bf16<1>(variant, s, t, cur, inv_u, g.omega, other, q4, q4_other));
        else if (precision == OC_FP32)
            OC_CHECK((launch_uu_node<float, 1>(variant, nq, s, s.values_sp.ptr, t, cur, inv_u, g.omega, other)));
        else
            OC_CHECK((launch_uu_node<double, 1>(variant, nq, s, s.values.ptr, t, cur, inv_u, g.omega, other)));
        std::swap(cur, other);
        std::swap(q4, q4_other);
        return OC_NO_ERROR;
    };
    q2_jacobi_kernel<<<grid1(nq), 256, 0, rt.stream>>>(nq, inv_u, t, cur, g.omega, 1, q4);
    for (int k = 1; k < g.pre; ++k) {
        if (fused) {
            OC_CHECK(sweep0());
            continue;
        }
        OC_CHECK(residual0(t, cur, g.r0.ptr));
        q2_jacobi_kernel<<<grid1(nq), 256, 0, rt.stream>>>(nq, inv_u, g.r0.ptr, cur, g.omega, 0, q4);
    }
    OC_CHECK(residual0(t, cur, g.r0.ptr));
    MgLevel &L1 = g.levels[1];
    restrict_q2_kernel<<<group_grid(L1.nn), 256, 0, rt.stream>>>(M, g.n.ptr, s.q_off.ptr, nq, L1.off.ptr, L1.nn,
                                                            L1.freed.ptr, g.r0.ptr, L1.b.ptr);
    OC_HIP_CHECK(hipGetLastError());
    // down the stencil levels
    for (int l = 1; l + 1 < nl; ++l) {
        MgLevel &F = g.levels[l], &C = g.levels[l + 1];
        if (g.stencil_pre <= 1 && F.As.ptr) {
            stencil_first_residual_kernel<<<group_grid(F.nn), 256, 0, rt.stream>>>(
                M, F.dim.ptr, F.off.ptr, F.nn, F.As.ptr, F.freed.ptr, F.b.ptr, F.Dinv.ptr, g.omega, F.x.ptr, F.r.ptr);
        } else {
            stencil_jacobi_kernel<<<grid1(F.nn), 256, 0, rt.stream>>>(F.nn, F.Dinv.ptr, F.b.ptr, F.x.ptr, g.omega, 1);
            for (int k = 1; k < g.stencil_pre; ++k) OC_CHECK(stencil_sweep(M, F, g.omega));
            stencil_apply_kernel<0><<<group_grid(F.nn), 256, 0, rt.stream>>>(M, F.dim.ptr, F.off.ptr, F.nn, F.A.ptr,
                                                                            F.freed.ptr, F.b.ptr, F.x.ptr, nullptr,
                                                                            0.0, F.r.ptr);
        }
        restrict_stencil_kernel<<<group_grid(C.nn), 256, 0, rt.stream>>>(M, F.dim.ptr, F.off.ptr, C.dim.ptr, C.off.ptr,
                                                                   C.nn, C.freed.ptr, F.r.ptr, C.b.ptr);
        OC_HIP_CHECK(hipGetLastError());
    }
    MgLevel &Lc = g.levels[nl - 1];
    coarse_apply_kernel<<<static_cast<unsigned>((3 * static_cast<int64_t>(Lc.nn) * 64 + 255) / 256), 256, 0, rt.stream>>>(
        M, Lc.dim.ptr, Lc.off.ptr, Lc.nn, g.coarse_inv.ptr, g.coarse_stride, Lc.b.ptr, Lc.x.ptr);
    OC_HIP_CHECK(hipGetLastError());
    // up
    for (int l = nl - 2; l >= 1; --l) {
        MgLevel &F = g.levels[l], &C = g.levels[l + 1];
        prolong_stencil_kernel<<<grid1(F.nn), 256, 0, rt.stream>>>(M, F.dim.ptr, F.off.ptr, F.nn, C.dim.ptr,
                                                                  C.off.ptr, F.freed.ptr, C.x.ptr, F.x.ptr);
        for (int k = 0; k < g.stencil_post; ++k) OC_CHECK(stencil_sweep(M, F, g.omega));
        OC_HIP_CHECK(hipGetLastError());
    }
    prolong_q2_kernel<<<grid1(nq), 256, 0, rt.stream>>>(M, g.n.ptr, s.q_off.ptr, nq, L1.off.ptr, s.free_mask.ptr,
                                                        L1.x.ptr, cur, q4);
    OC_HIP_CHECK(hipGetLastError());
    for (int k = 0; k < post; ++k) {
        if (fused) {
            OC_CHECK(sweep0());
            continue;
        }
        OC_CHECK(residual0(t, cur, g.r0.ptr));
        q2_jacobi_kernel<<<grid1(nq), 256, 0, rt.stream>>>(nq, inv_u, g.r0.ptr, cur, g.omega, 0, q4);
    }
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

}  // namespace oc

using namespace oc;

extern "C" int oc_multigrid_structure_set(oc_equations_set h, const int *n_per_mesh)
{
    OC_CHECK(require_runtime());
    if (!h) return set_error(OC_ERR_INVALID_HANDLE, "equations set handle is NULL");
    EquationsSet &s = h->s;
    if (s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN || !n_per_mesh)
        return set_error(OC_ERR_INVALID_ARGUMENT, "multigrid structure: a Mooney-Rivlin set and n per mesh");
    const int M = s.n_meshes;
    // the connectivity must be that of regular generated n^3 meshes
    std::vector<int32_t> cq(static_cast<size_t>(s.n_elements) * 27), cl(static_cast<size_t>(s.n_elements) * 8);
    OC_HIP_CHECK(hipMemcpy(cq.data(), s.conn.ptr, cq.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    OC_HIP_CHECK(hipMemcpy(cl.data(), s.conn_l.ptr, cl.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int m = 0; m < M; ++m) {
        const int64_t n = n_per_mesh[m];
        const int64_t T = 2 * n + 1, N = n + 1;
        if (n < 1 || n * n * n != s.h_e_off[m + 1] - s.h_e_off[m] || T * T * T != s.h_q_off[m + 1] - s.h_q_off[m] ||
            N * N * N != s.h_l_off[m + 1] - s.h_l_off[m])
            return set_error(OC_ERR_INVALID_ARGUMENT, "mesh %d is not a regular mesh of %lld^3 elements", m,
                             static_cast<long long>(n));
        for (int64_t e = 0; e < n * n * n; ++e) {
            const int64_t ex = e % n, ey = (e / n) % n, ez = e / (n * n);
            const int64_t ge = s.h_e_off[m] + e;
            for (int a = 0; a < 27; ++a) {
                const int64_t q = s.h_q_off[m] + (2 * ex + a % 3) + T * ((2 * ey + (a / 3) % 3) + T * (2 * ez + a / 9));
                if (cq[ge * 27 + a] != q)
                    return set_error(OC_ERR_INVALID_ARGUMENT, "mesh %d element %lld: not the regular node numbering",
                                     m, static_cast<long long>(e));
            }
            for (int k = 0; k < 8; ++k) {
                const int64_t p = s.h_l_off[m] + (ex + (k & 1)) + N * ((ey + ((k >> 1) & 1)) + N * (ez + (k >> 2)));
                if (cl[ge * 8 + k] != p)
                    return set_error(OC_ERR_INVALID_ARGUMENT, "mesh %d element %lld: not the regular pressure numbering",
                                     m, static_cast<long long>(e));
            }
        }
    }
    auto g = new (std::nothrow) Multigrid();
    if (!g) return set_error(OC_ERR_ALLOCATION, "out of host memory");
    std::unique_ptr<Multigrid> guard(g);
    g->h_n.assign(n_per_mesh, n_per_mesh + M);
    OC_CHECK(g->n.upload(g->h_n.data(), M));
    OC_CHECK(g->e_off.upload(s.h_e_off.data(), s.h_e_off.size()));
    // levels: [0] placeholder, [1] trilinear (n+1)^3, then coarsening until every dim <= 3
    g->levels.resize(2);
    std::vector<int32_t> dims(M);
    for (int m = 0; m < M; ++m) dims[m] = g->h_n[m] + 1;
    while (true) {
        MgLevel &L = g->levels.back();
        L.h_dim = dims;
        L.h_off.assign(M + 1, 0);
        for (int m = 0; m < M; ++m) L.h_off[m + 1] = L.h_off[m] + dims[m] * dims[m] * dims[m];
        L.nn = L.h_off[M];
        OC_CHECK(L.dim.upload(L.h_dim.data(), M));
        OC_CHECK(L.off.upload(L.h_off.data(), M + 1));
        OC_CHECK(L.A.alloc(static_cast<size_t>(L.nn) * 27 * 9));
        OC_CHECK(L.Dinv.alloc(static_cast<size_t>(L.nn) * 9));
        OC_CHECK(L.freed.alloc(static_cast<size_t>(L.nn) * 3));
        OC_CHECK(L.x.alloc(static_cast<size_t>(L.nn) * 3));
        OC_CHECK(L.b.alloc(static_cast<size_t>(L.nn) * 3));
        OC_CHECK(L.r.alloc(static_cast<size_t>(L.nn) * 3));
        if (*std::max_element(dims.begin(), dims.end()) <= 3) break;
        OC_CHECK(L.As.alloc(static_cast<size_t>(L.nn) * 27 * 9));
        for (int m = 0; m < M; ++m) dims[m] = (dims[m] - 1 + 1) / 2 + 1;  // n -> ceil(n/2)
        g->levels.emplace_back();
    }
    int maxnd = 0;
    for (int m = 0; m < M; ++m) maxnd = std::max(maxnd, 3 * dims[m] * dims[m] * dims[m]);
    if (maxnd > 81) return set_error(OC_ERR_UNSUPPORTED, "coarsest multigrid level too large");
    g->coarse_stride = maxnd * maxnd;
    OC_CHECK(g->coarse_inv.alloc(static_cast<size_t>(M) * g->coarse_stride));
    OC_CHECK(g->ke1.alloc(static_cast<size_t>(s.n_elements) * 576));
    OC_CHECK(g->r0.alloc(static_cast<size_t>(3) * s.n_nodes));
    OC_CHECK(g->t0.alloc(static_cast<size_t>(3) * s.n_nodes));
    OC_CHECK(g->x0.alloc(static_cast<size_t>(3) * s.n_nodes));
    delete s.mg;
    s.mg = guard.release();
    if (s.step_graph) {  // recorded without the multigrid setup
        (void)hipGraphExecDestroy(s.step_graph);
        s.step_graph = nullptr;
    }
    return OC_NO_ERROR;
}

extern "C" int oc_multigrid_options_set(oc_equations_set h, int pre_sweeps, int post_sweeps, double omega)
{
    OC_CHECK(require_runtime());
    if (!h) return set_error(OC_ERR_INVALID_HANDLE, "equations set handle is NULL");
    if (!h->s.mg) return set_error(OC_ERR_INVALID_ARGUMENT, "no multigrid structure (oc_multigrid_structure_set)");
    if (pre_sweeps < 1 || post_sweeps < 0 || !(omega > 0.0 && omega < 2.0))
        return set_error(OC_ERR_INVALID_ARGUMENT, "multigrid options: pre >= 1, post >= 0, 0 < omega < 2");
    h->s.mg->pre = pre_sweeps;
    h->s.mg->post = post_sweeps;
    h->s.mg->omega = omega;
    return OC_NO_ERROR;
}
