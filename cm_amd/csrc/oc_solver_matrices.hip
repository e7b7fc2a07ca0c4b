// oc_solver_matrices.hip -- the solver system of an equations set on the
// device: the equations matrix (Laplace) or the Jacobian (Mooney-Rivlin)
// restricted to the free DOFs, and the matching vector.
//
// The reference builds its solver system from the equations matrices through
// the solver mapping (SOLVER_MATRICES_STATIC_ASSEMBLE, solver_routines.f90:
// 13745-14771; SOLVER_MATRIX_JACOBIAN_MATRIX_ADD, solver_matrices_routines.f90:
// 1258-1483, and its equations-matrix counterpart):
//  * an equations row or column enters the solver system only when its DOF is
//    free (INCLUDE_ROW / INCLUDE_COLUMN, solver_mapping_routines.f90:434-520);
//  * solver rows and columns number the free DOFs in ascending order (one
//    equations set on one rank);
//  * the coupling coefficients of a single equations set are 1 and the solver
//    matrix starts at zero, so each solver entry is 0 + J(r, c);
//  * the solver vector of a linear set is the RHS with the fixed columns moved
//    over: for solver row r, in ascending column order, RHS(r) at c = r and
//    -K(r, c) u_c for every fixed column c, where u_c is the DOF's current
//    value and terms with |u_c| or |K(r, c)| below ZERO_TOLERANCE are skipped
//    (solver_routines.f90:14380-14620);
//  * the solver vector of a nonlinear set is the residual of the free rows.
//
// Layout: index[d] is the solver number of DOF d (-1 when fixed), rows[i] the
// equations row of solver row i.  One wavefront per solver row compacts the
// row: lane k takes entry k of a 64-entry chunk and a ballot of the free
// flags gives every free entry its place, so the solver CSR is written once,
// in order, with no scan pass over the nonzeros.
#include "oc_internal.h"

#include <algorithm>

namespace oc {

namespace {

constexpr double kZeroTolerance = 5.0 * 2.220446049250313e-16;  // ZERO_TOLERANCE, constants.f90:68
constexpr int kRowsPerBlock = 4;                                // one wavefront per row

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// free columns of every solver row
__global__ void __launch_bounds__(64 * kRowsPerBlock)
solver_count_kernel(int n, const int *__restrict__ rows, const int64_t *__restrict__ row_ptr,
                    const int *__restrict__ col_idx, const int *__restrict__ index, int *__restrict__ counts)
{
    const int i = blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;  // whole wavefronts
    const int r = rows[i];
    int c = 0;
    for (int64_t k = row_ptr[r] + lane; k < row_ptr[r + 1]; k += 64) c += index[col_idx[k]] >= 0;
    c = wave_sum(c);
    if (lane == 0) counts[i] = c;
}

// solver row i = equations row rows[i] without its fixed columns; with
// kStructure the solver column numbers are written as well.  A wavefront
// takes eight 64-entry chunks per pass (one pass for rows of up to 512
// entries) and issues their column and value loads before the solver-number
// lookups, so the row's loads overlap.
template <typename T, bool kStructure>
__global__ void __launch_bounds__(64 * kRowsPerBlock)
solver_matrix_kernel(int n, const int *__restrict__ rows, const int64_t *__restrict__ row_ptr,
                     const int *__restrict__ col_idx, const int *__restrict__ index,
                     const int64_t *__restrict__ s_row_ptr, const T *__restrict__ values, T *__restrict__ s_values,
                     int *__restrict__ s_col_idx)
{
    constexpr int kChunks = 8;
    const int i = blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;  // whole wavefronts
    const int r = rows[i];
    const int64_t k1 = row_ptr[r + 1];
    int64_t out = s_row_ptr[i];
    const uint64_t below = (uint64_t(1) << lane) - 1;
    for (int64_t k0 = row_ptr[r]; k0 < k1; k0 += 64 * kChunks) {
        int c[kChunks];
        T v[kChunks];
#pragma unroll
        for (int u = 0; u < kChunks; ++u) {
            const int64_t k = k0 + 64 * u + lane;
            c[u] = -1;
            if (k < k1) {
                c[u] = col_idx[k];
                v[u] = values[k];
            }
        }
        int sc[kChunks];
#pragma unroll
        for (int u = 0; u < kChunks; ++u) sc[u] = c[u] >= 0 ? index[c[u]] : -1;
#pragma unroll
        for (int u = 0; u < kChunks; ++u) {
            const uint64_t m = __ballot(sc[u] >= 0);
            if (sc[u] >= 0) {
                const int64_t p = out + __popcll(m & below);
                s_values[p] = T(0) + v[u];  // the solver matrix starts at zero
                if (kStructure) s_col_idx[p] = sc[u];
            }
            out += __popcll(m);
        }
    }
}

// Solver vector: with `lift` (linear sets) RHS(r) and the fixed columns in
// ascending column order, one thread per solver row; otherwise the residual.
__global__ void __launch_bounds__(256)
solver_vector_kernel(int n, const int *__restrict__ rows, const int64_t *__restrict__ row_ptr,
                     const int *__restrict__ col_idx, const int *__restrict__ index,
                     const double *__restrict__ values, const double *__restrict__ rhs,
                     const double *__restrict__ dofs, bool lift, double *__restrict__ out)
{
#pragma clang fp contract(off)  // each product rounded before its add, as the reference
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = rows[i];
    double acc = 0.0;
    if (!lift) {
        acc = acc + rhs[r];
    } else {
        for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            const int c = col_idx[k];
            if (c == r) {
                acc = acc + rhs[r];
            } else if (index[c] < 0) {
                const double u = dofs[c], K = values[k];
                if (fabs(u) >= kZeroTolerance && fabs(K) >= kZeroTolerance) acc = acc + (-1.0 * K) * u;
            }
        }
    }
    out[i] = acc;
}

}  // namespace

// Free-DOF numbering and the solver CSR structure, rebuilt when the
// boundary conditions change.
static int solver_structure(EquationsSet &s)
{
    SolverSystem &S = s.solver;
    if (S.built && S.bc_version == s.bc_version) return OC_NO_ERROR;
    Runtime &rt = runtime();
    std::vector<uint8_t> mask(static_cast<size_t>(s.n_rows), 1);
    if (s.n_bc > 0) {
        OC_HIP_CHECK(hipMemcpyAsync(mask.data(), s.free_mask.ptr, mask.size(), hipMemcpyDeviceToHost, rt.stream));
        OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    }
    std::vector<int32_t> index(mask.size()), rows;
    rows.reserve(mask.size());
    for (int d = 0; d < s.n_rows; ++d) {
        index[d] = mask[d] ? static_cast<int32_t>(rows.size()) : -1;
        if (mask[d]) rows.push_back(d);
    }
    S.n = static_cast<int>(rows.size());
    OC_CHECK(S.index.upload(index.data(), index.size()));
    OC_CHECK(S.rows.upload(rows.data(), rows.size()));
    std::vector<int64_t> ptr(static_cast<size_t>(S.n) + 1, 0);
    if (S.n > 0) {
        DevBuf<int> counts;
        OC_CHECK(counts.alloc(static_cast<size_t>(S.n)));
        const unsigned blocks = static_cast<unsigned>((S.n + kRowsPerBlock - 1) / kRowsPerBlock);
        solver_count_kernel<<<blocks, 64 * kRowsPerBlock, 0, rt.stream>>>(S.n, S.rows.ptr, s.row_ptr.ptr,
                                                                          s.col_idx.ptr, S.index.ptr, counts.ptr);
        OC_HIP_CHECK(hipGetLastError());
        std::vector<int> h(static_cast<size_t>(S.n));
        OC_HIP_CHECK(hipMemcpyAsync(h.data(), counts.ptr, h.size() * sizeof(int), hipMemcpyDeviceToHost, rt.stream));
        OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
        for (int i = 0; i < S.n; ++i) ptr[i + 1] = ptr[i] + h[i];
    }
    S.nnz = ptr.back();
    OC_CHECK(S.row_ptr.upload(ptr.data(), ptr.size()));
    OC_CHECK(S.col_idx.alloc(static_cast<size_t>(S.nnz)));
    S.values.release();
    S.values_sp.release();
    OC_CHECK(S.vector.alloc(static_cast<size_t>(S.n)));
    S.columns_written = false;
    S.built = true;
    S.bc_version = s.bc_version;
    return OC_NO_ERROR;
}

template <typename T>
static int solver_values(EquationsSet &s, const T *values, DevBuf<T> &out)
{
    SolverSystem &S = s.solver;
    if (!out.ptr) OC_CHECK(out.alloc(static_cast<size_t>(S.nnz)));
    if (S.n == 0) return OC_NO_ERROR;
    Runtime &rt = runtime();
    const unsigned blocks = static_cast<unsigned>((S.n + kRowsPerBlock - 1) / kRowsPerBlock);
    if (!S.columns_written) {
        solver_matrix_kernel<T, true><<<blocks, 64 * kRowsPerBlock, 0, rt.stream>>>(
            S.n, S.rows.ptr, s.row_ptr.ptr, s.col_idx.ptr, S.index.ptr, S.row_ptr.ptr, values, out.ptr, S.col_idx.ptr);
        S.columns_written = true;
    } else {
        solver_matrix_kernel<T, false><<<blocks, 64 * kRowsPerBlock, 0, rt.stream>>>(
            S.n, S.rows.ptr, s.row_ptr.ptr, s.col_idx.ptr, S.index.ptr, S.row_ptr.ptr, values, out.ptr, nullptr);
    }
    OC_HIP_CHECK(hipGetLastError());
    return OC_NO_ERROR;
}

int solver_matrices_assemble(EquationsSet &s, int precision)
{
    const bool linear = s.cls == OC_EQUATIONS_SET_LAPLACE;
    if (precision != OC_FP64 && precision != OC_FP32)
        return set_error(OC_ERR_INVALID_ARGUMENT, "precision must be OC_FP64 or OC_FP32");
    if (linear && precision != OC_FP64)
        return set_error(OC_ERR_INVALID_ARGUMENT, "a linear equations matrix is FP64 only");
    if (precision == OC_FP64 && !s.values.ptr)
        return set_error(OC_ERR_INVALID_ARGUMENT, "no FP64 matrix has been assembled");
    if (precision == OC_FP32 && !s.values_sp.ptr)
        return set_error(OC_ERR_INVALID_ARGUMENT, "no FP32 Jacobian has been assembled");
    OC_CHECK(solver_structure(s));
    SolverSystem &S = s.solver;
    KernelTimer timer(precision == OC_FP64 ? "solver_matrix_fp64" : "solver_matrix_fp32");
    if (precision == OC_FP64) OC_CHECK(solver_values<double>(s, s.values.ptr, S.values));
    else OC_CHECK(solver_values<float>(s, s.values_sp.ptr, S.values_sp));
    if (S.n > 0) {
        const double *vec = linear ? s.rhs.ptr : s.residual.ptr;
        solver_vector_kernel<<<static_cast<unsigned>((S.n + 255) / 256), 256, 0, runtime().stream>>>(
            S.n, S.rows.ptr, s.row_ptr.ptr, s.col_idx.ptr, S.index.ptr, s.values.ptr, vec, s.dofs.ptr, linear,
            S.vector.ptr);
        OC_HIP_CHECK(hipGetLastError());
    }
    S.precision = precision;
    return OC_NO_ERROR;
}

}  // namespace oc
