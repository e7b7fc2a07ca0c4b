// oc_mlp.hip -- the learned warm start: a per-node MLP trained and evaluated
// entirely by this library's kernels (FP32 GEMMs on MFMA 16x16x4, GELU,
// deterministic reductions, Adam).  It has no counterpart in the reference;
// an answer is accepted only by the Newton solve's own FP64 residual test.
#include <algorithm>
#include <cmath>
#include <vector>

#include "oc_internal.h"

namespace oc {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---- GEMM: C[MxN] = op(A)[MxK] op(B)[KxN] + beta C, row-major -----------------
// op(A)(m,k) = transA ? A[k*lda + m] : A[m*lda + k]
// op(B)(k,n) = transB ? B[n*ldb + k] : B[k*ldb + n]
// 64x64 block tile, four waves of 32x32 (2x2 MFMA tiles), K staged by 16.
constexpr int kTM = 64, kTN = 64, kTK = 16;

__global__ void __launch_bounds__(256) gemm_f32_kernel(int M, int N, int K, int transA, int transB,
                                                      const float *__restrict__ A, int lda,
                                                      const float *__restrict__ B, int ldb, float beta,
                                                      float *__restrict__ C, int ldc)
{
    __shared__ float As[kTK][kTM + 4];
    __shared__ float Bs[kTK][kTN + 4];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += kTK) {
        for (int q = t; q < kTK * kTM; q += 256) {
            int kk, mm;
            if (transA) { kk = q / kTM; mm = q % kTM; }
            else { mm = q / kTK; kk = q % kTK; }
            const int m = m0 + mm, k = k0 + kk;
            float v = 0.f;
            if (m < M && k < K) v = transA ? A[static_cast<int64_t>(k) * lda + m] : A[static_cast<int64_t>(m) * lda + k];
            As[kk][mm] = v;
        }
        for (int q = t; q < kTK * kTN; q += 256) {
            int kk, nn;
            if (transB) { nn = q / kTK; kk = q % kTK; }
            else { kk = q / kTN; nn = q % kTN; }
            const int n = n0 + nn, k = k0 + kk;
            float v = 0.f;
            if (n < N && k < K) v = transB ? B[static_cast<int64_t>(n) * ldb + k] : B[static_cast<int64_t>(k) * ldb + n];
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kTK / 4; ++ks) {
            const int kr = ks * 4 + (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const float a = As[kr][wm + i * 16 + (lane & 15)];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const float b = Bs[kr][wn + j * 16 + (lane & 15)];
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i][j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + r, n = n0 + wn + j * 16 + (lane & 15);
                if (m < M && n < N) {
                    float *c = C + static_cast<int64_t>(m) * ldc + n;
                    *c = acc[i][j][r] + (beta != 0.f ? beta * *c : 0.f);
                }
            }
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

// ---- BF16 inference ------------------------------------------------------------------
// bf16 = the upper 16 bits of an IEEE float; conversion rounds to nearest even
// (NaN stays a quiet NaN).
typedef short bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

// W (rows x cols, FP32, row-major) -> Wb (rows_p x cols_p, bf16, zero padded)
__global__ void __launch_bounds__(256) weights_bf16_kernel(int rows, int cols, int rows_p, int cols_p,
                                                          const float *__restrict__ W, uint16_t *__restrict__ Wb)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < rows_p * cols_p; q += gridDim.x * blockDim.x) {
        const int r = q / cols_p, c = q % cols_p;
        Wb[q] = (r < rows && c < cols) ? f32_to_bf16(W[r * cols + c]) : static_cast<uint16_t>(0);
    }
}

// x (n x cols, FP32) -> xb (n x cols_p, bf16, zero padded)
__global__ void __launch_bounds__(256) rows_bf16_kernel(int64_t n, int cols, int cols_p, const float *__restrict__ x,
                                                       uint16_t *__restrict__ xb)
{
    for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < n * cols_p;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t r = q / cols_p;
        const int c = static_cast<int>(q % cols_p);
        xb[q] = c < cols ? f32_to_bf16(x[r * cols + c]) : static_cast<uint16_t>(0);
    }
}

// One layer Y = A W^T + b on mfma_f32_16x16x16bf16_1k with FP32 accumulation.
// A: n x K bf16, W: N x K bf16 (K, N multiples of 16, at most 128), staged in
// LDS.  A block is four waves of 16 rows; lane l holds A[l%16][k0 + 4(l/16) ..
// +3] and W[16j + l%16][same k], and D[4(l/16) + r][16j + l%16].  Hidden
// layers (out_bf) store bf16(GELU(y)) for all N columns; otherwise the layer
// stores y (gelu_f: GELU(y)) in FP32 for its n_valid columns.
constexpr int kBfMax = 128;
__global__ void __launch_bounds__(256) mlp_layer_bf16_kernel(int n, int K, int N, int n_valid,
                                                            const uint16_t *__restrict__ A,
                                                            const uint16_t *__restrict__ Wb,
                                                            const float *__restrict__ bias,
                                                            uint16_t *__restrict__ out_bf, float *__restrict__ out_f,
                                                            int gelu_f)
{
    __shared__ __align__(16) uint16_t Ws[kBfMax][kBfMax + 8];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int q = t; q < N * K; q += 256) Ws[q / K][q % K] = Wb[q];
    __syncthreads();
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * 64 + wave * 16;
    const int64_t arow = row0 + (lane & 15);
    const int kq = 4 * (lane >> 4);
    f32x4 acc[kBfMax / 16];
#pragma unroll
    for (int j = 0; j < kBfMax / 16; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a = bf16x4{0, 0, 0, 0};
        if (arow < n) a = *reinterpret_cast<const bf16x4 *>(A + arow * K + k0 + kq);
#pragma unroll
        for (int j = 0; j < kBfMax / 16; ++j) {
            if (16 * j < N) {
                const bf16x4 b = *reinterpret_cast<const bf16x4 *>(&Ws[16 * j + (lane & 15)][k0 + kq]);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kBfMax / 16; ++j) {
        if (16 * j >= N) continue;
        const int col = 16 * j + (lane & 15);
        const float bc = col < n_valid ? bias[col] : 0.f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t row = row0 + 4 * (lane >> 4) + r;
            if (row >= n) continue;
            const float v = acc[j][r] + bc;
            if (out_bf) out_bf[row * N + col] = f32_to_bf16(gelu(v));
            else if (col < n_valid) out_f[row * n_valid + col] = gelu_f ? gelu(v) : v;
        }
    }
}
__device__ __forceinline__ float gelu_grad(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// Z += b (row broadcast); H = GELU(Z) when act, else H = Z
__global__ void __launch_bounds__(256) bias_act_kernel(int64_t total, int width, float *__restrict__ Z,
                                                      const float *__restrict__ b, float *__restrict__ H, int act)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float z = Z[i] + b[i % width];
        Z[i] = z;
        H[i] = act ? gelu(z) : z;
    }
}

__global__ void __launch_bounds__(256) gelu_backward_kernel(int64_t total, const float *__restrict__ dH,
                                                           const float *__restrict__ Z, float *__restrict__ dZ)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        dZ[i] = dH[i] * gelu_grad(Z[i]);
}

// db[h] = sum_s dZ[s*width + h]  (one block per column, fixed tree)
__global__ void __launch_bounds__(256) column_sum_kernel(int n, int width, const float *__restrict__ dZ,
                                                        float *__restrict__ db)
{
    __shared__ float s[256];
    const int h = blockIdx.x;
    float acc = 0.f;
    for (int r = threadIdx.x; r < n; r += 256) acc += dZ[static_cast<int64_t>(r) * width + h];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[h] = s[0];
}

// dY = 2 (Y - T) / count; per-block partial sums of (Y - T)^2
__global__ void __launch_bounds__(256) mse_kernel(int64_t count, const float *__restrict__ Y,
                                                 const float *__restrict__ Tt, float *__restrict__ dY,
                                                 float *__restrict__ partial)
{
    __shared__ float s[256];
    float acc = 0.f;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float d = Y[i] - Tt[i];
        dY[i] = 2.0f * d / static_cast<float>(count);
        acc += d * d;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256) adam_kernel(int n, float *__restrict__ p, const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                                  float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// minibatch gather with on-device seeded sampling and input normalisation
__global__ void __launch_bounds__(256) gather_batch_kernel(int batch, int n_samples, int n_in, int n_out,
                                                          uint64_t seed, int step, const float *__restrict__ X,
                                                          const float *__restrict__ Tt,
                                                          const float *__restrict__ shift,
                                                          const float *__restrict__ scale, float *__restrict__ bx,
                                                          float *__restrict__ bt)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const uint64_t h = mix64(seed + 0x9E3779B97F4A7C15ull * (static_cast<uint64_t>(step) * 1048576ull + b + 1));
    const int64_t s = static_cast<int64_t>(h % static_cast<uint64_t>(n_samples));
    for (int i = 0; i < n_in; ++i) bx[b * n_in + i] = (X[s * n_in + i] - shift[i]) * scale[i];
    for (int i = 0; i < n_out; ++i) bt[b * n_out + i] = Tt[s * n_out + i];
}

__global__ void __launch_bounds__(256) normalise_kernel(int n, int n_in, const float *__restrict__ X,
                                                       const float *__restrict__ shift,
                                                       const float *__restrict__ scale, float *__restrict__ out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    for (int i = 0; i < n_in; ++i) out[s * n_in + i] = (X[static_cast<int64_t>(s) * n_in + i] - shift[i]) * scale[i];
}

// warm-start features of every quadratic node: X, z - X, load increment, c10, c01
// mesh of union node `node` given the per-mesh node offsets (M + 1 entries)
__device__ __forceinline__ int mesh_of(const int *__restrict__ off, int M, int node)
{
    int lo = 0, hi = M - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= node) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) warm_features_kernel(int nq, const double *__restrict__ coords,
                                                           const double *__restrict__ dofs, int M,
                                                           const int *__restrict__ q_off,
                                                           const double *__restrict__ dload_mesh,
                                                           const double *__restrict__ node_mat,
                                                           const float *__restrict__ shift,
                                                           const float *__restrict__ scale, float *__restrict__ out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
        const double dload = dload_mesh[mesh_of(q_off, M, i)];
        float f[9];
        for (int c = 0; c < 3; ++c) {
            const double X = coords[static_cast<int64_t>(c) * nq + i];
            f[c] = static_cast<float>(X);
            f[3 + c] = static_cast<float>(dofs[static_cast<int64_t>(c) * nq + i] - X);
        }
        f[6] = static_cast<float>(dload);
        f[7] = static_cast<float>(node_mat[2 * i]);
        f[8] = static_cast<float>(node_mat[2 * i + 1]);
        for (int k = 0; k < 9; ++k) out[static_cast<int64_t>(i) * 9 + k] = (f[k] - shift[k]) * scale[k];
    }
}

// z += dload * y[:, 0:3] on the free displacement DOFs of every quadratic node
__global__ void __launch_bounds__(256) warm_apply_kernel(int nq, const float *__restrict__ y, int ldy, int M,
                                                        const int *__restrict__ q_off,
                                                        const double *__restrict__ dload_mesh,
                                                        const uint8_t *__restrict__ free_mask,
                                                        double *__restrict__ dofs)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
        const double dload = dload_mesh[mesh_of(q_off, M, i)];
        for (int c = 0; c < 3; ++c) {
            const int64_t d = static_cast<int64_t>(c) * nq + i;
            if (free_mask == nullptr || free_mask[d])
                dofs[d] += dload * static_cast<double>(y[static_cast<int64_t>(i) * ldy + c]);
        }
    }
}

// quadratic node coinciding with each pressure (trilinear) node: element
// corner k of the linear basis is local node 2*(k&1) + 6*((k>>1)&1) + 18*(k>>2)
// of the triquadratic one (xi1 fastest in both).  Every writer of one entry
// writes the same value.
__global__ void __launch_bounds__(256) linear_to_quadratic_kernel(int n_elements, const int *__restrict__ conn_q,
                                                                 const int *__restrict__ conn_l,
                                                                 int *__restrict__ map)
{
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n_elements * 8; t += gridDim.x * blockDim.x) {
        const int e = t >> 3, k = t & 7;
        const int a = 2 * (k & 1) + 6 * ((k >> 1) & 1) + 18 * (k >> 2);
        map[conn_l[t]] = conn_q[static_cast<int64_t>(e) * 27 + a];
    }
}

// p += dload * (c10 + c01) * y[quad(node), 3] on the free pressure DOFs (the
// network's pressure output is in units of c10 + c01)
__global__ void __launch_bounds__(256) warm_apply_pressure_kernel(int nq, int nl, const int *__restrict__ map,
                                                                 const float *__restrict__ y, int M,
                                                                 const int *__restrict__ q_off,
                                                                 const double *__restrict__ dload_mesh,
                                                                 const double *__restrict__ node_mat,
                                                                 const uint8_t *__restrict__ free_mask,
                                                                 double *__restrict__ dofs)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += gridDim.x * blockDim.x) {
        const int64_t d = static_cast<int64_t>(3) * nq + i;
        const int q = map[i];
        const double dload = dload_mesh[mesh_of(q_off, M, q)];
        if (free_mask == nullptr || free_mask[d])
            dofs[d] += dload * (node_mat[2 * q] + node_mat[2 * q + 1]) * static_cast<double>(y[static_cast<int64_t>(q) * 4 + 3]);
    }
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096))); }

}  // namespace

struct Mlp {
    int n_in, n_hidden, n_layers_hidden, n_out;
    std::vector<int> in_dim, out_dim;  // per linear layer
    std::vector<int64_t> w_off, b_off;
    int64_t n_params = 0;
    DevBuf<float> params, grads, adam_m, adam_v, shift, scale;
    int adam_t = 0;
    int capacity = 0;
    std::vector<DevBuf<float>> pre, act;  // pre[l]: n x out_l ; act[l]: n x in_l (act[0] = input)
    DevBuf<float> dY, dA, dB, partial, out;
    DevBuf<float> data_x, data_t;
    DevBuf<double> dload;  // per-mesh load increments of the last warm start
    int n_data = 0;
    // inference precision (OC_FP32 or OC_BF16); bf16 weights are converted
    // from params whenever params changed since the last conversion
    int infer_precision = OC_FP32;
    uint64_t params_version = 1, bf16_version = 0;
    DevBuf<uint16_t> wbf, xb, hb0, hb1;
    std::vector<int64_t> wbf_off;
    int bf_capacity = 0;
    static int pad16(int v) { return (v + 15) / 16 * 16; }

    int layers() const { return static_cast<int>(in_dim.size()); }

    int ensure_capacity(int n)
    {
        if (n <= capacity) return OC_NO_ERROR;
        const int L = layers();
        pre.resize(L);
        act.resize(L);
        for (int l = 0; l < L; ++l) {
            OC_CHECK(pre[l].alloc(static_cast<size_t>(n) * out_dim[l]));
            OC_CHECK(act[l].alloc(static_cast<size_t>(n) * in_dim[l]));
        }
        int wmax = std::max(n_hidden, std::max(n_in, n_out));
        OC_CHECK(dY.alloc(static_cast<size_t>(n) * wmax));
        OC_CHECK(dA.alloc(static_cast<size_t>(n) * wmax));
        OC_CHECK(dB.alloc(static_cast<size_t>(n) * wmax));
        OC_CHECK(out.alloc(static_cast<size_t>(n) * n_out));
        OC_CHECK(partial.alloc(4096));
        capacity = n;
        return OC_NO_ERROR;
    }

    int gemm(int M, int N, int K, int ta, int tb, const float *A, int lda, const float *B, int ldb, float beta,
             float *C, int ldc)
    {
        dim3 grid((N + kTN - 1) / kTN, (M + kTM - 1) / kTM);
        gemm_f32_kernel<<<grid, 256, 0, runtime().stream>>>(M, N, K, ta, tb, A, lda, B, ldb, beta, C, ldc);
        OC_HIP_CHECK(hipGetLastError());
        return OC_NO_ERROR;
    }

    // act[0] must hold the normalised input; result in pre[L-1]
    int forward(int n)
    {
        KernelTimer timer("mlp_forward");
        const int L = layers();
        Runtime &rt = runtime();
        for (int l = 0; l < L; ++l) {
            OC_CHECK(gemm(n, out_dim[l], in_dim[l], 0, 1, act[l].ptr, in_dim[l], params.ptr + w_off[l], in_dim[l], 0.f,
                          pre[l].ptr, out_dim[l]));
            float *h = l + 1 < L ? act[l + 1].ptr : out.ptr;
            const int64_t total = static_cast<int64_t>(n) * out_dim[l];
            bias_act_kernel<<<blocks_for(total), 256, 0, rt.stream>>>(total, out_dim[l], pre[l].ptr,
                                                                      params.ptr + b_off[l], h, l + 1 < L ? 1 : 0);
            OC_HIP_CHECK(hipGetLastError());
        }
        return OC_NO_ERROR;
    }

    // bf16 copies of the weights (converted when the parameters changed) and
    // bf16 activation buffers for n rows
    int prepare_bf16(int n)
    {
        const int L = layers();
        Runtime &rt = runtime();
        if (bf16_version != params_version) {
            wbf_off.assign(L + 1, 0);
            for (int l = 0; l < L; ++l)
                wbf_off[l + 1] = wbf_off[l] + static_cast<int64_t>(pad16(out_dim[l])) * pad16(in_dim[l]);
            if (wbf.n < static_cast<size_t>(wbf_off[L])) OC_CHECK(wbf.alloc(static_cast<size_t>(wbf_off[L])));
            for (int l = 0; l < L; ++l) {
                const int rp = pad16(out_dim[l]), cp = pad16(in_dim[l]);
                weights_bf16_kernel<<<blocks_for(static_cast<int64_t>(rp) * cp), 256, 0, rt.stream>>>(
                    out_dim[l], in_dim[l], rp, cp, params.ptr + w_off[l], wbf.ptr + wbf_off[l]);
                OC_HIP_CHECK(hipGetLastError());
            }
            bf16_version = params_version;
        }
        if (n > bf_capacity) {
            OC_CHECK(xb.alloc(static_cast<size_t>(n) * pad16(n_in)));
            OC_CHECK(hb0.alloc(static_cast<size_t>(n) * pad16(n_hidden)));
            OC_CHECK(hb1.alloc(static_cast<size_t>(n) * pad16(n_hidden)));
            bf_capacity = n;
        }
        return OC_NO_ERROR;
    }

    // act[0] must hold the normalised input; result in out (FP32).  BF16
    // operands and activations, FP32 accumulation, bias and GELU.
    int forward_bf16(int n)
    {
        KernelTimer timer("mlp_forward_bf16");
        const int L = layers();
        Runtime &rt = runtime();
        OC_CHECK(prepare_bf16(n));
        rows_bf16_kernel<<<blocks_for(static_cast<int64_t>(n) * pad16(n_in)), 256, 0, rt.stream>>>(
            n, n_in, pad16(n_in), act[0].ptr, xb.ptr);
        OC_HIP_CHECK(hipGetLastError());
        const uint16_t *a = xb.ptr;
        uint16_t *h = hb0.ptr, *spare = hb1.ptr;
        const unsigned blocks = static_cast<unsigned>((static_cast<int64_t>(n) + 63) / 64);
        for (int l = 0; l < L; ++l) {
            const bool last = l + 1 == L;
            mlp_layer_bf16_kernel<<<blocks, 256, 0, rt.stream>>>(n, pad16(in_dim[l]), pad16(out_dim[l]), out_dim[l], a,
                                                                 wbf.ptr + wbf_off[l], params.ptr + b_off[l],
                                                                 last ? nullptr : h, last ? out.ptr : nullptr, 0);
            OC_HIP_CHECK(hipGetLastError());
            a = h;
            std::swap(h, spare);
        }
        return OC_NO_ERROR;
    }

    // BF16 hidden layers between an FP32 first and last layer: the input
    // features and the output increment keep FP32; the hidden activations
    // after the first layer are bf16, and the last hidden layer writes its
    // GELU in FP32 for the FP32 output layer.
    int forward_bf16_hidden(int n)
    {
        KernelTimer timer("mlp_forward_bf16_hidden");
        const int L = layers();
        if (L < 3) return forward(n);
        Runtime &rt = runtime();
        OC_CHECK(prepare_bf16(n));
        auto fp32_layer = [&](int l, const float *in, float *h, int act) -> int {
            OC_CHECK(gemm(n, out_dim[l], in_dim[l], 0, 1, in, in_dim[l], params.ptr + w_off[l], in_dim[l], 0.f,
                          pre[l].ptr, out_dim[l]));
            const int64_t total = static_cast<int64_t>(n) * out_dim[l];
            bias_act_kernel<<<blocks_for(total), 256, 0, rt.stream>>>(total, out_dim[l], pre[l].ptr,
                                                                      params.ptr + b_off[l], h, act);
            OC_HIP_CHECK(hipGetLastError());
            return OC_NO_ERROR;
        };
        OC_CHECK(fp32_layer(0, act[0].ptr, act[1].ptr, 1));
        rows_bf16_kernel<<<blocks_for(static_cast<int64_t>(n) * pad16(n_hidden)), 256, 0, rt.stream>>>(
            n, n_hidden, pad16(n_hidden), act[1].ptr, hb0.ptr);
        OC_HIP_CHECK(hipGetLastError());
        const uint16_t *a = hb0.ptr;
        uint16_t *h = hb1.ptr, *spare = hb0.ptr;
        const unsigned blocks = static_cast<unsigned>((static_cast<int64_t>(n) + 63) / 64);
        for (int l = 1; l + 1 < L; ++l) {
            const bool to_f32 = l + 2 == L;  // feeds the FP32 output layer
            mlp_layer_bf16_kernel<<<blocks, 256, 0, rt.stream>>>(n, pad16(in_dim[l]), pad16(out_dim[l]), out_dim[l], a,
                                                                 wbf.ptr + wbf_off[l], params.ptr + b_off[l],
                                                                 to_f32 ? nullptr : h, to_f32 ? act[L - 1].ptr : nullptr,
                                                                 1);
            OC_HIP_CHECK(hipGetLastError());
            a = h;
            std::swap(h, spare);
        }
        return fp32_layer(L - 1, act[L - 1].ptr, out.ptr, 0);
    }

    int infer(int n)
    {
        if (infer_precision == OC_BF16) return forward_bf16(n);
        if (infer_precision == OC_BF16_HIDDEN) return forward_bf16_hidden(n);
        return forward(n);
    }

    // after forward(n) with targets T on device: gradients into grads, loss partials into partial
    int backward(int n, const float *T, int *n_partials)
    {
        KernelTimer timer("mlp_backward");
        const int L = layers();
        Runtime &rt = runtime();
        const int64_t cnt = static_cast<int64_t>(n) * n_out;
        const unsigned nb = std::min<unsigned>(blocks_for(cnt), 1024);
        mse_kernel<<<nb, 256, 0, rt.stream>>>(cnt, out.ptr, T, dY.ptr, partial.ptr);
        OC_HIP_CHECK(hipGetLastError());
        *n_partials = static_cast<int>(nb);
        float *dZ = dY.ptr, *spare = dA.ptr;
        for (int l = L - 1; l >= 0; --l) {
            // dW_l = dZ^T act_l ; db_l = colsum(dZ)
            OC_CHECK(gemm(out_dim[l], in_dim[l], n, 1, 0, dZ, out_dim[l], act[l].ptr, in_dim[l], 0.f,
                          grads.ptr + w_off[l], in_dim[l]));
            column_sum_kernel<<<out_dim[l], 256, 0, rt.stream>>>(n, out_dim[l], dZ, grads.ptr + b_off[l]);
            OC_HIP_CHECK(hipGetLastError());
            if (l == 0) break;
            // dAct_l = dZ W_l ; dZ_{l-1} = dAct_l * gelu'(pre_{l-1})
            OC_CHECK(gemm(n, in_dim[l], out_dim[l], 0, 0, dZ, out_dim[l], params.ptr + w_off[l], in_dim[l], 0.f,
                          spare, in_dim[l]));
            const int64_t total = static_cast<int64_t>(n) * in_dim[l];
            float *next = (dZ == dB.ptr) ? dY.ptr : dB.ptr;
            gelu_backward_kernel<<<blocks_for(total), 256, 0, rt.stream>>>(total, spare, pre[l - 1].ptr, next);
            OC_HIP_CHECK(hipGetLastError());
            dZ = next;
        }
        return OC_NO_ERROR;
    }

    int adam(float lr)
    {
        ++adam_t;
        ++params_version;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, static_cast<float>(adam_t)), c2 = 1.f - std::pow(b2, static_cast<float>(adam_t));
        adam_kernel<<<static_cast<unsigned>((n_params + 255) / 256), 256, 0, runtime().stream>>>(
            static_cast<int>(n_params), params.ptr, grads.ptr, adam_m.ptr, adam_v.ptr, lr, b1, b2, 1e-8f, c1, c2);
        OC_HIP_CHECK(hipGetLastError());
        return OC_NO_ERROR;
    }
};

}  // namespace oc

using namespace oc;

struct oc_mlp_s {
    Mlp m;
};

static uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" {

int oc_mlp_create(int n_inputs, int n_hidden, int n_hidden_layers, int n_outputs, unsigned long long seed, oc_mlp *mlp)
{
    OC_CHECK(require_runtime());
    if (!mlp || n_inputs < 1 || n_hidden < 1 || n_hidden_layers < 1 || n_outputs < 1)
        return set_error(OC_ERR_INVALID_ARGUMENT, "oc_mlp_create: invalid arguments");
    auto *h = new (std::nothrow) oc_mlp_s();
    if (!h) return set_error(OC_ERR_ALLOCATION, "oc_mlp_create: out of host memory");
    Mlp &m = h->m;
    m.n_in = n_inputs;
    m.n_hidden = n_hidden;
    m.n_layers_hidden = n_hidden_layers;
    m.n_out = n_outputs;
    m.in_dim.push_back(n_inputs);
    m.out_dim.push_back(n_hidden);
    for (int l = 1; l < n_hidden_layers; ++l) {
        m.in_dim.push_back(n_hidden);
        m.out_dim.push_back(n_hidden);
    }
    m.in_dim.push_back(n_hidden);
    m.out_dim.push_back(n_outputs);
    int64_t off = 0;
    for (size_t l = 0; l < m.in_dim.size(); ++l) {
        m.w_off.push_back(off);
        off += static_cast<int64_t>(m.in_dim[l]) * m.out_dim[l];
        m.b_off.push_back(off);
        off += m.out_dim[l];
    }
    m.n_params = off;
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases (seeded)
    std::vector<float> p(static_cast<size_t>(off), 0.f);
    uint64_t st = seed;
    for (size_t l = 0; l < m.in_dim.size(); ++l) {
        const float bound = 1.0f / std::sqrt(static_cast<float>(m.in_dim[l]));
        for (int64_t k = 0; k < static_cast<int64_t>(m.in_dim[l]) * m.out_dim[l]; ++k) {
            const double u = static_cast<double>(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            p[m.w_off[l] + k] = static_cast<float>((2.0 * u - 1.0) * bound);
        }
    }
    std::vector<float> zero(static_cast<size_t>(off), 0.f), sh(n_inputs, 0.f), sc(n_inputs, 1.f);
    int rc = m.params.upload(p.data(), p.size());
    if (rc == OC_NO_ERROR) rc = m.grads.upload(zero.data(), zero.size());
    if (rc == OC_NO_ERROR) rc = m.adam_m.upload(zero.data(), zero.size());
    if (rc == OC_NO_ERROR) rc = m.adam_v.upload(zero.data(), zero.size());
    if (rc == OC_NO_ERROR) rc = m.shift.upload(sh.data(), sh.size());
    if (rc == OC_NO_ERROR) rc = m.scale.upload(sc.data(), sc.size());
    if (rc != OC_NO_ERROR) {
        delete h;
        return rc;
    }
    *mlp = h;
    return OC_NO_ERROR;
}

int oc_mlp_destroy(oc_mlp mlp)
{
    if (!mlp) return set_error(OC_ERR_INVALID_HANDLE, "MLP handle is NULL");
    if (runtime().stream) (void)hipStreamSynchronize(runtime().stream);
    delete mlp;
    return OC_NO_ERROR;
}

int oc_mlp_parameter_count(oc_mlp mlp, int *count)
{
    if (!mlp || !count) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    *count = static_cast<int>(mlp->m.n_params);
    return OC_NO_ERROR;
}

int oc_mlp_parameters_get(oc_mlp mlp, float *parameters)
{
    OC_CHECK(require_runtime());
    if (!mlp || !parameters) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_HIP_CHECK(hipStreamSynchronize(runtime().stream));
    OC_HIP_CHECK(hipMemcpy(parameters, mlp->m.params.ptr, sizeof(float) * mlp->m.n_params, hipMemcpyDeviceToHost));
    return OC_NO_ERROR;
}

int oc_mlp_parameters_set(oc_mlp mlp, const float *parameters)
{
    OC_CHECK(require_runtime());
    if (!mlp || !parameters) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_HIP_CHECK(hipStreamSynchronize(runtime().stream));
    OC_HIP_CHECK(hipMemcpy(mlp->m.params.ptr, parameters, sizeof(float) * mlp->m.n_params, hipMemcpyHostToDevice));
    ++mlp->m.params_version;
    return OC_NO_ERROR;
}

int oc_mlp_inference_precision_set(oc_mlp mlp, int precision)
{
    if (!mlp) return set_error(OC_ERR_INVALID_HANDLE, "MLP handle is NULL");
    if (precision != OC_FP32 && precision != OC_BF16 && precision != OC_BF16_HIDDEN)
        return set_error(OC_ERR_INVALID_ARGUMENT, "inference precision must be OC_FP32, OC_BF16 or OC_BF16_HIDDEN");
    const Mlp &m = mlp->m;
    if (precision != OC_FP32 && (m.n_in > kBfMax || m.n_hidden > kBfMax || m.n_out > kBfMax))
        return set_error(OC_ERR_UNSUPPORTED, "BF16 inference needs layer widths of at most %d", kBfMax);
    mlp->m.infer_precision = precision;
    return OC_NO_ERROR;
}

int oc_mlp_normalisation_set(oc_mlp mlp, const float *shift, const float *scale)
{
    OC_CHECK(require_runtime());
    if (!mlp || !shift || !scale) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_HIP_CHECK(hipStreamSynchronize(runtime().stream));
    OC_HIP_CHECK(hipMemcpy(mlp->m.shift.ptr, shift, sizeof(float) * mlp->m.n_in, hipMemcpyHostToDevice));
    OC_HIP_CHECK(hipMemcpy(mlp->m.scale.ptr, scale, sizeof(float) * mlp->m.n_in, hipMemcpyHostToDevice));
    return OC_NO_ERROR;
}

int oc_mlp_forward(oc_mlp mlp, int n, const float *x, float *y)
{
    OC_CHECK(require_runtime());
    if (!mlp || !x || !y || n < 1) return set_error(OC_ERR_INVALID_ARGUMENT, "oc_mlp_forward: invalid arguments");
    Mlp &m = mlp->m;
    OC_CHECK(m.ensure_capacity(n));
    Runtime &rt = runtime();
    DevBuf<float> raw;
    OC_CHECK(raw.upload(x, static_cast<size_t>(n) * m.n_in));
    normalise_kernel<<<blocks_for(n), 256, 0, rt.stream>>>(n, m.n_in, raw.ptr, m.shift.ptr, m.scale.ptr, m.act[0].ptr);
    OC_HIP_CHECK(hipGetLastError());
    OC_CHECK(m.infer(n));
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    OC_HIP_CHECK(hipMemcpy(y, m.out.ptr, sizeof(float) * n * m.n_out, hipMemcpyDeviceToHost));
    return OC_NO_ERROR;
}

int oc_mlp_gradient(oc_mlp mlp, int n, const float *x, const float *t, float *loss, float *gradient)
{
    OC_CHECK(require_runtime());
    if (!mlp || !x || !t || !loss || !gradient || n < 1)
        return set_error(OC_ERR_INVALID_ARGUMENT, "oc_mlp_gradient: invalid arguments");
    Mlp &m = mlp->m;
    OC_CHECK(m.ensure_capacity(n));
    Runtime &rt = runtime();
    DevBuf<float> raw, tt;
    OC_CHECK(raw.upload(x, static_cast<size_t>(n) * m.n_in));
    OC_CHECK(tt.upload(t, static_cast<size_t>(n) * m.n_out));
    normalise_kernel<<<blocks_for(n), 256, 0, rt.stream>>>(n, m.n_in, raw.ptr, m.shift.ptr, m.scale.ptr, m.act[0].ptr);
    OC_HIP_CHECK(hipGetLastError());
    OC_CHECK(m.forward(n));
    int np = 0;
    OC_CHECK(m.backward(n, tt.ptr, &np));
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
    std::vector<float> part(np);
    OC_HIP_CHECK(hipMemcpy(part.data(), m.partial.ptr, sizeof(float) * np, hipMemcpyDeviceToHost));
    double acc = 0.0;
    for (float v : part) acc += v;
    *loss = static_cast<float>(acc / (static_cast<double>(n) * m.n_out));
    OC_HIP_CHECK(hipMemcpy(gradient, m.grads.ptr, sizeof(float) * m.n_params, hipMemcpyDeviceToHost));
    return OC_NO_ERROR;
}

int oc_mlp_dataset_set(oc_mlp mlp, int n, const float *x, const float *t)
{
    OC_CHECK(require_runtime());
    if (!mlp || !x || !t || n < 1) return set_error(OC_ERR_INVALID_ARGUMENT, "oc_mlp_dataset_set: invalid arguments");
    Mlp &m = mlp->m;
    OC_CHECK(m.data_x.upload(x, static_cast<size_t>(n) * m.n_in));
    OC_CHECK(m.data_t.upload(t, static_cast<size_t>(n) * m.n_out));
    m.n_data = n;
    return OC_NO_ERROR;
}

int oc_mlp_train(oc_mlp mlp, int steps, int batch, float learning_rate, unsigned long long seed, float *mean_loss)
{
    OC_CHECK(require_runtime());
    if (!mlp || steps < 1 || batch < 1 || !mean_loss) return set_error(OC_ERR_INVALID_ARGUMENT, "oc_mlp_train: invalid arguments");
    Mlp &m = mlp->m;
    if (m.n_data < 1) return set_error(OC_ERR_INVALID_ARGUMENT, "oc_mlp_train: no dataset");
    OC_CHECK(m.ensure_capacity(batch));
    Runtime &rt = runtime();
    DevBuf<float> bt, losses;
    OC_CHECK(bt.alloc(static_cast<size_t>(batch) * m.n_out));
    std::vector<float> part;
    double total = 0.0;
    for (int s = 0; s < steps; ++s) {
        gather_batch_kernel<<<blocks_for(batch), 256, 0, rt.stream>>>(batch, m.n_data, m.n_in, m.n_out, seed, s,
                                                                      m.data_x.ptr, m.data_t.ptr, m.shift.ptr,
                                                                      m.scale.ptr, m.act[0].ptr, bt.ptr);
        OC_HIP_CHECK(hipGetLastError());
        OC_CHECK(m.forward(batch));
        int np = 0;
        OC_CHECK(m.backward(batch, bt.ptr, &np));
        OC_CHECK(m.adam(learning_rate));
        if (s == steps - 1 || (s % 64) == 63) {
            part.resize(np);
            OC_HIP_CHECK(hipMemcpyAsync(part.data(), m.partial.ptr, sizeof(float) * np, hipMemcpyDeviceToHost, rt.stream));
            OC_HIP_CHECK(hipStreamSynchronize(rt.stream));
            double acc = 0.0;
            for (float v : part) acc += v;
            total = acc / (static_cast<double>(batch) * m.n_out);
        }
    }
    *mean_loss = static_cast<float>(total);
    return OC_NO_ERROR;
}

int oc_mlp_warm_start(oc_mlp mlp, oc_equations_set h, double load_increment)
{
    OC_CHECK(require_runtime());
    if (!h) return set_error(OC_ERR_INVALID_HANDLE, "NULL handle");
    std::vector<double> d(h->s.n_meshes, load_increment);
    return oc_mlp_warm_start_batch(mlp, h, d.data());
}

int oc_mlp_warm_start_batch(oc_mlp mlp, oc_equations_set h, const double *load_increments)
{
    OC_CHECK(require_runtime());
    if (!mlp || !h) return set_error(OC_ERR_INVALID_HANDLE, "NULL handle");
    if (!load_increments) return set_error(OC_ERR_INVALID_ARGUMENT, "load_increments is NULL");
    Mlp &m = mlp->m;
    EquationsSet &s = h->s;
    if (s.cls != OC_EQUATIONS_SET_MOONEY_RIVLIN || m.n_in != 9 || (m.n_out != 3 && m.n_out != 4))
        return set_error(OC_ERR_INVALID_ARGUMENT, "warm start needs a Mooney-Rivlin set and a 9-in, 3- or 4-out MLP");
    const int nq = s.n_nodes, nl = s.n_nodes_l, M = s.n_meshes;
    OC_CHECK(m.ensure_capacity(nq));
    Runtime &rt = runtime();
    OC_HIP_CHECK(hipStreamSynchronize(rt.stream));  // the increments buffer may still be read by a previous call
    OC_CHECK(m.dload.upload(load_increments, static_cast<size_t>(M)));
    // the plain start (for the Newton fallback) is the DOFs before the warm start
    if (!s.plain_start.ptr) OC_CHECK(s.plain_start.alloc(static_cast<size_t>(s.n_rows)));
    OC_HIP_CHECK(hipMemcpyAsync(s.plain_start.ptr, s.dofs.ptr, sizeof(double) * s.n_rows, hipMemcpyDeviceToDevice,
                                rt.stream));
    s.plain_start_valid = true;
    KernelTimer timer("mlp_warm_start");
    if (!s.node_mat.ptr) OC_CHECK(mr_node_materials(s));
    warm_features_kernel<<<blocks_for(nq), 256, 0, rt.stream>>>(nq, s.coords.ptr, s.dofs.ptr, M, s.q_off.ptr,
                                                                m.dload.ptr, s.node_mat.ptr, m.shift.ptr, m.scale.ptr,
                                                                m.act[0].ptr);
    OC_HIP_CHECK(hipGetLastError());
    OC_CHECK(m.infer(nq));
    warm_apply_kernel<<<blocks_for(nq), 256, 0, rt.stream>>>(nq, m.out.ptr, m.n_out, M, s.q_off.ptr, m.dload.ptr,
                                                             s.free_mask.ptr, s.dofs.ptr);
    OC_HIP_CHECK(hipGetLastError());
    if (m.n_out == 4) {
        if (!s.lin_to_quad.ptr) {
            OC_CHECK(s.lin_to_quad.alloc(nl));
            linear_to_quadratic_kernel<<<blocks_for(static_cast<int64_t>(s.n_elements) * 8), 256, 0, rt.stream>>>(
                s.n_elements, s.conn.ptr, s.conn_l.ptr, s.lin_to_quad.ptr);
            OC_HIP_CHECK(hipGetLastError());
        }
        warm_apply_pressure_kernel<<<blocks_for(nl), 256, 0, rt.stream>>>(nq, nl, s.lin_to_quad.ptr, m.out.ptr, M,
                                                                          s.q_off.ptr, m.dload.ptr, s.node_mat.ptr,
                                                                          s.free_mask.ptr, s.dofs.ptr);
        OC_HIP_CHECK(hipGetLastError());
    }
    return OC_NO_ERROR;
}

}  // extern "C"
