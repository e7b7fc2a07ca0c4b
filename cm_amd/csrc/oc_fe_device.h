// oc_fe_device.h -- device-side finite element helpers for gfx950:
// basis-table layout, 3x3 algebra, metrics, and the Mooney-Rivlin stress and
// spatial elasticity tensor.  All FP64; every routine names the reference
// routine whose arithmetic it restates.  3x3 matrices are column-major,
// A[i + 3*j] == A(i+1, j+1) of the Fortran source.
#pragma once

#include <hip/hip_runtime.h>

namespace oc {

// Basis tables: value and first derivatives of each basis function at each
// Gauss point, GAUSS_BASIS_FNS(ns, {1,2,4,7}, ng) of BASIS_QUADRATURE_CREATE
// (basis_routines.f90:2770-2890).  Layout [g][s][4] with nu 0 = value,
// 1..3 = d/dxi1..3; Gauss points xi1-fastest; basis nodes xi1-fastest.
struct BasisTables {
    double w27[27];              // 3x3x3 Gauss weights
    double w8[8];                // 2x2x2 Gauss weights
    double quad27[27 * 27 * 4];  // triquadratic at 3x3x3
    double lin27[27 * 8 * 4];    // trilinear at 3x3x3
    double lin8[8 * 8 * 4];      // trilinear at 2x2x2
};

// Builds the tables on the host (restating GAUSS_LEGENDRE,
// basis_routines.f90:5826-5875, and LAGRANGE_*_EVALUATE :7015-7137).
void build_basis_tables(BasisTables &t);
// Device copy owned by the runtime (built at oc_initialise).
const BasisTables *device_basis_tables();

// XCD-aware block order.  MI355X deals the workgroups of a launch out to its
// eight XCDs round-robin (block b runs on XCD b % 8), and each XCD has its own
// L2.  Row kernels whose neighbouring rows gather the same iterate entries
// take their block index through this bijection of [0, nb): XCD x then works
// through one contiguous share of the blocks, so the gathers of neighbouring
// rows meet in one L2 instead of being fetched once per XCD.  The order of the
// work within a row, and so every result, is unchanged.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    constexpr int kXcds = 8;
    const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// TENSOR_TO_VOIGT3, constants.f90:284 (0-based: 11,22,33,12,13,23).
__device__ __forceinline__ int t2v(int i, int j) { return i == j ? i : 2 + i + j; }

// INVERT_FULL_DP, 3x3 case, maths.f90:1455-1521 (cofactors over det;
// zero matrix when |det| <= ZERO_TOLERANCE = 5*eps, constants.f90:68).
__device__ __forceinline__ double invert3(const double *A, double *B)
{
#define a(i, j) A[(i - 1) + 3 * (j - 1)]
#define b(i, j) B[(i - 1) + 3 * (j - 1)]
    const double d = a(1, 1) * a(2, 2) * a(3, 3) + a(1, 2) * a(2, 3) * a(3, 1) + a(1, 3) * a(2, 1) * a(3, 2) -
                     a(1, 1) * a(3, 2) * a(2, 3) - a(2, 1) * a(1, 2) * a(3, 3) - a(3, 1) * a(2, 2) * a(1, 3);
    if (fabs(d) > 5.0 * 2.220446049250313e-16) {
        b(1, 1) = (a(2, 2) * a(3, 3) - a(3, 2) * a(2, 3)) / d;
        b(2, 1) = (a(2, 3) * a(3, 1) - a(3, 3) * a(2, 1)) / d;
        b(3, 1) = (a(2, 1) * a(3, 2) - a(3, 1) * a(2, 2)) / d;
        b(1, 2) = (a(3, 2) * a(1, 3) - a(1, 2) * a(3, 3)) / d;
        b(2, 2) = (a(3, 3) * a(1, 1) - a(1, 3) * a(3, 1)) / d;
        b(3, 2) = (a(3, 1) * a(1, 2) - a(1, 1) * a(3, 2)) / d;
        b(1, 3) = (a(1, 2) * a(2, 3) - a(2, 2) * a(1, 3)) / d;
        b(2, 3) = (a(1, 3) * a(2, 1) - a(2, 3) * a(1, 1)) / d;
        b(3, 3) = (a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2)) / d;
    } else {
        for (int i = 0; i < 9; ++i) B[i] = 0.0;
    }
    return d;
#undef a
#undef b
}

// INVERT_FULL_DP as invert3, with one division: the cofactors times 1/det
// (the same zero-tolerance rule; entries differ from invert3's by an ulp).
__device__ __forceinline__ double invert3_rcp(const double *A, double *B)
{
#define a(i, j) A[(i - 1) + 3 * (j - 1)]
#define b(i, j) B[(i - 1) + 3 * (j - 1)]
    const double d = a(1, 1) * a(2, 2) * a(3, 3) + a(1, 2) * a(2, 3) * a(3, 1) + a(1, 3) * a(2, 1) * a(3, 2) -
                     a(1, 1) * a(3, 2) * a(2, 3) - a(2, 1) * a(1, 2) * a(3, 3) - a(3, 1) * a(2, 2) * a(1, 3);
    if (fabs(d) > 5.0 * 2.220446049250313e-16) {
        const double r = 1.0 / d;
        b(1, 1) = (a(2, 2) * a(3, 3) - a(3, 2) * a(2, 3)) * r;
        b(2, 1) = (a(2, 3) * a(3, 1) - a(3, 3) * a(2, 1)) * r;
        b(3, 1) = (a(2, 1) * a(3, 2) - a(3, 1) * a(2, 2)) * r;
        b(1, 2) = (a(3, 2) * a(1, 3) - a(1, 2) * a(3, 3)) * r;
        b(2, 2) = (a(3, 3) * a(1, 1) - a(1, 3) * a(3, 1)) * r;
        b(3, 2) = (a(3, 1) * a(1, 2) - a(1, 1) * a(3, 2)) * r;
        b(1, 3) = (a(1, 2) * a(2, 3) - a(2, 2) * a(1, 3)) * r;
        b(2, 3) = (a(1, 3) * a(2, 1) - a(2, 3) * a(1, 1)) * r;
        b(3, 3) = (a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2)) * r;
    } else {
        for (int i = 0; i < 9; ++i) B[i] = 0.0;
    }
    return d;
#undef a
#undef b
}

// Per-Gauss-point kinematics of a triquadratic element
// (FiniteElasticity_FiniteElementResidualEvaluate :1491-1532 and
// FiniteElasticity_FiniteElementJacobianEvaluate :954-989):
//   dz/dxi, dX/dxi interpolated (FIELD_INTERPOLATE_GAUSS), J_z, J_X,
//   DXI_DX = inv(dz/dxi), F = dz/dxi * inv(dX/dxi)
//   (FiniteElasticity_GaussDeformationGradientTensor :3368-3425 with
//   Coordinates_MaterialSystemCalculate, no fibres), DPHIDZ.
struct GaussKinematics {
    double F[9];
    double jz, jx, Jznu;
    double dxidz[9];
};

// Triquadratic first derivatives at the 27 Gauss points, transposed for LDS:
// dphi_n/dxi_ni at g is tab[(ni * 27 + n) * kTabStride + g].  The odd stride
// keeps both access patterns nearly free of bank conflicts: lanes that differ
// in g (one thread per Gauss point) and lanes that differ in n.
constexpr int kTabStride = 29;
constexpr int kTabSize = 3 * 27 * kTabStride;

// ze, Xe: element parameters [c*27 + n]; tab: the transposed derivative table.
__device__ __forceinline__ void gauss_kinematics(const double *tab, int g, const double *ze, const double *Xe,
                                                 GaussKinematics &k)
{
    double dz[9], dX[9], dxidnu[9];
    // one derivative direction at a time, the three components together (each
    // sum still runs over n in ascending order); the loops are kept rolled so
    // that the unrolled loads do not take hundreds of registers
#pragma unroll 1
    for (int ni = 0; ni < 3; ++ni) {
        double sz[3] = {0.0, 0.0, 0.0}, sX[3] = {0.0, 0.0, 0.0};
#pragma unroll 3
        for (int n = 0; n < 27; ++n) {
            const double f = tab[(ni * 27 + n) * kTabStride + g];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                sz[x] += f * ze[x * 27 + n];
                sX[x] += f * Xe[x * 27 + n];
            }
        }
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            dz[x + 3 * ni] = sz[x];
            dX[x + 3 * ni] = sX[x];
        }
    }
    // volume Jacobians of COORDINATE_METRICS_CALCULATE (coordinate_routines.f90:
    // 647-840), sqrt|det(dx^T dx)|, taken as |det dx| from the inversions
    k.jz = fabs(invert3_rcp(dz, k.dxidz));
    k.jx = fabs(invert3_rcp(dX, dxidnu));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            k.F[i + 3 * j] = dz[i + 3 * 0] * dxidnu[0 + 3 * j] + dz[i + 3 * 1] * dxidnu[1 + 3 * j] +
                             dz[i + 3 * 2] * dxidnu[2 + 3 * j];
    k.Jznu = k.jz / k.jx;
}

// t(6x6) of FINITE_ELASTICITY_PUSH_STRESS_TENSOR / _ELASTICITY_TENSOR
// (finite_elasticity_routines.f90:1196-1206, :1235-1245); t[i + 6j].
__device__ __forceinline__ void voigt_push_matrix(const double *F, double *t)
{
    const int va[6] = {0, 1, 2, 0, 0, 1}, vb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            double v = F[va[i] + 3 * va[j]] * F[vb[i] + 3 * vb[j]];
            if (j >= 3) v = v + F[va[i] + 3 * vb[j]] * F[vb[i] + 3 * va[j]];
            t[i + 6 * j] = v;
        }
}

// Modified right Cauchy-Green tensor and the MR isochoric fictitious 2nd
// Piola tensor, Voigt (:4365-4384, :643-679).
__device__ __forceinline__ void mr_fictitious_piola(const double *F, double Jznu, double c10, double c01,
                                                    double *mod, double *S)
{
    const double s = rcbrt(Jznu);  // Jznu^(-1/3)
    for (int i = 0; i < 9; ++i) mod[i] = F[i] * s;
    double azl[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            azl[i + 3 * j] = mod[0 + 3 * i] * mod[0 + 3 * j] + mod[1 + 3 * i] * mod[1 + 3 * j] +
                             mod[2 + 3 * i] * mod[2 + 3 * j];
    const double I1 = azl[0] + azl[4] + azl[8];
    const double t1 = -2.0 * c01, t2 = 2.0 * (c10 + I1 * c01);
    S[0] = t1 * azl[0] + t2;
    S[1] = t1 * azl[4] + t2;
    S[2] = t1 * azl[8] + t2;
    S[3] = t1 * azl[1];
    S[4] = t1 * azl[2];
    S[5] = t1 * azl[5];
}

// FINITE_ELASTICITY_GAUSS_STRESS_TENSOR, MR branch
// (finite_elasticity_routines.f90:4342-4410): Cauchy stress, Voigt.
__device__ __forceinline__ void mr_cauchy_stress(const double *F, double Jznu, double p, double c10, double c01,
                                                 double *sigma)
{
    double mod[9], S[6], t[36];
    mr_fictitious_piola(F, Jznu, c10, c01, mod, S);
    voigt_push_matrix(mod, t);
    const double rJ = 1.0 / Jznu;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) v += t[i + 6 * j] * S[j];
        sigma[i] = v * rJ;
    }
    const double ot = (sigma[0] + sigma[1] + sigma[2]) / 3.0;
    for (int i = 0; i < 3; ++i) sigma[i] = sigma[i] - ot + p;
}

// FINITE_ELASTICITY_GAUSS_CAUCHY_TENSOR, INCOMPRESSIBLE_MOONEY_RIVLIN branch
// (finite_elasticity_routines.f90:3476-3481, 3522-3549, 4316-4322): AZL =
// F^T F, AZU = inv(AZL) with I3 = det AZL and Jznu = sqrt(I3); I1 = tr AZL,
// I2 = (I1^2 - tr AZL^2)/2; S = 2 Jznu^(-2/3) ((c1 + c2 I1) I - c2 AZL
// - (c1 I1 + 2 c2 I2 - 3/2 p Jznu^(5/3))/3 AZU); sigma = F S F^T / Jznu
// (Voigt).  Returns Jznu.
__device__ __forceinline__ double imr_cauchy_stress(const double *F, double p, double c10, double c01,
                                                    double *sigma)
{
    double azl[9], azu[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            azl[i + 3 * j] = F[0 + 3 * i] * F[0 + 3 * j] + F[1 + 3 * i] * F[1 + 3 * j] + F[2 + 3 * i] * F[2 + 3 * j];
    const double J = sqrt(invert3(azl, azu));
    const double I1 = azl[0] + azl[4] + azl[8];
    double sq[3];
    for (int i = 0; i < 3; ++i) sq[i] = azl[i] * azl[3 * i] + azl[i + 3] * azl[1 + 3 * i] + azl[i + 6] * azl[2 + 3 * i];
    const double I2 = 0.5 * (I1 * I1 - sq[0] - sq[1] - sq[2]);
    const double coef = 2.0 * pow(J, -2.0 / 3.0);
    const double t = (c10 * I1 + 2.0 * c01 * I2 - 1.5 * p * pow(J, 5.0 / 3.0)) / 3.0;
    double S[9], FS[9];
    for (int k = 0; k < 9; ++k) S[k] = coef * (((k % 4) == 0 ? c10 + c01 * I1 : 0.0) - c01 * azl[k] - t * azu[k]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            FS[i + 3 * j] = F[i] * S[3 * j] + F[i + 3] * S[1 + 3 * j] + F[i + 6] * S[2 + 3 * j];
    const int va[6] = {0, 1, 2, 0, 0, 1}, vb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const int i = va[a], j = vb[a];
        sigma[a] = (FS[i] * F[j] + FS[i + 3] * F[j + 3] + FS[i + 6] * F[j + 6]) / J;
    }
    return J;
}

// FINITE_ELASTICITY_GAUSS_ELASTICITY_TENSOR, MR branch
// (finite_elasticity_routines.f90:609-740): spatial elasticity tensor c
// (6x6 column-major) and Cauchy stress sigma (Voigt, pressure included).
__device__ __forceinline__ void mr_elasticity_tensor(const double *F, double Jznu, double p, double c10, double c01,
                                                     double *c, double *sigma)
{
    const double ONETHIRD = 1.0 / 3.0, TWOTHIRDS = 2.0 / 3.0;
    double mod[9], S[6], t[36];
    mr_fictitious_piola(F, Jznu, c10, c01, mod, S);
    voigt_push_matrix(mod, t);
    const double rJ = 1.0 / Jznu;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) v += t[i + 6 * j] * S[j];
        sigma[i] = v * rJ;
    }
    // fictitious material tangent: 4 c01 off-diagonal in the normal block,
    // -2 c01 on the shear diagonal (:686-697); push-forward t E t^T / J.
    const double e1 = 4.0 * c01, e2 = -2.0 * c01;
    double tE[36];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        tE[i + 0] = t[i + 6 * 1] * e1 + t[i + 6 * 2] * e1;
        tE[i + 6] = t[i + 6 * 0] * e1 + t[i + 6 * 2] * e1;
        tE[i + 12] = t[i + 6 * 0] * e1 + t[i + 6 * 1] * e1;
        tE[i + 18] = t[i + 18] * e2;
        tE[i + 24] = t[i + 24] * e2;
        tE[i + 30] = t[i + 30] * e2;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) v += tE[i + 6 * k] * t[j + 6 * k];
            c[i + 6 * j] = v * rJ;
        }
    const double trace = sigma[0] + sigma[1] + sigma[2];
    for (int i = 0; i < 3; ++i) sigma[i] = sigma[i] - ONETHIRD * trace;
    const double ttr = TWOTHIRDS * trace;
    for (int i = 0; i < 3; ++i) c[i + 6 * i] = c[i + 6 * i] + ttr;
    for (int i = 3; i < 6; ++i) c[i + 6 * i] = c[i + 6 * i] + ttr * 0.5;
    // deviatoric projection of the normal block and the coupling block
    // (:705-717), DEV = I - 1/3 11^T on the normal block.
    double a[9], b[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += c[i + 6 * k] * (k == j ? TWOTHIRDS : -ONETHIRD);
            a[i + 3 * j] = v;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += (i == k ? TWOTHIRDS : -ONETHIRD) * a[k + 3 * j];
            b[i + 3 * j] = v;
        }
    double tmp[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += (i == k ? TWOTHIRDS : -ONETHIRD) * c[k + 6 * (j + 3)];
            tmp[i + 3 * j] = v;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            c[i + 6 * j] = b[i + 3 * j];
            c[i + 6 * (j + 3)] = tmp[i + 3 * j];
            c[(j + 3) + 6 * i] = tmp[i + 3 * j];
        }
    // - (2/3 1 (x) sigma_dev + sigma_dev (x) 2/3 1)  (:719-724)
    for (int j = 0; j < 6; ++j)
        for (int i = 0; i < 6; ++i) {
            const double ui = i < 3 ? TWOTHIRDS : 0.0, uj = j < 3 ? TWOTHIRDS : 0.0;
            c[i + 6 * j] = c[i + 6 * j] - (ui * sigma[j] + uj * sigma[i]);
        }
    // volumetric part (:726-739)
    for (int i = 0; i < 3; ++i) sigma[i] = sigma[i] + p;
    for (int i = 0; i < 6; ++i) c[i + 6 * i] = c[i + 6 * i] - p;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) c[i + 6 * j] = c[i + 6 * j] + p;
}

// Position of column `col` in CSR row `row` (columns sorted ascending);
// the assembly guarantees presence, so a miss returns -1 and is skipped.
__device__ __forceinline__ int64_t csr_slot(const int64_t *__restrict__ row_ptr, const int *__restrict__ col_idx,
                                            int row, int col)
{
    int64_t lo = row_ptr[row], hi = row_ptr[row + 1] - 1;
    while (lo <= hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int v = col_idx[mid];
        if (v == col) return mid;
        if (v < col) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// Partial dot product of CSR entries [k0, k1) with x for lane `lane` of a row
// handled by kLanes lanes: four entries in flight per lane, fixed order.
template <typename VT, int kLanes>
__device__ __forceinline__ double csr_row_dot(int64_t k0, int64_t k1, int lane, const int *__restrict__ col,
                                              const VT *__restrict__ val, const double *__restrict__ x)
{
    double a0 = 0.0, a1 = 0.0;
    int64_t k = k0 + lane;
    for (; k + 3 * kLanes < k1; k += 4 * kLanes) {
        const int c0 = col[k], c1 = col[k + kLanes], c2 = col[k + 2 * kLanes], c3 = col[k + 3 * kLanes];
        const double v0 = static_cast<double>(val[k]), v1 = static_cast<double>(val[k + kLanes]);
        const double v2 = static_cast<double>(val[k + 2 * kLanes]), v3 = static_cast<double>(val[k + 3 * kLanes]);
        a0 += v0 * x[c0];
        a1 += v1 * x[c1];
        a0 += v2 * x[c2];
        a1 += v3 * x[c3];
    }
    for (; k < k1; k += kLanes) a0 += static_cast<double>(val[k]) * x[col[k]];
    return a0 + a1;
}

// Node i's three displacement rows times x over their displacement columns
// (rows with the node layout: row c of node i holds three blocks of `deg`
// entries, block c' over the same sorted neighbour nodes j, from s[c]):
// a[c] += sum_p sum_c' K[(c, i), (c', j_p)] x[c' nq + j_p].  kLanes lanes per
// node; lane `lane` takes neighbours p = lane + kLanes (u + kU k), and the
// loads of its kU neighbours are all issued before their products, so each
// lane keeps kU independent index -> vector load chains in flight.
template <typename VT, int kLanes, int kU>
__device__ __forceinline__ void node_rows_uu(const int64_t (&s)[3], int deg, int lane, int64_t nq64,
                                             const int *__restrict__ col_idx, const VT *__restrict__ vals,
                                             const double *__restrict__ x, double (&a)[3])
{
    for (int p0 = lane; p0 < deg; p0 += kLanes * kU) {
        int j[kU];
        bool ok[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            ok[u] = p0 + u * kLanes < deg;
            j[u] = ok[u] ? col_idx[s[0] + p0 + u * kLanes] : 0;
        }
        double v[kU][9];
#pragma unroll
        for (int u = 0; u < kU; ++u)
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    v[u][3 * r + c] = ok[u] ? static_cast<double>(vals[s[r] + c * deg + p0 + u * kLanes]) : 0.0;
        double xv[kU][3];
#pragma unroll
        for (int u = 0; u < kU; ++u)
#pragma unroll
            for (int c = 0; c < 3; ++c) xv[u][c] = ok[u] ? x[c * nq64 + j[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < kU; ++u)
#pragma unroll
            for (int r = 0; r < 3; ++r)
                a[r] += v[u][3 * r] * xv[u][0] + v[u][3 * r + 1] * xv[u][1] + v[u][3 * r + 2] * xv[u][2];
    }
}

}  // namespace oc
