/*
 * oc_oracle.c -- FP64 CPU restatement of OpenCMISS-Iron's finite element
 * assembly path.  TEST INFRASTRUCTURE ONLY: this file is the parity checker
 * for the HIP product library and the timed CPU baseline.  The product path
 * never links, loads or calls it.
 *
 * Every function names the reference file:line it restates.  Fortran
 * column-major 3x3 matrices are stored as A[i + 3*j] == A(i+1,j+1).
 * Loop orders follow the reference loops so the summation order matches
 * the reference element by element.
 */
#include "oc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ZERO_TOL (5.0 * 2.220446049250313e-16) /* constants.f90:68 */

/* ------------------------------------------------------------------------ */
/* splitmix64 uniform in [0,1): the seeded stream the mesh generator uses.  */
double oco_uniform(uint64_t seed, uint64_t index)
{
    uint64_t z = seed + (index + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------ */
/* GAUSS_LEGENDRE, basis_routines.f90:5826-5875 (closed-form table on [0,1]) */
int oco_gauss_legendre(int n, double *x, double *w)
{
    if (n < 1 || n > 4) return -1;
    const double s3 = 1.0 / sqrt(3.0), s6 = sqrt(0.6);
    const double a = sqrt((3.0 + 2.0 * sqrt(6.0 / 5.0)) / 7.0);
    const double b = sqrt((3.0 - 2.0 * sqrt(6.0 / 5.0)) / 7.0);
    const double xig[10] = {0.5,
                            (-s3 + 1.0) / 2.0, (s3 + 1.0) / 2.0,
                            (-s6 + 1.0) / 2.0, 0.5, (s6 + 1.0) / 2.0,
                            (-a + 1.0) / 2.0, (-b + 1.0) / 2.0, (b + 1.0) / 2.0, (a + 1.0) / 2.0};
    const double wig[10] = {1.0,
                            0.5, 0.5,
                            2.5 / 9.0, 4.0 / 9.0, 2.5 / 9.0,
                            (18.0 - sqrt(30.0)) / 72.0, (18.0 + sqrt(30.0)) / 72.0,
                            (18.0 + sqrt(30.0)) / 72.0, (18.0 - sqrt(30.0)) / 72.0};
    const int start[4] = {0, 1, 3, 6};
    for (int i = 0; i < n; ++i) {
        x[i] = xig[start[n - 1] + i];
        w[i] = wig[start[n - 1] + i];
    }
    return 0;
}

/* LAGRANGE_LINEAR_EVALUATE / LAGRANGE_QUADRATIC_EVALUATE,
 * basis_routines.f90:7015-7137.  node 1-based, deriv 0 = value, 1 = d/dxi. */
static double lagrange1d(int order, int node, int deriv, double xi)
{
    if (order == 1) {
        if (deriv == 0) return node == 1 ? 1.0 - xi : xi;
        return node == 1 ? -1.0 : 1.0;
    }
    if (deriv == 0) {
        switch (node) {
        case 1: return 1.0 - 3.0 * xi + 2.0 * xi * xi;
        case 2: return 4.0 * xi * (1.0 - xi);
        default: return xi * (xi + xi - 1.0);
        }
    }
    switch (node) {
    case 1: return 4.0 * xi - 3.0;
    case 2: return 4.0 - 8.0 * xi;
    default: return 4.0 * xi - 1.0;
    }
}

/* BASIS_QUADRATURE_CREATE, basis_routines.f90:2770-2890, with
 * BASIS_LHTP_BASIS_EVALUATE_DP (:2251-2381).  Gauss points ordered
 * ng = i + (j-1 + (k-1)*n2)*n1 (xi1 fastest); basis nodes xi1 fastest
 * (node position index, basis_routines.f90:1594-1630). */
int oco_basis_tables(int order, int ng1, double *weights, double *positions, double *fns)
{
    if (order < 1 || order > 2) return -1;
    double gx[4], gw[4];
    if (oco_gauss_legendre(ng1, gx, gw)) return -1;
    const int nx = order + 1, ns = nx * nx * nx;
    int g = 0;
    for (int k = 0; k < ng1; ++k)
        for (int j = 0; j < ng1; ++j)
            for (int i = 0; i < ng1; ++i, ++g) {
                const double xi[3] = {gx[i], gx[j], gx[k]};
                weights[g] = gw[i] * gw[j] * gw[k];
                positions[3 * g + 0] = xi[0];
                positions[3 * g + 1] = xi[1];
                positions[3 * g + 2] = xi[2];
                int s = 0;
                for (int n3 = 1; n3 <= nx; ++n3)
                    for (int n2 = 1; n2 <= nx; ++n2)
                        for (int n1 = 1; n1 <= nx; ++n1, ++s) {
                            const int pos[3] = {n1, n2, n3};
                            for (int nu = 0; nu < 4; ++nu) {
                                double v = 1.0;
                                for (int d = 0; d < 3; ++d)
                                    v *= lagrange1d(order, pos[d], nu == d + 1, xi[d]);
                                fns[((size_t)g * ns + s) * 4 + nu] = v;
                            }
                        }
            }
    return 0;
}

/* INVERT_FULL_DP 3x3 case, maths.f90:1455-1521 (cofactors / det). */
int oco_invert3(const double *A, double *B, double *det)
{
#define a(i, j) A[(i - 1) + 3 * (j - 1)]
#define b(i, j) B[(i - 1) + 3 * (j - 1)]
    double d = a(1, 1) * a(2, 2) * a(3, 3) + a(1, 2) * a(2, 3) * a(3, 1) + a(1, 3) * a(2, 1) * a(3, 2) -
               a(1, 1) * a(3, 2) * a(2, 3) - a(2, 1) * a(1, 2) * a(3, 3) - a(3, 1) * a(2, 2) * a(1, 3);
    *det = d;
    if (fabs(d) > ZERO_TOL) {
        b(1, 1) = (a(2, 2) * a(3, 3) - a(3, 2) * a(2, 3)) / d;
        b(2, 1) = (a(2, 3) * a(3, 1) - a(3, 3) * a(2, 1)) / d;
        b(3, 1) = (a(2, 1) * a(3, 2) - a(3, 1) * a(2, 2)) / d;
        b(1, 2) = (a(3, 2) * a(1, 3) - a(1, 2) * a(3, 3)) / d;
        b(2, 2) = (a(3, 3) * a(1, 1) - a(1, 3) * a(3, 1)) / d;
        b(3, 2) = (a(3, 1) * a(1, 2) - a(1, 1) * a(3, 2)) / d;
        b(1, 3) = (a(1, 2) * a(2, 3) - a(2, 2) * a(1, 3)) / d;
        b(2, 3) = (a(1, 3) * a(2, 1) - a(2, 3) * a(1, 1)) / d;
        b(3, 3) = (a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2)) / d;
        return 0;
    }
    for (int i = 0; i < 9; ++i) B[i] = 0.0; /* FLAG_WARNING path */
    return 1;
#undef a
#undef b
}

/* ------------------------------------------------------------------------ */
/* Regular generated mesh, generated_mesh_routines.f90:1386-1612.          */
int64_t oco_regular_num_nodes(int order, int n1, int n2, int n3)
{
    return (int64_t)(order * n1 + 1) * (order * n2 + 1) * (order * n3 + 1);
}

int oco_regular_elements(int order, int n1, int n2, int n3, int32_t *elem_nodes)
{
    if (order < 1 || order > 2) return -1;
    const int nx = order + 1, nen = nx * nx * nx;
    const int64_t T1 = order * n1 + 1, T2 = order * n2 + 1;
    int64_t e = 0;
    for (int e3 = 0; e3 < n3; ++e3)
        for (int e2 = 0; e2 < n2; ++e2)
            for (int e1 = 0; e1 < n1; ++e1, ++e) {
                const int64_t np = (int64_t)e1 * order + (int64_t)e2 * T1 * order + (int64_t)e3 * T1 * T2 * order;
                int k = 0;
                for (int c = 0; c < nx; ++c)
                    for (int b = 0; b < nx; ++b)
                        for (int a = 0; a < nx; ++a, ++k)
                            elem_nodes[e * nen + k] = (int32_t)(np + a + b * T1 + c * T1 * T2);
            }
    return 0;
}

/* GeneratedMesh_RegularGeometricParametersCalculate,
 * generated_mesh_routines.f90:3360-3445: x = origin + (i-1)*extent/(T-1).
 * Jitter of interior nodes is the synthetic-input rule of this project. */
int oco_regular_coordinates(int order, int n1, int n2, int n3, double jitter, uint64_t seed,
                            double *coords)
{
    const int64_t T[3] = {order * n1 + 1, order * n2 + 1, order * n3 + 1};
    const int64_t N = T[0] * T[1] * T[2];
    const double delta[3] = {1.0 / (double)(T[0] - 1), 1.0 / (double)(T[1] - 1), 1.0 / (double)(T[2] - 1)};
    const double h[3] = {1.0 / n1, 1.0 / n2, 1.0 / n3};
    for (int64_t p = 0; p < N; ++p) {
        const int64_t idx[3] = {p % T[0], (p / T[0]) % T[1], p / (T[0] * T[1])};
        const int interior = idx[0] > 0 && idx[0] < T[0] - 1 && idx[1] > 0 && idx[1] < T[1] - 1 &&
                             idx[2] > 0 && idx[2] < T[2] - 1;
        for (int c = 0; c < 3; ++c) {
            double v = 0.0 + (double)idx[c] * delta[c];
            if (interior && jitter != 0.0) {
                const double u = oco_uniform(seed, (uint64_t)(3 * p + c));
                v += (2.0 * u - 1.0) * jitter * h[c];
            }
            coords[c * N + p] = v;
        }
    }
    return 0;
}

/* Trilinear component of a triquadratic regular mesh: corner nodes take the
 * quadratic user numbers (GeneratedMesh_RegularComponentNodesToUserNumbers,
 * generated_mesh_routines.f90:5155-5180) and the component's mesh nodes are
 * the sorted global numbers (MeshTopologyNodesCalculate,
 * mesh_routines.f90:9006-9060), so corner (i,j,k) has rank
 * i + j*(n1+1) + k*(n1+1)*(n2+1). */
int oco_regular_pressure_elements(int n1, int n2, int n3, int32_t *elem_nodes_l)
{
    return oco_regular_elements(1, n1, n2, n3, elem_nodes_l);
}

/* ------------------------------------------------------------------------ */
/* CSR pattern: EquationsMatrix_StructureCalculate / JacobianMatrix_Structure
 * Calculate, equations_matrices_routines.f90:4572-4780, 5048-5230, with
 * LIST_REMOVE_DUPLICATES (sort + unique, lists.f90:2647-2690).  Row DOF
 * (component c, node np): union over the elements surrounding np of every
 * DOF of every component of those elements.                               */
static int cmp_i32(const void *a, const void *b)
{
    const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

int oco_pattern(int ncomp, int nelem, const int32_t *nn, const int32_t *nen,
                const int32_t *const *conn, int64_t *nnz_out, int64_t *row_ptr, int32_t *col_idx)
{
    if (ncomp < 1 || ncomp > 8) return -1;
    int64_t off[9] = {0};
    int maxcol = 0;
    for (int c = 0; c < ncomp; ++c) {
        off[c + 1] = off[c] + nn[c];
        maxcol += nen[c];
    }
    const int64_t ndof = off[ncomp];
    /* node -> surrounding elements, per component (ascending element order) */
    int64_t *sptr[8];
    int32_t *sel[8];
    int maxsurr = 0;
    for (int c = 0; c < ncomp; ++c) {
        sptr[c] = (int64_t *)calloc((size_t)nn[c] + 1, sizeof(int64_t));
        for (int64_t e = 0; e < nelem; ++e)
            for (int k = 0; k < nen[c]; ++k) sptr[c][conn[c][e * nen[c] + k] + 1]++;
        for (int64_t i = 0; i < nn[c]; ++i) {
            if (sptr[c][i + 1] > maxsurr) maxsurr = (int)sptr[c][i + 1];
            sptr[c][i + 1] += sptr[c][i];
        }
        sel[c] = (int32_t *)malloc(sizeof(int32_t) * (size_t)(sptr[c][nn[c]] + 1));
        int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)nn[c]);
        memcpy(fill, sptr[c], sizeof(int64_t) * (size_t)nn[c]);
        /* ascending element order, as the sorted surrounding-element lists of
         * mesh_routines.f90:5929-6011 */
        for (int64_t e = 0; e < nelem; ++e)
            for (int k = 0; k < nen[c]; ++k) sel[c][fill[conn[c][e * nen[c] + k]]++] = (int32_t)e;
        free(fill);
    }
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (size_t)maxsurr * (size_t)maxcol + 1);
    int64_t nnz = 0;
    if (row_ptr) row_ptr[0] = 0;
    for (int c = 0; c < ncomp; ++c) {
        for (int64_t np = 0; np < nn[c]; ++np) {
            int cnt = 0;
            for (int64_t s = sptr[c][np]; s < sptr[c][np + 1]; ++s) {
                const int64_t e = sel[c][s];
                for (int c2 = 0; c2 < ncomp; ++c2)
                    for (int k = 0; k < nen[c2]; ++k)
                        buf[cnt++] = (int32_t)(off[c2] + conn[c2][e * nen[c2] + k]);
            }
            qsort(buf, (size_t)cnt, sizeof(int32_t), cmp_i32);
            int u = 0;
            for (int i = 0; i < cnt; ++i)
                if (u == 0 || buf[i] != buf[u - 1]) buf[u++] = buf[i];
            if (col_idx) memcpy(col_idx + nnz, buf, sizeof(int32_t) * (size_t)u);
            nnz += u;
            if (row_ptr) row_ptr[off[c] + np + 1] = nnz;
        }
    }
    (void)ndof;
    *nnz_out = nnz;
    free(buf);
    for (int c = 0; c < ncomp; ++c) {
        free(sptr[c]);
        free(sel[c]);
    }
    return 0;
}

/* MATRIX_VALUES_ADD_DP2, CSR branch, matrix_vector.f90:2839-2878:
 * bisection down to a 10-wide window (threshold :172), then linear scan. */
int oco_csr_add(const int64_t *row_ptr, const int32_t *col_idx, double *values, int nrow,
                const int32_t *rows, int ncol, const int32_t *cols, const double *Ke)
{
    for (int i = 0; i < nrow; ++i) {
        const int32_t r = rows[i];
        const int64_t rs = row_ptr[r], re = row_ptr[r + 1];
        int64_t lo = rs, up = re;
        int32_t prev = -1;
        for (int j = 0; j < ncol; ++j) {
            const int32_t c = cols[j];
            if (c <= prev) lo = rs;
            else up = re;
            prev = c;
            while (up - lo > 10) {
                const int64_t mid = (up + lo) / 2;
                if (col_idx[mid] > c) up = mid;
                else lo = mid;
            }
            int64_t loc = -1;
            const int64_t last = up < re ? up : re - 1;
            for (int64_t k = lo; k <= last; ++k)
                if (col_idx[k] == c) {
                    loc = k;
                    lo = k + 1;
                    break;
                }
            if (loc < 0) return -2; /* "Row .. and column .. does not exist in the matrix." */
            values[loc] += Ke[(size_t)i * ncol + j];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Shared per-basis tables built once (trilinear 2x2x2 for Laplace,
 * triquadratic and trilinear 3x3x3 for Mooney-Rivlin).                     */
typedef struct {
    int ready;
    double w[27], pos[81];
    double fq[27 * 27 * 4]; /* triquadratic at 3x3x3 */
    double fl[27 * 8 * 4];  /* trilinear at 3x3x3 */
    double wl[8], posl[24];
    double fl8[8 * 8 * 4];  /* trilinear at 2x2x2 */
} tables_t;

static tables_t T;

static void tables_init(void)
{
    if (T.ready) return;
    oco_basis_tables(2, 3, T.w, T.pos, T.fq);
    double w2[27], p2[81];
    oco_basis_tables(1, 3, w2, p2, T.fl);
    oco_basis_tables(1, 2, T.wl, T.posl, T.fl8);
    T.ready = 1;
}

/* COORDINATE_METRICS_CALCULATE (rectangular Cartesian, 3 xi, volume
 * Jacobian), coordinate_routines.f90:647-840:
 * GL = DX_DXI^T DX_DXI; GU = inv(GL); J = sqrt|det GL|; DXI_DX = inv(DX_DXI). */
typedef struct {
    double dx_dxi[9], gl[9], gu[9], jac, dxi_dx[9];
} metrics_t;

static void metrics_calc(const double *dx_dxi, metrics_t *m)
{
    memcpy(m->dx_dxi, dx_dxi, sizeof(double) * 9);
    for (int mi = 0; mi < 3; ++mi)
        for (int ni = 0; ni < 3; ++ni)
            m->gl[mi + 3 * ni] = dx_dxi[0 + 3 * mi] * dx_dxi[0 + 3 * ni] + dx_dxi[1 + 3 * mi] * dx_dxi[1 + 3 * ni] +
                                 dx_dxi[2 + 3 * mi] * dx_dxi[2 + 3 * ni];
    double det_gl, det_dx;
    oco_invert3(m->gl, m->gu, &det_gl);
    m->jac = sqrt(fabs(det_gl));
    oco_invert3(m->dx_dxi, m->dxi_dx, &det_dx);
}

/* FIELD_INTERPOLATE_GAUSS first derivatives (field_routines.f90:6071-,
 * BASIS_INTERPOLATE_GAUSS_DP basis_routines.f90:1123-1185):
 * DX_DXI(x,ni) = sum_ns GAUSS_BASIS_FNS(ns, S_ni, g) * param(ns, x). */
static void interp_dxi(const double *fns, int ns, int g, const double *xe, double *dx_dxi)
{
    for (int x = 0; x < 3; ++x)
        for (int ni = 0; ni < 3; ++ni) {
            double s = 0.0;
            for (int n = 0; n < ns; ++n) s += fns[((size_t)g * ns + n) * 4 + 1 + ni] * xe[x * ns + n];
            dx_dxi[x + 3 * ni] = s;
        }
}

/* ------------------------------------------------------------------------ */
/* LAPLACE_EQUATION_FINITE_ELEMENT_CALCULATE, standard subtype,
 * Laplace_equations_routines.f90:474-571 (2x2x2 Gauss, trilinear).        */
int oco_laplace_element(const double *xe, double *Ke)
{
    tables_init();
    for (int i = 0; i < 64; ++i) Ke[i] = 0.0;
    for (int g = 0; g < 8; ++g) {
        double dx[9];
        metrics_t m;
        interp_dxi(T.fl8, 8, g, xe, dx);
        metrics_calc(dx, &m);
        const double rwg = m.jac * T.wl[g];
        for (int ms = 0; ms < 8; ++ms)
            for (int ns = 0; ns < 8; ++ns) {
                double pm[3], pn[3];
                for (int ni = 0; ni < 3; ++ni) {
                    pm[ni] = T.fl8[((size_t)g * 8 + ms) * 4 + 1 + ni];
                    pn[ni] = T.fl8[((size_t)g * 8 + ns) * 4 + 1 + ni];
                }
                double sum = 0.0;
                for (int mi = 0; mi < 3; ++mi)
                    for (int ni = 0; ni < 3; ++ni) sum = sum + pm[mi] * pn[ni] * m.gu[mi + 3 * ni];
                Ke[ms * 8 + ns] = Ke[ms * 8 + ns] + sum * rwg;
            }
    }
    return 0;
}

/* EQUATIONS_SET_ASSEMBLE_STATIC_LINEAR_FEM, equations_set_routines.f90:
 * 1295-1492: zero, then elements 1..E in order, element calculate, add. */
int oco_laplace_assemble(int nelem, const int32_t *elem_nodes, int nnode, const double *coords,
                         const int64_t *row_ptr, const int32_t *col_idx, double *values, double *rhs)
{
    const int64_t nnz = row_ptr[nnode];
    memset(values, 0, sizeof(double) * (size_t)nnz);
    if (rhs) memset(rhs, 0, sizeof(double) * (size_t)nnode);
    double xe[24], Ke[64];
    for (int64_t e = 0; e < nelem; ++e) {
        const int32_t *en = elem_nodes + e * 8;
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 8; ++k) xe[c * 8 + k] = coords[(int64_t)c * nnode + en[k]];
        oco_laplace_element(xe, Ke);
        int rc = oco_csr_add(row_ptr, col_idx, values, 8, en, 8, en, Ke);
        if (rc) return rc;
        /* RHS element vector is 0 (:542) */
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Mooney-Rivlin constitutive routines.                                      */
static const int V2T[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}}; /* constants.f90:285 */
static const int T2V[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};                 /* constants.f90:284 */

/* t(6x6) of FINITE_ELASTICITY_PUSH_STRESS_TENSOR / _ELASTICITY_TENSOR,
 * finite_elasticity_routines.f90:1196-1206 and :1235-1245.  t[i + 6j]. */
static void push_matrix(const double *F, double *t)
{
#define f(i, j) F[(i) + 3 * (j)]
    for (int j = 0; j < 6; ++j)
        for (int i = 0; i < 6; ++i) {
            const int ai = V2T[i][0], bi = V2T[i][1], aj = V2T[j][0], bj = V2T[j][1];
            double v = f(ai, aj) * f(bi, bj);
            if (j >= 3) v = v + f(ai, bj) * f(bi, aj);
            t[i + 6 * j] = v;
        }
#undef f
}

/* MOD_DZDNU = F*J^(-1/3); AZL = MOD^T MOD (:4365-4367, :646-648). */
static void modified_c(const double *F, double Jznu, double *mod, double *azl)
{
    const double s = pow(Jznu, -1.0 / 3.0);
    for (int i = 0; i < 9; ++i) mod[i] = F[i] * s;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += mod[k + 3 * i] * mod[k + 3 * j];
            azl[i + 3 * j] = v;
        }
}

/* Isochoric fictitious 2nd Piola tensor, Voigt (:4375-4384 / :670-679). */
static void mr_piola(const double *azl, double c10, double c01, double *S)
{
    const double I1 = azl[0] + azl[4] + azl[8];
    const double t1 = -2.0 * c01, t2 = 2.0 * (c10 + I1 * c01);
    S[0] = t1 * azl[0] + t2;
    S[1] = t1 * azl[4] + t2;
    S[2] = t1 * azl[8] + t2;
    S[3] = t1 * azl[1]; /* AZL(2,1) */
    S[4] = t1 * azl[2]; /* AZL(3,1) */
    S[5] = t1 * azl[5]; /* AZL(3,2) */
}

/* FINITE_ELASTICITY_GAUSS_STRESS_TENSOR, MR branch,
 * finite_elasticity_routines.f90:4342-4410, push :1221-1253. */
int oco_mr_stress(const double *F, double Jznu, double p, double c10, double c01, double *sigma)
{
    double mod[9], azl[9], S[6], t[36];
    modified_c(F, Jznu, mod, azl);
    mr_piola(azl, c10, c01, S);
    push_matrix(mod, t);
    for (int i = 0; i < 6; ++i) {
        double v = 0.0;
        for (int j = 0; j < 6; ++j) v += t[i + 6 * j] * S[j];
        sigma[i] = v / Jznu;
    }
    const double ot = (sigma[0] + sigma[1] + sigma[2]) / 3.0;
    for (int i = 0; i < 3; ++i) sigma[i] = sigma[i] - ot + p;
    return 0;
}

/* FINITE_ELASTICITY_GAUSS_ELASTICITY_TENSOR, MR branch,
 * finite_elasticity_routines.f90:609-740, push :1182-1214.
 * ctan is 6x6 column-major (ctan[i + 6j] = ELASTICITY_TENSOR(i+1,j+1)). */
int oco_mr_elasticity(const double *F, double Jznu, double p, double c10, double c01,
                      double *ctan, double *sigma)
{
    const double ONETHIRD = 1.0 / 3.0, TWOTHIRDS = 2.0 / 3.0;
    const double UNITY_DIAG[6] = {1.0, 1.0, 1.0, 0.5, 0.5, 0.5};
    const double TWOTHIRDS_UNITY[6] = {TWOTHIRDS, TWOTHIRDS, TWOTHIRDS, 0.0, 0.0, 0.0};
    double dev[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) dev[i + 3 * j] = (i == j) ? TWOTHIRDS : -ONETHIRD;
    double mod[9], azl[9], S[6], t[36], E[36];
    modified_c(F, Jznu, mod, azl);
    mr_piola(azl, c10, c01, S);
    for (int i = 0; i < 36; ++i) E[i] = 0.0;
    const double te1 = 4.0 * c01, te2 = -2.0 * c01;
    E[1 + 6 * 0] = te1; E[2 + 6 * 0] = te1; E[0 + 6 * 1] = te1;
    E[2 + 6 * 1] = te1; E[0 + 6 * 2] = te1; E[1 + 6 * 2] = te1;
    E[3 + 6 * 3] = te2; E[4 + 6 * 4] = te2; E[5 + 6 * 5] = te2;
    push_matrix(mod, t);
    /* stress push */
    double s2[6];
    for (int i = 0; i < 6; ++i) {
        double v = 0.0;
        for (int j = 0; j < 6; ++j) v += t[i + 6 * j] * S[j];
        s2[i] = v / Jznu;
    }
    /* elasticity push: MATMUL(MATMUL(t,E),TRANSPOSE(t))/Jznu */
    double tE[36], c[36];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double v = 0.0;
            for (int k = 0; k < 6; ++k) v += t[i + 6 * k] * E[k + 6 * j];
            tE[i + 6 * j] = v;
        }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double v = 0.0;
            for (int k = 0; k < 6; ++k) v += tE[i + 6 * k] * t[j + 6 * k];
            c[i + 6 * j] = v / Jznu;
        }
    const double trace = s2[0] + s2[1] + s2[2];
    for (int i = 0; i < 3; ++i) s2[i] = s2[i] - ONETHIRD * trace;
    const double ttr = TWOTHIRDS * trace;
    for (int i = 0; i < 6; ++i) c[i + 6 * i] = c[i + 6 * i] + ttr * UNITY_DIAG[i];
    /* c(1:3,1:3) = DEV*c(1:3,1:3)*DEV */
    double a[9], b[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += c[i + 6 * k] * dev[k + 3 * j];
            a[i + 3 * j] = v;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += dev[i + 3 * k] * a[k + 3 * j];
            b[i + 3 * j] = v;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) c[i + 6 * j] = b[i + 3 * j];
    /* TEMP = DEV*c(1:3,4:6); c(1:3,4:6)=TEMP; c(4:6,1:3)=TEMP^T */
    double tmp[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += dev[i + 3 * k] * c[k + 6 * (j + 3)];
            tmp[i + 3 * j] = v;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            c[i + 6 * (j + 3)] = tmp[i + 3 * j];
            c[(j + 3) + 6 * i] = tmp[i + 3 * j];
        }
    for (int j = 0; j < 6; ++j)
        for (int i = 0; i < 6; ++i)
            c[i + 6 * j] = c[i + 6 * j] - (TWOTHIRDS_UNITY[i] * s2[j] + TWOTHIRDS_UNITY[j] * s2[i]);
    for (int i = 0; i < 3; ++i) s2[i] = s2[i] + p;
    for (int i = 0; i < 6; ++i) c[i + 6 * i] = c[i + 6 * i] - p;
    c[1 + 6 * 0] += p; c[2 + 6 * 0] += p; c[0 + 6 * 1] += p;
    c[2 + 6 * 1] += p; c[0 + 6 * 2] += p; c[1 + 6 * 2] += p;
    memcpy(ctan, c, sizeof(double) * 36);
    memcpy(sigma, s2, sizeof(double) * 6);
    return 0;
}

/* Per-Gauss-point kinematics shared by residual and Jacobian:
 * interpolate z and X (:1493-1500 / :955-963), metrics, F = dz/dxi * dxi/dX
 * (FiniteElasticity_GaussDeformationGradientTensor :3368-3425 with
 * Coordinates_MaterialSystemCalculate coordinate_routines.f90:3883-3987,
 * no fibre field), Jznu = J_z/J_X, DPHIDZ (:1509-1525), pressure P. */
typedef struct {
    double F[9], Jznu, jz, jx, P, D[3 * 27]; /* D[mh*27+ns] = DPHIDZ(mh,ns) */
} gauss_state_t;

static void gauss_state(int g, const double *Xe, const double *ze, const double *pe, gauss_state_t *s)
{
    double dz[9], dX[9], dxidnu[9], det;
    metrics_t mz, mX;
    interp_dxi(T.fq, 27, g, ze, dz);
    metrics_calc(dz, &mz);
    interp_dxi(T.fq, 27, g, Xe, dX);
    metrics_calc(dX, &mX);
    oco_invert3(dX, dxidnu, &det);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += dz[i + 3 * k] * dxidnu[k + 3 * j];
            s->F[i + 3 * j] = v;
        }
    s->jz = mz.jac;
    s->jx = mX.jac;
    s->Jznu = mz.jac / mX.jac;
    for (int ns = 0; ns < 27; ++ns)
        for (int mh = 0; mh < 3; ++mh) {
            double sum = 0.0;
            for (int mi = 0; mi < 3; ++mi) sum = sum + mz.dxi_dx[mi + 3 * mh] * T.fq[((size_t)g * 27 + ns) * 4 + 1 + mi];
            s->D[mh * 27 + ns] = sum;
        }
    double P = 0.0;
    for (int n = 0; n < 8; ++n) P += T.fl[((size_t)g * 8 + n) * 4 + 0] * pe[n];
    s->P = P;
}

/* FiniteElasticity_FiniteElementResidualEvaluate, MR branch,
 * finite_elasticity_routines.f90:1488-1575. */
int oco_mr_element_residual(const double *Xe, const double *ze, const double *pe,
                            double c10, double c01, double *Re)
{
    tables_init();
    for (int i = 0; i < 89; ++i) Re[i] = 0.0;
    for (int g = 0; g < 27; ++g) {
        gauss_state_t s;
        gauss_state(g, Xe, ze, pe, &s);
        const double jgw = s.jz * T.w[g];
        double sig[6], jc[9];
        oco_mr_stress(s.F, s.Jznu, s.P, c10, c01, sig);
        for (int nh = 0; nh < 3; ++nh)
            for (int mh = 0; mh < 3; ++mh) jc[mh + 3 * nh] = jgw * sig[T2V[mh][nh]];
        int mhs = 0;
        for (int mh = 0; mh < 3; ++mh)
            for (int ms = 0; ms < 27; ++ms, ++mhs) {
                double dot = 0.0;
                for (int i = 0; i < 3; ++i) dot += s.D[i * 27 + ms] * jc[i + 3 * mh];
                Re[mhs] = Re[mhs] + dot;
            }
        const double jgwx = s.jx * T.w[g];
        const double tt = jgwx * (s.Jznu - 1.0);
        for (int ms = 0; ms < 8; ++ms, ++mhs) Re[mhs] = Re[mhs] + tt * T.fl[((size_t)g * 8 + ms) * 4 + 0];
    }
    return 0;
}

/* FiniteElasticity_FiniteElementJacobianEvaluate, analytic,
 * finite_elasticity_routines.f90:842-1175 (lower triangle, then mirror
 * :1149-1153).  Je row-major 89x89. */
int oco_mr_element_jacobian(const double *Xe, const double *ze, const double *pe,
                            double c10, double c01, double *Je)
{
    tables_init();
    const int ne = 89;
    for (int i = 0; i < ne * ne; ++i) Je[i] = 0.0;
#define JE(m, n) Je[(m) * ne + (n)]
    for (int g = 0; g < 27; ++g) {
        gauss_state_t s;
        gauss_state(g, Xe, ze, pe, &s);
        const double jgw = s.jz * T.w[g];
        double ct[36], sig[6], cau[9];
        oco_mr_elasticity(s.F, s.Jznu, s.P, c10, c01, ct, sig);
        for (int nh = 0; nh < 3; ++nh)
            for (int mh = 0; mh < 3; ++mh) cau[mh + 3 * nh] = sig[T2V[mh][nh]];
        /* 1) diagonal blocks mh == nh, ms >= ns (:1001-1017) */
        for (int nh = 0; nh < 3; ++nh) {
            double sub[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    sub[i + 3 * j] = jgw * (ct[T2V[i][nh] + 6 * T2V[j][nh]] + cau[i + 3 * j]);
            for (int ns = 0; ns < 27; ++ns) {
                double tv[3];
                for (int i = 0; i < 3; ++i) {
                    double v = 0.0;
                    for (int j = 0; j < 3; ++j) v += sub[i + 3 * j] * s.D[j * 27 + ns];
                    tv[i] = v;
                }
                const int nhs = nh * 27 + ns;
                for (int ms = ns; ms < 27; ++ms) {
                    double dot = 0.0;
                    for (int i = 0; i < 3; ++i) dot += s.D[i * 27 + ms] * tv[i];
                    JE(nh * 27 + ms, nhs) = JE(nh * 27 + ms, nhs) + dot;
                }
            }
        }
        /* 2) off-diagonal blocks mh > nh (:1019-1037) */
        static const int od1[3] = {0, 0, 1}, od2[3] = {1, 2, 2};
        for (int oh = 0; oh < 3; ++oh) {
            const int nh = od1[oh], mh = od2[oh];
            double sub[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) sub[i + 3 * j] = jgw * ct[T2V[i][mh] + 6 * T2V[j][nh]];
            for (int ns = 0; ns < 27; ++ns) {
                double tv[3];
                for (int i = 0; i < 3; ++i) {
                    double v = 0.0;
                    for (int j = 0; j < 3; ++j) v += sub[i + 3 * j] * s.D[j * 27 + ns];
                    tv[i] = v;
                }
                for (int ms = 0; ms < 27; ++ms) {
                    double dot = 0.0;
                    for (int i = 0; i < 3; ++i) dot += s.D[i * 27 + ms] * tv[i];
                    JE(mh * 27 + ms, nh * 27 + ns) = JE(mh * 27 + ms, nh * 27 + ns) + dot;
                }
            }
        }
        /* 3) pressure rows, node based (:1039-1054) */
        for (int nh = 0; nh < 3; ++nh)
            for (int ns = 0; ns < 27; ++ns) {
                const double jd = jgw * s.D[nh * 27 + ns];
                for (int ms = 0; ms < 8; ++ms) {
                    const double phi = T.fl[((size_t)g * 8 + ms) * 4 + 0];
                    JE(81 + ms, nh * 27 + ns) = JE(81 + ms, nh * 27 + ns) + jd * phi;
                }
            }
    }
    for (int n = 1; n < ne; ++n)
        for (int m = 0; m < n; ++m) JE(m, n) = JE(n, m);
#undef JE
    return 0;
}

/* EquationsSet_FiniteElementJacobianEvaluateFD, equations_set_routines.f90:
 * 3032-3150: column k = (R(x + delta e_k) - R(x)) / delta, columns in
 * component -> node order. */
typedef int (*element_residual_fn)(const double *, const double *, const double *, double, double, double *);

static void element_jacobian_fd(element_residual_fn res, const double *Xe, const double *ze, const double *pe,
                                double c10, double c01, double delta, double *Je)
{
    double R0[89], R1[89], z[81], p[8];
    res(Xe, ze, pe, c10, c01, R0);
    memcpy(z, ze, sizeof z);
    memcpy(p, pe, sizeof p);
    for (int col = 0; col < 89; ++col) {
        double *v = col < 81 ? &z[col] : &p[col - 81];
        const double orig = *v;
        *v = orig + delta;
        res(Xe, z, p, c10, c01, R1);
        *v = orig;
        for (int r = 0; r < 89; ++r) Je[r * 89 + col] = (R1[r] - R0[r]) / delta;
    }
}

int oco_mr_element_jacobian_fd(const double *Xe, const double *ze, const double *pe,
                               double c10, double c01, double delta, double *Je)
{
    element_jacobian_fd(oco_mr_element_residual, Xe, ze, pe, c10, c01, delta, Je);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Incompressible Mooney-Rivlin subtype                                      */
/* (EQUATIONS_SET_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE): Cauchy stress from  */
/* the 2nd Piola tensor with the pressure inside it, the generic 4-component */
/* residual, and the forward-difference Jacobian the reference uses for it.  */

/* FINITE_ELASTICITY_GAUSS_CAUCHY_TENSOR, INCOMPRESSIBLE_MOONEY_RIVLIN branch,
 * finite_elasticity_routines.f90:3476-3481 (AZL = F^T F, AZU = AZL^-1,
 * I3 = det AZL, Jznu = I3^(1/2)), :3522-3549 (I1 = tr AZL,
 * I2 = (I1^2 - tr AZL^2)/2, S = 2 Jznu^(-2/3) ((c1 + c2 I1) I - c2 AZL
 * - (c1 I1 + 2 c2 I2 - 3/2 p Jznu^(5/3))/3 AZU)) and :4316-4322
 * (sigma = F S F^T / Jznu).  sigma in Voigt order; returns Jznu in *jznu. */
int oco_imr_stress(const double *F, double p, double c10, double c01, double *sigma, double *jznu)
{
    double azl[9], azu[9], i3, S[9], tmp[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += F[k + 3 * i] * F[k + 3 * j];
            azl[i + 3 * j] = v;
        }
    oco_invert3(azl, azu, &i3);
    const double J = pow(i3, 0.5);
    const double I1 = azl[0] + azl[4] + azl[8];
    double sq[3];
    for (int i = 0; i < 3; ++i) {
        double v = 0.0;
        for (int k = 0; k < 3; ++k) v += azl[i + 3 * k] * azl[k + 3 * i];
        sq[i] = v;
    }
    const double I2 = 0.5 * (I1 * I1 - sq[0] - sq[1] - sq[2]);
    const double coef = 2.0 * pow(J, -2.0 / 3.0);
    const double t = (c10 * I1 + 2.0 * c01 * I2 - 1.5 * p * pow(J, 5.0 / 3.0)) / 3.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            S[i + 3 * j] = coef * ((i == j ? c10 + c01 * I1 : 0.0) - c01 * azl[i + 3 * j] - t * azu[i + 3 * j]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += F[i + 3 * k] * S[k + 3 * j];
            tmp[i + 3 * j] = v;
        }
    for (int a = 0; a < 6; ++a) {
        const int i = V2T[a][0], j = V2T[a][1];
        double v = 0.0;
        for (int k = 0; k < 3; ++k) v += tmp[i + 3 * k] * F[j + 3 * k];
        sigma[a] = v / J;
    }
    *jznu = J;
    return 0;
}

/* FiniteElasticity_FiniteElementResidualEvaluate, the 4-component branch that
 * the incompressible MR subtype takes (finite_elasticity_routines.f90:
 * 1654-1761): per Gauss point R(c, m) += w Jxxi Jznu sigma(c, c2)
 * DFDZ(m, c2) for c2 = 1..3 (DFDZ = dphi/dxi . inv(dz/dxi), FINITE_ELASTICITY_
 * GAUSS_DFDZ :4572-4643) and R_p(m) += w Jxxi psi_m (Jznu - 1), with Jxxi the
 * undeformed volume Jacobian and Jznu from the stress routine. */
int oco_imr_element_residual(const double *Xe, const double *ze, const double *pe,
                             double c10, double c01, double *Re)
{
    tables_init();
    for (int i = 0; i < 89; ++i) Re[i] = 0.0;
    for (int g = 0; g < 27; ++g) {
        gauss_state_t s;
        gauss_state(g, Xe, ze, pe, &s);
        double sig[6], J;
        oco_imr_stress(s.F, s.P, c10, c01, sig, &J);
        const double wjj = T.w[g] * s.jx * J;
        int idx = 0;
        for (int c = 0; c < 3; ++c)
            for (int m = 0; m < 27; ++m, ++idx)
                for (int c2 = 0; c2 < 3; ++c2) Re[idx] = Re[idx] + wjj * sig[T2V[c][c2]] * s.D[c2 * 27 + m];
        for (int m = 0; m < 8; ++m, ++idx)
            Re[idx] = Re[idx] + T.w[g] * s.jx * T.fl[((size_t)g * 8 + m) * 4 + 0] * (J - 1.0);
    }
    return 0;
}

int oco_imr_element_jacobian_fd(const double *Xe, const double *ze, const double *pe,
                                double c10, double c01, double delta, double *Je)
{
    element_jacobian_fd(oco_imr_element_residual, Xe, ze, pe, c10, c01, delta, Je);
    return 0;
}

static void mr_gather(int64_t e, const int32_t *elem_q, const int32_t *elem_l, int nq,
                      const double *Xq, const double *dofs, double *Xe, double *ze, double *pe, int32_t *edofs)
{
    const int32_t *eq = elem_q + e * 27, *el = elem_l + e * 8;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 27; ++k) {
            Xe[c * 27 + k] = Xq[(int64_t)c * nq + eq[k]];
            ze[c * 27 + k] = dofs[(int64_t)c * nq + eq[k]];
            edofs[c * 27 + k] = (int32_t)((int64_t)c * nq + eq[k]);
        }
    for (int k = 0; k < 8; ++k) {
        pe[k] = dofs[(int64_t)3 * nq + el[k]];
        edofs[81 + k] = (int32_t)((int64_t)3 * nq + el[k]);
    }
}

/* EQUATIONS_SET_RESIDUAL_EVALUATE_STATIC_FEM, equations_set_routines.f90:
 * 5483-5629: zero, elements in order, element residual, vector add. */
static int residual_assemble(element_residual_fn fn, int nelem, const int32_t *elem_q, const int32_t *elem_l,
                             int nq, int nl, const double *Xq, const double *dofs, double c10, double c01,
                             double *res)
{
    const int64_t ndof = (int64_t)3 * nq + nl;
    memset(res, 0, sizeof(double) * (size_t)ndof);
    double Xe[81], ze[81], pe[8], Re[89];
    int32_t ed[89];
    for (int64_t e = 0; e < nelem; ++e) {
        mr_gather(e, elem_q, elem_l, nq, Xq, dofs, Xe, ze, pe, ed);
        fn(Xe, ze, pe, c10, c01, Re);
        for (int i = 0; i < 89; ++i) res[ed[i]] += Re[i];
    }
    return 0;
}

int oco_mr_residual(int nelem, const int32_t *elem_q, const int32_t *elem_l, int nq, int nl,
                    const double *Xq, const double *dofs, double c10, double c01, double *res)
{
    return residual_assemble(oco_mr_element_residual, nelem, elem_q, elem_l, nq, nl, Xq, dofs, c10, c01, res);
}

/* The residual driver for the incompressible MR subtype (same element loop). */
int oco_imr_residual(int nelem, const int32_t *elem_q, const int32_t *elem_l, int nq, int nl,
                     const double *Xq, const double *dofs, double c10, double c01, double *res)
{
    return residual_assemble(oco_imr_element_residual, nelem, elem_q, elem_l, nq, nl, Xq, dofs, c10, c01, res);
}

/* EQUATIONS_SET_JACOBIAN_EVALUATE_STATIC_FEM, equations_set_routines.f90:
 * 4878-5024, with EQUATIONS_MATRICES_JACOBIAN_ELEMENT_ADD (:3174-3224). */
int oco_mr_jacobian(int nelem, const int32_t *elem_q, const int32_t *elem_l, int nq, int nl,
                    const double *Xq, const double *dofs, double c10, double c01, int mode,
                    const int64_t *row_ptr, const int32_t *col_idx, double *values)
{
    const int64_t ndof = (int64_t)3 * nq + nl;
    memset(values, 0, sizeof(double) * (size_t)row_ptr[ndof]);
    double delta = 0.0;
    if (mode == 1 || mode == 2) {
        double s = 0.0;
        for (int64_t i = 0; i < ndof; ++i) s += dofs[i] * dofs[i];
        delta = (1.0 + sqrt(s)) * 1e-7;
    }
    double Xe[81], ze[81], pe[8];
    double *Je = (double *)malloc(sizeof(double) * 89 * 89);
    int32_t ed[89];
    int rc = 0;
    for (int64_t e = 0; e < nelem && !rc; ++e) {
        mr_gather(e, elem_q, elem_l, nq, Xq, dofs, Xe, ze, pe, ed);
        if (mode == 2) oco_imr_element_jacobian_fd(Xe, ze, pe, c10, c01, delta, Je);
        else if (mode == 1) oco_mr_element_jacobian_fd(Xe, ze, pe, c10, c01, delta, Je);
        else oco_mr_element_jacobian(Xe, ze, pe, c10, c01, Je);
        rc = oco_csr_add(row_ptr, col_idx, values, 89, ed, 89, ed, Je);
    }
    free(Je);
    return rc;
}

/* Multi-threaded CPU baseline (not a restatement of one reference routine):
 * residual + analytic Jacobian with the elements split into greedy colours
 * (no two elements of a colour share a node), each colour's elements in
 * parallel over OpenMP threads, every element added with oco_csr_add as in
 * the sequential loop.  Summation order differs from the reference's element
 * order, so values agree with oco_mr_jacobian to rounding only. */
int oco_mr_assemble_parallel(int nelem, const int32_t *elem_q, const int32_t *elem_l, int nq, int nl,
                             const double *Xq, const double *dofs, double c10, double c01,
                             const int64_t *row_ptr, const int32_t *col_idx, double *values, double *res,
                             int threads)
{
    const int64_t ndof = (int64_t)3 * nq + nl;
    uint64_t *used = (uint64_t *)calloc((size_t)nq, sizeof(uint64_t));
    int32_t *colour = (int32_t *)malloc(sizeof(int32_t) * (size_t)nelem);
    if (!used || !colour) {
        free(used);
        free(colour);
        return -2;
    }
    int ncol = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        uint64_t forbidden = 0;
        for (int k = 0; k < 27; ++k) forbidden |= used[elem_q[e * 27 + k]];
        if (~forbidden == 0) {
            free(used);
            free(colour);
            return -3;
        }
        const int c = __builtin_ctzll(~forbidden);
        colour[e] = c;
        if (c + 1 > ncol) ncol = c + 1;
        for (int k = 0; k < 27; ++k) used[elem_q[e * 27 + k]] |= (1ull << c);
    }
    free(used);
    memset(values, 0, sizeof(double) * (size_t)row_ptr[ndof]);
    memset(res, 0, sizeof(double) * (size_t)ndof);
    int rc = 0;
    for (int c = 0; c < ncol; ++c) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 8) reduction(| : rc)
        for (int64_t e = 0; e < nelem; ++e) {
            if (colour[e] != c) continue;
            double Xe[81], ze[81], pe[8], Re[89];
            double Je[89 * 89];
            int32_t ed[89];
            mr_gather(e, elem_q, elem_l, nq, Xq, dofs, Xe, ze, pe, ed);
            oco_mr_element_residual(Xe, ze, pe, c10, c01, Re);
            for (int i = 0; i < 89; ++i) res[ed[i]] += Re[i];
            oco_mr_element_jacobian(Xe, ze, pe, c10, c01, Je);
            rc |= oco_csr_add(row_ptr, col_idx, values, 89, ed, 89, ed, Je) != 0;
        }
    }
    free(colour);
    return rc ? -1 : 0;
}

/* EQUATIONS_SET_BOUNDARY_CONDITIONS_INCREMENT, equations_set_routines.f90:
 * 6560-6585: full value at the last step, else cur + (full-cur)/(N-i+1). */
int oco_bc_increment(int ndof, const int32_t *dof_idx, const double *full, double *current,
                     int iteration, int max_iterations)
{
    for (int i = 0; i < ndof; ++i) {
        const int32_t d = dof_idx[i];
        if (iteration == max_iterations) current[d] = full[i];
        else current[d] = current[d] + (full[i] - current[d]) / (double)(max_iterations - iteration + 1);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Solver matrices of one equations set on one rank.                         */

/* ZERO_TOLERANCE, constants.f90:68 */
#define OCO_ZERO_TOLERANCE (5.0 * DBL_EPSILON)

int oco_solver_system(int n_rows, const int64_t *row_ptr, const int32_t *col_idx, const double *values,
                      const uint8_t *free_mask, const double *vec, const double *dofs, int linear, int64_t *nnz,
                      int64_t *out_row_ptr, int32_t *out_col_idx, double *out_values, double *out_vector)
{
    /* solver mapping: equations row/column d -> solver number, free DOFs in
     * order (INCLUDE_ROW / INCLUDE_COLUMN, solver_mapping_routines.f90:434-520) */
    int32_t *sidx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_rows > 0 ? n_rows : 1));
    if (!sidx) return -1;
    int n = 0;
    for (int d = 0; d < n_rows; ++d) sidx[d] = free_mask[d] ? n++ : -1;
    /* structure (SOLVER_MATRIX_STRUCTURE_CALCULATE): the mapped columns of every
     * mapped row, ascending */
    int64_t total = 0;
    for (int r = 0; r < n_rows; ++r) {
        if (sidx[r] < 0) continue;
        for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) total += sidx[col_idx[k]] >= 0;
    }
    *nnz = total;
    if (!out_row_ptr) {
        free(sidx);
        return n;
    }
    out_row_ptr[0] = 0;
    for (int r = 0, i = 0; r < n_rows; ++r) {
        if (sidx[r] < 0) continue;
        int64_t p = out_row_ptr[i];
        for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
            if (sidx[col_idx[k]] >= 0) out_col_idx[p++] = sidx[col_idx[k]];
        out_row_ptr[++i] = p;
    }
    /* matrix: zero, then for every equations row and stored column
     * VALUE = ALPHA * J * row coupling * column coupling added at the solver
     * entry (CSR branch of SOLVER_MATRIX_JACOBIAN_MATRIX_ADD; the add locates
     * the entry as MATRIX_VALUES_ADD_DP does) */
    for (int64_t k = 0; k < total; ++k) out_values[k] = 0.0;
    const double alpha = 1.0, row_c = 1.0, col_c = 1.0;
    for (int r = 0; r < n_rows; ++r) {
        if (sidx[r] < 0) continue;
        for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            const int32_t sc = sidx[col_idx[k]];
            if (sc < 0) continue;
            const double value = alpha * values[k] * row_c * col_c;
            const int32_t sr = sidx[r];
            int rc = oco_csr_add(out_row_ptr, out_col_idx, out_values, 1, &sr, 1, &sc, &value);
            if (rc) {
                free(sidx);
                return rc;
            }
        }
    }
    /* vector */
    for (int i = 0; i < n; ++i) out_vector[i] = 0.0;
    if (!linear) {
        for (int r = 0; r < n_rows; ++r)
            if (sidx[r] >= 0) out_vector[sidx[r]] += vec[r] * row_c;
        free(sidx);
        return n;
    }
    /* rows storing each column, ascending (the Dirichlet sparsity indices) */
    int64_t *cptr = (int64_t *)calloc((size_t)n_rows + 1, sizeof(int64_t));
    int64_t nz = row_ptr[n_rows];
    int32_t *crow = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nz > 0 ? nz : 1));
    int64_t *cpos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nz > 0 ? nz : 1));
    if (!cptr || !crow || !cpos) {
        free(sidx);
        free(cptr);
        free(crow);
        free(cpos);
        return -1;
    }
    for (int64_t k = 0; k < nz; ++k) cptr[col_idx[k] + 1]++;
    for (int c = 0; c < n_rows; ++c) cptr[c + 1] += cptr[c];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_rows > 0 ? n_rows : 1));
    for (int c = 0; c < n_rows; ++c) fill[c] = cptr[c];
    for (int d = 0; d < n_rows; ++d)
        for (int64_t k = row_ptr[d]; k < row_ptr[d + 1]; ++k) {
            crow[fill[col_idx[k]]] = d;
            cpos[fill[col_idx[k]]++] = k;
        }
    /* solver_routines.f90:14380-14620: in ascending equations row r, the RHS
     * of a free row, then for a fixed DOF r the column r times its value moved
     * to the right-hand side of every free row storing it */
    for (int r = 0; r < n_rows; ++r) {
        if (sidx[r] >= 0) out_vector[sidx[r]] += vec[r] * row_c;
        if (sidx[r] >= 0) continue;
        const double u = dofs[r];
        if (!(fabs(u) >= OCO_ZERO_TOLERANCE)) continue;
        for (int64_t q = cptr[r]; q < cptr[r + 1]; ++q) {
            const int32_t d = crow[q];
            const double K = values[cpos[q]];
            if (!(fabs(K) >= OCO_ZERO_TOLERANCE) || sidx[d] < 0) continue;
            out_vector[sidx[d]] += -1.0 * K * u * row_c;
        }
    }
    free(fill);
    free(cptr);
    free(crow);
    free(cpos);
    free(sidx);
    return n;
}
