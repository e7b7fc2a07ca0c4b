// oc_setup.cpp -- host-side setup for liboc_asm: the CSR structure of the
// equations matrices.  Runs once per equations set (not on the per-assembly
// path).
#include <algorithm>
#include <cstring>
#include <thread>

#include "oc_internal.h"

namespace oc {

namespace {

struct NodeElements {
    std::vector<int64_t> ptr;
    std::vector<int32_t> elems;
};

// Surrounding elements of every node of one mesh component, ascending element
// order (mesh_routines.f90:5929-6011 keeps these lists sorted).
NodeElements surrounding_elements(int n_nodes, int n_elements, int nen, const int32_t *conn)
{
    NodeElements ne;
    ne.ptr.assign(static_cast<size_t>(n_nodes) + 1, 0);
    for (int64_t e = 0; e < n_elements; ++e)
        for (int k = 0; k < nen; ++k) ne.ptr[conn[e * nen + k] + 1]++;
    for (int i = 0; i < n_nodes; ++i) ne.ptr[i + 1] += ne.ptr[i];
    ne.elems.resize(ne.ptr[n_nodes]);
    std::vector<int64_t> fill(ne.ptr.begin(), ne.ptr.end() - 1);
    for (int64_t e = 0; e < n_elements; ++e)
        for (int k = 0; k < nen; ++k) ne.elems[fill[conn[e * nen + k]]++] = static_cast<int32_t>(e);
    return ne;
}

int worker_count()
{
    unsigned hc = std::thread::hardware_concurrency();
    if (hc == 0) hc = 4;
    return static_cast<int>(std::min<unsigned>(hc, 16));
}

}  // namespace

// Equations-matrix structure (equations_matrices_routines.f90:4572-4780 for
// linear matrices, :5048-5230 for the Jacobian): the columns of row DOF
// (component c, node np) are every DOF of every component of the elements
// around np, sorted ascending and unique (LIST_REMOVE_DUPLICATES,
// lists.f90:2647-2690).  0-based output.
int build_pattern(int ncomp, int n_elements, const int32_t *nn, const int32_t *nen, const int32_t *const *conn,
                  std::vector<int64_t> &row_ptr, std::vector<int32_t> &col_idx)
{
    if (ncomp < 1 || ncomp > 4) return set_error(OC_ERR_INVALID_ARGUMENT, "pattern: %d components", ncomp);
    std::vector<int64_t> off(ncomp + 1, 0);
    for (int c = 0; c < ncomp; ++c) off[c + 1] = off[c] + nn[c];
    const int64_t ndof = off[ncomp];
    if (ndof >= INT32_MAX) return set_error(OC_ERR_PATTERN, "pattern: %lld DOFs exceed int32", (long long)ndof);
    std::vector<NodeElements> surr(ncomp);
    for (int c = 0; c < ncomp; ++c) surr[c] = surrounding_elements(nn[c], n_elements, nen[c], conn[c]);

    // Two passes over rows, parallel over contiguous row ranges: count, then fill.
    std::vector<int64_t> counts(ndof + 1, 0);
    const int nw = worker_count();
    auto row_columns = [&](int64_t row, std::vector<int32_t> &buf, std::vector<int32_t> &stamp) {
        int c = 0;
        while (row >= off[c + 1]) ++c;
        const int64_t np = row - off[c];
        buf.clear();
        const int32_t tag = static_cast<int32_t>(row);
        for (int64_t s = surr[c].ptr[np]; s < surr[c].ptr[np + 1]; ++s) {
            const int64_t e = surr[c].elems[s];
            for (int c2 = 0; c2 < ncomp; ++c2)
                for (int k = 0; k < nen[c2]; ++k) {
                    const int32_t col = static_cast<int32_t>(off[c2] + conn[c2][e * nen[c2] + k]);
                    if (stamp[col] != tag) {
                        stamp[col] = tag;
                        buf.push_back(col);
                    }
                }
        }
        std::sort(buf.begin(), buf.end());
    };
    auto run = [&](bool fill_pass) {
        std::vector<std::thread> th;
        for (int w = 0; w < nw; ++w)
            th.emplace_back([&, w]() {
                std::vector<int32_t> buf, stamp(ndof, -1);
                const int64_t r0 = ndof * w / nw, r1 = ndof * (w + 1) / nw;
                for (int64_t r = r0; r < r1; ++r) {
                    row_columns(r, buf, stamp);
                    if (!fill_pass) counts[r + 1] = static_cast<int64_t>(buf.size());
                    else std::memcpy(col_idx.data() + row_ptr[r], buf.data(), buf.size() * sizeof(int32_t));
                }
            });
        for (auto &t : th) t.join();
    };
    run(false);
    for (int64_t r = 0; r < ndof; ++r) counts[r + 1] += counts[r];
    row_ptr.assign(counts.begin(), counts.end());
    col_idx.resize(counts[ndof]);
    run(true);
    return OC_NO_ERROR;
}

}  // namespace oc
