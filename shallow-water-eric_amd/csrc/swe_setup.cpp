// swe_setup.cpp -- host-side P1 set-up of the shallow-water step (the part of
// construct_matrices, SOURCES/update.f90:18-132, that the step consumes).
//
// Bitwise contract with the reference:
//  * pattern of st_csr (SFEMaNS/FEMSUB/st_sparsekit.f90:788-863): each row holds its node and every
//    node sharing an element with it, ascending, unique;
//  * geometry of gauss_points_2d_p1 (gauss_points_2d_p1.f90:106-132) with the 3-point rule of
//    element_2d (347-393);
//  * consistent mass of qs_00_M (SFEMaNS/FEMSUB_OBJET/fem_sparsekit.f90:31-51) and c_ij of
//    update.f90:106-124, each entry accumulated in ascending element order (Gauss points in order
//    inside an element, which is the order the reference's element loops produce for one entry);
//  * lumped(i) = sequential row sum of the mass (update.f90:35-38).
// Rows are independent once the node->element adjacency is sorted by element, so rows are
// assembled in parallel without changing any entry's summation order.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "swe_amd.h"
#include "swe_p1.h"

namespace {

using swe_p1::ElemGeom;
using swe_p1::P1Rule;
using swe_p1::element_geometry;

template <class F>
void parallel_rows(int np, F&& f) {
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (np < 20000) nt = 1;
    std::vector<std::thread> th;
    int chunk = (np + (int)nt - 1) / (int)nt;
    for (unsigned t = 0; t < nt; ++t) {
        int b = (int)t * chunk, e = std::min(np, b + chunk);
        if (b >= e) break;
        th.emplace_back([=, &f] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

struct Adjacency {
    std::vector<int64_t> start;  // node -> first slot
    std::vector<int> elem;       // elements containing the node, ascending
};

Adjacency node_elements(int np, int me, const int* jj1) {
    Adjacency A;
    A.start.assign((size_t)np + 1, 0);
    for (int64_t q = 0; q < 3 * (int64_t)me; ++q) A.start[(size_t)jj1[q] - 1 + 1]++;
    for (int i = 0; i < np; ++i) A.start[i + 1] += A.start[i];
    A.elem.resize((size_t)A.start[np]);
    std::vector<int64_t> fill(A.start.begin(), A.start.end() - 1);
    for (int m = 0; m < me; ++m)        // ascending m => each list is sorted
        for (int a = 0; a < 3; ++a) A.elem[(size_t)fill[(size_t)jj1[3 * (size_t)m + a] - 1]++] = m;
    return A;
}

bool check_mesh(int np, int me, const int* jj) {
    if (np <= 0 || me <= 0 || !jj) return false;
    for (int64_t q = 0; q < 3 * (int64_t)me; ++q)
        if (jj[q] < 1 || jj[q] > np) return false;
    return true;
}

// row pattern: sorted unique nodes of the elements containing i (itself included)
void row_pattern(const Adjacency& A, const int* jj1, int i, std::vector<int>& out) {
    out.clear();
    for (int64_t s = A.start[i]; s < A.start[i + 1]; ++s) {
        const int* nd = jj1 + 3 * (size_t)A.elem[(size_t)s];
        for (int a = 0; a < 3; ++a) out.push_back(nd[a] - 1);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

}  // namespace

extern "C" long swe_assemble_nnz(int np, int me, const int* jj) {
    if (!check_mesh(np, me, jj)) return -1;
    Adjacency A = node_elements(np, me, jj);
    std::vector<long> cnt((size_t)np, 0);
    parallel_rows(np, [&](int b, int e) {
        std::vector<int> row;
        for (int i = b; i < e; ++i) {
            row_pattern(A, jj, i, row);
            cnt[(size_t)i] = (long)row.size();
        }
    });
    long nnz = 0;
    for (long c : cnt) nnz += c;
    return nnz;
}

extern "C" int swe_assemble(int np, int me, const int* jj, const double* rr, int* ia, int* ja,
                            int* diag, double* mass, double* lumped, double* cx, double* cy) {
    if (!check_mesh(np, me, jj) || !rr || !ia || !ja || !diag || !mass || !lumped || !cx || !cy)
        return SWE_EINVAL;
    Adjacency A = node_elements(np, me, jj);
    // pattern (1-based like st_csr)
    std::vector<int> cnt((size_t)np);
    parallel_rows(np, [&](int b, int e) {
        std::vector<int> row;
        for (int i = b; i < e; ++i) {
            row_pattern(A, jj, i, row);
            cnt[(size_t)i] = (int)row.size();
        }
    });
    ia[0] = 1;
    for (int i = 0; i < np; ++i) ia[i + 1] = ia[i] + cnt[(size_t)i];
    const P1Rule R;
    parallel_rows(np, [&](int b, int e) {
        std::vector<int> row;
        for (int i = b; i < e; ++i) {
            row_pattern(A, jj, i, row);
            const int p0 = ia[i] - 1;
            const int n = (int)row.size();
            for (int q = 0; q < n; ++q) {
                ja[p0 + q] = row[(size_t)q] + 1;
                mass[p0 + q] = 0.0;
                cx[p0 + q] = 0.0;
                cy[p0 + q] = 0.0;
            }
            auto slot = [&](int j) {
                return p0 + (int)(std::lower_bound(row.begin(), row.end(), j) - row.begin());
            };
            for (int64_t s = A.start[i]; s < A.start[i + 1]; ++s) {
                const int m = A.elem[(size_t)s];
                int nd[3] = {jj[3 * (size_t)m] - 1, jj[3 * (size_t)m + 1] - 1, jj[3 * (size_t)m + 2] - 1};
                const int ni = (nd[0] == i) ? 0 : (nd[1] == i) ? 1 : 2;
                ElemGeom G;
                element_geometry(R, rr, nd, G);
                // mass: qs_00_M visits l, then nj, then ni; for this row ni is fixed
                for (int l = 0; l < 3; ++l) {
                    const double al = 1.0 * G.rj[l];
                    for (int nj = 0; nj < 3; ++nj) {
                        const int p = slot(nd[nj]);
                        mass[p] = mass[p] + R.ww[nj][l] * al * R.ww[ni][l];
                    }
                }
                // c_ij: x(d) = SUM_l dw(d,nj,l) * ww(ni,l) * rj(l)
                for (int nj = 0; nj < 3; ++nj) {
                    double x[2];
                    for (int d = 0; d < 2; ++d) {
                        double sacc = 0.0;
                        for (int l = 0; l < 3; ++l) sacc = sacc + G.dw[d][nj][l] * R.ww[ni][l] * G.rj[l];
                        x[d] = sacc;
                    }
                    const int p = slot(nd[nj]);
                    cx[p] = cx[p] + x[0];
                    cy[p] = cy[p] + x[1];
                }
            }
            double sl = 0.0;
            for (int q = 0; q < n; ++q) sl = sl + mass[p0 + q];
            lumped[i] = sl;
            diag[i] = slot(i) + 1;
        }
    });
    return SWE_OK;
}
