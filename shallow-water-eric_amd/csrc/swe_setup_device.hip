// swe_setup_device.hip -- the P1 set-up of construct_matrices (SOURCES/update.f90:18-132) on the
// GPU: the pattern of st_csr (SFEMaNS/FEMSUB/st_sparsekit.f90:788-863), the consistent mass of
// qs_00_M (SFEMaNS/FEMSUB_OBJET/fem_sparsekit.f90:31-51), c_ij (update.f90:106-124), the lumped
// mass (update.f90:35-38) and the diagonal positions.
//
// Same contract as the host assembly in swe_setup.cpp, bit for bit: every entry is accumulated in
// ascending element order (Gauss points in order inside an element) with the shared element
// geometry of swe_p1.h, and lumped(i) is the sequential row sum.  One lane per row:
//   1. node -> element lists: counts (integer atomics), exclusive scan, fill, then each list
//      sorted ascending in its lane, so the lists do not depend on the order the atomics ran in;
//   2. row pattern: the sorted unique nodes of the row's elements, counted, scanned into ia;
//   3. row fill: ja, then mass and c_ij accumulated element by element, lumped, diag.
// A node may belong to at most kMaxElems elements (rows of at most 3*kMaxElems candidates);
// larger fans are reported as SWE_EINVAL and can be assembled on the host.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_scan.hpp>

#include <cstdint>
#include <cstring>

#include "swe_amd.h"
#include "swe_engine.h"
#include "swe_p1.h"

namespace {

constexpr int kMaxElems = 32;
constexpr int kMaxRow = 3 * kMaxElems;

using swe_p1::ElemGeom;
using swe_p1::P1Rule;
using swe_p1::element_geometry;

__global__ void k_count_elems(const int* __restrict__ jj, int64_t n3, int64_t* __restrict__ cnt) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n3) atomicAdd((unsigned long long*)&cnt[jj[q] - 1], 1ull);
}

__global__ void k_fill_elems(const int* __restrict__ jj, int me, const int64_t* __restrict__ start,
                             int* __restrict__ fill, int* __restrict__ elem) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3 * (int64_t)me) return;
    const int n = jj[q] - 1;
    const int k = atomicAdd(&fill[n], 1);
    elem[start[n] + k] = (int)(q / 3);
}

// Sort each node's element list ascending; flags a node in more than kMaxElems elements.
__global__ void k_sort_elems(int np, const int64_t* __restrict__ start, int* __restrict__ elem,
                             int* __restrict__ too_many) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    const int64_t b = start[i];
    const int n = (int)(start[i + 1] - b);
    if (n > kMaxElems) {
        atomicOr(too_many, 1);
        return;
    }
    for (int a = 1; a < n; ++a) {
        const int v = elem[b + a];
        int c = a - 1;
        while (c >= 0 && elem[b + c] > v) {
            elem[b + c + 1] = elem[b + c];
            --c;
        }
        elem[b + c + 1] = v;
    }
}

// Sorted unique nodes (0-based) of the elements of row i; returns their number.
__device__ int row_pattern(int i, const int64_t* __restrict__ start, const int* __restrict__ elem,
                           const int* __restrict__ jj, int* row) {
    int n = 0;
    for (int64_t s = start[i]; s < start[i + 1]; ++s) {
        const int m = elem[s];
        for (int a = 0; a < 3; ++a) {
            const int v = jj[3 * (int64_t)m + a] - 1;
            // insert v keeping row[0..n) sorted and unique
            int c = n - 1;
            bool dup = false;
            while (c >= 0 && row[c] >= v) {
                if (row[c] == v) {
                    dup = true;
                    break;
                }
                --c;
            }
            if (dup) continue;
            for (int t = n; t > c + 1; --t) row[t] = row[t - 1];
            row[c + 1] = v;
            ++n;
        }
    }
    return n;
}

__global__ void k_row_count(int np, const int64_t* __restrict__ start, const int* __restrict__ elem,
                            const int* __restrict__ jj, int64_t* __restrict__ cnt) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    int row[kMaxRow];
    cnt[i] = row_pattern(i, start, elem, jj, row);
}

__device__ __forceinline__ int slot_of(const int* row, int n, int j) {
    int lo = 0, hi = n;   // first position with row[pos] >= j
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (row[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_row_fill(int np, const int64_t* __restrict__ start, const int* __restrict__ elem,
                           const int* __restrict__ jj, const double* __restrict__ rr, const int64_t* __restrict__ ia0,
                           int* __restrict__ ja, int* __restrict__ diag, double* __restrict__ mass,
                           double* __restrict__ lumped, double* __restrict__ cx, double* __restrict__ cy) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    int row[kMaxRow];
    const int n = row_pattern(i, start, elem, jj, row);
    const int64_t p0 = ia0[i];
    for (int q = 0; q < n; ++q) {
        ja[p0 + q] = row[q] + 1;
        mass[p0 + q] = 0.0;
        cx[p0 + q] = 0.0;
        cy[p0 + q] = 0.0;
    }
    const P1Rule R;
    for (int64_t s = start[i]; s < start[i + 1]; ++s) {
        const int m = elem[s];
        const int nd[3] = {jj[3 * (int64_t)m] - 1, jj[3 * (int64_t)m + 1] - 1, jj[3 * (int64_t)m + 2] - 1};
        const int ni = (nd[0] == i) ? 0 : (nd[1] == i) ? 1 : 2;
        ElemGeom G;
        element_geometry(R, rr, nd, G);
        // mass: qs_00_M visits l, then nj, then ni; for this row ni is fixed
        for (int l = 0; l < 3; ++l) {
            const double al = 1.0 * G.rj[l];
            for (int nj = 0; nj < 3; ++nj) {
                const int64_t p = p0 + slot_of(row, n, nd[nj]);
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
            const int64_t p = p0 + slot_of(row, n, nd[nj]);
            cx[p] = cx[p] + x[0];
            cy[p] = cy[p] + x[1];
        }
    }
    double sl = 0.0;
    for (int q = 0; q < n; ++q) sl = sl + mass[p0 + q];
    lumped[i] = sl;
    diag[i] = (int)(p0 + slot_of(row, n, i)) + 1;
}

unsigned grid_of(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

struct swe_assembly {
    int np = 0;
    int64_t nnz = 0;
    hipStream_t stream = nullptr;
    int64_t* ia0 = nullptr;   // 0-based row starts, np+1
    int* ja = nullptr;
    int* diag = nullptr;
    double *mass = nullptr, *lumped = nullptr, *cx = nullptr, *cy = nullptr;
};

namespace {

void release(swe_assembly* a) {
    if (!a) return;
    void* bufs[] = {a->ia0, a->ja, a->diag, a->mass, a->lumped, a->cx, a->cy};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (a->stream) (void)hipStreamDestroy(a->stream);
    delete a;
}

template <class T>
hipError_t dalloc(T** p, size_t n) {
    return hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
}

hipError_t exclusive_scan(const int64_t* in, int64_t* out, int n, hipStream_t s) {
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s);
    if (e != hipSuccess) return e;
    void* tmp = nullptr;
    if ((e = hipMalloc(&tmp, bytes ? bytes : 1)) != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(tmp);
    return e != hipSuccess ? e : e2;
}

}  // namespace

extern "C" int swe_assemble_device(int np, int me, const int* jj, const double* rr, swe_assembly** out,
                                   long long* nnz) {
    swe::clear_error();
    if (!out || !nnz || !jj || !rr || np <= 0 || me <= 0) return swe::fail(SWE_EINVAL, "null or empty argument");
    *out = nullptr;
    *nnz = 0;
    const int64_t n3 = 3 * (int64_t)me;
    for (int64_t q = 0; q < n3; ++q)
        if (jj[q] < 1 || jj[q] > np) return swe::fail(SWE_EINVAL, "jj(%lld)=%d out of range", (long long)q + 1, jj[q]);
    swe_assembly* a = new swe_assembly();
    a->np = np;
    int *djj = nullptr, *elem = nullptr, *fill = nullptr, *flag = nullptr;
    double* drr = nullptr;
    int64_t *cnt = nullptr, *start = nullptr;
    int too_many = 0;
    hipError_t e = hipStreamCreate(&a->stream);
    hipStream_t s = a->stream;
    if (e == hipSuccess) e = dalloc(&djj, (size_t)n3);
    if (e == hipSuccess) e = dalloc(&drr, 2 * (size_t)np);
    if (e == hipSuccess) e = dalloc(&elem, (size_t)n3);
    if (e == hipSuccess) e = dalloc(&fill, (size_t)np);
    if (e == hipSuccess) e = dalloc(&flag, 1);
    if (e == hipSuccess) e = dalloc(&cnt, (size_t)np + 1);
    if (e == hipSuccess) e = dalloc(&start, (size_t)np + 1);
    if (e == hipSuccess) e = dalloc(&a->ia0, (size_t)np + 1);
    if (e == hipSuccess) e = hipMemcpyAsync(djj, jj, sizeof(int) * (size_t)n3, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(drr, rr, sizeof(double) * 2 * (size_t)np, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, sizeof(int64_t) * ((size_t)np + 1), s);
    if (e == hipSuccess) e = hipMemsetAsync(fill, 0, sizeof(int) * (size_t)np, s);
    if (e == hipSuccess) e = hipMemsetAsync(flag, 0, sizeof(int), s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_count_elems, dim3(grid_of(n3, 256)), dim3(256), 0, s, djj, n3, cnt);
        e = exclusive_scan(cnt, start, np + 1, s);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_elems, dim3(grid_of(n3, 256)), dim3(256), 0, s, djj, me, start, fill, elem);
        hipLaunchKernelGGL(k_sort_elems, dim3(grid_of(np, 256)), dim3(256), 0, s, np, start, elem, flag);
        e = hipMemcpyAsync(&too_many, flag, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    int rc = SWE_OK;
    if (e == hipSuccess && too_many) {
        rc = swe::fail(SWE_EINVAL, "a node belongs to more than %d elements; use swe_assemble", kMaxElems);
    } else if (e == hipSuccess) {
        hipLaunchKernelGGL(k_row_count, dim3(grid_of(np, 128)), dim3(128), 0, s, np, start, elem, djj, cnt);
        e = hipMemsetAsync(cnt + np, 0, sizeof(int64_t), s);
        if (e == hipSuccess) e = exclusive_scan(cnt, a->ia0, np + 1, s);
        int64_t total = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&total, a->ia0 + np, sizeof(int64_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess && total >= (int64_t)INT32_MAX) {
            rc = swe::fail(SWE_EINVAL, "nnz = %lld does not fit the 32-bit row pointers", (long long)total);
        } else if (e == hipSuccess) {
            a->nnz = total;
            e = dalloc(&a->ja, (size_t)total);
            if (e == hipSuccess) e = dalloc(&a->diag, (size_t)np);
            if (e == hipSuccess) e = dalloc(&a->mass, (size_t)total);
            if (e == hipSuccess) e = dalloc(&a->lumped, (size_t)np);
            if (e == hipSuccess) e = dalloc(&a->cx, (size_t)total);
            if (e == hipSuccess) e = dalloc(&a->cy, (size_t)total);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_row_fill, dim3(grid_of(np, 128)), dim3(128), 0, s, np, start, elem, djj, drr,
                                   a->ia0, a->ja, a->diag, a->mass, a->lumped, a->cx, a->cy);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipStreamSynchronize(s);
        }
    }
    void* tmp[] = {djj, drr, elem, fill, flag, cnt, start};
    for (void* b : tmp)
        if (b) (void)hipFree(b);
    if (e != hipSuccess) rc = swe::fail(SWE_EHIP, "device assembly: %s", hipGetErrorString(e));
    if (rc != SWE_OK) {
        release(a);
        return rc;
    }
    *out = a;
    *nnz = (long long)a->nnz;
    return SWE_OK;
}

extern "C" int swe_assembly_get(swe_assembly* a, int* ia, int* ja, int* diag, double* mass, double* lumped,
                                double* cx, double* cy) {
    if (!a) return swe::fail(SWE_EINVAL, "null assembly");
    const size_t nz = (size_t)a->nnz, np = (size_t)a->np;
    hipStream_t s = a->stream;
    hipError_t e = hipSuccess;
    if (ia) {
        // 1-based 32-bit row pointers, as st_csr builds them
        int64_t* h = new int64_t[np + 1];
        e = hipMemcpyAsync(h, a->ia0, sizeof(int64_t) * (np + 1), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess)
            for (size_t i = 0; i <= np; ++i) ia[i] = (int)h[i] + 1;
        delete[] h;
    }
    if (e == hipSuccess && ja) e = hipMemcpyAsync(ja, a->ja, sizeof(int) * nz, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && diag) e = hipMemcpyAsync(diag, a->diag, sizeof(int) * np, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && mass) e = hipMemcpyAsync(mass, a->mass, sizeof(double) * nz, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && lumped) e = hipMemcpyAsync(lumped, a->lumped, sizeof(double) * np, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && cx) e = hipMemcpyAsync(cx, a->cx, sizeof(double) * nz, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && cy) e = hipMemcpyAsync(cy, a->cy, sizeof(double) * nz, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return swe::fail(SWE_EHIP, "assembly read-back: %s", hipGetErrorString(e));
    return SWE_OK;
}

extern "C" int swe_assembly_destroy(swe_assembly* a) {
    release(a);
    return SWE_OK;
}
