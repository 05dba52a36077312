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
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <cmath>
#include <cstdint>
#include <cstring>

#include "swe_amd.h"
#include "swe_engine.h"
#include "swe_p1.h"

namespace {

constexpr int kMaxElems = 32;
constexpr int kMaxPattern = 3 * kMaxElems;

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
    int row[kMaxPattern];
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
    int row[kMaxPattern];
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
    double* rr = nullptr;     // node coordinates rr(2,np), kept for the locality order of the layout
};

namespace {

// Device memory of the set-up is stream-ordered (hipMallocAsync / hipFreeAsync on the set-up's own
// stream): unlike hipFree, releasing it does not wait for the work queued on other streams, so a
// batch can admit meshes while its steps run.
template <class T>
hipError_t dalloc(T** p, size_t n, hipStream_t s) {
    return hipMallocAsync((void**)p, sizeof(T) * (n ? n : 1), s);
}

void dfree(void* p, hipStream_t s) {
    if (p) (void)hipFreeAsync(p, s);
}

void release(swe_assembly* a) {
    if (!a) return;
    void* bufs[] = {a->ia0, a->ja, a->diag, a->mass, a->lumped, a->cx, a->cy, a->rr};
    for (void* b : bufs) dfree(b, a->stream);
    if (a->stream) {
        (void)hipStreamSynchronize(a->stream);
        (void)hipStreamDestroy(a->stream);
    }
    delete a;
}

// Exclusive prefix sum on stream s (ordered with the stream's other work; nothing waits here).
hipError_t exclusive_scan(const int64_t* in, int64_t* out, int n, hipStream_t s) {
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s);
    if (e != hipSuccess) return e;
    void* tmp = nullptr;
    if ((e = hipMallocAsync(&tmp, bytes ? bytes : 1, s)) != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, s);
    dfree(tmp, s);
    return e;
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
    if (e == hipSuccess) e = dalloc(&djj, (size_t)n3, s);
    if (e == hipSuccess) e = dalloc(&drr, 2 * (size_t)np, s);
    if (e == hipSuccess) e = dalloc(&elem, (size_t)n3, s);
    if (e == hipSuccess) e = dalloc(&fill, (size_t)np, s);
    if (e == hipSuccess) e = dalloc(&flag, 1, s);
    if (e == hipSuccess) e = dalloc(&cnt, (size_t)np + 1, s);
    if (e == hipSuccess) e = dalloc(&start, (size_t)np + 1, s);
    if (e == hipSuccess) e = dalloc(&a->ia0, (size_t)np + 1, s);
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
            e = dalloc(&a->ja, (size_t)total, s);
            if (e == hipSuccess) e = dalloc(&a->diag, (size_t)np, s);
            if (e == hipSuccess) e = dalloc(&a->mass, (size_t)total, s);
            if (e == hipSuccess) e = dalloc(&a->lumped, (size_t)np, s);
            if (e == hipSuccess) e = dalloc(&a->cx, (size_t)total, s);
            if (e == hipSuccess) e = dalloc(&a->cy, (size_t)total, s);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_row_fill, dim3(grid_of(np, 128)), dim3(128), 0, s, np, start, elem, djj, drr,
                                   a->ia0, a->ja, a->diag, a->mass, a->lumped, a->cx, a->cy);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipStreamSynchronize(s);
        }
    }
    a->rr = drr;
    void* tmp[] = {djj, elem, fill, flag, cnt, start};
    for (void* b : tmp) dfree(b, s);
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

// =============================================================================================
// The sliced layout of an assembled mesh, built on the GPU (the same internal numbering and slots
// as build_layout on the host: rows stably sorted by length inside each 64-row window of the
// caller's numbering, one 64-row slice per window).
// =============================================================================================
namespace {

__global__ void k_lay_len(int np, const int64_t* __restrict__ ia0, int* __restrict__ lenc, int* __restrict__ bad) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    const int64_t l = ia0[i + 1] - ia0[i];
    lenc[i] = (int)l;
    if (l > swe::kMaxRow || l < 1) atomicOr(bad, 1);
}

// One wave per window of the base order (base[p], or p when base is null): each lane's rank among
// the window's valid rows by (length, position in the base order).
__global__ void k_lay_sort(int np, int nwin, const int* __restrict__ lenc, const int* __restrict__ base,
                           int* __restrict__ perm) {
    const int w = (int)(blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const int pos = w * 64 + lane;
    const bool valid = w < nwin && pos < np;
    const int i = valid ? (base ? base[pos] : pos) : 0;
    const int li = valid ? lenc[i] : 0x7fffffff;
    int rank = 0;
    for (int m = 0; m < 64; ++m) {
        const int lm = __shfl(li, m, 64);
        const bool vm = w < nwin && (w * 64 + m) < np;
        rank += (vm && (lm < li || (lm == li && m < lane))) ? 1 : 0;
    }
    if (valid) perm[w * 64 + rank] = i;
}

// ---- locality order: nodes sorted along a Hilbert curve of their coordinates ------------------
// Orderable image of a double: unsigned comparison of the images orders finite doubles.
__device__ __forceinline__ unsigned long long ord_of(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
double double_of_ord(unsigned long long o) {
    const unsigned long long b = (o >> 63) ? (o & ~0x8000000000000000ull) : ~o;
    double v;
    std::memcpy(&v, &b, sizeof v);
    return v;
}

// box = {min x, max x, min y, max y} as orderable images (initialised to ~0, 0, ~0, 0).
__global__ void k_bbox(int np, const double* __restrict__ rr, unsigned long long* __restrict__ box) {
    unsigned long long v[4] = {~0ull, 0ull, ~0ull, 0ull};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long x = ord_of(rr[2 * i]), y = ord_of(rr[2 * i + 1]);
        v[0] = (x < v[0]) ? x : v[0];
        v[1] = (x > v[1]) ? x : v[1];
        v[2] = (y < v[2]) ? y : v[2];
        v[3] = (y > v[3]) ? y : v[3];
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 4; ++q) {
            const unsigned long long o = __shfl_xor(v[q], off, 64);
            v[q] = (q & 1) ? ((o > v[q]) ? o : v[q]) : ((o < v[q]) ? o : v[q]);
        }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(box + 0, v[0]);
        atomicMax(box + 1, v[1]);
        atomicMin(box + 2, v[2]);
        atomicMax(box + 3, v[3]);
    }
}

constexpr int kCurveBits = 20;   // per coordinate: 2^20 cells along each side of the bounding box

// Position of each node along a Hilbert curve through the 2^20 x 2^20 cells of the bounding box
// (the classic bit-by-bit conversion: quadrant, then rotate / reflect the lower bits).
__global__ void k_curve_key(int np, const double* __restrict__ rr, double x0, double sx, double y0, double sy,
                            unsigned long long* __restrict__ key, int* __restrict__ idx) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    const double top = (double)((1 << kCurveBits) - 1);
    double fx = (rr[2 * (size_t)i] - x0) * sx, fy = (rr[2 * (size_t)i + 1] - y0) * sy;
    fx = (fx > 0.0) ? ((fx < top) ? fx : top) : 0.0;
    fy = (fy > 0.0) ? ((fy < top) ? fy : top) : 0.0;
    unsigned x = (unsigned)fx, y = (unsigned)fy;
    unsigned long long d = 0;
    for (unsigned sq = 1u << (kCurveBits - 1); sq > 0; sq >>= 1) {
        const unsigned rx = (x & sq) ? 1u : 0u, ry = (y & sq) ? 1u : 0u;
        d += (unsigned long long)sq * sq * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = (x & ~(sq - 1)) + (sq - 1 - (x & (sq - 1)));
                y = (y & ~(sq - 1)) + (sq - 1 - (y & (sq - 1)));
            }
            const unsigned t = x;
            x = y;
            y = t;
        }
    }
    key[i] = d;
    idx[i] = i;
}

__global__ void k_inverse(int np, const int* __restrict__ order, int* __restrict__ pos) {
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p < np) pos[order[p]] = p;
}

// Sum over the off-diagonal entries of |pos(i) - pos(j)| (pos null: the caller's numbering).
__global__ void k_offset_sum(int np, const int64_t* __restrict__ ia0, const int* __restrict__ ja,
                             const int* __restrict__ pos, unsigned long long* __restrict__ sum) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    unsigned long long acc = 0;
    if (i < np) {
        const int pi = pos ? pos[i] : i;
        for (int64_t p = ia0[i]; p < ia0[i + 1]; ++p) {
            const int j = ja[p] - 1;
            const int pj = pos ? pos[j] : j;
            acc += (unsigned long long)((pi > pj) ? pi - pj : pj - pi);
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd(sum, acc);
}

__global__ void k_lay_iperm(int np, int np_pad, const int* __restrict__ perm, const int* __restrict__ lenc,
                            int* __restrict__ iperm, int* __restrict__ lenr) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= np_pad) return;
    if (r < np) {
        const int i = perm[r];
        iperm[i] = r;
        lenr[r] = lenc[i];
    } else {
        lenr[r] = 0;
    }
}

// Per slice: 64 x its longest row (slots) and its off-diagonal entries (compact-list length).
__global__ void k_lay_slices(int nsl, int np, const int* __restrict__ lenr, int64_t* __restrict__ slots,
                             int64_t* __restrict__ ecnt, int* __restrict__ wmax) {
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= nsl) return;
    int w = 0;
    int64_t e = 0;
    for (int l = 0; l < 64; ++l) {
        const int r = s * 64 + l;
        const int L = lenr[r];
        w = (L > w) ? L : w;
        if (r < np) e += L - 1;
    }
    slots[s] = 64 * (int64_t)w;
    ecnt[s] = e;
    atomicMax(wmax, w);
}

struct LayOut {
    uint8_t* len8;
    uint8_t* dpos8;
    int* perm;
    double* m;
    double* z;
    int* col;
    int16_t* dcol;
    double2* c;
};

// Per internal row: its slots (columns, offsets, c_ij) and node data, as build_layout fills them.
__global__ void k_lay_fill(int np, int np_pad, int64_t r0, int64_t p0, const int* __restrict__ perm,
                           const int* __restrict__ iperm, const int* __restrict__ lenr,
                           const int64_t* __restrict__ sptr, const int64_t* __restrict__ ia0,
                           const int* __restrict__ ja, const double* __restrict__ cx, const double* __restrict__ cy,
                           const double* __restrict__ lumped, const double* __restrict__ bath, LayOut o,
                           int64_t* __restrict__ csr2slot) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= np_pad) return;
    if (r >= np) {
        o.len8[r0 + r] = 0;
        o.dpos8[r0 + r] = 0;
        return;
    }
    const int i = perm[r];
    const int L = lenr[r];
    const int64_t base = sptr[r >> 6] + (r & 63);
    const int64_t pi = ia0[i];
    int dp = 0;
    for (int q = 0; q < L; ++q) {
        const int64_t p = pi + q;
        const int j = ja[p] - 1;
        const int jr = iperm[j];
        const int64_t slot = base + 64 * (int64_t)q;
        o.col[p0 + slot] = (int)(r0 + jr);
        const int off = jr - r;
        o.dcol[p0 + slot] = (int16_t)((off > swe::kColFar && off <= 32767) ? off : swe::kColFar);
        o.c[p0 + slot] = make_double2(cx[p], cy[p]);
        if (csr2slot) csr2slot[p] = slot;
        if (j == i) dp = q;
    }
    o.len8[r0 + r] = (uint8_t)L;
    o.dpos8[r0 + r] = (uint8_t)dp;
    o.perm[r0 + r] = i;
    o.m[r0 + r] = lumped[i];
    o.z[r0 + r] = bath[i];
}

// One wave per slice: its compact list of off-diagonal slots, entry by entry, lanes in order.
__global__ void k_lay_elist(int nsl, int np, int64_t r0, int64_t e0, const int* __restrict__ lenr,
                            const uint8_t* __restrict__ dpos8, const int64_t* __restrict__ sptr,
                            const int64_t* __restrict__ eoff, uint16_t* __restrict__ elist) {
    const int s = (int)(blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6));
    if (s >= nsl) return;
    const int lane = (int)(threadIdx.x & 63);
    const int r = s * 64 + lane;
    const int L = (r < np) ? lenr[r] : 0;
    const int dp = (r < np) ? (int)dpos8[r0 + r] : 0;
    const int w = (int)((sptr[s + 1] - sptr[s]) / 64);
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int64_t off = eoff[s];
    for (int q = 0; q < w; ++q) {
        const bool active = q < L && q != dp;
        const unsigned long long mask = __ballot(active);
        if (active) elist[e0 + off + __popcll(mask & below)] = (uint16_t)(q * 64 + lane);
        off += __popcll(mask);
    }
}

__global__ void k_lay_slice_info(int nsl, int64_t s0, int64_t p0, int64_t e0, int inst,
                                 const int64_t* __restrict__ sptr, const int64_t* __restrict__ eoff,
                                 swe::SliceInfo* __restrict__ sl) {
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= nsl) return;
    swe::SliceInfo v;
    v.sbase = p0 + sptr[s];
    v.ebase = e0 + eoff[s];
    v.inst = inst;
    v.ecount = (int)(eoff[s + 1] - eoff[s]);
    v.pad_[0] = 0;
    v.pad_[1] = 0;
    sl[s0 + s] = v;
}

__global__ void k_lay_bc(int nbc, int nh, int nux, int64_t r0, int64_t b0, int inst, const int* __restrict__ nodes1,
                         const int* __restrict__ iperm, swe::BcEntry* __restrict__ bc) {
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= nbc) return;
    const int comp = (q < nh) ? 0 : (q < nh + nux) ? 1 : 2;
    bc[b0 + q] = swe::BcEntry{(int)(r0 + iperm[nodes1[q] - 1]), inst, comp, (int)(b0 + q)};
}

}  // namespace

namespace {

hipError_t read_u64(const unsigned long long* d, unsigned long long* h, int n, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(h, d, sizeof(unsigned long long) * (size_t)n, hipMemcpyDeviceToHost, s);
    return (e == hipSuccess) ? hipStreamSynchronize(s) : e;
}

// The base order of the layout.  *base stays null for the caller's numbering; otherwise it lists the
// nodes along a Hilbert curve of their coordinates, the device counterpart of the host
// layout's reverse Cuthill-McKee order, under the same SWE_REORDER policy: 1 when the caller's
// numbering scatters neighbours (mean |i-j| above 2 sqrt(np)) and the curve brings them closer, 2
// always.  Only the processing order changes; results do not depend on it.
hipError_t locality_order(const swe_assembly* a, hipStream_t s, int** base) {
    *base = nullptr;
    const int mode = swe::reorder_mode();
    const int np = a->np;
    if (mode == 0 || np < 2 || a->nnz <= np) return hipSuccess;
    unsigned long long* acc = nullptr;   // [0] offset sum, caller's order; [1] curve order; [2..5] box
    unsigned long long* keys = nullptr;
    int *idx = nullptr, *order = nullptr, *pos = nullptr;
    unsigned long long* keys2 = nullptr;
    void* tmp = nullptr;
    unsigned long long h[6] = {0, 0, ~0ull, 0ull, ~0ull, 0ull};
    const unsigned bl = (unsigned)((np + 255) / 256);
    const double offdiag = (double)(a->nnz - np);
    hipError_t e = dalloc(&acc, 6, s);
    if (e == hipSuccess) e = hipMemcpyAsync(acc, h, sizeof h, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_offset_sum, dim3(bl), dim3(256), 0, s, np, (const int64_t*)a->ia0, (const int*)a->ja,
                           (const int*)nullptr, acc);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = read_u64(acc, h, 1, s);
    const double mean0 = (double)h[0] / offdiag;
    bool take = false;
    if (e == hipSuccess && (mode == 2 || mean0 > 2.0 * std::sqrt((double)np))) {
        hipLaunchKernelGGL(k_bbox, dim3(bl < 1024u ? bl : 1024u), dim3(256), 0, s, np, (const double*)a->rr, acc + 2);
        e = hipGetLastError();
        if (e == hipSuccess) e = read_u64(acc + 2, h + 2, 4, s);
        const double x0 = double_of_ord(h[2]), x1 = double_of_ord(h[3]);
        const double y0 = double_of_ord(h[4]), y1 = double_of_ord(h[5]);
        const double top = (double)((1 << kCurveBits) - 1);
        const double sx = (x1 > x0) ? top / (x1 - x0) : 0.0, sy = (y1 > y0) ? top / (y1 - y0) : 0.0;
        if (e == hipSuccess) e = dalloc(&keys, (size_t)np, s);
        if (e == hipSuccess) e = dalloc(&keys2, (size_t)np, s);
        if (e == hipSuccess) e = dalloc(&idx, (size_t)np, s);
        if (e == hipSuccess) e = dalloc(&order, (size_t)np, s);
        if (e == hipSuccess) e = dalloc(&pos, (size_t)np, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_curve_key, dim3(bl), dim3(256), 0, s, np, (const double*)a->rr, x0, sx, y0, sy, keys, idx);
            e = hipGetLastError();
        }
        size_t bytes = 0;
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, keys2, idx, order, np, 0, 2 * kCurveBits, s);
        if (e == hipSuccess) e = hipMallocAsync(&tmp, bytes ? bytes : 1, s);
        if (e == hipSuccess)   // a radix sort is stable: equal keys keep the caller's order
            e = hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys, keys2, idx, order, np, 0, 2 * kCurveBits, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_inverse, dim3(bl), dim3(256), 0, s, np, (const int*)order, pos);
            hipLaunchKernelGGL(k_offset_sum, dim3(bl), dim3(256), 0, s, np, (const int64_t*)a->ia0,
                               (const int*)a->ja, (const int*)pos, acc + 1);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = read_u64(acc + 1, h + 1, 1, s);
        take = e == hipSuccess && (mode == 2 || h[1] < h[0]);
    }
    void* bufs[] = {acc, keys, keys2, idx, pos, tmp};
    for (void* b : bufs) dfree(b, s);
    if (take) *base = order;
    else dfree(order, s);
    return e;
}

}  // namespace

namespace swe {

void free_device_layout(DeviceLayout& DL) {
    void* bufs[] = {DL.perm, DL.iperm, DL.len, DL.sptr, DL.eoff};
    for (void* b : bufs) dfree(b, DL.stream);
    DL = DeviceLayout{};
}

int build_layout_device(const swe_assembly* a, DeviceLayout& DL) {
    if (!a) return fail(SWE_EINVAL, "null assembly");
    DL = DeviceLayout{};
    const int np = a->np;
    DL.np = np;
    DL.nnz = a->nnz;
    const int nsl = (np + 63) / 64, np_pad = nsl * 64;
    DL.nsl = nsl;
    hipStream_t s = a->stream;
    DL.stream = s;
    int* lenc = nullptr;
    int* flag = nullptr;
    int64_t *slots = nullptr, *ecnt = nullptr;
    hipError_t e = dalloc(&DL.perm, (size_t)np, s);
    if (e == hipSuccess) e = dalloc(&DL.iperm, (size_t)np, s);
    if (e == hipSuccess) e = dalloc(&DL.len, (size_t)np_pad, s);
    if (e == hipSuccess) e = dalloc(&DL.sptr, (size_t)nsl + 1, s);
    if (e == hipSuccess) e = dalloc(&DL.eoff, (size_t)nsl + 1, s);
    if (e == hipSuccess) e = dalloc(&lenc, (size_t)np, s);
    if (e == hipSuccess) e = dalloc(&flag, 2, s);
    if (e == hipSuccess) e = dalloc(&slots, (size_t)nsl + 1, s);
    if (e == hipSuccess) e = dalloc(&ecnt, (size_t)nsl + 1, s);
    if (e == hipSuccess) e = hipMemsetAsync(flag, 0, 2 * sizeof(int), s);
    if (e == hipSuccess) e = hipMemsetAsync(slots + nsl, 0, sizeof(int64_t), s);
    if (e == hipSuccess) e = hipMemsetAsync(ecnt + nsl, 0, sizeof(int64_t), s);
    int* base = nullptr;
    if (e == hipSuccess) e = locality_order(a, s, &base);
    int hflag[2] = {0, 0};
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_lay_len, dim3(grid_of(np, 256)), dim3(256), 0, s, np, (const int64_t*)a->ia0, lenc, flag);
        hipLaunchKernelGGL(k_lay_sort, dim3(grid_of(nsl, 4)), dim3(256), 0, s, np, nsl, (const int*)lenc,
                           (const int*)base, DL.perm);
        hipLaunchKernelGGL(k_lay_iperm, dim3(grid_of(np_pad, 256)), dim3(256), 0, s, np, np_pad, (const int*)DL.perm,
                           (const int*)lenc, DL.iperm, DL.len);
        hipLaunchKernelGGL(k_lay_slices, dim3(grid_of(nsl, 256)), dim3(256), 0, s, nsl, np, (const int*)DL.len, slots,
                           ecnt, flag + 1);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = exclusive_scan(slots, DL.sptr, nsl + 1, s);
    if (e == hipSuccess) e = exclusive_scan(ecnt, DL.eoff, nsl + 1, s);
    if (e == hipSuccess) e = hipMemcpyAsync(hflag, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&DL.nslot, DL.sptr + nsl, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&DL.ne, DL.eoff + nsl, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    void* tmp[] = {lenc, flag, slots, ecnt, base};
    for (void* b : tmp) dfree(b, s);
    if (e != hipSuccess) {
        free_device_layout(DL);
        return fail(SWE_EHIP, "device layout: %s", hipGetErrorString(e));
    }
    if (hflag[0]) {
        free_device_layout(DL);
        return fail(SWE_EINVAL, "a row has more than %d entries", kMaxRow);
    }
    DL.wmax = hflag[1];
    return SWE_OK;
}

int pool_add_assembled(Pool* P, const swe_assembly* a, const DeviceLayout& DL, const double* bath,
                       const int* const lists[3], const int nlist[3], const swe_params& prm, const double* un,
                       double t0, double tfinal, long long max_steps, int* id, int64_t* csr2slot) {
    if (!a || !bath || !id) return fail(SWE_EINVAL, "null argument");
    Mode md;
    int rc = mode_of(prm, md);
    if (rc) return rc;
    if (P->mode_set && !same_mode(md, P->mode)) return fail(SWE_EBADOPT, "this instance's scheme differs from the pool's");
    if ((rc = check_sgn_lists(md, nlist[1], nlist[2]))) return rc;
    if (md.sgn && (rc = ensure_sgn_buffers(P))) return rc;
    const int np = DL.np;
    for (int c = 0; c < 3; ++c) {
        if (nlist[c] < 0 || (nlist[c] > 0 && !lists[c])) return fail(SWE_EINVAL, "Dirichlet list %d is malformed", c);
        for (int q = 0; q < nlist[c]; ++q)
            if (lists[c][q] < 1 || lists[c][q] > np) return fail(SWE_EINVAL, "Dirichlet node %d out of range", lists[c][q]);
    }
    int k = -1;
    for (int q = 0; q < P->inst_cap; ++q)
        if (!P->inst[(size_t)q].used) { k = q; break; }
    const int64_t nbc = (int64_t)nlist[0] + nlist[1] + nlist[2];
    const int64_t s0 = (k < 0) ? -1 : P->a_slices.alloc(DL.nsl);
    const int64_t p0 = (s0 < 0) ? -1 : P->a_slots.alloc(DL.nslot);
    const int64_t e0 = (p0 < 0) ? -1 : P->a_elist.alloc(DL.ne);
    const int64_t b0 = (e0 < 0) ? -1 : P->a_bc.alloc(nbc);
    if (k < 0 || s0 < 0 || p0 < 0 || e0 < 0 || b0 < 0) {
        if (s0 >= 0) P->a_slices.release(s0, DL.nsl);
        if (p0 >= 0) P->a_slots.release(p0, DL.nslot);
        if (e0 >= 0) P->a_elist.release(e0, DL.ne);
        return fail(SWE_EFULL, "the pool has no room for a mesh of %d nodes now", np);
    }
    const int64_t r0 = s0 * 64;
    const int np_pad = DL.nsl * 64;
    // the admission runs on the pool's copy stream, as an upload does: the steps queued on the compute
    // stream go on meanwhile (the new rows are not in their launch lists yet)
    hipStream_t s = P->copy;
    double* dbath = nullptr;
    int* dnodes = nullptr;
    hipError_t e = dalloc(&dbath, (size_t)np, s);
    if (e == hipSuccess) e = dalloc(&dnodes, (size_t)nbc, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dbath, bath, sizeof(double) * (size_t)np, hipMemcpyHostToDevice, s);
    for (int c = 0, q = 0; c < 3 && e == hipSuccess; q += nlist[c], ++c)
        if (nlist[c] > 0)
            e = hipMemcpyAsync(dnodes + q, lists[c], sizeof(int) * (size_t)nlist[c], hipMemcpyHostToDevice, s);
    // ranges freed by a removal are rewritten only after the removal's kernels
    if (e == hipSuccess) e = hipStreamWaitEvent(s, P->released, 0);
    if (e == hipSuccess && un) {
        const int nc = md.sgn ? 5 : 3;
        double* st = md.sgn ? P->stage5 : P->stage3;
        e = hipMemcpyAsync(st + nc * r0, un, sizeof(double) * nc * (size_t)np, hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess) {
        LayOut o{P->len, P->dpos, P->perm, P->m, P->z, P->col, P->dcol, P->c};
        hipLaunchKernelGGL(k_lay_fill, dim3(grid_of(np_pad, 256)), dim3(256), 0, s, np, np_pad, r0, p0,
                           (const int*)DL.perm, (const int*)DL.iperm, (const int*)DL.len, (const int64_t*)DL.sptr,
                           (const int64_t*)a->ia0, (const int*)a->ja, (const double*)a->cx, (const double*)a->cy,
                           (const double*)a->lumped, (const double*)dbath, o, csr2slot);
        hipLaunchKernelGGL(k_lay_elist, dim3(grid_of(DL.nsl, 4)), dim3(256), 0, s, DL.nsl, np, r0, e0,
                           (const int*)DL.len, (const uint8_t*)P->dpos, (const int64_t*)DL.sptr,
                           (const int64_t*)DL.eoff, P->elist);
        hipLaunchKernelGGL(k_lay_slice_info, dim3(grid_of(DL.nsl, 256)), dim3(256), 0, s, DL.nsl, s0, p0, e0, k,
                           (const int64_t*)DL.sptr, (const int64_t*)DL.eoff, P->sl);
        if (nbc > 0)
            hipLaunchKernelGGL(k_lay_bc, dim3(grid_of(nbc, 256)), dim3(256), 0, s, (int)nbc, nlist[0], nlist[1], r0, b0,
                               k, (const int*)dnodes, (const int*)DL.iperm, P->bc);
        e = hipGetLastError();
    }
    dfree(dbath, s);
    dfree(dnodes, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        P->a_slices.release(s0, DL.nsl);
        P->a_slots.release(p0, DL.nslot);
        P->a_elist.release(e0, DL.ne);
        P->a_bc.release(b0, nbc);
        return fail(SWE_EHIP, "admission of an assembled mesh: %s", hipGetErrorString(e));
    }
    InstHost I;
    I.np = np;
    I.nnz = (int)DL.nnz;
    I.np_pad = np_pad;
    I.wmax = DL.wmax;
    I.slice0 = s0;
    I.nslices = DL.nsl;
    I.slot0 = p0;
    I.nslots = DL.nslot;
    I.e0 = e0;
    I.ne = DL.ne;
    I.bc0 = b0;
    I.nbc = nbc;
    I.nh = nlist[0];
    I.nux = nlist[1];
    I.nuy = nlist[2];
    if ((rc = pool_activate(P, k, I, prm, un != nullptr, t0, tfinal, max_steps))) return rc;
    *id = k;
    return SWE_OK;
}

}  // namespace swe
