// swe_meshgen.cpp -- unstructured P1 meshes of the square [0,L]^2 for the benchmark workloads:
// a seeded Poisson-disk point set and its Delaunay triangulation.
//
// The reference's own meshes are unstructured triangulations written by the external `triangle`
// program (MESHES/SGN_CANAL/files_from_triangle/rectangle.sh) and renumbered by the loader in
// first-touch element order (SFEMaNS/FEMSUB_OBJET/prep_mesh.f90:884-916).  This generator makes
// meshes of that kind at any size (16.8M nodes in seconds), host-only C++, no dependencies.
//
// Points.  Boundary points are spaced b = L/floor(L/r) along the four sides (corners included).
// Interior points: dart throwing on a background grid of cells of side r/sqrt(2) (at most one point
// per cell), in passes; each pass visits the cells in 9 phases (cell (i,j) is in phase
// (i mod 3, j mod 3)) and draws one uniform candidate in every empty cell of the phase, kept when no
// point lies within r.  Two cells of one phase are 3 cells apart, so their candidates never
// conflict and never read a cell the other may write: the phases run in parallel and the result
// depends on the seed only (each candidate's random numbers are a hash of seed, pass and cell).
// Every point is at least r from every other, and an interior point is at least sqrt(3)/2 r from
// the boundary lines.
//
// Triangulation.  Coordinates are quantized to integers on a 2^26 lattice of the square (the output
// coordinates are those lattice values times L/2^26, exactly), so the orientation test is exact in
// 64-bit and the in-circle test exact in 128-bit integers.  Bowyer-Watson insertion along a
// Hilbert curve inside a super-triangle: walk to the triangle that holds the point, grow the cavity
// of triangles whose circumcircle strictly contains it, fan the cavity's boundary to the point.  The
// boundary points make every side segment a Gabriel edge (b < sqrt(2) r), so the triangles inside
// the square are exactly those not touching the super-triangle, and the result is the Delaunay
// triangulation of the point set (cocircular ties broken by insertion order).
//
// Numbering: nodes in insertion (Hilbert) order, elements in the order of their final slots.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

namespace {

typedef __int128 i128;

struct GenMesh {
    std::vector<double> rr;       // (np, 2)
    std::vector<int> jj, neigh;   // (me, 3), 0-based, counter-clockwise; neigh: opposite vertex n, -1 on the boundary
    std::vector<int> jjs, neighs, sides;   // boundary edges (mes, 2), owning element, side tag
    double r = 0.0;
};

// ---- random numbers: a hash of (seed, pass, cell) ------------------------------------------------
inline uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}
inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

template <class F>
void parallel_rows(int n, int nthreads, F&& f) {
    if (nthreads <= 1 || n < 64) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        const int b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=, &f] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

// ---- Poisson-disk points ---------------------------------------------------------------------------
struct Grid {
    int nc = 0;
    double s = 0.0;
    std::vector<double> px, py;   // per cell; NaN when empty
    bool has(int64_t c) const { return px[(size_t)c] == px[(size_t)c]; }
};

int cell_of(double x, double s, int nc) { return std::min(nc - 1, std::max(0, (int)(x / s))); }

void poisson_points(double L, double r, uint64_t seed, int passes, int nthreads, std::vector<double>& xs,
                    std::vector<double>& ys) {
    Grid G;
    G.s = r / std::sqrt(2.0);
    G.nc = (int)std::ceil(L / G.s);
    const int nc = G.nc;
    const size_t ncell = (size_t)nc * (size_t)nc;
    G.px.assign(ncell, std::nan(""));
    G.py.assign(ncell, std::nan(""));
    const double r2 = r * r;
    // boundary points, spacing b in [r, sqrt(2) r)
    const int nb = std::max(1, (int)std::floor(L / r));
    const double b = L / nb;
    auto put = [&](double x, double y) {
        const size_t c = (size_t)cell_of(y, G.s, nc) * nc + cell_of(x, G.s, nc);
        G.px[c] = x;
        G.py[c] = y;
    };
    for (int k = 0; k < nb; ++k) {
        put(k * b, 0.0);             // south, from the south-west corner
        put(L, k * b);               // east, from the south-east corner
        put(L - k * b, L);           // north, from the north-east corner
        put(0.0, L - k * b);         // west, from the north-west corner
    }
    for (int pass = 0; pass < passes; ++pass) {
        for (int phase = 0; phase < 9; ++phase) {
            const int pi = phase % 3, pj = phase / 3;
            const int nrows = (nc - pj + 2) / 3;
            parallel_rows(nrows, nthreads, [&](int rb, int re) {
                for (int rj = rb; rj < re; ++rj) {
                    const int j = pj + 3 * rj;
                    for (int i = pi; i < nc; i += 3) {
                        const size_t c = (size_t)j * nc + i;
                        if (G.has((int64_t)c)) continue;
                        const uint64_t h = mix64(seed ^ mix64(((uint64_t)pass << 48) ^ (uint64_t)c));
                        const double x = (i + unit(h)) * G.s;
                        const double y = (j + unit(mix64(h))) * G.s;
                        if (!(x > 0.0 && x < L && y > 0.0 && y < L)) continue;
                        bool ok = true;
                        for (int dj = -2; dj <= 2 && ok; ++dj) {
                            const int jj = j + dj;
                            if (jj < 0 || jj >= nc) continue;
                            for (int di = -2; di <= 2; ++di) {
                                const int ii = i + di;
                                if (ii < 0 || ii >= nc) continue;
                                const size_t q = (size_t)jj * nc + ii;
                                if (!G.has((int64_t)q)) continue;
                                const double dx = G.px[q] - x, dy = G.py[q] - y;
                                if (dx * dx + dy * dy < r2) {
                                    ok = false;
                                    break;
                                }
                            }
                        }
                        if (ok) {
                            G.px[c] = x;
                            G.py[c] = y;
                        }
                    }
                }
            });
        }
    }
    xs.clear();
    ys.clear();
    for (size_t c = 0; c < ncell; ++c)
        if (G.has((int64_t)c)) {
            xs.push_back(G.px[c]);
            ys.push_back(G.py[c]);
        }
}

// ---- Hilbert curve order ---------------------------------------------------------------------------
// Position of (x, y) along a Hilbert curve over a 65536 x 65536 grid (the classic iterative form:
// quadrant digits from the top, the sub-square rotated or reflected into the curve's frame).
uint32_t hilbert16(uint32_t x, uint32_t y) {
    const uint32_t n = 1u << 16;
    uint32_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

// stable LSD radix sort of indices 0..n-1 by 32-bit keys
std::vector<int> radix_order(const std::vector<uint32_t>& key) {
    const size_t n = key.size();
    std::vector<int> a(n), b(n);
    for (size_t k = 0; k < n; ++k) a[k] = (int)k;
    for (int shift = 0; shift < 32; shift += 16) {
        std::vector<size_t> cnt((1u << 16) + 1, 0);
        for (size_t k = 0; k < n; ++k) ++cnt[((key[(size_t)a[k]] >> shift) & 0xffffu) + 1];
        for (size_t q = 1; q < cnt.size(); ++q) cnt[q] += cnt[q - 1];
        for (size_t k = 0; k < n; ++k) b[cnt[(key[(size_t)a[k]] >> shift) & 0xffffu]++] = a[k];
        a.swap(b);
    }
    return a;
}

// ---- Delaunay triangulation --------------------------------------------------------------------------
struct Tri {
    int v[3];   // counter-clockwise
    int n[3];   // triangle across the edge opposite v[k], -1 if none
};

struct Delaunay {
    std::vector<int64_t> X, Y;
    std::vector<Tri> T;
    std::vector<int> mark;
    int stamp = 0;
    int last = 0;
    uint32_t rng = 2463534242u;

    int64_t orient(int a, int b, int c) const {
        return (X[b] - X[a]) * (Y[c] - Y[a]) - (Y[b] - Y[a]) * (X[c] - X[a]);
    }
    // > 0 when d lies strictly inside the circumcircle of the counter-clockwise triangle (a, b, c)
    bool incircle(int a, int b, int c, int d) const {
        const int64_t adx = X[a] - X[d], ady = Y[a] - Y[d];
        const int64_t bdx = X[b] - X[d], bdy = Y[b] - Y[d];
        const int64_t cdx = X[c] - X[d], cdy = Y[c] - Y[d];
        const int64_t al = adx * adx + ady * ady, bl = bdx * bdx + bdy * bdy, cl = cdx * cdx + cdy * cdy;
        const i128 det = (i128)al * (i128)(bdx * cdy - cdx * bdy) + (i128)bl * (i128)(cdx * ady - adx * cdy) +
                         (i128)cl * (i128)(adx * bdy - bdx * ady);
        return det > 0;
    }

    int locate(int p) {
        int t = last;
        for (;;) {
            rng ^= rng << 13;
            rng ^= rng >> 17;
            rng ^= rng << 5;
            const int k0 = (int)(rng % 3u);
            int next = -1;
            for (int q = 0; q < 3; ++q) {
                const int k = (k0 + q) % 3;
                const Tri& tr = T[(size_t)t];
                if (orient(tr.v[(k + 1) % 3], tr.v[(k + 2) % 3], p) < 0) {
                    next = tr.n[k];
                    break;
                }
            }
            if (next < 0) return t;
            t = next;
        }
    }

    struct BEdge {
        int a, b, out, slot;
    };
    std::vector<int> cav, stack;
    std::vector<BEdge> bnd;

    void insert(int p) {
        const int t0 = locate(p);
        ++stamp;
        cav.clear();
        stack.clear();
        bnd.clear();
        mark[(size_t)t0] = stamp;
        cav.push_back(t0);
        stack.push_back(t0);
        while (!stack.empty()) {
            const int c = stack.back();
            stack.pop_back();
            for (int k = 0; k < 3; ++k) {
                const int nb = T[(size_t)c].n[k];
                if (nb >= 0 && mark[(size_t)nb] == stamp) continue;
                if (nb >= 0) {
                    const Tri& o = T[(size_t)nb];
                    if (incircle(o.v[0], o.v[1], o.v[2], p)) {
                        mark[(size_t)nb] = stamp;
                        cav.push_back(nb);
                        stack.push_back(nb);
                        continue;
                    }
                }
                bnd.push_back({T[(size_t)c].v[(k + 1) % 3], T[(size_t)c].v[(k + 2) % 3], nb, c});
            }
        }
        // bnd holds |cav| + 2 edges (the cavity is star-shaped from p): the cavity's slots are reused
        // and two slots are appended
        const size_t m = bnd.size();
        for (size_t q = 0; q < m; ++q) {
            if (q < cav.size()) {
                bnd[q].slot = cav[q];
            } else {
                bnd[q].slot = (int)T.size();
                T.push_back(Tri{});
                mark.push_back(0);
            }
        }
        // the outside triangle's link to the old cavity triangle now points at the new one; the
        // outside edge is matched by its endpoints, since one old triangle may own several edges
        for (size_t q = 0; q < m; ++q) {
            const BEdge& e = bnd[q];
            if (e.out < 0) continue;
            Tri& o = T[(size_t)e.out];
            for (int k = 0; k < 3; ++k)
                if (o.v[(k + 1) % 3] == e.b && o.v[(k + 2) % 3] == e.a) o.n[k] = e.slot;
        }
        for (size_t q = 0; q < m; ++q) {
            const BEdge& e = bnd[q];
            Tri& t = T[(size_t)e.slot];
            t.v[0] = e.a;
            t.v[1] = e.b;
            t.v[2] = p;
            t.n[2] = e.out;
        }
        for (size_t q = 0; q < m; ++q) {
            // across (b, p): the new triangle that starts at b; it sees this one across (p, b)
            const BEdge& e = bnd[q];
            int s = -1;
            for (size_t w = 0; w < m; ++w)
                if (bnd[w].a == e.b) {
                    s = bnd[w].slot;
                    break;
                }
            T[(size_t)e.slot].n[0] = s;
            T[(size_t)s].n[1] = e.slot;
        }
        last = bnd[0].slot;
    }
};

}  // namespace

extern "C" {

typedef struct swe_genmesh swe_genmesh;

// Poisson-disk points with minimum distance r on [0,L]^2 (passes of dart throwing, seeded) and
// their Delaunay triangulation.  Returns the number of nodes (< 0 on failure); *out receives the
// mesh, to be read with swe_genmesh_get and freed with swe_genmesh_free.  nthreads <= 0: all cores.
long long swe_genmesh_poisson_delaunay(double L, double r, unsigned long long seed, int passes, int nthreads,
                                       swe_genmesh** out) {
    if (!out || !(L > 0.0) || !(r > 0.0) || r > L / 2 || passes < 1) return -1;
    *out = nullptr;
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    GenMesh* M = new (std::nothrow) GenMesh();
    if (!M) return -2;
    try {
        std::vector<double> xs, ys;
        poisson_points(L, r, (uint64_t)seed, passes, nthreads, xs, ys);
        const int64_t S = (int64_t)1 << 26;
        const size_t n = xs.size();
        if (n + 3 > (size_t)INT32_MAX / 3) {
            delete M;
            return -3;
        }
        std::vector<int64_t> qx(n), qy(n);
        std::vector<uint32_t> key(n);
        for (size_t k = 0; k < n; ++k) {
            qx[k] = std::llround(xs[k] / L * (double)S);
            qy[k] = std::llround(ys[k] / L * (double)S);
            key[k] = hilbert16((uint32_t)std::min<int64_t>(qx[k] >> 10, 65535),
                               (uint32_t)std::min<int64_t>(qy[k] >> 10, 65535));
        }
        const std::vector<int> order = radix_order(key);
        Delaunay D;
        D.X.resize(n + 3);
        D.Y.resize(n + 3);
        for (size_t k = 0; k < n; ++k) {
            D.X[k] = qx[(size_t)order[k]];
            D.Y[k] = qy[(size_t)order[k]];
        }
        const int s0 = (int)n, s1 = s0 + 1, s2 = s0 + 2;   // super-triangle around [0,S]^2
        D.X[(size_t)s0] = -3 * S, D.Y[(size_t)s0] = -3 * S;
        D.X[(size_t)s1] = 6 * S, D.Y[(size_t)s1] = -3 * S;
        D.X[(size_t)s2] = -3 * S, D.Y[(size_t)s2] = 6 * S;
        D.T.reserve(2 * n + 8);
        D.mark.reserve(2 * n + 8);
        D.T.push_back(Tri{{s0, s1, s2}, {-1, -1, -1}});
        D.mark.push_back(0);
        for (size_t k = 0; k < n; ++k) D.insert((int)k);
        // keep the triangles of the square, renumbered in slot order
        std::vector<int> id(D.T.size(), -1);
        int me = 0;
        for (size_t t = 0; t < D.T.size(); ++t) {
            const Tri& tr = D.T[t];
            if (tr.v[0] < s0 && tr.v[1] < s0 && tr.v[2] < s0) id[t] = me++;
        }
        M->r = r;
        M->rr.resize(2 * n);
        const double h = L / (double)S;
        for (size_t k = 0; k < n; ++k) {
            M->rr[2 * k] = (double)D.X[k] * h;
            M->rr[2 * k + 1] = (double)D.Y[k] * h;
        }
        M->jj.resize(3 * (size_t)me);
        M->neigh.resize(3 * (size_t)me);
        for (size_t t = 0; t < D.T.size(); ++t) {
            if (id[t] < 0) continue;
            const Tri& tr = D.T[t];
            const size_t e = (size_t)id[t];
            for (int k = 0; k < 3; ++k) {
                M->jj[3 * e + k] = tr.v[k];
                const int nb = tr.n[k];
                M->neigh[3 * e + k] = (nb >= 0) ? id[(size_t)nb] : -1;
                if (M->neigh[3 * e + k] < 0) {
                    const int a = tr.v[(k + 1) % 3], b = tr.v[(k + 2) % 3];
                    int side = 0;
                    if (D.Y[a] == 0 && D.Y[b] == 0) side = 1;
                    else if (D.X[a] == S && D.X[b] == S) side = 2;
                    else if (D.Y[a] == S && D.Y[b] == S) side = 3;
                    else if (D.X[a] == 0 && D.X[b] == 0) side = 4;
                    if (side == 0) {   // a boundary edge off the square's sides: not a valid mesh
                        delete M;
                        return -4;
                    }
                    M->jjs.push_back(a);
                    M->jjs.push_back(b);
                    M->neighs.push_back((int)e);
                    M->sides.push_back(side);
                }
            }
        }
    } catch (const std::bad_alloc&) {
        delete M;
        return -2;
    }
    *out = (swe_genmesh*)M;
    return (long long)(M->rr.size() / 2);
}

// Sizes of a generated mesh: nodes, elements, boundary edges.
int swe_genmesh_sizes(const swe_genmesh* g, int* np, int* me, int* mes) {
    const GenMesh* M = (const GenMesh*)g;
    if (!M) return -1;
    if (np) *np = (int)(M->rr.size() / 2);
    if (me) *me = (int)(M->jj.size() / 3);
    if (mes) *mes = (int)M->sides.size();
    return 0;
}

// Copy a generated mesh out (0-based): rr(np,2), jj(me,3), neigh(me,3), jjs(mes,2), neighs(mes),
// sides(mes) with the tags 1 south, 2 east, 3 north, 4 west.  Any pointer may be NULL.
int swe_genmesh_get(const swe_genmesh* g, double* rr, int* jj, int* neigh, int* jjs, int* neighs, int* sides) {
    const GenMesh* M = (const GenMesh*)g;
    if (!M) return -1;
    if (rr) std::memcpy(rr, M->rr.data(), sizeof(double) * M->rr.size());
    if (jj) std::memcpy(jj, M->jj.data(), sizeof(int) * M->jj.size());
    if (neigh) std::memcpy(neigh, M->neigh.data(), sizeof(int) * M->neigh.size());
    if (jjs) std::memcpy(jjs, M->jjs.data(), sizeof(int) * M->jjs.size());
    if (neighs) std::memcpy(neighs, M->neighs.data(), sizeof(int) * M->neighs.size());
    if (sides) std::memcpy(sides, M->sides.data(), sizeof(int) * M->sides.size());
    return 0;
}

void swe_genmesh_free(swe_genmesh* g) { delete (GenMesh*)g; }

}  // extern "C"
