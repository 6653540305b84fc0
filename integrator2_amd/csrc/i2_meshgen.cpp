// i2_meshgen.cpp -- deterministic synthetic surface meshes for the workloads
// the library is measured on: geodesic icospheres (optionally mapped to an
// ellipsoid), tori, and the seeded normal jitter both use.
//
// Conventions: 0-based vertex indices, triangles oriented so that
// (v2 - v1) x (v3 - v1) points out of the enclosed volume (the orientation
// Mesh3D's normals assume, src/Mesh3d.cu:23-32), shared midpoints deduplicated
// so that neighbour classification by vertex index (src/Mesh3d.cu:93-142) sees
// the true topology.  Random draws use std::mt19937_64 (fully specified by the
// C++ standard) and Box-Muller with glibc libm, so the same bytes come out on
// every x86-64 host.
#include "../../include/integrator2_mi355x.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <random>
#include <vector>

namespace {

struct P3 { double x, y, z; };
inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 crs(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dt(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 unit(P3 a) { double l = std::sqrt(dt(a, a)); return {a.x / l, a.y / l, a.z / l}; }

struct Tri { int a, b, c; };

// Standard normal from two 53-bit uniforms (Box-Muller, cosine branch).
double normal01(std::mt19937_64& g) {
    const double u1 = (double)(g() >> 11) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(g() >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
}

double mean_edge_length(const std::vector<P3>& v, const std::vector<Tri>& t) {
    std::map<std::pair<int, int>, int> seen;
    double sum = 0.0;
    long cnt = 0;
    for (const Tri& f : t) {
        const int e[3][2] = {{f.a, f.b}, {f.b, f.c}, {f.c, f.a}};
        for (auto& ed : e) {
            auto key = std::make_pair(std::min(ed[0], ed[1]), std::max(ed[0], ed[1]));
            if (seen.emplace(key, 1).second) {
                P3 d = sub(v[key.first], v[key.second]);
                sum += std::sqrt(dt(d, d));
                ++cnt;
            }
        }
    }
    return cnt ? sum / (double)cnt : 0.0;
}

void icosphere(int level, std::vector<P3>& v, std::vector<Tri>& t) {
    const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
    v.clear(); t.clear();
    // (0, +-1, +-phi) and its cyclic permutations
    for (int s1 = -1; s1 <= 1; s1 += 2)
        for (int s2 = -1; s2 <= 1; s2 += 2) {
            v.push_back(unit({0.0, (double)s1, s2 * phi}));
            v.push_back(unit({(double)s1, s2 * phi, 0.0}));
            v.push_back(unit({s2 * phi, 0.0, (double)s1}));
        }
    // faces: vertex triples at mutual edge distance (the 30 shortest chords)
    double emin = 1e300;
    for (size_t i = 0; i < v.size(); ++i)
        for (size_t j = i + 1; j < v.size(); ++j) { P3 d = sub(v[i], v[j]); emin = std::min(emin, dt(d, d)); }
    auto adj = [&](int i, int j) { P3 d = sub(v[i], v[j]); return std::fabs(dt(d, d) - emin) < 1e-9; };
    for (int i = 0; i < 12; ++i)
        for (int j = i + 1; j < 12; ++j)
            for (int k = j + 1; k < 12; ++k)
                if (adj(i, j) && adj(j, k) && adj(i, k)) {
                    P3 n = crs(sub(v[j], v[i]), sub(v[k], v[i]));
                    P3 c = {v[i].x + v[j].x + v[k].x, v[i].y + v[j].y + v[k].y, v[i].z + v[j].z + v[k].z};
                    if (dt(n, c) > 0) t.push_back({i, j, k}); else t.push_back({i, k, j});
                }
    for (int l = 0; l < level; ++l) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            P3 m = unit({0.5 * (v[a].x + v[b].x), 0.5 * (v[a].y + v[b].y), 0.5 * (v[a].z + v[b].z)});
            v.push_back(m);
            int idx = (int)v.size() - 1;
            mid.emplace(key, idx);
            return idx;
        };
        std::vector<Tri> nt;
        nt.reserve(t.size() * 4);
        for (const Tri& f : t) {
            int ab = midpoint(f.a, f.b), bc = midpoint(f.b, f.c), ca = midpoint(f.c, f.a);
            nt.push_back({f.a, ab, ca});
            nt.push_back({ab, f.b, bc});
            nt.push_back({ca, bc, f.c});
            nt.push_back({ab, bc, ca});
        }
        t.swap(nt);
    }
}

int emit(const std::vector<P3>& v, const std::vector<Tri>& t, double* xyz, int32_t* tri, int64_t* nV, int64_t* nT) {
    if (nV) *nV = (int64_t)v.size();
    if (nT) *nT = (int64_t)t.size();
    if (xyz) for (size_t q = 0; q < v.size(); ++q) { xyz[3 * q] = v[q].x; xyz[3 * q + 1] = v[q].y; xyz[3 * q + 2] = v[q].z; }
    if (tri) for (size_t q = 0; q < t.size(); ++q) { tri[3 * q] = t[q].a; tri[3 * q + 1] = t[q].b; tri[3 * q + 2] = t[q].c; }
    return I2_OK;
}

}  // namespace

extern "C" {

int i2_meshgen_icosphere(int level, double ax, double ay, double az, double jitter_rel, uint64_t seed,
                         double* xyz, int32_t* tri, int64_t* nV, int64_t* nT) {
    if (level < 0 || level > 9 || !(ax > 0) || !(ay > 0) || !(az > 0) || jitter_rel < 0) return I2_ERR_INVALID;
    std::vector<P3> v;
    std::vector<Tri> t;
    icosphere(level, v, t);
    for (P3& p : v) { p.x *= ax; p.y *= ay; p.z *= az; }
    if (jitter_rel > 0) {
        const double sd = jitter_rel * mean_edge_length(v, t);
        std::mt19937_64 g(seed);
        for (P3& p : v) {
            // unit normal of the ellipsoid x^2/ax^2 + y^2/ay^2 + z^2/az^2 = 1
            P3 n = unit({p.x / (ax * ax), p.y / (ay * ay), p.z / (az * az)});
            const double s = sd * normal01(g);
            p.x += s * n.x; p.y += s * n.y; p.z += s * n.z;
        }
    }
    return emit(v, t, xyz, tri, nV, nT);
}

int i2_meshgen_torus(int nmajor, int nminor, double R, double r, double jitter_rel, uint64_t seed,
                     double* xyz, int32_t* tri, int64_t* nV, int64_t* nT) {
    if (nmajor < 3 || nminor < 3 || !(R > r) || !(r > 0) || jitter_rel < 0) return I2_ERR_INVALID;
    const double TWO_PI = 6.283185307179586;
    std::vector<P3> v((size_t)nmajor * nminor), nrm((size_t)nmajor * nminor);
    std::vector<Tri> t;
    t.reserve((size_t)2 * nmajor * nminor);
    for (int a = 0; a < nmajor; ++a)
        for (int b = 0; b < nminor; ++b) {
            const double th = TWO_PI * a / nmajor, ph = TWO_PI * b / nminor;
            const double rr = R + r * std::cos(ph);
            v[(size_t)a * nminor + b] = {rr * std::cos(th), rr * std::sin(th), r * std::sin(ph)};
            nrm[(size_t)a * nminor + b] = {std::cos(th) * std::cos(ph), std::sin(th) * std::cos(ph), std::sin(ph)};
        }
    auto id = [&](int a, int b) { return ((a % nmajor) * nminor) + (b % nminor); };
    for (int a = 0; a < nmajor; ++a)
        for (int b = 0; b < nminor; ++b) {
            const int v00 = id(a, b), v10 = id(a + 1, b), v11 = id(a + 1, b + 1), v01 = id(a, b + 1);
            t.push_back({v00, v10, v11});
            t.push_back({v00, v11, v01});
        }
    if (jitter_rel > 0) {
        const double sd = jitter_rel * mean_edge_length(v, t);
        std::mt19937_64 g(seed);
        for (size_t q = 0; q < v.size(); ++q) {
            const double s = sd * normal01(g);
            v[q].x += s * nrm[q].x; v[q].y += s * nrm[q].y; v[q].z += s * nrm[q].z;
        }
    }
    return emit(v, t, xyz, tri, nV, nT);
}

// Draws of the 64-mesh ragged batch: subdivision level uniform on
// {2,3,4,5} and three semi-axes uniform on (0.5, 3.0), from mt19937_64(seed),
// in the order level, ax, ay, az per mesh.
int i2_meshgen_batch_spec(int nmeshes, uint64_t seed, int32_t* levels, double* axes3) {
    if (nmeshes < 0 || !levels || !axes3) return I2_ERR_INVALID;
    std::mt19937_64 g(seed);
    for (int m = 0; m < nmeshes; ++m) {
        levels[m] = 2 + (int)(g() % 4ull);
        for (int k = 0; k < 3; ++k) axes3[3 * m + k] = 0.5 + 2.5 * ((double)(g() >> 11) * (1.0 / 9007199254740992.0));
    }
    return I2_OK;
}

}  // extern "C"
