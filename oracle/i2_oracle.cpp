// i2_oracle.cpp -- C ABI around the CPU restatement in i2_oracle.hpp.
//
// TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
// library never links or calls it.
//
// Mesh preparation restates src/Mesh3d.cu:23-142 (normals, areas, neighbour
// classification by shared vertex indices); the per-type adaptive loop
// restates src/evaluators/evaluatorJ3DK.cu:895-1012 together with
// src/NumericalIntegrator3d.cu:37-173,400-499 and
// src/evaluators/evaluator3d.cu:76-99,297-341; o_refinement_counts restates
// the determineCellsToBeRefined call sequence (evaluatorJ3DK.cu:962-1010,
// NumericalIntegrator3d.cu:461-499) and is pinned by its closed forms
// (tests/test_cpu_api.py).
#include "i2_oracle.hpp"

#include <omp.h>
#include <cstring>
#include <cstdio>
#include <chrono>

using namespace i2o;

namespace {

struct Mesh {
    int64_t nV = 0, nT = 0;
    std::vector<V3<double>> v;
    std::vector<I3> t;
    std::vector<V3<double>> normal;
    std::vector<double> area;
    std::vector<I3> lists[3];   // unordered (i<j, k) lists per type, lexicographic order
};

int common_vertices(const I3& a, const I3& b) {
    int c = 0;
    if (a.x == b.x || a.x == b.y || a.x == b.z) ++c;
    if (a.y == b.x || a.y == b.y || a.y == b.z) ++c;
    if (a.z == b.x || a.z == b.y || a.z == b.z) ++c;
    return c;
}

// Mesh3d.cu:23-32 (normal) and :62-71 (area)
void prepare(Mesh& m) {
    m.normal.resize(m.nT);
    m.area.resize(m.nT);
    for (int64_t c = 0; c < m.nT; ++c) {
        const I3 t = m.t[c];
        const V3<double> v12 = m.v[t.y] - m.v[t.x];
        const V3<double> v13 = m.v[t.z] - m.v[t.x];
        m.normal[c] = normalize(cross(v12, v13));
        m.area[c] = vlen(cross(v12, v13)) * 0.5;
    }
    // Mesh3d.cu:93-142: pairs (i<j) by number of shared vertex indices;
    // 0 -> disjoint, 1 -> vertex-sharing, 2 -> edge-sharing.  Candidate
    // neighbours come from vertex->triangle incidence; every other j > i is disjoint.
    std::vector<std::vector<int>> inc(m.nV);
    for (int64_t c = 0; c < m.nT; ++c) {
        inc[m.t[c].x].push_back((int)c); inc[m.t[c].y].push_back((int)c); inc[m.t[c].z].push_back((int)c);
    }
    for (int k = 0; k < 3; ++k) m.lists[k].clear();
    for (int64_t i = 0; i < m.nT; ++i) {
        std::vector<int> cand;
        for (int k = 0; k < 3; ++k) for (int j : inc[comp(m.t[i], k)]) if (j > i) cand.push_back(j);
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        size_t ci = 0;
        for (int64_t j = i + 1; j < m.nT; ++j) {
            int cnt = 0;
            if (ci < cand.size() && cand[ci] == j) { cnt = common_vertices(m.t[i], m.t[j]); ++ci; }
            if (cnt == 0) m.lists[DISJOINT].push_back({(int)i, (int)j, 0});
            else if (cnt == 1) m.lists[VERTEX_SHARING].push_back({(int)i, (int)j, 0});
            else if (cnt == 2) m.lists[EDGE_SHARING].push_back({(int)i, (int)j, 0});
        }
    }
    for (int k = 0; k < 3; ++k)
        for (size_t q = 0; q < m.lists[k].size(); ++q) m.lists[k][q].z = (int)q;
}

template <class T>
V3<T> vert(const Mesh& m, int idx, const V3<T>* override_v, int ov_index) {
    if (override_v && idx == ov_index) return *override_v;
    const V3<double>& p = m.v[idx];
    return {T(p.x), T(p.y), T(p.z)};
}

// Geometry of one ordered task (i, j); vertices may carry dual parts.
template <class T> struct TaskGeom {
    V3<T> I[3], J[3];      // original vertex positions of K_i and K_j (reference order)
    V3<T> nI, nJ;
    T areaI;
};

template <class T>
TaskGeom<T> task_geom(const Mesh& m, int i, int j, const V3<T>* ov, int ovi) {
    TaskGeom<T> g;
    const I3 ti = m.t[i], tj = m.t[j];
    for (int k = 0; k < 3; ++k) { g.I[k] = vert<T>(m, comp(ti, k), ov, ovi); g.J[k] = vert<T>(m, comp(tj, k), ov, ovi); }
    const V3<T> a12 = g.I[1] - g.I[0], a13 = g.I[2] - g.I[0];
    const V3<T> b12 = g.J[1] - g.J[0], b13 = g.J[2] - g.J[0];
    g.nI = normalize(cross(a12, a13));
    g.nJ = normalize(cross(b12, b13));
    g.areaI = vlen(cross(a12, a13)) * T(0.5);
    return g;
}

template <class T>
PairEval<T> pair_eval(const Mesh& m, int type, int i, int j, const TaskGeom<T>& g) {
    PairEval<T> E;
    E.type = type;
    E.A = g.J[0]; E.B = g.J[1]; E.C = g.J[2];
    if (type == EDGE_SHARING) {
        int s1, s2; shifts_edge(m.t[i], m.t[j], &s1, &s2);
        E.eJA = g.J[s2 % 3]; E.eJB = g.J[(s2 + 1) % 3]; E.eJC = g.J[(s2 + 2) % 3];
    } else if (type == VERTEX_SHARING) {
        int s1, s2; shifts_vertex(m.t[i], m.t[j], &s1, &s2);
        E.vc = vertex_ctx(g.J[s2 % 3], g.J[(s2 + 1) % 3], g.J[(s2 + 2) % 3], g.nI, g.nJ, g.areaI);
    }
    return E;
}

template <class T>
PairVerts<T> rotated_verts(const Mesh& m, int type, int i, int j, const TaskGeom<T>& g) {
    int s1, s2;
    if (type == EDGE_SHARING) shifts_edge(m.t[i], m.t[j], &s1, &s2);
    else shifts_vertex(m.t[i], m.t[j], &s1, &s2);
    PairVerts<T> p;
    p.IA = g.I[s1 % 3]; p.IB = g.I[(s1 + 1) % 3]; p.IC = g.I[(s1 + 2) % 3];
    p.JA = g.J[s2 % 3]; p.JB = g.J[(s2 + 1) % 3]; p.JC = g.J[(s2 + 2) % 3];
    return p;
}

template <class T>
Q4<T> singular_integral(const Mesh& m, int type, int i, int j, const TaskGeom<T>& g) {
    if (type == DISJOINT) return {T(0.0), T(0.0), T(0.0), T(0.0)};
    PairVerts<T> p = rotated_verts(m, type, i, j, g);
    if (type == EDGE_SHARING) return integrate_sing_edge(p, g.nI, g.nJ, g.areaI);
    return integrate_sing_vertex(p, g.nI, g.nJ, g.areaI);
}

template <class T>
Q4<T> task_level_value(const Mesh& m, const Rule& R, int type, int i, int j, int level, const TaskGeom<T>& g) {
    PairEval<T> E = pair_eval(m, type, i, j, g);
    SubPanel<T> P{g.I[0], g.I[1], g.I[2], g.areaI};
    return level_value(R, P, level, E);
}

// Counter-based normal draws that define the weights of the scalar loss
// L = sum w_ij Theta_ij + sum v_ij . J_ij.  The product library uses the same
// definition (splitmix64 finaliser of (seed, pair id, stream) and Box-Muller).
inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double unit53(uint64_t h) { return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
inline double normal_draw(uint64_t seed, uint64_t pair_id, uint64_t stream) {
    uint64_t base = mix64(seed ^ mix64(pair_id * 4ull + stream));
    double u1 = unit53(mix64(base ^ 0x1ull));
    double u2 = unit53(mix64(base ^ 0x2ull));
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(TWO_PI * u2);
}

}  // namespace

extern "C" {

void* o_create(const double* xyz, int64_t nV, const int32_t* tri, int64_t nT) {
    Mesh* m = new Mesh;
    m->nV = nV; m->nT = nT;
    m->v.resize(nV); m->t.resize(nT);
    for (int64_t q = 0; q < nV; ++q) m->v[q] = {xyz[3 * q], xyz[3 * q + 1], xyz[3 * q + 2]};
    for (int64_t q = 0; q < nT; ++q) m->t[q] = {tri[3 * q], tri[3 * q + 1], tri[3 * q + 2]};
    prepare(*m);
    return m;
}
void o_destroy(void* h) { delete (Mesh*)h; }

int64_t o_pair_count(void* h, int type) { return (int64_t)((Mesh*)h)->lists[type].size(); }

void o_get_list(void* h, int type, int32_t* out3) {
    const Mesh& m = *(Mesh*)h;
    for (size_t q = 0; q < m.lists[type].size(); ++q) {
        out3[3 * q] = m.lists[type][q].x; out3[3 * q + 1] = m.lists[type][q].y; out3[3 * q + 2] = m.lists[type][q].z;
    }
}
void o_get_normals(void* h, double* out3) {
    const Mesh& m = *(Mesh*)h;
    for (int64_t c = 0; c < m.nT; ++c) { out3[3 * c] = m.normal[c].x; out3[3 * c + 1] = m.normal[c].y; out3[3 * c + 2] = m.normal[c].z; }
}
void o_get_measures(void* h, double* out) {
    const Mesh& m = *(Mesh*)h;
    for (int64_t c = 0; c < m.nT; ++c) out[c] = m.area[c];
}

// Point function entry points used by the pinning tests.
void o_theta_psi(const double* pts, int64_t n, const double* tri9, double* out4) {
    V3<double> A{tri9[0], tri9[1], tri9[2]}, B{tri9[3], tri9[4], tri9[5]}, C{tri9[6], tri9[7], tri9[8]};
    for (int64_t q = 0; q < n; ++q) {
        Q4<double> r = theta_psi(V3<double>{pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]}, A, B, C);
        out4[4 * q] = r.x; out4[4 * q + 1] = r.y; out4[4 * q + 2] = r.z; out4[4 * q + 3] = r.w;
    }
}
// Singular part of task (i, j) of the given type at n points (the subtracted function).
void o_singular_part(void* h, int type, int i, int j, const double* pts, int64_t n, double* out4) {
    const Mesh& m = *(Mesh*)h;
    TaskGeom<double> g = task_geom<double>(m, i, j, nullptr, -1);
    PairEval<double> E = pair_eval(m, type, i, j, g);
    for (int64_t q = 0; q < n; ++q) {
        V3<double> p{pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
        Q4<double> r{0, 0, 0, 0};
        if (type == EDGE_SHARING) r = sing_edge(p, E.eJA, E.eJB, E.eJC);
        else if (type == VERTEX_SHARING) r = sing_vertex(p, E.vc);
        out4[4 * q] = r.x; out4[4 * q + 1] = r.y; out4[4 * q + 2] = r.z; out4[4 * q + 3] = r.w;
    }
}
// Closed-form integral over K_i of the singular part of task (i, j); returns
// the special-case branches it took (BR_* bit mask).
unsigned o_singular_integral(void* h, int type, int i, int j, double* out4) {
    const Mesh& m = *(Mesh*)h;
    branch_flags() = 0;
    TaskGeom<double> g = task_geom<double>(m, i, j, nullptr, -1);
    Q4<double> r = singular_integral(m, type, i, j, g);
    out4[0] = r.x; out4[1] = r.y; out4[2] = r.z; out4[3] = r.w;
    return branch_flags();
}
// Closed-form singular integrals of ordered tasks evaluated twice: in double
// (the reference's precision) and in long double (x87, 64-bit significand).
// Their difference measures the FP64 rounding noise of the reference's
// formula at each pair; it is large where the closed forms approach their
// removable singularities (sin psi or sin delta -> 0, dihedral xi -> 0).
void o_singular_integrals(void* h, int type, const int32_t* ij, int64_t n, double* out_d4, double* out_ld4) {
    const Mesh& m = *(Mesh*)h;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < n; ++q) {
        const int i = ij[2 * q], j = ij[2 * q + 1];
        TaskGeom<double> g = task_geom<double>(m, i, j, nullptr, -1);
        Q4<double> r = singular_integral(m, type, i, j, g);
        TaskGeom<long double> gl = task_geom<long double>(m, i, j, nullptr, -1);
        Q4<long double> rl = singular_integral(m, type, i, j, gl);
        out_d4[4 * q] = r.x; out_d4[4 * q + 1] = r.y; out_d4[4 * q + 2] = r.z; out_d4[4 * q + 3] = r.w;
        out_ld4[4 * q] = (double)rl.x; out_ld4[4 * q + 1] = (double)rl.y; out_ld4[4 * q + 2] = (double)rl.z; out_ld4[4 * q + 3] = (double)rl.w;
    }
}

// FP64 rounding noise of the reference's formulas for whole tasks: the
// reported value of task q (numerical part at level rlevel[q] plus the closed
// singular part) evaluated in double and in long double; returns
// ||I_double - I_longdouble||_inf / ||I_longdouble||_inf per task.  Used by the
// parity gate for tasks whose quadrature points fall near the extension of an
// edge of K_j (1 + cos -> 0 in thetaPsi) or whose closed forms are near their
// removable singularities.
void o_task_noise(void* h, int type, const int32_t* ij, const int8_t* rlevel, int64_t n, const double* L12, const double* w,
                  int nPts, int order, double* out, double* out_j) {
    const Mesh& m = *(Mesh*)h;
    Rule R = make_rule(L12, w, nPts, order);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < n; ++q) {
        const int i = ij[2 * q], j = ij[2 * q + 1];
        TaskGeom<double> g = task_geom<double>(m, i, j, nullptr, -1);
        Q4<double> v = task_level_value(m, R, type, i, j, rlevel[q], g) + singular_integral(m, type, i, j, g);
        TaskGeom<long double> gl = task_geom<long double>(m, i, j, nullptr, -1);
        Q4<long double> vl = task_level_value(m, R, type, i, j, rlevel[q], gl) + singular_integral(m, type, i, j, gl);
        const double d = std::max(std::max(std::fabs(v.x - (double)vl.x), std::fabs(v.y - (double)vl.y)),
                                  std::max(std::fabs(v.z - (double)vl.z), std::fabs(v.w - (double)vl.w)));
        const double s = std::max(std::max(std::fabs((double)vl.x), std::fabs((double)vl.y)),
                                  std::max(std::fabs((double)vl.z), std::fabs((double)vl.w)));
        out[q] = d / std::max(s, 1e-300);
        if (out_j) {
            // the same for J, with the vertex-sharing 2*pi*p wrap taken from the double value
            int p = 0;
            V3<double> J = (type == VERTEX_SHARING) ? finalize_vertex(v, g.nJ, g.areaI, &p) : finalize_plain(v, g.nJ);
            V3<long double> Jl;
            if (type == VERTEX_SHARING) {
                const long double refT = (long double)TWO_PI * gl.areaI;
                Jl = (long double)RECIPROCAL_FOUR_PI * ((vl.w + (long double)(2.0 * p) * refT) * gl.nJ + cross(vecof(vl), gl.nJ));
            } else Jl = finalize_plain(vl, gl.nJ);
            const double dj = std::max(std::fabs(J.x - (double)Jl.x), std::max(std::fabs(J.y - (double)Jl.y), std::fabs(J.z - (double)Jl.z)));
            const double sj = std::max(std::fabs((double)Jl.x), std::max(std::fabs((double)Jl.y), std::fabs((double)Jl.z)));
            out_j[q] = dj / std::max(sj, 1e-300);
        }
    }
}

// Level-L value of the numerical part of task (i, j): sum over the 4^L sub-panels.
void o_level_value(void* h, int type, int i, int j, const double* L12, const double* w, int nPts, int order, int level, double* out4) {
    const Mesh& m = *(Mesh*)h;
    Rule R = make_rule(L12, w, nPts, order);
    TaskGeom<double> g = task_geom<double>(m, i, j, nullptr, -1);
    Q4<double> r = task_level_value(m, R, type, i, j, level, g);
    out4[0] = r.x; out4[1] = r.y; out4[2] = r.z; out4[3] = r.w;
}

// Runs ordered tasks (i_q, j_q) of one type through the reference's
// numerical integration with error control, the analytic singular part and
// the finalize step.
//   mode 0: fixed refinement level `level` (intended semantics: all 4^level
//           sub-panels; NumericalIntegrator3d.cu:343-362 builds wrong tasks
//           for level >= 2, which is not reproduced);
//   mode 1: automatic control with maxLevel = `level`, Runge tolerance `tol`.
// quirk != 0 reproduces the reference's buffer swap (evaluatorJ3DK.cu:976-990):
// a task converged at iteration c reports I_c when K - c is even, else I_{c-1},
// K being the number of loop iterations run for this type.  K_override >= 0
// replaces K (used when only a sample of a type's tasks is evaluated).
// Outputs per task: integrals (4), results J (3), conv = convergence iteration
// c (int8), and optionally all level values I_0..I_c (levels4, stride 4*(maxL+1)).
// prec != 0 evaluates every level value and the singular part in long double
// and rounds each once to double, so that the Runge decisions are taken on
// values free of the FP64 cancellation of the reference formula near the
// extension of an edge line; the criterion itself stays in double.
int o_run_tasks_prec(void* h, int type, const int32_t* ij, int64_t n, const double* L12, const double* w, int nPts,
                     int order, int mode, int level, double tol, int quirk, int K_override, int nthreads, int prec,
                     double* out_integrals4, double* out_results3, int8_t* out_conv, double* out_levels4, int* K_out,
                     uint8_t* out_unconv) {
    const Mesh& m = *(Mesh*)h;
    Rule R = make_rule(L12, w, nPts, order);
    const double pow2p = (double)(1 << order);
    const int maxL = level;
    std::vector<int8_t> conv(n);
    std::vector<Q4<double>> finalv(n), prevv(n);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int K = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : K)
    for (int64_t q = 0; q < n; ++q) {
        const int i = ij[2 * q], j = ij[2 * q + 1];
        TaskGeom<double> g = task_geom<double>(m, i, j, nullptr, -1);
        PairEval<double> E = pair_eval(m, type, i, j, g);
        SubPanel<double> P{g.I[0], g.I[1], g.I[2], g.areaI};
        TaskGeom<long double> gl{};
        PairEval<long double> El{};
        if (prec) { gl = task_geom<long double>(m, i, j, nullptr, -1); El = pair_eval(m, type, i, j, gl); }
        const SubPanel<long double> Pl{gl.I[0], gl.I[1], gl.I[2], gl.areaI};
        // level value in double, or in long double rounded to double once
        auto lv = [&](int L) -> Q4<double> {
            if (!prec) return level_value(R, P, L, E);
            const Q4<long double> v = level_value(R, Pl, L, El);
            return {(double)v.x, (double)v.y, (double)v.z, (double)v.w};
        };
        if (mode == 0) {
            Q4<double> v = lv(level);
            finalv[q] = v; prevv[q] = v; conv[q] = (int8_t)level;
            if (out_levels4) { double* o = out_levels4 + (size_t)q * 4 * (maxL + 1) + 4 * level; o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }
            continue;
        }
        Q4<double> prev = lv(0), cur = prev;
        if (out_levels4) { double* o = out_levels4 + (size_t)q * 4 * (maxL + 1); o[0] = prev.x; o[1] = prev.y; o[2] = prev.z; o[3] = prev.w; }
        int c = 0;
        bool still = true;   // unconverged after the last comparison made (no comparison when maxL = 0)
        for (int it = 1; it <= maxL; ++it) {
            cur = lv(it);
            if (out_levels4) { double* o = out_levels4 + (size_t)q * 4 * (maxL + 1) + 4 * it; o[0] = cur.x; o[1] = cur.y; o[2] = cur.z; o[3] = cur.w; }
            c = it;
            still = runge_unconverged(cur, prev, pow2p, tol);
            if (!still) break;
            if (it < maxL) prev = cur;
        }
        if (out_unconv) out_unconv[q] = still ? 1 : 0;
        if (c == 0) { finalv[q] = prev; prevv[q] = prev; }
        else { finalv[q] = cur; prevv[q] = prev; }
        conv[q] = (int8_t)c;
        if (c > K) K = c;
    }
    if (mode == 0) K = level;
    if (K_override >= 0) K = K_override;
    if (K_out) *K_out = K;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t q = 0; q < n; ++q) {
        const int i = ij[2 * q], j = ij[2 * q + 1];
        Q4<double> v = finalv[q];
        if (mode == 1 && quirk && conv[q] >= 1 && ((K - conv[q]) & 1)) v = prevv[q];
        TaskGeom<double> g = task_geom<double>(m, i, j, nullptr, -1);
        if (prec) {
            const Q4<long double> s = singular_integral(m, type, i, j, task_geom<long double>(m, i, j, nullptr, -1));
            v = v + Q4<double>{(double)s.x, (double)s.y, (double)s.z, (double)s.w};
        } else {
            v = v + singular_integral(m, type, i, j, g);
        }
        V3<double> J = (type == VERTEX_SHARING) ? finalize_vertex(v, g.nJ, g.areaI) : finalize_plain(v, g.nJ);
        if (out_integrals4) { double* o = out_integrals4 + 4 * q; o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }
        if (out_results3) { double* o = out_results3 + 3 * q; o[0] = J.x; o[1] = J.y; o[2] = J.z; }
        if (out_conv) out_conv[q] = conv[q];
    }
    return 0;
}

int o_run_tasks(void* h, int type, const int32_t* ij, int64_t n, const double* L12, const double* w, int nPts, int order,
                int mode, int level, double tol, int quirk, int K_override, int nthreads,
                double* out_integrals4, double* out_results3, int8_t* out_conv, double* out_levels4, int* K_out) {
    return o_run_tasks_prec(h, type, ij, n, L12, w, nPts, order, mode, level, tol, quirk, K_override, nthreads, 0,
                            out_integrals4, out_results3, out_conv, out_levels4, K_out, nullptr);
}

// Per-cell refinement counts of one type (NumericalIntegrator3D::
// getRefinementsRequired), by the call sequence of EvaluatorJ3DK::
// numericalIntegration (evaluatorJ3DK.cu:962-1010): determineCellsToBeRefined
// (NumericalIntegrator3d.cu:461-499) runs once on the full task list before the
// loop and once after every comparison that leaves tasks in the rest list, and
// adds one to each cell that is control panel i of a task in that list.  Task q
// is in the rest list after the comparison at iteration `it` while it has not
// converged: it < conv[q], or it == conv[q] with the task unconverged after its
// last comparison (unconv[q], from o_run_tasks_prec).
void o_refinement_counts(int nT, const int32_t* ij, int64_t n, const int8_t* conv, const uint8_t* unconv, int maxL,
                         uint8_t* out) {
    std::vector<int64_t> rest(n);
    for (int64_t q = 0; q < n; ++q) rest[q] = q;
    std::fill(out, out + nT, (uint8_t)0);
    std::vector<uint8_t> flag(nT);
    auto determine = [&](const std::vector<int64_t>& r) {
        std::fill(flag.begin(), flag.end(), (uint8_t)0);
        for (int64_t q : r) flag[ij[2 * q]] = 1;
        for (int i = 0; i < nT; ++i) out[i] = (uint8_t)(out[i] + flag[i]);
    };
    if (n > 0) determine(rest);
    int iteration = 0;
    while (!rest.empty() && iteration < maxL) {
        ++iteration;
        std::vector<int64_t> next;
        for (int64_t q : rest)
            if (conv[q] > iteration || (conv[q] == iteration && unconv[q])) next.push_back(q);
        rest.swap(next);
        if (!rest.empty()) determine(rest);
    }
}

// evaluator3d.cu:45-57 -- symmetry error for a task list laid out as
// [(i,j) x n] ++ [(j,i) x n].
void o_symmetry(const double* results3, int64_t n, double* out_delta) {
    for (int64_t q = 0; q < n; ++q) {
        V3<double> a{results3[3 * q], results3[3 * q + 1], results3[3 * q + 2]};
        V3<double> b{results3[3 * (n + q)], results3[3 * (n + q) + 1], results3[3 * (n + q) + 2]};
        double d = norm1(a + b) / std::max(norm1(a), norm1(b));
        out_delta[q] = d; out_delta[n + q] = d;
    }
}

double o_normal_draw(uint64_t seed, uint64_t pair_id, uint64_t stream) { return normal_draw(seed, pair_id, stream); }

}  // extern "C"

namespace {
// Loss term w Theta + v . J of task (i, j) at level rl with the wrap count p,
// geometry in scalar T with vertex `vtx` replaced by *moved (vtx < 0: none).
template <class T>
T task_loss(const Mesh& m, const Rule& R, int type, int i, int j, int rl, int p, const V3<T>* moved, int vtx,
            double wq, const double vq[3]) {
    TaskGeom<T> g = task_geom<T>(m, i, j, moved, vtx);
    Q4<T> v = task_level_value(m, R, type, i, j, rl, g) + singular_integral(m, type, i, j, g);
    V3<T> J;
    if (type == VERTEX_SHARING) {
        const T refT = T(TWO_PI) * g.areaI;
        J = T(RECIPROCAL_FOUR_PI) * ((v.w + T(2.0 * p) * refT) * g.nJ + cross(vecof(v), g.nJ));
    } else J = finalize_plain(v, g.nJ);
    return T(wq) * v.w + T(vq[0]) * J.x + T(vq[1]) * J.y + T(vq[2]) * J.z;
}
template <class R_> V3<R_> moved_vertex(const Mesh& m, int vtx, int axis, R_ h) {
    V3<R_> x{R_(m.v[vtx].x), R_(m.v[vtx].y), R_(m.v[vtx].z)};
    if (axis == 0) x.x += h; else if (axis == 1) x.y += h; else x.z += h;
    return x;
}
template <class R_> V3<DualT<R_>> seeded_vertex(const Mesh& m, int vtx, int axis) {
    V3<DualT<R_>> x{DualT<R_>(R_(m.v[vtx].x)), DualT<R_>(R_(m.v[vtx].y)), DualT<R_>(R_(m.v[vtx].z))};
    if (axis == 0) x.x.d = 1; else if (axis == 1) x.y.d = 1; else x.z.d = 1;
    return x;
}
}  // namespace

extern "C" {

// Loss contribution of ordered tasks (i_q, j_q) of one type, each at its
// frozen reported level rlevel[q], with vertex `vtx` moved by `h` along
// `axis` (vtx < 0: no move).  Weights: w = N(0,1) stream 0, v = N(0,1)^3
// streams 1..3, pair id = i*nT + j.  The 2*pi*p wrap of vertex-sharing
// tasks is frozen at its unperturbed value.  dual != 0 returns dL/dx of the
// same sum (forward mode) instead of its value.  prec = 1 evaluates the
// geometry and formulas in long double (64-bit significand): the finite
// differences of the adjoint's check use it so that their rounding noise,
// divided by the step, stays far below the gate.
double o_loss_tasks_prec(void* h, int type, const int32_t* ij, const int8_t* rlevel, int64_t n,
                         const double* L12, const double* w, int nPts, int order, uint64_t seed,
                         int vtx, int axis, double hstep, int dual, int prec, int nthreads) {
    const Mesh& m = *(Mesh*)h;
    Rule R = make_rule(L12, w, nPts, order);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    long double total = 0.0L;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t q = 0; q < n; ++q) {
        const int i = ij[2 * q], j = ij[2 * q + 1];
        const uint64_t pid = (uint64_t)i * (uint64_t)m.nT + (uint64_t)j;
        const double wq = normal_draw(seed, pid, 0);
        const double vq[3] = {normal_draw(seed, pid, 1), normal_draw(seed, pid, 2), normal_draw(seed, pid, 3)};
        // unperturbed value fixes the 2*pi*p wrap
        TaskGeom<double> g0 = task_geom<double>(m, i, j, nullptr, -1);
        Q4<double> v0 = task_level_value(m, R, type, i, j, rlevel[q], g0) + singular_integral(m, type, i, j, g0);
        int p = 0;
        if (type == VERTEX_SHARING) finalize_vertex(v0, g0.nJ, g0.areaI, &p);
        const int rl = rlevel[q];
        if (!dual && !prec) {
            V3<double> mv = vtx >= 0 ? moved_vertex<double>(m, vtx, axis, hstep) : V3<double>{0, 0, 0};
            total += task_loss<double>(m, R, type, i, j, rl, p, vtx >= 0 ? &mv : nullptr, vtx, wq, vq);
        } else if (!dual) {
            V3<long double> mv = vtx >= 0 ? moved_vertex<long double>(m, vtx, axis, (long double)hstep)
                                          : V3<long double>{0, 0, 0};
            total += task_loss<long double>(m, R, type, i, j, rl, p, vtx >= 0 ? &mv : nullptr, vtx, wq, vq);
        } else if (!prec) {
            V3<Dual> mv = seeded_vertex<double>(m, vtx, axis);
            total += task_loss<Dual>(m, R, type, i, j, rl, p, &mv, vtx, wq, vq).d;
        } else {
            V3<DualL> mv = seeded_vertex<long double>(m, vtx, axis);
            total += task_loss<DualL>(m, R, type, i, j, rl, p, &mv, vtx, wq, vq).d;
        }
    }
    return (double)total;
}
double o_loss_tasks(void* h, int type, const int32_t* ij, const int8_t* rlevel, int64_t n,
                    const double* L12, const double* w, int nPts, int order, uint64_t seed,
                    int vtx, int axis, double hstep, int dual, int nthreads) {
    return o_loss_tasks_prec(h, type, ij, rlevel, n, L12, w, nPts, order, seed, vtx, axis, hstep, dual, 0, nthreads);
}

// Monte Carlo arithmetic noise of tasks at their reported levels: the
// largest deviation over `nrep` randomly rounded evaluations from the plain
// FP64 value, relative to the task's integrals (out) and to its J (out_j).
void o_task_noise_mca(void* h, int type, const int32_t* ij, const int8_t* rlevel, int64_t n, const double* L12,
                      const double* w, int nPts, int order, int nrep, double* out, double* out_j) {
    const Mesh& m = *(Mesh*)h;
    Rule R = make_rule(L12, w, nPts, order);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < n; ++q) {
        const int i = ij[2 * q], j = ij[2 * q + 1];
        TaskGeom<double> g = task_geom<double>(m, i, j, nullptr, -1);
        const Q4<double> v = task_level_value(m, R, type, i, j, rlevel[q], g) + singular_integral(m, type, i, j, g);
        int p = 0;
        const V3<double> J = (type == VERTEX_SHARING) ? finalize_vertex(v, g.nJ, g.areaI, &p) : finalize_plain(v, g.nJ);
        const double s = std::max(std::max(std::fabs(v.x), std::fabs(v.y)), std::max(std::fabs(v.z), std::fabs(v.w)));
        const double sj = std::max(std::fabs(J.x), std::max(std::fabs(J.y), std::fabs(J.z)));
        double d = 0.0, dj = 0.0;
        for (int r = 0; r < nrep; ++r) {
            mca_state() = 0x1234567ull * (uint64_t)(q + 1) + 0x9E37ull * (uint64_t)(r + 1);
            TaskGeom<MCA> gm = task_geom<MCA>(m, i, j, nullptr, -1);
            const Q4<MCA> vm = task_level_value(m, R, type, i, j, rlevel[q], gm) + singular_integral(m, type, i, j, gm);
            const MCA refT = MCA(TWO_PI) * gm.areaI;
            const V3<MCA> Jm = (type == VERTEX_SHARING)
                                   ? MCA(RECIPROCAL_FOUR_PI) * ((vm.w + MCA(2.0 * p) * refT) * gm.nJ + cross(vecof(vm), gm.nJ))
                                   : finalize_plain(vm, gm.nJ);
            d = std::max(d, std::max(std::max(std::fabs(vm.x.v - v.x), std::fabs(vm.y.v - v.y)),
                                     std::max(std::fabs(vm.z.v - v.z), std::fabs(vm.w.v - v.w))));
            dj = std::max(dj, std::max(std::fabs(Jm.x.v - J.x), std::max(std::fabs(Jm.y.v - J.y), std::fabs(Jm.z.v - J.z))));
        }
        out[q] = d / std::max(s, 1e-300);
        if (out_j) out_j[q] = dj / std::max(sj, 1e-300);
    }
}

// Integrals (numerical part at level rlevel[q] plus the closed singular part)
// and J of ordered tasks at given levels, for comparing values when the
// refinement decisions of the compared runs are imposed.
// prec = 1: the same in long double (J with the wrap count of the double value).
void o_tasks_at_levels_prec(void* h, int type, const int32_t* ij, const int8_t* rlevel, int64_t n, const double* L12,
                            const double* w, int nPts, int order, int prec, double* out4, double* out3) {
    const Mesh& m = *(Mesh*)h;
    Rule R = make_rule(L12, w, nPts, order);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < n; ++q) {
        const int i = ij[2 * q], j = ij[2 * q + 1];
        TaskGeom<double> g = task_geom<double>(m, i, j, nullptr, -1);
        const Q4<double> v = task_level_value(m, R, type, i, j, rlevel[q], g) + singular_integral(m, type, i, j, g);
        int p = 0;
        const V3<double> J = (type == VERTEX_SHARING) ? finalize_vertex(v, g.nJ, g.areaI, &p) : finalize_plain(v, g.nJ);
        if (!prec) {
            out4[4 * q] = v.x; out4[4 * q + 1] = v.y; out4[4 * q + 2] = v.z; out4[4 * q + 3] = v.w;
            out3[3 * q] = J.x; out3[3 * q + 1] = J.y; out3[3 * q + 2] = J.z;
            continue;
        }
        TaskGeom<long double> gl = task_geom<long double>(m, i, j, nullptr, -1);
        const Q4<long double> vl = task_level_value(m, R, type, i, j, rlevel[q], gl) + singular_integral(m, type, i, j, gl);
        const long double refT = (long double)TWO_PI * gl.areaI;
        const V3<long double> Jl = (long double)RECIPROCAL_FOUR_PI *
                                   ((vl.w + (long double)(2.0 * p) * refT) * gl.nJ + cross(vecof(vl), gl.nJ));
        out4[4 * q] = (double)vl.x; out4[4 * q + 1] = (double)vl.y; out4[4 * q + 2] = (double)vl.z; out4[4 * q + 3] = (double)vl.w;
        out3[3 * q] = (double)Jl.x; out3[3 * q + 1] = (double)Jl.y; out3[3 * q + 2] = (double)Jl.z;
    }
}
void o_tasks_at_levels(void* h, int type, const int32_t* ij, const int8_t* rlevel, int64_t n, const double* L12,
                       const double* w, int nPts, int order, double* out4, double* out3) {
    o_tasks_at_levels_prec(h, type, ij, rlevel, n, L12, w, nPts, order, 0, out4, out3);
}

// Wall-clock helper for the CPU baseline.
double o_now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
int o_max_threads() { return omp_get_max_threads(); }

}  // extern "C"
