/*
 * integrator2.hpp -- the reference's C++ API over the C ABI of
 * libintegrator2_mi355x.so (header-only, host C++17; g++ or hipcc).
 *
 * Class names, member names, argument meaning and enum values are those of
 * andreyypopov/integrator2:
 *   src/Mesh3d.cuh                     Mesh3D, neighbour_type_enum, exportMeshToObj,
 *                                      exportMeshToVtk, neighborTypeString
 *   src/QuadratureFormula3d.cuh        QuadratureFormula3D, qf3D1 ... qf3D13
 *   src/NumericalIntegrator3d.cuh      NumericalIntegrator3D, error_control_type_enum
 *   src/evaluators/evaluator3d.cuh     Evaluator3D, output_format_enum
 *   src/evaluators/evaluatorJ3DK.cuh   EvaluatorJ3DK
 *   src/common/device_vector.cuh       deviceVector (here a view, see below)
 *   src/common/cuda_memory.cuh         copy_d2h
 * so the reference's driver (tests/integrator3D/main.cu) compiles against
 * this one header in place of the three it includes.
 *
 * Differences, each deliberate:
 *   - deviceVector<T> returned by the getters is a view of device memory that
 *     the library owns (data, size, capacity as in the reference, 64-bit
 *     sizes); it is never freed by the caller and is valid until the owning
 *     object runs again or is destroyed.
 *   - HIP and library errors throw integrator2_error instead of exiting the
 *     process (src/common/cuda_helper.cuh:41-54); loadMeshFromFile and
 *     outputResultsToFile keep their bool results.
 *   - The per-level machinery of the reference's integrator
 *     (prepareTasksAndMesh, gatherResults, refineMesh, resetMesh,
 *     determineCellsToBeRefined, getRefinedTasks, getResults) has no
 *     counterpart: refinement runs inside the library's kernels, so
 *     Evaluator3D's integrateOver*Neighbors are not virtual extension points.
 *   - Added: getIntegrals / getResults / getErrors / getConvergenceLevels
 *     (protected device buffers in the reference), setErrorControl (run-time
 *     maxLevel and tolerance, compile-time constants in the reference),
 *     setPath / setTuning (plain or batched kernels).
 *   - Objects are bound to HIP device 0 unless given a device.
 */
#ifndef INTEGRATOR2_MI355X_HPP
#define INTEGRATOR2_MI355X_HPP

#include <hip/hip_vector_types.h>   // int3, double2, double3, double4

#include <array>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../integrator2_mi355x.h"

using Point3 = double3;   // src/common/cuda_math.cuh

enum class neighbour_type_enum {
    simple_neighbors = I2_SIMPLE_NEIGHBORS,
    attached_neighbors = I2_ATTACHED_NEIGHBORS,
    not_neighbors = I2_NOT_NEIGHBORS,
    undefined = -1
};

enum class output_format_enum { plainText = I2_OUTPUT_PLAIN_TEXT, csv = I2_OUTPUT_CSV };

enum class error_control_type_enum { fixed_refinement_level = 0, automatic_error_control = 1 };

class integrator2_error : public std::runtime_error {
public:
    integrator2_error(int status, const char* what) : std::runtime_error(what ? what : "integrator2 error"), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

namespace i2cpp {
inline void check(int status) {
    if (status != I2_OK) throw integrator2_error(status, i2_last_error());
}
inline int type_index(neighbour_type_enum t) {
    const int k = static_cast<int>(t);
    if (k < 0 || k > 2) throw integrator2_error(I2_ERR_INVALID, "undefined neighbour type");
    return k;
}
}  // namespace i2cpp

// View of a device array owned by the library (src/common/device_vector.cuh).
template <class T>
struct deviceVector {
    T* data = nullptr;
    int64_t size = 0;
    int64_t capacity = 0;
    size_t bytes() const { return static_cast<size_t>(size) * sizeof(T); }
};

// copy_d2h (src/common/cuda_memory.cuh:104-109), synchronous.
template <class T>
inline void copy_d2h(const T* src, const T* dst, size_t elements_num) {
    if (elements_num)
        i2cpp::check(i2_copy_to_host(const_cast<T*>(dst), src, static_cast<int64_t>(elements_num * sizeof(T))));
    else
        std::printf("Zero device-to-host memory copy requested\n");
}

// ---------------------------------------------------------------- QuadratureFormula3D
struct QuadratureFormula3D {
    const std::vector<double2> coordinates;   // (L1, L2); L3 = 1 - L1 - L2
    const std::vector<double> weights;
    const int order;                          // Runge factor 2^order
};

namespace i2cpp {
inline QuadratureFormula3D library_rule(int n) {
    std::vector<double> L12(2 * n), w(n);
    int order = 0;
    check(i2_quadrature_rule(n, L12.data(), w.data(), &order));
    std::vector<double2> c(n);
    for (int k = 0; k < n; ++k) c[k] = make_double2(L12[2 * k], L12[2 * k + 1]);
    return QuadratureFormula3D{c, w, order};
}
}  // namespace i2cpp

// The reference's rules (Cowper 1973), same decimal literals.
inline const QuadratureFormula3D qf3D1 = i2cpp::library_rule(1);
inline const QuadratureFormula3D qf3D3 = i2cpp::library_rule(3);
inline const QuadratureFormula3D qf3D4 = i2cpp::library_rule(4);
inline const QuadratureFormula3D qf3D6 = i2cpp::library_rule(6);
inline const QuadratureFormula3D qf3D7 = i2cpp::library_rule(7);
inline const QuadratureFormula3D qf3D9 = i2cpp::library_rule(9);
inline const QuadratureFormula3D qf3D12 = i2cpp::library_rule(12);
inline const QuadratureFormula3D qf3D13 = i2cpp::library_rule(13);

// ---------------------------------------------------------------- Mesh3D
inline std::string neighborTypeString(neighbour_type_enum neighborType) {
    switch (neighborType) {
        case neighbour_type_enum::simple_neighbors: return "SimpleNeighbors";
        case neighbour_type_enum::attached_neighbors: return "AttachedNeighbors";
        case neighbour_type_enum::not_neighbors: return "NotNeighbors";
        default: return std::string();
    }
}

class Mesh3D {
public:
    explicit Mesh3D(int device = 0) { i2cpp::check(i2_mesh_create(&m_, device)); }
    virtual ~Mesh3D() { i2_mesh_destroy(m_); }
    Mesh3D(const Mesh3D&) = delete;
    Mesh3D& operator=(const Mesh3D&) = delete;

    // .dat reader of src/Mesh3d.cu:159-217 (declared counts, 1-based indices, type-203 cells).
    bool loadMeshFromFile(const std::string& filename, double scale = 1.0) {
        if (i2_mesh_load_file(m_, filename.c_str(), scale) != I2_OK) {
            std::printf("Error while opening the file\n");
            return false;
        }
        int64_t nV = 0, nT = 0, pairs[3];
        i2cpp::check(i2_mesh_sizes(m_, &nV, &nT, pairs));
        std::printf("Loaded mesh with %lld vertices and %lld cells\n", (long long)nV, (long long)nT);
        return true;
    }
    // Vertices and cells from host arrays (no counterpart: the reference only loads files).
    void setMesh(const std::vector<Point3>& vertices, const std::vector<int3>& cells) {
        i2cpp::check(i2_mesh_set_data(m_, reinterpret_cast<const double*>(vertices.data()), (int64_t)vertices.size(),
                                      reinterpret_cast<const int32_t*>(cells.data()), (int64_t)cells.size()));
    }

    void prepareMesh() {
        i2cpp::check(i2_mesh_prepare(m_));
        int64_t nV = 0, nT = 0, pairs[3];
        i2cpp::check(i2_mesh_sizes(m_, &nV, &nT, pairs));
        std::printf("Found %lld pairs of simple neighbors and %lld pairs of attached neighbors, %lld pairs are not neighbors\n",
                    (long long)pairs[0], (long long)pairs[1], (long long)pairs[2]);
    }

    const deviceVector<int3>& getSimpleNeighbors() const { return view(4, simple_); }
    const deviceVector<int3>& getAttachedNeighbors() const { return view(5, attached_); }
    // The disjoint (i < j, k) list is written out on the first call (the kernels decode it on the fly).
    const deviceVector<int3>& getNotNeighbors() const { return view(6, not_); }
    const deviceVector<Point3>& getVertices() const { return view(0, vertices_); }
    const deviceVector<int3>& getCells() const { return view(1, cells_); }
    const deviceVector<Point3>& getCellNormals() const { return view(2, normals_); }
    const deviceVector<double>& getCellMeasures() const { return view(3, measures_); }

    i2_mesh* handle() const { return m_; }

private:
    template <class T>
    const deviceVector<T>& view(int what, deviceVector<T>& v) const {
        const void* p = nullptr;
        int64_t n = 0;
        i2cpp::check(i2_mesh_device_array(m_, what, &p, &n));
        v.data = static_cast<T*>(const_cast<void*>(p));
        v.size = v.capacity = n;
        return v;
    }
    i2_mesh* m_ = nullptr;
    mutable deviceVector<int3> simple_, attached_, not_, cells_;
    mutable deviceVector<Point3> vertices_, normals_;
    mutable deviceVector<double> measures_;
};

// exportMeshToObj (src/Mesh3d.cu:264-276)
inline void exportMeshToObj(const std::string& filename, const std::vector<Point3>& vertices, const std::vector<int3>& cells) {
    i2cpp::check(i2_export_mesh_obj(filename.c_str(), reinterpret_cast<const double*>(vertices.data()), (int64_t)vertices.size(),
                                    reinterpret_cast<const int32_t*>(cells.data()), (int64_t)cells.size()));
    std::printf("Mesh saved to %s\n", filename.c_str());
}

// exportMeshToVtk (src/Mesh3d.cu:278-343); an empty count vector writes no array for that type.
inline void exportMeshToVtk(const std::string& filename, const std::vector<Point3>& vertices, const std::vector<int3>& cells,
                            const std::array<std::vector<unsigned char>, 3>& refinementsRequired) {
    const uint8_t* r[3];
    for (int k = 0; k < 3; ++k) {
        const auto& v = refinementsRequired[k];
        if (!v.empty() && v.size() != cells.size()) throw integrator2_error(I2_ERR_INVALID, "refinement counts need one entry per cell");
        r[k] = v.empty() ? nullptr : v.data();
    }
    i2cpp::check(i2_export_mesh_vtk(filename.c_str(), reinterpret_cast<const double*>(vertices.data()), (int64_t)vertices.size(),
                                    reinterpret_cast<const int32_t*>(cells.data()), (int64_t)cells.size(), r[0], r[1], r[2]));
    std::printf("Mesh saved to %s\n", filename.c_str());
}

// ---------------------------------------------------------------- NumericalIntegrator3D
class NumericalIntegrator3D {
public:
    NumericalIntegrator3D(const Mesh3D& mesh_, const QuadratureFormula3D& qf_) : mesh(mesh_), qf(qf_) {
        const int n = static_cast<int>(qf_.weights.size());
        if (qf_.coordinates.size() != qf_.weights.size()) throw integrator2_error(I2_ERR_INVALID, "rule coordinates and weights differ in length");
        std::vector<double> L12(2 * n);
        for (int k = 0; k < n; ++k) { L12[2 * k] = qf_.coordinates[k].x; L12[2 * k + 1] = qf_.coordinates[k].y; }
        i2cpp::check(i2_integrator_create(&it_, mesh_.handle(), L12.data(), qf_.weights.data(), n, qf_.order));
    }
    virtual ~NumericalIntegrator3D() { i2_integrator_destroy(it_); }
    NumericalIntegrator3D(const NumericalIntegrator3D&) = delete;
    NumericalIntegrator3D& operator=(const NumericalIntegrator3D&) = delete;

    void setFixedRefinementLevel(int refinementLevel = 0) {
        i2cpp::check(i2_integrator_set_fixed_refinement_level(it_, refinementLevel));
        errorControlType = error_control_type_enum::fixed_refinement_level;
    }
    // Automatic error control with run-time maxLevel and tolerance (the reference's
    // MAX_REFINE_LEVEL 5 and EPS_INTEGRATION 1e-5 are the defaults).
    void setErrorControl(int maxLevel = 5, double tolerance = 1e-5, bool reproduceBufferSwap = true) {
        i2cpp::check(i2_integrator_set_error_control(it_, maxLevel, tolerance, reproduceBufferSwap ? 1 : 0));
        errorControlType = error_control_type_enum::automatic_error_control;
    }

    int getGaussPointsNumber() const { return static_cast<int>(qf.weights.size()); }
    int getQuadratureFormulaOrder() const { return qf.order; }
    error_control_type_enum getErrorControlType() const { return errorControlType; }

    // Fixed-level mode: the mesh split to that level (built on the first call).
    const deviceVector<Point3>& getRefinedVertices() const { refined(); return refinedVertices; }
    const deviceVector<int3>& getRefinedCells() const { refined(); return refinedCells; }
    const deviceVector<double>& getRefinedCellMeasures() const { refined(); return refinedCellMeasures; }

    // Per original cell, refinement passes of the last automatic run (size 0 otherwise).
    deviceVector<unsigned char>* getRefinementsRequired(neighbour_type_enum neighborType) {
        const int t = i2cpp::type_index(neighborType);
        const void* p = nullptr;
        int64_t n = 0;
        i2cpp::check(i2_integrator_refinements_required(it_, t, &p, nullptr, &n));
        refinementsRequired[t].data = static_cast<unsigned char*>(const_cast<void*>(p));
        refinementsRequired[t].size = refinementsRequired[t].capacity = n;
        return &refinementsRequired[t];
    }

    i2_integrator* handle() const { return it_; }

private:
    void refined() const {
        const void *v = nullptr, *c = nullptr, *s = nullptr;
        int64_t nV = 0, nT = 0;
        i2cpp::check(i2_integrator_refined_mesh(it_, &v, &c, &s, &nV, &nT));
        refinedVertices.data = static_cast<Point3*>(const_cast<void*>(v));
        refinedVertices.size = refinedVertices.capacity = nV;
        refinedCells.data = static_cast<int3*>(const_cast<void*>(c));
        refinedCells.size = refinedCells.capacity = nT;
        refinedCellMeasures.data = static_cast<double*>(const_cast<void*>(s));
        refinedCellMeasures.size = refinedCellMeasures.capacity = nT;
    }
    const Mesh3D& mesh;
    const QuadratureFormula3D& qf;
    i2_integrator* it_ = nullptr;
    error_control_type_enum errorControlType = error_control_type_enum::automatic_error_control;
    mutable deviceVector<Point3> refinedVertices;
    mutable deviceVector<int3> refinedCells;
    mutable deviceVector<double> refinedCellMeasures;
    deviceVector<unsigned char> refinementsRequired[3];
};

// ---------------------------------------------------------------- Evaluator3D / EvaluatorJ3DK
class Evaluator3D {
public:
    Evaluator3D(const Mesh3D& mesh_, NumericalIntegrator3D& numIntegrator_) : mesh(mesh_), numIntegrator(numIntegrator_) {
        i2cpp::check(i2_evaluator_create(&ev_, mesh_.handle(), numIntegrator_.handle()));
    }
    virtual ~Evaluator3D() { i2_evaluator_destroy(ev_); }
    Evaluator3D(const Evaluator3D&) = delete;
    Evaluator3D& operator=(const Evaluator3D&) = delete;

    // All ordered pairs [(i, j, k)] ++ [(j, i, n + k)] per type (src/evaluators/evaluator3d.cu:120-204).
    virtual void runAllPairs(bool checkCorrectness = false) {
        i2cpp::check(i2_evaluator_run_all_pairs(ev_, checkCorrectness ? 1 : 0));
        report_times();
    }

    // User task lists (src/evaluators/evaluator3d.cu:213-288); all empty: nothing is run.
    void runPairs(const std::vector<int3>& userSimpleNeighborsTasks, const std::vector<int3>& userAttachedNeighborsTasks,
                  const std::vector<int3>& userNotNeighborsTasks) {
        if (userSimpleNeighborsTasks.empty() && userAttachedNeighborsTasks.empty() && userNotNeighborsTasks.empty()) return;
        i2cpp::check(i2_evaluator_run_pairs(
            ev_, reinterpret_cast<const int32_t*>(userSimpleNeighborsTasks.data()), (int64_t)userSimpleNeighborsTasks.size(),
            reinterpret_cast<const int32_t*>(userAttachedNeighborsTasks.data()), (int64_t)userAttachedNeighborsTasks.size(),
            reinterpret_cast<const int32_t*>(userNotNeighborsTasks.data()), (int64_t)userNotNeighborsTasks.size()));
        report_times();
    }

    // SimpleNeighbors / AttachedNeighbors / NotNeighbors .dat or .csv in the current directory
    // (src/evaluators/evaluator3d.cu:354-457).
    bool outputResultsToFile(neighbour_type_enum neighborType, output_format_enum outputFormat) const {
        const int t = i2cpp::type_index(neighborType);
        if (i2_evaluator_output_results_to_file(ev_, t, static_cast<int>(outputFormat), nullptr) != I2_OK) {
            std::printf("Error while opening the file\n");
            return false;
        }
        int64_t n = 0;
        i2cpp::check(i2_evaluator_task_count(ev_, t, &n));
        std::printf("%lld results saved to file %s%s\n", (long long)n, neighborTypeString(neighborType).c_str(),
                    outputFormat == output_format_enum::csv ? ".csv" : ".dat");
        return true;
    }

    const deviceVector<int3>* getTasks(neighbour_type_enum neighborType) const {
        if (neighborType == neighbour_type_enum::undefined) return nullptr;
        const int t = i2cpp::type_index(neighborType);
        const void* p = nullptr;
        int64_t n = 0;
        i2cpp::check(i2_evaluator_device_tasks(ev_, t, &p, &n));
        tasks_[t].data = static_cast<int3*>(const_cast<void*>(p));
        tasks_[t].size = tasks_[t].capacity = n;
        return &tasks_[t];
    }

    // ---- added accessors (host copies, task order of getTasks)
    std::vector<double4> getIntegrals(neighbour_type_enum neighborType) const {
        std::vector<double4> v(count(neighborType));
        i2cpp::check(i2_evaluator_get_results(ev_, i2cpp::type_index(neighborType), nullptr, reinterpret_cast<double*>(v.data()),
                                              nullptr, nullptr, nullptr));
        return v;
    }
    std::vector<Point3> getResults(neighbour_type_enum neighborType) const {
        std::vector<Point3> v(count(neighborType));
        i2cpp::check(i2_evaluator_get_results(ev_, i2cpp::type_index(neighborType), nullptr, nullptr,
                                              reinterpret_cast<double*>(v.data()), nullptr, nullptr));
        return v;
    }
    std::vector<double> getErrors(neighbour_type_enum neighborType) const {
        std::vector<double> v(count(neighborType));
        i2cpp::check(i2_evaluator_get_results(ev_, i2cpp::type_index(neighborType), nullptr, nullptr, nullptr, v.data(), nullptr));
        return v;
    }
    std::vector<int8_t> getConvergenceLevels(neighbour_type_enum neighborType) const {
        std::vector<int8_t> v(count(neighborType));
        i2cpp::check(i2_evaluator_get_results(ev_, i2cpp::type_index(neighborType), nullptr, nullptr, nullptr, nullptr, v.data()));
        return v;
    }
    // 0: plain per-type, per-level launches; 1: batched persistent kernels (runAllPairs).
    void setPath(int path) { i2cpp::check(i2_evaluator_set_path(ev_, path)); }
    void setTuning(int pairsPerWorkgroup, int pointsPerLane) { i2cpp::check(i2_evaluator_set_tuning(ev_, pairsPerWorkgroup, pointsPerLane)); }

    i2_evaluator* handle() const { return ev_; }

protected:
    const Mesh3D& mesh;
    NumericalIntegrator3D& numIntegrator;

private:
    int64_t count(neighbour_type_enum neighborType) const {
        int64_t n = 0;
        i2cpp::check(i2_evaluator_task_count(ev_, i2cpp::type_index(neighborType), &n));
        return n;
    }
    void report_times() const {
        static const char* names[3] = {"Simple neighbors integration", "Attached neighbors integration", "Non-neighbors integration"};
        for (int t = 0; t < 3; ++t) {
            double ms = 0.0;
            int64_t refined = 0;
            if (i2_evaluator_type_stats(ev_, t, &ms, &refined) == I2_OK && count(static_cast<neighbour_type_enum>(t)) > 0)
                std::printf("Time for %s: %6.3f ms\n", names[t], ms);
        }
    }
    i2_evaluator* ev_ = nullptr;
    mutable deviceVector<int3> tasks_[3];
};

// EvaluatorJ3DK (src/evaluators/evaluatorJ3DK.cuh:151-185): the gradient of the
// Newtonian potential's panel-pair integrals, (Psi, Theta) and J per ordered pair.
class EvaluatorJ3DK : public Evaluator3D {
public:
    EvaluatorJ3DK(const Mesh3D& mesh_, NumericalIntegrator3D& numIntegrator_) : Evaluator3D(mesh_, numIntegrator_) {}
};

#endif  // INTEGRATOR2_MI355X_HPP
