// i2_rules_io.cpp -- host-side pieces of the drop-in: the triangle rules of
// QuadratureFormula3D and the mesh export of Mesh3D.
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>

#include "../../include/integrator2_mi355x.h"

namespace i2 {
void set_error(const char* fmt, ...);
}

namespace {

// Gauss-point tables on a triangle (Cowper 1973, as the decimal literals of
// src/QuadratureFormula3d.cuh:34-213): (L1, L2) per point, L3 = 1 - L1 - L2.
// The literals are kept digit for digit: results at 1e-11 depend on them.
// qf3D1: 1 point, order 2
static const double kL12_1[] = {
    0.3333333333333333, 0.3333333333333333,
};
static const double kW_1[] = {
    1.0,
};
// qf3D3: 3 points, order 2
static const double kL12_3[] = {
    0.1666666666666667, 0.1666666666666667,
    0.6666666666666667, 0.1666666666666667,
    0.1666666666666667, 0.6666666666666667,
};
static const double kW_3[] = {
    0.3333333333333333,
    0.3333333333333333,
    0.3333333333333333,
};
// qf3D4: 4 points, order 3
static const double kL12_4[] = {
    0.3333333333333333, 0.3333333333333333,
    0.6, 0.2,
    0.2, 0.6,
    0.2, 0.2,
};
static const double kW_4[] = {
    -0.5625,
    0.5208333333333333,
    0.5208333333333333,
    0.5208333333333333,
};
// qf3D6: 6 points, order 4
static const double kL12_6[] = {
    0.816847572980459, 0.091576213509771,
    0.091576213509771, 0.816847572980459,
    0.091576213509771, 0.091576213509771,
    0.445948490915965, 0.108103018168070,
    0.108103018168070, 0.445948490915965,
    0.445948490915965, 0.445948490915965,
};
static const double kW_6[] = {
    0.109951743655322,
    0.109951743655322,
    0.109951743655322,
    0.223381589678011,
    0.223381589678011,
    0.223381589678011,
};
// qf3D7: 7 points, order 5
static const double kL12_7[] = {
    0.3333333333333333, 0.3333333333333333,
    0.101286507323456, 0.797426985353087,
    0.797426985353087, 0.101286507323456,
    0.101286507323456, 0.101286507323456,
    0.470142064105115, 0.059715871789770,
    0.059715871789770, 0.470142064105115,
    0.470142064105115, 0.470142064105115,
};
static const double kW_7[] = {
    0.225,
    0.125939180544827,
    0.125939180544827,
    0.125939180544827,
    0.132394152788506,
    0.132394152788506,
    0.132394152788506,
};
// qf3D9: 9 points, order 5
static const double kL12_9[] = {
    0.437525248383384, 0.124949503233232,
    0.124949503233232, 0.437525248383384,
    0.437525248383384, 0.437525248383384,
    0.797112651860071, 0.165409927389841,
    0.165409927389841, 0.797112651860071,
    0.797112651860071, 0.037477420750088,
    0.037477420750088, 0.797112651860071,
    0.165409927389841, 0.037477420750088,
    0.037477420750088, 0.165409927389841,
};
static const double kW_9[] = {
    0.205950504760887,
    0.205950504760887,
    0.205950504760887,
    0.063691414286223,
    0.063691414286223,
    0.063691414286223,
    0.063691414286223,
    0.063691414286223,
    0.063691414286223,
};
// qf3D12: 12 points, order 6
static const double kL12_12[] = {
    0.873821971016996, 0.063089014491502,
    0.063089014491502, 0.873821971016996,
    0.063089014491502, 0.063089014491502,
    0.501426509658179, 0.249286745170910,
    0.249286745170910, 0.501426509658179,
    0.249286745170910, 0.249286745170910,
    0.636502499121399, 0.310352451033785,
    0.310352451033785, 0.636502499121399,
    0.636502499121399, 0.053145049844816,
    0.053145049844816, 0.636502499121399,
    0.310352451033785, 0.053145049844816,
    0.053145049844816, 0.310352451033785,
};
static const double kW_12[] = {
    0.050844906370207,
    0.050844906370207,
    0.050844906370207,
    0.116786275726379,
    0.116786275726379,
    0.116786275726379,
    0.082851075618374,
    0.082851075618374,
    0.082851075618374,
    0.082851075618374,
    0.082851075618374,
    0.082851075618374,
};
// qf3D13: 13 points, order 7
static const double kL12_13[] = {
    0.333333333333333, 0.333333333333333,
    0.479308067841923, 0.260345966079038,
    0.260345966079038, 0.479308067841923,
    0.260345966079038, 0.260345966079038,
    0.869739794195598, 0.065130102902216,
    0.065130102902216, 0.869739794195598,
    0.065130102902216, 0.065130102902216,
    0.638444188569809, 0.312865496004875,
    0.312865496004875, 0.638444188569809,
    0.638444188569809, 0.048690315425316,
    0.048690315425316, 0.638444188569809,
    0.312865496004875, 0.048690315425316,
    0.048690315425316, 0.312865496004875,
};
static const double kW_13[] = {
    -0.149570044467670,
    0.175615257433204,
    0.175615257433204,
    0.175615257433204,
    0.053347235608839,
    0.053347235608839,
    0.053347235608839,
    0.077113760890257,
    0.077113760890257,
    0.077113760890257,
    0.077113760890257,
    0.077113760890257,
    0.077113760890257,
};

struct RuleRef {
    int n;
    const double* L12;
    const double* w;
    int order;
};
const RuleRef kRules[] = {
    {1, kL12_1, kW_1, 2},   {3, kL12_3, kW_3, 2},   {4, kL12_4, kW_4, 3},     {6, kL12_6, kW_6, 4},
    {7, kL12_7, kW_7, 5},   {9, kL12_9, kW_9, 5},   {12, kL12_12, kW_12, 6},  {13, kL12_13, kW_13, 7},
};

const char* kTypeNames[3] = {"SimpleNeighbors", "AttachedNeighbors", "NotNeighbors"};

bool tri_in_range(const int32_t* tri, int64_t nT, int64_t nV) {
    for (int64_t q = 0; q < 3 * nT; ++q)
        if (tri[q] < 0 || tri[q] >= nV) return false;
    return true;
}

}  // namespace

extern "C" {

int i2_quadrature_rule(int nPts, double* L12, double* w, int* order) {
    for (const RuleRef& r : kRules) {
        if (r.n != nPts) continue;
        if (L12) std::memcpy(L12, r.L12, sizeof(double) * 2 * r.n);
        if (w) std::memcpy(w, r.w, sizeof(double) * r.n);
        if (order) *order = r.order;
        return I2_OK;
    }
    i2::set_error("no triangle rule with %d points (1, 3, 4, 6, 7, 9, 12 or 13)", nPts);
    return I2_ERR_INVALID;
}

// exportMeshToObj (src/Mesh3d.cu:264-276): "v x y z" lines, then "f a b c"
// with 1-based indices; default stream formatting (6 significant digits).
int i2_export_mesh_obj(const char* path, const double* xyz, int64_t nV, const int32_t* tri, int64_t nT) {
    if (!path || nV < 0 || nT < 0 || (nV && !xyz) || (nT && !tri)) { i2::set_error("invalid argument"); return I2_ERR_INVALID; }
    if (!tri_in_range(tri, nT, nV)) { i2::set_error("cell references a vertex outside [0, %lld)", (long long)nV); return I2_ERR_INVALID; }
    std::ofstream f(path);
    if (!f.is_open()) { i2::set_error("cannot open %s for writing", path); return I2_ERR_IO; }
    for (int64_t v = 0; v < nV; ++v) f << "v " << xyz[3 * v] << " " << xyz[3 * v + 1] << " " << xyz[3 * v + 2] << "\n";
    for (int64_t c = 0; c < nT; ++c) f << "f " << tri[3 * c] + 1 << " " << tri[3 * c + 1] + 1 << " " << tri[3 * c + 2] + 1 << "\n";
    f.flush();
    if (!f.good()) { i2::set_error("write to %s failed", path); return I2_ERR_IO; }
    return I2_OK;
}

// exportMeshToVtk (src/Mesh3d.cu:278-343): VTK XML PolyData with ASCII
// points, 0-based connectivity and offsets, and one Int32 cell array
// "<type>Refinements" per non-null refinement-count array (nT bytes each).
int i2_export_mesh_vtk(const char* path, const double* xyz, int64_t nV, const int32_t* tri, int64_t nT,
                       const uint8_t* simpleRefinements, const uint8_t* attachedRefinements, const uint8_t* notRefinements) {
    if (!path || nV < 0 || nT < 0 || (nV && !xyz) || (nT && !tri)) { i2::set_error("invalid argument"); return I2_ERR_INVALID; }
    if (!tri_in_range(tri, nT, nV)) { i2::set_error("cell references a vertex outside [0, %lld)", (long long)nV); return I2_ERR_INVALID; }
    const uint8_t* ref[3] = {simpleRefinements, attachedRefinements, notRefinements};
    std::ofstream f(path);
    if (!f.is_open()) { i2::set_error("cannot open %s for writing", path); return I2_ERR_IO; }
    f << "<?xml version=\"1.0\" ?> \n";
    f << "<VTKFile type=\"PolyData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    f << "  <PolyData>\n";
    f << "    <Piece NumberOfPoints=\"" << nV << "\" NumberOfPolys=\"" << nT << "\">\n";
    f << "      <Points>\n";
    f << "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Format=\"ascii\">\n";
    f << "        ";
    for (int64_t v = 0; v < nV; ++v) f << xyz[3 * v] << " " << xyz[3 * v + 1] << " " << xyz[3 * v + 2] << " ";
    f << "\n";
    f << "        </DataArray>\n";
    f << "      </Points>\n";
    f << "      <Polys>\n";
    f << "        <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n";
    f << "          ";
    for (int64_t c = 0; c < 3 * nT; ++c) f << tri[c] << " ";
    f << "\n";
    f << "        </DataArray>\n";
    f << "        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n";
    f << "          ";
    for (int64_t c = 0; c < nT; ++c) f << 3 * (c + 1) << " ";
    f << "\n";
    f << "        </DataArray>\n";
    f << "      </Polys>\n";
    if (ref[0] || ref[1] || ref[2]) {
        f << "      <CellData>\n";
        for (int k = 0; k < 3; ++k) {
            if (!ref[k]) continue;
            f << "        <DataArray type=\"Int32\" Name=\"" << kTypeNames[k] << "Refinements\" Format=\"ascii\">\n";
            f << "          ";
            for (int64_t c = 0; c < nT; ++c) f << (int)ref[k][c] << " ";
            f << "\n";
            f << "        </DataArray>\n";
        }
        f << "      </CellData>\n";
    }
    f << "    </Piece>\n";
    f << "  </PolyData>\n";
    f << "</VTKFile>\n";
    f.flush();
    if (!f.good()) { i2::set_error("write to %s failed", path); return I2_ERR_IO; }
    return I2_OK;
}

}  // extern "C"
