// integrator2test3D -- command-line driver of the MI355X integrator with the
// options of the reference's driver (tests/integrator3D/main.cu):
//   -f, --meshfile=<file>   input mesh (.dat: counts, vertices, type-203 cells)
//   -s, --scale=<x>         mesh scale factor
//   -r, --refine=<n>        fixed refinement level n (default: adaptive error control)
//   -c, --checkresults      (i, j) / (j, i) symmetry check
//   --exportresults         results to SimpleNeighbors.dat etc.
//   --exporttocsv           results to SimpleNeighbors.csv etc.
//   --exporttoobj           original (and refined) mesh to OBJ
//   --exporttovtk           original (and refined) mesh to VTK, with the
//                           per-cell refinement counts of the adaptive run
//   -h, --help
// Added: --maxlevel=<n> and --tolerance=<x> for the adaptive run (compile-time
// constants in the reference) and --batched (the persistent batched kernels).
// Written against the reference's class API (include/integrator2/integrator2.hpp).
#include <integrator2/integrator2.hpp>

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

namespace {

struct Options {
    bool help = false;
    std::string meshfile;
    double scale = 1.0;
    bool refine = false;
    int refineLevel = -1;
    bool check = false, exportResults = false, exportCsv = false, exportObj = false, exportVtk = false;
    bool batched = false;
    int maxLevel = -1;
    double tolerance = -1.0;
};

void usage() {
    std::cout << "USAGE: integrator2test3D [options]\n\n"
                 "Options:\n"
                 "  -h, --help                Print usage and exit.\n"
                 "  -f <arg>, --meshfile=<arg> Input mesh file name.\n"
                 "  -s <arg>, --scale=<arg>   Mesh scale factor.\n"
                 "  --exporttoobj             Export original and refined meshes to OBJ files.\n"
                 "  --exporttovtk             Export original and refined meshes to VTK (VTP) files.\n"
                 "  --exporttocsv             Export results of integration to csv files.\n"
                 "  --exportresults           Export results of integration to text files.\n"
                 "  -r <arg>, --refine=<arg>  Refine the whole mesh N times.\n"
                 "  -c, --checkresults        Check correctness of pairs of results.\n"
                 "  --maxlevel=<arg>          Maximum refinement level of the adaptive procedure.\n"
                 "  --tolerance=<arg>         Runge tolerance of the adaptive procedure.\n"
                 "  --batched                 Use the batched persistent kernels.\n";
}

bool numeric(const char* s, double* out) {
    if (!s || !*s) return false;
    char* end = nullptr;
    *out = std::strtod(s, &end);
    return end != s && *end == 0;
}

// Returns false (after a message) on a malformed command line.
bool parse(int argc, char** argv, Options* o) {
    for (int a = 1; a < argc; ++a) {
        std::string arg = argv[a];
        std::string name, value;
        bool has_value = false;
        if (arg.rfind("--", 0) == 0) {
            const size_t eq = arg.find('=');
            name = arg.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
            if (eq != std::string::npos) { value = arg.substr(eq + 1); has_value = true; }
        } else if (arg.size() == 2 && arg[0] == '-') {
            name = arg.substr(1);
        } else if (arg.size() > 2 && arg[0] == '-') {
            name = arg.substr(1, 1);
            value = arg.substr(2);
            has_value = true;
        } else {
            std::fprintf(stderr, "Unknown option '%s'\n", arg.c_str());
            return false;
        }
        auto take = [&](const char* what) -> bool {
            if (has_value) return true;
            if (a + 1 < argc) { value = argv[++a]; has_value = true; return true; }
            std::fprintf(stderr, "Option '%s' requires %s argument\n", name.c_str(), what);
            return false;
        };
        double x = 0.0;
        if (name == "h" || name == "help") o->help = true;
        else if (name == "f" || name == "meshfile") {
            if (!take("a non-empty") || value.empty()) { std::fprintf(stderr, "Option '%s' requires a non-empty argument\n", name.c_str()); return false; }
            o->meshfile = value;
        } else if (name == "s" || name == "scale") {
            if (!take("a numeric") || !numeric(value.c_str(), &x)) { std::fprintf(stderr, "Option '%s' requires a numeric argument\n", name.c_str()); return false; }
            o->scale = x;
        } else if (name == "r" || name == "refine") {
            if (!take("a numeric") || !numeric(value.c_str(), &x)) { std::fprintf(stderr, "Option '%s' requires a numeric argument\n", name.c_str()); return false; }
            o->refine = true;
            o->refineLevel = static_cast<int>(x);
        } else if (name == "maxlevel") {
            if (!take("a numeric") || !numeric(value.c_str(), &x)) { std::fprintf(stderr, "Option '%s' requires a numeric argument\n", name.c_str()); return false; }
            o->maxLevel = static_cast<int>(x);
        } else if (name == "tolerance") {
            if (!take("a numeric") || !numeric(value.c_str(), &x)) { std::fprintf(stderr, "Option '%s' requires a numeric argument\n", name.c_str()); return false; }
            o->tolerance = x;
        } else if (name == "c" || name == "checkresults") o->check = true;
        else if (name == "exportresults") o->exportResults = true;
        else if (name == "exporttocsv") o->exportCsv = true;
        else if (name == "exporttoobj") o->exportObj = true;
        else if (name == "exporttovtk") o->exportVtk = true;
        else if (name == "batched") o->batched = true;
        else {
            std::fprintf(stderr, "Unknown option '%s'\n", arg.c_str());
            return false;
        }
    }
    return true;
}

template <class T>
std::vector<T> to_host(const deviceVector<T>& v) {
    std::vector<T> h(static_cast<size_t>(v.size));
    if (v.size) copy_d2h(v.data, h.data(), h.size());
    return h;
}

int run(const Options& opt) {
    Mesh3D mesh;
    if (!mesh.loadMeshFromFile(opt.meshfile, opt.scale)) return EXIT_FAILURE;
    mesh.prepareMesh();

    NumericalIntegrator3D numIntegrator(mesh, qf3D13);
    EvaluatorJ3DK evaluator(mesh, numIntegrator);
    if (opt.batched) evaluator.setPath(1);

    if (opt.refine) {
        numIntegrator.setFixedRefinementLevel(opt.refineLevel);
        if (opt.refineLevel)
            std::printf("Using fixed refinement level equal to %d\n", opt.refineLevel);
        else
            std::printf("Using original mesh without refinement\n");
    } else {
        if (opt.maxLevel >= 0 || opt.tolerance > 0.0)
            numIntegrator.setErrorControl(opt.maxLevel >= 0 ? opt.maxLevel : 5, opt.tolerance > 0.0 ? opt.tolerance : 1e-5);
        std::printf("Using adaptive error control procedure\n");
    }

    std::vector<Point3> vertices;
    std::vector<int3> cells;
    if (opt.exportObj || opt.exportVtk) {
        vertices = to_host(mesh.getVertices());
        cells = to_host(mesh.getCells());
        if (opt.exportObj) exportMeshToObj("OriginalMesh.obj", vertices, cells);
        if (opt.exportVtk && opt.refine) exportMeshToVtk("OriginalMesh.vtp", vertices, cells, {});
    }

    evaluator.runAllPairs(opt.check);

    if (opt.exportResults || opt.exportCsv) {
        const output_format_enum fmt = opt.exportCsv ? output_format_enum::csv : output_format_enum::plainText;
        evaluator.outputResultsToFile(neighbour_type_enum::simple_neighbors, fmt);
        evaluator.outputResultsToFile(neighbour_type_enum::attached_neighbors, fmt);
        evaluator.outputResultsToFile(neighbour_type_enum::not_neighbors, fmt);
    }

    if ((opt.exportObj || opt.exportVtk) && opt.refine && opt.refineLevel > 0) {
        const std::vector<Point3> rv = to_host(numIntegrator.getRefinedVertices());
        const std::vector<int3> rc = to_host(numIntegrator.getRefinedCells());
        if (opt.exportObj) exportMeshToObj("RefinedMesh.obj", rv, rc);
        if (opt.exportVtk) exportMeshToVtk("RefinedMesh.vtp", rv, rc, {});
    }

    if (opt.exportVtk && !opt.refine) {
        std::array<std::vector<unsigned char>, 3> refinements;
        for (int k = 0; k < 3; ++k) {
            const deviceVector<unsigned char>* r = numIntegrator.getRefinementsRequired(neighbour_type_enum(k));
            if (r->size) refinements[k] = to_host(*r);
        }
        exportMeshToVtk("OriginalMesh.vtp", vertices, cells, refinements);
    }
    return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char** argv) {
    Options opt;
    if (!parse(argc, argv, &opt)) return EXIT_FAILURE;
    if (opt.help || argc == 1) {
        usage();
        return EXIT_SUCCESS;
    }
    if (opt.meshfile.empty()) {
        std::printf("No input file with mesh specified. Exiting\n");
        return EXIT_FAILURE;
    }
    try {
        return run(opt);
    } catch (const integrator2_error& e) {
        std::fprintf(stderr, "integrator2: %s\n", e.what());
        return EXIT_FAILURE;
    }
}
