// Test-infrastructure driver: runs the reference's own Morse-Smale complex
// class on a .grad, compiled from the headers where they lie, to write golden
// vectors of the .grad consumer.  It never ships and never travels to a GPU
// machine.
//
// Call sequence follows extractmsc/extractmsc.cxx:467-552 (load the values
// and the .grad, LazyMaximumVertexLabeling, MorseSmaleComplexBasic with the
// same arc settings, ComputeFromGrad, then the base basins and mountains by
// fillGeometry of every minimum and maximum).  extractmsc itself does not
// build here (it needs Eigen and a SurfaceStatistics type the tree lacks).
//
// usage: ref_msc X Y Z in.raw in.grad out_prefix
//   writes out_prefix.nodes (int64 cell id and int64 dimension per node, in
//   node order), out_prefix.basins and out_prefix.mounts (int32 per vertex:
//   the node id of the minimum whose ascending manifold holds the vertex, and
//   of the maximum whose descending manifold holds the hex at the vertex, -1
//   where none).
#include <cstdio>
#include <set>
#include <string>
#include <vector>

#include "gi_topological_gradient_using_algorithms.h"
#include "gi_numeric_integrator_expanding_region_stop.h"
#include "gi_max_vertex_labeling.h"
#include "gi_topological_max_vertex_mesh_function.h"
#include "gi_morse_smale_complex_basic.h"

using namespace GInt;

typedef RegularGrid3D GridType;
typedef TopologicalRegularGrid3D MeshType;
typedef RegularGridTrilinearFunction GridFuncType;
typedef DiscreteGradientLabeling<MeshType> GradType;
typedef LazyMaximumVertexLabeling<MeshType, GridFuncType> MaxVLType;
typedef TopologicalMaxVertexMeshFunction<MeshType, MaxVLType, GridFuncType, float> TopoFuncType;
typedef MorseSmaleComplexBasic<float, MeshType, TopoFuncType, GradType> MscType;

int main(int argc, char** argv) {
    if (argc < 7) {
        fprintf(stderr, "usage: %s X Y Z in.raw in.grad out_prefix\n", argv[0]);
        return 2;
    }
    const int X = atoi(argv[1]), Y = atoi(argv[2]), Z = atoi(argv[3]);
    const std::string out = argv[6];
    GridType* grid = new GridType(Vec3i(X, Y, Z), Vec3b(0, 0, 0));
    GridFuncType* func = new GridFuncType(grid);
    func->LoadImageFromFile(argv[4]);
    MeshType* mesh = new MeshType(grid);
    MaxVLType* maxv = new MaxVLType(mesh, func);
    maxv->ComputeOutput();
    TopoFuncType* topo = new TopoFuncType();
    topo->setMeshAndFuncAndMaxVLabeling(mesh, func, maxv);
    GradType* grad = new GradType(mesh);
    grad->load_from_file(argv[5]);

    MscType* msc = new MscType(grad, mesh, topo);
    msc->SetBuildArcGeometry(Vec3b(false, false, false));
    msc->SetBuildArcs(Vec3b(true, false, true));
    msc->ComputeFromGrad();

    const INDEX_TYPE V = (INDEX_TYPE)X * Y * Z;
    std::vector<int32_t> basins(V, -1), mounts(V, -1);
    std::vector<int64_t> nodes;
    MscType::LivingNodesIterator it(msc);
    for (it.begin(); it.valid(); it.advance()) {
        const int n = it.value();
        const int dim = msc->getNode(n).dim;
        if ((int)nodes.size() < 2 * (n + 1)) nodes.resize(2 * (n + 1), -1);
        nodes[2 * n] = (int64_t)msc->getNode(n).cellindex;
        nodes[2 * n + 1] = dim;
        if (dim != 0 && dim != 3) continue;
        std::set<INDEX_TYPE> man;
        msc->fillGeometry(n, man, dim == 0);
        for (INDEX_TYPE id : man)
            if (mesh->dimension(id) == dim) (dim == 0 ? basins : mounts)[mesh->VertexNumberFromCellID(id)] = n;
    }
    FILE* f = fopen((out + ".nodes").c_str(), "wb");
    fwrite(nodes.data(), sizeof(int64_t), nodes.size(), f);
    fclose(f);
    f = fopen((out + ".basins").c_str(), "wb");
    fwrite(basins.data(), sizeof(int32_t), basins.size(), f);
    fclose(f);
    f = fopen((out + ".mounts").c_str(), "wb");
    fwrite(mounts.data(), sizeof(int32_t), mounts.size(), f);
    fclose(f);
    printf("nodes %zu\n", nodes.size() / 2);
    return 0;
}
