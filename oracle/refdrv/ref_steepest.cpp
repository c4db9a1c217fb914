// Test-infrastructure driver: runs the reference's own lower-star gradient
// classes, compiled from the headers where they lie, to produce golden
// vectors.  It never ships and never travels to a GPU machine.
//
// Call sequence follows steepest_lstar/steepest_lstar.cxx:89-258
// (load -> RegularGridMaxMinVertexLabeling3D::ComputeOutput ->
//  SlidingWindowRobinsNoalloc::ComputePairing_sliding ->
//  TopologicalGradientUsingAlgorithms::setAscendingManifoldDimensions ->
//  DiscreteGradientLabeling::outputToFile).
//
// Build note: include/gi_modified_robins.h:1005-1013 holds an unused loop that
// passes a Vec2l to TopologicalRegularGrid3D::cellid2Coords(INDEX_TYPE, Vec3l&),
// which GCC rejects when MyRobinsNoalloc::HomotopyExpand is instantiated.  The
// templates are therefore instantiated with RefMesh, a subclass that adds a
// no-op overload for exactly that call; the loop's result is never read, so
// the computation is unchanged.  No reference source is modified or copied.
//
// usage: ref_steepest X Y Z in.raw out_prefix [threads] [dump_stages]
//   writes out_prefix.grad (the reference .grad bytes) and, with
//   dump_stages=1, out_prefix.maxv / .minv (27-tid byte per cell) and
//   out_prefix.pairs (the gradient bytes before the ldir pass and flag pass).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gi_topological_gradient_using_algorithms.h"
#include "gi_numeric_integrator_expanding_region_stop.h"
#include "gi_max_vertex_labeling.h"
#include "gi_topological_max_vertex_mesh_function.h"
#include "gi_modified_robins.h"
#include "gi_fast_robins_noalloc.h"

using namespace GInt;

struct RefMesh : public TopologicalRegularGrid3D {
    using TopologicalRegularGrid3D::TopologicalRegularGrid3D;
    using TopologicalRegularGrid3D::cellid2Coords;
    // Only the dead loop at gi_modified_robins.h:1005-1013 calls this.
    void cellid2Coords(INDEX_TYPE, Vec2l& v) const { v = Vec2l(0, 0); }
};

typedef RegularGrid3D GridType;
typedef RegularGridTrilinearFunction GridFuncType;
typedef RefMesh MeshType;
typedef DiscreteGradientLabeling<MeshType> GradType;
typedef RegularGridMaxMinVertexLabeling3D<MeshType, GridFuncType> MaxVLType;
typedef TopologicalMaxVertexMeshFunction<MeshType, MaxVLType, GridFuncType, float> TopoFuncType;
typedef SlidingWindowRobinsNoalloc<GridType, GridFuncType, MeshType, MaxVLType, GradType> NewRobinsType;

static double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int main(int argc, char** argv) {
    if (argc < 6) {
        fprintf(stderr, "usage: %s X Y Z in.raw out_prefix [threads] [dump_stages]\n", argv[0]);
        return 2;
    }
    int X = atoi(argv[1]), Y = atoi(argv[2]), Z = atoi(argv[3]);
    std::string in = argv[4], out = argv[5];
    int threads = argc > 6 ? atoi(argv[6]) : omp_get_max_threads();
    int dump = argc > 7 ? atoi(argv[7]) : 0;
    omp_set_num_threads(threads);

    GridType* grid = new GridType(Vec3l((long long)X, (long long)Y, (long long)Z), Vec3b(0, 0, 0));
    GridFuncType* func = new GridFuncType(grid);
    func->LoadImageFromFile(in.c_str());
    MeshType* mesh = new MeshType(grid);

    auto t0 = std::chrono::steady_clock::now();
    MaxVLType* maxv = new MaxVLType(mesh, func);
    maxv->ComputeOutput();
    double t_maxv = ms_since(t0);
    TopoFuncType* topofunc = new TopoFuncType();
    topofunc->setMeshAndFuncAndMaxVLabeling(mesh, func, maxv);

    GradType* grad = new GradType(mesh);
    grad->ClearAllGradient();
    auto t1 = std::chrono::steady_clock::now();
    NewRobinsType* robins = new NewRobinsType(grid, func, mesh, maxv, grad);
    robins->ComputePairing_sliding();
    double t_pair = ms_since(t1);

    const INDEX_TYPE C = mesh->numCells();
    if (dump) {
        std::vector<unsigned char> mx(C), mn(C);
        for (INDEX_TYPE c = 0; c < C; c++) {
            mx[c] = maxv->GetUncompressedMaxVal(c);
            mn[c] = maxv->GetUncompressedMinVal(c);
        }
        FILE* f = fopen((out + ".maxv").c_str(), "wb"); fwrite(mx.data(), 1, C, f); fclose(f);
        f = fopen((out + ".minv").c_str(), "wb"); fwrite(mn.data(), 1, C, f); fclose(f);
        f = fopen((out + ".pairs").c_str(), "wb");
        fwrite(grad->m_dgrad->LabelArray(), 1, C, f);
        fclose(f);
    }

    auto t2 = std::chrono::steady_clock::now();
    TopologicalGradientUsingAlgorithms<MeshType, TopoFuncType, GradType>* algs =
        new TopologicalGradientUsingAlgorithms<MeshType, TopoFuncType, GradType>(topofunc, mesh, grad);
    algs->setAscendingManifoldDimensions();
    double t_ldir = ms_since(t2);

    grad->outputToFile((out + ".grad").c_str());
    printf("REFTIME threads=%d maxv_ms=%.3f pair_ms=%.3f ldir_ms=%.3f total_ms=%.3f\n",
           threads, t_maxv, t_pair, t_ldir, t_maxv + t_pair + t_ldir);
    return 0;
}
