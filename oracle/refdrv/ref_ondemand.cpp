// Test-infrastructure driver: runs the reference's accurate ascending
// 3-manifold labelling (the needASC2 branch of ondemandaccurate), compiled from
// the headers where they lie, to produce golden vectors.  Never ships.
//
// Call sequence follows ondemandaccurate/ondemandaccurate.cxx:728-1061 with
// need_ASC_3 = 1 and the other accuracy flags 0:
//   max/min labelling, sliding lower-star pairing (base gradient),
//   ComputeGradFromImage(1), SimplifiedExtremumGraph (MINS) +
//   SimplifyToThreshold, GridSimplifiedExtremalRegionBuilder,
//   NumericIntegratorPathCompressingToTerminal<AdaptiveEulerAdvector3D<-1>>,
//   IsolatedCCRegionRemoverNEW, VertexLabelingToBoundaryLabeling::ComputeMINBoundary,
//   constrained MyRobinsNoalloc<...,4,6> on affected vertices,
//   setAscendingManifoldDimensions + outputToFile.
// The same RefMesh subclass as ref_steepest.cpp is used for the dead loop at
// include/gi_modified_robins.h:1005-1013.
//
// usage: ref_ondemand X Y Z in.raw out_prefix err gthr iters presimp [threads]
// writes (int32 per vertex) out_prefix.certain, .integrated, .cleaned;
//        (int8 per cell) out_prefix.restriction; out_prefix.grad.
#include <chrono>
#include <cstdio>
#include <queue>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "gi_numeric_integrator_expanding_region_stop.h"
#include "gi_labeling_to_bounary_labeling.h"
#include "gi_topological_gradient_using_algorithms.h"
#include "gi_isolated_region_remover.h"
#include "gi_max_vertex_labeling.h"
#include "gi_topological_max_vertex_mesh_function.h"
#include "gi_modified_robins.h"
#include "gi_fast_robins_noalloc.h"
#include "gi_extrema_region_builder.h"
#include "gi_numeric_integrator_path_compressing.h"
#include "gi_adaptive_euler_advector.h"

using namespace GInt;

struct RefMesh : public TopologicalRegularGrid3D {
    using TopologicalRegularGrid3D::TopologicalRegularGrid3D;
    using TopologicalRegularGrid3D::cellid2Coords;
    void cellid2Coords(INDEX_TYPE, Vec2l& v) const { v = Vec2l(0, 0); }
};

typedef RegularGrid3D GridType;
typedef RegularGridTrilinearFunction GridFuncType;
typedef RefMesh MeshType;
typedef IndexCompareLessThan<GridFuncType> ComparerASC;
typedef NumericIntegratorPathCompressingToTerminal<AdaptiveEulerAdvector3D<GridFuncType, -1>, GridFuncType> IntegratorTypeASC;
typedef DiscreteGradientLabeling<MeshType> GradType;
typedef RegularGridMaxMinVertexLabeling3D<MeshType, GridFuncType> MaxVLType;
typedef MyRobinsNoalloc<MeshType, MaxVLType, GradType, 4, 6> RobinsType;
typedef TopologicalMaxVertexMeshFunction<MeshType, MaxVLType, GridFuncType, float> MeshFuncType;
typedef SlidingWindowRobinsNoalloc<GridType, GridFuncType, MeshType, MaxVLType, GradType> NewRobinsType;

static void dump_int(const std::string& path, const DenseLabeling<int>* l) { l->OutputToIntFile(path.c_str()); }

int main(int argc, char** argv) {
    if (argc < 10) {
        fprintf(stderr, "usage: %s X Y Z in.raw out_prefix err gthr iters presimp [threads]\n", argv[0]);
        return 2;
    }
    int X = atoi(argv[1]), Y = atoi(argv[2]), Z = atoi(argv[3]);
    std::string in = argv[4], out = argv[5];
    float err = (float)atof(argv[6]), gthr = (float)atof(argv[7]);
    int iters = atoi(argv[8]);
    float presimp = (float)atof(argv[9]);
    int threads = argc > 10 ? atoi(argv[10]) : 1;
    omp_set_num_threads(threads);

    GridType* grid = new GridType(Vec3l((long long)X, (long long)Y, (long long)Z), Vec3b(0, 0, 0));
    GridFuncType* func = new GridFuncType(grid);
    func->LoadImageFromFile(in.c_str());
    auto t0 = std::chrono::steady_clock::now();
    MeshType* mesh = new MeshType(grid);
    MaxVLType* maxv = new MaxVLType(mesh, func);
    maxv->ComputeOutput();
    MeshFuncType* topo_func = new MeshFuncType();
    topo_func->setMeshAndFuncAndMaxVLabeling(mesh, func, maxv);

    GradType* base_grad = new GradType(mesh);
    base_grad->ClearAllGradient();
    NewRobinsType* first_robins = new NewRobinsType(grid, func, mesh, maxv, base_grad);
    first_robins->ComputePairing_sliding();

    func->ComputeGradFromImage(1);
    DenseLabeling<char>* restriction = new DenseLabeling<char>(mesh->numCells());
    restriction->SetAll(0);
    RobinsType* constrained = new RobinsType(mesh, maxv, restriction, base_grad);

    SimplifiedExtremumGraph<MeshType, MeshFuncType, GradType>* ext_graph =
        new SimplifiedExtremumGraph<MeshType, MeshFuncType, GradType>(mesh, topo_func, base_grad);
    ext_graph->SetMode(SimplifiedExtremumGraph<MeshType, MeshFuncType, GradType>::EXTGRAPHMODE::MINS);
    ext_graph->ComputeMinMapFromGradient(presimp);

    VertexLabelingToBoundaryLabeling<int, MaxVLType>* edge_map =
        new VertexLabelingToBoundaryLabeling<int, MaxVLType>(mesh, restriction);
    edge_map->InitializeFirst();

    std::unordered_map<INDEX_TYPE, INT_TYPE> extmap;
    for (auto p : ext_graph->mMinGraph->mCellIndexToListIdMap)
        extmap[p.first] = ext_graph->mMinGraph->Representative(p.second);
    GridSimplifiedExtremalRegionBuilder<ComparerASC, GridFuncType, MeshType>* builder =
        new GridSimplifiedExtremalRegionBuilder<ComparerASC, GridFuncType, MeshType>(func, grid, mesh);
    builder->BeginIntegration(extmap);
    dump_int(out + ".certain", builder->GetOutputLabels());

    auto ti = std::chrono::steady_clock::now();
    IntegratorTypeASC* integrator = new IntegratorTypeASC(func, grid, err, gthr, iters);
    integrator->BeginIntegration(builder->GetIdMap(), builder->GetOutputLabels(), true);
    double t_integ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ti).count();
    dump_int(out + ".integrated", integrator->GetOutputLabels());

    IsolatedCCRegionRemoverNEW<ComparerASC, GridFuncType>* cleaner =
        new IsolatedCCRegionRemoverNEW<ComparerASC, GridFuncType>(func, integrator->GetOutputLabels());
    cleaner->ComputeOutput();
    dump_int(out + ".cleaned", integrator->GetOutputLabels());

    edge_map->ComputeMINBoundary(integrator->GetOutputLabels());
    restriction->OutputToFile((out + ".restriction").c_str());

    // conforming gradient (ondemandaccurate.cxx:1014-1052)
    long long fixed = 0;
    MeshType::DCellsIterator verts(mesh, 0);
    for (verts.begin(); verts.valid(); verts.advance()) {
        INDEX_TYPE v = verts.value();
        bool hasdiff = restriction->GetLabel(v) > 0;
        if (!hasdiff) {
            MeshType::AdjacentCellsIterator it(mesh);
            for (it.begin(v); it.valid(); it.advance()) {
                INDEX_TYPE c = it.value();
                if (restriction->GetLabel(c) > 0 && maxv->Cell2HighestVertex(c) == v) { hasdiff = true; break; }
            }
        }
        if (hasdiff) { constrained->ComputeLowerStar(v); fixed++; }
    }
    TopologicalGradientUsingAlgorithms<MeshType, MeshFuncType, GradType>* algs =
        new TopologicalGradientUsingAlgorithms<MeshType, MeshFuncType, GradType>(topo_func, mesh, base_grad);
    algs->setAscendingManifoldDimensions();
    base_grad->outputToFile((out + ".grad").c_str());
    double t_all = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    printf("REFTIME threads=%d integrate_ms=%.3f total_ms=%.3f fixed=%lld\n", threads, t_integ, t_all, fixed);
    return 0;
}
