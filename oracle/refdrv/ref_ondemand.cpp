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
// usage: ref_ondemand X Y Z in.raw out_prefix err gthr iters presimp [threads [asc1 dsc1 asc2 [dsc2]]]
// writes (int32 per vertex) out_prefix.certain, .integrated, .cleaned (when asc2);
//        .certain_dsc, .integrated_dsc, .cleaned_dsc (when dsc2: the needDSC2
//        branch, ondemandaccurate.cxx:913-991, after the needASC2 one);
//        (int8 per cell) out_prefix.restriction; out_prefix.grad.
//
// Built with -DREF_LINES (target ref_lines) it also runs the needASC1 /
// needDSC1 branches (ondemandaccurate.cxx:1075-1091): the digitizing
// streamline integrators of gi_numeric_streamline_integrator.h:526-1072 from
// the 2-saddles up and from the 1-saddles down, driven in the order of
// ReIntegrateUpFrom2Saddles / ReIntegrateDownFrom1Saddles
// (ondemandaccurate.cxx:254-500).  That build is compiled without
// optimisation: DigitizeSegmentInternal (gi_numeric_streamline_integrator.h:566-700)
// can flow off its end without returning a value, which an optimising C++
// compiler may treat as unreachable.
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
#ifdef REF_LINES
#include "gi_numeric_streamline_integrator.h"
#endif

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
typedef IndexCompareGreaterThan<GridFuncType> ComparerDSC;
typedef NumericIntegratorPathCompressingToTerminal<AdaptiveEulerAdvector3D<GridFuncType, 1>, GridFuncType> IntegratorTypeDSC;
typedef DiscreteGradientLabeling<MeshType> GradType;
typedef RegularGridMaxMinVertexLabeling3D<MeshType, GridFuncType> MaxVLType;
typedef MyRobinsNoalloc<MeshType, MaxVLType, GradType, 4, 6> RobinsType;
typedef TopologicalMaxVertexMeshFunction<MeshType, MaxVLType, GridFuncType, float> MeshFuncType;
typedef SlidingWindowRobinsNoalloc<GridType, GridFuncType, MeshType, MaxVLType, GradType> NewRobinsType;

#ifdef REF_LINES
typedef DigitizingNumericStreamlineIntegrator3dASC<MeshType, GridFuncType, AdaptiveEulerAdvector3D<GridFuncType, 1> >
    LinesASC;
typedef DigitizingNumericStreamlineIntegrator3dDSC<MeshType, GridFuncType, AdaptiveEulerAdvector3D<GridFuncType, -1> >
    LinesDSC;

// One thread's run of the saddle loop: saddles of dimension sdim by (key,
// id) from the largest, key = value (up) or -value (down); each saddle's
// seeds from a short breadth-first search over V-path steps; one streamline
// per seed cell of the target dimension; then the saddle itself is labelled.
template <class Integ, class NeighbourIt>
static long long lines_from_saddles(MeshType* mesh, MeshFuncType* topo_func, GradType* grad, Integ* integ, int sdim,
                                    bool up) {
    std::vector<std::pair<float, INDEX_TYPE> > sad;
    MeshType::DCellsIterator it(mesh, sdim);
    for (it.begin(); it.valid(); it.advance()) {
        const INDEX_TYPE c = it.value();
        if (grad->getCritical(c)) sad.push_back(std::make_pair(up ? topo_func->cellValue(c) : -1 * topo_func->cellValue(c), c));
    }
    std::sort(sad.begin(), sad.end());
    long long lines = 0;
    for (INDEX_TYPE k = (INDEX_TYPE)sad.size() - 1; k >= 0; k--) {
        const INDEX_TYPE s = sad[k].second;
        std::vector<INDEX_TYPE> result;
        std::queue<INDEX_TYPE> q;
        std::set<INDEX_TYPE> seen;
        q.push(s);
        for (int counter = 4; !q.empty() && counter >= 0; counter--) {
            const INDEX_TYPE cur = q.front();
            q.pop();
            seen.insert(cur);
            NeighbourIt nb(mesh);
            for (nb.begin(cur); nb.valid(); nb.advance()) {
                const INDEX_TYPE t = nb.value();
                if (seen.count(t)) continue;
                result.push_back(t);
                seen.insert(t);
                if (!grad->getCritical(t)) {
                    const INDEX_TYPE p = grad->getPair(t);
                    seen.insert(p);
                    q.push(p);
                }
            }
        }
        for (INDEX_TYPE t : result) {
            if (mesh->dimension(t) != (up ? 3 : 0)) continue;
            std::vector<Vec3d> points;
            std::vector<INDEX_TYPE> dline;
            Vec3d seed;
            mesh->centroid(t, seed);
            seed = seed * 0.5;
            integ->IntegrateStreamline(seed, points, dline);
            lines++;
        }
        integ->set_label(s);
    }
    return lines;
}
#endif

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
    const int need_asc1 = argc > 13 ? atoi(argv[11]) : 0, need_dsc1 = argc > 13 ? atoi(argv[12]) : 0;
    const int need_asc2 = argc > 13 ? atoi(argv[13]) : 1;
    const int need_dsc2 = argc > 14 ? atoi(argv[14]) : 0;
#ifndef REF_LINES
    if (need_asc1 || need_dsc1) {
        fprintf(stderr, "1-manifolds need the ref_lines build\n");
        return 2;
    }
#endif
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

    double t_integ = 0;
    SimplifiedExtremumGraph<MeshType, MeshFuncType, GradType>* ext_graph = nullptr;
    VertexLabelingToBoundaryLabeling<int, MaxVLType>* edge_map = nullptr;
    if (need_asc2 || need_dsc2) {
        // ondemandaccurate.cxx:806-827
        ext_graph = new SimplifiedExtremumGraph<MeshType, MeshFuncType, GradType>(mesh, topo_func, base_grad);
        typedef SimplifiedExtremumGraph<MeshType, MeshFuncType, GradType> EG;
        ext_graph->SetMode(need_asc2 && need_dsc2 ? EG::EXTGRAPHMODE::BOTH
                                                  : (need_asc2 ? EG::EXTGRAPHMODE::MINS : EG::EXTGRAPHMODE::MAXS));
        ext_graph->ComputeMinMapFromGradient(presimp);
        edge_map = new VertexLabelingToBoundaryLabeling<int, MaxVLType>(mesh, restriction);
        edge_map->InitializeFirst();
    }
    if (need_asc2) {

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
    t_integ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ti).count();
    dump_int(out + ".integrated", integrator->GetOutputLabels());

    IsolatedCCRegionRemoverNEW<ComparerASC, GridFuncType>* cleaner =
        new IsolatedCCRegionRemoverNEW<ComparerASC, GridFuncType>(func, integrator->GetOutputLabels());
    cleaner->ComputeOutput();
    dump_int(out + ".cleaned", integrator->GetOutputLabels());

    edge_map->ComputeMINBoundary(integrator->GetOutputLabels());
    }
    if (need_dsc2) {
        // ondemandaccurate.cxx:913-991
        std::unordered_map<INDEX_TYPE, INT_TYPE> extmap;
        for (auto p : ext_graph->mMaxGraph->mCellIndexToListIdMap)
            extmap[maxv->Cell2HighestVertex(p.first)] = ext_graph->mMaxGraph->Representative(p.second);
        GridSimplifiedExtremalRegionBuilder<ComparerDSC, GridFuncType, MeshType>* builder =
            new GridSimplifiedExtremalRegionBuilder<ComparerDSC, GridFuncType, MeshType>(func, grid, mesh);
        builder->BeginIntegration(extmap);
        dump_int(out + ".certain_dsc", builder->GetOutputLabels());
        auto ti = std::chrono::steady_clock::now();
        IntegratorTypeDSC* integrator = new IntegratorTypeDSC(func, grid, err, gthr, iters);
        integrator->BeginIntegration(builder->GetIdMap(), builder->GetOutputLabels(), true);
        t_integ += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ti).count();
        dump_int(out + ".integrated_dsc", integrator->GetOutputLabels());
        IsolatedCCRegionRemoverNEW<ComparerDSC, GridFuncType>* cleaner =
            new IsolatedCCRegionRemoverNEW<ComparerDSC, GridFuncType>(func, integrator->GetOutputLabels());
        cleaner->ComputeOutput();
        dump_int(out + ".cleaned_dsc", integrator->GetOutputLabels());
        edge_map->ComputeMAXBoundary(integrator->GetOutputLabels(), maxv);
    }
    long long lines_asc = 0, lines_dsc = 0;
#ifdef REF_LINES
    if (need_asc1) {
        LinesASC* la = new LinesASC(grid, func, mesh, err, gthr, iters);
        la->SetDigitizingTarget(restriction, 4);
        lines_asc = lines_from_saddles<LinesASC, MeshType::CofacetsIterator>(mesh, topo_func, base_grad, la, 2, true);
    }
    if (need_dsc1) {
        LinesDSC* ld = new LinesDSC(grid, func, mesh, err, gthr, iters);
        ld->SetDigitizingTarget(restriction, 5);
        lines_dsc = lines_from_saddles<LinesDSC, MeshType::FacetsIterator>(mesh, topo_func, base_grad, ld, 1, false);
    }
#endif
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
    printf("REFTIME threads=%d integrate_ms=%.3f total_ms=%.3f fixed=%lld lines_asc=%lld lines_dsc=%lld\n", threads,
           t_integ, t_all, fixed, lines_asc, lines_dsc);
    return 0;
}
