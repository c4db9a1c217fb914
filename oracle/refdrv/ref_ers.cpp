// Test-infrastructure driver: runs the reference's own
// NumericIntegratorExpandingRegionStop (no program of the reference
// instantiates it; this driver does, with the advector and comparer the
// ascending-manifold programs use), compiled from the headers where they lie,
// to write golden vectors.  It never ships and never travels to a GPU
// machine.
//
// Call sequence: load the values, ComputeGradFromImage(1) (as
// ondemandaccurate/ondemandaccurate.cxx:795 does), then
// NumericIntegratorExpandingRegionStop<AdaptiveEulerAdvector3D<F, -1>,
// IndexCompareLessThan<F>>(func, grid, err, gthr, iters).BeginIntegration()
// (gi_numeric_integrator_expanding_region_stop.h:204-440) with one OpenMP
// thread, whose pass order is the vertex order.
//
// usage: ref_ers X Y Z in.raw err gthr iters out_prefix [threads]
//   (threads: 1 by default; more show how much the reference's own result
//   depends on its OpenMP schedule) writes out_prefix.labels (int64 destination per vertex: the vertex id of
//   its minimum, or -1) and prints the number of extrema.
#include <cstdio>
#include <string>
#include <vector>

#include "gi_topological_gradient_using_algorithms.h"
#include "gi_numeric_integrator_expanding_region_stop.h"

using namespace GInt;

typedef RegularGridTrilinearFunction GridFuncType;
typedef NumericIntegratorExpandingRegionStop<AdaptiveEulerAdvector3D<GridFuncType, -1>, IndexCompareLessThan<GridFuncType> >
    IntegratorType;

int main(int argc, char** argv) {
    if (argc < 9) {
        fprintf(stderr, "usage: %s X Y Z in.raw err gthr iters out_prefix\n", argv[0]);
        return 2;
    }
    const int X = atoi(argv[1]), Y = atoi(argv[2]), Z = atoi(argv[3]);
    const float err = (float)atof(argv[5]), gthr = (float)atof(argv[6]);
    const int iters = atoi(argv[7]);
    const std::string out = argv[8];
    omp_set_num_threads(argc > 9 ? atoi(argv[9]) : 1);
    RegularGrid3D* grid = new RegularGrid3D(Vec3i(X, Y, Z), Vec3b(0, 0, 0));
    GridFuncType* func = new GridFuncType(grid);
    func->LoadImageFromFile(argv[4]);
    func->ComputeGradFromImage(1);
    IntegratorType* integ = new IntegratorType(func, grid, err, gthr, iters);
    integ->BeginIntegration();
    DenseLabeling<INDEX_TYPE>* lab = integ->GetOutputLabels();
    const INDEX_TYPE V = (INDEX_TYPE)X * Y * Z;
    std::vector<int64_t> labels(V);
    for (INDEX_TYPE i = 0; i < V; i++) labels[i] = (int64_t)lab->GetLabel(i);
    FILE* fo = fopen((out + ".labels").c_str(), "wb");
    fwrite(labels.data(), sizeof(int64_t), labels.size(), fo);
    fclose(fo);
    printf("EXTREMA %zu\n", integ->GetExtrema().size());
    return 0;
}
