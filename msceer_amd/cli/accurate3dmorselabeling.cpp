// Accurate3DMorseLabeling X Y Z filename error_threshold grad_threshold
//                         maxnumiter PresimpThesh AscOrDsc(0,1) [parallelism] [outputdebug]
//
// Drop-in for accurate3dmorselabeling/accurate3dmorselabeling.cxx:82-343:
// accurate ascending 3-manifold labels of the volume (of its negation with
// AscOrDsc = 1, as the reference negates the function, :180-181), written as
// <filename>._int_XxYxZ.raw (int32 per vertex, :333-335), and their region
// boundary kinds as <filename>._edges_MxMyMz.raw (int8 per cell, :337-343),
// with the timing file <filename>.<parallelism:%03d>.gtime.txt.  Region ids
// are the vertex id of each region's lowest minimum (the reference numbers
// them in its hash map's iteration order).  Returns 0 on success, as the
// reference's main does.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "msceer.h"

int main(int argc, char** argv) {
    if (argc < 11) {
        printf("Usage: X Y Z filename error_threshold grad_threshold maxnumiter PresimpThesh AscOrDsc(0,1) "
               "[parallelism=ompmaxnumthreads] [outputdebug=0] [integrationinteraltimer=0]\n");
        return 0;
    }
    const int X = atoi(argv[1]), Y = atoi(argv[2]), Z = atoi(argv[3]);
    const std::string filename = argv[4];
    const float err = (float)atof(argv[5]), gthr = (float)atof(argv[6]);
    const int iters = atoi(argv[7]);
    const float presimp = (float)atof(argv[8]);
    const int asc_or_dsc = atoi(argv[9]);
    const int par = atoi(argv[10]);
    printf("dims=(%d,%d,%d)\nfile=%s\nintegration parameters: e=%f, gt=%f, il=%d, ps=%f, ad=%d\npar=%d\n", X, Y, Z,
           filename.c_str(), err, gthr, iters, presimp, asc_or_dsc, par);
    const size_t nv = (size_t)X * Y * Z;
    const size_t nc = (size_t)(2 * X - 1) * (2 * Y - 1) * (2 * Z - 1);
    std::vector<float> values(nv);
    FILE* fin = fopen(filename.c_str(), "rb");
    if (!fin || fread(values.data(), sizeof(float), nv, fin) != nv) {
        fprintf(stderr, "could not read %zu floats from %s\n", nv, filename.c_str());
        if (fin) fclose(fin);
        return 2;
    }
    fclose(fin);
    if (asc_or_dsc == 1)
        for (float& v : values) v *= -1;
    msceer_ctx* ctx = nullptr;
    int rc = msceer_create(&ctx, 0);
    if (rc) { fprintf(stderr, "msceer_create: %s\n", msceer_strerror(rc)); return 2; }
    std::vector<int> labels(nv);
    std::vector<signed char> kinds(nc);
    msceer_label_stats st{};
    rc = msceer_ondemand_asc(ctx, values.data(), X, Y, Z, err, gthr, iters, presimp, nullptr, nullptr, labels.data(),
                             &st, MSCEER_TIMING);
    if (!rc) rc = msceer_region_boundary_kind(ctx, labels.data(), X, Y, Z, reinterpret_cast<int8_t*>(kinds.data()), 0);
    msceer_destroy(ctx);
    if (rc) { fprintf(stderr, "msceer: %s\n", msceer_strerror(rc)); return 2; }
    char name[2048];
    snprintf(name, sizeof(name), "%s.%03d.gtime.txt", filename.c_str(), par);
    if (FILE* ft = fopen(name, "w")) {
        // task lines "name start end duration" in ms, as StartTask/EndTaskAndRecord write them
        const float parts[4] = {st.ms_lstar, st.ms_regions, st.ms_integrate, st.ms_cleanup};
        const char* names[4] = {"Topological SlidingRobins", "Topological SimExtRBASC",
                                "NumericalTracing NewIntegrationASC", "Removed CleanerASC"};
        float t = 0.f;
        for (int i = 0; i < 4; i++) {
            fprintf(ft, "%s %d %d %d\n", names[i], (int)t, (int)(t + parts[i]), (int)parts[i]);
            t += parts[i];
        }
        fclose(ft);
    }
    snprintf(name, sizeof(name), "%s._int_%dx%dx%d.raw", filename.c_str(), X, Y, Z);
    printf("outputting labels:\n%s\n", name);
    FILE* fl = fopen(name, "wb");
    if (!fl || fwrite(labels.data(), sizeof(int), nv, fl) != nv) {
        fprintf(stderr, "could not write %s\n", name);
        if (fl) fclose(fl);
        return 2;
    }
    fclose(fl);
    snprintf(name, sizeof(name), "%s._edges_%dx%dx%d.raw", filename.c_str(), 2 * X - 1, 2 * Y - 1, 2 * Z - 1);
    printf("outputting labels:\n%s\n", name);
    FILE* fe = fopen(name, "wb");
    if (!fe || fwrite(kinds.data(), 1, nc, fe) != nc) {
        fprintf(stderr, "could not write %s\n", name);
        if (fe) fclose(fe);
        return 2;
    }
    fclose(fe);
    return 0;
}
