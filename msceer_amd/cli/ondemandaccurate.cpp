// OnDemandAccurate X Y Z filename error_threshold grad_threshold maxnumiter
//                  needASC1 needDSC1 needASC2 needDSC2 PresimpThesh [parallelism] [outputdebug]
//
// Drop-in for the volume path of ondemandaccurate/ondemandaccurate.cxx:704-1063
// with needASC2 (accurate ascending 3-manifolds): writes <filename>.grad, the
// timing file <filename>.<parallelism:%03d>.gtime.txt and, with outputdebug
// set, the per-vertex int32 labels <filename>._int_XxYxZ.raw (the
// accurate3dmorselabeling output format).  Without any accuracy flag it
// writes the plain lower-star .grad, as the reference does.  The 1-manifold
// (needASC1/needDSC1) and descending (needDSC2) branches are not built here
// and are rejected with a message.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "msceer.h"

int main(int argc, char** argv) {
    if (argc < 13) {
        printf("Usage: X Y Z filename error_threshold grad_threshold maxnumiter needASC1 needDSC1 needASC2 "
               "needDSC2 PresimpThesh [parallelism=ompmaxnumthreads] [outputdebug=0]\n");
        return 0;
    }
    const int X = atoi(argv[1]), Y = atoi(argv[2]), Z = atoi(argv[3]);
    const std::string filename = argv[4];
    const float err = (float)atof(argv[5]), gthr = (float)atof(argv[6]);
    const int iters = atoi(argv[7]);
    const int a1 = atoi(argv[8]), d1 = atoi(argv[9]), a2 = atoi(argv[10]), d2 = atoi(argv[11]);
    const float presimp = (float)atof(argv[12]);
    const int par = argc >= 14 ? atoi(argv[13]) : -1;
    const int dbg = argc >= 15 ? atoi(argv[14]) : 0;
    printf("dims=(%d,%d,%d)\nfile=%s\nintegration parameters: e=%f, gt=%f, il=%d\n"
           "ondemandacc: a1=%d, d1=%d, a2=%d, d2=%d, ps=%f\npar=%d\n",
           X, Y, Z, filename.c_str(), err, gthr, iters, a1, d1, a2, d2, presimp, par);
    if (a1 || d1 || d2) {
        fprintf(stderr, "only the ascending 3-manifold branch (needASC2) is available\n");
        return 2;
    }
    const size_t nv = (size_t)X * Y * Z;
    std::vector<float> values(nv);
    FILE* fin = fopen(filename.c_str(), "rb");
    if (!fin || fread(values.data(), sizeof(float), nv, fin) != nv) {
        fprintf(stderr, "could not read %zu floats from %s\n", nv, filename.c_str());
        if (fin) fclose(fin);
        return 2;
    }
    fclose(fin);
    msceer_ctx* ctx = nullptr;
    int rc = msceer_create(&ctx, 0);
    if (rc) { fprintf(stderr, "msceer_create: %s\n", msceer_strerror(rc)); return 2; }
    const size_t nc = (size_t)(2 * X - 1) * (2 * Y - 1) * (2 * Z - 1);
    std::vector<unsigned char> grad(nc);
    std::vector<int> labels(nv);
    msceer_label_stats st{};
    char tname[2048];
    snprintf(tname, sizeof(tname), "%s.%03d.gtime.txt", filename.c_str(), par);
    FILE* ft = fopen(tname, "w");
    if (a2) {
        rc = msceer_ondemand_asc(ctx, values.data(), X, Y, Z, err, gthr, iters, presimp, nullptr, grad.data(),
                                 labels.data(), &st, MSCEER_TIMING);
        if (rc) { fprintf(stderr, "msceer_ondemand_asc: %s\n", msceer_strerror(rc)); msceer_destroy(ctx); return 2; }
        // task lines "name start end duration" in ms, as StartTask/EndTaskAndRecord write them
        const float parts[5] = {st.ms_lstar, st.ms_regions, st.ms_integrate, st.ms_cleanup, st.ms_grad};
        const char* names[5] = {"Topological Robins0", "Topological SimExtRBASC", "NumericalTracing NewIntegrationASC",
                                "Removed CleanerASC", "Topological ConformingGrad"};
        float t = 0.f;
        for (int i = 0; i < 5; i++) {
            if (ft) fprintf(ft, "%s %d %d %d\n", names[i], (int)t, (int)(t + parts[i]), (int)parts[i]);
            printf("TIMING: %s %d %d %d\n", names[i], (int)t, (int)(t + parts[i]), (int)parts[i]);
            t += parts[i];
        }
        printf("built minmaxgraph: %lld minima, %lld saddles -> %lld regions\n", (long long)st.minima,
               (long long)st.saddles, (long long)st.regions);
        printf("re-integrated %lld vertices (%lld traced in FP64)\n", (long long)st.reintegrated_pass3,
               (long long)st.reintegrated_fp64);
        if (dbg) {
            char lname[2048];
            snprintf(lname, sizeof(lname), "%s._int_%dx%dx%d.raw", filename.c_str(), X, Y, Z);
            if (FILE* fl = fopen(lname, "wb")) { fwrite(labels.data(), sizeof(int), nv, fl); fclose(fl); }
        }
    } else {
        printf("no accuraccy needed, outputting\n");
        rc = msceer_lstar_grad(ctx, values.data(), X, Y, Z, grad.data(), 0);
        if (rc) { fprintf(stderr, "msceer_lstar_grad: %s\n", msceer_strerror(rc)); msceer_destroy(ctx); return 2; }
    }
    if (ft) fclose(ft);
    const std::string gradname = filename + ".grad";
    FILE* fg = fopen(gradname.c_str(), "wb");
    if (!fg || fwrite(grad.data(), 1, nc, fg) != nc) {
        fprintf(stderr, "could not write %s\n", gradname.c_str());
        if (fg) fclose(fg);
        msceer_destroy(ctx);
        return 2;
    }
    fclose(fg);
    printf("outputting to file %s\n", gradname.c_str());
    msceer_destroy(ctx);
    return 1;
}
