// OnDemandAccurate X Y Z filename error_threshold grad_threshold maxnumiter
//                  needASC1 needDSC1 needASC2 needDSC2 PresimpThesh [parallelism] [outputdebug]
//
// Drop-in for the volume path of ondemandaccurate/ondemandaccurate.cxx:704-1157
// with any of needASC1 / needDSC1 (accurate 1-manifolds from the saddles),
// needASC2 and needDSC2 (accurate ascending / descending 3-manifolds):
// writes <filename>.grad, the timing file
// <filename>.<parallelism:%03d>.gtime.txt and, with outputdebug set, the
// int8 restriction labels boundaries_<Mx>x<My>x<Mz>.raw in the working
// directory (the name ondemandaccurate.cxx:1006-1011 gives them) and the
// per-vertex int32 labels in the accurate3dmorselabeling output format:
// <filename>._int_XxYxZ.raw (needASC2) and <filename>.dsc._int_XxYxZ.raw
// (needDSC2; the reference keeps these in memory only).  Without any
// accuracy flag it writes the plain lower-star .grad, as the reference does.
// The timing line of the descending branch covers its region builder,
// integration and cleanup together.
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
    std::vector<int> labels(nv), labels_dsc(d2 ? nv : 0);
    msceer_label_stats st{};
    char tname[2048];
    snprintf(tname, sizeof(tname), "%s.%03d.gtime.txt", filename.c_str(), par);
    FILE* ft = fopen(tname, "w");
    if (a1 || d1 || a2 || d2) {
        const uint32_t need = (a1 ? MSCEER_NEED_ASC1 : 0u) | (d1 ? MSCEER_NEED_DSC1 : 0u) |
                              (a2 ? MSCEER_NEED_ASC2 : 0u) | (d2 ? MSCEER_NEED_DSC2 : 0u);
        std::vector<signed char> restr(dbg ? nc : 0);
        rc = msceer_ondemand2(ctx, values.data(), X, Y, Z, err, gthr, iters, presimp, need, nullptr, grad.data(),
                              a2 ? labels.data() : nullptr, d2 ? labels_dsc.data() : nullptr,
                              dbg ? reinterpret_cast<int8_t*>(restr.data()) : nullptr, &st, MSCEER_TIMING);
        if (rc) { fprintf(stderr, "msceer_ondemand2: %s\n", msceer_strerror(rc)); msceer_destroy(ctx); return 2; }
        // task lines "name start end duration" in ms, as StartTask/EndTaskAndRecord write them
        const float parts[7] = {st.ms_lstar,   st.ms_regions, st.ms_integrate, st.ms_cleanup,
                                st.ms_dsc,     st.ms_lines,   st.ms_grad};
        const char* names[7] = {"Topological Robins0",      "Topological SimExtRBASC",
                                "NumericalTracing NewIntegrationASC", "Removed CleanerASC",
                                "NumericalTracing NewIntegrationDSC", "Numerical SaddleInt",
                                "Topological ConformingGrad"};
        float t = 0.f;
        for (int i = 0; i < 7; i++) {
            if ((i >= 1 && i <= 3 && !a2) || (i == 4 && !d2) || (i == 5 && !(a1 || d1))) continue;
            if (ft) fprintf(ft, "%s %d %d %d\n", names[i], (int)t, (int)(t + parts[i]), (int)parts[i]);
            printf("TIMING: %s %d %d %d\n", names[i], (int)t, (int)(t + parts[i]), (int)parts[i]);
            t += parts[i];
        }
        if (a2) {
            printf("built minmaxgraph: %lld minima, %lld saddles -> %lld regions\n", (long long)st.minima,
                   (long long)st.saddles, (long long)st.regions);
            printf("re-integrated %lld vertices (%lld traced in FP64)\n", (long long)st.reintegrated_pass3,
                   (long long)st.reintegrated_fp64);
        }
        if (d2)
            printf("built maxgraph: %lld maxima, %lld 2-saddles -> %lld descending regions\n", (long long)st.maxima,
                   (long long)st.saddles2, (long long)st.regions_dsc);
        if (a1 || d1)
            printf("1-manifold streamlines: %lld from 2-saddles, %lld from 1-saddles (%lld re-traced in order)\n",
                   (long long)st.lines_asc, (long long)st.lines_dsc, (long long)st.lines_replayed);
        if (dbg && a2) {
            char lname[2048];
            snprintf(lname, sizeof(lname), "%s._int_%dx%dx%d.raw", filename.c_str(), X, Y, Z);
            if (FILE* fl = fopen(lname, "wb")) { fwrite(labels.data(), sizeof(int), nv, fl); fclose(fl); }
        }
        if (dbg && d2) {
            char lname[2048];
            snprintf(lname, sizeof(lname), "%s.dsc._int_%dx%dx%d.raw", filename.c_str(), X, Y, Z);
            if (FILE* fl = fopen(lname, "wb")) { fwrite(labels_dsc.data(), sizeof(int), nv, fl); fclose(fl); }
        }
        if (dbg) {
            char bname[2048];
            snprintf(bname, sizeof(bname), "boundaries_%dx%dx%d.raw", 2 * X - 1, 2 * Y - 1, 2 * Z - 1);
            if (FILE* fb = fopen(bname, "wb")) { fwrite(restr.data(), 1, nc, fb); fclose(fb); }
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
