// Steepest_lstar X Y Z filename per_x per_y per_z [numthreads]
//
// Drop-in for steepest_lstar/steepest_lstar.cxx:52-266: reads a raw float32
// x-fastest volume, computes the lower-star discrete gradient with ldir on
// the GPU and writes <filename>.grad and <filename>.steep.gtime.txt.  The
// periodic flags are parsed and ignored, as in the reference (its cell
// complex is always non-periodic).  numthreads is accepted for argument
// compatibility; the work runs on GPU 0.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "msceer.h"

int main(int argc, char** argv) {
    if (argc < 8) {
        printf("Usage: X Y Z filename per_x per_y per_z  [numthreads=max_on_machine]\n");
        return 0;
    }
    const int X = atoi(argv[1]), Y = atoi(argv[2]), Z = atoi(argv[3]);
    const std::string filename = argv[4];
    const int px = atoi(argv[5]), py = atoi(argv[6]), pz = atoi(argv[7]);
    const int threads = argc >= 9 ? atoi(argv[8]) : 0;
    printf("%d, %d, %d, %d, %d, %d, %d\n", X, Y, Z, px, py, pz, threads);

    const size_t nv = (size_t)X * Y * Z;
    std::vector<float> values(nv);
    FILE* fin = fopen(filename.c_str(), "rb");
    if (!fin || fread(values.data(), sizeof(float), nv, fin) != nv) {
        fprintf(stderr, "could not read %zu floats from %s\n", nv, filename.c_str());
        if (fin) fclose(fin);
        return 2;
    }
    fclose(fin);
    printf("loaded %s...\n", filename.c_str());

    msceer_ctx* ctx = nullptr;
    int rc = msceer_create(&ctx, 0);
    if (rc) { fprintf(stderr, "msceer_create: %s\n", msceer_strerror(rc)); return 2; }
    const size_t nc = (size_t)(2 * X - 1) * (2 * Y - 1) * (2 * Z - 1);
    std::vector<unsigned char> grad(nc);
    const auto t0 = std::chrono::steady_clock::now();
    rc = msceer_lstar_grad(ctx, values.data(), X, Y, Z, grad.data(), MSCEER_TIMING);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc) { fprintf(stderr, "msceer_lstar_grad: %s\n", msceer_strerror(rc)); msceer_destroy(ctx); return 2; }
    msceer_timing t;
    msceer_last_timing(ctx, &t);
    const long ms = (long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
    printf("robins1 %ld\n", (long)(t.lstar_ms + 0.5f));
    printf("computed robins\n");
    for (int d = 0; d < 4; d++) printf("index-%d=%lld\n", d, (long long)t.critical[d]);

    const std::string timing = filename + ".steep.gtime.txt";
    if (FILE* ft = fopen(timing.c_str(), "w")) {
        fprintf(ft, "Total Overall %d %ld %ld\n", 0, ms, ms);
        fclose(ft);
    }
    printf("Total Overall %d %ld %ld\n", 0, ms, ms);
    const std::string gradname = filename + ".grad";
    printf("writing output\n");
    FILE* fg = fopen(gradname.c_str(), "wb");
    if (!fg || fwrite(grad.data(), 1, nc, fg) != nc) {
        fprintf(stderr, "could not write %s\n", gradname.c_str());
        if (fg) fclose(fg);
        msceer_destroy(ctx);
        return 2;
    }
    fclose(fg);
    printf("   -- Output to %s.. %zu values of size 1\n", gradname.c_str(), nc);
    msceer_destroy(ctx);
    return 1;  // the reference's main returns 1 on success
}
