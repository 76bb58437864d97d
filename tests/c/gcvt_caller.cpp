// A reference-style C++ caller: centroidalVoronoi (source/gcvt.h:124-139)
// with its host part replaced by cvt_ref_init_sites (the reference's gcvt.h
// does not compile on glibc) and the GPU part by the gCVT shim, which keeps
// the reference's signature and C++ linkage (include/gcvt_compat.h).
// "host" runs the initialiser only (no GPU needed) and checks that both
// shims resolved; with no argument the whole call runs.
#include <cstdio>
#include <cstring>
#include <vector>

#include "cvt.h"
#include "gcvt_compat.h"

static void centroidalVoronoi(short *Voronoi, float *density, bool *constrainMask, int vertices,
                              int imageSize, int depth, int maxIter, bool run_gpu)
{
    const int rc = cvt_ref_init_sites(reinterpret_cast<int16_t *>(Voronoi), density,
                                      reinterpret_cast<const uint8_t *>(constrainMask), imageSize,
                                      vertices, nullptr, nullptr);
    if (rc != CVT_OK) {
        std::printf("init failed: %s\n", cvt_last_error());
        return;
    }
    if (run_gpu) gCVT(Voronoi, density, constrainMask, imageSize, depth, maxIter);
}

int main(int argc, char **argv)
{
    const bool host_only = argc > 1 && std::strcmp(argv[1], "host") == 0;
    const int n = 512;
    std::vector<float> density((size_t)n * n, 1.0f);
    std::vector<short> vor(2 * (size_t)n * n);
    bool *mask = new bool[(size_t)n * n]();
    void (*g)(short *, float *, bool *, int, int, int) = &gCVT;
    void (*d)(double *, double *, int, int *, int, float *, double, int) = &discretization_d;
    if (!g || !d) return 1;
    centroidalVoronoi(vor.data(), density.data(), mask, 1, n, 1, 10, !host_only);
    for (int p = 0; p < n * n; ++p)
        if (vor[2 * p] != -32768) {
            std::printf("first site %d %d\n", vor[2 * p], vor[2 * p + 1]);
            break;
        }
    delete[] mask;
    return 0;
}
