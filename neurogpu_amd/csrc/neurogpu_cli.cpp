/*
 * neuroGPU [i] -- the reference's executable contract on top of libneurogpu
 * (CANON/kernel.cu:8-34, CANON/CudaStuff.cu:518-587).
 *
 * Run from a bin/ directory next to Data/ (or set NGPU_DATA_DIR).  With an
 * argument i it reads Stim_raw<i>.csv and times<i>.csv, runs on device
 * i mod (device count) and writes Data/VHotP<i>.dat; without one it reads
 * Stim_raw.csv and writes Data/VHotP0.dat.  Data/RunTimes.csv receives
 * "NSets,ms".  Unlike the reference, failures exit non-zero with a message.
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/neurogpu.h"

int main(int argc, char **argv)
{
    const char *env = std::getenv("NGPU_DATA_DIR");
    std::string dir = env ? env : "../Data";
    std::string suffix;
    int idx = 0;
    if (argc > 1) {
        suffix = argv[1];
        idx = std::atoi(argv[1]);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        std::fprintf(stderr, "neuroGPU: no HIP device\n");
        return 3;
    }
    ngpu_ctx *ctx = nullptr;
    if (ngpu_open(dir.c_str(), suffix.c_str(), -1, idx % ndev, &ctx) != NGPU_OK) {
        std::fprintf(stderr, "neuroGPU: %s\n", ngpu_last_error(nullptr));
        return 2;
    }
    ngpu_stats st{};
    if (ngpu_run(ctx, nullptr, &st) != NGPU_OK) {
        std::fprintf(stderr, "neuroGPU: %s\n", ngpu_last_error(ctx));
        ngpu_close(ctx);
        return 1;
    }
    std::printf("total took %f ms (kernel %f ms, %.3e comp-steps/s)\n", st.ms_total, st.ms_kernel, st.comp_steps_per_s);
    std::string out = dir + "/VHotP" + std::to_string(idx) + ".dat";
    int rc = ngpu_write_vhot(ctx, out.c_str());
    if (rc != NGPU_OK) std::fprintf(stderr, "neuroGPU: %s\n", ngpu_last_error(ctx));
    FILE *fp = std::fopen((dir + "/RunTimes.csv").c_str(), "w");
    if (fp) {
        std::fprintf(fp, "%d,%f\n", st.nsets, st.ms_total);
        std::fclose(fp);
    }
    ngpu_close(ctx);
    return rc == NGPU_OK ? 0 : 1;
}
