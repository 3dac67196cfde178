// mcac_run <params.ini> [output_dir] — runs one aggregation on the MI355X path
// (the reference's `MCAC <params.ini>`, src/main.cpp:26-56).  Exit codes are
// the reference ErrorCodes.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "../../../include/mcac_sim.h"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <params.ini> [output_dir]\n", argv[0]);
        return 4;
    }
    std::ifstream f(argv[1]);
    if (!f) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    char err[512] = {0};
    mcac_sim *sim = nullptr;
    int rc = mcac_sim_create(ss.str().c_str(), argc > 2 ? argv[2] : nullptr, 0, &sim, err, sizeof(err));
    if (rc) {
        std::fprintf(stderr, "%s\n", err);
        return rc;
    }
    auto t0 = std::chrono::steady_clock::now();
    rc = mcac_sim_init(sim);
    if (rc) {
        std::fprintf(stderr, "init failed: %d\n", rc);
        mcac_sim_destroy(sim);
        return rc;
    }
    int64_t done = 0, total = 0;
    while (!mcac_sim_finished(sim)) {
        rc = mcac_sim_run(sim, 100000, &done);
        total += done;
        if (rc || done == 0) break;  // an error, or a run that can make no further step
        std::printf("steps %lld  aggregates %.0f  time %.6e s\n", static_cast<long long>(total),
                    mcac_sim_scalar(sim, "n_agg"), mcac_sim_scalar(sim, "time"));
        std::fflush(stdout);
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Final number of aggregates : %.0f\nFinal residence time=%.6e s\nsteps %lld in %.3f s\n",
                mcac_sim_scalar(sim, "n_agg"), mcac_sim_scalar(sim, "time"), static_cast<long long>(total), secs);
    mcac_sim_destroy(sim);
    return rc;
}
