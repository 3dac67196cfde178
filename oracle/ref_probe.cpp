// Probe linked against the reference's own tools.cpp and verlet.cpp (built by
// oracle/Makefile.ref).  Writes golden vectors as JSON on stdout; doubles are
// written as C99 hex floats so they round-trip exactly.
#include "tools/tools.hpp"
#include "verlet/verlet.hpp"
#include <cstdio>
#include <cstdlib>
#include <vector>

static void hex(double v) { std::printf("\"%a\"", v); }

int main() {
    std::printf("{\n\"rand\": [\n");
    const int seeds[] = {0, 1, 42, 12345, 2147483647};
    for (int si = 0; si < 5; si++) {
        mcac::init_random(seeds[si]);
        std::printf("{\"seed\": %d, \"uniform\": [", seeds[si]);
        for (int i = 0; i < 500; i++) { if (i) std::printf(","); hex(mcac::random()); }
        std::printf("], \"direction\": [");
        for (int i = 0; i < 200; i++) {
            auto d = mcac::random_direction();
            if (i) std::printf(",");
            std::printf("["); hex(d[0]); std::printf(","); hex(d[1]); std::printf(","); hex(d[2]); std::printf("]");
        }
        std::printf("]}%s\n", si < 4 ? "," : "");
    }
    std::printf("],\n\"inverf\": [");
    const int M = 801;
    for (int i = 0; i < M; i++) {
        double p = -1.0 + 2.0 * i / (M - 1);
        if (i) std::printf(",");
        std::printf("["); hex(p); std::printf(","); hex(mcac::inverf(p)); std::printf(","); hex(mcac::inverfc(p + 1.0)); std::printf("]");
    }
    std::printf("],\n\"verlet\": [\n");
    // Verlet queries over random memberships: labels in random cells, random sources,
    // sweep vectors and distances (including ones that span the whole grid).
    srand(7);
    auto u = []() { return rand() / (double)RAND_MAX; };
    int ncase = 0;
    const size_t divs[] = {3, 5, 10};
    for (size_t nd : divs) {
        for (int c = 0; c < 40; c++) {
            double width = 1e-6 * (1.0 + 10.0 * u());
            mcac::Verlet v(nd, width);
            int nlab = 1 + rand() % 200;
            std::vector<size_t> cells(3 * nlab);
            for (int l = 0; l < nlab; l++) {
                for (int a = 0; a < 3; a++) cells[3 * l + a] = rand() % nd;
                v.add(l, {cells[3 * l], cells[3 * l + 1], cells[3 * l + 2]});
            }
            std::array<double, 3> src{u() * width, u() * width, u() * width};
            double scale = (c % 4 == 0) ? width : width / nd;
            std::array<double, 3> vec{(2 * u() - 1) * scale, (2 * u() - 1) * scale, (2 * u() - 1) * scale};
            double dist = u() * scale;
            auto out = v.get_neighborhood(src, vec, dist);
            if (ncase++) std::printf(",\n");
            std::printf("{\"n_div\": %zu, \"width\": ", nd); hex(width);
            std::printf(", \"cells\": [");
            for (int l = 0; l < 3 * nlab; l++) std::printf("%s%zu", l ? "," : "", cells[l]);
            std::printf("], \"src\": ["); hex(src[0]); std::printf(","); hex(src[1]); std::printf(","); hex(src[2]);
            std::printf("], \"vec\": ["); hex(vec[0]); std::printf(","); hex(vec[1]); std::printf(","); hex(vec[2]);
            std::printf("], \"dist\": "); hex(dist);
            std::printf(", \"out\": [");
            for (size_t i = 0; i < out.size(); i++) std::printf("%s%zu", i ? "," : "", out[i]);
            std::printf("]}");
        }
    }
    std::printf("\n]\n}\n");
    return 0;
}
