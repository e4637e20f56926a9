// Driver over the reference's own header-only gdt code (test infrastructure).
// Built by oracle/Makefile against the headers where they lie under the
// reference checkout: common/gdt/gdt/random/random.h:55-89 (LCG<16>) and
// common/gdt/gdt/math/vec.h:288-316 (cross, dot, normalize, length).
// Prints known-answer vectors as JSON on stdout.
#include "gdt/math/vec.h"
#include "gdt/random/random.h"
#include <cstdio>
#include <cstring>

static unsigned bits(float f) { unsigned u; std::memcpy(&u, &f, 4); return u; }

int main() {
    std::printf("{\"lcg\": [");
    const unsigned seeds[][2] = {{0, 0}, {1, 0}, {0, 1}, {65535, 7}, {256 * 256 - 1, 15},
                                 {1920u * 1080u - 1u, 63}, {123456789u, 4095u}, {4294967295u, 4294967295u}};
    const int ns = sizeof(seeds) / sizeof(seeds[0]);
    for (int s = 0; s < ns; s++) {
        gdt::LCG<16> r;
        r.init(seeds[s][0], seeds[s][1]);
        std::printf("%s{\"v0\": %u, \"v1\": %u, \"state0\": %u, \"draws\": [", s ? ", " : "", seeds[s][0],
                    seeds[s][1], r.state);
        for (int k = 0; k < 16; k++) {
            float f = r();
            std::printf("%s[%u, %u]", k ? ", " : "", bits(f), r.state);
        }
        std::printf("]}");
    }
    std::printf("], \"vec\": [");
    const float vs[][6] = {{1, 2, 3, -4, 5, 0.5f},   {0.1f, 0.2f, 0.3f, 3, -1, 2},
                           {1e-3f, 7, -2, 0, 0, 1},  {-0.577f, 0.577f, 0.577f, 0.3f, 0.3f, -0.9f},
                           {3e4f, -2e4f, 1, 1e-4f, 2e-4f, 3e-4f}};
    const int nv = sizeof(vs) / sizeof(vs[0]);
    for (int i = 0; i < nv; i++) {
        gdt::vec3f a(vs[i][0], vs[i][1], vs[i][2]), b(vs[i][3], vs[i][4], vs[i][5]);
        gdt::vec3f c = gdt::cross(a, b), n = gdt::normalize(a);
        float d = gdt::dot(a, b), l = gdt::length(a);
        std::printf("%s{\"a\": [%u, %u, %u], \"b\": [%u, %u, %u], \"cross\": [%u, %u, %u], "
                    "\"normalize\": [%u, %u, %u], \"dot\": %u, \"length\": %u}",
                    i ? ", " : "", bits(a.x), bits(a.y), bits(a.z), bits(b.x), bits(b.y), bits(b.z), bits(c.x),
                    bits(c.y), bits(c.z), bits(n.x), bits(n.y), bits(n.z), bits(d), bits(l));
    }
    std::printf("]}\n");
    return 0;
}
