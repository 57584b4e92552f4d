// Host harness for tests/test_rounding_words.py: writes the stochastic-
// rounding operands (one per field) of every cell of an nx x ny x nz box
// for `steps` consecutive step keys, as the kernels derive them
// (fslbm_store.h: sr_salt, sr_plane, sr_words, sr_key), to stdout as
// little-endian uint32 in the order step, z, y, x, field.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fslbm_store.h"

int main(int argc, char **argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s nx ny nz steps seed\n", argv[0]);
        return 2;
    }
    const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]);
    const int steps = std::atoi(argv[4]);
    const unsigned seed = (unsigned)std::strtoul(argv[5], nullptr, 10);
    std::vector<uint32_t> out((size_t)nx * ny * 4);
    for (int s = 0; s < steps; s++) {
        const uint32_t key = fslbm::sr_key(seed + (uint32_t)s);
        for (int z = 0; z < nz; z++) {
            const fslbm::SrPlane k = fslbm::sr_plane(key, (uint32_t)z);
            for (int c = 0; c < nx * ny; c++) {
                const fslbm::SrWords w = fslbm::sr_words(fslbm::sr_salt((uint32_t)c), k);
                for (int f = 0; f < 4; f++) out[4 * (size_t)c + f] = w.r[f];
            }
            std::fwrite(out.data(), 4, out.size(), stdout);
        }
    }
    return 0;
}
