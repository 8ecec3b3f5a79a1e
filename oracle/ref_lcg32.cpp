// ref_lcg32.cpp -- golden-vector harness around the reference lcg32 (TEST INFRASTRUCTURE ONLY).
//
// Built by oracle/Makefile from the reference sources where they lie (random/src/lcg32.cpp is
// compiled as part of this translation unit so that LCG32_State<64> can be instantiated on the
// host; the reference instantiates it only inside its HIP build, random/src/lcg32.cpp:131-134).
// Output: JSON on stdout.
//
//   ref_lcg32 streams <W> <seed> <updates>   states after Init and after every Update
//   ref_lcg32 buffered <seed> <n> <n> ...     LCG32<CPU>::NextInteger(vector) calls of the given sizes
//   ref_lcg32 padd <T> ...                   bits of p_add for temperatures (swendsen_wang.cpp:100)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "random/src/lcg32.cpp" // found through -I$(REF)

namespace cp
{
    template class LCG32_State<64>;
}

using namespace cp;

template <std::uint32_t W>
static void Streams(const std::uint32_t seed, const std::uint32_t updates)
{
    LCG32_State<W> s(seed);
    std::printf("{\"w\": %u, \"seed\": %u, \"states\": [", W, seed);
    for (std::uint32_t k = 0; k <= updates; ++k)
    {
        if (k > 0)
            s.Update();
        std::printf("%s[", k ? ", " : "");
        for (std::uint32_t i = 0; i < W; ++i)
            std::printf("%s%u", i ? ", " : "", s[i]);
        std::printf("]");
    }
    std::printf("]}\n");
}

int main(int argc, char** argv)
{
    if (argc < 2)
        return 1;
    if (!std::strcmp(argv[1], "streams") && argc == 5)
    {
        const auto w = std::atoi(argv[2]);
        const auto seed = static_cast<std::uint32_t>(std::strtoul(argv[3], nullptr, 10));
        const auto updates = static_cast<std::uint32_t>(std::atoi(argv[4]));
        if (w == 16)
            Streams<16>(seed, updates);
        else if (w == 64)
            Streams<64>(seed, updates);
        else
            return 1;
        return 0;
    }
    if (!std::strcmp(argv[1], "buffered") && argc >= 3)
    {
        LCG32<DeviceName::CPU> rng(static_cast<std::uint32_t>(std::strtoul(argv[2], nullptr, 10)));
        std::printf("{\"w\": %u, \"calls\": [", CPU::WavefrontSize<std::uint32_t>());
        for (int k = 3; k < argc; ++k)
        {
            std::vector<std::uint32_t> v(std::atoi(argv[k]));
            rng.NextInteger(v);
            std::printf("%s[", k > 3 ? ", " : "");
            for (std::size_t i = 0; i < v.size(); ++i)
                std::printf("%s%u", i ? ", " : "", v[i]);
            std::printf("]");
        }
        std::printf("]}\n");
        return 0;
    }
    if (!std::strcmp(argv[1], "padd"))
    {
        std::printf("{");
        for (int k = 2; k < argc; ++k)
        {
            // The temperature reaches the strategy as a float (ising_2d.cpp:42,86); the expression is
            // the one of swendsen_wang.cpp:100, compiled here with the reference's flags.
            const float temperature = static_cast<float>(std::atof(argv[k]));
            const float p_add = 1.0f - static_cast<float>(std::exp(-2.0f / temperature));
            std::uint32_t bits;
            std::memcpy(&bits, &p_add, 4);
            std::printf("%s\"%s\": %u", k > 2 ? ", " : "", argv[k], bits);
        }
        std::printf("}\n");
        return 0;
    }
    return 1;
}
