// dropin_check -- the MI355X engine driven through the reference's own Lattice<2>, device selection and
// strategy types (mi355x_reference.hpp), compiled against the reference's headers and sources.
//
//   dropin_check <target> <N0> <N1> <T> <warmup> <sweeps>
//
// Selects the target by name with mi355x::GetDevice, runs `warmup` sweeps and then `sweeps` sweeps with a
// measurement every 10 (through mi355x::DispatchCall, Lattice<2>::Update<SwendsenWang_2D, LCG32> and
// Lattice<2>::GetEnergyAndMagnetization, as the reference's driver selects and calls them), and prints one
// JSON line: the sums of e, e^2 and |m| over the measurements, the last (e, m), and for the last
// configuration the energy and magnetization recomputed from the lattice's host array by the reference's CPU
// reduction (they agree when the host array holds the GPU's spins).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>

#include "mi355x_reference.hpp"

using namespace cp;

int main(int argc, char** argv)
{
    if (argc != 7)
    {
        std::cerr << "usage: dropin_check <target> <N0> <N1> <T> <warmup> <sweeps>" << std::endl;
        return 2;
    }
    try
    {
        const std::string target_name = argv[1];
        const std::uint32_t n0 = std::atoi(argv[2]), n1 = std::atoi(argv[3]);
        const float temperature = static_cast<float>(std::atof(argv[4]));
        const std::uint32_t warmup = std::atoi(argv[5]), sweeps = std::atoi(argv[6]);
        constexpr std::uint32_t n_sep = 10;

        Lattice<2> lattice({n0, n1});
        auto target = mi355x::GetDevice(target_name);
        const auto update = [&](auto& device, const std::uint32_t iterations) {
            for (std::uint32_t i = 0; i < iterations; ++i)
                lattice.Update<SwendsenWang_2D, LCG32>(temperature, device);
        };
        const auto measure = [&](auto& device) { return lattice.GetEnergyAndMagnetization(device); };

        mi355x::DispatchCall(target_name, *target, update, warmup);
        double sum_e = 0.0, sum_e2 = 0.0, sum_m = 0.0, last_e = 0.0, last_m = 0.0;
        std::uint32_t samples = 0;
        for (std::uint32_t i = 0; i < sweeps; i += n_sep)
        {
            mi355x::DispatchCall(target_name, *target, update, n_sep);
            const auto [e, m] = mi355x::DispatchCall(target_name, *target, measure);
            sum_e += e;
            sum_e2 += e * e;
            sum_m += std::abs(m);
            last_e = e;
            last_m = m;
            ++samples;
        }
        // the last configuration, reduced on the host from the lattice's own array
        CPU cpu(1, 0);
        const auto [host_e, host_m] = lattice.GetEnergyAndMagnetization(cpu);
        std::printf("{\"target\": \"%s\", \"n0\": %u, \"n1\": %u, \"samples\": %u, \"sum_e\": %.17g, \"sum_e2\": %.17g, "
                    "\"sum_abs_m\": %.17g, \"last_e\": %.17g, \"last_m\": %.17g, \"host_e\": %.17g, \"host_m\": %.17g}\n",
                    target_name.c_str(), n0, n1, samples, sum_e, sum_e2, sum_m, last_e, last_m, host_e, host_m);
        return 0;
    }
    catch (const std::exception& e)
    {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
}
