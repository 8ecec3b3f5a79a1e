// ising2d -- command-line driver of the MI355X Swendsen-Wang engine.
//
// Same flags, defaults and report lines as the reference driver ising/src/ising_2d.cpp:26-162
// (--extent=N0xN1, --num_sweeps, --temperature, --algorithm=swendsen_wang, --rng=lcg32, --target), with
// the target mi355x. Additional flags: --num_warmup (default 1000, as the reference's fixed N_Warmup),
// --seed (lcg32 seed), --tile_rows (tile height hint), --measure=host|device. With --measure=host the loop
// is the reference's: N_Sep sweeps, then GetEnergyAndMagnetization and FP64 accumulation on the host. With
// --measure=device (default) the samples are reduced and accumulated on the GPU (swmi_run).
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>

#include "swmi.hpp"

using namespace cp;

namespace
{
    constexpr std::uint32_t kSep = 10;  // ising_2d.cpp:27

    struct Options
    {
        std::uint32_t n0 = 32, n1 = 32;             // ising_2d.cpp:31-32
        float temperature = 2.2691853142130221F;    // ising_2d.cpp:33
        std::uint32_t n_sweeps = kSep * 10000;      // ising_2d.cpp:34
        std::uint32_t n_warmup = 1000;              // ising_2d.cpp:26
        std::string algorithm = "swendsen_wang";
        std::string rng = "lcg32";
        std::string target = "mi355x";
        std::string measure = "device";
        std::uint32_t seed = 0;
        std::uint32_t tile_rows = 0;
        std::string flip_rule = "parity";
    };

    void Help(const char* prog)
    {
        std::cout << "Usage: " << prog << " [options]\n"
                  << "Options:\n"
                  << "  --help, -h         Show this help message\n"
                  << "  --algorithm        Set algorithm (default: swendsen_wang)\n"
                  << "  --extent=N0xN1     Set lattice dimensions\n"
                  << "  --num_sweeps       Set number of MC sweeps (default: 100000)\n"
                  << "  --num_warmup       Set number of warmup sweeps (default: 1000)\n"
                  << "  --rng=<name>       Set RNG type (default: lcg32)\n"
                  << "  --target=<name>    Set target device (default: mi355x)\n"
                  << "      Supported target devices: mi355x\n"
                  << "  --temperature      Set temperature (default: 2.2691853)\n"
                  << "  --seed             Set the lcg32 seed (default: 0)\n"
                  << "  --tile_rows        Tile height hint in 64-site rows (default: engine choice)\n"
                  << "  --flip_rule=<rule> parity (reference: flip iff the root index is odd) or random (default: parity)\n"
                  << "  --measure=<where>  Accumulate samples on the device or the host (default: device)\n";
    }

    Options Parse(int argc, char** argv)
    {
        Options o;
        for (int i = 1; i < argc; ++i)
        {
            const std::string arg = argv[i];
            const auto pos = arg.find('=');
            if (pos == std::string::npos)
                continue;
            const std::string key = arg.substr(0, pos), value = arg.substr(pos + 1);
            if (key == "--extent")
            {
                if (const auto x = value.find('x'); x != std::string::npos)
                {
                    o.n0 = static_cast<std::uint32_t>(std::atoi(value.substr(0, x).c_str()));
                    o.n1 = static_cast<std::uint32_t>(std::atoi(value.substr(x + 1).c_str()));
                }
            }
            else if (key == "--num_sweeps")
                o.n_sweeps = static_cast<std::uint32_t>(std::atoi(value.c_str()));
            else if (key == "--num_warmup")
                o.n_warmup = static_cast<std::uint32_t>(std::atoi(value.c_str()));
            else if (key == "--temperature")
                o.temperature = static_cast<float>(std::atof(value.c_str()));
            else if (key == "--algorithm")
                o.algorithm = value;
            else if (key == "--rng")
                o.rng = value;
            else if (key == "--target")
                o.target = value;
            else if (key == "--measure")
                o.measure = value;
            else if (key == "--seed")
                o.seed = static_cast<std::uint32_t>(std::strtoul(value.c_str(), nullptr, 10));
            else if (key == "--tile_rows")
                o.tile_rows = static_cast<std::uint32_t>(std::atoi(value.c_str()));
            else if (key == "--flip_rule")
                o.flip_rule = value;
        }
        return o;
    }

    int Run(const Options& o)
    {
        if (o.rng != "lcg32")
            throw std::runtime_error("Unknown RNG name.");
        if (o.algorithm != "swendsen_wang")
            throw std::runtime_error("Unknown algorithm.");
        if (o.measure != "device" && o.measure != "host")
            throw std::runtime_error("Unknown measurement mode.");
        if (o.flip_rule != "parity" && o.flip_rule != "random")
            throw std::runtime_error("Unknown flip rule.");
        auto device = GetDevice(o.target);
        auto& gpu = static_cast<MI355X&>(*device);

        Lattice<2> lattice({o.n0, o.n1}, o.seed, o.tile_rows, o.flip_rule == "random" ? SWMI_FLIP_RANDOM : SWMI_FLIP_PARITY);
        // ising_2d.cpp:101 prints "AMD GPU" for every target that is not the CPU; the device details go to
        // stderr so that stdout holds the reference's report lines only
        std::cout << "Target name: " << "AMD GPU" << std::endl;
        std::clog << "device: MI355X (" << gpu.Info().arch << ", " << gpu.Concurrency() << " CUs)" << std::endl;
        std::cout << "WARMUP: " << o.n_warmup << " sweeps" << std::endl;
        std::cout << "MEASUREMENT: " << o.n_sweeps << " sweeps" << std::endl;

        for (std::uint32_t i = 0; i < o.n_warmup; ++i)
            lattice.Update<SwendsenWang_2D, LCG32>(o.temperature, gpu);
        lattice.Synchronize();

        double energy = 0.0, energy_squared = 0.0, magnetization = 0.0;
        const auto t0 = std::chrono::high_resolution_clock::now();
        if (o.measure == "host")
        {
            for (std::uint32_t i = 0; i < o.n_sweeps; i += kSep)
            {
                for (std::uint32_t k = 0; k < kSep; ++k)
                    lattice.Update<SwendsenWang_2D, LCG32>(o.temperature, gpu);
                const auto [e, m] = lattice.GetEnergyAndMagnetization(gpu);
                energy += e;
                energy_squared += e * e;
                magnetization += std::abs(m);
            }
        }
        else
        {
            const auto mom = lattice.Measure(o.temperature, o.n_sweeps, kSep);
            energy = mom[1];
            energy_squared = mom[2];
            magnetization = mom[3];
        }
        lattice.Synchronize();
        const auto t1 = std::chrono::high_resolution_clock::now();
        const double elapsed_s = std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count() * 1.0E-6;

        // report lines of ising_2d.cpp:150-161, with the stream's default precision. As there, the sums are
        // divided by n_sweeps / N_Sep (integer division), the number of groups when N_Sep divides n_sweeps.
        std::cout << "Update time per spin: "
                  << elapsed_s * 1.0E9 / (static_cast<std::int64_t>(o.n_sweeps) * o.n0 * o.n1) << " ns" << std::endl;
        const auto samples = o.n_sweeps / kSep;
        const auto heat_capacity = ((energy_squared / samples) - (energy / samples) * (energy / samples)) *
                                   lattice.NumSites() / (o.temperature * o.temperature);
        std::cout << "Internal energy per spin: " << energy / samples << std::endl;
        std::cout << "Heat capacity per spin: " << heat_capacity << std::endl;
        std::cout << "Absolute Magnetization per spin: " << magnetization / samples << std::endl;
        return 0;
    }
}

int main(int argc, char** argv)
{
    if (argc > 1 && (std::string(argv[1]) == "--help" || std::string(argv[1]) == "-h"))
    {
        Help(argv[0]);
        return 0;
    }
    try
    {
        return Run(Parse(argc, argv));
    }
    catch (const std::exception& e)
    {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
}
