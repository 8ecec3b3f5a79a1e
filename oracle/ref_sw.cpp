// ref_sw.cpp -- harness around the reference CPU Swendsen-Wang path (TEST INFRASTRUCTURE ONLY and the
// CPU baseline of bench.py). Linked by oracle/Makefile against the reference sources compiled where
// they lie: ising/src/{lattice,swendsen_wang,swendsen_wang_cpu}.cpp, random/src/lcg32.cpp,
// utility/device/helper.cpp.
//
//   ref_sw dump <N0> <N1> <T> <out.bin> <k1> <k2> ...
//       hot start (Lattice<2> ctor, srand48(1)), then sweeps through
//       Lattice<2>::Update<SwendsenWang_2D, LCG32>(T, cpu); after the cumulative sweep counts k1 < k2 < ..
//       appends the N0*N1 spin bytes and the int64 pair (sum s(s_r+s_d), sum s) to out.bin.
//   ref_sw run <N0> <N1> <T> <warmup> <sweeps>
//       the measurement protocol of ising/src/ising_2d.cpp:120-162 (observables every 10 sweeps inside
//       the timed region) with a configurable warmup, printing the reference report lines.
//   ref_sw series <N0> <N1> <T> <warmup> <sweeps> <out.bin>
//       the same protocol, writing the int64 pair (sum s(s_r+s_d), sum s) of every sample to out.bin, so
//       that error bars can be computed from the reference CPU run (the driver prints means only).
// NUM_THREADS selects the thread count as in the reference (utility/device/device_cpu.hpp:23).
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "swendsen_wang.hpp"

using namespace cp;

static void Sums(Lattice<2>& lattice, std::int64_t& e, std::int64_t& m)
{
    const auto n0 = lattice.Extent()[0], n1 = lattice.Extent()[1];
    const auto* s = lattice.RawPointer();
    e = 0;
    m = 0;
    for (std::uint32_t y = 0; y < n1; ++y)
        for (std::uint32_t x = 0; x < n0; ++x)
        {
            const int v = 2 * s[y * n0 + x] - 1;
            e += v * ((2 * s[y * n0 + (x + 1) % n0] - 1) + (2 * s[((y + 1) % n1) * n0 + x] - 1));
            m += v;
        }
}

int main(int argc, char** argv)
{
    if (argc >= 6 && !std::strcmp(argv[1], "dump"))
    {
        const std::uint32_t n0 = std::atoi(argv[2]), n1 = std::atoi(argv[3]);
        const float temperature = static_cast<float>(std::atof(argv[4]));
        std::FILE* out = std::fopen(argv[5], "wb");
        if (!out)
            return 1;
        Lattice<2> lattice({n0, n1});
        CPU cpu;
        std::uint32_t done = 0;
        for (int k = 6; k < argc; ++k)
        {
            const std::uint32_t target = std::atoi(argv[k]);
            for (; done < target; ++done)
                lattice.Update<SwendsenWang_2D, LCG32>(temperature, cpu);
            std::int64_t sums[2];
            Sums(lattice, sums[0], sums[1]);
            std::fwrite(lattice.RawPointer(), 1, static_cast<std::size_t>(n0) * n1, out);
            std::fwrite(sums, sizeof(std::int64_t), 2, out);
        }
        std::fclose(out);
        return 0;
    }
    if (argc == 7 && !std::strcmp(argv[1], "run"))
    {
        const std::uint32_t n0 = std::atoi(argv[2]), n1 = std::atoi(argv[3]);
        const float temperature = static_cast<float>(std::atof(argv[4]));
        const std::uint32_t n_warmup = std::atoi(argv[5]), n_sweeps = std::atoi(argv[6]);
        constexpr std::uint32_t n_sep = 10; // ising_2d.cpp:27
        Lattice<2> lattice({n0, n1});
        CPU cpu;
        std::cout << "Target name: CPU" << std::endl;
        std::cout << "WARMUP: " << n_warmup << " sweeps" << std::endl;
        std::cout << "MEASUREMENT: " << n_sweeps << " sweeps" << std::endl;
        for (std::uint32_t i = 0; i < n_warmup; ++i)
            lattice.Update<SwendsenWang_2D, LCG32>(temperature, cpu);
        double energy = 0.0, energy_squared = 0.0, magnetization = 0.0;
        const auto start = std::chrono::high_resolution_clock::now();
        for (std::uint32_t i = 0; i < n_sweeps; i += n_sep)
        {
            for (std::uint32_t k = 0; k < n_sep; ++k)
                lattice.Update<SwendsenWang_2D, LCG32>(temperature, cpu);
            auto [e, m] = lattice.GetEnergyAndMagnetization(cpu);
            energy += e;
            energy_squared += e * e;
            magnetization += std::abs(m);
        }
        const auto end = std::chrono::high_resolution_clock::now();
        const double elapsed = std::chrono::duration_cast<std::chrono::microseconds>(end - start).count() * 1.0e-6;
        const auto samples = n_sweeps / n_sep;
        const auto heat_capacity = ((energy_squared / samples) - (energy / samples) * (energy / samples)) *
            lattice.NumSites() / (temperature * temperature);
        std::cout << "Update time per spin: " << elapsed * 1.0e9 / (static_cast<std::int64_t>(n_sweeps) * n0 * n1) << " ns" << std::endl;
        std::cout << "Internal energy per spin: " << energy / samples << std::endl;
        std::cout << "Heat capacity per spin: " << heat_capacity << std::endl;
        std::cout << "Absolute Magnetization per spin: " << magnetization / samples << std::endl;
        std::cout << "Elapsed seconds: " << elapsed << std::endl;
        std::cout << "Threads: " << cpu.Concurrency() << std::endl;
        return 0;
    }
    if (argc == 8 && !std::strcmp(argv[1], "series"))
    {
        const std::uint32_t n0 = std::atoi(argv[2]), n1 = std::atoi(argv[3]);
        const float temperature = static_cast<float>(std::atof(argv[4]));
        const std::uint32_t n_warmup = std::atoi(argv[5]), n_sweeps = std::atoi(argv[6]);
        constexpr std::uint32_t n_sep = 10; // ising_2d.cpp:27
        std::FILE* out = std::fopen(argv[7], "wb");
        if (!out)
            return 1;
        Lattice<2> lattice({n0, n1});
        CPU cpu;
        for (std::uint32_t i = 0; i < n_warmup; ++i)
            lattice.Update<SwendsenWang_2D, LCG32>(temperature, cpu);
        for (std::uint32_t i = 0; i < n_sweeps; i += n_sep)
        {
            for (std::uint32_t k = 0; k < n_sep; ++k)
                lattice.Update<SwendsenWang_2D, LCG32>(temperature, cpu);
            std::int64_t sums[2];
            Sums(lattice, sums[0], sums[1]);
            std::fwrite(sums, sizeof(std::int64_t), 2, out);
        }
        std::fclose(out);
        return 0;
    }
    std::fprintf(stderr, "usage: ref_sw dump N0 N1 T out.bin k1 [k2 ..] | ref_sw run N0 N1 T warmup sweeps | "
                         "ref_sw series N0 N1 T warmup sweeps out.bin\n");
    return 1;
}
