// swmi.hpp -- C++ host mirror of the reference's Monte Carlo operator interface, backed by the MI355X
// engine's C ABI (include/swmi.h).
//
// Names and meaning follow flwende/computational_physics so that code written against
//   Lattice<2>::Update<SwendsenWang_2D, LCG32>(temperature, target)        (ising/include/lattice.hpp:46-58)
//   Lattice<2>::GetEnergyAndMagnetization(target) -> (-E/N, M/N)           (ising/include/lattice.hpp:66-69)
//   LatticeMonteCarloAlgorithm<Dim, RNG, Target>::Update(lattice, T)      (ising/include/lattice_mc.hpp:13-21)
//   AbstractDevice {IsOffloadDevice, Concurrency, Name, DeviceId}          (utility/device/device.hpp:11-42)
// reads the same with the target MI355X. Errors are std::runtime_error, as in the reference; the messages
// are the engine's (swmi_last_error). There is no CPU fallback: without a gfx950 device the MI355X
// constructor throws.
#pragma once

#include <array>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "swmi.h"

namespace cp
{
    enum class DeviceName : std::uint32_t
    {
        CPU = 1,
        AMD_GPU = 2,
        MI355X = 3
    };

    inline void SwmiCheck(const int status)
    {
        if (status != SWMI_OK)
            throw std::runtime_error(std::string("MI355X engine: ") + swmi_last_error());
    }

    class AbstractDevice
    {
    protected:
        DeviceName device_name{};
        std::uint32_t device_id{};

        AbstractDevice(const DeviceName name, const std::uint32_t id) : device_name(name), device_id(id) {}

    public:
        virtual ~AbstractDevice() noexcept = default;
        virtual bool IsOffloadDevice() const noexcept = 0;
        virtual std::uint32_t Concurrency() const noexcept = 0;
        DeviceName Name() const noexcept { return device_name; }
        std::uint32_t DeviceId() const noexcept { return device_id; }
    };

    // One MI355X (gfx950). Concurrency() is the number of compute units.
    class MI355X final : public AbstractDevice
    {
        swmi_device_info info{};

    public:
        explicit MI355X(const std::uint32_t id = 0) : AbstractDevice(DeviceName::MI355X, id)
        {
            if (id != 0)
                throw std::runtime_error("MI355X engine: only device 0 of the process is used");
            SwmiCheck(swmi_get_device_info(&info));
        }
        static constexpr DeviceName StaticName() noexcept { return DeviceName::MI355X; }
        bool IsOffloadDevice() const noexcept override { return true; }
        std::uint32_t Concurrency() const noexcept override { return info.compute_units; }
        std::uint32_t WavefrontSize() const noexcept { return info.wavefront_size; }
        const swmi_device_info& Info() const noexcept { return info; }
    };

    // RNG tag: the lcg32 generator (random/include/random/lcg32.hpp). On the MI355X it is the engine's
    // per-site layout, 64-site groups of LCG32_State<64>.
    template <DeviceName Target>
    class LCG32
    {
    };

    template <std::uint32_t Dimension>
    class Lattice;

    template <std::uint32_t Dimension, template <DeviceName> typename RNG, DeviceName Target>
    class LatticeMonteCarloAlgorithm
    {
    public:
        virtual ~LatticeMonteCarloAlgorithm() noexcept = default;
        virtual void Update(Lattice<Dimension>& lattice, float temperature) = 0;
    };

    template <template <DeviceName> typename RNG, DeviceName Target>
    class SwendsenWang_2D;

    // 2D spin lattice: uint8 spins in {0,1}, row-major, hot start srand48(1) (ising/src/lattice.cpp:15-27).
    // The host copy is authoritative until the first device update; afterwards the device state is, and
    // SyncHost() brings it back.
    template <>
    class Lattice<2> final
    {
    public:
        using Spin = std::uint8_t;

    private:
        std::array<std::uint32_t, 2> extent{};
        std::size_t num_sites{};
        std::vector<Spin> spins;
        swmi_engine* engine = nullptr;
        float engine_temperature = 0.0f;
        std::uint32_t seed = 0;
        std::uint32_t tile_rows = 0;
        std::uint32_t flip_rule = SWMI_FLIP_PARITY;

        template <template <DeviceName> typename, DeviceName>
        friend class SwendsenWang_2D;

        // Creates the engine on first use and hands it the host spins.
        swmi_engine* Engine(const float temperature)
        {
            if (!engine)
            {
                swmi_lattice_config c{};
                c.n0 = extent[0];
                c.n1 = extent[1];
                c.temperature = temperature;
                c.seed = seed;
                c.spin_seed = 1;
                c.tile_rows = tile_rows;
                c.flip_rule = flip_rule;
                c.max_samples = 0;
                SwmiCheck(swmi_create(&c, 1, &engine));
                SwmiCheck(swmi_upload_spins(engine, 0, spins.data()));
                engine_temperature = temperature;
            }
            else if (temperature != engine_temperature)
            {
                SwmiCheck(swmi_set_temperature(engine, 0, temperature));
                engine_temperature = temperature;
            }
            return engine;
        }

    public:
        // seed: the engine's lcg32 seed; tile_rows: tile height hint (results do not depend on it);
        // rule: SWMI_FLIP_PARITY (the reference's) or SWMI_FLIP_RANDOM.
        explicit Lattice(const std::array<std::uint32_t, 2>& ext, const std::uint32_t rng_seed = 0,
                         const std::uint32_t tile_rows_hint = 0, const std::uint32_t rule = SWMI_FLIP_PARITY)
            : extent(ext), num_sites(static_cast<std::size_t>(ext[0]) * ext[1]), spins(num_sites), seed(rng_seed),
              tile_rows(tile_rows_hint), flip_rule(rule)
        {
            // lattice.cpp:24-26: srand48(1); spin = drand48() < 0.5 ? 0 : 1
            unsigned short xs[3] = {0x330E, 1, 0};
            for (std::size_t i = 0; i < num_sites; ++i)
                spins[i] = erand48(xs) < 0.5 ? 0 : 1;
        }
        Lattice(const Lattice&) = delete;
        Lattice& operator=(const Lattice&) = delete;
        ~Lattice()
        {
            if (engine)
                (void)swmi_destroy(engine);
        }

        const auto& Extent() const noexcept { return extent; }
        std::size_t NumSites() const noexcept { return num_sites; }
        Spin* RawPointer() noexcept { return spins.data(); }
        const Spin* RawPointer() const noexcept { return spins.data(); }
        Spin operator[](const std::int32_t index) const { return spins[index]; }

        // Lattice<2>::Update<Strategy, RNG>(T, target) (lattice.hpp:46-58): one sweep on the target.
        template <template <template <DeviceName> typename, DeviceName> typename Strategy,
                  template <DeviceName> typename RNG, typename Target>
        void Update(const float temperature, Target& target)
        {
            static_assert(Target::StaticName() == DeviceName::MI355X, "this build drives the MI355X only");
            Strategy<RNG, Target::StaticName()> strategy(target);
            strategy.Update(*this, temperature);
        }

        // (-E/N, M/N) of the current configuration (lattice.cpp:29-79), reduced on the device.
        std::pair<double, double> GetEnergyAndMagnetization(MI355X&)
        {
            // before the first update the engine is created at the reference temperature; the next
            // Update sets its own
            swmi_engine* e = engine ? engine : Engine(2.2691853142130221f);
            std::int64_t esum = 0, msum = 0;
            SwmiCheck(swmi_observables(e, 0, &esum, &msum));
            return {-static_cast<double>(esum) / num_sites, static_cast<double>(msum) / num_sites};
        }

        // Device measurement loop (ising_2d.cpp:126-162 on the device): n_sweeps sweeps, a sample every
        // n_sep; returns {samples, sum e, sum e^2, sum |m|, sum m^2, sum m^4} of this call.
        std::array<double, 6> Measure(const float temperature, const std::uint32_t n_sweeps, const std::uint32_t n_sep)
        {
            swmi_engine* e = Engine(temperature);
            SwmiCheck(swmi_reset_measurements(e));
            SwmiCheck(swmi_run(e, n_sweeps, n_sep));
            std::array<double, 6> mom{};
            SwmiCheck(swmi_get_moments(e, 0, mom.data()));
            return mom;
        }

        void SyncHost()
        {
            if (engine)
                SwmiCheck(swmi_download_spins(engine, 0, spins.data()));
        }

        void Synchronize()
        {
            if (engine)
                SwmiCheck(swmi_synchronize(engine));
        }
    };

    // SwendsenWang_2D<LCG32, MI355X>: Update is one sweep of the engine (swendsen_wang.cpp:94-112:
    // bonds with p_add = 1 - exp(-2/T), labelling, merge, resolve, parity flip).
    template <>
    class SwendsenWang_2D<LCG32, DeviceName::MI355X> final : public LatticeMonteCarloAlgorithm<2, LCG32, DeviceName::MI355X>
    {
    public:
        explicit SwendsenWang_2D(MI355X&) {}
        void Update(Lattice<2>& lattice, const float temperature) override
        {
            if ((lattice.Extent()[0] % 2) || (lattice.Extent()[1] % 2))
                throw std::runtime_error("SwendsenWang_2D implementation requires evenly sized lattice.");
            SwmiCheck(swmi_sweep(lattice.Engine(temperature), 1));
        }
    };

    // Selection by name (utility/device/helper.hpp:16-29): only "mi355x" is served here.
    inline std::unique_ptr<AbstractDevice> GetDevice(const std::string& name)
    {
        if (name == "mi355x")
            return std::make_unique<MI355X>();
        throw std::runtime_error("Unknown target device.");
    }
}
