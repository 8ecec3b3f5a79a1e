// mi355x_reference.hpp -- the MI355X engine behind the reference's own device and strategy types.
//
// This header is compiled inside flwende/computational_physics, after its headers and with its include
// paths (-I utility -I random/include -I ising/include), plus -I <this repo>/include, and linked with
// -lswmi. It extends the reference's types rather than declaring its own:
//   * a third DeviceName value, MI355X_DEVICE = 3. The enum's underlying type is fixed (std::uint32_t,
//     utility/device/device.hpp:13-17), so 3 is a value of the type without editing the enum;
//   * class MI355X : AbstractDevice (utility/device/device.hpp:22-42) with the static Name() the reference's
//     templates call as Target::Name() (ising/include/lattice.hpp:48-52), and Device<MI355X_DEVICE>::Type
//     for the default GetDevice<Target>() (lattice.hpp:85-90);
//   * the explicit specialisation SwendsenWang_2D<LCG32, MI355X_DEVICE> (ising/include/swendsen_wang.hpp:27)
//     that Lattice<2>::Update<SwendsenWang_2D, LCG32>(T, target) instantiates as its function-local static
//     strategy (lattice.hpp:46-58);
//   * the explicit specialisation Lattice<2>::GetEnergyAndMagnetization<MI355X> (declared at lattice.hpp:
//     66-69, specialised per target at ising/src/lattice.cpp:29-79);
//   * mi355x::GetDevice / mi355x::DispatchCall: the reference's string selection (utility/device/helper.hpp:
//     16-29, helper.cpp:10-20) with the "mi355x" branch, delegating every other name to the reference's.
//
// State. The strategy is one static per process (lattice.hpp:56), so the engines live in the device: one
// engine per Lattice object, created on its first Update from the lattice's host spins (the reference's
// srand48(1) hot start, or whatever the caller wrote there). GetEnergyAndMagnetization reduces on the GPU
// and copies the spins back into the lattice's host array, as the reference's GPU target does
// (lattice.cpp:69-79), so the host copy is current after every measurement. The lcg32 streams are the
// engine's per-site layout (seed 0), not the reference CPU's per-thread streams.
//
// Errors are std::runtime_error with the engine's message, as the reference reports errors.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>

#include "device/device.hpp"
#include "lattice.hpp"
#include "lattice_mc.hpp"
#include "random/random.hpp"
#include "swendsen_wang.hpp"
#include "swmi.h"

namespace cp
{
    inline constexpr DeviceName MI355X_DEVICE = static_cast<DeviceName>(3);

    inline void SwmiCall(const int status)
    {
        if (status != SWMI_OK)
            throw std::runtime_error(std::string("MI355X: ") + swmi_last_error());
    }

    class MI355X final : public AbstractDevice
    {
        struct Run
        {
            swmi_engine* engine = nullptr;
            float temperature = 0.0f;
        };

        swmi_device_info info{};
        std::map<const void*, Run> runs;   // by Lattice object
        std::mutex mutex;

    public:
        MI355X() : MI355X(0) {}

        explicit MI355X(const std::uint32_t device_id) : AbstractDevice(MI355X_DEVICE, device_id)
        {
            if (device_id != 0)
                throw std::runtime_error("MI355X: the engine runs on device 0 of the process");
            SwmiCall(swmi_get_device_info(&info));
        }

        MI355X(const MI355X&) = delete;
        MI355X& operator=(const MI355X&) = delete;

        ~MI355X() override
        {
            for (auto& r : runs)
                (void)swmi_destroy(r.second.engine);
        }

        static constexpr auto Name() noexcept { return MI355X_DEVICE; }

        template <typename T>
        static constexpr auto WavefrontSize() noexcept
        {
            return 64u;
        }

        bool IsOffloadDevice() const noexcept override { return true; }

        std::uint32_t Concurrency() const noexcept override { return info.compute_units; }

        const swmi_device_info& Info() const noexcept { return info; }

        // One Swendsen-Wang sweep of `lattice` at `temperature` (its engine is created on first use).
        void Sweep(Lattice<2>& lattice, const float temperature)
        {
            std::lock_guard<std::mutex> lock(mutex);
            Run& r = runs[&lattice];
            if (!r.engine)
            {
                swmi_lattice_config c{};
                c.n0 = lattice.Extent()[0];
                c.n1 = lattice.Extent()[1];
                c.temperature = temperature;
                c.seed = 0;
                c.spin_seed = 1;
                c.flip_rule = SWMI_FLIP_PARITY;
                swmi_engine* e = nullptr;
                const int status = swmi_create(&c, 1, &e);
                if (status != SWMI_OK)
                {
                    runs.erase(&lattice);
                    SwmiCall(status);
                }
                r.engine = e;
                r.temperature = temperature;
                SwmiCall(swmi_upload_spins(e, 0, lattice.RawPointer()));
            }
            else if (temperature != r.temperature)
            {
                SwmiCall(swmi_set_temperature(r.engine, 0, temperature));
                r.temperature = temperature;
            }
            SwmiCall(swmi_sweep(r.engine, 1));
        }

        // (E, M) of the lattice on the GPU, or false if no sweep ran on it yet; the host array receives the spins.
        bool Observe(Lattice<2>& lattice, std::int64_t& energy, std::int64_t& magnetization)
        {
            std::lock_guard<std::mutex> lock(mutex);
            const auto it = runs.find(&lattice);
            if (it == runs.end())
                return false;
            SwmiCall(swmi_observables(it->second.engine, 0, &energy, &magnetization));
            SwmiCall(swmi_download_spins(it->second.engine, 0, lattice.RawPointer()));
            return true;
        }

        // Drops the lattice's engine (its host array keeps the last spins observed).
        void Release(const Lattice<2>& lattice)
        {
            std::lock_guard<std::mutex> lock(mutex);
            const auto it = runs.find(&lattice);
            if (it != runs.end())
            {
                (void)swmi_destroy(it->second.engine);
                runs.erase(it);
            }
        }

        void Synchronize()
        {
            std::lock_guard<std::mutex> lock(mutex);
            for (auto& r : runs)
                SwmiCall(swmi_synchronize(r.second.engine));
        }
    };

    template <>
    struct Device<MI355X_DEVICE>
    {
        using Type = MI355X;
    };

    // The strategy the function-local static of Lattice<2>::Update instantiates for the MI355X target.
    template <>
    class SwendsenWang_2D<LCG32, MI355X_DEVICE> : public LatticeMonteCarloAlgorithm<2, LCG32, MI355X_DEVICE>
    {
        MI355X& target;

    public:
        static constexpr std::uint32_t Dimension = 2;

        explicit SwendsenWang_2D(MI355X& target) : target(target) {}

        void Update(Lattice<2>& lattice, const float temperature) override
        {
            const auto& extent = lattice.Extent();
            if (extent[0] % 2 || extent[1] % 2)   // swendsen_wang.cpp:47-48
                throw std::runtime_error("SwendsenWang_2D implementation requires evenly sized lattice.");
            target.Sweep(lattice, temperature);
        }
    };

    template <>
    template <>
    inline std::pair<double, double> Lattice<2>::GetEnergyAndMagnetization<MI355X>(MI355X& gpu)
    {
        std::int64_t energy = 0, magnetization = 0;
        if (!gpu.Observe(*this, energy, magnetization))   // no sweep on the GPU yet: the host spins are current
            return GetEnergyAndMagnetization(GetDevice<DeviceName::CPU>());
        return {-1.0 * energy / NumSites(), 1.0 * magnetization / NumSites()};
    }

    namespace mi355x
    {
        // GetDevice(name) of utility/device/helper.cpp:10-20 with "mi355x".
        inline std::unique_ptr<AbstractDevice> GetDevice(const std::string& device_name)
        {
            if (device_name == "mi355x")
                return std::make_unique<MI355X>();
            return cp::GetDevice(device_name);
        }

        // DispatchCall(name, device, func, args...) of utility/device/helper.hpp:16-29 with "mi355x".
        template <typename Func, typename... Args>
        auto DispatchCall(const std::string& device_name, AbstractDevice& device, Func&& func, Args&&... args)
        {
            if (device_name == "mi355x")
                return func(static_cast<MI355X&>(device), std::forward<Args>(args)...);
            return cp::DispatchCall(device_name, device, std::forward<Func>(func), std::forward<Args>(args)...);
        }
    }
}
