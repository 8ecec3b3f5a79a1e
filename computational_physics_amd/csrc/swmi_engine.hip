// swmi_engine.hip -- host side of the C ABI (include/swmi.h).
//
// An engine owns a set of lattices (one for the plain path, many for a batch), their device buffers,
// one HIP stream and the block work list over all of them. Each sweep is three launches over the whole
// work list, so a batch of lattices of any sizes costs the same number of launches as one lattice.
// Errors never cross the ABI as exceptions: every entry point returns a status and records a message
// for swmi_last_error().
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "swmi_internal.h"
#include "../../include/swmi.h"

namespace
{
    thread_local std::string g_error;

    struct SwmiError : std::runtime_error
    {
        int code;
        SwmiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
    };

    void Check(hipError_t e, const char* what)
    {
        if (e != hipSuccess)
            throw SwmiError(SWMI_ERR_HIP, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
    }

    template <typename F>
    int Guard(F&& f)
    {
        try
        {
            f();
            return SWMI_OK;
        }
        catch (const SwmiError& e)
        {
            g_error = e.what();
            return e.code;
        }
        catch (const std::exception& e)
        {
            g_error = e.what();
            return SWMI_ERR_INVALID;
        }
    }

    // p_add exactly as the reference computes it (ising/src/swendsen_wang.cpp:100), all in float with
    // glibc expf, and the smallest draw u for which 2^-32f * (float)u < p_add fails.
    float BondProbability(const float temperature)
    {
        const float q = -2.0f / temperature;
        const float e = std::exp(q);
        return 1.0f - e;
    }

    uint64_t BondThreshold(const float p_add)
    {
        uint64_t lo = 0, hi = 1ull << 32;
        while (lo < hi)
        {
            const uint64_t mid = lo + (hi - lo) / 2;
            const bool bond = mid < (1ull << 32) && (2.3283064370807974E-10F * static_cast<float>(static_cast<uint32_t>(mid))) < p_add;
            if (bond)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo;
    }

    template <typename T>
    struct DeviceBuffer
    {
        T* ptr = nullptr;
        size_t count = 0;
        DeviceBuffer() = default;
        DeviceBuffer(const DeviceBuffer&) = delete;
        DeviceBuffer& operator=(const DeviceBuffer&) = delete;
        ~DeviceBuffer() { reset(); }
        void allocate(size_t n)
        {
            reset();
            if (n == 0)
                return;
            if (hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)) != hipSuccess)
            {
                ptr = nullptr;
                (void)hipGetLastError();
                throw SwmiError(SWMI_ERR_MEMORY, "device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed");
            }
            count = n;
        }
        void reset()
        {
            if (ptr)
                (void)hipFree(ptr);
            ptr = nullptr;
            count = 0;
        }
    };

    struct Lattice
    {
        swmi_lattice_config cfg{};
        uint64_t n = 0;
        uint32_t n_groups = 0;
        uint32_t tw = 0, rpw = 0, twr = 0, tiles_x = 0, blocks_y = 0, nwr_y = 0, n_blocks = 0, edge_cap = 0;
        float p_add = 0.0f;
        uint64_t threshold = 0;
        DeviceBuffer<uint8_t> spins;
        DeviceBuffer<uint32_t> rng;
        DeviceBuffer<uint64_t> masks;
        DeviceBuffer<uint32_t> rot;
        DeviceBuffer<uint16_t> local;
        DeviceBuffer<uint32_t> parent;
        DeviceBuffer<uint64_t> bbits;
        DeviceBuffer<uint32_t> state;
        DeviceBuffer<int32_t> partials;
        DeviceBuffer<uint16_t> edge_list;
        DeviceBuffer<uint32_t> edge_count;
        DeviceBuffer<uint8_t> bond_dump;
        DeviceBuffer<uint8_t> bond_in;
    };
}

struct swmi_engine
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<Lattice>> lattices;
    std::vector<swmi::LatDesc> descs;
    DeviceBuffer<swmi::LatDesc> d_desc;
    DeviceBuffer<uint32_t> d_block_lat;
    uint32_t total_blocks = 0;
    uint32_t max_twr = 0;
    bool all_tw64 = true;
    bool recording = false;   // some lattice records its bonds (selects the inspection build of k_bond_label)
    // single-lattice work lists for the inspection hooks
    DeviceBuffer<swmi::LatDesc> d_one_desc;
    DeviceBuffer<uint32_t> d_zero_blocks;

    ~swmi_engine()
    {
        if (stream)
        {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    Lattice& At(uint32_t i)
    {
        if (i >= lattices.size())
            throw SwmiError(SWMI_ERR_INVALID, "lattice index " + std::to_string(i) + " out of range");
        return *lattices[i];
    }

    swmi::LatDesc Desc(const Lattice& l, uint32_t block_base) const
    {
        swmi::LatDesc d{};
        d.spins = l.spins.ptr;
        d.rng = l.rng.ptr;
        d.masks = l.masks.ptr;
        d.rot = l.rot.ptr;
        d.local = l.local.ptr;
        d.parent = l.parent.ptr;
        d.bbits = l.bbits.ptr;
        d.state = l.state.ptr;
        d.bond_dump = l.bond_dump.ptr;
        d.bond_in = nullptr;
        d.partials = l.partials.ptr;
        d.edge_list = l.edge_list.ptr;
        d.edge_count = l.edge_count.ptr;
        d.edge_cap = l.edge_cap;
        d.n0 = l.cfg.n0;
        d.n1 = l.cfg.n1;
        d.tw = l.tw;
        d.rpw = l.rpw;
        d.twr = l.twr;
        d.tiles_x = l.tiles_x;
        d.blocks_y = l.blocks_y;
        d.nwr_y = l.nwr_y;
        d.threshold = static_cast<uint32_t>(l.threshold);
        d.block_base = block_base;
        d.n_blocks = l.n_blocks;
        d.flip_rule = l.cfg.flip_rule;
        return d;
    }

    void UploadDescriptors()
    {
        descs.clear();
        uint32_t base = 0;
        max_twr = 1;
        all_tw64 = true;
        std::vector<uint32_t> block_lat;
        for (uint32_t i = 0; i < lattices.size(); ++i)
        {
            const Lattice& l = *lattices[i];
            descs.push_back(Desc(l, base));
            block_lat.insert(block_lat.end(), l.n_blocks, i);
            base += l.n_blocks;
            max_twr = std::max(max_twr, l.twr);
            all_tw64 = all_tw64 && l.tw == 64;
        }
        total_blocks = base;
        if (d_desc.count != descs.size())
            d_desc.allocate(descs.size());
        if (d_block_lat.count != block_lat.size())
            d_block_lat.allocate(block_lat.size());
        Check(hipMemcpyAsync(d_desc.ptr, descs.data(), descs.size() * sizeof(swmi::LatDesc), hipMemcpyHostToDevice, stream),
              "descriptor upload");
        Check(hipMemcpyAsync(d_block_lat.ptr, block_lat.data(), block_lat.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream),
              "work list upload");
        Check(hipStreamSynchronize(stream), "descriptor upload");
    }

    void LaunchCheck(const char* what) { Check(hipGetLastError(), what); }

    void Sweeps(uint32_t n)
    {
        for (uint32_t i = 0; i < n; ++i)
        {
            swmi::LaunchBondLabel(d_desc.ptr, d_block_lat.ptr, total_blocks, max_twr, all_tw64, recording, stream);
            swmi::LaunchMerge(d_desc.ptr, d_block_lat.ptr, total_blocks, stream);
            swmi::LaunchFlip(d_desc.ptr, d_block_lat.ptr, total_blocks, max_twr, swmi::kFlip | swmi::kAdvance, stream);
        }
        LaunchCheck("sweep launch");
    }

    void Sample()
    {
        swmi::LaunchObservables(d_desc.ptr, d_block_lat.ptr, total_blocks, stream);
        swmi::LaunchFinalizeSamples(d_desc.ptr, static_cast<uint32_t>(lattices.size()), stream);
        LaunchCheck("sample launch");
    }

    // Work list holding only lattice i (block_base 0), for the inspection hooks.
    void OneLattice(const Lattice& l, const uint8_t* bond_in)
    {
        swmi::LatDesc d = Desc(l, 0);
        d.bond_in = bond_in;
        if (!d_one_desc.ptr)
            d_one_desc.allocate(1);
        if (d_zero_blocks.count < l.n_blocks)
        {
            d_zero_blocks.allocate(l.n_blocks);
            Check(hipMemsetAsync(d_zero_blocks.ptr, 0, l.n_blocks * sizeof(uint32_t), stream), "work list");
        }
        Check(hipMemcpyAsync(d_one_desc.ptr, &d, sizeof(d), hipMemcpyHostToDevice, stream), "descriptor upload");
        Check(hipStreamSynchronize(stream), "descriptor upload");
    }
};

namespace
{
    void ValidateConfig(const swmi_lattice_config& c)
    {
        const auto bad = [](const std::string& m) { throw SwmiError(SWMI_ERR_INVALID, m); };
        if (c.n0 < 2 || c.n1 < 2 || (c.n0 % 2) || (c.n1 % 2))
            bad("SwendsenWang_2D implementation requires evenly sized lattice (got " + std::to_string(c.n0) + " x " +
                std::to_string(c.n1) + ")");
        if (!((c.n0 % 64 == 0) || (64 % c.n0 == 0)))
            bad("extent[0] must be a multiple of 64 or a divisor of 64 (got " + std::to_string(c.n0) + ")");
        if (c.n0 < 64 && (c.n1 % (64 / c.n0)))
            bad("n0*n1 must be a multiple of 64 (got " + std::to_string(c.n0) + " x " + std::to_string(c.n1) + ")");
        if (static_cast<uint64_t>(c.n0) * c.n1 >= (1ull << 31))
            bad("lattice too large: n0*n1 must be < 2^31");
        if (!(c.temperature > 0.0f) || !std::isfinite(c.temperature))
            bad("temperature must be positive and finite");
        if (c.flip_rule != SWMI_FLIP_PARITY)
            bad("unknown flip rule " + std::to_string(c.flip_rule));
    }

    void SetupLattice(Lattice& l, const swmi_lattice_config& c, hipStream_t stream)
    {
        l.cfg = c;
        l.n = static_cast<uint64_t>(c.n0) * c.n1;
        l.n_groups = static_cast<uint32_t>(l.n / 64);
        l.tw = std::min<uint32_t>(c.n0, 64);
        l.rpw = 64 / l.tw;
        l.tiles_x = c.n0 / l.tw;
        l.nwr_y = c.n1 / l.rpw;
        uint32_t twr = c.tile_rows ? c.tile_rows : 16u;
        twr = std::min<uint32_t>(twr, 64u);
        twr = std::min<uint32_t>(twr, l.tw);      // lattice rows per tile <= 64 (right-edge bit mask)
        twr = std::min<uint32_t>(twr, l.nwr_y);
        l.twr = std::max<uint32_t>(twr, 1u);
        l.blocks_y = (l.nwr_y + swmi::kWavesPerBlock * l.twr - 1) / (swmi::kWavesPerBlock * l.twr);
        l.n_blocks = l.tiles_x * l.blocks_y;
        // distinct clusters reaching a block's boundary <= boundary sites (2 columns + 2 rows)
        l.edge_cap = std::min<uint32_t>(l.twr * 64u, 2u * l.tw + 2u * swmi::kWavesPerBlock * l.twr * l.rpw);
        l.p_add = BondProbability(c.temperature);
        l.threshold = BondThreshold(l.p_add);

        l.spins.allocate(l.n);
        l.rng.allocate(l.n);
        l.masks.allocate(3ull * l.n_groups);
        l.rot.allocate(l.n_groups);
        l.local.allocate(l.n);
        l.parent.allocate(l.n);
        l.bbits.allocate(static_cast<size_t>(swmi::kBbitWords) * l.n_blocks);
        const size_t state_words = swmi::kLatStateWords + 4ull * c.max_samples;
        l.state.allocate(state_words);
        l.partials.allocate(2ull * swmi::kWavesPerBlock * l.n_blocks);
        l.edge_list.allocate(static_cast<size_t>(swmi::kWavesPerBlock) * l.n_blocks * l.edge_cap);
        l.edge_count.allocate(static_cast<size_t>(swmi::kWavesPerBlock) * l.n_blocks);
        Check(hipMemsetAsync(l.edge_count.ptr, 0, l.edge_count.count * 4, stream), "edge list init");
        Check(hipMemsetAsync(l.state.ptr, 0, state_words * 4, stream), "state init");
        Check(hipMemcpyAsync(reinterpret_cast<uint8_t*>(l.state.ptr) + 8, &c.max_samples, 4, hipMemcpyHostToDevice, stream),
              "state init");
        swmi::LaunchHotStart(l.spins.ptr, l.n, c.spin_seed, stream);
        swmi::LaunchInitRng(l.spins.ptr, l.rng.ptr, l.masks.ptr, l.rot.ptr, l.n_groups, c.seed, stream);
        swmi::LaunchIota(l.parent.ptr, l.n, stream);
        Check(hipMemsetAsync(l.local.ptr, 0, l.n * sizeof(uint16_t), stream), "label init");
        Check(hipGetLastError(), "initialisation launch");
    }

    void RequireGfx950(int device)
    {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= device)
        {
            (void)hipGetLastError();
            throw SwmiError(SWMI_ERR_NO_DEVICE, "no HIP device available (the MI355X engine needs a gfx950 GPU)");
        }
        hipDeviceProp_t prop{};
        Check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw SwmiError(SWMI_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", the engine is built for gfx950");
    }
}

extern "C" {

const char* swmi_last_error(void) { return g_error.c_str(); }

const char* swmi_version(void) { return "swmi 0.1 (gfx950)"; }

int swmi_get_device_info(swmi_device_info* out)
{
    return Guard([&] {
        if (!out)
            throw SwmiError(SWMI_ERR_INVALID, "null output");
        RequireGfx950(0);
        hipDeviceProp_t prop{};
        Check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
        std::memset(out, 0, sizeof(*out));
        std::snprintf(out->name, sizeof(out->name), "%s", prop.name);
        std::snprintf(out->arch, sizeof(out->arch), "%s", prop.gcnArchName);
        out->compute_units = static_cast<uint32_t>(prop.multiProcessorCount);
        out->wavefront_size = static_cast<uint32_t>(prop.warpSize);
        out->total_memory = prop.totalGlobalMem;
        out->l2_bytes = static_cast<uint32_t>(prop.l2CacheSize);
        out->clock_khz = static_cast<uint32_t>(prop.clockRate);
    });
}

int swmi_create(const swmi_lattice_config* configs, uint32_t n_lattices, swmi_engine** out)
{
    return Guard([&] {
        if (!configs || !out || n_lattices == 0)
            throw SwmiError(SWMI_ERR_INVALID, "swmi_create needs at least one lattice config and an output handle");
        *out = nullptr;
        for (uint32_t i = 0; i < n_lattices; ++i)
            ValidateConfig(configs[i]);
        RequireGfx950(0);
        Check(hipSetDevice(0), "hipSetDevice");
        auto e = std::make_unique<swmi_engine>();
        Check(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (uint32_t i = 0; i < n_lattices; ++i)
        {
            e->lattices.push_back(std::make_unique<Lattice>());
            SetupLattice(*e->lattices.back(), configs[i], e->stream);
        }
        e->UploadDescriptors();
        Check(hipStreamSynchronize(e->stream), "initialisation");
        *out = e.release();
    });
}

int swmi_destroy(swmi_engine* engine)
{
    return Guard([&] { delete engine; });
}

int swmi_num_lattices(const swmi_engine* engine, uint32_t* out)
{
    return Guard([&] {
        if (!engine || !out)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        *out = static_cast<uint32_t>(engine->lattices.size());
    });
}

int swmi_synchronize(swmi_engine* engine)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        Check(hipStreamSynchronize(engine->stream), "synchronize");
    });
}

int swmi_upload_spins(swmi_engine* engine, uint32_t lattice, const uint8_t* spins)
{
    return Guard([&] {
        if (!engine || !spins)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = engine->At(lattice);
        for (uint64_t i = 0; i < l.n; ++i)
            if (spins[i] > 1)
                throw SwmiError(SWMI_ERR_INVALID, "spins must be 0 or 1");
        DeviceBuffer<uint8_t> tmp;
        tmp.allocate(l.n);
        Check(hipMemcpyAsync(tmp.ptr, spins, l.n, hipMemcpyHostToDevice, engine->stream), "spin upload");
        swmi::LaunchSetSpins(l.spins.ptr, tmp.ptr, l.n, engine->stream);
        engine->LaunchCheck("spin upload");
        Check(hipStreamSynchronize(engine->stream), "spin upload");
    });
}

int swmi_download_spins(swmi_engine* engine, uint32_t lattice, uint8_t* spins)
{
    return Guard([&] {
        if (!engine || !spins)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = engine->At(lattice);
        Check(hipMemcpyAsync(spins, l.spins.ptr, l.n, hipMemcpyDeviceToHost, engine->stream), "spin download");
        Check(hipStreamSynchronize(engine->stream), "spin download");
        for (uint64_t i = 0; i < l.n; ++i)   // bits 1-3 carry the lcg32 selectors
            spins[i] &= 1u;
    });
}

int swmi_set_temperature(swmi_engine* engine, uint32_t lattice, float temperature)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        Lattice& l = engine->At(lattice);
        if (!(temperature > 0.0f) || !std::isfinite(temperature))
            throw SwmiError(SWMI_ERR_INVALID, "temperature must be positive and finite");
        l.cfg.temperature = temperature;
        l.p_add = BondProbability(temperature);
        l.threshold = BondThreshold(l.p_add);
        Check(hipStreamSynchronize(engine->stream), "set temperature");
        engine->UploadDescriptors();
    });
}

int swmi_get_threshold(const swmi_engine* engine, uint32_t lattice, float* p_add, uint64_t* threshold)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        if (lattice >= engine->lattices.size())
            throw SwmiError(SWMI_ERR_INVALID, "lattice index out of range");
        const Lattice& l = *engine->lattices[lattice];
        if (p_add)
            *p_add = l.p_add;
        if (threshold)
            *threshold = l.threshold;
    });
}

int swmi_sweep(swmi_engine* engine, uint32_t n_sweeps)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        engine->Sweeps(n_sweeps);
    });
}

int swmi_sample(swmi_engine* engine)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        engine->Sample();
    });
}

int swmi_run(swmi_engine* engine, uint32_t n_sweeps, uint32_t n_sep)
{
    return Guard([&] {
        if (!engine || n_sep == 0)
            throw SwmiError(SWMI_ERR_INVALID, "null engine or n_sep == 0");
        // ising/src/ising_2d.cpp:132-146: n_sep sweeps, then a sample, while i < n_sweeps
        for (uint32_t i = 0; i < n_sweeps; i += n_sep)
        {
            engine->Sweeps(n_sep);
            engine->Sample();
        }
    });
}

int swmi_observables(swmi_engine* engine, uint32_t lattice, int64_t* energy, int64_t* magnetization)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        Lattice& l = engine->At(lattice);
        swmi::LaunchObservables(engine->d_desc.ptr, engine->d_block_lat.ptr, engine->total_blocks, engine->stream);
        engine->LaunchCheck("observables launch");
        std::vector<int32_t> p(2ull * swmi::kWavesPerBlock * l.n_blocks);
        Check(hipMemcpyAsync(p.data(), l.partials.ptr, p.size() * 4, hipMemcpyDeviceToHost, engine->stream), "observables");
        Check(hipStreamSynchronize(engine->stream), "observables");
        int64_t e = 0, m = 0;
        for (size_t i = 0; i < p.size() / 2; ++i)
        {
            e += p[2 * i];
            m += p[2 * i + 1];
        }
        if (energy)
            *energy = e;
        if (magnetization)
            *magnetization = m;
    });
}

int swmi_get_moments(swmi_engine* engine, uint32_t lattice, double* moments)
{
    return Guard([&] {
        if (!engine || !moments)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = engine->At(lattice);
        swmi::LatState st{};
        Check(hipMemcpyAsync(&st, l.state.ptr, sizeof(st), hipMemcpyDeviceToHost, engine->stream), "moments");
        Check(hipStreamSynchronize(engine->stream), "moments");
        std::memcpy(moments, st.moments, sizeof(st.moments));
    });
}

int swmi_get_series(swmi_engine* engine, uint32_t lattice, int64_t* energy, int64_t* magnetization, uint32_t max, uint32_t* count)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        Lattice& l = engine->At(lattice);
        swmi::LatState st{};
        Check(hipMemcpyAsync(&st, l.state.ptr, sizeof(st), hipMemcpyDeviceToHost, engine->stream), "series");
        Check(hipStreamSynchronize(engine->stream), "series");
        const uint32_t stored = std::min(st.n_samples, st.series_cap);
        const uint32_t n = std::min(stored, max);
        if (count)
            *count = stored;
        if (n == 0)
            return;
        std::vector<int64_t> buf(2ull * n);
        Check(hipMemcpyAsync(buf.data(), l.state.ptr + swmi::kLatStateWords, buf.size() * 8, hipMemcpyDeviceToHost, engine->stream),
              "series");
        Check(hipStreamSynchronize(engine->stream), "series");
        for (uint32_t i = 0; i < n; ++i)
        {
            if (energy)
                energy[i] = buf[2 * i];
            if (magnetization)
                magnetization[i] = buf[2 * i + 1];
        }
    });
}

int swmi_reset_measurements(swmi_engine* engine)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        for (auto& l : engine->lattices)
        {
            // n_samples (word 1) and the six FP64 moments (bytes 16..63)
            Check(hipMemsetAsync(reinterpret_cast<uint8_t*>(l->state.ptr) + 4, 0, 4, engine->stream), "reset");
            Check(hipMemsetAsync(reinterpret_cast<uint8_t*>(l->state.ptr) + 16, 0, 48, engine->stream), "reset");
        }
        Check(hipStreamSynchronize(engine->stream), "reset");
    });
}

int swmi_record_bonds(swmi_engine* engine, int enable)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        Check(hipStreamSynchronize(engine->stream), "record bonds");
        for (auto& l : engine->lattices)
        {
            if (enable && !l->bond_dump.ptr)
                l->bond_dump.allocate(l->n);
            if (!enable)
                l->bond_dump.reset();
        }
        engine->recording = enable != 0;
        engine->UploadDescriptors();
    });
}

int swmi_download_bonds(swmi_engine* engine, uint32_t lattice, uint8_t* bonds)
{
    return Guard([&] {
        if (!engine || !bonds)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = engine->At(lattice);
        if (!l.bond_dump.ptr)
            throw SwmiError(SWMI_ERR_STATE, "bond recording is not enabled");
        Check(hipMemcpyAsync(bonds, l.bond_dump.ptr, l.n, hipMemcpyDeviceToHost, engine->stream), "bond download");
        Check(hipStreamSynchronize(engine->stream), "bond download");
    });
}

int swmi_label_bonds(swmi_engine* engine, uint32_t lattice, const uint8_t* bonds)
{
    return Guard([&] {
        if (!engine || !bonds)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = engine->At(lattice);
        if (!l.bond_in.ptr)
            l.bond_in.allocate(l.n);
        Check(hipMemcpyAsync(l.bond_in.ptr, bonds, l.n, hipMemcpyHostToDevice, engine->stream), "bond upload");
        engine->OneLattice(l, l.bond_in.ptr);
        swmi::LaunchBondLabel(engine->d_one_desc.ptr, engine->d_zero_blocks.ptr, l.n_blocks, l.twr, l.tw == 64, true, engine->stream);
        swmi::LaunchMerge(engine->d_one_desc.ptr, engine->d_zero_blocks.ptr, l.n_blocks, engine->stream);
        swmi::LaunchFlip(engine->d_one_desc.ptr, engine->d_zero_blocks.ptr, l.n_blocks, l.twr, swmi::kWriteLabels, engine->stream);
        engine->LaunchCheck("label launch");
        Check(hipStreamSynchronize(engine->stream), "label");
    });
}

int swmi_download_labels(swmi_engine* engine, uint32_t lattice, uint32_t* labels)
{
    return Guard([&] {
        if (!engine || !labels)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = engine->At(lattice);
        engine->OneLattice(l, nullptr);
        swmi::LaunchFlip(engine->d_one_desc.ptr, engine->d_zero_blocks.ptr, l.n_blocks, l.twr, swmi::kWriteLabels, engine->stream);
        engine->LaunchCheck("resolve launch");
        Check(hipMemcpyAsync(labels, l.parent.ptr, l.n * 4, hipMemcpyDeviceToHost, engine->stream), "label download");
        Check(hipStreamSynchronize(engine->stream), "label download");
    });
}

int swmi_rng_export(swmi_engine* engine, uint32_t lattice, uint32_t* states, uint64_t* masks, uint32_t* rot, uint32_t* iteration)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        Lattice& l = engine->At(lattice);
        if (states)
            Check(hipMemcpyAsync(states, l.rng.ptr, l.n * 4, hipMemcpyDeviceToHost, engine->stream), "rng export");
        if (masks)
            Check(hipMemcpyAsync(masks, l.masks.ptr, 3ull * l.n_groups * 8, hipMemcpyDeviceToHost, engine->stream), "rng export");
        if (rot)
            Check(hipMemcpyAsync(rot, l.rot.ptr, 4ull * l.n_groups, hipMemcpyDeviceToHost, engine->stream), "rng export");
        if (iteration)
            Check(hipMemcpyAsync(iteration, l.state.ptr, 4, hipMemcpyDeviceToHost, engine->stream), "rng export");
        Check(hipStreamSynchronize(engine->stream), "rng export");
    });
}

int swmi_get_error_flags(swmi_engine* engine, uint32_t lattice, uint32_t* flags)
{
    return Guard([&] {
        if (!engine || !flags)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = engine->At(lattice);
        Check(hipMemcpyAsync(flags, l.state.ptr + 3, 4, hipMemcpyDeviceToHost, engine->stream), "error flags");
        Check(hipStreamSynchronize(engine->stream), "error flags");
    });
}

int swmi_time_kernels(swmi_engine* engine, uint32_t n_sweeps, float* ms)
{
    return Guard([&] {
        if (!engine || !ms || n_sweeps == 0)
            throw SwmiError(SWMI_ERR_INVALID, "null argument or zero sweeps");
        hipEvent_t ev[4];
        for (auto& e : ev)
            Check(hipEventCreate(&e), "event create");
        double total[3] = {0, 0, 0};
        for (uint32_t i = 0; i < n_sweeps; ++i)
        {
            Check(hipEventRecord(ev[0], engine->stream), "event");
            swmi::LaunchBondLabel(engine->d_desc.ptr, engine->d_block_lat.ptr, engine->total_blocks, engine->max_twr, engine->all_tw64,
                                  engine->recording, engine->stream);
            Check(hipEventRecord(ev[1], engine->stream), "event");
            swmi::LaunchMerge(engine->d_desc.ptr, engine->d_block_lat.ptr, engine->total_blocks, engine->stream);
            Check(hipEventRecord(ev[2], engine->stream), "event");
            swmi::LaunchFlip(engine->d_desc.ptr, engine->d_block_lat.ptr, engine->total_blocks, engine->max_twr, swmi::kFlip | swmi::kAdvance, engine->stream);
            Check(hipEventRecord(ev[3], engine->stream), "event");
            Check(hipEventSynchronize(ev[3]), "event sync");
            for (int k = 0; k < 3; ++k)
            {
                float t = 0;
                Check(hipEventElapsedTime(&t, ev[k], ev[k + 1]), "event elapsed");
                total[k] += t;
            }
        }
        for (auto& e : ev)
            (void)hipEventDestroy(e);
        for (int k = 0; k < 3; ++k)
            ms[k] = static_cast<float>(total[k] / n_sweeps);
    });
}

int swmi_copy_calibration(size_t bytes, uint32_t repeats, float* ms)
{
    return Guard([&] {
        if (bytes < 16 || repeats == 0)
            throw SwmiError(SWMI_ERR_INVALID, "bytes >= 16 and repeats > 0 required");
        RequireGfx950(0);
        Check(hipSetDevice(0), "hipSetDevice");
        const uint64_t n16 = bytes / 16;
        DeviceBuffer<uint4> src, dst;
        src.allocate(n16);
        dst.allocate(n16);
        Check(hipMemset(src.ptr, 1, n16 * 16), "memset");
        hipEvent_t a, b;
        Check(hipEventCreate(&a), "event");
        Check(hipEventCreate(&b), "event");
        swmi::LaunchCopy(src.ptr, dst.ptr, n16, nullptr);
        Check(hipEventRecord(a, nullptr), "event");
        for (uint32_t i = 0; i < repeats; ++i)
            swmi::LaunchCopy(src.ptr, dst.ptr, n16, nullptr);
        Check(hipEventRecord(b, nullptr), "event");
        Check(hipEventSynchronize(b), "event");
        float t = 0;
        Check(hipEventElapsedTime(&t, a, b), "event");
        if (ms)
            *ms = t / repeats;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    });
}

}  // extern "C"
