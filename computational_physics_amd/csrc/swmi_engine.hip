// swmi_engine.hip -- host side of the C ABI (include/swmi.h): engines, work lists, graphs, autotuning.
//
// Errors never cross the ABI as exceptions: every entry point returns a status and records a message
// for swmi_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "swmi_host.h"

namespace swmi
{
    namespace
    {
        thread_local std::string g_error;

        std::mutex g_tune_mutex;
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> g_tuned;   // (n0, n1) -> tile rows

        // Sweeps per captured graph of plain sweeps; a sweep loop of fewer sweeps runs eagerly. Loops of at least
        // two long units replay long graphs (each replay boundary costs a few microseconds of idle GPU), the
        // rest of the loop short ones.
        constexpr uint32_t kGraphSweeps = 16;
        constexpr uint32_t kLongGraphSweeps = 128;
        // Captured measurement graphs hold groups of (n_sep sweeps + sample) up to about this many sweeps.
        constexpr uint32_t kGraphGroupSweeps = 64;
        constexpr size_t kMaxGraphs = 6;
    }

    void SetLastError(const std::string& m) { g_error = m; }

    void Check(hipError_t e, const char* what)
    {
        if (e != hipSuccess)
            throw SwmiError(SWMI_ERR_HIP, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
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

    void ValidateConfig(const swmi_lattice_config& c)
    {
        const auto bad = [](const std::string& m) { throw SwmiError(SWMI_ERR_INVALID, m); };
        if (c.n0 < 2 || c.n1 < 2 || (c.n0 % 2) || (c.n1 % 2))
            bad("SwendsenWang_2D implementation requires evenly sized lattice (got " + std::to_string(c.n0) + " x " +
                std::to_string(c.n1) + ")");
        if (static_cast<uint64_t>(PitchOf(c.n0, c.n1)) * c.n1 >= (1ull << 31))
            bad("lattice too large: n0*n1 (with rows padded to " + std::to_string(PitchOf(c.n0, c.n1)) + " sites) must be < 2^31");
        // the kernels address a block's part of the per-site arrays through 32-bit buffer views
        const Geometry g = GeometryOf(c.n0, c.n1, 1);
        if (BlockSpanBytes(g) >= (1ull << 31))
            bad("lattice rows too long: " + std::to_string(c.n0) + " sites per row exceed the kernels' 32-bit block views");
        if (!(c.temperature > 0.0f) || !std::isfinite(c.temperature))
            bad("temperature must be positive and finite");
        if (c.flip_rule != SWMI_FLIP_PARITY && c.flip_rule != SWMI_FLIP_RANDOM)
            bad("unknown flip rule " + std::to_string(c.flip_rule));
        if (c.max_samples > (1u << 28))
            bad("max_samples too large");
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

    uint32_t DefaultTileRows(const uint32_t n0, const uint32_t n1)
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        const auto it = g_tuned.find({n0, n1});
        return it == g_tuned.end() ? 16u : it->second;
    }

    uint32_t TwrClass(const uint32_t twr)
    {
        return twr <= 8 ? 8u : (twr <= 16 ? 16u : (twr <= 32 ? 32u : 64u));
    }

    uint32_t PitchOf(const uint32_t n0, const uint32_t n1)
    {
        if (n0 % 64u == 0u)
            return n0;
        if (n0 > 64u)
            return (n0 + 63u) / 64u * 64u;
        for (uint32_t p = 2u; p < 64u; p *= 2u)
            if (p >= n0 && n1 % (64u / p) == 0u)
                return p;
        return 64u;
    }

    // Bytes of a block's part of the lcg32 state array (from its first to its last site), the largest
    // buffer view a kernel builds.
    uint64_t BlockSpanBytes(const Geometry& g)
    {
        const uint64_t rows = std::min<uint64_t>(static_cast<uint64_t>(kWavesPerBlock) * g.twr, g.nwr_y);
        return ((rows - 1u) * 64ull * g.tiles_x + 64u) * 4u;
    }

    Geometry GeometryOf(const uint32_t n0, const uint32_t n1, const uint32_t tile_rows)
    {
        Geometry g;
        g.pitch = PitchOf(n0, n1);
        g.tw = std::min<uint32_t>(g.pitch, 64);
        g.rpw = 64 / g.tw;
        g.tiles_x = g.pitch / g.tw;
        g.nwr_y = n1 / g.rpw;
        // last tile column: lanes lx < n0 - (tiles_x - 1) * tw of every lattice row hold sites
        const uint32_t lastw = n0 - (g.tiles_x - 1u) * g.tw;
        g.elast = lastw - 1u;
        const uint64_t row_mask = lastw == 64u ? ~0ull : ((1ull << lastw) - 1ull);
        g.vlast = 0;
        for (uint32_t j = 0; j < g.rpw; ++j)
            g.vlast |= row_mask << (j * g.tw);
        uint32_t twr = tile_rows ? tile_rows : DefaultTileRows(n0, n1);
        twr = std::min<uint32_t>(twr, 64u);
        twr = std::min<uint32_t>(twr, g.tw);      // lattice rows per tile <= 64 (right-edge bit mask)
        twr = std::min<uint32_t>(twr, g.nwr_y);
        g.twr = std::max<uint32_t>(twr, 1u);
        // blocks of very long rows: lower tiles until a block's buffer views fit 31 bits
        while (g.twr > 1u && BlockSpanBytes(g) >= (1ull << 31))
            g.twr /= 2u;
        g.blocks_y = (g.nwr_y + kWavesPerBlock * g.twr - 1) / (kWavesPerBlock * g.twr);
        g.n_blocks = g.tiles_x * g.blocks_y;
        // distinct clusters reaching a block's boundary <= boundary sites (2 columns + 2 rows)
        g.edge_cap = std::min<uint32_t>(g.twr * 64u, 2u * g.tw + 2u * kWavesPerBlock * g.twr * g.rpw);
        g.ecol_stride = kWavesPerBlock * g.twr * g.rpw;
        g.halo_stride = g.ecol_stride + 64u;
        g.qcap = kWavesPerBlock * g.twr * 64u;
        g.fbw = (g.qcap + kWavesPerBlock * g.edge_cap + 31u) / 32u;
        g.fsw = (kWavesPerBlock * g.edge_cap + 31u) / 32u;
        // the bond pass loads a block's slot flip bits one word per thread of its 256
        if (g.fsw > 256u)
            throw SwmiError(SWMI_ERR_INVALID, "block slot bitmap of " + std::to_string(g.fsw) + " words exceeds 256");
        return g;
    }

    LatDesc GeometryDesc(const Geometry& g, const uint32_t n0, const uint32_t n1)
    {
        LatDesc d{};
        d.edge_cap = g.edge_cap;
        d.ecol_stride = g.ecol_stride;
        d.n0 = g.pitch;
        d.n0v = n0;
        d.n1 = n1;
        d.elast = g.elast;
        d.vlast = g.vlast;
        d.tw = g.tw;
        d.rpw = g.rpw;
        d.twr = g.twr;
        d.tiles_x = g.tiles_x;
        d.blocks_y = g.blocks_y;
        d.nwr_y = g.nwr_y;
        d.n_blocks = g.n_blocks;
        d.halo_stride = g.halo_stride;
        d.qcap = g.qcap;
        d.fbw = g.fbw;
        d.fsw = g.fsw;
        return d;
    }

    uint64_t Lattice::ResidentBytes(const swmi_lattice_config& c)
    {
        const Geometry g = GeometryOf(c.n0, c.n1, c.tile_rows);
        const uint64_t n = static_cast<uint64_t>(g.pitch) * c.n1, groups = n / 64;
        return n * (1 + 4 + 2) + groups * (3 * 8 + 4) + static_cast<uint64_t>(kBbitWords) * 8 * g.n_blocks +
               4ull * (kLatStateWords + 4ull * c.max_samples) + 4ull * 2 * kWavesPerBlock * g.n_blocks +
               static_cast<uint64_t>(kWavesPerBlock) * g.n_blocks * (8ull * g.edge_cap + 4) + 8ull * g.ecol_stride * g.n_blocks +
               4ull * 2 * 64 * g.n_blocks + 4ull * g.fsw * g.n_blocks + 2ull * g.halo_stride * g.n_blocks;
    }

    bool Lattice::SameShape(const swmi_lattice_config& c) const
    {
        const Geometry h = GeometryOf(c.n0, c.n1, c.tile_rows);
        return cfg.n0 == c.n0 && cfg.n1 == c.n1 && g.twr == h.twr && cfg.max_samples == c.max_samples;
    }

    uint64_t Lattice::Bytes() const
    {
        return spins.bytes() + rng.bytes() + masks.bytes() + rot.bytes() + local.bytes() + parent.bytes() + cpar.bytes() +
               bbits.bytes() + state.bytes() + partials.bytes() + edge_count.bytes() + ecol.bytes() + erow.bytes() + fbits.bytes() +
               halo.bytes() + bond_dump.bytes() + bond_in.bytes();
    }

    void Lattice::Allocate(const swmi_lattice_config& c)
    {
        cfg = c;
        g = GeometryOf(c.n0, c.n1, c.tile_rows);
        n = static_cast<uint64_t>(c.n0) * c.n1;
        np = static_cast<uint64_t>(g.pitch) * c.n1;
        n_groups = static_cast<uint32_t>(np / 64);
        spins.allocate(np);
        rng.allocate(np);
        masks.allocate(3ull * n_groups);
        rot.allocate(n_groups);
        local.allocate(np);
        parent.reset();
        cpar.allocate(static_cast<size_t>(kWavesPerBlock) * g.n_blocks * g.edge_cap);
        bbits.allocate(static_cast<size_t>(kBbitWords) * g.n_blocks);
        state_words = kLatStateWords + 4u * c.max_samples;
        state.allocate(state_words);
        partials.allocate(2ull * kWavesPerBlock * g.n_blocks);
        edge_count.allocate(static_cast<size_t>(kWavesPerBlock) * g.n_blocks);
        ecol.allocate(2ull * g.ecol_stride * g.n_blocks);
        erow.allocate(2ull * 64 * g.n_blocks);
        fbits.allocate(static_cast<size_t>(g.fsw) * g.n_blocks);
        halo.allocate(2ull * g.halo_stride * g.n_blocks);
    }

    void Lattice::Initialize(const swmi_lattice_config& c, hipStream_t stream)
    {
        cfg.temperature = c.temperature;
        cfg.seed = c.seed;
        cfg.spin_seed = c.spin_seed;
        cfg.flip_rule = c.flip_rule;
        p_add = BondProbability(c.temperature);
        threshold = BondThreshold(p_add);
        bond_dump.reset();
        Check(hipMemsetAsync(edge_count.ptr, 0, edge_count.bytes(), stream), "edge list init");
        LaunchInitState(state.ptr, state_words, c.max_samples, 0xFFFFFFFFu, 0u, 0u, stream);
        if (np != n)   // padding columns hold spin 0
            Check(hipMemsetAsync(spins.ptr, 0, np, stream), "spin init");
        LaunchHotStart(spins.ptr, n, c.spin_seed, cfg.n0, g.pitch, stream);
        LaunchInitRng(spins.ptr, rng.ptr, masks.ptr, rot.ptr, n_groups, c.seed, stream);
        Check(hipMemsetAsync(local.ptr, 0, local.bytes(), stream), "label init");
        Check(hipGetLastError(), "initialisation launch");
    }
}

using namespace swmi;

swmi_engine::swmi_engine()
{
    // SWMI_GRAPHS=0 launches every kernel one by one (same results): for tools that trace each launch
    if (const char* g = std::getenv("SWMI_GRAPHS"))
        use_graphs = std::strcmp(g, "0") != 0;
    RequireGfx950(0);
    Check(hipSetDevice(0), "hipSetDevice");
    Check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
}

swmi_engine::~swmi_engine()
{
    if (stream)
        (void)hipStreamSynchronize(stream);
    for (auto s : side_streams)
        (void)hipStreamSynchronize(s);
    DropGraphs();
    for (auto s : side_streams)
        (void)hipStreamDestroy(s);
    for (auto e : join_events)
        (void)hipEventDestroy(e);
    if (fork_event)
        (void)hipEventDestroy(fork_event);
    if (stream)
        (void)hipStreamDestroy(stream);
}

void swmi_engine::EnsureSideStreams()
{
    if (!fork_event && classes.size() > 1)
        Check(hipEventCreateWithFlags(&fork_event, hipEventDisableTiming), "event create");
    while (side_streams.size() + 1 < classes.size())
    {
        hipStream_t s = nullptr;
        Check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        side_streams.push_back(s);
        hipEvent_t e = nullptr;
        Check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
        join_events.push_back(e);
    }
}

Lattice& swmi_engine::At(uint32_t i)
{
    if (i >= lattices.size() || !lattices[i])
        throw SwmiError(SWMI_ERR_INVALID, "lattice index " + std::to_string(i) + " out of range");
    return *lattices[i];
}

uint64_t swmi_engine::TotalSites() const
{
    uint64_t s = 0;
    for (const auto& l : lattices)
        if (l)
            s += l->n;
    return s;
}

LatDesc swmi_engine::Desc(const Lattice& l, uint32_t block_base) const
{
    LatDesc d = GeometryDesc(l.g, l.cfg.n0, l.cfg.n1);
    d.spins = l.spins.ptr;
    d.rng = l.rng.ptr;
    d.masks = l.masks.ptr;
    d.rot = l.rot.ptr;
    d.local = l.local.ptr;
    d.parent = l.parent.ptr;
    d.cpar = l.cpar.ptr;
    d.bbits = l.bbits.ptr;
    d.state = l.state.ptr;
    d.bond_dump = l.bond_dump.ptr;
    d.bond_in = nullptr;
    d.partials = l.partials.ptr;
    d.edge_count = l.edge_count.ptr;
    d.ecol = l.ecol.ptr;
    d.erow = l.erow.ptr;
    d.threshold = static_cast<uint32_t>(l.threshold);
    d.block_base = block_base;
    d.flip_rule = l.cfg.flip_rule;
    d.seed = l.cfg.seed;
    d.fbits = l.fbits.ptr;
    d.halo = l.halo.ptr;
    return d;
}

// Callers make sure no replay of these execs is in flight (the stream has been synchronised).
void swmi_engine::DropGraphs()
{
    for (auto& g : graphs)
        if (g.exec)
            (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

void swmi_engine::Rebuild()
{
    DropGraphs();
    // slots are compacted: the descriptor array has no holes
    lattices.erase(std::remove(lattices.begin(), lattices.end(), nullptr), lattices.end());
    // class of a lattice: (64-wide tiles, k_bond_label tile-height bound)
    std::vector<KernelClass> next;
    std::vector<LatDesc> d(lattices.size());
    for (uint32_t i = 0; i < lattices.size(); ++i)
    {
        const Lattice& l = *lattices[i];
        const bool tw64 = l.g.tw == 64;
        const bool pad = l.np != l.n;
        // padded lattices have two tile-height classes (16 and 64)
        const uint32_t t = pad ? (l.g.twr <= 16 ? 16u : 64u) : TwrClass(l.g.twr);
        auto it = std::find_if(next.begin(), next.end(),
                               [&](const KernelClass& k) { return k.tw64 == tw64 && k.pad == pad && k.twr_t == t; });
        if (it == next.end())
        {
            next.emplace_back();
            it = next.end() - 1;
            it->tw64 = tw64;
            it->pad = pad;
            it->twr_t = t;
        }
        d[i] = Desc(l, static_cast<uint32_t>(it->block_lat.size()));
        if (it->n_lat++ == 0)
            it->first_slot = i;
        it->block_lat.insert(it->block_lat.end(), l.g.n_blocks, i);
        it->max_twr = std::max(it->max_twr, l.g.twr);
    }
    // reuse the device work-list buffers of the previous classes where the sizes allow
    for (auto& k : next)
    {
        k.total_blocks = static_cast<uint32_t>(k.block_lat.size());
        auto old = std::find_if(classes.begin(), classes.end(), [&](KernelClass& o) {
            return o.d_block_lat.ptr && o.d_block_lat.count >= k.block_lat.size();
        });
        if (old != classes.end())
        {
            std::swap(k.d_block_lat.ptr, old->d_block_lat.ptr);
            std::swap(k.d_block_lat.count, old->d_block_lat.count);
        }
        else
            k.d_block_lat.allocate(std::max<size_t>(k.block_lat.size(), 1));
    }
    descs.swap(d);
    if (d_desc.count < std::max<size_t>(descs.size(), 1))
        d_desc.allocate(std::max<size_t>(descs.size(), 1));
    classes.swap(next);   // the old buffers die with `next` (after the stream has used them: hipFree synchronises)
    if (!descs.empty())
        Check(hipMemcpyAsync(d_desc.ptr, descs.data(), descs.size() * sizeof(LatDesc), hipMemcpyHostToDevice, stream),
              "descriptor upload");
    for (auto& k : classes)
        Check(hipMemcpyAsync(k.d_block_lat.ptr, k.block_lat.data(), k.block_lat.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                             stream),
              "work list upload");
    EnsureSideStreams();
}

void swmi_engine::LaunchCheck(const char* what) { Check(hipGetLastError(), what); }

void swmi_engine::EnqueueSweeps(uint32_t n)
{
    if (n == 0 || classes.empty())
        return;
    const size_t nc = classes.size();
    if (nc > 1)
    {
        Check(hipEventRecord(fork_event, stream), "fork");
        for (size_t c = 1; c < nc; ++c)
            Check(hipStreamWaitEvent(side_streams[c - 1], fork_event, 0), "fork");
    }
    for (size_t c = 0; c < nc; ++c)
    {
        const KernelClass& k = classes[c];
        hipStream_t s = ClassStream(c);
        for (uint32_t i = 0; i < n; ++i)
        {
            LaunchBondLabel(k.Descs(d_desc.ptr), k.WorkList(), k.total_blocks, k.max_twr, k.tw64, k.pad, recording, s);
            LaunchMerge(k.Descs(d_desc.ptr), k.WorkList(), k.total_blocks, s);
            LaunchResolve(k.Descs(d_desc.ptr), k.WorkList(), k.total_blocks, s);
        }
    }
    for (size_t c = 1; c < nc; ++c)
    {
        Check(hipEventRecord(join_events[c - 1], side_streams[c - 1]), "join");
        Check(hipStreamWaitEvent(stream, join_events[c - 1], 0), "join");
    }
    LaunchCheck("sweep launch");
}

void swmi_engine::EnqueueMaterialize()
{
    for (auto& k : classes)
        LaunchMaterialize(k.Descs(d_desc.ptr), k.WorkList(), k.total_blocks, stream);
    LaunchCheck("flip launch");
}

void swmi_engine::EnqueueSample(uint32_t mode)
{
    if (descs.empty())
        return;
    EnqueueMaterialize();
    for (auto& k : classes)
        LaunchObservables(k.Descs(d_desc.ptr), k.WorkList(), k.total_blocks, mode, stream);
    LaunchFinalizeSamples(d_desc.ptr, static_cast<uint32_t>(descs.size()), mode, stream);
    LaunchCheck("sample launch");
}

void swmi_engine::Run(uint32_t n, uint32_t sep, int32_t sample_mode)
{
    if (classes.empty() || n == 0)
        return;
    const bool sampled = sample_mode >= 0;
    if (sampled && sep > kGraphGroupSweeps)
    {
        // long groups: graphs of plain sweeps, eager samples
        for (uint32_t i = 0; i < n; i += sep)
        {
            Run(sep, 0, -1);
            EnqueueSample(static_cast<uint32_t>(sample_mode));
        }
        return;
    }
    // one replay unit: `unit` sweeps, or `groups` groups of (sep sweeps + sample)
    const uint32_t groups = sampled ? std::max<uint32_t>(1u, kGraphGroupSweeps / sep) : 1u;
    const uint32_t unit = sampled ? groups * sep : (n >= 2u * kLongGraphSweeps ? kLongGraphSweeps : kGraphSweeps);
    const uint32_t total_groups = sampled ? (n + sep - 1) / sep : 0u;   // ising_2d.cpp:132: i += N_Sep while i < n
    const uint32_t replays = sampled ? total_groups / groups : n / unit;
    const auto enqueue_unit = [&](uint32_t g_count) {
        if (sampled)
            for (uint32_t g = 0; g < g_count; ++g)
            {
                EnqueueSweeps(sep);
                EnqueueSample(static_cast<uint32_t>(sample_mode));
            }
        else
            EnqueueSweeps(g_count);
    };
    if (use_graphs && replays >= 2)
    {
        GraphEntry* entry = nullptr;
        for (auto& g : graphs)
            if (g.sweeps == unit && g.sample == (sampled ? sample_mode : -1) && g.sep == (sampled ? sep : 0u))
                entry = &g;
        if (!entry)
        {
            if (graphs.size() >= kMaxGraphs)
            {
                // earlier replays of these execs may still run on the stream (the side-stream branches
                // are joined into it inside each graph): destroy them only once they have finished
                Check(hipStreamSynchronize(stream), "graph eviction");
                DropGraphs();
            }
            hipGraph_t graph = nullptr;
            Check(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "graph capture");
            try
            {
                enqueue_unit(sampled ? groups : unit);
            }
            catch (...)
            {
                (void)hipStreamEndCapture(stream, &graph);
                if (graph)
                    (void)hipGraphDestroy(graph);
                throw;
            }
            Check(hipStreamEndCapture(stream, &graph), "graph capture");
            GraphEntry ge;
            ge.sweeps = unit;
            ge.sample = sampled ? sample_mode : -1;
            ge.sep = sampled ? sep : 0u;
            const hipError_t err = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            Check(err, "graph instantiate");
            graphs.push_back(ge);
            entry = &graphs.back();
        }
        for (uint32_t r = 0; r < replays; ++r)
            Check(hipGraphLaunch(entry->exec, stream), "graph launch");
        if (sampled)
            enqueue_unit(total_groups - replays * groups);
        else
            Run(n - replays * unit, 0, -1);   // short graphs or eager sweeps
        return;
    }
    enqueue_unit(sampled ? total_groups : n);
}

// Descriptor of lattice l alone (block_base 0) for the inspection hooks, which launch with a null work list.
void swmi_engine::OneLattice(const Lattice& l, const uint8_t* bond_in)
{
    LatDesc d = Desc(l, 0);
    d.bond_in = bond_in;
    if (!d_one_desc.ptr)
        d_one_desc.allocate(1);
    Check(hipMemcpyAsync(d_one_desc.ptr, &d, sizeof(d), hipMemcpyHostToDevice, stream), "descriptor upload");
    Check(hipStreamSynchronize(stream), "descriptor upload");
}

void swmi_engine::MaterializeOne(const Lattice& l)
{
    OneLattice(l, nullptr);
    LaunchMaterialize(d_one_desc.ptr, nullptr, l.g.n_blocks, stream);
    LaunchCheck("flip launch");
}

namespace
{
    // Row-major host values of a lattice's sites from its padded per-site array (row pitch g.pitch).
    template <typename T, typename F>
    void UnpadRows(const Lattice& l, const T* padded, T* out, F&& f)
    {
        const uint64_t n0 = l.cfg.n0, pitch = l.g.pitch;
        for (uint64_t y = 0; y < l.cfg.n1; ++y)
            for (uint64_t x = 0; x < n0; ++x)
                out[y * n0 + x] = f(padded[y * pitch + x]);
    }

    swmi_engine* Eng(swmi_engine* e)
    {
        if (!e)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        return e;
    }

    // Average device time of a plain sweep of an engine holding these lattices (HIP events).
    float TimeSweeps(const swmi_lattice_config* cs, const uint32_t n, const uint32_t sweeps)
    {
        swmi_engine e;
        for (uint32_t i = 0; i < n; ++i)
        {
            e.lattices.push_back(std::make_unique<Lattice>());
            e.lattices.back()->Allocate(cs[i]);
            e.lattices.back()->Initialize(cs[i], e.stream);
        }
        e.Rebuild();
        e.Run(4, 0, -1);
        hipEvent_t a, b;
        Check(hipEventCreate(&a), "event");
        Check(hipEventCreate(&b), "event");
        Check(hipEventRecord(a, e.stream), "event");
        e.Run(sweeps, 0, -1);
        Check(hipEventRecord(b, e.stream), "event");
        Check(hipEventSynchronize(b), "event");
        float ms = 0;
        Check(hipEventElapsedTime(&ms, a, b), "event");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return ms / sweeps;
    }
}

extern "C" {

const char* swmi_last_error(void) { return g_error.c_str(); }

const char* swmi_version(void) { return "swmi 0.2 (gfx950)"; }

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
        auto e = std::make_unique<swmi_engine>();
        for (uint32_t i = 0; i < n_lattices; ++i)
        {
            e->lattices.push_back(std::make_unique<Lattice>());
            e->lattices.back()->Allocate(configs[i]);
            e->lattices.back()->Initialize(configs[i], e->stream);
        }
        e->Rebuild();
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
    return Guard([&] { Check(hipStreamSynchronize(Eng(engine)->stream), "synchronize"); });
}

int swmi_upload_spins(swmi_engine* engine, uint32_t lattice, const uint8_t* spins)
{
    return Guard([&] {
        if (!spins)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = Eng(engine)->At(lattice);
        for (uint64_t i = 0; i < l.n; ++i)
            if (spins[i] > 1)
                throw SwmiError(SWMI_ERR_INVALID, "spins must be 0 or 1");
        DeviceBuffer<uint8_t> tmp;
        tmp.allocate(l.n);
        Check(hipMemcpyAsync(tmp.ptr, spins, l.n, hipMemcpyHostToDevice, engine->stream), "spin upload");
        LaunchSetSpins(l.spins.ptr, tmp.ptr, l.n, l.cfg.n0, l.g.pitch, engine->stream);
        LaunchSetApplied(l.state.ptr, engine->stream);   // the uploaded spins are the chain's state
        engine->LaunchCheck("spin upload");
        Check(hipStreamSynchronize(engine->stream), "spin upload");
    });
}

int swmi_download_spins(swmi_engine* engine, uint32_t lattice, uint8_t* spins)
{
    return Guard([&] {
        if (!spins)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = Eng(engine)->At(lattice);
        engine->MaterializeOne(l);
        std::vector<uint8_t> buf(l.np);
        Check(hipMemcpyAsync(buf.data(), l.spins.ptr, l.np, hipMemcpyDeviceToHost, engine->stream), "spin download");
        Check(hipStreamSynchronize(engine->stream), "spin download");
        UnpadRows(l, buf.data(), spins, [](uint8_t v) { return static_cast<uint8_t>(v & 1u); });   // bits 3-5: selectors
    });
}

int swmi_set_temperature(swmi_engine* engine, uint32_t lattice, float temperature)
{
    return Guard([&] {
        Lattice& l = Eng(engine)->At(lattice);
        if (!(temperature > 0.0f) || !std::isfinite(temperature))
            throw SwmiError(SWMI_ERR_INVALID, "temperature must be positive and finite");
        Check(hipStreamSynchronize(engine->stream), "set temperature");
        l.cfg.temperature = temperature;
        l.p_add = BondProbability(temperature);
        l.threshold = BondThreshold(l.p_add);
        engine->Rebuild();
        Check(hipStreamSynchronize(engine->stream), "set temperature");
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

int swmi_get_layout(const swmi_engine* engine, uint32_t lattice, uint32_t* pitch, uint32_t* n_groups)
{
    return Guard([&] {
        if (!engine)
            throw SwmiError(SWMI_ERR_INVALID, "null engine");
        if (lattice >= engine->lattices.size())
            throw SwmiError(SWMI_ERR_INVALID, "lattice index out of range");
        const Lattice& l = *engine->lattices[lattice];
        if (pitch)
            *pitch = l.g.pitch;
        if (n_groups)
            *n_groups = l.n_groups;
    });
}

int swmi_get_tile_rows(const swmi_engine* engine, uint32_t lattice, uint32_t* tile_rows)
{
    return Guard([&] {
        if (!engine || !tile_rows)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        if (lattice >= engine->lattices.size())
            throw SwmiError(SWMI_ERR_INVALID, "lattice index out of range");
        *tile_rows = engine->lattices[lattice]->g.twr;
    });
}

int swmi_sweep(swmi_engine* engine, uint32_t n_sweeps)
{
    return Guard([&] { Eng(engine)->Run(n_sweeps, 0, -1); });
}

int swmi_sample(swmi_engine* engine)
{
    return Guard([&] { Eng(engine)->EnqueueSample(kSampleAlways); });
}

int swmi_run(swmi_engine* engine, uint32_t n_sweeps, uint32_t n_sep)
{
    return Guard([&] {
        if (n_sep == 0)
            throw SwmiError(SWMI_ERR_INVALID, "n_sep == 0");
        // ising/src/ising_2d.cpp:132-146: n_sep sweeps, then a sample, while i < n_sweeps
        Eng(engine)->Run(n_sweeps, n_sep, kSampleAlways);
    });
}

int swmi_set_schedule(swmi_engine* engine, uint32_t lattice, uint32_t sweeps, uint32_t warmup, uint32_t n_sep)
{
    return Guard([&] {
        Lattice& l = Eng(engine)->At(lattice);
        if (n_sep == 0 || (sweeps && warmup > sweeps) || sweeps > (1u << 30))
            throw SwmiError(SWMI_ERR_INVALID, "schedule needs n_sep > 0, warmup <= sweeps and sweeps <= 2^30");
        LaunchSetSchedule(l.state.ptr, sweeps, warmup, n_sep, engine->stream);
        engine->LaunchCheck("schedule launch");
    });
}

int swmi_run_scheduled(swmi_engine* engine, uint32_t n_sweeps, uint32_t n_sep)
{
    return Guard([&] {
        if (n_sep == 0)
            throw SwmiError(SWMI_ERR_INVALID, "n_sep == 0");
        Eng(engine)->Run(n_sweeps, n_sep, kSampleDue);
    });
}

int swmi_set_graphs(swmi_engine* engine, int enable)
{
    return Guard([&] {
        Check(hipStreamSynchronize(Eng(engine)->stream), "set graphs");
        engine->use_graphs = enable != 0;
        engine->DropGraphs();
    });
}

int swmi_observables(swmi_engine* engine, uint32_t lattice, int64_t* energy, int64_t* magnetization)
{
    return Guard([&] {
        Lattice& l = Eng(engine)->At(lattice);
        engine->MaterializeOne(l);
        LaunchObservables(engine->d_one_desc.ptr, nullptr, l.g.n_blocks, kSampleAlways, engine->stream);
        engine->LaunchCheck("observables launch");
        std::vector<int32_t> p(2ull * kWavesPerBlock * l.g.n_blocks);
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
        if (!moments)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = Eng(engine)->At(lattice);
        LatState st{};
        Check(hipMemcpyAsync(&st, l.state.ptr, sizeof(st), hipMemcpyDeviceToHost, engine->stream), "moments");
        Check(hipStreamSynchronize(engine->stream), "moments");
        std::memcpy(moments, st.moments, sizeof(st.moments));
    });
}

int swmi_get_progress(swmi_engine* engine, uint32_t lattice, uint32_t* sweeps_done, uint32_t* n_samples)
{
    return Guard([&] {
        Lattice& l = Eng(engine)->At(lattice);
        LatState st{};
        Check(hipMemcpyAsync(&st, l.state.ptr, sizeof(st), hipMemcpyDeviceToHost, engine->stream), "progress");
        Check(hipStreamSynchronize(engine->stream), "progress");
        if (sweeps_done)
            *sweeps_done = st.iteration / 2u;
        if (n_samples)
            *n_samples = st.n_samples;
    });
}

int swmi_get_series(swmi_engine* engine, uint32_t lattice, int64_t* energy, int64_t* magnetization, uint32_t max, uint32_t* count)
{
    return Guard([&] {
        Lattice& l = Eng(engine)->At(lattice);
        LatState st{};
        Check(hipMemcpyAsync(&st, l.state.ptr, sizeof(st), hipMemcpyDeviceToHost, engine->stream), "series");
        Check(hipStreamSynchronize(engine->stream), "series");
        const uint32_t stored = std::min(st.n_samples, st.series_cap);
        const uint32_t n = std::min(stored, max);
        if (count)
            *count = stored;
        if (n == 0)
            return;
        std::vector<int64_t> buf(2ull * n);
        Check(hipMemcpyAsync(buf.data(), l.state.ptr + kLatStateWords, buf.size() * 8, hipMemcpyDeviceToHost, engine->stream), "series");
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
        for (auto& l : Eng(engine)->lattices)
        {
            if (!l)
                continue;
            Check(hipMemsetAsync(reinterpret_cast<uint8_t*>(l->state.ptr) + 4 * kStSamples, 0, 4, engine->stream), "reset");
            Check(hipMemsetAsync(reinterpret_cast<uint8_t*>(l->state.ptr) + 4 * kStMoments, 0, sizeof(LatState::moments), engine->stream),
                  "reset");
        }
        Check(hipStreamSynchronize(engine->stream), "reset");
    });
}

int swmi_record_bonds(swmi_engine* engine, int enable)
{
    return Guard([&] {
        Check(hipStreamSynchronize(Eng(engine)->stream), "record bonds");
        for (auto& l : engine->lattices)
        {
            if (!l)
                continue;
            if (enable && !l->bond_dump.ptr)
                l->bond_dump.allocate(l->np);
            if (!enable)
                l->bond_dump.reset();
        }
        engine->recording = enable != 0;
        engine->Rebuild();
        Check(hipStreamSynchronize(engine->stream), "record bonds");
    });
}

int swmi_download_bonds(swmi_engine* engine, uint32_t lattice, uint8_t* bonds)
{
    return Guard([&] {
        if (!bonds)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = Eng(engine)->At(lattice);
        if (!l.bond_dump.ptr)
            throw SwmiError(SWMI_ERR_STATE, "bond recording is not enabled");
        std::vector<uint8_t> buf(l.np);
        Check(hipMemcpyAsync(buf.data(), l.bond_dump.ptr, l.np, hipMemcpyDeviceToHost, engine->stream), "bond download");
        Check(hipStreamSynchronize(engine->stream), "bond download");
        UnpadRows(l, buf.data(), bonds, [](uint8_t v) { return v; });
    });
}

int swmi_label_bonds(swmi_engine* engine, uint32_t lattice, const uint8_t* bonds)
{
    return Guard([&] {
        if (!bonds)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = Eng(engine)->At(lattice);
        for (uint64_t i = 0; i < l.n; ++i)
            if (bonds[i] > 3)
                throw SwmiError(SWMI_ERR_INVALID, "bond bytes must be 0..3");
        if (!l.bond_in.ptr)
            l.bond_in.allocate(l.np);
        if (l.np == l.n)
            Check(hipMemcpyAsync(l.bond_in.ptr, bonds, l.n, hipMemcpyHostToDevice, engine->stream), "bond upload");
        else
        {
            DeviceBuffer<uint8_t> tmp;
            tmp.allocate(l.n);
            Check(hipMemcpyAsync(tmp.ptr, bonds, l.n, hipMemcpyHostToDevice, engine->stream), "bond upload");
            LaunchPadBytes(l.bond_in.ptr, tmp.ptr, l.np, l.cfg.n0, l.g.pitch, engine->stream);
            Check(hipStreamSynchronize(engine->stream), "bond upload");
        }
        engine->MaterializeOne(l);   // the injected pass replaces the labels a pending flip needs
        if (!l.parent.ptr)
            l.parent.allocate(l.np);
        engine->OneLattice(l, l.bond_in.ptr);
        LaunchBondLabel(engine->d_one_desc.ptr, nullptr, l.g.n_blocks, l.g.twr, l.g.tw == 64, l.np != l.n, true,
                        engine->stream);
        LaunchMerge(engine->d_one_desc.ptr, nullptr, l.g.n_blocks, engine->stream);
        LaunchWriteLabels(engine->d_one_desc.ptr, nullptr, l.g.n_blocks, l.g.twr, engine->stream);
        engine->LaunchCheck("label launch");
        Check(hipStreamSynchronize(engine->stream), "label");
    });
}

int swmi_download_labels(swmi_engine* engine, uint32_t lattice, uint32_t* labels)
{
    return Guard([&] {
        if (!labels)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = Eng(engine)->At(lattice);
        if (!l.parent.ptr)
            l.parent.allocate(l.np);
        engine->OneLattice(l, nullptr);
        LaunchWriteLabels(engine->d_one_desc.ptr, nullptr, l.g.n_blocks, l.g.twr, engine->stream);
        engine->LaunchCheck("resolve launch");
        if (l.np == l.n)
            Check(hipMemcpyAsync(labels, l.parent.ptr, l.n * 4, hipMemcpyDeviceToHost, engine->stream), "label download");
        else
        {
            DeviceBuffer<uint32_t> tmp;
            tmp.allocate(l.n);
            LaunchUnpadLabels(tmp.ptr, l.parent.ptr, l.n, l.cfg.n0, l.g.pitch, engine->stream);
            engine->LaunchCheck("label unpad launch");
            Check(hipMemcpyAsync(labels, tmp.ptr, l.n * 4, hipMemcpyDeviceToHost, engine->stream), "label download");
            Check(hipStreamSynchronize(engine->stream), "label download");
        }
        Check(hipStreamSynchronize(engine->stream), "label download");
    });
}

int swmi_rng_export(swmi_engine* engine, uint32_t lattice, uint32_t* states, uint64_t* masks, uint32_t* rot, uint32_t* iteration)
{
    return Guard([&] {
        Lattice& l = Eng(engine)->At(lattice);
        if (states)
            Check(hipMemcpyAsync(states, l.rng.ptr, l.np * 4, hipMemcpyDeviceToHost, engine->stream), "rng export");
        if (masks)
            Check(hipMemcpyAsync(masks, l.masks.ptr, 3ull * l.n_groups * 8, hipMemcpyDeviceToHost, engine->stream), "rng export");
        if (rot)
            Check(hipMemcpyAsync(rot, l.rot.ptr, 4ull * l.n_groups, hipMemcpyDeviceToHost, engine->stream), "rng export");
        if (iteration)
            Check(hipMemcpyAsync(iteration, l.state.ptr, 4, hipMemcpyDeviceToHost, engine->stream), "rng export");
        Check(hipStreamSynchronize(engine->stream), "rng export");
    });
}

namespace
{
    struct CheckpointHeader
    {
        char magic[8];           // "SWMICKP1"
        uint32_t n0, n1;
        uint32_t seed, flip_rule;
        float temperature;
        uint32_t state_words;
        uint64_t n;              // padded sites (row pitch * n1)
        uint32_t n_groups;
        uint32_t reserved[5];
    };
    static_assert(sizeof(CheckpointHeader) == 64, "checkpoint header is 64 bytes");

    uint64_t CheckpointBytes(const Lattice& l)
    {
        return sizeof(CheckpointHeader) + l.np * (1 + 4) + l.n_groups * (3ull * 8 + 4) + 4ull * l.state_words;
    }
}

int swmi_checkpoint_size(swmi_engine* engine, uint32_t lattice, uint64_t* bytes)
{
    return Guard([&] {
        if (!bytes)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        *bytes = CheckpointBytes(Eng(engine)->At(lattice));
    });
}

int swmi_checkpoint(swmi_engine* engine, uint32_t lattice, void* buffer, uint64_t bytes)
{
    return Guard([&] {
        Lattice& l = Eng(engine)->At(lattice);
        if (!buffer || bytes < CheckpointBytes(l))
            throw SwmiError(SWMI_ERR_INVALID, "checkpoint buffer too small (" + std::to_string(CheckpointBytes(l)) + " bytes needed)");
        engine->MaterializeOne(l);   // the checkpoint holds the chain's spins, with no flip pending
        CheckpointHeader h{};
        std::memcpy(h.magic, "SWMICKP1", 8);
        h.n0 = l.cfg.n0;
        h.n1 = l.cfg.n1;
        h.seed = l.cfg.seed;
        h.flip_rule = l.cfg.flip_rule;
        h.temperature = l.cfg.temperature;
        h.state_words = l.state_words;
        h.n = l.np;
        h.n_groups = l.n_groups;
        uint8_t* out = static_cast<uint8_t*>(buffer);
        std::memcpy(out, &h, sizeof(h));
        uint8_t* q = out + sizeof(h);
        const auto get = [&](const void* src, size_t nb) {
            Check(hipMemcpyAsync(q, src, nb, hipMemcpyDeviceToHost, engine->stream), "checkpoint");
            q += nb;
        };
        get(l.spins.ptr, l.np);
        get(l.rng.ptr, 4 * l.np);
        get(l.masks.ptr, 24ull * l.n_groups);
        get(l.rot.ptr, 4ull * l.n_groups);
        get(l.state.ptr, 4ull * l.state_words);
        Check(hipStreamSynchronize(engine->stream), "checkpoint");
    });
}

int swmi_restore(swmi_engine* engine, uint32_t lattice, const void* buffer, uint64_t bytes)
{
    return Guard([&] {
        Lattice& l = Eng(engine)->At(lattice);
        if (!buffer || bytes < sizeof(CheckpointHeader))
            throw SwmiError(SWMI_ERR_INVALID, "checkpoint buffer too small");
        CheckpointHeader h{};
        std::memcpy(&h, buffer, sizeof(h));
        if (std::memcmp(h.magic, "SWMICKP1", 8) != 0)
            throw SwmiError(SWMI_ERR_INVALID, "not an swmi checkpoint");
        if (h.n0 != l.cfg.n0 || h.n1 != l.cfg.n1 || h.state_words != l.state_words || h.n != l.np || h.n_groups != l.n_groups)
            throw SwmiError(SWMI_ERR_INVALID, "checkpoint of a " + std::to_string(h.n0) + " x " + std::to_string(h.n1) +
                                                  " lattice (series words " + std::to_string(h.state_words) +
                                                  ") does not fit this lattice");
        if (bytes < CheckpointBytes(l))
            throw SwmiError(SWMI_ERR_INVALID, "checkpoint truncated");
        swmi_lattice_config c = l.cfg;
        c.seed = h.seed;
        c.flip_rule = h.flip_rule;
        c.temperature = h.temperature;
        ValidateConfig(c);
        Check(hipStreamSynchronize(engine->stream), "restore");
        const uint8_t* q = static_cast<const uint8_t*>(buffer) + sizeof(h);
        const auto put = [&](void* dst, size_t nb) {
            Check(hipMemcpyAsync(dst, q, nb, hipMemcpyHostToDevice, engine->stream), "restore");
            q += nb;
        };
        put(l.spins.ptr, l.np);
        put(l.rng.ptr, 4 * l.np);
        put(l.masks.ptr, 24ull * l.n_groups);
        put(l.rot.ptr, 4ull * l.n_groups);
        put(l.state.ptr, 4ull * l.state_words);
        LaunchSetApplied(l.state.ptr, engine->stream);   // a checkpoint's spins hold the chain's state
        Check(hipStreamSynchronize(engine->stream), "restore");
        l.cfg.seed = c.seed;
        l.cfg.flip_rule = c.flip_rule;
        l.cfg.temperature = c.temperature;
        l.p_add = BondProbability(c.temperature);
        l.threshold = BondThreshold(l.p_add);
        engine->Rebuild();
        Check(hipStreamSynchronize(engine->stream), "restore");
    });
}

int swmi_error_bars(swmi_engine* engine, uint32_t lattice, uint32_t n_bins, double* out)
{
    return Guard([&] {
        if (!out || n_bins < 2)
            throw SwmiError(SWMI_ERR_INVALID, "null output or fewer than 2 bins");
        Lattice& l = Eng(engine)->At(lattice);
        LatState st{};
        Check(hipMemcpyAsync(&st, l.state.ptr, sizeof(st), hipMemcpyDeviceToHost, engine->stream), "error bars");
        Check(hipStreamSynchronize(engine->stream), "error bars");
        const uint32_t stored = std::min(st.n_samples, st.series_cap);
        const uint32_t per = stored / n_bins;
        if (per == 0)
            throw SwmiError(SWMI_ERR_STATE, "not enough samples in the series for " + std::to_string(n_bins) + " bins");
        const uint32_t used = per * n_bins;
        std::vector<int64_t> ser(2ull * used);
        Check(hipMemcpyAsync(ser.data(), l.state.ptr + kLatStateWords, ser.size() * 8, hipMemcpyDeviceToHost, engine->stream),
              "error bars");
        Check(hipStreamSynchronize(engine->stream), "error bars");
        const double n = static_cast<double>(l.n), t = l.cfg.temperature;
        std::vector<double> be(n_bins, 0.0), be2(n_bins, 0.0), bm(n_bins, 0.0);
        double s1 = 0, s2 = 0, sm = 0;
        for (uint32_t i = 0; i < used; ++i)
        {
            const double e = -static_cast<double>(ser[2 * i]) / n, m = std::fabs(static_cast<double>(ser[2 * i + 1]) / n);
            be[i / per] += e;
            be2[i / per] += e * e;
            bm[i / per] += m;
            s1 += e;
            s2 += e * e;
            sm += m;
        }
        const double mean_e = s1 / used, mean_m = sm / used;
        double var_be = 0, var_bm = 0;
        for (uint32_t b = 0; b < n_bins; ++b)
        {
            var_be += (be[b] / per - mean_e) * (be[b] / per - mean_e);
            var_bm += (bm[b] / per - mean_m) * (bm[b] / per - mean_m);
        }
        var_be /= (n_bins - 1);
        var_bm /= (n_bins - 1);
        const double var_e = s2 / used - mean_e * mean_e;
        const double c_all = var_e * n / (t * t);
        // jackknife over bins for C/N
        double cj_mean = 0;
        std::vector<double> cj(n_bins);
        for (uint32_t b = 0; b < n_bins; ++b)
        {
            const double m1 = (s1 - be[b]) / (used - per), m2 = (s2 - be2[b]) / (used - per);
            cj[b] = (m2 - m1 * m1) * n / (t * t);
            cj_mean += cj[b] / n_bins;
        }
        double cj_var = 0;
        for (uint32_t b = 0; b < n_bins; ++b)
            cj_var += (cj[b] - cj_mean) * (cj[b] - cj_mean);
        out[0] = mean_e;
        out[1] = std::sqrt(var_be / n_bins);
        out[2] = mean_m;
        out[3] = std::sqrt(var_bm / n_bins);
        out[4] = c_all;
        out[5] = std::sqrt(cj_var * (n_bins - 1) / n_bins);
        out[6] = var_e > 0 ? 0.5 * var_be * per / var_e : 0.0;
        out[7] = used;
    });
}

int swmi_get_error_flags(swmi_engine* engine, uint32_t lattice, uint32_t* flags)
{
    return Guard([&] {
        if (!flags)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        Lattice& l = Eng(engine)->At(lattice);
        Check(hipMemcpyAsync(flags, l.state.ptr + kStErr, 4, hipMemcpyDeviceToHost, engine->stream), "error flags");
        Check(hipStreamSynchronize(engine->stream), "error flags");
    });
}

int swmi_time_kernels(swmi_engine* engine, uint32_t n_sweeps, float* ms)
{
    return Guard([&] {
        if (!ms || n_sweeps == 0)
            throw SwmiError(SWMI_ERR_INVALID, "null argument or zero sweeps");
        Eng(engine);
        hipEvent_t ev[4];
        for (auto& e : ev)
            Check(hipEventCreate(&e), "event create");
        double total[3] = {0, 0, 0};
        for (uint32_t i = 0; i < n_sweeps; ++i)
        {
            Check(hipEventRecord(ev[0], engine->stream), "event");
            for (auto& k : engine->classes)
                LaunchBondLabel(k.Descs(engine->d_desc.ptr), k.WorkList(), k.total_blocks, k.max_twr, k.tw64, k.pad, engine->recording,
                                engine->stream);
            Check(hipEventRecord(ev[1], engine->stream), "event");
            for (auto& k : engine->classes)
                LaunchMerge(k.Descs(engine->d_desc.ptr), k.WorkList(), k.total_blocks, engine->stream);
            Check(hipEventRecord(ev[2], engine->stream), "event");
            for (auto& k : engine->classes)
                LaunchResolve(k.Descs(engine->d_desc.ptr), k.WorkList(), k.total_blocks, engine->stream);
            Check(hipEventRecord(ev[3], engine->stream), "event");
            Check(hipEventSynchronize(ev[3]), "event sync");
            for (int k = 0; k < 3; ++k)
            {
                float t = 0;
                Check(hipEventElapsedTime(&t, ev[k], ev[k + 1]), "event elapsed");
                total[k] += t;
            }
        }
        engine->LaunchCheck("timed sweep launch");
        for (auto& e : ev)
            (void)hipEventDestroy(e);
        for (int k = 0; k < 3; ++k)
            ms[k] = static_cast<float>(total[k] / n_sweeps);
    });
}

int swmi_autotune(const swmi_lattice_config* config, uint32_t sweeps, uint32_t* best_tile_rows, float* ms_per_sweep)
{
    return Guard([&] {
        if (!config || sweeps == 0)
            throw SwmiError(SWMI_ERR_INVALID, "null config or zero sweeps");
        ValidateConfig(*config);
        constexpr uint32_t kCand[7] = {1, 2, 4, 8, 16, 32, 64};
        uint32_t best = 0, prev = 0;
        float best_ms = 0.0f;
        for (uint32_t i = 0; i < 7; ++i)
        {
            if (ms_per_sweep)
                ms_per_sweep[i] = -1.0f;
            const uint32_t twr = GeometryOf(config->n0, config->n1, kCand[i]).twr;
            if (twr == prev)
                continue;   // clamped to a height already measured
            prev = twr;
            swmi_lattice_config c = *config;
            c.tile_rows = twr;
            c.max_samples = 0;
            const float ms = TimeSweeps(&c, 1, sweeps);
            if (ms_per_sweep)
                ms_per_sweep[i] = ms;
            if (best == 0 || ms < best_ms)
            {
                best = twr;
                best_ms = ms;
            }
        }
        {
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            g_tuned[{config->n0, config->n1}] = best;
        }
        if (best_tile_rows)
            *best_tile_rows = best;
    });
}

int swmi_autotune_batch(const swmi_lattice_config* configs, uint32_t n_lattices, uint32_t sweeps, uint32_t* best_tile_rows,
                        float* ms_per_sweep)
{
    return Guard([&] {
        if (!configs || n_lattices == 0 || sweeps == 0)
            throw SwmiError(SWMI_ERR_INVALID, "null configs, no lattices or zero sweeps");
        for (uint32_t i = 0; i < n_lattices; ++i)
            ValidateConfig(configs[i]);
        constexpr uint32_t kCand[7] = {1, 2, 4, 8, 16, 32, 64};
        std::vector<swmi_lattice_config> cs(configs, configs + n_lattices);
        uint32_t best = 0;
        float best_ms = 0.0f;
        for (uint32_t i = 0; i < 7; ++i)
        {
            for (auto& c : cs)
            {
                c.tile_rows = kCand[i];
                c.max_samples = 0;
            }
            const float ms = TimeSweeps(cs.data(), n_lattices, sweeps);
            if (ms_per_sweep)
                ms_per_sweep[i] = ms;
            if (best == 0 || ms < best_ms)
            {
                best = kCand[i];
                best_ms = ms;
            }
        }
        if (best_tile_rows)
            *best_tile_rows = best;
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
        LaunchCopy(src.ptr, dst.ptr, n16, nullptr);
        Check(hipEventRecord(a, nullptr), "event");
        for (uint32_t i = 0; i < repeats; ++i)
            LaunchCopy(src.ptr, dst.ptr, n16, nullptr);
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
