// swmi_host.h -- host-side objects of the engine, shared by the engine (swmi_engine.hip) and the
// serving loop (swmi_server.cpp). Not part of the C ABI.
//
// An engine owns a set of lattices, their device buffers, one HIP stream and the block work lists over
// them. Lattices are grouped into kernel classes by the k_bond_label specialisation they need (64-wide
// tiles or not, tile-height bound 8/16/32/64); a sweep is, for every class, one k_bond_label, one k_merge
// and one k_resolve launch over the class's work list, so a batch of lattices of any sizes costs a handful
// of launches per sweep. Repeated launch sequences are captured once as hipGraphs and replayed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "swmi_internal.h"
#include "../../include/swmi.h"

namespace swmi
{
    struct SwmiError : std::runtime_error
    {
        int code;
        SwmiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
    };

    void Check(hipError_t e, const char* what);
    void SetLastError(const std::string& m);

    // Runs f; exceptions become a status code and the thread's last error message.
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
            SetLastError(e.what());
            return e.code;
        }
        catch (const std::exception& e)
        {
            SetLastError(e.what());
            return SWMI_ERR_INVALID;
        }
    }

    float BondProbability(float temperature);
    uint64_t BondThreshold(float p_add);
    void ValidateConfig(const swmi_lattice_config& c);
    void RequireGfx950(int device);

    template <typename T>
    struct DeviceBuffer
    {
        T* ptr = nullptr;
        size_t count = 0;
        DeviceBuffer() = default;
        DeviceBuffer(const DeviceBuffer&) = delete;
        DeviceBuffer& operator=(const DeviceBuffer&) = delete;
        DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), count(o.count)
        {
            o.ptr = nullptr;
            o.count = 0;
        }
        DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
        {
            if (this != &o)
            {
                reset();
                ptr = o.ptr;
                count = o.count;
                o.ptr = nullptr;
                o.count = 0;
            }
            return *this;
        }
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
        size_t bytes() const { return count * sizeof(T); }
    };

    // Tile geometry of a lattice (see swmi_internal.h for the meaning of the fields). pitch: the padded row
    // width; elast, vlast: the last tile column's last-site lane and site lanes.
    struct Geometry
    {
        uint32_t pitch = 0, tw = 0, rpw = 0, twr = 0, tiles_x = 0, blocks_y = 0, nwr_y = 0, n_blocks = 0, edge_cap = 0,
                 ecol_stride = 0, halo_stride = 0, elast = 0, qcap = 0, fbw = 0, fsw = 0;
        uint64_t vlast = 0;
    };
    // Row pitch of an n0 x n1 lattice: n0 when the wave64 tiling takes it as it is, else the next width it
    // takes (the next multiple of 64 above 64; below, the smallest power of two p >= n0 with n1 % (64/p) == 0).
    uint32_t PitchOf(uint32_t n0, uint32_t n1);
    // Bytes of a block's part of the lcg32 state array, the largest buffer view the kernels build.
    uint64_t BlockSpanBytes(const Geometry& g);
    Geometry GeometryOf(uint32_t n0, uint32_t n1, uint32_t tile_rows);
    // Descriptor with the geometry fields of a lattice filled in (buffers null).
    LatDesc GeometryDesc(const Geometry& g, uint32_t n0, uint32_t n1);
    // Tile height for tile_rows == 0: the autotuned value for this extent if one was measured in this
    // process, else 16 wave rows.
    uint32_t DefaultTileRows(uint32_t n0, uint32_t n1);
    // k_bond_label specialisation bound for a tile height: 8, 16, 32 or 64.
    uint32_t TwrClass(uint32_t twr);

    struct Lattice
    {
        swmi_lattice_config cfg{};
        uint64_t n = 0;            // sites, n0 * n1
        uint64_t np = 0;           // padded sites, pitch * n1 (the per-site arrays hold np entries)
        uint32_t n_groups = 0;     // np / 64
        Geometry g;
        float p_add = 0.0f;
        uint64_t threshold = 0;
        uint32_t state_words = 0;
        DeviceBuffer<uint8_t> spins;
        DeviceBuffer<uint32_t> rng;
        DeviceBuffer<uint64_t> masks;
        DeviceBuffer<uint32_t> rot;
        DeviceBuffer<uint16_t> local;
        DeviceBuffer<uint32_t> parent;       // inspection labels, allocated by the first call that needs them
        DeviceBuffer<uint64_t> cpar;
        DeviceBuffer<uint64_t> bbits;
        DeviceBuffer<uint32_t> state;
        DeviceBuffer<int32_t> partials;
        DeviceBuffer<uint32_t> edge_count;
        DeviceBuffer<uint32_t> ecol;
        DeviceBuffer<uint32_t> erow;
        DeviceBuffer<uint32_t> fbits;
        DeviceBuffer<uint8_t> halo;
        DeviceBuffer<uint8_t> bond_dump;
        DeviceBuffer<uint8_t> bond_in;

        // Device bytes of a lattice of this configuration (what Allocate takes).
        static uint64_t ResidentBytes(const swmi_lattice_config& c);
        void Allocate(const swmi_lattice_config& c);
        // Hot start, lcg32 initialisation and fresh counters, enqueued on `stream` (no host sync, no
        // host memory read after return).
        void Initialize(const swmi_lattice_config& c, hipStream_t stream);
        bool SameShape(const swmi_lattice_config& c) const;
        uint64_t Bytes() const;
    };

    struct KernelClass
    {
        bool tw64 = true;
        bool pad = false;                    // padded lattices (the PAD specialisation of k_bond_label)
        uint32_t twr_t = 16;                 // k_bond_label TWR_MAX
        uint32_t max_twr = 1;
        std::vector<uint32_t> block_lat;     // block -> lattice slot
        DeviceBuffer<uint32_t> d_block_lat;
        uint32_t total_blocks = 0;
        uint32_t n_lat = 0;                  // lattices in the class
        uint32_t first_slot = 0;             // descriptor slot of the class's first lattice
        // The kernels' arguments for this class: a class of one lattice passes that lattice's descriptor and a
        // null work list (the kernels then skip the work-list load), a larger class all descriptors and its list.
        const LatDesc* Descs(const LatDesc* all) const { return n_lat == 1 ? all + first_slot : all; }
        const uint32_t* WorkList() const { return n_lat == 1 ? nullptr : d_block_lat.ptr; }
    };

    struct GraphEntry
    {
        uint32_t sweeps = 0;       // sweeps per replay
        int32_t sample = -1;       // -1: no sample; else the SampleMode taken after every `sep` sweeps
        uint32_t sep = 0;
        hipGraphExec_t exec = nullptr;
    };
}

struct swmi_engine
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<swmi::Lattice>> lattices;
    std::vector<swmi::LatDesc> descs;
    swmi::DeviceBuffer<swmi::LatDesc> d_desc;
    std::vector<swmi::KernelClass> classes;
    bool recording = false;   // some lattice records its bonds (selects the inspection build of k_bond_label)
    bool use_graphs = true;
    std::vector<swmi::GraphEntry> graphs;   // valid for the current work lists only
    // kernel classes are independent: class c > 0 runs on side stream c - 1, forked from and joined
    // back into the engine stream around every sweep loop (parallel branches in a captured graph)
    std::vector<hipStream_t> side_streams;
    std::vector<hipEvent_t> join_events;
    hipEvent_t fork_event = nullptr;
    // single-lattice descriptor for the inspection hooks (launched with a null work list)
    swmi::DeviceBuffer<swmi::LatDesc> d_one_desc;

    swmi_engine();
    ~swmi_engine();
    swmi_engine(const swmi_engine&) = delete;
    swmi_engine& operator=(const swmi_engine&) = delete;

    swmi::Lattice& At(uint32_t i);
    swmi::LatDesc Desc(const swmi::Lattice& l, uint32_t block_base) const;
    // Rebuilds descriptors and class work lists after a change of the lattice set, a temperature or the
    // recording mode, and drops the captured graphs. The uploads are ordered on the engine stream; the
    // host copies stay alive until the next rebuild.
    void Rebuild();
    void DropGraphs();
    void LaunchCheck(const char* what);
    void EnqueueSweeps(uint32_t n);
    // applies the pending flip of every lattice (engine stream), so that their spin arrays hold the chains
    void EnqueueMaterialize();
    void EnqueueSample(uint32_t mode);
    // side streams and events for every class beyond the first
    void EnsureSideStreams();
    hipStream_t ClassStream(size_t c) const { return c == 0 ? stream : side_streams[c - 1]; }
    // n sweeps (with a sample of `sample_mode` after every `sep` sweeps when sample_mode >= 0), through
    // captured graphs when enabled.
    void Run(uint32_t n, uint32_t sep, int32_t sample_mode);
    void OneLattice(const swmi::Lattice& l, const uint8_t* bond_in);
    // OneLattice, then applies lattice l's pending flip
    void MaterializeOne(const swmi::Lattice& l);
    uint64_t TotalSites() const;
};
