// swmi_noalloc.hip -- swmi_label_bonds_noalloc: the labelling of swmi_label_bonds for callers that must not
// allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).
//
// The caller passes the bond field and the label array in device memory. Every other buffer the kernels
// need (block-local labels, the union-find over edge-root slots, block-edge bond bits, slot counts, edge
// columns and rows, per-lattice counters and the lattice descriptor) is carved out of one file-scope device array, `scratch`,
// which every GPU holds for the life of the process. For a lattice that needs no row padding the last kernel
// writes the canonical labels straight into the caller's label array; a padded lattice gets its bond field
// padded into scratch and its labels written into scratch, and a last kernel writes the unpadded labels. The call enqueues kernels on the caller's stream and returns: no
// allocation, no memset, no host synchronisation.
//
// Every call on a GPU shares that GPU's copy of the array, so two calls that overlap on one GPU (two
// streams, two threads) race; calls on one stream are ordered and do not.
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <string>

#include "swmi_host.h"

namespace swmi
{
    namespace
    {
        constexpr uint64_t kNoAllocMaxSites = 1ull << 22;   // 4,194,304 sites, e.g. 2048 x 2048
        constexpr size_t kScratchBytes = 48ull << 20;       // 48 MiB per GPU
        constexpr int kMaxDevices = 64;
        constexpr uint32_t kNoAllocTileRows = 16;
    }
}

// The scratch array. In host code `scratch` names a host-side stand-in, not device memory: host code uses
// only the device address that hipGetSymbolAddress returns for the current device (ScratchOf below).
static __device__ __align__(256) uint8_t scratch[swmi::kScratchBytes];

namespace swmi
{
    namespace
    {
        std::mutex g_scratch_mutex;
        std::array<uint8_t*, kMaxDevices> g_scratch{};   // device address of `scratch`, per device

        // Address of the current device's scratch array, looked up once per device (with that device
        // current), after checking that the device is a gfx950.
        uint8_t* ScratchOf(const int device)
        {
            std::lock_guard<std::mutex> lock(g_scratch_mutex);
            if (!g_scratch[device])
            {
                RequireGfx950(device);
                void* ptr = nullptr;
                Check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
                if (!ptr)
                    throw SwmiError(SWMI_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
                g_scratch[device] = static_cast<uint8_t*>(ptr);
            }
            return g_scratch[device];
        }

        // Consecutive 256-byte aligned pieces of the scratch array. With a null base it only adds up the bytes.
        struct Carve
        {
            uint8_t* base = nullptr;
            size_t used = 0;
            template <typename T>
            T* Take(const size_t count)
            {
                used = (used + 255u) & ~static_cast<size_t>(255u);
                T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
                used += count * sizeof(T);
                return p;
            }
        };

        // The label-only descriptor over a carve (spins alias the bond field: the injected pass reads spin
        // bytes but uses none of them; it reads no lcg32 state and writes no spin, halo or flip bit).
        LatDesc NoAllocDesc(const Geometry& g, const uint32_t n0, const uint32_t n1, const uint8_t* bonds, uint32_t* labels,
                            Carve& c, LatDesc** d_desc)
        {
            LatDesc d = GeometryDesc(g, n0, n1);
            const uint64_t n = static_cast<uint64_t>(g.pitch) * n1;   // padded sites
            const bool pad = g.pitch != n0;
            *d_desc = c.Take<LatDesc>(1);
            d.state = c.Take<uint32_t>(kLatStateWords);
            d.local = c.Take<uint16_t>(n);
            d.cpar = c.Take<uint64_t>(static_cast<size_t>(kWavesPerBlock) * g.n_blocks * g.edge_cap);
            d.bbits = c.Take<uint64_t>(static_cast<size_t>(kBbitWords) * g.n_blocks);
            d.edge_count = c.Take<uint32_t>(static_cast<size_t>(kWavesPerBlock) * g.n_blocks);
            d.ecol = c.Take<uint32_t>(2ull * g.ecol_stride * g.n_blocks);
            d.erow = c.Take<uint32_t>(2ull * 64 * g.n_blocks);
            d.parent = pad ? c.Take<uint32_t>(n) : labels;
            d.bond_in = pad ? c.Take<uint8_t>(n) : bonds;
            d.spins = const_cast<uint8_t*>(d.bond_in);
            return d;
        }
    }
}

using namespace swmi;

extern "C" int swmi_label_bonds_noalloc(const uint8_t* bonds, uint32_t n0, uint32_t n1, uint32_t* labels, void* stream)
{
    return Guard([&] {
        if (!bonds || !labels)
            throw SwmiError(SWMI_ERR_INVALID, "null bond or label array");
        swmi_lattice_config c{};
        c.n0 = n0;
        c.n1 = n1;
        c.temperature = 1.0f;
        ValidateConfig(c);
        if (static_cast<uint64_t>(n0) * n1 > kNoAllocMaxSites)
            throw SwmiError(SWMI_ERR_INVALID, "swmi_label_bonds_noalloc takes at most " + std::to_string(kNoAllocMaxSites) +
                                                  " sites (got " + std::to_string(n0) + " x " + std::to_string(n1) + ")");
        const Geometry g = GeometryOf(n0, n1, kNoAllocTileRows);
        Carve size;
        LatDesc* d_desc = nullptr;
        (void)NoAllocDesc(g, n0, n1, bonds, labels, size, &d_desc);
        if (size.used > kScratchBytes)
            throw SwmiError(SWMI_ERR_INVALID, "lattice needs " + std::to_string(size.used) + " bytes of scratch, more than the " +
                                                  std::to_string(kScratchBytes) + " the static array holds");
        int device = 0;
        Check(hipGetDevice(&device), "hipGetDevice");
        if (device < 0 || device >= kMaxDevices)
            throw SwmiError(SWMI_ERR_NO_DEVICE, "device index " + std::to_string(device) + " out of range");
        Carve carve;
        carve.base = ScratchOf(device);
        const LatDesc d = NoAllocDesc(g, n0, n1, bonds, labels, carve, &d_desc);
        hipStream_t s = static_cast<hipStream_t>(stream);
        const bool pad = g.pitch != n0;
        const uint64_t np = static_cast<uint64_t>(g.pitch) * n1, n = static_cast<uint64_t>(n0) * n1;
        // descriptor and fresh counters are written by a kernel (no memset, no host copy); the kernels run with
        // a null work list (every block belongs to the one lattice)
        LaunchNoAllocSetup(d, d_desc, s);
        if (pad)
            LaunchPadBytes(const_cast<uint8_t*>(d.bond_in), bonds, np, n0, g.pitch, s);
        LaunchBondLabel(d_desc, nullptr, g.n_blocks, g.twr, g.tw == 64, pad, true, s);
        LaunchMerge(d_desc, nullptr, g.n_blocks, s);
        LaunchWriteLabels(d_desc, nullptr, g.n_blocks, g.twr, s);
        if (pad)
            LaunchUnpadLabels(labels, d.parent, n, n0, g.pitch, s);
        Check(hipGetLastError(), "label launch");
    });
}
