// swmi_internal.h -- device data layout shared by the kernels and the engine.
//
// HBM layout of one lattice (n = n0*n1 sites, n % 64 == 0, groups of 64 sites):
//   spins   uint8[n]        {0,1}, row-major
//   rng     uint32[n]       per-site lcg32 state (lane i of group g = site 64g+i)
//   masks   uint64[3*n/64]  initial (a,c) selector of each lane as three bit planes
//   rot     uint32[n/64]    rotation offset of each group's (a,c) arrays
//   labels  uint32[n]       union-find parents; after a sweep, site -> canonical (minimum) site index
//   bbits   uint64[2*tiles] tile-boundary bond bits (down edge, right edge)
//
// Tiling. A "wave row" is 64 consecutive row-major sites, which is exactly one lcg32 group:
//   n0 >= 64 (n0 % 64 == 0): a wave row is 64 columns of one lattice row;
//   n0 <  64 (64 % n0 == 0): a wave row holds rpw = 64/n0 whole lattice rows.
// A tile is `twr` wave rows stacked vertically (tw = min(n0,64) columns wide) and is labelled by one
// wavefront: lane l of wave row r is local site r*64 + l. Local order is row-major, so the local
// minimum maps to the global minimum.
#pragma once

#include <cstdint>

namespace swmi
{
    constexpr uint32_t kWave = 64;
    constexpr uint32_t kWavesPerBlock = 4;   // tiles per workgroup
    constexpr uint32_t kNumXcd = 8;          // MI355X: 8 XCDs, workgroups dispatched round-robin

    struct LatDesc
    {
        uint8_t* spins;
        uint32_t* rng;
        const uint64_t* masks;
        uint32_t* rot;
        uint32_t* labels;
        uint64_t* bbits;
        uint32_t* state;         // per-lattice device counters, see LatState
        uint8_t* bond_dump;      // per-site bond bytes of the last sweep (nullptr: not recorded)
        const uint8_t* bond_in;  // injected bond bytes (label-only mode), nullptr otherwise
        int32_t* partials;       // 2 int32 (E, M) per tile for observables
        uint32_t n0, n1;
        uint32_t tw;             // tile width in sites: min(n0, 64)
        uint32_t rpw;            // lattice rows per wave row: 64 / tw
        uint32_t twr;            // wave rows per tile
        uint32_t tiles_x;        // n0 / tw
        uint32_t tiles_y;        // ceil(nwr_y / twr)
        uint32_t nwr_y;          // wave rows per column of tiles: n1 / rpw
        uint32_t threshold;      // bond iff draw u < threshold
        uint32_t tile_base;      // first tile of this lattice in the engine work list
        uint32_t n_tiles;
        uint32_t flip_rule;
    };

    // Per-lattice device counters and FP64 moments (uint32 words; moments start at word 4).
    struct LatState
    {
        uint32_t iteration;      // lcg32 Update() count, the same for every group of the lattice
        uint32_t n_samples;      // samples taken
        uint32_t series_cap;     // capacity of the (E, M) series
        uint32_t error_flags;    // set by kernels on a violated invariant (label out of range)
        double moments[6];       // samples, sum e, sum e^2, sum |m|, sum m^2, sum m^4
        // followed by int64 series[2 * series_cap]
    };
    constexpr uint32_t kLatStateWords = sizeof(LatState) / 4;

    enum KernelFlags : uint32_t
    {
        kFlip = 1u,          // flip spins by root parity
        kWriteLabels = 2u,   // write canonical labels back
        kAdvance = 4u        // advance the lcg32 iteration count by 2 (end of a sweep)
    };

    // Launch wrappers (swmi_kernels.hip). Work lists: tile_lat[t] = lattice of global tile t.
    void LaunchBondLabel(const LatDesc* d_desc, const uint32_t* d_tile_lat, uint32_t total_tiles, uint32_t max_twr,
                         hipStream_t stream);
    void LaunchMerge(const LatDesc* d_desc, const uint32_t* d_tile_lat, uint32_t total_tiles, hipStream_t stream);
    void LaunchFlip(const LatDesc* d_desc, const uint32_t* d_tile_lat, uint32_t total_tiles, uint32_t flags,
                    hipStream_t stream);
    void LaunchObservables(const LatDesc* d_desc, const uint32_t* d_tile_lat, uint32_t total_tiles, hipStream_t stream);
    void LaunchFinalizeSamples(const LatDesc* d_desc, uint32_t n_lattices, hipStream_t stream);
    void LaunchInitRng(uint32_t* rng, uint64_t* masks, uint32_t* rot, uint32_t n_groups, uint32_t seed, hipStream_t stream);
    void LaunchHotStart(uint8_t* spins, uint64_t n, uint32_t spin_seed, hipStream_t stream);
    void LaunchCopy(const uint4* src, uint4* dst, uint64_t n16, hipStream_t stream);
    void LaunchIota(uint32_t* labels, uint64_t n, hipStream_t stream);

    constexpr uint32_t kErrLabelRange = 1u;  // a label chase left [0, n) or ran longer than n steps
}
