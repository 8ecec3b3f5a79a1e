// swmi_internal.h -- device data layout shared by the kernels and the engine.
//
// Padding. A lattice of width n0v is stored with a row pitch n0 >= n0v that the wave64 tiling takes: n0v
// itself when it is a multiple of 64 or a divisor of 64 (with n1 a multiple of 64/n0v), else the next
// multiple of 64 (n0v > 64) or the smallest power of two p >= n0v with n1 % (64/p) == 0. Site (x, y) is
// padded site y*n0 + x; the columns x in [n0v, n0) are padding: their lanes draw (they belong to an lcg32
// group) but bond to nothing, and every kernel that reduces or reports sites skips them. Below, n is the
// padded site count n0*n1 and "site" means padded site.
//
// HBM layout of one lattice (n = n0*n1 padded sites, n % 64 == 0, groups of 64 sites):
//   spins    uint8[n]        bit 0: spin {0,1}; bits 3-5: the site's current lcg32 parameter selector
//   rng      uint32[n]       per-site lcg32 state (lane i of group g = site 64g+i)
//   masks    uint64[3*n/64]  initial (a,c) selector of each lane as three bit planes (read-only)
//   rot      uint32[n/64]    rotation offset of each group's (a,c) arrays (updated on rotating sweeps)
//   local    uint16[n]       block label of each site: the block-local index q of its cluster's root, or for a
//                            cluster that reaches the block edge 0x8000 | (qcap + s), s = the root's slot
//   cpar     uint64[4*edge_cap*blocks]  global union-find over the roots of edge-reaching clusters, one
//                            entry per slot: slot s of block b is c = 4*edge_cap*b + s, its entry
//                            (key << 32) | parent slot, key = the root's site index (the order of the
//                            union-find: a parent's key never exceeds its child's, so a tree's root is the
//                            minimum site of the component). Each block numbers its edge roots densely, in
//                            the order their first edge marks land
//   edge_count uint32[4*blocks]  word 4b: slots in use in block b
//   bbits    uint64[5*blocks] block-edge bond bits: down edge, then right edge (64 rows per word)
//   ecol     uint32[2*ecol_stride*blocks] (label << 16) | root index q of each block's first (side 0) and
//                            last (side 1) column, one entry per lattice row of the block, for the merge
//   erow     uint32[2*64*blocks] the same for the block's first (side 0) and last (side 1) wave row
//   fbits    uint32[fsw*blocks] per block: the flip bit of every edge-reaching cluster by its slot, written by
//                            k_resolve, read by the next sweep's k_bond_label (the bits of the clusters inside a
//                            block follow from the flip rule at their roots and are not stored)
//   parent   uint32[n]       inspection only: canonical (minimum) site index of every site (k_write_labels)
//   halo     uint8[2*(ecol_stride+64)*blocks] two sides (by sweep parity) of, per block, the spins of its
//                            first column (one per lattice row) and of its first wave row: k_bond_label
//                            writes them into the sweep's side, k_resolve flips them in place, and in the
//                            next sweep the block's left and upper neighbours read them as halos while the
//                            block's own spins still wait for their flip
//
// Deferred flip. A sweep is k_bond_label, k_merge, k_resolve. k_resolve decides every cluster's flip but
// leaves the spins as they were; the next sweep's k_bond_label flips each site as it loads it (from its
// label and the block's flip bits) and stores the flipped spin. LatState::applied is the iteration at
// which the spin array holds the chain's state; applied != iteration means the last sweep's flip is still
// pending, and every reader of spins (observables, downloads, checkpoints, the inspection hooks) first
// runs k_flip, which applies it.
//
// Tiling. A "wave row" is 64 consecutive row-major sites, which is exactly one lcg32 group:
//   n0 >= 64 (n0 % 64 == 0): a wave row is 64 columns of one lattice row;
//   n0 <  64 (64 % n0 == 0): a wave row holds rpw = 64/n0 whole lattice rows.
// A tile is `twr` wave rows (tw = min(n0,64) columns wide) and is labelled by one wavefront; a block
// is kWavesPerBlock tiles stacked vertically and is labelled by one workgroup. Block-local index
// q = (wave row within the block)*64 + lane is row-major, so the local minimum maps to the global
// minimum: site(q) = block_base + (q >> 6)*stride + (q & 63), stride = 64*tiles_x.
#pragma once

#include <cstddef>
#include <cstdint>

namespace swmi
{
    constexpr uint32_t kWave = 64;
    constexpr uint32_t kWavesPerBlock = 4;   // tiles per block (workgroup)
    constexpr uint32_t kNumXcd = 8;          // MI355X: 8 XCDs, workgroups dispatched round-robin
    constexpr uint16_t kEdgeFlag = 0x8000u;  // cluster reaches the block edge
    constexpr uint32_t kRightWords = kWavesPerBlock;
    constexpr uint32_t kBbitWords = 1 + kRightWords;
    constexpr uint32_t kSelShift = 3;        // spin byte: selector k in bits 3-5 (k << 3 is its 8-byte table offset)
    constexpr uint32_t kSelMask = 7u << kSelShift;

    struct LatDesc
    {
        uint8_t* spins;
        uint32_t* rng;
        const uint64_t* masks;
        uint32_t* rot;
        uint16_t* local;
        uint32_t* parent;        // inspection output of k_write_labels
        uint64_t* cpar;          // slot union-find
        uint64_t* bbits;         // kBbitWords words per block
        uint32_t* state;         // per-lattice device counters, see LatState
        uint8_t* bond_dump;      // per-site bond bytes of the last sweep (nullptr: not recorded)
        const uint8_t* bond_in;  // injected bond bytes (label-only mode), nullptr otherwise
        int32_t* partials;       // 2 int32 (E, M) per tile for observables
        uint32_t* edge_count;    // slots in use, word kWavesPerBlock * block
        uint32_t* ecol;          // block edge columns: [block][side][lattice row in block] of (label << 16) | q
        uint32_t* erow;          // block edge wave rows: [block][side][lane] of (label << 16) | q
        uint32_t ecol_stride;    // lattice rows per block: kWavesPerBlock * twr * rpw
        uint32_t edge_cap;       // slots of one tile: min(tile sites, block boundary sites)
        uint32_t qcap;           // block-local indices of a block: kWavesPerBlock * twr * 64
        uint32_t fbw;            // flip-bit words per block: (qcap + 4 * edge_cap) / 32, rounded up
        uint32_t fsw;            // slot flip-bit words of a block: 4 * edge_cap / 32, rounded up
        uint32_t n0, n1;         // row pitch (padded width) and height
        uint32_t n0v;            // lattice width: columns [n0v, n0) are padding
        uint32_t elast;          // lane-in-tile (lx) of a row's last site in the last tile column
        uint64_t vlast;          // lanes of a wave row of the last tile column that hold sites
        uint32_t tw;             // tile width in sites: min(n0, 64)
        uint32_t rpw;            // lattice rows per wave row: 64 / tw
        uint32_t twr;            // wave rows per tile
        uint32_t tiles_x;        // n0 / tw
        uint32_t blocks_y;       // ceil(nwr_y / (kWavesPerBlock * twr))
        uint32_t nwr_y;          // wave rows per column of tiles: n1 / rpw
        uint32_t threshold;      // bond iff draw u < threshold
        uint32_t block_base;     // first block of this lattice in the engine work list
        uint32_t n_blocks;       // tiles_x * blocks_y
        uint32_t flip_rule;      // SWMI_FLIP_PARITY (0) or SWMI_FLIP_RANDOM (1)
        uint32_t seed;           // the lattice's lcg32 seed (keys the random flip rule)
        uint32_t* fbits;         // fsw words per block
        uint8_t* halo;           // two sides of n_blocks * halo_stride bytes: first column, then first wave row
        uint32_t halo_stride;    // ecol_stride + 64
    };

    // Per-lattice device counters and FP64 moments (uint32 words; moments start at word kStMoments).
    // Sweep schedule: k_bond_label advances a lattice only while iteration < it_end (two lcg32 updates per
    // sweep) and records the decision in `active`, which k_merge and k_resolve of the same sweep read (the
    // iteration itself changes inside k_resolve). Conditional samples (sample_mode kSampleDue) are taken
    // when iteration > last_sample, iteration > it_meas and (iteration - it_meas) % sep2 == 0, i.e. after
    // every n_sep-th sweep past the warmup of the lattice's own schedule.
    struct LatState
    {
        uint32_t iteration;      // lcg32 Update() count, the same for every group of the lattice
        uint32_t n_samples;      // samples taken
        uint32_t series_cap;     // capacity of the (E, M) series
        uint32_t error_flags;    // set by kernels on a violated invariant (label out of range)
        uint32_t it_end;         // iteration at which the lattice stops (0xFFFFFFFF: never)
        uint32_t active;         // 1 if the current sweep advances the lattice
        uint32_t it_meas;        // iteration at the end of the warmup
        uint32_t sep2;           // 2 * n_sep
        uint32_t last_sample;    // iteration of the last sample (0: none)
        uint32_t flip_key;       // random flip rule: key of the current sweep (written by k_bond_label)
        uint32_t applied;        // iteration at which the spin array holds the chain's state
        uint32_t halo_side;      // side of the halo array the current sweep writes ((iteration / 2) & 1)
        double moments[6];       // samples, sum e, sum e^2, sum |m|, sum m^2, sum m^4
        // followed by int64 series[2 * series_cap]
    };
    constexpr uint32_t kLatStateWords = sizeof(LatState) / 4;
    constexpr uint32_t kStIteration = 0, kStSamples = 1, kStErr = 3, kStItEnd = 4, kStActive = 5, kStItMeas = 6,
                       kStSep2 = 7, kStLastSample = 8, kStFlipKey = 9, kStApplied = 10, kStHaloSide = 11,
                       kStMoments = 12;
    static_assert(offsetof(LatState, moments) == 4 * kStMoments, "LatState layout");

    enum SampleMode : uint32_t
    {
        kSampleAlways = 0u,  // every lattice, now
        kSampleDue = 1u      // only lattices whose schedule calls for a sample at this iteration
    };


    constexpr uint32_t kErrLabelRange = 1u;  // a label chase left [0, n) or met an unflagged root

    // Launch wrappers (swmi_kernels.hip). Work list: block_lat[b] = lattice of global block b; a null work list
    // means every block belongs to descriptor 0.
    // all_tw64: every lattice of the work list has n0 % 64 == 0 (64-wide tiles); selects the specialised kernel
    // debug: compiles in injected bonds (LatDesc::bond_in) and bond recording (LatDesc::bond_dump)
    // pad: some lattice of the work list is padded (selects the PAD specialisation, correct for every lattice)
    void LaunchBondLabel(const LatDesc* d_desc, const uint32_t* d_block_lat, uint32_t total_blocks, uint32_t max_twr,
                         bool all_tw64, bool pad, bool debug, hipStream_t stream);
    void LaunchMerge(const LatDesc* d_desc, const uint32_t* d_block_lat, uint32_t total_blocks, hipStream_t stream);
    // end of a sweep: flip bits and halos of every block, then iteration += 2
    void LaunchResolve(const LatDesc* d_desc, const uint32_t* d_block_lat, uint32_t total_blocks, hipStream_t stream);
    // apply a pending flip to the spin array (no-op for lattices without one)
    void LaunchMaterialize(const LatDesc* d_desc, const uint32_t* d_block_lat, uint32_t total_blocks, hipStream_t stream);
    // canonical label (minimum site index) of every site into parent[] (inspection)
    void LaunchWriteLabels(const LatDesc* d_desc, const uint32_t* d_block_lat, uint32_t total_blocks, uint32_t max_twr,
                           hipStream_t stream);
    // a lattice's spin array was replaced: no flip is pending
    void LaunchSetApplied(uint32_t* state, hipStream_t stream);
    // allocation-free labelling: descriptor d into d_desc, fresh d.state
    void LaunchNoAllocSetup(const LatDesc& d, LatDesc* d_desc, hipStream_t stream);
    void LaunchObservables(const LatDesc* d_desc, const uint32_t* d_block_lat, uint32_t total_blocks, uint32_t sample_mode,
                           hipStream_t stream);
    void LaunchFinalizeSamples(const LatDesc* d_desc, uint32_t n_lattices, uint32_t sample_mode, hipStream_t stream);
    void LaunchInitState(uint32_t* state, uint32_t words, uint32_t series_cap, uint32_t it_end, uint32_t it_meas, uint32_t sep2,
                         hipStream_t stream);
    void LaunchSetSchedule(uint32_t* state, uint32_t sweeps, uint32_t warmup, uint32_t n_sep, hipStream_t stream);
    void LaunchInitRng(uint8_t* spins, uint32_t* rng, uint64_t* masks, uint32_t* rot, uint32_t n_groups, uint32_t seed,
                       hipStream_t stream);
    // n row-major spins (width n0v) into the padded spin bytes (pitch), keeping the selector bits
    void LaunchSetSpins(uint8_t* spins, const uint8_t* src, uint64_t n, uint32_t n0v, uint32_t pitch, hipStream_t stream);
    // row-major bytes (width n0v) to np padded bytes (pitch), padding 0
    void LaunchPadBytes(uint8_t* dst, const uint8_t* src, uint64_t np, uint32_t n0v, uint32_t pitch, hipStream_t stream);
    // padded canonical labels to the n row-major labels of the row-major sites
    void LaunchUnpadLabels(uint32_t* dst, const uint32_t* src, uint64_t n, uint32_t n0v, uint32_t pitch, hipStream_t stream);
    void LaunchHotStart(uint8_t* spins, uint64_t n, uint32_t spin_seed, uint32_t n0v, uint32_t pitch, hipStream_t stream);
    void LaunchCopy(const uint4* src, uint4* dst, uint64_t n16, hipStream_t stream);
    void LaunchIota(uint32_t* labels, uint64_t n, hipStream_t stream);
}
