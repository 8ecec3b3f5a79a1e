// swmi_kernels.hip -- gfx950 kernels of the Swendsen-Wang engine.
//
// One sweep = three launches over the engine's block work list (all lattices of the engine at once):
//   k_bond_label  per block (one workgroup, 4 wavefronts, one tile each): applies the previous sweep's
//                 cluster flip to each site as it loads it (label -> bit of the block's flip bitmap) and
//                 stores the flipped spin; two lcg32 updates per site; bond activation against the integer
//                 threshold on 64-bit ballot masks (scalar unit); row runs from the masks, unions in an LDS
//                 union-find (atomicMin keeps parent <= child, so roots are cluster minima); block-local
//                 uint16 labels with an edge flag; the block-edge bond bits and the block's halo spins.
//   k_merge       per block (two waves): unions across block edges by atomicMin on the uint32 parent
//                 array in HBM, over the roots of edge-reaching clusters only.
//   k_resolve     per block (two waves): the listed edge-cluster roots are chased in the global parent
//                 array (one chase per root); the block's flip bitmap (the flip rule applied to every
//                 cluster's canonical root) is written out and the block's halo spins are flipped in place.
//                 Advances the lcg32 iteration count; the spin array waits for the next k_bond_label.
// k_flip applies a pending flip to the spin array; it runs only before something reads the spins.
// Spin bytes: bit 0 is the spin, bits 3-5 hold the site's current lcg32 parameter selector, i.e. the
// index k of the Numerical Recipes pair (a_k, c_k) that lane i of its group uses after the group's
// rotations (sel0[(i + rot) & 63]). The byte is read every sweep anyway, so the generator needs no other
// per-group data except on the sweeps that rotate (one in 7.5).
// Reference behaviour: ising/src/swendsen_wang_amd_gpu.cpp:78-467 (four kernels per sweep) and the CPU
// labeller ising/src/swendsen_wang_cpu.cpp:13-205; the kernels here are written for wave64 CDNA4.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "swmi_internal.h"

namespace swmi
{
    __device__ __forceinline__ uint32_t Uniform(const uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

    // The buffers are reached through pointers stored in the lattice descriptor, which the compiler sees
    // as generic (FLAT) addresses; FLAT accesses wait on both the memory and the LDS counters. Every
    // buffer access goes through a global-address-space view instead.
    template <typename T>
    using gptr = __attribute__((address_space(1))) T*;

    template <typename T>
    __device__ __forceinline__ gptr<T> Glob(T* p)
    {
        return (gptr<T>)(p);
    }

    // plain vector types (HIP's uint2/uint4 classes cannot be assigned through address-space-1 pointers)
    using u32x2 = uint32_t __attribute__((ext_vector_type(2)));
    using u32x4 = uint32_t __attribute__((ext_vector_type(4)));

    // Buffer views (raw: stride 0, byte offsets). The hardware checks a lane's offset against the view's
    // size and drops a store past it, which lets a lane opt out of a store without an exec mask.
    using Rsrc = __amdgpu_buffer_rsrc_t;
    constexpr uint32_t kNoStore = 0x80000000u;   // a lane offset past every view
    template <typename T>
    __device__ __forceinline__ Rsrc MakeRsrc(T* base, const uint32_t bytes)
    {
        return __builtin_amdgcn_make_buffer_rsrc(static_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
    }

    // bit `lane` of a wave-uniform mask, as a per-lane predicate (v_cndmask on the SGPR pair)
    __device__ __forceinline__ bool LaneBit(const uint64_t mask) { return __builtin_amdgcn_inverse_ballot_w64(mask); }

    // 64-bit shifts by 0..64 (a shift by 64 yields 0; the hardware would use the count modulo 64)
    __device__ __forceinline__ uint64_t Shr(const uint64_t v, const uint32_t s) { return s >= 64u ? 0ull : v >> s; }
    __device__ __forceinline__ uint64_t Shl(const uint64_t v, const uint32_t s) { return s >= 64u ? 0ull : v << s; }

    __device__ __forceinline__ uint32_t ReadLane(const uint32_t v, const uint32_t lane)
    {
        return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(lane)));
    }

    // Numerical Recipes pair k of random/src/lcg32.cpp:18-24.
    __device__ __forceinline__ void ParamsOf(const uint32_t k, uint32_t& a, uint32_t& c)
    {
        const bool b0 = (k & 1u) != 0u, b1 = (k & 2u) != 0u, b2 = (k & 4u) != 0u;
        const uint32_t a01 = b0 ? 2891336453u : 1372383749u;
        const uint32_t a23 = b0 ? 32310901u : 2024337845u;
        const uint32_t c01 = b0 ? 1640531513u : 1289706101u;
        const uint32_t c23 = b0 ? 626627237u : 797082193u;
        a = b2 ? 29943829u : (b1 ? a23 : a01);
        c = b2 ? 1013904223u : (b1 ? c23 : c01);
    }

    // Offset added to the rotation at update `it` (random/src/lcg32.cpp:56-64): state[it & 63] + (it even).
    __device__ __forceinline__ uint32_t ShuffleShift(const uint32_t st, const uint32_t it)
    {
        return ReadLane(st, it & 63u) + ((it & 1u) ? 0u : 1u);
    }

    // ---- random flip rule: MurmurHash3 (Austin Appleby, public domain). The cluster with canonical root r
    // flips iff the top bit of Fmix32(r ^ key) is set; key = MurmurHash3_x86_32 of the 4-byte lcg32
    // iteration count before the sweep, seeded with the lattice seed. ----
    __device__ __forceinline__ uint32_t Fmix32(uint32_t h)
    {
        h ^= h >> 16;
        h *= 0x85EBCA6Bu;
        h ^= h >> 13;
        h *= 0xC2B2AE35u;
        h ^= h >> 16;
        return h;
    }

    __device__ __forceinline__ uint32_t FlipKey(const uint32_t seed, const uint32_t iteration)
    {
        uint32_t k = iteration * 0xCC9E2D51u;
        k = (k << 15) | (k >> 17);
        k *= 0x1B873593u;
        uint32_t h = seed ^ k;
        h = (h << 13) | (h >> 19);
        h = h * 5u + 0xE6546B64u;
        return Fmix32(h ^ 4u);
    }

    // ---- union-find in LDS. Nodes are named by their byte offset into the workgroup's parent array P
    // (so no instruction scales an index into an address); parents satisfy P[x] <= x, so roots are cluster
    // minima. ----
    __device__ __forceinline__ uint32_t* At(uint32_t* P, const uint32_t off)
    {
        return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(P) + off);
    }

    __device__ __forceinline__ uint32_t LdsLoad(const uint32_t* p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }

    // An empty asm statement that reads v: v's load completes before this point, so it cannot be sunk past a
    // later branch (an early exit) and is waited for together with the loads issued beside it.
    __device__ __forceinline__ void Issued(const uint32_t v)
    {
        asm volatile("" ::"v"(v));
    }

    __device__ __forceinline__ void Issued(const uint64_t v)
    {
        asm volatile("" ::"v"(v));
    }

    // An empty asm statement that needs the uniform value v in scalar registers here: a descriptor field read
    // this way is fetched at this point, with the others read beside it, not at its first use.
    template <typename T>
    __device__ __forceinline__ void HeldScalar(const T v)
    {
        asm volatile("" ::"s"(v));
    }

    __device__ __forceinline__ void LdsMin(uint32_t* p, const uint32_t v)
    {
        (void)__hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }

    // ---- union-find in HBM over the roots of edge-reaching clusters, one node per slot (every block numbers
    // its edge roots densely, so a block's nodes share a few cache lines). Entry of slot c: (key << 32) | parent
    // slot, where a root's key is its site index and a non-root holds its parent's key. Links are made with
    // agent-scope 64-bit atomicMin (the 8 XCD L2s are not coherent with each other for plain stores), which
    // orders by key first: a parent's key never exceeds its child's, so a tree's root is the minimum site of
    // its component. Chase loads are relaxed atomic loads (never hoisted by the compiler); a stale value is an
    // earlier parent, still an ancestor, and every link is decided by an atomicMin that returns the true value,
    // so staleness only costs retries and longer paths. ----
    __device__ __forceinline__ uint64_t SlotEntry(const uint32_t key, const uint32_t slot)
    {
        return (static_cast<uint64_t>(key) << 32) | slot;
    }

    __device__ __forceinline__ uint64_t CachedLoad64(gptr<uint64_t> p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }

    // Loads of the merge's chases. FRESH reads the entry as it is now, past the CU's cache (agent scope);
    // otherwise the load may hit a line the CU cached earlier in the kernel. Fresh loads cost memory traffic
    // and save retries: they pay where the merge is latency-bound, i.e. where all its blocks are resident at
    // once (LaunchMerge chooses).
    template <bool FRESH>
    __device__ __forceinline__ uint64_t ChaseLoad64(gptr<uint64_t> p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, FRESH ? __HIP_MEMORY_SCOPE_AGENT : __HIP_MEMORY_SCOPE_WORKGROUP);
    }

    __device__ __forceinline__ uint64_t SlotLinkMin(gptr<uint64_t> p, const uint64_t v)
    {
        return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // Union of the clusters of slots a and b (keys ka, kb), the loop of the reference Merge
    // (swendsen_wang_amd_gpu.cpp:368-401) on slot entries. The ends start as the listed roots k_bond_label
    // wrote (entry = (own key << 32) | self), so the first link is tried at once, without chasing: when a
    // concurrent union got there first, atomicMin returns the parent found and the loop continues from it.
    // Both ends are chased together (their loads in flight at once); a start node more than one hop from its
    // root is pointed at the root (the atomic returns nothing, so the wave does not wait for it). A slot out
    // of range or a parent with a larger key can only come from a broken invariant, which is reported through
    // the lattice's error flag instead of being followed.
    template <bool FRESH>
    __device__ void SlotUnion(gptr<uint64_t> E, uint32_t a, uint32_t ka, uint32_t b, uint32_t kb, const uint32_t nslots,
                              gptr<uint32_t> err)
    {
        if (a < nslots && b < nslots && ka != kb)
        {
            if (ka > kb)
            {
                const uint32_t t = a, tk = ka;
                a = b;
                ka = kb;
                b = t;
                kb = tk;
            }
            const uint64_t old = SlotLinkMin(&E[b], SlotEntry(ka, a));
            if (old == SlotEntry(kb, b) || static_cast<uint32_t>(old) == a)
                return;
            b = static_cast<uint32_t>(old);
            kb = static_cast<uint32_t>(old >> 32);
        }
        for (;;)
        {
            if (a >= nslots || b >= nslots)
                break;
            const uint32_t a0 = a, b0 = b;
            uint64_t ea = ChaseLoad64<FRESH>(&E[a]), eb = ChaseLoad64<FRESH>(&E[b]);
            uint32_t ha = 0, hb = 0;
            bool bad = false;
            while (static_cast<uint32_t>(ea) != a || static_cast<uint32_t>(eb) != b)
            {
                const uint32_t pa = static_cast<uint32_t>(ea), qa = static_cast<uint32_t>(ea >> 32);
                const uint32_t pb = static_cast<uint32_t>(eb), qb = static_cast<uint32_t>(eb >> 32);
                if (pa >= nslots || qa > ka || pb >= nslots || qb > kb)
                {
                    bad = true;
                    break;
                }
                if (pa != a)
                {
                    a = pa;
                    ka = qa;
                    ea = ChaseLoad64<FRESH>(&E[a]);
                    ++ha;
                }
                if (pb != b)
                {
                    b = pb;
                    kb = qb;
                    eb = ChaseLoad64<FRESH>(&E[b]);
                    ++hb;
                }
            }
            if (bad)
                break;
            if (ha > 1)
                (void)SlotLinkMin(&E[a0], SlotEntry(ka, a));
            if (hb > 1)
                (void)SlotLinkMin(&E[b0], SlotEntry(kb, b));
            if (a == b)
                return;
            if (ka > kb)
            {
                const uint32_t t = a, tk = ka;
                a = b;
                ka = kb;
                b = t;
                kb = tk;
            }
            const uint64_t old = SlotLinkMin(&E[b], SlotEntry(ka, a));
            if (old == SlotEntry(kb, b))
                return;
            b = static_cast<uint32_t>(old);
            kb = static_cast<uint32_t>(old >> 32);
        }
        (void)__hip_atomic_fetch_or(err, kErrLabelRange, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // Key of the canonical root (the minimum site) of the cluster of slot c, whose entry e the caller loaded;
    // c_max bounds the slots. A chase that leaves the slots or meets a parent with a larger key clears ok.
    __device__ __forceinline__ uint32_t SlotRootKey(gptr<uint64_t> E, uint32_t c, uint64_t e, const uint32_t c_max, bool& ok)
    {
        uint32_t k = static_cast<uint32_t>(e >> 32);
        while (static_cast<uint32_t>(e) != c)
        {
            const uint32_t p = static_cast<uint32_t>(e), kp = static_cast<uint32_t>(e >> 32);
            if (p >= c_max || kp > k)
            {
                ok = false;
                break;
            }
            c = p;
            k = kp;
            e = CachedLoad64(&E[c]);
        }
        return static_cast<uint32_t>(e >> 32);
    }

    // ---- work decomposition ----
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of blocks
    // so that blocks that share cache lines (spin rows, halos) meet in the same L2.
    __device__ __forceinline__ uint32_t RemapBlock(const uint32_t units)
    {
        const uint32_t per = (units + kNumXcd - 1) / kNumXcd;
        return (blockIdx.x % kNumXcd) * per + blockIdx.x / kNumXcd;
    }

    // Lattice of block b in the work list; a work list of one lattice is passed as a null pointer, which
    // saves the dependent load of the block's entry before the descriptor.
    __device__ __forceinline__ uint32_t LatticeOf(const uint32_t* __restrict__ block_lat, const uint32_t b)
    {
        return block_lat ? block_lat[b] : 0u;
    }

    struct Block
    {
        uint32_t lb;       // block index within its lattice
        uint32_t sx, sy;   // block column, block row
        uint32_t wyb;      // first wave row of the block
        uint32_t nwr;      // wave rows in the block
        uint32_t stride;   // sites between consecutive wave rows (64 * tiles_x)
        uint32_t base;     // first site of the block
    };

    __device__ __forceinline__ Block DecodeBlock(const LatDesc& d, const uint32_t b)
    {
        Block bl;
        bl.lb = b - d.block_base;
        bl.sx = bl.lb % d.tiles_x;
        bl.sy = bl.lb / d.tiles_x;
        bl.wyb = bl.sy * kWavesPerBlock * d.twr;
        bl.nwr = min(kWavesPerBlock * d.twr, d.nwr_y - bl.wyb);
        bl.stride = 64u * d.tiles_x;
        bl.base = bl.wyb * bl.stride + bl.sx * 64u;
        return bl;
    }

    __device__ __forceinline__ uint32_t BlockBase(const LatDesc& d, const uint32_t sx, const uint32_t sy)
    {
        return sy * kWavesPerBlock * d.twr * 64u * d.tiles_x + sx * 64u;
    }

    __device__ __forceinline__ uint32_t SiteOf(const uint32_t base, const uint32_t stride, const uint32_t q)
    {
        return base + (q >> 6) * stride + (q & 63u);
    }

    // ---- the flip rule (reference FlipClusters_Kernel, swendsen_wang_amd_gpu.cpp:445-467) ----
    // Row-major site index of a padded site index p (d.n0 is the row pitch, d.n0v the lattice width).
    __device__ __forceinline__ uint32_t RealSite(const LatDesc& d, const uint32_t p)
    {
        return d.n0v == d.n0 ? p : (p / d.n0) * d.n0v + p % d.n0;
    }

    // The flip rule for the cluster whose canonical root is the padded site `root`. The pitch and the width
    // are even, so a padded index and its row-major index have the same parity.
    __device__ __forceinline__ uint32_t FlipBit(const LatDesc& d, const uint32_t root, const bool random_rule, const uint32_t key)
    {
        return random_rule ? (Fmix32(RealSite(d, root) ^ key) >> 31) : (root & 1u);
    }

    // A block's flip bitmap has one bit per label index: bit q (q < qcap) for the cluster rooted at block-local
    // index q that stays inside the block, the rule applied to the root's own site; bit qcap + s for the
    // edge-reaching cluster whose root holds slot s, the rule applied to its canonical root (set by the
    // resolve). Word w of the rule part: bit j for q = 32w + j (block bases and row strides are even, so under
    // the parity rule q's parity is its site's parity).
    __device__ __forceinline__ uint32_t RuleWord(const LatDesc& d, const Block& bl, const uint32_t w, const bool random_rule,
                                                 const uint32_t key)
    {
        if (!random_rule)
            return 0xAAAAAAAAu;
        uint32_t v = 0;
        for (uint32_t j = 0; j < 32u; ++j)
            v |= FlipBit(d, SiteOf(bl.base, bl.stride, 32u * w + j), true, key) << j;
        return v;
    }

    // =============================================================================================
    // k_bond_label
    // =============================================================================================
    // Down bonds are deduplicated before any union: lanes l-1 and l lie in the same "overlap interval"
    // when they share their row run above and their row run below (mask C); only the first down bond
    // of every interval can join two different clusters. With M = C << 1 (continuation) and
    // S1 = (D << 1) & M, the positions that follow an earlier bond of the same interval are
    // S1 | (M & ((M + S1) ^ M)) (carries of S1 run through the ones of M).
    __device__ __forceinline__ uint64_t FirstBondPerInterval(const uint64_t D, const uint64_t C, const uint64_t rowstart)
    {
        const uint64_t M = (C << 1) & ~rowstart;
        const uint64_t S1 = (D << 1) & M;
        const uint64_t later = S1 | (M & ((M + S1) ^ M));
        return D & ~later;
    }

    // Union of the pairs (a, b) held by the lanes, all lanes in lockstep. Each step moves the ends that
    // are not roots up (two parents, halving the path); when both are roots the larger is linked under the
    // smaller with atomicMin, and a lost race continues with the value found (the loop of the reference
    // Merge, swendsen_wang_amd_gpu.cpp:368-401). Lanes with nothing to do hold a == b; only lanes that
    // link issue the atomic (idle lanes would all hit the same word). a and b are byte offsets into P.
    __device__ __forceinline__ void LdsUnionLockstep(uint32_t* P, uint32_t a, uint32_t b)
    {
        while (__ballot(a != b))
        {
            // lanes that are done (a == b) only read P[a]; they follow no pointer and write nothing, so
            // their a may name a node another wave has not written yet
            const bool act = a != b;
            const uint32_t pa = LdsLoad(At(P, a));
            const uint32_t pb = LdsLoad(At(P, b));
            const bool roots = (pa == a) && (pb == b);
            const bool up = act && !roots;
            const uint32_t lo = min(a, b), hi = max(a, b);
            uint32_t old = hi;
            if (act && roots)
                old = __hip_atomic_fetch_min(At(P, hi), lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const bool linked = roots && old == hi;
            // ends that are not roots move two steps up and point their node at the grandparent (path
            // halving, monotone like every write to P)
            const uint32_t ga = up ? LdsLoad(At(P, pa)) : 0u;
            const uint32_t gb = up ? LdsLoad(At(P, pb)) : 0u;
            if (up && ga != pa)
                LdsMin(At(P, a), ga);
            if (up && gb != pb)
                LdsMin(At(P, b), gb);
            if (act)
            {
                a = roots ? (linked ? hi : lo) : ga;
                b = roots ? (linked ? hi : old) : gb;
            }
        }
    }

    // Per wavefront (tile) the rows are processed top to bottom with the loads of the next kCh rows in
    // flight together. Bonds are computed on wave-uniform 64-bit masks: S (spin up), U1/U2 (draw below
    // threshold), B (bond to the right neighbour inside the tile), D (bond to the site below).
    // 64-wide tiles keep the previous row's roots resolved in registers (Rp). A row run's label is the
    // minimum of its first site and the roots its down bonds reach (one LDS atomicMin per bonded lane into
    // the run's slot); a bond to any other root links that root under the minimum with one atomicMin, and
    // only a link that loses a race goes through the lockstep union.
    // Narrow tiles record one union per overlap interval of consecutive rows in a per-wave LDS queue,
    // worked off 64 at a time whenever it fills.
    // Nodes of the LDS union-find are byte offsets into P. After the final find, the root of every
    // cluster that reaches the block edge carries kPFlag in its own entry, so a site's label (root index,
    // plus kEdgeFlag) is P[root] >> 2.
    // DBG: the inspection modes (injected bonds instead of generated ones, bond recording) are compiled in;
    // the sweep kernel proper has neither branch.
    // PAD: the lattice is stored with a row pitch d.n0 above its width d.n0v; in the last tile column only
    // the lanes up to d.elast hold sites (mask d.vlast), the rest are padding: they bond to nothing, their
    // draws are unused, and the row's last site bonds across the lattice edge instead of lane tw - 1.
    template <uint32_t TWR_MAX, bool TW64, bool DBG, bool PAD>
    __device__ __forceinline__ void BondLabel(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ block_lat,
                                              const uint32_t total_blocks)
    {
        // All loads of a chunk of rows are issued before its first store: on this hardware the memory
        // counter that a load waits on also counts the stores issued before it.
        constexpr uint32_t kCh = TWR_MAX < 16 ? TWR_MAX : 16;
        constexpr uint32_t kRows = TWR_MAX < 16 ? TWR_MAX : 16;   // rows resolved together at the end
        constexpr uint32_t kSites = kWavesPerBlock * TWR_MAX * 64;
        constexpr uint32_t kQw = 128u;                             // per-wave LDS words of Q
        constexpr uint32_t kPFlag = 1u << 17;
        static_assert(kSites * 4u <= (1u << 16), "byte offsets must fit the 16-bit halves of a queue entry");
        static_assert((kPFlag >> 2) == kEdgeFlag, "a flagged root entry shifts into a flagged label");
        __shared__ uint32_t P[kSites];
        // per wave: 64-wide tiles keep the run minima of the current row in words 0-63 and give each lane a
        // word of its own in 64-127 (the target of LDS operations a lane has no use for, so that they need
        // no exec mask); narrow tiles keep their union queue here
        __shared__ uint32_t Q[kWavesPerBlock * kQw];
        __shared__ uint32_t FB[2u * kSites / 32u];   // flip bits of a pending flip, by label index (root or qcap + slot)
        __shared__ uint64_t s_first_B[kWavesPerBlock];
        __shared__ uint64_t AC[8];              // Numerical Recipes pairs (c << 32 | a) by selector
        __shared__ uint32_t s_nslot;            // slots given to the block's edge-reaching clusters

        const uint32_t b = RemapBlock(total_blocks);
        if (b >= total_blocks)
            return;
        if (threadIdx.x < 8u)
        {
            uint32_t a, c;
            ParamsOf(threadIdx.x, a, c);
            AC[threadIdx.x] = (static_cast<uint64_t>(c) << 32) | a;
        }
        if (threadIdx.x == 0)
            s_nslot = 0u;
        const LatDesc& d = descs[LatticeOf(block_lat, b)];
        const Block bl = DecodeBlock(d, b);
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t lane4 = lane * 4u;
        const uint32_t w = Uniform(threadIdx.x >> 6);

        const uint32_t n0 = d.n0, n1 = d.n1, twr = d.twr;
        const uint32_t tw = TW64 ? 64u : d.tw, rpw = TW64 ? 1u : d.rpw;
        const uint32_t lx = lane & (tw - 1u);
        const uint32_t lrow = TW64 ? 0u : lane / tw;
        const uint32_t x = bl.sx * tw + lx;
        const bool lastcol = PAD && bl.sx + 1u == d.tiles_x;
        const uint64_t vmask = lastcol ? d.vlast : ~0ull;   // lanes that hold sites
        const uint32_t elane = lastcol ? d.elast : tw - 1u;  // lx of a lattice row's last site in this tile
        const bool rbound = lx == elane;
        const bool injected = DBG && d.bond_in != nullptr;
        const uint32_t thr = d.threshold;
        const gptr<uint32_t> stp = Glob(d.state);
        const gptr<uint8_t> dump = DBG ? Glob(d.bond_dump) : nullptr;
        const gptr<uint8_t> spins = Glob(d.spins);
        const gptr<uint32_t> rotp = Glob(d.rot);
        const gptr<const uint8_t> bond_in = Glob(d.bond_in);

        const uint32_t r_begin = w * twr;
        const uint32_t nrow = bl.nwr > r_begin ? min(twr, bl.nwr - r_begin) : 0u;
        // buffer views of this block's part of the per-site arrays (row offsets in scalar registers, the lane
        // offset in a vector register, no per-row vector addresses)
        const uint32_t span = (bl.nwr - 1u) * bl.stride + 64u;   // sites from the block's first to its last
        const Rsrc rs_spins = MakeRsrc(d.spins + bl.base, span);
        const Rsrc rs_rng = MakeRsrc(d.rng + bl.base, span * 4u);
        const Rsrc rs_local = MakeRsrc(d.local + bl.base, span * 2u);
        const gptr<const uint32_t> fbits_g = Glob(static_cast<const uint32_t*>(d.fbits)) + bl.lb * d.fsw;
        const uint32_t nwaves = (bl.nwr + twr - 1u) / twr;
        const bool last_wave = w + 1u == nwaves;
        const uint32_t sx_r = bl.sx + 1u == d.tiles_x ? 0u : bl.sx + 1u;   // block column to the right
        const uint32_t sy_b = bl.sy + 1u == d.blocks_y ? 0u : bl.sy + 1u;   // block row below
        // the two sides of the halo array (by sweep parity)
        const gptr<const uint8_t> halo_s0 = Glob(static_cast<const uint8_t*>(d.halo));
        const gptr<const uint8_t> halo_s1 = halo_s0 + d.n_blocks * d.halo_stride;

        // Everything the block reads that needs only the descriptor is loaded together, before any of it is
        // used, so that the round trips overlap: the lattice's counters (read before this kernel stores any
        // state), the block's flip bits (they matter only under a pending flip: the edge-reaching clusters'
        // bits from k_resolve; before them in FB the rule words of the clusters inside the block, constant
        // under the parity rule) and every candidate for the spin below the tile (which one is used depends on
        // the counters). The descriptor fields these loads need are fetched first, in one batch, and the loads
        // carry no condition (in-range addresses, values unused where they do not apply), so that the compiler
        // neither fetches a field at its first use nor waits for every load at the end of a branch.
        HeldScalar(d.state);
        HeldScalar(d.fbits);
        HeldScalar(d.halo);
        HeldScalar(d.ecol_stride);
        HeldScalar(d.halo_stride);
        HeldScalar(d.qcap);
        HeldScalar(d.fsw);
        HeldScalar(d.n_blocks);
        const uint32_t it0 = injected ? 0u : stp[kStIteration];
        const uint32_t it_end = injected ? 0u : stp[kStItEnd];
        const uint32_t applied = injected ? 0u : stp[kStApplied];
        const uint32_t y_last = (bl.wyb + r_begin + (nrow > 0 ? nrow : 1u) - 1u) * rpw + lrow;
        const uint32_t y_below = nrow > 0 ? ((y_last + 1u == n1) ? 0u : y_last + 1u) : 0u;
        const uint32_t below = (r_begin + nrow) * bl.stride + lx;   // the site below the tile, inside the block
        const uint32_t below_halo = (sy_b * d.tiles_x + bl.sx) * d.halo_stride + d.ecol_stride + lx;
        const uint32_t hd_spin = spins[y_below * n0 + x];
        // (the label-only pass over injected bonds has no flip bits and no halos: those loads belong to the sweep,
        // where `injected` is false at compile time)
        uint32_t fbw = 0, hd_lab = 0, hd_sv = 0, hd_h0 = 0, hd_h1 = 0;
        if (!injected)
        {
            // fsw <= 256: a block has at most 2 * 64 + 2 * kWavesPerBlock * 64 boundary sites, 4 * edge_cap / 32
            // words of slots (the engine checks the bound), so one word per thread
            fbw = fbits_g[min(static_cast<uint32_t>(threadIdx.x), d.fsw - 1u)];
            hd_lab = __builtin_amdgcn_raw_buffer_load_b16(rs_local, below * 2u, 0, 0);
            hd_sv = __builtin_amdgcn_raw_buffer_load_b8(rs_spins, below, 0, 0);
            hd_h0 = halo_s0[below_halo];
            hd_h1 = halo_s1[below_halo];
            Issued(fbw);   // not sunk into the branch that stores it
        }

        // sweep schedule: a lattice past its last sweep is left alone by the whole sweep
        const bool act = injected || it0 < it_end;
        if (bl.lb == 0 && threadIdx.x == 0)
        {
            stp[kStActive] = act ? 1u : 0u;
            if (act && d.flip_rule != 0u)
                stp[kStFlipKey] = FlipKey(d.seed, it0);   // an inactive lattice keeps the key of its pending flip
            if (act && !injected)
                stp[kStHaloSide] = (it0 >> 1) & 1u;
        }
        if (!act)
            return;
        // The previous sweep's flip is still pending (applied != iteration): it is applied here, each site
        // as it is loaded, from its label and the block's flip bits; the flipped spins are stored as the
        // rows are processed. That flip belongs to the sweep that started at iteration it0 - 2.
        const bool pending = !injected && applied != it0;
        const uint32_t it1 = it0 + 1u;
        const uint32_t it2 = it1 + 1u;
        const bool rot1 = !injected && (it1 % 15u == 0u);
        const bool rot2 = !injected && (it2 % 15u == 0u);

        // halos: this sweep writes the block's first column and first wave row into its side of the halo
        // array; the other side holds the neighbours' flipped halos of the pending flip
        const uint32_t hside = (it0 >> 1) & 1u;
        const gptr<const uint8_t> halo_g = hside ? halo_s0 : halo_s1;
        const Rsrc rs_halo = MakeRsrc(d.halo + (hside * d.n_blocks + bl.lb) * d.halo_stride, d.halo_stride);
        const uint32_t hcol_off = (lx == 0u && !injected) ? lrow : kNoStore;   // first column: one lane per lattice row
        uint32_t* const Qw = Q + w * kQw;
        uint32_t* const Pl = At(P, lane4);   // this lane's entry of wave row 0
        uint32_t* const dummy = Qw + 64u + lane;
        // a full queue (>= 64 entries) is worked off right away; the tail moves to the front
        const auto drain64 = [&](uint32_t& qlen) {
            const uint32_t e = Qw[lane];
            LdsUnionLockstep(P, e >> 16, e & 0xFFFFu);
            const uint32_t rem = qlen - 64u;
            const uint32_t t = Qw[64u + lane];
            if (lane < rem)
                Qw[lane] = t;
            qlen = rem;
        };

        // lane masks: first column of a lattice row, lanes with a right neighbour inside the wave row,
        // lanes whose down neighbour sits in the same wave row
        uint64_t rowstart = 1;
        if (!TW64)
            for (uint32_t j = tw; j < 64u; j += tw)
                rowstart |= 1ull << j;
        const uint64_t redge = rowstart << elane;   // the last site of every lattice row of the wave row
        const uint64_t in_right = PAD ? (vmask & ~redge) : ~redge;
        const uint64_t low = TW64 ? 0ull : ((1ull << (64u - tw)) - 1ull);
        const uint64_t lanes_below = (1ull << lane) - 1ull;

        // pending flip: the block's flip bits (loaded into FB above), one per label index (k_resolve wrote the
        // bit of every cluster: by root index, or by qcap + slot for an edge-reaching one), so a site's flip is
        // the bit at its label's index
        const auto flip_of = [&](const uint32_t lab) { return (LdsLoad(&FB[(lab & 0x7FFFu) >> 5]) >> (lab & 31u)) & 1u; };

        // spins below the tile (down halo), needed last but fetched first. Under a pending flip, below the
        // block the neighbour's halo holds the flipped spin; inside the block the spin and label of the site
        // below are carried across the barrier in the lane's spare LDS word (writing them there needs the
        // loaded values, so these loads complete before the wave that owns the site may store its flipped
        // spin) and flipped after it, with the block's flip bits.
        uint32_t halo_down = 0;
        const bool carry = nrow > 0 && pending && !last_wave;
        // The block's flip bits and rule words, the carried site below and the barrier that publishes them: run
        // by every wave after the loads of its first chunk of rows are issued, so that this work and the
        // barrier overlap their round trip.
        const auto meet = [&]() {
            if (!injected)
            {
                if (threadIdx.x < d.fsw)
                    FB[d.qcap / 32u + threadIdx.x] = fbw;
                if (d.flip_rule == 0u)
                    for (uint32_t i = threadIdx.x; i < bl.nwr * 2u; i += 256u)   // bl.nwr * 64 root indices, 32 per word
                        FB[i] = RuleWord(d, bl, i, false, 0u);
                else if (pending)   // the random rule's words, with the key of the pending flip's sweep
                    for (uint32_t i = threadIdx.x; i < bl.nwr * 2u; i += 256u)
                        FB[i] = RuleWord(d, bl, i, true, FlipKey(d.seed, it0 - 2u));
            }
            if (nrow > 0)
            {
                if (!pending)
                    halo_down = hd_spin;
                else if (carry)
                    __hip_atomic_store(dummy, (hd_lab << 8) | (hd_sv & 0xFFu), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                else
                    halo_down = hside ? hd_h0 : hd_h1;
            }
            __syncthreads();   // AC, FB, the carried site below
            if (carry)
            {
                const uint32_t v = LdsLoad(dummy);
                const uint32_t lab = v >> 8;
                halo_down = (v & 0xFFu) ^ flip_of(lab);
            }
        };

        uint64_t last_B = 0, last_D = 0, down_bits = 0, right_bits = 0;
        uint32_t qlen = 0;
        uint32_t Rp = 0;   // 64-wide tiles: root of this lane's site in the previous row
        uint64_t prev_S = 0, prev_U2 = 0, prev_B = 0, prev_Ball = 0;
        uint32_t prev_site = 0;
        uint32_t colv = 0;   // 64-wide tiles: lane r collects the first-column spin of row r
        {
            // the chunk loop exists twice: sweeps without a group rotation (13 in 15) carry no rotation code
            // (and twice more for the sweeps that apply a pending flip on load, the rule after any sweep)
            const auto chunks = [&](auto rotating, auto flipping) {
            constexpr bool kRot = decltype(rotating)::value;
            constexpr bool kPend = decltype(flipping)::value;
            // rows loaded together: a flip on load also keeps the labels in flight, so half as many
            constexpr uint32_t kC = kPend && kCh > 1u ? kCh / 2u : kCh;
            // what one chunk of rows loads per lane: the spin byte, the lcg32 state and, under a pending flip,
            // the label of its site in each row; one right-halo spin and one rotation offset per row (lane i:
            // row r0 + i of the chunk)
            struct Chunk
            {
                uint32_t sp[kC], aux[kC], lab[kC], h, g_rot;
            };
            // The loads carry no condition (rows past the tile re-read its last row, values unused), so that
            // they issue back to back and their latencies overlap.
            const auto load = [&](const uint32_t r0, Chunk& c) {
                const uint32_t lrow_ch = r_begin + min(r0 + (lane % kC), nrow - 1u);
                // right halo of a 64-wide tile: under a pending flip the right neighbour's flipped first column
                c.h = 0;
                if (TW64)
                    c.h = pending ? halo_g[(bl.sy * d.tiles_x + sx_r) * d.halo_stride + lrow_ch]
                                  : spins[(bl.wyb + lrow_ch) * n0 + sx_r * 64u];
                // rotation offsets of the chunk's groups, needed on rotating sweeps only
                c.g_rot = kRot ? rotp[(bl.base + lrow_ch * bl.stride) >> 6] : 0u;
#pragma unroll
                for (uint32_t i = 0; i < kC; ++i)
                {
                    // row base in scalar registers, the lane offset in a vector register
                    const uint32_t wr = r_begin + min(r0 + i, nrow - 1u);
                    const uint32_t rowoff = Uniform(wr * bl.stride);
                    c.sp[i] = __builtin_amdgcn_raw_buffer_load_b8(rs_spins, lane, rowoff, 0);
                    c.aux[i] = injected ? (bond_in + bl.base + rowoff)[lane] : __builtin_amdgcn_raw_buffer_load_b32(rs_rng, lane4, rowoff * 4u, 0);
                    c.lab[i] = kPend ? __builtin_amdgcn_raw_buffer_load_b16(rs_local, lane * 2u, rowoff * 2u, 0) : 0u;
                }
            };
            const auto process = [&](const uint32_t r0, Chunk& c) {
                uint32_t(&sp)[kC] = c.sp;
                const uint32_t(&aux)[kC] = c.aux;
                const uint32_t(&lab)[kC] = c.lab;
                const uint32_t h = c.h, g_rot = c.g_rot;
                if (kPend)
                {
#pragma unroll
                    for (uint32_t i = 0; i < kC; ++i)
                        sp[i] ^= flip_of(lab[i]);
                }
                const uint64_t HR = TW64 ? __ballot((h & 1u) != 0u) : 0ull;

#pragma unroll
                for (uint32_t i = 0; i < kC; ++i)
                {
                    const uint32_t r = r0 + i;
                    if (r < nrow)
                    {
                        const uint32_t wr = r_begin + r;
                        const uint32_t site = bl.base + wr * bl.stride + lane;
                        const uint64_t S = __ballot((sp[i] & 1u) != 0u);
                        if (TW64)
                            colv = LaneBit(1ull << r) ? static_cast<uint32_t>(S) : colv;   // bit 0 of S: lane 0's spin
                        else
                            __builtin_amdgcn_raw_buffer_store_b8(static_cast<uint8_t>(sp[i] & 1u), rs_halo, hcol_off, wr * rpw, 0);
                        if (wr == 0 && !injected)
                            __builtin_amdgcn_raw_buffer_store_b8(static_cast<uint8_t>(sp[i] & 1u), rs_halo, d.ecol_stride + lane, 0, 0);
                        uint64_t Ball, U2;
                        if (!injected)
                        {
                            // two LCG32_State<64>::Update of the group of this wave row
                            // (random/src/lcg32.cpp:51-84); a rotation by `shift` moves lane
                            // (i + shift) & 63's selector to lane i
                            uint32_t ks = sp[i] & kSelMask, st = aux[i], shift = 0;
                            if (kRot && rot1)
                            {
                                shift = ShuffleShift(st, it1) & 63u;
                                ks = static_cast<uint32_t>(__shfl(static_cast<int>(ks), static_cast<int>((lane + shift) & 63u)));
                            }
                            uint64_t ac = AC[ks >> kSelShift];
                            st = static_cast<uint32_t>(ac) * st + static_cast<uint32_t>(ac >> 32);
                            const uint32_t u1 = st;
                            if (kRot && rot2)
                            {
                                const uint32_t s2 = ShuffleShift(st, it2) & 63u;
                                shift += s2;
                                ks = static_cast<uint32_t>(__shfl(static_cast<int>(ks), static_cast<int>((lane + s2) & 63u)));
                                ac = AC[ks >> kSelShift];
                            }
                            st = static_cast<uint32_t>(ac) * st + static_cast<uint32_t>(ac >> 32);
                            __builtin_amdgcn_raw_buffer_store_b32(st, rs_rng, lane4, Uniform(wr * bl.stride) * 4u, 0);
                            if (kRot)
                            {
                                spins[site] = static_cast<uint8_t>((sp[i] & 1u) | ks);
                                if (lane == 0)
                                    rotp[site >> 6] = (ReadLane(g_rot, i) + shift) & 63u;
                            }
                            else if (kPend)
                                __builtin_amdgcn_raw_buffer_store_b8(static_cast<uint8_t>(sp[i]), rs_spins, lane, Uniform(wr * bl.stride), 0);
                            const uint64_t U1 = __ballot(u1 < thr);
                            U2 = __ballot(st < thr);
                            // right neighbour's spin: next lane, or for a row's last site the halo (64-wide
                            // tiles) or the first column of the same lattice row (narrow tiles: one tile
                            // column)
                            uint64_t RS;
                            if (!PAD)
                                RS = TW64 ? ((S >> 1) | (((HR >> i) & 1ull) << 63))
                                          : (((S >> 1) & in_right) | ((S << (tw - 1u)) & ~in_right));
                            else
                                RS = TW64 ? (((S >> 1) & ~redge) | (((HR >> i) & 1ull) << elane))
                                          : (((S >> 1) & ~redge) | ((S << elane) & redge));
                            Ball = ~(S ^ RS) & U1;
                        }
                        else
                        {
                            Ball = __ballot((aux[i] & 1u) != 0u);
                            U2 = __ballot((aux[i] & 2u) != 0u);
                        }
                        if (PAD)
                        {
                            Ball &= vmask;
                            U2 &= vmask;
                        }
                        const uint64_t B = Ball & in_right;
                        const uint64_t edge = Ball & redge;
                        if (TW64)
                            right_bits |= (PAD ? ((edge >> elane) & 1ull) : (edge >> 63)) << r;
                        else
                            for (uint32_t j = 0; j < rpw; ++j)
                                right_bits |= ((edge >> (j * tw + elane)) & 1ull) << (r * rpw + j);
                        if (r == 0)
                            s_first_B[w] = B;

                        const uint32_t row_b = wr * 256u;   // byte offset of the wave row in P
                        if (TW64)
                        {
                            // slot of this lane's row run: the number of run starts after lane 0 up to this
                            // lane; the first lane of the run initialises it with its own node
                            uint32_t* const Mk = Qw + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(~B >> 32),
                                                                               __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(~B), 0u));
                            const uint32_t self = row_b + lane4;
                            __hip_atomic_store(LaneBit(~(B << 1)) ? Mk : dummy, self, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            // the row's entries start as roots (overwritten below), so chases end on them
                            __hip_atomic_store(At(P, self), self, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            const uint64_t D = r == 0 ? 0ull : (injected ? prev_U2 : (~(prev_S ^ S) & prev_U2));
                            if (r > 0)
                                LdsMin(LaneBit(D) ? Mk : dummy, Rp);
                            uint32_t root = LdsLoad(Mk);
                            // bonds to a root other than the run's minimum link that root under it
                            const uint64_t mg = D & __ballot(Rp != root);
                            if (mg != 0ull)
                            {
                                const uint32_t old = __hip_atomic_fetch_min(LaneBit(mg) ? At(P, Rp) : dummy, root, __ATOMIC_RELAXED,
                                                                            __HIP_MEMORY_SCOPE_WORKGROUP);
                                // Rp was no longer a root (another lane linked it first): join the two trees
                                const uint64_t more = mg & __ballot(old != Rp && old != root);
                                if (more != 0ull)
                                {
                                    const bool mo = LaneBit(more);
                                    LdsUnionLockstep(P, mo ? root : 0u, mo ? old : 0u);
                                }
                                // the run's minimum may itself have been linked under another root
                                for (;;)
                                {
                                    const uint32_t pr = LdsLoad(At(P, root));
                                    if (!__ballot(pr != root))
                                        break;
                                    root = pr;
                                }
                            }
                            if (dump && r > 0)
                                dump[prev_site] = static_cast<uint8_t>((LaneBit(prev_Ball) ? 1u : 0u) | (LaneBit(D) ? 2u : 0u));
                            *At(Pl, row_b) = root;
                            Rp = root;
                        }
                        else
                        {
                            // every site starts under the first site of its row run
                            const uint64_t below = ~B & lanes_below;
                            const uint32_t run_start = below ? 64u - static_cast<uint32_t>(__clzll(below)) : 0u;
                            *At(Pl, row_b) = row_b + run_start * 4u;
                            if (r > 0)
                            {
                                // bonds of the previous wave row to the row below: queue one union per overlap
                                // interval
                                const uint64_t SD = (Shr(prev_S, tw) & low) | Shl(S, 64u - tw);
                                const uint64_t D = injected ? prev_U2 : (~(prev_S ^ SD) & prev_U2);
                                const uint64_t tgtB = Shr(prev_B, tw) | Shl(B, 64u - tw);
                                const uint64_t K = FirstBondPerInterval(D, prev_B & tgtB, rowstart);
                                if (LaneBit(K))
                                {
                                    const uint32_t top = row_b - 256u + lane4;
                                    const uint32_t bot = (lane < 64u - tw) ? top + tw * 4u : row_b + lane4 - (64u - tw) * 4u;
                                    const uint32_t pos = qlen + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(K >> 32),
                                                                                         __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(K), 0u));
                                    Qw[pos] = (top << 16) | bot;
                                }
                                qlen += static_cast<uint32_t>(__popcll(K));
                                if (qlen >= 64u)
                                    drain64(qlen);
                                if (dump)
                                    dump[prev_site] = static_cast<uint8_t>((LaneBit(prev_Ball) ? 1u : 0u) | (LaneBit(D) ? 2u : 0u));
                            }
                        }
                        prev_S = S;
                        prev_U2 = U2;
                        prev_B = B;
                        prev_Ball = Ball;
                        prev_site = site;
                    }
                }
            };
            // a tile of at most kC rows is one chunk, starting at row 0; the first chunk's loads go out before
            // the barrier (a wave without rows loads nothing and only meets the others there)
            Chunk c0;
            if (nrow > 0)
                load(0u, c0);
            meet();
            if (nrow == 0)
                return;
            process(0u, c0);
            for (uint32_t r0 = kC; r0 < (TWR_MAX <= kC ? 0u : nrow); r0 += kC)
            {
                Chunk c;
                load(r0, c);
                process(r0, c);
            }
            };
            if (rot1 || rot2)
            {
                if (pending)
                    chunks(std::true_type{}, std::true_type{});
                else
                    chunks(std::true_type{}, std::false_type{});
            }
            else if (pending)
                chunks(std::false_type{}, std::true_type{});
            else
                chunks(std::false_type{}, std::false_type{});
        }
        if (nrow > 0)
        {
            if (TW64 && !injected)   // first column of the tile's rows (rpw == 1): one store per wave
                __builtin_amdgcn_raw_buffer_store_b8(static_cast<uint8_t>(colv & 1u), rs_halo, lane < nrow ? r_begin + lane : kNoStore, 0, 0);

            // bonds out of the last wave row: into the same wave row (tw < 64), into the next tile of the
            // block, or across the block edge
            {
                const uint32_t wr = r_begin + nrow - 1u;
                // halo spins arrive in the lanes that need them (lanes >= 64 - tw)
                const uint64_t H = __ballot((halo_down & 1u) != 0u);
                const uint64_t SD = TW64 ? H : ((Shr(prev_S, tw) & low) | (H & ~low));
                const uint64_t D = injected ? prev_U2 : (~(prev_S ^ SD) & prev_U2);
                if (!TW64)
                {
                    const uint64_t Din = D & low;
                    const uint64_t K = FirstBondPerInterval(Din, prev_B & Shr(prev_B, tw), rowstart);
                    if (LaneBit(K))
                    {
                        const uint32_t pos = qlen + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(K >> 32),
                                                                             __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(K), 0u));
                        const uint32_t top = wr * 256u + lane4;
                        Qw[pos] = (top << 16) | (top + tw * 4u);
                    }
                    qlen += static_cast<uint32_t>(__popcll(K));
                    if (qlen >= 64u)
                        drain64(qlen);
                }
                if (dump)
                    dump[prev_site] = static_cast<uint8_t>((LaneBit(prev_Ball) ? 1u : 0u) | (LaneBit(D) ? 2u : 0u));
                last_B = prev_B;
                if (last_wave)
                    down_bits = D & ~low;
                else
                    last_D = D & ~low;
            }

            // the rest of the queued unions of this tile (fewer than 64)
            if (!TW64 && qlen > 0)
            {
                const uint32_t e = (lane < qlen) ? Qw[lane] : 0u;
                LdsUnionLockstep(P, e >> 16, e & 0xFFFFu);
            }
        }
        __syncthreads();

        // unions across the tile edges inside the block
        if (nrow > 0 && !last_wave)
        {
            const uint32_t wr = r_begin + nrow - 1u;
            const uint64_t Bbot = s_first_B[w + 1u];
            const uint64_t tgtB = TW64 ? Bbot : Shl(Bbot, 64u - tw);
            const bool go = LaneBit(FirstBondPerInterval(last_D, last_B & tgtB, rowstart));
            const uint32_t top = wr * 256u + lane4;
            LdsUnionLockstep(P, go ? top : 0u, go ? top + 256u - (64u - tw) * 4u : 0u);
        }
        __syncthreads();

        // every site's block-local root (kRows chases in flight per lane). A tile of at most kRows rows
        // keeps its roots in registers; taller tiles write them back into P.
        const bool one_chunk = TWR_MAX <= kRows || nrow <= kRows;
        uint32_t xs[kRows];
        // (a wave of a partial block may own no rows: its part of P was never written and is not chased)
        for (uint32_t r0 = 0; r0 < (TWR_MAX <= kRows ? (nrow > 0 ? 1u : 0u) : nrow); r0 += kRows)
        {
#pragma unroll
            for (uint32_t i = 0; i < kRows; ++i)
                xs[i] = (r_begin + ((r0 + i < nrow) ? r0 + i : r0)) * 256u + lane4;
            // two steps per round; a round ends the chase once a step moved no lane of any row. The
            // differences are summed in vector registers (16 values below 2^17 cannot wrap to zero). Parents
            // only decrease, so a chase ends within kSites steps; more rounds mean a broken forest, which is
            // reported instead of followed.
            for (uint32_t round = 0;; ++round)
            {
                if (round * 2u > kSites)
                {
                    if (lane == 0)
                        (void)__hip_atomic_fetch_or(stp + kStErr, kErrLabelRange, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                uint32_t ys[kRows];
                uint32_t moved = 0;
#pragma unroll
                for (uint32_t i = 0; i < kRows; ++i)
                {
                    ys[i] = LdsLoad(At(P, xs[i]));
                    moved += ys[i] ^ xs[i];
                }
                if (!__ballot(moved != 0u))
                    break;
                moved = 0;
#pragma unroll
                for (uint32_t i = 0; i < kRows; ++i)
                {
                    xs[i] = LdsLoad(At(P, ys[i]));
                    moved += xs[i] ^ ys[i];
                }
                if (!__ballot(moved != 0u))
                    break;
            }
            if (!one_chunk)
            {
#pragma unroll
                for (uint32_t i = 0; i < kRows; ++i)
                    if (r0 + i < nrow)
                        *At(Pl, (r_begin + r0 + i) * 256u) = xs[i];
            }
        }
        // (taller tiles: a flagged entry is a root's own, first flagged by the edge marks, later holding its
        // slot label; either way the site is its own root)
        const auto root_of = [&](const uint32_t r) {
            if (one_chunk)
                return xs[r];
            const uint32_t v = LdsLoad(At(Pl, (r_begin + r) * 256u));
            return v >= kPFlag ? (r_begin + r) * 256u + lane4 : v;
        };
        __syncthreads();   // no chase may meet a flagged entry

        // Sites on the block edge flag their root (lanes off the edge flag their own word of Q). The edge
        // lanes of a row are `side`, plus every lane of the block's first and last lattice rows. The marks of
        // up to kRows rows are issued together and return the root's entry as it was: the lane that found it
        // unflagged (the cluster's first mark) gives the cluster the next slot of the block (one LDS counter,
        // one addition per wave and chunk of rows), writes the slot's union-find entry (a root keyed by the
        // root's site) and makes the root's entry in P kPFlag | (qcap + slot) << 2, so that after the barrier
        // every site of the cluster reads the label 0x8000 | (qcap + slot). A block has at most as many edge
        // roots as boundary sites, within its slot capacity; more would be a broken invariant, reported
        // through the error flag.
        const uint64_t side = TW64 ? (PAD ? (1ull | (1ull << elane)) : 0x8000000000000001ull) : __ballot(lx == 0 || rbound);
        const uint64_t top_first = r_begin == 0 ? (TW64 ? vmask : (__ballot(lrow == 0) & vmask)) : 0ull;
        const uint64_t bottom_last = last_wave ? (TW64 ? vmask : (__ballot(lrow == rpw - 1u) & vmask)) : 0ull;
        const uint32_t qcap = d.qcap, bslots = kWavesPerBlock * d.edge_cap;
        const Rsrc rs_cpar = MakeRsrc(d.cpar + bslots * bl.lb, bslots * 8u);
        // full tiles (the rule) take the row count and the first/last row masks at compile time
        const bool full = TWR_MAX <= kRows && nrow == kRows;
        const auto mark_rows = [&](auto full_t, const uint32_t r0) {
            constexpr bool kFull = decltype(full_t)::value;
            uint32_t old[kRows], root[kRows];
#pragma unroll
            for (uint32_t i = 0; i < kRows; ++i)
            {
                old[i] = kPFlag;
                root[i] = 0u;
                const uint32_t r = r0 + i;
                if (kFull || r < nrow)
                {
                    const bool last = kFull ? i + 1u == kRows : r + 1u == nrow;
                    const uint64_t m = side | (r == 0u ? top_first : 0ull) | (last ? bottom_last : 0ull);
                    root[i] = root_of(r);
                    const uint32_t o = __hip_atomic_fetch_or(LaneBit(m) ? At(P, root[i]) : dummy, kPFlag, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP);
                    old[i] = LaneBit(m) ? o : kPFlag;
                }
            }
            uint64_t F[kRows];
            uint32_t total = 0;
#pragma unroll
            for (uint32_t i = 0; i < kRows; ++i)
            {
                F[i] = __ballot(old[i] < kPFlag);
                total += static_cast<uint32_t>(__popcll(F[i]));
            }
            if (total == 0u)
                return;
            uint32_t base = 0;
            if (lane == 0u)
                base = __hip_atomic_fetch_add(&s_nslot, total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            base = __builtin_amdgcn_readfirstlane(base);
#pragma unroll
            for (uint32_t i = 0; i < kRows; ++i)
            {
                const bool first = LaneBit(F[i]);
                const uint32_t pos = base + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(F[i] >> 32),
                                                                      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(F[i]), 0u));
                __hip_atomic_store(first ? At(P, root[i]) : dummy, kPFlag | ((qcap + min(pos, bslots - 1u)) << 2),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                const u32x2 entry = {bslots * bl.lb + pos, SiteOf(bl.base, bl.stride, root[i] >> 2)};   // (site << 32) | slot
                __builtin_amdgcn_raw_buffer_store_b64(entry, rs_cpar, first ? pos * 8u : kNoStore, 0, 0);
                base += static_cast<uint32_t>(__popcll(F[i]));
            }
        };
        if (full)
            mark_rows(std::true_type{}, 0u);
        else if (one_chunk)
            mark_rows(std::false_type{}, 0u);
        else
            for (uint32_t r0 = 0; r0 < nrow; r0 += kRows)
                mark_rows(std::false_type{}, r0);
        __syncthreads();   // every slot label in place

        // labels out: the block-local labels, and (label << 16) | root index of the block's first and last
        // column and first and last wave row for the merge and the halo flips. The outputs go through buffer
        // views of this block's part of each array; a lane with nothing to store gets an offset past the
        // view, and the hardware drops that store (no exec mask).
        const Rsrc rs_ecol = MakeRsrc(d.ecol + 2u * bl.lb * d.ecol_stride, 8u * d.ecol_stride);
        const uint32_t ecol_off = LaneBit(side) ? ((rbound ? d.ecol_stride : 0u) + r_begin * rpw + lrow) * 4u : kNoStore;
        const Rsrc rs_erow = MakeRsrc(d.erow + 2u * 64u * bl.lb, 2u * 64u * 4u);
        const auto emit = [&](const uint32_t r, const uint32_t root) {
            const uint32_t wr = r_begin + r;
            const uint32_t rowoff = wr * bl.stride;   // sites from the block's first site
            const uint32_t lab = LdsLoad(At(P, root)) >> 2;
            const uint32_t e = (lab << 16) | (root >> 2);
            __builtin_amdgcn_raw_buffer_store_b16(static_cast<uint16_t>(lab), rs_local, lane * 2u, rowoff * 2u, 0);
            __builtin_amdgcn_raw_buffer_store_b32(e, rs_ecol, ecol_off, r * rpw * 4u, 0);
            if (wr == 0u)
                __builtin_amdgcn_raw_buffer_store_b32(e, rs_erow, lane4, 0, 0);
            if (wr + 1u == bl.nwr)
                __builtin_amdgcn_raw_buffer_store_b32(e, rs_erow, lane4, 256u, 0);
        };
        const auto rows_do = [&](auto full_t, auto&& f) {
            constexpr bool kFull = decltype(full_t)::value;
#pragma unroll
            for (uint32_t i = 0; i < kRows; ++i)
                if (kFull || i < nrow)
                    f(i, xs[i]);
        };
        const auto for_rows = [&](auto&& f) {
            if (full)
                rows_do(std::true_type{}, f);
            else if (one_chunk)
                rows_do(std::false_type{}, f);
            else
                for (uint32_t r = 0; r < nrow; ++r)
                    f(r, root_of(r));
        };
        for_rows(emit);
        const gptr<uint64_t> bb = Glob(d.bbits) + kBbitWords * bl.lb;
        if (lane == 0)
        {
            if (last_wave)
                bb[0] = down_bits;
            bb[1u + w] = right_bits;
        }
        if (threadIdx.x == 0)
        {
            const uint32_t n = LdsLoad(&s_nslot);
            Glob(d.edge_count)[kWavesPerBlock * bl.lb] = min(n, bslots);
            if (n > bslots)
                (void)__hip_atomic_fetch_or(stp + kStErr, kErrLabelRange, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    // 64-wide tiles of up to 16 rows: eight workgroups per CU (the register budget of 8 waves per SIMD is
    // asked for explicitly; the compiler keeps the excess scalar registers in vector lanes). Narrow tiles
    // keep the default budget (the cap would spill vector registers to scratch), and taller tiles are
    // limited by their LDS anyway.
    template <uint32_t TWR_MAX, bool DBG>
    __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_bond_label(
        const LatDesc* __restrict__ descs, const uint32_t* __restrict__ block_lat, const uint32_t total_blocks)
    {
        BondLabel<TWR_MAX, true, DBG, false>(descs, block_lat, total_blocks);
    }

    template <uint32_t TWR_MAX, bool TW64, bool DBG, bool PAD>
    __global__ __launch_bounds__(256) void k_bond_label_any(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ block_lat,
                                                            const uint32_t total_blocks)
    {
        BondLabel<TWR_MAX, TW64, DBG, PAD>(descs, block_lat, total_blocks);
    }

    // =============================================================================================
    // k_merge: unions across block edges (reference MergeLabels_Kernel, swendsen_wang_amd_gpu.cpp:280-329)
    // =============================================================================================
    // One two-wave workgroup per block: wavefront 0 takes the right edge (one lane per lattice row),
    // wavefront 1 the down edge (one lane per column). Two waves per block keep every block of a 4096^2
    // lattice resident at once, so the latency-bound union chains of all blocks overlap.
    // Bit `lane` of `bits`: a bond from a site of this block to a site of block lb_b (first site base_b), whose
    // edge entries (ecol / erow: (label << 16) | root index q) are va and vb. Both sites belong to clusters
    // that reach their block's edge, so their labels name slots.
    // Per wave, the distinct slots of one edge's root pairs in LDS: S[p] = lattice-wide slot + 1 of local node p
    // (0: free), K[p] its key; L[p] = (key of its parent << 32) | local parent (a root holds its own key and
    // position). Inserted by open addressing over 128 positions (at most 128 slots: every insert ends).
    __device__ __forceinline__ uint32_t LocalNode(uint32_t* S, uint32_t* K, uint64_t* L, const uint32_t c, const uint32_t k, const bool want)
    {
        uint32_t p = Fmix32(c * 0x85EBCA6Bu) & 127u;
        bool open = want;
        while (__ballot(open))
        {
            if (open)
            {
                uint32_t found = 0u;
                if (__hip_atomic_compare_exchange_strong(&S[p], &found, c + 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP))
                {
                    K[p] = k;
                    __hip_atomic_store(&L[p], SlotEntry(k, p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    open = false;
                }
                else if (found == c + 1u)
                    open = false;
                else
                    p = (p + 1u) & 127u;
            }
        }
        return p;
    }

    __device__ __forceinline__ uint64_t LocalEntry(uint64_t* L, const uint32_t p)
    {
        return __hip_atomic_load(&L[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }

    template <bool FRESH, bool LOCAL>
    __device__ __forceinline__ void MergeEdge(const LatDesc& d, const Block& bl, const uint64_t bits, const uint32_t va,
                                              const uint32_t vb, const uint32_t lb_b, const uint32_t base_b, gptr<uint32_t> err,
                                              uint64_t* H, uint32_t* S, uint32_t* K, uint32_t* W)
    {
        const uint32_t lane = threadIdx.x & 63u;
        const bool bit = LaneBit(bits);
        const uint32_t slots = kWavesPerBlock * d.edge_cap;
        uint32_t ka = 0, kb = 0, ca = 0, cb = 0;
        if (bit)
        {
            const uint32_t la = va >> 16, lb = vb >> 16;
            const uint32_t sa = (la & 0x7FFFu) - d.qcap, sb = (lb & 0x7FFFu) - d.qcap;
            if (!(la & kEdgeFlag) || !(lb & kEdgeFlag) || sa >= slots || sb >= slots)
                (void)__hip_atomic_fetch_or(err, kErrLabelRange, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ka = SiteOf(bl.base, bl.stride, va & 0xFFFFu);
            kb = SiteOf(base_b, bl.stride, vb & 0xFFFFu);
            ca = bl.lb * slots + min(sa, slots - 1u);
            cb = lb_b * slots + min(sb, slots - 1u);
        }
        // Without local unions, one lattice-wide union per distinct pair of roots (below T_c most bonds of an edge
        // join the same two clusters): the lanes insert their pairs into the wave's LDS hash set H (128 slots for
        // at most 64 pairs, linear probing, so every insert ends), and the lane whose insert lands leads the union
        // for its pair. With them, every bonded lane takes part in the local unions below, where a repeated pair
        // costs nothing.
        const uint32_t lo = min(ka, kb), hi = max(ka, kb);
        const bool want = bit && ka != kb;
        bool leader = LOCAL && want;
        if (!LOCAL && __ballot(want))
        {
            H[lane] = 0ull;
            H[64u + lane] = 0ull;
            const uint64_t key = (static_cast<uint64_t>(hi) << 32) | lo;   // hi > lo: never 0 (the empty slot)
            uint32_t slot = Fmix32(lo ^ (hi * 0x9E3779B9u)) & 127u;
            bool open = want;
            while (__ballot(open))
            {
                if (open)
                {
                    uint64_t found = 0ull;
                    if (__hip_atomic_compare_exchange_strong(&H[slot], &found, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP))
                    {
                        leader = true;
                        open = false;
                    }
                    else if (found == key)
                        open = false;
                    else
                        slot = (slot + 1u) & 127u;
                }
            }
        }
        if constexpr (!LOCAL)
        {
            if (leader)
                SlotUnion<FRESH>(Glob(d.cpar), ca, ka, cb, kb, d.n_blocks * slots, err);
            return;
        }
        if (!__ballot(leader))
            return;
        // The pairs of one edge often share a slot (a cluster of one block meets several of its neighbour's), and
        // lanes that link the same root race on one entry of the slot union-find. So the pairs are first united in
        // LDS, over their distinct slots (H holds the local entries here), and each slot that is not its local
        // component's root is then linked under that root: one lattice-wide union per lane, no two on one slot.
        S[lane] = 0u;
        S[64u + lane] = 0u;
        uint32_t a = LocalNode(S, K, H, ca, ka, leader);
        uint32_t b = LocalNode(S, K, H, cb, kb, leader);
        bool open = leader;
        while (__ballot(open))
        {
            if (open)
            {
                uint64_t ea = LocalEntry(H, a), eb = LocalEntry(H, b);
                while (static_cast<uint32_t>(ea) != a)
                {
                    a = static_cast<uint32_t>(ea) & 127u;
                    ea = LocalEntry(H, a);
                }
                while (static_cast<uint32_t>(eb) != b)
                {
                    b = static_cast<uint32_t>(eb) & 127u;
                    eb = LocalEntry(H, b);
                }
                if (a == b)
                    open = false;
                else
                {
                    if ((ea >> 32) > (eb >> 32))   // roots hold their own keys; link the larger under the smaller
                    {
                        const uint32_t t = a;
                        a = b;
                        b = t;
                        const uint64_t te = ea;
                        ea = eb;
                        eb = te;
                    }
                    const uint64_t old = __hip_atomic_fetch_min(&H[b], SlotEntry(static_cast<uint32_t>(ea >> 32), a),
                                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (old == eb)
                        open = false;
                    else   // b had been linked under old's parent (which this link may have displaced): unite with it
                        b = static_cast<uint32_t>(old) & 127u;
                }
            }
        }
        // the local nodes that are not their component's root, listed one per lane
        uint32_t item[2], root[2];
        bool has[2];
#pragma unroll
        for (uint32_t h = 0; h < 2u; ++h)
        {
            const uint32_t p = lane + 64u * h;
            uint32_t r = p;
            has[h] = false;
            if (S[p] != 0u)
            {
                uint64_t e = LocalEntry(H, r);
                while (static_cast<uint32_t>(e) != r)
                {
                    r = static_cast<uint32_t>(e) & 127u;
                    e = LocalEntry(H, r);
                }
                has[h] = r != p;
            }
            item[h] = p;
            root[h] = r;
        }
        const uint64_t f0 = __ballot(has[0]), f1 = __ballot(has[1]);
        const uint32_t n0 = static_cast<uint32_t>(__popcll(f0)), n = n0 + static_cast<uint32_t>(__popcll(f1));
        const uint32_t i0 = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(f0 >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(f0), 0u));
        const uint32_t i1 = n0 + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(f1 >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(f1), 0u));
        // at most one per local link, so at most one per pair: n <= 64
        if (has[0] && i0 < 64u)
            W[i0] = (root[0] << 8) | item[0];
        if (has[1] && i1 < 64u)
            W[i1] = (root[1] << 8) | item[1];
        if (lane < n && lane < 64u)
        {
            const uint32_t w = W[lane], p = w & 0xFFu, r = w >> 8;
            SlotUnion<FRESH>(Glob(d.cpar), S[p] - 1u, K[p], S[r] - 1u, K[r], d.n_blocks * slots, err);
        }
        if (n > 64u && lane == 0u)
            (void)__hip_atomic_fetch_or(err, kErrLabelRange, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    template <bool FRESH, bool LOCAL>
    __global__ __launch_bounds__(128) void k_merge(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ block_lat,
                                                   const uint32_t total_blocks)
    {
        const uint32_t b = RemapBlock(total_blocks);
        if (b >= total_blocks)
            return;
        const LatDesc& d = descs[LatticeOf(block_lat, b)];
        // the lattice's active flag is loaded with the edge data (an inactive lattice links nothing)
        const gptr<const uint32_t> active = Glob(static_cast<const uint32_t*>(d.state)) + kStActive;
        const Block bl = DecodeBlock(d, b);
        const gptr<uint64_t> bb = Glob(d.bbits) + kBbitWords * bl.lb;
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t w = Uniform(threadIdx.x >> 6);
        const uint32_t tw = d.tw, rpw = d.rpw, twr = d.twr;
        const gptr<uint32_t> err = Glob(d.state) + kStErr;
        __shared__ uint64_t HS[2][128];   // per wave: the hash set of root pairs (small launches) or the local entries
        __shared__ uint32_t SS[2][128], SK[2][128], SW[2][64];   // per wave: local slots, their keys, the union list

        if (w == 0)
        {
            // right edge: the block's lattice rows, 64 per pass (lane = row); bit r of word 1 + t is row
            // t * twr * rpw + r; entries from the compact edge columns, loaded with the bond words
            const uint32_t rows = bl.nwr * rpw, per_tile = twr * rpw;
            const uint32_t sx_r = (bl.sx + 1u == d.tiles_x) ? 0u : bl.sx + 1u;
            const uint32_t lb_r = bl.sy * d.tiles_x + sx_r;
            const uint32_t base_r = BlockBase(d, sx_r, bl.sy);
            const gptr<const uint32_t> ec = Glob(static_cast<const uint32_t*>(d.ecol));
            const gptr<const uint32_t> ea = ec + (2u * bl.lb + 1u) * d.ecol_stride;
            const gptr<const uint32_t> eb = ec + (2u * lb_r) * d.ecol_stride;
            const uint64_t r0 = bb[1], r1 = bb[2], r2 = bb[3], r3 = bb[4];
            for (uint32_t p0 = 0; p0 < rows; p0 += 64u)
            {
                const uint32_t row = min(p0 + lane, rows - 1u);
                const uint32_t va = ea[row], vb = eb[row];
                const uint32_t act = *active;
                Issued(r0);
                Issued(r1);
                Issued(r2);
                Issued(r3);
                Issued(va);
                Issued(vb);
                const uint32_t t = row / per_tile, r = row - t * per_tile;
                const uint64_t word = t == 0u ? r0 : (t == 1u ? r1 : (t == 2u ? r2 : r3));
                const uint64_t bits = __ballot(p0 + lane < rows && ((word >> r) & 1ull) != 0ull);
                if (bits && act != 0u)
                    MergeEdge<FRESH, LOCAL>(d, bl, bits, va, vb, lb_r, base_r, err, HS[0], SS[0], SK[0], SW[0]);
            }
        }
        else
        {
            // down edge: the lanes of the block's last lattice row (the last lattice row of its last wave row)
            // against the first lattice row of the block below (lanes 0 .. tw - 1 of its first wave row)
            const uint32_t sy_b = bl.sy + 1u == d.blocks_y ? 0u : bl.sy + 1u;
            const uint32_t lb_b = sy_b * d.tiles_x + bl.sx;
            const gptr<const uint32_t> er = Glob(static_cast<const uint32_t*>(d.erow));
            const uint32_t va = er[2u * 64u * bl.lb + 64u + lane], vb = er[2u * 64u * lb_b + (lane & (tw - 1u))];
            const uint64_t down_bits = bb[0];
            const uint32_t act = *active;
            Issued(va);
            Issued(vb);
            if (down_bits && act != 0u)
                MergeEdge<FRESH, LOCAL>(d, bl, down_bits, va, vb, lb_b, BlockBase(d, bl.sx, sy_b), err, HS[1], SS[1], SK[1], SW[1]);
        }
    }

    // =============================================================================================
    // k_resolve / k_flip: resolve + flip (reference ResolveLabels_Kernel / FlipClusters_Kernel,
    // swendsen_wang_amd_gpu.cpp:403-467), split into the decision (k_resolve, end of every sweep) and its
    // application to the spin array (the next sweep's k_bond_label, or k_flip when something reads spins)
    // =============================================================================================
    // Bitmap index of a label: the root index, or qcap + slot for an edge-reaching cluster. An index outside
    // the block's range (a broken invariant) clears ok and reads bit 0.
    __device__ __forceinline__ uint32_t LabelBit(const uint32_t lab, const uint32_t qmax, const uint32_t qcap, const uint32_t slots,
                                                 bool& ok)
    {
        const uint32_t i = lab & 0x7FFFu;
        const bool good = (lab & kEdgeFlag) ? (i >= qcap && i < qcap + slots) : (i < qmax);
        if (!good)
            ok = false;
        return good ? i : 0u;
    }

    // flip bit of a site with label lab from the block's bitmap
    __device__ __forceinline__ uint32_t FlipOfLabel(const uint32_t lab, const uint32_t* fb, const uint32_t qmax, const uint32_t qcap,
                                                    const uint32_t slots, bool& ok)
    {
        const uint32_t i = LabelBit(lab, qmax, qcap, slots, ok);
        return (fb[i >> 5] >> (i & 31u)) & 1u;
    }

    // The block's edge-root slots (the dense prefix k_bond_label filled), for i in [first, total) step
    // `stride`: f(slot within the block, lattice-wide slot).
    template <typename F>
    __device__ __forceinline__ void ForBlockSlots(const LatDesc& d, const Block& bl, const uint32_t first, const uint32_t stride, F&& f)
    {
        const uint32_t slots = kWavesPerBlock * d.edge_cap;
        const uint32_t total = min(Glob(d.edge_count)[kWavesPerBlock * bl.lb], slots);
        for (uint32_t i = first; i < total; i += stride)
            f(i, slots * bl.lb + i);
    }

    // End of a sweep, per block (two waves): the slots of the block's edge-reaching clusters are chased in the
    // slot union-find, and each slot's bit becomes the flip rule applied to its canonical root; then the
    // block's flip bitmap and the flipped spins of its first column and first wave row (the halos its left
    // and upper neighbours read in the next sweep) are written out, and the lattice's counters move on:
    // applied = the iteration whose spins the array holds, iteration += 2. The spin array itself is left for
    // the next sweep to flip.
    __global__ __launch_bounds__(128) void k_resolve(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ block_lat,
                                                     const uint32_t total_blocks)
    {
        constexpr uint32_t kThreads = 128;
        constexpr uint32_t kMaxSites = kWavesPerBlock * 64u * 64u;
        __shared__ uint32_t sw[kMaxSites / 32u];   // flip bits of the block's slots (at most one slot per site)
        const uint32_t b = RemapBlock(total_blocks);
        if (b >= total_blocks)
            return;
        const LatDesc& d = descs[LatticeOf(block_lat, b)];
        const gptr<uint32_t> stp = Glob(d.state);
        const Block bl = DecodeBlock(d, b);
        const uint32_t qmax = bl.nwr * 64u, qcap = d.qcap, slots = kWavesPerBlock * d.edge_cap;
        const gptr<const uint32_t> erow0 = Glob(static_cast<const uint32_t*>(d.erow)) + 2u * 64u * bl.lb;
        const gptr<const uint32_t> ecol0 = Glob(static_cast<const uint32_t*>(d.ecol)) + 2u * bl.lb * d.ecol_stride;
        bool ok = true;

        // halo sites: entry t < rows is the first column at lattice row t of the block, entry rows + j the
        // first wave row at lane j. Their labels and the block's slot count need only the descriptor and are
        // loaded with the lattice's counters; their spins (written into the halo by this sweep's
        // k_bond_label) follow with the first entries of the slot chases. The halo is flipped in place.
        const uint32_t rows = bl.nwr * d.rpw;
        // rows <= kWavesPerBlock * 64 (a tile has twr <= tw wave rows of rpw = 64 / tw lattice rows)
        constexpr uint32_t kHaloPer = (kWavesPerBlock * 64u + 64u + kThreads - 1u) / kThreads;
        // (straight-line loads from a selected address, every one in range, so that they are all in flight
        // at once; hoff is the entry's offset in the halo)
        uint32_t hs[kHaloPer], hl[kHaloPer], hoff[kHaloPer];
#pragma unroll
        for (uint32_t k = 0; k < kHaloPer; ++k)
        {
            const uint32_t t = threadIdx.x + k * kThreads;
            const bool col = t < rows;
            const uint32_t j = min(t - rows, 63u);
            hoff[k] = col ? t : d.ecol_stride + j;
            hl[k] = *(col ? ecol0 + t : erow0 + j) >> 16;
        }
        const uint32_t nslot = min(Glob(d.edge_count)[kWavesPerBlock * bl.lb], slots);
        const uint32_t active = stp[kStActive], hside = stp[kStHaloSide], fkey = stp[kStFlipKey];
#pragma unroll
        for (uint32_t k = 0; k < kHaloPer; ++k)
            Issued(hl[k]);
        Issued(nslot);
        Issued(hside);
        Issued(fkey);
        if (active == 0u)
            return;
        const bool random_rule = d.flip_rule != 0u;
        const uint32_t key = random_rule ? fkey : 0u;
        const gptr<uint8_t> halo = Glob(d.halo) + (hside * d.n_blocks + bl.lb) * d.halo_stride;
        const gptr<uint64_t> E = Glob(d.cpar);
        const uint32_t c_max = d.n_blocks * slots, c0 = slots * bl.lb;
        const uint64_t e_first = threadIdx.x < nslot ? CachedLoad64(&E[c0 + threadIdx.x]) : 0ull;
#pragma unroll
        for (uint32_t k = 0; k < kHaloPer; ++k)
            hs[k] = halo[hoff[k]];
        // each slot's canonical root, one slot per lane; a wave assembles the flip bits of 64 slots per pass
        // with a ballot and writes them as whole words, into LDS for the halo flips and into the block's fbits
        // (words past the last slot are left as they were: no label names a slot past the count)
        const gptr<uint32_t> fbits_g = Glob(d.fbits) + bl.lb * d.fsw;
        const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
        for (uint32_t s0 = 0; s0 < nslot; s0 += kThreads)
        {
            const uint32_t i = s0 + threadIdx.x;
            uint32_t bit = 0;
            if (i < nslot)
                bit = FlipBit(d, SlotRootKey(E, c0 + i, i == threadIdx.x ? e_first : CachedLoad64(&E[c0 + i]), c_max, ok),
                              random_rule, key);
            const uint64_t m = __ballot(bit != 0u);
            const uint32_t wd = (s0 + w * 64u) >> 5;
            if (lane == 0u && wd < d.fsw)
            {
                sw[wd] = static_cast<uint32_t>(m);
                fbits_g[wd] = static_cast<uint32_t>(m);
            }
            if (lane == 32u && wd + 1u < d.fsw)
            {
                sw[wd + 1u] = static_cast<uint32_t>(m >> 32);
                fbits_g[wd + 1u] = static_cast<uint32_t>(m >> 32);
            }
        }
        __syncthreads();

        // the flip of a halo site: its slot's bit, or for a cluster inside the block (a padding lane of the
        // first wave row) the rule at its root; a label out of range clears ok
        const auto halo_flip = [&](const uint32_t lab) -> uint32_t {
            const uint32_t i = lab & 0x7FFFu;
            if (lab & kEdgeFlag)
            {
                if (i < qcap || i - qcap >= nslot)
                {
                    ok = false;
                    return 0u;
                }
                return (sw[(i - qcap) >> 5] >> ((i - qcap) & 31u)) & 1u;
            }
            if (i >= qmax)
            {
                ok = false;
                return 0u;
            }
            return FlipBit(d, SiteOf(bl.base, bl.stride, i), random_rule, key);
        };
#pragma unroll
        for (uint32_t k = 0; k < kHaloPer; ++k)
        {
            const uint32_t t = threadIdx.x + k * kThreads;
            if (t < rows + 64u)
                halo[hoff[k]] = static_cast<uint8_t>((hs[k] ^ halo_flip(hl[k])) & 1u);
        }
        if (!ok)
            (void)__hip_atomic_fetch_or(stp + kStErr, kErrLabelRange, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (bl.lb == 0 && threadIdx.x == 0)
        {
            const uint32_t it = stp[kStIteration];
            stp[kStApplied] = it;
            stp[kStIteration] = it + 2u;
        }
    }

    // Applies a pending flip to the spin array (lattices without one return at once): per block (two
    // waves, every block of a 4096^2 lattice resident at once), the flip bits written by k_resolve and the
    // labels of the last sweep, four sites per lane. k_set_applied follows it.
    __global__ __launch_bounds__(128) void k_flip(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ block_lat,
                                                  const uint32_t total_blocks)
    {
        constexpr uint32_t kThreads = 128;
        constexpr uint32_t kPass = kThreads / 16;   // wave rows per pass of the workgroup (16 lanes per wave row)
        constexpr uint32_t kCh = 8;                 // passes loaded together
        constexpr uint32_t kMaxSites = kWavesPerBlock * 64u * 64u;
        __shared__ uint32_t fb[2u * kMaxSites / 32u];
        const uint32_t b = RemapBlock(total_blocks);
        if (b >= total_blocks)
            return;
        const LatDesc& d = descs[LatticeOf(block_lat, b)];
        const gptr<uint32_t> stp = Glob(d.state);
        if (stp[kStApplied] == stp[kStIteration])
            return;
        const Block bl = DecodeBlock(d, b);
        const gptr<uint16_t> local = Glob(d.local);
        const gptr<uint8_t> spins = Glob(d.spins);
        const uint32_t qmax = bl.nwr * 64u, qcap = d.qcap, slots = kWavesPerBlock * d.edge_cap;
        const uint32_t sub = threadIdx.x >> 4;
        const uint32_t col = (threadIdx.x & 15u) * 4u;
        bool ok = true;

        u32x2 lv[kCh];
        uint32_t sv[kCh];
        const auto load = [&](const uint32_t r0) {
#pragma unroll
            for (uint32_t k = 0; k < kCh; ++k)
            {
                const uint32_t r = r0 + sub + k * kPass;
                lv[k] = u32x2{0u, 0u};
                sv[k] = 0u;
                if (r < bl.nwr)
                {
                    const uint32_t site = bl.base + r * bl.stride + col;
                    lv[k] = *(gptr<const u32x2>)(local + site);
                    sv[k] = *(gptr<const uint32_t>)(spins + site);
                }
            }
        };
        load(0u);
        // the edge-reaching clusters' bits from k_resolve, and the rule words of the clusters inside the block
        // (the pending flip belongs to the sweep that started at iteration `applied`)
        const gptr<const uint32_t> fbits_g = Glob(static_cast<const uint32_t*>(d.fbits)) + bl.lb * d.fsw;
        for (uint32_t i = threadIdx.x; i < d.fsw; i += kThreads)
            fb[qcap / 32u + i] = fbits_g[i];
        const bool random_rule = d.flip_rule != 0u;
        const uint32_t key = random_rule ? FlipKey(d.seed, stp[kStApplied]) : 0u;
        for (uint32_t i = threadIdx.x; i < bl.nwr * 2u; i += kThreads)
            fb[i] = RuleWord(d, bl, i, random_rule, key);
        __syncthreads();

        for (uint32_t r0 = 0;;)
        {
#pragma unroll
            for (uint32_t k = 0; k < kCh; ++k)
            {
                const uint32_t r = r0 + sub + k * kPass;
                if (r < bl.nwr)
                {
                    const uint32_t site = bl.base + r * bl.stride + col;
                    const uint32_t lab[4] = {lv[k].x & 0xFFFFu, lv[k].x >> 16, lv[k].y & 0xFFFFu, lv[k].y >> 16};
                    uint32_t flip = 0;
#pragma unroll
                    for (uint32_t e = 0; e < 4; ++e)
                        flip |= FlipOfLabel(lab[e], fb, qmax, qcap, slots, ok) << (8u * e);
                    *(gptr<uint32_t>)(spins + site) = sv[k] ^ flip;
                }
            }
            r0 += kCh * kPass;
            if (r0 >= bl.nwr)
                break;
            load(r0);
        }
        if (!ok)
            (void)__hip_atomic_fetch_or(stp + kStErr, kErrLabelRange, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // After k_flip over a work list: every lattice's spin array holds its chain's state.
    __global__ __launch_bounds__(256) void k_set_applied(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ block_lat,
                                                         const uint32_t total_blocks)
    {
        const uint32_t b = blockIdx.x * 256u + threadIdx.x;
        if (b >= total_blocks)
            return;
        const LatDesc& d = descs[LatticeOf(block_lat, b)];
        if (b == d.block_base)
            d.state[kStApplied] = d.state[kStIteration];
    }

    __global__ void k_set_applied_one(uint32_t* __restrict__ state)
    {
        state[kStApplied] = state[kStIteration];
    }

    // Inspection kernel: the canonical label (minimum site index) of every site, written to parent[]: the
    // canonical root of every slot of the block (kept in LDS, one word per slot), then every site's label.
    __global__ __launch_bounds__(256) void k_write_labels(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ block_lat,
                                                          const uint32_t total_blocks)
    {
        constexpr uint32_t kPass = 16;   // wave rows per pass of the workgroup (16 lanes per wave row)
        constexpr uint32_t kCh = 4;      // passes loaded together
        extern __shared__ uint32_t R[];  // kWavesPerBlock * edge_cap entries (at most the block's sites)
        const uint32_t b = RemapBlock(total_blocks);
        if (b >= total_blocks)
            return;
        const LatDesc& d = descs[LatticeOf(block_lat, b)];
        const Block bl = DecodeBlock(d, b);
        const gptr<uint32_t> L = Glob(d.parent);
        const gptr<uint16_t> local = Glob(d.local);
        const uint32_t qmax = bl.nwr * 64u, qcap = d.qcap, slots = kWavesPerBlock * d.edge_cap;
        const uint32_t sub = threadIdx.x >> 4;
        const uint32_t col = (threadIdx.x & 15u) * 4u;
        bool ok = true;

        u32x2 lv[kCh];
        const auto load = [&](const uint32_t r0) {
#pragma unroll
            for (uint32_t k = 0; k < kCh; ++k)
            {
                const uint32_t r = r0 + sub + k * kPass;
                lv[k] = u32x2{0u, 0u};
                if (r < bl.nwr)
                    lv[k] = *(gptr<const u32x2>)(local + bl.base + r * bl.stride + col);
            }
        };
        load(0u);

        // phase 1: canonical root of every slot in use
        const gptr<uint64_t> E = Glob(d.cpar);
        const uint32_t c_max = d.n_blocks * slots;
        ForBlockSlots(d, bl, threadIdx.x, 256u, [&](const uint32_t slot, const uint32_t c) { R[slot] = SlotRootKey(E, c, CachedLoad64(&E[c]), c_max, ok); });
        __syncthreads();

        // phase 2: every site's label, four sites per lane
        for (uint32_t r0 = 0;;)
        {
#pragma unroll
            for (uint32_t k = 0; k < kCh; ++k)
            {
                const uint32_t r = r0 + sub + k * kPass;
                if (r < bl.nwr)
                {
                    const uint32_t site = bl.base + r * bl.stride + col;
                    const uint32_t lab[4] = {lv[k].x & 0xFFFFu, lv[k].x >> 16, lv[k].y & 0xFFFFu, lv[k].y >> 16};
                    uint32_t root[4];
#pragma unroll
                    for (uint32_t e = 0; e < 4; ++e)
                    {
                        const uint32_t i = LabelBit(lab[e], qmax, qcap, slots, ok);
                        root[e] = (lab[e] & kEdgeFlag) ? R[i >= qcap ? i - qcap : 0u] : SiteOf(bl.base, bl.stride, i);
                    }
                    *(gptr<u32x4>)(L + site) = u32x4{root[0], root[1], root[2], root[3]};
                }
            }
            r0 += kCh * kPass;
            if (r0 >= bl.nwr)
                break;
            load(r0);
        }
        if (!ok)
            (void)__hip_atomic_fetch_or(Glob(d.state) + kStErr, kErrLabelRange, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // =============================================================================================
    // Observables: per-tile partial sums of sum s(s_right + s_down) and sum s (lattice.cpp:29-67)
    // =============================================================================================
    // A conditional sample is due when the lattice has advanced since its last sample, is past its warmup
    // and sits on its own n_sep grid.
    __device__ __forceinline__ bool SampleDue(const uint32_t* st)
    {
        const uint32_t it = st[kStIteration], meas = st[kStItMeas], sep2 = st[kStSep2];
        return it > st[kStLastSample] && it > meas && sep2 != 0u && (it - meas) % sep2 == 0u;
    }

    __global__ __launch_bounds__(256) void k_observables(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ block_lat,
                                                         const uint32_t total_blocks, const uint32_t sample_mode)
    {
        const uint32_t b = RemapBlock(total_blocks);
        if (b >= total_blocks)
            return;
        const LatDesc& d = descs[LatticeOf(block_lat, b)];
        if (sample_mode == kSampleDue && !SampleDue(d.state))
            return;
        const Block bl = DecodeBlock(d, b);
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t w = Uniform(threadIdx.x >> 6);
        const uint32_t r_begin = w * d.twr;
        const uint32_t nrow = bl.nwr > r_begin ? min(d.twr, bl.nwr - r_begin) : 0u;
        const uint32_t pitch = d.n0, n0v = d.n0v, n1 = d.n1, tw = d.tw, rpw = d.rpw;
        const uint32_t x = bl.sx * tw + (lane & (tw - 1u));
        // padding columns (x >= n0v) hold no site; the right neighbour of column n0v - 1 is column 0
        const bool valid = x < n0v;
        const int32_t nv = __popcll(__ballot(valid));
        const uint32_t xr = (x + 1u >= n0v) ? 0u : x + 1u;
        int32_t e = 0, m = 0;
        for (uint32_t r = 0; r < nrow; ++r)
        {
            const uint32_t y = (bl.wyb + r_begin + r) * rpw + lane / tw;
            const uint32_t s = d.spins[y * pitch + x] & 1u;
            const uint32_t sr = d.spins[y * pitch + xr] & 1u;
            const uint32_t sd = d.spins[((y + 1u == n1) ? 0u : y + 1u) * pitch + x] & 1u;
            e += 2 * (__popcll(__ballot(valid && s == sr)) + __popcll(__ballot(valid && s == sd))) - 2 * nv;
            m += 2 * __popcll(__ballot(valid && s != 0)) - nv;
        }
        if (lane == 0)
        {
            const uint32_t t = kWavesPerBlock * bl.lb + w;
            d.partials[2u * t] = e;
            d.partials[2u * t + 1u] = m;
        }
    }

    // One workgroup per lattice: reduce the tile partials in a fixed order, append (E, M) to the series
    // and accumulate the FP64 moments (the only FP64 on the device).
    __global__ __launch_bounds__(256) void k_finalize_samples(const LatDesc* __restrict__ descs, const uint32_t sample_mode)
    {
        __shared__ long long red_e[256], red_m[256];
        const LatDesc& d = descs[blockIdx.x];
        if (sample_mode == kSampleDue && !SampleDue(d.state))
            return;
        long long e = 0, m = 0;
        for (uint32_t i = threadIdx.x; i < kWavesPerBlock * d.n_blocks; i += 256u)
        {
            e += d.partials[2u * i];
            m += d.partials[2u * i + 1u];
        }
        red_e[threadIdx.x] = e;
        red_m[threadIdx.x] = m;
        __syncthreads();
        for (uint32_t s = 128; s > 0; s >>= 1)
        {
            if (threadIdx.x < s)
            {
                red_e[threadIdx.x] += red_e[threadIdx.x + s];
                red_m[threadIdx.x] += red_m[threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0)
        {
            LatState* st = reinterpret_cast<LatState*>(d.state);
            const double n = static_cast<double>(d.n0v) * static_cast<double>(d.n1);
            const double ev = -static_cast<double>(red_e[0]) / n;
            const double mv = static_cast<double>(red_m[0]) / n;
            st->moments[0] += 1.0;
            st->moments[1] += ev;
            st->moments[2] += ev * ev;
            st->moments[3] += fabs(mv);
            st->moments[4] += mv * mv;
            st->moments[5] += (mv * mv) * (mv * mv);
            if (st->n_samples < st->series_cap)
            {
                long long* series = reinterpret_cast<long long*>(d.state + kLatStateWords);
                series[2u * st->n_samples] = red_e[0];
                series[2u * st->n_samples + 1u] = red_m[0];
            }
            st->n_samples += 1u;
            st->last_sample = st->iteration;
        }
    }

    // Fresh per-lattice state: counters and moments zero, no sweep limit unless it_end says otherwise.
    __global__ __launch_bounds__(256) void k_init_state(uint32_t* __restrict__ state, const uint32_t words, const uint32_t series_cap,
                                                        const uint32_t it_end, const uint32_t it_meas, const uint32_t sep2)
    {
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x)
        {
            uint32_t v = 0;
            if (i == 2u)
                v = series_cap;
            else if (i == kStItEnd)
                v = it_end;
            else if (i == kStActive)
                v = 1u;
            else if (i == kStItMeas)
                v = it_meas;
            else if (i == kStSep2)
                v = sep2;
            state[i] = v;
        }
    }

    // Schedule relative to the current iteration: `sweeps` more sweeps (0: no limit), the first `warmup` of
    // them unsampled, then a conditional sample after every n_sep-th.
    __global__ void k_set_schedule(uint32_t* __restrict__ state, const uint32_t sweeps, const uint32_t warmup, const uint32_t n_sep)
    {
        if (threadIdx.x != 0)
            return;
        const uint32_t it = state[kStIteration];
        state[kStItEnd] = sweeps ? it + 2u * sweeps : 0xFFFFFFFFu;
        state[kStItMeas] = it + 2u * warmup;
        state[kStSep2] = 2u * n_sep;
        state[kStLastSample] = it;
    }

    // Set-up of the allocation-free labelling (swmi_label_bonds_noalloc): the descriptor and fresh counters,
    // written by this kernel into the scratch memory.
    __global__ __launch_bounds__(64) void k_noalloc_setup(const LatDesc d, LatDesc* __restrict__ desc_out)
    {
        const uint32_t i = threadIdx.x;
        if (i == 0)
            *desc_out = d;
        if (i < kLatStateWords)
            d.state[i] = i == kStItEnd ? 0xFFFFFFFFu : (i == kStActive ? 1u : 0u);
    }

    // =============================================================================================
    // Initialisation
    // =============================================================================================
    // LCG32_State<64>::Init (random/src/lcg32.cpp:31-49) of every group, one thread per group, with the
    // POSIX drand48 recurrence in 48-bit integer arithmetic and IEEE double products.
    // The selector of lane i (bits 1-3 of the spin byte of site 64g+i) starts as sel0[i]; run after the
    // spins are set.
    __global__ __launch_bounds__(256) void k_init_rng(uint8_t* __restrict__ spins, uint32_t* __restrict__ rng, uint64_t* __restrict__ masks,
                                                      uint32_t* __restrict__ rot, const uint32_t n_groups, const uint32_t seed)
    {
        const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
        if (g >= n_groups)
            return;
        const uint64_t kA = 0x5DEECE66Dull, kC = 0xBull, kMask = (1ull << 48) - 1ull;
        uint64_t xs = (static_cast<uint64_t>(seed * 0x9E3779B1u + g + 1u) << 16) | 0x330Eull;
        uint64_t m0 = 0, m1 = 0, m2 = 0;
        for (uint32_t i = 0; i < 64u; ++i)
        {
            xs = (kA * xs + kC) & kMask;
            const double dr = static_cast<double>(xs) * 0x1p-48;
            const uint32_t k = static_cast<uint32_t>(__dmul_rn(1000.0, dr)) % 5u;
            m0 |= static_cast<uint64_t>(k & 1u) << i;
            m1 |= static_cast<uint64_t>((k >> 1) & 1u) << i;
            m2 |= static_cast<uint64_t>((k >> 2) & 1u) << i;
        }
        xs = (kA * xs + kC) & kMask;
        const double dr = static_cast<double>(xs) * 0x1p-48;
        uint32_t a, c;
        ParamsOf(static_cast<uint32_t>((m0 & 1ull) | ((m1 & 1ull) << 1) | ((m2 & 1ull) << 2)), a, c);
        uint32_t st = a * (static_cast<uint32_t>(__dmul_rn(4026531839.0, dr)) + 1u) + c;
        uint32_t* out = rng + static_cast<size_t>(g) * 64u;
        uint8_t* sp = spins + static_cast<size_t>(g) * 64u;
        out[0] = st;
        for (uint32_t i = 0; i < 64u; ++i)
        {
            const uint32_t k = static_cast<uint32_t>(((m0 >> i) & 1ull) | (((m1 >> i) & 1ull) << 1) | (((m2 >> i) & 1ull) << 2));
            sp[i] = static_cast<uint8_t>((sp[i] & 1u) | (k << kSelShift));
            if (i > 0)
            {
                ParamsOf(k, a, c);
                st = a * st + c;
                out[i] = st;
            }
        }
        masks[3u * g] = m0;
        masks[3u * g + 1u] = m1;
        masks[3u * g + 2u] = m2;
        rot[g] = 0;
    }

    // Hot start (ising/src/lattice.cpp:24-26): site i gets the (i+1)-th drand48 of srand48(spin_seed);
    // drand48() < 0.5 <=> bit 47 of the 48-bit state is clear. Each thread jumps ahead to its chunk. Site
    // i = y*n0v + x is stored at y*pitch + x.
    __global__ __launch_bounds__(256) void k_hot_start(uint8_t* __restrict__ spins, const uint64_t n, const uint32_t spin_seed,
                                                       const uint32_t n0v, const uint32_t pitch)
    {
        constexpr uint32_t kChunk = 64;
        const uint64_t first = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kChunk;
        if (first >= n)
            return;
        const uint64_t kMask = (1ull << 48) - 1ull;
        // (ja, jc) = affine map x -> A x + C raised to the power `first`
        uint64_t ja = 1, jc = 0, ba = 0x5DEECE66Dull, bc = 0xBull;
        for (uint64_t e = first; e; e >>= 1)
        {
            if (e & 1ull)
            {
                ja = (ba * ja) & kMask;
                jc = (ba * jc + bc) & kMask;
            }
            bc = (ba * bc + bc) & kMask;
            ba = (ba * ba) & kMask;
        }
        uint64_t xs = (static_cast<uint64_t>(spin_seed) << 16) | 0x330Eull;
        xs = (ja * xs + jc) & kMask;
        const uint64_t end = min(n, first + kChunk);
        for (uint64_t i = first; i < end; ++i)
        {
            xs = (0x5DEECE66Dull * xs + 0xBull) & kMask;
            const uint64_t at = n0v == pitch ? i : (i / n0v) * pitch + i % n0v;
            spins[at] = static_cast<uint8_t>((xs >> 47) & 1ull);
        }
    }

    // Upload of n row-major spins: replace bit 0 of every spin byte (site i = y*n0v + x at y*pitch + x), keep
    // the lcg32 selector bits.
    __global__ __launch_bounds__(256) void k_set_spins(uint8_t* __restrict__ spins, const uint8_t* __restrict__ src, const uint64_t n,
                                                       const uint32_t n0v, const uint32_t pitch)
    {
        const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i < n)
        {
            const uint64_t at = n0v == pitch ? i : (i / n0v) * pitch + i % n0v;
            spins[at] = static_cast<uint8_t>((spins[at] & 0xFEu) | (src[i] & 1u));
        }
    }

    // Row-major bytes (n0v wide) to the padded layout (pitch wide, padding columns 0), and back.
    __global__ __launch_bounds__(256) void k_pad_bytes(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, const uint64_t np,
                                                       const uint32_t n0v, const uint32_t pitch)
    {
        const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (p < np)
        {
            const uint64_t x = p % pitch;
            dst[p] = x < n0v ? src[(p / pitch) * n0v + x] : 0u;
        }
    }

    // Canonical padded labels (root's padded index per padded site) to row-major labels of row-major sites.
    __global__ __launch_bounds__(256) void k_unpad_labels(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, const uint64_t n,
                                                          const uint32_t n0v, const uint32_t pitch)
    {
        const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i < n)
        {
            const uint32_t root = src[(i / n0v) * pitch + i % n0v];
            dst[i] = (root / pitch) * n0v + root % pitch;
        }
    }

    __global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ labels, const uint64_t n)
    {
        const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i < n)
            labels[i] = static_cast<uint32_t>(i);
    }

    __global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, const uint64_t n16)
    {
        for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16;
             i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
            dst[i] = src[i];
    }

    // =============================================================================================
    // launch wrappers
    // =============================================================================================
    static inline uint32_t XcdGrid(const uint32_t units) { return ((units + kNumXcd - 1) / kNumXcd) * kNumXcd; }

    template <bool TW64, bool DBG>
    static void LaunchBondLabelT(const LatDesc* d_desc, const uint32_t* d_block_lat, const uint32_t total_blocks,
                                 const uint32_t max_twr, const bool pad, hipStream_t stream)
    {
        const uint32_t g = XcdGrid(total_blocks);
        const auto launch = [&](auto kernel) { kernel<<<g, 256, 0, stream>>>(d_desc, d_block_lat, total_blocks); };
        if (pad)   // padded lattices: two tile-height bounds
        {
            if (max_twr <= 16)
                launch(k_bond_label_any<16, TW64, DBG, true>);
            else
                launch(k_bond_label_any<64, TW64, DBG, true>);
        }
        else if (max_twr <= 8)
            TW64 ? launch(k_bond_label<8, DBG>) : launch(k_bond_label_any<8, false, DBG, false>);
        else if (max_twr <= 16)
            TW64 ? launch(k_bond_label<16, DBG>) : launch(k_bond_label_any<16, false, DBG, false>);
        else if (max_twr <= 32)
            launch(k_bond_label_any<32, TW64, DBG, false>);
        else
            launch(k_bond_label_any<64, TW64, DBG, false>);
    }

    void LaunchBondLabel(const LatDesc* d_desc, const uint32_t* d_block_lat, const uint32_t total_blocks, const uint32_t max_twr,
                         const bool all_tw64, const bool pad, const bool debug, hipStream_t stream)
    {
        if (debug)
        {
            if (all_tw64)
                LaunchBondLabelT<true, true>(d_desc, d_block_lat, total_blocks, max_twr, pad, stream);
            else
                LaunchBondLabelT<false, true>(d_desc, d_block_lat, total_blocks, max_twr, pad, stream);
        }
        else if (all_tw64)
            LaunchBondLabelT<true, false>(d_desc, d_block_lat, total_blocks, max_twr, pad, stream);
        else
            LaunchBondLabelT<false, false>(d_desc, d_block_lat, total_blocks, max_twr, pad, stream);
    }

    // Two-wave merge workgroups resident at once on MI355X: 256 CUs x 16 (8 waves per SIMD).
    constexpr uint32_t kOneRoundMergeBlocks = 4096;
    // Launches of at least this many blocks unite each edge's pairs in LDS before the lattice-wide unions: many
    // blocks race on the memory-side atomics there. Smaller launches see few races and only the added latency
    // (measured at 512 to 2048 blocks: 1024^2 and 2048^2 lattices are faster without).
    constexpr uint32_t kLocalUnionBlocks = 4096;

    template <bool FRESH>
    void LaunchMergeAs(const LatDesc* d_desc, const uint32_t* d_block_lat, const uint32_t total_blocks, hipStream_t stream)
    {
        if (total_blocks >= kLocalUnionBlocks)
            k_merge<FRESH, true><<<XcdGrid(total_blocks), 128, 0, stream>>>(d_desc, d_block_lat, total_blocks);
        else
            k_merge<FRESH, false><<<XcdGrid(total_blocks), 128, 0, stream>>>(d_desc, d_block_lat, total_blocks);
    }

    void LaunchMerge(const LatDesc* d_desc, const uint32_t* d_block_lat, const uint32_t total_blocks, hipStream_t stream)
    {
        if (total_blocks <= kOneRoundMergeBlocks)
            LaunchMergeAs<true>(d_desc, d_block_lat, total_blocks, stream);
        else
            LaunchMergeAs<false>(d_desc, d_block_lat, total_blocks, stream);
    }

    void LaunchResolve(const LatDesc* d_desc, const uint32_t* d_block_lat, const uint32_t total_blocks, hipStream_t stream)
    {
        k_resolve<<<XcdGrid(total_blocks), 128, 0, stream>>>(d_desc, d_block_lat, total_blocks);
    }

    void LaunchMaterialize(const LatDesc* d_desc, const uint32_t* d_block_lat, const uint32_t total_blocks, hipStream_t stream)
    {
        k_flip<<<XcdGrid(total_blocks), 128, 0, stream>>>(d_desc, d_block_lat, total_blocks);
        k_set_applied<<<(total_blocks + 255u) / 256u, 256, 0, stream>>>(d_desc, d_block_lat, total_blocks);
    }

    void LaunchWriteLabels(const LatDesc* d_desc, const uint32_t* d_block_lat, const uint32_t total_blocks, const uint32_t max_twr,
                           hipStream_t stream)
    {
        const size_t lds = static_cast<size_t>(kWavesPerBlock) * max_twr * 64u * sizeof(uint32_t);
        k_write_labels<<<XcdGrid(total_blocks), 256, lds, stream>>>(d_desc, d_block_lat, total_blocks);
    }

    void LaunchSetApplied(uint32_t* state, hipStream_t stream) { k_set_applied_one<<<1, 1, 0, stream>>>(state); }

    void LaunchNoAllocSetup(const LatDesc& d, LatDesc* d_desc, hipStream_t stream)
    {
        static_assert(kLatStateWords <= 64, "one wave writes the counters");
        k_noalloc_setup<<<1, 64, 0, stream>>>(d, d_desc);
    }

    void LaunchObservables(const LatDesc* d_desc, const uint32_t* d_block_lat, const uint32_t total_blocks, const uint32_t sample_mode,
                           hipStream_t stream)
    {
        k_observables<<<XcdGrid(total_blocks), 256, 0, stream>>>(d_desc, d_block_lat, total_blocks, sample_mode);
    }

    void LaunchFinalizeSamples(const LatDesc* d_desc, const uint32_t n_lattices, const uint32_t sample_mode, hipStream_t stream)
    {
        k_finalize_samples<<<n_lattices, 256, 0, stream>>>(d_desc, sample_mode);
    }

    void LaunchInitState(uint32_t* state, const uint32_t words, const uint32_t series_cap, const uint32_t it_end, const uint32_t it_meas,
                         const uint32_t sep2, hipStream_t stream)
    {
        k_init_state<<<1, 256, 0, stream>>>(state, words, series_cap, it_end, it_meas, sep2);
    }

    void LaunchSetSchedule(uint32_t* state, const uint32_t sweeps, const uint32_t warmup, const uint32_t n_sep, hipStream_t stream)
    {
        k_set_schedule<<<1, 64, 0, stream>>>(state, sweeps, warmup, n_sep);
    }

    void LaunchInitRng(uint8_t* spins, uint32_t* rng, uint64_t* masks, uint32_t* rot, const uint32_t n_groups, const uint32_t seed,
                       hipStream_t stream)
    {
        k_init_rng<<<(n_groups + 255) / 256, 256, 0, stream>>>(spins, rng, masks, rot, n_groups, seed);
    }

    void LaunchSetSpins(uint8_t* spins, const uint8_t* src, const uint64_t n, const uint32_t n0v, const uint32_t pitch,
                        hipStream_t stream)
    {
        k_set_spins<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, stream>>>(spins, src, n, n0v, pitch);
    }

    void LaunchPadBytes(uint8_t* dst, const uint8_t* src, const uint64_t np, const uint32_t n0v, const uint32_t pitch,
                        hipStream_t stream)
    {
        k_pad_bytes<<<static_cast<uint32_t>((np + 255) / 256), 256, 0, stream>>>(dst, src, np, n0v, pitch);
    }

    void LaunchUnpadLabels(uint32_t* dst, const uint32_t* src, const uint64_t n, const uint32_t n0v, const uint32_t pitch,
                           hipStream_t stream)
    {
        k_unpad_labels<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, stream>>>(dst, src, n, n0v, pitch);
    }

    void LaunchHotStart(uint8_t* spins, const uint64_t n, const uint32_t spin_seed, const uint32_t n0v, const uint32_t pitch,
                        hipStream_t stream)
    {
        const uint64_t threads = (n + 63) / 64;
        k_hot_start<<<static_cast<uint32_t>((threads + 255) / 256), 256, 0, stream>>>(spins, n, spin_seed, n0v, pitch);
    }

    void LaunchIota(uint32_t* labels, const uint64_t n, hipStream_t stream)
    {
        k_iota<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, stream>>>(labels, n);
    }

    void LaunchCopy(const uint4* src, uint4* dst, const uint64_t n16, hipStream_t stream)
    {
        k_copy<<<2048, 256, 0, stream>>>(src, dst, n16);
    }
}
