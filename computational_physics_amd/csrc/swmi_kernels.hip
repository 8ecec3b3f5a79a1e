// swmi_kernels.hip -- gfx950 kernels of the Swendsen-Wang engine.
//
// One sweep = three launches over the engine's tile work list (all lattices of the engine at once):
//   k_bond_label  per tile (one wavefront): two lcg32 updates per site, bond activation against the
//                 integer threshold, row runs from 64-bit ballots, vertical unions in an LDS
//                 union-find (atomicMin keeps parent <= child, so roots are cluster minima),
//                 tile-local labels and the tile-boundary bond bits to HBM.
//   k_merge       per tile: unions across tile edges by atomicMin on the uint32 labels in HBM.
//   k_flip        per tile: chase each site's label to its root, spin ^= root & 1; advances the lcg32
//                 iteration count.
// Reference behaviour: ising/src/swendsen_wang_amd_gpu.cpp:78-467 (four kernels per sweep) and the CPU
// labeller ising/src/swendsen_wang_cpu.cpp:13-205; the kernels here are written for wave64 CDNA4.
#include <hip/hip_runtime.h>

#include "swmi_internal.h"

namespace swmi
{
    // Numerical Recipes (a, c) pairs of random/src/lcg32.cpp:18-24.
    __device__ __forceinline__ void LcgParams(const uint32_t k, uint32_t& a, uint32_t& c)
    {
        a = 29943829u;
        c = 1013904223u;
        a = (k == 3u) ? 32310901u : a;
        c = (k == 3u) ? 626627237u : c;
        a = (k == 2u) ? 2024337845u : a;
        c = (k == 2u) ? 797082193u : c;
        a = (k == 1u) ? 2891336453u : a;
        c = (k == 1u) ? 1640531513u : c;
        a = (k == 0u) ? 1372383749u : a;
        c = (k == 0u) ? 1289706101u : c;
    }

    __device__ __forceinline__ uint32_t Uniform(const uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

    // One LCG32_State<64>::Update (random/src/lcg32.cpp:51-84) of the group held by this wavefront:
    // lane i owns state[i]; a[i] = Parameters[sel0[(i + rot) & 63]].
    __device__ __forceinline__ uint32_t LcgUpdate(uint32_t st, uint32_t& rot, const uint32_t it, const uint32_t lane,
                                                  const uint64_t m0, const uint64_t m1, const uint64_t m2)
    {
        if (it % 15u == 0u)
        {
            const uint32_t shift = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(st), static_cast<int>(it & 63u))) +
                ((it & 1u) ? 0u : 1u);
            rot = (rot + shift) & 63u;
        }
        const uint32_t j = (lane + rot) & 63u;
        const uint32_t k = static_cast<uint32_t>(((m0 >> j) & 1ull) | (((m1 >> j) & 1ull) << 1) | (((m2 >> j) & 1ull) << 2));
        uint32_t a, c;
        LcgParams(k, a, c);
        return a * st + c;
    }

    // ---- union-find in LDS (parents satisfy P[x] <= x) ----
    __device__ __forceinline__ uint32_t LdsLoad(const uint32_t* p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }

    __device__ __forceinline__ uint32_t LdsFind(const uint32_t* P, uint32_t x)
    {
        uint32_t p = LdsLoad(&P[x]);
        while (p != x)
        {
            x = p;
            p = LdsLoad(&P[x]);
        }
        return x;
    }

    // Same loop as the reference Merge (swendsen_wang_amd_gpu.cpp:368-401), started from roots.
    __device__ void LdsUnion(uint32_t* P, uint32_t a, uint32_t b)
    {
        for (;;)
        {
            a = LdsFind(P, a);
            b = LdsFind(P, b);
            if (a == b)
                return;
            if (a > b)
            {
                const uint32_t t = a;
                a = b;
                b = t;
            }
            const uint32_t old = __hip_atomic_fetch_min(&P[b], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            if (old == b)
                return;
            b = old;
        }
    }

    // ---- union-find in HBM (agent scope: the 8 XCD L2s are not coherent with each other) ----
    __device__ __forceinline__ uint32_t GlobalLoad(const uint32_t* p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // Parents satisfy L[x] <= x, so a chase ends within n steps; a value outside [0, n) can only come
    // from a broken invariant, which is reported through the lattice's error flag instead of being
    // dereferenced.
    __device__ __forceinline__ bool GlobalFind(const uint32_t* L, uint32_t& x, const uint32_t n)
    {
        uint32_t p = GlobalLoad(&L[x]);
        while (p != x)
        {
            if (p >= n || p > x)
                return false;
            x = p;
            p = GlobalLoad(&L[x]);
        }
        return true;
    }

    __device__ void GlobalUnion(uint32_t* L, uint32_t a, uint32_t b, const uint32_t n, uint32_t* err)
    {
        for (;;)
        {
            if (a >= n || b >= n || !GlobalFind(L, a, n) || !GlobalFind(L, b, n))
            {
                atomicOr(err, kErrLabelRange);
                return;
            }
            if (a == b)
                return;
            if (a > b)
            {
                const uint32_t t = a;
                a = b;
                b = t;
            }
            const uint32_t old = __hip_atomic_fetch_min(&L[b], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old == b)
                return;
            b = old;
        }
    }

    // ---- work decomposition ----
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of tiles
    // so that tiles that share cache lines (spin rows, halos) meet in the same L2.
    __device__ __forceinline__ bool TileOfWave(const uint32_t nblocks, const uint32_t total_tiles, uint32_t& t)
    {
        const uint32_t per = (nblocks + kNumXcd - 1) / kNumXcd;
        const uint32_t b = (blockIdx.x % kNumXcd) * per + blockIdx.x / kNumXcd;
        if (b >= nblocks)
            return false;
        t = Uniform(b * kWavesPerBlock + (threadIdx.x >> 6));
        return t < total_tiles;
    }

    struct Tile
    {
        uint32_t lt;       // tile index within its lattice
        uint32_t sx;       // tile column
        uint32_t wy0;      // first wave row
        uint32_t nrow;     // wave rows in this tile
        uint32_t stride;   // sites between consecutive wave rows (64 * tiles_x)
        uint32_t base0;    // first site of the tile
    };

    __device__ __forceinline__ Tile DecodeTile(const LatDesc& d, const uint32_t t)
    {
        Tile tl;
        tl.lt = t - d.tile_base;
        tl.sx = tl.lt % d.tiles_x;
        tl.wy0 = (tl.lt / d.tiles_x) * d.twr;
        tl.nrow = min(d.twr, d.nwr_y - tl.wy0);
        tl.stride = 64u * d.tiles_x;
        tl.base0 = tl.wy0 * tl.stride + tl.sx * 64u;
        return tl;
    }

    // =============================================================================================
    // k_bond_label
    // =============================================================================================
    template <uint32_t TWR_MAX>
    __global__ __launch_bounds__(256) void k_bond_label(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ tile_lat,
                                                        const uint32_t total_tiles, const uint32_t nblocks)
    {
        __shared__ uint32_t lds_parent[kWavesPerBlock][TWR_MAX * 64];
        uint32_t t;
        if (!TileOfWave(nblocks, total_tiles, t))
            return;
        const uint32_t lane = threadIdx.x & 63u;
        const LatDesc& d = descs[tile_lat[t]];
        const Tile tl = DecodeTile(d, t);
        uint32_t* P = lds_parent[threadIdx.x >> 6];

        const uint32_t n0 = d.n0, n1 = d.n1, tw = d.tw, rpw = d.rpw;
        const uint32_t lx = lane & (tw - 1u);
        const uint32_t lrow = lane / tw;
        const uint32_t x = tl.sx * tw + lx;
        const bool injected = d.bond_in != nullptr;
        const uint32_t thr = d.threshold;
        const uint32_t it0 = injected ? 0u : d.state[0];
        uint8_t* const dump = d.bond_dump;

        // lane masks: first column of a lattice row, lanes with an in-tile right neighbour, lanes whose
        // down neighbour sits in the same wave row
        uint64_t rowstart = 0;
        for (uint32_t j = 0; j < 64u; j += tw)
            rowstart |= 1ull << j;
        const uint64_t in_right = ~(rowstart << (tw - 1u));
        const uint64_t low = (tw == 64u) ? 0ull : ((1ull << (64u - tw)) - 1ull);
        const bool same_row_down = lane < 64u - tw;

        uint32_t prev_s = 0, prev_u2 = 0, prev_site = 0;
        bool prev_bx = false;
        uint64_t prev_B = 0, right_bits = 0;

        for (uint32_t r = 0; r < tl.nrow; ++r)
        {
            const uint32_t row_base = tl.base0 + r * tl.stride;
            const uint32_t site = row_base + lane;
            const uint32_t y = (tl.wy0 + r) * rpw + lrow;
            const uint32_t s = d.spins[site];
            uint32_t u1 = 0, u2 = 0;
            if (!injected)
            {
                const uint32_t grp = row_base >> 6;
                uint32_t rot = d.rot[grp];
                const uint32_t rot0 = rot;
                const uint64_t m0 = d.masks[3u * grp], m1 = d.masks[3u * grp + 1u], m2 = d.masks[3u * grp + 2u];
                uint32_t st = d.rng[site];
                st = LcgUpdate(st, rot, it0 + 1u, lane, m0, m1, m2);
                u1 = st;
                st = LcgUpdate(st, rot, it0 + 2u, lane, m0, m1, m2);
                u2 = st;
                d.rng[site] = st;
                if (rot != rot0 && lane == 0)
                    d.rot[grp] = rot;
            }

            // bond to (x+1, y)
            uint32_t rs = static_cast<uint32_t>(__shfl(static_cast<int>(s), static_cast<int>(lane + 1u)));
            const bool rbound = lx == tw - 1u;
            if (rbound)
                rs = d.spins[y * n0 + ((x + 1u == n0) ? 0u : x + 1u)];
            const bool bx = injected ? ((d.bond_in[site] & 1u) != 0) : ((s == rs) && (u1 < thr));
            const uint64_t b_all = __ballot(bx);
            const uint64_t B = b_all & in_right;
            const uint64_t b_edge = b_all & ~in_right;
            if (tw == 64u)
                right_bits |= (b_edge >> 63) << r;
            else
                for (uint32_t j = 0; j < rpw; ++j)
                    right_bits |= ((b_edge >> (j * tw + tw - 1u)) & 1ull) << (r * rpw + j);

            // every site starts under the first site of its row run
            const uint64_t below = ~B & ((1ull << lane) - 1ull);
            const uint32_t run_start = below ? 64u - static_cast<uint32_t>(__clzll(below)) : 0u;
            P[r * 64u + lane] = r * 64u + run_start;

            if (r > 0)
            {
                // bonds (x, y) -> (x, y+1) of wave row r-1
                const uint32_t sd_same = static_cast<uint32_t>(__shfl(static_cast<int>(prev_s), static_cast<int>(lane + tw)));
                const uint32_t sd_next = static_cast<uint32_t>(__shfl(static_cast<int>(s), static_cast<int>(lane - (64u - tw))));
                const uint32_t sd = same_row_down ? sd_same : sd_next;
                const bool by = injected ? ((d.bond_in[prev_site] & 2u) != 0) : ((prev_s == sd) && (prev_u2 < thr));
                const uint64_t D = __ballot(by);
                const uint64_t tgtB = (tw == 64u) ? B : ((prev_B >> tw) | (B << (64u - tw)));
                const uint64_t redundant = ((D & prev_B & tgtB) << 1) & ~rowstart;
                if (((D & ~redundant) >> lane) & 1ull)
                {
                    const uint32_t tq = same_row_down ? (r - 1u) * 64u + lane + tw : r * 64u + lane - (64u - tw);
                    LdsUnion(P, (r - 1u) * 64u + lane, tq);
                }
                if (dump)
                    dump[prev_site] = static_cast<uint8_t>((prev_bx ? 1u : 0u) | (by ? 2u : 0u));
            }
            prev_s = s;
            prev_u2 = u2;
            prev_B = B;
            prev_site = site;
            prev_bx = bx;
        }

        // bonds out of the last wave row: into the same wave row (tw < 64) or across the tile edge
        uint64_t down_bits;
        {
            const uint32_t r = tl.nrow - 1u;
            const uint32_t y = (tl.wy0 + r) * rpw + lrow;
            const uint32_t sd_same = static_cast<uint32_t>(__shfl(static_cast<int>(prev_s), static_cast<int>(lane + tw)));
            uint32_t sd = sd_same;
            if (!same_row_down)
                sd = d.spins[((y + 1u == n1) ? 0u : y + 1u) * n0 + x];
            const bool by = injected ? ((d.bond_in[prev_site] & 2u) != 0) : ((prev_s == sd) && (prev_u2 < thr));
            const uint64_t D = __ballot(by);
            const uint64_t Din = D & low;
            down_bits = D & ~low;
            const uint64_t tgtB = (tw == 64u) ? 0ull : (prev_B >> tw);
            const uint64_t redundant = ((Din & prev_B & tgtB) << 1) & ~rowstart;
            if (((Din & ~redundant) >> lane) & 1ull)
                LdsUnion(P, r * 64u + lane, r * 64u + lane + tw);
            if (dump)
                dump[prev_site] = static_cast<uint8_t>((prev_bx ? 1u : 0u) | (by ? 2u : 0u));
        }

        // tile-local roots -> global site indices (monotone map)
        for (uint32_t r = 0; r < tl.nrow; ++r)
        {
            const uint32_t q = LdsFind(P, r * 64u + lane);
            d.labels[tl.base0 + r * tl.stride + lane] = tl.base0 + (q >> 6) * tl.stride + (q & 63u);
        }
        if (lane == 0)
        {
            d.bbits[2u * tl.lt] = down_bits;
            d.bbits[2u * tl.lt + 1u] = right_bits;
        }
    }

    // =============================================================================================
    // k_merge: unions across tile edges (reference MergeLabels_Kernel, swendsen_wang_amd_gpu.cpp:280-329)
    // =============================================================================================
    __global__ __launch_bounds__(256) void k_merge(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ tile_lat,
                                                   const uint32_t total_tiles, const uint32_t nblocks)
    {
        uint32_t t;
        if (!TileOfWave(nblocks, total_tiles, t))
            return;
        const uint32_t lane = threadIdx.x & 63u;
        const LatDesc& d = descs[tile_lat[t]];
        const Tile tl = DecodeTile(d, t);
        const uint64_t down_bits = d.bbits[2u * tl.lt];
        const uint64_t right_bits = d.bbits[2u * tl.lt + 1u];
        if ((down_bits | right_bits) == 0ull)
            return;
        uint32_t* const L = d.labels;
        const uint32_t n0 = d.n0, n1 = d.n1, tw = d.tw, rpw = d.rpw;
        const uint32_t lx = lane & (tw - 1u);
        const uint32_t n = n0 * n1;
        uint32_t* const err = d.state + 3;

        // down edge: lanes of the last lattice row of the last wave row
        {
            const uint32_t r = tl.nrow - 1u;
            const uint32_t y = (tl.wy0 + r) * rpw + lane / tw;
            const uint32_t x = tl.sx * tw + lx;
            const bool bit = ((down_bits >> lane) & 1ull) != 0;
            uint32_t a = 0, b = 0;
            if (bit)
            {
                a = GlobalLoad(&L[tl.base0 + r * tl.stride + lane]);
                b = GlobalLoad(&L[((y + 1u == n1) ? 0u : y + 1u) * n0 + x]);
            }
            const uint32_t ap = static_cast<uint32_t>(__shfl(static_cast<int>(a), static_cast<int>(lane - 1u)));
            const uint32_t bp = static_cast<uint32_t>(__shfl(static_cast<int>(b), static_cast<int>(lane - 1u)));
            const bool dup = lx != 0 && ((down_bits >> (lane - 1u)) & 1ull) && ap == a && bp == b;
            if (bit && !dup && a != b)
                GlobalUnion(L, a, b, n, err);
        }
        // right edge: lane k = lattice row k of the tile
        {
            const uint32_t rows = tl.nrow * rpw;
            const uint32_t k = lane;
            const bool bit = k < rows && ((right_bits >> k) & 1ull) != 0;
            const uint32_t y = tl.wy0 * rpw + k;
            const uint32_t xr = tl.sx * tw + tw - 1u;
            uint32_t a = 0, b = 0;
            if (bit)
            {
                a = GlobalLoad(&L[y * n0 + xr]);
                b = GlobalLoad(&L[y * n0 + ((xr + 1u == n0) ? 0u : xr + 1u)]);
            }
            const uint32_t ap = static_cast<uint32_t>(__shfl(static_cast<int>(a), static_cast<int>(lane - 1u)));
            const uint32_t bp = static_cast<uint32_t>(__shfl(static_cast<int>(b), static_cast<int>(lane - 1u)));
            const bool dup = k > 0 && ((right_bits >> (k - 1u)) & 1ull) && ap == a && bp == b;
            if (bit && !dup && a != b)
                GlobalUnion(L, a, b, n, err);
        }
    }

    // =============================================================================================
    // k_flip: resolve + flip (reference ResolveLabels_Kernel / FlipClusters_Kernel,
    // swendsen_wang_amd_gpu.cpp:403-467)
    // =============================================================================================
    __global__ __launch_bounds__(256) void k_flip(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ tile_lat,
                                                  const uint32_t total_tiles, const uint32_t nblocks, const uint32_t flags)
    {
        uint32_t t;
        if (!TileOfWave(nblocks, total_tiles, t))
            return;
        const uint32_t lane = threadIdx.x & 63u;
        const LatDesc& d = descs[tile_lat[t]];
        const Tile tl = DecodeTile(d, t);
        uint32_t* const L = d.labels;
        for (uint32_t r = 0; r < tl.nrow; ++r)
        {
            const uint32_t site = tl.base0 + r * tl.stride + lane;
            uint32_t x = L[site];
            bool ok = x <= site;
            uint32_t p = ok ? L[x] : x;
            while (p != x)
            {
                if (p > x)
                {
                    ok = false;
                    break;
                }
                x = p;
                p = L[x];
            }
            if (!ok)
            {
                atomicOr(d.state + 3, kErrLabelRange);
                x = site;
            }
            if (flags & kFlip)
                d.spins[site] ^= static_cast<uint8_t>(x & 1u);
            if (flags & kWriteLabels)
                L[site] = x;
        }
        if ((flags & kAdvance) && tl.lt == 0 && lane == 0)
            d.state[0] += 2u;
    }

    // =============================================================================================
    // Observables: per-tile partial sums of sum s(s_right + s_down) and sum s (lattice.cpp:29-67)
    // =============================================================================================
    __global__ __launch_bounds__(256) void k_observables(const LatDesc* __restrict__ descs, const uint32_t* __restrict__ tile_lat,
                                                         const uint32_t total_tiles, const uint32_t nblocks)
    {
        uint32_t t;
        if (!TileOfWave(nblocks, total_tiles, t))
            return;
        const uint32_t lane = threadIdx.x & 63u;
        const LatDesc& d = descs[tile_lat[t]];
        const Tile tl = DecodeTile(d, t);
        const uint32_t n0 = d.n0, n1 = d.n1, tw = d.tw, rpw = d.rpw;
        const uint32_t x = tl.sx * tw + (lane & (tw - 1u));
        int32_t e = 0, m = 0;
        for (uint32_t r = 0; r < tl.nrow; ++r)
        {
            const uint32_t y = (tl.wy0 + r) * rpw + lane / tw;
            const uint32_t s = d.spins[y * n0 + x];
            const uint32_t sr = d.spins[y * n0 + ((x + 1u == n0) ? 0u : x + 1u)];
            const uint32_t sd = d.spins[((y + 1u == n1) ? 0u : y + 1u) * n0 + x];
            e += 2 * (__popcll(__ballot(s == sr)) + __popcll(__ballot(s == sd))) - 128;
            m += 2 * __popcll(__ballot(s != 0)) - 64;
        }
        if (lane == 0)
        {
            d.partials[2u * tl.lt] = e;
            d.partials[2u * tl.lt + 1u] = m;
        }
    }

    // One workgroup per lattice: reduce the tile partials in a fixed order, append (E, M) to the series
    // and accumulate the FP64 moments (the only FP64 on the device).
    __global__ __launch_bounds__(256) void k_finalize_samples(const LatDesc* __restrict__ descs)
    {
        __shared__ long long red_e[256], red_m[256];
        const LatDesc& d = descs[blockIdx.x];
        long long e = 0, m = 0;
        for (uint32_t i = threadIdx.x; i < d.n_tiles; i += 256u)
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
            const double n = static_cast<double>(d.n0) * static_cast<double>(d.n1);
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
        }
    }

    // =============================================================================================
    // Initialisation
    // =============================================================================================
    // LCG32_State<64>::Init (random/src/lcg32.cpp:31-49) of every group, one thread per group, with the
    // POSIX drand48 recurrence in 48-bit integer arithmetic and IEEE double products.
    __global__ __launch_bounds__(256) void k_init_rng(uint32_t* __restrict__ rng, uint64_t* __restrict__ masks, uint32_t* __restrict__ rot,
                                                      const uint32_t n_groups, const uint32_t seed)
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
        LcgParams(static_cast<uint32_t>((m0 & 1ull) | ((m1 & 1ull) << 1) | ((m2 & 1ull) << 2)), a, c);
        uint32_t st = a * (static_cast<uint32_t>(__dmul_rn(4026531839.0, dr)) + 1u) + c;
        uint32_t* out = rng + static_cast<size_t>(g) * 64u;
        out[0] = st;
        for (uint32_t i = 1; i < 64u; ++i)
        {
            LcgParams(static_cast<uint32_t>(((m0 >> i) & 1ull) | (((m1 >> i) & 1ull) << 1) | (((m2 >> i) & 1ull) << 2)), a, c);
            st = a * st + c;
            out[i] = st;
        }
        masks[3u * g] = m0;
        masks[3u * g + 1u] = m1;
        masks[3u * g + 2u] = m2;
        rot[g] = 0;
    }

    // Hot start (ising/src/lattice.cpp:24-26): site i gets the (i+1)-th drand48 of srand48(spin_seed);
    // drand48() < 0.5 <=> bit 47 of the 48-bit state is clear. Each thread jumps ahead to its chunk.
    __global__ __launch_bounds__(256) void k_hot_start(uint8_t* __restrict__ spins, const uint64_t n, const uint32_t spin_seed)
    {
        constexpr uint32_t kChunk = 64;
        const uint64_t first = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kChunk;
        if (first >= n)
            return;
        const uint64_t kMask = (1ull << 48) - 1ull;
        // (ja, jc) = affine map x -> A x + C raised to the power first
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
            spins[i] = static_cast<uint8_t>((xs >> 47) & 1ull);
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
    static inline uint32_t Blocks(const uint32_t total_tiles) { return (total_tiles + kWavesPerBlock - 1) / kWavesPerBlock; }
    static inline uint32_t Grid(const uint32_t nblocks) { return ((nblocks + kNumXcd - 1) / kNumXcd) * kNumXcd; }

    void LaunchBondLabel(const LatDesc* d_desc, const uint32_t* d_tile_lat, const uint32_t total_tiles, const uint32_t max_twr,
                         hipStream_t stream)
    {
        const uint32_t nb = Blocks(total_tiles);
        if (max_twr <= 8)
            k_bond_label<8><<<Grid(nb), 256, 0, stream>>>(d_desc, d_tile_lat, total_tiles, nb);
        else if (max_twr <= 16)
            k_bond_label<16><<<Grid(nb), 256, 0, stream>>>(d_desc, d_tile_lat, total_tiles, nb);
        else if (max_twr <= 32)
            k_bond_label<32><<<Grid(nb), 256, 0, stream>>>(d_desc, d_tile_lat, total_tiles, nb);
        else
            k_bond_label<64><<<Grid(nb), 256, 0, stream>>>(d_desc, d_tile_lat, total_tiles, nb);
    }

    void LaunchMerge(const LatDesc* d_desc, const uint32_t* d_tile_lat, const uint32_t total_tiles, hipStream_t stream)
    {
        const uint32_t nb = Blocks(total_tiles);
        k_merge<<<Grid(nb), 256, 0, stream>>>(d_desc, d_tile_lat, total_tiles, nb);
    }

    void LaunchFlip(const LatDesc* d_desc, const uint32_t* d_tile_lat, const uint32_t total_tiles, const uint32_t flags,
                    hipStream_t stream)
    {
        const uint32_t nb = Blocks(total_tiles);
        k_flip<<<Grid(nb), 256, 0, stream>>>(d_desc, d_tile_lat, total_tiles, nb, flags);
    }

    void LaunchObservables(const LatDesc* d_desc, const uint32_t* d_tile_lat, const uint32_t total_tiles, hipStream_t stream)
    {
        const uint32_t nb = Blocks(total_tiles);
        k_observables<<<Grid(nb), 256, 0, stream>>>(d_desc, d_tile_lat, total_tiles, nb);
    }

    void LaunchFinalizeSamples(const LatDesc* d_desc, const uint32_t n_lattices, hipStream_t stream)
    {
        k_finalize_samples<<<n_lattices, 256, 0, stream>>>(d_desc);
    }

    void LaunchInitRng(uint32_t* rng, uint64_t* masks, uint32_t* rot, const uint32_t n_groups, const uint32_t seed,
                       hipStream_t stream)
    {
        k_init_rng<<<(n_groups + 255) / 256, 256, 0, stream>>>(rng, masks, rot, n_groups, seed);
    }

    void LaunchHotStart(uint8_t* spins, const uint64_t n, const uint32_t spin_seed, hipStream_t stream)
    {
        const uint64_t threads = (n + 63) / 64;
        k_hot_start<<<static_cast<uint32_t>((threads + 255) / 256), 256, 0, stream>>>(spins, n, spin_seed);
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
