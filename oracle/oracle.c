/*
 * oracle.c -- CPU restatement of the reference Swendsen-Wang path (flwende/computational_physics).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the MI355X engine. The product never links or calls it.
 * Built with plain C11 and IEEE arithmetic (no -ffast-math): every operation whose rounding matters
 * (drand48 scaling, the float bond probability) is spelled out.
 *
 * Pinning (what checks each function, so that a plausible mistake fails one of them):
 *   or_srand48/or_drand48   glibc drand48/srand48 called through ctypes (tests/test_oracle.py)
 *   or_lcg32_*              golden streams of the compiled reference LCG32_State<16>/<64> and
 *                           LCG32<CPU>::NextInteger calls (tests/golden/lcg32_streams.npz,
 *                           lcg32_buffered.json, written by tests/golden/make_golden.py)
 *   or_hot_start            sweep-0 snapshots of the compiled reference Lattice<2> (refcpu.npz)
 *   or_p_add/or_threshold   reference expression compiled with the reference flags (padd.json) and an
 *                           exhaustive 2^32 check of the threshold against the float compare
 *   or_label                brute-force BFS / scipy connected components (min site index per cluster)
 *                           and the compiled reference trajectories
 *   or_refcpu_*             bit-exact spin trajectories of the compiled reference CPU path at
 *                           NUM_THREADS in {1,2,3,4,5,32} (tests/golden/refcpu.npz)
 *   or_observables          brute-force numpy sums
 *   or_ps_*                 composition of the pinned pieces above, plus the rotation invariant
 *                           checked against a full or_lcg32 per group
 *   or_murmur3_32/or_fmix32 MurmurHash3_x86_32 published test vectors (SMHasher); or_flip_random is
 *                           their composition with the pinned labeller
 *   or_refcpu_measure       report lines of the compiled reference driver (tests/golden/protocol.json)
 *                           and the reference's published sample (README.md:66-82)
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------------------
 * drand48 family, POSIX: X_{n+1} = (0x5DEECE66D * X_n + 0xB) mod 2^48, srand48 sets the high 32 bits
 * of X to the (32-bit) seed and the low 16 bits to 0x330E; drand48 returns X_{n+1} / 2^48.
 * Used by the reference at ising/src/lattice.cpp:24-26 and random/src/lcg32.cpp:35-43.
 * ------------------------------------------------------------------------------------------------ */
static const uint64_t RAND48_A = 0x5DEECE66DULL;
static const uint64_t RAND48_C = 0xBULL;
static const uint64_t RAND48_MASK = (1ULL << 48) - 1;

void or_srand48(uint64_t* x, uint32_t seed)
{
    *x = (((uint64_t)seed) << 16) | 0x330EULL;
}

double or_drand48(uint64_t* x)
{
    *x = (RAND48_A * (*x) + RAND48_C) & RAND48_MASK;
    return ldexp((double)(*x), -48); /* exact: 48 significant bits fit in a double */
}

/* ------------------------------------------------------------------------------------------------
 * lcg32 (random/src/lcg32.cpp:16-84). Five Numerical Recipes (a, c) pairs; each lane draws one at
 * Init; with RANDOM_SHUFFLE_STATE (defined by the reference build, ising/src/CMakeLists.txt:11) the
 * (a, c) arrays are rotated every 15 updates by an offset read from the state itself.
 * ------------------------------------------------------------------------------------------------ */
static const uint32_t LCG_PARAM[5][2] = {
    {1372383749U, 1289706101U},
    {2891336453U, 1640531513U},
    {2024337845U, 797082193U},
    {32310901U, 626627237U},
    {29943829U, 1013904223U},
};
enum { LCG_SHUFFLE_DISTANCE = 15 };

/* LCG32_State<W>::Init, random/src/lcg32.cpp:31-49. */
void or_lcg32_init(or_lcg32* s, uint32_t w, uint32_t seed)
{
    uint64_t x;
    memset(s, 0, sizeof(*s));
    s->w = w;
    or_srand48(&x, seed);
    for (uint32_t i = 0; i < w; ++i)
    {
        const uint32_t select = ((uint32_t)(1000.0 * or_drand48(&x))) % 5U;
        s->a[i] = LCG_PARAM[select][0];
        s->c[i] = LCG_PARAM[select][1];
    }
    s->state[0] = s->a[0] * ((uint32_t)(4026531839.0 /* 0xEFFFFFFF */ * or_drand48(&x)) + 1U) + s->c[0];
    for (uint32_t i = 1; i < w; ++i)
        s->state[i] = s->a[i] * s->state[i - 1] + s->c[i];
    s->iteration = 0;
}

/* LCG32_State<W>::Update, random/src/lcg32.cpp:51-84 (RANDOM_SHUFFLE_STATE branch). */
void or_lcg32_update(or_lcg32* s)
{
    const uint32_t w = s->w, m = w - 1;
    if ((++s->iteration) % LCG_SHUFFLE_DISTANCE == 0)
    {
        const uint32_t shift = s->state[s->iteration & m] + ((s->iteration & 1U) ? 0U : 1U);
        uint32_t buf[64];
        for (uint32_t i = 0; i < w; ++i)
            buf[i] = s->a[(i + shift) & m];
        memcpy(s->a, buf, w * sizeof(uint32_t));
        for (uint32_t i = 0; i < w; ++i)
            buf[i] = s->c[(i + shift) & m];
        memcpy(s->c, buf, w * sizeof(uint32_t));
    }
    for (uint32_t i = 0; i < w; ++i)
        s->state[i] = s->a[i] * s->state[i] + s->c[i];
}

/* LCG32<CPU>: random/include/random/lcg32.hpp:103-140 and random/src/lcg32.cpp:86-129. The member
 * 'current' starts at 0, so the first scalar draw returns state[1] of the initial state. */
void or_lcg32_cpu_init(or_lcg32_cpu* g, uint32_t w, uint32_t seed)
{
    or_lcg32_init(&g->s, w, seed);
    g->current = 0;
}

uint32_t or_lcg32_cpu_next_integer(or_lcg32_cpu* g) /* lcg32.cpp:86-95 */
{
    if ((++g->current) == g->s.w)
    {
        or_lcg32_update(&g->s);
        g->current = 0;
    }
    return g->s.state[g->current];
}

void or_lcg32_cpu_next_integers(or_lcg32_cpu* g, uint32_t* out, size_t n) /* lcg32.cpp:103-118 */
{
    const size_t w = g->s.w;
    const size_t i_max = (n / w) * w;
    for (size_t i = 0; i < i_max; i += w)
    {
        or_lcg32_update(&g->s);
        for (size_t ii = 0; ii < w; ++ii)
            out[i + ii] = g->s.state[ii];
    }
    g->current = (uint32_t)(w - 1);
    for (size_t i = i_max; i < n; ++i)
        out[i] = or_lcg32_cpu_next_integer(g);
}

void or_lcg32_cpu_next_reals(or_lcg32_cpu* g, float* out, size_t n) /* lcg32.cpp:120-129 */
{
    uint32_t* u = (uint32_t*)malloc(n * sizeof(uint32_t));
    or_lcg32_cpu_next_integers(g, u, n);
    for (size_t i = 0; i < n; ++i)
        out[i] = 2.3283064370807974E-10F * (float)u[i]; /* 2^-32, exact scaling */
    free(u);
}

/* ------------------------------------------------------------------------------------------------
 * Lattice hot start, ising/src/lattice.cpp:15-27.
 * ------------------------------------------------------------------------------------------------ */
void or_hot_start(uint8_t* spins, size_t n, uint32_t seed)
{
    uint64_t x;
    or_srand48(&x, seed);
    for (size_t i = 0; i < n; ++i)
        spins[i] = or_drand48(&x) < 0.5 ? 0 : 1;
}

/* ------------------------------------------------------------------------------------------------
 * Bond probability, ising/src/swendsen_wang.cpp:100: p_add = 1.0f - (float)std::exp(-2.0f / T), all
 * in float (std::exp(float) is glibc expf). The reference compares r = 2^-32f * (float)u against
 * p_add (swendsen_wang_selflabeling.hpp:85,100, swendsen_wang_cpu.cpp:78,90); the comparison is
 * monotone in u, so it equals u < threshold for the threshold found here.
 * ------------------------------------------------------------------------------------------------ */
float or_p_add(float temperature)
{
    const float q = -2.0f / temperature;
    const float e = expf(q);
    return 1.0f - e;
}

static int bond_draw(uint32_t u, float p_add)
{
    const float r = 2.3283064370807974E-10F * (float)u;
    return r < p_add;
}

uint64_t or_threshold(float p_add)
{
    /* smallest u in [0, 2^32] with !(r(u) < p_add); r is non-decreasing in u */
    uint64_t lo = 0, hi = 1ULL << 32;
    while (lo < hi)
    {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (mid < (1ULL << 32) && bond_draw((uint32_t)mid, p_add))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

uint64_t or_threshold_mismatches(float p_add, uint64_t threshold)
{
    uint64_t bad = 0;
    for (uint64_t u = 0; u < (1ULL << 32); ++u)
        bad += (uint64_t)(bond_draw((uint32_t)u, p_add) != (u < threshold));
    return bad;
}

/* ------------------------------------------------------------------------------------------------
 * Observables, ising/src/lattice.cpp:29-67: e = sum s (s_right + s_down), m = sum s, with s = 2b-1.
 * ------------------------------------------------------------------------------------------------ */
void or_observables(uint32_t n0, uint32_t n1, const uint8_t* spins, int64_t* energy, int64_t* magnetization)
{
    int64_t e = 0, m = 0;
    for (uint32_t y = 0; y < n1; ++y)
    {
        for (uint32_t x = 0; x < n0; ++x)
        {
            const int s = 2 * spins[(size_t)y * n0 + x] - 1;
            const int sr = 2 * spins[(size_t)y * n0 + (x + 1) % n0] - 1;
            const int sd = 2 * spins[(size_t)((y + 1) % n1) * n0 + x] - 1;
            e += s * (sr + sd);
            m += s;
        }
    }
    *energy = e;
    *magnetization = m;
}

/* ------------------------------------------------------------------------------------------------
 * Labeller: the reference CPU labeller on a given bond field.
 *   tile-local relaxation  swendsen_wang_selflabeling.hpp:106-192 (steps 3-5)
 *   cross-tile merge       swendsen_wang_cpu.cpp:50-105 (MergeLabels), :124-156 (Merge),
 *                          utility/atomic/atomic.hpp:12-33 (AtomicMin, here single-threaded)
 *   resolve                swendsen_wang_cpu.cpp:158-181
 * ------------------------------------------------------------------------------------------------ */
static uint32_t atomic_min_seq(uint32_t* ptr, uint32_t desired) /* atomic.hpp:12-33 */
{
    const uint32_t old = *ptr;
    if (old <= desired)
        return old;
    *ptr = desired;
    return old;
}

static void merge_seq(uint32_t* ptr, uint32_t a, uint32_t b) /* swendsen_wang_cpu.cpp:124-156 */
{
    for (;;)
    {
        const uint32_t c = atomic_min_seq(&ptr[b], a);
        if (c == a)
            break;
        if (c > a)
            b = c;
        if (c < a)
        {
            b = a;
            a = c;
        }
    }
}

static void tile_relax(uint32_t n0, uint32_t i, uint32_t j, uint32_t ii_max, uint32_t jj_max,
                       const uint8_t* bonds, uint32_t* labels, uint32_t* l, uint32_t* c, uint32_t* tmp)
{
    /* local bond bits: 0x2 = in-tile +x bond, 0x4 = in-tile +y bond (selflabeling.hpp:72-104) */
    for (uint32_t jj = 0; jj < jj_max; ++jj)
    {
        for (uint32_t ii = 0; ii < ii_max; ++ii)
        {
            const uint8_t b = bonds[(size_t)(j + jj) * n0 + i + ii];
            uint32_t v = 0;
            if (ii < ii_max - 1 && (b & 1))
                v |= 0x2;
            if (jj < jj_max - 1 && (b & 2))
                v |= 0x4;
            l[jj * ii_max + ii] = v;
            c[jj * ii_max + ii] = jj * ii_max + ii; /* step 3 */
        }
    }
    /* step 4 */
    int break_loop = 0;
    while (!break_loop)
    {
        break_loop = 1;
        for (uint32_t jj = 0; jj < jj_max; ++jj)
        {
            int label_changes = 1;
            while (label_changes)
            {
                for (uint32_t ii = 0; ii < ii_max; ++ii)
                    tmp[ii] = l[jj * ii_max + ii] & 0x2;
                label_changes = 0;
                for (uint32_t ii = 0; ii + 1 < ii_max; ++ii)
                {
                    if (tmp[ii])
                    {
                        const uint32_t a = c[jj * ii_max + ii], b = c[jj * ii_max + ii + 1];
                        if (a != b)
                        {
                            const uint32_t ab = a < b ? a : b;
                            c[jj * ii_max + ii] = ab;
                            c[jj * ii_max + ii + 1] = ab;
                            label_changes = 1;
                        }
                    }
                }
                if (label_changes)
                    break_loop = 0;
            }
            if (jj == jj_max - 1)
                continue;
            uint32_t counter = 0;
            for (uint32_t ii = 0; ii < ii_max; ++ii)
            {
                if (l[jj * ii_max + ii] & 0x4)
                {
                    const uint32_t a = c[jj * ii_max + ii], b = c[(jj + 1) * ii_max + ii];
                    if (a != b)
                    {
                        const uint32_t ab = a < b ? a : b;
                        c[jj * ii_max + ii] = ab;
                        c[(jj + 1) * ii_max + ii] = ab;
                        ++counter;
                    }
                }
            }
            if (counter)
                break_loop = 0;
        }
    }
    /* step 5: local -> global (selflabeling.hpp:182-192) */
    for (uint32_t jj = 0; jj < jj_max; ++jj)
    {
        for (uint32_t ii = 0; ii < ii_max; ++ii)
        {
            const uint32_t v = c[jj * ii_max + ii];
            const uint32_t a = v % ii_max, b = v / ii_max;
            labels[(size_t)(j + jj) * n0 + i + ii] = (j + b) * n0 + (i + a);
        }
    }
}

void or_label(uint32_t n0, uint32_t n1, uint32_t tile_x, uint32_t tile_y, const uint8_t* bonds, uint32_t* labels)
{
    const uint32_t ntx = (n0 + tile_x - 1) / tile_x, nty = (n1 + tile_y - 1) / tile_y;
    uint32_t* l = (uint32_t*)malloc((size_t)tile_x * tile_y * sizeof(uint32_t));
    uint32_t* c = (uint32_t*)malloc((size_t)tile_x * tile_y * sizeof(uint32_t));
    uint32_t* tmp = (uint32_t*)malloc((size_t)tile_x * sizeof(uint32_t));
    /* AssignLabels, swendsen_wang_cpu.cpp:13-48 */
    for (uint32_t n = 0; n < ntx * nty; ++n)
    {
        const uint32_t j = (n / ntx) * tile_y, i = (n % ntx) * tile_x;
        const uint32_t ii_max = tile_x < n0 - i ? tile_x : n0 - i;
        const uint32_t jj_max = tile_y < n1 - j ? tile_y : n1 - j;
        tile_relax(n0, i, j, ii_max, jj_max, bonds, labels, l, c, tmp);
    }
    /* MergeLabels, swendsen_wang_cpu.cpp:50-105: bottom row +y bonds, then right column +x bonds */
    for (uint32_t n = 0; n < ntx * nty; ++n)
    {
        const uint32_t j = (n / ntx) * tile_y, i = (n % ntx) * tile_x;
        const uint32_t ii_max = tile_x < n0 - i ? tile_x : n0 - i;
        const uint32_t jj_max = tile_y < n1 - j ? tile_y : n1 - j;
        for (uint32_t ii = 0; ii < ii_max; ++ii)
        {
            const size_t p = (size_t)(j + jj_max - 1) * n0 + i + ii;
            if (bonds[p] & 2)
            {
                const size_t q = (size_t)((j + jj_max) % n1) * n0 + i + ii;
                const uint32_t a = labels[p], b = labels[q];
                if (a != b)
                    merge_seq(labels, a, b);
            }
        }
        for (uint32_t jj = 0; jj < jj_max; ++jj)
        {
            const size_t p = (size_t)(j + jj) * n0 + i + ii_max - 1;
            if (bonds[p] & 1)
            {
                const size_t q = (size_t)(j + jj) * n0 + (i + ii_max) % n0;
                const uint32_t a = labels[p], b = labels[q];
                if (a != b)
                    merge_seq(labels, a, b);
            }
        }
    }
    /* ResolveLabels, swendsen_wang_cpu.cpp:158-181 */
    const size_t n = (size_t)n0 * n1;
    for (size_t s = 0; s < n; ++s)
    {
        uint32_t v = labels[s];
        while (v != labels[v])
            v = labels[v];
        labels[s] = v;
    }
    free(l);
    free(c);
    free(tmp);
}

/* FlipClusters, swendsen_wang_cpu.cpp:183-205: flip iff the canonical root index is odd. */
void or_flip(size_t n, uint8_t* spins, const uint32_t* labels)
{
    for (size_t i = 0; i < n; ++i)
        spins[i] ^= (uint8_t)(labels[i] & 1U);
}

/* ------------------------------------------------------------------------------------------------
 * The engine's random flip rule (no counterpart in the reference, whose rule is the root parity,
 * swendsen_wang_cpu.cpp:183-205): the cluster with canonical root r flips iff the top bit of
 * fmix32(r ^ key) is set, key = MurmurHash3_x86_32 of the 4-byte little-endian lcg32 iteration count
 * before the sweep, seeded with the lattice's lcg32 seed. fmix32 and MurmurHash3_x86_32 are Austin
 * Appleby's public-domain MurmurHash3 (SMHasher); they are pinned by its published test vectors.
 * ------------------------------------------------------------------------------------------------ */
uint32_t or_fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6BU;
    h ^= h >> 13;
    h *= 0xC2B2AE35U;
    h ^= h >> 16;
    return h;
}

static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

uint32_t or_murmur3_32(const uint8_t* data, size_t len, uint32_t seed)
{
    const uint32_t c1 = 0xCC9E2D51U, c2 = 0x1B873593U;
    uint32_t h = seed;
    const size_t nblocks = len / 4;
    for (size_t i = 0; i < nblocks; ++i)
    {
        const uint8_t* b = data + 4 * i;
        uint32_t k = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
        k *= c1;
        k = rotl32(k, 15);
        k *= c2;
        h ^= k;
        h = rotl32(h, 13);
        h = h * 5U + 0xE6546B64U;
    }
    const uint8_t* tail = data + 4 * nblocks;
    uint32_t k1 = 0;
    switch (len & 3U)
    {
    case 3:
        k1 ^= (uint32_t)tail[2] << 16;
        /* fall through */
    case 2:
        k1 ^= (uint32_t)tail[1] << 8;
        /* fall through */
    case 1:
        k1 ^= tail[0];
        k1 *= c1;
        k1 = rotl32(k1, 15);
        k1 *= c2;
        h ^= k1;
        break;
    default:
        break;
    }
    h ^= (uint32_t)len;
    return or_fmix32(h);
}

uint32_t or_flip_key(uint32_t seed, uint32_t iteration)
{
    const uint8_t b[4] = {(uint8_t)iteration, (uint8_t)(iteration >> 8), (uint8_t)(iteration >> 16), (uint8_t)(iteration >> 24)};
    return or_murmur3_32(b, 4, seed);
}

void or_flip_random(size_t n, uint8_t* spins, const uint32_t* labels, uint32_t key)
{
    for (size_t i = 0; i < n; ++i)
        spins[i] ^= (uint8_t)(or_fmix32(labels[i] ^ key) >> 31);
}

/* ------------------------------------------------------------------------------------------------
 * The reference CPU path SwendsenWang_2D<LCG32, CPU>.
 *   Initialize   swendsen_wang.cpp:41-91: NUM_THREADS generators, generator t seeded with t+1
 *   Update       swendsen_wang.cpp:94-112
 *   tile order   swendsen_wang_cpu.cpp:13-48: thread t owns tiles [t*chunk, min((t+1)*chunk, n)) in
 *                row-major tile order, chunk = ceil(n_tiles / NUM_THREADS)
 *   bond draws   selflabeling.hpp:64-104: per tile row one NextReal(tile_x) for the +x bonds (the last
 *                column never bonds in-tile), then per row but the last one NextReal(tile_x) for +y;
 *                swendsen_wang_cpu.cpp:68-104: per tile one NextReal(tile_x + tile_y), buf[ii] for
 *                the bottom-row +y bonds and buf[tile_x + jj] for the right-column +x bonds.
 * The merge order does not change the canonical labels, so threads are replayed one after another.
 * ------------------------------------------------------------------------------------------------ */
struct or_refcpu
{
    uint32_t n0, n1, w, num_threads, tile_x, tile_y;
    or_lcg32_cpu* rng;
    uint8_t* bonds;
    uint32_t* labels;
};

or_refcpu* or_refcpu_create(uint32_t n0, uint32_t n1, uint32_t w, uint32_t num_threads, uint32_t tile_x, uint32_t tile_y)
{
    if ((n0 % 2) || (n1 % 2) || w == 0 || w > 64 || (w & (w - 1)) || num_threads == 0)
        return NULL; /* swendsen_wang.cpp:47-48 rejects odd extents */
    or_refcpu* r = (or_refcpu*)calloc(1, sizeof(or_refcpu));
    r->n0 = n0;
    r->n1 = n1;
    r->w = w;
    r->num_threads = num_threads;
    r->tile_x = tile_x;
    r->tile_y = tile_y;
    r->rng = (or_lcg32_cpu*)malloc(num_threads * sizeof(or_lcg32_cpu));
    for (uint32_t t = 0; t < num_threads; ++t)
        or_lcg32_cpu_init(&r->rng[t], w, t + 1);
    r->bonds = (uint8_t*)malloc((size_t)n0 * n1);
    r->labels = (uint32_t*)malloc((size_t)n0 * n1 * sizeof(uint32_t));
    return r;
}

void or_refcpu_destroy(or_refcpu* r)
{
    if (!r)
        return;
    free(r->rng);
    free(r->bonds);
    free(r->labels);
    free(r);
}

void or_refcpu_sweep(or_refcpu* r, uint8_t* spins, float temperature, uint8_t* bonds_out, uint32_t* labels_out)
{
    const uint32_t n0 = r->n0, n1 = r->n1, tx = r->tile_x, ty = r->tile_y;
    const float p_add = or_p_add(temperature);
    const uint32_t ntx = (n0 + tx - 1) / tx, nty = (n1 + ty - 1) / ty, n_total = ntx * nty;
    const uint32_t chunk = (n_total + r->num_threads - 1) / r->num_threads;
    const size_t n = (size_t)n0 * n1;
    float* buf = (float*)malloc((size_t)(tx + ty) * sizeof(float));
    memset(r->bonds, 0, n);
#define SPIN(x, y) spins[(size_t)(y) * n0 + (x)]
    /* AssignLabels: in-tile bonds */
    for (uint32_t t = 0; t < r->num_threads; ++t)
    {
        const uint32_t start = t * chunk, end = (start + chunk < n_total) ? start + chunk : n_total;
        for (uint32_t k = start; k < end; ++k)
        {
            const uint32_t j = (k / ntx) * ty, i = (k % ntx) * tx;
            const uint32_t ii_max = tx < n0 - i ? tx : n0 - i, jj_max = ty < n1 - j ? ty : n1 - j;
            for (uint32_t jj = 0; jj < jj_max; ++jj)
            {
                or_lcg32_cpu_next_reals(&r->rng[t], buf, tx);
                for (uint32_t ii = 0; ii + 1 < ii_max; ++ii)
                    if (SPIN(i + ii, j + jj) == SPIN(i + ii + 1, j + jj) && buf[ii] < p_add)
                        r->bonds[(size_t)(j + jj) * n0 + i + ii] |= 1;
            }
            for (uint32_t jj = 0; jj + 1 < jj_max; ++jj)
            {
                or_lcg32_cpu_next_reals(&r->rng[t], buf, tx);
                for (uint32_t ii = 0; ii < ii_max; ++ii)
                    if (SPIN(i + ii, j + jj) == SPIN(i + ii, j + jj + 1) && buf[ii] < p_add)
                        r->bonds[(size_t)(j + jj) * n0 + i + ii] |= 2;
            }
        }
    }
    /* MergeLabels: tile-boundary bonds */
    for (uint32_t t = 0; t < r->num_threads; ++t)
    {
        const uint32_t start = t * chunk, end = (start + chunk < n_total) ? start + chunk : n_total;
        for (uint32_t k = start; k < end; ++k)
        {
            const uint32_t j = (k / ntx) * ty, i = (k % ntx) * tx;
            const uint32_t ii_max = tx < n0 - i ? tx : n0 - i, jj_max = ty < n1 - j ? ty : n1 - j;
            or_lcg32_cpu_next_reals(&r->rng[t], buf, tx + ty);
            for (uint32_t ii = 0; ii < ii_max; ++ii)
                if (buf[ii] < p_add && SPIN(i + ii, j + jj_max - 1) == SPIN(i + ii, (j + jj_max) % n1))
                    r->bonds[(size_t)(j + jj_max - 1) * n0 + i + ii] |= 2;
            for (uint32_t jj = 0; jj < jj_max; ++jj)
                if (buf[tx + jj] < p_add && SPIN(i + ii_max - 1, j + jj) == SPIN((i + ii_max) % n0, j + jj))
                    r->bonds[(size_t)(j + jj) * n0 + i + ii_max - 1] |= 1;
        }
    }
#undef SPIN
    free(buf);
    or_label(n0, n1, tx, ty, r->bonds, r->labels);
    or_flip(n, spins, r->labels);
    if (bonds_out)
        memcpy(bonds_out, r->bonds, n);
    if (labels_out)
        memcpy(labels_out, r->labels, n * sizeof(uint32_t));
}

/* ------------------------------------------------------------------------------------------------
 * Per-site lcg32 layout of the MI355X engine (no reference counterpart: it is the north-star layout,
 * built from LCG32_State<64> of random/src/lcg32.cpp:31-84). A rotation of (a, c) by s followed by a
 * rotation by t is a rotation by s + t, so a group keeps only its initial selectors and an offset.
 * ------------------------------------------------------------------------------------------------ */
uint32_t or_ps_pitch(uint32_t n0, uint32_t n1)
{
    if (n0 % 64U == 0U)
        return n0;
    if (n0 > 64U)
        return (n0 + 63U) / 64U * 64U;
    for (uint32_t p = 2U; p < 64U; p *= 2U)
        if (p >= n0 && n1 % (64U / p) == 0U)
            return p;
    return 64U;
}

uint32_t or_group_seed(uint32_t seed, uint32_t g)
{
    return seed * 0x9E3779B1U + g + 1U; /* seed 0 -> g + 1, the reference's "state i seeded with i+1" */
}

static uint32_t param_index(uint32_t a)
{
    for (uint32_t k = 0; k < 5; ++k)
        if (LCG_PARAM[k][0] == a)
            return k;
    abort();
}

void or_ps_init(uint32_t n_groups, uint32_t seed, uint32_t* states, uint64_t* masks, uint32_t* rot)
{
    or_lcg32 s;
    for (uint32_t g = 0; g < n_groups; ++g)
    {
        or_lcg32_init(&s, 64, or_group_seed(seed, g));
        uint64_t m0 = 0, m1 = 0, m2 = 0;
        for (uint32_t i = 0; i < 64; ++i)
        {
            const uint64_t k = param_index(s.a[i]);
            m0 |= (k & 1ULL) << i;
            m1 |= ((k >> 1) & 1ULL) << i;
            m2 |= ((k >> 2) & 1ULL) << i;
            states[(size_t)g * 64 + i] = s.state[i];
        }
        masks[3 * (size_t)g + 0] = m0;
        masks[3 * (size_t)g + 1] = m1;
        masks[3 * (size_t)g + 2] = m2;
        rot[g] = 0;
    }
}

void or_ps_rng_update(uint32_t n_groups, uint32_t* states, const uint64_t* masks, uint32_t* rot, uint32_t iteration_before)
{
    const uint32_t it = iteration_before + 1U;
    for (uint32_t g = 0; g < n_groups; ++g)
    {
        uint32_t* st = states + (size_t)g * 64;
        if (it % LCG_SHUFFLE_DISTANCE == 0)
        {
            const uint32_t shift = st[it & 63U] + ((it & 1U) ? 0U : 1U);
            rot[g] = (rot[g] + shift) & 63U;
        }
        const uint64_t m0 = masks[3 * (size_t)g], m1 = masks[3 * (size_t)g + 1], m2 = masks[3 * (size_t)g + 2];
        for (uint32_t i = 0; i < 64; ++i)
        {
            const uint32_t j = (i + rot[g]) & 63U;
            const uint32_t k = (uint32_t)(((m0 >> j) & 1ULL) | (((m1 >> j) & 1ULL) << 1) | (((m2 >> j) & 1ULL) << 2));
            st[i] = LCG_PARAM[k][0] * st[i] + LCG_PARAM[k][1];
        }
    }
}

void or_ps_bonds(uint32_t n0, uint32_t n1, const uint8_t* spins, const uint32_t* u1, const uint32_t* u2,
                 uint64_t threshold, uint8_t* bonds)
{
    for (uint32_t y = 0; y < n1; ++y)
    {
        for (uint32_t x = 0; x < n0; ++x)
        {
            const size_t i = (size_t)y * n0 + x;
            uint8_t b = 0;
            if (spins[i] == spins[(size_t)y * n0 + (x + 1) % n0] && (uint64_t)u1[i] < threshold)
                b |= 1;
            if (spins[i] == spins[(size_t)((y + 1) % n1) * n0 + x] && (uint64_t)u2[i] < threshold)
                b |= 2;
            bonds[i] = b;
        }
    }
}

void or_ps_sweep(uint32_t n0, uint32_t n1, uint64_t threshold, uint8_t* spins, uint32_t* states,
                 const uint64_t* masks, uint32_t* rot, uint32_t* iteration, uint8_t* bonds_out, uint32_t* labels_out)
{
    or_ps_sweep_rule(n0, n1, threshold, spins, states, masks, rot, iteration, bonds_out, labels_out, 0U, 0U);
}

void or_ps_sweep_rule(uint32_t n0, uint32_t n1, uint64_t threshold, uint8_t* spins, uint32_t* states,
                      const uint64_t* masks, uint32_t* rot, uint32_t* iteration, uint8_t* bonds_out, uint32_t* labels_out,
                      uint32_t flip_rule, uint32_t seed)
{
    const uint32_t key = or_flip_key(seed, *iteration);
    const size_t n = (size_t)n0 * n1;
    const uint32_t pitch = or_ps_pitch(n0, n1);
    const size_t npad = (size_t)pitch * n1;
    const uint32_t n_groups = (uint32_t)(npad / 64);
    uint32_t* u1 = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint32_t* u2 = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint8_t* bonds = (uint8_t*)malloc(n);
    uint32_t* labels = (uint32_t*)malloc(n * sizeof(uint32_t));
    /* the draws of site (x, y) are lane x % 64 of padded group (y * pitch + x) / 64; lanes of the padding
     * columns x >= n0 advance with their group and decide nothing */
    or_ps_rng_update(n_groups, states, masks, rot, *iteration);
    for (uint32_t y = 0; y < n1; ++y)
        for (uint32_t x = 0; x < n0; ++x)
            u1[(size_t)y * n0 + x] = states[(size_t)y * pitch + x];
    or_ps_rng_update(n_groups, states, masks, rot, *iteration + 1U);
    for (uint32_t y = 0; y < n1; ++y)
        for (uint32_t x = 0; x < n0; ++x)
            u2[(size_t)y * n0 + x] = states[(size_t)y * pitch + x];
    *iteration += 2U;
    or_ps_bonds(n0, n1, spins, u1, u2, threshold, bonds);
    or_label(n0, n1, 16, 8, bonds, labels);
    if (flip_rule == 1U)
        or_flip_random(n, spins, labels, key);
    else
        or_flip(n, spins, labels);
    if (bonds_out)
        memcpy(bonds_out, bonds, n);
    if (labels_out)
        memcpy(labels_out, labels, n * sizeof(uint32_t));
    free(u1);
    free(u2);
    free(bonds);
    free(labels);
}

/* ------------------------------------------------------------------------------------------------
 * The reference driver's measurement protocol (ising/src/ising_2d.cpp:122-162) on the restated CPU
 * path: warmup sweeps, then n_sep sweeps and a sample (e = -E/N, m = M/N, lattice.cpp:29-67) while
 * i < sweeps; FP64 sums of e, e^2 and |m|; C = (<e^2> - <e>^2) * N / (T*T with T float).
 * out = {<e>, C, <|m|>}.
 * ------------------------------------------------------------------------------------------------ */
void or_refcpu_measure(or_refcpu* r, uint8_t* spins, float temperature, uint32_t warmup, uint32_t sweeps, uint32_t n_sep,
                       double* out)
{
    const uint32_t n0 = r->n0, n1 = r->n1;
    const double n = (double)n0 * (double)n1;
    for (uint32_t i = 0; i < warmup; ++i)
        or_refcpu_sweep(r, spins, temperature, NULL, NULL);
    double energy = 0.0, energy_squared = 0.0, magnetization = 0.0;
    for (uint32_t i = 0; i < sweeps; i += n_sep)
    {
        for (uint32_t k = 0; k < n_sep; ++k)
            or_refcpu_sweep(r, spins, temperature, NULL, NULL);
        int64_t e = 0, m = 0;
        or_observables(n0, n1, spins, &e, &m);
        const double ev = -1.0 * (double)e / n, mv = 1.0 * (double)m / n;
        energy += ev;
        energy_squared += ev * ev;
        magnetization += fabs(mv);
    }
    const double samples = (double)(sweeps / n_sep);
    const float t2 = temperature * temperature;
    out[0] = energy / samples;
    out[1] = ((energy_squared / samples) - (energy / samples) * (energy / samples)) * n / (double)t2;
    out[2] = magnetization / samples;
}
