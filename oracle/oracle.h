/*
 * oracle.h -- CPU restatement of the reference Swendsen-Wang path (flwende/computational_physics).
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (the HIP library, the C++ host adapter, the
 * Python package) links, loads or calls this code. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every function states the reference file:line it restates. Paths are relative to the reference
 * repository root.
 */
#ifndef SWMI_ORACLE_H
#define SWMI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- POSIX drand48 family (used by ising/src/lattice.cpp:24-26 and random/src/lcg32.cpp:35-43) ---- */
void or_srand48(uint64_t* x, uint32_t seed);
double or_drand48(uint64_t* x);

/* ---- lcg32: LCG32_State<W> (random/include/random/lcg32.hpp:26-58, random/src/lcg32.cpp:31-84) ---- */
typedef struct
{
    uint32_t w;          /* lane count W (16 on an AVX-512 host, 64 on a wave64 GPU) */
    uint32_t iteration;  /* number of Update() calls so far */
    uint32_t state[64];
    uint32_t a[64];
    uint32_t c[64];
} or_lcg32;

void or_lcg32_init(or_lcg32* s, uint32_t w, uint32_t seed);
void or_lcg32_update(or_lcg32* s);

/* Buffered CPU generator LCG32<CPU> (random/src/lcg32.cpp:86-129). */
typedef struct
{
    or_lcg32 s;
    uint32_t current;
} or_lcg32_cpu;

void or_lcg32_cpu_init(or_lcg32_cpu* g, uint32_t w, uint32_t seed);
uint32_t or_lcg32_cpu_next_integer(or_lcg32_cpu* g);
void or_lcg32_cpu_next_integers(or_lcg32_cpu* g, uint32_t* out, size_t n);
void or_lcg32_cpu_next_reals(or_lcg32_cpu* g, float* out, size_t n);

/* ---- Lattice hot start (ising/src/lattice.cpp:15-27): srand48(seed); s[i] = drand48() < 0.5 ? 0 : 1 ---- */
void or_hot_start(uint8_t* spins, size_t n, uint32_t seed);

/* ---- bond probability and its exact integer threshold (ising/src/swendsen_wang.cpp:100) ---- */
float or_p_add(float temperature);
/* Smallest u with (2^-32f * (float)u) >= p_add, so that "r < p_add" <=> "u < threshold". */
uint64_t or_threshold(float p_add);
/* Number of draws u in [0, 2^32) for which (u < threshold) differs from the reference's float compare. */
uint64_t or_threshold_mismatches(float p_add, uint64_t threshold);

/* ---- labeller, flip, observables ----
 * Bond bytes: bit 0 = bond to (x+1 mod N0, y), bit 1 = bond to (x, y+1 mod N1). Site index y*N0+x. */
void or_label(uint32_t n0, uint32_t n1, uint32_t tile_x, uint32_t tile_y, const uint8_t* bonds, uint32_t* labels);
void or_flip(size_t n, uint8_t* spins, const uint32_t* labels);
/* engine random flip rule: MurmurHash3 finaliser of (root ^ key), key = MurmurHash3_x86_32(iteration, seed) */
uint32_t or_fmix32(uint32_t h);
uint32_t or_murmur3_32(const uint8_t* data, size_t len, uint32_t seed);
uint32_t or_flip_key(uint32_t seed, uint32_t iteration);
void or_flip_random(size_t n, uint8_t* spins, const uint32_t* labels, uint32_t key);
void or_observables(uint32_t n0, uint32_t n1, const uint8_t* spins, int64_t* energy, int64_t* magnetization);

/* ---- the reference CPU path, SwendsenWang_2D<LCG32, CPU> (ising/src/swendsen_wang.cpp:41-112,
 *      ising/src/swendsen_wang_cpu.cpp, ising/include/swendsen_wang_selflabeling.hpp) ---- */
typedef struct or_refcpu or_refcpu;
or_refcpu* or_refcpu_create(uint32_t n0, uint32_t n1, uint32_t w, uint32_t num_threads, uint32_t tile_x, uint32_t tile_y);
void or_refcpu_destroy(or_refcpu* r);
/* One sweep. bonds_out / labels_out may be NULL. */
void or_refcpu_sweep(or_refcpu* r, uint8_t* spins, float temperature, uint8_t* bonds_out, uint32_t* labels_out);
/* the reference driver's measurement loop (ising_2d.cpp:122-162): out = {<e>, C/N, <|m|>} */
void or_refcpu_measure(or_refcpu* r, uint8_t* spins, float temperature, uint32_t warmup, uint32_t sweeps, uint32_t n_sep,
                       double* out);

/* ---- the per-site lcg32 layout of the MI355X engine ----
 * The lattice is laid out with a row pitch P >= n0 (or_ps_pitch): site (x, y) is padded site y*P + x, and
 * the columns x in [n0, P) are padding. Group g = padded sites [64g, 64g+64), n_groups = P*n1/64; lane i of
 * the group = padded site 64g+i, so for P = n0 group g is row-major sites [64g, 64g+64). Each group is one
 * LCG32_State<64> seeded with or_group_seed(seed, g). Its (a,c) arrays are held as the initial
 * parameter selector of every lane (3 bit planes, masks[3g+0..2]) plus a rotation offset rot[g]:
 * a[i] = Parameters[sel0[(i + rot) & 63]].a. The iteration count is the same for every group. Padding
 * lanes advance with their group (the rotation may read their state) and decide no bond. */
/* P = n0 if n0 % 64 == 0; the next multiple of 64 if n0 > 64; else the smallest power of two p >= n0 with
 * n1 % (64/p) == 0 (64 at most), so that a 64-site group holds 64/p whole padded rows. */
uint32_t or_ps_pitch(uint32_t n0, uint32_t n1);
uint32_t or_group_seed(uint32_t seed, uint32_t g);
void or_ps_init(uint32_t n_groups, uint32_t seed, uint32_t* states, uint64_t* masks, uint32_t* rot);
/* One Update() of every group, iteration_before -> iteration_before + 1. */
void or_ps_rng_update(uint32_t n_groups, uint32_t* states, const uint64_t* masks, uint32_t* rot, uint32_t iteration_before);
/* Bond field from spins and two draws per site (u1 -> +x, u2 -> +y). */
void or_ps_bonds(uint32_t n0, uint32_t n1, const uint8_t* spins, const uint32_t* u1, const uint32_t* u2,
                 uint64_t threshold, uint8_t* bonds);
/* One sweep of the engine's layout (states, masks and rot of the P*n1/64 groups; spins, bonds and labels in
 * the unpadded row-major layout): two updates, bonds, labels, flip. *iteration += 2.
 * flip_rule 0: spin ^= root & 1 (reference). bonds_out / labels_out may be NULL. */
void or_ps_sweep(uint32_t n0, uint32_t n1, uint64_t threshold, uint8_t* spins, uint32_t* states,
                 const uint64_t* masks, uint32_t* rot, uint32_t* iteration, uint8_t* bonds_out, uint32_t* labels_out);
void or_ps_sweep_rule(uint32_t n0, uint32_t n1, uint64_t threshold, uint8_t* spins, uint32_t* states,
                      const uint64_t* masks, uint32_t* rot, uint32_t* iteration, uint8_t* bonds_out, uint32_t* labels_out,
                      uint32_t flip_rule, uint32_t seed);

#ifdef __cplusplus
}
#endif

#endif
