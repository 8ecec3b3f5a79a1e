/*
 * swmi.h -- C ABI of the MI355X-native Swendsen-Wang engine for the 2D Ising model.
 *
 * Drop-in for the Ising2D --algorithm=swendsen_wang --rng=lcg32 path of flwende/computational_physics.
 * Plain C: opaque handles, plain pointers and sizes, int status codes (SWMI_OK = 0), no exceptions and
 * no torch types. The last error message of the calling thread is swmi_last_error().
 *
 * Conventions shared with the reference:
 *   spins   uint8 in {0,1} (1 = up), row-major, site index y*n0 + x          (ising/include/lattice.hpp:24)
 *   labels  uint32, the minimum row-major site index of the cluster          (ising/src/swendsen_wang_cpu.cpp:124-181)
 *   flip    spin ^= label & 1 (flip_rule SWMI_FLIP_PARITY)                   (ising/src/swendsen_wang_cpu.cpp:183-205)
 *   energy  E = sum s(s_right + s_down), M = sum s with s = 2*spin - 1;
 *           per-spin values are -E/N and M/N                                 (ising/src/lattice.cpp:29-67)
 *   bonds   p_add = 1.0f - (float)expf(-2.0f / T), bond iff spins equal and r < p_add,
 *           evaluated as u < threshold on the uint32 draw u                  (ising/src/swendsen_wang.cpp:100)
 * lcg32 layout (new, the engine's own): the lattice is held with its rows padded to a pitch P >= n0
 * (swmi_get_layout): P = n0 if n0 is a multiple of 64 or a divisor of 64 with n1 a multiple of 64/n0,
 * else the next multiple of 64 (n0 > 64) or the smallest power of two p >= n0 with n1 % (64/p) == 0.
 * Site (x, y) is padded site y*P + x. Group g = padded sites [64g, 64g+64) is one LCG32_State<64>
 * (random/src/lcg32.cpp:31-84) seeded with seed*0x9E3779B1 + g + 1; each sweep draws twice per site,
 * the first draw decides the bond to (x+1, y), the second the bond to (x, y+1). The lanes of padding
 * columns draw with their group and decide nothing. Every array that crosses the ABI (spins, bonds,
 * labels) is unpadded row-major; only swmi_rng_export and checkpoints hold the padded layout.
 */
#ifndef SWMI_H
#define SWMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum
{
    SWMI_OK = 0,
    SWMI_ERR_INVALID = 1,     /* bad argument (extent, temperature, index, size) */
    SWMI_ERR_HIP = 2,         /* HIP runtime error */
    SWMI_ERR_NO_DEVICE = 3,   /* no usable gfx950 device */
    SWMI_ERR_MEMORY = 4,      /* device memory or budget exhausted */
    SWMI_ERR_STATE = 5        /* call not valid in the current state */
};

enum
{
    SWMI_FLIP_PARITY = 0, /* flip the cluster iff its canonical root index is odd (reference rule) */
    SWMI_FLIP_RANDOM = 1  /* flip the cluster with root r iff the top bit of fmix32(r ^ key) is set, key =
                           * MurmurHash3_x86_32(lcg32 iteration before the sweep, 4 bytes LE; seed) --
                           * one random bit per cluster, site 0 not frozen (no counterpart in the reference) */
};

typedef struct swmi_engine swmi_engine;

typedef struct
{
    uint32_t n0;          /* extent in x (fast index); even (swendsen_wang.cpp:45-48), at least 2 */
    uint32_t n1;          /* extent in y; even, at least 2; P*n1 < 2^31 for the row pitch P */
    float temperature;    /* T > 0 */
    uint32_t seed;        /* lcg32 group seeds: seed*0x9E3779B1 + g + 1 (seed 0: g + 1) */
    uint32_t spin_seed;   /* hot start: srand48(spin_seed); spin = drand48() < 0.5 ? 0 : 1 (1: reference) */
    uint32_t tile_rows;   /* tile height in 64-site wave rows (0: default); results do not depend on it */
    uint32_t flip_rule;   /* SWMI_FLIP_PARITY or SWMI_FLIP_RANDOM */
    uint32_t max_samples; /* capacity of the on-device (E, M) time series (0: no series, moments only) */
} swmi_lattice_config;

typedef struct
{
    char name[64];
    char arch[32];
    uint32_t compute_units;
    uint32_t wavefront_size;
    uint64_t total_memory;
    uint32_t l2_bytes;
    uint32_t clock_khz;
} swmi_device_info;

/* ---- errors and device ---- */
const char* swmi_last_error(void);
/* Replaces AMD_GPU / HipGetDeviceProperty (utility/device/device_amd_gpu.hpp:15-87,
 * utility/hip/hip.cpp:26-101): device 0 properties. */
int swmi_get_device_info(swmi_device_info* out);
/* Library version string. */
const char* swmi_version(void);

/* ---- engine lifecycle: one engine owns one or more lattices (a batch); calls on one engine are
 * serialised on its own HIP stream. Replaces the Lattice<2> constructor (ising/src/lattice.cpp:15-27),
 * Lattice::InitializeGpuSpins (lattice.cpp:81-97) and SwendsenWang_2D::Initialize
 * (ising/src/swendsen_wang.cpp:41-91, swendsen_wang_amd_gpu.cpp:22-53). ---- */
int swmi_create(const swmi_lattice_config* configs, uint32_t n_lattices, swmi_engine** out);
int swmi_destroy(swmi_engine* engine);
int swmi_num_lattices(const swmi_engine* engine, uint32_t* out);
int swmi_synchronize(swmi_engine* engine);

/* ---- state transfer (host buffers of n0*n1 bytes) ---- */
int swmi_upload_spins(swmi_engine* engine, uint32_t lattice, const uint8_t* spins);
int swmi_download_spins(swmi_engine* engine, uint32_t lattice, uint8_t* spins);

/* ---- temperature: p_add and its integer threshold are derived on the host
 * (ising/src/swendsen_wang.cpp:100). ---- */
int swmi_set_temperature(swmi_engine* engine, uint32_t lattice, float temperature);
int swmi_get_threshold(const swmi_engine* engine, uint32_t lattice, float* p_add, uint64_t* threshold);

/* ---- sweeps: every lattice of the engine advances n_sweeps Swendsen-Wang sweeps. Replaces
 * LatticeMonteCarloAlgorithm::Update (ising/include/lattice_mc.hpp:18) as implemented by
 * SwendsenWang_2D::Update (ising/src/swendsen_wang.cpp:94-112). Asynchronous on the engine stream. ---- */
int swmi_sweep(swmi_engine* engine, uint32_t n_sweeps);
/* Launch mode: captured hipGraphs of 16 sweeps (or of sample groups) replayed for long sweep loops
 * (enable = 1, the default) or every kernel launched one by one (0). Results are identical. The
 * reference launches every kernel and synchronises after each (swendsen_wang_amd_gpu.cpp:253-492).
 * New engines start with graphs unless the environment variable SWMI_GRAPHS is "0". */
int swmi_set_graphs(swmi_engine* engine, int enable);
/* Tile height (in 64-site wave rows) the lattice runs with. */
int swmi_get_tile_rows(const swmi_engine* engine, uint32_t lattice, uint32_t* tile_rows);
/* Row pitch P of the lattice's padded layout and its number of lcg32 groups, P*n1/64 (the sizes of the
 * swmi_rng_export arrays). The reference stores the lattice unpadded (ising/include/lattice.hpp:24); P = n0
 * for every extent that needs no padding. */
int swmi_get_layout(const swmi_engine* engine, uint32_t lattice, uint32_t* pitch, uint32_t* n_groups);
/* Tile-height autotuner: times plain sweeps of this configuration for each tile height in
 * {1, 2, 4, 8, 16, 32, 64} (clamped to the lattice; ms_per_sweep[i] = -1 for a height clamped to one already
 * measured), returns the fastest and remembers it for this extent, so that later lattices created with
 * tile_rows = 0 use it. Spins, labels and random streams do not depend on the tile height. The
 * reference fixes the tile from the wavefront size (ising/src/swendsen_wang.cpp:16-31). */
int swmi_autotune(const swmi_lattice_config* config, uint32_t sweeps, uint32_t* best_tile_rows, float* ms_per_sweep);
/* Batch autotuner: times plain sweeps of an engine holding all n_lattices lattices with one tile height
 * for all of them (each of {1, 2, 4, 8, 16, 32, 64}, clamped per lattice) and returns the height with the
 * highest aggregate throughput. A batch prefers taller tiles than its small lattices would alone: its
 * blocks already fill the device, and taller tiles mean fewer block edges to merge. */
int swmi_autotune_batch(const swmi_lattice_config* configs, uint32_t n_lattices, uint32_t sweeps, uint32_t* best_tile_rows,
                        float* ms_per_sweep);
/* Per-lattice schedule, relative to the lattice's current sweep count: it advances `sweeps` more sweeps
 * (0: no limit) and then stays unchanged while the other lattices of the engine go on; swmi_run_scheduled
 * samples it after every n_sep-th sweep past the first `warmup`. The measurement protocol is that of
 * ising/src/ising_2d.cpp:122-148 (warmup, then a sample every N_Sep sweeps), per lattice. */
int swmi_set_schedule(swmi_engine* engine, uint32_t lattice, uint32_t sweeps, uint32_t warmup, uint32_t n_sep);
/* n_sweeps sweeps of the engine with the scheduled samples (every lattice on its own schedule). */
int swmi_run_scheduled(swmi_engine* engine, uint32_t n_sweeps, uint32_t n_sep);
/* Sweeps done by the lattice since it was created, and samples taken. */
int swmi_get_progress(swmi_engine* engine, uint32_t lattice, uint32_t* sweeps_done, uint32_t* n_samples);

/* ---- observables. swmi_sample appends one sample (E, M of the current configuration) per lattice to
 * its on-device series and FP64 moments; swmi_run runs n_sweeps sweeps with a sample after every
 * n_sep-th sweep (the measurement loop of ising/src/ising_2d.cpp:126-148, all on the device). ---- */
int swmi_sample(swmi_engine* engine);
int swmi_run(swmi_engine* engine, uint32_t n_sweeps, uint32_t n_sep);
/* Sums of the current configuration. Replaces Lattice<2>::GetEnergyAndMagnetization
 * (ising/include/lattice.hpp:66-69, ising/src/lattice.cpp:29-79) without the full-lattice copy. */
int swmi_observables(swmi_engine* engine, uint32_t lattice, int64_t* energy, int64_t* magnetization);
/* moments[0..6] = {samples, sum e, sum e^2, sum |m|, sum m^2, sum m^4} with e = -E/N, m = M/N (FP64). */
int swmi_get_moments(swmi_engine* engine, uint32_t lattice, double* moments);
/* Copies up to max samples of the (E, M) series; *count receives the number stored. */
int swmi_get_series(swmi_engine* engine, uint32_t lattice, int64_t* energy, int64_t* magnetization,
                    uint32_t max, uint32_t* count);
int swmi_reset_measurements(swmi_engine* engine);

/* ---- inspection hooks (parity tests) ---- */
/* Bond bytes per site (bit 0: bond to (x+1,y), bit 1: bond to (x,y+1)) drawn by the last sweep;
 * recorded only while enabled. */
int swmi_record_bonds(swmi_engine* engine, int enable);
int swmi_download_bonds(swmi_engine* engine, uint32_t lattice, uint8_t* bonds);
/* Label the given bond field (same byte format) without touching spins or lcg32 state, and leave the
 * canonical labels for swmi_download_labels. */
int swmi_label_bonds(swmi_engine* engine, uint32_t lattice, const uint8_t* bonds);
/* Canonical labels of the last labelling (after a sweep: the clusters that sweep flipped). */
int swmi_download_labels(swmi_engine* engine, uint32_t lattice, uint32_t* labels);
/* Allocation-free labelling. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Labels the bond field `bonds` (device memory, n0*n1 bytes, the format of swmi_label_bonds) of an
 * n0 x n1 lattice (the extents swmi_lattice_config allows, at most 2^22 sites) and writes the canonical
 * labels into `labels` (device memory, n0*n1 uint32), exactly as swmi_label_bonds + swmi_download_labels.
 * The kernels are enqueued on `stream` (a hipStream_t; NULL is the default stream) and the call returns
 * without waiting; it allocates nothing, issues no memset and copies nothing to or from the host. Every
 * buffer beyond the two arrays comes from a 48 MiB static device array that each GPU holds until the
 * process ends, shared by all calls on that GPU: calls that overlap on one GPU (other streams, other
 * threads) race, calls on one stream are ordered. An unpadded lattice needs about 2.5 bytes of it per
 * site; one whose rows are padded (swmi_get_layout) about 7 per padded site, and one that needs more than
 * the array holds is rejected with the amount it needs. Errors in the arguments are reported before
 * anything is enqueued. */
int swmi_label_bonds_noalloc(const uint8_t* bonds, uint32_t n0, uint32_t n1, uint32_t* labels, void* stream);
/* Per-site lcg32 states (n_groups*64, padded layout), selector bit planes (3*n_groups), rotations (n_groups)
 * and the iteration count of the lattice (n_groups: swmi_get_layout). */
int swmi_rng_export(swmi_engine* engine, uint32_t lattice, uint32_t* states, uint64_t* masks, uint32_t* rot,
                    uint32_t* iteration);

/* ---- checkpoint / resume of a lattice's Markov chain: spins (with the lcg32 selector bits), per-site
 * lcg32 states, selector planes, rotations, counters, FP64 moments and the (E, M) series. Labels are
 * per-sweep scratch and are not saved. A restored lattice continues bit for bit as the saved one would
 * have. The reference has no checkpoint (its state lives in a function-local static,
 * ising/include/lattice.hpp:56). Format: a 64-byte header ("SWMICKP1", extents, seed, flip rule,
 * temperature, word counts), then the arrays in that order, little endian. ---- */
int swmi_checkpoint_size(swmi_engine* engine, uint32_t lattice, uint64_t* bytes);
int swmi_checkpoint(swmi_engine* engine, uint32_t lattice, void* buffer, uint64_t bytes);
/* The target lattice must have the same extents and series capacity; its temperature, seed and flip rule
 * become the checkpoint's. */
int swmi_restore(swmi_engine* engine, uint32_t lattice, const void* buffer, uint64_t bytes);

/* ---- error bars from the on-device (E, M) series (max_samples > 0): out[0..7] = {<e>, err <e> (binning
 * into n_bins bins), <|m|>, err <|m|>, C/N = N (<e^2> - <e>^2) / T^2, jackknife err C/N over the bins,
 * tau_int of e (samples, from the binned variance: 0.5 * var_bin_mean * n_per_bin / var_sample), samples
 * used}, e = -E/N, m = M/N. The reference prints the means only (ising/src/ising_2d.cpp:152-162). ---- */
int swmi_error_bars(swmi_engine* engine, uint32_t lattice, uint32_t n_bins, double* out);

/* Invariant violations seen by the kernels (bit 0: a label chase left [0, n)); 0 in a correct run. */
int swmi_get_error_flags(swmi_engine* engine, uint32_t lattice, uint32_t* flags);

/* ---- measurement support ---- */
/* Average device time per kernel over n_sweeps sweeps, measured with HIP events on the engine stream:
 * ms[0] bond+label (which also applies the previous sweep's flip), ms[1] merge, ms[2] resolve (the flip
 * decision of every cluster). */
int swmi_time_kernels(swmi_engine* engine, uint32_t n_sweeps, float* ms);
/* Device-to-device copy of `bytes` bytes (calibration kernel for HBM byte counters). */
int swmi_copy_calibration(size_t bytes, uint32_t repeats, float* ms);

/* ---- serving: continuous batching of independent jobs. No counterpart in the reference, which runs one
 * lattice per process (ising/include/lattice.hpp:56); each job follows the reference's measurement
 * protocol (ising/src/ising_2d.cpp:122-148): `warmup` sweeps, then `sweeps` sweeps with a sample after
 * every n_sep-th, FP64 moments as swmi_get_moments. As in the reference's loop (`for (i = 0; i < n_sweeps;
 * i += N_Sep)`, ising_2d.cpp:132), the measured sweeps come in whole groups of n_sep: a job runs
 * ceil(sweeps / n_sep) groups and takes one sample after each. A worker thread owns the GPU: jobs are
 * admitted as device memory (memory_budget) and slots (max_jobs) free up, run in size-class lanes (one HIP
 * stream each) in chunks of chunk_sweeps sweeps together with the other resident jobs, and retire after
 * exactly warmup + ceil(sweeps / n_sep) * n_sep sweeps. A job's result is the same as when it runs
 * alone. ---- */
typedef struct swmi_server swmi_server;

typedef struct
{
    uint64_t memory_budget;  /* device bytes for resident jobs and pooled buffers (0: 200 GiB) */
    uint32_t max_jobs;       /* most resident jobs (0: 4096) */
    uint32_t chunk_sweeps;   /* sweeps per scheduling step of a lane, a multiple of n_sep (0: 50) */
    uint32_t n_sep;          /* sweeps between samples of every job (0: 10) */
    uint32_t n_lanes;        /* size-class lanes, 1..4 (0: 4: <= 128^2, <= 512^2, <= 2048^2, larger) */
} swmi_server_config;

typedef struct
{
    uint32_t n0, n1;         /* extents, as swmi_lattice_config */
    float temperature;
    uint32_t seed;           /* lcg32 seed */
    uint32_t spin_seed;      /* hot start seed */
    uint32_t warmup;         /* unsampled sweeps first, a multiple of n_sep */
    uint32_t sweeps;         /* measured sweeps after the warmup (> 0) */
    uint32_t flip_rule;      /* SWMI_FLIP_PARITY or SWMI_FLIP_RANDOM */
    uint64_t tag;            /* caller's value, returned with the result */
} swmi_job;

typedef struct
{
    uint64_t job_id, tag;
    uint32_t n0, n1;
    float temperature;
    uint32_t sweeps_done;    /* warmup + ceil(sweeps / n_sep) * n_sep */
    uint32_t error_flags;    /* as swmi_get_error_flags */
    uint32_t status;         /* SWMI_OK, or the code that ended the job */
    double moments[6];       /* as swmi_get_moments, over the samples after the warmup */
    int64_t energy;          /* E and M of the final configuration */
    int64_t magnetization;
    double t_submit, t_admit, t_finish;   /* seconds on the server clock */
} swmi_job_result;

typedef struct
{
    uint64_t submitted, admitted, finished;
    uint64_t queued, active;
    uint64_t bytes_in_use, bytes_pooled;
    uint64_t spin_updates;   /* sum over finished jobs of (warmup + sweeps) * n0 * n1 */
    double clock;            /* server clock now, seconds since creation */
} swmi_server_stats;

int swmi_server_create(const swmi_server_config* config, swmi_server** out);
/* Waits for the chunks in flight, then drops queued and resident jobs (without results) and frees the
 * device memory. */
int swmi_server_destroy(swmi_server* server);
/* Thread-safe; returns immediately. */
int swmi_server_submit(swmi_server* server, const swmi_job* job, uint64_t* job_id);
/* Takes up to max finished results (in completion order). */
int swmi_server_poll(swmi_server* server, swmi_job_result* out, uint32_t max, uint32_t* count);
/* Blocks until every submitted job has finished or timeout_ms has passed. */
int swmi_server_wait(swmi_server* server, uint32_t timeout_ms, uint32_t* unfinished);
int swmi_server_get_stats(swmi_server* server, swmi_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif
