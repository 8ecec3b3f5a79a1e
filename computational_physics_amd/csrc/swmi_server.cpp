// swmi_server.cpp -- continuous-batching request loop of the engine (swmi_server_* in include/swmi.h).
//
// Jobs (L, T, seed, warmup, sweeps) are submitted from any thread. A worker thread owns the GPU side:
// jobs are sorted by lattice size into lanes; every lane is an engine with its own HIP stream and work
// list and runs chunks of `chunk_sweeps` sweeps over all of its resident jobs, so small lattices are not
// paced by large ones and the lanes' kernels overlap on the device. Between two chunks of a lane the
// worker retires the jobs that finished (each job stops at exactly warmup + ceil(sweeps / n_sep) * n_sep
// sweeps through its on-device schedule, whatever the chunk boundaries), admits queued jobs while the device-memory budget
// allows, and rebuilds the lane's work list. Device buffers of retired jobs are pooled by shape and
// reused. A job's result does not depend on the other jobs: spins, random streams and moments are per
// lattice and the moments are accumulated in the same order as a run alone.
//
// The reference runs one lattice per process (ising/include/lattice.hpp:56); this loop has no
// counterpart there. Its per-job protocol is the reference's measurement loop (ising/src/ising_2d.cpp:
// 122-148): warmup sweeps, then a sample after every n_sep-th sweep, FP64 moments of e = -E/N and m = M/N.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "swmi_host.h"

using namespace swmi;

namespace
{
    // measured sweeps of a job: whole groups of n_sep sweeps covering `sweeps`
    uint32_t MeasuredSweeps(const uint32_t sweeps, const uint32_t n_sep) { return ((sweeps + n_sep - 1) / n_sep) * n_sep; }

    struct PendingJob
    {
        uint64_t id = 0;
        swmi_job job{};
        double t_submit = 0;
    };

    struct ActiveJob
    {
        PendingJob p;
        Lattice* lat = nullptr;
        uint64_t bytes = 0;
        uint32_t total = 0;      // warmup + sweeps
        uint32_t run = 0;        // sweeps the lane has run since admission (the job stops at `total`)
        double t_admit = 0;
    };

    struct Lane
    {
        std::unique_ptr<swmi_engine> eng;
        std::deque<PendingJob> queue;      // guarded by the server mutex
        std::vector<ActiveJob> active;     // worker thread only
        hipEvent_t done = nullptr;
        bool busy = false;
        bool dirty = false;                // membership changed since the last rebuild
        uint32_t chunk = 0;                // sweeps of the chunk in flight
    };

    constexpr uint64_t kDefaultBudget = 200ull << 30;
}

struct swmi_server
{
    swmi_server_config cfg{};
    std::vector<Lane> lanes;
    std::vector<std::unique_ptr<Lattice>> pool;   // retired lattices, reusable by shape (worker only)
    uint64_t pooled_bytes = 0, active_bytes = 0;  // worker only (mirrored into stats under the mutex)

    std::mutex mu;
    std::condition_variable wake;       // worker: new job or stop
    std::condition_variable finished;   // waiters: a job finished
    std::deque<swmi_job_result> results;
    swmi_server_stats stats{};
    uint64_t next_id = 1;
    bool stop = false;
    int fatal = SWMI_OK;
    std::string fatal_msg;

    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::thread worker;

    double Now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    uint32_t LaneOf(const swmi_job& j) const
    {
        // size classes by sites: <= 128^2, <= 512^2, <= 2048^2, larger (4 lanes); fewer lanes merge the top
        const uint64_t n = static_cast<uint64_t>(j.n0) * j.n1;
        uint32_t k = n <= (1ull << 14) ? 0u : (n <= (1ull << 18) ? 1u : (n <= (1ull << 22) ? 2u : 3u));
        return std::min<uint32_t>(k, static_cast<uint32_t>(lanes.size()) - 1u);
    }

    static swmi_lattice_config ConfigOf(const swmi_job& j)
    {
        swmi_lattice_config c{};
        c.n0 = j.n0;
        c.n1 = j.n1;
        c.temperature = j.temperature;
        c.seed = j.seed;
        c.spin_seed = j.spin_seed;
        c.tile_rows = 0;
        c.flip_rule = j.flip_rule;
        c.max_samples = 0;
        return c;
    }

    void Loop();
    bool Service(Lane& lane);
    void Retire(Lane& lane, ActiveJob& a);
    bool Admit(Lane& lane);
    std::unique_ptr<Lattice> TakeLattice(const swmi_lattice_config& c, uint64_t need);
    void Fail(int code, const std::string& msg);
};

std::unique_ptr<Lattice> swmi_server::TakeLattice(const swmi_lattice_config& c, const uint64_t need)
{
    for (auto it = pool.begin(); it != pool.end(); ++it)
        if ((*it)->SameShape(c))
        {
            std::unique_ptr<Lattice> l = std::move(*it);
            pool.erase(it);
            pooled_bytes -= l->Bytes();
            return l;
        }
    // make room: release pooled buffers of other shapes (largest first)
    std::sort(pool.begin(), pool.end(), [](const auto& a, const auto& b) { return a->Bytes() < b->Bytes(); });
    while (active_bytes + pooled_bytes + need > cfg.memory_budget && !pool.empty())
    {
        pooled_bytes -= pool.back()->Bytes();
        pool.pop_back();
    }
    if (active_bytes + pooled_bytes + need > cfg.memory_budget)
        return nullptr;
    auto l = std::make_unique<Lattice>();
    l->Allocate(c);
    return l;
}

bool swmi_server::Admit(Lane& lane)
{
    bool any = false;
    for (;;)
    {
        PendingJob p;
        {
            std::lock_guard<std::mutex> lock(mu);
            if (lane.queue.empty() || stop)
                break;
            uint32_t resident = 0;
            for (const auto& l : lanes)
                resident += static_cast<uint32_t>(l.active.size());
            if (resident >= cfg.max_jobs)
                break;
            p = lane.queue.front();
        }
        const swmi_lattice_config c = ConfigOf(p.job);
        const uint64_t need = Lattice::ResidentBytes(c);
        std::unique_ptr<Lattice> l = TakeLattice(c, need);
        if (!l)
            break;   // wait for memory (FIFO per lane)
        {
            std::lock_guard<std::mutex> lock(mu);
            lane.queue.pop_front();
            stats.queued -= 1;
            stats.active += 1;
            stats.admitted += 1;
        }
        l->Initialize(c, lane.eng->stream);
        // the measured sweeps run in whole groups of n_sep, as the reference's loop `for (i = 0; i < n_sweeps;
        // i += N_Sep)` does (ising/src/ising_2d.cpp:132): ceil(sweeps / n_sep) groups, a sample after each
        const uint32_t measured = MeasuredSweeps(p.job.sweeps, cfg.n_sep);
        LaunchSetSchedule(l->state.ptr, p.job.warmup + measured, p.job.warmup, cfg.n_sep, lane.eng->stream);
        Check(hipGetLastError(), "job initialisation");
        ActiveJob a;
        a.p = p;
        a.lat = l.get();
        a.bytes = l->Bytes();
        a.total = p.job.warmup + measured;
        a.t_admit = Now();
        active_bytes += a.bytes;
        lane.eng->lattices.push_back(std::move(l));
        lane.active.push_back(a);
        lane.dirty = true;
        any = true;
    }
    return any;
}

void swmi_server::Retire(Lane& lane, ActiveJob& a)
{
    swmi_engine& e = *lane.eng;
    swmi_job_result r{};
    r.job_id = a.p.id;
    r.tag = a.p.job.tag;
    r.n0 = a.p.job.n0;
    r.n1 = a.p.job.n1;
    r.temperature = a.p.job.temperature;
    LatState st{};
    Check(hipMemcpyAsync(&st, a.lat->state.ptr, sizeof(st), hipMemcpyDeviceToHost, e.stream), "job result");
    // final configuration's (E, M), reduced over this lattice only
    e.MaterializeOne(*a.lat);
    LaunchObservables(e.d_one_desc.ptr, nullptr, a.lat->g.n_blocks, kSampleAlways, e.stream);
    Check(hipGetLastError(), "job observables");
    std::vector<int32_t> part(2ull * kWavesPerBlock * a.lat->g.n_blocks);
    Check(hipMemcpyAsync(part.data(), a.lat->partials.ptr, part.size() * 4, hipMemcpyDeviceToHost, e.stream), "job result");
    Check(hipStreamSynchronize(e.stream), "job result");
    for (size_t i = 0; i < part.size() / 2; ++i)
    {
        r.energy += part[2 * i];
        r.magnetization += part[2 * i + 1];
    }
    r.sweeps_done = st.iteration / 2u;
    r.error_flags = st.error_flags;
    r.status = (r.sweeps_done == a.total && st.error_flags == 0) ? SWMI_OK : SWMI_ERR_STATE;
    std::memcpy(r.moments, st.moments, sizeof(r.moments));
    r.t_submit = a.p.t_submit;
    r.t_admit = a.t_admit;
    r.t_finish = Now();
    // buffers back to the pool
    for (auto& u : e.lattices)
        if (u.get() == a.lat)
        {
            pooled_bytes += u->Bytes();
            pool.push_back(std::move(u));
            break;
        }
    active_bytes -= a.bytes;
    lane.dirty = true;
    {
        std::lock_guard<std::mutex> lock(mu);
        results.push_back(r);
        stats.active -= 1;
        stats.finished += 1;
        stats.spin_updates += static_cast<uint64_t>(a.total) * r.n0 * r.n1;
    }
    finished.notify_all();
}

bool swmi_server::Service(Lane& lane)
{
    bool progressed = false;
    if (lane.busy)
    {
        const hipError_t q = hipEventQuery(lane.done);
        if (q == hipErrorNotReady)
            return false;
        Check(q, "lane chunk");
        lane.busy = false;
        progressed = true;
        for (auto& a : lane.active)
            a.run += lane.chunk;
        std::vector<ActiveJob> keep;
        for (auto& a : lane.active)
        {
            if (a.run >= a.total)
                Retire(lane, a);
            else
                keep.push_back(a);
        }
        lane.active.swap(keep);
    }
    {
        // shutting down: no new chunk; resident jobs end unfinished (their results are not reported)
        std::lock_guard<std::mutex> lock(mu);
        if (stop)
            return progressed;
    }
    progressed = Admit(lane) || progressed;
    if (lane.dirty)
    {
        lane.eng->Rebuild();
        lane.dirty = false;
        std::lock_guard<std::mutex> lock(mu);
        stats.bytes_in_use = active_bytes;
        stats.bytes_pooled = pooled_bytes;
    }
    if (!lane.active.empty())
    {
        // a chunk no longer than the longest remaining job, rounded up to the sample grid
        uint32_t left = 0;
        for (const auto& a : lane.active)
            left = std::max(left, a.total - a.run);
        uint32_t chunk = std::min(cfg.chunk_sweeps, ((left + cfg.n_sep - 1) / cfg.n_sep) * cfg.n_sep);
        lane.chunk = std::max(chunk, cfg.n_sep);
        lane.eng->Run(lane.chunk, cfg.n_sep, kSampleDue);
        Check(hipEventRecord(lane.done, lane.eng->stream), "lane event");
        lane.busy = true;
        progressed = true;
    }
    return progressed;
}

void swmi_server::Fail(const int code, const std::string& msg)
{
    std::lock_guard<std::mutex> lock(mu);
    fatal = code;
    fatal_msg = msg;
    // every unfinished job ends with the error
    for (auto& lane : lanes)
    {
        for (auto& a : lane.active)
        {
            swmi_job_result r{};
            r.job_id = a.p.id;
            r.tag = a.p.job.tag;
            r.n0 = a.p.job.n0;
            r.n1 = a.p.job.n1;
            r.status = static_cast<uint32_t>(code);
            r.t_submit = a.p.t_submit;
            r.t_finish = Now();
            results.push_back(r);
            stats.finished += 1;
        }
        lane.active.clear();
        for (auto& p : lane.queue)
        {
            swmi_job_result r{};
            r.job_id = p.id;
            r.tag = p.job.tag;
            r.n0 = p.job.n0;
            r.n1 = p.job.n1;
            r.status = static_cast<uint32_t>(code);
            r.t_submit = p.t_submit;
            r.t_finish = Now();
            results.push_back(r);
            stats.finished += 1;
        }
        lane.queue.clear();
    }
    stats.active = 0;
    stats.queued = 0;
    finished.notify_all();
}

void swmi_server::Loop()
{
    try
    {
        Check(hipSetDevice(0), "hipSetDevice");
        for (;;)
        {
            bool progressed = false;
            for (auto& lane : lanes)
                progressed = Service(lane) || progressed;
            bool any_busy = false;
            for (const auto& lane : lanes)
                any_busy = any_busy || lane.busy;
            std::unique_lock<std::mutex> lock(mu);
            if (stop && !any_busy)
                break;
            if (!progressed)
            {
                // idle lanes sleep until a submission; busy lanes are polled every 20 us
                if (any_busy)
                    wake.wait_for(lock, std::chrono::microseconds(20));
                else
                    wake.wait_for(lock, std::chrono::milliseconds(5), [&] {
                        if (stop)
                            return true;
                        for (const auto& l : lanes)
                            if (!l.queue.empty())
                                return true;
                        return false;
                    });
            }
        }
    }
    catch (const std::exception& ex)
    {
        for (auto& lane : lanes)
            if (lane.eng && lane.eng->stream)
                (void)hipStreamSynchronize(lane.eng->stream);
        const auto* se = dynamic_cast<const SwmiError*>(&ex);
        Fail(se ? se->code : SWMI_ERR_HIP, ex.what());
    }
}

extern "C" {

int swmi_server_create(const swmi_server_config* config, swmi_server** out)
{
    return Guard([&] {
        if (!out)
            throw SwmiError(SWMI_ERR_INVALID, "null output");
        *out = nullptr;
        swmi_server_config c{};
        if (config)
            c = *config;
        if (c.memory_budget == 0)
            c.memory_budget = kDefaultBudget;
        if (c.max_jobs == 0)
            c.max_jobs = 4096;
        if (c.n_sep == 0)
            c.n_sep = 10;
        if (c.chunk_sweeps == 0)
            c.chunk_sweeps = 50;
        if (c.n_lanes == 0)
            c.n_lanes = 4;
        if (c.n_lanes > 4 || c.chunk_sweeps % c.n_sep != 0)
            throw SwmiError(SWMI_ERR_INVALID, "server needs 1..4 lanes and chunk_sweeps a multiple of n_sep");
        RequireGfx950(0);
        auto s = std::make_unique<swmi_server>();
        s->cfg = c;
        s->lanes.resize(c.n_lanes);
        for (auto& lane : s->lanes)
        {
            lane.eng = std::make_unique<swmi_engine>();
            lane.eng->use_graphs = false;   // the lanes' work lists change between chunks
            Check(hipEventCreateWithFlags(&lane.done, hipEventDisableTiming), "lane event");
        }
        swmi_server* raw = s.get();
        s->worker = std::thread([raw] { raw->Loop(); });
        *out = s.release();
    });
}

int swmi_server_destroy(swmi_server* server)
{
    return Guard([&] {
        if (!server)
            return;
        {
            std::lock_guard<std::mutex> lock(server->mu);
            server->stop = true;
        }
        server->wake.notify_all();
        if (server->worker.joinable())
            server->worker.join();
        for (auto& lane : server->lanes)
        {
            if (lane.eng && lane.eng->stream)
                (void)hipStreamSynchronize(lane.eng->stream);
            if (lane.done)
                (void)hipEventDestroy(lane.done);
        }
        delete server;
    });
}

int swmi_server_submit(swmi_server* server, const swmi_job* job, uint64_t* job_id)
{
    return Guard([&] {
        if (!server || !job)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        const swmi_lattice_config c = swmi_server::ConfigOf(*job);
        ValidateConfig(c);
        if (job->sweeps == 0 || job->sweeps > (1u << 30) || job->warmup % server->cfg.n_sep != 0 ||
            static_cast<uint64_t>(job->warmup) + MeasuredSweeps(job->sweeps, server->cfg.n_sep) > (1u << 30))
            throw SwmiError(SWMI_ERR_INVALID, "job needs sweeps > 0, warmup a multiple of n_sep and warmup + sweeps <= 2^30");
        if (Lattice::ResidentBytes(c) > server->cfg.memory_budget)
            throw SwmiError(SWMI_ERR_MEMORY, "job larger than the server's memory budget");
        std::lock_guard<std::mutex> lock(server->mu);
        if (server->fatal != SWMI_OK)
            throw SwmiError(server->fatal, "server stopped after an error: " + server->fatal_msg);
        if (server->stop)
            throw SwmiError(SWMI_ERR_STATE, "server is shutting down");
        PendingJob p;
        p.id = server->next_id++;
        p.job = *job;
        p.t_submit = server->Now();
        server->lanes[server->LaneOf(*job)].queue.push_back(p);
        server->stats.submitted += 1;
        server->stats.queued += 1;
        if (job_id)
            *job_id = p.id;
        server->wake.notify_all();
    });
}

int swmi_server_poll(swmi_server* server, swmi_job_result* out, uint32_t max, uint32_t* count)
{
    return Guard([&] {
        if (!server || (!out && max))
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lock(server->mu);
        uint32_t n = 0;
        while (n < max && !server->results.empty())
        {
            out[n++] = server->results.front();
            server->results.pop_front();
        }
        if (count)
            *count = n;
    });
}

int swmi_server_wait(swmi_server* server, uint32_t timeout_ms, uint32_t* unfinished)
{
    return Guard([&] {
        if (!server)
            throw SwmiError(SWMI_ERR_INVALID, "null server");
        std::unique_lock<std::mutex> lock(server->mu);
        server->finished.wait_for(lock, std::chrono::milliseconds(timeout_ms),
                                  [&] { return server->stats.finished >= server->stats.submitted; });
        if (unfinished)
            *unfinished = static_cast<uint32_t>(server->stats.submitted - server->stats.finished);
        if (server->fatal != SWMI_OK)
            throw SwmiError(server->fatal, "server stopped after an error: " + server->fatal_msg);
    });
}

int swmi_server_get_stats(swmi_server* server, swmi_server_stats* out)
{
    return Guard([&] {
        if (!server || !out)
            throw SwmiError(SWMI_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lock(server->mu);
        *out = server->stats;
        out->clock = server->Now();
    });
}

}  // extern "C"
