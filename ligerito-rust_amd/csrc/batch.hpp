// Batched proves and the request server (see batch.cpp).
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace lig {

// One proving lane: its own stream, pool, pinned staging and configurations.
template <class H> struct Worker {
    Ctx ctx;
    std::map<int, std::unique_ptr<Config<H>>> cfgs;
    explicit Worker(int dev);
    ~Worker();
    Config<H>& config(int log_n);
    // poly: 2^log_n elements in host memory, or device memory if on_device
    void prove_one(int log_n, const void* poly, bool on_device, uint32_t seed, Proof<H>& out);
    // count proves of 2^log_n elements each in lockstep (prove_gang;
    // GF(2^32), one recursive step)
    void prove_gang_jobs(int log_n, const void* const* polys, bool on_device, const uint32_t* seeds, int count, Proof<H>* out);
    // device bytes the worker holds between proves: its pool (cached
    // blocks) and the tables of its prover configurations
    size_t retained() const;
    // returns the cached blocks to the device and drops every configuration
    // but the one for keep_log_n (-1: all); only while the worker is idle
    void trim(int keep_log_n);
};

// Device bytes one prove of 2^log_n elements allocates at most, from an
// empty pool (upper bound used for admission).
size_t prove_footprint(int field_bits, int log_n);

template <class H> struct Batch {
    int device;
    std::vector<std::unique_ptr<Worker<H>>> workers;
    Batch(int dev, int n_streams);
    void run(size_t n_jobs, const int* log_n, const void* const* polys, bool on_device, const uint32_t* seeds,
             std::vector<Proof<H>>& out);
};

template <class H> struct Server {
    struct Req {
        int log_n = 0;
        const void* poly = nullptr;
        uint32_t seed = 0;
        size_t footprint = 0;
        bool done = false;
        int status = 0;
        std::string message;
        Proof<H> proof;
        std::chrono::steady_clock::time_point t_submit, t_start, t_done;
    };
    int device;
    uint64_t hbm_budget;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Req> reqs;
    uint64_t last_ticket = 0;
    bool stopping = false;
    // set while warm() runs: workers admit no request
    bool warming = false;
    // the warm-up's sizes, input, per-worker tasks and outcome (under mu)
    std::vector<int> warm_sizes;
    const void* warm_input = nullptr;
    std::vector<char> warm_todo;
    size_t warm_left = 0, warm_done = 0;
    bool warm_full = false;
    int warm_code = 0;
    std::string warm_msg;
    std::vector<std::unique_ptr<Worker<H>>> workers;
    // per worker: running a request, and the device bytes charged for it
    // (what the worker held when the request started plus the request's
    // footprint bound)
    std::vector<char> busy;
    std::vector<size_t> charge;
    std::vector<std::thread> threads;

    Server(int dev, int n_slots, uint64_t budget);
    ~Server();
    uint64_t submit(int log_n, const void* poly, uint32_t seed);
    // Each worker, on its own thread, proves a zero input of each size as a
    // request from pinned host memory would be proved (configuration, pool
    // blocks, staging, upload path), while no request runs; a worker stops
    // at the first size the budget cannot hold.  Returns the number of
    // (worker, size) pairs warmed.
    size_t warm(const int* log_ns, size_t count);
    // 1 with the proof when done, 0 while pending (wait == false), or the
    // request's negative status; a finished ticket is released.
    int take(uint64_t ticket, bool wait, Proof<H>& out, double* latency_ms, std::string& msg);
    // device bytes charged against the budget right now (every worker's
    // retained memory, or its charge while it runs a request)
    size_t charged() const;

private:
    // Decides (under mu) whether request r may start on worker wi, trimming
    // idle workers' caches when that makes it fit; sets charge[wi].
    bool admit(size_t wi, const Req& r);
    void loop(size_t wi);
    void warm_worker(size_t wi);
};

}  // namespace lig
