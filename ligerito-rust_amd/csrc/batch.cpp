// Many independent proves on one device: a batch call and a request server.
//
// Neither exists in the reference (its prover runs one proof at a time on
// the host).  Both keep the single-proof prover (engine.cpp) unchanged and
// run several of them at once, each on a worker with its own HIP stream,
// device-memory pool, pinned staging and cached prover configurations, so
// the kernels of different proofs overlap on the GPU while any one proof
// waits on its host-side transcript steps.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <thread>
#include <type_traits>
#include <unordered_map>

#include "batch.hpp"

namespace lig {

template <class H> Worker<H>::Worker(int dev) {
    ctx.device = dev;
    LIG_HIP(hipSetDevice(dev));
    LIG_HIP(hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking));
    ctx.spin_wait = false;  // many workers share the host cores
}

template <class H> Worker<H>::~Worker() {
    if (!ctx.stream) return;
    (void)hipSetDevice(ctx.device);
    (void)hipStreamSynchronize(ctx.stream);
    ctx.release_side();
    cfgs.clear();
    ctx.pool.drain();
    (void)hipStreamDestroy(ctx.stream);
    ctx.stream = nullptr;
}

template <class H> Config<H>& Worker<H>::config(int log_n) {
    auto it = cfgs.find(log_n);
    if (it != cfgs.end()) return *it->second;
    std::unique_ptr<Config<H>> c(new Config<H>());
    c->init(&ctx, log_n);
    Config<H>& ref = *c;
    cfgs.emplace(log_n, std::move(c));
    return ref;
}

template <class H> void Worker<H>::prove_one(int log_n, const void* poly, bool on_device, uint32_t seed, Proof<H>& out) {
    using T = typename H::T;
    LIG_HIP(hipSetDevice(ctx.device));
    Config<H>& cfg = config(log_n);
    const size_t n = (size_t)1 << log_n;
    PBuf<T> tmp;
    const T* d = (const T*)poly;
    if (!on_device) {
        tmp = PBuf<T>(&ctx.pool, n);
        ctx.upload(tmp.p, poly, n * sizeof(T));
        d = tmp.p;
    }
    prove<H>(cfg, d, seed, out);
}

template <class H> size_t Worker<H>::retained() const {
    size_t b = ctx.pool.held();
    for (const auto& kv : cfgs) b += kv.second->device_bytes();
    return b;
}

template <class H> void Worker<H>::trim(int keep_log_n) {
    (void)hipSetDevice(ctx.device);
    for (auto it = cfgs.begin(); it != cfgs.end();) {
        if (it->first == keep_log_n) ++it;
        else it = cfgs.erase(it);
    }
    ctx.pool.drain();
}

size_t prove_footprint(int field_bits, int log_n) {
    // input copy N, message N and codeword 4N elements of the first commit,
    // folds, the recursive commits and sumcheck buffers below N elements,
    // the first Merkle tree (2 * N/16 nodes of 32 B, 4N bytes) and the
    // recursive trees; 64 MiB cover the configuration tables and the small
    // buffers of the induction and the openings
    const size_t n = (size_t)1 << log_n, fb = (size_t)field_bits / 8;
    return 7 * n * fb + (n / 8) * 32 + ((size_t)64 << 20);
}

// ---------------------------------------------------------------- batch --
template <class H> Batch<H>::Batch(int dev, int n_streams) : device(dev) {
    LIG_REQUIRE(n_streams >= 1 && n_streams <= 64, LIG_E_INVALID, "n_streams must be in [1, 64]");
    for (int i = 0; i < n_streams; i++) workers.emplace_back(new Worker<H>(dev));
}

// Gangs: same-size GF(2^32) jobs are proved in lockstep by one worker
// (prove_gang), every kernel launched once for the gang; at most
// GANG_BATCH_MAX jobs and 2^GANG_MAX_ELEMS_LOG input elements per gang.
// Measured on MI355X, 36 jobs of one size on 12 workers: 2^16 jobs 47.5 ->
// 24.5 us each, 2^19 60.3 -> 40.4, 2^20 81.2 -> 61.8, 2^22 unchanged (190
// us: bound by the encode); config D 21.0 -> 16.4 ms.  Gangs of 8 or 16
// jobs, or of 2^23 / 2^26 elements, measured the same within noise: a
// gang's host steps (queries, node lists) run on its one worker thread, so
// larger gangs trade fewer launches for longer host phases.  Eight 2^24
// proves: 6.7 ms without gangs, 6.4 ms in gangs of two (means of five runs).
constexpr int GANG_BATCH_MAX = 4, GANG_MAX_ELEMS_LOG = 25;

static bool gangs_enabled() {
    static const bool on = [] {
        const char* e = getenv("LIG_BATCH_GANG");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <class H>
void Worker<H>::prove_gang_jobs(int log_n, const void* const* polys, bool on_device, const uint32_t* seeds, int count,
                                Proof<H>* out) {
    using T = typename H::T;
    if constexpr (std::is_same<H, H32>::value) {
        LIG_HIP(hipSetDevice(ctx.device));
        Config<H>& cfg = config(log_n);
        const size_t n = (size_t)1 << log_n;
        PBuf<T> tmp;
        std::vector<const T*> d(count);
        if (!on_device) {
            tmp = PBuf<T>(&ctx.pool, n * (size_t)count);
            for (int j = 0; j < count; j++) {
                ctx.upload(tmp.p + j * n, polys[j], n * sizeof(T));
                d[j] = tmp.p + j * n;
            }
        } else {
            for (int j = 0; j < count; j++) d[j] = (const T*)polys[j];
        }
        prove_gang<H>(cfg, d.data(), seeds, count, out);
    } else {
        throw Error(LIG_E_INVALID, "gang proves are GF(2^32)");
    }
}

template <class H>
void Batch<H>::run(size_t n_jobs, const int* log_n, const void* const* polys, bool on_device, const uint32_t* seeds,
                   std::vector<Proof<H>>& out) {
    for (size_t i = 0; i < n_jobs; i++) {
        LIG_REQUIRE(polys[i], LIG_E_INVALID, "NULL polynomial");
        LIG_REQUIRE(log_n[i] >= 16 && log_n[i] <= 27, LIG_E_INVALID, "log_n must be in [16, 27]");
    }
    out.assign(n_jobs, Proof<H>());
    // largest first, so the tail of the batch is made of short proofs
    std::vector<size_t> order(n_jobs);
    for (size_t i = 0; i < n_jobs; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return log_n[a] > log_n[b]; });

    // work items in that order: a gang of same-size jobs or a single job
    std::vector<std::vector<size_t>> items;
    for (size_t k = 0; k < n_jobs;) {
        size_t e = k;
        while (e < n_jobs && log_n[order[e]] == log_n[order[k]]) e++;
        const int L = log_n[order[k]];
        size_t cap = 1;
        if (std::is_same<H, H32>::value && gangs_enabled()) {
            cap = L >= GANG_MAX_ELEMS_LOG ? 1 : ((size_t)1 << GANG_MAX_ELEMS_LOG) >> L;
            cap = std::min<size_t>(cap, GANG_BATCH_MAX);
        }
        // gangs of about equal size
        const size_t n_same = e - k, n_items = (n_same + cap - 1) / cap;
        for (size_t g = 0, at = k; g < n_items; g++) {
            const size_t sz = n_same / n_items + (g < n_same % n_items ? 1 : 0);
            items.emplace_back(order.begin() + (long)at, order.begin() + (long)(at + sz));
            at += sz;
        }
        k = e;
    }

    std::atomic<size_t> next{0};
    std::atomic<bool> failed{false};
    std::mutex err_mu;
    int err_code = LIG_OK;
    std::string err_msg;
    auto fail = [&](int code, const std::string& msg) {
        std::lock_guard<std::mutex> g(err_mu);
        if (!failed.exchange(true)) {
            err_code = code;
            err_msg = msg;
        }
    };
    auto body = [&](Worker<H>* w) {
        try {
            for (;;) {
                if (failed.load()) return;
                const size_t it = next.fetch_add(1);
                if (it >= items.size()) return;
                const std::vector<size_t>& g = items[it];
                if (g.size() == 1) {
                    w->prove_one(log_n[g[0]], polys[g[0]], on_device, seeds[g[0]], out[g[0]]);
                    continue;
                }
                std::vector<const void*> ps(g.size());
                std::vector<uint32_t> ss(g.size());
                std::vector<Proof<H>> pr(g.size());
                for (size_t i = 0; i < g.size(); i++) {
                    ps[i] = polys[g[i]];
                    ss[i] = seeds[g[i]];
                }
                w->prove_gang_jobs(log_n[g[0]], ps.data(), on_device, ss.data(), (int)g.size(), pr.data());
                for (size_t i = 0; i < g.size(); i++) out[g[i]] = std::move(pr[i]);
            }
        } catch (const Error& e) {
            fail(e.code, e.what());
        } catch (const std::bad_alloc&) {
            fail(LIG_E_OOM, "host allocation failed");
        } catch (const std::exception& e) {
            fail(LIG_E_ASSERT, e.what());
        }
    };
    const size_t nw = std::min(workers.size(), std::max<size_t>(items.size(), 1));
    std::vector<std::thread> threads;
    for (size_t i = 1; i < nw; i++) threads.emplace_back(body, workers[i].get());
    body(workers[0].get());
    for (auto& t : threads) t.join();
    if (failed.load()) throw Error(err_code, err_msg);
}

// --------------------------------------------------------------- server --
template <class H> Server<H>::Server(int dev, int n_slots, uint64_t budget) : device(dev), hbm_budget(budget) {
    LIG_REQUIRE(n_slots >= 1 && n_slots <= 64, LIG_E_INVALID, "n_slots must be in [1, 64]");
    for (int i = 0; i < n_slots; i++) workers.emplace_back(new Worker<H>(dev));
    busy.assign(n_slots, 0);
    charge.assign(n_slots, 0);
    warm_todo.assign(n_slots, 0);
    for (int i = 0; i < n_slots; i++) threads.emplace_back([this, i] { loop((size_t)i); });
}

template <class H> Server<H>::~Server() {
    {
        std::lock_guard<std::mutex> g(mu);
        stopping = true;
        // queued requests that were never admitted fail
        for (uint64_t t : queue) {
            auto& r = reqs[t];
            r.status = LIG_E_INVALID;
            r.message = "server stopped before the request was admitted";
            r.done = true;
        }
        queue.clear();
    }
    cv.notify_all();
    done_cv.notify_all();  // waiters on the failed requests
    for (auto& t : threads) t.join();
}

template <class H> uint64_t Server<H>::submit(int log_n, const void* poly, uint32_t seed) {
    LIG_REQUIRE(poly, LIG_E_INVALID, "NULL polynomial");
    LIG_REQUIRE(log_n >= 16 && log_n <= 27, LIG_E_INVALID, "log_n must be in [16, 27]");
    std::lock_guard<std::mutex> g(mu);
    LIG_REQUIRE(!stopping, LIG_E_INVALID, "server is stopping");
    const uint64_t t = ++last_ticket;
    Req& r = reqs[t];
    r.log_n = log_n;
    r.poly = poly;
    r.seed = seed;
    r.footprint = prove_footprint(H::BYTES * 8, log_n);
    r.t_submit = std::chrono::steady_clock::now();
    queue.push_back(t);
    cv.notify_all();
    return t;
}

template <class H> size_t Server<H>::warm(const int* log_ns, size_t count) {
    using T = typename H::T;
    LIG_REQUIRE(log_ns || count == 0, LIG_E_INVALID, "NULL size list");
    std::vector<int> sizes(log_ns, log_ns + count);
    for (int L : sizes) LIG_REQUIRE(L >= 16 && L <= 27, LIG_E_INVALID, "log_n must be in [16, 27]");
    if (sizes.empty()) return 0;
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    // one zero input of the largest size, in pinned host memory as the
    // requests that matter most for latency bring theirs
    const size_t bytes = ((size_t)1 << sizes.back()) * sizeof(T);
    LIG_HIP(hipSetDevice(device));
    void* input = nullptr;
    LIG_HIP(hipHostMalloc(&input, bytes, hipHostMallocDefault));
    memset(input, 0, bytes);
    size_t warmed = 0;
    int code = 0;
    std::string msg;
    {
        std::unique_lock<std::mutex> lk(mu);
        if (stopping || warming) {
            lk.unlock();
            (void)hipHostFree(input);
            LIG_REQUIRE(false, LIG_E_INVALID, stopping ? "server is stopping" : "server is already warming");
        }
        warming = true;
        // running requests finish; queued ones wait until the warm-up ends
        cv.wait(lk, [&] { return std::none_of(busy.begin(), busy.end(), [](char b) { return b != 0; }); });
        warm_sizes = sizes;
        warm_input = input;
        warm_todo.assign(workers.size(), 1);
        warm_left = workers.size();
        warm_done = 0;
        warm_full = false;
        warm_code = 0;
        warm_msg.clear();
        cv.notify_all();
        done_cv.wait(lk, [&] { return warm_left == 0; });
        warming = false;
        warm_input = nullptr;
        warmed = warm_done;
        code = warm_code;
        msg = warm_msg;
    }
    cv.notify_all();
    (void)hipHostFree(input);
    if (code) throw Error(code, msg);
    return warmed;
}

template <class H> void Server<H>::warm_worker(size_t wi) {
    Worker<H>& w = *workers[wi];
    size_t n = 0;
    int code = 0;
    std::string msg;
    for (int L : warm_sizes) {
        const size_t fp = prove_footprint(H::BYTES * 8, L);
        {
            std::lock_guard<std::mutex> g(mu);
            if (warm_full || charged() + fp > hbm_budget) {
                warm_full = true;
                break;
            }
            charge[wi] = w.retained() + fp;
            busy[wi] = 1;
        }
        try {
            Proof<H> proof;
            w.prove_one(L, warm_input, false, 0, proof);
            n++;
        } catch (const Error& e) {
            code = e.code;
            msg = e.what();
        } catch (const std::bad_alloc&) {
            code = LIG_E_OOM;
            msg = "host allocation failed";
        } catch (const std::exception& e) {
            code = LIG_E_ASSERT;
            msg = e.what();
        }
        {
            std::lock_guard<std::mutex> g(mu);
            busy[wi] = 0;
            charge[wi] = 0;
        }
        if (code) break;
    }
    {
        std::lock_guard<std::mutex> g(mu);
        // what the worker keeps must fit beside what the others hold
        if (charged() > hbm_budget) {
            w.trim(-1);
            warm_full = true;
        }
        warm_done += n;
        if (code && !warm_code) {
            warm_code = code;
            warm_msg = msg;
        }
        warm_left--;
    }
    done_cv.notify_all();
}

template <class H> size_t Server<H>::charged() const {
    size_t b = 0;
    for (size_t j = 0; j < workers.size(); j++) b += busy[j] ? charge[j] : workers[j]->retained();
    return b;
}

template <class H> bool Server<H>::admit(size_t wi, const Req& r) {
    // Device bytes if r starts on worker wi now: the running workers'
    // charges, the idle workers' retained memory and wi's retained memory
    // plus the request's footprint bound.  Idle memory can be given back:
    // wi keeps only the configuration r will use, the others everything.
    Worker<H>& w = *workers[wi];
    size_t running = 0, idle = 0;
    bool others_busy = false;
    for (size_t j = 0; j < workers.size(); j++) {
        if (j == wi) continue;
        if (busy[j]) {
            running += charge[j];
            others_busy = true;
        } else {
            idle += workers[j]->retained();
        }
    }
    auto fits = [&](size_t mine) { return running + idle + mine <= hbm_budget; };
    if (!fits(w.retained() + r.footprint)) {
        // the least wi can hold is the configuration tables r reuses
        auto kept = w.cfgs.find(r.log_n);
        const size_t own_min = kept == w.cfgs.end() ? 0 : kept->second->device_bytes();
        // a request that cannot fit even with every idle byte returned
        // waits for running ones, unless nothing runs (then it runs alone)
        if (others_busy && running + own_min + r.footprint > hbm_budget) return false;
        w.trim(r.log_n);
        for (size_t j = 0; j < workers.size() && !fits(w.retained() + r.footprint); j++) {
            if (j == wi || busy[j]) continue;
            const size_t before = workers[j]->retained();
            workers[j]->trim(-1);
            idle -= before - workers[j]->retained();
        }
    }
    charge[wi] = w.retained() + r.footprint;
    busy[wi] = 1;
    return true;
}

template <class H> void Server<H>::loop(size_t wi) {
    Worker<H>* w = workers[wi].get();
    for (;;) {
        uint64_t t;
        int log_n;
        const void* poly;
        uint32_t seed;
        {
            std::unique_lock<std::mutex> lk(mu);
            // first come, first admitted: the head waits until it fits
            cv.wait(lk, [&] {
                return stopping || warm_todo[wi] || (!warming && !queue.empty() && admit(wi, reqs[queue.front()]));
            });
            if (warm_todo[wi]) {
                warm_todo[wi] = 0;
                lk.unlock();
                warm_worker(wi);
                continue;
            }
            if (stopping && queue.empty()) return;
            if (queue.empty()) continue;
            t = queue.front();
            queue.pop_front();
            Req& r = reqs[t];
            r.t_start = std::chrono::steady_clock::now();
            log_n = r.log_n;
            poly = r.poly;
            seed = r.seed;
        }
        Proof<H> proof;
        int status = LIG_OK;
        std::string msg;
        try {
            w->prove_one(log_n, poly, false, seed, proof);
        } catch (const Error& e) {
            status = e.code;
            msg = e.what();
        } catch (const std::bad_alloc&) {
            status = LIG_E_OOM;
            msg = "host allocation failed";
        } catch (const std::exception& e) {
            status = LIG_E_ASSERT;
            msg = e.what();
        }
        {
            std::lock_guard<std::mutex> g(mu);
            busy[wi] = 0;
            charge[wi] = 0;
            Req& r = reqs[t];
            r.proof = std::move(proof);
            r.status = status;
            r.message = msg;
            r.t_done = std::chrono::steady_clock::now();
            r.done = true;
        }
        cv.notify_all();
        done_cv.notify_all();
    }
}

template <class H> int Server<H>::take(uint64_t ticket, bool wait, Proof<H>& out, double* latency_ms, std::string& msg) {
    std::unique_lock<std::mutex> lk(mu);
    auto it = reqs.find(ticket);
    LIG_REQUIRE(it != reqs.end(), LIG_E_INVALID, "unknown or already taken ticket");
    if (!it->second.done) {
        if (!wait) return 0;
        // another caller may take (and erase) the same ticket meanwhile
        done_cv.wait(lk, [&] {
            auto jt = reqs.find(ticket);
            return jt == reqs.end() || jt->second.done;
        });
        it = reqs.find(ticket);
        LIG_REQUIRE(it != reqs.end(), LIG_E_INVALID, "ticket taken by another caller");
    }
    Req& r = it->second;
    const int status = r.status;
    msg = r.message;
    if (latency_ms) latency_ms[0] = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
    if (status == LIG_OK) out = std::move(r.proof);
    reqs.erase(it);
    return status == LIG_OK ? 1 : status;
}

template struct Worker<H32>;
template struct Worker<H128>;
template struct Batch<H32>;
template struct Batch<H128>;
template struct Server<H32>;
template struct Server<H128>;

}  // namespace lig
