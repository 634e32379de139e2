// Serving: a request loop over resident compiled circuits with memory-budget admission and continuous
// batching.
//
// A request is (sampler, shots, request seed). Its output rows (bit-packed detection events and
// observables) live in device memory from admission until the caller releases them. Admission takes
// queued requests first-fit in arrival order (within a window of the oldest ones) while their rows fit
// the memory budget and fewer than max_active requests run. Every step builds one work list from the
// running requests and runs it in one launch of the symptom batch kernel: three quarters of the step's
// work go to the requests with the least work left (so a small request that arrives behind large ones
// finishes within a step or two), the rest to the oldest request (so large ones keep moving). A
// request's random stream is keyed by its sampler and its request seed, and a tile's randomness by the
// tile's index inside the request, so a request's result does not depend on what it was batched with
// or how its tiles were split across steps.
#include <chrono>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>

#include "sampler_impl.h"

namespace {

// Request rows are carved out of device segments that are allocated on demand (8 GB, or the request's
// size if larger) up to the budget and kept until the server is destroyed: first-fit over each
// segment's free list of (offset, size) extents, merging neighbours on release, so steady-state
// serving calls neither hipMalloc nor hipFree (hipFree synchronizes the device).
struct Arena {
    struct Segment {
        uint8_t *base;
        size_t cap;
        std::map<size_t, size_t> free_ext;   // offset -> size
    };
    std::vector<Segment> segs;
    size_t total = 0;
    static constexpr size_t SEGMENT = (size_t)8 << 30;

    static size_t round(size_t n) { return (std::max<size_t>(n, 1) + 255) / 256 * 256; }

    // Returns null when nothing fits and no segment can be added within `limit` bytes in total.
    uint8_t *get(size_t n, size_t limit) {
        for (Segment &g : segs) {
            for (auto it = g.free_ext.begin(); it != g.free_ext.end(); ++it) {
                if (it->second >= n) {
                    const size_t off = it->first, sz = it->second;
                    g.free_ext.erase(it);
                    if (sz > n) g.free_ext[off + n] = sz - n;
                    return g.base + off;
                }
            }
        }
        const size_t want = std::max(n, std::min(SEGMENT, limit > total ? limit - total : 0));
        if (total + want > limit || want < n) return nullptr;
        Segment g{nullptr, want, {}};
        check(hipMalloc(&g.base, want), "hipMalloc (serving arena)");
        total += want;
        if (want > n) g.free_ext[n] = want - n;
        segs.push_back(std::move(g));
        return segs.back().base;
    }

    void put(uint8_t *p, size_t n) {
        for (Segment &g : segs) {
            if (p < g.base || p >= g.base + g.cap) continue;
            size_t off = (size_t)(p - g.base), sz = n;
            auto next = g.free_ext.lower_bound(off);
            if (next != g.free_ext.end() && off + sz == next->first) {
                sz += next->second;
                next = g.free_ext.erase(next);
            }
            if (next != g.free_ext.begin()) {
                auto prev = std::prev(next);
                if (prev->first + prev->second == off) {
                    prev->second += sz;
                    return;
                }
            }
            g.free_ext[off] = sz;
            return;
        }
    }

    ~Arena() {
        for (Segment &g : segs) (void)hipFree(g.base);
    }
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct mfs_server {
    enum State { QUEUED, ACTIVE, DONE };
    struct Req {
        mfs_sampler *s = nullptr;
        uint64_t shots = 0;
        State state = QUEUED;
        SymDevParams P{};       // whole-request launch parameters (tile 0 at the start of the rows)
        size_t lds = 0;
        uint64_t tiles = 0, next_tile = 0;
        double tile_cost = 0;   // estimated work per tile (bytes written + 16 bytes per hit)
        uint8_t *dets = nullptr, *obs = nullptr;
        size_t det_row = 0, obs_row = 0;
        size_t bytes = 0;       // arena bytes held against the budget
        double t_submit = 0, t_admit = 0, t_done = 0;
    };

    uint64_t budget;
    double step_cost;           // work per step (same unit as tile_cost)
    size_t max_active = 64;     // requests running at once
    size_t admit_window = 64;   // queued requests considered for first-fit admission
    uint64_t in_use = 0, peak_in_use = 0, steps = 0, next_id = 1;
    std::unordered_map<uint64_t, Req> reqs;
    std::deque<uint64_t> queue;
    std::vector<uint64_t> active, finished;
    Arena arena;
    hipStream_t stream = nullptr;
    DevBuf d_jobs, d_prefix, d_counter;
    SymDevParams *h_jobs = nullptr;   // pinned staging for the work list
    uint64_t *h_prefix = nullptr;
    size_t h_cap = 0;
    std::map<size_t, uint32_t> grid_for_lds;
    // Device time of the serving launches (HIP events around each) and the row bytes they wrote.
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double kernel_ms = 0;
    uint64_t bytes_written = 0;

    mfs_server(uint64_t budget_bytes, double step_cost_bytes) : budget(budget_bytes), step_cost(step_cost_bytes) {
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
        check(hipEventCreate(&ev0), "event");
        check(hipEventCreate(&ev1), "event");
        d_counter.ensure(4);
    }

    ~mfs_server() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (auto &kv : reqs) release_memory(kv.second);
        if (h_jobs) (void)hipHostFree(h_jobs);
        if (h_prefix) (void)hipHostFree(h_prefix);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void release_memory(Req &r) {
        if (r.dets) arena.put(r.dets, r.bytes);
        r.dets = r.obs = nullptr;
        if (r.state != QUEUED) in_use -= r.bytes;
    }

    uint64_t submit(mfs_sampler *s, uint64_t shots, uint64_t seed) {
        if (!s) throw std::invalid_argument("sampler is null");
        if (!shots) throw std::invalid_argument("shots must be positive");
        if (!s->use_symptoms(false)) {
            throw std::invalid_argument("the server runs symptom-engine circuits (deterministic detectors) only: " +
                                        s->sym_why);
        }
        check(hipStreamSynchronize(s->stream), "symptom tables");
        Req r;
        r.s = s;
        r.shots = shots;
        r.det_row = (s->prog.num_detectors + 7) / 8;
        r.obs_row = (s->prog.num_observables + 7) / 8;
        const size_t det_bytes = ((size_t)shots * r.det_row + 255) / 256 * 256;
        const size_t obs_bytes = (size_t)shots * r.obs_row;
        r.bytes = Arena::round(det_bytes + obs_bytes);
        if (r.bytes > budget) throw std::invalid_argument("request output exceeds the server's memory budget");
        static uint8_t probe;
        const auto L = s->layout(MFS_BIT_PACKED | MFS_DEVICE_OUTPUT | (r.obs_row ? MFS_SEPARATE_OBSERVABLES : 0));
        auto sl = s->sym_prepare(L, shots, &probe, r.det_row, r.obs_row ? &probe : nullptr, r.obs_row, 0);
        if (!sl.scan_ok) throw std::invalid_argument("circuit too large for the serving kernel's tile positions");
        // The request's own stream: the sampler's key mixed with the request seed.
        sl.P.seed = splitmix64(sl.P.seed ^ splitmix64(seed ^ 0x7E9A3C15D2B84F61ull));
        r.P = sl.P;
        r.lds = sl.lds_scan;
        r.tiles = sl.tiles;
        double hits = 0;
        for (const SymClass &cl : s->plan.classes) hits += (double)cl.nsites * cl.p;
        r.tile_cost = (double)r.P.shots_per_tile * ((double)(r.det_row + r.obs_row) + 16.0 * hits) + 4096.0;
        r.t_submit = now_s();
        const uint64_t id = next_id++;
        reqs.emplace(id, r);
        queue.push_back(id);
        return id;
    }

    // Admits queued requests first-fit in arrival order while their rows fit the budget.
    void admit() {
        size_t looked = 0;
        for (auto it = queue.begin(); it != queue.end() && active.size() < max_active && looked < admit_window;) {
            looked++;
            Req &r = reqs.at(*it);
            if (in_use + r.bytes > budget) {
                ++it;
                continue;
            }
            uint8_t *base = arena.get(r.bytes, budget);
            if (!base) {   // fragmented: wait for releases
                ++it;
                continue;
            }
            r.dets = base;
            r.obs = base + ((size_t)r.shots * r.det_row + 255) / 256 * 256;
            r.P.dets.ptr = r.dets;
            r.P.dets.stride = r.det_row;
            r.P.obs.ptr = r.obs;
            r.P.obs.stride = r.obs_row;
            in_use += r.bytes;
            peak_in_use = std::max(peak_in_use, in_use);
            r.state = ACTIVE;
            r.t_admit = now_s();
            active.push_back(*it);
            it = queue.erase(it);
        }
    }

    uint32_t grid(size_t lds, uint64_t tiles) {
        auto it = grid_for_lds.find(lds);
        if (it == grid_for_lds.end()) {
            int dev = 0, cus = 0;
            check(hipGetDevice(&dev), "device");
            check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device");
            it = grid_for_lds.emplace(lds, (uint32_t)std::max(cus, 1) * (uint32_t)sym_batch_blocks_per_cu(lds)).first;
        }
        return (uint32_t)std::min<uint64_t>(it->second, tiles);
    }

    // One step: admit, take a fair share of tiles from every active request, run them in one launch,
    // wait, and retire the requests that are complete. Returns false when there was nothing to do.
    bool step() {
        admit();
        if (active.empty()) return false;
        // Shares: three quarters of the step by least remaining work, the rest to the oldest request.
        std::vector<std::pair<uint64_t, uint64_t>> take;   // (request id, tiles this step)
        std::vector<uint64_t> order(active);
        std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
            const Req &x = reqs.at(a), &y = reqs.at(b);
            return (double)(x.tiles - x.next_tile) * x.tile_cost < (double)(y.tiles - y.next_tile) * y.tile_cost;
        });
        double left = 0.75 * step_cost;
        auto give = [&](uint64_t id, double budget_left) -> double {
            Req &r = reqs.at(id);
            uint64_t already = 0;
            for (auto &t : take)
                if (t.first == id) already = t.second;
            const uint64_t want = r.tiles - r.next_tile - already;
            if (!want) return 0;
            const uint64_t n = std::min<uint64_t>(want, std::max<uint64_t>(1, (uint64_t)(budget_left / r.tile_cost)));
            bool found = false;
            for (auto &t : take)
                if (t.first == id) {
                    t.second += n;
                    found = true;
                }
            if (!found) take.push_back({id, n});
            return (double)n * r.tile_cost;
        };
        for (uint64_t id : order) {
            if (left <= 0) break;
            left -= give(id, left);
        }
        left += 0.25 * step_cost;
        for (uint64_t id : active) {   // oldest first
            if (left <= 0) break;
            left -= give(id, left);
        }
        // Work list: one job per request chunk.
        if (take.size() > h_cap) {
            if (h_jobs) (void)hipHostFree(h_jobs);
            if (h_prefix) (void)hipHostFree(h_prefix);
            h_cap = std::max<size_t>(take.size(), 2 * h_cap);
            check(hipHostMalloc((void **)&h_jobs, h_cap * sizeof(SymDevParams), hipHostMallocDefault), "pinned");
            check(hipHostMalloc((void **)&h_prefix, (h_cap + 1) * sizeof(uint64_t), hipHostMallocDefault), "pinned");
            d_jobs.ensure(h_cap * sizeof(SymDevParams));
            d_prefix.ensure((h_cap + 1) * sizeof(uint64_t));
        }
        size_t lds = 64;
        h_prefix[0] = 0;
        for (size_t k = 0; k < take.size(); k++) {
            Req &r = reqs.at(take[k].first);
            SymDevParams P = r.P;
            const uint64_t a = r.next_tile;
            const uint64_t s0 = a * P.shots_per_tile;
            P.tile0 = a;
            P.shots = r.shots - s0;
            P.dets.ptr = r.dets + s0 * r.det_row;
            P.obs.ptr = r.obs + s0 * r.obs_row;
            h_jobs[k] = P;
            h_prefix[k + 1] = h_prefix[k] + take[k].second;
            bytes_written += std::min<uint64_t>(take[k].second * P.shots_per_tile, P.shots) * (r.det_row + r.obs_row);
            lds = std::max(lds, r.lds);
            r.next_tile += take[k].second;
        }
        const uint64_t tiles = h_prefix[take.size()];
        check(hipMemcpyAsync(d_jobs.p, h_jobs, take.size() * sizeof(SymDevParams), hipMemcpyHostToDevice, stream),
              "upload work list");
        check(hipMemcpyAsync(d_prefix.p, h_prefix, (take.size() + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, stream),
              "upload work list");
        const uint32_t g = grid(lds, tiles);
        check(hipEventRecord(ev0, stream), "event");
        check(launch_sym_batch(d_jobs.as<SymDevParams>(), d_prefix.as<uint64_t>(), nullptr, (uint32_t)take.size(),
                               d_counter.as<uint32_t>(), g, lds, stream), "serving step");
        check(hipEventRecord(ev1, stream), "event");
        check(hipStreamSynchronize(stream), "serving step");
        float ms = 0;
        check(hipEventElapsedTime(&ms, ev0, ev1), "event");
        kernel_ms += ms;
        steps++;
        const double t = now_s();
        std::vector<uint64_t> still;
        for (uint64_t id : active) {
            Req &r = reqs.at(id);
            if (r.next_tile == r.tiles) {
                r.state = DONE;
                r.t_done = t;
                finished.push_back(id);
            } else {
                still.push_back(id);
            }
        }
        active.swap(still);
        return true;
    }

    Req &get(uint64_t id) {
        auto it = reqs.find(id);
        if (it == reqs.end()) throw std::out_of_range("unknown request id");
        return it->second;
    }

    void release(uint64_t id) {
        Req &r = get(id);
        if (r.state == ACTIVE) throw std::invalid_argument("request is still running");
        release_memory(r);
        if (r.state == QUEUED) {
            for (auto it = queue.begin(); it != queue.end(); ++it) {
                if (*it == id) {
                    queue.erase(it);
                    break;
                }
            }
        }
        reqs.erase(id);
    }
};

extern "C" {

int mfs_server_create(uint64_t budget_bytes, double step_cost_bytes, mfs_server **out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        if (!budget_bytes) throw std::invalid_argument("budget must be positive");
        *out = new mfs_server(budget_bytes, step_cost_bytes > 0 ? step_cost_bytes : 2e9);
    });
}

void mfs_server_destroy(mfs_server *v) { delete v; }

int mfs_server_submit(mfs_server *v, mfs_sampler *s, uint64_t shots, uint64_t seed, uint64_t *out_id) {
    return guarded([&] {
        const uint64_t id = v->submit(s, shots, seed);
        if (out_id) *out_id = id;
    });
}

int mfs_server_step(mfs_server *v, int *did_work) {
    return guarded([&] {
        const bool w = v->step();
        if (did_work) *did_work = w ? 1 : 0;
    });
}

int mfs_server_poll(mfs_server *v, uint64_t *ids, uint32_t cap, uint32_t *n) {
    return guarded([&] {
        const uint32_t k = (uint32_t)std::min<size_t>(cap, v->finished.size());
        for (uint32_t i = 0; i < k; i++) ids[i] = v->finished[i];
        v->finished.erase(v->finished.begin(), v->finished.begin() + k);
        if (n) *n = k;
    });
}

int mfs_server_result(mfs_server *v, uint64_t id, uint8_t **dets, size_t *dets_stride, uint8_t **obs,
                      size_t *obs_stride, uint64_t *shots, double times[3]) {
    return guarded([&] {
        auto &r = v->get(id);
        if (r.state != mfs_server::DONE) throw std::invalid_argument("request has not finished");
        if (dets) *dets = r.dets;
        if (dets_stride) *dets_stride = r.det_row;
        if (obs) *obs = r.obs_row ? r.obs : nullptr;
        if (obs_stride) *obs_stride = r.obs_row;
        if (shots) *shots = r.shots;
        if (times) {
            times[0] = r.t_submit;
            times[1] = r.t_admit;
            times[2] = r.t_done;
        }
    });
}

int mfs_server_copy_result(mfs_server *v, uint64_t id, uint8_t *dets, size_t dets_stride, uint8_t *obs,
                           size_t obs_stride) {
    return guarded([&] {
        auto &r = v->get(id);
        if (r.state != mfs_server::DONE) throw std::invalid_argument("request has not finished");
        if (dets && r.det_row)
            check(copy_rows_async(dets, dets_stride, r.dets, r.det_row, r.det_row, r.shots, hipMemcpyDeviceToHost,
                                   v->stream), "copy result");
        if (obs && r.obs_row)
            check(copy_rows_async(obs, obs_stride, r.obs, r.obs_row, r.obs_row, r.shots, hipMemcpyDeviceToHost,
                                   v->stream), "copy result");
        check(hipStreamSynchronize(v->stream), "copy result");
    });
}

int mfs_server_release(mfs_server *v, uint64_t id) {
    return guarded([&] { v->release(id); });
}

int mfs_server_stats(mfs_server *v, uint64_t out[8]) {
    return guarded([&] {
        out[0] = v->steps;
        out[1] = v->queue.size();
        out[2] = v->active.size();
        out[3] = v->in_use;
        out[4] = v->peak_in_use;
        out[5] = v->budget;
        out[6] = (uint64_t)(v->kernel_ms * 1000.0);
        out[7] = v->bytes_written;
    });
}

int mfs_server_replay(mfs_server *v, uint64_t n, mfs_sampler *const *samplers, const uint64_t *shots,
                      const uint64_t *seeds, const double *arrival_s, double *latency_ms, double *wall_s) {
    return guarded([&] {
        for (uint64_t i = 1; i < n; i++) {
            if (arrival_s[i] < arrival_s[i - 1]) throw std::invalid_argument("arrival times must be sorted");
        }
        std::unordered_map<uint64_t, uint64_t> index;   // request id -> trace position
        std::vector<double> arrive_at(n);
        const double t0 = now_s();
        uint64_t next = 0, done = 0;
        std::vector<uint64_t> ids(256);
        while (done < n) {
            const double t = now_s() - t0;
            while (next < n && arrival_s[next] <= t) {
                const uint64_t id = v->submit(samplers[next], shots[next], seeds[next]);
                index[id] = next;
                arrive_at[next] = t0 + arrival_s[next];
                next++;
            }
            const bool worked = v->step();
            for (uint64_t id : v->finished) {
                const uint64_t k = index.at(id);
                latency_ms[k] = (v->reqs.at(id).t_done - arrive_at[k]) * 1e3;
                v->release(id);
                done++;
            }
            v->finished.clear();
            if (!worked && next < n) {
                const double wait = arrival_s[next] - (now_s() - t0);
                if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
            }
        }
        if (wall_s) *wall_s = now_s() - t0;
    });
}

}  // extern "C"
