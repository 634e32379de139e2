// Batches: many sampling jobs (each a sampler, a shot count and device output rows) run by one
// persistent kernel. Jobs of one batch may come from different circuits. Each job draws the same
// random stream as mfs_sample on its sampler would at the same point of the sampler's stream, with the
// same tile size, so a batch run equals the separate launches bit for bit.
#include "sampler_impl.h"

struct mfs_batch {
    struct Job {
        mfs_sampler *s;
        uint64_t shots;
        uint8_t *dets;
        size_t dets_stride;
        uint8_t *obs;
        size_t obs_stride;
    };
    std::vector<Job> jobs;
    int tile_kb = 0;                 // LDS output image per workgroup (0: each sampler's own setting)
    int persistent = 1;              // 1: persistent workgroups claiming tiles; 0: one workgroup per tile
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf d_jobs, d_prefix, d_counter, d_advance;
    // Occupancy levels: jobs grouped by the workgroups per CU their own LDS allows, one persistent
    // kernel per level (each at the occupancy a plain launch of its jobs would get).
    struct Level {
        size_t lds;
        uint32_t grid;
        uint32_t first, count;   // jobs [first, first + count) of the uploaded list
        uint64_t tiles;
    };
    std::vector<Level> levels;
    // Prepared state.
    bool prepared = false;
    std::vector<uint32_t> sym_jobs;      // job indices run by the persistent kernel, in launch order
    std::vector<uint32_t> frame_jobs;    // job indices run as their own launches
    std::vector<std::pair<mfs_sampler *, uint64_t>> advance;  // RNG tiles each sampler uses per run (all its jobs)
    std::vector<std::pair<mfs_sampler *, uint64_t>> expect;   // sampler tile counters the params assume
    uint64_t total_tiles = 0;
    size_t lds = 0;      // largest level's LDS
    uint32_t grid = 0;   // workgroups over all levels
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~mfs_batch() {
        drop_graph();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    void invalidate() {
        prepared = false;
        drop_graph();
    }

    void prepare() {
        invalidate();
        sym_jobs.clear();
        frame_jobs.clear();
        advance.clear();
        expect.clear();
        // Per sampler: the tile counter at the start of the run and the tiles its jobs take.
        std::vector<std::pair<mfs_sampler *, uint64_t>> next_tile;
        auto slot = [&](mfs_sampler *s) -> uint64_t & {
            for (auto &kv : next_tile)
                if (kv.first == s) return kv.second;
            next_tile.push_back({s, s->tiles_used});
            expect.push_back({s, s->tiles_used});
            return next_tile.back().second;
        };
        struct Item {
            SymDevParams P;
            uint64_t tiles;
            double cost;
            uint32_t job;
            size_t lds;
            int per_cu;
            mfs_sampler *s;
        };
        std::vector<Item> items;
        for (uint32_t j = 0; j < jobs.size(); j++) {
            Job &jb = jobs[j];
            mfs_sampler *s = jb.s;
            const bool sym = s->use_symptoms(false);
            if (!sym) {
                frame_jobs.push_back(j);
                continue;
            }
            check(hipStreamSynchronize(s->stream), "symptom tables");
            const uint32_t flags = MFS_BIT_PACKED | MFS_DEVICE_OUTPUT | (jb.obs ? MFS_SEPARATE_OBSERVABLES : 0);
            const auto L = s->layout(flags);
            auto r = s->sym_prepare(L, jb.shots, jb.dets, jb.dets_stride, jb.obs, jb.obs_stride, 0, tile_kb);
            if (!r.scan_ok || !s->sym_scan) {
                frame_jobs.push_back(j);
                continue;
            }
            if (r.lds_scan + 64 > MAX_LDS) throw std::invalid_argument("a batch job's tile does not fit in LDS");
            uint64_t &t0 = slot(s);
            r.P.tile0 = t0;
            t0 += r.tiles;
            double hits = 0;
            for (const SymClass &cl : s->plan.classes) hits += (double)cl.nsites * cl.p;
            const double S = r.P.shots_per_tile;
            // Per-tile cost estimate: hits applied plus output bytes written, plus a fixed overhead.
            const double cost = S * (hits * 16.0 + r.P.det_row_bytes + r.P.obs_row_bytes) + 4096.0;
            items.push_back({r.P, r.tiles, cost, j, std::max<size_t>(r.lds_scan, 64),
                             sym_batch_blocks_per_cu(std::max<size_t>(r.lds_scan, 64)), s});
        }
        for (auto &kv : next_tile) {
            uint64_t start = 0;
            for (auto &e : expect)
                if (e.first == kv.first) start = e.second;
            advance.push_back({kv.first, kv.second - start});
        }
        // Levels by occupancy (highest first); inside a level, heaviest tiles first so the tail of the
        // run is made of light tiles.
        std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
            return a.per_cu != b.per_cu ? a.per_cu > b.per_cu : a.cost > b.cost;
        });
        std::vector<SymDevParams> params;
        std::vector<uint64_t> prefix, adv;
        levels.clear();
        total_tiles = 0;
        lds = 64;
        for (size_t i = 0; i < items.size();) {
            size_t e = i;
            Level lv{0, 0, (uint32_t)i, 0, 0};
            std::vector<uint64_t> pre{0};
            while (e < items.size() && items[e].per_cu == items[i].per_cu) {
                params.push_back(items[e].P);
                pre.push_back(pre.back() + items[e].tiles);
                uint64_t a = 0;
                for (auto &kv : advance)
                    if (kv.first == items[e].s) a = kv.second;
                adv.push_back(a);
                sym_jobs.push_back(items[e].job);
                lv.lds = std::max(lv.lds, items[e].lds);
                e++;
            }
            lv.count = (uint32_t)(e - i);
            lv.tiles = pre.back();
            if (lv.tiles >= (1ull << 32)) throw std::invalid_argument("too many tiles in one batch");
            prefix.insert(prefix.end(), pre.begin(), pre.end());
            total_tiles += lv.tiles;
            lds = std::max(lds, lv.lds);
            levels.push_back(lv);
            i = e;
        }
        if (!stream) {
            check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
            check(hipEventCreate(&ev0), "event");
            check(hipEventCreate(&ev1), "event");
        }
        grid = 0;
        if (!params.empty()) {
            d_jobs.ensure(params.size() * sizeof(SymDevParams));
            d_prefix.ensure(prefix.size() * sizeof(uint64_t));
            d_advance.ensure(adv.size() * sizeof(uint64_t));
            d_counter.ensure(levels.size() * 4);
            check(hipMemcpyAsync(d_jobs.p, params.data(), params.size() * sizeof(SymDevParams), hipMemcpyHostToDevice,
                                 stream), "upload batch");
            check(hipMemcpyAsync(d_prefix.p, prefix.data(), prefix.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                                 stream), "upload batch");
            check(hipMemcpyAsync(d_advance.p, adv.data(), adv.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                                 stream), "upload batch");
            check(hipStreamSynchronize(stream), "upload batch");
            int dev = 0, cus = 0;
            check(hipGetDevice(&dev), "device");
            check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device");
            for (Level &lv : levels) {
                const uint64_t resident = (uint64_t)std::max(cus, 1) * (uint64_t)sym_batch_blocks_per_cu(lv.lds);
                lv.grid = persistent ? (uint32_t)std::min<uint64_t>(resident, lv.tiles) : (uint32_t)lv.tiles;
                grid += lv.grid;
            }
        }
        prepared = true;
    }

    // Re-prepares when a sampler's tile counter moved outside this batch (a plain call in between).
    void sync_counters() {
        if (!prepared) {
            prepare();
            return;
        }
        for (auto &e : expect) {
            if (e.first->tiles_used != e.second) {
                prepare();
                return;
            }
        }
    }

    void enqueue_sym() {
        // Level l's jobs start at levels[l].first; its prefix array (count + 1 entries) at first + l.
        for (size_t l = 0; l < levels.size(); l++) {
            const Level &lv = levels[l];
            check(launch_sym_batch(d_jobs.as<SymDevParams>() + lv.first, d_prefix.as<uint64_t>() + lv.first + l,
                                   d_advance.as<uint64_t>() + lv.first, lv.count, d_counter.as<uint32_t>() + l,
                                   lv.grid, lv.lds, stream),
                  "batch launch");
        }
    }

    // One run: the persistent kernel (plain launches or a graph replay), then the frame-engine jobs.
    void run_once(bool use_graph) {
        sync_counters();
        if (grid) {
            if (use_graph) {
                if (!exec) {
                    check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "graph capture");
                    enqueue_sym();
                    check(hipStreamEndCapture(stream, &graph), "graph capture");
                    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
                }
                check(hipGraphLaunch(exec, stream), "graph launch");
            } else {
                enqueue_sym();
            }
        }
        for (auto &a : advance) a.first->tiles_used += a.second;
        for (auto &e : expect) {
            for (auto &a : advance)
                if (a.first == e.first) e.second += a.second;
        }
        for (uint32_t j : frame_jobs) {
            Job &jb = jobs[j];
            const uint32_t flags = MFS_BIT_PACKED | MFS_DEVICE_OUTPUT | (jb.obs ? MFS_SEPARATE_OBSERVABLES : 0);
            jb.s->sample(jb.shots, flags, jb.dets, jb.dets_stride, jb.obs, jb.obs_stride);
            // A sampler with jobs on both paths: its device tile bases are now stale.
            for (auto &e : expect)
                if (e.first == jb.s) invalidate();
        }
    }
};

extern "C" {

int mfs_batch_create(mfs_batch **out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        *out = new mfs_batch();
    });
}

void mfs_batch_destroy(mfs_batch *b) { delete b; }

int mfs_batch_add(mfs_batch *b, mfs_sampler *s, uint64_t shots, uint8_t *dets, size_t dets_stride, uint8_t *obs,
                  size_t obs_stride) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("sampler is null");
        if (!dets && shots) throw std::invalid_argument("dets is null");
        const size_t drow = (s->prog.num_detectors + 7) / 8, orow = (s->prog.num_observables + 7) / 8;
        if (dets_stride < drow) throw std::invalid_argument("dets_stride is smaller than a packed detector row");
        if (obs && obs_stride < orow) throw std::invalid_argument("obs_stride is smaller than a packed observable row");
        b->jobs.push_back({s, shots, dets, dets_stride, obs, obs_stride});
        b->invalidate();
    });
}

int mfs_batch_set_tile_kb(mfs_batch *b, int kb) {
    return guarded([&] {
        if (kb < 0 || kb > 128) throw std::invalid_argument("batch tile size must be 0..128 KB");
        b->tile_kb = kb;
        b->invalidate();
    });
}

int mfs_batch_run(mfs_batch *b, int use_graph) {
    return guarded([&] {
        b->run_once(use_graph != 0);
        check(hipStreamSynchronize(b->stream ? b->stream : nullptr), "batch run");
    });
}

int mfs_batch_time(mfs_batch *b, int use_graph, int warmup, int iters, double out_ms[2]) {
    return guarded([&] {
        if (iters < 1 || warmup < 0) throw std::invalid_argument("iters must be >= 1 and warmup >= 0");
        for (int i = 0; i < warmup; i++) b->run_once(use_graph != 0);
        b->sync_counters();
        check(hipStreamSynchronize(b->stream), "warmup");
        const auto t0 = std::chrono::steady_clock::now();
        check(hipEventRecord(b->ev0, b->stream), "event");
        for (int i = 0; i < iters; i++) b->run_once(use_graph != 0);
        check(hipEventRecord(b->ev1, b->stream), "event");
        check(hipEventSynchronize(b->ev1), "event");
        const auto t1 = std::chrono::steady_clock::now();
        float ms = 0;
        check(hipEventElapsedTime(&ms, b->ev0, b->ev1), "elapsed");
        out_ms[0] = ms / iters;
        out_ms[1] = std::chrono::duration<double, std::milli>(t1 - t0).count() / iters;
    });
}

int mfs_batch_autotune(mfs_batch *b, const int *candidates, int n, int iters, int *best_kb, double *best_ms) {
    return guarded([&] {
        if (!candidates || n < 1) throw std::invalid_argument("no candidate tile sizes");
        double best = 1e300;
        int pick = candidates[0], pick_sched = 1;
        for (int i = 0; i < n; i++) {
            if (candidates[i] < 1 || candidates[i] > 128) throw std::invalid_argument("candidate tile sizes must be 1..128 KB");
            for (int sched = 1; sched >= 0; sched--) {
                b->tile_kb = candidates[i];
                b->persistent = sched;
                b->invalidate();
                b->run_once(false);
                check(hipStreamSynchronize(b->stream), "autotune");
                const auto t0 = std::chrono::steady_clock::now();
                for (int k = 0; k < std::max(iters, 1); k++) b->run_once(false);
                check(hipStreamSynchronize(b->stream), "autotune");
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
                                  std::max(iters, 1);
                if (ms < best) {
                    best = ms;
                    pick = candidates[i];
                    pick_sched = sched;
                }
            }
        }
        b->tile_kb = pick;
        b->persistent = pick_sched;
        b->invalidate();
        // Plain calls on these samplers use the chosen tiles too, so they stay comparable to the batch.
        for (auto &jb : b->jobs) jb.s->sym_tile_kb = pick;
        if (best_kb) *best_kb = pick;
        if (best_ms) *best_ms = best;
    });
}

int mfs_batch_set_schedule(mfs_batch *b, int persistent) {
    return guarded([&] {
        b->persistent = persistent ? 1 : 0;
        b->invalidate();
    });
}

int mfs_batch_info(mfs_batch *b, uint64_t out[8]) {
    return guarded([&] {
        b->sync_counters();
        out[0] = b->jobs.size();
        out[1] = b->sym_jobs.size();
        out[2] = b->frame_jobs.size();
        out[3] = b->total_tiles;
        out[4] = b->grid;   // workgroups over all occupancy levels
        out[5] = b->lds;
        out[6] = (uint64_t)b->persistent;
        out[7] = (uint64_t)b->tile_kb;
    });
}

}  // extern "C"
