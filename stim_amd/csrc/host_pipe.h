// Double-buffered delivery of batch outputs from device memory to the caller's host rows.
//
// A call that returns rows in host memory runs in batches. Batch i's kernels write into device slot
// i % 2; a copy engine moves the slot's rows to pinned host buffer i % 2 on a second stream; host
// threads then copy them from the pinned buffer to the caller's rows (pageable, any row stride). While
// batch i's rows cross PCIe and are copied on the host, batch i + 1's kernels run into the other slot.
// Per slot, three orderings keep the buffers consistent:
//   - the DMA of batch i waits for batch i's kernels (event `produced`, a wait on the copy stream);
//   - batch i + 2's kernels wait for batch i's DMA out of the slot (event `copied`, a wait on the
//     compute stream: the host does not block);
//   - batch i + 2's DMA into the pinned buffer waits for the host copy of batch i out of it (a join of
//     that batch's host threads, before the DMA is queued).
// Both streams have the default priority.
#pragma once

#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace sa {

// Copies `rows` rows of `width` bytes between pitched buffers. Contiguous rows (both pitches equal to
// the width) go as one linear copy: the runtime splits a pitched copy into per-row transfers, which
// costs microseconds per row.
inline hipError_t copy_rows_async(void *dst, size_t dst_pitch, const void *src, size_t src_pitch, size_t width,
                                  size_t rows, hipMemcpyKind kind, hipStream_t stream) {
    if (width == 0 || rows == 0) return hipSuccess;
    if (dst_pitch == width && src_pitch == width) return hipMemcpyAsync(dst, src, width * rows, kind, stream);
    return hipMemcpy2DAsync(dst, dst_pitch, src, src_pitch, width, rows, kind, stream);
}

class HostPipe {
   public:
    // One output of a batch: `rows` rows of `row` bytes at `dev` (stride `dev_stride`) go to `host`
    // (stride `host_stride`); when `newline_at` >= 0 each host row also gets '\n' at that byte.
    struct Part {
        const uint8_t *dev;
        size_t dev_stride;
        size_t row;
        uint8_t *host;
        size_t host_stride;
        long newline_at;
    };

    HostPipe() = default;
    HostPipe(const HostPipe &) = delete;
    HostPipe &operator=(const HostPipe &) = delete;
    ~HostPipe() {
        drain_noexcept();
        for (int s = 0; s < 2; s++) {
            if (dev_[s]) (void)hipFree(dev_[s]);
            if (pin_[s]) (void)hipHostFree(pin_[s]);
            if (produced_[s]) (void)hipEventDestroy(produced_[s]);
            if (copied_[s]) (void)hipEventDestroy(copied_[s]);
        }
        if (copy_) (void)hipStreamDestroy(copy_);
    }

    // Host threads per batch copy: the CPUs this process may run on, at most 8.
    static int host_threads() {
        cpu_set_t set;
        int n = 1;
        if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
        return std::max(1, std::min(n, 8));
    }

    // Prepares a call on `compute`: both device slots hold at least `dev_bytes`, both pinned buffers
    // at least `host_bytes`. Nothing of an earlier call is in flight (finish() or drain ended it).
    void begin(hipStream_t compute, size_t dev_bytes, size_t host_bytes) {
        compute_ = compute;
        if (!copy_) {
            chk(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking), "copy stream");
            for (int s = 0; s < 2; s++) {
                chk(hipEventCreateWithFlags(&produced_[s], hipEventDisableTiming), "event");
                chk(hipEventCreateWithFlags(&copied_[s], hipEventDisableTiming), "event");
            }
        }
        for (int s = 0; s < 2; s++) {
            if (dev_cap_[s] < dev_bytes) {
                if (dev_[s]) chk(hipFree(dev_[s]), "hipFree");
                dev_[s] = nullptr;
                dev_cap_[s] = 0;
                chk(hipMalloc(reinterpret_cast<void **>(&dev_[s]), dev_bytes), "hipMalloc (host pipeline slot)");
                dev_cap_[s] = dev_bytes;
            }
            if (pin_cap_[s] < host_bytes) {
                if (pin_[s]) chk(hipHostFree(pin_[s]), "hipHostFree");
                pin_[s] = nullptr;
                pin_cap_[s] = 0;
                chk(hipHostMalloc(reinterpret_cast<void **>(&pin_[s]), host_bytes, hipHostMallocDefault),
                    "hipHostMalloc (host pipeline buffer)");
                pin_cap_[s] = host_bytes;
            }
            used_[s] = false;
        }
        batch_ = 0;
        pending_ = -1;
        nthreads_ = host_threads();
    }

    // Device slot of the next batch, with its capacity. The compute stream first waits until the DMA
    // of the batch that last used the slot has read it.
    uint8_t *next_slot(size_t *cap) {
        const int s = batch_ & 1;
        if (used_[s]) chk(hipStreamWaitEvent(compute_, copied_[s], 0), "wait for slot");
        *cap = dev_cap_[s];
        return dev_[s];
    }

    // Queues the DMA of the batch whose kernels were just queued into next_slot()'s slot, then starts
    // the host copy of the batch before it (waiting for that batch's DMA to end).
    void submit(const std::vector<Part> &parts, uint64_t rows) {
        const int s = batch_ & 1;
        size_t need = 0;
        for (const Part &p : parts) need += p.row * rows;
        if (need > pin_cap_[s]) throw std::logic_error("host pipeline batch exceeds its pinned buffer");
        join(s);
        chk(hipEventRecord(produced_[s], compute_), "record");
        chk(hipStreamWaitEvent(copy_, produced_[s], 0), "wait for kernels");
        Batch &b = batches_[s];
        b.parts.clear();
        b.rows = rows;
        size_t off = 0;
        for (const Part &p : parts) {
            if (p.row && rows) {
                chk(copy_rows_async(pin_[s] + off, p.row, p.dev, p.dev_stride, p.row, rows, hipMemcpyDeviceToHost,
                                    copy_),
                    "copy to pinned buffer");
            }
            b.parts.push_back(Part{pin_[s] + off, p.row, p.row, p.host, p.host_stride, p.newline_at});
            off += p.row * rows;
        }
        chk(hipEventRecord(copied_[s], copy_), "record");
        used_[s] = true;
        if (pending_ >= 0) start_host_copy(pending_);
        pending_ = s;
        batch_++;
    }

    // Delivers the last batch and waits for every host copy.
    void finish() {
        if (pending_ >= 0) start_host_copy(pending_);
        pending_ = -1;
        join(0);
        join(1);
    }

    // Ends whatever is in flight without reporting errors (used when a call is unwinding).
    void drain_noexcept() {
        for (int s = 0; s < 2; s++) join(s);
        if (copy_) (void)hipStreamSynchronize(copy_);
        pending_ = -1;
    }

   private:
    struct Batch {
        std::vector<Part> parts;   // dev = pinned source rows (dense), dev_stride = row
        uint64_t rows = 0;
        std::vector<std::thread> workers;
    };

    static void chk(hipError_t e, const char *what) {
        if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    }

    static void copy_rows(const Batch &b, uint64_t r0, uint64_t r1) {
        for (const Part &p : b.parts) {
            if (!p.row && p.newline_at < 0) continue;
            const uint8_t *src = p.dev + r0 * p.row;
            uint8_t *dst = p.host + r0 * p.host_stride;
            if (p.host_stride == p.row && p.newline_at < 0) {
                std::memcpy(dst, src, (r1 - r0) * p.row);
                continue;
            }
            for (uint64_t r = r0; r < r1; r++, src += p.row, dst += p.host_stride) {
                std::memcpy(dst, src, p.row);
                if (p.newline_at >= 0) dst[p.newline_at] = '\n';
            }
        }
    }

    void start_host_copy(int s) {
        chk(hipEventSynchronize(copied_[s]), "copy to pinned buffer");
        Batch &b = batches_[s];
        const uint64_t rows = b.rows;
        size_t bytes = 0;
        for (const Part &p : b.parts) bytes += p.row * rows;
        // Threads only where each gets at least 1 MiB.
        const uint64_t k = std::max<uint64_t>(1, std::min<uint64_t>(nthreads_, bytes >> 20));
        if (k == 1) {
            copy_rows(b, 0, rows);
            return;
        }
        for (uint64_t t = 0; t < k; t++) {
            const uint64_t r0 = rows * t / k, r1 = rows * (t + 1) / k;
            b.workers.emplace_back([&b, r0, r1] { copy_rows(b, r0, r1); });
        }
    }

    void join(int s) {
        for (std::thread &t : batches_[s].workers) t.join();
        batches_[s].workers.clear();
    }

    hipStream_t compute_ = nullptr, copy_ = nullptr;
    hipEvent_t produced_[2] = {nullptr, nullptr}, copied_[2] = {nullptr, nullptr};
    uint8_t *dev_[2] = {nullptr, nullptr};
    uint8_t *pin_[2] = {nullptr, nullptr};
    size_t dev_cap_[2] = {0, 0}, pin_cap_[2] = {0, 0};
    bool used_[2] = {false, false};
    Batch batches_[2];
    uint64_t batch_ = 0;
    int pending_ = -1;
    int nthreads_ = 1;
};

}  // namespace sa
