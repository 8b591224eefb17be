// table_writer.cpp -- background writes of resident tables to files in the
// reference layout (include/fdtomoc_amd.h, fdgpu_tables_write_async): the
// file step of FDtomoC's sphfd (src/sphfd.c:5702-5803) taken off the solve's
// critical path.
//
// A worker thread per context owns a copy stream and two pinned buffers.  For
// each queued table it starts the device-to-host copy into one buffer, then
// writes the previous table's file from the other, so the copies, the file
// writes and the next fdgpu_solve on the compute stream all overlap.  Jobs
// capture the table's device pointer and its header inputs at submission;
// the table cannot be released (FDGPU_EBUSY) while a job reads it.
#include <hip/hip_runtime.h>
#include <string.h>

#include "fdgpu_internal.h"
#include "sphfd_io.h"

using namespace fdgpu_detail;

namespace {

// Writes one job's file from host memory (the header as fdgpu_table_write
// derives it).
int write_job(const WriteJob &j, const float *host) {
    fdio::VHead head;
    if (j.has_head) {
        memcpy(&head, j.head, sizeof head);
    } else {
        memset(&head, 0, sizeof head);
        head = fdio::table_header(to_values(j.par), j.prep, head, false);
    }
    return fdio::write_table(j.path.c_str(), j.par, j.prep, &head, host);
}

void worker(fdgpu_ctx *c, Writer *W) {
    (void)hipSetDevice(c->device);
    WriteJob prev;
    int prev_slot = -1;   // buffer holding (or receiving) the previous table, -1: none
    // completes one job's bookkeeping
    auto settle = [&](int table, int rc) {
        std::lock_guard<std::mutex> lk(W->mu);
        if (rc && !W->err) W->err = rc;
        if (--W->busy[table] == 0) W->busy.erase(table);
        if (--W->inflight == 0) W->idle.notify_all();
    };
    // waits for the previous table's copy and writes its file
    auto finish_prev = [&]() {
        int rc = hipEventSynchronize(W->done[prev_slot]) == hipSuccess ? FDGPU_OK : FDGPU_EHIP;
        if (!rc) rc = write_job(prev, W->pinned[prev_slot]);
        settle(prev.table, rc);
        prev_slot = -1;
    };
    for (;;) {
        WriteJob job;
        bool have = false;
        {
            std::unique_lock<std::mutex> lk(W->mu);
            // with a copy in flight, do not sleep: write its file first
            if (prev_slot < 0) W->cv.wait(lk, [&] { return W->stop || !W->q.empty(); });
            if (!W->q.empty()) {
                job = std::move(W->q.front());
                W->q.pop_front();
                have = true;
            } else if (prev_slot < 0) {
                return;   // stopped and drained
            }
        }
        int rc = FDGPU_OK, slot = 0;
        if (have) {
            const size_t bytes = sizeof(float) * (size_t)job.n;
            if (bytes > W->pinned_bytes) {
                // larger buffers: finish the previous table first, then replace both
                if (prev_slot >= 0) finish_prev();
                W->pinned_bytes = 0;
                for (int b = 0; b < 2; b++) {
                    if (W->pinned[b]) (void)hipHostFree(W->pinned[b]);
                    W->pinned[b] = nullptr;
                    if (hipHostMalloc((void **)&W->pinned[b], bytes, hipHostMallocDefault) != hipSuccess)
                        rc = FDGPU_ENOMEM;
                }
                if (!rc) W->pinned_bytes = bytes;
            }
            slot = prev_slot < 0 ? 0 : 1 - prev_slot;
            if (!rc && (hipMemcpyAsync(W->pinned[slot], job.d_time, bytes, hipMemcpyDeviceToHost, W->stream) !=
                            hipSuccess ||
                        hipEventRecord(W->done[slot], W->stream) != hipSuccess))
                rc = FDGPU_EHIP;
        }
        // the previous table's file, while this one copies
        if (prev_slot >= 0) finish_prev();
        if (have) {
            if (rc) {
                settle(job.table, rc);
            } else {
                prev = std::move(job);
                prev_slot = slot;
            }
        }
    }
}

}  // namespace

namespace fdgpu_detail {

bool writer_busy(fdgpu_ctx *c, int table) {
    Writer *W = c->writer.get();
    if (!W) return false;
    std::lock_guard<std::mutex> lk(W->mu);
    return W->busy.count(table) != 0;
}

void writer_shutdown(fdgpu_ctx *c) {
    Writer *W = c->writer.get();
    if (!W) return;
    {
        std::lock_guard<std::mutex> lk(W->mu);
        W->stop = true;
    }
    W->cv.notify_all();
    if (W->th.joinable()) W->th.join();
    (void)hipSetDevice(c->device);
    for (int b = 0; b < 2; b++) {
        if (W->pinned[b]) (void)hipHostFree(W->pinned[b]);
        if (W->done[b]) (void)hipEventDestroy(W->done[b]);
    }
    if (W->stream) (void)hipStreamDestroy(W->stream);
    c->writer.reset();
}

}  // namespace fdgpu_detail

extern "C" int fdgpu_tables_write_async(fdgpu_ctx *c, const int *tables, const char *const *paths,
                                        const void *const *vel_headers, int n) {
    if (!c || n < 0 || (n > 0 && (!tables || !paths))) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    std::vector<WriteJob> jobs((size_t)n);
    for (int k = 0; k < n; k++) {
        auto it = c->tables.find(tables[k]);
        if (it == c->tables.end() || it->second.from_file || !paths[k]) return FDGPU_EINPUT;
        WriteJob &j = jobs[(size_t)k];
        j.table = tables[k];
        j.d_time = it->second.d_time;
        j.n = it->second.n;
        j.par = it->second.par;
        j.prep = it->second.prep;
        j.has_head = vel_headers && vel_headers[k];
        if (j.has_head) memcpy(j.head, vel_headers[k], sizeof j.head);
        j.path = paths[k];
    }
    if (!c->writer) {
        std::unique_ptr<Writer> W(new Writer());
        if (hipSetDevice(c->device) != hipSuccess ||
            hipStreamCreateWithFlags(&W->stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&W->done[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&W->done[1], hipEventDisableTiming) != hipSuccess)
            return FDGPU_EHIP;
        c->writer = std::move(W);
        c->writer->th = std::thread(worker, c, c->writer.get());
    }
    Writer *W = c->writer.get();
    {
        std::lock_guard<std::mutex> wl(W->mu);
        for (auto &j : jobs) {
            W->busy[j.table]++;
            W->inflight++;
            W->q.push_back(std::move(j));
        }
    }
    W->cv.notify_all();
    return FDGPU_OK;
}

extern "C" int fdgpu_write_wait(fdgpu_ctx *c) {
    if (!c) return FDGPU_EINPUT;
    Writer *W = c->writer.get();
    if (!W) return FDGPU_OK;
    std::unique_lock<std::mutex> lk(W->mu);
    W->idle.wait(lk, [&] { return W->inflight == 0; });
    const int rc = W->err;
    W->err = 0;
    return rc;
}
