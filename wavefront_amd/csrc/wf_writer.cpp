// Snapshot writer thread (see wf_writer.h).
#include "wf_writer.h"

#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <cstdlib>
#include <cstring>

#include "wf_internal.h"

namespace wf {

SnapshotWriter::~SnapshotWriter() { (void)finish(); }

int SnapshotWriter::open(const char *path, uint64_t snapshot_bytes, int device,
                         uint64_t chunk_bytes, int nbuf)
{
    if (!path || snapshot_bytes == 0 || snapshot_bytes % kAlign)
        return fail(WF_ERR_INVALID, "snapshot writer: bad path or size");
    if (nbuf < 2) nbuf = 2;
    if (nbuf > kMaxBuf) nbuf = kMaxBuf;
    chunk_bytes = (chunk_bytes + kAlign - 1) / kAlign * kAlign;
    if (chunk_bytes > snapshot_bytes) chunk_bytes = snapshot_bytes;
    if (chunk_bytes > kMaxWrite) chunk_bytes = kMaxWrite;
    opened_ = true;
    snap_ = snapshot_bytes;
    chunk_ = chunk_bytes;
    device_ = device;
    // O_DIRECT as the reference does (src/engine.rs:582, 688-693); file
    // systems that refuse it (tmpfs) get ordinary buffered writes.
    fd_ = ::open(path, O_WRONLY | O_CREAT | O_TRUNC | O_DIRECT, 0644);
    direct_ = fd_ >= 0;
    if (fd_ < 0 && (errno == EINVAL || errno == EOPNOTSUPP))
        fd_ = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd_ < 0) return fail(WF_ERR_IO, "cannot open snapshot file %s: %s", path, std::strerror(errno));
    if (hipSetDevice(device) != hipSuccess)
        return fail(WF_ERR_HIP, "snapshot writer: hipSetDevice(%d) failed", device);
    for (auto &st : stream_)
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
            return fail(WF_ERR_HIP, "snapshot writer: stream creation failed");
    nbuf_ = nbuf;
    for (int i = 0; i < nbuf_; ++i) {
        if (hipHostMalloc(&pinned_[i], chunk_, hipHostMallocDefault) != hipSuccess)
            return fail(WF_ERR_OOM, "snapshot writer: pinned buffer of %llu bytes",
                        (unsigned long long)chunk_);
        if (hipEventCreateWithFlags(&copied_[i], hipEventDisableTiming) != hipSuccess)
            return fail(WF_ERR_HIP, "snapshot writer: event creation failed");
    }
    stop_ = false;
    io_stop_ = false;
    error_ = 0;
    pipelined_ = false;
    if (direct_) {
        void *a = nullptr, *b = nullptr;
        if (::posix_memalign(&a, kAlign, chunk_) == 0 && ::posix_memalign(&b, kAlign, chunk_) == 0) {
            stage_[0] = a;
            stage_[1] = b;
            stage_free_[0] = stage_free_[1] = true;
            pipelined_ = true;
            io_thread_ = std::thread([this] { io_run(); });
        } else {
            std::free(a);  // buffered path below: copy and write in turn
        }
    }
    thread_ = std::thread([this] { run(); });
    return WF_OK;
}

int SnapshotWriter::submit(int slot, const void *dev, hipEvent_t packed)
{
    {
        std::unique_lock<std::mutex> lk(mu_);
        if (error_) return fail(error_, "%s", error_msg_.c_str());
    }
    // The wait is queued here, on the caller's thread and before the caller
    // can begin a stream capture, so the writer thread never touches the
    // caller's stream or events.  One copy stream per slot keeps a later
    // snapshot's wait from holding back an earlier snapshot's copies.
    if (hipStreamWaitEvent(stream_[slot], packed, 0) != hipSuccess)
        return fail(WF_ERR_HIP, "snapshot writer: hipStreamWaitEvent failed");
    std::unique_lock<std::mutex> lk(mu_);
    busy_[slot] = true;
    jobs_.push_back(Job{slot, dev, next_offset_});
    next_offset_ += snap_;
    cv_.notify_all();
    return WF_OK;
}

int SnapshotWriter::wait_free(int slot)
{
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return !busy_[slot] || error_; });
    if (error_) return fail(error_, "%s", error_msg_.c_str());
    return WF_OK;
}

int SnapshotWriter::write_all(const void *buf, uint64_t len, uint64_t offset)
{
    if (fault_injected("short_write")) {
        // forced from the host for the error-path tests: report a short write
        const ssize_t half = (ssize_t)(len / 2);
        return fail(WF_ERR_IO,
                    "short snapshot write: wrote %zd of %llu requested bytes at offset %llu "
                    "(fault injected)",
                    half, (unsigned long long)len, (unsigned long long)offset);
    }
    const char *p = static_cast<const char *>(buf);
    while (len > 0) {
        const uint64_t req = len < kMaxWrite ? len : kMaxWrite;
        const char *src = p;
        if (bounce_) {
            std::memcpy(bounce_, p, req);
            src = static_cast<const char *>(bounce_);
        }
        const ssize_t w = ::pwrite(fd_, src, req, (off_t)offset);
        if (w < 0) {
            const int err = errno;
            if (err == EINTR) continue;
            if (err == EFAULT && direct_ && !bounce_) {
                // O_DIRECT cannot pin driver-mapped (pinned HIP) memory:
                // stage each chunk through an ordinary page-aligned buffer
                void *b = nullptr;
                if (::posix_memalign(&b, kAlign, chunk_) == 0) {
                    bounce_ = b;
                    continue;
                }
            }
            if (err == EINVAL && direct_) {
                // the file system took O_DIRECT at open but refuses it here
                const int fl = ::fcntl(fd_, F_GETFL);
                if (fl != -1 && ::fcntl(fd_, F_SETFL, fl & ~O_DIRECT) == 0) {
                    direct_ = false;
                    continue;
                }
            }
            return fail(WF_ERR_IO, "snapshot write failed at offset %llu: %s",
                        (unsigned long long)offset, std::strerror(err));
        }
        if ((uint64_t)w != req)
            return fail(WF_ERR_IO,
                        "short snapshot write: wrote %zd of %llu requested bytes at offset %llu",
                        w, (unsigned long long)req, (unsigned long long)offset);
        p += req;
        len -= req;
        offset += req;
        written_ += req;
    }
    return WF_OK;
}

void SnapshotWriter::run()
{
    (void)hipSetDevice(device_);
    for (;;) {
        Job job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
            if (jobs_.empty()) return;
            job = jobs_.front();
        }
        int rc = WF_OK;
        std::string msg;
        hipStream_t st = stream_[job.slot];
        // copy chunk k+1 while chunk k is written
        const uint64_t nchunks = (snap_ + chunk_ - 1) / chunk_;
        auto issue = [&](uint64_t k) {
            const uint64_t off = k * chunk_;
            const uint64_t len = snap_ - off < chunk_ ? snap_ - off : chunk_;
            const int b = (int)(k % (uint64_t)nbuf_);
            if (hipMemcpyAsync(pinned_[b], static_cast<const char *>(job.dev) + off, len,
                               hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipEventRecord(copied_[b], st) != hipSuccess)
                return false;
            return true;
        };
        uint64_t issued = 0;
        for (; rc == WF_OK && issued < nchunks && issued < (uint64_t)nbuf_ - 1; ++issued)
            if (!issue(issued)) {
                rc = WF_ERR_HIP;
                msg = "snapshot writer: copy failed";
            }
        for (uint64_t k = 0; rc == WF_OK && k < nchunks; ++k) {
            const int b = (int)(k % (uint64_t)nbuf_);
            if (hipEventSynchronize(copied_[b]) != hipSuccess) {
                rc = WF_ERR_HIP;
                msg = "snapshot writer: copy failed";
                break;
            }
            if (issued < nchunks) {
                if (!issue(issued)) {
                    rc = WF_ERR_HIP;
                    msg = "snapshot writer: copy failed";
                    break;
                }
                ++issued;
            }
            const uint64_t off = k * chunk_;
            const uint64_t len = snap_ - off < chunk_ ? snap_ - off : chunk_;
            if (pipelined_) {
                // hand the chunk to the I/O thread through a free staging buffer
                int j = -1;
                {
                    std::unique_lock<std::mutex> lk(mu_);
                    cv_.wait(lk, [&] { return stage_free_[0] || stage_free_[1] || error_; });
                    if (error_) {
                        rc = error_;
                        msg = error_msg_;
                        break;
                    }
                    j = stage_free_[0] ? 0 : 1;
                    stage_free_[j] = false;
                }
                std::memcpy(stage_[j], pinned_[b], len);
                std::unique_lock<std::mutex> lk(mu_);
                writes_.push_back(WriteJob{j, len, job.offset + off});
                cv_.notify_all();
            } else {
                rc = write_all(pinned_[b], len, job.offset + off);
                if (rc) msg = wf_last_error();
            }
        }
        if (rc == WF_OK) (void)hipStreamSynchronize(st);
        {
            std::unique_lock<std::mutex> lk(mu_);
            jobs_.pop_front();
            busy_[job.slot] = false;
            if (rc && !error_) {
                error_ = rc;
                error_msg_ = msg;
            }
            cv_.notify_all();
        }
        if (rc) {
            // drop the remaining jobs: the file is already incomplete
            std::unique_lock<std::mutex> lk(mu_);
            while (!jobs_.empty()) {
                busy_[jobs_.front().slot] = false;
                jobs_.pop_front();
            }
            cv_.notify_all();
        }
    }
}

// Writes the staged chunks in the order they were handed over; after an
// error the rest are dropped (the file is already incomplete).
void SnapshotWriter::io_run()
{
    for (;;) {
        WriteJob wj;
        bool skip;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return io_stop_ || !writes_.empty(); });
            if (writes_.empty()) return;
            wj = writes_.front();
            writes_.pop_front();
            skip = error_ != 0;
        }
        int rc = WF_OK;
        std::string msg;
        if (!skip) {
            rc = write_all(stage_[wj.buf], wj.len, wj.offset);
            if (rc) msg = wf_last_error();
        }
        std::unique_lock<std::mutex> lk(mu_);
        stage_free_[wj.buf] = true;
        if (rc && !error_) {
            error_ = rc;
            error_msg_ = msg;
        }
        cv_.notify_all();
    }
}

int SnapshotWriter::finish()
{
    if (!opened_) return WF_OK;
    opened_ = false;
    {
        std::unique_lock<std::mutex> lk(mu_);
        stop_ = true;
        cv_.notify_all();
    }
    if (thread_.joinable()) thread_.join();
    {
        std::unique_lock<std::mutex> lk(mu_);
        io_stop_ = true;
        cv_.notify_all();
    }
    if (io_thread_.joinable()) io_thread_.join();
    for (auto &b : stage_) {
        std::free(b);
        b = nullptr;
    }
    pipelined_ = false;
    for (auto &st : stream_) {
        if (!st) continue;
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
        st = nullptr;
    }
    for (int i = 0; i < nbuf_; ++i) {
        if (pinned_[i]) (void)hipHostFree(pinned_[i]);
        if (copied_[i]) (void)hipEventDestroy(copied_[i]);
        pinned_[i] = nullptr;
        copied_[i] = nullptr;
    }
    nbuf_ = 0;
    std::free(bounce_);
    bounce_ = nullptr;
    int rc = error_;
    if (fd_ >= 0) {
        if (::close(fd_) != 0 && !rc) {
            rc = WF_ERR_IO;
            error_msg_ = std::string("closing snapshot file: ") + std::strerror(errno);
        }
        fd_ = -1;
    }
    if (rc) return fail(rc, "%s", error_msg_.c_str());
    return WF_OK;
}

}  // namespace wf
