// Snapshot egress: device snapshot buffers -> pinned host chunks -> file.
//
// The reference streams every snapshot from two page-aligned buffers through
// io_uring O_DIRECT writes in chunks of at most 1 GiB, and turns a short
// write into an error (src/engine.rs:424-640, 678-807).  Here a writer thread
// owns the file and a copy stream per device buffer: each submitted snapshot's
// stream waits for the event that marks the buffer packed; the thread copies it out in
// chunks through a ring of pinned host buffers (copy of chunk i+1 overlaps
// the write of chunk i), writes each chunk at its file offset, and finally
// releases the device buffer.  The timestep loop only blocks when it wants
// to pack into a device buffer whose previous snapshot is still being
// written.
//
// O_DIRECT cannot write from pinned HIP memory on this platform (the kernel
// cannot pin driver-mapped pages: EFAULT), so with O_DIRECT each chunk is
// copied into one of two page-aligned host buffers and written by a second
// thread: the copy of chunk i+1 overlaps the write of chunk i, and the file
// is written at the disk's speed instead of at copy + write.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <string>
#include <thread>

namespace wf {

class SnapshotWriter {
public:
    // Largest single write request (the reference's MAX_DIRECT_IO_WRITE_BYTES).
    static constexpr uint64_t kMaxWrite = 1ull << 30;
    static constexpr uint64_t kAlign = 4096;

    SnapshotWriter() = default;
    SnapshotWriter(const SnapshotWriter &) = delete;
    SnapshotWriter &operator=(const SnapshotWriter &) = delete;
    ~SnapshotWriter();

    // Opens (truncates) the file, allocates nbuf pinned chunks of chunk_bytes
    // and starts the writer thread on `device`.  snapshot_bytes must be a
    // multiple of 4096 (it is: cells * 24 with cells a multiple of 512).
    int open(const char *path, uint64_t snapshot_bytes, int device, uint64_t chunk_bytes = 256ull << 20,
             int nbuf = 3);
    // Queues device buffer `slot` (ready once `packed` completes) as the next
    // snapshot in the file.
    int submit(int slot, const void *dev, hipEvent_t packed);
    // Blocks until device buffer `slot` is no longer being read.
    int wait_free(int slot);
    // Drains every queued snapshot, stops the thread and closes the file;
    // returns the first error (WF_ERR_IO for short or failed writes).  Only
    // the first call after open() does anything.
    int finish();
    bool direct_io() const { return direct_; }
    uint64_t bytes_written() const { return written_; }

private:
    struct Job {
        int slot;
        const void *dev;
        uint64_t offset;
    };
    void run();
    void io_run();
    int write_all(const void *buf, uint64_t len, uint64_t offset);

    bool opened_ = false;
    int fd_ = -1;
    bool direct_ = false;
    int device_ = 0;
    uint64_t snap_ = 0, chunk_ = 0, next_offset_ = 0, written_ = 0;
    hipStream_t stream_[2] = {nullptr, nullptr};  // copy stream per device slot
    static constexpr int kMaxBuf = 4;
    int nbuf_ = 0;
    void *pinned_[kMaxBuf] = {};
    void *bounce_ = nullptr;  // page-aligned staging when O_DIRECT refuses pinned memory
    // the O_DIRECT path: two page-aligned buffers handed to the I/O thread
    struct WriteJob {
        int buf;
        uint64_t len, offset;
    };
    bool pipelined_ = false;
    void *stage_[2] = {nullptr, nullptr};
    bool stage_free_[2] = {true, true};
    std::deque<WriteJob> writes_;
    bool io_stop_ = false;
    std::thread io_thread_;
    hipEvent_t copied_[kMaxBuf] = {};
    std::thread thread_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Job> jobs_;
    bool busy_[2] = {false, false};
    bool stop_ = false;
    int error_ = 0;
    std::string error_msg_;
};

}  // namespace wf
