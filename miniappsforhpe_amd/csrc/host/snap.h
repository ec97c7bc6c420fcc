// Snapshot store of the forward wavefield for the imaging condition: host-RAM LIFO (file_snap_orig), its factory,
// the token hand-off between the solver thread and a copy thread, and rtm_snap_gpu, which moves the device snapshot
// volumes to pinned host volumes and back on a stream of its own.
//
// Reference: snap_helpers/rtm_snap.h:14-44, snap_helpers/rtm_snap_gpu.h:16-87, snap_helpers/rtm_snap_gpu.cu:7-226,
// snap_helpers/helper_thread_synchronizer.h:21-45 + .cpp:4-37, snap_helpers/file_snap.h:17-38,
// snap_helpers/file_snap_orig.cpp:26-355 (memory strategy), snap_helpers/file_snap_factory.cpp:19-61,
// utils/sys_info_linux.h:31-106.
//
// Steps that do not fit the host-RAM budget spill to a temporary file in the scratch folder, written and read with
// POSIX AIO through one page-aligned staging buffer (O_DIRECT where the file system allows it), interleaved with the
// in-memory steps as in snap_helpers/file_snap_orig.cpp:289-313 and file_async.cpp:15-116. The reference compiles
// that path but guards it off (EMSL_VERIFY(false), file_snap_orig.cpp:267, :298) and aborts instead; here it runs,
// and a store that fits neither memory nor the scratch file system throws.
// Compressed snapshots (useZFP): each device volume is coded by the gfx950 fixed-rate ZFP codec into its pinned host
// buffer and the store keeps the coded words (snap_helpers/rtm_snap_gpu.cu:228-280); the zfp types are those of
// zfp_fixed_rate.h.
#pragma once

#include <aio.h>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "comm.h"
#include "grid.h"
#include "zfp_fixed_rate.h"

class helper_thread_synchronizer
{
public:
    void main_thread_wait();
    void main_thread_release();
    void helper_thread_wait();
    void helper_thread_release();

private:
    enum class holder
    {
        main_thread,
        helper_thread
    } token_holder = holder::main_thread;
    std::condition_variable cv;
    std::mutex mutex;
};

class sys_info_linux
{
public:
    double get_host_memory_mb() const;    // MemTotal of /proc/meminfo
    double get_process_memory_mb() const; // VmRSS of /proc/self/status
};

class file_snap
{
public:
    virtual ~file_snap() = default;
    // one step = the concatenation of the buffers; steps are read back last-in first-out
    virtual void write_step(std::vector<float*>& data_in, std::vector<int>& size_in) = 0;
    virtual void read_step(std::vector<float*>& data_in, std::vector<int>& size_in) = 0;
    virtual void write_step(float* data) = 0;
    virtual void read_step(float* data) = 0;
};

// Asynchronous file I/O with a small set of request slots (snap_helpers/file_async.h:16-39). Each slot holds at most
// one outstanding request; wait() blocks until it completes and checks its byte count.
class file_async
{
public:
    static constexpr int nbufs = 8;
    // flags as for open(2); when O_DIRECT is refused by the file system (EINVAL) the file is opened without it
    file_async(const char* name, int flags);
    ~file_async(); // waits for every outstanding request, then closes
    file_async(const file_async&) = delete;
    file_async& operator=(const file_async&) = delete;
    void seek_beg(off_t offset, int bufp = 0);
    void write(void* buf, size_t nbytes, int bufp = 0);
    void read(void* buf, size_t nbytes, int bufp = 0);
    void wait(int bufp = 0);
    void mark_temporary(); // unlink now: the data lives until the descriptor is closed
    bool direct() const { return direct_; }

private:
    std::string name;
    struct aiocb aioreq[nbufs];
    bool pending[nbufs] = {};
    int fd = -1;
    bool direct_ = false;
};

class file_snap_orig : public file_snap
{
public:
    // nval floats per step, nt steps, budget = mem_perc * MemTotal - resident memory of this process; the steps that
    // do not fit go to scratch/snap_scratch.<pid>.<n>. max_mem_steps >= 0 caps the in-memory steps (tests, or callers
    // that keep host RAM for other work).
    file_snap_orig(MPI_Comm comm, const char* scratch, long nval, float mem_perc, int nt, int max_mem_steps = -1);
    ~file_snap_orig() override;
    void write_step(std::vector<float*>& data_in, std::vector<int>& size_in) override;
    void read_step(std::vector<float*>& data_in, std::vector<int>& size_in) override;
    void write_step(float* data) override;
    void read_step(float* data) override;
    int steps_in_memory() const { return (int)memBuff.size(); }
    int steps_on_disk() const { return disk_steps_; }
    int time_index() const { return time_index_; }
    bool direct_io() const { return fd && fd->direct(); }
    // one character per step: '.' memory, '#' disk (print_strategy, snap_helpers/file_snap_orig.cpp:186-208)
    std::string strategy_string() const;

private:
    enum class Destination
    {
        DISK,
        MEMORY
    };
    struct StrategyType
    {
        Destination dest;
        size_t offset; // byte offset in the scratch file (DISK)
        float* buff;   // step buffer (MEMORY)
    };
    long nval;
    long nvalPad = 0; // nval rounded up to a whole number of pages (O_DIRECT transfer size)
    int nt;
    int time_index_ = 0;
    int disk_steps_ = 0;
    std::string fname;
    std::vector<StrategyType> strategy;
    std::vector<std::vector<float>> memBuff;
    struct free_deleter
    {
        void operator()(float* p) const { free(p); }
    };
    std::unique_ptr<float, free_deleter> tempBuff; // page-aligned staging buffer of nvalPad floats
    std::unique_ptr<file_async> fd;                // destroyed first: it waits for requests into tempBuff
};

struct file_snap_factory
{
    static file_snap* create(MPI_Comm comm, char* scratch, uint64_t nval, float mem_perc, int nt, int snap_type,
                             int max_mem_steps = -1);
};

class rtm_snap
{
public:
    virtual ~rtm_snap() = default;
    virtual void start_read_uncompressed_fwi_snapshot(std::vector<cart_volume<float>*>& corrBuffList,
                                                      std::vector<int>& corr_buff_size, file_snap* fd_snap_p) = 0;
    virtual void finish_read_uncompressed_fwi_snapshot() = 0;
    virtual void start_write_uncompressed_fwi_snapshot(std::vector<cart_volume<float>*>& corrBuffList,
                                                       std::vector<int>& corr_buff_size, file_snap* fd_snap_p) = 0;
    virtual void lock_for_write_uncompressed_fwi_snapshot() = 0;
    virtual void start_read_compressed_fwi_snapshot(std::vector<cart_volume<float>*>& corrBuffList,
                                                    std::vector<int>& corr_buff_size, file_snap* fd_snap_p,
                                                    std::vector<float*>& zipped_corr_buff,
                                                    std::vector<zfp_field*>& zfpFields, zfp_stream* zfpStream) = 0;
    virtual void finish_read_compressed_fwi_snapshot() = 0;
    virtual void start_write_compressed_fwi_snapshot(std::vector<cart_volume<float>*>& corrBuffList,
                                                     std::vector<int>& corr_buff_size, file_snap* fd_snap_p,
                                                     std::vector<float*>& zipped_corr_buff,
                                                     std::vector<zfp_field*>& zfpFields, zfp_stream* zfpStream) = 0;
    virtual void lock_for_write_compressed_fwi_snapshot() = 0;
};

// The compressed-snapshot state a driver owns (simulator2/simulator2SetUp.cu:285-323): one zfp stream in fixed-rate
// mode at bitPerFloat bits per value, one 3-D field per snapshot volume over its valid region (nvalid extents, ntot
// strides, from the volume's first element), one pinned host buffer per volume of zfp_stream_maximum_size bytes.
struct snapshot_zfp_state
{
    zfp_stream* stream = nullptr;
    std::vector<zfp_field*> fields;
    std::vector<float*> zipped;
    // sizes[i] receives volume i's buffer size in floats
    void init(float bitPerFloat, const std::vector<cart_volume<float>*>& vols, std::vector<int>& sizes);
    snapshot_zfp_state() = default;
    snapshot_zfp_state(const snapshot_zfp_state&) = delete;
    snapshot_zfp_state& operator=(const snapshot_zfp_state&) = delete;
    ~snapshot_zfp_state();
};

class rtm_snap_gpu : public rtm_snap
{
public:
    // corrBuffList: the device volumes of one snapshot step. write waits for writeSnapshotsCompleteEvent (recorded by
    // the solver after it filled them), read waits for readSnapshotsCompleteEvent (recorded after it consumed them).
    // With useZFP only the compressed methods may be called (and without it only the uncompressed ones).
    rtm_snap_gpu(std::vector<cart_volume<float>*>& corrBuffList, hipEvent_t writeSnapshotsCompleteEvent,
                 hipEvent_t readSnapshotsCompleteEvent, bool useZFP);
    ~rtm_snap_gpu() override;

    void start_read_uncompressed_fwi_snapshot(std::vector<cart_volume<float>*>& gpu_corrBuffList,
                                              std::vector<int>& corr_buff_size, file_snap* fd_snap_p) override;
    void finish_read_uncompressed_fwi_snapshot() override;
    void start_write_uncompressed_fwi_snapshot(std::vector<cart_volume<float>*>& gpu_corrBuffList,
                                               std::vector<int>& corr_buff_size, file_snap* fd_snap_p) override;
    void lock_for_write_uncompressed_fwi_snapshot() override;
    // corr_buff_size[i]: floats of zipped_corr_buff[i] (at least the coded size of volume i's valid region)
    void start_read_compressed_fwi_snapshot(std::vector<cart_volume<float>*>& gpu_corrBuffList,
                                            std::vector<int>& corr_buff_size, file_snap* fd_snap_p,
                                            std::vector<float*>& zipped_corr_buff, std::vector<zfp_field*>& zfpFields,
                                            zfp_stream* zfpStream) override;
    void finish_read_compressed_fwi_snapshot() override;
    void start_write_compressed_fwi_snapshot(std::vector<cart_volume<float>*>& gpu_corrBuffList,
                                             std::vector<int>& corr_buff_size, file_snap* fd_snap_p,
                                             std::vector<float*>& zipped_corr_buff, std::vector<zfp_field*>& zfpFields,
                                             zfp_stream* zfpStream) override;
    void lock_for_write_compressed_fwi_snapshot() override;

private:
    enum class snapshot_command
    {
        write,
        read,
        close
    };
    void helper_thread_loop();
    void helper_thread_write();
    void helper_thread_read();
    void helper_thread_write_compressed();
    void helper_thread_read_compressed();
    void check_list(const std::vector<cart_volume<float>*>& l, const std::vector<int>& sizes) const;
    void check_zipped(const std::vector<cart_volume<float>*>& l, const std::vector<int>& sizes,
                      const std::vector<float*>& zipped, const std::vector<zfp_field*>& fields, zfp_stream* z) const;
    void wait_token(); // main_thread_wait, then rethrow a helper-thread failure

    hipStream_t helperStream = nullptr;
    hipEvent_t writeSnapshotsCompleteEvent_;
    hipEvent_t readSnapshotsCompleteEvent_;
    helper_thread_synchronizer sync;
    // shared with the helper thread; written by the main thread only between main_thread_wait and release
    snapshot_command command = snapshot_command::close;
    file_snap* fd_snap_p = nullptr;
    std::vector<int> corrBuffSize;
    std::vector<cart_volume<float>*> gpu_corrBuffList;
    std::vector<cart_volume_regular_host*> cpu_corrBuffList; // empty with useZFP
    size_t nvol = 0;
    bool useZFP_ = false;
    std::vector<float*> zipped_corr_buff;
    std::vector<zfp_field*> zfpFields;
    zfp_stream* zfpStream = nullptr;
    std::string helper_error; // first exception raised on the helper thread, rethrown on the main thread
    std::thread helper_thread;
};
