#include "snap.h"

#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

// ---- helper_thread_synchronizer: a token that one of the two threads holds (starts with the main thread) ----------

void
helper_thread_synchronizer::main_thread_wait()
{
    std::unique_lock<std::mutex> lock(mutex);
    cv.wait(lock, [this] { return token_holder == holder::main_thread; });
}

void
helper_thread_synchronizer::main_thread_release()
{
    std::lock_guard<std::mutex> lock(mutex);
    if (token_holder != holder::main_thread)
        throw std::logic_error("main thread released a token it does not hold");
    token_holder = holder::helper_thread;
    cv.notify_one();
}

void
helper_thread_synchronizer::helper_thread_wait()
{
    std::unique_lock<std::mutex> lock(mutex);
    cv.wait(lock, [this] { return token_holder == holder::helper_thread; });
}

void
helper_thread_synchronizer::helper_thread_release()
{
    std::lock_guard<std::mutex> lock(mutex);
    if (token_holder != holder::helper_thread)
        throw std::logic_error("helper thread released a token it does not hold");
    token_holder = holder::main_thread;
    cv.notify_one();
}

// ---- sys_info_linux ------------------------------------------------------------------------------------------------

namespace {

double
kb_field(const char* path, const char* key)
{
    std::ifstream f(path);
    std::string line;
    const size_t klen = std::strlen(key);
    while (std::getline(f, line)) {
        if (line.compare(0, klen, key) == 0) {
            std::istringstream is(line.substr(klen));
            double kb = 0;
            is >> kb;
            return kb;
        }
    }
    return 0.0;
}

} // namespace

double
sys_info_linux::get_host_memory_mb() const
{
    return kb_field("/proc/meminfo", "MemTotal:") / 1024.0;
}

double
sys_info_linux::get_process_memory_mb() const
{
    return kb_field("/proc/self/status", "VmRSS:") / 1024.0;
}

// ---- file_snap_orig (memory strategy) -------------------------------------------------------------------------------

file_snap_orig::file_snap_orig(MPI_Comm, const char*, long nval_, float mem_perc, int nt_) : nval(nval_), nt(nt_)
{
    if (nval < 0 || nt < 0)
        throw std::runtime_error("file_snap_orig: negative size");
    // snap_helpers/file_snap_orig.cpp:219-287: snapshots kept in host RAM up to mem_perc of the host total minus what
    // this process already holds; the rest would go to local scratch, which this build does not provide
    sys_info_linux info;
    const double avail_mb = mem_perc * info.get_host_memory_mb() - info.get_process_memory_mb();
    const double snap_mb = (double)nval * sizeof(float) / (1024.0 * 1024.0);
    int mem_snaps = nt;
    if (snap_mb > 0.0)
        mem_snaps = (int)std::max(0.0, std::min((double)nt, avail_mb / snap_mb));
    if (mem_snaps < nt && snap_mb > 0.0) {
        std::ostringstream m;
        m << "file_snap_orig: " << nt << " snapshots of " << snap_mb << " MB need " << nt * snap_mb
          << " MB of host memory but " << avail_mb << " MB are available";
        throw std::runtime_error(m.str());
    }
    if (nval > 0) {
        memBuff.resize(nt);
        for (auto& b : memBuff)
            b.resize(nval);
    }
}

void
file_snap_orig::write_step(std::vector<float*>& data_in, std::vector<int>& size_in)
{
    if (time_index_ >= nt)
        throw std::runtime_error("file_snap_orig: more steps written than reserved");
    if (data_in.size() != size_in.size())
        throw std::runtime_error("file_snap_orig: buffer and size lists differ");
    long at = 0;
    for (size_t i = 0; i < data_in.size(); ++i) {
        if (at + size_in[i] > nval)
            throw std::runtime_error("file_snap_orig: step larger than reserved");
        std::memcpy(memBuff[time_index_].data() + at, data_in[i], (size_t)size_in[i] * sizeof(float));
        at += size_in[i];
    }
    ++time_index_;
}

void
file_snap_orig::read_step(std::vector<float*>& data_in, std::vector<int>& size_in)
{
    if (time_index_ <= 0)
        throw std::runtime_error("file_snap_orig: more steps read than written");
    if (data_in.size() != size_in.size())
        throw std::runtime_error("file_snap_orig: buffer and size lists differ");
    --time_index_;
    long at = 0;
    for (size_t i = 0; i < data_in.size(); ++i) {
        if (at + size_in[i] > nval)
            throw std::runtime_error("file_snap_orig: step larger than reserved");
        std::memcpy(data_in[i], memBuff[time_index_].data() + at, (size_t)size_in[i] * sizeof(float));
        at += size_in[i];
    }
}

void
file_snap_orig::write_step(float* data)
{
    std::vector<float*> d{ data };
    std::vector<int> n{ (int)nval };
    write_step(d, n);
}

void
file_snap_orig::read_step(float* data)
{
    std::vector<float*> d{ data };
    std::vector<int> n{ (int)nval };
    read_step(d, n);
}

file_snap*
file_snap_factory::create(MPI_Comm comm, char* scratch, uint64_t nval, float mem_perc, int nt, int snap_type)
{
    if (snap_type != 0)
        throw std::runtime_error("file_snap_factory: only the original (uncompressed) snapshot type is supported");
    return new file_snap_orig(comm, scratch, (long)nval, mem_perc, nt);
}

// ---- rtm_snap_gpu ----------------------------------------------------------------------------------------------------

rtm_snap_gpu::rtm_snap_gpu(std::vector<cart_volume<float>*>& corrBuffList, hipEvent_t writeEvent,
                           hipEvent_t readEvent, bool useZFP)
    : writeSnapshotsCompleteEvent_(writeEvent), readSnapshotsCompleteEvent_(readEvent)
{
    if (useZFP)
        throw std::runtime_error("rtm_snap_gpu: ZFP compression is not supported");
    // pinned host twins without x padding (snap_helpers/rtm_snap_gpu.cu:13-21)
    for (auto* v : corrBuffList) {
        auto* r = v->as<cart_volume_regular>();
        axis ax1(r->ax1()->o, r->ax1()->d, r->ax1()->n, r->ax1()->nghost);
        cpu_corrBuffList.push_back(new cart_volume_regular_host(&ax1, r->ax2(), r->ax3(), true));
    }
    tti::hip_check(hipStreamCreateWithFlags(&helperStream, hipStreamNonBlocking), "rtm_snap_gpu: stream");
    helper_thread = std::thread(&rtm_snap_gpu::helper_thread_loop, this);
}

rtm_snap_gpu::~rtm_snap_gpu()
{
    sync.main_thread_wait();
    command = snapshot_command::close;
    sync.main_thread_release();
    helper_thread.join();
    for (auto* v : cpu_corrBuffList)
        delete v;
    (void)hipStreamDestroy(helperStream);
}

void
rtm_snap_gpu::check_list(const std::vector<cart_volume<float>*>& l, const std::vector<int>& sizes) const
{
    if (l.size() != cpu_corrBuffList.size() || sizes.size() != l.size())
        throw std::runtime_error("rtm_snap_gpu: snapshot list does not match the one given at construction");
    for (size_t i = 0; i < l.size(); ++i)
        if ((size_t)sizes[i] != cpu_corrBuffList[i]->nelems())
            throw std::runtime_error("rtm_snap_gpu: buffer size differs from the valid region of its volume");
}

void
rtm_snap_gpu::start_read_uncompressed_fwi_snapshot(std::vector<cart_volume<float>*>& l, std::vector<int>& sizes,
                                                   file_snap* fd)
{
    check_list(l, sizes);
    sync.main_thread_wait();
    if (!helper_error.empty())
        throw std::runtime_error(helper_error);
    gpu_corrBuffList = l;
    corrBuffSize = sizes;
    command = snapshot_command::read;
    fd_snap_p = fd;
    sync.main_thread_release();
}

void
rtm_snap_gpu::finish_read_uncompressed_fwi_snapshot()
{
    sync.main_thread_wait();
    if (!helper_error.empty())
        throw std::runtime_error(helper_error);
}

void
rtm_snap_gpu::start_write_uncompressed_fwi_snapshot(std::vector<cart_volume<float>*>& l, std::vector<int>& sizes,
                                                    file_snap* fd)
{
    // the caller holds the token since lock_for_write_uncompressed_fwi_snapshot
    check_list(l, sizes);
    sync.main_thread_wait();
    gpu_corrBuffList = l;
    corrBuffSize = sizes;
    command = snapshot_command::write;
    fd_snap_p = fd;
    sync.main_thread_release();
}

void
rtm_snap_gpu::lock_for_write_uncompressed_fwi_snapshot()
{
    sync.main_thread_wait();
    if (!helper_error.empty())
        throw std::runtime_error(helper_error);
}

void
rtm_snap_gpu::helper_thread_loop()
{
    while (true) {
        sync.helper_thread_wait();
        const snapshot_command c = command;
        if (c == snapshot_command::close) {
            sync.helper_thread_release();
            return;
        }
        if (helper_error.empty()) {
            try {
                if (c == snapshot_command::write)
                    helper_thread_write();
                else
                    helper_thread_read();
            } catch (const std::exception& e) {
                helper_error = e.what();
            }
        }
        sync.helper_thread_release();
    }
}

void
rtm_snap_gpu::helper_thread_write()
{
    // snap_helpers/rtm_snap_gpu.cu:186-205: device volumes (valid region) -> pinned host -> store
    tti::hip_check(hipStreamWaitEvent(helperStream, writeSnapshotsCompleteEvent_, 0), "rtm_snap_gpu: wait");
    for (size_t i = 0; i < cpu_corrBuffList.size(); ++i)
        gpu_corrBuffList[i]->as<cart_volume_regular_gpu>()->copyToAsync(cpu_corrBuffList[i], helperStream, false);
    tti::hip_check(hipStreamSynchronize(helperStream), "rtm_snap_gpu: D2H");
    std::vector<float*> data;
    for (auto* h : cpu_corrBuffList)
        data.push_back(h->getData());
    fd_snap_p->write_step(data, corrBuffSize);
}

void
rtm_snap_gpu::helper_thread_read()
{
    // snap_helpers/rtm_snap_gpu.cu:207-226: store -> pinned host -> device volumes, after the previous step's
    // consumers are done (readSnapshotsCompleteEvent)
    tti::hip_check(hipStreamWaitEvent(helperStream, readSnapshotsCompleteEvent_, 0), "rtm_snap_gpu: wait");
    std::vector<float*> data;
    for (auto* h : cpu_corrBuffList)
        data.push_back(h->getData());
    fd_snap_p->read_step(data, corrBuffSize);
    for (size_t i = 0; i < cpu_corrBuffList.size(); ++i)
        gpu_corrBuffList[i]->as<cart_volume_regular_gpu>()->copyFromAsync(cpu_corrBuffList[i], helperStream, false);
    tti::hip_check(hipStreamSynchronize(helperStream), "rtm_snap_gpu: H2D");
}
