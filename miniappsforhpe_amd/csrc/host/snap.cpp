#include "snap.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <sys/statvfs.h>
#include <unistd.h>

#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdlib>
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

// ---- file_async ------------------------------------------------------------------------------------------------------

file_async::file_async(const char* name_, int flags) : name(name_)
{
    std::memset(aioreq, 0, sizeof(aioreq));
    direct_ = (flags & O_DIRECT) != 0;
    fd = ::open(name_, flags, S_IRUSR | S_IWUSR);
    if (fd < 0 && errno == EINVAL && direct_) {
        // tmpfs, overlay and some network file systems refuse O_DIRECT: keep going through the page cache
        direct_ = false;
        fd = ::open(name_, flags & ~O_DIRECT, S_IRUSR | S_IWUSR);
    }
    if (fd < 0)
        throw std::runtime_error("file_async: open " + name + " failed: " + std::strerror(errno));
}

file_async::~file_async()
{
    for (int b = 0; b < nbufs; ++b) {
        if (!pending[b])
            continue;
        const struct aiocb* list[1] = { &aioreq[b] };
        while (aio_error(&aioreq[b]) == EINPROGRESS)
            aio_suspend(list, 1, nullptr);
        aio_return(&aioreq[b]);
    }
    if (fd >= 0)
        ::close(fd);
}

void
file_async::seek_beg(off_t offset, int bufp)
{
    if (bufp < 0 || bufp >= nbufs)
        throw std::runtime_error("file_async: bad request slot");
    aioreq[bufp].aio_fildes = fd;
    aioreq[bufp].aio_offset = offset;
}

void
file_async::write(void* buf, size_t nbytes, int bufp)
{
    if (bufp < 0 || bufp >= nbufs || pending[bufp])
        throw std::runtime_error("file_async: request slot busy");
    aioreq[bufp].aio_fildes = fd;
    aioreq[bufp].aio_buf = buf;
    aioreq[bufp].aio_nbytes = nbytes;
    if (aio_write(&aioreq[bufp]) != 0)
        throw std::runtime_error("file_async: aio_write of " + std::to_string(nbytes) + " bytes to " + name +
                                 " failed: " + std::strerror(errno));
    pending[bufp] = true;
}

void
file_async::read(void* buf, size_t nbytes, int bufp)
{
    if (bufp < 0 || bufp >= nbufs || pending[bufp])
        throw std::runtime_error("file_async: request slot busy");
    aioreq[bufp].aio_fildes = fd;
    aioreq[bufp].aio_buf = buf;
    aioreq[bufp].aio_nbytes = nbytes;
    if (aio_read(&aioreq[bufp]) != 0)
        throw std::runtime_error("file_async: aio_read from " + name + " failed: " + std::strerror(errno));
    pending[bufp] = true;
}

void
file_async::wait(int bufp)
{
    // snap_helpers/file_async.cpp:83-109 polls every millisecond for up to 1000 s; aio_suspend blocks instead
    if (bufp < 0 || bufp >= nbufs || !pending[bufp])
        return;
    const struct aiocb* list[1] = { &aioreq[bufp] };
    const struct timespec limit = { 1000, 0 };
    int err;
    while ((err = aio_error(&aioreq[bufp])) == EINPROGRESS) {
        if (aio_suspend(list, 1, &limit) != 0 && errno == EAGAIN)
            throw std::runtime_error("file_async: a request on " + name + " exceeds 1000 seconds");
    }
    const ssize_t got = aio_return(&aioreq[bufp]);
    pending[bufp] = false;
    if (err != 0)
        throw std::runtime_error("file_async: I/O on " + name + " failed: " + std::strerror(err));
    if ((size_t)got != aioreq[bufp].aio_nbytes)
        throw std::runtime_error("file_async: I/O on " + name + " moved " + std::to_string(got) + " of " +
                                 std::to_string(aioreq[bufp].aio_nbytes) + " bytes");
}

void
file_async::mark_temporary()
{
    ::unlink(name.c_str());
}

// ---- file_snap_orig ------------------------------------------------------------------------------------------------

namespace {
std::atomic<int> g_snap_files{ 0 };
}

file_snap_orig::file_snap_orig(MPI_Comm, const char* scratch, long nval_, float mem_perc, int nt_, int max_mem_steps)
    : nval(nval_), nt(nt_)
{
    if (nval < 0 || nt < 0)
        throw std::runtime_error("file_snap_orig: negative size");
    strategy.resize(nt);
    // snap_helpers/file_snap_orig.cpp:219-287: snapshots are kept in host RAM up to mem_perc of the host total minus
    // what this process already holds; the rest go to local scratch
    sys_info_linux info;
    const double avail_mb = mem_perc * info.get_host_memory_mb() - info.get_process_memory_mb();
    const double snap_mb = (double)nval * sizeof(float) / (1024.0 * 1024.0);
    int mem_snaps = nt;
    if (snap_mb > 0.0)
        mem_snaps = (int)std::max(0.0, std::min((double)nt, avail_mb / snap_mb));
    if (max_mem_steps >= 0)
        mem_snaps = std::min(mem_snaps, max_mem_steps);
    if (nval == 0)
        return;
    const long page = sysconf(_SC_PAGESIZE);
    nvalPad = (nval + page - 1) / page * page;
    // Spread the disk steps evenly among the memory steps (snap_helpers/file_snap_orig.cpp:299-313): step idx goes to
    // memory when floor(ratio * idx) advances, which happens mem_snaps times for ratio = mem_snaps / nt. One rank per
    // host here, so the reference's per-rank rotation of the pattern is zero. Buffers are allocated once the scratch
    // space is known to suffice.
    const double ratio = nt > 0 ? (double)mem_snaps / (double)nt : 0.0;
    int mem_counter = 0;
    for (int idx = 0; idx < nt; ++idx) {
        if (std::floor(ratio * (idx - 1)) == std::floor(ratio * idx) || mem_counter == mem_snaps) {
            strategy[idx] = { Destination::DISK, (size_t)disk_steps_ * (size_t)nvalPad * sizeof(float), nullptr };
            ++disk_steps_;
        } else {
            strategy[idx] = { Destination::MEMORY, 0, nullptr };
            ++mem_counter;
        }
    }
    auto allocate_memory_steps = [&] {
        memBuff.resize(mem_counter);
        int m = 0;
        for (auto& st : strategy)
            if (st.dest == Destination::MEMORY) {
                memBuff[m].resize(nval);
                st.buff = memBuff[m++].data();
            }
    };
    if (disk_steps_ == 0) {
        allocate_memory_steps();
        return;
    }
    std::string dir = scratch && *scratch ? scratch : ".";
    std::ostringstream fn;
    fn << dir << "/snap_scratch." << getpid() << "." << g_snap_files++;
    fname = fn.str();
    const double need_mb = (double)disk_steps_ * nvalPad * sizeof(float) / (1024.0 * 1024.0);
    struct statvfs vfs;
    const double free_mb =
        statvfs(dir.c_str(), &vfs) == 0 ? (double)vfs.f_bavail * vfs.f_frsize / (1024.0 * 1024.0) : 0.0;
    if (need_mb > free_mb) {
        std::ostringstream m;
        m << "file_snap_orig: " << nt << " snapshots of " << snap_mb << " MB: " << mem_counter
          << " fit in host memory (" << avail_mb << " MB available); the other " << disk_steps_ << " need " << need_mb
          << " MB of local scratch in " << dir << " but " << free_mb << " MB are available";
        throw std::runtime_error(m.str());
    }
    void* staging = nullptr;
    if (posix_memalign(&staging, (size_t)page, (size_t)nvalPad * sizeof(float)) != 0)
        throw std::runtime_error("file_snap_orig: cannot allocate the staging buffer");
    tempBuff.reset(static_cast<float*>(staging));
    std::memset(staging, 0, (size_t)nvalPad * sizeof(float));
    ::remove(fname.c_str());
    fd.reset(new file_async(fname.c_str(), O_RDWR | O_CREAT | O_DIRECT));
    fd->mark_temporary();
    allocate_memory_steps();
}

file_snap_orig::~file_snap_orig()
{
    fd.reset(); // waits for outstanding requests before the staging buffer goes
}

std::string
file_snap_orig::strategy_string() const
{
    std::string s;
    for (const auto& st : strategy)
        s.push_back(st.dest == Destination::MEMORY ? '.' : '#');
    return s;
}

void
file_snap_orig::write_step(std::vector<float*>& data_in, std::vector<int>& size_in)
{
    if (time_index_ >= nt)
        throw std::runtime_error("file_snap_orig: more steps written than reserved");
    if (data_in.size() != size_in.size())
        throw std::runtime_error("file_snap_orig: buffer and size lists differ");
    long total = 0;
    for (int n : size_in)
        total += n;
    if (total > nval)
        throw std::runtime_error("file_snap_orig: step larger than reserved");
    if (nval == 0) {
        ++time_index_;
        return;
    }
    const StrategyType& st = strategy[time_index_];
    float* dst = st.buff;
    if (st.dest == Destination::DISK) {
        // the staging buffer is reused: wait for the previous disk write (file_snap_orig.cpp:63-75)
        fd->wait();
        dst = tempBuff.get();
    }
    long at = 0;
    for (size_t i = 0; i < data_in.size(); ++i) {
        std::memcpy(dst + at, data_in[i], (size_t)size_in[i] * sizeof(float));
        at += size_in[i];
    }
    if (st.dest == Destination::DISK) {
        fd->seek_beg((off_t)st.offset);
        fd->write(tempBuff.get(), (size_t)nvalPad * sizeof(float));
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
    long total = 0;
    for (int n : size_in)
        total += n;
    if (total > nval)
        throw std::runtime_error("file_snap_orig: step larger than reserved");
    --time_index_;
    if (nval == 0)
        return;
    const StrategyType& st = strategy[time_index_];
    const float* src = st.buff;
    if (st.dest == Destination::DISK) {
        // The staging buffer holds this step: the first disk step read back is the last one written (still in the
        // buffer once its write completes), every later one was prefetched by the previous disk read
        // (file_snap_orig.cpp:88-110).
        fd->wait();
        src = tempBuff.get();
    }
    long at = 0;
    for (size_t i = 0; i < data_in.size(); ++i) {
        std::memcpy(data_in[i], src + at, (size_t)size_in[i] * sizeof(float));
        at += size_in[i];
    }
    if (st.dest == Destination::DISK) {
        for (int i = time_index_ - 1; i >= 0; --i) {
            if (strategy[i].dest == Destination::DISK) {
                fd->seek_beg((off_t)strategy[i].offset);
                fd->read(tempBuff.get(), (size_t)nvalPad * sizeof(float));
                break;
            }
        }
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
file_snap_factory::create(MPI_Comm comm, char* scratch, uint64_t nval, float mem_perc, int nt, int snap_type,
                          int max_mem_steps)
{
    if (snap_type != 0)
        throw std::runtime_error("file_snap_factory: only the original (uncompressed) snapshot type is supported");
    return new file_snap_orig(comm, scratch, (long)nval, mem_perc, nt, max_mem_steps);
}

// ---- snapshot_zfp_state --------------------------------------------------------------------------------------------

void
snapshot_zfp_state::init(float bitPerFloat, const std::vector<cart_volume<float>*>& vols, std::vector<int>& sizes)
{
    stream = zfp_stream_open(nullptr);
    if (zfp_stream_set_rate(stream, bitPerFloat, zfp_type_float, 3, 0) <= 0.0)
        throw std::runtime_error("snapshot_zfp_state: invalid ZFP rate");
    zfp_stream_set_execution(stream, zfp_exec_cuda);
    sizes.clear();
    for (auto* v : vols) {
        auto* r = v->as<cart_volume_regular>();
        zfp_field* f = zfp_field_3d(nullptr, zfp_type_float, (unsigned)r->ax1()->nvalid, (unsigned)r->ax2()->nvalid,
                                    (unsigned)r->ax3()->nvalid);
        zfp_field_set_stride_3d(f, 1, r->ax1()->ntot, (ptrdiff_t)r->ax1()->ntot * r->ax2()->ntot);
        fields.push_back(f);
        const size_t nbyte = zfp_stream_maximum_size(stream, f);
        const size_t nfloat = (nbyte + sizeof(float) - 1) / sizeof(float);
        if (nbyte == 0 || nfloat > (size_t)INT32_MAX)
            throw std::runtime_error("snapshot_zfp_state: volume too large for a compressed snapshot buffer");
        sizes.push_back((int)nfloat);
        float* buf = nullptr;
        tti::hip_check(hipHostMalloc((void**)&buf, nfloat * sizeof(float), hipHostMallocDefault),
                       "snapshot_zfp_state: pinned buffer");
        zipped.push_back(buf);
    }
}

snapshot_zfp_state::~snapshot_zfp_state()
{
    for (auto* b : zipped)
        (void)hipHostFree(b);
    for (auto* f : fields)
        zfp_field_free(f);
    zfp_stream_close(stream);
}

// ---- rtm_snap_gpu ----------------------------------------------------------------------------------------------------

rtm_snap_gpu::rtm_snap_gpu(std::vector<cart_volume<float>*>& corrBuffList, hipEvent_t writeEvent,
                           hipEvent_t readEvent, bool useZFP)
    : writeSnapshotsCompleteEvent_(writeEvent), readSnapshotsCompleteEvent_(readEvent), nvol(corrBuffList.size()),
      useZFP_(useZFP)
{
    // pinned host twins without x padding (snap_helpers/rtm_snap_gpu.cu:13-21); the compressed path codes straight
    // from the device volumes into the caller's pinned buffers and needs none
    if (!useZFP)
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
rtm_snap_gpu::wait_token()
{
    sync.main_thread_wait();
    if (!helper_error.empty())
        throw std::runtime_error(helper_error);
}

void
rtm_snap_gpu::check_list(const std::vector<cart_volume<float>*>& l, const std::vector<int>& sizes) const
{
    if (useZFP_)
        throw std::runtime_error("rtm_snap_gpu: built for compressed snapshots, the uncompressed methods are not "
                                 "available");
    if (l.size() != cpu_corrBuffList.size() || sizes.size() != l.size())
        throw std::runtime_error("rtm_snap_gpu: snapshot list does not match the one given at construction");
    for (size_t i = 0; i < l.size(); ++i)
        if ((size_t)sizes[i] != cpu_corrBuffList[i]->nelems())
            throw std::runtime_error("rtm_snap_gpu: buffer size differs from the valid region of its volume");
}

void
rtm_snap_gpu::check_zipped(const std::vector<cart_volume<float>*>& l, const std::vector<int>& sizes,
                           const std::vector<float*>& zipped, const std::vector<zfp_field*>& fields,
                           zfp_stream* z) const
{
    if (!useZFP_)
        throw std::runtime_error("rtm_snap_gpu: built for uncompressed snapshots, the compressed methods are not "
                                 "available");
    if (l.size() != nvol || sizes.size() != nvol || zipped.size() != nvol || fields.size() != nvol || !z)
        throw std::runtime_error("rtm_snap_gpu: compressed snapshot lists do not match the volumes");
    for (size_t i = 0; i < nvol; ++i) {
        const zfp_field* f = fields[i];
        if (!f || !zipped[i])
            throw std::runtime_error("rtm_snap_gpu: missing zfp field or buffer");
        const size_t need = tti_zfp_stream_bytes(z->rate, (int)f->nx, (int)f->ny, (int)f->nz);
        if (need == 0 || (size_t)sizes[i] * sizeof(float) < need)
            throw std::runtime_error("rtm_snap_gpu: compressed buffer smaller than the coded volume");
    }
}

void
rtm_snap_gpu::start_read_uncompressed_fwi_snapshot(std::vector<cart_volume<float>*>& l, std::vector<int>& sizes,
                                                   file_snap* fd)
{
    check_list(l, sizes);
    wait_token();
    gpu_corrBuffList = l;
    corrBuffSize = sizes;
    command = snapshot_command::read;
    fd_snap_p = fd;
    sync.main_thread_release();
}

void
rtm_snap_gpu::finish_read_uncompressed_fwi_snapshot()
{
    wait_token();
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
    wait_token();
}

void
rtm_snap_gpu::start_read_compressed_fwi_snapshot(std::vector<cart_volume<float>*>& l, std::vector<int>& sizes,
                                                 file_snap* fd, std::vector<float*>& zipped,
                                                 std::vector<zfp_field*>& fields, zfp_stream* z)
{
    // snap_helpers/rtm_snap_gpu.cu:53-74
    check_zipped(l, sizes, zipped, fields, z);
    wait_token();
    gpu_corrBuffList = l;
    zfpFields = fields;
    zfpStream = z;
    zipped_corr_buff = zipped;
    corrBuffSize = sizes;
    command = snapshot_command::read;
    fd_snap_p = fd;
    sync.main_thread_release();
}

void
rtm_snap_gpu::finish_read_compressed_fwi_snapshot()
{
    wait_token();
}

void
rtm_snap_gpu::start_write_compressed_fwi_snapshot(std::vector<cart_volume<float>*>& l, std::vector<int>& sizes,
                                                  file_snap* fd, std::vector<float*>& zipped,
                                                  std::vector<zfp_field*>& fields, zfp_stream* z)
{
    // snap_helpers/rtm_snap_gpu.cu:104-125; the caller holds the token since lock_for_write_compressed_fwi_snapshot
    check_zipped(l, sizes, zipped, fields, z);
    sync.main_thread_wait();
    gpu_corrBuffList = l;
    zfpFields = fields;
    zipped_corr_buff = zipped;
    zfpStream = z;
    corrBuffSize = sizes;
    command = snapshot_command::write;
    fd_snap_p = fd;
    sync.main_thread_release();
}

void
rtm_snap_gpu::lock_for_write_compressed_fwi_snapshot()
{
    wait_token();
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
                    useZFP_ ? helper_thread_write_compressed() : helper_thread_write();
                else
                    useZFP_ ? helper_thread_read_compressed() : helper_thread_read();
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

void
rtm_snap_gpu::helper_thread_write_compressed()
{
    // snap_helpers/rtm_snap_gpu.cu:228-252: each device volume coded into its pinned buffer, then the store. The codec
    // runs on this object's non-blocking stream (the reference uses the default stream, which would also wait for
    // the solver's unrelated work).
    tti::hip_check(hipStreamWaitEvent(helperStream, writeSnapshotsCompleteEvent_, 0), "rtm_snap_gpu: wait");
    zfp_stream_set_hip_stream(zfpStream, helperStream);
    for (size_t i = 0; i < nvol; ++i) {
        zfp_field_set_pointer(zfpFields[i], gpu_corrBuffList[i]->as<cart_volume_regular_gpu>()->getData());
        bitstream* bs = stream_open(zipped_corr_buff[i], (size_t)corrBuffSize[i] * sizeof(float));
        zfp_stream_set_bit_stream(zfpStream, bs);
        zfp_stream_rewind(zfpStream);
        const size_t n = zfp_compress(zfpStream, zfpFields[i]);
        stream_close(bs);
        if (!n)
            throw std::runtime_error("rtm_snap_gpu: ZFP compression failed");
    }
    fd_snap_p->write_step(zipped_corr_buff, corrBuffSize);
}

void
rtm_snap_gpu::helper_thread_read_compressed()
{
    // snap_helpers/rtm_snap_gpu.cu:254-280: store -> pinned buffers -> decoded into the device volumes, after the
    // previous step's consumers are done
    tti::hip_check(hipStreamWaitEvent(helperStream, readSnapshotsCompleteEvent_, 0), "rtm_snap_gpu: wait");
    fd_snap_p->read_step(zipped_corr_buff, corrBuffSize);
    zfp_stream_set_hip_stream(zfpStream, helperStream);
    for (size_t i = 0; i < nvol; ++i) {
        zfp_field_set_pointer(zfpFields[i], gpu_corrBuffList[i]->as<cart_volume_regular_gpu>()->getData());
        bitstream* bs = stream_open(zipped_corr_buff[i], (size_t)corrBuffSize[i] * sizeof(float));
        zfp_stream_set_bit_stream(zfpStream, bs);
        zfp_stream_rewind(zfpStream);
        const size_t n = zfp_decompress(zfpStream, zfpFields[i]);
        stream_close(bs);
        if (!n)
            throw std::runtime_error("rtm_snap_gpu: ZFP decompression failed");
    }
    tti::hip_check(hipStreamSynchronize(helperStream), "rtm_snap_gpu: decode");
}
