// Device time of driver phases without a synchronisation per phase.
//
// The reference times each phase with an event pair and waits for the second event before the next phase
// (simulator2/simulator2SetUp.cu:547-599), which keeps the host from queueing ahead of the GPU. Here each phase
// records a pair from a fixed pool on the null stream; the pairs are resolved together (one wait on the latest
// event) when the pool is full or at flush(), so the host runs up to `pairs` phases ahead.
#pragma once

#include "grid.h"

#include <hip/hip_runtime.h>

#include <vector>

namespace tti {

class PhaseTimer
{
  public:
    explicit PhaseTimer(int pairs = 64) : ev_(2 * pairs, nullptr), acc_(pairs, nullptr) {}
    ~PhaseTimer()
    {
        for (auto e : ev_)
            if (e)
                (void)hipEventDestroy(e);
    }
    PhaseTimer(const PhaseTimer&) = delete;
    PhaseTimer& operator=(const PhaseTimer&) = delete;

    // adds the device time of the work f() queues on the null stream to `acc` (seconds) at the next flush
    template <class F>
    void time(double& acc, F&& f)
    {
        if (ev_[0] == nullptr) // the events are created on first use (the drivers validate their arguments first)
            for (auto& e : ev_)
                hip_check(hipEventCreate(&e), "PhaseTimer: hipEventCreate");
        if (used_ == acc_.size())
            flush();
        hip_check(hipEventRecord(ev_[2 * used_], nullptr), "PhaseTimer: hipEventRecord");
        f();
        hip_check(hipEventRecord(ev_[2 * used_ + 1], nullptr), "PhaseTimer: hipEventRecord");
        acc_[used_++] = &acc;
    }

    // waits for the latest recorded event and adds every pending pair's elapsed time to its accumulator
    void flush()
    {
        if (used_ == 0)
            return;
        hip_check(hipEventSynchronize(ev_[2 * used_ - 1]), "PhaseTimer: hipEventSynchronize");
        for (size_t k = 0; k < used_; ++k) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, ev_[2 * k], ev_[2 * k + 1]), "PhaseTimer: hipEventElapsedTime");
            *acc_[k] += ms * 1e-3;
        }
        used_ = 0;
    }

  private:
    std::vector<hipEvent_t> ev_;
    std::vector<double*> acc_;
    size_t used_ = 0;
};

} // namespace tti
