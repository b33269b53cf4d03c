// mp_internal.hpp -- host-side structures shared by the model loader, the
// batch planner and the HIP driver.
#pragma once

#include <atomic>
#include <cstdint>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "mp_pairhmm.h"
#include "mp_poa.h"

namespace mp {

// Runs f(t) for t = 0 .. n-1 on n host threads (t = 0 on the calling thread)
// and joins every thread it started before returning.  An exception thrown by
// any f (std::bad_alloc, say), or a thread that cannot be started, is caught
// where it happens and the first one is rethrown here once all threads have
// ended, so no joinable std::thread is ever destroyed.  Work handed out from a
// shared counter is still finished by the threads that run; a worker that
// waits for others must also stop when `failed()` turns true.
class ThreadPoolRun {
public:
    template <class F> void run(int n, F &&f) {
        std::vector<std::thread> th;
        auto guarded = [&](int t) {
            try {
                f(t);
            } catch (...) {
                record(std::current_exception());
            }
        };
        try {
            if (n > 1) th.reserve((size_t)n - 1);
            for (int t = 1; t < n; t++) th.emplace_back(guarded, t);
        } catch (...) {
            record(std::current_exception());
        }
        guarded(0);
        for (auto &x : th) x.join();
        if (first_) std::rethrow_exception(first_);
    }
    bool failed() const { return failed_.load(std::memory_order_relaxed); }

private:
    void record(std::exception_ptr e) {
        std::lock_guard<std::mutex> lk(mu_);
        if (!first_) first_ = e;
        failed_.store(true);
    }
    std::mutex mu_;
    std::exception_ptr first_;
    std::atomic<bool> failed_{false};
};

template <class F> void parallel_run(int n, F &&f) {
    ThreadPoolRun pool;
    pool.run(n, std::forward<F>(f));
}

constexpr int kMaxRepeat = 51;            // MAXIMUM_REPEAT_LENGTH, margin.h:123
constexpr int kRepeatEntries = 5 * kMaxRepeat * kMaxRepeat;

// Transition slots, natural log, in the order of StateMachine3 (stateMachine.c:490-500).
enum Trans { kMatchContinue, kMatchFromGapX, kMatchFromGapY, kGapOpenX, kGapOpenY,
             kGapExtendX, kGapExtendY, kGapSwitchToX, kGapSwitchToY };

// Five-state machine (no counterpart in the reference): M, short gap X, short
// gap Y, long gap X, long gap Y.  A gap state of either length emits like the
// three-state gap of its side.
enum State5 { kS5Match, kS5ShortX, kS5ShortY, kS5LongX, kS5LongY };

struct Machine {
    int nStates = 3;
    double trans5[25] = {};  // natural log, [from * 5 + to], when nStates == 5
    double trans[9];
    double match[16], gapX[4], gapY[4];
    int rle = 0;
    int strand = 1;
};

void set_error(const std::string &msg);

}  // namespace mp

struct mp_model {
    mp::Machine mach[2];          // [0] reverse strand, [1] forward strand (same nStates and transitions)
    std::vector<double> repeatLog10;  // [5][51][51] log10, [base][underlying][observed]
    mp_band_params p;
    bool endPinned = false;       // end-state vector from endTrans instead of the transitions
    double endTrans[9] = {};
};

namespace mp {

// Geometry of one sub-problem (after large-gap splitting).
struct Band {
    std::vector<int32_t> L, R;    // xmyL / xmyR per anti-diagonal 0..D
};

struct Window {
    int32_t tb, lo, from;
};

// band_constructDynamic / band_construct semantics (pairwiseAligner.c:120-226).
// Returns false when a diagonal is invalid (bad anchors).
bool build_band(const int64_t *anchors, int64_t n, int64_t lX, int64_t lY, int64_t expansion, bool dynamic,
                int32_t *L, int32_t *R);
// The same with anchor coordinates read as (x - xo, y - yo): a sub-problem's
// anchors taken in place from the pair's.
bool build_band(const int64_t *anchors, int64_t n, int64_t xo, int64_t yo, int64_t lX, int64_t lY,
                int64_t expansion, bool dynamic, int32_t *L, int32_t *R);
// Traceback windows of getPosteriorProbsWithBanding (pairwiseAligner.c:715-783).
void build_windows(const int32_t *L, const int32_t *R, int64_t D, const mp_band_params &p,
                   std::vector<Window> &out);
// getSplitPoints (pairwiseAligner.c:899-926): (x1, y1, x2, y2) tuples.
void split_points(const int64_t *anchors, int64_t n, int64_t lX, int64_t lY, int64_t maxArea, bool raggedL,
                  bool raggedR, std::vector<int64_t> &out);
// The same with anchor x read as x - xo (a cropped reference).
void split_points(const int64_t *anchors, int64_t n, int64_t xo, int64_t lX, int64_t lY, int64_t maxArea,
                  bool raggedL, bool raggedR, std::vector<int64_t> &out);
// Anchor validity: strictly increasing, inside the sequences, even expansion.
bool anchors_valid(const int64_t *anchors, int64_t n, int64_t lX, int64_t lY);

// The augment loop of poa_realign over the pairs of a batch (mp_poa.cpp):
// pair i's lists are lists[l][starts[l][i] .. starts[l][i + 1]).
mp_status poa_augment_pairs(mp_poa *const *poas, int64_t nPoas, const int64_t *pairPoa, const int64_t *pairReadNo,
                            const mp_rle_string *reads, const uint8_t *strand, int64_t npairs,
                            const mp_aligned_pair *const lists[3], const int64_t *const starts[3], bool cmp,
                            int threads);

}  // namespace mp
