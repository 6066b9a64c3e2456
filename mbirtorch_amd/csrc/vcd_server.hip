// Continuous-batching reconstruction engine.  Admitted VCD jobs occupy slots;
// a native worker thread advances every active job by one subset step per
// round (mbir_vcd_run), each job on its own HIP stream, so the steps of
// different reconstructions overlap on the device.  Jobs submitted while
// others run are admitted as soon as a slot frees; finished jobs are handed
// back by tag.  The library allocates no device memory: every job's state
// and scratch come from the caller (see mbir_vcd_job).
#include "common.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace {

struct Job {
  mbir_vcd_job job;
  const int32_t* schedule;
  const int32_t* iter_offsets;
  int32_t num_iters;
  double* stats;
  int64_t tag;
  int iter = 0;
  int step = 0;
  int status = MBIR_OK;
  bool issued = false;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  bool waited = false;  // completion established by a stream wait, not the event
  Job job;
};

}  // namespace

struct mbir_server {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> pending;
  std::vector<Slot> slots;
  std::deque<std::pair<int64_t, int>> finished;
  int device = 0;
  int in_flight = 0;  // pending + in slots
  bool stop = false;
  std::thread worker;
};

namespace {

// The job's last launch is queued: record its completion event.  If the
// record fails the completion cannot be polled, so wait for the stream here.
void mark_issued(Slot& sl) {
  sl.job.issued = true;
  sl.waited = false;
  if (hipEventRecord(sl.done, sl.stream) != hipSuccess) {
    if (hipStreamSynchronize(sl.stream) != hipSuccess && sl.job.status == MBIR_OK)
      sl.job.status = MBIR_ERR_LAUNCH;
    sl.waited = true;
  }
}

// Whether an issued job has finished on the device (an event error ends the
// job with MBIR_ERR_LAUNCH rather than leaving it pending).
bool slot_done(Slot& sl) {
  if (sl.waited) return true;
  const hipError_t q = hipEventQuery(sl.done);
  if (q == hipErrorNotReady) return false;
  if (q != hipSuccess && sl.job.status == MBIR_OK) sl.job.status = MBIR_ERR_LAUNCH;
  return true;
}

// Issue the next subset step of a job (or its end-of-iteration statistics
// for an empty iteration); marks the job issued after its last step.
void issue_step(Slot& sl) {
  Job& j = sl.job;
  const int S = j.job.prm.num_slices;
  if (j.iter >= j.num_iters || j.status != MBIR_OK) {
    mark_issued(sl);
    return;
  }
  const int begin = j.iter_offsets[j.iter], end = j.iter_offsets[j.iter + 1];
  double* stats = j.stats + (size_t)j.iter * (S + 6);
  const int at = begin + j.step;
  const bool last = at + 1 >= end;
  j.status = mbir_vcd_run(&j.job, at < end ? j.schedule + at : nullptr, at < end ? 1 : 0, stats,
                          last ? 1 : 0, (mbir_stream_t)sl.stream);
  if (last) {
    ++j.iter;
    j.step = 0;
  } else {
    ++j.step;
  }
  if (j.iter >= j.num_iters || j.status != MBIR_OK) mark_issued(sl);
}

void worker_loop(mbir_server* sv) {
  // On a device it cannot select, the worker fails every job it takes.
  const bool device_ok = hipSetDevice(sv->device) == hipSuccess;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      for (Slot& sl : sv->slots) {
        if (!sl.busy && !sv->pending.empty()) {
          sl.job = sv->pending.front();
          sv->pending.pop_front();
          sl.busy = true;
          if (!device_ok) sl.job.status = MBIR_ERR_LAUNCH;
        }
      }
      bool any = false;
      for (const Slot& sl : sv->slots) any = any || sl.busy;
      if (!any) {
        if (sv->stop) return;
        sv->cv.wait(lk, [sv] { return sv->stop || !sv->pending.empty(); });
        continue;
      }
    }
    bool progressed = false;
    for (Slot& sl : sv->slots) {
      if (sl.busy && !sl.job.issued) {
        issue_step(sl);
        progressed = true;
      }
    }
    for (Slot& sl : sv->slots) {
      if (sl.busy && sl.job.issued && slot_done(sl)) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->finished.emplace_back(sl.job.tag, sl.job.status);
        sl.busy = false;
        --sv->in_flight;
        sv->cv.notify_all();
      }
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

}  // namespace

extern "C" {

mbir_server* mbir_server_create(int32_t max_active) {
  if (max_active < 1 || max_active > 64) return nullptr;
  mbir_server* sv = new mbir_server();
  if (hipGetDevice(&sv->device) != hipSuccess) {
    delete sv;
    return nullptr;
  }
  sv->slots.resize(max_active);
  for (Slot& sl : sv->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      for (Slot& s2 : sv->slots) {
        if (s2.stream) (void)hipStreamDestroy(s2.stream);
        if (s2.done) (void)hipEventDestroy(s2.done);
      }
      delete sv;
      return nullptr;
    }
  }
  sv->worker = std::thread(worker_loop, sv);
  return sv;
}

int mbir_server_submit(mbir_server* sv, const mbir_vcd_job* job, const int32_t* schedule,
                       const int32_t* iter_offsets, int32_t num_iters, double* stats,
                       int64_t tag) {
  if (!sv || !job || !iter_offsets || !stats || num_iters < 0) return MBIR_ERR_ARG;
  if (iter_offsets[num_iters] > 0 && !schedule) return MBIR_ERR_ARG;
  for (int i = 0; i < num_iters; ++i)
    if (iter_offsets[i + 1] < iter_offsets[i]) return MBIR_ERR_ARG;
  Job j;
  j.job = *job;
  j.schedule = schedule;
  j.iter_offsets = iter_offsets;
  j.num_iters = num_iters;
  j.stats = stats;
  j.tag = tag;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return MBIR_ERR_ARG;
    sv->pending.push_back(j);
    ++sv->in_flight;
  }
  sv->cv.notify_all();
  return MBIR_OK;
}

int32_t mbir_server_poll(mbir_server* sv, int64_t* tags, int32_t* status, int32_t max_jobs,
                         int32_t timeout_ms) {
  if (!sv || !tags || !status || max_jobs < 1) return MBIR_ERR_ARG;
  std::unique_lock<std::mutex> lk(sv->mu);
  if (sv->finished.empty() && timeout_ms > 0)
    sv->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                    [sv] { return !sv->finished.empty(); });
  int n = 0;
  while (n < max_jobs && !sv->finished.empty()) {
    tags[n] = sv->finished.front().first;
    status[n] = sv->finished.front().second;
    sv->finished.pop_front();
    ++n;
  }
  return n;
}

int32_t mbir_server_in_flight(mbir_server* sv) {
  if (!sv) return 0;
  std::lock_guard<std::mutex> lk(sv->mu);
  return sv->in_flight;
}

void mbir_server_destroy(mbir_server* sv) {
  if (!sv) return;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv.wait(lk, [sv] { return sv->in_flight == 0; });
    sv->stop = true;
  }
  sv->cv.notify_all();
  sv->worker.join();
  for (Slot& sl : sv->slots) {
    (void)hipStreamDestroy(sl.stream);
    (void)hipEventDestroy(sl.done);
  }
  delete sv;
}

}  // extern "C"
