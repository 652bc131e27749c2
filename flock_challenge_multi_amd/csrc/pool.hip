// Batched and served proving: a pool of host worker threads, each with its own
// HIP stream, proving BLAKE3 R1CS batches of any size through one device memory
// budget.
//
// The reference proves one batch per call (`prove_ligerito`, CP/prover.rs:373-472)
// and its harness runs ranked workers one proof at a time
// (benchmark-tools/harness/src/main.rs:360-407).  On one MI355X several proofs run
// at once: while one waits on a read-back for its Fiat-Shamir transcript, another
// one's kernels use the device, and small proofs (m = 24..28) do not fill it alone.
//
//   admission   the pool owns one device arena of `hbm_budget` bytes.  A job
//               starts when a view of its setup's footprint (blake3_arena_bytes)
//               can be carved from it; jobs are admitted in submission order (a
//               job that does not fit holds back the ones behind it, so a large
//               proof is never starved by a stream of small ones).
//   setups      one flk_blake3_setup per batch size class (n_blocks_log), made on
//               first use (the R1CS, its statement digest and the ladder) or ahead
//               of time with flk_prover_pool_prepare.
//   results     each job's proof bytes stay in the pool until flk_prover_pool_wait
//               copies them out; the pool records submit, start and finish times.
//
// A job's bytes equal those of flk_blake3_prove on the same input: the prove is
// the same code on its own stream and arena view, and the transcript is per job.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

namespace {

using clk = std::chrono::steady_clock;

struct Job {
  uint64_t id = 0;
  const flk_compression* blocks = nullptr;
  size_t n_blocks = 0, n_blocks_log = 0;
  std::string domain;
  std::vector<uint8_t> proof;
  int rc = FLK_OK;
  std::string err;  // the worker's flk_last_error() when rc != FLK_OK
  bool done = false;
  clk::time_point t_submit, t_start, t_done;
};

size_t n_blocks_log_of(size_t n) {
  size_t nbl = 3;  // min_n_blocks_log: ceil(log2(max(n, 8))) (CP/r1cs_hashes/blake3.rs:1854-1858)
  while ((size_t(1) << nbl) < n) ++nbl;
  return nbl;
}

}  // namespace

struct flk_prover_pool {
  int profile = 0;
  int device = 0;
  size_t budget = 0;
  flk::Arena* arena = nullptr;

  std::mutex mu;  // queue, jobs, counters
  std::condition_variable cv;
  std::deque<Job*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  uint64_t next_id = 1;
  bool stop = false;
  bool head_waiting = false;  // a worker holds the head job until memory frees up
  size_t reserved = 0, peak_reserved = 0, running = 0, peak_running = 0, completed = 0;

  std::mutex setup_mu;  // setups (creation is slow; kept out of `mu`)
  std::map<size_t, flk_blake3_setup*> setups;

  std::vector<std::thread> workers;

  int setup_for(size_t nbl, flk_blake3_setup** out) {
    std::lock_guard<std::mutex> lk(setup_mu);
    auto it = setups.find(nbl);
    if (it != setups.end()) {
      *out = it->second;
      return FLK_OK;
    }
    flk_blake3_setup* s = nullptr;
    if (int rc = flk_blake3_setup_new(size_t(1) << nbl, profile, &s)) return rc;
    setups[nbl] = s;
    *out = s;
    return FLK_OK;
  }

  void worker_main() {
    (void)hipSetDevice(device);
    hipStream_t st = nullptr;
    const bool have_stream = hipStreamCreate(&st) == hipSuccess;  // default priority
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || (!queue.empty() && !head_waiting); });
        if (stop) break;
        j = queue.front();
        queue.pop_front();
        head_waiting = true;  // nobody else is admitted before this job
      }
      flk_blake3_setup* s = nullptr;
      int rc = have_stream ? setup_for(j->n_blocks_log, &s)
                           : flk::set_error(FLK_E_DEVICE, "prover pool: hipStreamCreate failed");
      flk::Arena* view = nullptr;
      size_t need = 0;
      if (rc == FLK_OK) {
        need = flk::blake3_arena_bytes(s);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          int vr = FLK_OK;
          view = flk::arena_view_new(arena, need, &vr);
          if (view || running == 0 || stop) break;
          cv.wait(lk);  // a finishing job frees its view
        }
        if (!view)
          rc = flk::set_error(FLK_E_NOMEM, "prover pool: a job of %zu bytes does not fit the %zu-byte budget", need,
                              budget);
        else {
          reserved += need;
          ++running;
          peak_reserved = std::max(peak_reserved, reserved);
          peak_running = std::max(peak_running, running);
        }
        head_waiting = false;
        cv.notify_all();
      } else {
        std::lock_guard<std::mutex> lk(mu);
        head_waiting = false;
        cv.notify_all();
      }
      j->t_start = clk::now();
      if (rc == FLK_OK) {
        size_t bound = 0, len = 0;
        rc = flk_blake3_proof_bound(s, &bound);
        if (rc == FLK_OK) {
          j->proof.resize(bound);
          rc = flk::blake3_prove_in(s, view, j->blocks, j->n_blocks,
                                    reinterpret_cast<const uint8_t*>(j->domain.data()), j->domain.size(),
                                    j->proof.data(), bound, &len, nullptr, st);
          j->proof.resize(rc == FLK_OK ? len : 0);
        }
      }
      if (view) flk::arena_delete(view);
      {
        std::lock_guard<std::mutex> lk(mu);
        if (view) {
          reserved -= need;
          --running;
        }
        j->rc = rc;
        if (rc != FLK_OK) j->err = flk_last_error();
        j->t_done = clk::now();
        j->done = true;
        ++completed;
        cv.notify_all();
      }
    }
    if (st) (void)hipStreamDestroy(st);
  }
};

extern "C" {

int flk_prover_pool_new(int profile, size_t workers, size_t hbm_budget, flk_prover_pool** out) {
  if (int rc = flk::check_enabled()) return rc;
  if (!out) return flk::set_error(FLK_E_SHAPE, "prover_pool_new: null pointer");
  *out = nullptr;
  if (profile < 0 || profile > 2) return flk::set_error(FLK_E_SHAPE, "prover_pool_new: profile %d", profile);
  if (workers < 1 || workers > 16) return flk::set_error(FLK_E_SHAPE, "prover_pool_new: workers=%zu (1..16)", workers);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return flk::device_failure(e, "prover_pool_new: hipGetDevice");
  if (hbm_budget == 0) {  // 60% of the device's free memory
    size_t free_b = 0, total_b = 0;
    e = hipMemGetInfo(&free_b, &total_b);
    if (e != hipSuccess) return flk::device_failure(e, "prover_pool_new: hipMemGetInfo");
    hbm_budget = free_b / 10 * 6;
  }
  std::unique_ptr<flk_prover_pool> p(new (std::nothrow) flk_prover_pool());
  if (!p) return flk::set_error(FLK_E_NOMEM, "prover_pool_new: out of host memory");
  int rc = FLK_OK;
  p->arena = flk::arena_new(hbm_budget, &rc);
  if (!p->arena) return rc;
  p->profile = profile;
  p->device = dev;
  p->budget = hbm_budget;
  try {
    for (size_t i = 0; i < workers; ++i) p->workers.emplace_back([q = p.get()] { q->worker_main(); });
  } catch (const std::exception&) {
    {
      std::lock_guard<std::mutex> lk(p->mu);
      p->stop = true;
    }
    p->cv.notify_all();
    for (auto& t : p->workers) t.join();
    flk::arena_delete(p->arena);
    return flk::set_error(FLK_E_NOMEM, "prover_pool_new: could not start the worker threads");
  }
  *out = p.release();
  return FLK_OK;
}

void flk_prover_pool_free(flk_prover_pool* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    p->stop = true;  // running jobs finish; queued ones are dropped
  }
  p->cv.notify_all();
  for (auto& t : p->workers) t.join();
  for (auto& kv : p->setups) flk_blake3_setup_free(kv.second);
  flk::arena_delete(p->arena);
  delete p;
}

int flk_prover_pool_prepare(flk_prover_pool* p, size_t n_blocks) {
  if (!p || n_blocks < 1) return flk::set_error(FLK_E_SHAPE, "prover_pool_prepare: null pool or n_blocks = 0");
  flk_blake3_setup* s = nullptr;
  return p->setup_for(n_blocks_log_of(n_blocks), &s);
}

int flk_prover_pool_submit(flk_prover_pool* p, const flk_compression* blocks, size_t n_blocks, const uint8_t* domain,
                           size_t domain_len, uint64_t* ticket) {
  if (int rc = flk::check_enabled()) return rc;
  if (!p || !ticket || (n_blocks && !blocks) || (domain_len && !domain))
    return flk::set_error(FLK_E_SHAPE, "prover_pool_submit: null pointer");
  if (n_blocks < 1) return flk::set_error(FLK_E_SHAPE, "prover_pool_submit: n_blocks must be >= 1");
  const size_t nbl = n_blocks_log_of(n_blocks);
  if (nbl < 8 || nbl > 21)
    return flk::set_error(FLK_E_SHAPE, "prover_pool_submit: %zu compressions (m = %zu) outside m = 22..35", n_blocks,
                          14 + nbl);
  std::unique_ptr<Job> j(new (std::nothrow) Job());
  if (!j) return flk::set_error(FLK_E_NOMEM, "prover_pool_submit: out of host memory");
  j->blocks = blocks;
  j->n_blocks = n_blocks;
  j->n_blocks_log = nbl;
  j->domain.assign(reinterpret_cast<const char*>(domain), domain_len);
  j->t_submit = clk::now();
  {
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->stop) return flk::set_error(FLK_E_SHAPE, "prover_pool_submit: the pool is shutting down");
    j->id = p->next_id++;
    *ticket = j->id;
    p->queue.push_back(j.get());
    p->jobs[j->id] = std::move(j);
  }
  p->cv.notify_all();
  return FLK_OK;
}

int flk_prover_pool_wait(flk_prover_pool* p, uint64_t ticket, uint8_t* proof_out, size_t cap, size_t* proof_len,
                         double* times) {
  if (!p || !proof_len) return flk::set_error(FLK_E_SHAPE, "prover_pool_wait: null pointer");
  std::unique_lock<std::mutex> lk(p->mu);
  auto it = p->jobs.find(ticket);
  if (it == p->jobs.end()) return flk::set_error(FLK_E_SHAPE, "prover_pool_wait: unknown ticket %llu",
                                                 (unsigned long long)ticket);
  Job* j = it->second.get();
  p->cv.wait(lk, [&] { return j->done || p->stop; });
  if (!j->done) return flk::set_error(FLK_E_SHAPE, "prover_pool_wait: the pool stopped before the job ran");
  *proof_len = j->proof.size();
  if (times) {
    times[0] = std::chrono::duration<double>(j->t_start - j->t_submit).count();
    times[1] = std::chrono::duration<double>(j->t_done - j->t_start).count();
    times[2] = std::chrono::duration<double>(j->t_done - j->t_submit).count();
  }
  if (j->rc != FLK_OK) {
    const int rc = j->rc;
    const std::string msg = j->err;
    p->jobs.erase(it);
    return flk::set_error(rc, "prover pool job: %s", msg.c_str());
  }
  if (!proof_out || cap < j->proof.size())
    return flk::set_error(FLK_E_SHAPE, "prover_pool_wait: %zu-byte proof, %zu-byte buffer (the job is kept)",
                          j->proof.size(), cap);
  std::memcpy(proof_out, j->proof.data(), j->proof.size());
  p->jobs.erase(it);
  return FLK_OK;
}

int flk_prover_pool_stats(flk_prover_pool* p, size_t* budget, size_t* peak_reserved, size_t* peak_running,
                          size_t* completed) {
  if (!p) return flk::set_error(FLK_E_SHAPE, "prover_pool_stats: null pointer");
  std::lock_guard<std::mutex> lk(p->mu);
  if (budget) *budget = p->budget;
  if (peak_reserved) *peak_reserved = p->peak_reserved;
  if (peak_running) *peak_running = p->peak_running;
  if (completed) *completed = p->completed;
  return FLK_OK;
}

}  // extern "C"
