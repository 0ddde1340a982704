// hn_serve.h -- persistent request loop over the batched engine (serving).
//
// Requests (a structure + an energy grid + solve options) are submitted from
// any thread and served by W worker threads, each driving its own engine and
// HIP stream, so chunks of different requests run on the device at the same
// time.  A request is admitted when its structure fits the HBM budget; its
// energies are then cut into chunks (batches) sized so that one chunk is a
// bounded amount of work, and every worker repeatedly takes the next chunk of
// the admitted request chosen by the policy (FCFS: oldest first; LAS: least
// GPU time received so far).  New requests therefore join the running set
// between chunks (continuous batching at chunk granularity) instead of
// waiting for the requests ahead of them to finish.  Each item's result is
// computed by the same kernels whatever chunk it is in, so a request's result
// is the same alone and in any mix.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hn_engine.h"

namespace hn {

struct ServeOpts {
  int workers = 2;              // engines / streams
  int policy = 0;               // 0 FCFS, 1 least attained service
  int max_chunk = 0;            // items per chunk cap (0: automatic)
  double chunk_flops = 2e12;    // target algorithmic work per chunk (automatic sizing)
  int min_chunk = 64;           // items per chunk at least (keeps large-n chunks wide)
  size_t mem_budget = 0;        // bytes for structures + workspaces (0: 85 % of free HBM)
  size_t worker_ws = 0;         // workspace bytes per worker (0: budget / (4 workers))
};

// What a request is solved on.
struct ReqStruct {
  // generated slab (seed, n, nbl, k): built in HBM at admission and freed at completion
  int gen = 0;
  unsigned long long seed = 0;
  int n = 0, nbl = 0, overlap = 1;
  double kx = 0.0, ky = 0.0;
  // or a caller-owned resident structure (slot `slot`), not freed by the server
  DevStruct *resident = nullptr;
  int slot = 0;
};

struct ReqResult {
  std::vector<double> T;
  std::vector<int> ncyc, status;
  double t_queue = 0.0;    // s, submit -> admitted
  double t_latency = 0.0;  // s, submit -> results on the host
  double t_gpu = 0.0;      // s, summed chunk time on the workers
  int chunks = 0;
  std::string error;
};

class Server {
 public:
  Server(int device, const ServeOpts &o);
  ~Server();
  unsigned long long submit(const ReqStruct &rs, const std::vector<double> &E, const SolveOpts &o);
  bool done(unsigned long long id);
  // blocks until the request finished; removes it from the server
  ReqResult wait(unsigned long long id);
  // submitted, completed, admitted (running), queued
  void stats(long long out[4]);

 private:
  enum State { QUEUED, ADMITTING, RUNNING, DONE };
  struct Req {
    unsigned long long id = 0;
    ReqStruct rs;
    std::vector<double> E;
    SolveOpts o;
    State state = QUEUED;
    std::unique_ptr<DevStruct> owned;  // generated structure
    DevStruct *st = nullptr;
    void *mem = nullptr;               // device block: structure (if generated) + E, T, ncyc, status
    size_t footprint = 0;
    double *dE = nullptr, *dT = nullptr;
    int *dN = nullptr, *dS = nullptr, *dSl = nullptr;  // ncyc, status, slot per item
    int nE = 0, chunk = 1, next = 0, inflight = 0, finished = 0, ncont = 0;
    std::chrono::steady_clock::time_point t_submit, t_admit;
    ReqResult res;
  };

  int device_;
  ServeOpts o_;
  size_t budget_ = 0, used_ = 0;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  bool stop_ = false;
  unsigned long long next_id_ = 1;
  long long n_submitted_ = 0, n_completed_ = 0;
  std::unordered_map<unsigned long long, std::unique_ptr<Req>> reqs_;
  std::deque<Req *> queue_;     // waiting for admission, FCFS
  std::vector<Req *> running_;  // admitted, chunks left or in flight
  std::vector<std::thread> threads_;

  void worker(int w);
  size_t footprint(const Req &r) const;
  void admit(Req &r, Engine &eng);       // outside the lock
  void finish(Req &r, Engine &eng);      // outside the lock
};

}  // namespace hn
