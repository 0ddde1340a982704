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
//
// Generated structures stay resident in HBM after their requests (keyed by
// their generator arguments, least recently used evicted when admission needs
// the memory), and the workers share one surface Green's function cache, so a
// request on a structure served before skips the generation and, for the
// energies already solved, the decimation (the role of ln_cache's memory
// cache, ln_cache.F90:154-171, 213-295, with surface_green's load/store,
// contselfenergy.F90:161-173, 251-259).
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
  size_t sgf_cache = 0;         // bytes of the surface-function cache shared by the workers (0: none)
  int keep_structs = 1;         // generated structures stay resident after their requests (LRU)
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
  // requests served on an already resident generated structure, resident
  // generated structures, surface functions loaded from / stored into the cache
  void cache_stats(long long out[4]);

 private:
  enum State { QUEUED, ADMITTING, RUNNING, DONE };
  // a generated structure kept in HBM, keyed by its generator arguments
  struct GenKey {
    unsigned long long seed, kx, ky;
    int n, nbl, overlap;
    bool operator==(const GenKey &o) const {
      return seed == o.seed && kx == o.kx && ky == o.ky && n == o.n && nbl == o.nbl &&
             overlap == o.overlap;
    }
  };
  struct GenKeyHash {
    size_t operator()(const GenKey &k) const {
      size_t h = std::hash<unsigned long long>()(k.seed);
      h ^= std::hash<unsigned long long>()(k.kx * 1000003ULL ^ k.ky) + 0x9e3779b97f4a7c15ULL +
           (h << 6) + (h >> 2);
      h ^= (size_t)k.n * 7919u + (size_t)k.nbl * 104729u + (size_t)k.overlap;
      return h;
    }
  };
  struct Resident {
    std::unique_ptr<DevStruct> st;
    void *mem = nullptr;
    size_t bytes = 0;
    int refs = 0;          // admitted requests using it
    bool ready = false;    // generated
    unsigned long long last = 0;
  };
  struct Req {
    unsigned long long id = 0;
    ReqStruct rs;
    std::vector<double> E;
    SolveOpts o;
    State state = QUEUED;
    Resident *gst = nullptr;           // generated structure (resident, shared)
    DevStruct *st = nullptr;
    void *mem = nullptr;               // device block: E, T, ncyc, status, slot per item
    size_t footprint = 0;              // bytes of mem
    size_t struct_bytes = 0;           // bytes of the generated structure
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
  std::unordered_map<GenKey, std::unique_ptr<Resident>, GenKeyHash> residents_;
  unsigned long long res_clock_ = 0;
  long long struct_hits_ = 0;
  std::shared_ptr<SgfCache> sgf_;
  std::deque<Req *> queue_;     // waiting for admission, FCFS
  std::vector<Req *> running_;  // admitted, chunks left or in flight
  std::vector<std::thread> threads_;

  void worker(int w);
  size_t footprint(const Req &r) const;
  static GenKey gen_key(const ReqStruct &rs);
  // frees idle resident structures (least recently used first) until `need`
  // more bytes fit; their device memory is returned in `freed` (caller holds mu_)
  void evict_idle(size_t need, std::vector<std::unique_ptr<Resident>> &freed);
  void admit(Req &r, Engine &eng, bool build);  // outside the lock
  void finish(Req &r, Engine &eng);             // outside the lock
};

}  // namespace hn
