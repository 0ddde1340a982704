// hn_serve.hip -- persistent request loop (see hn_serve.h).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hn_gen.h"
#include "hn_serve.h"

namespace hn {

namespace {

using clk = std::chrono::steady_clock;

double secs(clk::time_point a, clk::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Shape of a generated slab: nbl PLs of n, contact 1 on the last block,
// contact 2 on block 1, contacts of PL size n.
struct SlabShape {
  std::vector<int> bsz, cblk, nc;
  int smode;
  SlabShape(int n, int nbl, int overlap)
      : bsz(nbl, n), cblk{nbl, 1}, nc{n, n}, smode(overlap ? (S_DIAG | S_OFFDIAG | S_CONTACT) : 0) {}
};

std::mutex g_prepare_mu;  // first-use checks of shared (resident) structures

}  // namespace

Server::Server(int device, const ServeOpts &o) : device_(device), o_(o) {
  if (o_.workers < 1 || o_.workers > 16) throw std::runtime_error("server: workers must be 1..16");
  HN_CHECK(hipSetDevice(device_));
  size_t free_b = 0, total_b = 0;
  HN_CHECK(hipMemGetInfo(&free_b, &total_b));
  budget_ = o_.mem_budget ? o_.mem_budget : (size_t)(0.85 * (double)free_b);
  if (!o_.worker_ws) o_.worker_ws = std::min<size_t>(budget_ / (4 * (size_t)o_.workers), 32ull << 30);
  if (budget_ <= o_.worker_ws * (size_t)o_.workers) throw std::runtime_error("server: HBM budget too small");
  budget_ -= o_.worker_ws * (size_t)o_.workers;  // the rest holds request structures and buffers
  if (o_.sgf_cache) {
    if (budget_ <= o_.sgf_cache) throw std::runtime_error("server: HBM budget too small for the cache");
    sgf_ = std::make_shared<SgfCache>(device_, o_.sgf_cache);
    budget_ -= o_.sgf_cache;
  }
  for (int w = 0; w < o_.workers; ++w) threads_.emplace_back(&Server::worker, this, w);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto &t : threads_) t.join();
  // requests never admitted or still holding device memory
  for (auto &kv : reqs_) {
    Req &r = *kv.second;
    if (r.mem) (void)hipFree(r.mem);
  }
  for (auto &kv : residents_) {
    kv.second->st.reset();
    if (kv.second->mem) (void)hipFree(kv.second->mem);
  }
}

unsigned long long Server::submit(const ReqStruct &rs, const std::vector<double> &E,
                                  const SolveOpts &o) {
  if (E.empty()) throw std::runtime_error("server: empty energy grid");
  if (!rs.gen && !rs.resident) throw std::runtime_error("server: request without a structure");
  if (rs.gen && (rs.n < 1 || rs.nbl < 2)) throw std::runtime_error("server: bad generated shape");
  if (rs.resident && (rs.slot < 0 || rs.slot >= rs.resident->nslots))
    throw std::runtime_error("server: slot out of range");
  auto r = std::make_unique<Req>();
  r->rs = rs;
  r->E = E;
  r->o = o;
  r->o.lanes = 1;  // the workers themselves are the concurrent streams
  if (sgf_ && r->o.sgf_mode == 0) r->o.sgf_mode = 1;  // load if cached, else compute and store
  r->nE = (int)E.size();
  r->t_submit = clk::now();
  std::lock_guard<std::mutex> g(mu_);
  if (stop_) throw std::runtime_error("server: stopped");
  r->id = next_id_++;
  r->footprint = footprint(*r);
  if (rs.gen) {
    SlabShape sh(rs.n, rs.nbl, rs.overlap);
    r->struct_bytes = align256(DevStruct::layout_bytes(rs.nbl, sh.bsz.data(), 2, sh.cblk.data(),
                                                       sh.nc.data(), sh.smode, 1));
  }
  Req *p = r.get();
  reqs_[p->id] = std::move(r);
  queue_.push_back(p);
  ++n_submitted_;
  cv_work_.notify_one();
  return p->id;
}

size_t Server::footprint(const Req &r) const {
  const int ncont = r.rs.gen ? 2 : std::max(r.rs.resident->ncont, 1);
  const size_t n = (size_t)r.nE;
  return align256(n * sizeof(double)) * 2 + align256(n * sizeof(int) * ncont) +
         2 * align256(n * sizeof(int));
}

Server::GenKey Server::gen_key(const ReqStruct &rs) {
  GenKey k;
  k.seed = rs.seed;
  std::memcpy(&k.kx, &rs.kx, sizeof(double));
  std::memcpy(&k.ky, &rs.ky, sizeof(double));
  k.n = rs.n;
  k.nbl = rs.nbl;
  k.overlap = rs.overlap;
  return k;
}

void Server::evict_idle(size_t need, std::vector<std::unique_ptr<Resident>> &freed) {
  while (used_ + need > budget_) {
    auto victim = residents_.end();
    for (auto it = residents_.begin(); it != residents_.end(); ++it)
      if (it->second->ready && it->second->refs == 0 &&
          (victim == residents_.end() || it->second->last < victim->second->last))
        victim = it;
    if (victim == residents_.end()) return;
    used_ -= victim->second->bytes;
    freed.push_back(std::move(victim->second));
    residents_.erase(victim);
  }
}

void Server::admit(Req &r, Engine &eng, bool build) {
  hipStream_t s = eng.stream();
  if (build) {
    SlabShape sh(r.rs.n, r.rs.nbl, r.rs.overlap);
    Resident &g = *r.gst;
    HN_CHECK(hipMalloc(&g.mem, g.bytes));
    g.st.reset(new DevStruct(r.rs.nbl, sh.bsz.data(), 2, sh.cblk.data(), sh.nc.data(), sh.smode, 1,
                             g.mem, s));
    gen_kslab(*g.st, 0, r.rs.seed, r.rs.kx, r.rs.ky, r.rs.overlap, s);
  }
  HN_CHECK(hipMallocAsync(&r.mem, r.footprint, s));
  char *p = (char *)r.mem;
  r.st = r.rs.gen ? r.gst->st.get() : r.rs.resident;
  r.ncont = std::max(r.st->ncont, 1);
  const size_t n = (size_t)r.nE;
  r.dE = (double *)p; p += align256(n * sizeof(double));
  r.dT = (double *)p; p += align256(n * sizeof(double));
  r.dN = (int *)p; p += align256(n * sizeof(int) * r.ncont);
  r.dS = (int *)p; p += align256(n * sizeof(int));
  r.dSl = (int *)p;
  HN_CHECK(hipMemcpyAsync(r.dE, r.E.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
  fill_int(r.dSl, r.nE, r.rs.gen ? 0 : r.rs.slot, s);
  HN_CHECK(hipStreamSynchronize(s));
  {
    std::lock_guard<std::mutex> g(g_prepare_mu);
    eng.prepare(*r.st);
  }
  // chunk: about chunk_flops of algorithmic work, within the worker workspace
  const int nm = r.st->nmax();
  const double n3 = (double)nm * nm * nm;
  const double per_item_flops = (56.0 * 2 * 25 + 24.0 * r.st->nbl + 64.0) * n3;
  long long c = std::max<long long>((long long)std::floor(o_.chunk_flops / per_item_flops),
                                    o_.min_chunk);
  const size_t ib = Engine::transmission_item_bytes(*r.st, 1, &r.o.ni, &r.o.nf) +
                    inv_work_bytes(nm, 1, 8);
  c = std::min<long long>(c, (long long)(o_.worker_ws / std::max<size_t>(ib, 1)));
  if (o_.max_chunk > 0) c = std::min<long long>(c, o_.max_chunk);
  c = std::max<long long>(1, std::min<long long>(c, std::min(r.nE, 4096)));
  r.chunk = (int)c;
  r.o.max_batch = r.chunk;
}

void Server::finish(Req &r, Engine &eng) {
  hipStream_t s = eng.stream();
  if (r.st && r.res.error.empty()) {
    r.res.T.resize(r.nE);
    r.res.ncyc.resize((size_t)r.nE * r.ncont);
    r.res.status.resize(r.nE);
    HN_CHECK(hipMemcpyAsync(r.res.T.data(), r.dT, sizeof(double) * r.nE, hipMemcpyDeviceToHost, s));
    HN_CHECK(hipMemcpyAsync(r.res.ncyc.data(), r.dN, sizeof(int) * r.nE * r.ncont,
                            hipMemcpyDeviceToHost, s));
    HN_CHECK(hipMemcpyAsync(r.res.status.data(), r.dS, sizeof(int) * r.nE, hipMemcpyDeviceToHost, s));
  }
  if (r.mem) HN_CHECK(hipFreeAsync(r.mem, s));
  HN_CHECK(hipStreamSynchronize(s));
  r.mem = nullptr;
  r.st = nullptr;
  r.res.t_latency = secs(r.t_submit, clk::now());
}

void Server::worker(int w) {
  (void)w;
  std::unique_ptr<Engine> eng;
  try {
    eng.reset(new Engine(device_));
    if (sgf_) eng->attach_sgf_cache(sgf_);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "hipnegf server: worker start failed: %s\n", e.what());
    return;
  }
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    // ---- admission (FCFS, by HBM footprint) ----
    if (!stop_ && !queue_.empty()) {
      Req *r = queue_.front();
      Resident *res = nullptr;
      bool build = false, wait_build = false;
      size_t need = r->footprint;
      if (r->rs.gen) {
        auto it = residents_.find(gen_key(r->rs));
        if (it != residents_.end()) {
          res = it->second.get();
          wait_build = !res->ready;  // another worker is generating it
        } else {
          build = true;
          need += r->struct_bytes;
        }
      }
      std::vector<std::unique_ptr<Resident>> freed;
      if (!wait_build && used_ + need > budget_) evict_idle(need, freed);
      const bool never = need > budget_;
      const bool fits = !wait_build && used_ + need <= budget_;
      if (!freed.empty()) {  // device memory of evicted structures, outside the lock
        lk.unlock();
        for (auto &f : freed) {
          f->st.reset();
          if (f->mem) (void)hipFree(f->mem);
        }
        freed.clear();
        lk.lock();
        continue;  // re-examine the queue
      }
      if (fits || never) {
        queue_.pop_front();
        r->state = ADMITTING;
        if (!never) {
          used_ += need;
          if (build) {
            auto nr = std::make_unique<Resident>();
            nr->bytes = r->struct_bytes;
            res = nr.get();
            residents_[gen_key(r->rs)] = std::move(nr);
          } else if (res) {
            ++struct_hits_;
          }
          if (res) {
            ++res->refs;
            res->last = ++res_clock_;
          }
          r->gst = res;
        }
        lk.unlock();
        std::string err;
        if (never) err = "request does not fit the HBM budget";
        else {
          try {
            admit(*r, *eng, build);
          } catch (const std::exception &e) {
            err = e.what();
          }
        }
        if (!err.empty()) {
          r->res.error = err;
          try { finish(*r, *eng); } catch (...) {}
        }
        lk.lock();
        r->t_admit = clk::now();
        r->res.t_queue = secs(r->t_submit, r->t_admit);
        if (build && r->gst) r->gst->ready = err.empty();
        if (!err.empty()) {
          if (!never) {
            used_ -= r->footprint;
            if (r->gst) {
              --r->gst->refs;
              if (build) {  // drop the structure that failed to build
                auto it = residents_.find(gen_key(r->rs));
                std::unique_ptr<Resident> gone = std::move(it->second);
                residents_.erase(it);
                used_ -= gone->bytes;
                lk.unlock();
                gone->st.reset();
                if (gone->mem) (void)hipFree(gone->mem);
                lk.lock();
              }
            }
          }
          r->state = DONE;
          ++n_completed_;
          cv_done_.notify_all();
        } else {
          r->state = RUNNING;
          running_.push_back(r);
        }
        cv_work_.notify_all();
        continue;
      }
    }
    // ---- next chunk of a running request ----
    Req *best = nullptr;
    for (Req *r : running_) {
      if (r->next >= r->nE) continue;
      if (!best) { best = r; continue; }
      if (o_.policy == 1 ? r->res.t_gpu < best->res.t_gpu : r->id < best->id) best = r;
    }
    if (best) {
      const int i0 = best->next, c = std::min(best->chunk, best->nE - i0);
      best->next += c;
      ++best->inflight;
      lk.unlock();
      const auto t0 = clk::now();
      std::string err;
      try {
        eng->transmission(*best->st, c, best->dE + i0, best->dSl + i0, best->o, best->dT + i0,
                          best->dN + (long long)i0 * best->ncont, best->dS + i0);
      } catch (const std::exception &e) {
        err = e.what();
      }
      const double dt = secs(t0, clk::now());
      lk.lock();
      --best->inflight;
      best->finished += c;
      best->res.t_gpu += dt;
      ++best->res.chunks;
      if (!err.empty()) {
        best->res.error = err;
        best->finished += best->nE - best->next;  // abandon the rest
        best->next = best->nE;
      }
      if (best->finished >= best->nE && best->inflight == 0) {
        running_.erase(std::find(running_.begin(), running_.end(), best));
        lk.unlock();
        try {
          finish(*best, *eng);
        } catch (const std::exception &e) {
          best->res.error = e.what();
        }
        lk.lock();
        used_ -= best->footprint;
        Resident *g = best->gst;
        best->gst = nullptr;
        best->state = DONE;
        ++n_completed_;
        if (g) {
          --g->refs;
          g->last = ++res_clock_;
          if (!o_.keep_structs && g->refs == 0) {  // not kept: free it now
            auto it = residents_.find(gen_key(best->rs));
            std::unique_ptr<Resident> gone = std::move(it->second);
            residents_.erase(it);
            used_ -= gone->bytes;
            lk.unlock();
            gone->st.reset();
            if (gone->mem) (void)hipFree(gone->mem);
            lk.lock();
          }
        }
        cv_done_.notify_all();
        cv_work_.notify_all();
      }
      continue;
    }
    if (stop_) break;
    cv_work_.wait(lk);
  }
}

bool Server::done(unsigned long long id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw std::runtime_error("server: unknown request id");
  return it->second->state == DONE;
}

ReqResult Server::wait(unsigned long long id) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw std::runtime_error("server: unknown request id");
  Req *r = it->second.get();
  cv_done_.wait(lk, [&] { return r->state == DONE || stop_; });
  if (r->state != DONE) throw std::runtime_error("server: stopped before the request finished");
  ReqResult out = std::move(r->res);
  reqs_.erase(it);
  return out;
}

void Server::cache_stats(long long out[4]) {
  std::lock_guard<std::mutex> g(mu_);
  out[0] = struct_hits_;
  out[1] = (long long)residents_.size();
  out[2] = 0;
  out[3] = 0;
  if (sgf_) {
    std::lock_guard<std::mutex> gc(sgf_->mu);
    out[2] = sgf_->hits;
    out[3] = sgf_->stores;
  }
}

void Server::stats(long long out[4]) {
  std::lock_guard<std::mutex> g(mu_);
  out[0] = n_submitted_;
  out[1] = n_completed_;
  out[2] = (long long)running_.size();
  out[3] = (long long)queue_.size();
}

}  // namespace hn
