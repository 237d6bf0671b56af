// Request server (the nrt_server_* entry points): continuous batching of
// independent simulations on one device.
//
// Builder threads take submitted requests in FIFO order once their footprint
// fits in free device memory (hipMemGetInfo, less what other builders are about
// to allocate) and build their device state (assembly, tree plan) on their own
// streams, concurrently with the device work. A loop thread owns one batch; each
// iteration it
//   1. moves built requests into the batch (or the large-request list),
//   2. enqueues steps of every running request: the batched launch for the
//      requests in batch slots (about k * 16384 element-steps each, so a launch
//      lasts about as long whatever the mix of sizes and small requests are not
//      held back by large ones), then k steps of each large request, in one
//      persistent-kernel launch per request
//      (more elements than big_threshold, which would hold a whole batch launch
//      back if one workgroup had to carry it), all on the batch stream,
//   3. waits for the launch and retires the requests it completed (final V and
//      probe trace copied out over the server's own stream; the slot's memory
//      back to the pool, which may let the FIFO head in).
// Latency of a request = completion time - submission time (steady clock).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include "batch.hpp"

namespace nrt {
namespace {

// Implicit slots take steps_per_launch steps per launch at this size and
// proportionally more when smaller (up to 8x), fewer when larger (at least 1).
constexpr long long kBudgetElems = 16384;

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// A submitted problem copied into host memory, with a view for the C ABI.
struct OwnedProblem {
  HostProblem q;
  nrt_problem view{};
  explicit OwnedProblem(const nrt_problem* p) : q(load_problem(p)) {
    view.num_elem = q.N;
    view.dT = q.dT;
    view.potential0 = q.V0.data();
    view.potential1 = q.V1.data();
    view.dx = q.dx.data();
    view.c_m = q.c_m.data();
    view.r_a = q.r_a.data();
    view.W = q.W.data();
    view.K = q.K.data();
    view.input = q.input.data();
    view.mother = q.mother.data();
    view.dr = q.dr.data();
    view.dl = q.dl.data();
    view.branching = q.branching.data();
    view.num_term = q.T;
    view.terminals = q.terminals.data();
    view.num_hh = q.H;
    view.hh_elem = q.hh_elem.data();
    view.hh_m = q.m.data();
    view.hh_h = q.h.data();
    view.hh_n = q.n.data();
    view.E_Na = q.E_Na.data();
    view.E_K = q.E_K.data();
    view.E_L = q.E_L.data();
    view.g_Na = q.g_Na.data();
    view.g_K = q.g_K.data();
    view.G_L = q.G_L.data();
    view.ls_Na = q.ls_Na.data();
    view.ls_K = q.ls_K.data();
    view.rest_pot = q.rest.data();
  }
};

struct Request {
  int64_t id = 0;
  std::unique_ptr<OwnedProblem> prob;
  int64_t num_steps = 0, a0 = 0, a1 = -1, every = -1;
  int32_t trace_elem = 0;
  double t_submit = 0, t_start = 0;
  int N() const { return prob->q.N; }
};

struct Result {
  nrt_request_result info{};
  std::vector<double> V, trace;
};

// A large request on a handle of its own (grid mode), stepped k at a time.
struct BigRun {
  Request req;
  nrt_handle* h = nullptr;
  DBuf<double> trace;
  int64_t done = 0;
  ~BigRun() {
    if (h) handle_destroy(h);
  }
};

}  // namespace
}  // namespace nrt

using namespace nrt;

// A request whose device state has been built by a builder thread.
struct Built {
  Request req;
  nrt_handle* h = nullptr;
  bool big = false;
  int status = NRT_OK;
  std::string err;
};

struct nrt_server {
  nrt_options opt{};
  nrt_server_config cfg{};
  nrt_batch* batch = nullptr;
  std::thread loop;
  std::vector<std::thread> builders;
  std::mutex mu;
  std::condition_variable cv;        // loop thread: submissions, built requests, stop
  std::condition_variable build_cv;  // builders: queue head admissible, stop
  std::condition_variable idle_cv;   // wait_idle
  // shared (guarded by mu)
  std::deque<Request> to_build;  // FIFO of submitted requests
  std::deque<Built> ready;       // built, waiting to join the batch
  int building = 0;              // requests being built
  int64_t reserved = 0;          // footprint of those, not yet visible as allocated
  int big_count = 0;             // large requests built, building or running
  int64_t active = 0;            // requests in the batch or running large
  std::map<int64_t, Result> results;
  std::deque<int64_t> unpolled;
  std::vector<double> latencies;
  nrt_server_stats stats{};
  bool stop = false;
  bool busy = false;
  int64_t next_id = 1;
  std::string last_error, loop_error;
  int loop_status = NRT_OK;
  double t_created = 0;
  // loop thread only
  std::map<int, Request> running;  // batch slot -> request
  std::vector<std::unique_ptr<BigRun>> big;
  hipStream_t side = nullptr;      // result copies of retired slots

  bool is_big(const Request& r) const {
    return opt.kind == NRT_IMPLICIT && r.N() > cfg.big_threshold;
  }

  int64_t footprint(const Request& r) const {
    return is_big(r) ? (int64_t)r.N() * 600 : batch_footprint(batch, r.N(), r.num_steps, r.every);
  }

  void finish_request(Request& r, int status, std::vector<double> V, std::vector<double> trace,
                      int64_t steps_done) {
    Result res;
    res.info.id = r.id;
    res.info.status = status;
    res.info.N = r.N();
    res.info.num_steps = r.num_steps;
    res.info.steps_done = steps_done;
    res.info.submit_s = r.t_submit - t_created;
    res.info.start_s = (r.t_start > 0 ? r.t_start : r.t_submit) - t_created;
    res.info.finish_s = now_s() - t_created;
    res.info.trace_samples = (int64_t)trace.size();
    if (cfg.keep_results) {
      res.V = std::move(V);
      res.trace = std::move(trace);
    }
    std::lock_guard<std::mutex> g(mu);
    latencies.push_back((res.info.finish_s - res.info.submit_s) * 1e3);
    if (status == NRT_OK) {
      stats.completed++;
      stats.compartment_steps += (int64_t)r.N() * steps_done;
    } else {
      stats.failed++;
    }
    unpolled.push_back(r.id);
    results.emplace(r.id, std::move(res));
  }

  // Builder threads: take the FIFO head when its footprint fits in free device
  // memory (less what other builders are about to allocate), build its device
  // state on their own stream, and hand it to the loop. When nothing at all is
  // running or being built, the head is taken anyway, and fails if it does not
  // fit: nothing will ever free memory for it.
  void build_loop() {
    hipStream_t stream = nullptr;
    try {
      NRT_HIP(hipSetDevice(opt.device));
      NRT_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    } catch (const Error& e) {
      std::lock_guard<std::mutex> g(mu);
      loop_status = e.status;
      loop_error = e.what();
      cv.notify_all();
      return;
    }
    while (true) {
      Request r;
      int64_t need = 0;
      bool big_req = false;
      {
        std::unique_lock<std::mutex> g(mu);
        while (true) {
          if (stop && to_build.empty()) {
            if (stream) (void)hipStreamDestroy(stream);
            return;
          }
          if (!to_build.empty()) {
            Request& head = to_build.front();
            big_req = is_big(head);
            need = footprint(head);
            size_t free_b = 0, total_b = 0;
            (void)hipMemGetInfo(&free_b, &total_b);
            const int64_t avail = (int64_t)free_b + pool_cached_bytes(opt.device) - reserved;
            const bool nothing_else = active == 0 && building == 0 && ready.empty();
            const bool fits = avail >= need + (512ll << 20);
            const bool big_ok = !big_req || big_count < cfg.max_big;
            if ((fits && big_ok) || nothing_else) break;
          }
          build_cv.wait(g);
        }
        r = std::move(to_build.front());
        to_build.pop_front();
        building++;
        reserved += need;
        if (big_req) big_count++;
      }
      Built b;
      b.big = big_req;
      try {
        if (big_req) {
          NRT_REQUIRE(cfg.max_big > 0, NRT_ERR_ADMISSION, "large requests are disabled");
          nrt_options o = opt;
          o.num_blocks = 0;  // grid mode for a tree this size
          b.h = handle_create(&r.prob->view, o, stream, 0);
        } else {
          batch_admission_check(batch, r.N(), r.num_steps, r.every, 0);
          b.h = handle_create(&r.prob->view, batch_slot_options(batch, r.N()), stream,
                              batch_plan_threads(batch));
        }
      } catch (const Error& e) {
        b.status = e.status;
        b.err = e.what();
      } catch (const std::exception& e) {
        b.status = NRT_ERR_INVALID;
        b.err = e.what();
      }
      b.req = std::move(r);
      std::lock_guard<std::mutex> g(mu);
      building--;
      reserved -= need;
      ready.push_back(std::move(b));
      cv.notify_all();
    }
  }

  // Loop thread: built requests join the batch (or the large-request list).
  void join_ready(std::deque<Built>& got) {
    for (Built& b : got) {
      Request& r = b.req;
      if (b.status == NRT_OK) {
        r.t_start = now_s();
        try {
          if (b.big) {
            auto br = std::make_unique<BigRun>();
            br->h = b.h;  // owned by br from here on
            b.h = nullptr;
            br->h->stream = batch_stream(batch);  // stepped in the loop's rhythm
            if (r.every >= 0 && r.num_steps / (r.every + 1) > 0)
              br->trace.alloc((size_t)(r.num_steps / (r.every + 1)));
            br->req = std::move(r);
            big.push_back(std::move(br));
          } else {
            nrt_handle* h = b.h;
            b.h = nullptr;  // owned by the batch from here on
            const int slot =
                batch_add_handle(batch, h, r.num_steps, r.a0, r.a1, r.trace_elem, r.every);
            running.emplace(slot, std::move(r));
          }
          std::lock_guard<std::mutex> g(mu);
          active++;
          continue;
        } catch (const Error& e) {
          b.status = e.status;
          b.err = e.what();
        }
      }
      if (b.big) {
        std::lock_guard<std::mutex> g(mu);
        big_count--;
      }
      finish_request(b.req, b.status, {}, {}, 0);
    }
    got.clear();
  }

  void launch_big(long long k) {
    for (auto& b : big) {
      const int64_t n = std::min<int64_t>(k, b->req.num_steps - b->done);
      if (n <= 0) continue;
      const Request& r = b->req;
      const int64_t every = r.every;
      double* tr = nullptr;
      int64_t gl0 = 0;
      if (every >= 0 && b->trace.p) {
        tr = b->trace.p + b->done / (every + 1);
        gl0 = b->done % (every + 1);
      }
      implicit_steps_async(b->h, b->done + 1, n, r.a0, r.a1, r.trace_elem, every, gl0, tr);
      b->done += n;
    }
  }

  // Completed batch slots of the launch just finished: their results come over the
  // server's own stream and their memory goes back to the pool.
  void retire_slots(std::vector<int>& done) {
    for (int slot : done) {
      auto it = running.find(slot);
      if (it == running.end()) continue;
      Request& r = it->second;
      std::vector<double> V, tr;
      int status = NRT_OK;
      int64_t steps = 0;
      batch_slot_collect(batch, slot, side, cfg.keep_results ? &V : nullptr,
                         cfg.keep_results ? &tr : nullptr, &status, &steps);
      finish_request(r, status, std::move(V), std::move(tr), steps);
      batch_slot_drop(batch, slot);
      running.erase(it);
    }
    if (!done.empty()) {
      std::lock_guard<std::mutex> g(mu);
      active -= (int64_t)done.size();
      build_cv.notify_all();  // memory freed: the FIFO head may fit now
    }
    done.clear();
  }

  // Large requests run on the batch stream; after the launch has finished, the
  // completed ones are retired in place.
  void retire_big() {
    int freed = 0;
    for (auto it = big.begin(); it != big.end();) {
      BigRun& b = **it;
      if (b.done < b.req.num_steps) {
        ++it;
        continue;
      }
      std::vector<double> V, tr;
      if (cfg.keep_results) {
        V.resize(b.req.N());
        implicit_snapshot(b.h, V.data());
        tr.resize(b.trace.n);
        if (b.trace.n)
          NRT_HIP(hipMemcpy(tr.data(), b.trace.p, sizeof(double) * b.trace.n,
                            hipMemcpyDeviceToHost));
      }
      int status = NRT_OK;
      int64_t inner = 0;
      implicit_slot_stats(b.h, &status, &inner);
      finish_request(b.req, status, std::move(V), std::move(tr), b.done);
      it = big.erase(it);
      freed++;
    }
    if (freed) {
      std::lock_guard<std::mutex> g(mu);
      big_count -= freed;
      active -= freed;
      build_cv.notify_all();
    }
  }

  void run() {
    try {
      NRT_HIP(hipSetDevice(opt.device));
      NRT_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
      std::deque<Built> got;
      std::vector<int> done;  // batch slots completed by the launch just finished
      while (true) {
        {
          std::unique_lock<std::mutex> g(mu);
          const bool has_work = !running.empty() || !big.empty();
          if (!has_work && ready.empty()) {
            if (to_build.empty() && building == 0) {
              busy = false;
              idle_cv.notify_all();
              if (stop) return;
            }
            cv.wait(g, [&] { return stop || !ready.empty() || loop_status != NRT_OK; });
            if (loop_status != NRT_OK) return;
            if (stop && ready.empty() && to_build.empty() && building == 0) return;
            if (ready.empty()) continue;
          }
          busy = true;
          got.swap(ready);
          stats.queued = (int64_t)to_build.size() + building;
          stats.running = (int64_t)(running.size() + big.size() + got.size());
        }
        const double t0 = now_s();
        join_ready(got);
        const long long k = cfg.steps_per_launch;
        const bool stepping = !running.empty() || !big.empty();
        if (stepping) {
          if (opt.kind == NRT_IMPLICIT) batch_launch_budget(batch, k, k * kBudgetElems, 8 * k);
          else batch_launch(batch, k);
          launch_big(k);
        }
        const double t1 = now_s();
        if (stepping) {
          batch_finish(batch);
          NRT_HIP(hipStreamSynchronize(batch_stream(batch)));
        }
        const double t2 = now_s();
        // The slots this launch completed are retired at once, before the next
        // launch: copies on another stream while a launch runs were found to wait
        // for it anyway (streams share the process's few hardware queues), so an
        // overlap saved nothing and delayed each result by a launch.
        for (auto& kv : running)
          if (batch_slot_complete(batch, kv.first)) done.push_back(kv.first);
        retire_slots(done);
        retire_big();
        const double t3 = now_s();
        std::lock_guard<std::mutex> g(mu);
        if (stepping) stats.launches++;
        stats.launch_ms += (t1 - t0) * 1e3;
        stats.retire_ms += (t3 - t2) * 1e3;
        stats.wait_ms += (t2 - t1) * 1e3;
      }
    } catch (const Error& e) {
      fail(e.status, e.what());
    } catch (const std::exception& e) {
      fail(NRT_ERR_HIP, e.what());
    }
  }

  void fail(int status, const std::string& what) {
    std::lock_guard<std::mutex> g(mu);
    loop_status = status;
    loop_error = what;
    busy = false;
    idle_cv.notify_all();
    build_cv.notify_all();
  }
};

namespace {
template <class F>
int guarded_s(nrt_server* s, F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    if (s) s->last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    if (s) s->last_error = "host allocation failed";
    return NRT_ERR_NOMEM;
  } catch (const std::exception& e) {
    if (s) s->last_error = e.what();
    return NRT_ERR_INVALID;
  }
}
}  // namespace

extern "C" {

void nrt_default_server_config(nrt_server_config* c) {
  if (!c) return;
  memset(c, 0, sizeof(*c));
  c->steps_per_launch = 16;
  c->big_threshold = 1 << 30;  // large slots run as virtual grids inside the batch
  c->max_big = 4;
  c->keep_results = 1;
  c->builders = 4;
}

int nrt_server_create(const nrt_options* opt, const nrt_server_config* cfg, nrt_server** out) {
  if (!out) return NRT_ERR_INVALID;
  *out = nullptr;
  auto* s = new nrt_server;
  int st = guarded_s(s, [&] {
    if (cfg) s->cfg = *cfg;
    else nrt_default_server_config(&s->cfg);
    NRT_REQUIRE(s->cfg.steps_per_launch >= 1 && s->cfg.max_big >= 0 &&
                    s->cfg.big_threshold >= 3 && s->cfg.builders >= 1 && s->cfg.builders <= 16,
                NRT_ERR_INVALID, "bad server configuration");
    s->opt = resolve_options(opt);
    s->opt.device = pick_device(s->opt.device);
    const int bst = nrt_batch_create(&s->opt, &s->batch);
    if (bst != NRT_OK) throw Error(bst, "batch: " + std::string(nrt_last_error(nullptr)));
    s->t_created = now_s();
    s->loop = std::thread([s] { s->run(); });
    for (int i = 0; i < s->cfg.builders; i++) s->builders.emplace_back([s] { s->build_loop(); });
    return NRT_OK;
  });
  if (st != NRT_OK) {
    if (s->batch) nrt_batch_destroy(s->batch);
    delete s;
    return st;
  }
  *out = s;
  return NRT_OK;
}

int nrt_server_submit(nrt_server* s, const nrt_problem* p, int64_t num_steps, int64_t a0,
                      int64_t a1, int32_t trace_elem, int64_t trace_every, int64_t* id) {
  if (!s) return NRT_ERR_INVALID;
  return guarded_s(s, [&] {
    NRT_REQUIRE(p, NRT_ERR_INVALID, "problem is NULL");
    NRT_REQUIRE(num_steps >= 0, NRT_ERR_INVALID, "num_steps < 0");
    NRT_REQUIRE(trace_every < 0 || (trace_elem >= 0 && trace_elem < p->num_elem),
                NRT_ERR_INVALID, "trace_elem out of range");
    Request r;
    r.prob = std::make_unique<OwnedProblem>(p);
    r.num_steps = num_steps;
    r.a0 = a0;
    r.a1 = a1;
    r.trace_elem = trace_elem;
    r.every = trace_every;
    r.t_submit = now_s();
    std::lock_guard<std::mutex> g(s->mu);
    NRT_REQUIRE(s->loop_status == NRT_OK, s->loop_status, "server stopped: " + s->loop_error);
    r.id = s->next_id++;
    if (id) *id = r.id;
    s->stats.submitted++;
    s->busy = true;
    s->to_build.push_back(std::move(r));
    s->build_cv.notify_all();
    return NRT_OK;
  });
}

int nrt_server_poll(nrt_server* s, nrt_request_result* out, int32_t capacity, int32_t* count) {
  if (!s) return NRT_ERR_INVALID;
  return guarded_s(s, [&] {
    NRT_REQUIRE(count && (capacity <= 0 || out), NRT_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> g(s->mu);
    int n = 0;
    while (n < capacity && !s->unpolled.empty()) {
      out[n++] = s->results.at(s->unpolled.front()).info;
      s->unpolled.pop_front();
    }
    *count = n;
    return s->loop_status;
  });
}

int nrt_server_fetch(nrt_server* s, int64_t id, double* v_out, double* trace_out,
                     int64_t trace_capacity) {
  if (!s) return NRT_ERR_INVALID;
  return guarded_s(s, [&] {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->results.find(id);
    NRT_REQUIRE(it != s->results.end(), NRT_ERR_INVALID,
                "request " + std::to_string(id) + " has no result (pending or fetched)");
    const Result& r = it->second;
    if (v_out && !r.V.empty()) memcpy(v_out, r.V.data(), sizeof(double) * r.V.size());
    if (trace_out && trace_capacity > 0)
      memcpy(trace_out, r.trace.data(),
             sizeof(double) * std::min<size_t>(r.trace.size(), (size_t)trace_capacity));
    const int status = r.info.status;
    s->results.erase(it);
    return status;
  });
}

int nrt_server_wait_idle(nrt_server* s, double timeout_s) {
  if (!s) return NRT_ERR_INVALID;
  return guarded_s(s, [&] {
    std::unique_lock<std::mutex> g(s->mu);
    const auto pred = [&] { return !s->busy || s->loop_status != NRT_OK; };
    if (timeout_s < 0) s->idle_cv.wait(g, pred);
    else if (!s->idle_cv.wait_for(g, std::chrono::duration<double>(timeout_s), pred))
      throw Error(NRT_ERR_NOT_CONVERGED, "server still busy after the timeout");
    return s->loop_status;
  });
}

int nrt_server_get_stats(nrt_server* s, nrt_server_stats* out) {
  if (!s) return NRT_ERR_INVALID;
  return guarded_s(s, [&] {
    NRT_REQUIRE(out, NRT_ERR_INVALID, "out is NULL");
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->stats;
    out->wall_s = now_s() - s->t_created;
    std::vector<double> l = s->latencies;
    if (!l.empty()) {
      std::sort(l.begin(), l.end());
      auto pct = [&](double q) {
        const double pos = q * (double)(l.size() - 1);
        const size_t i = (size_t)pos;
        const double f = pos - (double)i;
        return i + 1 < l.size() ? l[i] * (1 - f) + l[i + 1] * f : l[i];
      };
      out->p50_ms = pct(0.50);
      out->p99_ms = pct(0.99);
      double sum = 0;
      for (double x : l) sum += x;
      out->mean_ms = sum / (double)l.size();
    }
    nrt_batch_stats bs{};
    if (nrt_batch_get_stats(s->batch, &bs) == NRT_OK) out->batch_kernel_ms = bs.kernel_ms;
    return NRT_OK;
  });
}

const char* nrt_server_last_error(const nrt_server* s) { return s ? s->last_error.c_str() : ""; }

int nrt_server_destroy(nrt_server* s) {
  if (!s) return NRT_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv.notify_all();
    s->build_cv.notify_all();
  }
  for (auto& t : s->builders)
    if (t.joinable()) t.join();
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->cv.notify_all();
  }
  if (s->loop.joinable()) s->loop.join();
  for (Built& b : s->ready)
    if (b.h) handle_destroy(b.h);
  s->ready.clear();
  s->running.clear();
  if (s->side) (void)hipStreamDestroy(s->side);
  s->big.clear();
  if (s->batch) nrt_batch_destroy(s->batch);
  delete s;
  return NRT_OK;
}

}  // extern "C"
