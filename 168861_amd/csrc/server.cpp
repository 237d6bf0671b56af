// Request server (the nrt_server_* entry points): continuous batching of
// independent simulations on one device.
//
// A loop thread owns one batch. Each iteration it
//   1. takes newly submitted requests into its FIFO,
//   2. enqueues k steps of every running request: the batched launch for the
//      requests in batch slots, then one cooperative launch per large request
//      (more elements than big_threshold, which would hold a whole batch launch
//      back if one workgroup had to carry it), all on the batch stream,
//   3. admits waiting requests while that work runs: slot construction (device
//      assembly, the tree plan) overlaps the launch; admission stops at the first
//      request whose footprint does not fit in free device memory,
//   4. waits for the launch, then retires completed requests: final V and the
//      probe trace are copied out and the slot is freed.
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

struct nrt_server {
  nrt_options opt{};
  nrt_server_config cfg{};
  nrt_batch* batch = nullptr;
  std::thread loop;
  std::mutex mu;
  std::condition_variable cv, idle_cv;
  // shared with the callers (guarded by mu)
  std::deque<Request> inbox;
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
  std::deque<Request> waiting;
  std::map<int, Request> running;  // batch slot -> request
  std::vector<std::unique_ptr<BigRun>> big;

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

  // Admission in FIFO order; returns when the head does not fit.
  void admit() {
    while (!waiting.empty()) {
      Request& r = waiting.front();
      const bool large = opt.kind == NRT_IMPLICIT && r.N() > cfg.big_threshold;
      if (large) {
        if ((int)big.size() >= cfg.max_big) return;
        auto br = std::make_unique<BigRun>();
        try {
          size_t free_b = 0, total_b = 0;
          NRT_HIP(hipMemGetInfo(&free_b, &total_b));
          NRT_REQUIRE((int64_t)free_b > (int64_t)r.N() * 600 + (512ll << 20), NRT_ERR_ADMISSION,
                      "not enough free device memory");
          nrt_options o = opt;
          o.num_blocks = 0;  // grid mode for a tree this size
          br->h = handle_create(&r.prob->view, o, batch_stream(batch), 0);
          if (r.every >= 0 && r.num_steps / (r.every + 1) > 0)
            br->trace.alloc((size_t)(r.num_steps / (r.every + 1)));
        } catch (const Error& e) {
          if (e.status == NRT_ERR_ADMISSION && (!running.empty() || !big.empty())) return;
          finish_request(r, e.status, {}, {}, 0);
          waiting.pop_front();
          continue;
        }
        r.t_start = now_s();
        br->req = std::move(r);
        waiting.pop_front();
        big.push_back(std::move(br));
        continue;
      }
      int32_t slot = -1;
      const int st = nrt_batch_add(batch, &r.prob->view, r.num_steps, r.a0, r.a1, r.trace_elem,
                                   r.every, &slot);
      if (st == NRT_ERR_ADMISSION && (!running.empty() || !big.empty())) return;  // wait
      if (st != NRT_OK) {
        finish_request(r, st, {}, {}, 0);
        waiting.pop_front();
        continue;
      }
      r.t_start = now_s();
      running.emplace(slot, std::move(r));
      waiting.pop_front();
    }
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

  void retire() {
    for (auto it = running.begin(); it != running.end();) {
      if (!batch_slot_complete(batch, it->first)) {
        ++it;
        continue;
      }
      nrt_slot_info si{};
      if (nrt_batch_slot_info(batch, it->first, &si) != NRT_OK)
        throw Error(NRT_ERR_HIP, nrt_batch_last_error(batch));
      Request& r = it->second;
      std::vector<double> V, tr;
      if (cfg.keep_results) {
        V.resize(r.N());
        if (nrt_batch_snapshot(batch, it->first, V.data()) != NRT_OK)
          throw Error(NRT_ERR_HIP, nrt_batch_last_error(batch));
        int64_t n = 0;
        nrt_batch_trace(batch, it->first, nullptr, 0, &n);
        tr.resize((size_t)n);
        if (n > 0) nrt_batch_trace(batch, it->first, tr.data(), n, &n);
      }
      finish_request(r, si.status, std::move(V), std::move(tr), si.steps_done);
      nrt_batch_remove(batch, it->first);
      it = running.erase(it);
    }
    for (auto it = big.begin(); it != big.end();) {
      BigRun& b = **it;
      if (b.done < b.req.num_steps) {
        ++it;
        continue;
      }
      NRT_HIP(hipStreamSynchronize(b.h->stream));
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
    }
  }

  void run() {
    try {
      NRT_HIP(hipSetDevice(opt.device));
      while (true) {
        {
          std::unique_lock<std::mutex> g(mu);
          const bool idle = waiting.empty() && running.empty() && big.empty();
          if (idle) {
            busy = false;
            idle_cv.notify_all();
            cv.wait(g, [&] { return stop || !inbox.empty(); });
            if (stop && inbox.empty()) return;
          }
          busy = true;
          while (!inbox.empty()) {
            waiting.push_back(std::move(inbox.front()));
            inbox.pop_front();
          }
          stats.queued = (int64_t)waiting.size();
          stats.running = (int64_t)(running.size() + big.size());
        }
        const long long k = cfg.steps_per_launch;
        const bool work = !running.empty() || !big.empty();
        if (work) {
          batch_launch(batch, k);
          launch_big(k);
        }
        admit();  // overlaps the launch
        if (work) batch_finish(batch);
        NRT_HIP(hipStreamSynchronize(batch_stream(batch)));
        if (work) {
          std::lock_guard<std::mutex> g(mu);
          stats.launches++;
        }
        retire();
        if (!work && running.empty() && big.empty() && !waiting.empty()) {
          // nothing runs and the head still does not fit: it never will
          Request& r = waiting.front();
          finish_request(r, NRT_ERR_ADMISSION, {}, {}, 0);
          waiting.pop_front();
        }
      }
    } catch (const Error& e) {
      std::lock_guard<std::mutex> g(mu);
      loop_status = e.status;
      loop_error = e.what();
      busy = false;
      idle_cv.notify_all();
    } catch (const std::exception& e) {
      std::lock_guard<std::mutex> g(mu);
      loop_status = NRT_ERR_HIP;
      loop_error = e.what();
      busy = false;
      idle_cv.notify_all();
    }
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
  c->big_threshold = 1 << 17;
  c->max_big = 4;
  c->keep_results = 1;
}

int nrt_server_create(const nrt_options* opt, const nrt_server_config* cfg, nrt_server** out) {
  if (!out) return NRT_ERR_INVALID;
  *out = nullptr;
  auto* s = new nrt_server;
  int st = guarded_s(s, [&] {
    if (cfg) s->cfg = *cfg;
    else nrt_default_server_config(&s->cfg);
    NRT_REQUIRE(s->cfg.steps_per_launch >= 1 && s->cfg.max_big >= 0 && s->cfg.big_threshold >= 3,
                NRT_ERR_INVALID, "bad server configuration");
    s->opt = resolve_options(opt);
    s->opt.device = pick_device(s->opt.device);
    const int bst = nrt_batch_create(&s->opt, &s->batch);
    if (bst != NRT_OK) throw Error(bst, "batch: " + std::string(nrt_last_error(nullptr)));
    s->t_created = now_s();
    s->loop = std::thread([s] { s->run(); });
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
    s->inbox.push_back(std::move(r));
    s->cv.notify_one();
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
    const auto pred = [&] { return (!s->busy && s->inbox.empty()) || s->loop_status != NRT_OK; };
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
  }
  if (s->loop.joinable()) s->loop.join();
  s->running.clear();
  s->big.clear();
  if (s->batch) nrt_batch_destroy(s->batch);
  delete s;
  return NRT_OK;
}

}  // extern "C"
