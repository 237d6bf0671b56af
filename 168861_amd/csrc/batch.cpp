// Batches of independent neurites (the nrt_batch_* entry points). Each slot is a
// complete solver state (a handle without a stream of its own); the batch keeps a
// device table of slot descriptors that is rebuilt only when membership changes,
// and advances all slots with one launch per call (implicit, persistent over the
// k steps) or one graph of k steps (explicit).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include "batch.hpp"

namespace nrt {
// implicit.hip
size_t implicit_slot_bytes(int precision);
void implicit_slot_fill(nrt_handle* h, long long a0, long long a1, int trace_elem,
                        long long interval, double* trace, void* out);
size_t implicit_wgrec_bytes(int precision, int nwg);
void implicit_batch_expand(int precision, const void* table, const void* work,
                           const int2* wgmap, int nwg, void* wgrec, hipStream_t s);
void implicit_batch_launch(int precision, int krylov, int threads, const void* wgrec, int nwg,
                           hipStream_t s);
size_t implicit_slot_work_bytes();
void implicit_slot_work(void* out, long long done, long long steps, int vn);
int implicit_batch_capacity(int precision, int krylov, int threads);
void implicit_slot_stats(nrt_handle* h, int* status, int64_t* inner);
void implicit_slot_stats_on(nrt_handle* h, hipStream_t s, int* status, int64_t* inner);
void implicit_snapshot_on(nrt_handle* hnd, double* host_out, hipStream_t s);
int64_t implicit_device_bytes(const Implicit* e);
int64_t implicit_nnz(const Implicit* e);
// explicit.hip
size_t explicit_slot_bytes();
void explicit_slot_fill(nrt_handle* h, long long s0, long long num_steps, long long a0,
                        long long a1, int trace_elem, long long interval, double* trace,
                        void* out);
void explicit_batch_steps(const void* slots, const int2* blk, int nblk, int nslots,
                          const long long* gbase, int k, hipStream_t s);
const double* explicit_slot_v(nrt_handle* h, long long done);
int64_t explicit_device_bytes(const Explicit* e);
int explicit_batch_block_elems();

namespace {
// Workgroup sizes the autotuner tries, in this order, for each size class of slots.
constexpr int kTuneThreads[] = {256, 512, 1024, 128};
constexpr int kTuneCount = (int)(sizeof(kTuneThreads) / sizeof(int));
// Size classes (rows): each class runs as a launch of its own with its own
// workgroup size, so small trees and large ones each get the size that serves them
// best (rows per worker differ by an order of magnitude across a sweep).
constexpr int kClasses = 3;
inline int size_class(int N) { return N < 4096 ? 0 : (N < 16384 ? 1 : 2); }
// workers assumed by a slot's chunk choice: the size the workgroup autotuner
// picks for the FP64 sweep (4,096 trees: planned for 512, 2.16 ms per step; for
// 256, 2.06; with level-0 chunks of at most 16 rows, 2.00)
constexpr int kPlanThreads = 256;
// Footprint per element used for admission (bytes; the arrays a slot allocates,
// rounded up): implicit tree solve ~ 7 FP64 + 3 int32 vectors, 3 CSR entries, the
// plan and 8 tree work values over ~1.1 N level nodes; explicit ~ 6 FP64 + code.
constexpr int64_t kImplicitBytesPerElem = 420;
constexpr int64_t kExplicitBytesPerElem = 64;
constexpr int64_t kAdmissionReserve = 512ll << 20;  // left free for the runtime

struct Slot;
}  // namespace
int64_t batch_footprint(const nrt_batch* b, int N, int64_t num_steps, int64_t trace_every);
void batch_admission_check(const nrt_batch* b, int N, int64_t num_steps, int64_t trace_every,
                           int64_t reserved);
int batch_plan_threads(const nrt_batch* b);
int batch_add_handle(nrt_batch* b, nrt_handle* h, int64_t num_steps, int64_t a0, int64_t a1,
                     int32_t trace_elem, int64_t trace_every);
namespace {
struct Slot {
  nrt_handle* h = nullptr;
  int N = 0;
  long long s0 = 0;       // batch step at which it joined (explicit: its step = g - s0)
  long long done = 0;     // steps taken, counting the launch in flight
  long long num_steps = 0, a0 = 0, a1 = -1, interval = -1;
  int trace_elem = 0;
  DBuf<double> trace;
  int64_t bytes = 0;
  ~Slot() {
    if (h) handle_destroy(h);
  }
  long long left() const { return num_steps - done; }
};
}  // namespace
}  // namespace nrt

using namespace nrt;

struct nrt_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t adm = nullptr;  // slot construction, overlapping the step launches
  nrt_options opt{};
  std::vector<std::unique_ptr<Slot>> slots;  // index = slot id; null once removed
  long long g = 0;                           // batch steps done
  bool dirty = true;
  std::vector<int> order;  // slot ids in table order
  DBuf<uint8_t> table;
  DBuf<int2> blk;
  int nblk = 0;
  DBuf<uint8_t> work;  // implicit: steps and workgroups per table entry
  std::vector<uint8_t> work_h;
  DBuf<int2> wgmap;    // implicit: (table entry, index in its virtual grid) per workgroup
  DBuf<uint8_t> wgrec;  // implicit: each workgroup's slot record for the launch
  std::vector<int2> wgmap_h;
  std::vector<int> vn;  // workgroups per table entry in the launch being built
  std::map<int, int> capacity;  // resident workgroups per workgroup size
  DBuf<long long> gdev;  // explicit: the first batch step of the call being replayed
  long long gbase_h = 0;  // its host source (kept until the launch is finished)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long long graph_k = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool pending = false;  // a launch is in flight (finish() accounts for it)
  int64_t pend_cs = 0, pend_ss = 0, pend_launches = 0;
  int pend_threads = 0;
  // implicit: per size class, the workgroup size (0: still tuning), the tuning
  // record (threads, ms per compartment-step), the launch in flight and its events
  int cthreads[kClasses] = {0, 0, 0};
  std::vector<std::pair<int, double>> ctune[kClasses];
  // a class's first launch also fills and factors its slots (later launches reuse
  // the factors), so it is not a timing sample
  bool cwarm[kClasses] = {false, false, false};
  // implicit, set by nrt_batch_tune: chunk lengths new slots of a class are planned
  // with (0: the options' value / auto rule)
  int cchunk[kClasses] = {0, 0, 0}, cchunk_hi[kClasses] = {0, 0, 0};
  struct ClassLaunch {
    bool on;
    int th;
    int64_t cs;
  } pend_cls[kClasses] = {};
  hipEvent_t cev[kClasses][2] = {};
  hipStream_t side[kClasses] = {};  // streams of the class launches besides `stream`
  nrt_batch_stats stats{};
  std::string last_error;

  bool is_implicit() const { return opt.kind == NRT_IMPLICIT; }

  // Options a new slot of N rows is planned with.
  nrt_options slot_options(int N) const {
    nrt_options o = opt;
    const int c = size_class(N);
    if (is_implicit() && cchunk[c] > 0) {
      o.chunk = cchunk[c];
      o.chunk_hi = cchunk_hi[c];
    }
    return o;
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    graph_k = 0;
  }

  Slot& slot(int id) {
    NRT_REQUIRE(id >= 0 && id < (int)slots.size() && slots[id], NRT_ERR_INVALID,
                "no slot " + std::to_string(id));
    return *slots[id];
  }

  // Descriptor table of the slots that still have steps left.
  void rebuild() {
    order.clear();
    for (int id = 0; id < (int)slots.size(); id++)
      if (slots[id] && slots[id]->left() > 0) order.push_back(id);
    drop_graph();
    if (is_implicit()) {
      // largest trees first: the long workgroups start early (LPT order)
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int b) { return slots[a]->N > slots[b]->N; });
      const size_t sb = implicit_slot_bytes(opt.precision);
      std::vector<uint8_t> host(sb * order.size());
      for (size_t k = 0; k < order.size(); k++) {
        Slot& s = *slots[order[k]];
        implicit_slot_fill(s.h, s.a0, s.a1, s.trace_elem, s.interval, s.trace.p,
                           host.data() + k * sb);
      }
      table.upload(host, stream);
      work.alloc(order.size() * implicit_slot_work_bytes());
      work_h.resize(order.size() * implicit_slot_work_bytes());
      vn.assign(order.size(), 1);
    } else {
      const size_t sb = explicit_slot_bytes();
      std::vector<uint8_t> host(sb * order.size());
      std::vector<int2> bl;
      for (size_t k = 0; k < order.size(); k++) {
        Slot& s = *slots[order[k]];
        explicit_slot_fill(s.h, s.s0, s.num_steps, s.a0, s.a1, s.trace_elem, s.interval,
                           s.trace.p, host.data() + k * sb);
        for (int e0 = 0; e0 < s.N; e0 += explicit_batch_block_elems()) bl.push_back(make_int2((int)k, e0));
      }
      table.upload(host, stream);
      blk.upload(bl, stream);
      nblk = (int)bl.size();
      if (!gdev.p) gdev.alloc(1);
    }
    NRT_HIP(hipStreamSynchronize(stream));
    dirty = false;
  }

  int pick_threads(int c) {
    if (opt.batch_threads > 0) return opt.batch_threads;
    if (cthreads[c] > 0) return cthreads[c];
    return kTuneThreads[ctune[c].size()];
  }

  // Enqueues steps of every active slot and returns: k each (lockstep), or, with
  // elem_budget > 0 (implicit only), max(1, min(kmax, elem_budget / N)) each, so
  // that one launch takes about the same time whatever the sizes and small
  // slots do not wait on large ones. Slots added before finish() join after it.
  void launch(long long k, long long elem_budget = 0, long long kmax = 0) {
    NRT_REQUIRE(k >= 0, NRT_ERR_INVALID, "k < 0");
    NRT_REQUIRE(!pending, NRT_ERR_INVALID, "a launch is already in flight");
    NRT_REQUIRE(elem_budget <= 0 || is_implicit(), NRT_ERR_UNSUPPORTED,
                "explicit batches step in lockstep");
    if (dirty) rebuild();
    g += k;
    if (order.empty() || k == 0) return;
    int th[kClasses] = {0, 0, 0};
    std::vector<int2> seg;  // wgmap range of each class
    bool merged = false;    // one launch for all classes
    int dom = 0;            // the class holding the most rows
    if (is_implicit()) {
      int64_t rows[kClasses] = {0, 0, 0};
      for (int id : order) rows[size_class(slots[id]->N)] += slots[id]->N;
      for (int c = 1; c < kClasses; c++)
        if (rows[c] > rows[dom]) dom = c;
      for (int c = 0; c < kClasses; c++) th[c] = pick_threads(c);
      seg = plan_workgroups(th);
      // Classes run as launches of their own only while each fills at least half
      // the device's co-resident workgroups (a large sweep); otherwise (a few
      // requests of a server, say) one launch, with the dominant class's size.
      for (int c = 0; c < kClasses; c++) {
        const int n = seg[c].y - seg[c].x;
        if (n > 0 && n < capacity.at(th[c]) / 2) merged = true;
      }
      if (merged) {
        for (int c = 0; c < kClasses; c++) th[c] = th[dom];
        seg = plan_workgroups(th);
      }
    }
    int64_t cs = 0, ss = 0, ccs[kClasses] = {0, 0, 0};
    const size_t wb = implicit_slot_work_bytes();
    for (size_t i = 0; i < order.size(); i++) {
      Slot& s = *slots[order[i]];
      long long want = k;
      if (elem_budget > 0) {
        const long long v = is_implicit() ? vn[i] : 1;
        want = std::max(1ll, std::min(kmax > 0 ? kmax : k, elem_budget * v / std::max(1, s.N)));
      }
      const long long adv = std::min(want, s.left());
      if (is_implicit()) implicit_slot_work(work_h.data() + i * wb, s.done, adv, vn[i]);
      s.done += adv;
      ss += adv;
      cs += adv * s.N;
      ccs[size_class(s.N)] += adv * s.N;
    }
    if (is_implicit()) {
      NRT_HIP(hipMemcpyAsync(work.p, work_h.data(), work_h.size(), hipMemcpyHostToDevice,
                             stream));
      wgmap.upload(wgmap_h, stream);
      const int nwg = (int)wgmap_h.size();
      const size_t rb = implicit_wgrec_bytes(opt.precision, nwg);
      if (wgrec.n < rb) wgrec.alloc(rb);
      implicit_batch_expand(opt.precision, table.p, work.p, wgmap.p, nwg, wgrec.p, stream);
      const size_t rec1 = implicit_wgrec_bytes(opt.precision, 1);
      // One launch per size class present, running concurrently: the largest class
      // on the batch stream, the others on side streams forked from it and joined
      // back, so that no launch waits for another's tail. A slot spanning several
      // workgroups still finds them all resident: its launch alone fits the
      // co-resident limit, and the other launches' workgroups wait on nothing.
      pend_launches = 0;
      int64_t best = -1;
      NRT_HIP(hipEventRecord(ev0, stream));
      hipStream_t on = nullptr;
      if (merged) {  // the whole wgmap as the dominant class's launch
        seg.assign(kClasses, make_int2(0, 0));
        seg[dom] = make_int2(0, nwg);
        for (int c = 0; c < kClasses; c++)
          if (c != dom) {
            ccs[dom] += ccs[c];
            ccs[c] = 0;
          }
      }
      for (int c = kClasses - 1; c >= 0; c--) {
        pend_cls[c] = ClassLaunch{false, th[c], ccs[c]};
        const int n = seg[c].y - seg[c].x;
        if (n <= 0) continue;
        const hipStream_t cs_ = on == nullptr ? stream : side[c];
        if (cs_ != stream) NRT_HIP(hipStreamWaitEvent(cs_, ev0, 0));
        on = cs_;
        NRT_HIP(hipEventRecord(cev[c][0], cs_));
        implicit_batch_launch(opt.precision, opt.krylov, th[c], wgrec.p + rec1 * seg[c].x, n,
                              cs_);
        NRT_HIP(hipEventRecord(cev[c][1], cs_));
        pend_cls[c].on = true;
        pend_launches++;
        if (ccs[c] > best) {
          best = ccs[c];
          pend_threads = th[c];
        }
      }
      for (int c = 0; c < kClasses; c++)
        if (pend_cls[c].on) NRT_HIP(hipStreamWaitEvent(stream, cev[c][1], 0));
      NRT_HIP(hipEventRecord(ev1, stream));
    } else {
      NRT_REQUIRE(k <= (1ll << 30), NRT_ERR_INVALID, "k too large for one call");
      if (!gexec || graph_k != k) {
        drop_graph();
        NRT_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        explicit_batch_steps(table.p, blk.p, nblk, (int)order.size(), gdev.p, (int)k, stream);
        NRT_HIP(hipStreamEndCapture(stream, &graph));
        NRT_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        graph_k = k;
      }
      // the first batch step of this call (g already counts its k steps); the graph's
      // launches take it from device memory, set here before each replay
      gbase_h = g - k;
      NRT_HIP(hipMemcpyAsync(gdev.p, &gbase_h, sizeof(long long), hipMemcpyHostToDevice, stream));
      NRT_HIP(hipEventRecord(ev0, stream));
      NRT_HIP(hipGraphLaunch(gexec, stream));
      NRT_HIP(hipEventRecord(ev1, stream));
      pend_launches = k;
    }
    pend_cs = cs;
    pend_ss = ss;
    pending = true;
  }

  // Workgroups per slot: one per kRowsPerWorkgroup rows (at most kMaxSlotBlocks,
  // at most opt.num_blocks when set), and only while every workgroup of the launch
  // can be resident at once -- otherwise the largest grids are halved until they
  // fit, down to one workgroup per slot, which needs no co-residency.
  // Each size class is its own launch, so the co-residency limit applies per class
  // (with that class's workgroup size). Returns the wgmap range [x, y) of each class;
  // the table is in decreasing size, so a class is a contiguous run of entries, and
  // wgmap_h holds the classes largest first.
  std::vector<int2> plan_workgroups(const int (&th)[kClasses]) {
    const int vmax = opt.num_blocks > 0 ? std::min(opt.num_blocks, kMaxSlotBlocks) : kMaxSlotBlocks;
    std::vector<int> first(kClasses, (int)order.size()), last(kClasses, 0);
    for (size_t i = 0; i < order.size(); i++) {
      const int N = slots[order[i]]->N;
      const int c = size_class(N);
      first[c] = std::min(first[c], (int)i);
      last[c] = std::max(last[c], (int)i + 1);
      vn[i] = std::max(1, std::min(vmax, (N + kRowsPerWorkgroup - 1) / kRowsPerWorkgroup));
    }
    std::vector<int2> seg(kClasses, make_int2(0, 0));
    wgmap_h.clear();
    for (int c = kClasses - 1; c >= 0; c--) {
      seg[c] = make_int2((int)wgmap_h.size(), (int)wgmap_h.size());
      if (first[c] >= last[c]) continue;
      auto it = capacity.find(th[c]);
      if (it == capacity.end())
        it = capacity.emplace(th[c], implicit_batch_capacity(opt.precision, opt.krylov, th[c]))
                 .first;
      const int cap = it->second;
      long long total = 0;
      for (int i = first[c]; i < last[c]; i++) total += vn[i];
      while (total > cap) {
        int big = first[c];
        for (int i = first[c] + 1; i < last[c]; i++)
          if (vn[i] > vn[big]) big = i;
        if (vn[big] == 1) break;  // one workgroup per slot: no co-residency needed
        const int half = vn[big] / 2;
        total -= vn[big] - half;
        vn[big] = half;
      }
      if (total > cap)
        for (int i = first[c]; i < last[c]; i++) vn[i] = 1;
      for (int i = first[c]; i < last[c]; i++)
        for (int b = 0; b < vn[i]; b++) wgmap_h.push_back(make_int2(i, b));
      seg[c].y = (int)wgmap_h.size();
    }
    return seg;
  }

  // Waits for the launch in flight and accounts for it.
  void finish() {
    if (!pending) return;
    pending = false;
    float ms = 0;
    if (is_implicit()) {
      NRT_HIP(hipEventSynchronize(ev1));
      NRT_HIP(hipEventElapsedTime(&ms, ev0, ev1));  // the launches together
      for (int c = 0; c < kClasses; c++) {
        if (!pend_cls[c].on) continue;
        float mc = 0;  // this class's launch (overlapping the others)
        NRT_HIP(hipEventElapsedTime(&mc, cev[c][0], cev[c][1]));
        pend_cls[c].on = false;
        if (opt.batch_threads <= 0 && cthreads[c] == 0 && pend_cls[c].cs > 0 && !cwarm[c]) {
          cwarm[c] = true;
        } else if (opt.batch_threads <= 0 && cthreads[c] == 0 && pend_cls[c].cs > 0) {
          ctune[c].emplace_back(pend_cls[c].th, mc / (double)pend_cls[c].cs);
          if ((int)ctune[c].size() == kTuneCount)
            cthreads[c] = std::min_element(ctune[c].begin(), ctune[c].end(), [](auto& a, auto& b) {
                            return a.second < b.second;
                          })->first;
        }
        stats.class_threads[c] = pend_cls[c].th;
      }
      stats.threads = pend_threads;
    } else {
      NRT_HIP(hipEventSynchronize(ev1));
      NRT_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    }
    stats.launches += pend_launches;
    stats.kernel_ms += ms;
    stats.slot_steps += pend_ss;
    stats.compartment_steps += pend_cs;
    for (int id : order)
      if (slots[id] && slots[id]->left() <= 0) dirty = true;
  }

  void step(long long k) {
    launch(k);
    finish();
  }
};

namespace {
template <class F>
int guarded_b(nrt_batch* b, F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    if (b) b->last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    if (b) b->last_error = "host allocation failed";
    return NRT_ERR_NOMEM;
  } catch (const std::exception& e) {
    if (b) b->last_error = e.what();
    return NRT_ERR_INVALID;
  }
}
}  // namespace

extern "C" {

int nrt_batch_create(const nrt_options* opt, nrt_batch** out) {
  if (!out) return NRT_ERR_INVALID;
  *out = nullptr;
  auto* b = new nrt_batch;
  int st = guarded_b(b, [&] {
    b->opt = resolve_options(opt);
    NRT_REQUIRE(b->opt.kind == NRT_EXPLICIT || b->opt.krylov == NRT_DIRECT ||
                    (b->opt.krylov == NRT_BICGSTAB && b->opt.precond == NRT_PRECOND_TREE),
                NRT_ERR_UNSUPPORTED,
                "implicit batches use the tree solve (krylov = NRT_DIRECT) or BiCGSTAB "
                "preconditioned by it (NRT_BICGSTAB with NRT_PRECOND_TREE)");
    NRT_REQUIRE(b->opt.batch_threads == 0 || b->opt.batch_threads == 128 ||
                    b->opt.batch_threads == 256 || b->opt.batch_threads == 512 ||
                    b->opt.batch_threads == 1024,
                NRT_ERR_INVALID, "batch_threads must be 0, 128, 256, 512 or 1024");
    b->device = pick_device(b->opt.device);
    b->opt.device = b->device;
    NRT_HIP(hipSetDevice(b->device));
    NRT_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    NRT_HIP(hipStreamCreateWithFlags(&b->adm, hipStreamNonBlocking));
    NRT_HIP(hipEventCreate(&b->ev0));
    NRT_HIP(hipEventCreate(&b->ev1));
    for (int c = 0; c < kClasses; c++) {
      for (int e = 0; e < 2; e++) NRT_HIP(hipEventCreate(&b->cev[c][e]));
      if (b->opt.kind == NRT_IMPLICIT && c < kClasses - 1)
        NRT_HIP(hipStreamCreateWithFlags(&b->side[c], hipStreamNonBlocking));
    }
    return NRT_OK;
  });
  if (st != NRT_OK) {
    nrt_batch_destroy(b);
    return st;
  }
  *out = b;
  return NRT_OK;
}

int nrt_batch_add(nrt_batch* b, const nrt_problem* p, int64_t num_steps, int64_t a0,
                  int64_t a1, int32_t trace_elem, int64_t trace_every, int32_t* slot) {
  if (!b) return NRT_ERR_INVALID;
  return guarded_b(b, [&] {
    NRT_REQUIRE(p && slot, NRT_ERR_INVALID, "NULL argument");
    NRT_REQUIRE(num_steps >= 0, NRT_ERR_INVALID, "num_steps < 0");
    NRT_REQUIRE(p->num_elem >= 3, NRT_ERR_INVALID, "num_elem must be >= 3");
    NRT_REQUIRE(trace_every < 0 || (trace_elem >= 0 && trace_elem < p->num_elem),
                NRT_ERR_INVALID, "trace_elem out of range");
    NRT_HIP(hipSetDevice(b->device));
    batch_admission_check(b, p->num_elem, num_steps, trace_every, 0);
    // built on the admission stream (overlapping a launch in flight)
    nrt_handle* h = handle_create(p, b->slot_options(p->num_elem), b->adm, batch_plan_threads(b));
    *slot = batch_add_handle(b, h, num_steps, a0, a1, trace_elem, trace_every);
    return NRT_OK;
  });
}

int nrt_batch_step(nrt_batch* b, int64_t k) {
  if (!b) return NRT_ERR_INVALID;
  return guarded_b(b, [&] {
    NRT_HIP(hipSetDevice(b->device));
    b->step(k);
    return NRT_OK;
  });
}

int nrt_batch_snapshot(nrt_batch* b, int32_t slot, double* host_out) {
  if (!b) return NRT_ERR_INVALID;
  return guarded_b(b, [&] {
    NRT_REQUIRE(host_out, NRT_ERR_INVALID, "host_out is NULL");
    Slot& s = b->slot(slot);
    NRT_HIP(hipSetDevice(b->device));
    b->finish();
    if (b->is_implicit()) {
      implicit_snapshot(s.h, host_out);
    } else {
      const double* v = explicit_slot_v(s.h, s.done);
      NRT_HIP(hipMemcpyAsync(host_out, v, sizeof(double) * s.N, hipMemcpyDeviceToHost,
                             b->stream));
      NRT_HIP(hipStreamSynchronize(b->stream));
    }
    return NRT_OK;
  });
}

int nrt_batch_trace(nrt_batch* b, int32_t slot, double* host_out, int64_t capacity,
                    int64_t* count) {
  if (!b) return NRT_ERR_INVALID;
  return guarded_b(b, [&] {
    Slot& s = b->slot(slot);
    NRT_HIP(hipSetDevice(b->device));
    b->finish();
    const int64_t n =
        s.interval >= 0 ? std::min<int64_t>(s.done / (s.interval + 1), (int64_t)s.trace.n)
                        : 0;
    const int64_t c = std::min(n, std::max<int64_t>(capacity, 0));
    if (c > 0) {
      NRT_REQUIRE(host_out, NRT_ERR_INVALID, "host_out is NULL");
      NRT_HIP(hipSetDevice(b->device));
      NRT_HIP(hipMemcpyAsync(host_out, s.trace.p, sizeof(double) * c, hipMemcpyDeviceToHost,
                             b->stream));
      NRT_HIP(hipStreamSynchronize(b->stream));
    }
    if (count) *count = n;
    return NRT_OK;
  });
}

int nrt_batch_slot_info(nrt_batch* b, int32_t slot, nrt_slot_info* info) {
  if (!b) return NRT_ERR_INVALID;
  return guarded_b(b, [&] {
    NRT_REQUIRE(info, NRT_ERR_INVALID, "info is NULL");
    Slot& s = b->slot(slot);
    NRT_HIP(hipSetDevice(b->device));
    b->finish();
    memset(info, 0, sizeof(*info));
    info->N = s.N;
    info->steps_done = s.done;
    info->num_steps = s.num_steps;
    info->trace_samples = s.interval >= 0 ? s.done / (s.interval + 1) : 0;
    if (b->is_implicit()) {
      NRT_HIP(hipSetDevice(b->device));
      int st = 0;
      int64_t inner = 0;
      implicit_slot_stats(s.h, &st, &inner);
      info->status = st;
      info->inner_iters = inner;
      info->nnz = implicit_nnz(s.h->im);
    }
    return NRT_OK;
  });
}

int nrt_batch_remove(nrt_batch* b, int32_t slot) {
  if (!b) return NRT_ERR_INVALID;
  return guarded_b(b, [&] {
    b->slot(slot);
    NRT_HIP(hipSetDevice(b->device));
    b->finish();
    NRT_HIP(hipStreamSynchronize(b->stream));
    b->slots[slot].reset();
    b->dirty = true;
    return NRT_OK;
  });
}

int nrt_batch_get_stats(nrt_batch* b, nrt_batch_stats* out) {
  if (!b) return NRT_ERR_INVALID;
  return guarded_b(b, [&] {
    NRT_REQUIRE(out, NRT_ERR_INVALID, "out is NULL");
    *out = b->stats;
    out->active_slots = 0;
    out->device_bytes = 0;
    for (auto& s : b->slots) {
      if (!s) continue;
      if (s->left() > 0) out->active_slots++;
      out->device_bytes += s->bytes;
    }
    if (b->is_implicit()) {
      int best = -1;
      int64_t most = -1, per[kClasses] = {0, 0, 0};
      for (auto& s : b->slots)
        if (s) per[size_class(s->N)] += s->N;
      for (int c = 0; c < kClasses; c++) {
        out->class_threads[c] = b->pick_threads(c);
        if (per[c] > most) {
          most = per[c];
          best = c;
        }
      }
      out->threads = b->pick_threads(best);
    }
    return NRT_OK;
  });
}

// Launch-shape tuning: trial batches of one size class each, built from the
// caller's sample, over every (plan, workgroup size) pair; a plan under which any
// trial slot fails is not chosen.
int nrt_batch_tune(nrt_batch* b, const nrt_problem* const* sample, int32_t n, int64_t steps,
                   int64_t rows_per_class, nrt_batch_tune_result* out) {
  if (!b) return NRT_ERR_INVALID;
  return guarded_b(b, [&] {
    NRT_REQUIRE(out && (sample || n == 0) && n >= 0, NRT_ERR_INVALID, "NULL argument");
    NRT_REQUIRE(steps >= 1 && rows_per_class >= 1, NRT_ERR_INVALID,
                "steps and rows_per_class must be >= 1");
    NRT_REQUIRE(b->is_implicit(), NRT_ERR_UNSUPPORTED,
                "launch shapes are tuned for implicit batches only");
    NRT_HIP(hipSetDevice(b->device));
    b->finish();
    *out = nrt_batch_tune_result{};
    const auto t0 = std::chrono::steady_clock::now();
    // each class: its problems in increasing size, every k-th one so that the
    // sample spans the class within the row budget
    std::vector<const nrt_problem*> cls[kClasses];
    for (int i = 0; i < n; i++) {
      NRT_REQUIRE(sample[i] && sample[i]->num_elem >= 4, NRT_ERR_INVALID,
                  "sample problem " + std::to_string(i) + " is NULL or has < 4 elements");
      cls[size_class(sample[i]->num_elem)].push_back(sample[i]);
    }
    struct Plan {
      int chunk, chunk_hi;
    };
    std::vector<Plan> plans;
    if (b->opt.chunk > 0) plans.push_back({b->opt.chunk, b->opt.chunk_hi});
    else plans = {{0, 0}, {8, 8}, {12, 8}, {16, 8}, {24, 8}, {16, 5}, {16, 12}};
    std::vector<int> ths;
    if (b->opt.batch_threads > 0) ths.push_back(b->opt.batch_threads);
    else ths.assign(std::begin(kTuneThreads), std::end(kTuneThreads));
    const int64_t run_steps = 2 * steps * (int64_t)ths.size();
    for (int c = 0; c < kClasses; c++) {
      auto& v = cls[c];
      if (v.empty()) continue;
      std::stable_sort(v.begin(), v.end(), [](const nrt_problem* x, const nrt_problem* y) {
        return x->num_elem < y->num_elem;
      });
      int64_t total = 0;
      for (auto* p : v) total += p->num_elem;
      const size_t stride = (size_t)std::max<int64_t>(1, (total + rows_per_class - 1) / rows_per_class);
      std::vector<const nrt_problem*> pick;
      for (size_t i = std::min(stride / 2, v.size() - 1); i < v.size(); i += stride)
        pick.push_back(v[i]);
      for (auto* p : pick) out->rows[c] += p->num_elem;
      out->slots[c] = (int32_t)pick.size();
      double best = 0;
      for (const Plan& pl : plans) {
        nrt_options o = b->opt;
        o.chunk = pl.chunk;
        o.chunk_hi = pl.chunk_hi;
        o.batch_threads = 0;
        nrt_batch* t = nullptr;
        int st = nrt_batch_create(&o, &t);
        NRT_REQUIRE(st == NRT_OK, st, "trial batch: create failed");
        std::unique_ptr<nrt_batch, int (*)(nrt_batch*)> own(t, nrt_batch_destroy);
        for (auto* p : pick) {
          int32_t id = 0;
          st = nrt_batch_add(t, p, run_steps, 1, 20, 0, -1, &id);
          NRT_REQUIRE(st == NRT_OK, st, std::string("trial batch: ") + t->last_error);
        }
        std::vector<std::pair<int, double>> timed;
        for (int th : ths) {
          t->opt.batch_threads = th;  // the launch shape of this trial
          t->step(steps);             // untimed
          const double k0 = t->stats.kernel_ms;
          const int64_t cs0 = t->stats.compartment_steps;
          t->step(steps);
          const int64_t cs = t->stats.compartment_steps - cs0;
          timed.emplace_back(th, cs > 0 ? (t->stats.kernel_ms - k0) * 1e6 / (double)cs : 0.0);
          out->trials++;
        }
        bool ok = true;
        for (auto& sl : t->slots) {
          int sst = NRT_OK;
          int64_t inner = 0;
          implicit_slot_stats(sl->h, &sst, &inner);
          ok = ok && sst == NRT_OK;
        }
        if (!ok) continue;
        for (auto& [th, ns] : timed) {
          if (pl.chunk == 0 && th == 256) out->auto_ns_per_row_step[c] = ns;
          if (best == 0 || ns < best) {
            best = ns;
            out->threads[c] = th;
            out->chunk[c] = pl.chunk;
            out->chunk_hi[c] = pl.chunk_hi;
          }
        }
      }
      NRT_REQUIRE(best > 0, NRT_ERR_NUMERIC,
                  "every trial plan failed for size class " + std::to_string(c));
      out->ns_per_row_step[c] = best;
      b->cchunk[c] = out->chunk[c];
      b->cchunk_hi[c] = out->chunk_hi[c];
      b->cthreads[c] = out->threads[c];
      b->ctune[c].clear();
    }
    out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return NRT_OK;
  });
}

int nrt_batch_set_shape(nrt_batch* b, int32_t size_class, int32_t threads, int32_t chunk,
                        int32_t chunk_hi) {
  if (!b) return NRT_ERR_INVALID;
  return guarded_b(b, [&] {
    NRT_REQUIRE(b->is_implicit(), NRT_ERR_UNSUPPORTED, "launch shapes are for implicit batches");
    NRT_REQUIRE(size_class >= 0 && size_class < kClasses, NRT_ERR_INVALID,
                "size_class must be 0, 1 or 2");
    NRT_REQUIRE(threads == 0 || threads == 128 || threads == 256 || threads == 512 ||
                    threads == 1024,
                NRT_ERR_INVALID, "threads must be 0, 128, 256, 512 or 1024");
    NRT_REQUIRE(chunk >= 0 && chunk <= kMaxChunk && chunk_hi >= 0 && chunk_hi <= kMaxChunk,
                NRT_ERR_INVALID, "chunk lengths must be in [0, 4096]");
    b->cthreads[size_class] = threads;
    b->ctune[size_class].clear();
    b->cchunk[size_class] = chunk;
    b->cchunk_hi[size_class] = chunk > 0 ? chunk_hi : 0;
    return NRT_OK;
  });
}

const char* nrt_batch_last_error(const nrt_batch* b) { return b ? b->last_error.c_str() : ""; }

int nrt_batch_destroy(nrt_batch* b) {
  if (!b) return NRT_OK;
  if (b->stream) {
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
  }
  if (b->adm) (void)hipStreamSynchronize(b->adm);
  b->drop_graph();
  b->slots.clear();
  b->table.release();
  b->blk.release();
  b->gdev.release();
  for (int c = 0; c < kClasses; c++) {
    for (int e = 0; e < 2; e++)
      if (b->cev[c][e]) (void)hipEventDestroy(b->cev[c][e]);
    if (b->side[c]) (void)hipStreamDestroy(b->side[c]);
  }
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->adm) (void)hipStreamDestroy(b->adm);
  delete b;
  return NRT_OK;
}

}  // extern "C"

namespace nrt {
int64_t batch_footprint(const nrt_batch* b, int N, int64_t num_steps, int64_t trace_every) {
  const int64_t per = b->is_implicit() ? kImplicitBytesPerElem : kExplicitBytesPerElem;
  const int64_t samples = trace_every >= 0 ? num_steps / (trace_every + 1) : 0;
  return per * (int64_t)N + 8 * samples;
}

void batch_admission_check(const nrt_batch* b, int N, int64_t num_steps, int64_t trace_every,
                           int64_t reserved) {
  size_t free_b = 0, total_b = 0;
  NRT_HIP(hipMemGetInfo(&free_b, &total_b));
  const int64_t avail = (int64_t)free_b + pool_cached_bytes(b->device) - reserved;
  const int64_t need = batch_footprint(b, N, num_steps, trace_every);
  NRT_REQUIRE(avail >= need + kAdmissionReserve, NRT_ERR_ADMISSION,
              "slot of " + std::to_string(N) + " elements needs ~" + std::to_string(need >> 20) +
                  " MiB; " + std::to_string(avail >> 20) + " MiB free");
}

int batch_plan_threads(const nrt_batch* b) {
  return b->opt.batch_threads > 0 ? b->opt.batch_threads : kPlanThreads;
}

nrt_options batch_slot_options(const nrt_batch* b, int N) { return b->slot_options(N); }

int batch_add_handle(nrt_batch* b, nrt_handle* h, int64_t num_steps, int64_t a0, int64_t a1,
                     int32_t trace_elem, int64_t trace_every) {
  auto s = std::make_unique<Slot>();
  s->h = h;  // owned by the slot from here on, also if the rest throws
  h->stream = b->stream;  // every later operation on the slot runs in stream order
  s->N = h->N;
  s->s0 = b->g;
  s->num_steps = num_steps;
  s->a0 = a0;
  s->a1 = a1;
  s->interval = trace_every;
  s->trace_elem = trace_elem;
  const int64_t samples = trace_every >= 0 ? num_steps / (trace_every + 1) : 0;
  if (samples > 0) s->trace.alloc((size_t)samples);
  s->bytes = (b->is_implicit() ? implicit_device_bytes(h->im) : explicit_device_bytes(h->ex)) +
             8 * samples;
  const int id = (int)b->slots.size();
  b->slots.push_back(std::move(s));
  b->dirty = true;
  return id;
}

static bool in_flight(const nrt_batch* b, int slot) {
  return b->pending && std::find(b->order.begin(), b->order.end(), slot) != b->order.end();
}

void batch_slot_collect(nrt_batch* b, int slot, hipStream_t side, std::vector<double>* V,
                        std::vector<double>* trace, int* status, int64_t* steps_done) {
  Slot& s = b->slot(slot);
  NRT_REQUIRE(s.left() <= 0 && !in_flight(b, slot), NRT_ERR_INVALID,
              "slot " + std::to_string(slot) + " is still running");
  *status = NRT_OK;
  *steps_done = s.done;
  if (V) {
    V->resize(s.N);
    if (b->is_implicit()) {
      implicit_snapshot_on(s.h, V->data(), side);
    } else {
      NRT_HIP(hipMemcpyAsync(V->data(), explicit_slot_v(s.h, s.done), sizeof(double) * s.N,
                             hipMemcpyDeviceToHost, side));
    }
  }
  if (trace) {
    const int64_t n =
        s.interval >= 0 ? std::min<int64_t>(s.done / (s.interval + 1), (int64_t)s.trace.n) : 0;
    trace->resize((size_t)n);
    if (n > 0)
      NRT_HIP(hipMemcpyAsync(trace->data(), s.trace.p, sizeof(double) * n,
                             hipMemcpyDeviceToHost, side));
  }
  if (b->is_implicit()) {
    int64_t inner = 0;
    implicit_slot_stats_on(s.h, side, status, &inner);  // synchronizes `side`
  } else {
    NRT_HIP(hipStreamSynchronize(side));
  }
}

void batch_slot_drop(nrt_batch* b, int slot) {
  b->slot(slot);
  NRT_REQUIRE(!in_flight(b, slot), NRT_ERR_INVALID,
              "slot " + std::to_string(slot) + " is in the launch in flight");
  b->slots[slot].reset();  // its memory goes back to the pool, no device wait
  b->dirty = true;
}

void batch_launch(nrt_batch* b, long long k) { b->launch(k); }
void batch_launch_budget(nrt_batch* b, long long k, long long elem_budget, long long kmax) {
  b->launch(k, elem_budget, kmax);
}
void batch_finish(nrt_batch* b) { b->finish(); }
hipStream_t batch_stream(nrt_batch* b) { return b->stream; }
bool batch_slot_complete(nrt_batch* b, int slot) {
  const Slot& s = b->slot(slot);
  return s.left() <= 0;
}
}  // namespace nrt
