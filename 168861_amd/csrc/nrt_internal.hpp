// Internal definitions of the MI355X Neurite solver library (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include <stdexcept>
#include "../../include/neurite_hip.h"

namespace nrt {

// Error raised inside the library; converted to a status + message at the C ABI.
// Replaces the reference's cudaException (inc/cudaException.h:331-357): it carries
// the status, a description, and the file/function/line of the failing call.
struct Error : public std::runtime_error {
  int status;
  Error(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define NRT_HIP(call)                                                  \
  do {                                                                 \
    hipError_t e_ = (call);                                            \
    if (e_ != hipSuccess) ::nrt::throw_hip(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define NRT_REQUIRE(cond, status, msg)                                 \
  do {                                                                 \
    if (!(cond)) throw ::nrt::Error((status), (msg));                  \
  } while (0)

// Device memory pool (pool.cpp): a caching allocator by size class, per device.
// Freed blocks stay cached for the next allocation of their class instead of going
// back through hipFree, which waits for the device and costs up to ~0.2 ms per
// block; solvers, batch slots and served requests come and go at that rate. A
// block must be returned only once no queued work uses it.
void* pool_alloc(size_t bytes);                 // throws NRT_ERR_NOMEM
void pool_free(void* p, size_t bytes);
int64_t pool_cached_bytes(int device);          // cached, free for reuse
int64_t pool_trim(int device);                  // returns cached blocks to HIP

// Owning device buffer.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) pool_free(p, n * sizeof(T));
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    p = (T*)pool_alloc(count * sizeof(T));
    n = count;
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    alloc(count);
    if (count) NRT_HIP(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

// Topology exception record of the explicit step: every element whose
// neighbourhood is not the plain chain (mother j-1, right daughter j+1, no left
// daughter, no stimulus, no sealed-end copy) carries one.
struct ExRec {
  int32_t mother, dr, dl, hh;   // hh = -1 when passive
  int32_t tgt[3];               // sealed-end copies V_new[tgt] = V_new[self]
  int32_t flags;                // bit0 branching, bit1 stimulus
  double input;                 // stimulus current when active
};
static_assert(sizeof(ExRec) == 40, "ExRec layout");

// element code (32 bit): flag bits + index
constexpr uint32_t kCodeHH = 0x80000000u;    // plain chain element, index = HH index
constexpr uint32_t kCodeRec = 0x40000000u;   // index = exception record
constexpr uint32_t kCodeSkip = 0x20000000u;  // not computed (terminal/ghost)
constexpr uint32_t kCodeIdx = 0x1FFFFFFFu;
constexpr int kMaxChunk = 4096;  // longest chain chunk of the tree solve
constexpr int kMaxSlotBlocks = 64;  // workgroups of one batch slot's virtual grid
constexpr int kRowsPerWorkgroup = 24576;  // a batch slot gets a workgroup per this many rows

// Host-side view of an nrt_problem with validation helpers.
struct HostProblem {
  int N = 0, H = 0, T = 0;
  double dT = 0;
  std::vector<double> V0, V1, dx, c_m, r_a, W, K, input;
  std::vector<int32_t> mother, dr, dl;
  std::vector<uint8_t> branching;
  std::vector<int32_t> terminals, hh_elem;
  std::vector<double> m, h, n, E_Na, E_K, E_L, g_Na, g_K, G_L, ls_Na, ls_K, rest;
};
HostProblem load_problem(const nrt_problem* p);

struct Explicit;
struct Implicit;

}  // namespace nrt

struct nrt_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  nrt_options opt{};
  nrt_stats stats{};
  std::string last_error;
  bool input_active = false;
  nrt::Explicit* ex = nullptr;
  nrt::Implicit* im = nullptr;
  int N = 0;
  bool owns_stream = true;  // false for the slots of a batch (they use its stream)
};

namespace nrt {
// explicit.hip
Explicit* explicit_create(nrt_handle* h, const HostProblem& p);
void explicit_destroy(Explicit* e);
void explicit_step(nrt_handle* h, bool active, double* trace_dst, int32_t trace_elem);
void explicit_snapshot(nrt_handle* h, double* host_out);
void explicit_run(nrt_handle* h, int64_t first, int64_t num, int64_t a0, int64_t a1,
                  int32_t trace_elem, int64_t interval, double* trace_out, int64_t* ns);
// implicit.hip
// batch_threads > 0: a slot of a batch (one workgroup of that size per step).
Implicit* implicit_create(nrt_handle* h, const HostProblem& p, int batch_threads = 0);
void implicit_destroy(Implicit* e);
void implicit_update(nrt_handle* h, bool active);
int implicit_calculate(nrt_handle* h);
void implicit_snapshot(nrt_handle* h, double* host_out);
int implicit_run(nrt_handle* h, int64_t first, int64_t num, int64_t a0, int64_t a1,
                 int32_t trace_elem, int64_t interval, double* trace_out, int64_t* ns);

inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

// capi.cpp
int pick_device(int requested);
nrt_options resolve_options(const nrt_options* opt);  // defaults + validation
// A handle on its own stream (shared == nullptr) or on a batch's stream.
nrt_handle* handle_create(const nrt_problem* p, const nrt_options& o, hipStream_t shared,
                          int batch_threads);
void handle_destroy(nrt_handle* h);
}  // namespace nrt
