// C ABI of the solver library (include/neurite_hip.h). Every entry point
// converts internal exceptions into a status code plus a message, which is how
// the reference's cudaException (inc/cudaException.h:331-357, thrown from
// solverEx-gpu.cpp:435-438 and friends) is replaced at this boundary.
#include <cstring>
#include <cstdlib>
#include <mutex>
#include "nrt_internal.hpp"

namespace nrt {
int setup_on_device(const nrt_geometry* g, int device, nrt_setup_out* out);
int setup_noalloc_on_device(const nrt_geometry* g, int device, void* stream,
                            const nrt_setup_out* out, int32_t hh_capacity, double* dT_crit,
                            int32_t* num_hh);

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line) {
  (void)hipGetLastError();
  const char* base = strrchr(file, '/');
  throw Error(e == hipErrorOutOfMemory ? NRT_ERR_NOMEM : NRT_ERR_HIP,
              std::string(what) + " failed: " + hipGetErrorString(e) + " [" +
                  (base ? base + 1 : file) + ":" + std::to_string(line) + "]");
}

template <class T>
static std::vector<T> take(const T* p, int n, const char* name) {
  if (n <= 0) return {};
  NRT_REQUIRE(p != nullptr, NRT_ERR_INVALID, std::string("array '") + name + "' is NULL");
  return std::vector<T>(p, p + n);
}

HostProblem load_problem(const nrt_problem* p) {
  NRT_REQUIRE(p != nullptr, NRT_ERR_INVALID, "problem is NULL");
  HostProblem q;
  q.N = p->num_elem;
  q.H = p->num_hh;
  q.T = p->num_term;
  q.dT = p->dT;
  NRT_REQUIRE(q.N >= 3, NRT_ERR_INVALID, "num_elem must be >= 3");
  NRT_REQUIRE(q.H >= 0 && q.H <= q.N, NRT_ERR_INVALID, "num_hh out of range");
  NRT_REQUIRE(q.T >= 1, NRT_ERR_INVALID, "at least one terminal (element 0) is required");
  NRT_REQUIRE(q.dT > 0, NRT_ERR_INVALID, "dT must be positive");
  const int N = q.N, H = q.H;
  q.V0 = take(p->potential0, N, "potential0");
  q.V1 = p->potential1 ? take(p->potential1, N, "potential1") : q.V0;
  q.dx = take(p->dx, N, "dx");
  q.c_m = take(p->c_m, N, "c_m");
  q.r_a = take(p->r_a, N, "r_a");
  q.W = take(p->W, N, "W");
  q.K = take(p->K, N, "K");
  q.input = p->input ? take(p->input, N, "input") : std::vector<double>(N, 0.0);
  q.mother = take(p->mother, N, "mother");
  q.dr = take(p->dr, N, "dr");
  q.dl = take(p->dl, N, "dl");
  q.branching = p->branching ? take(p->branching, N, "branching") : std::vector<uint8_t>(N, 0);
  q.terminals = take(p->terminals, q.T, "terminals");
  q.hh_elem = take(p->hh_elem, H, "hh_elem");
  q.m = take(p->hh_m, H, "hh_m");
  q.h = take(p->hh_h, H, "hh_h");
  q.n = take(p->hh_n, H, "hh_n");
  q.E_Na = take(p->E_Na, H, "E_Na");
  q.E_K = take(p->E_K, H, "E_K");
  q.E_L = take(p->E_L, H, "E_L");
  q.g_Na = take(p->g_Na, H, "g_Na");
  q.g_K = take(p->g_K, H, "g_K");
  q.G_L = take(p->G_L, H, "G_L");
  q.ls_Na = take(p->ls_Na, H, "ls_Na");
  q.ls_K = take(p->ls_K, H, "ls_K");
  q.rest = take(p->rest_pot, H, "rest_pot");
  for (int t : q.terminals)
    NRT_REQUIRE(t >= 0 && t < N, NRT_ERR_TOPOLOGY, "terminal index out of range");
  for (int e : q.hh_elem) NRT_REQUIRE(e >= 0 && e < N, NRT_ERR_INVALID, "hh_elem out of range");
  for (int j = 0; j < N; j++)
    NRT_REQUIRE(q.dx[j] > 0 && q.r_a[j] > 0 && q.c_m[j] > 0, NRT_ERR_INVALID,
                "dx, r_a and c_m must be positive (element " + std::to_string(j) + ")");
  return q;
}

int pick_device(int requested) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    throw Error(NRT_ERR_NO_DEVICE, "no HIP device available");
  }
  int dev = requested;
  if (dev < 0) {
    const char* env = getenv("idGPU");  // solverEx-gpu.cpp:210-214
    dev = env ? atoi(env) : 0;
  }
  NRT_REQUIRE(dev >= 0 && dev < count, NRT_ERR_INVALID,
              "device " + std::to_string(dev) + " out of range (" + std::to_string(count) + ")");
  return dev;
}

nrt_options resolve_options(const nrt_options* opt) {
  nrt_options o;
  if (opt) o = *opt;
  else nrt_default_options(&o);
  NRT_REQUIRE(o.kind == NRT_EXPLICIT || o.kind == NRT_IMPLICIT, NRT_ERR_INVALID, "bad kind");
  NRT_REQUIRE(o.precision == NRT_FP64 || o.precision == NRT_FP32_REFINE, NRT_ERR_INVALID,
              "bad precision");
  NRT_REQUIRE(o.precond == NRT_PRECOND_NONE || o.precond == NRT_PRECOND_TREE, NRT_ERR_INVALID,
              "bad precond");
  NRT_REQUIRE(o.precond == NRT_PRECOND_NONE ||
                  (o.kind == NRT_IMPLICIT && o.krylov == NRT_BICGSTAB),
              NRT_ERR_UNSUPPORTED,
              "the tree-solve preconditioner is built for implicit BiCGSTAB (GMRES takes none)");
  NRT_REQUIRE(o.precision == NRT_FP64 ||
                  (o.kind == NRT_IMPLICIT &&
                   (o.krylov == NRT_DIRECT || o.krylov == NRT_GMRES ||
                    (o.krylov == NRT_BICGSTAB && o.precond == NRT_PRECOND_TREE))),
              NRT_ERR_UNSUPPORTED,
              "FP32 with FP64 refinement is available for the implicit tree solve, GMRES and "
              "BiCGSTAB preconditioned by the tree solve");
  NRT_REQUIRE(o.kind == NRT_EXPLICIT || o.krylov == NRT_BICGSTAB || o.krylov == NRT_GMRES ||
                  o.krylov == NRT_DIRECT,
              NRT_ERR_UNSUPPORTED, "implicit methods: NRT_BICGSTAB, NRT_GMRES or NRT_DIRECT");
  NRT_REQUIRE(o.kind == NRT_EXPLICIT || o.krylov != NRT_GMRES ||
                  (o.restart >= 1 && o.restart <= 512),
              NRT_ERR_INVALID, "GMRES restart must be in [1, 512]");
  NRT_REQUIRE(o.x0_mode == NRT_X0_PREVIOUS || o.x0_mode == NRT_X0_RHS ||
                  o.x0_mode == NRT_X0_ZERO,
              NRT_ERR_INVALID, "bad x0_mode");
  NRT_REQUIRE(o.chunk >= 0 && o.chunk <= kMaxChunk && o.chunk_hi >= 0 &&
                  o.chunk_hi <= kMaxChunk,
              NRT_ERR_INVALID, "tree chunk lengths must be in [0, 4096]");
  NRT_REQUIRE(o.tree_groups >= 0 && o.tree_groups <= 1024, NRT_ERR_INVALID,
              "tree_groups must be in [0, 1024]");
  return o;
}

nrt_handle* handle_create(const nrt_problem* p, const nrt_options& o, hipStream_t shared,
                          int batch_threads) {
  HostProblem q = load_problem(p);
  auto* h = new nrt_handle;
  try {
    h->opt = o;
    h->device = pick_device(o.device);
    h->N = q.N;
    h->stats.tree_form = -1;
    NRT_HIP(hipSetDevice(h->device));
    if (shared) {
      h->stream = shared;
      h->owns_stream = false;
    } else {
      NRT_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    }
    if (o.kind == NRT_EXPLICIT) h->ex = explicit_create(h, q);
    else h->im = implicit_create(h, q, batch_threads);
  } catch (...) {
    if (h->stream && h->owns_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    throw;
  }
  return h;
}

void handle_destroy(nrt_handle* h) {
  if (h->ex) explicit_destroy(h->ex);
  if (h->im) implicit_destroy(h->im);
  if (h->stream && h->owns_stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

}  // namespace nrt

using namespace nrt;

static thread_local std::string g_error;

template <class F>
static int guarded(nrt_handle* h, F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    (h ? h->last_error : g_error) = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    (h ? h->last_error : g_error) = "host allocation failed";
    return NRT_ERR_NOMEM;
  } catch (const std::exception& e) {
    (h ? h->last_error : g_error) = e.what();
    return NRT_ERR_INVALID;
  }
}

extern "C" {

int nrt_version(void) { return 1; }

int nrt_device_count(int* count) {
  return guarded(nullptr, [&] {
    NRT_REQUIRE(count, NRT_ERR_INVALID, "count is NULL");
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      *count = 0;
    }
    return NRT_OK;
  });
}

const char* nrt_last_error(const nrt_handle* h) {
  return h ? h->last_error.c_str() : g_error.c_str();
}

const char* nrt_status_string(int status) {
  switch (status) {
    case NRT_OK: return "ok";
    case NRT_ERR_INVALID: return "invalid argument";
    case NRT_ERR_HIP: return "HIP runtime error";
    case NRT_ERR_NOMEM: return "out of device memory";
    case NRT_ERR_TOPOLOGY: return "unsupported topology";
    case NRT_ERR_NOT_CONVERGED: return "Krylov solver did not converge";
    case NRT_ERR_NUMERIC: return "non-finite value in the solver";
    case NRT_ERR_UNSUPPORTED: return "unsupported option";
    case NRT_ERR_ADMISSION: return "not enough free device memory";
    case NRT_ERR_NO_DEVICE: return "no device";
    default: return "unknown status";
  }
}

void nrt_default_options(nrt_options* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->kind = NRT_EXPLICIT;
  o->precision = NRT_FP64;
  o->krylov = NRT_BICGSTAB;
  o->restart = 30;
  o->tol_mode = NRT_TOL_RELATIVE;
  o->tol = 1e-10;
  o->x0_mode = NRT_X0_PREVIOUS;
  o->check_every = 1;
  o->max_iter = 0;
  o->device = -1;
  o->num_blocks = 0;
  o->refine_tol = 1e-10;
  o->refine_max = 8;
}

int nrt_setup(const nrt_geometry* g, int device, nrt_setup_out* out) {
  return guarded(nullptr, [&] {
    NRT_REQUIRE(g && out, NRT_ERR_INVALID, "NULL argument");
    return setup_on_device(g, pick_device(device), out);
  });
}

int nrt_setup_noalloc(const nrt_geometry* g, int device, void* stream,
                      const nrt_setup_out* out, int32_t hh_capacity, double* dT_crit_out,
                      int32_t* num_hh_out) {
  return guarded(nullptr, [&] {
    return setup_noalloc_on_device(g, pick_device(device), stream, out, hh_capacity, dT_crit_out,
                                   num_hh_out);
  });
}

int nrt_create(const nrt_problem* p, const nrt_options* opt, nrt_handle** out) {
  return guarded(nullptr, [&] {
    NRT_REQUIRE(out != nullptr, NRT_ERR_INVALID, "out is NULL");
    *out = nullptr;
    *out = handle_create(p, resolve_options(opt), nullptr, 0);
    return NRT_OK;
  });
}

int nrt_update(nrt_handle* h, int input_active) {
  if (!h) return NRT_ERR_INVALID;
  return guarded(h, [&] {
    h->input_active = input_active != 0;
    return NRT_OK;
  });
}

int nrt_calculate(nrt_handle* h) {
  if (!h) return NRT_ERR_INVALID;
  return guarded(h, [&] {
    NRT_HIP(hipSetDevice(h->device));
    if (h->ex) {
      explicit_step(h, h->input_active, nullptr, 0);
      h->stats.steps++;
      return NRT_OK;
    }
    return implicit_calculate(h);
  });
}

int nrt_snapshot(nrt_handle* h, double* host_out) {
  if (!h) return NRT_ERR_INVALID;
  return guarded(h, [&] {
    NRT_REQUIRE(host_out, NRT_ERR_INVALID, "host_out is NULL");
    NRT_HIP(hipSetDevice(h->device));
    if (h->ex) explicit_snapshot(h, host_out);
    else implicit_snapshot(h, host_out);
    return NRT_OK;
  });
}

int nrt_run(nrt_handle* h, int64_t first_step, int64_t num_steps, int64_t active_first,
            int64_t active_last, int32_t trace_elem, int64_t trace_every, double* trace_out,
            int64_t* num_samples) {
  if (!h) return NRT_ERR_INVALID;
  return guarded(h, [&] {
    NRT_REQUIRE(num_steps >= 0, NRT_ERR_INVALID, "num_steps < 0");
    NRT_REQUIRE(trace_every < 0 || (trace_elem >= 0 && trace_elem < h->N), NRT_ERR_INVALID,
                "trace_elem out of range");
    NRT_HIP(hipSetDevice(h->device));
    if (num_samples) *num_samples = 0;
    if (num_steps == 0) return NRT_OK;
    if (h->ex) {
      explicit_run(h, first_step, num_steps, active_first, active_last, trace_elem, trace_every,
                   trace_out, num_samples);
      return NRT_OK;
    }
    return implicit_run(h, first_step, num_steps, active_first, active_last, trace_elem,
                        trace_every, trace_out, num_samples);
  });
}

int nrt_get_stats(nrt_handle* h, nrt_stats* st) {
  if (!h || !st) return NRT_ERR_INVALID;
  *st = h->stats;
  return NRT_OK;
}

int nrt_sync(nrt_handle* h) {
  if (!h) return NRT_ERR_INVALID;
  return guarded(h, [&] {
    NRT_HIP(hipStreamSynchronize(h->stream));
    return NRT_OK;
  });
}

int nrt_destroy(nrt_handle* h) {
  if (!h) return NRT_OK;
  int st = guarded(h, [&] {
    NRT_HIP(hipSetDevice(h->device));
    NRT_HIP(hipStreamSynchronize(h->stream));
    return NRT_OK;
  });
  handle_destroy(h);
  return st;
}

}  // extern "C"
