// Explicit (FTCS + forward-Euler HH) step, FP64 plain form, one fused kernel.
//
// Semantics follow the reference CPU solver, the general-topology one
// (solverEx-cpu.cpp:45-123 with hhRate_Const coupling.cpp:25-48); the reference
// GPU solver splits the same step into four launches (solverEx-gpu.cpp:87-194)
// and assumes chain numbering, which this kernel does not.
//
// Layout in HBM (structure of arrays, 8-byte values, 4-byte codes):
//   V[2][N]        double-buffered membrane potential
//   coefA[N]       dT/(c_m*dx)           (solverEx-cpu.cpp:96)
//   rdx[N]         r_a*dx                (divisor of the axial terms)
//   wdx[N], kdx[N] W*dx, K*dx of passive elements (constant in time)
//   code[N]        element class + index (plain chain / HH / record / skip)
//   recs[R]        exception records (branch points, daughters, stimulus
//                  sites, sources of sealed-end copies)
//   HH SoA [H]     m, h, n (state) and E_Na, E_K, E_L, g_Na, g_K, G_L,
//                  ls_Na, ls_K, rest, dx (parameters)
// Algorithmic bytes per step: 48 per passive element (V r/w, coefA, rdx, wdx,
// kdx), 152 per HH element (V r/w, coefA, rdx, m/h/n r/w, 9 parameters).
//
// Every operation keeps the reference's expression order and divisions, and the
// library is built with -ffp-contract=off, so each element update is the same
// IEEE double sequence as the reference CPU build; the only differing operation
// is exp() (device libm vs host libm, both within an ulp or so).
#include <algorithm>
#include <cstring>
#include "nrt_internal.hpp"

namespace nrt {

struct HHDev {
  double* m;
  double* h;
  double* n;
  const double* E_Na;
  const double* E_K;
  const double* E_L;
  const double* g_Na;
  const double* g_K;
  const double* G_L;
  const double* ls_Na;
  const double* ls_K;
  const double* rest;
  const double* dx;
};

struct Explicit {
  int N = 0, H = 0, R = 0;
  double dT = 0, dt_ms = 0;
  DBuf<double> V[2];
  DBuf<double> coefA, rdx, wdx, kdx;
  DBuf<uint32_t> code;
  DBuf<ExRec> recs;
  DBuf<double> m, h, n, E_Na, E_K, E_L, g_Na, g_K, G_L, ls_Na, ls_K, rest, hdx;
  DBuf<double> trace;
  int cur = 0;  // buffer holding the latest potential (the reference's prevT)
  int64_t trace_cap = 0;
  bool res_attr = false;  // the resident kernel's LDS limit is raised
  HHDev hd() {
    return HHDev{m.p, h.p, n.p, E_Na.p, E_K.p, E_L.p, g_Na.p, g_K.p, G_L.p, ls_Na.p, ls_K.p,
                 rest.p, hdx.p};
  }
};

// Forward-Euler HH gating of HH element i at potential v (solverEx-cpu.cpp:62-82,
// coupling.cpp:25-48, discretization.cpp:81-86); returns W*dx and K*dx.
__device__ __forceinline__ void hh_gate(const HHDev& hd, int i, double v, double dt_ms,
                                        double& wdx, double& kdx) {
  const double rest = hd.rest[i];
  double vna = v + hd.ls_Na[i];
  double vk = v + hd.ls_K[i];
  vna = (vna - rest) / 0.001;
  vk = (vk - rest) / 0.001;
  const double am = (2.5 - 0.1 * vna) / (exp(2.5 - 0.1 * vna) - 1);
  const double bm = 4 * exp(-vna / 18);
  const double ah = 0.07 * exp(-vna / 20);
  const double bh = 1 / (exp(3 - 0.1 * vna) + 1);
  const double an = (0.1 - 0.01 * vk) / (exp(1 - 0.1 * vk) - 1);
  const double bn = 0.125 * exp(-vk / 80);
  const double m0 = hd.m[i], h0 = hd.h[i], n0 = hd.n[i];
  const double m = m0 + dt_ms * (am * (1 - m0) - bm * m0);
  const double h = h0 + dt_ms * (ah * (1 - h0) - bh * h0);
  const double n = n0 + dt_ms * (an * (1 - n0) - bn * n0);
  hd.m[i] = m;
  hd.h[i] = h;
  hd.n[i] = n;
  const double gna = hd.g_Na[i] * m * m * m * h;
  const double gk = hd.g_K[i] * n * n * n * n;
  const double gl = hd.G_L[i];
  const double W = -(gna + gk + gl);
  const double K = gna * hd.E_Na[i] + gk * hd.E_K[i] + gl * hd.E_L[i];
  const double dx = hd.dx[i];
  wdx = W * dx;
  kdx = K * dx;
}

#define NRT_GLB __attribute__((address_space(1)))
template <class P>
__device__ __forceinline__ NRT_GLB P* glb(P* p) {
  return (NRT_GLB P*)p;
}
// Empty asm statements that take a value: the value (and so the load that makes it)
// is complete at this point, which keeps the compiler from sinking a load below a
// later branch.
template <class X>
__device__ __forceinline__ void ready_v(X x) {
  asm volatile("" ::"v"(x));
}
template <class X>
__device__ __forceinline__ void ready_s(X x) {
  asm volatile("" ::"s"(x));
}

// Element kinds of the explicit step (code >> 29; the low 29 bits are an index):
// the flag bits kCodeHH, kCodeRec, kCodeSkip alone or HH with Skip keep their
// meaning (chain element, HH chain element with its HH index, exception record,
// terminal/ghost); the three unused combinations carry an element with one
// exception and no record: a daughter (index = its mother), a branch point (index =
// its left daughter), a source of the sealed end right after it (no index).
constexpr uint32_t kExPlain = 0, kExSkip = 1, kExRecord = 2, kExBranch = 3, kExHH = 4,
                   kExSkipHH = 5, kExDaughter = 6, kExCopyNext = 7;

// Explicit step of element j of N: reads Vo, writes Vn[j] and its sealed-end copies.
// A plain chain element (no record) couples to j - 1 and j + 1; its operands are
// fetched with its code (ex_load), before the code is known, so a thread waits for
// memory once (the clamped neighbour indices equal j -/+ 1 wherever the element is
// plain). Pointers go through the global address space: the batch kernel reads them
// from its slot records, where the compiler cannot tell that they are global. The
// per-element constants and the new potential of a plain element stream through
// once a step (nontemporal); the potentials are read by the neighbours too (cached).
struct ExOps {
  uint32_t c;
  double v, vm, vp, rj, rp, aj, w0, k0;
};

// VO: where the potentials are read (global memory, or the LDS of the resident
// kernel); the constants always come from global memory.
template <class VO>
__device__ __forceinline__ ExOps ex_load_on(int j, int N, VO Vo, const double* __restrict__ coefA_,
                                            const double* __restrict__ rdx_,
                                            const double* __restrict__ wdx_,
                                            const double* __restrict__ kdx_,
                                            const uint32_t* __restrict__ code_) {
  const NRT_GLB double* __restrict__ rdx = glb(rdx_);
  const int jm = j > 0 ? j - 1 : j, jp = j + 1 < N ? j + 1 : j;
  ExOps o;
  o.c = __builtin_nontemporal_load(glb(code_) + j);
  o.v = Vo[j];
  o.vm = Vo[jm];
  o.vp = Vo[jp];
  o.rj = rdx[j];
  o.rp = rdx[jp];
  o.aj = __builtin_nontemporal_load(glb(coefA_) + j);
  o.w0 = __builtin_nontemporal_load(glb(wdx_) + j);
  o.k0 = __builtin_nontemporal_load(glb(kdx_) + j);
  return o;
}

__device__ __forceinline__ ExOps ex_load(int j, int N, const double* __restrict__ Vo_,
                                         const double* __restrict__ coefA_,
                                         const double* __restrict__ rdx_,
                                         const double* __restrict__ wdx_,
                                         const double* __restrict__ kdx_,
                                         const uint32_t* __restrict__ code_) {
  const NRT_GLB double* __restrict__ Vo = glb(Vo_);
  return ex_load_on(j, N, Vo, coefA_, rdx_, wdx_, kdx_, code_);
}

// The loads of ex_load complete here (not later, under the branches on the code).
__device__ __forceinline__ void ex_ready(const ExOps& o) {
  ready_v(o.c);
  ready_v(o.v);
  ready_v(o.vm);
  ready_v(o.vp);
  ready_v(o.rj);
  ready_v(o.rp);
  ready_v(o.aj);
  ready_v(o.w0);
  ready_v(o.k0);
}

// NT: the new potential of a plain element is stored nontemporally (global).
template <bool NT, class VO, class VN>
__device__ __forceinline__ void ex_finish_on(int j, const ExOps& o, VO Vo, VN Vn,
                                             const double* __restrict__ rdx_,
                                             const ExRec* __restrict__ recs_, const HHDev& hd,
                                             double dt_ms, bool active) {
  const NRT_GLB double* __restrict__ rdx = glb(rdx_);
  const uint32_t c = o.c, kind = c >> 29, idx = c & kCodeIdx;
  const double v = o.v;
  if (kind == kExSkip || kind == kExSkipHH) {
    // Terminal / ghost: its potential is a sealed-end copy written by its source,
    // but an HH terminal still advances its gates (solverEx-cpu.cpp:62).
    if (kind == kExSkipHH) {
      double wd, kd;
      hh_gate(hd, (int)idx, v, dt_ms, wd, kd);
    }
    return;
  }
  if (kind != kExRecord) {
    // A chain element, or one whose single exception sits in its code: a daughter
    // (its mother elsewhere), a branch point (its left daughter), the source of
    // the sealed end right after it. One load waits on the code at most, none on
    // a record; the operations are the record path's, in its order.
    double wd = o.w0, kd = o.k0;
    if (kind == kExHH) hh_gate(hd, (int)idx, v, dt_ms, wd, kd);
    const double vm = kind == kExDaughter ? Vo[idx] : o.vm;
    // (DL term +) DR term + mother term + W*dX*V + K*dX
    double s = (o.vp - v) / o.rp;
    if (kind == kExBranch) s = (Vo[idx] - v) / rdx[idx] + s;
    s = s + (vm - v) / o.rj;
    s = s + wd * v;
    s = s + kd;
    const double vn = v + o.aj * s;
    if constexpr (NT) __builtin_nontemporal_store(vn, &Vn[j]);
    else Vn[j] = vn;
    if (kind == kExCopyNext) Vn[j + 1] = vn;
    return;
  }
  const NRT_GLB ExRec& rec = glb(recs_)[c & kCodeIdx];
  const int mo = rec.mother, dr = rec.dr, hh = rec.hh, fl = rec.flags;
  const int dl = (fl & 1) ? rec.dl : -1;
  double wd = o.w0, kd = o.k0;
  if (hh >= 0) hh_gate(hd, hh, v, dt_ms, wd, kd);
  // branching*(DL term) + DR term + mother term + W*dX*V + K*dX + input
  const double tdr = (Vo[dr] - v) / rdx[dr];
  double s = (dl >= 0) ? ((Vo[dl] - v) / rdx[dl] + tdr) : tdr;
  s = s + (Vo[mo] - v) / o.rj;
  s = s + wd * v;
  s = s + kd;
  if ((fl & 2) && active) s = s + rec.input;
  const double vn = v + o.aj * s;
  Vn[j] = vn;
#pragma unroll
  for (int t = 0; t < 3; t++)
    if (rec.tgt[t] >= 0) Vn[rec.tgt[t]] = vn;
}

__device__ __forceinline__ void ex_finish(int j, const ExOps& o, const double* __restrict__ Vo_,
                                          double* __restrict__ Vn_,
                                          const double* __restrict__ rdx_,
                                          const ExRec* __restrict__ recs_, const HHDev& hd,
                                          double dt_ms, bool active) {
  const NRT_GLB double* __restrict__ Vo = glb(Vo_);
  NRT_GLB double* __restrict__ Vn = glb(Vn_);
  ex_finish_on<true>(j, o, Vo, Vn, rdx_, recs_, hd, dt_ms, active);
}

__device__ __forceinline__ void explicit_element(
    int j, int N, const double* __restrict__ Vo, double* __restrict__ Vn,
    const double* __restrict__ coefA, const double* __restrict__ rdx,
    const double* __restrict__ wdx, const double* __restrict__ kdx,
    const uint32_t* __restrict__ code, const ExRec* __restrict__ recs, const HHDev& hd,
    double dt_ms, bool active) {
  const ExOps o = ex_load(j, N, Vo, coefA, rdx, wdx, kdx, code);
  ex_ready(o);
  ex_finish(j, o, Vo, Vn, rdx, recs, hd, dt_ms, active);
}

// One explicit step for every element (thread per element).
__global__ __launch_bounds__(256) void k_explicit_fp64(
    int N, const double* __restrict__ Vo, double* __restrict__ Vn,
    const double* __restrict__ coefA, const double* __restrict__ rdx,
    const double* __restrict__ wdx, const double* __restrict__ kdx,
    const uint32_t* __restrict__ code, const ExRec* __restrict__ recs, HHDev hd, double dt_ms,
    int active) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  explicit_element(j, N, Vo, Vn, coefA, rdx, wdx, kdx, code, recs, hd, dt_ms, active != 0);
}

// Whole runs of a small neurite in one workgroup: both potentials live in LDS
// for the launch and the steps are separated by workgroup barriers instead of
// kernel launches (a launch per step costs more than a thousand elements' work).
// The element update is explicit_element's, operand for operand.
constexpr int kResBlock = 1024;
constexpr int kResMaxN = 8192;  // two potentials of 8 B: 128 KiB of LDS

__global__ __launch_bounds__(kResBlock) void k_explicit_resident(
    int N, const double* __restrict__ Vin, const double* __restrict__ Vother,
    double* __restrict__ Vout, const double* __restrict__ coefA, const double* __restrict__ rdx,
    const double* __restrict__ wdx, const double* __restrict__ kdx,
    const uint32_t* __restrict__ code, const ExRec* __restrict__ recs, HHDev hd, double dt_ms,
    long long first, long long num, long long a0, long long a1, int trace_elem,
    long long interval, double* __restrict__ trace) {
  extern __shared__ double sv[];
  typedef __attribute__((address_space(3))) double LdsD;
  LdsD* A = (LdsD*)sv;
  LdsD* B = (LdsD*)sv + N;
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  for (int j = tid; j < N; j += nt) {
    A[j] = Vin[j];
    B[j] = Vother[j];
  }
  __syncthreads();
  long long gl = 0, k = 0;
  for (long long i = first; i < first + num; i++) {
    const bool active = i >= a0 && i <= a1;
    for (int j = tid; j < N; j += nt) {
      const ExOps o = ex_load_on(j, N, (const LdsD*)A, coefA, rdx, wdx, kdx, code);
      ex_ready(o);
      ex_finish_on<false>(j, o, (const LdsD*)A, B, rdx, recs, hd, dt_ms, active);
    }
    __syncthreads();
    LdsD* t = A;
    A = B;
    B = t;
    if (interval >= 0 && ++gl > interval) {
      gl = 0;
      if (trace && tid == 0) trace[k] = A[trace_elem];
      k++;
    }
  }
  for (int j = tid; j < N; j += nt) Vout[j] = A[j];
}

Explicit* explicit_create(nrt_handle* hnd, const HostProblem& p) {
  hipStream_t s = hnd->stream;
  const int N = p.N, H = p.H;
  auto* e = new Explicit;
  try {
    e->N = N;
    e->H = H;
    e->dT = p.dT;
    e->dt_ms = p.dT * 1000;  // solverEx-cpu.cpp:51
    std::vector<int32_t> hh_of(N, -1);
    for (int i = 0; i < H; i++) {
      NRT_REQUIRE(hh_of[p.hh_elem[i]] < 0, NRT_ERR_INVALID, "duplicate HH element index");
      hh_of[p.hh_elem[i]] = i;
    }
    // Sealed-end copies (solverEx-cpu.cpp:112-116): terminals[0] copies element
    // terminals[0]+1, every other terminal copies its mother.
    std::vector<uint8_t> is_term(N, 0);
    for (int k = 0; k < p.T; k++) is_term[p.terminals[k]] = 1;
    std::vector<std::vector<int32_t>> targets(N);
    for (int k = 0; k < p.T; k++) {
      const int t = p.terminals[k];
      const int src = (k == 0) ? t + 1 : p.mother[t];
      NRT_REQUIRE(src >= 1 && src <= N - 2 && !is_term[src], NRT_ERR_TOPOLOGY,
                  "terminal " + std::to_string(t) + " copies element " + std::to_string(src) +
                      ", which is not an interior non-terminal element");
      auto& tg = targets[src];
      if (std::find(tg.begin(), tg.end(), t) == tg.end()) tg.push_back(t);
      NRT_REQUIRE(tg.size() <= 3, NRT_ERR_TOPOLOGY,
                  "more than three sealed ends copy element " + std::to_string(src));
    }
    NRT_REQUIRE((uint32_t)N <= kCodeIdx, NRT_ERR_INVALID,
                "more elements than an element code can index");
    std::vector<uint32_t> code(N, 0);
    std::vector<ExRec> recs;
    std::vector<double> coefA(N, 0.0), rdx(N), wdx(N, 0.0), kdx(N, 0.0);
    for (int j = 0; j < N; j++) rdx[j] = p.r_a[j] * p.dx[j];
    for (int j = 0; j < N; j++) {
      const int hh = hh_of[j];
      const bool computed = j >= 1 && j <= N - 2 && !is_term[j];
      if (!computed) {
        code[j] = kCodeSkip | (hh >= 0 ? (kCodeHH | (uint32_t)hh) : 0u);
        continue;
      }
      const int mo = p.mother[j], dr = p.dr[j], dl = p.dl[j];
      NRT_REQUIRE(mo >= 0 && mo < N && dr >= 0 && dr < N, NRT_ERR_TOPOLOGY,
                  "element " + std::to_string(j) + " has an out-of-range neighbour");
      NRT_REQUIRE(!p.branching[j] || (dl >= 0 && dl < N), NRT_ERR_TOPOLOGY,
                  "branching element " + std::to_string(j) + " has an out-of-range DL");
      coefA[j] = p.dT / (p.c_m[j] * p.dx[j]);
      if (hh < 0) {
        wdx[j] = p.W[j] * p.dx[j];
        kdx[j] = p.K[j] * p.dx[j];
      }
      const bool plain = mo == j - 1 && dr == j + 1 && !p.branching[j] && p.input[j] == 0.0 &&
                         targets[j].empty();
      if (plain) {
        code[j] = hh >= 0 ? (kCodeHH | (uint32_t)hh) : kExPlain << 29;
        continue;
      }
      // one exception, passive, no stimulus: in the code, no record
      if (hh < 0 && p.input[j] == 0.0 && dr == j + 1) {
        const bool br = p.branching[j] != 0, own = mo == j - 1;
        if (!br && !own && targets[j].empty()) {
          code[j] = (kExDaughter << 29) | (uint32_t)mo;
          continue;
        }
        if (br && own && targets[j].empty()) {
          code[j] = (kExBranch << 29) | (uint32_t)dl;
          continue;
        }
        if (!br && own && targets[j].size() == 1 && targets[j][0] == j + 1) {
          code[j] = kExCopyNext << 29;
          continue;
        }
      }
      ExRec r{};
      r.mother = mo;
      r.dr = dr;
      r.dl = dl;
      r.hh = hh;
      for (int t = 0; t < 3; t++) r.tgt[t] = t < (int)targets[j].size() ? targets[j][t] : -1;
      r.flags = (p.branching[j] ? 1 : 0) | (p.input[j] != 0.0 ? 2 : 0);
      r.input = p.input[j];
      NRT_REQUIRE(recs.size() < kCodeIdx, NRT_ERR_INVALID, "too many exception records");
      code[j] = kCodeRec | (uint32_t)recs.size();
      recs.push_back(r);
    }
    e->R = (int)recs.size();
    e->V[0].upload(p.V0, s);  // prevT (solverEx-cpu.cpp:26-27)
    e->V[1].upload(p.V1, s);  // currT
    e->cur = 0;
    e->coefA.upload(coefA, s);
    e->rdx.upload(rdx, s);
    e->wdx.upload(wdx, s);
    e->kdx.upload(kdx, s);
    e->code.upload(code, s);
    e->recs.upload(recs, s);
    std::vector<double> hdx(H);
    for (int i = 0; i < H; i++) hdx[i] = p.dx[p.hh_elem[i]];
    e->m.upload(p.m, s);
    e->h.upload(p.h, s);
    e->n.upload(p.n, s);
    e->E_Na.upload(p.E_Na, s);
    e->E_K.upload(p.E_K, s);
    e->E_L.upload(p.E_L, s);
    e->g_Na.upload(p.g_Na, s);
    e->g_K.upload(p.g_K, s);
    e->G_L.upload(p.G_L, s);
    e->ls_Na.upload(p.ls_Na, s);
    e->ls_K.upload(p.ls_K, s);
    e->rest.upload(p.rest, s);
    e->hdx.upload(hdx, s);
    NRT_HIP(hipStreamSynchronize(s));
  } catch (...) {
    delete e;
    throw;
  }
  return e;
}

void explicit_destroy(Explicit* e) { delete e; }

static void launch_step(nrt_handle* hnd, Explicit* e, bool active) {
  const int block = 256;
  const double* Vo = e->V[e->cur].p;
  double* Vn = e->V[1 - e->cur].p;
  hipLaunchKernelGGL(k_explicit_fp64, dim3(grid_for(e->N, block)), dim3(block), 0, hnd->stream,
                     e->N, Vo, Vn, e->coefA.p, e->rdx.p, e->wdx.p, e->kdx.p, e->code.p, e->recs.p,
                     e->hd(), e->dt_ms, active ? 1 : 0);
  NRT_HIP(hipGetLastError());
  e->cur = 1 - e->cur;  // flip (solverEx-cpu.cpp:119-121)
}

void explicit_step(nrt_handle* hnd, bool active, double* trace_dst, int32_t trace_elem) {
  Explicit* e = hnd->ex;
  launch_step(hnd, e, active);
  if (trace_dst)
    NRT_HIP(hipMemcpyAsync(trace_dst, e->V[e->cur].p + trace_elem, sizeof(double),
                           hipMemcpyDeviceToDevice, hnd->stream));
}

void explicit_snapshot(nrt_handle* hnd, double* host_out) {
  Explicit* e = hnd->ex;
  NRT_HIP(hipMemcpyAsync(host_out, e->V[e->cur].p, sizeof(double) * e->N,
                         hipMemcpyDeviceToHost, hnd->stream));
  NRT_HIP(hipStreamSynchronize(hnd->stream));
}

void explicit_run(nrt_handle* hnd, int64_t first, int64_t num, int64_t a0, int64_t a1,
                  int32_t trace_elem, int64_t interval, double* trace_out, int64_t* ns) {
  Explicit* e = hnd->ex;
  const bool tracing = interval >= 0;
  const int64_t max_samples = tracing ? num / (interval + 1) : 0;
  if (tracing && trace_out && max_samples > e->trace_cap) {
    e->trace.alloc((size_t)max_samples);
    e->trace_cap = max_samples;
  }
  hipEvent_t t0, t1;
  NRT_HIP(hipEventCreate(&t0));
  NRT_HIP(hipEventCreate(&t1));
  NRT_HIP(hipEventRecord(t0, hnd->stream));
  int64_t k = 0;
  if (e->N <= kResMaxN) {
    // one launch for the whole run; the result lands in the other buffer, which
    // becomes the current one (the flips of num steps, then one more)
    const size_t lds = sizeof(double) * 2 * (size_t)e->N;
    if (!e->res_attr) {
      NRT_HIP(hipFuncSetAttribute((const void*)k_explicit_resident,
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(sizeof(double) * 2 * kResMaxN)));
      e->res_attr = true;
    }
    hipLaunchKernelGGL(k_explicit_resident, dim3(1), dim3(kResBlock), lds, hnd->stream, e->N,
                       e->V[e->cur].p, e->V[1 - e->cur].p, e->V[1 - e->cur].p, e->coefA.p,
                       e->rdx.p, e->wdx.p, e->kdx.p, e->code.p, e->recs.p, e->hd(), e->dt_ms,
                       (long long)first, (long long)num, (long long)a0, (long long)a1,
                       tracing ? trace_elem : 0, (long long)interval,
                       tracing && trace_out ? e->trace.p : nullptr);
    NRT_HIP(hipGetLastError());
    e->cur = 1 - e->cur;
    k = max_samples;
  } else {
    int64_t gl = 0;
    for (int64_t i = first; i < first + num; i++) {
      const bool active = i >= a0 && i <= a1;
      double* dst = nullptr;
      if (tracing && ++gl > interval) {
        gl = 0;
        if (trace_out) dst = e->trace.p + k;
        k++;
      }
      explicit_step(hnd, active, dst, trace_elem);
    }
  }
  NRT_HIP(hipEventRecord(t1, hnd->stream));
  if (trace_out && k > 0)
    NRT_HIP(hipMemcpyAsync(trace_out, e->trace.p, sizeof(double) * k, hipMemcpyDeviceToHost,
                           hnd->stream));
  NRT_HIP(hipStreamSynchronize(hnd->stream));
  float ms = 0;
  NRT_HIP(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  hnd->stats.kernel_ms = ms;
  hnd->stats.kernel_launches = e->N <= kResMaxN ? 1 : num;
  hnd->stats.steps += num;
  if (ns) *ns = k;
}

// ------------------------------------------------------------ batch support
// Slots of an explicit batch: one thread per element over all slots, one launch
// per step; blocks map to (slot, first element). The batch step counter lives on
// the device so a captured sequence of steps can be replayed.
constexpr int kExEPT = 1;  // elements per thread of the batch kernel (256 threads a block)

struct ExSlot {
  int N;
  int cur0;  // buffer holding V when the slot joined
  double* V[2];
  const double* coefA;
  const double* rdx;
  const double* wdx;
  const double* kdx;
  const uint32_t* code;
  const ExRec* recs;
  HHDev hd;
  double dt_ms;
  long long s0, num_steps, a0, a1, interval;
  int trace_elem;
  double* trace;
};

// Probe samples of the slots whose trace is due after `done` of their steps
// (done = batch steps completed - s0), from the buffer holding that step's V.
__device__ __forceinline__ void ex_batch_probes(const ExSlot* __restrict__ slots, int nslots,
                                                long long steps_completed, int t0, int nt) {
  for (int si = t0; si < nslots; si += nt) {
    const ExSlot& S = slots[si];
    const long long done = steps_completed - S.s0;
    if (S.interval < 0 || !S.trace || done < 1 || done > S.num_steps) continue;
    if (done % (S.interval + 1) != 0) continue;
    const int cur = (S.cur0 + (int)(done & 1)) & 1;
    S.trace[done / (S.interval + 1) - 1] = S.V[cur][S.trace_elem];
  }
}

// Batch step gbase[0] + i. Block 0 also records the probes due after the step
// before (whose V this step only reads), so a step is one launch; the probes of a
// call's last step are recorded by k_explicit_batch_trace after it.
__global__ __launch_bounds__(256) void k_explicit_batch(const ExSlot* __restrict__ slots,
                                                        const int2* __restrict__ blk,
                                                        const long long* __restrict__ gbase,
                                                        int i, int nslots) {
  const int2 b = blk[blockIdx.x];
  const ExSlot& S = slots[b.x];
  // the slot's fields are read before the one exit test (not short-circuit), so the
  // scalar loads are in flight together
  const long long s0 = S.s0, nst = S.num_steps, g = *gbase + i, a0 = S.a0, a1 = S.a1;
  if (blockIdx.x == 0 && i > 0) ex_batch_probes(slots, nslots, g, (int)threadIdx.x, 256);
  const int N = S.N, cur0 = S.cur0;
  const double* V0 = S.V[0];
  const double* V1 = S.V[1];
  const double* coefA = S.coefA;
  const double* rdx = S.rdx;
  const double* wdx = S.wdx;
  const double* kdx = S.kdx;
  const uint32_t* code = S.code;
  const ExRec* recs = S.recs;
  ready_s(V0);
  ready_s(V1);
  ready_s(coefA);
  ready_s(rdx);
  ready_s(wdx);
  ready_s(kdx);
  ready_s(code);
  ready_s(recs);
  const long long done = g - s0;
  const int j0 = b.y + (int)threadIdx.x;
  if ((done < 0) | (done >= nst) | (j0 >= N)) return;
  const bool odd = ((cur0 + (int)(done & 1)) & 1) != 0;
  const double* Vo = odd ? V1 : V0;
  double* Vn = odd ? (double*)V0 : (double*)V1;
  const bool active = done + 1 >= a0 && done + 1 <= a1;
  // kExEPT elements per thread, 256 apart: every element's loads are issued before
  // any is waited for (an element past the slot's end loads element j0 again)
  ExOps o[kExEPT];
#pragma unroll
  for (int u = 0; u < kExEPT; u++) {
    const int j = j0 + u * 256;
    o[u] = ex_load(j < N ? j : j0, N, Vo, coefA, rdx, wdx, kdx, code);
  }
#pragma unroll
  for (int u = 0; u < kExEPT; u++) ex_ready(o[u]);
#pragma unroll
  for (int u = 0; u < kExEPT; u++) {
    const int j = j0 + u * 256;
    if (j < N) ex_finish(j, o[u], Vo, Vn, rdx, recs, S.hd, S.dt_ms, active);
  }
}

int explicit_batch_block_elems() { return 256 * kExEPT; }

// After a call's last step (batch step gbase[0] + k - 1): its probe samples.
__global__ void k_explicit_batch_trace(const ExSlot* __restrict__ slots, int nslots,
                                       const long long* __restrict__ gbase, int k) {
  ex_batch_probes(slots, nslots, *gbase + k, blockIdx.x * blockDim.x + threadIdx.x,
                  gridDim.x * blockDim.x);
}

size_t explicit_slot_bytes() { return sizeof(ExSlot); }

void explicit_slot_fill(nrt_handle* h, long long s0, long long num_steps, long long a0,
                        long long a1, int trace_elem, long long interval, double* trace,
                        void* out) {
  Explicit* e = h->ex;
  ExSlot S{};
  S.N = e->N;
  S.cur0 = e->cur;
  S.V[0] = e->V[0].p;
  S.V[1] = e->V[1].p;
  S.coefA = e->coefA.p;
  S.rdx = e->rdx.p;
  S.wdx = e->wdx.p;
  S.kdx = e->kdx.p;
  S.code = e->code.p;
  S.recs = e->recs.p;
  S.hd = e->hd();
  S.dt_ms = e->dt_ms;
  S.s0 = s0;
  S.num_steps = num_steps;
  S.a0 = a0;
  S.a1 = a1;
  S.interval = interval;
  S.trace_elem = trace_elem;
  S.trace = trace;
  std::memcpy(out, &S, sizeof(S));
}

// k batch steps from batch step gbase[0]: one launch of the element kernel per step
// over `nblk` blocks, then the last step's probes.
void explicit_batch_steps(const void* slots, const int2* blk, int nblk, int nslots,
                          const long long* gbase, int k, hipStream_t s) {
  const auto* S = (const ExSlot*)slots;
  for (int i = 0; i < k; i++)
    if (nblk > 0)
      hipLaunchKernelGGL(k_explicit_batch, dim3(nblk), dim3(256), 0, s, S, blk, gbase, i, nslots);
  if (nslots > 0 && k > 0)
    hipLaunchKernelGGL(k_explicit_batch_trace, dim3(grid_for(nslots, 256)), dim3(256), 0, s, S,
                       nslots, gbase, k);
  NRT_HIP(hipGetLastError());
}

// Buffer holding V of a slot after `done` of its steps.
const double* explicit_slot_v(nrt_handle* h, long long done) {
  Explicit* e = h->ex;
  return e->V[(e->cur + (int)(done & 1)) & 1].p;
}

int64_t explicit_device_bytes(const Explicit* e) {
  return (int64_t)e->N * (8 * 6 + 4) + (int64_t)e->R * 40 + (int64_t)e->H * 8 * 13;
}

}  // namespace nrt
