// Implicit (backward-Euler) step, FP64 plain form.
//
// Matrix: the CSR of the reference GPU solver (solverIm-gpu.cpp:199-369, row types
// of :31-37), assembled on the device: a row-length kernel, an exclusive scan for
// the row pointers, then a fill kernel. HH rows are refreshed every step with the
// Update_Mat formulas (solverIm-gpu.cpp:103-156); the right-hand side follows
// Update_B (:162-171).
//
// Solve: BiCGSTAB with the reference's recurrences (solverIm-gpu.cpp:725-843), run
// as ONE persistent kernel per call (a co-resident grid): the HH gating, HH-row refresh and
// right-hand side of each time step, then every Krylov iteration, with device
// resident scalars and no host round trip. Grid-wide dependencies use a bounded
// grid barrier; each iteration needs three (one per global reduction). Vectors that
// a neighbour block reads while the owner rewrites them are double buffered, and the
// off-block entries of p and s that an SpMV needs are recomputed from their
// defining recurrences (p = r + beta*(p - omega*v), s = r - alpha*v) rather than
// exchanged, which removes two barriers per iteration.
//
// Dot products: per-wave DPP reduction, per-block LDS reduction, then every block
// sums the per-block partials in the same fixed order, so all blocks hold
// bit-identical scalars and take identical control-flow decisions.
//
// Layout in HBM (FP64 unless noted): x(=pot), b, r0, r, p[2], v[2], s, t (N each);
// rowptr/col (int32), val; per element cmdx = c_m*dx, kdx = K*dx, rdx = r_a*dx,
// input, code (uint32); HH SoA as in the explicit step plus row type/position/DL.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <type_traits>
#include "nrt_internal.hpp"
#include "nrt_device.hpp"
#include "tree.hpp"

namespace nrt {

enum : int32_t {
  RT_NORMAL = 1, RT_BRANCH = 2, RT_TERM = 4, RT_NEXT_TERM = 8, RT_FIRST_NORMAL = 16,
  RT_FIRST_BRANCH = 32, RT_OTHER = 64
};

struct ImpHH {
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
  const int32_t* type;
  const int32_t* mat;
  const int32_t* dl;
  const int32_t* row;  // element (row) of each HH entry
};

struct ImpStats {
  unsigned long long iters_total;
  unsigned long long sweeps_total;  // refinement sweeps (FP32 forms)
  unsigned long long steps;
  int last_iters;
  int status;
  double last_rr;
  double last_bb;
  // first failure: step, iteration and the scalars of that iteration
  long long fail_step;
  int fail_iter;
  int restarts;
  double fail_rho, fail_apr, fail_w1, fail_w2, fail_rr;
  double fail_bb;    // |b|^2 of the failing step (refinement: |D^-1 b|^2)
  int fail_cycles;   // GMRES restart cycles / refinement sweeps of the failing step
};

struct ImpDev {
  int N;
  int H;
  int nblocks;  // workgroups working on this tree (its "virtual grid")
  int vb;       // this workgroup's index in that grid
  int rpb;  // rows per block
  double dT;
  double dt_ms;
  const int32_t* rowptr;
  const int32_t* col;
  double* val;
  const double* cmdx;
  const double* kdx;
  const double* input;
  const double* rdx;
  const double* dxv;
  const uint32_t* code;
  ImpHH hh;
  double* x;
  double* b;
  double* r0;
  double* r;
  double* p0;
  double* p1;
  double* v0;
  double* v1;
  double* s;
  double* t;
  double* part;
  unsigned* bar;
  unsigned* err;
  ImpStats* st;
  double tol;
  int tol_mode;
  int x0_mode;
  int check_every;
  int max_iter;
  unsigned long long spin_limit;
  long long first, num, a0, a1, interval;
  long long gl0;  // trace phase at step `first` (steps since the last sample)
  int trace_elem;  // element (BiCGSTAB) or its level-0 slot (tree solve)
  double* trace;
  // tree solve: the state in level-0 slot order (ns0 slots, padding included)
  int ns0;
  const double* scmdx;
  const double* skdx;
  const double* sinput;
  const uint32_t* scode;    // HH flag and index, or kCodeSkip for padding
  const int32_t* srowptr;   // the FP64 matrix with rows in slot order
  const int32_t* scol;      // column slots
  const uint8_t* srole;     // 0 diagonal, 1 coupling to the parent, 2 to a child
  const int32_t* hh_slot;   // slot of each HH entry's row
  // the tree solve's work arrays already hold this plan's matrix and factors (a
  // batch slot after its first launch): the launch skips the fill and the full
  // factorization (static chunks keep theirs between launches, dynamic ones are
  // eliminated again by every step)
  int keep_factors;
};

struct Implicit {
  int N = 0, H = 0, nnz = 0, nblocks = 0;
  double dT = 0;
  DBuf<int32_t> rowptr, col, len, scan_tmp;
  DBuf<double> val;
  DBuf<double> cmdx, kdx, input, rdx, dxv, Wc;
  DBuf<int32_t> mother, dr, dl;
  DBuf<uint8_t> branching;
  DBuf<uint32_t> code;
  DBuf<double> m, h, n, E_Na, E_K, E_L, g_Na, g_K, G_L, ls_Na, ls_K, rest;
  DBuf<int32_t> hh_type, hh_mat, hh_dl, hh_row;
  DBuf<double> x, b, r0, r, p[2], v[2], s, t;
  DBuf<double> part;
  DBuf<unsigned> bar;
  DBuf<ImpStats> st;
  DBuf<double> trace;
  int64_t trace_cap = 0;
  int64_t step = 1;
  // tree partition solve (method != 0)
  int method = 0;        // 0 Krylov, 1 exact FP64 tree solve, 2 FP64 refinement over FP32
  bool gmres = false;    // Krylov method: GMRES(m) (else BiCGSTAB)
  int gm = 0;            // GMRES restart length
  DBuf<double> gV, gR, gpart;
  DBuf<float> gV32, gW32, val32;  // mixed-precision GMRES: FP32 basis, vectors, matrix
  bool grid_mode = false;
  bool local_mode = false;  // tree solve in local mode (plan_h.groups workgroups)
  bool local_res = false;   // ... in its group-resident form
  size_t local_lds = 0;     // dynamic LDS of a local-mode launch
  size_t forest_lds_max = 0;  // largest staged final forest of a block launch (bytes)
  DBuf<int32_t> hh_goff;
  int threads = 256;
  TreePlanHost plan_h;
  int ns0 = 0;  // level-0 slots (the tree solve's vectors are in slot order)
  DBuf<double> scmdx, skdx, sinput;
  DBuf<uint32_t> scode;
  DBuf<int32_t> srowptr, scol, hh_slot;
  DBuf<uint8_t> srole;
  DBuf<int32_t> plan_buf;
  TreePlanDev plan{};
  DBuf<double> tw64, tw64alt;  // alt: the second top-input buffer (group-resident form)
  DBuf<float> tw32, tw32alt;
};

// ------------------------------------------------------------------ assembly
// Row length per the reference GPU assembly (solverIm-gpu.cpp:205-368).
__global__ void k_row_len(int N, const int32_t* __restrict__ dr, const int32_t* __restrict__ dl,
                          int32_t* __restrict__ len) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int l;
  if (i == 0 || i == N - 1) l = 1;
  else if (i == 1) l = 2 + (dl[1] != 0 ? 1 : 0);
  else if (i == N - 2) l = 2;
  else if (dr[i] == i) l = 1;
  else if (dr[i + 1] == i + 1) l = 2;
  else l = 3 + (dl[i] != 0 ? 1 : 0);
  len[i] = l;
}

// Exclusive scan of int32 in three passes (per-block sums, scan of the block sums
// in one block, per-block local scans plus offsets). out has n+1 entries.
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(int n, const int32_t* in, int32_t* sums) {
  __shared__ int32_t sh[kScanBlock / kWave];
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  int v = i < n ? in[i] : 0;
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kScanBlock / kWave; w++) s += sh[w];
    sums[blockIdx.x] = s;
  }
}
__device__ __forceinline__ int block_excl_scan(int v, int* sh, int& total) {
  // inclusive wave scan by shuffles, then wave offsets through LDS
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  int x = v;
  for (int o = 1; o < kWave; o <<= 1) {
    int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) sh[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); w++) {
      int t = sh[w];
      sh[w] = acc;
      acc += t;
    }
    sh[kScanBlock / kWave] = acc;
  }
  __syncthreads();
  int res = x - v + sh[wid];
  total = sh[kScanBlock / kWave];
  __syncthreads();
  return res;
}
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(int nb, int32_t* sums) {
  __shared__ int sh[kScanBlock / kWave + 1];
  int carry = 0;
  for (int base = 0; base < nb; base += kScanBlock) {
    const int i = base + threadIdx.x;
    int v = i < nb ? sums[i] : 0, total;
    int e = block_excl_scan(v, sh, total);
    if (i < nb) sums[i] = e + carry;
    carry += total;
  }
}
__global__ __launch_bounds__(kScanBlock) void k_scan_final(int n, const int32_t* in,
                                                            const int32_t* sums, int32_t* out) {
  __shared__ int sh[kScanBlock / kWave + 1];
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  int v = i < n ? in[i] : 0, total;
  int e = block_excl_scan(v, sh, total);
  if (i < n) out[i] = e + sums[blockIdx.x];
  if (i == n - 1) out[n] = e + sums[blockIdx.x] + v;
}

// Fill cols/values (solverIm-gpu.cpp:199-369) and the HH row metadata.
__global__ void k_fill(int N, double dT, const int32_t* __restrict__ rowptr,
                       int32_t* __restrict__ col, double* __restrict__ val,
                       const int32_t* __restrict__ mother, const int32_t* __restrict__ dr,
                       const int32_t* __restrict__ dl, const uint8_t* __restrict__ branching,
                       const double* __restrict__ rdx, const double* __restrict__ cmdx,
                       const double* __restrict__ dxv, const double* __restrict__ W,
                       const uint32_t* __restrict__ code, int32_t* hh_type, int32_t* hh_mat,
                       int32_t* hh_dl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int k = rowptr[i];
  const uint32_t c = code[i];
  const int hh = (c & kCodeHH) ? (int)(c & kCodeIdx) : -1;
  int type;
  int dlv = -1;
  if (i == 0 || i == N - 1) {
    col[k] = i;
    val[k] = 1;
    type = RT_OTHER;
  } else if (i != N - 2 && i != 1 && dr[i] == i) {
    col[k] = i;
    val[k] = 1;
    type = RT_TERM;
  } else {
    const int DR = dr[i], DL = dl[i];
    const double alpha = -1 / rdx[i];
    const double betaR = 1 / rdx[DR];
    const double betaL = (i == N - 2) ? 0.0 : (double)branching[i] / rdx[DL];
    const double betaMe = -W[i] * dxv[i];
    const double beta = (i == N - 2) ? cmdx[i] / dT + betaMe - alpha + betaR
                                     : cmdx[i] / dT + betaMe - alpha + betaL + betaR;
    const double delta = -betaR;
    if (i == 1) {
      col[k] = 1;
      val[k] = alpha + beta;
      col[k + 1] = 2;
      val[k + 1] = delta;
      if (DL != 0) {
        col[k + 2] = DL;
        val[k + 2] = -betaL;
        type = RT_FIRST_BRANCH;
        dlv = DL;
      } else {
        type = RT_FIRST_NORMAL;
      }
    } else if (i == N - 2 || dr[i + 1] == i + 1) {
      col[k] = mother[i];
      val[k] = alpha;
      col[k + 1] = i;
      val[k + 1] = beta + delta;
      type = RT_NEXT_TERM;
    } else {
      col[k] = mother[i];
      val[k] = alpha;
      col[k + 1] = i;
      val[k + 1] = beta;
      col[k + 2] = DR;
      val[k + 2] = delta;
      if (DL != 0) {
        col[k + 3] = DL;
        val[k + 3] = -betaL;
        type = RT_BRANCH;
        dlv = DL;
      } else {
        type = RT_NORMAL;
      }
    }
  }
  if (hh >= 0) {
    hh_type[hh] = type;
    hh_mat[hh] = k;
    hh_dl[hh] = dlv;
  }
}

// FP32 copy of the matrix with every row divided by its diagonal entry.
__global__ void k_rows_scaled_f32(int N, const int32_t* __restrict__ rowptr,
                                  const int32_t* __restrict__ col, const double* __restrict__ val,
                                  float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double dg = 1.0;
  for (int k = rowptr[i]; k < rowptr[i + 1]; k++)
    if (col[k] == i) dg = val[k];
  for (int k = rowptr[i]; k < rowptr[i + 1]; k++) out[k] = (float)(val[k] / dg);
}

// --------------------------------------------------------- persistent kernel
constexpr int kImpBlock = 256;
// breakdown restarts allowed per time step before the step reports NOT_CONVERGED
constexpr int kMaxRestarts = 64;

// LDS-staged row tiles of the Krylov kernels' SpMV y_i = sum_k val[k] f(col[k]) over
// this workgroup's rows [r_begin, r_end): the operand f (a stored vector, or one
// recomputed from its recurrence) is evaluated once per row of a tile and of its
// one-row halo and kept in LDS, and the rows read their neighbours from there; a
// column outside the tile (a branch point's far daughter, a daughter's far
// mother) is evaluated again. kSpmvTile is a multiple of the workgroup size, so
// every row keeps its thread and every thread its order of rows (the partial
// sums of the reductions fused into `row` do not change), and the products are
// summed in the row's column order as before. row(i, y_i, f(i)) consumes own rows.
constexpr int kSpmvTile = 2048;
static_assert(kSpmvTile % kImpBlock == 0, "tile rows keep their thread");
template <class V, class VT, class F, class ROW>
__device__ __forceinline__ void spmv_tiles(int N, int r_begin, int r_end,
                                           const int32_t* __restrict__ rowptr,
                                           const int32_t* __restrict__ col, const VT* val, V* sx,
                                           F f, ROW row) {
  const int tid = (int)threadIdx.x, bs = (int)blockDim.x;
  for (int t0 = r_begin; t0 < r_end; t0 += kSpmvTile) {
    const int t1 = min(r_end, t0 + kSpmvTile);
    const int lo = max(0, t0 - 1), hi = min(N, t1 + 1);
    for (int i = lo + tid; i < hi; i += bs) sx[i - lo] = f(i);
    __syncthreads();
    for (int i = t0 + tid; i < t1; i += bs) {
      V acc = 0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const int c = col[k];
        const V xc = (c >= lo && c < hi) ? sx[c - lo] : f(c);
        acc = acc + val[k] * xc;
      }
      row(i, acc, sx[i - lo]);
    }
    __syncthreads();
  }
}

// HH gating (solverIm-gpu.cpp:61-98): the per-entry parameters, and one step of
// the gates at potential v (updates m, h, n; returns W and K).
struct HHGate {
  double rest, ls_Na, ls_K, g_Na, g_K, G_L, E_Na, E_K, E_L;
};
__device__ __forceinline__ HHGate hh_gate_load(const ImpHH& hd, int i) {
  return HHGate{hd.rest[i], hd.ls_Na[i], hd.ls_K[i], hd.g_Na[i], hd.g_K[i],
                hd.G_L[i],  hd.E_Na[i],  hd.E_K[i],  hd.E_L[i]};
}
// The rate functions (coupling.cpp:25-48) in precision R: FP64 for the FP64 forms;
// FP32 (expf, single-precision divisions) for the FP32 forms, whose gate state,
// conductances and currents stay FP64.
template <class R>
struct HHRates {
  double am, bm, ah, bh, an, bn;
};
template <class R>
__device__ __forceinline__ HHRates<R> hh_rates(const HHGate& q, double v) {
  double vna = v + q.ls_Na;
  double vk = v + q.ls_K;
  vna = (vna - q.rest) / 0.001;
  vk = (vk - q.rest) / 0.001;
  if constexpr (sizeof(R) == 4) {
    const float un = (float)vna, uk = (float)vk;
    return HHRates<R>{(2.5f - 0.1f * un) / (expf(2.5f - 0.1f * un) - 1.0f),
                      4.0f * expf(-un / 18.0f),
                      0.07f * expf(-un / 20.0f),
                      1.0f / (expf(3.0f - 0.1f * un) + 1.0f),
                      (0.1f - 0.01f * uk) / (expf(1.0f - 0.1f * uk) - 1.0f),
                      0.125f * expf(-uk / 80.0f)};
  } else {
    return HHRates<R>{(2.5 - 0.1 * vna) / (exp(2.5 - 0.1 * vna) - 1), 4 * exp(-vna / 18),
                      0.07 * exp(-vna / 20), 1 / (exp(3 - 0.1 * vna) + 1),
                      (0.1 - 0.01 * vk) / (exp(1 - 0.1 * vk) - 1), 0.125 * exp(-vk / 80)};
  }
}
template <class R = double>
__device__ __forceinline__ void hh_gate_step(const HHGate& q, double v, double dt_ms, double& m,
                                             double& h, double& n, double& W, double& K) {
  const HHRates<R> rt = hh_rates<R>(q, v);
  const double am = rt.am, bm = rt.bm, ah = rt.ah, bh = rt.bh, an = rt.an, bn = rt.bn;
  const double m0 = m, h0 = h, n0 = n;
  m = m0 + dt_ms * (am * (1 - m0) - bm * m0);
  h = h0 + dt_ms * (ah * (1 - h0) - bh * h0);
  n = n0 + dt_ms * (an * (1 - n0) - bn * n0);
  const double gna = q.g_Na * m * m * m * h;
  const double gk = q.g_K * n * n * n * n;
  const double gl = q.G_L;
  W = -(gna + gk + gl);
  K = gna * q.E_Na + gk * q.E_K + gl * q.E_L;
}
template <class R = double>
__device__ __forceinline__ void imp_gate(const ImpHH& hd, int i, double v, double dt_ms,
                                         double& W, double& K) {
  double m = hd.m[i], h = hd.h[i], n = hd.n[i];
  hh_gate_step<R>(hh_gate_load(hd, i), v, dt_ms, m, h, n, W, K);
  hd.m[i] = m;
  hd.h[i] = h;
  hd.n[i] = n;
}

// Update_Mat (solverIm-gpu.cpp:103-156) for the HH row of element iv: the row's
// fixed coefficients, and its refresh for a new W (returns false for identity
// rows, left as they are, else true with the new diagonal).
struct HHRow {
  int type, im;
  double alpha, inv_rdx, betaR, betaL, dxv, c_dt;
};
__device__ __forceinline__ HHRow hh_row_load(const ImpDev& d, int hh, int iv) {
  HHRow q{};
  q.type = d.hh.type[hh];
  // identity rows (ghosts, sealed ends) keep their 1 and read no neighbour
  if (q.type == RT_OTHER || q.type == RT_TERM) return q;
  q.im = d.hh.mat[hh];
  q.alpha = -1 / d.rdx[iv];
  q.inv_rdx = 1 / d.rdx[iv];
  q.betaR = 1 / d.rdx[iv + 1];
  q.dxv = d.dxv[iv];
  q.c_dt = d.cmdx[iv] / d.dT;
  if (q.type == RT_BRANCH || q.type == RT_FIRST_BRANCH) q.betaL = 1 / d.rdx[d.hh.dl[hh]];
  return q;
}
__device__ __forceinline__ bool hh_row_refresh(const HHRow& q, double W, double* v,
                                               double& diag) {
  if (q.type == RT_OTHER || q.type == RT_TERM) return false;
  const int im = q.im;
  const double alpha = q.alpha;
  const double betaR = q.betaR;
  const double betaMe = -(W * q.dxv) + q.inv_rdx;
  const double beta = q.c_dt + betaMe + betaR;
  const double delta = -betaR;
  const double betaL = q.betaL;
  switch (q.type) {
    case RT_NEXT_TERM:
      diag = beta + delta;
      v[im] = alpha;
      v[im + 1] = diag;
      return true;
    case RT_BRANCH:
      diag = beta + betaL;
      v[im] = alpha;
      v[im + 1] = diag;
      v[im + 2] = delta;
      v[im + 3] = -betaL;
      return true;
    case RT_NORMAL:
      diag = beta;
      v[im] = alpha;
      v[im + 1] = diag;
      v[im + 2] = delta;
      return true;
    case RT_FIRST_NORMAL:
      diag = alpha + beta;
      v[im] = diag;
      v[im + 1] = delta;
      return true;
    case RT_FIRST_BRANCH:
      diag = alpha + beta + betaL;
      v[im] = diag;
      v[im + 1] = delta;
      v[im + 2] = -betaL;
      return true;
    default:
      return false;
  }
}
__device__ __forceinline__ bool imp_update_row(const ImpDev& d, int hh, int iv, double W,
                                               double& diag) {
  return hh_row_refresh(hh_row_load(d, hh, iv), W, d.val, diag);
}

// Grid-wide sum of NV values: block partials go to one of two alternating slots,
// one barrier, then every block sums all partials in the same order. A block can
// only refill a slot after passing the next reduction's barrier, by which time
// every block has read that slot, so one barrier per reduction suffices.
// Grid-wide maximum of NV non-negative values, as reduce_all.
template <int NV>
__device__ __forceinline__ bool reduce_all_max(const ImpDev& d, double (&v)[NV], double* lds,
                                               int& slot) {
  block_max<NV>(v, lds);
  double* part = d.part + (size_t)slot * d.nblocks * 4;
  slot ^= 1;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) part[d.vb * 4 + k] = v[k];
  }
  grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
  grid_partials_max<NV>(part, 4, d.nblocks, v, lds);
  return __hip_atomic_load(d.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
}

template <int NV>
__device__ __forceinline__ bool reduce_all(const ImpDev& d, double (&v)[NV], double* lds,
                                           int& slot) {
  block_sum<NV>(v, lds);
  double* part = d.part + (size_t)slot * d.nblocks * 4;
  slot ^= 1;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) part[d.vb * 4 + k] = v[k];
  }
  grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
  grid_partials_sum<NV>(part, 4, d.nblocks, v, lds);
  return __hip_atomic_load(d.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
}

__global__ __launch_bounds__(kImpBlock) void k_implicit_bicgstab_fp64(ImpDev d) {
  __shared__ double lds[4 * 16];
  __shared__ double sx[kSpmvTile + 2];
  d.vb = (int)blockIdx.x;
  const int r_begin = blockIdx.x * d.rpb;
  const int r_end = min(d.N, r_begin + d.rpb);
  const double* val = d.val;  // HH rows are rewritten through d.val in the update phase
  const int32_t* __restrict__ rowptr = d.rowptr;
  const int32_t* __restrict__ col = d.col;
  long long gl = 0, ksamp = 0;
  int pb = 0, slot = 0;
  for (long long step = d.first; step < d.first + d.num; step++) {
    const bool active = step >= d.a0 && step <= d.a1;
    // ---- update phase: gating, HH rows, right-hand side, initial guess
    for (int i = r_begin + threadIdx.x; i < r_end; i += kImpBlock) {
      const double v = d.x[i];
      const uint32_t c = d.code[i];
      double kd;
      if (c & kCodeHH) {
        const int hh = (int)(c & kCodeIdx);
        double W, K;
        imp_gate(d.hh, hh, v, d.dt_ms, W, K);
        double diag;
        (void)imp_update_row(d, hh, i, W, diag);
        kd = K * d.dxv[i];
      } else {
        kd = d.kdx[i];
      }
      double bi = d.cmdx[i] * v / d.dT + kd;
      if (active) bi = bi + d.input[i];
      d.b[i] = bi;
      if (d.x0_mode == NRT_X0_RHS) d.x[i] = bi;
    }
    grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
    // ---- r0 = b - A x0 ; r = r0 ; rho = r0.r0 ; bb = b.b
    double red2[2] = {0.0, 0.0};
    spmv_tiles<double>(
        d.N, r_begin, r_end, rowptr, col, val, sx, [&](int c) { return d.x[c]; },
        [&](int i, double ax, double) {
          const double bi = d.b[i];
          const double ri = bi + (-1.0) * ax;
          d.r0[i] = ri;
          d.r[i] = ri;
          red2[0] = red2[0] + ri * ri;
          red2[1] = red2[1] + bi * bi;
        });
    bool ok = reduce_all<2>(d, red2, lds, slot);
    double rho = red2[0], rr = red2[0];
    const double bb = red2[1];
    const double thr = d.tol_mode == NRT_TOL_RELATIVE ? d.tol * sqrt(bb) : d.tol;
    const double limit = thr * thr;
    int status = (rr <= limit) ? NRT_OK : NRT_ERR_NOT_CONVERGED;
    int it = 0, restarts = 0;
    double alpha = 0, omega = 0, beta = 0;
    double f_apr = 0, f_w1 = 0, f_w2 = 0;
    bool first_it = true, restart = false;
    while (ok && status == NRT_ERR_NOT_CONVERGED && it < d.max_iter) {
      double* __restrict__ pold = pb ? d.p1 : d.p0;
      double* __restrict__ pnew = pb ? d.p0 : d.p1;
      double* __restrict__ vold = pb ? d.v1 : d.v0;
      double* __restrict__ vnew = pb ? d.v0 : d.v1;
      const double nomega = -omega;
      // ---- phase A: p = r + beta*(p - omega*v) ; v = A p ; apr = v.r0
      // (after a breakdown of rho the shadow residual restarts as r0 = r)
      double redA[1] = {0.0};
      spmv_tiles<double>(
          d.N, r_begin, r_end, rowptr, col, val, sx,
          [&](int c) { return first_it ? d.r[c] : d.r[c] + beta * (pold[c] + nomega * vold[c]); },
          [&](int i, double acc, double pi) {
            if (restart) d.r0[i] = d.r[i];
            pnew[i] = pi;
            vnew[i] = acc;
            redA[0] = redA[0] + acc * d.r0[i];
          });
      restart = false;
      ok = reduce_all<1>(d, redA, lds, slot);
      if (!ok) break;
      f_apr = redA[0];
      if (!isfinite(redA[0])) {
        status = NRT_ERR_NUMERIC;
        break;
      }
      if (redA[0] == 0.0 || !isfinite(rho / redA[0])) {
        // Lanczos breakdown (v orthogonal to the shadow residual): restart
        it++;
        if (++restarts > kMaxRestarts) break;
        restart = true;
        first_it = true;
        rho = rr;
        continue;
      }
      alpha = rho / redA[0];
      const double nalpha = -(rho / redA[0]);
      // ---- phase B: s = r - alpha*v ; t = A s ; w1 = t.s ; w2 = t.t ; ss = s.s
      double redB[3] = {0.0, 0.0, 0.0};
      spmv_tiles<double>(
          d.N, r_begin, r_end, rowptr, col, val, sx,
          [&](int c) { return d.r[c] + nalpha * vnew[c]; },
          [&](int i, double acc, double si) {
            d.s[i] = si;
            d.t[i] = acc;
            redB[0] = redB[0] + acc * si;
            redB[1] = redB[1] + acc * acc;
            redB[2] = redB[2] + si * si;
          });
      ok = reduce_all<3>(d, redB, lds, slot);
      if (!ok) break;
      f_w1 = redB[0];
      f_w2 = redB[1];
      it++;
      if (!isfinite(redB[2]) || !isfinite(redB[1])) {
        status = NRT_ERR_NUMERIC;
        break;
      }
      if (redB[2] <= limit || redB[1] == 0.0) {
        // converged at the half step: x += alpha*p, r = s
        for (int i = r_begin + threadIdx.x; i < r_end; i += kImpBlock) {
          d.x[i] = d.x[i] + alpha * pnew[i];
          d.r[i] = d.s[i];
        }
        rr = redB[2];
        status = NRT_OK;
        break;
      }
      omega = redB[0] / redB[1];
      const double nw = -(redB[0] / redB[1]);
      // ---- phase C: x += alpha*p + omega*s ; r = s - omega*t ; p1 = r.r0 ; rr = r.r
      double redC[2] = {0.0, 0.0};
      for (int i = r_begin + threadIdx.x; i < r_end; i += kImpBlock) {
        const double si = d.s[i];
        double xi = d.x[i] + alpha * pnew[i];
        xi = xi + omega * si;
        d.x[i] = xi;
        const double ri = si + nw * d.t[i];
        d.r[i] = ri;
        redC[0] = redC[0] + ri * d.r0[i];
        redC[1] = redC[1] + ri * ri;
      }
      ok = reduce_all<2>(d, redC, lds, slot);
      if (!ok) break;
      beta = (alpha / omega) * (redC[0] / rho);
      rho = redC[0];
      rr = redC[1];
      pb ^= 1;
      first_it = false;
      if (!isfinite(rr)) {
        status = NRT_ERR_NUMERIC;
        break;
      }
      if ((it % d.check_every) == 0 && rr <= limit) {
        status = NRT_OK;
        break;
      }
      if (!isfinite(beta) || rho == 0.0 || omega == 0.0) {
        // breakdown of the bi-orthogonality: restart with r0 = r, p = r
        if (++restarts > kMaxRestarts) break;
        restart = true;
        first_it = true;
        rho = rr;
      }
    }
    if (!ok) status = NRT_ERR_HIP;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      d.st->iters_total += (unsigned long long)it;
      d.st->steps += 1;
      d.st->last_iters = it;
      d.st->last_rr = rr;
      d.st->last_bb = bb;
      d.st->restarts += restarts;
      if (d.st->status == NRT_OK && status != NRT_OK) {
        d.st->status = status;
        d.st->fail_step = step;
        d.st->fail_iter = it;
        d.st->fail_rho = rho;
        d.st->fail_apr = f_apr;
        d.st->fail_w1 = f_w1;
        d.st->fail_w2 = f_w2;
        d.st->fail_rr = rr;
        d.st->fail_bb = bb;
        d.st->fail_cycles = restarts;
      }
    }
    if (!ok) return;
    // ---- trace sample (neurite.cpp:241-249 cadence)
    if (d.interval >= 0 && ++gl > d.interval) {
      gl = 0;
      if (d.trace && d.trace_elem >= r_begin && d.trace_elem < r_end && threadIdx.x == 0)
        d.trace[ksamp] = d.x[d.trace_elem];
      ksamp++;
    }
    // the next step's update phase reads only rows this block owns
  }
}

// ------------------------------------------------------------------ GMRES(m)
// Restarted GMRES (the method of the reference CPU solver, MKL FGMRES(150) with
// x0 = 0 and an absolute 1e-14 stop, solverIm-cpu.cpp:192-272) as one persistent
// kernel on a co-resident grid: per step the same update phase as the BiCGSTAB kernel, then
// Arnoldi cycles of up to m steps. Orthogonalisation is classical Gram-Schmidt
// with one re-orthogonalisation (CGS2): each Arnoldi step takes one SpMV and three
// grid-wide reductions (the j+1 projections, the j+1 corrections, the norm), where
// modified Gram-Schmidt would take j+1 sequential ones. Every workgroup receives
// the same Hessenberg column (the reductions are bit-identical everywhere),
// applies the Givens rotations itself and keeps its own copy of the triangular
// factor, so the small least-squares problem needs no barrier of its own. The
// SpMV reads a neighbour's entries of v_{k+1} = w_k / h_{k+1,k} by recomputing
// them from w_k, double buffered by step parity, instead of waiting for them.
//
// T = float is the mixed-precision form: every cycle starts from the FP64
// residual r = b - A x (FP64 matrix values) and ends with the FP64 update
// x += V y, and the Arnoldi work between them (basis, SpMV, Gram-Schmidt) runs in
// FP32 with compensated inner products. A cycle is then one sweep of iterative
// refinement around an FP32 GMRES solve of the Jacobi-scaled system
// D^-1 A d = D^-1 r (D = diag A; the FP32 matrix copy holds the scaled rows): the
// unit rows of ghosts and sealed ends beside physical rows of ~1e-9 S give A a
// condition number near 1e9, beyond what FP32 refinement converges on, and the
// scaled system's is near 1e3. A cycle ends once its residual estimate has fallen
// by kGmresInnerRel32 (the reach of FP32 arithmetic) or after m steps; only the
// FP64 residual at the next cycle's start decides convergence, in the same scaled
// norm: |D^-1 (b - A x)| <= refine_tol |D^-1 b|. The Hessenberg column, the Givens
// rotations and the least-squares solve stay FP64 in both forms.
constexpr int kGmresMaxRestart = 512;
constexpr double kGmresInnerRel32 = 1e-6;

template <class T>
struct GmresDev {
  int m;           // restart length
  T* V;            // (m+1) x N basis (v_k at V + k N; only a block's own rows)
  T* w0;           // Arnoldi vectors, even and odd steps
  T* w1;
  double* R;       // per workgroup: packed R (m(m+1)/2), cs, sn (m), g (m+1), y (m)
  int64_t rstride;
  double* part;    // 2 sets x nblocks x (m+1) partial sums
  T* vals;         // matrix values of the Arnoldi SpMV (FP64: the matrix itself)
  double refine_tol;  // FP32: convergence of the FP64 residual, relative to |b|
};

// Inner-product accumulators of one thread. FP64: plain sums. FP32: the product
// error (one FMA) and the addition error (Knuth two-sum) are carried in a second
// word, so a thread's partial is accurate to about twice FP32 precision before the
// cross-thread sums, which run in FP64.
struct DotAcc64 {
  double s = 0;
  __device__ __forceinline__ void add(double a, double b) { s = s + a * b; }
  __device__ __forceinline__ double sum() const { return s; }
};
struct DotAcc32 {
  float s = 0, c = 0;
  __device__ __forceinline__ void add(float a, float b) {
#pragma clang fp contract(off)
    const float p = a * b;
    const float pe = fmaf(a, b, -p);  // a*b - p exactly
    float t, e;
    two_sum(s, p, t, e);
    s = t;
    c = c + (e + pe);
  }
  __device__ __forceinline__ double sum() const { return (double)s + (double)c; }
};
template <class T>
using DotAcc = typename std::conditional<std::is_same<T, float>::value, DotAcc32, DotAcc64>::type;

// h[0..cnt) = V_j . w over the whole vector (one grid barrier); every block ends
// with the same values in hsh.
template <class T>
__device__ __forceinline__ void gmres_dots(const ImpDev& d, const GmresDev<T>& g, int cnt,
                                           const T* __restrict__ w, int r_begin, int r_end,
                                           double* hsh, double* lds, int& gslot) {
  const int stride = g.m + 1;
  double* part = g.part + (size_t)gslot * d.nblocks * stride;
  gslot ^= 1;
  const size_t N = (size_t)d.N;
  for (int k0 = 0; k0 < cnt; k0 += 8) {
    DotAcc<T> acc[8];
    for (int i = r_begin + (int)threadIdx.x; i < r_end; i += (int)blockDim.x) {
      const T wi = w[i];
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (k0 + j < cnt) acc[j].add(g.V[(size_t)(k0 + j) * N + i], wi);
    }
    double sums[8];
#pragma unroll
    for (int j = 0; j < 8; j++) sums[j] = acc[j].sum();
    block_sum<8>(sums, lds);
    if (threadIdx.x == 0)
      for (int j = 0; j < 8 && k0 + j < cnt; j++) part[d.vb * stride + k0 + j] = sums[j];
  }
  grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    double sum = 0;
    for (int b = 0; b < d.nblocks; b++) sum = sum + part[b * stride + t];
    hsh[t] = sum;
  }
  __syncthreads();
}

template <class T>
__global__ __launch_bounds__(kImpBlock) void k_implicit_gmres(ImpDev d, GmresDev<T> g) {
  constexpr bool F32 = std::is_same<T, float>::value;
  __shared__ double lds[8 * 16];
  __shared__ double hsh[kGmresMaxRestart + 1];
  __shared__ double h2sh[kGmresMaxRestart + 1];
  __shared__ double ysh[kGmresMaxRestart];
  __shared__ double fsh[2];
  __shared__ double sx[kSpmvTile + 2];
  __shared__ T sxt[kSpmvTile + 2];
  d.vb = (int)blockIdx.x;
  const int r_begin = blockIdx.x * d.rpb;
  const int r_end = min(d.N, r_begin + d.rpb);
  const int bs = kImpBlock;
  const double* val = d.val;
  const T* vals = g.vals;
  const int32_t* __restrict__ rowptr = d.rowptr;
  const int32_t* __restrict__ col = d.col;
  const size_t N = (size_t)d.N;
  double* Rb = g.R + (size_t)blockIdx.x * g.rstride;  // this block's least-squares data
  const int m = g.m;
  double* Rp = Rb;
  double* cs = Rb + (size_t)m * (m + 1) / 2;
  double* sn = cs + m;
  double* gv = sn + m;
  double* yv = gv + m + 1;
  long long gl = 0, ksamp = 0;
  int slot = 0, gslot = 0;
  for (long long step = d.first; step < d.first + d.num; step++) {
    const bool active = step >= d.a0 && step <= d.a1;
    // ---- update phase: gating, HH rows, right-hand side, initial guess
    for (int i = r_begin + threadIdx.x; i < r_end; i += bs) {
      const double v = d.x[i];
      const uint32_t c = d.code[i];
      double kd;
      if (c & kCodeHH) {
        const int hh = (int)(c & kCodeIdx);
        double W, K;
        imp_gate<T>(d.hh, hh, v, d.dt_ms, W, K);
        double diag;
        if (imp_update_row(d, hh, i, W, diag) && F32)
          for (int k = rowptr[i]; k < rowptr[i + 1]; k++) g.vals[k] = (T)(val[k] / diag);
        kd = K * d.dxv[i];
      } else {
        kd = d.kdx[i];
      }
      double bi = d.cmdx[i] * v / d.dT + kd;
      if (active) bi = bi + d.input[i];
      d.b[i] = bi;
      if (d.x0_mode == NRT_X0_RHS) d.x[i] = bi;
      else if (d.x0_mode == NRT_X0_ZERO) d.x[i] = 0.0;
    }
    grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
    int it = 0, status = NRT_ERR_NOT_CONVERGED, cycles = 0;
    double rr = 0, bb = 0, thr = 0;
    bool ok = true;
    while (ok) {
      // ---- cycle start: r0 = b - A x, beta = |r0| (FP64 in both forms; FP32
      // cycles work on the Jacobi-scaled system, r0 = D^-1 (b - A x))
      double red[2] = {0.0, 0.0};
      spmv_tiles<double>(
          d.N, r_begin, r_end, rowptr, col, val, sx, [&](int c) { return d.x[c]; },
          [&](int i, double ax, double) {
            double dg = 1.0;
            if (F32)
              for (int k = rowptr[i]; k < rowptr[i + 1]; k++)
                if (col[k] == i) dg = val[k];
            double bi = d.b[i];
            double ri = bi - ax;
            if (F32) {
              ri = ri / dg;
              bi = bi / dg;
            }
            d.r0[i] = ri;
            red[0] = red[0] + ri * ri;
            red[1] = red[1] + bi * bi;
          });
      ok = reduce_all<2>(d, red, lds, slot);
      if (!ok) break;
      rr = red[0];
      if (cycles == 0) {
        bb = red[1];
        if (F32) thr = g.refine_tol * sqrt(bb);
        else thr = d.tol_mode == NRT_TOL_RELATIVE ? d.tol * sqrt(bb) : d.tol;
      }
      if (!isfinite(rr)) {
        status = NRT_ERR_NUMERIC;
        break;
      }
      const double beta = sqrt(rr);
      if (beta <= thr) {
        status = NRT_OK;
        break;
      }
      if (it >= d.max_iter) break;
      cycles++;
      if (threadIdx.x == 0) {
        gv[0] = beta;
        for (int j = 1; j <= m; j++) gv[j] = 0.0;
      }
      // an FP32 cycle stops where FP32 stops resolving its own residual
      const double stop = F32 ? fmax(thr, kGmresInnerRel32 * beta) : thr;
      // v_k = src / div, neighbours' entries recomputed: v_0 from the FP64
      // residual, v_k (k > 0) from the previous step's Arnoldi vector
      const T* prev = nullptr;
      double div = beta;
      int kk = 0;
      bool conv = false;
      for (int k = 0; k < m; k++) {
        T* __restrict__ wk = (k & 1) ? g.w1 : g.w0;
        const T divT = (T)div;
        auto vk = [&](int c) -> T { return k == 0 ? (T)(d.r0[c] / div) : prev[c] / divT; };
        it++;
        // v_k (own rows) and w = A v_k
        spmv_tiles<T>(d.N, r_begin, r_end, rowptr, col, vals, sxt, vk, [&](int i, T acc, T vi) {
          g.V[(size_t)k * N + i] = vi;
          wk[i] = acc;
        });
        // CGS2: h = V^T w; w -= V h; h2 = V^T w; w -= V h2; h += h2
        gmres_dots<T>(d, g, k + 1, wk, r_begin, r_end, hsh, lds, gslot);
        for (int i = r_begin + threadIdx.x; i < r_end; i += bs) {
          T wi = wk[i];
          for (int j = 0; j <= k; j++) wi = wi - (T)hsh[j] * g.V[(size_t)j * N + i];
          wk[i] = wi;
        }
        gmres_dots<T>(d, g, k + 1, wk, r_begin, r_end, h2sh, lds, gslot);
        DotAcc<T> na;
        for (int i = r_begin + threadIdx.x; i < r_end; i += bs) {
          T wi = wk[i];
          for (int j = 0; j <= k; j++) wi = wi - (T)h2sh[j] * g.V[(size_t)j * N + i];
          wk[i] = wi;
          na.add(wi, wi);
        }
        double nrm[1] = {na.sum()};
        ok = reduce_all<1>(d, nrm, lds, slot);
        if (!ok) break;
        const double hn = sqrt(nrm[0]);
        // Givens rotations on the new column (every block, identical values)
        if (threadIdx.x == 0) {
          double* Rk = Rp + (size_t)k * (k + 1) / 2;  // column k, rows 0..k
          for (int j = 0; j <= k; j++) Rk[j] = hsh[j] + h2sh[j];
          double hk1 = hn;
          for (int j = 0; j < k; j++) {
            const double a = Rk[j], b2 = Rk[j + 1];
            Rk[j] = cs[j] * a + sn[j] * b2;
            Rk[j + 1] = -sn[j] * a + cs[j] * b2;
          }
          const double a = Rk[k];
          const double den = sqrt(a * a + hk1 * hk1);
          cs[k] = den == 0 ? 1.0 : a / den;
          sn[k] = den == 0 ? 0.0 : hk1 / den;
          Rk[k] = den;
          gv[k + 1] = -sn[k] * gv[k];
          gv[k] = cs[k] * gv[k];
          fsh[0] = fabs(gv[k + 1]);
          fsh[1] = hn;
        }
        __syncthreads();
        const double res = fsh[0];
        kk = k + 1;
        if (!isfinite(res)) {
          status = NRT_ERR_NUMERIC;
          ok = false;
          break;
        }
        if (res <= stop || fsh[1] == 0.0) {
          conv = true;
          rr = res * res;
          break;
        }
        if (it >= d.max_iter) break;
        prev = wk;
        div = hn;
      }
      if (status == NRT_ERR_NUMERIC) {
        ok = true;
        break;
      }
      if (!ok) break;
      // ---- x += V y, R y = g (back substitution, every block; FP64 update)
      if (threadIdx.x == 0) {
        for (int j = kk - 1; j >= 0; j--) {
          double acc = gv[j];
          for (int l = j + 1; l < kk; l++) acc = acc - Rp[(size_t)l * (l + 1) / 2 + j] * yv[l];
          yv[j] = acc / Rp[(size_t)j * (j + 1) / 2 + j];
        }
        for (int j = 0; j < kk; j++) ysh[j] = yv[j];
      }
      __syncthreads();
      for (int i = r_begin + threadIdx.x; i < r_end; i += bs) {
        double xi = d.x[i];
        for (int j = 0; j < kk; j++) xi = xi + ysh[j] * (double)g.V[(size_t)j * N + i];
        d.x[i] = xi;
      }
      grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);  // x before the next residual
      if (__hip_atomic_load(d.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        ok = false;
        break;
      }
      // FP64: the cycle's residual estimate decides; FP32: the next FP64 residual
      if (conv && !F32) {
        status = NRT_OK;
        break;
      }
      if (it >= d.max_iter) break;
    }
    if (!ok) status = NRT_ERR_HIP;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      d.st->iters_total += (unsigned long long)it;
      d.st->sweeps_total += (unsigned long long)cycles;
      d.st->steps += 1;
      d.st->last_iters = it;
      d.st->last_rr = rr;
      d.st->last_bb = bb;
      d.st->restarts += cycles > 0 ? cycles - 1 : 0;
      if (d.st->status == NRT_OK && status != NRT_OK) {
        d.st->status = status;
        d.st->fail_step = step;
        d.st->fail_iter = it;
        d.st->fail_rr = rr;
        d.st->fail_bb = bb;
        d.st->fail_cycles = cycles;
      }
    }
    if (!ok) return;
    if (d.interval >= 0 && ++gl > d.interval) {
      gl = 0;
      if (d.trace && d.trace_elem >= r_begin && d.trace_elem < r_end && threadIdx.x == 0)
        d.trace[ksamp] = d.x[d.trace_elem];
      ksamp++;
    }
  }
}

// --------------------------------------------- tree partition solve kernels
// One persistent kernel per call, as the BiCGSTAB one, with the Krylov loop
// replaced by the exact tree partition solve (tree.hpp): either one FP64 solve
// per step (direct), or FP64 iterative refinement around FP32 solves whose
// residual b - A x and update x += d are the only FP64 vector work. Block mode:
// one workgroup owns the whole tree and synchronises with workgroup barriers;
// grid mode (large trees): every workgroup takes a share of the rows and chunks,
// grid barriers separate the levels, and workgroup 0 solves the compressed forest.
constexpr int kTreeBlock = 1024;
// Largest staged final forest (dynamic LDS of one workgroup, block launches), within
// the device's limit per workgroup.
constexpr size_t kForestLdsMax = 160 * 1024;
constexpr int kTreeArrays = 11;  // d, up, dn, r, x, ca, cb, cg, cf, ci, cw

template <class T>
struct TreeArgs {
  TreePlanDev P;
  TreeWork<T> w;
  int direct;
  int krylov;   // BiCGSTAB preconditioned by this tree solve (else direct / refinement)
  double refine_tol;
  int refine_max;
  int top_lds;  // the launch provides dynamic LDS: top_lds_bytes<T>() (grid launches) or
                // forest_lds_bytes<T>(P) (block launches)
  const int32_t* hh_goff;  // local mode: HH entries of group g are [hh_goff[g], +1)
  TreeWork<T> walt;        // group-resident form: the second top-input buffer
};

template <bool GRID>
__device__ __forceinline__ void tsync(const ImpDev& d) {
  if (GRID) grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
  else __syncthreads();
}

// Full elimination of every chunk: the factors that the steps reuse for chunks
// whose matrix values do not change (run once per launch, after the level-0 fill).
template <class T, bool GRID>
__device__ void tps_factor(const ImpDev& d, const TreePlanDev& P, const TreeWork<T>& w) {
  const int wid = GRID ? d.vb * (int)blockDim.x + (int)threadIdx.x : (int)threadIdx.x;
  const int nw = GRID ? d.nblocks * (int)blockDim.x : (int)blockDim.x;
  for (int lv = 0; lv < P.levels; lv++) {
    const int c1 = P.chunk_off[lv + 1];
    for (int ch = P.chunk_off[lv] + wid; ch < c1; ch += nw) tps_reduce_chunk(P, w, lv, ch);
    tsync<GRID>(d);
  }
}

// One solve. Dynamic chunks (holding a row whose values changed this step) are
// eliminated in full, the others sweep the right-hand side with their kept
// factors; all_static: the matrix did not change since the last solve (a later
// refinement sweep of the same step), every chunk reuses its factors.
// rs (direct solves): the right-hand side of the rows of static level-0 chunks is
// computed from the state as they are swept (the update phase wrote only the rows
// of P.dyn0).
template <class T, bool GRID>
__device__ __forceinline__ void tps_run(const ImpDev& d, const TreePlanDev& P,
                                        const TreeWork<T>& w, bool acc,
                        bool all_static, unsigned char* tl, double* out,
                        const TpsRhsUpdate<T>* rs = nullptr) {
  const int wid = GRID ? d.vb * (int)blockDim.x + (int)threadIdx.x : (int)threadIdx.x;
  const int nw = GRID ? d.nblocks * (int)blockDim.x : (int)blockDim.x;
  for (int lv = 0; lv < P.levels; lv++) {
    const int c1 = P.chunk_off[lv + 1];
    for (int ch = P.chunk_off[lv] + wid; ch < c1; ch += nw) {
      if (!all_static && P.chunk_dyn[ch]) tps_reduce_chunk(P, w, lv, ch);
      else if (lv == 0 && rs != nullptr) tps_reduce_chunk_static(P, w, 0, ch, *rs);
      else tps_reduce_chunk_static(P, w, lv, ch);
    }
    tsync<GRID>(d);
  }
  // Final forest. Split (grid launches): every workgroup eliminates the subtrees of
  // its groups below the cut height, workgroup 0 the heights above, then back down;
  // a subtree's nodes never leave their workgroup, so only the two hand-overs
  // need grid barriers. Unsplit: workgroup 0 takes every height.
  const int cut = GRID ? P.fin_cut : 0;
  if (cut > 0) {
    for (int h = 0; h < cut; h++) {
      for (int g = d.vb; g < P.sub_groups; g += d.nblocks)
        for (int k = P.sub_hoff[g * cut + h] + (int)threadIdx.x; k < P.sub_hoff[g * cut + h + 1];
             k += blockDim.x)
          tps_final_up(P, w, P.sub_order[k]);
      __syncthreads();
    }
    tsync<GRID>(d);
  }
  if (GRID && tl != nullptr && P.top_n > 0) {
    if (d.vb == 0) {
      const TopLds<T> L = TopLds<T>::carve(tl);
      const int ns = top_stage_count(P), nh = P.heights - cut;
      for (int j = threadIdx.x; j < ns; j += blockDim.x) top_stage(P, w, L, j);
      __syncthreads();
      for (int h = 0; h < nh; h++) {
        for (int j = L.hoff[h] + (int)threadIdx.x; j < L.hoff[h + 1]; j += blockDim.x) top_up(L, j);
        __syncthreads();
      }
      for (int h = nh - 1; h >= 0; h--) {
        for (int j = L.hoff[h] + (int)threadIdx.x; j < L.hoff[h + 1]; j += blockDim.x)
          top_down(L, j);
        __syncthreads();
      }
      for (int j = threadIdx.x; j < P.top_n; j += blockDim.x) top_store(P, w, L, j);
    }
  } else if (!GRID && tl != nullptr && P.ftask_hn > 0) {
    // one workgroup with room for the whole forest (tl holds forest_lds_bytes):
    // staged in LDS at task positions, a workgroup barrier per task height
    const ForestLds<T> L = ForestLds<T>::carve(tl, P);
    const int ns = forest_stage_count(P);
    for (int j = threadIdx.x; j < ns; j += blockDim.x) forest_stage(P, w, L, j);
    __syncthreads();
    for (int h = 0; h < P.ftask_hn; h++) {
      for (int t = L.thoff[h] + (int)threadIdx.x; t < L.thoff[h + 1]; t += blockDim.x)
        forest_up(L, t);
      __syncthreads();
    }
    for (int h = P.ftask_hn - 1; h >= 0; h--) {
      for (int t = L.thoff[h] + (int)threadIdx.x; t < L.thoff[h + 1]; t += blockDim.x)
        forest_down(L, t);
      __syncthreads();
    }
    for (int j = threadIdx.x; j < P.fin_n; j += blockDim.x) forest_store(P, w, L, j);
    __syncthreads();
  } else if ((!GRID || d.vb == 0) && cut == 0 && P.ftask_hn > 0) {
    // unsplit: chain tasks, a workgroup barrier per task height
    for (int h = 0; h < P.ftask_hn; h++) {
      for (int t = P.ftask_hoff[h] + (int)threadIdx.x; t < P.ftask_hoff[h + 1]; t += blockDim.x)
        tps_ftask_up(P, w, t);
      __syncthreads();
    }
    for (int h = P.ftask_hn - 1; h >= 0; h--) {
      for (int t = P.ftask_hoff[h] + (int)threadIdx.x; t < P.ftask_hoff[h + 1]; t += blockDim.x)
        tps_ftask_down(P, w, t);
      __syncthreads();
    }
  } else if (!GRID || d.vb == 0) {
    for (int h = cut; h < P.heights; h++) {
      for (int k = P.fin_hoff[h] + (int)threadIdx.x; k < P.fin_hoff[h + 1]; k += blockDim.x)
        tps_final_up(P, w, P.fin_order[k]);
      __syncthreads();
    }
    for (int h = P.heights - 1; h >= cut; h--) {
      for (int k = P.fin_hoff[h] + (int)threadIdx.x; k < P.fin_hoff[h + 1]; k += blockDim.x)
        tps_final_down(P, w, P.fin_order[k]);
      __syncthreads();
    }
  }
  if (cut > 0) {
    tsync<GRID>(d);
    for (int h = cut - 1; h >= 0; h--) {
      for (int g = d.vb; g < P.sub_groups; g += d.nblocks)
        for (int k = P.sub_hoff[g * cut + h] + (int)threadIdx.x; k < P.sub_hoff[g * cut + h + 1];
             k += blockDim.x)
          tps_final_down(P, w, P.sub_order[k]);
      __syncthreads();
    }
  }
  if (GRID) tsync<GRID>(d);
  for (int lv = P.levels - 1; lv >= 1; lv--) {
    const int c1 = P.chunk_off[lv + 1];
    for (int ch = P.chunk_off[lv] + wid; ch < c1; ch += nw)
      tps_back_chunk<OUT_LEVEL>(P, w, lv, ch, nullptr);
    tsync<GRID>(d);
  }
  // level 0 lands in the FP64 vector `out` directly (the potential: x = solution,
  // or x += correction; a preconditioned Krylov step: z = M^-1 r)
  if (P.levels == 0) {
    const int n0 = P.nslots[0];
    for (int v = wid; v < n0; v += nw) {
      if (acc) tps_final_out<OUT_ADD>(P, w, v, out);
      else tps_final_out<OUT_ASSIGN>(P, w, v, out);
    }
  } else {
    const int c1 = P.chunk_off[1];
    for (int ch = wid; ch < c1; ch += nw) {
      if (acc) tps_back_chunk<OUT_ADD>(P, w, 0, ch, out);
      else tps_back_chunk<OUT_ASSIGN>(P, w, 0, ch, out);
    }
  }
  tsync<GRID>(d);
}

template <int NV, bool GRID>
__device__ __forceinline__ bool tree_reduce(const ImpDev& d, double (&v)[NV], double* lds,
                                            int& slot) {
  if (GRID) return reduce_all<NV>(d, v, lds, slot);
  block_sum<NV>(v, lds);
  return true;
}

// One step's solve by BiCGSTAB, the reference GPU solver's method
// (solverIm-gpu.cpp:725-843: the same recurrences for alpha, omega and beta),
// right-preconditioned by the tree solve: p^ = M^-1 p and s^ = M^-1 s are tree
// solves of the step's matrix (the FP64 one is exact, so the half step converges
// in the first iteration; the FP32 one, M^-1 = the FP32 tree solve, takes a few),
// and x is updated with p^, s^. Vectors in level-0 slot order; x0 = the previous
// potential; the test ||b - A x||_2 <= tol (relative: tol ||b||_2) at the half and
// the full step of every iteration; a breakdown ((r0, v) = 0, rho = 0, omega = 0)
// restarts the shadow residual as r0 = r, at most kMaxRestarts times a step.
struct KrylovOut {
  int iters, solves, status, restarts;
  double rr, bb;
};
template <class T, bool GRID>
__device__ __forceinline__ KrylovOut tree_bicgstab(const ImpDev& d, const TreeArgs<T>& t,
                                                   double* lds, unsigned char* tl, int& slot) {
  const int r_begin = GRID ? d.vb * d.rpb : 0;
  const int r_end = GRID ? min(d.ns0, r_begin + d.rpb) : d.ns0;
  const int bs = blockDim.x;
  const double* __restrict__ val = d.val;
  const int32_t* __restrict__ rowptr = d.srowptr;
  const int32_t* __restrict__ col = d.scol;
  double* __restrict__ x = d.x;
  const double* __restrict__ b = d.b;
  double* __restrict__ r = d.r;
  double* __restrict__ r0 = d.r0;
  double* __restrict__ p = d.p0;
  double* __restrict__ v = d.v0;
  double* __restrict__ sv = d.s;
  double* __restrict__ tv = d.t;
  double* __restrict__ ph = d.p1;  // M^-1 p
  double* __restrict__ sh = d.v1;  // M^-1 s
  T* __restrict__ wr = t.w.r;
  KrylovOut o{0, 0, NRT_ERR_NOT_CONVERGED, 0, 0.0, 0.0};
  // r = b - A x ; r0 = r ; rho = (r0, r) ; ||b||^2
  double red2[2] = {0.0, 0.0};
  for (int i = r_begin + (int)threadIdx.x; i < r_end; i += bs) {
    double ax = 0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) ax = ax + val[k] * x[col[k]];
    const double bi = b[i];
    const double ri = bi + (-1.0) * ax;
    r[i] = ri;
    r0[i] = ri;
    red2[0] = red2[0] + ri * ri;
    red2[1] = red2[1] + bi * bi;
  }
  bool ok = tree_reduce<2, GRID>(d, red2, lds, slot);
  double rho = red2[0], rr = red2[0];
  o.bb = red2[1];
  const double thr = d.tol_mode == NRT_TOL_RELATIVE ? d.tol * sqrt(o.bb) : d.tol;
  const double limit = thr * thr;
  if (rr <= limit) o.status = NRT_OK;
  double alpha = 0, omega = 0, beta = 0;
  bool first = true, restart = false;
  while (ok && o.status == NRT_ERR_NOT_CONVERGED && o.iters < d.max_iter) {
    // p = r + beta (p - omega v), and the tree system's right-hand side
    const double nomega = -omega;
    for (int i = r_begin + (int)threadIdx.x; i < r_end; i += bs) {
      if (restart) r0[i] = r[i];
      const double pi = first ? r[i] : r[i] + beta * (p[i] + nomega * v[i]);
      p[i] = pi;
      wr[i] = (T)pi;
    }
    restart = false;
    tsync<GRID>(d);
    // the matrix changed (HH rows) only before the step's first solve
    tps_run<T, GRID>(d, t.P, t.w, false, o.solves > 0, tl, ph);
    o.solves++;
    // v = A p^ ; (r0, v)
    double redA[1] = {0.0};
    for (int i = r_begin + (int)threadIdx.x; i < r_end; i += bs) {
      double acc = 0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++) acc = acc + val[k] * ph[col[k]];
      v[i] = acc;
      redA[0] = redA[0] + acc * r0[i];
    }
    ok = tree_reduce<1, GRID>(d, redA, lds, slot);
    if (!ok) break;
    o.iters++;
    if (!isfinite(redA[0])) {
      o.status = NRT_ERR_NUMERIC;
      break;
    }
    if (redA[0] == 0.0 || !isfinite(rho / redA[0])) {
      if (++o.restarts > kMaxRestarts) break;
      restart = true;
      first = true;
      rho = rr;
      continue;
    }
    alpha = rho / redA[0];
    const double nalpha = -(rho / redA[0]);
    // s = r - alpha v ; ||s||^2 ; the tree system's right-hand side
    double red1[1] = {0.0};
    for (int i = r_begin + (int)threadIdx.x; i < r_end; i += bs) {
      const double si = r[i] + nalpha * v[i];
      sv[i] = si;
      wr[i] = (T)si;
      red1[0] = red1[0] + si * si;
    }
    ok = tree_reduce<1, GRID>(d, red1, lds, slot);
    if (!ok) break;
    if (!isfinite(red1[0])) {
      o.status = NRT_ERR_NUMERIC;
      break;
    }
    if (red1[0] <= limit) {
      // converged at the half step: x += alpha p^
      for (int i = r_begin + (int)threadIdx.x; i < r_end; i += bs) {
        x[i] = x[i] + alpha * ph[i];
        r[i] = sv[i];
      }
      rr = red1[0];
      o.status = NRT_OK;
      break;
    }
    tps_run<T, GRID>(d, t.P, t.w, false, true, tl, sh);
    o.solves++;
    // t = A s^ ; (t, s), (t, t)
    double redB[2] = {0.0, 0.0};
    for (int i = r_begin + (int)threadIdx.x; i < r_end; i += bs) {
      double acc = 0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++) acc = acc + val[k] * sh[col[k]];
      tv[i] = acc;
      redB[0] = redB[0] + acc * sv[i];
      redB[1] = redB[1] + acc * acc;
    }
    ok = tree_reduce<2, GRID>(d, redB, lds, slot);
    if (!ok) break;
    if (!isfinite(redB[0]) || !isfinite(redB[1]) || redB[1] == 0.0) {
      o.status = NRT_ERR_NUMERIC;
      break;
    }
    omega = redB[0] / redB[1];
    const double nw = -(redB[0] / redB[1]);
    // x += alpha p^ + omega s^ ; r = s - omega t ; (r0, r), (r, r)
    double redC[2] = {0.0, 0.0};
    for (int i = r_begin + (int)threadIdx.x; i < r_end; i += bs) {
      double xi = x[i] + alpha * ph[i];
      xi = xi + omega * sh[i];
      x[i] = xi;
      const double ri = sv[i] + nw * tv[i];
      r[i] = ri;
      redC[0] = redC[0] + ri * r0[i];
      redC[1] = redC[1] + ri * ri;
    }
    ok = tree_reduce<2, GRID>(d, redC, lds, slot);
    if (!ok) break;
    beta = (alpha / omega) * (redC[0] / rho);
    rho = redC[0];
    rr = redC[1];
    first = false;
    if (!isfinite(rr)) {
      o.status = NRT_ERR_NUMERIC;
      break;
    }
    if (rr <= limit) {
      o.status = NRT_OK;
      break;
    }
    if (!isfinite(beta) || rho == 0.0 || omega == 0.0) {
      if (++o.restarts > kMaxRestarts) break;
      restart = true;
      first = true;
      rho = rr;
    }
  }
  if (!ok) o.status = NRT_ERR_HIP;
  o.rr = rr;
  tsync<GRID>(d);  // x complete before the next step's update and residual
  return o;
}

// Update phase of a tree-solved step, over slots. HH elements are spread one per
// thread over the whole launch (their gating and row refresh are long dependent
// chains, so they must not queue up inside one wave); the other slots of the
// range follow. Each writes its right-hand side straight into the level-0 tree
// system (direct) or into b (refinement, whose FP64 residual reads it); an HH row
// also refreshes its CSR values and the level-0 diagonal.
template <class T, bool GRID>
__device__ __forceinline__ void tree_update(const ImpDev& d, const TreeArgs<T>& t, int r0, int r1,
                                            bool active) {
  const int tid = GRID ? d.vb * (int)blockDim.x + (int)threadIdx.x : (int)threadIdx.x;
  const int nt = GRID ? d.nblocks * (int)blockDim.x : (int)blockDim.x;
  const double* __restrict__ xv = d.x;
  const uint32_t* __restrict__ code = d.scode;
  const double* __restrict__ cmdx = d.scmdx;
  const double* __restrict__ kdx = d.skdx;
  const double* __restrict__ input = d.sinput;
  T* __restrict__ rt = t.w.r;
  double* __restrict__ b = d.b;
  const double dT = d.dT;
  for (int hh = tid; hh < d.H; hh += nt) {
    const int i = d.hh.row[hh];
    const int si = d.hh_slot[hh];
    const double v = xv[si];
    double W, K, diag;
    imp_gate<T>(d.hh, hh, v, d.dt_ms, W, K);
    if (imp_update_row(d, hh, i, W, diag)) t.w.d[si] = (T)diag;
    double bi = d.cmdx[i] * v / dT + K * d.dxv[i];
    if (active) bi = bi + d.input[i];
    if (t.direct) rt[si] = (T)bi;
    else b[si] = bi;
  }
  if (t.direct) {
    // direct solve: only the rows of dynamic chunks (the static ones are computed
    // by the sweep); the HH rows among them were written above
    const int32_t* __restrict__ dl = t.P.dyn0;
    for (int k = tid; k < t.P.ndyn0; k += nt) {
      const int i = dl[k];
      if (code[i] & (kCodeHH | kCodeSkip)) continue;
      double bi = cmdx[i] * xv[i] / dT + kdx[i];
      if (active) bi = bi + input[i];
      rt[i] = (T)bi;
    }
    return;
  }
  // the other slots, loads of kTpsBatch slots per thread issued together
  const int st = (int)blockDim.x;
  for (int i0 = r0 + (int)threadIdx.x; i0 < r1; i0 += kTpsBatch * st) {
    uint32_t bc[kTpsBatch];
    double bv[kTpsBatch], bm[kTpsBatch], bk[kTpsBatch], bin[kTpsBatch];
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int i = i0 + j * st;
      if (i < r1) {
        bc[j] = code[i];
        bv[j] = xv[i];
        bm[j] = cmdx[i];
        bk[j] = kdx[i];
        bin[j] = active ? input[i] : 0.0;
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int i = i0 + j * st;
      if (i < r1 && !(bc[j] & (kCodeHH | kCodeSkip))) {
        double bi = bm[j] * bv[j] / dT + bk[j];
        if (active) bi = bi + bin[j];
        if (t.direct) rt[i] = (T)bi;
        else b[i] = bi;
      }
    }
  }
}

// The form of each step's solve: taken from TreeArgs at run time (kSolveAny), or
// fixed at compile time, so that a kernel holds only the code of one form (the
// batch kernel: the registers of the Krylov and refinement loops would otherwise
// spill in the direct one).
constexpr int kSolveAny = -1, kSolveDirect = 0, kSolveRefine = 1, kSolveKrylov = 2;

// Steps d.first .. d.first+d.num-1 of one tree (GRID: the whole launch works on
// it; else one workgroup does).
template <class T, bool GRID, int MODE = kSolveAny>
__device__ __forceinline__ void tree_steps(const ImpDev& d, const TreeArgs<T>& t, double* lds,
                                           unsigned char* tl) {
  const bool direct = MODE == kSolveAny ? t.direct != 0 : MODE == kSolveDirect;
  const bool krylov = MODE == kSolveAny ? t.krylov != 0 : MODE == kSolveKrylov;
  // slot ranges (the tree solve's vectors are in level-0 slot order)
  const int r_begin = GRID ? d.vb * d.rpb : 0;
  const int r_end = GRID ? min(d.ns0, r_begin + d.rpb) : d.ns0;
  const int bs = blockDim.x;
  const double* val = d.val;
  const int32_t* __restrict__ rowptr = d.srowptr;
  const int32_t* __restrict__ col = d.scol;
  long long gl = d.gl0, ksamp = 0;
  int slot = 0;
  const double lim2 = t.refine_tol * t.refine_tol;
  // Level-0 tree system from the matrix. Off-diagonals never change; the diagonal
  // of HH rows is refreshed by the update phase of every step, by the thread that
  // owns the HH entry (not the slot), hence the barrier before the first update.
  if (!d.keep_factors) {
    for (int i = r_begin + (int)threadIdx.x; i < r_end; i += bs) {
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const T v = (T)val[k];
        const int role = d.srole[k];
        if (role == 0) t.w.d[i] = v;
        else if (role == 1) t.w.up[i] = v;
        else t.w.dn[col[k]] = v;
      }
    }
    tsync<GRID>(d);
    tps_factor<T, GRID>(d, t.P, t.w);
  }
  for (long long step = d.first; step < d.first + d.num; step++) {
    const bool active = step >= d.a0 && step <= d.a1;
    tree_update<T, GRID>(d, t, r_begin, r_end, active);
    tsync<GRID>(d);
    int sweeps = 0, status = NRT_OK;
    double rr = 0, bb = 0;
    bool ok = true;
    int iters = -1;
    if (direct) {
      const TpsRhsUpdate<T> rs{d.x, d.scmdx, d.skdx, d.sinput, d.dT, active};
      tps_run<T, GRID>(d, t.P, t.w, false, false, tl, d.x, &rs);
      sweeps = 1;
    } else if (krylov) {
      const KrylovOut ko = tree_bicgstab<T, GRID>(d, t, lds, tl, slot);
      iters = ko.iters;
      sweeps = ko.solves;
      status = ko.status;
      rr = ko.rr;
      bb = ko.bb;
      ok = status != NRT_ERR_HIP;
      if (d.vb == 0 && threadIdx.x == 0) d.st->restarts += ko.restarts;
    } else {
      while (true) {
        // FP64 residual r = b - A x into the level-0 right-hand side; the test is on
        // max norms, Jacobi-scaled: |D^-1 r|_inf <= refine_tol |D^-1 b|_inf, a bound
        // per element whatever the number of rows (rr, bb hold their squares)
        double red[2] = {0.0, 0.0};
        for (int i = r_begin + (int)threadIdx.x; i < r_end; i += bs) {
          double ax = 0, dg = 1.0;
          for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            ax = ax + val[k] * d.x[col[k]];
            if (d.srole[k] == 0) dg = val[k];
          }
          const double bi = d.b[i];
          const double ri = bi - ax;
          t.w.r[i] = (T)ri;
          red[0] = fmax(red[0], fabs(ri / dg));
          red[1] = fmax(red[1], fabs(bi / dg));
        }
        if (GRID) ok = reduce_all_max<2>(d, red, lds, slot);
        else block_max<2>(red, lds);
        if (!ok) break;
        rr = red[0] * red[0];
        bb = red[1] * red[1];
        if (!isfinite(rr)) {
          status = NRT_ERR_NUMERIC;
          break;
        }
        if (rr <= lim2 * bb) break;
        if (sweeps >= t.refine_max) {
          status = NRT_ERR_NOT_CONVERGED;
          break;
        }
        tps_run<T, GRID>(d, t.P, t.w, true, sweeps > 0, tl, d.x);
        sweeps++;
      }
    }
    if (GRID && __hip_atomic_load(d.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
      ok = false;
    if (!ok) status = NRT_ERR_HIP;
    if (d.vb == 0 && threadIdx.x == 0) {
      // Krylov form: iterations, and the tree solves as sweeps
      const int its = iters >= 0 ? iters : sweeps;
      d.st->iters_total += (unsigned long long)its;
      if (!direct) d.st->sweeps_total += (unsigned long long)sweeps;
      d.st->steps += 1;
      d.st->last_iters = its;
      d.st->last_rr = rr;
      d.st->last_bb = bb;
      if (d.st->status == NRT_OK && status != NRT_OK) {
        d.st->status = status;
        d.st->fail_step = step;
        d.st->fail_iter = its;
        d.st->fail_rr = rr;
        d.st->fail_bb = bb;
        d.st->fail_cycles = sweeps;
      }
    }
    if (!ok) return;
    if (d.interval >= 0 && ++gl > d.interval) {
      gl = 0;
      if (d.trace && d.trace_elem >= r_begin && d.trace_elem < r_end && threadIdx.x == 0)
        d.trace[ksamp] = d.x[d.trace_elem];
      ksamp++;
    }
  }
}

// ------------------------------------------------------------ local mode
// One workgroup per row group (TreePlanHost::groups): workgroup g reduces the
// columns of its group at every level and its local final nodes with workgroup
// barriers only; after a grid barrier workgroup 0 solves the top of the final
// forest (LDS-staged when it fits), and after a second one every workgroup
// back-substitutes its own nodes. The arithmetic per chunk and per node is the
// level-synchronous solve's.

// Factors of every chunk of the group (once per launch).
template <class T>
__device__ __forceinline__ void tps_factor_local(const TreePlanDev& P, const TreeWork<T>& w, int g) {
  for (int lv = 0; lv < P.levels; lv++) {
    const int32_t* gc = P.gcol + lv * (P.groups + 1);
    const int c1 = gc[g + 1];
    for (int c = gc[g] + (int)threadIdx.x; c < c1; c += blockDim.x)
      tps_reduce_chunk(P, w, lv, P.chunk_off[lv] + c);
    __syncthreads();
  }
}

// The top of the final forest, between the two grid barriers (workgroup 0).
template <class T>
__device__ __forceinline__ void local_top(const ImpDev& d, const TreePlanDev& P,
                                          const TreeWork<T>& w, unsigned char* tl) {
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
  if (d.vb == 0) {
    const int nh = P.heights - P.fin_cut;
    if (tl != nullptr) {
      const TopLds<T> L = TopLds<T>::carve(tl, gl_top_entries(P));
      const int ns = top_stage_count(P);
      for (int j = tid; j < ns; j += nt) top_stage(P, w, L, j);
      __syncthreads();
      // Heights below top_narrow (wide) take the workgroup and a barrier each;
      // from top_narrow up (at most a wave's nodes per height) wave 0 works alone,
      // its heights ordered by the wave's own LDS order (LDS operations of one wave
      // complete in issue order), with no barrier.
      const int hn = P.top_narrow;
      for (int h = 0; h < hn; h++) {
        for (int j = L.hoff[h] + tid; j < L.hoff[h + 1]; j += nt) top_up(L, j);
        __syncthreads();
      }
      if (tid < kWave) {
        for (int h = hn; h < nh; h++) {
          const int j = L.hoff[h] + tid;
          if (j < L.hoff[h + 1]) top_up(L, j);
          wave_lds_sync();
        }
        for (int h = nh - 1; h >= hn; h--) {
          const int j = L.hoff[h] + tid;
          if (j < L.hoff[h + 1]) top_down(L, j);
          wave_lds_sync();
        }
      }
      __syncthreads();
      for (int h = hn - 1; h >= 0; h--) {
        for (int j = L.hoff[h] + tid; j < L.hoff[h + 1]; j += nt) top_down(L, j);
        __syncthreads();
      }
      for (int j = tid; j < P.top_n; j += nt) top_store(P, w, L, j);
    } else {
      for (int h = 0; h < nh; h++) {
        for (int j = P.top_hoff[h] + tid; j < P.top_hoff[h + 1]; j += nt)
          tps_final_up(P, w, P.top_node[j]);
        __syncthreads();
      }
      for (int h = nh - 1; h >= 0; h--) {
        for (int j = P.top_hoff[h] + tid; j < P.top_hoff[h + 1]; j += nt)
          tps_final_down(P, w, P.top_node[j]);
        __syncthreads();
      }
    }
  }
  grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
}

template <class T>
__device__ __forceinline__ void tps_run_local(const ImpDev& d, const TreePlanDev& P, const TreeWork<T>& w,
                              bool acc, bool all_static, unsigned char* tl) {
  const int g = d.vb, G = P.groups, L = P.levels;
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  for (int lv = 0; lv < L; lv++) {
    const int32_t* gc = P.gcol + lv * (G + 1);
    const int c1 = gc[g + 1];
    for (int c = gc[g] + tid; c < c1; c += nt) {
      const int ch = P.chunk_off[lv] + c;
      if (!all_static && P.chunk_dyn[ch]) tps_reduce_chunk(P, w, lv, ch);
      else tps_reduce_chunk_static(P, w, lv, ch);
    }
    __syncthreads();
  }
  // the group's local final nodes by chain tasks (a barrier per task height)
  const int TH = P.ltask_hn;
  const int32_t* lt = P.ltask_hoff + g * TH;
  for (int h = 0; h < TH; h++) {
    const int k0 = lt[h], k1 = lt[h + 1];
    if (k0 == k1) continue;  // uniform over the workgroup
    for (int t = k0 + tid; t < k1; t += nt) tps_task_up(P, w, P.ltask_off, P.ltask_node, t);
    __syncthreads();
  }
  if (P.top_n > 0) local_top<T>(d, P, w, tl);
  for (int h = TH - 1; h >= 0; h--) {
    const int k0 = lt[h], k1 = lt[h + 1];
    if (k0 == k1) continue;
    for (int t = k0 + tid; t < k1; t += nt) tps_task_down(P, w, P.ltask_off, P.ltask_node, t);
    __syncthreads();
  }
  for (int lv = L - 1; lv >= 1; lv--) {
    const int32_t* gc = P.gcol + lv * (G + 1);
    const int c1 = gc[g + 1];
    for (int c = gc[g] + tid; c < c1; c += nt)
      tps_back_chunk<OUT_LEVEL>(P, w, lv, P.chunk_off[lv] + c, nullptr);
    __syncthreads();
  }
  const int c1 = P.gcol[g + 1];
  for (int c = P.gcol[g] + tid; c < c1; c += nt) {
    if (acc) tps_back_chunk<OUT_ADD>(P, w, 0, c, d.x);
    else tps_back_chunk<OUT_ASSIGN>(P, w, 0, c, d.x);
  }
  __syncthreads();
}

// Group-resident form: the group's levels above 0 and its local final nodes in
// LDS (lw), its plan metadata staged in LDS (M).

// A wide chunk (column j of level lv) eliminated by the whole workgroup: parallel
// cyclic reduction of its interior, each step reading every row before any row is
// written (a barrier between), interior row 1 + tid for thread tid.
template <class T>
__device__ __forceinline__ void wide_reduce(const GlIO<T>& io, const TreeWorkL<T>& lw,
                                            const GlMeta& M, int lv, int j) {
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  const int q = 4 * (M.lvl[7 * lv + 3] + j);
  const int m = M.col[q], es = M.col[q + 1], ee = M.col[q + 2];
  const int base = M.lvl[7 * lv + 2] + j, st = M.lvl[7 * lv + 1];
  for (int k = 1 + tid; k <= m - 2; k += nt) wide_init(lw, base, st, m, k);
  __syncthreads();
  const int ns = wide_steps(m - 2);
  const int k0 = 1 + tid;  // the plan keeps m - 2 <= 512 = blockDim
  for (int s = 0, h = 1; s < ns; s++, h *= 2) {
    WideRow<T> o0{};
    if (k0 <= m - 2) o0 = wide_step(lw, base, st, m, k0, h);
    __syncthreads();
    if (k0 <= m - 2) wide_put(lw, base, st, k0, o0);
    __syncthreads();
  }
  for (int k = 1 + tid; k <= m - 2; k += nt) wide_finish(lw, base, st, k);
  __syncthreads();
  if (tid == 0) wide_ends(lw, base, st, m, es, ee, io.ends(lw, M.scal[0]));
}

template <class T>
__device__ __forceinline__ void wide_back(const GlIO<T>& io, const TreeWorkL<T>& lw,
                                          const GlMeta& M, int lv, int j) {
  const int q = 4 * (M.lvl[7 * lv + 3] + j);
  const int m = M.col[q], es = M.col[q + 1], ee = M.col[q + 2];
  const int base = M.lvl[7 * lv + 2] + j, st = M.lvl[7 * lv + 1];
  const EndsLocal<T> E = io.ends(lw, M.scal[0]);
  wide_back_rows(lw, base, st, m, E.x(es), E.x(ee), (int)threadIdx.x, (int)blockDim.x);
}

// Levels 0 .. L-1 of the group: per-thread chunks, then its wide chunks.
template <class T>
__device__ __forceinline__ void res_levels(const GlIO<T>& io, const TreeWorkL<T>& lw,
                                           const GlMeta& M, int L, int how) {
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  for (int lv = 0; lv < L; lv++) {
    const int nc = M.lvl[7 * lv + 1];
    for (int j = tid; j < nc; j += nt) {
      if (lv == 0) glm_reduce<true>(io, lw, M, 0, j, how);
      else glm_reduce<false>(io, lw, M, lv, j, how);
    }
    __syncthreads();
    const int wn = M.lvl[7 * lv + 6];
    for (int i = 0; i < wn; i++) {
      wide_reduce(io, lw, M, lv, M.wcol[M.lvl[7 * lv + 5] + i]);
      __syncthreads();
    }
  }
}

template <class T>
__device__ __forceinline__ void tps_factor_res(const GlIO<T>& io, const TreeWorkL<T>& lw,
                                               const GlMeta& M, int L) {
  res_levels<T>(io, lw, M, L, kGlFull);
}

// A small top (at most a wave's nodes, at most kLaneKids children each) solved by
// wave 0 with one node per lane: the node's height, children and parent (local
// top indices) stay in registers for the launch, so a height costs one round of
// LDS reads. The arithmetic is top_up/top_down's.
constexpr int kLaneKids = 4;
struct TopLane {
  int h, nk, p;
  int k[kLaneKids];
};
__device__ __forceinline__ TopLane top_lane_load(const TreePlanDev& P, int j) {
  TopLane t{};
  t.h = -1;
  if (j >= P.top_n) return t;
  const int nh = P.heights - P.fin_cut;
  for (int h = 0; h < nh; h++)
    if (j >= P.top_hoff[h] && j < P.top_hoff[h + 1]) t.h = h;
  t.nk = P.top_cptr[j + 1] - P.top_cptr[j];
#pragma unroll
  for (int q = 0; q < kLaneKids; q++) t.k[q] = q < t.nk ? P.top_child[P.top_cptr[j] + q] : 0;
  t.p = P.top_par[j];
  return t;
}
// the top qualifies for the lane form
__device__ __forceinline__ bool top_lanes_ok(const TreePlanDev& P) {
  return P.top_n <= kWave && P.top_kids <= kLaneKids;
}
template <class T>
__device__ __forceinline__ void top_lane_solve(const TopLds<T>& L, const TopLane& t, int j,
                                               int nh) {
  for (int h = 0; h < nh; h++) {
    if (t.h == h) {
      T dv = L.d[j], rv = L.r[j];
#pragma unroll
      for (int q = 0; q < kLaneKids; q++) {
        if (q < t.nk) {
          const int c = t.k[q];
          const T f = L.dn[c] * tps_rcp(L.d[c]);
          dv = dv - f * L.up[c];
          rv = rv - f * L.r[c];
        }
      }
      L.d[j] = dv;
      L.r[j] = rv;
    }
    wave_lds_sync();
  }
  for (int h = nh - 1; h >= 0; h--) {
    if (t.h == h) {
      T acc = L.r[j];
      if (t.p >= 0) acc = acc - L.up[j] * L.x[t.p];
      L.x[j] = acc * tps_rcp(L.d[j]);
    }
    wave_lds_sync();
  }
}

// The top of the final forest in the group-resident form: after the grid barrier
// every workgroup stages this solve's top inputs (io.wt) and solves the top itself
// in its LDS, all writing the same solution to io.w.x; no second barrier. A
// workgroup cannot write the next-but-one solve's inputs into the buffer read
// here before every workgroup has passed the next solve's barrier, that is, has
// finished this solve.
template <class T>
__device__ __forceinline__ void local_top_res(const ImpDev& d, const TreePlanDev& P,
                                              const GlIO<T>& io, unsigned char* tl,
                                              const TopLane& lane) {
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
  const int nh = P.heights - P.fin_cut;
  if (tl != nullptr && top_lanes_ok(P)) {
    const TopLds<T> L = TopLds<T>::carve(tl, gl_top_entries(P));
    for (int j = tid; j < P.top_n + P.top_ext; j += nt) top_stage_values(P, io.wt, L, j);
    __syncthreads();
    if (tid < kWave) top_lane_solve(L, lane, tid, nh);
    __syncthreads();
    for (int j = tid; j < P.top_n; j += nt) top_store(P, io.w, L, j);
  } else if (tl != nullptr) {
    const TopLds<T> L = TopLds<T>::carve(tl, gl_top_entries(P));
    const int ns = top_stage_count(P);
    for (int j = tid; j < ns; j += nt) top_stage(P, io.wt, L, j);
    __syncthreads();
    const int hn = P.top_narrow;
    for (int h = 0; h < hn; h++) {
      for (int j = L.hoff[h] + tid; j < L.hoff[h + 1]; j += nt) top_up(L, j);
      __syncthreads();
    }
    if (tid < kWave) {
      for (int h = hn; h < nh; h++) {
        const int j = L.hoff[h] + tid;
        if (j < L.hoff[h + 1]) top_up(L, j);
        wave_lds_sync();
      }
      for (int h = nh - 1; h >= hn; h--) {
        const int j = L.hoff[h] + tid;
        if (j < L.hoff[h + 1]) top_down(L, j);
        wave_lds_sync();
      }
    }
    __syncthreads();
    for (int h = hn - 1; h >= 0; h--) {
      for (int j = L.hoff[h] + tid; j < L.hoff[h + 1]; j += nt) top_down(L, j);
      __syncthreads();
    }
    for (int j = tid; j < P.top_n; j += nt) top_store(P, io.w, L, j);
  }
  // (the group-resident form is only chosen with its top staged: tl is set)
  __syncthreads();
}

template <class T>
__device__ __forceinline__ void tps_run_res(const ImpDev& d, const TreePlanDev& P,
                                            const GlIO<T>& io, const TreeWorkL<T>& lw,
                                            const GlMeta& M, bool acc, bool all_static,
                                            unsigned char* tl, const TopLane& lane) {
  const int H = P.heights, L = P.levels, how = all_static ? kGlStatic : kGlStep;
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  res_levels<T>(io, lw, M, L, how);
  for (int h = 0; h < H; h++) {
    const int k0 = M.hoff[h], k1 = M.hoff[h + 1];
    if (k0 == k1) continue;  // uniform over the workgroup
    for (int j = k0 + tid; j < k1; j += nt) glm_final_up(io, lw, M, j);
    __syncthreads();
  }
  if (P.top_n > 0) local_top_res<T>(d, P, io, tl, lane);
  for (int h = H - 1; h >= 0; h--) {
    const int k0 = M.hoff[h], k1 = M.hoff[h + 1];
    if (k0 == k1) continue;
    for (int j = k0 + tid; j < k1; j += nt) glm_final_down(io, lw, M, j);
    __syncthreads();
  }
  for (int lv = L - 1; lv >= 1; lv--) {
    const int nc = M.lvl[7 * lv + 1];
    for (int j = tid; j < nc; j += nt) glm_back<false, OUT_LEVEL>(io, lw, M, lv, j, nullptr);
    const int wn = M.lvl[7 * lv + 6];
    for (int i = 0; i < wn; i++) wide_back(io, lw, M, lv, M.wcol[M.lvl[7 * lv + 5] + i]);
    __syncthreads();
  }
  const int nc0 = M.lvl[1];
  for (int j = tid; j < nc0; j += nt) {
    if (acc) glm_back<true, OUT_ADD>(io, lw, M, 0, j, d.x);
    else glm_back<true, OUT_ASSIGN>(io, lw, M, 0, j, d.x);
  }
  __syncthreads();
}

// Update phase over the group's own level-0 slots (its columns at every row
// position) and its HH entries; as tree_update otherwise.
template <class T>
__device__ __forceinline__ void tree_update_local(const ImpDev& d, const TreeArgs<T>& t,
                                                  bool active) {
  const int g = d.vb, tid = (int)threadIdx.x, nt = (int)blockDim.x;
  const double* __restrict__ xv = d.x;
  const uint32_t* __restrict__ code = d.scode;
  const double* __restrict__ cmdx = d.scmdx;
  const double* __restrict__ kdx = d.skdx;
  const double* __restrict__ input = d.sinput;
  T* __restrict__ rt = t.w.r;
  double* __restrict__ b = d.b;
  const double dT = d.dT;
  // HH rows are long dependent chains (gating); when they fill at most half of
  // the workgroup they get whole waves of their own and the plain slots the rest
  const int hh0 = t.hh_goff[g], nhh = t.hh_goff[g + 1] - hh0;
  const int hthr = nhh > 0 && 2 * nhh <= nt ? (nhh + kWave - 1) / kWave * kWave : 0;
  for (int hh = hh0 + tid; hh < hh0 + nhh && (hthr == 0 || tid < hthr); hh += hthr ? hthr : nt) {
    const int i = d.hh.row[hh];
    const int si = d.hh_slot[hh];
    const double v = xv[si];
    double W, K, diag;
    imp_gate<T>(d.hh, hh, v, d.dt_ms, W, K);
    if (imp_update_row(d, hh, i, W, diag)) t.w.d[si] = (T)diag;
    double bi = d.cmdx[i] * v / dT + K * d.dxv[i];
    if (active) bi = bi + d.input[i];
    if (t.direct) rt[si] = (T)bi;
    else b[si] = bi;
  }
  const int NC = t.P.nchunk[0], c0 = t.P.gcol[g], nc = t.P.gcol[g + 1] - c0;
  const int ns = nc * (t.P.nslots[0] / NC);
  const int ptid = tid - hthr, pnt = nt - hthr;
  for (int j0 = ptid; ptid >= 0 && j0 < ns; j0 += kTpsBatch * pnt) {
    int si[kTpsBatch];
    uint32_t bc[kTpsBatch];
    double bv[kTpsBatch], bm[kTpsBatch], bk[kTpsBatch], bin[kTpsBatch];
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int q = j0 + j * pnt;
      si[j] = q < ns ? (q / nc) * NC + c0 + q % nc : -1;
      if (si[j] >= 0) {
        bc[j] = code[si[j]];
        bv[j] = xv[si[j]];
        bm[j] = cmdx[si[j]];
        bk[j] = kdx[si[j]];
        bin[j] = active ? input[si[j]] : 0.0;
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      if (si[j] >= 0 && !(bc[j] & (kCodeHH | kCodeSkip))) {
        double bi = bm[j] * bv[j] / dT + bk[j];
        if (active) bi = bi + bin[j];
        if (t.direct) rt[si[j]] = (T)bi;
        else b[si[j]] = bi;
      }
    }
  }
}

// Steps d.first .. d.first+d.num-1 in local mode.
template <class T, bool RES>
__device__ __forceinline__ void tree_steps_local(const ImpDev& d, const TreeArgs<T>& t,
                                                 double* lds, unsigned char* dyn_lds) {
  const TreePlanDev& P = t.P;
  const int g = d.vb, tid = (int)threadIdx.x, nt = (int)blockDim.x;
  const int NC = P.nchunk[0], c0 = P.gcol[g], nc = P.gcol[g + 1] - c0;
  const int ns = nc * (P.nslots[0] / NC);
  const double* val = d.val;
  const int32_t* __restrict__ rowptr = d.srowptr;
  const int32_t* __restrict__ col = d.scol;
  long long gl = d.gl0, ksamp = 0;
  int slot = 0;
  const double lim2 = t.refine_tol * t.refine_tol;
  const bool owns_trace = d.trace && d.trace_elem >= 0 && d.trace_elem % NC >= c0 &&
                          d.trace_elem % NC < c0 + nc;
  // level-0 tree system of the own slots (a row's coupling to a child in another
  // group is written into that child's slot: a grid barrier before the factors)
  for (int q = tid; q < ns; q += nt) {
    const int i = (q / nc) * NC + c0 + q % nc;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      const T v = (T)val[k];
      const int role = d.srole[k];
      if (role == 0) t.w.d[i] = v;
      else if (role == 1) t.w.up[i] = v;
      else t.w.dn[col[k]] = v;
    }
  }
  // dynamic LDS: the staged top (when used), then (RES) the group's values and
  // metadata
  unsigned char* tl = t.top_lds ? dyn_lds : nullptr;
  const GlLayout lay = gl_layout<T>(P, t.top_lds != 0);
  TreeWorkL<T> lw{};
  GlMeta M{};
  if (RES) {
    lw = gl_carve<T>(dyn_lds + lay.top, P.lds_nodes);
    M = gl_meta_carve(dyn_lds + lay.top + lay.values, P);
    gl_meta_stage(P, g, M, tid, nt);
  }
  grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
  // group-resident form: the top inputs of solve k go to buffer k & 1 (the factor
  // phase fills both)
  if (RES) tps_factor_res<T>(GlIO<T>{t.w, t.w, t.walt, true}, lw, M, P.levels);
  else tps_factor_local<T>(P, t.w, g);
  int nsolve = 0;
  const TopLane lane = RES && top_lanes_ok(P) && tid < kWave ? top_lane_load(P, tid) : TopLane{};
  for (long long step = d.first; step < d.first + d.num; step++) {
    const bool active = step >= d.a0 && step <= d.a1;
    tree_update_local<T>(d, t, active);
    __syncthreads();
    int sweeps = 0, status = NRT_OK;
    double rr = 0, bb = 0;
    bool ok = true;
    if (t.direct) {
      if (RES) tps_run_res<T>(d, P, GlIO<T>{t.w, (nsolve++ & 1) ? t.walt : t.w, t.w, false}, lw, M,
                              false, false, tl, lane);
      else tps_run_local<T>(d, P, t.w, false, false, tl);
      sweeps = 1;
    } else {
      while (true) {
        // FP64 residual of the own rows; it reads x across group boundaries, so
        // every group's back-substitution must be complete
        grid_barrier(d.bar, d.nblocks, d.err, d.spin_limit);
        double red[2] = {0.0, 0.0};
        for (int q = tid; q < ns; q += nt) {
          const int i = (q / nc) * NC + c0 + q % nc;
          double ax = 0, dg = 1.0;
          for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            ax = ax + val[k] * d.x[col[k]];
            if (d.srole[k] == 0) dg = val[k];
          }
          const double bi = d.b[i];
          const double ri = bi - ax;
          t.w.r[i] = (T)ri;
          red[0] = fmax(red[0], fabs(ri / dg));
          red[1] = fmax(red[1], fabs(bi / dg));
        }
        ok = reduce_all_max<2>(d, red, lds, slot);
        if (!ok) break;
        rr = red[0] * red[0];
        bb = red[1] * red[1];
        if (!isfinite(rr)) {
          status = NRT_ERR_NUMERIC;
          break;
        }
        if (rr <= lim2 * bb) break;
        if (sweeps >= t.refine_max) {
          status = NRT_ERR_NOT_CONVERGED;
          break;
        }
        if (RES)
          tps_run_res<T>(d, P, GlIO<T>{t.w, (nsolve++ & 1) ? t.walt : t.w, t.w, false}, lw, M, true,
                         sweeps > 0, tl, lane);
        else tps_run_local<T>(d, P, t.w, true, sweeps > 0, tl);
        sweeps++;
      }
    }
    if (__hip_atomic_load(d.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) ok = false;
    if (!ok) status = NRT_ERR_HIP;
    if (g == 0 && tid == 0) {
      d.st->iters_total += (unsigned long long)sweeps;
      if (!t.direct) d.st->sweeps_total += (unsigned long long)sweeps;
      d.st->steps += 1;
      d.st->last_iters = sweeps;
      d.st->last_rr = rr;
      d.st->last_bb = bb;
      if (d.st->status == NRT_OK && status != NRT_OK) {
        d.st->status = status;
        d.st->fail_step = step;
        d.st->fail_iter = sweeps;
        d.st->fail_rr = rr;
        d.st->fail_bb = bb;
        d.st->fail_cycles = sweeps;
      }
    }
    if (!ok) return;
    if (d.interval >= 0 && ++gl > d.interval) {
      gl = 0;
      if (owns_trace && tid == 0) d.trace[ksamp] = d.x[d.trace_elem];
      ksamp++;
    }
  }
}

// Workgroup size of local mode: a group's levels have a few hundred chunks at most,
// and 512 threads leave 256 VGPRs to the FP64 sweeps (1024 would spill).
constexpr int kLocalBlock = 512;

template <class T, bool RES>
__global__ __launch_bounds__(kLocalBlock) void k_implicit_tree_local(ImpDev d, TreeArgs<T> t) {
  __shared__ double lds[4 * 16];
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  d.vb = (int)blockIdx.x;
  tree_steps_local<T, RES>(d, t, lds, dyn_lds);
}

template <class T, bool GRID>
__global__ __launch_bounds__(GRID ? kImpBlock : kTreeBlock) void k_implicit_tree(ImpDev d,
                                                                                 TreeArgs<T> t) {
  __shared__ double lds[4 * 16];
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  d.vb = GRID ? (int)blockIdx.x : 0;
  tree_steps<T, GRID>(d, t, lds, t.top_lds ? dyn_lds : nullptr);
}

// Batched form: independent trees ("slots"), one workgroup each, no grid-wide
// dependency. A slot's descriptor is static while the batch's membership does not
// change; each launch passes, per slot, the steps it has done and the steps it
// takes now (slots may take different numbers of steps in one launch).
template <class T>
struct BatchSlot {
  ImpDev d;
  TreeArgs<T> t;
  double* trace_base;  // probe samples, num_steps / (interval + 1) of them
};

// Steps of a slot in one launch: done before it, taken now, and the workgroups
// that take them (vn > 1: a virtual grid of co-resident workgroups with the
// slot's own grid barrier; the launch guarantees co-residency by its size).
struct SlotWork {
  long long done, steps;
  int vn, pad;
};

// Per-workgroup slot records of a launch: workgroup w of the step kernel runs
// wgmap[w] = (table entry, index in its virtual grid), and its record is that
// entry's with this launch's steps, trace position and grid filled in (num = 0:
// nothing to do). Written by a kernel of its own so that the step kernel reads its
// record through a read-only pointer at a workgroup-uniform address: scalar loads
// into SGPRs, cached, instead of per-lane copies in registers or LDS.
template <class T>
__global__ __launch_bounds__(64) void k_batch_expand(const BatchSlot<T>* __restrict__ slots,
                                                     const SlotWork* __restrict__ work,
                                                     const int2* __restrict__ wgmap,
                                                     BatchSlot<T>* __restrict__ wgrec) {
  const int2 m = wgmap[blockIdx.x];
  static_assert(sizeof(BatchSlot<T>) % 4 == 0, "slot record in words");
  constexpr int kWords = (int)(sizeof(BatchSlot<T>) / 4);
  const uint32_t* __restrict__ src = (const uint32_t*)&slots[m.x];
  uint32_t* __restrict__ dst = (uint32_t*)&wgrec[blockIdx.x];
  for (int i = (int)threadIdx.x; i < kWords; i += 64) dst[i] = src[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    const SlotWork wk = work[m.x];
    ImpDev& d = wgrec[blockIdx.x].d;
    d.first = wk.done + 1;
    d.num = wk.steps > 0 ? wk.steps : 0;
    if (d.interval >= 0) {
      d.gl0 = wk.done % (d.interval + 1);
      const BatchSlot<T>& S = slots[m.x];
      d.trace = S.trace_base ? S.trace_base + wk.done / (d.interval + 1) : nullptr;
    }
    d.vb = m.y;
    d.nblocks = wk.vn;
    d.keep_factors = wk.done > 0;  // an earlier launch filled and factored the arrays
    if (wk.vn > 1) d.rpb = (d.ns0 + wk.vn - 1) / wk.vn;
  }
}

// forest_lds: bytes of dynamic LDS the launch gives each workgroup; a one-workgroup
// slot whose final forest fits there solves it in LDS.
template <class T, int BS, int MODE>
__global__ __launch_bounds__(BS, 3) void k_batch_tree(const BatchSlot<T>* __restrict__ wgrec,
                                                      int forest_lds) {
  __shared__ double lds[4 * 16];
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  const BatchSlot<T>& rec = wgrec[blockIdx.x];
  if (rec.d.num <= 0) return;
  if (rec.d.nblocks > 1) {
    tree_steps<T, true, MODE>(rec.d, rec.t, lds, nullptr);
  } else {
    const size_t fb = forest_lds_bytes<T>(rec.t.P);
    tree_steps<T, false, MODE>(rec.d, rec.t, lds,
                               fb > 0 && fb <= (size_t)forest_lds ? dyn_lds : nullptr);
  }
}

// ---------------------------------------------------------------- host side
static void device_exclusive_scan(const int32_t* in, int32_t* out, int n, DBuf<int32_t>& tmp,
                                  hipStream_t s) {
  const int nb = grid_for(n, kScanBlock);
  tmp.alloc((size_t)nb);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, s, n, in, tmp.p);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, s, nb, tmp.p);
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kScanBlock), 0, s, n, in, tmp.p, out);
  NRT_HIP(hipGetLastError());
}

static int choose_blocks(nrt_handle* hnd, int N, const void* kernel = nullptr) {
  int dev = hnd->device, cus = 0, per_cu = 0;
  NRT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (kernel)
    NRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kImpBlock, 0));
  else
    NRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_implicit_bicgstab_fp64,
                                                         kImpBlock, 0));
  int max_coresident = std::max(1, cus * std::max(1, per_cu));
  int want = hnd->opt.num_blocks > 0 ? hnd->opt.num_blocks : grid_for(N, kImpBlock * 8);
  want = std::min(want, cus);  // one block per CU at most unless asked explicitly
  if (hnd->opt.num_blocks > 0) want = hnd->opt.num_blocks;
  NRT_REQUIRE(want <= max_coresident, NRT_ERR_INVALID,
              "num_blocks exceeds the co-resident limit " + std::to_string(max_coresident));
  return std::max(1, want);
}

Implicit* implicit_create(nrt_handle* hnd, const HostProblem& p, int batch_threads) {
  hipStream_t s = hnd->stream;
  const int N = p.N, H = p.H;
  NRT_REQUIRE(N >= 4, NRT_ERR_INVALID, "the implicit solver needs at least 4 elements");
  auto* e = new Implicit;
  try {
    e->N = N;
    e->H = H;
    e->dT = p.dT;
    {
      int per_block = 0;
      NRT_HIP(hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                    hnd->device));
      e->forest_lds_max = std::min(kForestLdsMax, (size_t)std::max(per_block, 0));
    }
    std::vector<int32_t> hh_of(N, -1);
    for (int i = 0; i < H; i++) {
      NRT_REQUIRE(hh_of[p.hh_elem[i]] < 0, NRT_ERR_INVALID, "duplicate HH element index");
      hh_of[p.hh_elem[i]] = i;
    }
    for (int i = 1; i < N - 1; i++) {
      NRT_REQUIRE(p.mother[i] >= 0 && p.mother[i] < N && p.dr[i] >= 0 && p.dr[i] < N &&
                      p.dl[i] >= 0 && p.dl[i] < N,
                  NRT_ERR_TOPOLOGY, "element " + std::to_string(i) + " has an out-of-range neighbour");
    }
    std::vector<double> cmdx(N), kdx(N, 0.0), rdx(N), Wc(N);
    std::vector<uint32_t> code(N, 0);
    for (int i = 0; i < N; i++) {
      cmdx[i] = p.c_m[i] * p.dx[i];
      rdx[i] = p.r_a[i] * p.dx[i];
      Wc[i] = p.W[i];
      if (hh_of[i] >= 0) code[i] = kCodeHH | (uint32_t)hh_of[i];
      else kdx[i] = p.K[i] * p.dx[i];
    }
    e->cmdx.upload(cmdx, s);
    e->kdx.upload(kdx, s);
    e->rdx.upload(rdx, s);
    e->dxv.upload(p.dx, s);
    e->Wc.upload(Wc, s);
    e->input.upload(p.input, s);
    e->mother.upload(p.mother, s);
    e->dr.upload(p.dr, s);
    e->dl.upload(p.dl, s);
    e->branching.upload(p.branching, s);
    e->code.upload(code, s);
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
    e->hh_type.alloc(H);
    e->hh_mat.alloc(H);
    e->hh_dl.alloc(H);
    e->hh_row.upload(p.hh_elem, s);
    // device assembly
    e->len.alloc(N);
    e->rowptr.alloc(N + 1);
    hipLaunchKernelGGL(k_row_len, dim3(grid_for(N, 256)), dim3(256), 0, s, N, e->dr.p, e->dl.p,
                       e->len.p);
    NRT_HIP(hipGetLastError());
    device_exclusive_scan(e->len.p, e->rowptr.p, N, e->scan_tmp, s);
    int32_t nnz = 0;
    NRT_HIP(hipMemcpyAsync(&nnz, e->rowptr.p + N, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NRT_HIP(hipStreamSynchronize(s));
    NRT_REQUIRE(nnz > 0 && nnz <= 4 * N, NRT_ERR_HIP, "device assembly produced a bad nnz");
    e->nnz = nnz;
    e->col.alloc(nnz);
    e->val.alloc(nnz);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256)), dim3(256), 0, s, N, p.dT, e->rowptr.p,
                       e->col.p, e->val.p, e->mother.p, e->dr.p, e->dl.p, e->branching.p,
                       e->rdx.p, e->cmdx.p, e->dxv.p, e->Wc.p, e->code.p, e->hh_type.p,
                       e->hh_mat.p, e->hh_dl.p);
    NRT_HIP(hipGetLastError());
    // Krylov vectors; x starts as potential[0] (solverIm-gpu.cpp:523-524). The tree
    // solve needs only x and b.
    const nrt_options& o = hnd->opt;
    e->x.upload(p.V0, s);
    e->b.alloc(N);
    NRT_HIP(hipMemsetAsync(e->b.p, 0, sizeof(double) * N, s));
    const bool mixed = o.precision == NRT_FP32_REFINE;
    const bool precond = o.precond == NRT_PRECOND_TREE;
    if (!precond && (o.krylov == NRT_BICGSTAB || o.krylov == NRT_GMRES)) {
      // GMRES: the residual (r0), and in FP64 its Arnoldi vectors (s, t)
      std::vector<DBuf<double>*> vecs{&e->r0};
      if (o.krylov == NRT_BICGSTAB) vecs = {&e->r0, &e->r, &e->p[0], &e->p[1], &e->v[0],
                                            &e->v[1], &e->s, &e->t};
      else if (!mixed) vecs = {&e->r0, &e->s, &e->t};
      for (DBuf<double>* vbuf : vecs) {
        vbuf->alloc(N);
        NRT_HIP(hipMemsetAsync(vbuf->p, 0, sizeof(double) * N, s));
      }
    }
    // 0 Krylov, 1 FP64 tree solve, 2 FP32 tree solve in FP64 refinement, 3 BiCGSTAB
    // preconditioned by the FP64 tree solve, 4 ... by the FP32 tree solve
    e->method = o.krylov == NRT_DIRECT ? (o.precision == NRT_FP64 ? 1 : 2)
                : precond              ? (o.precision == NRT_FP64 ? 3 : 4)
                                       : 0;
    const bool tree_f64 = e->method == 1 || e->method == 3;
    if (e->method != 0) {
      // plan of the tree partition solve from the assembled pattern
      std::vector<int32_t> hrp(N + 1), hcol(nnz);
      NRT_HIP(hipMemcpyAsync(hrp.data(), e->rowptr.p, sizeof(int32_t) * (N + 1),
                             hipMemcpyDeviceToHost, s));
      NRT_HIP(hipMemcpyAsync(hcol.data(), e->col.p, sizeof(int32_t) * nnz,
                             hipMemcpyDeviceToHost, s));
      NRT_HIP(hipStreamSynchronize(s));
      std::vector<int32_t> par;
      try {
        par = csr_parents(N, hrp.data(), hcol.data());
      } catch (const std::exception& ex) {
        throw Error(NRT_ERR_TOPOLOGY, std::string("tree solve: ") + ex.what());
      }
      for (int i = 0; i < N; i++)
        for (int k = hrp[i]; k < hrp[i + 1]; k++)
          NRT_REQUIRE(hcol[k] <= i || par[hcol[k]] == i, NRT_ERR_TOPOLOGY,
                      "tree solve: row " + std::to_string(i) + " couples to row " +
                          std::to_string(hcol[k]) + ", which does not couple back");
      // Workgroups: every reduction level costs a grid barrier (an L2 write-back), so
      // a mid-size tree runs best in one workgroup (measured on a 20,202-row cable
      // with reused factors: 1 -> 71 us, 4 -> 81 us per step) and only large trees
      // spread over the chip.
      int nb = o.num_blocks;
      if (nb == 0 && N <= (1 << 17)) nb = 1;
      if (batch_threads > 0) nb = 1;  // a batch slot: one workgroup of batch_threads
      // Local mode (tree_groups > 1): one workgroup of kLocalBlock threads per row
      // group, each reducing and back-substituting its own rows alone; falls back
      // to the other forms when the plan has no reduction level or the HH entries
      // are not in row order. Automatic for the tree solve from 16k rows: ~2,900
      // rows per workgroup (FP32: ~3,400), at most 64 workgroups, past which the top of the final
      // forest costs more than the groups save (measured per step, FP64: 20,202-row
      // cable 70 us in one workgroup; with level-0 chunks of one per thread, 4
      // groups 33.7 us, 7 groups 31.7, 8 groups 32.0, 12 groups 33.9; depth-14
      // tree of 902,037 rows 389 us in grid mode, 127 with 64 groups, 132 with 80,
      // 146 with 128; FP32 + refinement: 263 -> 138 us and 799 -> 327 us).
      // (the preconditioned Krylov form runs in block or grid launches)
      int G = batch_threads > 0 || precond ? 0 : o.tree_groups;
      if (G == 0 && batch_threads == 0 && !precond && o.num_blocks == 0 && N >= (1 << 14))
        G = std::min(64, e->method == 1 ? (N + 1450) / 2900 : (N + 1700) / 3400);
      if (G > 1) {
        for (int hh = 1; hh < H; hh++)
          if (p.hh_elem[hh] < p.hh_elem[hh - 1]) G = 0;
        G = std::min(G, N / 4);
      }
      e->local_mode = G > 1;
      if (e->local_mode) {
        const void* k = e->method == 1 ? (const void*)k_implicit_tree_local<double, false>
                                       : (const void*)k_implicit_tree_local<float, false>;
        const size_t shmem = e->method == 1 ? top_lds_bytes<double>() : top_lds_bytes<float>();  // upper bound
        int cus = 0, per_cu = 0;
        NRT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, hnd->device));
        NRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kLocalBlock, shmem));
        NRT_REQUIRE(G <= cus * std::max(1, per_cu), NRT_ERR_INVALID,
                    "tree_groups exceeds the co-resident limit " +
                        std::to_string(cus * std::max(1, per_cu)));
        nb = G;
      }
      e->grid_mode = nb != 1 && !e->local_mode;
      if (e->local_mode) {
        e->nblocks = G;
        e->threads = kLocalBlock;
      } else if (e->grid_mode) {
        const void* k = tree_f64 ? (const void*)k_implicit_tree<double, true>
                                 : (const void*)k_implicit_tree<float, true>;
        e->nblocks = o.num_blocks > 0 || nb <= 0 ? choose_blocks(hnd, N, k) : nb;
        e->threads = kImpBlock;
      } else {
        e->nblocks = 1;
        e->threads = batch_threads > 0 ? batch_threads : kTreeBlock;
      }
      // Level 0: one chunk per worker covers an unbranched cable in one round;
      // reduced levels: short chunks (the serial sweep is the latency, and each
      // level shrinks a chain by C/2). Grid launches take 128-row chunks: every
      // level costs grid barriers, and with chunks as long as a dendrite's
      // segments one level is left (depth-14 tree: 0.47 -> 0.40 ms per step).
      // Local mode: one level-0 chunk per thread of a group, down to 5 rows for
      // FP64 (a group with more chunks than threads sweeps two in a row: 20,202-row
      // cable with 8 groups and 4-row chunks 74 us per step, 5-row chunks 32 us),
      // 8 for FP32 (its refinement took 4 sweeps a step instead of 3 with 6- and
      // 7-row chunks on that cable: 176 us per step against 142).
      const int workers = e->nblocks * e->threads;
      int auto0 = e->grid_mode ? 128 : std::min(64, std::max(8, (N + workers - 1) / workers));
      if (e->local_mode)
        auto0 = std::min(32, std::max(e->method == 1 ? 5 : 8, (N + workers - 1) / workers));
      if (batch_threads > 0) auto0 = std::min(16, auto0);  // batch slots: at most 16 rows
      const int chunk = o.chunk > 0 ? o.chunk : auto0;
      // reduced levels: 8-row chunks; 5 in FP64 local mode with groups of at most
      // 4,096 rows (20,202-row cable, 7 groups: 32.05 against 32.53 us per step,
      // means of three runs each)
      const bool short_hi = e->local_mode && e->method == 1 && N <= G * 4096;
      const int chunk_hi = o.chunk_hi > 0 ? o.chunk_hi : (o.chunk > 0 ? o.chunk : short_hi ? 5 : 8);
      // HH rows change their diagonal every step; every other chunk keeps its factors
      std::vector<uint8_t> dyn(N, 0);
      for (int hh = 0; hh < H; hh++) dyn[p.hh_elem[hh]] = 1;
      std::vector<int32_t> rgrp;
      if (e->local_mode) {
        rgrp = row_groups(N, G);
        e->plan_h = build_tree_plan(par, chunk, chunk_hi, &dyn, &rgrp);
        if (e->plan_h.levels == 0) {  // nothing to reduce: one workgroup instead
          e->local_mode = false;
          e->nblocks = 1;
          e->threads = kTreeBlock;
        } else {
          try {
            check_local_plan(e->plan_h);
          } catch (const std::exception& ex) {
            throw Error(NRT_ERR_TOPOLOGY, std::string("tree solve, local mode: ") + ex.what());
          }
        }
      }
      if (!e->local_mode) e->plan_h = build_tree_plan(par, chunk, chunk_hi, &dyn);
      if (e->local_mode) {
        // group-resident form when the top (if staged), the group's values and its
        // metadata fit in LDS beside the kernel's own; else the levels stay in
        // global memory, where the plan keeps no wide chunk (their cooperative
        // elimination needs the group's LDS)
        const bool f64 = e->method == 1;
        int max_lds = 0, cus = 0, per_cu = 0;
        NRT_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                      hnd->device));
        NRT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, hnd->device));
        auto layout = [&]() {
          const TreePlanDev V = plan_view(e->plan_h, nullptr, nullptr);  // sizes only
          const bool top = e->plan_h.top_lds_ok != 0;
          return f64 ? gl_layout<double>(V, top) : gl_layout<float>(V, top);
        };
        GlLayout lay = layout();
        const void* kr = f64 ? (const void*)k_implicit_tree_local<double, true>
                             : (const void*)k_implicit_tree_local<float, true>;
        // (its top, when it has one, must be staged: every workgroup solves it)
        e->local_res = lay.total + 1024 <= (size_t)max_lds &&
                       (e->plan_h.top_n == 0 || e->plan_h.top_lds_ok);
        if (e->local_res) {
          NRT_HIP(hipFuncSetAttribute(kr, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lay.total));
          NRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kr, kLocalBlock, lay.total));
          e->local_res = G <= cus * per_cu;
        }
        if (!e->local_res && !e->plan_h.wide_col.empty()) {
          e->plan_h = build_tree_plan(par, chunk, chunk_hi, &dyn, &rgrp, false);
          try {
            check_local_plan(e->plan_h);
          } catch (const std::exception& ex) {
            throw Error(NRT_ERR_TOPOLOGY, std::string("tree solve, local mode: ") + ex.what());
          }
          lay = layout();
        }
        e->local_lds = e->local_res ? lay.total : lay.top;
        if (!e->local_res && lay.top > 0) {
          if (lay.top + 1024 > (size_t)max_lds) {  // solve the top from global memory
            e->plan_h.top_lds_ok = 0;
            e->local_lds = 0;
          } else {
            const void* kl = f64 ? (const void*)k_implicit_tree_local<double, false>
                                 : (const void*)k_implicit_tree_local<float, false>;
            NRT_HIP(hipFuncSetAttribute(kl, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lay.top));
          }
        }
      }
      // grid launches split the final forest over their workgroups; a batch slot
      // may run as a virtual grid of up to one workgroup per kRowsPerWorkgroup rows
      if (e->grid_mode)
        partition_final_forest(e->plan_h, e->nblocks, e->threads, true);
      else if (batch_threads > 0)
        partition_final_forest(e->plan_h,
                               std::min(kMaxSlotBlocks, (N + kRowsPerWorkgroup - 1) /
                                                            kRowsPerWorkgroup),
                               batch_threads);
      const TreePlanHost& P = e->plan_h;
      // one device buffer holding every plan array
      std::vector<int32_t> flat;
      plan_view(P, nullptr, &flat);
      e->plan_buf.upload(flat, s);
      e->plan = plan_view(P, e->plan_buf.p, nullptr);
      // Slot order: V, b, the per-element coefficients and the FP64 matrix are laid
      // out by level-0 slot, so the update, the residual and the back-substitution
      // read and write consecutive addresses (chunk k-th rows side by side).
      const int ns0 = P.nslots[0];
      e->ns0 = ns0;
      std::vector<double> hval(nnz);
      std::vector<int32_t> hhm(H);
      NRT_HIP(hipMemcpyAsync(hval.data(), e->val.p, sizeof(double) * nnz, hipMemcpyDeviceToHost,
                             s));
      if (H > 0)
        NRT_HIP(hipMemcpyAsync(hhm.data(), e->hh_mat.p, sizeof(int32_t) * H,
                               hipMemcpyDeviceToHost, s));
      NRT_HIP(hipStreamSynchronize(s));
      std::vector<double> sx(ns0, 0.0), scm(ns0, 0.0), skd(ns0, 0.0), sin_(ns0, 0.0);
      std::vector<uint32_t> scode(ns0, kCodeSkip);
      for (int i = 0; i < N; i++) {
        const int si = P.slot0[i];
        sx[si] = p.V0[i];
        scm[si] = cmdx[i];
        skd[si] = kdx[i];
        sin_[si] = p.input[i];
        scode[si] = code[i];
      }
      const SlotCsr sc = slot_order_csr(P, N, hrp.data(), hcol.data(), hval.data(), H,
                                        p.hh_elem.data(), hhm.data());
      e->x.upload(sx, s);
      e->b.alloc(ns0);
      NRT_HIP(hipMemsetAsync(e->b.p, 0, sizeof(double) * ns0, s));
      e->scmdx.upload(scm, s);
      e->skdx.upload(skd, s);
      e->sinput.upload(sin_, s);
      e->scode.upload(scode, s);
      e->srowptr.upload(sc.rowptr, s);
      e->scol.upload(sc.col, s);
      e->srole.upload(sc.role, s);
      e->val.upload(sc.val, s);  // the refreshed HH rows are written here
      e->hh_slot.upload(sc.hh_slot, s);
      e->hh_mat.upload(sc.hh_mat, s);
      if (e->local_mode) {
        // HH entries by group (they are in row order, groups are row ranges)
        std::vector<int32_t> hoff(G + 1, 0);
        for (int hh = 0; hh < H; hh++) hoff[rgrp[p.hh_elem[hh]] + 1]++;
        for (int g = 0; g < G; g++) hoff[g + 1] += hoff[g];
        e->hh_goff.upload(hoff, s);
      }
      e->rowptr.release();
      e->col.release();
      const size_t T = (size_t)P.total_nodes();
      if (precond) {
        // the Krylov vectors, in slot order: r, r0, p, v, s, t, M^-1 p, M^-1 s
        for (DBuf<double>* vbuf : {&e->r, &e->r0, &e->p[0], &e->v[0], &e->s, &e->t, &e->p[1],
                                   &e->v[1]}) {
          vbuf->alloc(ns0);
          NRT_HIP(hipMemsetAsync(vbuf->p, 0, sizeof(double) * ns0, s));
        }
      }
      if (tree_f64) {
        e->tw64.alloc(kTreeArrays * T);
        NRT_HIP(hipMemsetAsync(e->tw64.p, 0, sizeof(double) * kTreeArrays * T, s));
        if (e->local_res) {
          e->tw64alt.alloc(kTreeArrays * T);
          NRT_HIP(hipMemsetAsync(e->tw64alt.p, 0, sizeof(double) * kTreeArrays * T, s));
        }
      } else {
        e->tw32.alloc(kTreeArrays * T);
        NRT_HIP(hipMemsetAsync(e->tw32.p, 0, sizeof(float) * kTreeArrays * T, s));
        if (e->local_res) {
          e->tw32alt.alloc(kTreeArrays * T);
          NRT_HIP(hipMemsetAsync(e->tw32alt.p, 0, sizeof(float) * kTreeArrays * T, s));
        }
      }
    } else if (o.krylov == NRT_GMRES) {
      e->gmres = true;
      e->gm = std::min(o.restart, N);
      e->nblocks = choose_blocks(hnd, N, mixed ? (const void*)k_implicit_gmres<float>
                                               : (const void*)k_implicit_gmres<double>);
      e->threads = kImpBlock;
      const int m = e->gm;
      if (mixed) {
        e->gV32.alloc((size_t)(m + 1) * N);
        e->gW32.alloc((size_t)2 * N);
        NRT_HIP(hipMemsetAsync(e->gW32.p, 0, sizeof(float) * 2 * N, s));
        e->val32.alloc(nnz);
        hipLaunchKernelGGL(k_rows_scaled_f32, dim3(grid_for(N, 256)), dim3(256), 0, s, N,
                           (const int32_t*)e->rowptr.p, (const int32_t*)e->col.p,
                           (const double*)e->val.p, e->val32.p);
        NRT_HIP(hipGetLastError());
      } else {
        e->gV.alloc((size_t)(m + 1) * N);
      }
      const size_t rstride = (size_t)m * (m + 1) / 2 + 4 * (size_t)m + 1;
      e->gR.alloc(rstride * e->nblocks);
      e->gpart.alloc((size_t)2 * e->nblocks * (m + 1));
    } else {
      e->nblocks = choose_blocks(hnd, N);
      e->threads = kImpBlock;
    }
    // partial sums: two alternating sets of 4 per workgroup (a batch slot may
    // run as a virtual grid of up to kMaxSlotBlocks workgroups)
    e->part.alloc((size_t)std::max(e->nblocks, batch_threads > 0 ? kMaxSlotBlocks : 1) * 4 * 2);
    e->bar.alloc(4);
    NRT_HIP(hipMemsetAsync(e->bar.p, 0, sizeof(unsigned) * 4, s));
    e->st.alloc(1);
    NRT_HIP(hipMemsetAsync(e->st.p, 0, sizeof(ImpStats), s));
    NRT_HIP(hipStreamSynchronize(s));
    // assembly-only inputs
    for (DBuf<int32_t>* ib : {&e->len, &e->scan_tmp, &e->mother, &e->dr, &e->dl}) ib->release();
    e->Wc.release();
    e->branching.release();
  } catch (...) {
    delete e;
    throw;
  }
  return e;
}

void implicit_destroy(Implicit* e) { delete e; }

// Device view of a handle's state (step window and trace left for the caller).
static ImpDev imp_dev(nrt_handle* hnd) {
  Implicit* e = hnd->im;
  const nrt_options& o = hnd->opt;
  ImpDev d{};
  d.N = e->N;
  d.H = e->H;
  d.nblocks = e->nblocks;
  d.rpb = grid_for(e->method != 0 ? e->ns0 : e->N, e->nblocks);
  d.dT = e->dT;
  d.dt_ms = e->dT * 1000;
  d.rowptr = e->rowptr.p;
  d.col = e->col.p;
  d.val = e->val.p;
  d.cmdx = e->cmdx.p;
  d.kdx = e->kdx.p;
  d.input = e->input.p;
  d.rdx = e->rdx.p;
  d.dxv = e->dxv.p;
  d.code = e->code.p;
  d.hh = ImpHH{e->m.p, e->h.p, e->n.p, e->E_Na.p, e->E_K.p, e->E_L.p, e->g_Na.p, e->g_K.p,
               e->G_L.p, e->ls_Na.p, e->ls_K.p, e->rest.p, e->hh_type.p, e->hh_mat.p, e->hh_dl.p,
               e->hh_row.p};
  d.x = e->x.p;
  d.b = e->b.p;
  d.r0 = e->r0.p;
  d.r = e->r.p;
  d.p0 = e->p[0].p;
  d.p1 = e->p[1].p;
  d.v0 = e->v[0].p;
  d.v1 = e->v[1].p;
  d.s = e->s.p;
  d.t = e->t.p;
  d.part = e->part.p;
  d.bar = e->bar.p;
  d.err = e->bar.p + 2;
  d.st = e->st.p;
  d.tol = o.tol;
  d.tol_mode = o.tol_mode;
  d.x0_mode = o.x0_mode;
  d.check_every = std::max(1, o.check_every);
  d.max_iter = o.max_iter > 0 ? o.max_iter : e->N;
  d.spin_limit = 1ull << 28;
  if (const char* sl = getenv("NRT_SPIN_LIMIT")) d.spin_limit = strtoull(sl, nullptr, 10);
  d.interval = -1;
  d.trace_elem = 0;
  d.ns0 = e->ns0;
  d.scmdx = e->scmdx.p;
  d.skdx = e->skdx.p;
  d.sinput = e->sinput.p;
  d.scode = e->scode.p;
  d.srowptr = e->srowptr.p;
  d.scol = e->scol.p;
  d.srole = e->srole.p;
  d.hh_slot = e->hh_slot.p;
  return d;
}

// The probe's index in the vectors a launch works on: the element itself, or its
// level-0 slot for the tree solve.
static int32_t probe_index(const Implicit* e, int32_t elem, int64_t interval) {
  if (interval < 0 || elem < 0 || elem >= e->N) return 0;  // no probe: never read
  return e->method != 0 ? e->plan_h.slot0[elem] : elem;
}

template <class T>
static TreeWork<T> tree_work(T* b, size_t n) {
  return TreeWork<T>{b,         b + n,     b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n,
                     b + 6 * n, b + 7 * n, b + 8 * n, b + 9 * n, b + 10 * n};
}

template <class T>
static TreeArgs<T> tree_args(Implicit* e, const nrt_options& o, T* b, T* alt = nullptr) {
  const size_t T_ = (size_t)e->plan_h.total_nodes();
  const TreeWork<T> w = tree_work(b, T_);
  return TreeArgs<T>{e->plan,
                     w,
                     e->method == 1 ? 1 : 0,
                     e->method >= 3 ? 1 : 0,
                     o.refine_tol,
                     o.refine_max,
                     (e->grid_mode && e->plan_h.top_n > 0) ||
                             (e->local_mode && e->plan_h.top_lds_ok) ||
                             (!e->grid_mode && !e->local_mode &&
                              forest_lds_bytes<T>(e->plan) > 0 &&
                              forest_lds_bytes<T>(e->plan) <= e->forest_lds_max)
                         ? 1
                         : 0,
                     e->local_mode ? e->hh_goff.p : nullptr,
                     alt ? tree_work(alt, T_) : w};
}

// Persistent kernels are launched as ordinary kernels: their grids are sized to be
// co-resident (choose_blocks: at most the occupancy limit, normally one workgroup
// per CU), and their grid barrier is bounded, so a grid that could not all be
// resident ends in an error status instead of a wait without end. A cooperative
// launch is not used: it takes the device's cooperative queue, which one process
// at a time can hold, so a second process on the same GPU (a driver run beside a
// test run, two simulations) would be refused its launches.
static int launch_persistent(nrt_handle* hnd, int64_t first, int64_t num, int64_t a0, int64_t a1,
                             int32_t trace_elem, int64_t interval, double* trace_dev,
                             int64_t gl0 = 0) {
  Implicit* e = hnd->im;
  const nrt_options& o = hnd->opt;
  ImpDev d = imp_dev(hnd);
  d.first = first;
  d.num = num;
  d.a0 = a0;
  d.a1 = a1;
  d.interval = interval;
  d.gl0 = gl0;
  d.trace_elem = probe_index(e, trace_elem, interval);
  d.trace = trace_dev;
  if (e->method == 0 && e->gmres) {
    const int m = e->gm;
    const int64_t rstride = (int64_t)m * (m + 1) / 2 + 4 * (int64_t)m + 1;
    if (o.precision == NRT_FP32_REFINE) {
      GmresDev<float> g{m, e->gV32.p, e->gW32.p, e->gW32.p + e->N, e->gR.p, rstride,
                        e->gpart.p, e->val32.p, o.refine_tol};
      void* args[] = {&d, &g};
      NRT_HIP(hipLaunchKernel((const void*)k_implicit_gmres<float>, dim3(e->nblocks),
                                         dim3(kImpBlock), args, 0, hnd->stream));
    } else {
      GmresDev<double> g{m, e->gV.p, e->t.p, e->s.p, e->gR.p, rstride, e->gpart.p, e->val.p,
                         0.0};
      void* args[] = {&d, &g};
      NRT_HIP(hipLaunchKernel((const void*)k_implicit_gmres<double>, dim3(e->nblocks),
                                         dim3(kImpBlock), args, 0, hnd->stream));
    }
    return NRT_OK;
  }
  if (e->method == 0) {
    void* args[] = {&d};
    NRT_HIP(hipLaunchKernel((const void*)k_implicit_bicgstab_fp64, dim3(e->nblocks),
                                       dim3(kImpBlock), args, 0, hnd->stream));
    return NRT_OK;
  }
  auto launch = [&](auto targs, const void* kgrid, const void* kblock, const void* klocal) {
    void* args[] = {&d, &targs};
    using TT = typename std::remove_pointer<decltype(targs.w.d)>::type;
    const size_t shmem = targs.top_lds ? top_lds_bytes<TT>() : 0;
    if (e->local_mode)
      NRT_HIP(hipLaunchKernel(klocal, dim3(e->nblocks), dim3(kLocalBlock), args, e->local_lds,
                              hnd->stream));
    else if (e->grid_mode)
      NRT_HIP(hipLaunchKernel(kgrid, dim3(e->nblocks), dim3(kImpBlock), args, shmem,
                                         hnd->stream));
    else
      NRT_HIP(hipLaunchKernel(kblock, dim3(1), dim3(kTreeBlock), args,
                              targs.top_lds ? forest_lds_bytes<TT>(targs.P) : 0, hnd->stream));
  };
  if (e->method == 1 || e->method == 3)
    launch(tree_args<double>(e, o, e->tw64.p, e->tw64alt.p), (const void*)k_implicit_tree<double, true>,
           (const void*)k_implicit_tree<double, false>,
           (const void*)(e->local_res ? (const void*)k_implicit_tree_local<double, true>
                                      : (const void*)k_implicit_tree_local<double, false>));
  else
    launch(tree_args<float>(e, o, e->tw32.p, e->tw32alt.p), (const void*)k_implicit_tree<float, true>,
           (const void*)k_implicit_tree<float, false>,
           (const void*)(e->local_res ? (const void*)k_implicit_tree_local<float, true>
                                      : (const void*)k_implicit_tree_local<float, false>));
  return NRT_OK;
}

// Zero the grid barrier's words (arrivals, generation, error flag) after a launch
// that timed out in it; the copy follows the launch on the stream.
static void reset_barrier(Implicit* e, hipStream_t s) {
  static const unsigned zeros[4] = {0u, 0u, 0u, 0u};
  NRT_HIP(hipMemcpyAsync(e->bar.p, zeros, sizeof(zeros), hipMemcpyHostToDevice, s));
  NRT_HIP(hipStreamSynchronize(s));
}

// What a failed step reports, in the terms of its method.
static std::string failure_message(const Implicit* e, const nrt_options& o, const ImpStats& st) {
  char msg[320];
  const double rel = st.fail_bb > 0 ? sqrt(st.fail_rr / st.fail_bb) : 0.0;
  const char* what = st.status == NRT_ERR_NUMERIC ? "non-finite value" : "not converged";
  if (e->method == 0 && !e->gmres) {
    snprintf(msg, sizeof(msg),
             "BiCGSTAB step %lld, %s after %d iterations: |r| = %.3e (relative %.3e, tol %.1e "
             "%s); rho=%.3e (r,r0)-term=%.3e (t,s)=%.3e (t,t)=%.3e; breakdown restarts %d",
             st.fail_step, what, st.fail_iter, sqrt(st.fail_rr), rel, o.tol,
             o.tol_mode == NRT_TOL_RELATIVE ? "relative" : "absolute", st.fail_rho, st.fail_apr,
             st.fail_w1, st.fail_w2, st.fail_cycles);
  } else if (e->method == 0) {
    snprintf(msg, sizeof(msg),
             "GMRES(%d) step %lld, %s after %d Arnoldi steps in %d restart cycles: |r| = %.3e "
             "(relative %.3e; %s)",
             e->gm, st.fail_step, what, st.fail_iter, st.fail_cycles, sqrt(st.fail_rr), rel,
             o.precision == NRT_FP32_REFINE ? "Jacobi-scaled, refine_tol" : "tol");
  } else if (e->method >= 3) {
    snprintf(msg, sizeof(msg),
             "BiCGSTAB preconditioned by the %s tree solve, step %lld, %s after %d iterations "
             "(%d tree solves): |r| = %.3e (relative %.3e, tol %.1e %s)",
             e->method == 3 ? "FP64" : "FP32", st.fail_step, what, st.fail_iter,
             st.fail_cycles, sqrt(st.fail_rr), rel, o.tol,
             o.tol_mode == NRT_TOL_RELATIVE ? "relative" : "absolute");
  } else if (e->method == 2) {
    snprintf(msg, sizeof(msg),
             "FP32 tree solve in FP64 refinement, step %lld, %s after %d sweeps: scaled "
             "residual max norm %.3e (relative %.3e, refine_tol %.1e, refine_max %d)",
             st.fail_step, what, st.fail_cycles, sqrt(st.fail_rr), rel, o.refine_tol,
             o.refine_max);
  } else {
    snprintf(msg, sizeof(msg), "FP64 tree solve, step %lld: %s", st.fail_step, what);
  }
  return msg;
}

static int collect_stats(nrt_handle* hnd) {
  Implicit* e = hnd->im;
  ImpStats st{};
  unsigned err = 0;
  NRT_HIP(hipMemcpyAsync(&st, e->st.p, sizeof(ImpStats), hipMemcpyDeviceToHost, hnd->stream));
  NRT_HIP(hipMemcpyAsync(&err, e->bar.p + 2, sizeof(unsigned), hipMemcpyDeviceToHost, hnd->stream));
  NRT_HIP(hipStreamSynchronize(hnd->stream));
  hnd->stats.steps = (int64_t)st.steps;
  hnd->stats.krylov_iters = (int64_t)st.iters_total;
  hnd->stats.refine_sweeps =
      hnd->opt.precision == NRT_FP32_REFINE || e->method >= 3 ? (int64_t)st.sweeps_total : 0;
  hnd->stats.last_iters = st.last_iters;
  hnd->stats.last_residual = sqrt(st.last_rr);
  hnd->stats.last_rhs_norm = sqrt(st.last_bb);
  hnd->stats.last_status = st.status;
  hnd->stats.matrix_nnz = e->nnz;
  hnd->stats.num_blocks = e->nblocks;
  hnd->stats.tree_form = e->method == 0      ? -1
                         : e->local_mode     ? (e->local_res ? NRT_TREE_LOCAL_LDS : NRT_TREE_LOCAL)
                         : e->grid_mode      ? NRT_TREE_GRID
                                             : NRT_TREE_BLOCK;
  hnd->stats.restarts = st.restarts;
  if (err) {
    // A timed-out barrier leaves its arrival count and the error flag set; the
    // launch has drained (every workgroup leaves at the flag), so clear them for
    // the handle's next launch.
    reset_barrier(e, hnd->stream);
    throw Error(NRT_ERR_HIP,
                "grid barrier timed out inside the persistent kernel (its workgroups were not "
                "all resident at once: another launch held part of the device)");
  }
  // a reported failure is returned once, then cleared
  if (st.status != NRT_OK) {
    hnd->last_error = failure_message(e, hnd->opt, st);
    const int zero = 0;
    NRT_HIP(hipMemcpyAsync(&e->st.p->status, &zero, sizeof(int), hipMemcpyHostToDevice,
                           hnd->stream));
    NRT_HIP(hipStreamSynchronize(hnd->stream));
  }
  return st.status;
}

void implicit_update(nrt_handle* hnd, bool active) { hnd->input_active = active; }

int implicit_calculate(nrt_handle* hnd) {
  Implicit* e = hnd->im;
  const int64_t st = e->step;
  launch_persistent(hnd, st, 1, hnd->input_active ? st : st + 1, st, 0, -1, nullptr);
  e->step++;
  return collect_stats(hnd);
}

// V in element order, copied on stream s (the handle's own stream, or another
// one when the caller knows that no queued work writes V).
void implicit_snapshot_on(nrt_handle* hnd, double* host_out, hipStream_t s) {
  Implicit* e = hnd->im;
  if (e->method == 0) {
    NRT_HIP(hipMemcpyAsync(host_out, e->x.p, sizeof(double) * e->N, hipMemcpyDeviceToHost, s));
    NRT_HIP(hipStreamSynchronize(s));
    return;
  }
  std::vector<double> xs(e->ns0);  // slot order -> element order
  NRT_HIP(hipMemcpyAsync(xs.data(), e->x.p, sizeof(double) * e->ns0, hipMemcpyDeviceToHost, s));
  NRT_HIP(hipStreamSynchronize(s));
  const int32_t* s0 = e->plan_h.slot0.data();
  for (int i = 0; i < e->N; i++) host_out[i] = xs[s0[i]];
}

void implicit_snapshot(nrt_handle* hnd, double* host_out) {
  implicit_snapshot_on(hnd, host_out, hnd->stream);
}

int implicit_run(nrt_handle* hnd, int64_t first, int64_t num, int64_t a0, int64_t a1,
                 int32_t trace_elem, int64_t interval, double* trace_out, int64_t* ns) {
  Implicit* e = hnd->im;
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
  launch_persistent(hnd, first, num, a0, a1, trace_elem, interval,
                    (tracing && trace_out) ? e->trace.p : nullptr);
  NRT_HIP(hipEventRecord(t1, hnd->stream));
  const int64_t k = max_samples;
  if (trace_out && k > 0)
    NRT_HIP(hipMemcpyAsync(trace_out, e->trace.p, sizeof(double) * k, hipMemcpyDeviceToHost,
                           hnd->stream));
  NRT_HIP(hipStreamSynchronize(hnd->stream));
  float ms = 0;
  NRT_HIP(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  e->step = first + num;
  const int status = collect_stats(hnd);
  hnd->stats.kernel_ms = ms;
  hnd->stats.kernel_launches = 1;
  if (ns) *ns = k;
  return status;
}

// ------------------------------------------------------------ batch support
size_t implicit_slot_bytes(int precision) {
  return precision == NRT_FP64 ? sizeof(BatchSlot<double>) : sizeof(BatchSlot<float>);
}

void implicit_slot_fill(nrt_handle* h, long long a0, long long a1, int trace_elem,
                        long long interval, double* trace, void* out) {
  Implicit* e = h->im;
  NRT_REQUIRE(e->method != 0 && !e->grid_mode && !e->local_mode, NRT_ERR_UNSUPPORTED,
              "batch slots use the block-mode tree solve");
  ImpDev d = imp_dev(h);
  d.a0 = a0;
  d.a1 = a1;
  d.interval = interval;
  d.trace_elem = probe_index(e, trace_elem, interval);
  if (e->method == 1 || e->method == 3) {
    BatchSlot<double> S{d, tree_args<double>(e, h->opt, e->tw64.p), trace};
    std::memcpy(out, &S, sizeof(S));
  } else {
    BatchSlot<float> S{d, tree_args<float>(e, h->opt, e->tw32.p), trace};
    std::memcpy(out, &S, sizeof(S));
  }
}

size_t implicit_slot_work_bytes() { return sizeof(SlotWork); }

void implicit_slot_work(void* out, long long done, long long steps, int vn) {
  SlotWork w{done, steps, vn, 0};
  std::memcpy(out, &w, sizeof(w));
}

// Solve form of a batch's slots (all alike): the kernel instantiation to launch.
static int batch_mode(int precision, int krylov) {
  if (krylov == NRT_BICGSTAB) return kSolveKrylov;
  return precision == NRT_FP64 ? kSolveDirect : kSolveRefine;
}

// The k_batch_tree instantiation for (value type, solve form, workgroup size).
template <class T, int MODE>
static const void* batch_kernel_of(int threads) {
  if (threads == 128) return (const void*)k_batch_tree<T, 128, MODE>;
  if (threads == 256) return (const void*)k_batch_tree<T, 256, MODE>;
  if (threads == 512) return (const void*)k_batch_tree<T, 512, MODE>;
  return (const void*)k_batch_tree<T, 1024, MODE>;
}
static const void* batch_kernel(int precision, int krylov, int threads) {
  const int mode = batch_mode(precision, krylov);
  if (precision == NRT_FP64)
    return mode == kSolveKrylov ? batch_kernel_of<double, kSolveKrylov>(threads)
                                : batch_kernel_of<double, kSolveDirect>(threads);
  return mode == kSolveKrylov ? batch_kernel_of<float, kSolveKrylov>(threads)
                              : batch_kernel_of<float, kSolveRefine>(threads);
}

// Dynamic LDS a batch launch gives each workgroup for a staged final forest: as much
// as leaves the kernel's occupancy (set by its registers) unchanged, at most
// kForestLdsMax. Per device and kernel, computed once.
static size_t batch_forest_lds(int precision, int krylov, int threads) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int>, size_t> known;
  int dev = 0;
  NRT_HIP(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, precision, batch_mode(precision, krylov), threads);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = known.find(key);
    if (it != known.end()) return it->second;
  }
  const void* k = batch_kernel(precision, krylov, threads);
  int per0 = 0, per_cu_lds = 0, per_block = 0;
  NRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per0, k, threads, 0));
  NRT_HIP(hipDeviceGetAttribute(&per_cu_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                dev));
  NRT_HIP(hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  hipFuncAttributes fa{};
  NRT_HIP(hipFuncGetAttributes(&fa, k));
  size_t cap = 0;
  if (per0 > 0 && per_cu_lds > 0) {
    const size_t share = (size_t)per_cu_lds / (size_t)per0;
    cap = share > fa.sharedSizeBytes ? share - fa.sharedSizeBytes : 0;
    const size_t blk = (size_t)std::max(per_block, 0);
    cap = std::min(cap, blk > fa.sharedSizeBytes ? blk - fa.sharedSizeBytes : 0);
    cap = std::min(cap, kForestLdsMax) & ~(size_t)1023;
    int per = 0;
    while (cap > 0) {
      NRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, threads, cap));
      if (per >= per0) break;
      cap -= 1024;
    }
  }
  std::lock_guard<std::mutex> g(mu);
  known[key] = cap;
  return cap;
}

// Workgroups of the batched kernel that can be resident at once (an otherwise
// idle device): virtual grids of several workgroups are only formed within it.
int implicit_batch_capacity(int precision, int krylov, int threads) {
  int dev = 0, cus = 0, per = 0;
  NRT_HIP(hipGetDevice(&dev));
  NRT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  NRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per, batch_kernel(precision, krylov, threads), threads,
      batch_forest_lds(precision, krylov, threads)));
  return cus * std::max(per, 0);
}

size_t implicit_wgrec_bytes(int precision, int nwg) {
  return (size_t)std::max(nwg, 0) * implicit_slot_bytes(precision);
}

void implicit_batch_expand(int precision, const void* table, const void* work,
                           const int2* wgmap, int nwg, void* wgrec, hipStream_t s) {
  if (nwg <= 0) return;
  const auto* w = (const SlotWork*)work;
  if (precision == NRT_FP64)
    hipLaunchKernelGGL(k_batch_expand<double>, dim3(nwg), dim3(64), 0, s,
                       (const BatchSlot<double>*)table, w, wgmap, (BatchSlot<double>*)wgrec);
  else
    hipLaunchKernelGGL(k_batch_expand<float>, dim3(nwg), dim3(64), 0, s,
                       (const BatchSlot<float>*)table, w, wgmap, (BatchSlot<float>*)wgrec);
  NRT_HIP(hipGetLastError());
}

void implicit_batch_launch(int precision, int krylov, int threads, const void* wgrec, int nwg,
                           hipStream_t s) {
  if (nwg <= 0) return;
  int lds = (int)batch_forest_lds(precision, krylov, threads);
  void* args[] = {(void*)&wgrec, (void*)&lds};
  NRT_HIP(hipLaunchKernel(batch_kernel(precision, krylov, threads), dim3(nwg), dim3(threads), args,
                          (size_t)lds, s));
}

// Steps first..first+num-1 enqueued on the handle's stream without waiting; the
// trace continues a run whose phase is gl0 (steps since its last sample).
void implicit_steps_async(nrt_handle* h, int64_t first, int64_t num, int64_t a0, int64_t a1,
                          int32_t trace_elem, int64_t interval, int64_t gl0, double* trace_dev) {
  launch_persistent(h, first, num, a0, a1, trace_elem, interval, trace_dev, gl0);
  h->im->step = first + num;
}

// Status and inner-iteration count of a slot (its failure is reported, not cleared).
void implicit_slot_stats_on(nrt_handle* h, hipStream_t s, int* status, int64_t* inner) {
  Implicit* e = h->im;
  ImpStats st{};
  NRT_HIP(hipMemcpyAsync(&st, e->st.p, sizeof(ImpStats), hipMemcpyDeviceToHost, s));
  NRT_HIP(hipStreamSynchronize(s));
  *status = st.status;
  *inner = (int64_t)st.iters_total;
}

void implicit_slot_stats(nrt_handle* h, int* status, int64_t* inner) {
  implicit_slot_stats_on(h, h->stream, status, inner);
}

int64_t implicit_nnz(const Implicit* e) { return e->nnz; }

// Bytes of device memory a handle holds (approximate: its arrays).
int64_t implicit_device_bytes(const Implicit* e) {
  const int64_t N = e->N, T = e->plan_h.total_nodes(), S = e->ns0;
  return N * (8 * 5 + 4) + S * (8 * 5 + 4) + (int64_t)e->nnz * 13 + (S + 1) * 4 +
         (int64_t)e->H * 8 * 13 + (int64_t)e->plan_buf.n * 4 +
         T * kTreeArrays * (e->method == 1 || e->method == 3 ? 8 : 4) +
         (e->method >= 3 ? S * 8 * 8 : 0);
}

}  // namespace nrt
