// Setup chain on the device: raw geometry (what a space.cpp builder produces) to
// the derived per-element arrays the solvers consume, one thread per element.
//
// Follows, expression for expression:
//   dX' = dX*(1+eps)                                  init.cpp:329-331
//   eps_s = sqrt(1+eps)-1                             discretization.h:50-53
//   HH: d' = d/sqrt(1+eps); r_a, r_m, c_m             init.cpp:239-241, discretization.h:61-65
//       E_Na, E_K damage                              discretization.h:216-227
//       g_L, g_Na, g_K; left shifts; g/sqrt(1+eps)    discretization.h:208-213,230-238; init.cpp:250-251
//       m_inf, h_inf, n_inf, E_L balancing            init.cpp:280-297, coupling.cpp:25-48
//   CT: d' = d/sqrt(1+eps); r_a, r_m, c_m; myelin     init.cpp:184-196, discretization.h:126-151
//       W = -1/r_m, K = Vr/r_m                        discretization.cpp:63-68
//   critical dT per element, min over 1..N-2          init.cpp:138-174
// The constant pi is the reference's 3.14159265 (neurite.cpp:72).
#include <cstring>
#include <mutex>
#include "nrt_internal.hpp"
#include "nrt_device.hpp"

namespace nrt {

constexpr double kPi = 3.14159265;

struct SetupDev {
  int N;
  const double* diameter;
  const double* dx0;
  const double* epsilon;
  const int32_t* material;
  const nrt_material* mats;
  const int32_t* hh_index;  // -1 for passive elements
  int myelinate;
  double EMAX;
  double* dx;
  double* c_m;
  double* r_a;
  double* r_m;
  double* W;
  double* K;
  double* crit;
  double* rest;
  double* m;
  double* h;
  double* n;
  double* E_Na;
  double* E_K;
  double* E_L;
  double* g_Na;
  double* g_K;
  double* G_L;
  double* ls_Na;
  double* ls_K;
  double* hh_rest;
  unsigned long long* crit_min_bits;
};

// Derived arrays of element k; hi is its index among the HH elements (-1 when
// passive). HH outputs are written only below hh_cap (the capacity of the HH
// arrays; the host path sizes them exactly).
__device__ __forceinline__ void setup_element(const SetupDev& s, int k, int hi, int hh_cap) {
  const nrt_material mt = s.mats[s.material[k]];
  const double eps = s.epsilon[k];
  const double eps_s = sqrt(1 + eps) - 1;
  const double dx = s.dx0[k] * (1 + eps);
  s.dx[k] = dx;
  s.rest[k] = mt.rest_pot;
  double d = s.diameter[k];
  d = d / (sqrt(1 + eps));
  const double ra = 4 * mt.rho_a / (kPi * d * d);
  double rm = (mt.rho_m * mt.thickness) / (kPi * d);
  double cm = (mt.per_m * kPi * d) / (mt.thickness);
  double crit;
  if (mt.kind == 1) {
    const double rest = mt.rest_pot;
    const double ENa0 = mt.dE_Na + rest, EK0 = mt.dE_K + rest;
    double ENa, EK, lsNa, lsK;
    if (eps_s <= s.EMAX) {
      ENa = ENa0 - pow((eps_s / s.EMAX), mt.n_damage) * (ENa0);
      EK = EK0 - pow((eps_s / s.EMAX), mt.n_damage) * (EK0);
      lsNa = pow(eps_s / s.EMAX, mt.n_damage) * (ENa0);
      lsK = pow(eps_s / s.EMAX, mt.n_damage) * (EK0);
    } else {
      ENa = 0;
      EK = 0;
      lsNa = ENa0;
      lsK = EK0;
    }
    const double gL = mt.sigma_L * kPi * d / mt.thickness;
    double gNa = mt.sigma_Na * kPi * d / mt.thickness;
    double gK = mt.sigma_K * kPi * d / mt.thickness;
    gNa = gNa / (sqrt(1 + eps));
    gK = gK / (sqrt(1 + eps));
    // hhIconds: rates at the resting potential
    double vna = rest + lsNa, vk = rest + lsK;
    vna = (vna - rest) / 0.001;
    vk = (vk - rest) / 0.001;
    const double am = (2.5 - 0.1 * vna) / (exp(2.5 - 0.1 * vna) - 1);
    const double bm = 4 * exp(-vna / 18);
    const double ah = 0.07 * exp(-vna / 20);
    const double bh = 1 / (exp(3 - 0.1 * vna) + 1);
    const double an = (0.1 - 0.01 * vk) / (exp(1 - 0.1 * vk) - 1);
    const double bn = 0.125 * exp(-vk / 80);
    const double m = am / (am + bm), h = ah / (ah + bh), n = an / (an + bn);
    const double vlost =
        (gNa * m * m * m * h * (rest - ENa) + gK * n * n * n * n * (rest - EK) + gL * rest) / gL;
    if (hi >= 0 && hi < hh_cap) {
      s.m[hi] = m;
      s.h[hi] = h;
      s.n[hi] = n;
      s.E_Na[hi] = ENa;
      s.E_K[hi] = EK;
      s.E_L[hi] = vlost;
      s.g_Na[hi] = gNa;
      s.g_K[hi] = gK;
      s.G_L[hi] = gL;
      s.ls_Na[hi] = lsNa;
      s.ls_K[hi] = lsK;
      s.hh_rest[hi] = rest;
    }
    s.W[k] = 0;
    s.K[k] = 0;
    crit = (2 * ra * cm * dx * dx) / 4;
  } else {
    if (s.myelinate && mt.layers_my != 0) {
      double temp = 0;
      for (int i = 1; i <= mt.layers_my; i++)
        temp = temp + 1 / (d + mt.thickness * 2 + 2 * (i - 1) * mt.thickness_my);
      const double c_my_T = mt.thickness_my * temp / (mt.per_my * kPi);
      const double c_my = 1 / c_my_T;
      const double r_my = (mt.ro_my * mt.thickness_my * temp) / kPi;
      rm = rm + r_my;
      cm = (cm * c_my) / (c_my + cm);
    }
    s.W[k] = -1 / rm;
    s.K[k] = mt.rest_pot / rm;
    crit = (2 * ra * cm * dx * dx) / (4 + (ra * dx * dx) / rm);
  }
  s.r_a[k] = ra;
  s.r_m[k] = rm;
  s.c_m[k] = cm;
  s.crit[k] = crit;
  // critical_time_step_selection: min over 1..N-2 (positive doubles order as
  // their bit patterns, so an integer atomic min is exact and order-free)
  if (k >= 1 && k <= s.N - 2) {
    unsigned long long bits = (unsigned long long)__double_as_longlong(crit);
    if (crit > 0) atomicMin(s.crit_min_bits, bits);
  }
}

__global__ void k_setup(SetupDev s) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= s.N) return;
  setup_element(s, k, s.hh_index[k], 0x7fffffff);
}

int setup_on_device(const nrt_geometry* g, int device, nrt_setup_out* out) {
  NRT_HIP(hipSetDevice(device));
  const int N = g->num_elem;
  NRT_REQUIRE(N >= 3, NRT_ERR_INVALID, "geometry needs at least 3 elements");
  NRT_REQUIRE(g->diameter && g->dx0 && g->epsilon && g->material && g->materials,
              NRT_ERR_INVALID, "geometry arrays missing");
  std::vector<int32_t> hh_index(N, -1);
  int H = 0;
  for (int k = 0; k < N; k++) {
    const int mi = g->material[k];
    NRT_REQUIRE(mi >= 0 && mi < g->num_materials, NRT_ERR_INVALID, "material index out of range");
    if (g->materials[mi].kind == 1) hh_index[k] = H++;
  }
  hipStream_t st;
  NRT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  DBuf<double> diam, dx0, eps, dx, cm, ra, rm, W, K, crit, rest;
  DBuf<int32_t> mat, hidx;
  DBuf<nrt_material> mats;
  DBuf<double> m, h, n, ENa, EK, EL, gNa, gK, GL, lsNa, lsK, hrest;
  DBuf<unsigned long long> cmin;
  diam.upload(g->diameter, N, st);
  dx0.upload(g->dx0, N, st);
  eps.upload(g->epsilon, N, st);
  mat.upload(g->material, N, st);
  mats.upload(g->materials, g->num_materials, st);
  hidx.upload(hh_index, st);
  for (DBuf<double>* b : {&dx, &cm, &ra, &rm, &W, &K, &crit, &rest}) b->alloc(N);
  for (DBuf<double>* b : {&m, &h, &n, &ENa, &EK, &EL, &gNa, &gK, &GL, &lsNa, &lsK, &hrest})
    b->alloc(std::max(H, 1));
  cmin.alloc(1);
  const unsigned long long init = ~0ull;
  NRT_HIP(hipMemcpyAsync(cmin.p, &init, sizeof(init), hipMemcpyHostToDevice, st));
  SetupDev s{N, diam.p, dx0.p, eps.p, mat.p, mats.p, hidx.p, g->myelinate, g->E_MAX,
             dx.p, cm.p, ra.p, rm.p, W.p, K.p, crit.p, rest.p, m.p, h.p, n.p, ENa.p, EK.p, EL.p,
             gNa.p, gK.p, GL.p, lsNa.p, lsK.p, hrest.p, cmin.p};
  hipLaunchKernelGGL(k_setup, dim3(grid_for(N, 256)), dim3(256), 0, st, s);
  NRT_HIP(hipGetLastError());
  auto down = [&](double* dst, const DBuf<double>& src, int n_) {
    if (dst && n_ > 0)
      NRT_HIP(hipMemcpyAsync(dst, src.p, sizeof(double) * n_, hipMemcpyDeviceToHost, st));
  };
  down(out->dx, dx, N);
  down(out->c_m, cm, N);
  down(out->r_a, ra, N);
  down(out->r_m, rm, N);
  down(out->W, W, N);
  down(out->K, K, N);
  down(out->crit_dT, crit, N);
  down(out->rest, rest, N);
  down(out->hh_m, m, H);
  down(out->hh_h, h, H);
  down(out->hh_n, n, H);
  down(out->E_Na, ENa, H);
  down(out->E_K, EK, H);
  down(out->E_L, EL, H);
  down(out->g_Na, gNa, H);
  down(out->g_K, gK, H);
  down(out->G_L, GL, H);
  down(out->ls_Na, lsNa, H);
  down(out->ls_K, lsK, H);
  down(out->hh_rest, hrest, H);
  unsigned long long bits = 0;
  NRT_HIP(hipMemcpyAsync(&bits, cmin.p, sizeof(bits), hipMemcpyDeviceToHost, st));
  NRT_HIP(hipStreamSynchronize(st));
  if (out->hh_elem)
    for (int k = 0; k < N; k++)
      if (hh_index[k] >= 0) out->hh_elem[hh_index[k]] = k;
  double cmin_v;
  std::memcpy(&cmin_v, &bits, sizeof(double));
  out->dT_crit = (bits == ~0ull) ? 0.0 : cmin_v;
  out->num_hh = H;
  return NRT_OK;
}

// ------------------------------------------------ setup without allocation
// nrt_setup_noalloc: the same per-element setup over arrays the caller already
// holds on the device, enqueued on the caller's stream, allocating nothing. What
// nrt_setup computes on the host (each element's index among the HH elements) is
// a device scan here: per-block counts of HH elements, an exclusive scan of those
// counts, then each element's rank within its block. The block counts, the
// critical-step minimum and a bad-material flag live in `scratch`, one fixed
// array per device: calls that overlap on one device race on it.
constexpr int kSetupBlock = 256;
constexpr int kSetupMaxBlocks = 1 << 18;  // 67,108,864 elements at most
constexpr size_t kSetupScratchHead = 256;   // min bits (8 B), bad-material flag (4 B)
constexpr size_t kSetupScratchBytes = kSetupScratchHead + sizeof(int32_t) * kSetupMaxBlocks;
static __device__ __align__(256) unsigned char scratch[kSetupScratchBytes];

struct SetupScratch {
  unsigned long long* cmin;
  unsigned* bad;
  int32_t* sums;
};

// Clear what a call reads before writing it (a kernel rather than a memset, so the
// sequence can be captured in a graph).
__global__ void k_setup_init(SetupScratch w) {
  if (threadIdx.x == 0) {
    *w.cmin = ~0ull;
    *w.bad = 0u;
  }
}

__device__ __forceinline__ int setup_is_hh(const SetupDev& s, int nmat, int k, unsigned* bad) {
  const int mi = s.material[k];
  if (mi < 0 || mi >= nmat) {
    atomicOr(bad, 1u);
    return -1;
  }
  return s.mats[mi].kind == 1 ? 1 : 0;
}

__global__ __launch_bounds__(kSetupBlock) void k_setup_count(SetupDev s, int nmat,
                                                             SetupScratch w) {
  __shared__ int wsum[kSetupBlock / kWave];
  const int k = blockIdx.x * kSetupBlock + threadIdx.x;
  const int hh = k < s.N ? setup_is_hh(s, nmat, k, w.bad) : 0;
  const unsigned long long b = __ballot(hh == 1);
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int i = 0; i < kSetupBlock / kWave; i++) t += wsum[i];
    w.sums[blockIdx.x] = t;
  }
}

// Exclusive scan of the nb block counts in place (one workgroup), the total to
// *num_hh (-1 when an element named a material out of range).
constexpr int kSetupScan = 1024;
__global__ __launch_bounds__(kSetupScan) void k_setup_scan(int nb, SetupScratch w,
                                                           int32_t* num_hh) {
  __shared__ int sh[kSetupScan / kWave + 1];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  int carry = 0;
  for (int base = 0; base < nb; base += kSetupScan) {
    const int i = base + threadIdx.x;
    const int v = i < nb ? w.sums[i] : 0;
    int x = v;
    for (int o = 1; o < kWave; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == kWave - 1) sh[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0;
      for (int q = 0; q < kSetupScan / kWave; q++) {
        const int t = sh[q];
        sh[q] = acc;
        acc += t;
      }
      sh[kSetupScan / kWave] = acc;
    }
    __syncthreads();
    if (i < nb) w.sums[i] = carry + sh[wid] + x - v;
    carry += sh[kSetupScan / kWave];
    __syncthreads();
  }
  if (threadIdx.x == 0) *num_hh = *w.bad ? -1 : carry;
}

__global__ __launch_bounds__(kSetupBlock) void k_setup_scanned(SetupDev s, int nmat,
                                                               SetupScratch w, int hh_cap,
                                                               int32_t* hh_elem) {
  __shared__ int woff[kSetupBlock / kWave];
  const int k = blockIdx.x * kSetupBlock + threadIdx.x;
  const int hh = k < s.N ? setup_is_hh(s, nmat, k, w.bad) : 0;
  const unsigned long long b = __ballot(hh == 1);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) woff[wid] = __popcll(b);
  __syncthreads();
  int off = w.sums[blockIdx.x];
  for (int i = 0; i < wid; i++) off += woff[i];
  if (k >= s.N || hh < 0) return;
  const int hi = hh == 1 ? off + __popcll(b & ((1ull << lane) - 1ull)) : -1;
  setup_element(s, k, hi, hh_cap);
  if (hi >= 0 && hi < hh_cap) hh_elem[hi] = k;
}

__global__ void k_setup_finish(SetupScratch w, double* dT_crit) {
  if (threadIdx.x == 0) {
    const unsigned long long bits = *w.cmin;
    *dT_crit = bits == ~0ull ? 0.0 : __longlong_as_double((long long)bits);
  }
}

// The device address of `scratch` on each device, taken once per device with that
// device current (the array's name is only a host-side stand-in).
static std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

static SetupScratch scratch_on(int device) {
  NRT_REQUIRE(device >= 0 && device < kMaxDevices, NRT_ERR_INVALID, "device index too large");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    NRT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NRT_REQUIRE(p != nullptr, NRT_ERR_HIP, "no device address for the setup scratch");
    g_scratch_ptr[device] = (unsigned char*)p;
  }
  unsigned char* b = g_scratch_ptr[device];
  return SetupScratch{(unsigned long long*)b, (unsigned*)(b + 8),
                      (int32_t*)(b + kSetupScratchHead)};
}

int setup_noalloc_on_device(const nrt_geometry* g, int device, void* stream,
                            const nrt_setup_out* out, int32_t hh_capacity, double* dT_crit,
                            int32_t* num_hh) {
  NRT_REQUIRE(g && out && dT_crit && num_hh, NRT_ERR_INVALID, "NULL argument");
  const int64_t N = g->num_elem;
  NRT_REQUIRE(N >= 3, NRT_ERR_INVALID, "geometry needs at least 3 elements");
  NRT_REQUIRE(N <= (int64_t)kSetupMaxBlocks * kSetupBlock, NRT_ERR_INVALID,
              "nrt_setup_noalloc takes at most 67,108,864 elements (its fixed scratch holds "
              "one count per 256 elements); use nrt_setup");
  NRT_REQUIRE(g->diameter && g->dx0 && g->epsilon && g->material && g->materials &&
                  g->num_materials >= 1,
              NRT_ERR_INVALID, "geometry arrays missing");
  NRT_REQUIRE(out->dx && out->c_m && out->r_a && out->r_m && out->W && out->K && out->crit_dT &&
                  out->rest,
              NRT_ERR_INVALID, "per-element output arrays missing");
  NRT_REQUIRE(hh_capacity >= 0, NRT_ERR_INVALID, "hh_capacity < 0");
  NRT_REQUIRE(hh_capacity == 0 ||
                  (out->hh_elem && out->hh_m && out->hh_h && out->hh_n && out->E_Na &&
                   out->E_K && out->E_L && out->g_Na && out->g_K && out->G_L && out->ls_Na &&
                   out->ls_K && out->hh_rest),
              NRT_ERR_INVALID, "HH output arrays missing");
  NRT_HIP(hipSetDevice(device));
  const SetupScratch w = scratch_on(device);
  hipStream_t st = (hipStream_t)stream;
  SetupDev s{(int)N,      g->diameter, g->dx0,      g->epsilon, g->material, g->materials,
             nullptr,     g->myelinate, g->E_MAX,   out->dx,    out->c_m,    out->r_a,
             out->r_m,    out->W,       out->K,     out->crit_dT, out->rest, out->hh_m,
             out->hh_h,   out->hh_n,    out->E_Na,  out->E_K,   out->E_L,    out->g_Na,
             out->g_K,    out->G_L,     out->ls_Na, out->ls_K,  out->hh_rest, w.cmin};
  const int nb = grid_for(N, kSetupBlock);
  hipLaunchKernelGGL(k_setup_init, dim3(1), dim3(64), 0, st, w);
  hipLaunchKernelGGL(k_setup_count, dim3(nb), dim3(kSetupBlock), 0, st, s, g->num_materials, w);
  hipLaunchKernelGGL(k_setup_scan, dim3(1), dim3(kSetupScan), 0, st, nb, w, num_hh);
  hipLaunchKernelGGL(k_setup_scanned, dim3(nb), dim3(kSetupBlock), 0, st, s, g->num_materials, w,
                     (int)hh_capacity, out->hh_elem);
  hipLaunchKernelGGL(k_setup_finish, dim3(1), dim3(64), 0, st, w, dT_crit);
  NRT_HIP(hipGetLastError());
  return NRT_OK;
}

}  // namespace nrt
