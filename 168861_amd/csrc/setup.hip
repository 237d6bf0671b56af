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
#include "nrt_internal.hpp"

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

__global__ void k_setup(SetupDev s) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= s.N) return;
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
    const int hi = s.hh_index[k];
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

}  // namespace nrt
