// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of Neurite's (ModelDBRepository/168861) cable-equation solver
// path, used as the checker for the MI355X HIP library. Only tests/, the smoke
// entry point and the CPU-baseline leg of bench.py may load this library; the
// product path never links or calls it.
//
// What each function follows (reference file:line) and what pins it:
//   orc_setup              init.cpp:180-298,320-333; discretization.h:50-238;
//                          discretization.cpp:19-104; init.cpp:138-174
//                          -> pinned bitwise against oracle/_ref builds of the
//                             reference setup chain (tests/golden/*.json).
//   orc_ex_*               solverEx-cpu.cpp:19-129, coupling.cpp:25-48
//                          -> pinned bitwise against oracle/_ref trajectories.
//   orc_im_* variant CPU   solverIm-cpu.cpp:29-183 (createMatrix), :192-272 (FGMRES)
//                          -> matrix pinned by dense brute-force assembly and by the
//                             GPU variant; solution pinned by the exact tree solve
//                             (FGMRES iterates of MKL: parity unpinned, MKL absent).
//   orc_im_* variant GPU   solverIm-gpu.cpp:61-171 (Update_*), :199-369 (assembly),
//                          :725-843 (BiCGSTAB)
//                          -> matrix pinned against the CPU variant (same matrix up
//                             to FP op order) and dense assembly; solves pinned by
//                             the exact tree solve and scipy spsolve.
//   orc_tree_solve         exact Hines-style elimination of the assembled system
//                          -> pinned against numpy.linalg.solve / scipy spsolve.
//   orc_run                neurite.cpp:224-255 (time loop, trace cadence)
//                          -> pinned bitwise against oracle/_ref .dat traces.
//
// Build: g++ -O3 (no -ffast-math, no FMA contraction on x86-64 baseline), the same
// arithmetic as the reference's own CPU build (Makefile:5).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include "../include/neurite_hip.h"

namespace {

const double kPi = 3.14159265;   // neurite.cpp:72 (not M_PI)
const double kScale = 0.001;     // coupling.cpp:30, configuration.h:36

struct Rates { double am, bm, ah, bh, an, bn; };

// coupling.cpp:25-48 (hhRate_Const), expression for expression.
inline Rates hh_rates(double rest_pot, double potential, double ls_na, double ls_k) {
  double vna = potential + ls_na;
  double vk = potential + ls_k;
  vna = (vna - rest_pot) / kScale;
  vk = (vk - rest_pot) / kScale;
  Rates r;
  r.am = (2.5 - 0.1 * vna) / (exp(2.5 - 0.1 * vna) - 1);
  r.bm = 4 * exp(-vna / 18);
  r.ah = 0.07 * exp(-vna / 20);
  r.bh = 1 / (exp(3 - 0.1 * vna) + 1);
  r.an = (0.1 - 0.01 * vk) / (exp(1 - 0.1 * vk) - 1);
  r.bn = 0.125 * exp(-vk / 80);
  return r;
}

// Host copy of an nrt_problem (what the reference solver constructors copy).
struct Prob {
  int N = 0;
  double dT = 0;
  std::vector<double> V0, V1, dx, c_m, r_a, W, K, input;
  std::vector<int> mother, dr, dl;
  std::vector<uint8_t> branching;
  std::vector<int> terminals;
  std::vector<int> hh_elem;
  std::vector<double> m, h, n, E_Na, E_K, E_L, g_Na, g_K, G_L, ls_Na, ls_K, rest, G_Na, G_K, hhW;
};

template <class T>
std::vector<T> vec(const T* p, int n) {
  return p ? std::vector<T>(p, p + n) : std::vector<T>(n, T(0));
}

Prob load(const nrt_problem* p) {
  Prob q;
  q.N = p->num_elem;
  q.dT = p->dT;
  int N = q.N, H = p->num_hh;
  q.V0 = vec(p->potential0, N);
  q.V1 = vec(p->potential1, N);
  q.dx = vec(p->dx, N);
  q.c_m = vec(p->c_m, N);
  q.r_a = vec(p->r_a, N);
  q.W = vec(p->W, N);
  q.K = vec(p->K, N);
  q.input = vec(p->input, N);
  q.mother = vec(p->mother, N);
  q.dr = vec(p->dr, N);
  q.dl = vec(p->dl, N);
  q.branching = vec(p->branching, N);
  q.terminals = vec(p->terminals, p->num_term);
  q.hh_elem = vec(p->hh_elem, H);
  q.m = vec(p->hh_m, H);
  q.h = vec(p->hh_h, H);
  q.n = vec(p->hh_n, H);
  q.E_Na = vec(p->E_Na, H);
  q.E_K = vec(p->E_K, H);
  q.E_L = vec(p->E_L, H);
  q.g_Na = vec(p->g_Na, H);
  q.g_K = vec(p->g_K, H);
  q.G_L = vec(p->G_L, H);
  q.ls_Na = vec(p->ls_Na, H);
  q.ls_K = vec(p->ls_K, H);
  q.rest = vec(p->rest_pot, H);
  q.G_Na.assign(H, 0.0);
  q.G_K.assign(H, 0.0);
  q.hhW.assign(H, 0.0);
  return q;
}

// One HH gating step for HH element i (solverEx-cpu.cpp:62-82 and
// solverIm-cpu.cpp:506-529): forward Euler on m, h, n then W, K.
inline void hh_step(Prob& q, int i, double v_prev, double dt_ms, double* W, double* K) {
  Rates r = hh_rates(q.rest[i], v_prev, q.ls_Na[i], q.ls_K[i]);
  double m0 = q.m[i], h0 = q.h[i], n0 = q.n[i];
  double m = m0 + dt_ms * (r.am * (1 - m0) - r.bm * m0);
  double h = h0 + dt_ms * (r.ah * (1 - h0) - r.bh * h0);
  double n = n0 + dt_ms * (r.an * (1 - n0) - r.bn * n0);
  double gna = q.g_Na[i] * m * m * m * h;
  double gk = q.g_K[i] * n * n * n * n;
  q.G_Na[i] = gna;
  q.G_K[i] = gk;
  *W = -(gna + gk + q.G_L[i]);                              // discretization.cpp:81-83
  *K = gna * q.E_Na[i] + gk * q.E_K[i] + q.G_L[i] * q.E_L[i];  // discretization.cpp:84-86
  q.m[i] = m;
  q.h[i] = h;
  q.n[i] = n;
}

// ------------------------------------------------------------------ explicit
struct Explicit {
  Prob q;
  std::vector<double> cur, prev;
  bool active = false;
};

// ------------------------------------------------------------------ implicit
struct Csr {
  std::vector<int> rowptr, col;
  std::vector<double> val;
};

enum { VAR_CPU = 0, VAR_GPU = 1 };
enum { M_GMRES = 0, M_BICGSTAB = 1, M_DIRECT = 2 };

struct Implicit {
  Prob q;
  int variant = VAR_GPU;
  int method = M_BICGSTAB;
  int restart = 150;
  double tol = 1e-14;
  int tol_mode = NRT_TOL_ABSOLUTE;
  int x0_mode = NRT_X0_RHS;
  int check_every = 30;
  int max_iter = 0;
  bool active = false;
  // CPU variant state (solverIm-cpu.cpp)
  std::vector<double> cur, prev;
  // GPU variant state (solverIm-gpu.cpp)
  std::vector<double> pot, b;
  Csr A;
  // GPU variant HH metadata (type, indMat, indVec_DL)
  std::vector<int> hh_type, hh_mat, hh_dl;
  int last_iters = 0;
  double last_res = 0, last_bnorm = 0;
  int64_t total_iters = 0;
};

// GPU variant row types (solverIm-gpu.cpp:31-37).
enum { T_NORMAL = 1, T_BRANCH = 2, T_TERM = 4, T_NEXT_TERM = 8, T_FIRST_NORMAL = 16,
       T_FIRST_BRANCH = 32, T_OTHER = 64 };

inline double rdx(const Prob& q, int e) { return q.r_a[e] * q.dx[e]; }

// solverIm-gpu.cpp:199-369 (prevBoundaries, coreMatrix, postBoundaries).
void gpu_assemble(Implicit& s) {
  Prob& q = s.q;
  int N = q.N;
  Csr& A = s.A;
  A.rowptr.assign(N + 1, 0);
  A.col.clear();
  A.val.clear();
  std::vector<int> hh_of(N, -1);
  for (size_t i = 0; i < q.hh_elem.size(); i++) hh_of[q.hh_elem[i]] = (int)i;
  s.hh_type.assign(q.hh_elem.size(), 0);
  s.hh_mat.assign(q.hh_elem.size(), 0);
  s.hh_dl.assign(q.hh_elem.size(), -1);
  auto W = [&](int e) { return hh_of[e] >= 0 ? q.hhW[hh_of[e]] : q.W[e]; };
  auto tag = [&](int e, int type, int mat, int dlv) {
    if (hh_of[e] >= 0) {
      s.hh_type[hh_of[e]] = type;
      s.hh_mat[hh_of[e]] = mat;
      s.hh_dl[hh_of[e]] = dlv;
    }
  };
  auto push = [&](double v, int c) {
    A.val.push_back(v);
    A.col.push_back(c);
  };
  // row 0
  tag(0, T_OTHER, 0, -1);
  A.rowptr[0] = 0;
  push(1, 0);
  // row 1
  {
    int e = 1;
    double alpha = -1 / (q.r_a[e] * q.dx[e]);
    double betaR = 1 / (q.r_a[q.dr[e]] * q.dx[q.dr[e]]);
    double betaL = q.branching[e] / (q.r_a[q.dl[e]] * q.dx[q.dl[e]]);
    double betaMe = -W(e) * q.dx[e];
    double beta = (q.c_m[e] * q.dx[e]) / q.dT + betaMe - alpha + betaL + betaR;
    double delta = -betaR;
    A.rowptr[1] = 1;
    push(alpha + beta, 1);
    push(delta, 2);
    if (q.dl[1] != 0) {
      push(-betaL, q.dl[1]);
      // solverIm-gpu.cpp:240 stores indVec_DL = 1 here, so Update_Mat would take
      // betaL from element 1 instead of its left daughter; the CPU solver
      // (solverIm-cpu.cpp:231-262) uses the daughter, which this follows.
      tag(1, T_FIRST_BRANCH, 1, q.dl[1]);
    } else {
      tag(1, T_FIRST_NORMAL, 1, -1);
    }
  }
  // rows 2 .. N-3
  for (int i = 2; i < N - 2; i++) {
    int nnz = (int)A.val.size();
    A.rowptr[i] = nnz;
    if (q.dr[i] == i) {
      push(1, i);
      tag(i, T_TERM, nnz, -1);
    } else {
      double alpha = -1 / (q.r_a[i] * q.dx[i]);
      double betaR = 1 / (q.r_a[q.dr[i]] * q.dx[q.dr[i]]);
      double betaL = q.branching[i] / (q.r_a[q.dl[i]] * q.dx[q.dl[i]]);
      double betaMe = -W(i) * q.dx[i];
      double beta = (q.c_m[i] * q.dx[i]) / q.dT + betaMe - alpha + betaL + betaR;
      double delta = -betaR;
      push(alpha, q.mother[i]);
      if (q.dr[i + 1] == i + 1) {
        push(beta + delta, i);
        tag(i, T_NEXT_TERM, nnz, -1);
      } else {
        push(beta, i);
        push(delta, q.dr[i]);
        if (q.dl[i] != 0) {
          push(-betaL, q.dl[i]);
          tag(i, T_BRANCH, nnz, q.dl[i]);
        } else {
          tag(i, T_NORMAL, nnz, -1);
        }
      }
    }
  }
  // row N-2
  {
    int e = N - 2;
    double alpha = -1 / (q.r_a[e] * q.dx[e]);
    double betaR = 1 / (q.r_a[q.dr[e]] * q.dx[q.dr[e]]);
    double betaMe = -W(e) * q.dx[e];
    double beta = (q.c_m[e] * q.dx[e]) / q.dT + betaMe - alpha + betaR;
    double delta = -betaR;
    int nnz = (int)A.val.size();
    tag(e, T_NEXT_TERM, nnz, -1);
    A.rowptr[e] = nnz;
    push(alpha, q.mother[e]);
    push(beta + delta, e);
  }
  // row N-1
  {
    int nnz = (int)A.val.size();
    tag(N - 1, T_OTHER, nnz, -1);
    A.rowptr[N - 1] = nnz;
    push(1, N - 1);
    A.rowptr[N] = (int)A.val.size();
  }
}

// solverIm-gpu.cpp:103-156 (Update_Mat) for HH element i.
void gpu_update_mat(Implicit& s, int i) {
  Prob& q = s.q;
  int iv = q.hh_elem[i];
  int im = s.hh_mat[i];
  double* v = s.A.val.data();
  double alpha = -1 / (q.r_a[iv] * q.dx[iv]);
  double betaR = 1 / (q.r_a[iv + 1] * q.dx[iv + 1]);
  double betaMe = -(q.hhW[i] * q.dx[iv]) + 1 / (q.r_a[iv] * q.dx[iv]);
  double beta = (q.c_m[iv] * q.dx[iv]) / q.dT + betaMe + betaR;
  double delta = -betaR;
  double betaL;
  switch (s.hh_type[i]) {
    case T_NEXT_TERM:
      v[im] = alpha;
      v[im + 1] = beta + delta;
      break;
    case T_BRANCH:
      betaL = 1 / (q.r_a[s.hh_dl[i]] * q.dx[s.hh_dl[i]]);
      v[im] = alpha;
      v[im + 1] = beta + betaL;
      v[im + 2] = delta;
      v[im + 3] = -betaL;
      break;
    case T_NORMAL:
      v[im] = alpha;
      v[im + 1] = beta;
      v[im + 2] = delta;
      break;
    case T_FIRST_NORMAL:
      v[im] = alpha + beta;
      v[im + 1] = delta;
      break;
    case T_FIRST_BRANCH:
      betaL = 1 / (q.r_a[s.hh_dl[i]] * q.dx[s.hh_dl[i]]);
      v[im] = alpha + beta + betaL;
      v[im + 1] = delta;
      v[im + 2] = -betaL;
      break;
    default:
      break;
  }
}

// solverIm-cpu.cpp:29-183 (createMatrix): full CSR + RHS into cur.
void cpu_create_matrix(Implicit& s) {
  Prob& q = s.q;
  int N = q.N;
  Csr& A = s.A;
  const double* in = s.active ? q.input.data() : nullptr;
  auto inp = [&](int j) { return in ? in[j] : 0.0; };
  std::vector<int> hh_of(N, -1);
  for (size_t i = 0; i < q.hh_elem.size(); i++) hh_of[q.hh_elem[i]] = (int)i;
  auto W = [&](int e) { return hh_of[e] >= 0 ? q.hhW[hh_of[e]] : q.W[e]; };
  auto K = [&](int e) { return hh_of[e] >= 0 ? q.K[e] : q.K[e]; };
  A.rowptr.assign(N + 1, 0);
  A.col.clear();
  A.val.clear();
  auto push = [&](double v, int c) {
    A.val.push_back(v);
    A.col.push_back(c);
  };
  auto coeffs = [&](int e, double& alpha, double& beta, double& gamma, double& delta) {
    double br = q.branching[e];
    int DL = q.dl[e], DR = q.dr[e];
    gamma = -br / (q.r_a[DL] * q.dx[DL]);
    beta = (q.c_m[e] * q.dx[e]) / q.dT - W(e) * q.dx[e] + br / (q.r_a[DL] * q.dx[DL]) +
           1 / (q.r_a[DR] * q.dx[DR]) + 1 / (q.r_a[e] * q.dx[e]);
    alpha = -1 / (q.r_a[e] * q.dx[e]);
    delta = -1 / (q.r_a[DR] * q.dx[DR]);
  };
  auto rhs = [&](int j) {
    return q.c_m[j] * q.dx[j] * s.prev[j] / q.dT + K(j) * q.dx[j] + inp(j);
  };
  double alpha, beta, gamma, delta;
  push(1, 0);
  A.rowptr[0] = 0;
  // row 1
  coeffs(1, alpha, beta, gamma, delta);
  beta = beta + alpha;
  A.rowptr[1] = (int)A.val.size();
  push(beta, 1);
  if (q.dl[1] == 0) {
    push(delta, q.dr[1]);
  } else if (q.dr[1] < q.dl[1]) {
    push(delta, q.dr[1]);
    push(gamma, q.dl[1]);
  } else {
    push(gamma, q.dl[1]);
    push(delta, q.dr[1]);
  }
  s.cur[1] = rhs(1);
  int jedit = 2;
  bool salto = false;
  int j;
  for (j = 2; j < N - 2; j++) {
    if (salto) {
      salto = false;
      A.rowptr[jedit] = (int)A.val.size();
      push(1, j);
      jedit++;
    } else {
      coeffs(j, alpha, beta, gamma, delta);
      A.rowptr[jedit] = (int)A.val.size();
      push(alpha, q.mother[j]);
      if (q.dr[j + 1] == j + 1) {
        salto = true;
        push(beta + delta, j);
      } else {
        push(beta, j);
        push(delta, q.dr[j]);
        if (q.dl[j] != 0) push(gamma, q.dl[j]);
      }
      s.cur[j] = rhs(j);
      jedit++;
    }
  }
  // row N-2 (the loop variable j == N-2 here)
  int e = N - 2;
  coeffs(e, alpha, beta, gamma, delta);
  beta = beta + delta;
  A.rowptr[jedit] = (int)A.val.size();
  jedit++;
  push(alpha, q.mother[e]);
  push(beta, j);
  if (q.dl[e] != 0) push(gamma, q.dl[e]);
  s.cur[e] = rhs(e);
  A.rowptr[jedit] = (int)A.val.size();
  jedit++;
  push(1, N - 1);
  A.rowptr[jedit] = (int)A.val.size();
}

void spmv(const Csr& A, const double* x, double* y, int N) {
  for (int i = 0; i < N; i++) {
    double acc = 0;
    for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; k++) acc += A.val[k] * x[A.col[k]];
    y[i] = acc;
  }
}

double dot(const double* a, const double* b, int N) {
  double acc = 0;
  for (int i = 0; i < N; i++) acc += a[i] * b[i];
  return acc;
}

// Exact elimination of a tree-structured system (parent index < child index,
// one sub-diagonal entry per row). Identity rows are leaves of their own.
// Returns 0 on success, -1 when the sparsity is not a forest of that form.
int tree_solve(int N, const int* rowptr, const int* col, const double* val,
               const double* rhs, double* x) {
  std::vector<double> d(N), r(rhs, rhs + N);
  std::vector<int> parent(N, -1);
  std::vector<double> up(N, 0.0), down(N, 0.0);  // A[c][p], A[p][c]
  for (int i = 0; i < N; i++) {
    d[i] = 0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      int c = col[k];
      if (c == i) d[i] += val[k];
      else if (c < i) {
        if (parent[i] >= 0 && parent[i] != c) return -1;
        parent[i] = c;
        up[i] += val[k];
      }
    }
  }
  for (int i = 0; i < N; i++) {
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      int c = col[k];
      if (c > i) {
        if (parent[c] != i) return -1;  // coupling must mirror the child's row
        down[c] += val[k];
      }
    }
  }
  for (int c = N - 1; c >= 0; c--) {
    int p = parent[c];
    if (p < 0) continue;
    double f = down[c] / d[c];
    d[p] -= f * up[c];
    r[p] -= f * r[c];
  }
  for (int c = 0; c < N; c++) {
    int p = parent[c];
    double acc = r[c];
    if (p >= 0) acc -= up[c] * x[p];
    x[c] = acc / d[c];
  }
  return 0;
}

// Restarted GMRES, modified Gram-Schmidt, Givens rotations; x0 = 0; stops when
// the estimated residual norm <= tol (abs) after an Arnoldi step; at most
// max_iter iterations (solverIm-cpu.cpp:192-272: restart min(150,N), ipar[4]=N,
// user test dpar[4] <= 1e-14). Result overwrites rhs (ipar[12]=1).
int gmres(const Csr& A, int N, double* rhs, int restart, double tol, int max_iter,
          int* iters_out, double* res_out) {
  int m = std::min(restart, N);
  std::vector<double> x(N, 0.0), r(N), w(N);
  std::vector<double> V((size_t)(m + 1) * N), H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  int it = 0;
  double beta = 0;
  int status = NRT_ERR_NOT_CONVERGED;
  while (true) {
    spmv(A, x.data(), r.data(), N);
    for (int i = 0; i < N; i++) r[i] = rhs[i] - r[i];
    beta = std::sqrt(dot(r.data(), r.data(), N));
    if (beta <= tol) {
      status = NRT_OK;
      break;
    }
    if (it >= max_iter) break;
    for (int i = 0; i < N; i++) V[i] = r[i] / beta;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k;
    bool done = false;
    for (k = 0; k < m && it < max_iter; k++) {
      it++;
      spmv(A, &V[(size_t)k * N], w.data(), N);
      for (int j = 0; j <= k; j++) {
        double hjk = dot(w.data(), &V[(size_t)j * N], N);
        H[(size_t)j * m + k] = hjk;
        for (int i = 0; i < N; i++) w[i] -= hjk * V[(size_t)j * N + i];
      }
      double hk1 = std::sqrt(dot(w.data(), w.data(), N));
      H[(size_t)(k + 1) * m + k] = hk1;
      if (hk1 != 0)
        for (int i = 0; i < N; i++) V[(size_t)(k + 1) * N + i] = w[i] / hk1;
      for (int j = 0; j < k; j++) {
        double a = H[(size_t)j * m + k], b = H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = cs[j] * a + sn[j] * b;
        H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * b;
      }
      double a = H[(size_t)k * m + k], b = H[(size_t)(k + 1) * m + k];
      double den = std::sqrt(a * a + b * b);
      cs[k] = den == 0 ? 1 : a / den;
      sn[k] = den == 0 ? 0 : b / den;
      H[(size_t)k * m + k] = den;
      H[(size_t)(k + 1) * m + k] = 0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      if (std::fabs(g[k + 1]) <= tol || hk1 == 0) {
        k++;
        done = true;
        break;
      }
    }
    for (int j = k - 1; j >= 0; j--) {
      double acc = g[j];
      for (int l = j + 1; l < k; l++) acc -= H[(size_t)j * m + l] * y[l];
      y[j] = acc / H[(size_t)j * m + j];
    }
    for (int j = 0; j < k; j++)
      for (int i = 0; i < N; i++) x[i] += y[j] * V[(size_t)j * N + i];
    if (done) {
      spmv(A, x.data(), r.data(), N);
      for (int i = 0; i < N; i++) r[i] = rhs[i] - r[i];
      beta = std::sqrt(dot(r.data(), r.data(), N));
      status = NRT_OK;
      break;
    }
  }
  for (int i = 0; i < N; i++) rhs[i] = x[i];
  *iters_out = it;
  *res_out = beta;
  return status;
}

// solverIm-gpu.cpp:725-843 restated on the CPU (same vector recurrences and
// scalar formulas; residual tested every check_every iterations).
int bicgstab(const Csr& A, int N, const double* b, double* x, double tol, int tol_mode,
             int check_every, int max_iter, int* iters_out, double* res_out, double* bnorm_out) {
  std::vector<double> R0(b, b + N), AP(N), R(N), P(N), AS(N);
  double bnorm = std::sqrt(dot(b, b, N));
  *bnorm_out = bnorm;
  double thr = tol_mode == NRT_TOL_RELATIVE ? tol * bnorm : tol;
  double limit = thr * thr;
  spmv(A, x, AP.data(), N);
  for (int i = 0; i < N; i++) R0[i] = R0[i] + (-1.0) * AP[i];
  R = R0;
  P = R0;
  double p0 = dot(R0.data(), R0.data(), N), p1 = 0;
  double result = p0;
  int it = 0;
  int status = NRT_ERR_NOT_CONVERGED;
  if (result <= limit) {
    status = NRT_OK;
  } else {
    while (it < max_iter) {
      for (int j = 0; j < check_every; j++) {
        spmv(A, P.data(), AP.data(), N);
        double apr = dot(AP.data(), R0.data(), N);
        double alpha = p0 / apr, nalpha = -(p0 / apr);
        for (int i = 0; i < N; i++) R[i] = R[i] + nalpha * AP[i];
        spmv(A, R.data(), AS.data(), N);
        double w1 = dot(AS.data(), R.data(), N);
        double w2 = dot(AS.data(), AS.data(), N);
        double w = w1 / w2, nw = -(w1 / w2);
        for (int i = 0; i < N; i++) x[i] = x[i] + alpha * P[i];
        for (int i = 0; i < N; i++) x[i] = x[i] + w * R[i];
        for (int i = 0; i < N; i++) R[i] = R[i] + nw * AS[i];
        p1 = dot(R.data(), R0.data(), N);
        double beta = (alpha / w) * (p1 / p0);
        AS = P;
        for (int i = 0; i < N; i++) AS[i] = AS[i] + nw * AP[i];
        P = R;
        for (int i = 0; i < N; i++) P[i] = P[i] + beta * AS[i];
        std::swap(p0, p1);
        it++;
      }
      result = dot(R.data(), R.data(), N);
      if (!std::isfinite(result)) {
        status = NRT_ERR_NUMERIC;
        break;
      }
      if (result <= limit) {
        status = NRT_OK;
        break;
      }
    }
  }
  *iters_out = it;
  *res_out = std::sqrt(result);
  return status;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ setup chain
// init.cpp:320-333 (dX*(1+eps)), :235-254 (electrical_properties),
// :260-267 (setting_rest_pot), :273-298 (hhIconds), :180-198 (cable_init_constants),
// :138-174 (critical time step). Fills the same struct the device setup fills.
int orc_setup(const nrt_geometry* g, nrt_setup_out* o) {
  int N = g->num_elem;
  int H = 0;
  for (int k = 0; k < N; k++)
    if (g->materials[g->material[k]].kind == 1) H++;
  o->num_hh = H;
  std::vector<double> d(N), eps(N), eps_s(N);
  for (int k = 0; k < N; k++) {
    eps[k] = g->epsilon[k];
    eps_s[k] = sqrt(1 + eps[k]) - 1;                       // discretization.h:50-53
    o->dx[k] = g->dx0[k] * (1 + eps[k]);                   // init.cpp:330
    d[k] = g->diameter[k];
  }
  int hi = 0;
  for (int k = 0; k < N; k++) {
    const nrt_material& mt = g->materials[g->material[k]];
    if (mt.kind != 1) continue;
    d[k] = d[k] / (sqrt(1 + eps[k]));                       // init.cpp:239
    o->r_a[k] = 4 * mt.rho_a / (kPi * d[k] * d[k]);        // discretization.h:61-65
    o->r_m[k] = (mt.rho_m * mt.thickness) / (kPi * d[k]);
    o->c_m[k] = (mt.per_m * kPi * d[k]) / (mt.thickness);
    double rest = mt.rest_pot;
    double ENa0 = mt.dE_Na + rest, EK0 = mt.dE_K + rest;  // discretization.cpp:79-82
    double ENa, EK, lsNa, lsK;
    double EMAX = g->E_MAX;
    if (eps_s[k] <= EMAX) {                                 // discretization.h:216-227
      ENa = ENa0 - pow((eps_s[k] / EMAX), mt.n_damage) * (ENa0);
      EK = EK0 - pow((eps_s[k] / EMAX), mt.n_damage) * (EK0);
    } else {
      ENa = 0;
      EK = 0;
    }
    double gL = mt.sigma_L * kPi * d[k] / mt.thickness;    // discretization.h:208-213
    double gNa = mt.sigma_Na * kPi * d[k] / mt.thickness;
    double gK = mt.sigma_K * kPi * d[k] / mt.thickness;
    if (eps_s[k] <= EMAX) {                                 // discretization.h:230-238
      lsNa = pow(eps_s[k] / EMAX, mt.n_damage) * (ENa0);
      lsK = pow(eps_s[k] / EMAX, mt.n_damage) * (EK0);
    } else {
      lsNa = ENa0;
      lsK = EK0;
    }
    gNa = gNa / (sqrt(1 + eps[k]));                         // init.cpp:250-251
    gK = gK / (sqrt(1 + eps[k]));
    // hhIconds (init.cpp:280-297)
    Rates r = hh_rates(rest, rest, lsNa, lsK);
    double m = r.am / (r.am + r.bm);
    double h = r.ah / (r.ah + r.bh);
    double n = r.an / (r.an + r.bn);
    double vlost = (gNa * m * m * m * h * (rest - ENa) + gK * n * n * n * n * (rest - EK) +
                    gL * rest) / gL;
    o->hh_elem[hi] = k;
    o->hh_m[hi] = m;
    o->hh_h[hi] = h;
    o->hh_n[hi] = n;
    o->E_Na[hi] = ENa;
    o->E_K[hi] = EK;
    o->E_L[hi] = vlost;
    o->g_Na[hi] = gNa;
    o->g_K[hi] = gK;
    o->G_L[hi] = gL;
    o->ls_Na[hi] = lsNa;
    o->ls_K[hi] = lsK;
    o->hh_rest[hi] = rest;
    o->W[k] = 0;
    o->K[k] = 0;
    // critical_time_step_class (init.cpp:141)
    o->crit_dT[k] = (2 * o->r_a[k] * o->c_m[k] * o->dx[k] * o->dx[k]) / 4;
    hi++;
  }
  for (int k = 0; k < N; k++) {
    const nrt_material& mt = g->materials[g->material[k]];
    o->rest[k] = mt.rest_pot;
    if (mt.kind == 1) continue;
    d[k] = d[k] / (sqrt(1 + eps[k]));                       // init.cpp:184
    double ra = 4 * mt.rho_a / (kPi * d[k] * d[k]);
    double rm = (mt.rho_m * mt.thickness) / (kPi * d[k]);
    double cm = (mt.per_m * kPi * d[k]) / (mt.thickness);
    if (g->myelinate) {                                      // discretization.h:126-151
      if (mt.layers_my != 0) {
        double temp = 0;
        for (int i = 1; i <= mt.layers_my; i++)
          temp = temp + 1 / (d[k] + mt.thickness * 2 + 2 * (i - 1) * mt.thickness_my);
        double c_my_T = mt.thickness_my * temp / (mt.per_my * kPi);
        double c_my = 1 / c_my_T;
        double r_my = (mt.ro_my * mt.thickness_my * temp) / kPi;
        rm = rm + r_my;
        cm = (cm * c_my) / (c_my + cm);
      }
    }
    o->r_a[k] = ra;
    o->r_m[k] = rm;
    o->c_m[k] = cm;
    o->W[k] = -1 / rm;                                       // discretization.cpp:63-68
    o->K[k] = mt.rest_pot / rm;
    o->crit_dT[k] = (2 * ra * cm * o->dx[k] * o->dx[k]) / (4 + (ra * o->dx[k] * o->dx[k]) / rm);
  }
  // critical_time_step_selection (init.cpp:165-174)
  double t = N > 1 ? o->crit_dT[1] : 0;
  for (int i = 2; i < N - 1; i++)
    if (t > o->crit_dT[i]) t = o->crit_dT[i];
  o->dT_crit = t;
  return 0;
}

// The stimulus window of neurite.cpp:227-230 as an inclusive step range.
void orc_active_range(double dT, int64_t num_steps_T, double init_time, double end_input,
                      int64_t* first, int64_t* last) {
  *first = -1;
  *last = -2;
  for (int64_t i = 1; i < num_steps_T - 1; i++) {
    double t = i * dT;
    if ((t >= init_time) && (t <= end_input)) {
      if (*first < 0) *first = i;
      *last = i;
    }
  }
}

// -------------------------------------------------------------- explicit API
void* orc_ex_create(const nrt_problem* p) {
  Explicit* s = new Explicit;
  s->q = load(p);
  s->cur = s->q.V1;   // solverEx-cpu.cpp:24-27
  s->prev = s->q.V0;
  return s;
}

void orc_ex_update(void* h, int active) {
  Explicit* s = (Explicit*)h;
  Prob& q = s->q;
  s->active = active != 0;
  const double dt = q.dT * 1000;
  for (size_t i = 0; i < q.hh_elem.size(); i++) {
    int e = q.hh_elem[i];
    double W, K;
    hh_step(q, (int)i, s->prev[e], dt, &W, &K);
    q.W[e] = W;
    q.K[e] = K;
  }
}

void orc_ex_calculate(void* h) {
  Explicit* s = (Explicit*)h;
  Prob& q = s->q;
  const double* in = s->active ? q.input.data() : nullptr;
  const double* P = s->prev.data();
  double* C = s->cur.data();
  int N = q.N;
  for (int j = 1; j < N - 1; j++) {                  // solverEx-cpu.cpp:92-108
    double V = P[j];
    int DL = q.dl[j], DR = q.dr[j], MO = q.mother[j];
    double inp = in ? in[j] : 0.0;
    C[j] = V + (q.dT / (q.c_m[j] * q.dx[j])) *
                   ((double)q.branching[j] * ((P[DL] - V) / (q.r_a[DL] * q.dx[DL])) +
                    (P[DR] - V) / (q.r_a[DR] * q.dx[DR]) +
                    (P[MO] - V) / (q.r_a[j] * q.dx[j]) + q.W[j] * q.dx[j] * V +
                    q.K[j] * q.dx[j] + inp);
  }
  C[q.terminals[0]] = C[q.terminals[0] + 1];        // solverEx-cpu.cpp:112-116
  for (size_t k = 1; k < q.terminals.size(); k++) C[q.terminals[k]] = C[q.mother[q.terminals[k]]];
  std::swap(s->cur, s->prev);
}

void orc_ex_snapshot(void* h, double* out) {
  Explicit* s = (Explicit*)h;
  memcpy(out, s->prev.data(), sizeof(double) * s->q.N);
}

void orc_ex_gates(void* h, double* m, double* hh, double* n) {
  Explicit* s = (Explicit*)h;
  size_t H = s->q.hh_elem.size();
  for (size_t i = 0; i < H; i++) {
    m[i] = s->q.m[i];
    hh[i] = s->q.h[i];
    n[i] = s->q.n[i];
  }
}

void orc_ex_destroy(void* h) { delete (Explicit*)h; }

// -------------------------------------------------------------- implicit API
// variant: 0 = CPU createMatrix (+ GMRES/direct), 1 = GPU assembly (+ BiCGSTAB/direct)
// method: 0 = GMRES(restart), 1 = BiCGSTAB, 2 = direct tree solve
void* orc_im_create(const nrt_problem* p, int variant, int method, int restart, double tol,
                    int tol_mode, int x0_mode, int check_every, int max_iter) {
  Implicit* s = new Implicit;
  s->q = load(p);
  s->variant = variant;
  s->method = method;
  s->restart = restart > 0 ? restart : 150;
  s->tol = tol;
  s->tol_mode = tol_mode;
  s->x0_mode = x0_mode;
  s->check_every = check_every > 0 ? check_every : 30;
  s->max_iter = max_iter > 0 ? max_iter : s->q.N;
  Prob& q = s->q;
  for (size_t i = 0; i < q.hh_elem.size(); i++) q.hhW[i] = q.W[q.hh_elem[i]];
  if (variant == VAR_CPU) {
    s->cur = q.V1;     // solverIm-cpu.cpp:463-466
    s->prev = q.V0;
  } else {
    s->pot = q.V0;     // solverIm-gpu.cpp:523-524
    s->b.assign(q.N, 0.0);
    gpu_assemble(*s);
  }
  return s;
}

int orc_im_update(void* h, int active) {
  Implicit* s = (Implicit*)h;
  Prob& q = s->q;
  s->active = active != 0;
  const double dt = q.dT * 1000;
  if (s->variant == VAR_CPU) {
    for (size_t i = 0; i < q.hh_elem.size(); i++) {
      int e = q.hh_elem[i];
      double W, K;
      hh_step(q, (int)i, s->prev[e], dt, &W, &K);
      q.hhW[i] = W;
      q.K[e] = K;
    }
    cpu_create_matrix(*s);
  } else {
    for (size_t i = 0; i < q.hh_elem.size(); i++) {   // Update_HH (solverIm-gpu.cpp:61-98)
      int e = q.hh_elem[i];
      double W, K;
      hh_step(q, (int)i, s->pot[e], dt, &W, &K);
      q.hhW[i] = W;
      q.K[e] = K;
    }
    for (size_t i = 0; i < q.hh_elem.size(); i++) gpu_update_mat(*s, (int)i);
    const double* in = s->active ? q.input.data() : nullptr;
    for (int i = 0; i < q.N; i++) {                   // Update_B (solverIm-gpu.cpp:162-171)
      double inp = in ? in[i] : 0.0;
      s->b[i] = q.c_m[i] * q.dx[i] * s->pot[i] / q.dT + q.K[i] * q.dx[i] + inp;
    }
  }
  return 0;
}

int orc_im_calculate(void* h) {
  Implicit* s = (Implicit*)h;
  Prob& q = s->q;
  int N = q.N;
  int st = NRT_OK;
  if (s->variant == VAR_CPU) {
    if (s->method == M_DIRECT) {
      std::vector<double> x(N);
      if (tree_solve(N, s->A.rowptr.data(), s->A.col.data(), s->A.val.data(), s->cur.data(),
                     x.data()) != 0)
        return NRT_ERR_TOPOLOGY;
      s->cur = x;
      s->last_iters = 0;
    } else {
      st = gmres(s->A, N, s->cur.data(), s->restart, s->tol, s->max_iter, &s->last_iters,
                 &s->last_res);
    }
    std::swap(s->cur, s->prev);
  } else {
    std::vector<double> x(N);
    if (s->method == M_DIRECT) {
      if (tree_solve(N, s->A.rowptr.data(), s->A.col.data(), s->A.val.data(), s->b.data(),
                     x.data()) != 0)
        return NRT_ERR_TOPOLOGY;
      s->last_iters = 0;
    } else {
      x = (s->x0_mode == NRT_X0_RHS) ? s->b : s->pot;
      st = bicgstab(s->A, N, s->b.data(), x.data(), s->tol, s->tol_mode, s->check_every,
                    s->max_iter, &s->last_iters, &s->last_res, &s->last_bnorm);
    }
    s->pot = x;
  }
  s->total_iters += s->last_iters;
  return st;
}

void orc_im_snapshot(void* h, double* out) {
  Implicit* s = (Implicit*)h;
  const std::vector<double>& v = s->variant == VAR_CPU ? s->prev : s->pot;
  memcpy(out, v.data(), sizeof(double) * s->q.N);
}

// Copies the last assembled matrix (and right-hand side) out; returns nnz.
int orc_im_matrix(void* h, int* rowptr, int* col, double* val, double* rhs) {
  Implicit* s = (Implicit*)h;
  int N = s->q.N;
  int nnz = (int)s->A.val.size();
  if (rowptr) memcpy(rowptr, s->A.rowptr.data(), sizeof(int) * (N + 1));
  if (col) memcpy(col, s->A.col.data(), sizeof(int) * nnz);
  if (val) memcpy(val, s->A.val.data(), sizeof(double) * nnz);
  if (rhs) {
    const std::vector<double>& b = s->variant == VAR_CPU ? s->cur : s->b;
    memcpy(rhs, b.data(), sizeof(double) * N);
  }
  return nnz;
}

void orc_im_stats(void* h, int* last_iters, double* last_res, int64_t* total_iters) {
  Implicit* s = (Implicit*)h;
  *last_iters = s->last_iters;
  *last_res = s->last_res;
  *total_iters = s->total_iters;
}

void orc_im_destroy(void* h) { delete (Implicit*)h; }

// Exact solve of an assembled tree system (see tree_solve).
int orc_tree_solve(int N, const int* rowptr, const int* col, const double* val,
                   const double* rhs, double* x) {
  return tree_solve(N, rowptr, col, val, rhs, x);
}

// ------------------------------------------------------------------ time loop
// neurite.cpp:224-255: steps i = first .. first+num-1; the stimulus is on for
// a0 <= i <= a1; a sample of V[trace_elem] is taken when ++GL_cont > interval.
// kind: 0 explicit, 1 implicit. Returns the first non-OK status.
int orc_run(void* h, int kind, int64_t first, int64_t num, int64_t a0, int64_t a1,
            int trace_elem, int64_t interval, double* trace, int64_t* ns) {
  int64_t gl = 0, k = 0;
  int N = kind == 0 ? ((Explicit*)h)->q.N : ((Implicit*)h)->q.N;
  std::vector<double> snap(N);
  int status = NRT_OK;
  for (int64_t i = first; i < first + num; i++) {
    int active = (i >= a0 && i <= a1) ? 1 : 0;
    if (kind == 0) {
      orc_ex_update(h, active);
      orc_ex_calculate(h);
    } else {
      orc_im_update(h, active);
      int st = orc_im_calculate(h);
      if (st != NRT_OK && status == NRT_OK) status = st;
    }
    if (interval >= 0 && ++gl > interval) {
      gl = 0;
      if (trace) {
        if (kind == 0) orc_ex_snapshot(h, snap.data());
        else orc_im_snapshot(h, snap.data());
        trace[k] = snap[trace_elem];
      }
      k++;
    }
  }
  if (ns) *ns = k;
  return status;
}

}  // extern "C"
