// Drop-in replacement of SolverExGPU / SolverImGPU (see solverHip.h).
//
// The constructors read the discretization objects through their public getters,
// exactly the arrays the reference GPU constructors copy
// (solverEx-gpu.cpp:327-399, solverIm-gpu.cpp:527-595) plus the topology the CPU
// solver walks (solverEx-cpu.cpp:92-116), and hand them to nrt_create. Errors
// surface as nrt::hipSolverError (the cudaException replacement).
#include <vector>
#include "solverHip.h"

namespace {

typedef char type_must_be_double[sizeof(TYPE) == sizeof(double) ? 1 : -1];

struct Arrays {
  std::vector<double> dx, c_m, r_a, W, K, input;
  std::vector<int32_t> mother, dr, dl, terminals, hh_elem;
  std::vector<uint8_t> branching;
  std::vector<double> m, h, n, E_Na, E_K, E_L, g_Na, g_K, G_L, ls_Na, ls_K, rest;
};

nrt_problem extract(Arrays& a, TYPE dT, TYPE** potential, int num_term, int* ind_terminals,
                    int num_HH, Hodgkin_Huxley** HH, int N, discretization** e) {
  a.dx.resize(N); a.c_m.resize(N); a.r_a.resize(N); a.W.resize(N); a.K.resize(N);
  a.input.resize(N); a.mother.resize(N); a.dr.resize(N); a.dl.resize(N); a.branching.resize(N);
  for (int i = 0; i < N; i++) {
    a.dx[i] = e[i]->get_dX();
    a.c_m[i] = e[i]->get_c_m();
    a.r_a[i] = e[i]->get_r_a();
    a.W[i] = e[i]->get_W();
    a.K[i] = e[i]->get_K();
    a.input[i] = e[i]->get_input_current();
    a.mother[i] = e[i]->get_mother();
    a.dr[i] = e[i]->get_DR();
    a.dl[i] = e[i]->get_DL();
    a.branching[i] = e[i]->get_branching() ? 1 : 0;
  }
  a.terminals.assign(ind_terminals, ind_terminals + num_term);
  a.hh_elem.resize(num_HH); a.m.resize(num_HH); a.h.resize(num_HH); a.n.resize(num_HH);
  a.E_Na.resize(num_HH); a.E_K.resize(num_HH); a.E_L.resize(num_HH); a.g_Na.resize(num_HH);
  a.g_K.resize(num_HH); a.G_L.resize(num_HH); a.ls_Na.resize(num_HH); a.ls_K.resize(num_HH);
  a.rest.resize(num_HH);
  for (int i = 0; i < num_HH; i++) {
    a.hh_elem[i] = HH[i]->get_element();
    a.m[i] = HH[i]->get_m();
    a.h[i] = HH[i]->get_h();
    a.n[i] = HH[i]->get_n();
    a.E_Na[i] = HH[i]->get_E_Na();
    a.E_K[i] = HH[i]->get_E_K();
    a.E_L[i] = HH[i]->get_E_L();
    a.g_Na[i] = HH[i]->get_g_Na();
    a.g_K[i] = HH[i]->get_g_K();
    a.G_L[i] = HH[i]->get_G_L();
    a.ls_Na[i] = HH[i]->get_Left_Shift_Na();
    a.ls_K[i] = HH[i]->get_Left_Shift_K();
    a.rest[i] = HH[i]->get_rest_pot();
  }
  nrt_problem p;
  p.num_elem = N;
  p.dT = dT;
  p.potential0 = potential[0];
  p.potential1 = potential[1];
  p.dx = &a.dx[0]; p.c_m = &a.c_m[0]; p.r_a = &a.r_a[0]; p.W = &a.W[0]; p.K = &a.K[0];
  p.input = &a.input[0]; p.mother = &a.mother[0]; p.dr = &a.dr[0]; p.dl = &a.dl[0];
  p.branching = &a.branching[0];
  p.num_term = num_term;
  p.terminals = &a.terminals[0];
  p.num_hh = num_HH;
  bool h = num_HH > 0;
  p.hh_elem = h ? &a.hh_elem[0] : 0; p.hh_m = h ? &a.m[0] : 0; p.hh_h = h ? &a.h[0] : 0;
  p.hh_n = h ? &a.n[0] : 0; p.E_Na = h ? &a.E_Na[0] : 0; p.E_K = h ? &a.E_K[0] : 0;
  p.E_L = h ? &a.E_L[0] : 0; p.g_Na = h ? &a.g_Na[0] : 0; p.g_K = h ? &a.g_K[0] : 0;
  p.G_L = h ? &a.G_L[0] : 0; p.ls_Na = h ? &a.ls_Na[0] : 0; p.ls_K = h ? &a.ls_K[0] : 0;
  p.rest_pot = h ? &a.rest[0] : 0;
  return p;
}

nrt::HipSolverBase* make(int kind, TYPE dT, TYPE** potential, int num_term, int* term,
                         int num_HH, Hodgkin_Huxley** HH, int N, discretization** e) {
  Arrays a;
  nrt_problem p = extract(a, dT, potential, num_term, term, num_HH, HH, N, e);
  if (kind == NRT_EXPLICIT) return new nrt::SolverExHIP(p);
  return new nrt::SolverImHIP(p);
}

}  // namespace

SolverExHIP::SolverExHIP(TYPE dT, TYPE** potential, int num_term, int* ind_terminals,
                         TYPE input_val, int num_HH_elements, Hodgkin_Huxley** HH_elements,
                         int num_elem, discretization** element)
    : impl_(make(NRT_EXPLICIT, dT, potential, num_term, ind_terminals, num_HH_elements,
                 HH_elements, num_elem, element)) {
  (void)input_val;  // unused by the reference as well
}
void SolverExHIP::update(bool input_active) { impl_->update(input_active); }
void SolverExHIP::calculate() { impl_->calculate(); }
void SolverExHIP::snapshot(TYPE* potential) { impl_->snapshot(potential); }
SolverExHIP::~SolverExHIP() { delete impl_; }

SolverImHIP::SolverImHIP(TYPE dT, TYPE** potential, int num_term, int* ind_terminals,
                         TYPE input_val, int num_HH_elements, Hodgkin_Huxley** HH_elements,
                         int num_elem, discretization** element)
    : impl_(make(NRT_IMPLICIT, dT, potential, num_term, ind_terminals, num_HH_elements,
                 HH_elements, num_elem, element)) {
  (void)input_val;
}
void SolverImHIP::update(bool input_active) { impl_->update(input_active); }
void SolverImHIP::calculate() { impl_->calculate(); }
void SolverImHIP::snapshot(TYPE* potential) { impl_->snapshot(potential); }
SolverImHIP::~SolverImHIP() { delete impl_; }
