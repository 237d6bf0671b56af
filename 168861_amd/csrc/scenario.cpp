// Scenarios of the drop-in driver: the reference's command-line tokens, its
// geometry builders and file loaders, and the mechanical model that scales the
// measurement distance -- everything the reference main does before the solver
// exists (neurite.cpp:84-184). Host code only; no device work.
//
//   command line        input.cpp:25-149 (tokens), init.cpp:40-73 (e/i/g flags)
//   builder selection   init.cpp:84-130 (the `simulation` environment variable)
//   builders            space.cpp:37-798
//   mechanical model    mechanical_model.cpp:21-99, init.cpp:320-333
//   measurement point   init.cpp:339-358
//
// The builders produce the raw geometry (nrt_geometry) that nrt_setup turns into
// the solver's per-element state; their numbering, couplings and per-element
// parameters must be the reference's, so each follows its builder's steps (the
// quirks included: a cable's last element keeps its daughter index N, the
// symmetric tree's branch bookkeeping, the neuron loader's end-of-file handling).
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include "nrt_internal.hpp"

namespace nrt {
namespace {

// The reference's `input` struct (neurite.h:35-38) plus the globals the command
// line and the builders set (neurite.cpp:63-68), with main's defaults
// (neurite.cpp:111-122).
struct Params {
  double tspan = 0, xspan = 0, dist_mes = 0;
  double input_current = 0, input_time = 0, init_time = 0, end_input = 0;
  double epsilon = 0, nu = 0.00261813, dt_factor = 1, in_dist_mes = 0, E_MAX = 1;
  int IRE = 0, NRE = 0;
};

struct Scenario {
  std::string simulation, out_path, log_path;
  bool explicit_mode = false, implicit_mode = false, opengl = false;
  bool hh_model = false;
  Params in;
  int N = 0;
  std::vector<int32_t> mother, dr, dl, material, terminals;
  std::vector<uint8_t> branching;
  std::vector<double> diameter, dx0, epsilon, input;
  double epsilon_macro = 0;
  int mes_point = 0;
  std::string log;  // what the builders write to the log file
  // extension tokens (not in the reference; see nrt_scenario_info)
  int64_t max_steps = 0;
  int method = -1, precision = 0, restart = 0;
  double tol = 0;
  std::string vout_path;
  nrt_material mats[2];
  nrt_geometry geom{};
};

void logf(Scenario& s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void logf(Scenario& s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  s.log += buf;
}

// ------------------------------------------------------------- command line
// input.cpp:25-149 and init.cpp:40-55. Each token is tested against every
// pattern in turn, on its letters after an optional leading '-'; a pattern that
// takes a value consumes the next token when it starts with a digit (a file name
// always), and the later patterns of the same pass then test that consumed token
// (with the first token's '-' offset), as the reference's in-place index does.
// The first letter of each token is recorded in `args` (a consumed token's slot
// holds ' '); the solver flags are e, i and g anywhere in `args` up to its first
// NUL.
bool numeric(const char* a) { return a[0] > 47 && a[0] < 58; }
bool has(const char* t, const char* p) {
  for (; *p; ++p, ++t)
    if (*t != *p) return false;
  return true;
}

void parse_args(Scenario& s, int argc, const char* const* argv) {
  std::vector<char> args(argc, 0);
  for (int i = 1; i < argc; i++) {
    const int f = argv[i][0] == '-' ? 1 : 0;
    const int i0 = i;
    args[i - 1] = argv[i][f];
    auto t = [&]() { return argv[i] + f; };
    auto real = [&](double& v) {
      if (i != argc - 1 && numeric(argv[i + 1])) {
        v = std::atof(argv[i + 1]);
        args[i] = ' ';
        i++;
      }
    };
    auto integer = [&](int& v) {
      if (i != argc - 1 && numeric(argv[i + 1])) {
        v = std::atoi(argv[i + 1]);
        args[i] = ' ';
        i++;
      }
    };
    auto word = [&](std::string& v) {
      if (i != argc - 1) {
        v = argv[i + 1];
        args[i] = ' ';
        i++;
      }
    };
    if (has(t(), "xspan")) real(s.in.xspan);
    if (has(t(), "Tt")) real(s.in.tspan);
    if (has(t(), "nu")) real(s.in.nu);
    if (has(t(), "str")) real(s.in.epsilon);
    if (has(t(), "IRE")) integer(s.in.IRE);
    if (has(t(), "NRE")) integer(s.in.NRE);
    if (has(t(), "log")) word(s.log_path);
    if (has(t(), "dtf")) real(s.in.dt_factor);
    if (has(t(), "out")) word(s.out_path);
    if (has(t(), "mes")) real(s.in.in_dist_mes);
    if (has(t(), "MAXE")) real(s.in.E_MAX);
    // extensions of this build (tokens the reference does not have), tested only
    // on a token that no reference pattern consumed a value after
    if (i != i0) continue;
    if (has(t(), "steps")) {
      double v = 0;
      real(v);
      s.max_steps = (int64_t)v;
    }
    if (has(t(), "method")) {
      std::string w;
      word(w);
      s.method = w == "bicgstab" ? NRT_BICGSTAB : w == "gmres" ? NRT_GMRES
                 : w == "direct" ? NRT_DIRECT : -2;
      NRT_REQUIRE(s.method >= 0, NRT_ERR_INVALID, "method must be bicgstab, gmres or direct");
    }
    if (has(t(), "prec")) {
      std::string w;
      word(w);
      NRT_REQUIRE(w == "fp64" || w == "mixed", NRT_ERR_INVALID, "prec must be fp64 or mixed");
      s.precision = w == "mixed" ? NRT_FP32_REFINE : NRT_FP64;
    }
    if (has(t(), "restart")) integer(s.restart);
    if (has(t(), "tol")) real(s.tol);
    if (has(t(), "vout")) word(s.vout_path);
  }
  if (argc >= 1) args[argc - 1] = 0;
  for (char c : args) {
    if (c == 0) break;
    if (c == 'e') s.explicit_mode = true;
    if (c == 'i') s.implicit_mode = true;
    if (c == 'g') s.opengl = true;
  }
}

// ----------------------------------------------------------------- builders
// Element defaults of discretization.cpp:19-40: d = 3e-6 m, no strain, no left
// daughter, no branching, no stimulus.
struct Net {
  Scenario& s;
  void resize(int n) {
    s.N = n;
    s.mother.assign(n, 0);
    s.dr.assign(n, 0);
    s.dl.assign(n, 0);
    s.material.assign(n, 0);
    s.branching.assign(n, 0);
    s.diameter.assign(n, 3 * std::pow(10.0, -6));
    s.dx0.assign(n, 0);
    s.epsilon.assign(n, 0);
    s.input.assign(n, 0);
  }
  // Cable(k, dr, mom, dx) / Hodgkin_Huxley(k, dr, mom, dx) (discretization.cpp:49-97)
  void add(int k, int dr, int mom, double dx, bool hh, double eps) {
    s.dr[k] = dr;
    s.mother[k] = mom;
    s.dx0[k] = dx;
    s.material[k] = hh ? 1 : 0;
    s.epsilon[k] = eps;
  }
};

const double kMicro = 1e-6;

// Only_HH_Model_class (space.cpp:542-583): one HH element between two fictitious
// ones; the builder fixes the time span, the step factor and the stimulus.
void build_only_hh(Scenario& s) {
  Params& in = s.in;
  in.IRE = 1;
  in.NRE = 1;
  in.in_dist_mes = 4.2e-3;
  in.xspan = 6.3e-3;
  in.init_time = 0.003;
  in.input_time = 0.003;
  in.end_input = in.init_time + in.input_time;
  in.input_current = 0.8e-9;
  in.dt_factor = 0.0001;
  in.tspan = 0.025;
  Net n{s};
  n.resize(3);
  n.add(0, 1, -1, 2.1e-3, false, 0);
  n.add(1, 1, 0, 2.1e-3, true, 0);
  n.add(2, 2, 1, 2.1e-3, false, 0);
  s.terminals = {0, 2};
  s.input[1] = in.input_current;
}

// without_branching_CT_HH (space.cpp:646-798): nodes of Ranvier (NRE elements of
// 2.1 um / NRE) alternating with internodes (IRE elements of 0.8 mm / IRE); the
// strain is split between nodes and internodes by nu.
void build_hh_axon(Scenario& s) {
  Params& in = s.in;
  in.init_time = 0.04;
  in.input_time = 0.003;
  in.input_current = 0.04e-9;
  in.end_input = in.init_time + in.input_time;
  in.dist_mes = 0.010;
  const int cable_n = in.IRE < 1e-10 ? 100 : in.IRE;
  const int hh_n = in.NRE < 1e-10 ? 1 : in.NRE;
  if (in.xspan < 1e-10) in.xspan = 2;
  if (in.tspan < 1e-10) in.tspan = 0.040;
  const double node = 2.1e-6, inter = 0.0008;
  const int parts = int((in.xspan / (inter + node)) + 1);
  in.xspan = parts * (inter + node);
  const double hh_eps = (in.nu * in.xspan / (node * parts)) * in.epsilon;
  const double ct_eps = ((1 - in.nu) * in.xspan / (parts * inter)) * in.epsilon;
  const int N = parts * (cable_n + hh_n);
  const double cable_dx = (double)inter / cable_n;
  const double hh_dx = (double)node / hh_n;
  logf(s, "for discretization and simulation cabledx =%g  hhdx=%g steps cable %i steps hh=%i\n",
       cable_dx, hh_dx, cable_n, hh_n);
  logf(s, "distances cable=%.20f  hh=%.20f\n", cable_dx * cable_n, hh_dx * hh_n);
  NRT_REQUIRE(!(hh_dx > node), NRT_ERR_INVALID, "HH_axon: element larger than the node");
  NRT_REQUIRE(!(node > inter), NRT_ERR_INVALID, "HH_axon: node longer than the internode");
  // Walk the axon from element 1, alternating a node and an internode: an
  // element joins its section while the section's accumulated length plus half
  // an element stays within the section length. A node that fills up hands over
  // within the same pass (space.cpp:735-757).
  std::vector<uint8_t> is_hh(N, 0);
  int i = 1, channels = 0;
  bool nodal = true;
  double acc = 0;
  const double tol_hh = hh_dx / 2, tol_ct = cable_dx / 2;
  while (i < N) {
    if (nodal) {
      if (i < N) is_hh[i] = 1;
      channels++;
      acc = acc + hh_dx;
      i++;
      if (std::fabs(acc + tol_hh) > node) {
        nodal = false;
        acc = 0;
      }
    }
    if (!nodal) {
      i++;
      acc = acc + cable_dx;
      if (std::fabs(acc + tol_ct) > inter) {
        nodal = true;
        acc = 0;
      }
    }
  }
  logf(s, "Total numChannels %i. With nodal size=%f and %f as internodal distance.\n", channels,
       node, inter);
  Net n{s};
  n.resize(N);
  for (int k = 0; k < N; k++) {
    if (is_hh[k]) n.add(k, k + 1, k - 1, hh_dx, true, hh_eps);
    else n.add(k, k + 1, k - 1, cable_dx, false, ct_eps);
  }
  s.dr[N - 1] = N - 1;
  s.terminals = {0, N - 1};
  logf(s, "nu=%g\n", in.nu);
  logf(s, "nu of equilibrium:%g\n", node * channels / in.xspan);
  logf(s, "Setting the input current at the elements that you defined %g\n", in.input_current);
  s.input[1] = in.input_current;
}

// branching_Y_CT_HH (space.cpp:405-534) and branching_Y_CT (space.cpp:343-403):
// three parts of at least 50 elements; the first two run on as one line, the
// third starts at the end of the first (its left daughter) and the end of the
// second becomes a sealed end. With HH, each part is made of chunks of IRE cable
// elements followed by NRE HH elements; without, of single cable elements.
void build_y(Scenario& s, bool with_hh) {
  Params& in = s.in;
  in.input_time = 0.003;
  in.input_current = 0.8e-9;
  in.end_input = in.init_time + in.input_time;
  if (in.tspan < 1e-10) in.tspan = 0.025;
  int ct_n = 1, hh_n = 0;
  double ct_dx = 50e-6, hh_dx = 0;
  const double node = 2.1e-6, inter = 0.0008;
  if (with_hh) {
    ct_n = in.IRE < 1e-10 ? 10 : in.IRE;
    hh_n = in.NRE < 1e-10 ? 1 : in.NRE;
    ct_dx = (double)inter / ct_n;
    hh_dx = (double)node / hh_n;
    logf(s, "for discretization and simulation cabledx =%g  hhdx=%g, steps hh %i y ct %i\n",
         ct_dx, hh_dx, hh_n, ct_n);
    NRT_REQUIRE(!(hh_dx > node), NRT_ERR_INVALID, "HH_Y_branching: element larger than the node");
    logf(s, "total elements steps=%i\n", 0);
  }
  const int part_len = 50;
  // element count first: each part is whole chunks until it holds >= 50
  const int chunk = ct_n + hh_n;
  const int per_part = ((part_len + chunk - 1) / chunk) * chunk;
  const int N = 3 * per_part;
  Net n{s};
  n.resize(N);
  s.terminals.assign(3, 0);
  int k = 0, branch_at = 0;
  for (int part = 0; part < 3; part++) {
    int made = 0;
    while (made < part_len) {
      for (int c = 0; c < ct_n; c++) {
        made++;
        if (part == 2 && c == 0 && made == 1) {
          // first element of the third part: hangs off the end of the first
          s.terminals[1] = k - 1;
          n.add(k, k + 1, branch_at, ct_dx, false, in.epsilon);
          s.dl[branch_at] = k;
          s.branching[branch_at] = 1;
          s.dr[k - 1] = k - 1;
        } else {
          n.add(k, k + 1, k - 1, ct_dx, false, in.epsilon);
        }
        k++;
      }
      for (int h = 0; h < hh_n; h++) {
        n.add(k, k + 1, k - 1, hh_dx, true, in.epsilon);
        made++;
        k++;
      }
    }
    if (part == 0) branch_at = made - 1;
  }
  s.terminals[2] = N - 1;
  s.dr[N - 1] = N - 1;
  if (with_hh) {
    in.dist_mes = N * ct_dx / 10;
    s.input[4] = in.input_current;
    s.input[90] = in.input_current;
  } else {
    in.dist_mes = N * ct_dx / 10;
    s.input[1] = in.input_current;
  }
}

// without_branching_CT (space.cpp:593-637): 110 cable elements over xspan.
void build_ct_cable(Scenario& s) {
  Params& in = s.in;
  in.input_time = 0.003;
  in.input_current = 0.8e-9;
  in.end_input = in.init_time + in.input_time;
  if (in.xspan < 1e-10) in.xspan = 0.002;
  in.dist_mes = in.xspan / 10;
  if (in.tspan < 1e-10) in.tspan = 0.025;
  const int N = 110;
  const double dx = (double)in.xspan / N;
  Net n{s};
  n.resize(N);
  // every element's daughter is the next one; the last keeps index N (no sealed
  // end is set on it, space.cpp:628-634)
  for (int k = 0; k < N; k++) n.add(k, k + 1, k - 1, dx, false, in.epsilon);
  s.terminals = {0, N - 1};
  s.input[1] = in.input_current;
}

// general_symmetric_tree_CT (space.cpp:190-330): 15 fixed-length branches of
// 10 um elements built depth first. A branch ends in a branch point unless it is
// the deepest of its line, which then becomes a sealed end; after a sealed end,
// the most recent branch point whose left daughter is still free, searched back
// from the second-to-last recorded one, gets the next element as its left
// daughter, and the remaining depth of the new line is the search distance.
void build_symmetric_tree(Scenario& s) {
  Params& in = s.in;
  in.input_time = 0.003;
  in.input_current = 0.8e-9;
  in.init_time = 0;
  in.end_input = in.init_time + in.input_time;
  if (in.tspan < 1e-10) in.tspan = 0.025;
  const double dx = 10e-6;
  const int len[15] = {44, 62, 48, 50, 28, 50, 26, 32, 48, 50, 30, 32, 40, 55, 38};
  const int sections = 4;
  const int n_term = 1 * (int)std::pow(2, sections - 1) + 1;
  const int branches = (n_term - 1) * 2 - 1;
  int total = 0;
  for (int b = 0; b < branches; b++) total += len[b];
  Net n{s};
  n.resize(total);
  s.terminals.assign(n_term, 0);
  struct Fork {
    int elem;
    bool used;
  };
  std::vector<Fork> forks;  // the last element of every branch, in creation order
  int t = 1, k = 0, depth_left = sections, depth = 0, hang = 0;
  for (int b = 0; b < branches; b++) {
    for (int e = 0; e < len[b]; e++) {
      if (e == len[b] - 1) {
        if (depth == depth_left - 1) {
          s.terminals[t++] = k;
          n.add(k, k, k - 1, dx, false, in.epsilon);
          int back = 0;
          const int nf = (int)forks.size();
          for (int d = 1; d < nf; d++) {
            back++;
            Fork& f = forks[nf - 1 - d];
            if (!f.used) {
              f.used = true;
              hang = f.elem;
              s.dl[f.elem] = k + 1;
              forks.resize(nf - d);
              break;
            }
          }
          depth = 0;
          depth_left = back;
        } else {
          n.add(k, k + 1, k - 1, dx, false, in.epsilon);
          s.branching[k] = 1;
          depth++;
        }
      } else if (hang == 0) {
        n.add(k, k + 1, k - 1, dx, false, in.epsilon);
      } else {
        n.add(k, k + 1, hang, dx, false, in.epsilon);
        hang = 0;
      }
      k++;
      if (e + 1 == len[b] - 1) forks.push_back({k, false});
    }
  }
  in.dist_mes = dx * k / 10;
  logf(s, "total elements steps=%i === %i\n", k, total);
  s.input[1] = in.input_current;
}

// loading_neuron_passive_dendritic_tree (space.cpp:37-182): a segmented neuron
// from `neuron_for_neurite.txt` (per row: element, mother, right daughter,
// branching flag, left daughter, diameter [um], length [um]), one element per
// row after the fictitious element 0.
void build_neuron(Scenario& s, const std::string& dir) {
  Params& in = s.in;
  in.input_time = 0.003;
  in.input_current = 0.8e-9;
  in.end_input = in.init_time + in.input_time;
  const std::string fname = dir + "/neuron_for_neurite.txt";
  std::ifstream f(fname, std::ios::binary);
  NRT_REQUIRE(f.is_open(), NRT_ERR_INVALID, "cannot open " + fname);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  struct Row {
    int me, mother, dr, branching, dl;
    double diameter, length;
  };
  std::vector<Row> rows;
  std::istringstream is(text);
  Row r{};
  while (is >> r.me >> r.mother >> r.dr >> r.branching >> r.dl >> r.diameter >> r.length)
    rows.push_back(r);
  // The reference reads until end-of-file and drops the last read: when the file
  // ends in whitespace that read is an empty row (zeros, which also counts as a
  // terminal), otherwise it is the last real row (space.cpp:62-75).
  const bool trailing_ws = !text.empty() && std::isspace((unsigned char)text.back());
  const int reads = (int)rows.size() + (trailing_ws ? 1 : 0);
  int term_rows = 0;
  for (const Row& q : rows)
    if (q.me == q.dr) term_rows++;
  if (trailing_ws) term_rows++;  // the empty read: me == dR == 0
  const int nrows = reads - 1;
  NRT_REQUIRE(nrows >= 1, NRT_ERR_INVALID, fname + " holds no neuron");
  in.xspan = 0;  // the sum of the rows' lengths [um], for the log only
  for (const Row& q : rows) in.xspan += q.length;
  if (in.tspan < 1e-10) in.tspan = 0.025;
  const int N = nrows + 1;
  Net n{s};
  n.resize(N);
  s.terminals.assign(term_rows, 0);  // element 0 plus the terminal rows
  n.add(0, 1, -1, rows[0].length / 1 * kMicro, false, in.epsilon);
  s.diameter[0] = rows[0].diameter * kMicro;
  int t = 1;
  for (int i = 0; i < nrows; i++) {
    const Row& q = rows[i];
    const int pos = q.me;
    NRT_REQUIRE(pos >= 1 && pos < N && q.mother >= 0 && q.mother < N, NRT_ERR_TOPOLOGY,
                fname + ": element index out of range");
    const bool term = !q.branching && q.me == q.dr;
    n.add(pos, term ? pos : pos + 1, q.mother, q.length / 1 * kMicro, false, in.epsilon);
    if (q.branching) {
      s.dl[pos] = q.dl * 1 - 1 + 1;
      s.branching[pos] = 1;
    }
    s.diameter[pos] = q.diameter * kMicro;
    if (term) {
      NRT_REQUIRE(t < (int)s.terminals.size(), NRT_ERR_TOPOLOGY, fname + ": terminal count");
      s.terminals[t++] = pos;
    }
  }
  in.dist_mes = 100e-6;
  logf(s, "Total steps = %i ", N);
  s.input[1] = in.input_current;
}

// calculate_epsilon_macro (mechanical_model.cpp:21-99): macroscopic strain from
// the membrane's axial strain through `input_data.txt` (case, K, macroscopic and
// membrane strain at t0, lower bound of the membrane strain).
double epsilon_macro(double eps_membrane, const std::string& dir) {
  const std::string fname = dir + "/input_data.txt";
  FILE* fp = std::fopen(fname.c_str(), "r");
  NRT_REQUIRE(fp != nullptr, NRT_ERR_INVALID,
              "the mechanical model needs " + fname + " (case, K, strains)");
  double v[5];
  for (int j = 0; j < 5; j++) {
    if (std::fscanf(fp, "%lf", &v[j]) != 1) {
      std::fclose(fp);
      throw Error(NRT_ERR_INVALID, fname + ": expected five numbers");
    }
  }
  std::fclose(fp);
  const int cs = int(v[0] + 0.1);
  const double K = v[1], macro_t0 = v[2], memb_t0 = v[3], lo = v[4];
  NRT_REQUIRE(cs >= 1 && cs <= 6, NRT_ERR_INVALID, fname + ": case must be 1..6");
  NRT_REQUIRE(!(K > 1 || K < 0), NRT_ERR_INVALID, fname + ": K must be in [0, 1]");
  NRT_REQUIRE(!(eps_membrane > memb_t0 || eps_membrane < lo), NRT_ERR_INVALID,
              "membrane strain outside the mechanical model's range");
  const double constant = macro_t0 - K * memb_t0;
  return K * eps_membrane + constant;
}

void finish(Scenario& s, const std::string& dir) {
  Params& in = s.in;
  // mechanical_model_initialization (init.cpp:320-333): the measurement distance
  // follows the macroscopic strain; element lengths are strained in nrt_setup
  if (in.epsilon > 0) {
    s.epsilon_macro = epsilon_macro(in.epsilon, dir);
    in.dist_mes = in.dist_mes * (1 + s.epsilon_macro);
  }
  in.xspan = in.xspan * (1 + in.epsilon);
  // measurement_point (init.cpp:339-358) over the strained lengths
  double acc = 0;
  s.mes_point = 0;
  for (int k = 0; k < s.N; k++) {
    acc = acc + s.dx0[k] * (1 + s.epsilon[k]);
    if (acc >= in.dist_mes) {
      s.mes_point = k;
      break;
    }
  }
  if (s.mes_point == 0) s.mes_point = int(s.N / 2);
  // materials: the element classes' defaults (discretization.cpp:19-97)
  nrt_material m{};
  m.kind = 0;
  m.layers_my = 45;
  m.rho_a = 1.87;
  m.rho_m = 2.5e9;
  m.per_m = 4e-11;
  m.thickness = 4 * std::pow(10.0, -9);
  m.rest_pot = -0.0655;
  m.thickness_my = 18e-9;
  m.per_my = 1.08e-10;
  m.ro_my = 4.44e6;
  m.sigma_Na = 4.8e-6;
  m.sigma_K = 1.44e-6;
  m.sigma_L = 1.2e-8;
  m.dE_Na = 0.115;
  m.dE_K = -0.012;
  m.n_damage = 2;
  s.mats[0] = m;
  s.mats[1] = m;
  s.mats[1].kind = 1;
  nrt_geometry& g = s.geom;
  g.num_elem = s.N;
  g.mother = s.mother.data();
  g.dr = s.dr.data();
  g.dl = s.dl.data();
  g.branching = s.branching.data();
  g.diameter = s.diameter.data();
  g.dx0 = s.dx0.data();
  g.epsilon = s.epsilon.data();
  g.input = s.input.data();
  g.material = s.material.data();
  g.num_materials = 2;
  g.materials = s.mats;
  g.myelinate = s.hh_model ? 1 : 0;
  g.num_term = (int32_t)s.terminals.size();
  g.terminals = s.terminals.data();
  g.E_MAX = in.E_MAX;
}

}  // namespace
}  // namespace nrt

using namespace nrt;

struct nrt_scenario {
  Scenario s;
};

static thread_local std::string g_scenario_error;

extern "C" int nrt_scenario_create(int argc, const char* const* argv, const char* simulation,
                                   const char* working_directory, nrt_scenario** out) {
  try {
    NRT_REQUIRE(out != nullptr, NRT_ERR_INVALID, "out is NULL");
    *out = nullptr;
    NRT_REQUIRE(argc >= 1 && argv != nullptr, NRT_ERR_INVALID, "argv is empty");
    auto* sc = new nrt_scenario;
    Scenario& s = sc->s;
    try {
      const char* sim = simulation ? simulation : std::getenv("simulation");
      NRT_REQUIRE(sim != nullptr, NRT_ERR_INVALID,
                  "no scenario: set `simulation` (Only_HH_Model, HH_axon, HH_Y_branching, "
                  "CT_cable, CT_Y_branching, CT_symmetric_tree, CT_segmented_neuron)");
      const char* wd = working_directory ? working_directory : std::getenv("working_directory");
      const std::string dir = wd ? wd : "";
      s.simulation = sim;
      parse_args(s, argc, argv);
      NRT_REQUIRE(!(s.explicit_mode && s.implicit_mode), NRT_ERR_INVALID,
                  "both solvers selected (e and i)");
      NRT_REQUIRE(s.explicit_mode || s.implicit_mode, NRT_ERR_INVALID,
                  "no solver selected (e or i)");
      const std::string& v = s.simulation;
      auto need_dir = [&] {
        NRT_REQUIRE(wd != nullptr, NRT_ERR_INVALID, "`working_directory` is not set");
      };
      if (v == "Only_HH_Model") {
        NRT_REQUIRE(!s.implicit_mode, NRT_ERR_UNSUPPORTED,
                    "the implicit solver is not available for the isolated ion channel");
        s.hh_model = true;
        build_only_hh(s);
      } else if (v == "HH_Y_branching") {
        s.hh_model = true;
        build_y(s, true);
      } else if (v == "HH_axon") {
        s.hh_model = true;
        build_hh_axon(s);
      } else if (v == "CT_Y_branching") {
        build_y(s, false);
      } else if (v == "CT_cable") {
        build_ct_cable(s);
      } else if (v == "CT_symmetric_tree") {
        build_symmetric_tree(s);
      } else if (v == "CT_segmented_neuron") {
        need_dir();
        build_neuron(s, dir);
      } else {
        throw Error(NRT_ERR_INVALID, "unknown simulation `" + v + "`");
      }
      if (s.in.epsilon > 0) need_dir();
      finish(s, dir);
    } catch (...) {
      delete sc;
      throw;
    }
    *out = sc;
    return NRT_OK;
  } catch (const Error& e) {
    g_scenario_error = e.what();
    return e.status;
  } catch (const std::exception& e) {
    g_scenario_error = e.what();
    return NRT_ERR_INVALID;
  }
}

extern "C" int nrt_scenario_get(const nrt_scenario* sc, nrt_scenario_info* out) {
  if (!sc || !out) {
    g_scenario_error = "NULL argument";
    return NRT_ERR_INVALID;
  }
  const Scenario& s = sc->s;
  nrt_scenario_info i{};
  i.geometry = s.geom;
  i.explicit_mode = s.explicit_mode;
  i.implicit_mode = s.implicit_mode;
  i.viewer = s.opengl;
  i.tspan = s.in.tspan;
  i.xspan = s.in.xspan;
  i.dt_factor = s.in.dt_factor;
  i.nu = s.in.nu;
  i.epsilon = s.in.epsilon;
  i.E_MAX = s.in.E_MAX;
  i.init_time = s.in.init_time;
  i.end_input = s.in.end_input;
  i.input_current = s.in.input_current;
  i.dist_mes = s.in.dist_mes;
  i.epsilon_macro = s.epsilon_macro;
  i.mes_point = s.mes_point;
  i.out_path = s.out_path.empty() ? nullptr : s.out_path.c_str();
  i.log_path = s.log_path.empty() ? nullptr : s.log_path.c_str();
  i.builder_log = s.log.c_str();
  i.max_steps = s.max_steps;
  i.method = s.method;
  i.precision = s.precision;
  i.restart = s.restart;
  i.tol = s.tol;
  i.vout_path = s.vout_path.empty() ? nullptr : s.vout_path.c_str();
  *out = i;
  return NRT_OK;
}

extern "C" const char* nrt_scenario_last_error(void) { return g_scenario_error.c_str(); }

extern "C" void nrt_scenario_destroy(nrt_scenario* sc) { delete sc; }
