// neurite_hip: drop-in for the reference's `Neurite_solver-gpu` executable.
//
// Same command line (input.cpp:25-149), same environment variables
// (`simulation`, `working_directory`, `idGPU`), same outputs: the `.dat` file of
// "time V" lines at the measurement element every `interval` steps and the `.log`
// file (neurite.cpp:84-291). The reference main's sequence is kept: builder,
// mechanical model and measurement point (nrt_scenario_*), setup chain
// (nrt_setup, on the device), dT = critical step x dtf, numStepsT = int(Tt/dT),
// interval = int(numStepsT/1000), then steps i = 1 .. numStepsT-2 with the
// stimulus on while init_time <= i dT <= end_input. The time loop runs on the
// device (nrt_run) in pieces of whole trace intervals, so the sampled steps are
// the reference's; the explicit solver's V is bit-identical to the reference CPU
// solver's, so are the `.dat` lines.
//
// Differences: errors exit with status 1 (the reference exits 0); the `g` viewer
// flag is accepted and ignored (the OpenGL viewer is not part of this build);
// extension tokens (`steps`, `method`, `prec`, `restart`, `tol`, `vout`) are
// described in neurite_hip.h.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "../../../include/neurite_hip.h"

namespace {

int fail(const char* what, const char* msg) {
  std::fprintf(stderr, "ERROR. %s: %s\n", what, msg ? msg : "");
  return 1;
}

void usage(const char* prog) {
  std::printf("The solver:\n e: explicit mode\n i: implicit mode\nExamples:\n");
  std::printf(" %s e dtf 0.9 IRE 10 NRE 5 out output.txt log simulation.log\n", prog);
  std::printf(" %s i out output.txt log simulation.log\n", prog);
  std::printf("Environment: simulation=<Only_HH_Model|HH_axon|HH_Y_branching|CT_cable|"
              "CT_Y_branching|CT_symmetric_tree|CT_segmented_neuron>, working_directory, "
              "idGPU\n");
  std::printf("Extensions: steps N, method bicgstab|gmres|direct, prec fp64|mixed, "
              "restart M, tol T, vout FILE\n");
}

// Stimulus window of neurite.cpp:227-230 as an inclusive step range, by the
// reference's own comparisons on time_run = i * dT.
void active_range(double dT, long long n_steps, double t0, double t1, long long* a0,
                  long long* a1) {
  *a0 = -1;
  *a1 = -2;
  for (long long i = 1; i < n_steps - 1; i++) {
    const double t = i * dT;
    if (t >= t0 && t <= t1) {
      if (*a0 < 0) *a0 = i;
      *a1 = i;
    } else if (*a0 >= 0) {
      break;  // the window is one interval of steps
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5 || !std::strcmp(argv[1], "help")) {
    if (argc < 2 || std::strcmp(argv[1], "help"))
      std::printf("Your number of arguments is not enough. It has to be at least 5\n");
    usage(argv[0]);
    return 0;
  }
  const auto wall0 = std::chrono::steady_clock::now();
  nrt_scenario* sc = nullptr;
  if (nrt_scenario_create(argc, argv, nullptr, nullptr, &sc) != NRT_OK)
    return fail("scenario", nrt_scenario_last_error());
  nrt_scenario_info si;
  nrt_scenario_get(sc, &si);
  const nrt_geometry& g = si.geometry;
  const int N = g.num_elem;
  if (!si.out_path || !si.log_path) return fail("command line", "`out` and `log` are required");
  FILE* study = std::fopen(si.out_path, "w");
  FILE* standard = std::fopen(si.log_path, "w");
  if (!study || !standard) return fail("output files", "cannot open `out` or `log`");
  if (si.viewer) std::printf("The OpenGL viewer is not part of this build; running without it.\n");

  // log lines of the builder, then of the main (neurite.cpp:141-148)
  std::fputs(si.builder_log, standard);
  std::fprintf(standard, "in.nu=%f\n", si.nu);
  std::fprintf(standard, "in.ep=%f\n", si.epsilon);
  std::fprintf(standard, " Total time %g\n", si.tspan);
  std::fprintf(standard, "REAL The total length is =%g, dis_mes=%g\n", si.xspan, si.dist_mes);
  int H = 0;
  for (int k = 0; k < N; k++) H += g.materials[g.material[k]].kind == 1 ? 1 : 0;
  for (int k = 1; k < N - 1; k++) {
    const int hh = g.materials[g.material[k]].kind == 1;
    std::fprintf(standard, "element=%i type=%i REAL epsilon =%.10f, epsilon=%g\n", k, hh,
                 si.epsilon, g.epsilon[k]);
    std::fprintf(standard, " element size = %g \n", g.dx0[k] * (1 + g.epsilon[k]));
  }

  // setup chain on the device (init.cpp:138-298, neurite.cpp:162-169)
  std::vector<double> dx(N), c_m(N), r_a(N), r_m(N), W(N), K(N), crit(N), rest(N);
  const int Hb = H > 0 ? H : 1;
  std::vector<int32_t> hh_elem(Hb);
  std::vector<double> hm(Hb), hh(Hb), hn(Hb), ENa(Hb), EK(Hb), EL(Hb), gNa(Hb), gK(Hb), GL(Hb),
      lsNa(Hb), lsK(Hb), hrest(Hb);
  nrt_setup_out so{};
  so.dx = dx.data(); so.c_m = c_m.data(); so.r_a = r_a.data(); so.r_m = r_m.data();
  so.W = W.data(); so.K = K.data(); so.crit_dT = crit.data(); so.rest = rest.data();
  so.hh_elem = hh_elem.data(); so.hh_m = hm.data(); so.hh_h = hh.data(); so.hh_n = hn.data();
  so.E_Na = ENa.data(); so.E_K = EK.data(); so.E_L = EL.data(); so.g_Na = gNa.data();
  so.g_K = gK.data(); so.G_L = GL.data(); so.ls_Na = lsNa.data(); so.ls_K = lsK.data();
  so.hh_rest = hrest.data();
  if (nrt_setup(&g, -1, &so) != NRT_OK) return fail("setup", nrt_last_error(nullptr));
  double dT = so.dT_crit;
  std::fprintf(standard, "stability dT=%g\n", dT);
  dT = dT * si.dt_factor;
  std::fprintf(standard, "applied factor %f dT=%g\n", si.dt_factor, dT);
  std::fprintf(standard, "total_time t span %f\n", si.tspan);
  const double steps_f = si.tspan / dT;
  const unsigned numStepsT = (unsigned)int(steps_f);
  std::printf("Total time Steps=%u\n", numStepsT);
  const int interval = int(numStepsT / 1000);
  std::fprintf(standard, "intervals =%i\n", interval);
  std::fprintf(standard, "TotalTime=%g,Total_length=%g\n", numStepsT * dT, si.xspan);
  std::fprintf(standard, "NumStepsT=%i,NumStepsX=%i\n", (int)numStepsT, N);

  // the solver: the reference's GPU methods by default (explicit FTCS; implicit
  // BiCGSTAB from x0 = b, ||r|| <= 1e-14 A checked every 30 iterations)
  nrt_options o;
  nrt_default_options(&o);
  o.kind = si.explicit_mode ? NRT_EXPLICIT : NRT_IMPLICIT;
  if (o.kind == NRT_IMPLICIT) {
    o.krylov = si.method >= 0 ? si.method : NRT_BICGSTAB;
    o.precision = si.precision;
    o.tol_mode = NRT_TOL_ABSOLUTE;
    o.tol = si.tol > 0 ? si.tol : 1e-14;
    if (o.krylov == NRT_BICGSTAB) {
      o.x0_mode = NRT_X0_RHS;
      o.check_every = 30;
    } else if (o.krylov == NRT_GMRES) {
      o.x0_mode = NRT_X0_ZERO;
      o.restart = si.restart > 0 ? si.restart : 150;
      if (si.tol > 0) o.refine_tol = si.tol;
    } else if (si.tol > 0) {
      o.refine_tol = si.tol;
    }
  }
  nrt_problem p{};
  p.num_elem = N;
  p.dT = dT;
  p.potential0 = rest.data();  // setting_rest_pot (init.cpp:260-267)
  p.potential1 = rest.data();
  p.dx = dx.data(); p.c_m = c_m.data(); p.r_a = r_a.data(); p.W = W.data(); p.K = K.data();
  p.input = g.input;
  p.mother = g.mother; p.dr = g.dr; p.dl = g.dl; p.branching = g.branching;
  p.num_term = g.num_term; p.terminals = g.terminals;
  p.num_hh = H; p.hh_elem = hh_elem.data();
  p.hh_m = hm.data(); p.hh_h = hh.data(); p.hh_n = hn.data();
  p.E_Na = ENa.data(); p.E_K = EK.data(); p.E_L = EL.data();
  p.g_Na = gNa.data(); p.g_K = gK.data(); p.G_L = GL.data();
  p.ls_Na = lsNa.data(); p.ls_K = lsK.data(); p.rest_pot = hrest.data();
  nrt_handle* h = nullptr;
  std::printf("The solver is being created\n");
  if (nrt_create(&p, &o, &h) != NRT_OK) return fail("solver", nrt_last_error(nullptr));

  // time loop: steps 1 .. numStepsT-2 (neurite.cpp:224-255), on the device in
  // pieces of whole trace intervals so that GL_cont's cadence carries over
  long long total = (long long)numStepsT - 2;
  if (total < 0) total = 0;
  if (si.max_steps > 0 && si.max_steps < total) total = si.max_steps;
  long long a0, a1;
  active_range(dT, (long long)numStepsT, si.init_time, si.end_input, &a0, &a1);
  const long long every = (long long)interval + 1;
  long long piece = (total / 20 / every + 1) * every;
  std::printf("The simulation begins...\n");
  std::vector<double> tr;
  long long done = 0;
  int status = NRT_OK;
  while (done < total) {
    const long long n = piece < total - done ? piece : total - done;
    tr.assign((size_t)(n / every + 1), 0.0);
    int64_t ns = 0;
    status = nrt_run(h, 1 + done, n, a0, a1, si.mes_point, interval, tr.data(), &ns);
    if (status != NRT_OK && status != NRT_ERR_NOT_CONVERGED) break;
    for (int64_t k = 0; k < ns; k++) {
      const long long i = done + (k + 1) * every;  // the step of the k-th sample
      const double time_run = i * dT;
      std::fprintf(study, "%.20f %.20f\n", time_run, tr[k]);
    }
    done += n;
    std::printf("%.2f%% [Time = %g]\n", 100.0 * done / (numStepsT > 0 ? numStepsT : 1),
                done * dT);
    if (status == NRT_ERR_NOT_CONVERGED)
      std::fprintf(stderr, "WARNING. %s\n", nrt_last_error(h));
  }
  if (status != NRT_OK && status != NRT_ERR_NOT_CONVERGED) {
    const std::string msg = nrt_last_error(h);
    nrt_destroy(h);
    return fail("time loop", msg.c_str());
  }
  std::vector<double> V(N);
  nrt_snapshot(h, V.data());
  if (si.vout_path) {
    FILE* fv = std::fopen(si.vout_path, "w");
    if (!fv) return fail("vout", si.vout_path);
    for (int k = 0; k < N; k++) std::fprintf(fv, "%.20f\n", V[k]);
    std::fclose(fv);
  }
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
  std::printf("\n The simulation has ended. Real_time=%f seconds (%lld steps of %d elements)\n",
              secs, done, N);
  nrt_destroy(h);
  nrt_scenario_destroy(sc);
  std::fclose(study);
  std::fclose(standard);
  return 0;
}
