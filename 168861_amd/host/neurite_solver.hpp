// C++ host layer over the C ABI (include/neurite_hip.h).
//
//   nrt::Solver         the reference's solver interface (inc/solver.h:12-24):
//                       update(bool), calculate(), snapshot(double*), virtual dtor
//   nrt::hipSolverError replaces cudaException (inc/cudaException.h:331-357): it
//                       carries the library status, a description, the call's
//                       parameters and the file/function/line that raised it
//   nrt::SolverExHIP    explicit solver   (replaces SolverExGPU, solverEx-gpu.h:79-85)
//   nrt::SolverImHIP    implicit solver   (replaces SolverImGPU, solverIm-gpu.h:111-118)
//
// Header-only, C++98-compatible, so it builds inside the reference tree with the
// reference's own flags as well as in C++17 code.
#ifndef NEURITE_SOLVER_HPP
#define NEURITE_SOLVER_HPP

#include <cstdio>
#include <exception>
#include <sstream>
#include <string>

#include "../../include/neurite_hip.h"

namespace nrt {

class Solver {
 public:
  virtual void update(bool input_active) = 0;
  virtual void calculate() = 0;
  virtual void snapshot(double* potential) = 0;
  virtual ~Solver() {}
};

class hipSolverError : public std::exception {
 public:
  hipSolverError(const char* file, const char* func, int line, const std::string& desc,
                 int status, const std::string& params)
      : file_(file), func_(func), line_(line), desc_(desc), status_(status), params_(params) {
    std::ostringstream os;
    os << file_ << ":" << line_ << " [" << func_ << "] " << desc_ << " -> "
       << nrt_status_string(status_) << " (" << status_ << ")";
    if (!params_.empty()) os << " {" << params_ << "}";
    what_ = os.str();
  }
  virtual ~hipSolverError() throw() {}
  virtual const char* what() const throw() { return what_.c_str(); }
  int status() const { return status_; }
  const std::string& params() const { return params_; }
  void print() const { fprintf(stderr, "ERROR: %s\n", what_.c_str()); }

 private:
  std::string file_, func_;
  int line_;
  std::string desc_;
  int status_;
  std::string params_;
  std::string what_;
};

#define NRT_CHECK_CALL(h, call)                                                        \
  do {                                                                                 \
    int st_ = (call);                                                                  \
    if (st_ != NRT_OK)                                                                 \
      throw ::nrt::hipSolverError(__FILE__, __FUNCTION__, __LINE__, #call, st_,        \
                                  nrt_last_error(h));                                  \
  } while (0)

// Common owner of an nrt_handle.
class HipSolverBase : public Solver {
 public:
  virtual void update(bool input_active) { NRT_CHECK_CALL(h_, nrt_update(h_, input_active ? 1 : 0)); }
  virtual void calculate() { NRT_CHECK_CALL(h_, nrt_calculate(h_)); }
  virtual void snapshot(double* potential) { NRT_CHECK_CALL(h_, nrt_snapshot(h_, potential)); }
  virtual ~HipSolverBase() { nrt_destroy(h_); }

  // Device-resident time loop (neurite.cpp:224-255 without host round trips).
  int64_t run(int64_t first_step, int64_t num_steps, int64_t active_first, int64_t active_last,
              int32_t trace_elem, int64_t trace_every, double* trace_out) {
    int64_t n = 0;
    NRT_CHECK_CALL(h_, nrt_run(h_, first_step, num_steps, active_first, active_last, trace_elem,
                               trace_every, trace_out, &n));
    return n;
  }
  nrt_stats stats() const {
    nrt_stats s;
    nrt_get_stats(h_, &s);
    return s;
  }
  nrt_handle* handle() { return h_; }

 protected:
  HipSolverBase(const nrt_problem& p, const nrt_options& o) : h_(0) {
    int st = nrt_create(&p, &o, &h_);
    if (st != NRT_OK) {
      std::ostringstream params;
      params << "num_elem=" << p.num_elem << " num_hh=" << p.num_hh << " dT=" << p.dT;
      throw hipSolverError(__FILE__, "nrt_create", __LINE__, nrt_last_error(0), st,
                           params.str());
    }
  }
  nrt_handle* h_;

 private:
  HipSolverBase(const HipSolverBase&);
  HipSolverBase& operator=(const HipSolverBase&);
};

inline nrt_options make_options(int kind) {
  nrt_options o;
  nrt_default_options(&o);
  o.kind = kind;
  return o;
}

class SolverExHIP : public HipSolverBase {
 public:
  explicit SolverExHIP(const nrt_problem& p) : HipSolverBase(p, make_options(NRT_EXPLICIT)) {}
  SolverExHIP(const nrt_problem& p, nrt_options o) : HipSolverBase(p, (o.kind = NRT_EXPLICIT, o)) {}
};

class SolverImHIP : public HipSolverBase {
 public:
  explicit SolverImHIP(const nrt_problem& p) : HipSolverBase(p, make_options(NRT_IMPLICIT)) {}
  SolverImHIP(const nrt_problem& p, nrt_options o) : HipSolverBase(p, (o.kind = NRT_IMPLICIT, o)) {}
};

}  // namespace nrt

#endif  // NEURITE_SOLVER_HPP
