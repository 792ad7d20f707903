// Internal view of a context (qg_ctx) shared by the C-ABI translation units:
// the member API (qg_api.cpp) and the ensemble / tuner (qg_ensemble.cpp).
// Not part of the public interface (include/qgybj.h).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "qg_common.hpp"
#include "qgybj.h"

struct qg_ctx {
  qg::Dev d{};
  hipStream_t stream = nullptr;
  // side stream for the B -> A inversions, which only the next step's update
  // reads: evB marks B^{n+1} written, evA marks A^{n+1} written
  hipStream_t stream2 = nullptr;
  hipEvent_t evB = nullptr, evA = nullptr;
  int device = 0;
  int iter = 0;
  int p_valid = 0;        // I0, I2, I4 hold the inverse-y transforms of q, BR, BI
  int clear_pending = 0;  // steps that still zero the modes outside the box
  std::string err;
  std::vector<void*> allocs;
  uint32_t timing = 0;  // bit k: time kernel id k
  int timing_every = 1;  // time every timing_every-th launch of a kernel
  int tseen[qg::K_NKERNELS] = {};  // launches of each kernel since qg_timing_sample
  struct Ev {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Ev> events;
  double tms[qg::K_NKERNELS] = {};
  int tcnt[qg::K_NKERNELS] = {};
  // stream capture (ensemble graphs): no timing events while capturing, and
  // evA is waited on only once it has been recorded in the current capture
  int capturing = 0;
  int evA_live = 1;
  int generation = 0;  // bumped whenever captured launches would change
  double2* stage = nullptr;  // contiguous staging for host transfers (one field)
  // kept B tendencies (qg_keep_tendencies): buffers allocated on first use,
  // attached to d.Z / d.tR / d.tI while keep_tend is set; tend_valid once a
  // step has filled them since they were attached or the state was set
  int keep_tend = 0;
  int tend_valid = 0;
  double2* kz[2] = {nullptr, nullptr};
  double2* kt[2] = {nullptr, nullptr};
  // work fields of qg_omega (5 spectral + 4 physical), allocated on its first call
  double2* omega_blk = nullptr;
  // spectra output and per-chunk partials of qg_hspec_waves, allocated on its first call
  double* spec_blk = nullptr;
};

namespace qgi {

int fail(qg_ctx* c, int code, const std::string& msg);
int fail_hip(qg_ctx* c, hipError_t e, const char* what);
// advance the context one step on its streams (no synchronisation)
int step_once(qg_ctx* c);
// make the context's device current
int check_ctx(qg_ctx* c);
size_t field_elems(const qg::Geo& g);
// host (iktx, ikty, nz) complex arrays <-> pitched device fields via `stage`
// (d2h synchronises the stream)
hipError_t h2d_field(const qg::Geo& g, double2* dst, const double* src, int nz, hipStream_t st, double2* stage);
hipError_t d2h_field(const qg::Geo& g, double* dst, const double2* src, int nz, hipStream_t st, double2* stage);
// true once the step sequence no longer changes from step to step (after the
// Euler step, the first full-width transforms and the clearing steps)
inline bool steady(const qg_ctx* c) { return c->iter >= 1 && c->p_valid && c->clear_pending == 0; }

}  // namespace qgi

#define HIPCHK(c, call)                                          \
  do {                                                           \
    hipError_t _e = (call);                                      \
    if (_e != hipSuccess) return qgi::fail_hip((c), _e, #call);  \
  } while (0)
