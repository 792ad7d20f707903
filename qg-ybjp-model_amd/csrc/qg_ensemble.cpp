// Ensembles of contexts stepped together from one captured HIP graph, and the
// kernel-shape tuner.  Neither has a counterpart in the reference, which runs
// one member per MPI job with compile-time shapes.
//
// Ensemble: every context (a member or a batch of same-size members) keeps
// its own streams; a step of the ensemble forks from one origin stream into
// every context's stream, runs that context's kernel sequence, and joins back.
// Once every context is in its steady step sequence, K such steps are captured
// into one graph and replayed, so a call of qg_ensemble_step launches one
// graph per K steps and the contexts' kernels overlap on the device.
//
// Tuner: coordinate search over the kernel variants of one context, each trial
// timing whole steps from the same saved state with HIP events; the state is
// restored afterwards.
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "qg_common.hpp"
#include "qg_ctx.hpp"
#include "qgybj.h"

struct qg_ens {
  std::vector<qg_ctx*> m;
  int device = 0;
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  std::map<int, hipGraphExec_t> graphs;  // by steps per replay
  std::vector<int> gen;                  // context generations the graphs were captured at
  int kmax = 8;
  std::string err;
};

namespace {

int efail(qg_ens* e, int code, const std::string& msg) {
  if (e) e->err = msg;
  return code;
}
#define EHIP(e, call)                                                                                   \
  do {                                                                                                  \
    hipError_t _r = (call);                                                                             \
    if (_r != hipSuccess) return efail((e), QG_ERR_HIP, std::string("HIP error '") + hipGetErrorString(_r) + \
                                                          "' in " #call);                               \
  } while (0)

void drop_graphs(qg_ens* e) {
  for (auto& kv : e->graphs) hipGraphExecDestroy(kv.second);
  e->graphs.clear();
}

// Capture k steady steps of every context into one executable graph.
int capture(qg_ens* e, int k, hipGraphExec_t* out) {
  *out = nullptr;
  EHIP(e, hipStreamBeginCapture(e->origin, hipStreamCaptureModeThreadLocal));
  EHIP(e, hipEventRecord(e->fork, e->origin));
  int rc = QG_OK;
  for (size_t i = 0; i < e->m.size() && rc == QG_OK; ++i) {
    qg_ctx* c = e->m[i];
    if (hipStreamWaitEvent(c->stream, e->fork, 0) != hipSuccess) { rc = QG_ERR_HIP; break; }
    const int iter0 = c->iter;
    c->capturing = 1;
    c->evA_live = 0;
    for (int s = 0; s < k && rc == QG_OK; ++s) rc = qgi::step_once(c);
    // join the side stream back into the context's stream
    if (rc == QG_OK && c->evA_live && hipStreamWaitEvent(c->stream, c->evA, 0) != hipSuccess) rc = QG_ERR_HIP;
    c->capturing = 0;
    c->evA_live = 1;
    c->iter = iter0;
    if (rc != QG_OK) {
      e->err = "capture of context " + std::to_string(i) + " failed: " + c->err;
      break;
    }
    if (hipEventRecord(e->join[i], c->stream) != hipSuccess ||
        hipStreamWaitEvent(e->origin, e->join[i], 0) != hipSuccess)
      rc = QG_ERR_HIP;
  }
  hipGraph_t graph = nullptr;
  hipError_t r = hipStreamEndCapture(e->origin, &graph);
  if (rc != QG_OK) {
    if (graph) hipGraphDestroy(graph);
    if (e->err.empty()) e->err = "stream capture failed";
    return rc;
  }
  if (r != hipSuccess) return efail(e, QG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(r));
  r = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (r != hipSuccess) return efail(e, QG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(r));
  return QG_OK;
}

int replay(qg_ens* e, int k, int times) {
  auto it = e->graphs.find(k);
  if (it == e->graphs.end()) {
    hipGraphExec_t g = nullptr;
    int rc = capture(e, k, &g);
    if (rc != QG_OK) return rc;
    it = e->graphs.emplace(k, g).first;
  }
  for (int t = 0; t < times; ++t) EHIP(e, hipGraphLaunch(it->second, e->origin));
  return QG_OK;
}

}  // namespace

extern "C" {

int qg_ensemble_create(qg_ctx* const* members, int32_t n, qg_ens** out) {
  if (!members || !out || n < 1) return QG_ERR_ARG;
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (!members[i]) return QG_ERR_ARG;
    for (int j = 0; j < i; ++j)
      if (members[j] == members[i]) return QG_ERR_ARG;
  }
  qg_ens* e = new qg_ens();
  e->m.assign(members, members + n);
  e->device = members[0]->device;
  for (int i = 1; i < n; ++i)
    if (members[i]->device != e->device) {
      delete e;
      return QG_ERR_ARG;
    }
  hipError_t r = hipSetDevice(e->device);
  if (r == hipSuccess) r = hipStreamCreateWithFlags(&e->origin, hipStreamNonBlocking);
  if (r == hipSuccess) r = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
  e->join.assign(n, nullptr);
  for (int i = 0; i < n && r == hipSuccess; ++i) r = hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming);
  if (r != hipSuccess) {
    qg_ensemble_destroy(e);
    return QG_ERR_HIP;
  }
  for (auto* c : e->m) e->gen.push_back(c->generation);
  *out = e;
  return QG_OK;
}

int qg_ensemble_destroy(qg_ens* e) {
  if (!e) return QG_OK;
  hipSetDevice(e->device);
  if (e->origin) hipStreamSynchronize(e->origin);
  drop_graphs(e);
  for (auto ev : e->join)
    if (ev) hipEventDestroy(ev);
  if (e->fork) hipEventDestroy(e->fork);
  if (e->origin) hipStreamDestroy(e->origin);
  delete e;
  return QG_OK;
}

const char* qg_ensemble_error(const qg_ens* e) { return e ? e->err.c_str() : "null ensemble"; }

int qg_ensemble_set_graph_steps(qg_ens* e, int32_t k) {
  if (!e || k < 1 || k > 64) return QG_ERR_ARG;
  if (k != e->kmax) drop_graphs(e);
  e->kmax = k;
  return QG_OK;
}

int qg_ensemble_step(qg_ens* e, int32_t nsteps) {
  if (!e || nsteps < 0) return QG_ERR_ARG;
  EHIP(e, hipSetDevice(e->device));
  e->err.clear();
  // graphs captured with other kernel variants are stale
  for (size_t i = 0; i < e->m.size(); ++i)
    if (e->gen[i] != e->m[i]->generation) {
      drop_graphs(e);
      for (size_t j = 0; j < e->m.size(); ++j) e->gen[j] = e->m[j]->generation;
      break;
    }
  for (auto* c : e->m) {
    EHIP(e, hipMemsetAsync(c->d.flag, 0, sizeof(int), c->stream));
    EHIP(e, hipStreamSynchronize(c->stream));
  }
  int done = 0;
  // eager steps until every context runs its steady sequence
  while (done < nsteps) {
    bool all = true;
    for (auto* c : e->m) all = all && qgi::steady(c);
    if (all) break;
    for (size_t i = 0; i < e->m.size(); ++i) {
      int rc = qgi::step_once(e->m[i]);
      if (rc != QG_OK) return efail(e, rc, "context " + std::to_string(i) + ": " + e->m[i]->err);
    }
    ++done;
  }
  for (auto* c : e->m) {
    EHIP(e, hipStreamWaitEvent(c->stream, c->evA, 0));
    EHIP(e, hipStreamSynchronize(c->stream));
  }
  const int rest = nsteps - done;
  if (rest > 0) {
    const int k = rest < e->kmax ? rest : e->kmax;
    int rc = replay(e, k, rest / k);
    if (rc == QG_OK && rest % k) rc = replay(e, rest % k, 1);
    if (rc != QG_OK) return rc;
    EHIP(e, hipStreamSynchronize(e->origin));
    for (auto* c : e->m) {
      c->iter += rest;
      // evA was last recorded inside the capture: give it an ordinary
      // (already complete) record for later eager steps
      EHIP(e, hipEventRecord(c->evA, c->stream2));
    }
  }
  int bad = -1;
  for (size_t i = 0; i < e->m.size(); ++i) {
    int flag = 0;
    EHIP(e, hipMemcpy(&flag, e->m[i]->d.flag, sizeof(int), hipMemcpyDeviceToHost));
    if (flag && bad < 0) bad = (int)i;
  }
  if (bad >= 0) return efail(e, QG_ERR_NONFINITE, "non-finite value in context " + std::to_string(bad));
  return QG_OK;
}

// ---------------------------------------------------------------------------
// tuner
// ---------------------------------------------------------------------------
int qg_tune(qg_ctx* c, int32_t steps, int32_t* best, double* best_ms) {
  int rc = qgi::check_ctx(c);
  if (rc) return rc;
  if (steps < 1 || steps > 1000) return qgi::fail(c, QG_ERR_ARG, "qg_tune: steps out of range");
  qg::Dev& d = c->d;
  const size_t S = qgi::field_elems(d.g);
  // everything a step carries from one step to the next
  double2* keep[] = {d.psi, d.q, d.qo, d.BR, d.BI, d.BRo, d.BIo, d.AR, d.AI, d.I[0], d.I[2], d.I[4]};
  const int nk = sizeof(keep) / sizeof(keep[0]);
  HIPCHK(c, hipStreamWaitEvent(c->stream, c->evA, 0));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double2* snap = nullptr;
  hipError_t r = hipMalloc(&snap, (size_t)nk * S * sizeof(double2));
  if (r != hipSuccess) return qgi::fail(c, QG_ERR_NOMEM, "qg_tune: no memory for the state snapshot");
  for (int i = 0; i < nk; ++i)
    hipMemcpyAsync(snap + (size_t)i * S, keep[i], S * sizeof(double2), hipMemcpyDeviceToDevice, c->stream);
  const int iter0 = c->iter, pv0 = c->p_valid, cp0 = c->clear_pending;
  const uint32_t timing0 = c->timing;
  c->timing = 0;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipEventCreate(&t0);
  hipEventCreate(&t1);

  auto restore = [&]() -> hipError_t {
    for (int i = 0; i < nk; ++i) {
      hipError_t e2 = hipMemcpyAsync(keep[i], snap + (size_t)i * S, S * sizeof(double2), hipMemcpyDeviceToDevice,
                                     c->stream);
      if (e2 != hipSuccess) return e2;
    }
    c->iter = iter0;
    c->p_valid = pv0;
    c->clear_pending = cp0;
    return hipSuccess;
  };
  // one trial: restore, reach the steady sequence, time `steps` steps
  auto trial = [&](double* ms) -> int {
    if (restore() != hipSuccess) return QG_ERR_HIP;
    for (int s = 0; s < 2 || !qgi::steady(c); ++s) {
      int e2 = qgi::step_once(c);
      if (e2) return e2;
    }
    if (hipEventRecord(t0, c->stream) != hipSuccess) return QG_ERR_HIP;
    for (int s = 0; s < steps; ++s) {
      int e2 = qgi::step_once(c);
      if (e2) return e2;
    }
    if (hipStreamWaitEvent(c->stream, c->evA, 0) != hipSuccess) return QG_ERR_HIP;
    if (hipEventRecord(t1, c->stream) != hipSuccess) return QG_ERR_HIP;
    if (hipEventSynchronize(t1) != hipSuccess) return QG_ERR_HIP;
    float f = 0.f;
    if (hipEventElapsedTime(&f, t0, t1) != hipSuccess) return QG_ERR_HIP;
    *ms = (double)f / steps;
    return QG_OK;
  };

  // variant ranges: ps_inv 2 shapes, the other transform kernels 4, the
  // solve 4, the A placement 2
  const int hi[QG_NVARIANTS] = {1, 3, 3, 3, 3, 3, 1};
  qg::Tune cur = d.tune;
  auto get = [&](const qg::Tune& t, int i) { return i < 5 ? t.e16[i] : (i == 5 ? t.solve : t.a_main); };
  auto set = [&](qg::Tune& t, int i, int v) {
    if (i < 5) t.e16[i] = v;
    else if (i == 5) t.solve = v;
    else t.a_main = v;
  };
  double best_t = 0.0;
  d.tune = cur;
  rc = trial(&best_t);
  for (int i = 0; i < QG_NVARIANTS && rc == QG_OK; ++i) {
    const int v0 = get(cur, i);
    for (int v = 0; v <= hi[i] && rc == QG_OK; ++v) {
      if (v == v0) continue;
      qg::Tune t = cur;
      set(t, i, v);
      d.tune = t;
      double ms = 0.0;
      rc = trial(&ms);
      if (rc == QG_OK && ms < best_t) {
        best_t = ms;
        cur = t;
      }
    }
  }
  d.tune = cur;
  c->generation++;
  c->tend_valid = 0;  // kept tendencies are those of a trial step
  hipError_t rr = restore();
  if (rr == hipSuccess) rr = hipStreamSynchronize(c->stream);
  hipFree(snap);
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  c->timing = timing0;
  if (rc != QG_OK) return rc;
  if (rr != hipSuccess) return qgi::fail_hip(c, rr, "qg_tune restore");
  if (best)
    for (int i = 0; i < QG_NVARIANTS; ++i) best[i] = get(cur, i);
  if (best_ms) *best_ms = best_t;
  return QG_OK;
}

}  // extern "C"
