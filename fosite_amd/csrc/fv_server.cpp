// fv_server.cpp -- continuous batching of independent simulation runs.
//
// No reference counterpart (the reference runs one simulation per process).
// Requests (a configuration, its initial data and a stop time) queue on the
// host.  They are admitted in arrival order while their device footprint
// (fv_device_bytes) fits the memory budget (or past a head that does not
// fit, with backfilling); an admitted run lives on the
// device until it reaches its stop time and joins the batch of its element
// type (FP32 or FP64).  One scheduling round:
//
//   1. each non-empty batch advances all its members by up to `round_steps`
//      steps (fv_batch_run, one HIP graph per batch), the two batches from
//      two host threads on their own streams, so their kernels overlap;
//   2. meanwhile the calling thread admits queued requests that fit: it
//      creates their contexts, uploads the initial data and takes the first
//      step on each context's own stream.  With backfilling, runs that fit
//      may pass a queue head that does not, for a bounded number of rounds;
//   3. after both batches returned, finished runs are copied back and their
//      memory released, and the runs admitted in 2 join their batches.
//
// Every member of a batch computes exactly what it computes alone, so a
// served result is bit-identical to the same run made alone with fv_run.
// Only public entry points of the library are used.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fosite_fv.h"

namespace {

enum { Q_QUEUED = 0, Q_RUNNING = 1, Q_DONE = 2, Q_FAILED = 3, Q_TAKEN = 4 };

struct Request {
  fv_config cfg{};
  std::vector<double> pvar, bcc, fcs, w;  // initial data until admission
  double bytes = 0;
  double t_submit = 0, t_admit = -1, t_done = -1;
  int state = Q_QUEUED;
  int passed = 0;  // rounds this request, at the head of the queue, was passed over
  fv_ctx *ctx = nullptr;
  long steps = 0;
  std::vector<double> cvar;  // result
  fv_scalars sc{};
  std::string err;
};

bool finished(const fv_scalars &s, const fv_config &c) {
  return s.brk || std::fabs(c.stoptime - s.time) <= 1.0E-05 * c.stoptime;
}

}  // namespace

struct fv_server {
  fv_server_opts opt{};
  std::string err;
  std::chrono::steady_clock::time_point t0;
  std::vector<std::unique_ptr<Request>> req;  // by id
  std::deque<int> queue;
  fv_batch *batch[2] = {};
  std::vector<int> members[2];  // request ids in batch member order
  double used = 0, peak_used = 0;
  int peak_running = 0;
  long rounds = 0;
  double busy_ms = 0;  // summed wall time of the batch rounds
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

int sfail(fv_server *s, const std::string &m, int code = -1) {
  s->err = m;
  return code;
}

// create, load and first-step one request's context (calling thread)
void prepare(fv_server *s, Request &r) {
  fv_ctx *ctx = nullptr;
  if (fv_create(&ctx, &r.cfg)) {
    r.state = Q_FAILED;
    r.err = "fv_create failed";
    return;
  }
  int brk = 0, rc = 0;
  if (!r.bcc.empty()) rc = fv_set_sound_speeds(ctx, r.bcc.data(), r.fcs.data());
  if (!rc && !r.w.empty() && r.cfg.fargo == 2) rc = fv_set_background_velocity(ctx, r.w.data());
  if (!rc) rc = fv_set_pvar(ctx, r.pvar.data());
  if (!rc) rc = fv_first_step(ctx, &brk);
  if (rc) {
    r.state = Q_FAILED;
    r.err = fv_last_error(ctx);
    fv_destroy(ctx);
    return;
  }
  r.ctx = ctx;
  r.t_admit = s->now();
  std::vector<double>().swap(r.pvar);
  std::vector<double>().swap(r.bcc);
  std::vector<double>().swap(r.fcs);
  std::vector<double>().swap(r.w);
}

// copy a finished run back and release its device memory
void retire(fv_server *s, Request &r) {
  int ni = 0, nj = 0, v = 0;
  fv_dims(r.ctx, &ni, &nj, &v);
  r.cvar.resize((size_t)ni * nj * v);
  if (fv_get_state(r.ctx, r.cvar.data(), nullptr) || fv_get_scalars(r.ctx, &r.sc)) {
    r.state = Q_FAILED;
    r.err = fv_last_error(r.ctx);
  } else {
    r.state = Q_DONE;
  }
  fv_destroy(r.ctx);  // also leaves its batch
  r.ctx = nullptr;
  s->used -= r.bytes;
  r.t_done = s->now();
}

}  // namespace

extern "C" int fv_server_create(fv_server **out, const fv_server_opts *o) {
  if (!out || !o) return -1;
  *out = nullptr;
  if (!(o->hbm_budget_bytes > 0) || o->round_steps < 1 || o->round_steps > 512 || o->backfill < 0) {
    std::fprintf(stderr, "fv_server_create: need hbm_budget_bytes > 0, 1 <= round_steps <= 512, backfill >= 0\n");
    return -1;
  }
  fv_server *s = new (std::nothrow) fv_server();
  if (!s) return -1;
  s->opt = *o;
  s->t0 = std::chrono::steady_clock::now();
  fv_batch_opts bo{o->autotune, 0, o->round_steps, 1};  // both batches of a class tune once
  for (int d = 0; d < 2; d++)
    if (fv_batch_create(&s->batch[d], &bo)) {
      fv_server_destroy(s);
      return -3;
    }
  *out = s;
  return 0;
}

extern "C" int fv_server_submit(fv_server *s, const fv_config *cfg, const double *pvar, const double *bccsound,
                                const double *fcsound, const double *w, int *id) {
  if (!s || !cfg || !pvar || !id) return -1;
  double bytes = 0;
  if (fv_device_bytes(cfg, &bytes)) return sfail(s, "unsupported configuration");
  if (bytes > s->opt.hbm_budget_bytes) return sfail(s, "run needs more device memory than the whole budget", -2);
  auto r = std::make_unique<Request>();
  r->cfg = *cfg;
  r->bytes = bytes;
  const size_t N = (size_t)(cfg->inum + 4) * (cfg->jnum + 4), V = cfg->physics == 17 ? 4 : 3;
  r->pvar.assign(pvar, pvar + V * N);
  if (bccsound && fcsound) {
    r->bcc.assign(bccsound, bccsound + N);
    r->fcs.assign(fcsound, fcsound + 4 * N);
  }
  if (w) r->w.assign(w, w + cfg->inum + 4);
  r->t_submit = s->now();
  *id = (int)s->req.size();
  s->req.push_back(std::move(r));
  s->queue.push_back(*id);
  return 0;
}

extern "C" int fv_server_step(fv_server *s, int *running, int *queued) {
  if (!s) return -1;
  const int K = s->opt.round_steps;
  // 1. advance both batches (one host thread each; FV_SERVER_SERIAL=1 runs
  // them one after the other in the calling thread, before admission)
  int rc_b[2] = {0, 0};
  std::vector<int> taken[2];
  std::thread th[2];
  const double t_round = s->now();
  static const bool serial = std::getenv("FV_SERVER_SERIAL") && std::atoi(std::getenv("FV_SERVER_SERIAL"));
  for (int d = 0; d < 2; d++) {
    if (s->members[d].empty()) continue;
    taken[d].assign(s->members[d].size(), 0);
    if (serial)
      rc_b[d] = fv_batch_run(s->batch[d], K, taken[d].data());
    else
      th[d] = std::thread([s, d, K, &rc_b, &taken] { rc_b[d] = fv_batch_run(s->batch[d], K, taken[d].data()); });
  }
  // 2. admit queued requests that fit while the batches run: in arrival
  // order, and with backfilling past a head that does not fit until that
  // head has been passed over in opt.backfill rounds
  std::vector<int> ready;
  size_t k = 0;
  while (k < s->queue.size()) {
    const int id = s->queue[k];
    Request &r = *s->req[id];
    if (s->used + r.bytes > s->opt.hbm_budget_bytes) {
      if (k == 0 && ++r.passed > s->opt.backfill) break;  // the head waits; no one passes it
      k++;
      continue;
    }
    s->queue.erase(s->queue.begin() + (long)k);
    prepare(s, r);
    if (r.state == Q_FAILED) continue;
    s->used += r.bytes;
    s->peak_used = std::max(s->peak_used, s->used);
    ready.push_back(id);
  }
  bool ran = false;
  for (int d = 0; d < 2; d++)
    if (th[d].joinable()) {
      th[d].join();
      ran = true;
    }
  if (ran) {
    s->rounds++;
    s->busy_ms += 1e3 * (s->now() - t_round);
  }
  for (int d = 0; d < 2; d++)
    if (rc_b[d]) return sfail(s, std::string("batch round failed: ") + fv_batch_last_error(s->batch[d]), rc_b[d]);
  // 3. retire finished runs, then let the admitted ones join
  for (int d = 0; d < 2; d++) {
    std::vector<int> keep;
    for (size_t k = 0; k < s->members[d].size(); k++) {
      Request &r = *s->req[s->members[d][k]];
      r.steps += taken[d].empty() ? 0 : taken[d][k];
      fv_scalars sc;
      fv_get_scalars(r.ctx, &sc);
      if (finished(sc, r.cfg))
        retire(s, r);
      else
        keep.push_back(s->members[d][k]);
    }
    s->members[d].swap(keep);
  }
  for (int id : ready) {
    Request &r = *s->req[id];
    const int d = r.cfg.dtype == 1 ? 0 : 1;
    if (fv_batch_add(s->batch[d], r.ctx)) {
      r.state = Q_FAILED;
      r.err = fv_batch_last_error(s->batch[d]);
      fv_destroy(r.ctx);
      r.ctx = nullptr;
      s->used -= r.bytes;
      continue;
    }
    r.state = Q_RUNNING;
    s->members[d].push_back(id);
  }
  const int run = (int)(s->members[0].size() + s->members[1].size());
  s->peak_running = std::max(s->peak_running, run);
  if (running) *running = run;
  if (queued) *queued = (int)s->queue.size();
  return 0;
}

extern "C" int fv_server_poll(const fv_server *s, int id, fv_request_info *o) {
  if (!s || !o || id < 0 || id >= (int)s->req.size()) return -1;
  const Request &r = *s->req[id];
  o->state = r.state;
  o->steps = r.steps;
  o->t_submit = r.t_submit;
  o->t_admit = r.t_admit;
  o->t_done = r.t_done;
  o->device_bytes = r.bytes;
  return 0;
}

extern "C" int fv_server_result(fv_server *s, int id, double *cvar, fv_scalars *sc) {
  if (!s || id < 0 || id >= (int)s->req.size()) return -1;
  Request &r = *s->req[id];
  if (r.state == Q_FAILED) return sfail(s, "request " + std::to_string(id) + " failed: " + r.err, -5);
  if (r.state != Q_DONE) return sfail(s, "request " + std::to_string(id) + " has no result (yet)", -4);
  if (cvar) std::memcpy(cvar, r.cvar.data(), sizeof(double) * r.cvar.size());
  if (sc) *sc = r.sc;
  std::vector<double>().swap(r.cvar);
  r.state = Q_TAKEN;
  return 0;
}

extern "C" int fv_server_get_stats(const fv_server *s, fv_server_stats *o) {
  if (!s || !o) return -1;
  std::memset(o, 0, sizeof(*o));
  for (auto &r : s->req) {
    o->submitted++;
    if (r->state == Q_DONE || r->state == Q_TAKEN) o->completed++;
    if (r->state == Q_FAILED) o->failed++;
  }
  o->running = (int)(s->members[0].size() + s->members[1].size());
  o->queued = (int)s->queue.size();
  o->peak_running = s->peak_running;
  o->device_bytes = s->used;
  o->peak_device_bytes = s->peak_used;
  o->rounds = s->rounds;
  o->busy_ms = s->busy_ms;
  o->clock = s->now();
  return 0;
}

extern "C" int fv_server_destroy(fv_server *s) {
  if (!s) return 0;
  for (auto &r : s->req)
    if (r->ctx) fv_destroy(r->ctx);
  for (auto *b : s->batch)
    if (b) fv_batch_destroy(b);
  delete s;
  return 0;
}

extern "C" const char *fv_server_last_error(const fv_server *s) { return s ? s->err.c_str() : "null server"; }
