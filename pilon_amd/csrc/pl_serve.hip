// pl_serve.hip — the serving form: a request loop with continuous batching.
//
// Part of the libpilonhip.so translation unit (included by pl_engine.hip, it
// uses its lanes and batch pipeline).  Polish requests (a region of a
// resident contig + read set) arrive at any time through pl_serve_submit.
// Each lane — an independent batch pipeline with its own HIP stream and
// device buffers — runs on a host worker thread; an idle lane takes every
// waiting request that fits beside the requests in flight (estimated device
// bytes against the budget, FIFO, up to max_batch) and runs them as one
// batch, so requests that arrive while others are on the GPU join the next
// free lane and batches of different lanes overlap on the device.  One
// request's result is the same alone or in any batch: regions of a batch
// never share a counter (every output index belongs to one region).
//
// The request body is GenomeFile.processRegions' loop body for one region
// (GenomeFile.scala:101-120) on resident data; the serving loop itself has
// no counterpart in the reference.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

using serve_clock = std::chrono::steady_clock;

struct ServeReq {
  int64_t ticket, tag;
  JobReq job;
  int64_t est;
  serve_clock::time_point t_submit;
};

struct ServeDone {
  pl_serve_result res;
  JobReq job;
  bool keep_pile = false, keep_fix = false;   // parts not yet fetched
  std::vector<int32_t> pile;
  std::vector<pl_call> calls;
  std::vector<plh::Change> changes;
};

// Device bytes one region of G positions can need inside a lane: pileup and
// calls (80 B/position), window and sort tables, and its share of the
// evidence buffers at their growth rates, plus a fixed floor for the
// smallest buffer sizes.
int64_t estimate_bytes(int64_t G, int32_t W) {
  const int64_t per_pos = 4 * PL_NFIELDS + (int64_t)sizeof(pl_call);        // pileup + calls
  const int64_t wins = (G + W - 1) / W;
  const int64_t rec = G / 16 * ((int64_t)sizeof(IndelRec) + 16);            // records + sort keys/values
  const int64_t cand = G / 256 * (2 * (int64_t)sizeof(Cand) + 16 + 8);      // candidates, sort, defer, mutated
  return G * per_pos + wins * ((int64_t)sizeof(WinDev) + 20) + rec + cand + ((int64_t)8 << 20);
}

}  // namespace

struct pl_server {
  pl_ctx* c = nullptr;
  pl_serve_opts o;
  int64_t budget = 0;
  std::vector<std::unique_ptr<Lane>> lanes;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ServeReq> pending;
  std::deque<int64_t> ready;                        // tickets done, not yet polled
  std::unordered_map<int64_t, ServeDone> done;      // results not yet polled / fetched
  int64_t inflight = 0;                             // estimated bytes of running requests
  int64_t next_ticket = 1;
  int64_t n_submitted = 0, n_finished = 0;
  bool stop = false;
};

namespace {

void serve_worker(pl_server* S, Lane* L) {
  (void)hipSetDevice(S->c->device);
  pl_ctx* c = S->c;
  for (;;) {
    std::vector<ServeReq> batch;
    int64_t est = 0;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      // closing drains what is already queued, then the worker exits
      S->cv_work.wait(lk, [&] {
        return (S->stop && S->pending.empty()) ||
               (!S->pending.empty() && S->inflight + S->pending.front().est <= S->budget);
      });
      if (S->pending.empty()) return;
      while (!S->pending.empty() && (int32_t)batch.size() < S->o.max_batch &&
             S->inflight + est + S->pending.front().est <= S->budget) {
        est += S->pending.front().est;
        batch.push_back(S->pending.front());
        S->pending.pop_front();
      }
      S->inflight += est;
    }
    const serve_clock::time_point t_start = serve_clock::now();
    std::vector<JobReq> q;
    q.reserve(batch.size());
    for (const ServeReq& r : batch) q.push_back(r.job);
    pl_status bs = run_jobs(c, L, q, S->o.window, 1, S->o.block, false);
    if (!bs) bs = fetch_batch_meta(L);
    std::vector<ServeDone> out(batch.size());
    for (size_t j = 0; j < batch.size(); j++) {
      ServeDone& d = out[j];
      std::memset(&d.res, 0, sizeof d.res);
      d.res.ticket = batch[j].ticket;
      d.res.tag = batch[j].tag;
      d.res.est_bytes = batch[j].est;
      d.res.batch_size = (int32_t)batch.size();
      d.res.batch_device_ms = bs ? 0.0 : L->last_ms;
      d.job = batch[j].job;
      pl_status s = bs;
      if (!s) s = job_status(L, j);
      if (!s) s = job_changes(c, L, j, &d.res.stats, (S->o.copy_out & PL_SERVE_KEEP_FIX) ? &d.changes : nullptr);
      d.keep_fix = !s && (S->o.copy_out & PL_SERVE_KEEP_FIX);
      d.keep_pile = !s && (S->o.copy_out & PL_SERVE_KEEP_PILEUP);
      if (!s && (S->o.copy_out & PL_SERVE_KEEP_PILEUP)) {
        const int64_t sz = L->hjobs[j].size;
        d.pile.resize((size_t)PL_NFIELDS * sz);
        d.calls.resize((size_t)sz);
        s = job_export(L, j, d.pile.data(), d.calls.data());
      }
      d.res.status = s;
      if (s) d.keep_fix = d.keep_pile = false;
    }
    // a lane keeps its buffers while they fit its share of the budget (no
    // free and allocate again between batches of varying size); above that
    // share they are kept only near what its batches need
    const int64_t held = (int64_t)L->device_bytes();
    if (held > S->budget / (int64_t)S->lanes.size() && held > 2 * est + ((int64_t)64 << 20)) {
      (void)hipStreamSynchronize(L->stream);
      drop_graph(L);
      L->free_buffers();
    }
    const serve_clock::time_point t_end = serve_clock::now();
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->inflight -= est;
      for (size_t j = 0; j < batch.size(); j++) {
        ServeDone& d = out[j];
        d.res.queue_ms = std::chrono::duration<double, std::milli>(t_start - batch[j].t_submit).count();
        d.res.latency_ms = std::chrono::duration<double, std::milli>(t_end - batch[j].t_submit).count();
        const int64_t t = d.res.ticket;
        S->done.emplace(t, std::move(d));
        S->ready.push_back(t);
      }
      S->n_finished += (int64_t)batch.size();
    }
    S->cv_work.notify_all();
    S->cv_done.notify_all();
  }
}

}  // namespace

extern "C" {

pl_status pl_serve_open(pl_ctx* c, const pl_serve_opts* opts, pl_server** out) {
  if (!c || !out) return fail(PL_E_INVAL, "null argument");
  HIPCHK(hipSetDevice(c->device));
  std::unique_ptr<pl_server> S(new pl_server());
  S->c = c;
  std::memset(&S->o, 0, sizeof S->o);
  if (opts) S->o = *opts;
  if (S->o.lanes <= 0) S->o.lanes = 2;
  if (S->o.lanes > 16) return fail(PL_E_INVAL, "at most 16 lanes");
  if (S->o.max_batch <= 0) S->o.max_batch = 256;
  if (S->o.max_queue <= 0) S->o.max_queue = 65536;
  if (S->o.window <= 0) S->o.window = PL_DEFAULT_WINDOW;
  if (S->o.block <= 0) S->o.block = PL_DEFAULT_BLOCK;
  if (S->o.window % 64 || S->o.window > 2048) return fail(PL_E_INVAL, "window %d must be a multiple of 64 up to 2048", S->o.window);
  if (S->o.block != 256 && S->o.block != 512 && S->o.block != 1024) return fail(PL_E_INVAL, "block must be 256, 512 or 1024");
  if (window_lds_total(S->o.window, S->o.block) > 160 * 1024) return fail(PL_E_INVAL, "window/block exceed the LDS of a CU");
  S->budget = S->o.mem_budget;
  if (S->budget <= 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    S->budget = (int64_t)(fr / 4);
  }
  for (int i = 0; i < S->o.lanes; i++) {
    S->lanes.emplace_back(new Lane());
    pl_status s = lane_init(S->lanes.back().get());
    if (s) { for (auto& l : S->lanes) lane_destroy(l.get()); return s; }
  }
  c->serving++;
  for (auto& l : S->lanes) S->workers.emplace_back(serve_worker, S.get(), l.get());
  *out = S.release();
  return PL_OK;
}

pl_status pl_serve_submit(pl_server* S, const pl_region_desc* d, int64_t tag, int64_t* ticket) {
  if (!S || !d) return fail(PL_E_INVAL, "null argument");
  pl_ctx* c = S->c;
  if (d->readset_id < 0 || d->readset_id >= (int32_t)c->rsets.size() || !c->rsets[d->readset_id] ||
      c->rsets[d->readset_id]->contig != d->contig_id)
    return fail(PL_E_INVAL, "read set %d is not a set of contig %d", d->readset_id, d->contig_id);
  const Contig& ct = c->contigs.at(d->contig_id);
  if (d->start < 1 || d->stop < d->start || d->stop > ct.len)
    return fail(PL_E_INVAL, "region %d-%d outside contig of length %lld", d->start, d->stop, (long long)ct.len);
  ServeReq r;
  r.job = {d->readset_id, d->start, d->stop};
  r.tag = tag;
  r.est = estimate_bytes((int64_t)d->stop - d->start + 1, S->o.window);
  r.t_submit = serve_clock::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return fail(PL_E_INVAL, "server is closing");
    if (r.est > S->budget)
      return fail(PL_E_REJECTED, "request needs ~%lld bytes, budget is %lld", (long long)r.est, (long long)S->budget);
    if ((int64_t)S->pending.size() >= S->o.max_queue) return fail(PL_E_BUSY, "%d requests waiting", S->o.max_queue);
    r.ticket = S->next_ticket++;
    S->pending.push_back(r);
    S->n_submitted++;
  }
  S->cv_work.notify_one();
  if (ticket) *ticket = r.ticket;
  return PL_OK;
}

pl_status pl_serve_poll(pl_server* S, pl_serve_result* out, int32_t max, int32_t wait_ms, int32_t* n_out) {
  if (!S || !out || !n_out || max <= 0) return fail(PL_E_INVAL, "bad argument");
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->ready.empty() && wait_ms != 0) {
    auto pred = [&] { return !S->ready.empty() || S->n_finished == S->n_submitted; };
    if (wait_ms < 0) S->cv_done.wait(lk, pred);
    else S->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
  }
  int32_t n = 0;
  while (n < max && !S->ready.empty()) {
    const int64_t t = S->ready.front();
    S->ready.pop_front();
    auto it = S->done.find(t);
    if (it == S->done.end()) continue;
    out[n++] = it->second.res;
    if (!it->second.keep_pile && !it->second.keep_fix) S->done.erase(it);
  }
  *n_out = n;
  return PL_OK;
}

pl_status pl_serve_fetch(pl_server* S, int64_t ticket, int32_t* pile, pl_call* calls) {
  if (!S) return fail(PL_E_INVAL, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  if (!(S->o.copy_out & PL_SERVE_KEEP_PILEUP)) return fail(PL_E_INVAL, "server was opened without PL_SERVE_KEEP_PILEUP");
  auto it = S->done.find(ticket);
  if (it == S->done.end() || !it->second.keep_pile)
    return fail(PL_E_INVAL, "no kept result for ticket %lld (not done, not polled, or fetched)", (long long)ticket);
  if (pile) std::memcpy(pile, it->second.pile.data(), 4 * it->second.pile.size());
  if (calls) std::memcpy(calls, it->second.calls.data(), sizeof(pl_call) * it->second.calls.size());
  it->second.keep_pile = false;
  std::vector<int32_t>().swap(it->second.pile);
  std::vector<pl_call>().swap(it->second.calls);
  if (!it->second.keep_fix) S->done.erase(it);
  return PL_OK;
}

pl_status pl_serve_drain(pl_server* S) {
  if (!S) return fail(PL_E_INVAL, "null argument");
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_done.wait(lk, [&] { return S->n_finished == S->n_submitted; });
  return PL_OK;
}

void pl_serve_close(pl_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (std::thread& t : S->workers) t.join();
  (void)hipSetDevice(S->c->device);
  for (auto& l : S->lanes) lane_destroy(l.get());
  S->c->serving--;
  delete S;
}

}  // extern "C"
