// Serving loop (include/mrfield.h, mrf_server_*): resident boards, admission
// by an HBM label budget, and continuous batching of requests at relaxation-
// round granularity on one worker thread.
//
// The reference serves /solve one request per call, building the grid and
// routing it synchronously (crates/mr-server/src/lib.rs:362-453).  Here the
// board planes stay on the device, and every request's fields join the one
// running relaxation at the next round boundary (driver_admit), relax
// together with the fields of every other admitted request on any resident
// board, and leave as soon as their own fields have converged
// (driver_finished).  Each request's paths are re-validated on the host like
// the provider's (path_ok), on a validator thread beside the worker so that the
// device keeps relaxing meanwhile; a result is exactly what mrf_solve returns
// for the same nets alone.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "../../include/mrfield.h"
#include "engine.h"
#include "hostcheck.h"
#include "hostpool.h"

using mrf::Error;
using mrf::kMax;
using namespace mrf::host;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct ServeBoard {
  uint32_t id = 0;
  int slot = -1;
  mrf::DevBoard dev;
  mrf_dims dims{};
  std::vector<uint32_t> cost, xe, ye, ve, add;
  std::vector<uint8_t> vf, board, va;
  std::vector<int32_t> halo;
  uint32_t halo_pen = 0;
  uint32_t users = 0;  // queued + running requests
  HostGrid grid() const {
    HostGrid g;
    g.d = dims;
    g.cost = cost.data();
    g.vf = vf.empty() ? nullptr : vf.data();
    g.board = board.empty() ? nullptr : board.data();
    g.xe = xe.data();
    g.ye = ye.data();
    g.ve = ve.data();
    g.va = va.data();
    g.add = add.empty() ? nullptr : add.data();
    g.halo = halo.empty() ? nullptr : halo.data();
    g.halo_pen = halo_pen;
    return g;
  }
};

struct ServeReq {
  uint64_t id = 0;
  ServeBoard* board = nullptr;
  std::vector<mrf_net> nets;  // pads point into plans
  std::vector<NetPlan> plans;
  std::vector<mrf_window> wins;
  std::vector<int32_t> ids;  // empty: all -1
  bool retry_full = false;
  // current pass: indices of nets being solved and their windows
  std::vector<size_t> idx;
  std::vector<mrf_window> pass_wins;
  bool retry_pass = false;
  // the current pass runs on wide labels: its packed run failed the
  // certificate (a label came within one step of the packed limit)
  bool wide = false;
  int group = -1;
  int state = 0;  // 0 queued, 1 running, 2 done
  int code = MRF_OK;
  std::string msg;
  Batch result;
  Clock::time_point t_submit, t_admit, t_done;
  bool admitted_once = false;
  uint32_t rounds = 0, peers_max = 0;
  uint32_t checks = 0;      // passes handed to the validator and not yet checked
  bool check_failed = false;
};

// One finished pass of a request: its device paths, read back before the
// pass's device memory is released, waiting for host re-validation.
struct PassCheck {
  ServeReq* r = nullptr;
  std::vector<size_t> idx;
  std::vector<mrf_window> wins;
  std::vector<uint8_t> routed;
  std::vector<uint32_t> cost, cells;
  std::vector<uint64_t> off;
  bool final = false;  // the request completes when this pass is checked
  bool again = false;  // the pass ran on packed labels that failed the certificate: solve it again, wide
};

}  // namespace

struct mrf_server {
  mrf::Ctx c;
  mrf::RoundDriver d;
  std::thread worker, validator;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_check;
  std::deque<PassCheck*> checks;
  bool stop = false;
  bool broken = false;
  std::string broken_msg;
  std::deque<ServeReq*> queue;
  std::vector<ServeReq*> running;
  std::unordered_map<uint64_t, ServeReq*> reqs;
  std::unordered_map<uint32_t, ServeBoard*> boards;
  uint32_t next_board = 1;
  uint64_t next_req = 1;
  uint32_t max_inflight = 256;
  uint32_t admit_below = 0;  // admission pacing: active tiles under which the device takes more work
  uint64_t admit_small = 0;  // requests of at most this many tiles are admitted despite the pacing
  size_t label_bytes = 0;
  mrf_server_stats st{};
};

namespace {

uint64_t pass_tiles(const mrf_server* S, const ServeReq& r) {
  const uint32_t tw = S->c.shape.tw, th = S->c.shape.th;
  uint64_t t = 0;
  for (size_t k = 0; k < r.idx.size(); ++k) {
    const mrf_window& w = r.pass_wins[k];
    t += (uint64_t)((w.x1 - w.x0 + tw) / tw) * ((w.y1 - w.y0 + th) / th);
  }
  return t;
}

void set_err(mrf_error* err, int code, const std::string& msg, uint32_t net = 0) {
  if (!err) return;
  err->code = code;
  err->net = net;
  snprintf(err->message, sizeof(err->message), "%s", msg.c_str());
}

// Tiles of a request's current pass (its fields' windows at the server's
// tile shape).
uint64_t pass_tiles(const mrf_server* S, const ServeReq& r);

std::vector<mrf::FieldReq> field_reqs(const ServeReq& r) {
  std::vector<mrf::FieldReq> out(r.idx.size());
  for (size_t k = 0; k < r.idx.size(); ++k) {
    const size_t i = r.idx[k];
    const mrf_net& n = r.nets[i];
    mrf::FieldReq& q = out[k];
    q.x0 = r.pass_wins[k].x0;
    q.y0 = r.pass_wins[k].y0;
    q.x1 = r.pass_wins[k].x1;
    q.y1 = r.pass_wins[k].y1;
    q.sources = {n.src};
    q.dst = n.dst;
    q.pads = r.plans[i].pads;
    q.via_pads = r.plans[i].vpads;
    q.own_id = r.ids.empty() ? -1 : r.ids[i];
    q.forced = {n.src, n.dst};
    q.flags = mrf::F_FORCE;
    q.board = (uint32_t)r.board->slot;
  }
  return out;
}

// Reads back the paths of the request's current pass (the group's device
// memory is still held) and sets up the retry pass when one is needed: the
// nets the windowed pass left unrouted, searched again on the full board.
PassCheck* collect_pass(mrf_server* S, ServeReq* r) {
  PassCheck* pc = new PassCheck();
  pc->r = r;
  pc->idx = r->idx;
  pc->wins = r->pass_wins;
  if (!mrf::driver_paths(S->d, r->group, pc->routed, pc->cost, pc->off, pc->cells)) {
    r->wide = true;
    pc->again = true;
    return pc;
  }
  r->wide = false;  // a later pass (the full-board retry) may pack again
  std::vector<size_t> unrouted;
  for (size_t k = 0; k < r->idx.size(); ++k)
    if (!pc->routed[k]) unrouted.push_back(r->idx[k]);
  pc->final = true;
  if (!r->retry_pass && r->retry_full && !unrouted.empty()) {
    const mrf_window full{0, 0, r->board->dims.w - 1, r->board->dims.h - 1};
    r->idx = unrouted;
    r->pass_wins.assign(unrouted.size(), full);
    r->retry_pass = true;
    pc->final = false;
  }
  return pc;
}

using mrf::host::parallel_for;

// Host re-validation of one pass (path_ok, as mrf_solve does), on up to 16
// threads; stores the checked paths in the request's result.  Returns false
// when any path fails.
bool check_pass(PassCheck& pc) {
  ServeReq* r = pc.r;
  const HostGrid g = r->board->grid();
  std::vector<uint8_t> bad(pc.idx.size(), 0);
  parallel_for(pc.idx.size(), 4, [&](size_t a, size_t b) {
    for (size_t k = a; k < b; ++k) {
      if (!pc.routed[k]) continue;
      const size_t i = pc.idx[k];
      const uint32_t* p = pc.cells.data() + pc.off[k];
      const size_t len = pc.off[k + 1] - pc.off[k];
      const int32_t own = r->ids.empty() ? -1 : r->ids[i];
      if (!path_ok(g, r->nets[i], r->plans[i], pc.wins[k], own, p, len, pc.cost[k], false)) {
        bad[k] = 1;
        continue;
      }
      r->result.routed[i] = 1;
      r->result.cost[i] = pc.cost[k];
      r->result.paths[i].assign(p, p + len);
    }
  });
  for (uint8_t v : bad)
    if (v) return false;
  return true;
}

void complete(mrf_server* S, ServeReq* r, int code, const std::string& msg) {
  r->state = 2;
  r->code = code;
  r->msg = msg;
  r->t_done = Clock::now();
  if (r->board && r->board->users) r->board->users--;
  if (code == MRF_OK)
    S->st.completed++;
  else
    S->st.failed++;
}

// Fail everything on the device and in the queue (worker error path).
void fail_all(mrf_server* S, const std::string& msg) {
  std::lock_guard<std::mutex> lk(S->mu);
  S->broken = true;
  S->broken_msg = msg;
  for (ServeReq* r : S->running) complete(S, r, MRF_BACKEND_UNAVAILABLE, msg);
  S->running.clear();
  for (ServeReq* r : S->queue) complete(S, r, MRF_BACKEND_UNAVAILABLE, msg);
  S->queue.clear();
  S->cv_done.notify_all();
}

// Re-validates finished passes off the worker thread, so the device keeps
// relaxing the other requests meanwhile, and completes each request when its
// final pass is checked.  A request that failed elsewhere (state 2) is only
// released: the waiter holds it until its passes here are done.
void validator_loop(mrf_server* S) {
  for (;;) {
    PassCheck* pc = nullptr;
    bool skip = false;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_check.wait(lk, [&] { return S->stop || !S->checks.empty(); });
      if (S->stop) return;
      pc = S->checks.front();
      S->checks.pop_front();
      skip = pc->r->state == 2;
    }
    ServeReq* r = pc->r;
    bool ok = true;
    std::string why;
    if (!skip) {
      try {
        ok = check_pass(*pc);
        if (!ok) why = "device path failed host re-validation";
      } catch (const std::exception& e) {
        ok = false;
        why = e.what();
      }
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (!ok && !r->check_failed) {
      r->check_failed = true;
      r->msg = why;
    }
    if (pc->final && r->state != 2) {
      if (r->check_failed)
        complete(S, r, MRF_BACKEND_UNAVAILABLE, r->msg);
      else
        complete(S, r, MRF_OK, "ok");
    }
    r->checks--;
    delete pc;
    S->cv_done.notify_all();
  }
}

void worker_loop(mrf_server* S) {
  std::vector<int> fin;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty() || !S->running.empty(); });
      if (S->stop) return;
    }
    try {
      // Admission at the round boundary.  Requests go in arrival order while
      // the budget and the slots allow and the running requests leave the
      // device room (active tiles of the next round, plus the seeds of
      // requests admitted just now).  A small request (at most admit_small
      // tiles over its pass) is also admitted when the device is paced or
      // when the requests ahead of it wait for memory: it relaxes beside a
      // large request instead of queueing behind it.
      std::vector<ServeReq*> unfit;  // waiting for memory since this boundary
      for (;;) {
        ServeReq* r = nullptr;
        {
          std::lock_guard<std::mutex> lk(S->mu);
          if (S->queue.empty() || S->running.size() >= S->max_inflight) break;
          const bool paced = !S->running.empty() && mrf::driver_next_tiles(S->d) >= S->admit_below;
          bool head = true;  // the first request not waiting for memory
          for (ServeReq* q : S->queue) {
            if (std::find(unfit.begin(), unfit.end(), q) != unfit.end()) continue;
            if ((head && !paced) || pass_tiles(S, *q) <= S->admit_small) {
              r = q;
              break;
            }
            head = false;
          }
        }
        if (!r) break;
        auto unqueue = [&]() { S->queue.erase(std::find(S->queue.begin(), S->queue.end(), r)); };
        std::vector<mrf::FieldReq> reqs = field_reqs(*r);
        const uint64_t words = mrf::driver_group_words(S->d, reqs);
        if (words > S->d.la.cap()) {
          std::lock_guard<std::mutex> lk(S->mu);
          unqueue();
          complete(S, r, MRF_BACKEND_UNAVAILABLE, "request needs more label memory than the server budget");
          S->cv_done.notify_all();
          continue;
        }
        int g = -1;
        if (!mrf::driver_admit(S->d, reqs, g, !r->wide)) {
          std::lock_guard<std::mutex> lk(S->mu);
          if (!S->running.empty()) {  // wait for memory to free up; a smaller one may fit
            // a pass solved again on wide labels waits for the packed ones
            // to drain: nothing else is admitted before it
            if (r->wide) break;
            unfit.push_back(r);
            continue;
          }
          // nothing else holds device memory: the request can never fit
          unqueue();
          complete(S, r, MRF_BACKEND_UNAVAILABLE, "request exceeds the server's item, field or pool capacity");
          S->cv_done.notify_all();
          continue;
        }
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->queue.front() != r) S->st.admitted_ahead++;
        unqueue();
        r->group = g;
        r->state = 1;
        if (!r->admitted_once) {
          r->t_admit = Clock::now();
          r->admitted_once = true;
        }
        S->running.push_back(r);
        S->st.inflight_peak = std::max<uint32_t>(S->st.inflight_peak, (uint32_t)S->running.size());
        S->st.label_bytes_peak = std::max<size_t>(S->st.label_bytes_peak, (size_t)S->d.la.used() * 8);
      }
      // One round over every admitted request's active tiles.
      size_t n_running;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        n_running = S->running.size();
      }
      mrf::driver_step(S->d);
      // Requests whose fields have all converged leave now; the device's
      // visit counters and kernel times are read at these points only.
      mrf::driver_finished(S->d, fin);
      if (!fin.empty()) mrf::driver_sync_stats(S->d);
      std::vector<ServeReq*> run;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        run = S->running;
        S->st.rounds = S->d.rounds;
        S->st.tile_visits = S->d.tile_visits;
        S->st.kernel_ms = S->d.kernel_ms;
      }
      for (ServeReq* r : run) {
        r->rounds++;
        r->peers_max = std::max<uint32_t>(r->peers_max, (uint32_t)n_running - 1);
      }
      for (int g : fin) {
        ServeReq* r = nullptr;
        for (ServeReq* x : run)
          if (x->group == g) r = x;
        if (!r) {
          mrf::driver_release(S->d, g);
          continue;
        }
        PassCheck* pc = nullptr;
        int code = MRF_OK;
        std::string msg;
        try {
          pc = collect_pass(S, r);
        } catch (const Error& e) {
          code = e.code();
          msg = e.what();
        }
        mrf::driver_release(S->d, g);
        std::lock_guard<std::mutex> lk(S->mu);
        S->running.erase(std::find(S->running.begin(), S->running.end(), r));
        r->group = -1;
        if (!pc) {
          complete(S, r, code, msg);
          S->cv_done.notify_all();
          continue;
        }
        if (pc->again) {
          // the same pass again on wide labels, ahead of later arrivals
          delete pc;
          r->state = 0;
          S->queue.push_front(r);
          S->cv_done.notify_all();
          continue;
        }
        if (!pc->final) {
          r->state = 0;
          S->queue.push_front(r);  // the retry pass goes ahead of later arrivals
        }
        r->checks++;
        S->checks.push_back(pc);
        S->cv_check.notify_one();
        S->cv_done.notify_all();
      }
    } catch (const std::exception& e) {
      fail_all(S, std::string("serving loop failed: ") + e.what());
      // Leave the device state as it is; later submits fail closed.
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stop; });
      return;
    }
  }
}

// Stops and joins the worker and validator threads (either may not have started).
void stop_threads(mrf_server* S) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  S->cv_check.notify_all();
  if (S->worker.joinable()) S->worker.join();
  if (S->validator.joinable()) S->validator.join();
}

}  // namespace

extern "C" {

int mrf_server_create(int device, const mrf_server_config* cfg, mrf_server** out, mrf_error* err) {
  if (!out) return MRF_BACKEND_UNAVAILABLE;
  *out = nullptr;
  mrf_server* S = new (std::nothrow) mrf_server();
  if (!S) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, "out of host memory");
    return MRF_BACKEND_UNAVAILABLE;
  }
  try {
    mrf::ctx_init(S->c, device);
    mrf_server_config k{};
    if (cfg) k = *cfg;
    S->c.shape.tw = k.tile_w ? k.tile_w : 16;
    S->c.shape.th = k.tile_h ? k.tile_h : 16;
    S->c.shape.threads = k.threads ? k.threads : 64;
    S->c.shape.max_local = 32;
    if (S->c.shape.threads < 64 || S->c.shape.threads > 256 || S->c.shape.threads % 64)
      throw Error(MRF_BACKEND_UNAVAILABLE, "threads must be 64, 128, 192 or 256");
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) throw Error(MRF_BACKEND_UNAVAILABLE, "hipMemGetInfo failed");
    size_t lb = k.label_bytes ? k.label_bytes : std::min<size_t>(total_b / 3, size_t(96) << 30);
    if (lb > free_b / 10 * 9) throw Error(MRF_BACKEND_UNAVAILABLE, "label budget exceeds free device memory");
    S->label_bytes = lb;
    S->max_inflight = k.max_inflight ? k.max_inflight : 256;
    S->admit_below = k.admit_below_tiles ? k.admit_below_tiles : 6u * (uint32_t)S->c.cus;
    S->admit_small = k.admit_small_tiles ? k.admit_small_tiles : 16ull * S->admit_below;
    mrf::driver_init(S->d, S->c, lb, k.max_boards ? k.max_boards : 64, k.max_layers ? k.max_layers : 8,
                     S->max_inflight);
    S->st.label_bytes = lb;
    S->worker = std::thread(worker_loop, S);
    S->validator = std::thread(validator_loop, S);
  } catch (const Error& e) {
    stop_threads(S);
    mrf::driver_free(S->d);
    mrf::ctx_free(S->c);
    delete S;
    set_err(err, e.code(), e.what());
    return e.code();
  } catch (const std::exception& e) {
    stop_threads(S);
    mrf::driver_free(S->d);
    mrf::ctx_free(S->c);
    delete S;
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
  *out = S;
  if (err) set_err(err, MRF_OK, "ok");
  return MRF_OK;
}

void mrf_server_destroy(mrf_server* S) {
  if (!S) return;
  stop_threads(S);
  for (PassCheck* pc : S->checks) delete pc;
  for (auto& kv : S->reqs) delete kv.second;
  for (auto& kv : S->boards) {
    mrf::board_free(kv.second->dev);
    delete kv.second;
  }
  mrf::driver_free(S->d);
  mrf::ctx_free(S->c);
  delete S;
}

int mrf_server_add_board(mrf_server* S, mrf_dims dims, const mrf_grid* grid, const mrf_edges* edges,
                         const mrf_pricing* pr, uint32_t* board_id, mrf_error* err) {
  if (!S || !board_id) return MRF_BACKEND_UNAVAILABLE;
  ServeBoard* B = new (std::nothrow) ServeBoard();
  if (!B) return MRF_BACKEND_UNAVAILABLE;
  try {
    check_grid(dims, grid);
    check_edges(dims, edges);
    const size_t n = dims_len(dims);
    B->dims = dims;
    B->cost.assign(grid->cost, grid->cost + n);
    if (grid->n_via_forbidden) B->vf.assign(grid->via_forbidden, grid->via_forbidden + n);
    if (grid->n_board_constraint) B->board.assign(grid->board_constraint, grid->board_constraint + n);
    B->xe.assign(edges->x, edges->x + edges->nx);
    B->ye.assign(edges->y, edges->y + edges->ny);
    B->ve.assign(edges->via, edges->via + edges->nvia);
    B->va.assign(edges->via_allowed, edges->via_allowed + edges->nvia);
    B->xe.push_back(0);
    B->ye.push_back(0);
    B->ve.push_back(0);
    B->va.push_back(0);
    if (pr) {
      if (pr->n_additive) {
        if (pr->n_additive != n || !pr->additive) throw Error(MRF_MALFORMED_GRID, "additive plane length does not match dims");
        B->add.assign(pr->additive, pr->additive + n);
      }
      if (pr->n_halo_owner) {
        if (pr->n_halo_owner != n || !pr->halo_owner) throw Error(MRF_MALFORMED_GRID, "halo plane length does not match dims");
        B->halo.assign(pr->halo_owner, pr->halo_owner + n);
        B->halo_pen = pr->halo_penalty;
      }
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->broken) throw Error(MRF_BACKEND_UNAVAILABLE, S->broken_msg);
    // The worker only reads the board table between rounds under this lock's
    // protocol: uploads go on the server's stream, which the worker also
    // drives, so they are ordered before any round that uses the board.
    mrf::board_upload(S->c, B->dev, dims.w, dims.h, dims.layers, B->cost.data(), B->vf.empty() ? nullptr : B->vf.data(),
                      B->board.empty() ? nullptr : B->board.data(), B->xe.data(), B->ye.data(), B->ve.data(),
                      B->va.data(), B->add.empty() ? nullptr : B->add.data(),
                      B->halo.empty() ? nullptr : B->halo.data(), B->halo_pen);
    B->slot = mrf::driver_add_board(S->d, B->dev);
    if (B->slot < 0) {
      mrf::board_free(B->dev);
      throw Error(MRF_BACKEND_UNAVAILABLE, "no free board slot");
    }
    B->id = S->next_board++;
    S->boards[B->id] = B;
    *board_id = B->id;
  } catch (const Error& e) {
    delete B;
    set_err(err, e.code(), e.what());
    return e.code();
  } catch (const std::exception& e) {
    delete B;
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
  if (err) set_err(err, MRF_OK, "ok");
  return MRF_OK;
}

int mrf_server_remove_board(mrf_server* S, uint32_t board_id, mrf_error* err) {
  if (!S) return MRF_BACKEND_UNAVAILABLE;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->boards.find(board_id);
  if (it == S->boards.end()) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, "unknown board");
    return MRF_BACKEND_UNAVAILABLE;
  }
  if (it->second->users) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, "board is in use by a queued or running request");
    return MRF_BACKEND_UNAVAILABLE;
  }
  mrf::driver_remove_board(S->d, it->second->slot);
  mrf::board_free(it->second->dev);
  delete it->second;
  S->boards.erase(it);
  if (err) set_err(err, MRF_OK, "ok");
  return MRF_OK;
}

int mrf_server_submit(mrf_server* S, uint32_t board_id, const mrf_net* nets, size_t n_nets,
                      const mrf_window* windows, int retry_full, const int32_t* net_ids, uint64_t* request_id,
                      mrf_error* err) {
  if (!S || !request_id || (n_nets && !nets)) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, "null server, nets or request id");
    return MRF_BACKEND_UNAVAILABLE;
  }
  ServeReq* r = new (std::nothrow) ServeReq();
  if (!r) return MRF_BACKEND_UNAVAILABLE;
  try {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->broken) throw Error(MRF_BACKEND_UNAVAILABLE, S->broken_msg);
    auto it = S->boards.find(board_id);
    if (it == S->boards.end()) throw Error(MRF_BACKEND_UNAVAILABLE, "unknown board");
    ServeBoard* B = it->second;
    const HostGrid g = B->grid();
    r->board = B;
    r->nets.assign(nets, nets + n_nets);
    r->plans.resize(n_nets);
    for (size_t i = 0; i < n_nets; ++i) check_net(g, nets[i], i);
    for (size_t i = 0; i < n_nets; ++i) {
      r->plans[i].pads = sorted_unique(nets[i].pads, nets[i].n_pads);
      r->plans[i].vpads = sorted_unique(nets[i].via_pads, nets[i].n_via_pads);
      // the request owns its pad lists: point the copied nets at them
      r->nets[i].pads = r->plans[i].pads.data();
      r->nets[i].n_pads = r->plans[i].pads.size();
      r->nets[i].via_pads = r->plans[i].vpads.data();
      r->nets[i].n_via_pads = r->plans[i].vpads.size();
    }
    const mrf_window full{0, 0, B->dims.w - 1, B->dims.h - 1};
    if (windows) {
      for (size_t i = 0; i < n_nets; ++i) check_window(B->dims, windows[i], i);
      r->wins.assign(windows, windows + n_nets);
    } else {
      r->wins.assign(n_nets, full);
    }
    if (net_ids) r->ids.assign(net_ids, net_ids + n_nets);
    r->retry_full = retry_full != 0 && windows != nullptr;
    r->idx.resize(n_nets);
    for (size_t i = 0; i < n_nets; ++i) r->idx[i] = i;
    r->pass_wins = r->wins;
    r->result.routed.assign(n_nets, 0);
    r->result.cost.assign(n_nets, kMax);
    r->result.paths.assign(n_nets, {});
    r->id = S->next_req++;
    r->t_submit = Clock::now();
    B->users++;
    S->reqs[r->id] = r;
    S->st.submitted++;
    if (n_nets == 0) {
      r->t_admit = r->t_submit;
      complete(S, r, MRF_OK, "ok");
    } else {
      S->queue.push_back(r);
    }
    *request_id = r->id;
  } catch (const Error& e) {
    delete r;
    set_err(err, e.code(), e.what(), e.net());
    return e.code();
  }
  S->cv_work.notify_all();
  if (err) set_err(err, MRF_OK, "ok");
  return MRF_OK;
}

int mrf_server_wait(mrf_server* S, uint64_t request_id, int64_t timeout_ms, mrf_paths* out,
                    mrf_request_stats* stats, mrf_error* err) {
  if (out) memset(out, 0, sizeof(*out));
  if (!S || !out) return MRF_BACKEND_UNAVAILABLE;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(request_id);
  if (it == S->reqs.end()) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, "unknown request");
    return MRF_BACKEND_UNAVAILABLE;
  }
  ServeReq* r = it->second;
  // done, and no pass of it still with the validator
  auto ready = [&] { return r->state == 2 && r->checks == 0; };
  if (timeout_ms < 0) {
    S->cv_done.wait(lk, ready);
  } else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
    set_err(err, MRF_NOT_READY, "request has not finished");
    return MRF_NOT_READY;
  }
  S->reqs.erase(it);
  lk.unlock();
  if (stats) {
    stats->queue_ms = ms_between(r->t_submit, r->admitted_once ? r->t_admit : r->t_done);
    stats->solve_ms = r->admitted_once ? ms_between(r->t_admit, r->t_done) : 0.0;
    stats->total_ms = ms_between(r->t_submit, r->t_done);
    stats->rounds = r->rounds;
    stats->peers_max = r->peers_max;
  }
  int code = r->code;
  std::string msg = r->msg;
  if (code == MRF_OK) {
    try {
      export_paths(r->result, out);
    } catch (const std::exception& e) {
      code = MRF_BACKEND_UNAVAILABLE;
      msg = e.what();
    }
  }
  delete r;
  set_err(err, code, msg);
  return code;
}

int mrf_server_get_stats(mrf_server* S, mrf_server_stats* out) {
  if (!S || !out) return MRF_BACKEND_UNAVAILABLE;
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->st;
  return MRF_OK;
}

}  // extern "C"
