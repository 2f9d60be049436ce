// pd_serve.hip -- the request loop of served mode over a job pool, for C and
// C++ hosts (pd_serve_* in include/peridem_gpu.h).  Host code only, on top of
// the pool entry points: requests are laid out on worker threads
// (pd_job_prepare) while the pool steps, admitted in arrival order while
// their footprint fits (pd_add_prepared), stepped in chunks that end at the
// earliest budget end among the resident jobs (pd_step), and reported when
// their budget is used up (or the library stopped them); the caller reads a
// finished job's results from its pool slot and releases it.  No
// counterpart in the reference, which runs one system per process
// (run/simulate3d.cpp:142).

#include <chrono>
#include <condition_variable>
#include <deque>
#include <future>
#include <memory>

namespace {

// A request's particle arrays, owned, so that submit can return before the
// layout is done.
struct ReqPart {
  pd_particle_desc d{};
  std::vector<double> pos, vol, disp, vel, acc, ext;
  std::vector<uint32_t> conn, clamped, bdry;
};

template <class T>
static const T* own(std::vector<T>& v, const T* src, size_t n) {
  if (!src) return nullptr;
  v.assign(src, src + n);
  return v.data();
}

struct Request {
  uint64_t id = 0;
  uint64_t nsteps = 0;
  double t_submit = 0;
  pd_run_desc run{};
  std::vector<ReqPart> parts;
  std::shared_future<std::pair<int, pd_job_blob*>> blob;   // status, layout
};

struct Resident {
  uint64_t id;
  double t_submit, t_admit;
};

}  // namespace

struct pd_server {
  pd_ctx* pool = nullptr;
  uint32_t chunk = 256;
  uint32_t workers = 2;
  std::chrono::steady_clock::time_point t0;
  uint64_t next_id = 1;
  std::deque<std::unique_ptr<Request>> queue;   // arrival order
  std::map<uint32_t, Resident> live;            // pool job slot -> request
  std::deque<pd_serve_event> done;              // finished, not yet returned by poll
  uint64_t steps_run = 0;
  std::string err;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  // start the layout of the first `workers` queued requests
  void prefetch() {
    uint32_t k = 0;
    for (auto& r : queue) {
      if (k++ >= workers) break;
      if (r->blob.valid()) continue;
      Request* q = r.get();
      std::vector<pd_particle_desc> descs(q->parts.size());
      for (size_t i = 0; i < descs.size(); i++) descs[i] = q->parts[i].d;
      q->blob = std::async(std::launch::async, [q, descs]() {
                  pd_job_blob* b = nullptr;
                  const int rc = pd_job_prepare(&q->run, descs.data(), (uint32_t)descs.size(), &b);
                  return std::make_pair(rc, b);
                }).share();
    }
  }
  void finish(uint64_t id, uint32_t job, int status, double ts, double ta) {
    pd_serve_event e{};
    e.request = id;
    e.job = job;
    e.status = status;
    e.t_submit = ts;
    e.t_admit = ta;
    e.t_done = now();
    done.push_back(e);
  }
};

extern "C" {

int pd_serve_create(pd_ctx* pool, uint32_t chunk, uint32_t prepare_threads, pd_server** out) {
  if (!pool || !out || !chunk) return PD_E_ARG;
  if (!pool->reserved) return fail(pool, PD_E_STATE, "pd_serve_create before pd_reserve");
  pd_server* s = new pd_server();
  s->pool = pool;
  s->chunk = chunk;
  s->workers = std::max<uint32_t>(1, std::min<uint32_t>(prepare_threads ? prepare_threads : 2, 16));
  s->t0 = std::chrono::steady_clock::now();
  *out = s;
  return PD_OK;
}

int pd_serve_submit(pd_server* s, const pd_run_desc* run, const pd_particle_desc* parts, uint32_t nparts,
                    uint64_t nsteps, uint64_t* request) {
  if (!s || !run || !parts || !nparts || !nsteps) return PD_E_ARG;
  auto r = std::make_unique<Request>();
  r->id = s->next_id++;
  r->nsteps = nsteps;
  r->t_submit = s->now();
  r->run = *run;
  r->parts.resize(nparts);
  for (uint32_t i = 0; i < nparts; i++) {
    const pd_particle_desc& p = parts[i];
    ReqPart& q = r->parts[i];
    q.d = p;
    const size_t n3 = 3ull * p.nnodes;
    q.d.pos = own(q.pos, p.pos, n3);
    q.d.vol = own(q.vol, p.vol, p.nnodes);
    q.d.disp = own(q.disp, p.disp, n3);
    q.d.vel = own(q.vel, p.vel, n3);
    q.d.acc = own(q.acc, p.acc, n3);
    q.d.extforce = own(q.ext, p.extforce, n3);
    q.d.conn = own(q.conn, p.conn, 2 * (size_t)p.nconn);
    q.d.clamped = own(q.clamped, p.clamped, p.nclamped);
    q.d.bdry = own(q.bdry, p.bdry, p.nbdry);
  }
  if (request) *request = r->id;
  s->queue.push_back(std::move(r));
  s->prefetch();
  return PD_OK;
}

// Admit, step one chunk, harvest; report up to cap finished requests.
int pd_serve_poll(pd_server* s, pd_serve_event* events, uint32_t cap, uint32_t* n) {
  if (!s || !n || (cap && !events)) return PD_E_ARG;
  *n = 0;
  pd_ctx* c = s->pool;
  int rc;
  // admission in arrival order while the next request is laid out and fits;
  // with nothing resident, wait for its layout rather than spin
  while (!s->queue.empty()) {
    Request& q = *s->queue.front();
    if (!q.blob.valid()) s->prefetch();
    if (!s->live.empty() && q.blob.wait_for(std::chrono::seconds(0)) != std::future_status::ready) break;
    const std::pair<int, pd_job_blob*> b = q.blob.get();
    if (b.first != PD_OK) {   // the layout failed: reported as such
      s->finish(q.id, UINT32_MAX, b.first, q.t_submit, s->now());
      s->queue.pop_front();
      s->prefetch();
      continue;
    }
    uint32_t job = 0;
    rc = pd_add_prepared(c, b.second, q.nsteps, &job);
    if (rc == PD_E_NOMEM && !s->live.empty()) break;   // wait for room
    pd_job_blob_free(b.second);
    if (rc != PD_OK) {   // does not fit an empty pool, or another error
      s->finish(q.id, UINT32_MAX, rc, q.t_submit, s->now());
    } else {
      s->live[job] = Resident{q.id, q.t_submit, s->now()};
    }
    s->queue.pop_front();
    s->prefetch();
  }
  if (!s->live.empty()) {
    uint64_t step = s->chunk;
    for (auto& kv : s->live) {
      uint64_t left = 0;
      if (pd_job_remaining(c, kv.first, &left) == PD_OK && left) step = std::min(step, left);
    }
    rc = pd_step(c, (uint32_t)step);
    // a job the library stopped (PD_E_OVERFLOW) leaves the others stepped;
    // any other failure is the caller's
    if (rc != PD_OK && rc != PD_E_OVERFLOW) return rc;
    s->steps_run += step;
    for (auto it = s->live.begin(); it != s->live.end();) {
      uint64_t left = 0;
      const int st = pd_job_remaining(c, it->first, &left);
      if (st == PD_OK && left) {
        ++it;
        continue;
      }
      s->finish(it->second.id, it->first, st, it->second.t_submit, it->second.t_admit);
      it = s->live.erase(it);
    }
  }
  while (*n < cap && !s->done.empty()) {
    events[(*n)++] = s->done.front();
    s->done.pop_front();
  }
  return PD_OK;
}

int pd_serve_pending(pd_server* s, uint64_t* queued, uint64_t* resident, uint64_t* unreported) {
  if (!s) return PD_E_ARG;
  if (queued) *queued = s->queue.size();
  if (resident) *resident = s->live.size();
  if (unreported) *unreported = s->done.size();
  return PD_OK;
}

int pd_serve_release(pd_server* s, uint32_t job) {
  if (!s) return PD_E_ARG;
  return pd_remove_job(s->pool, job);
}

void pd_serve_destroy(pd_server* s) {
  if (!s) return;
  for (auto& r : s->queue)
    if (r->blob.valid()) {
      const auto b = r->blob.get();
      if (b.first == PD_OK) pd_job_blob_free(b.second);
    }
  delete s;
}

}  // extern "C"
