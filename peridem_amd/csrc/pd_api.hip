// pd_api.hip -- host side of the C-ABI declared in include/peridem_gpu.h.
//
// Device state is a pool of job slots.  A job is one independent system (the
// particles of one reference run, its contact and wall parameters, its own
// step counter and step budget).  A job occupies whole 64-node bond slices
// and one contiguous range of bond entries; every step launch covers all
// resident jobs, and a job that has used its budget stays frozen in place
// until it is read back and removed.  The plain reference interface
// (pd_add_particle + pd_commit) is one job per instance value with an
// unlimited budget.
//
// Rows are built in the reference's conn2NArr order (load_hdf5.h:27-37) and
// laid out as 64-node slices.  The per-step launch sequence runs on one HIP
// stream; runs of steps replay a captured HIP graph of kGraphSteps steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/peridem_gpu.h"
#include "pd_device.h"
#include "pd_kernels.hip"

using namespace pd;

namespace {

constexpr uint64_t kUnlimited = ~0ull;
constexpr uint64_t kTuneMinNodes = 1u << 18;  // autotune pools of at least this many nodes
constexpr int kDefaultBondShape = 4;          // unstaged, pipelined, 4 entries per batch (always usable)
constexpr int kSmallPoolShape = 10;           // staged, pipelined and unrolled twice, 4 entries per batch
constexpr int kQuadShape = 14;                // four lanes per node slot, 2 entries per lane and batch
// Pools below this many resident nodes fill too few SIMDs for one lane per
// node: the automatic small-pool shape is then kQuadShape (two-sphere impact
// of 8,338 nodes 73.0 -> 56.4 us per step, 25,461 nodes 57.6 -> 49.7; even
// at ~37k nodes, slower above).
constexpr uint64_t kQuadMaxNodes = 1u << 15;
constexpr double kSkinFrac = 0.3;             // candidate-list skin / R_c
constexpr int kGraphSteps = 8;  // steps per captured graph
constexpr int kProfRing = 16;   // event sets of profiled steps in flight
constexpr int kProfLag = 8;     // a profiled step's marks are read this many steps later
// device bytes per node slot (X, U, V, Vold, A, F, E, pid, nflag, deg, wmask, ckey,
// csorted, kCellsPerSlot cell-table keys) and per bond entry
// node arrays + the bond kernel's record copies and both displacement
// buffers + block tables (kLdsCap per 256 nodes); per entry the u32 and u16
// forms.  The candidate lists (4 B per node and list entry) are sized for the
// pool as a whole.
constexpr uint64_t kBytesPerNode = 32 + 32 + 5 * 24 + 4 + 1 + 4 + 1 + 4 + 4 + 8 * kCellsPerSlot + 80 + 36 + 2;
constexpr uint64_t kBytesPerEntry = 6;

struct HostPart {
  uint32_t n = 0;
  std::vector<double> pos;
  std::vector<uint32_t> clamped;
  double snot = 0;
  int movable = 1, breakable = 1, stoppable = 1;
  uint64_t off = 0;         // first node slot
  uint32_t slot = 0;        // particle slot
  uint64_t nbonds = 0;      // directed entries
  std::vector<double> cl0;  // clamped nodes: disp, vel, acc when the state was set
};

struct Job {
  bool used = false;
  pd_run_desc run{};
  std::vector<HostPart> parts;
  uint64_t node0 = 0, nslices = 0;   // node slots [node0, node0 + 64*nslices)
  uint64_t bond0 = 0, nentries = 0;  // bond entries [bond0, bond0 + nentries)
  uint64_t remaining = 0, steps_done = 0, steps_since_set = 0;
  uint64_t movable_bonds = 0, total_bonds = 0, nodes = 0;
  bool reaction_valid = false;
  bool vc_on = false;        // a prescribed-velocity particle is set and switched on
  int64_t vc_part = -1;      // its particle index in the job
  bool lds_ok = true;        // every block table fits kLdsCap
  std::string failed;        // why the job was stopped (empty: it was not)
};

// first-fit allocator over [0, cap)
struct Ranges {
  std::map<uint64_t, uint64_t> fr;  // start -> length
  void init(uint64_t cap) {
    fr.clear();
    if (cap) fr[0] = cap;
  }
  bool alloc(uint64_t n, uint64_t* at) {
    if (n == 0) {
      *at = 0;
      return true;
    }
    for (auto it = fr.begin(); it != fr.end(); ++it)
      if (it->second >= n) {
        *at = it->first;
        uint64_t rest = it->second - n, s = it->first + n;
        fr.erase(it);
        if (rest) fr[s] = rest;
        return true;
      }
    return false;
  }
  void release(uint64_t at, uint64_t n) {
    if (!n) return;
    auto it = fr.emplace(at, n).first;
    auto nx = std::next(it);
    if (nx != fr.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      fr.erase(nx);
    }
    if (it != fr.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        fr.erase(it);
      }
    }
  }
};

// rows of one job in conn2NArr order, node indices relative to the job
struct JobRows {
  uint64_t nodes = 0, nslices = 0, entries = 0;
  std::vector<uint32_t> deg;     // row length
  std::vector<uint64_t> rstart;
  std::vector<uint32_t> rows;
  std::vector<uint32_t> width;   // per slice
  std::vector<uint64_t> poff;    // per particle: first node (job-relative)
  // lanes of a 256-node block hold its nodes by decreasing row length (in
  // classes of 8), so a slice's width (its longest row) wastes few entries on
  // shorter rows
  std::vector<uint8_t> lane_node;  // per slot: offset in the block of the node it holds
  std::vector<uint8_t> node_lane;  // per node: offset in the block of its slot
  uint64_t node_of(uint64_t t) const { return (t & ~(uint64_t)255) + lane_node[t]; }
};

}  // namespace

struct pd_ctx {
  pd_run_desc run0{};
  // pd_add_particle copies, owned here until pd_commit
  std::vector<pd_particle_desc> pending;
  std::vector<std::vector<double>> keep_d;
  std::vector<std::vector<uint32_t>> keep_u;
  std::vector<std::pair<uint32_t, uint32_t>> plain;  // plain particle -> (job, local)
  bool committed = false, reserved = false;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Dev d{};
  std::vector<void*> allocs;
  uint64_t cap_slices = 0, cap_entries = 0, cap_parts = 0, cap_jobs = 0, cap_pairs = 0;
  std::vector<Job> jobs;
  Ranges node_free, bond_free;
  std::vector<uint32_t> part_free;
  pd_counters cnt{};
  double last_ms = 0;
  int prof = 0;                 // PD_PROFILE_PHASES or PD_PROFILE_BOND while profiling
  hipEvent_t pev[kProfRing][PD_NPHASE + 1] = {};    // phase marks of profiled steps
  pd_profile acc{};
  bool use_graph = true, graph_dirty = true;
  uint64_t hwm_slices = 0;      // end of the highest node range in use
  int bond_shape = -1;          // index into bond_shapes() or kWideShape; -1: tune at the next step
  bool small_auto = false;      // bond_shape is the automatic small-pool choice
  int arith = PD_ARITH_BITWISE; // bond-term arithmetic (pd_set_arithmetic)
  uint32_t lds_bad = 0;         // resident jobs whose block tables exceed kLdsCap
  uint64_t tuned_nodes = 0;     // resident nodes when the shape was tuned
  double tune_ms[kBondShapes] = {};
  uint32_t* conn_buf = nullptr;   // connectivity read-back: pairs of one particle, then their count
  uint64_t conn_cap = 0;          // in pairs
  char* stage = nullptr;          // pinned host staging of read-backs
  size_t stage_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_par = 0;            // displacement buffer the captured graph starts from
  // the two displacement buffers; ucur holds the current displacement
  double2* uxy[2] = {nullptr, nullptr};
  double* uz[2] = {nullptr, nullptr};
  int ucur = 0;
  uint64_t list_regrows = 0;    // candidate-list capacity increases
};

static int fail(pd_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(c, call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail((c), PD_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                   \
  } while (0)

template <typename T>
static int dalloc(pd_ctx* c, T** p, size_t n) {
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess)
    return fail(c, e == hipErrorOutOfMemory ? PD_E_NOMEM : PD_E_HIP, "hipMalloc(%zu) failed: %s",
                bytes, hipGetErrorString(e));
  c->allocs.push_back((void*)*p);
  return PD_OK;
}

static inline unsigned grid_of(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

static void void_part(DevPart& q, uint32_t void_inst) {
  memset(&q, 0, sizeof q);
  q.inst = void_inst;
  q.rho = q.rinv = 1.0;
}

static void void_inst(DevInst& q) {
  memset(&q, 0, sizeof q);
  q.dt = 1.0;
  q.vc_slot = kNoKey;
}

static void fill_inst(const pd_run_desc& r, DevInst& D) {
  D.Rc = r.contact_rad;
  D.Kn = r.normal_stiffness;
  D.mu = r.friction_coefficient;
  D.zeta = r.damping_ratio;
  // candidate lists: radius R_c plus a skin (the 1e-9 covers the few ulp by
  // which differently ordered distance formulas differ); rebuilt when a node
  // has moved by half the skin, less a relative 1e-6
  const double skin = kSkinFrac * r.contact_rad;
  D.rb = r.contact_rad * (1.0 + 1e-9) + skin + 1e-300;
  D.cell = D.rb;
  const double h = 0.5 * skin * (1.0 - 1e-6);
  D.du_lim2 = h * h;
  D.dirty = 1;
  D.dt = r.dt;
  D.allow_wall = r.allow_wall;
  D.allow_damping = r.allow_damping;
  D.allow_friction = r.allow_friction;
  D.self_contact = r.self_contact;
  D.enable_fracture = r.enable_fracture;
  D.gradient_extforce = r.gradient_extforce;
  D.compute_reaction = r.compute_wall_reaction;
  D.extforce_maxstep = r.extforce_maxstep;
  D.enable_torque = r.enable_torque;
  D.nl_bdry_only = r.nl_bdry_only;
  D.vc_slot = kNoKey;
}

// Cells per axis of a job's contact grid: its nodes' box at placement (the
// particles' boxes) widened by an eighth of its extent and four cells on
// each side, and no wider than the walls' box (plus a cell and a half) when
// the job has walls, in cells of edge `cell`; scaled down together until the
// grid fits `share` keys.  Any size is correct (positions wrap around the axes); one that spans
// the job keeps its cells on distinct keys.
static void job_grid(const double* lrtb, uint32_t np, const pd_run_desc& r, double cell, uint64_t share,
                     uint32_t gn[3]) {
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (uint32_t i = 0; i < np; i++)
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], lrtb[6 * i + k]);
      hi[k] = std::max(hi[k], lrtb[6 * i + 3 + k]);
    }
  double e[3];
  for (int k = 0; k < 3; k++) {
    double n = 1.25 * (hi[k] - lo[k]) / cell + 8.0;
    if (r.allow_wall) n = std::min(n, (r.wall_lrtb[3 + k] - r.wall_lrtb[k]) / cell + 3.0);
    if (!(n >= 1.0)) n = 1.0;
    e[k] = std::min(n, 1.0e6);
  }
  uint64_t n[3];
  for (int k = 0; k < 3; k++) n[k] = (uint64_t)e[k];
  while (n[0] * n[1] * n[2] > share) {
    const double f = std::cbrt((double)share / (double)(n[0] * n[1] * n[2])) * 0.999;
    for (int k = 0; k < 3; k++) n[k] = std::max<uint64_t>(1, (uint64_t)((double)n[k] * f));
    if (n[0] * n[1] * n[2] > share)   // rounding: shrink the longest axis
      *std::max_element(n, n + 3) -= 1;
  }
  for (int k = 0; k < 3; k++) gn[k] = (uint32_t)n[k];
}

static int check_run(pd_ctx* c, const pd_run_desc& r) {
  if (!(r.dt > 0)) return fail(c, PD_E_ARG, "dt must be positive");
  return PD_OK;
}

template <typename F>
static void parallel_for(uint64_t n, uint64_t work, F f);

static int check_particles(pd_ctx* c, const pd_particle_desc* P, uint32_t n) {
  // Connectivity entries, checked in parallel: first bad entry per particle
  std::vector<uint64_t> bad(n, ~0ull);
  uint64_t work = 0;
  for (uint32_t i = 0; i < n; i++) work += P[i].conn ? P[i].nconn : 0;
  parallel_for(n, work, [&](uint64_t i) {
    const pd_particle_desc& d = P[i];
    if (!d.conn) return;
    for (uint64_t l = 0; l < 2 * d.nconn; l++)
      if (d.conn[l] >= d.nnodes) {
        bad[i] = l;
        return;
      }
  });
  for (uint32_t i = 0; i < n; i++) {
    const pd_particle_desc& d = P[i];
    if (d.nnodes == 0 || !d.pos || !d.vol) return fail(c, PD_E_ARG, "particle %u needs nnodes, pos, vol", i);
    if (d.torque_axis > 2) return fail(c, PD_E_ARG, "particle %u: torque_axis must be 0, 1 or 2", i);
    if (d.nconn && !d.conn) return fail(c, PD_E_ARG, "particle %u: nconn > 0 but conn is NULL", i);
    if (bad[i] != ~0ull)
      return fail(c, PD_E_ARG, "particle %u: Connectivity entry %llu out of range", i, (unsigned long long)bad[i]);
    if (d.nclamped && !d.clamped) return fail(c, PD_E_ARG, "particle %u: clamped is NULL", i);
    for (uint32_t q = 0; q < d.nclamped; q++)
      if (d.clamped[q] >= d.nnodes) return fail(c, PD_E_ARG, "particle %u: clamped node out of range", i);
    if (d.nbdry && !d.bdry) return fail(c, PD_E_ARG, "particle %u: bdry is NULL", i);
    for (uint32_t q = 0; q < d.nbdry; q++)
      if (d.bdry[q] >= d.nnodes) return fail(c, PD_E_ARG, "particle %u: boundary node out of range", i);
  }
  return PD_OK;
}

// f(i) for i < n on up to 16 host threads (one when the job is small);
// items are handed out one at a time.
// Job layouts in progress on this process: a caller that lays out several
// jobs at once on its own threads gets correspondingly fewer threads per
// layout, instead of every layout starting its own full set.
static std::atomic<int> g_layouts{0};

template <typename F>
static void parallel_for(uint64_t n, uint64_t work, F f) {
  unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  nt = std::max(1u, nt / (unsigned)std::max(1, g_layouts.load(std::memory_order_relaxed)));
  if (work < (1u << 20)) nt = 1;
  nt = (unsigned)std::min<uint64_t>(nt, n);
  if (nt <= 1) {
    for (uint64_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<uint64_t> next(0);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&]() {
      for (uint64_t i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto& x : th) x.join();
}

// conn2NArr rows of a job (load_hdf5.h:27-37) and its slice widths.  Rows
// of different particles are independent and built in parallel.
static void build_rows(const pd_particle_desc* P, uint32_t np, JobRows& R) {
  R.poff.resize(np);
  uint64_t nodes = 0, work = 0;
  for (uint32_t i = 0; i < np; i++) {
    R.poff[i] = nodes;
    nodes += P[i].nnodes;
    work += P[i].nconn;
  }
  R.nodes = nodes;
  // whole node blocks (kBlockSlices slices)
  R.nslices = ((nodes + kSlice - 1) / kSlice + kBlockSlices - 1) / kBlockSlices * kBlockSlices;
  const uint64_t Np = R.nslices * kSlice;
  R.deg.assign(Np, 0);
  parallel_for(np, work, [&](uint64_t i) {
    uint32_t* dg = R.deg.data() + R.poff[i];
    for (uint64_t l = 0; l < 2 * P[i].nconn; l++) dg[P[i].conn[l]]++;
  });
  R.rstart.assign(Np + 1, 0);
  for (uint64_t g = 0; g < Np; g++) R.rstart[g + 1] = R.rstart[g] + R.deg[g];
  R.rows.resize(R.rstart[Np]);
  parallel_for(np, work, [&](uint64_t i) {
    const uint64_t o = R.poff[i];
    std::vector<uint64_t> fp(R.rstart.begin() + o, R.rstart.begin() + o + P[i].nnodes);
    for (uint64_t l = 0; l < P[i].nconn; l++) {
      const uint32_t a = P[i].conn[2 * l], b = P[i].conn[2 * l + 1];
      R.rows[fp[a]++] = (uint32_t)(o + b);
      R.rows[fp[b]++] = (uint32_t)(o + a);
    }
  });
  const uint64_t bn = (uint64_t)kBlockSlices * kSlice;
  R.lane_node.assign(Np, 0);
  R.node_lane.assign(Np, 0);
  for (uint64_t b0 = 0; b0 < Np; b0 += bn) {
    uint32_t ord[(uint64_t)kBlockSlices * kSlice];
    for (uint32_t i = 0; i < bn; i++) ord[i] = i;
    // row lengths in classes of 8: lanes with rows of similar length, still
    // mostly in node order (neighbouring lanes gather neighbouring records)
    std::stable_sort(ord, ord + bn, [&](uint32_t a, uint32_t b) { return R.deg[b0 + a] / 8 > R.deg[b0 + b] / 8; });
    for (uint32_t t = 0; t < bn; t++) {
      R.lane_node[b0 + t] = (uint8_t)ord[t];
      R.node_lane[b0 + ord[t]] = (uint8_t)t;
    }
  }
  R.width.assign(R.nslices, 0);
  for (uint64_t s = 0; s < R.nslices; s++) {
    uint32_t W = 0;
    for (int l = 0; l < kSlice; l++) W = std::max(W, R.deg[R.node_of(s * kSlice + l)]);
    R.width[s] = W;
  }
  R.entries = 0;
  for (uint64_t s = 0; s < R.nslices; s++) R.entries += (uint64_t)R.width[s] * kSlice;
}

static bool rows_sorted(const JobRows& R, uint64_t g0, uint64_t g1) {
  for (uint64_t g = g0; g < g1; g++)
    for (uint64_t k = R.rstart[g] + 1; k < R.rstart[g + 1]; k++)
      if (R.rows[k] < R.rows[k - 1]) return false;
  return true;
}

static bool any_contact_possible(const pd_ctx* c) {
  if (c->d.npairs) return true;
  for (auto& J : c->jobs)
    if (J.used && J.run.self_contact && J.run.enable_fracture) return true;
  return false;
}

static bool any_torque(pd_ctx* c) {
  for (auto& J : c->jobs)
    if (J.used && J.run.enable_torque) return true;
  return false;
}

static bool any_vc(pd_ctx* c) {
  for (auto& J : c->jobs)
    if (J.used && J.vc_on && J.vc_part >= 0) return true;
  return false;
}

static bool any_bdry_only(pd_ctx* c) {
  for (auto& J : c->jobs)
    if (J.used && J.run.nl_bdry_only) return true;
  return false;
}

static bool any_clamped(pd_ctx* c) {
  for (auto& J : c->jobs)
    if (J.used)
      for (auto& H : J.parts)
        if (H.movable && !H.clamped.empty()) return true;
  return false;
}

static bool any_reaction(pd_ctx* c) {
  for (auto& J : c->jobs)
    if (J.used && J.run.compute_wall_reaction && J.run.allow_wall) return true;
  return false;
}

// particle pairs (i<j) of the same job, in the reference's pair order
// (timeloop.h:243-255); upload is synchronous
static int upload_pairs(pd_ctx* c) {
  std::vector<uint32_t> pairs;
  for (auto& J : c->jobs) {
    if (!J.used) continue;
    for (size_t a = 0; a < J.parts.size(); a++)
      for (size_t b = a + 1; b < J.parts.size(); b++) {
        pairs.push_back(J.parts[a].slot);
        pairs.push_back(J.parts[b].slot);
      }
  }
  const uint64_t np = pairs.size() / 2;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (np > c->cap_pairs) {
    uint64_t cap = std::max<uint64_t>(np, 2 * c->cap_pairs);
    uint32_t* nb = nullptr;
    int rc = dalloc(c, &nb, 2 * cap);
    if (rc) return rc;
    if (c->d.pairs) {
      auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)c->d.pairs);
      if (it != c->allocs.end()) c->allocs.erase(it);
      HIPCHK(c, hipFree(c->d.pairs));
    }
    c->d.pairs = nb;
    c->cap_pairs = cap;
  }
  if (np) HIPCHK(c, hipMemcpy(c->d.pairs, pairs.data(), pairs.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  c->d.npairs = np;
  c->graph_dirty = true;
  return PD_OK;
}

static void recount(pd_ctx* c) {
  uint64_t nodes = 0, parts = 0, bonds = 0, hwm = 0;
  for (auto& J : c->jobs)
    if (J.used) {
      nodes += J.nodes;
      parts += J.parts.size();
      bonds += J.total_bonds;
      hwm = std::max<uint64_t>(hwm, J.node0 / kSlice + J.nslices);
    }
  c->hwm_slices = hwm;
  c->cnt.nodes = nodes;
  c->cnt.particles = parts;
  c->cnt.total_bonds = bonds;
}

extern "C" {

const char* pd_version(void) { return "peridem_gpu 0.2 (gfx950, FP64, bitwise reference order, job pool)"; }

int pd_create(pd_ctx** out, const pd_run_desc* run) {
  if (!out || !run) return PD_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PD_E_HIP;
  pd_ctx* c = new pd_ctx();
  c->run0 = *run;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return PD_E_HIP;
  }
  *out = c;
  return PD_OK;
}

int pd_add_particle(pd_ctx* c, const pd_particle_desc* d, uint32_t* index) {
  if (!c || !d) return PD_E_ARG;
  if (c->committed) return fail(c, PD_E_STATE, "pd_add_particle after pd_commit");
  int rc = check_particles(c, d, 1);
  if (rc) return rc;
  if (!c->pending.empty() && d->instance < c->pending.back().instance)
    return fail(c, PD_E_ARG, "particles must be added in instance order");
  // the caller keeps ownership of its arrays: copy them
  pd_particle_desc q = *d;
  const uint64_t n = d->nnodes;
  auto cpd = [&](const double* s, size_t k) -> const double* {
    if (!s) return nullptr;
    c->keep_d.emplace_back(s, s + k);
    return c->keep_d.back().data();
  };
  auto cpu = [&](const uint32_t* s, size_t k) -> const uint32_t* {
    if (!s || !k) return nullptr;
    c->keep_u.emplace_back(s, s + k);
    return c->keep_u.back().data();
  };
  q.pos = cpd(d->pos, 3 * n);
  q.vol = cpd(d->vol, n);
  q.disp = cpd(d->disp, 3 * n);
  q.vel = cpd(d->vel, 3 * n);
  q.acc = cpd(d->acc, 3 * n);
  q.extforce = cpd(d->extforce, 3 * n);
  q.conn = cpu(d->conn, 2 * d->nconn);
  q.clamped = cpu(d->clamped, d->nclamped);
  q.bdry = cpu(d->bdry, d->nbdry);
  if (index) *index = (uint32_t)c->pending.size();
  c->pending.push_back(q);
  return PD_OK;
}

int pd_job_footprint(const pd_particle_desc* parts, uint32_t nparts, uint64_t* node_slots,
                     uint64_t* bond_entries, uint64_t* bytes) {
  if (!parts || !nparts) return PD_E_ARG;
  if (check_particles(nullptr, parts, nparts)) return PD_E_ARG;
  JobRows R;
  build_rows(parts, nparts, R);
  if (node_slots) *node_slots = R.nslices * kSlice;
  if (bond_entries) *bond_entries = R.entries;
  if (bytes) *bytes = R.nslices * kSlice * kBytesPerNode + R.entries * kBytesPerEntry;
  return PD_OK;
}

int pd_reserve(pd_ctx* c, const pd_pool_desc* pool) {
  if (!c || !pool) return PD_E_ARG;
  if (c->reserved) return fail(c, PD_E_STATE, "pool already reserved");
  if (!pool->max_nodes || !pool->max_jobs || !pool->max_particles)
    return fail(c, PD_E_ARG, "pool needs nodes, particles and jobs");
  const uint64_t S = ((pool->max_nodes + kSlice - 1) / kSlice + kBlockSlices - 1) / kBlockSlices * kBlockSlices,
                 N = S * kSlice;
  if (N >= 0x7FFFFFFFull) return fail(c, PD_E_OVERFLOW, "more than 2^31-1 node slots");
  if (pool->max_particles >= 0x7FFFFFFFull || pool->max_jobs >= 0x7FFFFFFFull)
    return fail(c, PD_E_OVERFLOW, "too many particle or job slots");
  c->cap_slices = S;
  c->cap_entries = pool->max_bond_entries;
  c->cap_parts = pool->max_particles;
  c->cap_jobs = pool->max_jobs;
  const uint64_t P = c->cap_parts + 1, I = c->cap_jobs + 1;  // + the void slots
  // contact cell table: kCellsPerSlot keys per node slot, each job's grid in
  // the keys of its own node slots
  if (N * kCellsPerSlot >= 0xFFFFFFFFull)
    return fail(c, PD_E_ARG, "more than %u node slots", (unsigned)(0xFFFFFFFFull / kCellsPerSlot));
  const uint32_t T = (uint32_t)(N * kCellsPerSlot);
  Dev& d = c->d;
  d.N = N;
  d.S = S;
  d.P = P;
  d.ninst = I;
  d.T = T;
  d.npairs = 0;
  int rc;
#define ALLOC(f, n) \
  if ((rc = dalloc(c, &d.f, (n)))) return rc;
  ALLOC(X, N);
  ALLOC(U, N);
  ALLOC(XY, N);
  ALLOC(ZV, N);
  ALLOC(UXY, N);
  ALLOC(UZ, N);
  ALLOC(UXYo, N);
  ALLOC(UZo, N);
  c->uxy[0] = d.UXY;
  c->uz[0] = d.UZ;
  c->uxy[1] = d.UXYo;
  c->uz[1] = d.UZo;
  c->ucur = 0;
  ALLOC(bloc, std::max<uint64_t>(c->cap_entries, 1));
  ALLOC(btab, S / kBlockSlices * kLdsCap);
  ALLOC(binfo, S / kBlockSlices);
  ALLOC(lane_node, N);
  ALLOC(node_lane, N);
  ALLOC(V, 3 * N);
  ALLOC(Vold, 3 * N);
  ALLOC(A, 3 * N);
  ALLOC(F, 3 * N);
  ALLOC(E, 3 * N);
  ALLOC(pid, N);
  ALLOC(brank, N);
  ALLOC(blist, N);
  ALLOC(nflag, N);
  ALLOC(deg, N);
  ALLOC(soff, S);
  ALLOC(swid, S);
  ALLOC(bond, c->cap_entries);
  ALLOC(part, P);
  ALLOC(inst, I);
  ALLOC(lrtb, 6 * P);
  ALLOC(cent, 3 * P);
  ALLOC(lrtb_next, 6 * P);
  ALLOC(pactive, P);
  ALLOC(counters, 8);
  ALLOC(du2s, I * kDuSlots * kDuStride);
  ALLOC(jneed, I);
  ALLOC(flags, 1);
  ALLOC(ctl, 1);
  ALLOC(reaction, 18 * I);
  ALLOC(wmask, N);
  ALLOC(bflag, N / 256 + 1);
  d.ccap = kMaxCand;
  ALLOC(clist, N * (uint64_t)d.ccap);
  ALLOC(ccnt, N);
  ALLOC(UBXY, N);
  ALLOC(UBZ, N);
  ALLOC(cpos, N);
  ALLOC(cps, N);
  ALLOC(ckey, N);
  ALLOC(cstart, (size_t)T + 1);
  ALLOC(ccur, T);
  ALLOC(csorted, N);
  ALLOC(scan_tmp, T / kScanBlock + 2);
  ALLOC(rjobs, I);
  ALLOC(jkeyed, I);
#undef ALLOC
  hipStream_t st = c->stream;
  // every slot starts void: node -> void particle, void particle -> void job
  const uint32_t vp = (uint32_t)(P - 1), vi = (uint32_t)(I - 1);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(N, 256)), dim3(256), 0, st, d.pid, N, vp);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemsetAsync(d.X, 0, N * sizeof(double4), st));
  HIPCHK(c, hipMemsetAsync(d.U, 0, N * sizeof(double4), st));
  HIPCHK(c, hipMemsetAsync(d.XY, 0, N * sizeof(double2), st));
  HIPCHK(c, hipMemsetAsync(d.ZV, 0, N * sizeof(double2), st));
  HIPCHK(c, hipMemsetAsync(d.UXY, 0, N * sizeof(double2), st));
  HIPCHK(c, hipMemsetAsync(d.UZ, 0, N * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.UXYo, 0, N * sizeof(double2), st));
  HIPCHK(c, hipMemsetAsync(d.UZo, 0, N * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.V, 0, 3 * N * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.A, 0, 3 * N * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.F, 0, 3 * N * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.E, 0, 3 * N * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.nflag, 0, N, st));
  HIPCHK(c, hipMemsetAsync(d.brank, 0xFF, N * sizeof(uint32_t), st));
  HIPCHK(c, hipMemsetAsync(d.deg, 0, N * sizeof(uint32_t), st));
  HIPCHK(c, hipMemsetAsync(d.soff, 0, S * sizeof(uint64_t), st));
  HIPCHK(c, hipMemsetAsync(d.swid, 0, S * sizeof(uint32_t), st));
  HIPCHK(c, hipMemsetAsync(d.cstart, 0, ((size_t)T + 1) * sizeof(uint32_t), st));
  HIPCHK(c, hipMemsetAsync(d.rjobs, 0, I * sizeof(uint32_t), st));
  HIPCHK(c, hipMemsetAsync(d.jkeyed, 0, I * sizeof(uint32_t), st));
  if (c->cap_entries) HIPCHK(c, hipMemsetAsync(d.bond, 0xFF, c->cap_entries * sizeof(uint32_t), st));
  if (c->cap_entries) HIPCHK(c, hipMemsetAsync(d.bloc, 0xFF, c->cap_entries * sizeof(uint16_t), st));
  HIPCHK(c, hipMemsetAsync(d.btab, 0, S / kBlockSlices * kLdsCap * sizeof(uint32_t), st));
  HIPCHK(c, hipMemsetAsync(d.binfo, 0, S / kBlockSlices * sizeof(uint2), st));
  HIPCHK(c, hipMemsetAsync(d.lane_node, 0, N, st));
  HIPCHK(c, hipMemsetAsync(d.node_lane, 0, N, st));
  HIPCHK(c, hipMemsetAsync(d.lrtb, 0, 6 * P * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.pactive, 0, P * sizeof(uint32_t), st));
  HIPCHK(c, hipMemsetAsync(d.counters, 0, 8 * sizeof(unsigned long long), st));
  HIPCHK(c, hipMemsetAsync(d.du2s, 0, I * kDuSlots * kDuStride * sizeof(unsigned long long), st));
  HIPCHK(c, hipMemsetAsync(d.jneed, 0, I * sizeof(uint32_t), st));
  HIPCHK(c, hipMemsetAsync(d.flags, 0, sizeof(uint32_t), st));
  HIPCHK(c, hipMemsetAsync(d.ctl, 0, sizeof(DevCtl), st));
  HIPCHK(c, hipMemsetAsync(d.reaction, 0, 18 * I * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.wmask, 0, N, st));
  HIPCHK(c, hipMemsetAsync(d.ccnt, 0, N, st));
  HIPCHK(c, hipMemsetAsync(d.Vold, 0, 3 * N * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.UBXY, 0, N * sizeof(double2), st));
  HIPCHK(c, hipMemsetAsync(d.UBZ, 0, N * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.bflag, 0, N / 256 + 1, st));
  std::vector<DevPart> vparts(P);
  for (auto& q : vparts) void_part(q, vi);
  std::vector<DevInst> vinst(I);
  for (auto& q : vinst) void_inst(q);
  std::vector<unsigned long long> lnext(6 * P);
  for (uint64_t q = 0; q < P; q++)
    for (int k = 0; k < 3; k++) {
      lnext[6 * q + k] = ~0ull;
      lnext[6 * q + 3 + k] = 0ull;
    }
  HIPCHK(c, hipStreamSynchronize(st));
  HIPCHK(c, hipMemcpy(d.part, vparts.data(), P * sizeof(DevPart), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d.inst, vinst.data(), I * sizeof(DevInst), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d.lrtb_next, lnext.data(), lnext.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
  c->node_free.init(S);
  c->bond_free.init(c->cap_entries);
  c->part_free.clear();
  for (uint64_t q = c->cap_parts; q-- > 0;) c->part_free.push_back((uint32_t)q);
  c->jobs.assign(c->cap_jobs, Job());
  c->reserved = true;
  c->graph_dirty = true;
  return PD_OK;
}

}  // extern "C"

// A job laid out on the host, independent of where it will sit in a pool:
// node arrays, the particles' records and the bond slices with job-relative
// node indices.  Building it needs no context (pd_job_prepare may run on
// another thread while a pool steps); placing it is a few copies.
struct pd_job_blob {
  pd_run_desc run{};
  uint32_t np = 0;
  uint64_t nodes = 0, nslices = 0, entries = 0;
  std::vector<uint64_t> poff;     // per particle: first node (job-relative)
  std::vector<uint32_t> deg, width;
  std::vector<double4> X, U;
  std::vector<double> V, A, E;
  std::vector<uint32_t> lpid;     // per node slot: particle index in the job, ~0 = void
  std::vector<uint32_t> brank;    // per node slot: rank in its boundary list (nl_bdry_only jobs)
  std::vector<uint32_t> blist;    // boundary lists, job-relative node indices, particle by particle
  std::vector<uint64_t> bofs;     // per particle: first entry of its list in blist
  std::vector<uint8_t> nflag;
  std::vector<DevPart> dparts;    // node_off job-relative
  std::vector<double> lrtb;
  std::vector<HostPart> parts;    // off job-relative
  std::vector<uint64_t> soff;     // relative to the job's first bond entry
  std::vector<uint32_t> bond;     // job-relative node indices, kPad padding
  std::vector<uint16_t> bloc;     // the entries as 16 * positions in their block's table, kPad16 padding
  std::vector<uint32_t> btab;     // per block, kLdsCap slots: sorted job-relative node indices
  std::vector<uint2> binfo;       // per block: table length, position of the block's first node
  std::vector<uint8_t> lane_node, node_lane;   // see JobRows
  bool lds_ok = true;             // every block table fits kLdsCap
  uint32_t fl = 0;
  uint64_t movable_bonds = 0, total_bonds = 0;
};

namespace pd {
// bond entries of a newly placed job: job-relative -> pool node indices
__global__ __launch_bounds__(256) void k_rebase(uint32_t* b, uint64_t n, uint32_t g0) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && b[i] != kPad) b[i] += g0;
}
}  // namespace pd

static thread_local std::string t_err;
static int stage_reserve(pd_ctx* c, size_t bytes);
constexpr size_t kStageJobMax = 256ull << 20;   // largest job uploaded through the pinned stage

extern "C" {

int pd_job_prepare(const pd_run_desc* run, const pd_particle_desc* P, uint32_t np, pd_job_blob** out) {
  if (!run || !P || !np || !out) return PD_E_ARG;
  *out = nullptr;
  pd_ctx tmp;   // carries the error message only
  int rc;
  if ((rc = check_run(&tmp, *run)) || (rc = check_particles(&tmp, P, np))) {
    t_err = tmp.err;
    return rc;
  }
  g_layouts.fetch_add(1, std::memory_order_relaxed);
  struct Done {
    ~Done() { g_layouts.fetch_sub(1, std::memory_order_relaxed); }
  } done;
  pd_job_blob* B = new pd_job_blob();
  B->run = *run;
  B->np = np;
  JobRows R;
  build_rows(P, np, R);
  B->nodes = R.nodes;
  B->nslices = R.nslices;
  B->entries = R.entries;
  B->poff = R.poff;
  const uint64_t Np = R.nslices * kSlice;
  B->X.assign(Np, make_double4(0, 0, 0, 0));
  B->U.assign(Np, make_double4(0, 0, 0, 0));
  B->V.assign(3 * Np, 0.0);
  B->A.assign(3 * Np, 0.0);
  B->E.assign(3 * Np, 0.0);
  B->lpid.assign(Np, ~0u);
  B->brank.assign(Np, kNoKey);
  B->bofs.assign(np, 0);
  if (run->nl_bdry_only) {
    // boundary lists (load_hdf5.h bdry_nodes), in list order; a node listed
    // twice would be visited twice by the reference loops
    for (uint32_t i = 0; i < np; i++) {
      B->bofs[i] = B->blist.size();
      for (uint32_t r = 0; r < P[i].nbdry; r++) {
        const uint64_t g = R.poff[i] + P[i].bdry[r];
        if (B->brank[g] != kNoKey) {
          delete B;
          t_err = "nl_bdry_only with a node listed twice in bdry_nodes is not supported";
          return PD_E_UNSUPPORTED;
        }
        B->brank[g] = r;
        B->blist.push_back((uint32_t)g);
      }
    }
  }
  B->nflag.assign(Np, 0);
  B->dparts.resize(np);
  B->lrtb.resize(6 * np);
  // the particles' host-side records (serial), then the node arrays, the
  // fast-path domain checks and the bond slices (parallel)
  for (uint32_t i = 0; i < np; i++) {
    const pd_particle_desc& d = P[i];
    HostPart H;
    H.n = d.nnodes;
    H.pos.assign(d.pos, d.pos + 3ull * d.nnodes);
    if (d.nclamped) H.clamped.assign(d.clamped, d.clamped + d.nclamped);
    H.snot = d.snot;
    H.movable = d.movable;
    H.breakable = d.breakable;
    H.stoppable = d.stoppable;
    H.off = R.poff[i];
    H.nbonds = R.rstart[R.poff[i] + d.nnodes] - R.rstart[R.poff[i]];
    H.cl0.resize(9 * H.clamped.size());
    for (size_t q = 0; q < H.clamped.size(); q++)
      for (int k = 0; k < 3; k++) {
        uint32_t cn = H.clamped[q];
        H.cl0[9 * q + k] = d.disp ? d.disp[3 * cn + k] : 0.0;
        H.cl0[9 * q + 3 + k] = d.vel ? d.vel[3 * cn + k] : 0.0;
        H.cl0[9 * q + 6 + k] = d.acc ? d.acc[3 * cn + k] : 0.0;
      }
    B->total_bonds += H.nbonds;
    if (H.movable) B->movable_bonds += H.nbonds;
    B->parts.push_back(std::move(H));
  }
  const uint64_t work = R.rows.size();
  std::vector<uint8_t> pfl(np, 0);
  parallel_for(np, work, [&](uint64_t i) {
    const pd_particle_desc& d = P[i];
    DevPart& D = B->dparts[i];
    memset(&D, 0, sizeof D);
    D.node_off = R.poff[i];
    D.nnodes = d.nnodes;
    D.rho = d.rho;
    D.rinv = 1 / d.rho;
    D.cnot = d.cnot;
    D.snot = d.snot;
    D.movable = d.movable;
    D.breakable = d.breakable;
    D.stoppable = d.stoppable;
    D.rows_sorted = rows_sorted(R, R.poff[i], R.poff[i] + d.nnodes);
    D.tval = d.torque_val;
    D.taxis = d.torque_axis;
    D.nbdry = run->nl_bdry_only ? d.nbdry : 0;
    D.boff = B->bofs[i];
    bool slow = false;
    double o[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t v = 0; v < d.nnodes; v++) {
      const uint64_t r = R.poff[i] + v;
      const double u0 = d.disp ? d.disp[3 * v] : 0.0, u1 = d.disp ? d.disp[3 * v + 1] : 0.0,
                   u2 = d.disp ? d.disp[3 * v + 2] : 0.0;
      B->X[r] = make_double4(d.pos[3 * v], d.pos[3 * v + 1], d.pos[3 * v + 2], d.vol[v]);
      B->U[r] = make_double4(u0, u1, u2, 0.0);
      for (int k = 0; k < 3; k++) {
        B->V[3 * r + k] = d.vel ? d.vel[3 * v + k] : 0.0;
        B->A[3 * r + k] = d.acc ? d.acc[3 * v + k] : 0.0;
        B->E[3 * r + k] = d.extforce ? d.extforce[3 * v + k] : 0.0;
      }
      B->lpid[r] = (uint32_t)i;
      slow |= !(in_fast_range(d.pos[3 * v]) && in_fast_range(d.pos[3 * v + 1]) && in_fast_range(d.pos[3 * v + 2]));
      // initial lagged box: min/max of pos + disp (timeloop.h:389-393)
      const double cp[3] = {d.pos[3 * v] + u0, d.pos[3 * v + 1] + u1, d.pos[3 * v + 2] + u2};
      for (int k = 0; k < 3; k++) {
        if (v == 0) o[k] = o[3 + k] = cp[k];
        o[k] = (cp[k] < o[k]) ? cp[k] : o[k];
        o[3 + k] = (o[3 + k] < cp[k]) ? cp[k] : o[3 + k];
      }
    }
    for (uint32_t q = 0; q < d.nclamped; q++) B->nflag[R.poff[i] + d.clamped[q]] |= 1;
    for (int k = 0; k < 6; k++) B->lrtb[6 * i + k] = o[k];
    // fast-path domain of the bond arithmetic: reference bond lengths
    // (displacements are checked by the position kernel)
    for (uint64_t g = R.poff[i]; g < R.poff[i] + d.nnodes && !slow; g++)
      for (uint64_t k = R.rstart[g]; k < R.rstart[g + 1]; k++) {
        const uint64_t q = R.rows[k] - R.poff[i], v = g - R.poff[i];
        const double x0 = d.pos[3 * q] - d.pos[3 * v];
        const double x1 = d.pos[3 * q + 1] - d.pos[3 * v + 1];
        const double x2 = d.pos[3 * q + 2] - d.pos[3 * v + 2];
        const double l2 = (x0 * x0 + x1 * x1) + x2 * x2;
        if (!(l2 >= kFastLo && l2 <= kFastHi)) {
          slow = true;
          break;
        }
      }
    pfl[i] = slow;
  });
  for (uint32_t i = 0; i < np; i++)
    if (pfl[i]) B->fl = kFlagSafeMath;
  // bond slices, entry k of lane l at soff[s] + 64k + l
  B->soff.resize(R.nslices);
  B->bond.assign(R.entries, kPad);
  uint64_t at = 0;
  for (uint64_t s = 0; s < R.nslices; s++) {
    B->soff[s] = at;
    at += (uint64_t)R.width[s] * kSlice;
  }
  parallel_for(R.nslices, work, [&](uint64_t s) {
    for (uint64_t t = s * kSlice; t < (s + 1) * kSlice; t++) {
      const uint64_t g = R.node_of(t);
      uint32_t* col = B->bond.data() + B->soff[s] + (t % kSlice);
      for (uint64_t k = R.rstart[g]; k < R.rstart[g + 1]; k++) col[(k - R.rstart[g]) * kSlice] = R.rows[k];
    }
  });
  // node blocks: the sorted set of nodes each block's rows reference (its
  // own nodes included) and every entry's position in it
  const uint64_t nblk = R.nslices / kBlockSlices, bn = (uint64_t)kBlockSlices * kSlice;
  B->bloc.assign(R.entries, (uint16_t)kPad16);
  B->btab.assign(nblk * kLdsCap, 0);
  B->binfo.assign(nblk, make_uint2(0, 0));
  std::vector<uint8_t> over(nblk, 0);
  parallel_for(nblk, work, [&](uint64_t b) {
    const uint64_t n0 = b * bn, n1 = std::min<uint64_t>(R.nodes, n0 + bn);
    if (n0 >= n1) return;
    // per worker thread: a mark per job node (the block being built) and the
    // table position of each marked node
    thread_local std::vector<uint32_t> mark;
    thread_local std::vector<uint32_t> where;
    thread_local uint32_t epoch = 0;
    if (mark.size() < R.nodes || epoch == ~0u) {
      mark.assign(std::max<uint64_t>(R.nodes, mark.size()), 0);
      where.resize(mark.size());
      epoch = 0;
    }
    const uint32_t tag = ++epoch;
    std::vector<uint32_t> t;
    t.reserve(kLdsCap + bn);
    auto add = [&](uint32_t v) {
      if (mark[v] != tag) {
        mark[v] = tag;
        t.push_back(v);
      }
    };
    for (uint64_t g = n0; g < n1; g++) {
      add((uint32_t)g);
      for (uint64_t k = R.rstart[g]; k < R.rstart[g + 1]; k++) add(R.rows[k]);
    }
    if (t.size() > kLdsCap) {
      over[b] = 1;
      return;
    }
    std::sort(t.begin(), t.end());
    for (uint32_t i = 0; i < t.size(); i++) where[t[i]] = i;
    auto pos = [&](uint32_t v) { return where[v]; };
    for (uint64_t g = n0; g < n1; g++) {
      const uint64_t t = (g & ~(uint64_t)255) + R.node_lane[g];
      uint16_t* col = B->bloc.data() + B->soff[t / kSlice] + (t % kSlice);
      for (uint64_t k = R.rstart[g]; k < R.rstart[g + 1]; k++)
        col[(k - R.rstart[g]) * kSlice] = (uint16_t)(16 * pos(R.rows[k]));
    }
    std::copy(t.begin(), t.end(), B->btab.begin() + b * kLdsCap);
    B->binfo[b] = make_uint2((uint32_t)t.size(), pos((uint32_t)n0));
  });
  for (uint64_t b = 0; b < nblk; b++) B->lds_ok &= !over[b];
  B->lane_node = std::move(R.lane_node);
  B->node_lane = std::move(R.node_lane);
  B->deg = std::move(R.deg);
  B->width = std::move(R.width);
  *out = B;
  return PD_OK;
}

int pd_job_blob_footprint(const pd_job_blob* B, uint64_t* node_slots, uint64_t* bond_entries, uint64_t* bytes) {
  if (!B) return PD_E_ARG;
  if (node_slots) *node_slots = B->nslices * kSlice;
  if (bond_entries) *bond_entries = B->entries;
  if (bytes) *bytes = B->nslices * kSlice * kBytesPerNode + B->entries * kBytesPerEntry;
  return PD_OK;
}

void pd_job_blob_free(pd_job_blob* B) { delete B; }

int pd_add_prepared(pd_ctx* c, const pd_job_blob* B, uint64_t nsteps, uint32_t* job) {
  if (!c || !B) return PD_E_ARG;
  if (!c->reserved) return fail(c, PD_E_STATE, "pd_add_prepared before pd_reserve");
  int rc;
  if ((rc = check_run(c, B->run))) return rc;
  const uint32_t np = B->np;
  uint32_t jid = 0;
  while (jid < c->cap_jobs && c->jobs[jid].used) jid++;
  if (jid == c->cap_jobs)
    return fail(c, PD_E_NOMEM, "all %llu job slots are in use", (unsigned long long)c->cap_jobs);
  if (c->part_free.size() < np) return fail(c, PD_E_NOMEM, "not enough particle slots");
  if (B->nodes + c->cnt.nodes >= 0x7FFFFFFFull) return fail(c, PD_E_OVERFLOW, "more than 2^31-1 nodes");
  uint64_t s0 = 0, b0 = 0;
  if (!c->node_free.alloc(B->nslices, &s0))
    return fail(c, PD_E_NOMEM, "no free range of %llu node slices", (unsigned long long)B->nslices);
  if (!c->bond_free.alloc(B->entries, &b0)) {
    c->node_free.release(s0, B->nslices);
    return fail(c, PD_E_NOMEM, "no free range of %llu bond entries", (unsigned long long)B->entries);
  }
  Job& J = c->jobs[jid];
  J = Job();
  J.used = true;
  J.run = B->run;
  J.node0 = s0 * kSlice;
  J.nslices = B->nslices;
  J.bond0 = b0;
  J.nentries = B->entries;
  J.remaining = nsteps;
  J.nodes = B->nodes;
  J.movable_bonds = B->movable_bonds;
  J.total_bonds = B->total_bonds;
  J.lds_ok = B->lds_ok;
  if (!J.lds_ok) c->lds_bad++;
  const uint64_t g0 = J.node0, Np = B->nslices * kSlice;
  // place: particle slots, pool node indices, slice offsets
  std::vector<uint32_t> slot(np);
  for (uint32_t i = 0; i < np; i++) {
    slot[i] = c->part_free.back();
    c->part_free.pop_back();
    HostPart H = B->parts[i];
    H.off = g0 + B->poff[i];
    H.slot = slot[i];
    J.parts.push_back(std::move(H));
  }
  std::vector<uint32_t> pid(Np);
  const uint32_t vpart = (uint32_t)(c->d.P - 1);
  for (uint64_t r = 0; r < Np; r++) pid[r] = B->lpid[r] == ~0u ? vpart : slot[B->lpid[r]];
  std::vector<uint64_t> soff(B->nslices);
  for (uint64_t s = 0; s < B->nslices; s++) soff[s] = b0 + B->soff[s];
  std::vector<DevPart> dparts(B->dparts);
  for (uint32_t i = 0; i < np; i++) {
    dparts[i].node_off = g0 + B->poff[i];
    dparts[i].inst = jid;
    dparts[i].boff = g0 + B->bofs[i];
  }
  std::vector<uint32_t> blist(B->blist);
  for (auto& b : blist) b += (uint32_t)g0;
  DevInst I;
  memset(&I, 0, sizeof I);
  fill_inst(B->run, I);
  memcpy(I.wall, B->run.wall_lrtb, sizeof I.wall);
  memcpy(I.speed, B->run.wall_speed, sizeof I.speed);
  I.remaining = nsteps;
  I.node0 = g0;
  I.node1 = g0 + B->nodes;
  I.t = B->run.first_step ? B->run.first_step : 1;
  job_grid(B->lrtb.data(), np, B->run, I.cell, Np * kCellsPerSlot, I.gn);
  I.gtb = (uint32_t)(g0 * kCellsPerSlot);
  const double zero18[18] = {0};
  Dev& d = c->d;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // The job's arrays go out through the pinned stage when they fit
  // kStageJobMax (bounded by its footprint): one host copy each, then
  // asynchronous copies in stream order with the kernels between them and
  // one wait at the end; a larger job is copied synchronously.
  const size_t bound = (size_t)(Np * kBytesPerNode + B->entries * kBytesPerEntry) +
                       (size_t)np * (sizeof(DevPart) + 6 * sizeof(double) + 512) + sizeof(DevInst) + 64 * 256;
  const bool staged = bound <= kStageJobMax;
  if (staged && (rc = stage_reserve(c, bound))) return rc;
  size_t at = 0;
  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (!bytes) return PD_OK;
    if (staged && at + bytes <= c->stage_cap) {
      char* h = (char*)c->stage + at;
      memcpy(h, src, bytes);
      at += (bytes + 255) & ~(size_t)255;
      HIPCHK(c, hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, c->stream));
    } else {
      HIPCHK(c, hipStreamSynchronize(c->stream));   // after the copies and kernels before it
      HIPCHK(c, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    }
    return PD_OK;
  };
#define UP(dst, src, n) \
  if ((rc = up((dst), (src), (n) * sizeof(*(dst))))) return rc
  UP(d.X + g0, B->X.data(), Np);
  UP(d.U + g0, B->U.data(), Np);
  hipLaunchKernelGGL(k_split, dim3(grid_of(Np, 256)), dim3(256), 0, c->stream, d, g0, Np, 1);
  HIPCHK(c, hipGetLastError());
  UP(d.V + 3 * g0, B->V.data(), 3 * Np);
  UP(d.A + 3 * g0, B->A.data(), 3 * Np);
  UP(d.E + 3 * g0, B->E.data(), 3 * Np);
  HIPCHK(c, hipMemsetAsync(d.F + 3 * g0, 0, 3 * Np * sizeof(double), c->stream));
  UP(d.pid + g0, pid.data(), Np);
  UP(d.nflag + g0, B->nflag.data(), Np);
  UP(d.brank + g0, B->brank.data(), Np);
  if (!blist.empty()) UP(d.blist + g0, blist.data(), blist.size());
  UP(d.deg + g0, B->deg.data(), Np);
  UP(d.lane_node + g0, B->lane_node.data(), Np);
  UP(d.node_lane + g0, B->node_lane.data(), Np);
  UP(d.soff + s0, soff.data(), B->nslices);
  UP(d.swid + s0, B->width.data(), B->nslices);
  if (B->entries) {
    UP(d.bond + b0, B->bond.data(), B->entries);
    hipLaunchKernelGGL(k_rebase, dim3(grid_of(B->entries, 256)), dim3(256), 0, c->stream, d.bond + b0, B->entries,
                       (uint32_t)g0);
    HIPCHK(c, hipGetLastError());
    UP(d.bloc + b0, B->bloc.data(), B->entries);
  }
  {
    // block tables: job-relative -> pool node indices (the blocks start at
    // s0 / kBlockSlices, since every job occupies whole blocks)
    const uint64_t nb = B->nslices / kBlockSlices, t0 = s0 / kBlockSlices * (uint64_t)kLdsCap;
    UP(d.btab + t0, B->btab.data(), nb * kLdsCap);
    hipLaunchKernelGGL(k_rebase, dim3(grid_of(nb * kLdsCap, 256)), dim3(256), 0, c->stream, d.btab + t0,
                       nb * (uint64_t)kLdsCap, (uint32_t)g0);
    HIPCHK(c, hipGetLastError());
    UP(d.binfo + s0 / kBlockSlices, B->binfo.data(), nb);
  }
  for (uint32_t i = 0; i < np; i++) {
    UP(d.part + slot[i], &dparts[i], 1);
    UP(d.lrtb + 6 * slot[i], &B->lrtb[6 * i], 6);
  }
  UP(d.inst + jid, &I, 1);
  UP(d.reaction + 18 * jid, zero18, 18);
  HIPCHK(c, hipMemset(d.du2s + (uint64_t)jid * kDuSlots * kDuStride, 0,
                      kDuSlots * kDuStride * sizeof(unsigned long long)));
#undef UP
  HIPCHK(c, hipStreamSynchronize(c->stream));   // the stage is free again
  if (B->fl) {
    uint32_t cur = 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&cur, d.flags, sizeof cur, hipMemcpyDeviceToHost));
    cur |= B->fl;
    HIPCHK(c, hipMemcpy(d.flags, &cur, sizeof cur, hipMemcpyHostToDevice));
  }
  if ((rc = upload_pairs(c))) return rc;   // synchronises the stream
  recount(c);
  c->graph_dirty = true;
  if (job) *job = jid;
  return PD_OK;
}

int pd_add_job(pd_ctx* c, const pd_run_desc* run, const pd_particle_desc* P, uint32_t np, uint64_t nsteps,
               uint32_t* job) {
  if (!c || !run || !P || !np) return PD_E_ARG;
  if (!c->reserved) return fail(c, PD_E_STATE, "pd_add_job before pd_reserve");
  pd_job_blob* B = nullptr;
  int rc = pd_job_prepare(run, P, np, &B);
  if (rc) return fail(c, rc, "%s", t_err.c_str());
  rc = pd_add_prepared(c, B, nsteps, job);
  pd_job_blob_free(B);
  return rc;
}

int pd_remove_job(pd_ctx* c, uint32_t job) {
  if (!c) return PD_E_ARG;
  if (job >= c->jobs.size() || !c->jobs[job].used) return fail(c, PD_E_ARG, "no job %u", job);
  Job& J = c->jobs[job];
  Dev& d = c->d;
  hipStream_t st = c->stream;
  HIPCHK(c, hipStreamSynchronize(st));
  const uint64_t Np = J.nslices * kSlice;
  if (Np) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(Np, 256)), dim3(256), 0, st, d.pid + J.node0, Np,
                       (uint32_t)(d.P - 1));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemsetAsync(d.deg + J.node0, 0, Np * sizeof(uint32_t), st));
    HIPCHK(c, hipMemsetAsync(d.nflag + J.node0, 0, Np, st));
    HIPCHK(c, hipMemsetAsync(d.swid + J.node0 / kSlice, 0, J.nslices * sizeof(uint32_t), st));
  }
  if (J.nentries) HIPCHK(c, hipMemsetAsync(d.bond + J.bond0, 0xFF, J.nentries * sizeof(uint32_t), st));
  if (J.nentries) HIPCHK(c, hipMemsetAsync(d.bloc + J.bond0, 0xFF, J.nentries * sizeof(uint16_t), st));
  HIPCHK(c, hipMemsetAsync(d.binfo + J.node0 / kSlice / kBlockSlices, 0, J.nslices / kBlockSlices * sizeof(uint2), st));
  if (!J.lds_ok) c->lds_bad--;
  HIPCHK(c, hipStreamSynchronize(st));
  DevInst vi;
  void_inst(vi);
  HIPCHK(c, hipMemcpy(d.inst + job, &vi, sizeof vi, hipMemcpyHostToDevice));
  std::vector<DevPart> vps(J.parts.size());
  for (size_t i = 0; i < J.parts.size(); i++) {
    void_part(vps[i], (uint32_t)(d.ninst - 1));
    HIPCHK(c, hipMemcpy(d.part + J.parts[i].slot, &vps[i], sizeof(DevPart), hipMemcpyHostToDevice));
    c->part_free.push_back(J.parts[i].slot);
  }
  c->node_free.release(J.node0 / kSlice, J.nslices);
  c->bond_free.release(J.bond0, J.nentries);
  J = Job();
  int rc = upload_pairs(c);
  if (rc) return rc;
  recount(c);
  c->graph_dirty = true;
  return PD_OK;
}

int pd_commit(pd_ctx* c) {
  if (!c) return PD_E_ARG;
  if (c->committed) return fail(c, PD_E_STATE, "already committed");
  if (c->pending.empty()) return fail(c, PD_E_STATE, "no particles");
  int rc = check_run(c, c->run0);
  if (rc) return rc;
  // one job per instance value, particles in the order they were added
  std::vector<std::pair<size_t, size_t>> groups;
  for (size_t i = 0; i < c->pending.size(); i++)
    if (groups.empty() || c->pending[i].instance != c->pending[groups.back().first].instance)
      groups.push_back({i, i + 1});
    else
      groups.back().second = i + 1;
  uint64_t slots = 0, entries = 0;
  for (auto& gr : groups) {
    uint64_t ns = 0, ne = 0;
    if (pd_job_footprint(&c->pending[gr.first], (uint32_t)(gr.second - gr.first), &ns, &ne, nullptr))
      return fail(c, PD_E_ARG, "bad particle description");
    slots += ns;
    entries += ne;
  }
  if (slots >= 0x7FFFFFFFull) return fail(c, PD_E_OVERFLOW, "more than 2^31-1 nodes");
  if (!c->reserved) {
    pd_pool_desc pool{};
    pool.max_nodes = slots;
    pool.max_bond_entries = entries;
    pool.max_particles = c->pending.size();
    pool.max_jobs = groups.size();
    if ((rc = pd_reserve(c, &pool))) return rc;
  }
  for (auto& gr : groups) {
    uint32_t job = 0;
    if ((rc = pd_add_job(c, &c->run0, &c->pending[gr.first], (uint32_t)(gr.second - gr.first), kUnlimited,
                         &job)))
      return rc;
    for (size_t i = gr.first; i < gr.second; i++) c->plain.push_back({job, (uint32_t)(i - gr.first)});
  }
  c->pending.clear();
  c->keep_d.clear();
  c->keep_u.clear();
  c->committed = true;
  return PD_OK;
}

}  // extern "C"

// One explicit step of every resident job (timeloop.h:395-797).  Contains no
// memset and no host synchronisation unless profiling, so it can be captured.
// The pool as the step kernels see it: node slots up to the highest one in
// use, and the keys of the cell table that belong to them.
static Dev step_view(const pd_ctx* c) {
  Dev d = c->d;
  d.S = c->hwm_slices;
  d.N = d.S * kSlice;
  d.lists = any_contact_possible(c) ? 1u : 0u;
  d.T = (uint32_t)(d.N * kCellsPerSlot);
  return d;
}

// Node records of pools this large are addressed with 64-bit offsets.
static bool wide_pool(const pd_ctx* c) { return c->cap_slices * kSlice > kOff32MaxNodes; }

// The shape table of the context's arithmetic.
static const BondShape* shapes(const pd_ctx* c) {
  return c->arith == PD_ARITH_TOLERANCE ? bond_shapes_tol() : bond_shapes();
}

// A shape is usable if it is built in the context's arithmetic; the staged
// (LDS) shapes also need every resident job's block tables to fit.
static bool shape_usable(const pd_ctx* c, int v) {
  const BondShape& b = shapes(c)[v];
  return b.fn && (!b.lds || (c->lds_bad == 0 && !wide_pool(c)));
}

// The shape a launch uses: the selected one while usable, else the default.
// kWideShape (one past the tuned shapes) selects bond_wide on any pool.
static int effective_shape(const pd_ctx* c) {
  if (c->bond_shape == kWideShape) return kWideShape;
  int v = c->bond_shape < 0 ? kDefaultBondShape : c->bond_shape;
  if (c->small_auto && c->cnt.nodes < kQuadMaxNodes && shape_usable(c, kQuadShape)) v = kQuadShape;
  return shape_usable(c, v) ? v : kDefaultBondShape;
}

// Workgroups of a bond-kernel launch over Np node slots: a multiple of the
// XCD count (the bond kernels remap blocks), lanes() threads per slot.
static dim3 bond_grid(uint64_t Np, const BondShape& b) {
  return dim3((grid_of(Np, 256) * b.lanes() + kXcds - 1) / kXcds * kXcds);
}

static BondShape launch_shape(const pd_ctx* c) {
  const int v = effective_shape(c);
  return wide_pool(c) || v == kWideShape ? bond_wide() : shapes(c)[v];
}

// Phase times of one profiled step (waits for its last mark).
static int accumulate_marks(pd_ctx* c, hipEvent_t* ev) {
  if (c->prof == PD_PROFILE_BOND) {
    HIPCHK(c, hipEventSynchronize(ev[PD_PHASE_BOND + 1]));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, ev[PD_PHASE_BOND], ev[PD_PHASE_BOND + 1]));
    c->acc.ms[PD_PHASE_BOND] += ms;
    c->acc.launches[PD_PHASE_BOND] += 1;
    return PD_OK;
  }
  HIPCHK(c, hipEventSynchronize(ev[PD_NPHASE]));
  for (int k = 0; k < PD_NPHASE; k++) {
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
    c->acc.ms[k] += ms;
    c->acc.launches[k] += 1;
  }
  return PD_OK;
}

// One step's launches; with `ev` set, a mark is recorded before each phase
// and after the last (PD_PROFILE_BOND: only around the bond kernel).
static int launch_step(pd_ctx* c, int par, hipEvent_t* ev = nullptr) {
  Dev d = step_view(c);
  // the step reads the displacement of buffer par and writes buffer par ^ 1
  d.UXYo = c->uxy[par];
  d.UZo = c->uz[par];
  d.UXY = c->uxy[par ^ 1];
  d.UZ = c->uz[par ^ 1];
  hipStream_t st = c->stream;
  const uint64_t Np = d.S * kSlice;
  auto mark = [&](int k) -> int {
    if (ev && (c->prof == PD_PROFILE_PHASES || k == PD_PHASE_BOND || k == PD_PHASE_BOND + 1))
      HIPCHK(c, hipEventRecord(ev[k], st));
    return PD_OK;
  };
  int rc;
  if ((rc = mark(0))) return rc;
  hipLaunchKernelGGL(k_position, dim3(grid_of(Np, 256)), dim3(256), 0, st, d);
  if ((rc = mark(1))) return rc;
  // pair box test, wall prefilter and (for rebuilding jobs) cell keys
  hipLaunchKernelGGL(k_cellkey, dim3(grid_of(std::max({d.N, d.ninst, d.npairs, d.P}), 256)), dim3(256), 0, st, d);
  if (any_torque(c)) hipLaunchKernelGGL(k_centroid, dim3(d.P), dim3(64), 0, st, d);
  if (d.lists) {
    // every kernel of the list rebuild exits at once unless a job rebuilds
    const uint32_t nb = d.T / kScanBlock;
    hipLaunchKernelGGL(k_scan1, dim3(nb), dim3(kScanThreads), 0, st, d);
    hipLaunchKernelGGL(k_scan3, dim3(nb), dim3(kScanThreads), 0, st, d);
    hipLaunchKernelGGL(k_scatter, dim3(grid_of(d.N, 256)), dim3(256), 0, st, d);
    hipLaunchKernelGGL(k_build, dim3(grid_of(d.N, 256)), dim3(256), 0, st, d);
    if (any_bdry_only(c)) hipLaunchKernelGGL(k_build_bdry, dim3(grid_of(d.N, 256)), dim3(256), 0, st, d);
  }
  if ((rc = mark(2))) return rc;
  // the bond kernel sums each node's row, then adds the node's contact,
  // wall, external and torque terms and takes the Verlet velocity half;
  // grid padded to a multiple of the XCD count (the bond kernel remaps
  // blocks)
  const BondShape bs = launch_shape(c);
  hipLaunchKernelGGL(bs.fn, bond_grid(Np, bs), dim3(256), 0, st, d);
  if ((rc = mark(3))) return rc;
  if (any_reaction(c)) hipLaunchKernelGGL(k_reaction, dim3(d.ninst), dim3(64), 0, st, d);
  if ((rc = mark(4))) return rc;
  if (any_clamped(c)) hipLaunchKernelGGL(k_clamp, dim3(grid_of(Np, 256)), dim3(256), 0, st, d);
  if (any_vc(c)) {
    hipLaunchKernelGGL(k_pv_mean, dim3(d.ninst), dim3(64), 0, st, d);
    hipLaunchKernelGGL(k_pv_apply, dim3(grid_of(Np, 256)), dim3(256), 0, st, d);
  }
  // at least 64 blocks: after a rebuild they also clear the cell table
  hipLaunchKernelGGL(k_finalize, dim3(grid_of(std::max<uint64_t>({d.P, d.ninst, 64 * 256}), 256)), dim3(256), 0, st, d);
  if ((rc = mark(5))) return rc;
  HIPCHK(c, hipGetLastError());
  return PD_OK;
}

// Time every bond-kernel shape on the resident pool and keep the fastest.
// The timed launches are whole step launches (bond sums, breaks, contact
// terms, velocity half) writing into scratch copies of everything the kernel
// writes (forces, bond entries, velocities, accelerations, applied force,
// wall masks), so the state is untouched.  Pools too large for 32-bit record
// offsets have one shape only.
static int autotune(pd_ctx* c, uint32_t reps) {
  c->small_auto = false;
  if (wide_pool(c)) {
    c->bond_shape = kDefaultBondShape;
    return PD_OK;
  }
  Dev t = step_view(c);
  const uint64_t Np = t.S * kSlice;
  int rc = PD_OK;
  double *F = nullptr, *A = nullptr, *V = nullptr, *AFD = nullptr;
  uint32_t* B = nullptr;
  uint16_t* B16 = nullptr;
  uint8_t *WM = nullptr, *BF = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const uint64_t ne = std::max<uint64_t>(c->cap_entries, 1);
  auto release = [&]() {
    for (void* q : {(void*)F, (void*)B, (void*)B16, (void*)A, (void*)V, (void*)AFD, (void*)WM, (void*)BF}) {
      if (!q) continue;
      auto it = std::find(c->allocs.begin(), c->allocs.end(), q);
      if (it != c->allocs.end()) c->allocs.erase(it);
      (void)hipFree(q);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  auto run = [&]() -> int {
    if ((rc = dalloc(c, &F, 3 * Np)) || (rc = dalloc(c, &B, ne)) || (rc = dalloc(c, &B16, ne)) ||
        (rc = dalloc(c, &A, 3 * Np)) || (rc = dalloc(c, &V, 3 * Np)) ||
        (c->d.AFD && (rc = dalloc(c, &AFD, 3 * Np))) || (rc = dalloc(c, &WM, Np)) ||
        (rc = dalloc(c, &BF, Np / 256 + 1)))
      return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpyAsync(A, c->d.A, 3 * Np * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(V, c->d.V, 3 * Np * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    // the applied force exists once a prescribed velocity was set; only its
    // particle's nodes write it
    if (AFD) HIPCHK(c, hipMemcpyAsync(AFD, c->d.AFD, 3 * Np * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(WM, 0, Np, c->stream));
    HIPCHK(c, hipMemsetAsync(BF, 0, Np / 256 + 1, c->stream));
    if (c->cap_entries) {
      HIPCHK(c, hipMemcpyAsync(B, c->d.bond, c->cap_entries * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(B16, c->d.bloc, c->cap_entries * sizeof(uint16_t), hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    t.F = F;
    t.bond = B;
    t.bloc = B16;
    t.A = A;
    t.V = V;
    t.AFD = AFD;
    t.wmask = WM;
    t.bflag = BF;
    // mean time of `reps` launches of shape v
    auto timed = [&](int v, double* ms_out) -> int {
      const dim3 grid = bond_grid(Np, shapes(c)[v]);
      HIPCHK(c, hipEventRecord(e0, c->stream));
      for (uint32_t r = 0; r < reps; r++) hipLaunchKernelGGL(shapes(c)[v].fn, grid, dim3(256), 0, c->stream, t);
      HIPCHK(c, hipEventRecord(e1, c->stream));
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipEventSynchronize(e1));
      float ms = 0;
      HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
      *ms_out = ms / reps;
      return PD_OK;
    };
    // every usable shape once (a warm-up launch, then the timed ones); the
    // shapes within 15% of the fastest are timed again in three interleaved
    // rounds and keep their smallest time, so a clock still ramping up when
    // the first shapes ran does not decide.  The fastest wins; the default
    // shape if none is usable.
    // While the staged shapes are usable the unstaged ones (35-75% slower on
    // every workload measured) are represented by the default one alone.
    // The four-lanes-per-node shapes are for pools too small to fill the GPU
    // and are timed only on pools below the automatic tuning size.
    bool staged = false;
    for (int v = 0; v < kBondShapes; v++) staged |= shapes(c)[v].lds && shape_usable(c, v);
    double first_best = HUGE_VAL;
    for (int v = 0; v < kBondShapes; v++) {
      const bool quad = shapes(c)[v].lanes() > 1;
      if (!shape_usable(c, v) || (staged && !shapes(c)[v].lds && v != kDefaultBondShape && !quad) ||
          (quad && c->cnt.nodes >= kTuneMinNodes)) {
        c->tune_ms[v] = -1.0;   // not timed
        continue;
      }
      hipLaunchKernelGGL(shapes(c)[v].fn, bond_grid(Np, shapes(c)[v]), dim3(256), 0, c->stream, t);   // warm-up
      if ((rc = timed(v, &c->tune_ms[v]))) return rc;
      first_best = std::min(first_best, c->tune_ms[v]);
    }
    for (int round = 0; round < 3; round++)
      for (int v = 0; v < kBondShapes; v++) {
        if (c->tune_ms[v] < 0.0 || c->tune_ms[v] > 1.15 * first_best) continue;
        double ms = 0.0;
        if ((rc = timed(v, &ms))) return rc;
        c->tune_ms[v] = std::min(c->tune_ms[v], ms);
      }
    int best = kDefaultBondShape;
    double best_ms = HUGE_VAL;
    for (int v = 0; v < kBondShapes; v++)
      if (c->tune_ms[v] >= 0.0 && c->tune_ms[v] < best_ms) {
        best_ms = c->tune_ms[v];
        best = v;
      }
    c->bond_shape = best;
    return PD_OK;
  };
  rc = run();
  release();
  uint64_t nodes = 0;
  for (auto& J : c->jobs)
    if (J.used) nodes += J.nodes;
  c->tuned_nodes = nodes;
  c->graph_dirty = true;
  return rc;
}

static int drop_graph(pd_ctx* c) {
  if (c->gexec) {
    HIPCHK(c, hipGraphExecDestroy(c->gexec));
    c->gexec = nullptr;
  }
  if (c->graph) {
    HIPCHK(c, hipGraphDestroy(c->graph));
    c->graph = nullptr;
  }
  return PD_OK;
}

// Capture kGraphSteps steps into a graph; captured again whenever the launch
// sequence or its arguments change (jobs added or removed, switches).
static int ensure_graph(pd_ctx* c, int par) {
  if (c->gexec && !c->graph_dirty && c->graph_par == par) return PD_OK;
  int rc = drop_graph(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (int s = 0; s < kGraphSteps; s++) {
    rc = launch_step(c, par ^ (s & 1));
    if (rc) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
  }
  HIPCHK(c, hipStreamEndCapture(c->stream, &c->graph));
  HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
  c->graph_dirty = false;
  c->graph_par = par;   // kGraphSteps is even: a replay ends on the buffer it starts from
  return PD_OK;
}

static void set_current_disp(pd_ctx* c, int par) {
  c->ucur = par;
  c->d.UXY = c->uxy[par];
  c->d.UZ = c->uz[par];
  c->d.UXYo = c->uxy[par ^ 1];
  c->d.UZo = c->uz[par ^ 1];
}

// Launch nsteps steps from the current displacement buffer (graph replays
// and single launches), wait for them and read the control words.
static int run_steps(pd_ctx* c, uint32_t nsteps, DevCtl* ctl, double* ms_out) {
  int rc;
  // steps_left drives reaction_step(): the reaction is formed on the last
  // step of this call (and on each job's own last step)
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, &c->d.ctl->steps_left, (uint64_t)1, nsteps);
  HIPCHK(c, hipGetLastError());
  uint32_t done = 0;
  int par = c->ucur;
  if (c->prof) {
    // marks cannot be timed from inside a graph: profiled steps launch one
    // by one, and each step's marks are read kProfLag steps later, so the
    // host waits on work that finished long ago and the queue stays full
    for (; done < nsteps; done++, par ^= 1) {
      if (done >= (uint32_t)kProfLag && (rc = accumulate_marks(c, c->pev[(done - kProfLag) % kProfRing])))
        return rc;
      if ((rc = launch_step(c, par, c->pev[done % kProfRing]))) return rc;
    }
    for (uint32_t i = nsteps > (uint32_t)kProfLag ? nsteps - kProfLag : 0; i < nsteps; i++)
      if ((rc = accumulate_marks(c, c->pev[i % kProfRing]))) return rc;
  } else {
    if (c->use_graph && nsteps >= (uint32_t)kGraphSteps) {
      // a graph captured from the other buffer: one single step first
      if (c->gexec && !c->graph_dirty && c->graph_par != par) {
        if ((rc = launch_step(c, par))) return rc;
        par ^= 1;
        done++;
      }
      if (nsteps - done >= (uint32_t)kGraphSteps) {
        if ((rc = ensure_graph(c, par))) return rc;
        for (; done + kGraphSteps <= nsteps; done += kGraphSteps) HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
      }
    }
    for (; done < nsteps; done++, par ^= 1)
      if ((rc = launch_step(c, par))) return rc;
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *ms_out = ms;
  HIPCHK(c, hipMemcpy(ctl, c->d.ctl, sizeof *ctl, hipMemcpyDeviceToHost));
  return PD_OK;
}

// The pool halted inside a step: some candidate list was longer than the
// capacity.  The state is that of the end of the last complete step (the
// halted step's displacement went to the buffer that is not current); the
// step's box accumulators, pair flags, pair count and cell counts are
// cleared.  Jobs whose lists need more than kMaxCap entries are stopped
// there (PD_E_OVERFLOW names them, pd_job_remaining reports it); for the
// rest the lists grow and every job rebuilds its lists at its next step.
static int recover_overflow(pd_ctx* c, std::string* stopped) {
  std::vector<uint32_t> need(c->d.ninst);
  HIPCHK(c, hipMemcpy(need.data(), c->d.jneed, need.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  uint32_t grow = 0;
  int rc;
  for (uint32_t q = 0; q < c->jobs.size(); q++) {
    Job& J = c->jobs[q];
    if (!J.used || !need[q]) continue;
    if (need[q] > kMaxCap) {
      char buf[200];
      snprintf(buf, sizeof buf, "job %u stopped: a node has %u contact candidates within R_c + skin, more than %u",
               q, need[q], kMaxCap);
      J.failed = buf;
      J.remaining = 0;
      const uint64_t zero = 0;
      HIPCHK(c, hipMemcpy((char*)(c->d.inst + q) + offsetof(DevInst, remaining), &zero, sizeof zero,
                          hipMemcpyHostToDevice));
      if (!stopped->empty()) *stopped += "; ";
      *stopped += buf;
    } else {
      grow = std::max(grow, need[q]);
    }
  }
  uint32_t relist = 0;
  if (grow > c->d.ccap) {
    const uint32_t ncap = std::min<uint32_t>(kMaxCap, std::max<uint32_t>(grow, (grow + grow / 4 + 15) / 16 * 16));
    uint32_t* nl = nullptr;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = dalloc(c, &nl, c->d.N * (uint64_t)ncap))) return rc;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)c->d.clist);
    if (it != c->allocs.end()) c->allocs.erase(it);
    HIPCHK(c, hipFree(c->d.clist));
    c->d.clist = nl;
    c->d.ccap = ncap;
    c->list_regrows++;
    relist = 1;
    c->graph_dirty = true;
  }
  hipLaunchKernelGGL(k_rollback, dim3(256), dim3(256), 0, c->stream, c->d, c->d.T, relist);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PD_OK;
}

extern "C" {

int pd_step(pd_ctx* c, uint32_t nsteps) {
  if (!c) return PD_E_ARG;
  if (!c->reserved) return fail(c, PD_E_STATE, "pd_step before pd_commit / pd_reserve");
  if (nsteps == 0) return PD_OK;
  bool any = false;
  for (auto& J : c->jobs) any |= J.used && J.remaining > 0;
  if (!any) {
    // a plain run whose job was stopped keeps reporting why
    if (c->committed)
      for (auto& J : c->jobs)
        if (J.used && !J.failed.empty()) return fail(c, PD_E_OVERFLOW, "%s", J.failed.c_str());
    return PD_OK;
  }
  int rc;
  if (c->bond_shape < 0) {
    // small pools: the launch is too short for shape to matter more than
    // the tuning itself; keep the shape that is fastest on large pools
    if (c->cnt.nodes >= kTuneMinNodes) {
      if ((rc = autotune(c, 3))) return rc;
    } else {
      c->bond_shape = kSmallPoolShape;   // kDefaultBondShape while it is not usable
      c->small_auto = true;              // kQuadShape while the pool is below kQuadMaxNodes
    }
  }
  uint32_t left = nsteps;
  double total_ms = 0;
  std::string stopped;
  while (left) {
    DevCtl ctl{};
    double ms = 0;
    const int par0 = c->ucur;
    if ((rc = run_steps(c, left, &ctl, &ms))) return rc;
    total_ms += ms;
    const uint32_t done = ctl.halt ? left - ctl.steps_left : left;
    set_current_disp(c, par0 ^ (int)(done & 1));
    for (auto& J : c->jobs) {
      if (!J.used || !J.remaining) continue;
      const uint64_t k = std::min<uint64_t>(J.remaining, done);
      if (J.remaining != kUnlimited) J.remaining -= k;
      J.steps_done += k;
      J.steps_since_set += k;
      J.reaction_valid = J.run.compute_wall_reaction && J.run.allow_wall;
      c->cnt.bond_updates += J.movable_bonds * k;
    }
    c->cnt.steps += done;
    left -= done;
    if (!ctl.halt) break;
    const uint32_t cap0 = c->d.ccap;
    const size_t nstop = stopped.size();
    if ((rc = recover_overflow(c, &stopped))) return rc;
    if (c->d.ccap == cap0 && stopped.size() == nstop)
      return fail(c, PD_E_OVERFLOW, "contact candidate lists could not grow (capacity %u)", cap0);
    bool live = false;
    for (auto& J : c->jobs) live |= J.used && J.remaining > 0;
    if (!live) break;
  }
  c->last_ms = total_ms;
  if (!stopped.empty()) return fail(c, PD_E_OVERFLOW, "%s", stopped.c_str());
  return PD_OK;
}

int pd_bond_shape_count(void) { return kBondShapes; }

int pd_set_arithmetic(pd_ctx* c, int mode) {
  if (!c) return PD_E_ARG;
  if (mode != PD_ARITH_BITWISE && mode != PD_ARITH_TOLERANCE) return fail(c, PD_E_ARG, "no arithmetic mode %d", mode);
  if (mode != c->arith) {
    c->arith = mode;
    c->bond_shape = -1;   // the other table's shapes: tuned again at the next step
    c->small_auto = false;
    c->graph_dirty = true;
  }
  return PD_OK;
}

int pd_get_arithmetic(pd_ctx* c, int* mode) {
  if (!c || !mode) return PD_E_ARG;
  *mode = c->arith;
  return PD_OK;
}

int pd_bond_shape_info(int shape, int* unroll, int* pipelined, int* staged) {
  if (shape < 0 || shape > kWideShape) return PD_E_ARG;
  const BondShape b = shape == kWideShape ? bond_wide() : bond_shapes()[shape];
  if (unroll) *unroll = b.unroll;
  if (pipelined) *pipelined = b.pipe;
  if (staged) *staged = b.lds;
  return PD_OK;
}

int pd_set_bond_shape(pd_ctx* c, int shape) {
  if (!c) return PD_E_ARG;
  if (shape < -1 || shape > kWideShape) return fail(c, PD_E_ARG, "no bond shape %d", shape);
  c->bond_shape = shape;
  c->small_auto = false;
  c->graph_dirty = true;
  return PD_OK;
}

int pd_autotune(pd_ctx* c, uint32_t reps, int* best, double* ms) {
  if (!c) return PD_E_ARG;
  if (!c->reserved) return fail(c, PD_E_STATE, "pd_autotune before pd_commit / pd_reserve");
  int rc = autotune(c, reps ? reps : 3);
  if (rc) return rc;
  if (best) *best = c->bond_shape;
  if (ms)
    for (int v = 0; v < kBondShapes; v++) ms[v] = c->tune_ms[v];
  return PD_OK;
}

int pd_get_bond_shape(pd_ctx* c, int* shape) {
  if (!c || !shape) return PD_E_ARG;
  *shape = c->bond_shape < 0 ? -1 : effective_shape(c);
  return PD_OK;
}

int pd_set_graphs(pd_ctx* c, int on) {
  if (!c) return PD_E_ARG;
  c->use_graph = on != 0;
  return PD_OK;
}

int pd_job_remaining(pd_ctx* c, uint32_t job, uint64_t* remaining) {
  if (!c || !remaining) return PD_E_ARG;
  if (job >= c->jobs.size() || !c->jobs[job].used) return fail(c, PD_E_ARG, "no job %u", job);
  *remaining = c->jobs[job].remaining;
  if (!c->jobs[job].failed.empty()) return fail(c, PD_E_OVERFLOW, "%s", c->jobs[job].failed.c_str());
  return PD_OK;
}

}  // extern "C"

// read-modify-write of one job's device parameters
template <typename F>
static int edit_inst(pd_ctx* c, uint32_t job, F f) {
  DevInst D;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(&D, c->d.inst + job, sizeof D, hipMemcpyDeviceToHost));
  f(D);
  HIPCHK(c, hipMemcpy(c->d.inst + job, &D, sizeof D, hipMemcpyHostToDevice));
  return PD_OK;
}

static int locate(pd_ctx* c, uint32_t job, uint32_t p, HostPart** H) {
  if (job >= c->jobs.size() || !c->jobs[job].used || p >= c->jobs[job].parts.size())
    return fail(c, PD_E_ARG, "no particle %u in job %u", p, job);
  *H = &c->jobs[job].parts[p];
  return PD_OK;
}

// plain particle index -> (job, particle of the job)
static int plain(pd_ctx* c, uint32_t pi, uint32_t* job, uint32_t* p) {
  if (!c->committed) return fail(c, PD_E_STATE, "before pd_commit");
  if (pi >= c->plain.size()) return fail(c, PD_E_ARG, "no particle %u", pi);
  *job = c->plain[pi].first;
  *p = c->plain[pi].second;
  return PD_OK;
}

static void apply_flag(pd_run_desc& r, int flag, int value) {
  switch (flag) {
    case PD_FLAG_ENABLE_FRACTURE: r.enable_fracture = value; break;
    case PD_FLAG_SELF_CONTACT: r.self_contact = value; break;
    case PD_FLAG_ALLOW_DAMPING: r.allow_damping = value; break;
    case PD_FLAG_ALLOW_FRICTION: r.allow_friction = value; break;
    case PD_FLAG_ENABLE_TORQUE: r.enable_torque = value; break;
    case PD_FLAG_COMPUTE_WALL_REACTION: r.compute_wall_reaction = value; break;
  }
}

static int job_set_flag(pd_ctx* c, uint32_t j, int flag, int value) {
  Job& J = c->jobs[j];
  if (flag == PD_FLAG_VELOCITY_CONSTRAINT) {
    J.vc_on = value != 0;
    c->graph_dirty = true;
    return edit_inst(c, j, [&](DevInst& D) { D.vc_on = value != 0; });
  }
  pd_run_desc r = J.run;
  apply_flag(r, flag, value);
  int rc = check_run(c, r);
  if (rc) return rc;
  J.run = r;
  rc = edit_inst(c, j, [&](DevInst& D) {
    D.enable_fracture = r.enable_fracture;
    D.self_contact = r.self_contact;
    D.allow_damping = r.allow_damping;
    D.allow_friction = r.allow_friction;
    D.compute_reaction = r.compute_wall_reaction;
    D.enable_torque = r.enable_torque;
    D.dirty = 1;   // self-contact eligibility may have changed
  });
  c->graph_dirty = true;
  return rc;
}

static int edit_part(pd_ctx* c, uint32_t job, uint32_t p, int which, int value, double snot, bool set_snot) {
  HostPart* H;
  int rc = locate(c, job, p, &H);
  if (rc) return rc;
  Job& J = c->jobs[job];
  if (!set_snot) {
    const bool was = H->movable != 0;
    switch (which) {
      case PD_PFLAG_MOVABLE: H->movable = value; break;
      case PD_PFLAG_STOPPABLE: H->stoppable = value; break;
      case PD_PFLAG_BREAKABLE: H->breakable = value; break;
      default: return fail(c, PD_E_ARG, "unknown particle flag %d", which);
    }
    const bool now = H->movable != 0;
    if (was && !now) J.movable_bonds -= H->nbonds;
    if (!was && now) J.movable_bonds += H->nbonds;
  } else {
    H->snot = snot;
  }
  DevPart D;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(&D, c->d.part + H->slot, sizeof D, hipMemcpyDeviceToHost));
  D.movable = H->movable;
  D.stoppable = H->stoppable;
  D.breakable = H->breakable;
  D.snot = H->snot;
  HIPCHK(c, hipMemcpy(c->d.part + H->slot, &D, sizeof D, hipMemcpyHostToDevice));
  return edit_inst(c, job, [](DevInst& I) { I.dirty = 1; });
}

extern "C" {

int pd_set_flag(pd_ctx* c, int flag, int value) {
  if (!c) return PD_E_ARG;
  if (flag < PD_FLAG_ENABLE_FRACTURE || flag > PD_FLAG_VELOCITY_CONSTRAINT)
    return fail(c, PD_E_ARG, "unknown flag %d", flag);
  for (uint32_t j = 0; j < c->jobs.size(); j++)
    if (c->jobs[j].used) {
      int rc = job_set_flag(c, j, flag, value);
      if (rc) return rc;
    }
  apply_flag(c->run0, flag, value);
  return PD_OK;
}

int pd_job_set_flag(pd_ctx* c, uint32_t job, int flag, int value) {
  if (!c) return PD_E_ARG;
  if (flag < PD_FLAG_ENABLE_FRACTURE || flag > PD_FLAG_VELOCITY_CONSTRAINT)
    return fail(c, PD_E_ARG, "unknown flag %d", flag);
  if (job >= c->jobs.size() || !c->jobs[job].used) return fail(c, PD_E_ARG, "no job %u", job);
  return job_set_flag(c, job, flag, value);
}

int pd_set_wall_speed(pd_ctx* c, const double speed[6]) {
  if (!c || !speed) return PD_E_ARG;
  for (uint32_t j = 0; j < c->jobs.size(); j++)
    if (c->jobs[j].used) {
      int rc = edit_inst(c, j, [&](DevInst& D) { memcpy(D.speed, speed, sizeof D.speed); });
      if (rc) return rc;
      memcpy(c->jobs[j].run.wall_speed, speed, 6 * sizeof(double));
    }
  memcpy(c->run0.wall_speed, speed, 6 * sizeof(double));
  return PD_OK;
}

int pd_set_wall(pd_ctx* c, const double lrtb[6]) {
  if (!c || !lrtb) return PD_E_ARG;
  for (uint32_t j = 0; j < c->jobs.size(); j++)
    if (c->jobs[j].used) {
      int rc = edit_inst(c, j, [&](DevInst& D) { memcpy(D.wall, lrtb, sizeof D.wall); });
      if (rc) return rc;
    }
  memcpy(c->run0.wall_lrtb, lrtb, 6 * sizeof(double));
  return PD_OK;
}

int pd_set_velocity_constraint(pd_ctx* c, uint32_t pi, int fix_mean_x, double mean_x, int fix_angular,
                               double angular_vel) {
  if (!c) return PD_E_ARG;
  uint32_t job, lp;
  int rc = plain(c, pi, &job, &lp);
  if (rc) return rc;
  if (!c->d.AFD) {
    if ((rc = dalloc(c, &c->d.AFD, 3 * c->d.N))) return rc;
    HIPCHK(c, hipMemsetAsync(c->d.AFD, 0, 3 * c->d.N * sizeof(double), c->stream));
  }
  Job& J = c->jobs[job];
  J.vc_part = lp;
  const uint32_t slot = J.parts[lp].slot;
  c->graph_dirty = true;
  return edit_inst(c, job, [&](DevInst& D) {
    D.vc_slot = slot;
    D.vc_fix_mean_x = fix_mean_x != 0;
    D.vc_mean_x = mean_x;
    D.vc_fix_ang = fix_angular != 0;
    D.vc_omega = angular_vel;
  });
}

int pd_get_applied_force(pd_ctx* c, uint32_t pi, double* afd) {
  if (!c || !afd) return PD_E_ARG;
  uint32_t job, lp;
  int rc = plain(c, pi, &job, &lp);
  if (rc) return rc;
  if (c->jobs[job].vc_part != (int64_t)lp || !c->d.AFD)
    return fail(c, PD_E_ARG, "particle %u has no prescribed velocity", pi);
  const HostPart& H = c->jobs[job].parts[lp];
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(afd, c->d.AFD + 3 * H.off, 3ull * H.n * sizeof(double), hipMemcpyDeviceToHost));
  return PD_OK;
}

int pd_set_particle_flag(pd_ctx* c, uint32_t pi, int which, int value) {
  if (!c) return PD_E_ARG;
  if (!c->committed) {
    if (pi >= c->pending.size()) return fail(c, PD_E_ARG, "no particle %u", pi);
    pd_particle_desc& q = c->pending[pi];
    switch (which) {
      case PD_PFLAG_MOVABLE: q.movable = value; break;
      case PD_PFLAG_STOPPABLE: q.stoppable = value; break;
      case PD_PFLAG_BREAKABLE: q.breakable = value; break;
      default: return fail(c, PD_E_ARG, "unknown particle flag %d", which);
    }
    return PD_OK;
  }
  uint32_t j, p;
  int rc = plain(c, pi, &j, &p);
  return rc ? rc : edit_part(c, j, p, which, value, 0, false);
}

int pd_set_snot(pd_ctx* c, uint32_t pi, double snot) {
  if (!c) return PD_E_ARG;
  if (!c->committed) {
    if (pi >= c->pending.size()) return fail(c, PD_E_ARG, "no particle %u", pi);
    c->pending[pi].snot = snot;
    return PD_OK;
  }
  uint32_t j, p;
  int rc = plain(c, pi, &j, &p);
  return rc ? rc : edit_part(c, j, p, 0, 0, snot, true);
}

int pd_reset_partzero_y(pd_ctx* c, double wheel_rad) {
  if (!c) return PD_E_ARG;
  uint32_t job, lp;
  int rc = plain(c, 0, &job, &lp);
  if (rc) return rc;
  Job& J = c->jobs[job];
  if (lp != 0 || J.parts.size() < 2 || J.parts[1].n < 2)
    return fail(c, PD_E_ARG, "reset_partzero_y needs particle 0 and a particle 1 of at least two nodes");
  if (!(wheel_rad >= 0.0)) return fail(c, PD_E_ARG, "wheel_rad must be >= 0");
  HostPart& W = J.parts[0];
  const HostPart& B = J.parts[1];
  // the move's bits come back through the last counter slot (unused by steps)
  double* out = (double*)(c->d.counters + 7);
  hipLaunchKernelGGL(k_partzero, dim3(1), dim3(kResetThreads), 0, c->stream, c->d, W.slot, B.off,
                     J.node0 + J.nodes, wheel_rad, J.run.contact_rad, out);
  HIPCHK(c, hipGetLastError());
  double move = 0.0;
  HIPCHK(c, hipMemcpyAsync(&move, out, sizeof move, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < W.n; i++) W.pos[3 * (size_t)i + 1] += move;
  // particle 0 jumped: its job's candidate lists are rebuilt in the next step
  return edit_inst(c, job, [](DevInst& I) { I.dirty = 1; });
}

// Pinned host staging for read-backs, grown on demand: device-to-host copies
// into pinned memory run at the link's rate without the runtime's bounce
// buffers, and all of a call's copies go out before one wait.
static int stage_reserve(pd_ctx* c, size_t bytes) {
  if (bytes <= c->stage_cap) return PD_OK;
  const size_t cap = std::max<size_t>(bytes, 2 * c->stage_cap);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->stage) {
    HIPCHK(c, hipHostFree(c->stage));
    c->stage = nullptr;
    c->stage_cap = 0;
  }
  const hipError_t e = hipHostMalloc((void**)&c->stage, cap, hipHostMallocDefault);
  if (e != hipSuccess) {
    c->stage = nullptr;
    return fail(c, e == hipErrorOutOfMemory ? PD_E_NOMEM : PD_E_HIP, "hipHostMalloc(%zu) failed: %s", cap,
                hipGetErrorString(e));
  }
  c->stage_cap = cap;
  return PD_OK;
}

int pd_job_get_state(pd_ctx* c, uint32_t job, uint32_t pi, double* disp, double* currpos, double* vel,
                     double* acc, double* force) {
  if (!c) return PD_E_ARG;
  HostPart* Hp;
  int rc = locate(c, job, pi, &Hp);
  if (rc) return rc;
  const HostPart& p = *Hp;
  const Job& J = c->jobs[job];
  // every requested array goes to the pinned stage in one stream of copies:
  // U and X (double4 per node) for disp / CurrPos, then V, A, F (3 per node)
  const bool pos = disp || currpos;
  const size_t n4 = (size_t)p.n * sizeof(double4), n3 = 3ull * p.n * sizeof(double);
  if ((rc = stage_reserve(c, (pos ? 2 * n4 : 0) + 3 * n3))) return rc;
  const double4* U = (const double4*)c->stage;
  const double4* X = (const double4*)(c->stage + n4);
  char* rest = c->stage + (pos ? 2 * n4 : 0);
  if (pos) {
    // the displacement lives in the split copies; U is its transfer form
    hipLaunchKernelGGL(k_join, dim3(grid_of(p.n, 256)), dim3(256), 0, c->stream, c->d, p.off, (uint64_t)p.n);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->stage, c->d.U + p.off, n4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->stage + n4, c->d.X + p.off, n4, hipMemcpyDeviceToHost, c->stream));
  }
  double* const dst[3] = {vel, acc, force};
  const double* const src[3] = {c->d.V, c->d.A, c->d.F};
  for (int a = 0; a < 3; a++)
    if (dst[a])
      HIPCHK(c, hipMemcpyAsync(rest + a * n3, src[a] + 3 * p.off, n3, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int a = 0; a < 3; a++)
    if (dst[a]) memcpy(dst[a], rest + a * n3, n3);
  if (pos) {
    for (uint32_t i = 0; i < p.n; i++) {
      const double u[3] = {U[i].x, U[i].y, U[i].z};
      const double x[3] = {X[i].x, X[i].y, X[i].z};
      for (int k = 0; k < 3; k++) {
        if (disp) disp[3 * i + k] = u[k];
        if (currpos) currpos[3 * i + k] = x[k] + u[k];
      }
    }
    if (currpos && J.steps_since_set == 1) {
      // CurrPos of a clamped node is pos + the disp of the position half,
      // formed before the clamp reset disp (timeloop.h:539,767-774).  From
      // the second step on that disp is 0 + (dt*0 + h*0); only the first step
      // after the state was set sees the initial disp/vel/acc.
      const double dt = J.run.dt, h = (dt * dt) * 0.5;
      for (size_t q = 0; q < p.clamped.size(); q++) {
        const uint32_t i = p.clamped[q];
        for (int k = 0; k < 3; k++) {
          double u = p.cl0[9 * q + k] + ((dt * p.cl0[9 * q + 3 + k]) + (h * p.cl0[9 * q + 6 + k]));
          currpos[3 * i + k] = p.pos[3 * i + k] + u;
        }
      }
    }
  }
  return PD_OK;
}

int pd_get_state(pd_ctx* c, uint32_t pi, double* disp, double* currpos, double* vel, double* acc,
                 double* force) {
  if (!c) return PD_E_ARG;
  uint32_t j, p;
  int rc = plain(c, pi, &j, &p);
  return rc ? rc : pd_job_get_state(c, j, p, disp, currpos, vel, acc, force);
}

int pd_set_state(pd_ctx* c, uint32_t pi, const double* disp, const double* vel, const double* acc) {
  if (!c) return PD_E_ARG;
  uint32_t job, lp;
  int rc = plain(c, pi, &job, &lp);
  if (rc) return rc;
  HostPart& p = c->jobs[job].parts[lp];
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (disp) {
    std::vector<double4> U(p.n);
    for (uint32_t i = 0; i < p.n; i++) U[i] = make_double4(disp[3 * i], disp[3 * i + 1], disp[3 * i + 2], 0.0);
    HIPCHK(c, hipMemcpy(c->d.U + p.off, U.data(), p.n * sizeof(double4), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_split, dim3(grid_of(p.n, 256)), dim3(256), 0, c->stream, c->d, p.off, (uint64_t)p.n, 0);
    HIPCHK(c, hipGetLastError());
    // a resumed run starts from lrtb = min/max of pos + disp (timeloop.h:389-393)
    double o[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t v = 0; v < p.n; v++)
      for (int k = 0; k < 3; k++) {
        const double cp = p.pos[3 * v + k] + disp[3 * v + k];
        if (v == 0) o[k] = o[3 + k] = cp;
        o[k] = (cp < o[k]) ? cp : o[k];
        o[3 + k] = (o[3 + k] < cp) ? cp : o[3 + k];
      }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d.lrtb + 6 * (uint64_t)p.slot, o, sizeof o, hipMemcpyHostToDevice));
  }
  if (vel) HIPCHK(c, hipMemcpy(c->d.V + 3 * p.off, vel, 3ull * p.n * sizeof(double), hipMemcpyHostToDevice));
  if (acc) HIPCHK(c, hipMemcpy(c->d.A + 3 * p.off, acc, 3ull * p.n * sizeof(double), hipMemcpyHostToDevice));
  for (size_t q = 0; q < p.clamped.size(); q++)
    for (int k = 0; k < 3; k++) {
      const uint32_t i = p.clamped[q];
      if (disp) p.cl0[9 * q + k] = disp[3 * i + k];
      if (vel) p.cl0[9 * q + 3 + k] = vel[3 * i + k];
      if (acc) p.cl0[9 * q + 6 + k] = acc[3 * i + k];
    }
  c->jobs[job].steps_since_set = 0;
  return edit_inst(c, job, [](DevInst& I) { I.dirty = 1; });
}

int pd_job_get_connectivity(pd_ctx* c, uint32_t job, uint32_t pi, uint32_t* pairs, uint64_t cap, uint64_t* m) {
  if (!c || !m) return PD_E_ARG;
  HostPart* Hp;
  int rc = locate(c, job, pi, &Hp);
  if (rc) return rc;
  const HostPart& p = *Hp;
  // NbdArr2conn (timeloop.h:20-44) on the device: pairs (i, q) with i < q in
  // row order, compacted into a buffer of the context (grown to the largest
  // particle: its entries as pairs, then per node the pair offsets and the
  // total) and copied through the pinned stage into the caller's buffer
  const uint64_t need = p.nbonds + ((uint64_t)p.n + 2) / 2 + 1;   // in pairs
  if (need > c->conn_cap) {
    if (c->conn_buf) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)c->conn_buf);
      if (it != c->allocs.end()) c->allocs.erase(it);
      HIPCHK(c, hipFree(c->conn_buf));
      c->conn_buf = nullptr;
      c->conn_cap = 0;
    }
    if ((rc = dalloc(c, &c->conn_buf, 2 * need))) return rc;
    c->conn_cap = need;
  }
  uint32_t* offs = c->conn_buf + 2 * std::max<uint64_t>(p.nbonds, 1);   // p.n + 1 entries
  hipLaunchKernelGGL(k_conn_count, dim3(grid_of(p.n, 256)), dim3(256), 0, c->stream, c->d, p.off, p.n, offs);
  hipLaunchKernelGGL(k_conn_scan, dim3(1), dim3(kConnThreads), 0, c->stream, offs, p.n);
  hipLaunchKernelGGL(k_conn_write, dim3(grid_of(p.n, 256)), dim3(256), 0, c->stream, c->d, p.off, p.n, offs,
                     c->conn_buf);
  HIPCHK(c, hipGetLastError());
  if ((rc = stage_reserve(c, 2 * need * sizeof(uint32_t)))) return rc;
  // the count and, when the caller wants the pairs and can hold every
  // possible one (half the directed entries), all of them in one wait
  uint32_t* hs = (uint32_t*)c->stage;
  const uint64_t most = p.nbonds / 2;
  const bool whole = pairs && cap >= most && most;
  HIPCHK(c, hipMemcpyAsync(hs, offs + p.n, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  if (whole)
    HIPCHK(c, hipMemcpyAsync(hs + 2, c->conn_buf, 2ull * most * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const uint32_t n = hs[0];
  *m = n;
  if (pairs && cap < n)
    return fail(c, PD_E_ARG, "connectivity buffer holds %llu pairs, %llu needed", (unsigned long long)cap,
                (unsigned long long)n);
  if (pairs && n) {
    if (!whole) {
      HIPCHK(c, hipMemcpyAsync(hs + 2, c->conn_buf, 2ull * n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    memcpy(pairs, hs + 2, 2ull * n * sizeof(uint32_t));
  }
  return PD_OK;
}

int pd_get_connectivity(pd_ctx* c, uint32_t pi, uint32_t* pairs, uint64_t cap, uint64_t* m) {
  if (!c) return PD_E_ARG;
  uint32_t j, p;
  int rc = plain(c, pi, &j, &p);
  return rc ? rc : pd_job_get_connectivity(c, j, p, pairs, cap, m);
}

int pd_get_wall(pd_ctx* c, uint32_t job, double lrtb[6], double reaction[18]) {
  if (!c) return PD_E_ARG;
  if (!c->committed && !c->reserved) {
    if (lrtb) memcpy(lrtb, c->run0.wall_lrtb, 6 * sizeof(double));
    if (reaction) memset(reaction, 0, 18 * sizeof(double));
    return PD_OK;
  }
  if (job >= c->jobs.size() || !c->jobs[job].used) return fail(c, PD_E_ARG, "no job %u", job);
  DevInst D;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(&D, c->d.inst + job, sizeof D, hipMemcpyDeviceToHost));
  if (lrtb) memcpy(lrtb, D.wall, 6 * sizeof(double));
  if (reaction) {
    if (c->jobs[job].reaction_valid)
      HIPCHK(c, hipMemcpy(reaction, c->d.reaction + 18 * job, 18 * sizeof(double), hipMemcpyDeviceToHost));
    else
      memset(reaction, 0, 18 * sizeof(double));
  }
  return PD_OK;
}

int pd_get_counters(pd_ctx* c, pd_counters* out) {
  if (!c || !out) return PD_E_ARG;
  *out = c->cnt;
  if (c->reserved) {
    unsigned long long pw = 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&pw, c->d.counters, sizeof pw, hipMemcpyDeviceToHost));
    out->pair_computations = pw;
    unsigned long long rb = 0;
    HIPCHK(c, hipMemcpy(&rb, c->d.counters + 2, sizeof rb, hipMemcpyDeviceToHost));
    out->cell_rebuilds = rb;
    out->list_capacity = c->d.ccap;
    out->list_regrows = c->list_regrows;
    unsigned long long* dcount = c->d.counters + 1;
    HIPCHK(c, hipMemsetAsync(dcount, 0, sizeof(unsigned long long), c->stream));
    if (c->cap_entries) {
      hipLaunchKernelGGL(k_count_alive, dim3(std::min<uint64_t>(grid_of(c->cap_entries, 256), 4096)), dim3(256), 0,
                         c->stream, c->d, c->cap_entries, dcount);
      HIPCHK(c, hipGetLastError());
    }
    unsigned long long alive = 0;
    HIPCHK(c, hipMemcpyAsync(&alive, dcount, sizeof alive, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out->alive_bonds = alive;
  }
  return PD_OK;
}

int pd_set_profiling(pd_ctx* c, int on) {
  if (!c || on < 0 || on > PD_PROFILE_BOND) return PD_E_ARG;
  if (on && !c->pev[0][0])
    for (int s = 0; s < kProfRing; s++)
      for (int k = 0; k <= PD_NPHASE; k++) HIPCHK(c, hipEventCreate(&c->pev[s][k]));
  c->prof = on;
  c->acc = pd_profile{};
  return PD_OK;
}

int pd_get_profile(pd_ctx* c, pd_profile* out) {
  if (!c || !out) return PD_E_ARG;
  *out = c->acc;
  c->acc = pd_profile{};
  return PD_OK;
}

int pd_selftest_math(uint64_t n, uint64_t seed, uint64_t* mismatch_div, uint64_t* mismatch_sqrt) {
  if (!mismatch_div || !mismatch_sqrt) return PD_E_ARG;
  unsigned long long* dm = nullptr;
  if (hipMalloc((void**)&dm, 3 * sizeof(unsigned long long)) != hipSuccess) return PD_E_HIP;
  int rc = PD_OK;
  unsigned long long hm[3] = {0, 0, 0};
  if (hipMemset(dm, 0, 3 * sizeof(unsigned long long)) != hipSuccess) rc = PD_E_HIP;
  if (rc == PD_OK) {
    hipLaunchKernelGGL(k_selftest_math, dim3(2048), dim3(256), 0, 0, n, seed, dm);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(hm, dm, sizeof hm, hipMemcpyDeviceToHost) != hipSuccess)
      rc = PD_E_HIP;
  }
  (void)hipFree(dm);
  *mismatch_div = hm[0];
  *mismatch_sqrt = hm[1];
  return rc;
}

// ---------------------------------------------------------------------------
// The math self-test without device allocation.  Its counters live in the
// file-scope array `scratch` (one copy per GPU, each held until the process
// ends).  Host code never passes the array's name to a kernel: it takes the
// device address of the copy on the current GPU with hipGetSymbolAddress at
// the first call on that GPU and uses only that address afterwards.
constexpr int kScratchWords = 32;                     // 256 bytes; 3 are used
constexpr uint64_t kSelftestMaxSamples = 1ull << 40;  // per-thread counters are 32-bit
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];
static void* g_scratch_ptr[kScratchDevices];
static std::mutex g_scratch_mu;

static int scratch_on_current_device(unsigned long long** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) {
    t_err = "pd_selftest_math_noalloc: no usable current device";
    return PD_E_HIP;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      t_err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
      return PD_E_HIP;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = (unsigned long long*)g_scratch_ptr[dev];
  return PD_OK;
}

int pd_selftest_math_noalloc(uint64_t n, uint64_t seed, uint64_t* mismatch_div,
                             uint64_t* mismatch_sqrt, uint64_t* evaluated) {
  if (!mismatch_div || !mismatch_sqrt) {
    t_err = "pd_selftest_math_noalloc: null output pointer";
    return PD_E_ARG;
  }
  if (n > kSelftestMaxSamples) {
    char buf[160];
    snprintf(buf, sizeof buf, "pd_selftest_math_noalloc: n = %llu exceeds the maximum of %llu samples",
             (unsigned long long)n, (unsigned long long)kSelftestMaxSamples);
    t_err = buf;
    return PD_E_ARG;
  }
  unsigned long long* dm = nullptr;
  int rc = scratch_on_current_device(&dm);
  if (rc) return rc;
  // the array keeps the previous call's counts: clear them (fill kernel, no memset)
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, 0, (uint32_t*)dm, (uint64_t)6, 0u);
  hipLaunchKernelGGL(k_selftest_math, dim3(2048), dim3(256), 0, 0, n, seed, dm);
  unsigned long long hm[3] = {0, 0, 0};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(hm, dm, sizeof hm, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    t_err = std::string("pd_selftest_math_noalloc: ") + hipGetErrorString(e);
    return PD_E_HIP;
  }
  *mismatch_div = hm[0];
  *mismatch_sqrt = hm[1];
  if (evaluated) *evaluated = hm[2];
  return PD_OK;
}

int pd_math_mode(pd_ctx* c, int* safe) {
  if (!c || !safe) return PD_E_ARG;
  uint32_t fl = 0;
  if (c->reserved) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&fl, c->d.flags, sizeof fl, hipMemcpyDeviceToHost));
  }
  *safe = (fl & kFlagSafeMath) ? 1 : 0;
  return PD_OK;
}

int pd_last_step_ms(pd_ctx* c, double* ms) {
  if (!c || !ms) return PD_E_ARG;
  *ms = c->last_ms;
  return PD_OK;
}

const char* pd_last_error(pd_ctx* c) { return c ? c->err.c_str() : t_err.c_str(); }

void pd_destroy(pd_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)drop_graph(c);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->stage) (void)hipHostFree(c->stage);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (int s = 0; s < kProfRing; s++)
    for (int k = 0; k <= PD_NPHASE; k++)
      if (c->pev[s][k]) (void)hipEventDestroy(c->pev[s][k]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"

// served mode's request loop (host code over the pool entry points above)
#include "pd_serve.hip"
