// pd_api.hip -- host side of the C-ABI declared in include/peridem_gpu.h.
//
// Owns the device state of one context, builds the neighbour rows in the
// reference's conn2NArr order (load_hdf5.h:27-37), lays them out as 64-node
// bond slices and drives the per-step launch sequence on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/peridem_gpu.h"
#include "pd_device.h"
#include "pd_kernels.hip"

using namespace pd;

namespace {

struct HostPart {
  uint32_t n = 0, inst = 0;
  std::vector<double> pos, vol, disp, vel, acc, ext;
  std::vector<uint32_t> conn, clamped, bdry;
  double delta = 0, rho = 0, cnot = 0, snot = 0;
  int movable = 1, breakable = 1, stoppable = 1;
  uint32_t taxis = 0;
  double tval = 0;
  uint64_t off = 0;      // first global node
  uint64_t nbonds = 0;   // directed entries
  std::vector<double> cl0;  // clamped nodes: disp, vel, acc when the state was set
};

struct HostInst {
  pd_run_desc run;
  double wall[6], speed[6];
};

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace

struct pd_ctx {
  std::vector<HostPart> parts;
  std::vector<HostInst> insts;
  bool committed = false;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int device = 0;
  Dev d{};
  std::vector<void*> allocs;
  uint64_t* inst_nodes = nullptr;   // device [ninst+1]
  std::vector<uint64_t> inst_nodes_h;
  uint64_t nbond_entries = 0;       // slice entries incl. padding
  uint32_t t = 1;                   // next reference step index
  pd_counters cnt{};
  double last_ms = 0;
  bool reaction_valid = false;
  std::vector<uint64_t> movable_bonds;  // per particle directed bonds
  uint64_t steps_since_set = 0;
  bool prof = false;
  bool safe_math_static = false;
  hipEvent_t pev[PD_NPHASE + 1] = {};
  pd_profile acc{};
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

#define HIPCHK(c, call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((c), PD_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                               \
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

extern "C" {

const char* pd_version(void) { return "peridem_gpu 0.1 (gfx950, FP64, bitwise reference order)"; }

int pd_create(pd_ctx** out, const pd_run_desc* run) {
  if (!out || !run) return PD_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PD_E_HIP;
  pd_ctx* c = new pd_ctx();
  hipGetDevice(&c->device);
  HostInst I;
  I.run = *run;
  memcpy(I.wall, run->wall_lrtb, sizeof I.wall);
  memcpy(I.speed, run->wall_speed, sizeof I.speed);
  c->insts.push_back(I);
  c->t = run->first_step ? run->first_step : 1;
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
  if (d->nnodes == 0 || !d->pos || !d->vol) return fail(c, PD_E_ARG, "particle needs nnodes, pos, vol");
  if (d->instance >= c->insts.size()) return fail(c, PD_E_ARG, "unknown instance %u", d->instance);
  if (d->nconn && !d->conn) return fail(c, PD_E_ARG, "nconn > 0 but conn is NULL");
  HostPart p;
  uint32_t n = p.n = d->nnodes;
  p.inst = d->instance;
  auto cp3 = [n](const double* s, std::vector<double>& v, int k) {
    v.assign((size_t)n * k, 0.0);
    if (s) memcpy(v.data(), s, (size_t)n * k * sizeof(double));
  };
  cp3(d->pos, p.pos, 3);
  cp3(d->vol, p.vol, 1);
  cp3(d->disp, p.disp, 3);
  cp3(d->vel, p.vel, 3);
  cp3(d->acc, p.acc, 3);
  cp3(d->extforce, p.ext, 3);
  p.conn.assign(d->conn, d->conn + 2 * d->nconn);
  for (uint64_t l = 0; l < 2 * d->nconn; l++)
    if (p.conn[l] >= n) return fail(c, PD_E_ARG, "Connectivity entry %llu out of range", (unsigned long long)l);
  if (d->nclamped) p.clamped.assign(d->clamped, d->clamped + d->nclamped);
  for (auto q : p.clamped)
    if (q >= n) return fail(c, PD_E_ARG, "clamped node out of range");
  if (d->nbdry) p.bdry.assign(d->bdry, d->bdry + d->nbdry);
  for (auto q : p.bdry)
    if (q >= n) return fail(c, PD_E_ARG, "boundary node out of range");
  p.delta = d->delta;
  p.rho = d->rho;
  p.cnot = d->cnot;
  p.snot = d->snot;
  p.movable = d->movable;
  p.breakable = d->breakable;
  p.stoppable = d->stoppable;
  p.taxis = d->torque_axis;
  p.tval = d->torque_val;
  if (index) *index = (uint32_t)c->parts.size();
  c->parts.push_back(std::move(p));
  return PD_OK;
}

static int check_supported(pd_ctx* c) {
  for (auto& I : c->insts) {
    if (I.run.nl_bdry_only) return fail(c, PD_E_UNSUPPORTED, "nl_bdry_only is not built yet");
    if (I.run.enable_torque) return fail(c, PD_E_UNSUPPORTED, "enable_torque is not built yet");
  }
  return PD_OK;
}

static void fill_inst(const HostInst& H, DevInst& D) {
  const pd_run_desc& r = H.run;
  D.Rc = r.contact_rad;
  D.Kn = r.normal_stiffness;
  D.mu = r.friction_coefficient;
  D.zeta = r.damping_ratio;
  // cell edge slightly above R_c: the reference's pair distance
  // (u_j-u_i)+(X_j-X_i) differs from CurrPos_j-CurrPos_i by a few ulp
  D.cell = r.contact_rad * (1.0 + 1e-9) + 1e-300;
  memcpy(D.wall, H.wall, sizeof D.wall);
  memcpy(D.speed, H.speed, sizeof D.speed);
  D.allow_wall = r.allow_wall;
  D.allow_damping = r.allow_damping;
  D.allow_friction = r.allow_friction;
  D.self_contact = r.self_contact;
  D.enable_fracture = r.enable_fracture;
  D.gradient_extforce = r.gradient_extforce;
  D.extforce_maxstep = r.extforce_maxstep;
  D.pad0 = 0;
}

static void fill_part(const HostPart& H, DevPart& D, int rows_sorted) {
  D.node_off = H.off;
  D.nnodes = H.n;
  D.inst = H.inst;
  D.rho = H.rho;
  D.rinv = 1 / H.rho;
  D.cnot = H.cnot;
  D.snot = H.snot;
  D.movable = H.movable;
  D.breakable = H.breakable;
  D.stoppable = H.stoppable;
  D.rows_sorted = rows_sorted;
}

int pd_commit(pd_ctx* c) {
  if (!c) return PD_E_ARG;
  if (c->committed) return fail(c, PD_E_STATE, "already committed");
  if (c->parts.empty()) return fail(c, PD_E_STATE, "no particles");
  int rc = check_supported(c);
  if (rc) return rc;
  // particles of one instance must be contiguous (pair order within instance)
  for (size_t i = 1; i < c->parts.size(); i++)
    if (c->parts[i].inst < c->parts[i - 1].inst)
      return fail(c, PD_E_ARG, "particles must be added in instance order");
  uint64_t N = 0;
  for (auto& p : c->parts) {
    p.off = N;
    N += p.n;
  }
  if (N >= 0x7FFFFFFFull) return fail(c, PD_E_OVERFLOW, "more than 2^31-1 nodes");
  const uint64_t S = (N + kSlice - 1) / kSlice;
  const uint64_t Np = S * kSlice;

  // conn2NArr: rows in Connectivity file order (load_hdf5.h:27-37)
  std::vector<uint32_t> deg(Np, 0);
  for (auto& p : c->parts)
    for (size_t l = 0; l < p.conn.size(); l++) deg[p.off + p.conn[l]]++;
  std::vector<uint64_t> rstart(Np + 1, 0);
  for (uint64_t g = 0; g < Np; g++) rstart[g + 1] = rstart[g] + deg[g];
  std::vector<uint32_t> rows(rstart[Np]);
  {
    std::vector<uint64_t> fillp(rstart.begin(), rstart.end() - 1);
    for (auto& p : c->parts)
      for (size_t l = 0; l + 1 < p.conn.size(); l += 2) {
        uint64_t i = p.off + p.conn[l], j = p.off + p.conn[l + 1];
        rows[fillp[i]++] = (uint32_t)j;
        rows[fillp[j]++] = (uint32_t)i;
      }
  }
  // slices of 64 rows, entries interleaved (entry k of lane l at k*64+l)
  std::vector<uint64_t> soff(S + 1, 0);
  for (uint64_t s = 0; s < S; s++) {
    uint32_t W = 0;
    for (int l = 0; l < kSlice; l++) W = std::max(W, deg[s * kSlice + l]);
    soff[s + 1] = soff[s] + (uint64_t)W * kSlice;
  }
  c->nbond_entries = soff[S];
  std::vector<uint32_t> bond(soff[S], kPad);
  for (uint64_t g = 0; g < N; g++) {
    uint64_t s = g / kSlice, l = g % kSlice;
    for (uint32_t k = 0; k < deg[g]; k++) bond[soff[s] + (uint64_t)k * kSlice + l] = rows[rstart[g] + k];
  }
  // per particle: rows ascending? (enables binary search for bonded pairs)
  std::vector<DevPart> dp(c->parts.size());
  c->movable_bonds.assign(c->parts.size(), 0);
  for (size_t pi = 0; pi < c->parts.size(); pi++) {
    auto& p = c->parts[pi];
    int sorted = 1;
    uint64_t nb = 0;
    for (uint64_t g = p.off; g < p.off + p.n; g++) {
      nb += deg[g];
      for (uint64_t k = rstart[g] + 1; k < rstart[g + 1] && sorted; k++)
        if (rows[k] < rows[k - 1]) sorted = 0;
    }
    p.nbonds = nb;
    c->movable_bonds[pi] = nb;
    c->cnt.total_bonds += nb;
    fill_part(p, dp[pi], sorted);
  }
  std::vector<DevInst> di(c->insts.size());
  for (size_t i = 0; i < c->insts.size(); i++) fill_inst(c->insts[i], di[i]);

  // node arrays
  std::vector<double4> X(Np), U(Np);
  std::vector<double> V(3 * Np, 0.0), A(3 * Np, 0.0), E(3 * Np, 0.0);
  std::vector<uint32_t> pid(Np, 0);
  std::vector<uint8_t> nflag(Np, 0);
  for (size_t pi = 0; pi < c->parts.size(); pi++) {
    auto& p = c->parts[pi];
    for (uint32_t i = 0; i < p.n; i++) {
      uint64_t g = p.off + i;
      X[g] = make_double4(p.pos[3 * i], p.pos[3 * i + 1], p.pos[3 * i + 2], p.vol[i]);
      U[g] = make_double4(p.disp[3 * i], p.disp[3 * i + 1], p.disp[3 * i + 2], 0.0);
      for (int k = 0; k < 3; k++) {
        V[3 * g + k] = p.vel[3 * i + k];
        A[3 * g + k] = p.acc[3 * i + k];
        E[3 * g + k] = p.ext[3 * i + k];
      }
      pid[g] = (uint32_t)pi;
    }
    for (auto q : p.clamped) nflag[p.off + q] |= 1;
  }
  for (uint64_t g = N; g < Np; g++) {
    X[g] = make_double4(0, 0, 0, 0);
    U[g] = make_double4(0, 0, 0, 0);
    pid[g] = 0;
  }
  // initial lagged boxes: minmax(pos + disp) of every particle (timeloop.h:389-393)
  const size_t P = c->parts.size();
  std::vector<double> lrtb(6 * P);
  for (size_t pi = 0; pi < P; pi++) {
    auto& p = c->parts[pi];
    double o[6];
    for (int k = 0; k < 3; k++) o[k] = o[3 + k] = p.pos[k] + p.disp[k];
    for (uint32_t i = 0; i < p.n; i++)
      for (int k = 0; k < 3; k++) {
        double v = p.pos[3 * i + k] + p.disp[3 * i + k];
        o[k] = (v < o[k]) ? v : o[k];
        o[3 + k] = (o[3 + k] < v) ? v : o[3 + k];
      }
    for (int k = 0; k < 6; k++) lrtb[6 * pi + k] = o[k];
  }
  std::vector<unsigned long long> lnext(6 * P);
  for (size_t pi = 0; pi < P; pi++)
    for (int k = 0; k < 3; k++) {
      lnext[6 * pi + k] = ~0ull;
      lnext[6 * pi + 3 + k] = 0ull;
    }
  // particle pairs (i<j) within each instance, in the reference's pair order
  std::vector<uint32_t> pairs;
  c->inst_nodes_h.assign(c->insts.size() + 1, 0);
  {
    size_t i0 = 0;
    for (size_t i = 0; i <= P; i++) {
      if (i == P || (i > i0 && c->parts[i].inst != c->parts[i0].inst)) {
        for (size_t a = i0; a < i; a++)
          for (size_t b = a + 1; b < i; b++) {
            pairs.push_back((uint32_t)a);
            pairs.push_back((uint32_t)b);
          }
        i0 = i;
      }
    }
    for (size_t i = 0; i < P; i++) c->inst_nodes_h[c->parts[i].inst + 1] += c->parts[i].n;
    for (size_t k = 0; k < c->insts.size(); k++) c->inst_nodes_h[k + 1] += c->inst_nodes_h[k];
  }
  uint32_t T = 1024;
  while (T < 2 * N && T < (1u << 31)) T <<= 1;

  Dev& d = c->d;
  d.N = N;
  d.S = S;
  d.P = P;
  d.npairs = pairs.size() / 2;
  d.ninst = c->insts.size();
  d.T = T;
#define ALLOC(f, n) \
  if ((rc = dalloc(c, &d.f, (n)))) return rc;
  ALLOC(X, Np);
  ALLOC(U, Np);
  ALLOC(V, 3 * Np);
  ALLOC(A, 3 * Np);
  ALLOC(F, 3 * Np);
  ALLOC(E, 3 * Np);
  ALLOC(pid, Np);
  ALLOC(nflag, Np);
  ALLOC(deg, Np);
  ALLOC(soff, S + 1);
  ALLOC(bond, soff[S]);
  ALLOC(part, P);
  ALLOC(inst, di.size());
  ALLOC(lrtb, 6 * P);
  ALLOC(lrtb_next, 6 * P);
  ALLOC(pactive, P);
  ALLOC(pairs, pairs.size());
  ALLOC(counters, 8);
  ALLOC(err, 1);
  ALLOC(flags, 1);
  ALLOC(reaction, 18 * di.size());
  ALLOC(wmask, Np);
  ALLOC(bflag, Np / 256 + 1);
  ALLOC(ckey, Np);
  ALLOC(cstart, (size_t)T + 1);
  ALLOC(ccur, T);
  ALLOC(csorted, Np);
  ALLOC(scan_tmp, T / kScanBlock + 2);
#undef ALLOC
  if ((rc = dalloc(c, &c->inst_nodes, c->inst_nodes_h.size()))) return rc;
  hipStream_t st = c->stream;
#define UP(dst, src, n) HIPCHK(c, hipMemcpyAsync((dst), (src), (n) * sizeof(*(dst)), hipMemcpyHostToDevice, st))
  UP(d.X, X.data(), Np);
  UP(d.U, U.data(), Np);
  UP(d.V, V.data(), 3 * Np);
  UP(d.A, A.data(), 3 * Np);
  UP(d.E, E.data(), 3 * Np);
  UP(d.pid, pid.data(), Np);
  UP(d.nflag, nflag.data(), Np);
  UP(d.deg, deg.data(), Np);
  UP(d.soff, soff.data(), S + 1);
  if (soff[S]) UP(d.bond, bond.data(), soff[S]);
  UP(d.part, dp.data(), P);
  UP(d.inst, di.data(), di.size());
  UP(d.lrtb, lrtb.data(), 6 * P);
  UP(d.lrtb_next, lnext.data(), 6 * P);
  if (!pairs.empty()) UP(d.pairs, pairs.data(), pairs.size());
  UP(c->inst_nodes, c->inst_nodes_h.data(), c->inst_nodes_h.size());
#undef UP
  HIPCHK(c, hipMemsetAsync(d.F, 0, 3 * Np * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.pactive, 0, P * sizeof(uint32_t), st));
  HIPCHK(c, hipMemsetAsync(d.counters, 0, 8 * sizeof(unsigned long long), st));
  HIPCHK(c, hipMemsetAsync(d.err, 0, sizeof(uint32_t), st));
  {
    // fast-path domain of the bond arithmetic: coordinates and reference
    // bond lengths (displacements are checked by the position kernel)
    uint32_t fl = 0;
    for (uint64_t g = 0; g < N && !fl; g++)
      if (!(in_fast_range(X[g].x) && in_fast_range(X[g].y) && in_fast_range(X[g].z))) fl = kFlagSafeMath;
    for (uint64_t g = 0; g < N && !fl; g++)
      for (uint64_t k = rstart[g]; k < rstart[g + 1]; k++) {
        const double4& a = X[g];
        const double4& b = X[rows[k]];
        double x0 = b.x - a.x, x1 = b.y - a.y, x2 = b.z - a.z;
        double l2 = (x0 * x0 + x1 * x1) + x2 * x2;
        if (!(l2 >= kFastLo && l2 <= kFastHi)) {
          fl = kFlagSafeMath;
          break;
        }
      }
    c->safe_math_static = fl != 0;
    HIPCHK(c, hipMemcpyAsync(d.flags, &fl, sizeof fl, hipMemcpyHostToDevice, st));
  }
  HIPCHK(c, hipMemsetAsync(d.reaction, 0, 18 * di.size() * sizeof(double), st));
  HIPCHK(c, hipMemsetAsync(d.wmask, 0, Np, st));
  HIPCHK(c, hipMemsetAsync(d.bflag, 0, Np / 256 + 1, st));
  HIPCHK(c, hipStreamSynchronize(st));
  // host copies no longer needed
  for (auto& p : c->parts) {
    p.cl0.resize(9 * p.clamped.size());
    for (size_t q = 0; q < p.clamped.size(); q++)
      for (int k = 0; k < 3; k++) {
        p.cl0[9 * q + k] = p.disp[3 * p.clamped[q] + k];
        p.cl0[9 * q + 3 + k] = p.vel[3 * p.clamped[q] + k];
        p.cl0[9 * q + 6 + k] = p.acc[3 * p.clamped[q] + k];
      }
    std::vector<uint32_t>().swap(p.conn);
    std::vector<double>().swap(p.disp);
    std::vector<double>().swap(p.vel);
    std::vector<double>().swap(p.acc);
    std::vector<double>().swap(p.ext);
  }
  c->cnt.nodes = N;
  c->cnt.particles = P;
  c->committed = true;
  return PD_OK;
}

static bool any_contact_possible(pd_ctx* c) {
  if (c->d.npairs) return true;
  for (auto& I : c->insts)
    if (I.run.self_contact && I.run.enable_fracture) return true;
  return false;
}

static int launch_step(pd_ctx* c, bool want_reaction) {
  Dev& d = c->d;
  hipStream_t st = c->stream;
  const double dt = c->insts[0].run.dt;
  const uint64_t Np = d.S * kSlice;
  auto mark = [&](int k) -> int {
    if (c->prof) HIPCHK(c, hipEventRecord(c->pev[k], st));
    return PD_OK;
  };
  int rc;
  if ((rc = mark(0))) return rc;
  hipLaunchKernelGGL(k_position, dim3(grid_of(Np, 256)), dim3(256), 0, st, d, dt);
  if ((rc = mark(1))) return rc;
  if (d.npairs) hipLaunchKernelGGL(k_pairs, dim3(grid_of(d.npairs, 256)), dim3(256), 0, st, d);
  if (any_contact_possible(c)) {
    HIPCHK(c, hipMemsetAsync(d.cstart, 0, ((size_t)d.T + 1) * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_cellkey, dim3(grid_of(d.N, 256)), dim3(256), 0, st, d);
    uint32_t nb = grid_of(d.T, kScanBlock);
    hipLaunchKernelGGL(k_scan1, dim3(nb), dim3(kScanBlock), 0, st, d.cstart, d.T, d.scan_tmp);
    hipLaunchKernelGGL(k_scan2, dim3(1), dim3(kScanBlock), 0, st, d.scan_tmp, nb, d.cstart + d.T);
    hipLaunchKernelGGL(k_scan3, dim3(nb), dim3(kScanBlock), 0, st, d.cstart, d.T, d.scan_tmp, d.ccur);
    hipLaunchKernelGGL(k_scatter, dim3(grid_of(d.N, 256)), dim3(256), 0, st, d);
  }
  StepArgs a;
  a.d = d;
  a.dt = dt;
  const pd_run_desc& r0 = c->insts[0].run;
  double sc = (double)c->t / (double)r0.extforce_maxstep;
  if (sc > 1) sc = 1;
  a.ext_scale = sc;
  a.t = c->t;
  a.want_reaction = want_reaction ? 1u : 0u;
  if (want_reaction)
    hipLaunchKernelGGL(k_clear_bflag, dim3(grid_of(Np / 256 + 1, 256)), dim3(256), 0, st, d,
                       (uint64_t)(Np / 256 + 1));
  if ((rc = mark(2))) return rc;
  // grid padded to a multiple of the XCD count (k_bond remaps blocks)
  hipLaunchKernelGGL(k_bond, dim3((grid_of(Np, 256) + kXcds - 1) / kXcds * kXcds), dim3(256), 0, st, d);
  if ((rc = mark(3))) return rc;
  hipLaunchKernelGGL(k_contact, dim3(grid_of(Np, 256)), dim3(256), 0, st, a);
  if ((rc = mark(4))) return rc;
  if (want_reaction)
    hipLaunchKernelGGL(k_reaction, dim3(grid_of(d.ninst, 64)), dim3(64), 0, st, d, dt, c->inst_nodes);
  hipLaunchKernelGGL(k_update, dim3(grid_of(Np, 256)), dim3(256), 0, st, d, dt);
  hipLaunchKernelGGL(k_finalize, dim3(grid_of(std::max<uint64_t>(d.P, d.ninst), 256)), dim3(256), 0,
                     st, d, dt);
  if ((rc = mark(5))) return rc;
  HIPCHK(c, hipGetLastError());
  if (c->prof) {
    HIPCHK(c, hipEventSynchronize(c->pev[PD_NPHASE]));
    for (int k = 0; k < PD_NPHASE; k++) {
      float ms = 0;
      HIPCHK(c, hipEventElapsedTime(&ms, c->pev[k], c->pev[k + 1]));
      c->acc.ms[k] += ms;
      c->acc.launches[k] += 1;
    }
  }
  // host mirror of the wall planes (timeloop.h:784-797)
  for (auto& I : c->insts)
    for (int k = 0; k < 6; k++) I.wall[k] += I.speed[k] * dt;
  for (size_t pi = 0; pi < c->parts.size(); pi++)
    if (c->parts[pi].movable) c->cnt.bond_updates += c->movable_bonds[pi];
  c->t++;
  c->cnt.steps++;
  c->steps_since_set++;
  return PD_OK;
}

int pd_step(pd_ctx* c, uint32_t nsteps) {
  if (!c) return PD_E_ARG;
  if (!c->committed) return fail(c, PD_E_STATE, "pd_step before pd_commit");
  if (nsteps == 0) return PD_OK;
  int rc = check_supported(c);
  if (rc) return rc;
  bool want_reac = false;
  for (auto& I : c->insts) want_reac |= I.run.compute_wall_reaction && I.run.allow_wall;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (uint32_t s = 0; s < nsteps; s++)
    if ((rc = launch_step(c, want_reac && s + 1 == nsteps))) return rc;
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  c->reaction_valid = want_reac;
  uint32_t err = 0;
  HIPCHK(c, hipMemcpy(&err, c->d.err, sizeof err, hipMemcpyDeviceToHost));
  if (err & kErrCandOverflow)
    return fail(c, PD_E_OVERFLOW, "more than %d contact partners within R_c for a node", kMaxCand);
  return PD_OK;
}

int pd_set_flag(pd_ctx* c, int flag, int value) {
  if (!c) return PD_E_ARG;
  for (size_t i = 0; i < c->insts.size(); i++) {
    pd_run_desc& r = c->insts[i].run;
    switch (flag) {
      case PD_FLAG_ENABLE_FRACTURE: r.enable_fracture = value; break;
      case PD_FLAG_SELF_CONTACT: r.self_contact = value; break;
      case PD_FLAG_ALLOW_DAMPING: r.allow_damping = value; break;
      case PD_FLAG_ALLOW_FRICTION: r.allow_friction = value; break;
      case PD_FLAG_ENABLE_TORQUE: r.enable_torque = value; break;
      case PD_FLAG_COMPUTE_WALL_REACTION: r.compute_wall_reaction = value; break;
      default: return fail(c, PD_E_ARG, "unknown flag %d", flag);
    }
    if (c->committed) {
      DevInst D;
      fill_inst(c->insts[i], D);
      // wall planes live on the device; keep them
      HIPCHK(c, hipMemcpyAsync(&D.wall, &c->d.inst[i].wall, sizeof D.wall, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipMemcpy(&c->d.inst[i], &D, sizeof D, hipMemcpyHostToDevice));
    }
  }
  return PD_OK;
}

static int push_inst(pd_ctx* c, size_t i) {
  if (!c->committed) return PD_OK;
  DevInst D;
  fill_inst(c->insts[i], D);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(&c->d.inst[i], &D, sizeof D, hipMemcpyHostToDevice));
  return PD_OK;
}

int pd_set_wall_speed(pd_ctx* c, const double speed[6]) {
  if (!c || !speed) return PD_E_ARG;
  for (size_t i = 0; i < c->insts.size(); i++) {
    memcpy(c->insts[i].speed, speed, sizeof c->insts[i].speed);
    int rc = push_inst(c, i);
    if (rc) return rc;
  }
  return PD_OK;
}

int pd_set_wall(pd_ctx* c, const double lrtb[6]) {
  if (!c || !lrtb) return PD_E_ARG;
  for (size_t i = 0; i < c->insts.size(); i++) {
    memcpy(c->insts[i].wall, lrtb, sizeof c->insts[i].wall);
    int rc = push_inst(c, i);
    if (rc) return rc;
  }
  return PD_OK;
}

int pd_set_particle_flag(pd_ctx* c, uint32_t p, int which, int value) {
  if (!c || p >= c->parts.size()) return PD_E_ARG;
  auto& H = c->parts[p];
  switch (which) {
    case PD_PFLAG_MOVABLE: H.movable = value; break;
    case PD_PFLAG_STOPPABLE: H.stoppable = value; break;
    case PD_PFLAG_BREAKABLE: H.breakable = value; break;
    default: return fail(c, PD_E_ARG, "unknown particle flag %d", which);
  }
  if (c->committed) {
    DevPart D;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&D, &c->d.part[p], sizeof D, hipMemcpyDeviceToHost));
    D.movable = H.movable;
    D.stoppable = H.stoppable;
    D.breakable = H.breakable;
    HIPCHK(c, hipMemcpy(&c->d.part[p], &D, sizeof D, hipMemcpyHostToDevice));
  }
  return PD_OK;
}

int pd_set_snot(pd_ctx* c, uint32_t p, double snot) {
  if (!c || p >= c->parts.size()) return PD_E_ARG;
  c->parts[p].snot = snot;
  if (c->committed) {
    DevPart D;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&D, &c->d.part[p], sizeof D, hipMemcpyDeviceToHost));
    D.snot = snot;
    HIPCHK(c, hipMemcpy(&c->d.part[p], &D, sizeof D, hipMemcpyHostToDevice));
  }
  return PD_OK;
}

int pd_get_state(pd_ctx* c, uint32_t pi, double* disp, double* currpos, double* vel, double* acc,
                 double* force) {
  if (!c || pi >= c->parts.size()) return PD_E_ARG;
  if (!c->committed) return fail(c, PD_E_STATE, "pd_get_state before pd_commit");
  auto& p = c->parts[pi];
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (disp || currpos) {
    std::vector<double4> U(p.n), X(p.n);
    HIPCHK(c, hipMemcpy(U.data(), c->d.U + p.off, p.n * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(X.data(), c->d.X + p.off, p.n * sizeof(double4), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < p.n; i++) {
      double u[3] = {U[i].x, U[i].y, U[i].z};
      double x[3] = {X[i].x, X[i].y, X[i].z};
      for (int k = 0; k < 3; k++) {
        if (disp) disp[3 * i + k] = u[k];
        // CurrPos is pos+disp of the position half; clamped nodes keep it
        // although their disp was reset (timeloop.h:539,767-774)
        if (currpos) currpos[3 * i + k] = x[k] + u[k];
      }
    }
    if (currpos && c->steps_since_set == 1) {
      // CurrPos of a clamped node is pos + the disp of the position half,
      // formed before the clamp reset disp (timeloop.h:539,767-774).  From
      // the second step on that disp is 0 + (dt*0 + h*0); only the first step
      // after the state was set sees the initial disp/vel/acc.
      const double dt = c->insts[p.inst].run.dt, h = (dt * dt) * 0.5;
      for (size_t q = 0; q < p.clamped.size(); q++) {
        uint32_t i = p.clamped[q];
        for (int k = 0; k < 3; k++) {
          double u = p.cl0[9 * q + k] + ((dt * p.cl0[9 * q + 3 + k]) + (h * p.cl0[9 * q + 6 + k]));
          currpos[3 * i + k] = p.pos[3 * i + k] + u;
        }
      }
    }
  }
  auto down3 = [&](double* dst, const double* src) -> int {
    if (!dst) return PD_OK;
    HIPCHK(c, hipMemcpy(dst, src + 3 * p.off, 3ull * p.n * sizeof(double), hipMemcpyDeviceToHost));
    return PD_OK;
  };
  int rc;
  if ((rc = down3(vel, c->d.V))) return rc;
  if ((rc = down3(acc, c->d.A))) return rc;
  if ((rc = down3(force, c->d.F))) return rc;
  return PD_OK;
}

int pd_set_state(pd_ctx* c, uint32_t pi, const double* disp, const double* vel, const double* acc) {
  if (!c || pi >= c->parts.size()) return PD_E_ARG;
  if (!c->committed) return fail(c, PD_E_STATE, "pd_set_state before pd_commit");
  auto& p = c->parts[pi];
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (disp) {
    std::vector<double4> U(p.n);
    for (uint32_t i = 0; i < p.n; i++) U[i] = make_double4(disp[3 * i], disp[3 * i + 1], disp[3 * i + 2], 0.0);
    HIPCHK(c, hipMemcpy(c->d.U + p.off, U.data(), p.n * sizeof(double4), hipMemcpyHostToDevice));
  }
  if (vel) HIPCHK(c, hipMemcpy(c->d.V + 3 * p.off, vel, 3ull * p.n * sizeof(double), hipMemcpyHostToDevice));
  if (acc) HIPCHK(c, hipMemcpy(c->d.A + 3 * p.off, acc, 3ull * p.n * sizeof(double), hipMemcpyHostToDevice));
  for (size_t q = 0; q < p.clamped.size(); q++)
    for (int k = 0; k < 3; k++) {
      uint32_t i = p.clamped[q];
      if (disp) p.cl0[9 * q + k] = disp[3 * i + k];
      if (vel) p.cl0[9 * q + 3 + k] = vel[3 * i + k];
      if (acc) p.cl0[9 * q + 6 + k] = acc[3 * i + k];
    }
  c->steps_since_set = 0;
  return PD_OK;
}

// Download the bond rows of particle p (local indices, broken bit kept).
static int particle_rows(pd_ctx* c, uint32_t pi, std::vector<uint32_t>& deg,
                         std::vector<std::vector<uint32_t>>& rows) {
  auto& p = c->parts[pi];
  uint64_t s0 = p.off / kSlice, s1 = (p.off + p.n + kSlice - 1) / kSlice;
  std::vector<uint64_t> soff(s1 - s0 + 1);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(soff.data(), c->d.soff + s0, soff.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
  std::vector<uint32_t> ent(soff.back() - soff.front());
  if (!ent.empty())
    HIPCHK(c, hipMemcpy(ent.data(), c->d.bond + soff.front(), ent.size() * sizeof(uint32_t),
                        hipMemcpyDeviceToHost));
  deg.resize(p.n);
  HIPCHK(c, hipMemcpy(deg.data(), c->d.deg + p.off, p.n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  rows.assign(p.n, {});
  for (uint32_t i = 0; i < p.n; i++) {
    uint64_t g = p.off + i, s = g / kSlice, l = g % kSlice;
    uint64_t base = soff[s - s0] - soff.front();
    rows[i].resize(deg[i]);
    for (uint32_t k = 0; k < deg[i]; k++) {
      uint32_t e = ent[base + (uint64_t)k * kSlice + l];
      rows[i][k] = ((e & kIdxMask) - (uint32_t)p.off) | (e & kBroken);
    }
  }
  return PD_OK;
}

int pd_get_connectivity(pd_ctx* c, uint32_t pi, uint32_t* pairs, uint64_t cap, uint64_t* m) {
  if (!c || pi >= c->parts.size() || !m) return PD_E_ARG;
  if (!c->committed) return fail(c, PD_E_STATE, "pd_get_connectivity before pd_commit");
  std::vector<uint32_t> deg;
  std::vector<std::vector<uint32_t>> rows;
  int rc = particle_rows(c, pi, deg, rows);
  if (rc) return rc;
  uint64_t n = 0;
  for (uint32_t i = 0; i < rows.size(); i++)
    for (uint32_t e : rows[i])
      if (!(e & kBroken) && i < e) n++;
  *m = n;
  if (!pairs) return PD_OK;
  if (cap < n) return fail(c, PD_E_ARG, "connectivity buffer holds %llu pairs, %llu needed",
                           (unsigned long long)cap, (unsigned long long)n);
  uint64_t w = 0;
  for (uint32_t i = 0; i < rows.size(); i++)
    for (uint32_t e : rows[i])
      if (!(e & kBroken) && i < e) {
        pairs[2 * w] = i;
        pairs[2 * w + 1] = e;
        w++;
      }
  return PD_OK;
}

int pd_get_wall(pd_ctx* c, uint32_t instance, double lrtb[6], double reaction[18]) {
  if (!c || instance >= c->insts.size()) return PD_E_ARG;
  if (lrtb) memcpy(lrtb, c->insts[instance].wall, 6 * sizeof(double));
  if (reaction) {
    if (c->committed && c->reaction_valid) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipMemcpy(reaction, c->d.reaction + 18 * instance, 18 * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      for (int k = 0; k < 18; k++) reaction[k] = 0.0;
    }
  }
  return PD_OK;
}

int pd_get_counters(pd_ctx* c, pd_counters* out) {
  if (!c || !out) return PD_E_ARG;
  *out = c->cnt;
  if (c->committed) {
    unsigned long long pw = 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&pw, c->d.counters, sizeof pw, hipMemcpyDeviceToHost));
    out->pair_computations = pw;
    unsigned long long* dcount = c->d.counters + 1;
    HIPCHK(c, hipMemsetAsync(dcount, 0, sizeof(unsigned long long), c->stream));
    if (c->nbond_entries)
      hipLaunchKernelGGL(k_count_alive, dim3(std::min<uint64_t>(grid_of(c->nbond_entries, 256), 4096)),
                         dim3(256), 0, c->stream,
                         c->d, c->nbond_entries, dcount);
    unsigned long long alive = 0;
    HIPCHK(c, hipMemcpyAsync(&alive, dcount, sizeof alive, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out->alive_bonds = alive;
  }
  return PD_OK;
}

int pd_set_profiling(pd_ctx* c, int on) {
  if (!c) return PD_E_ARG;
  if (on && !c->pev[0])
    for (int k = 0; k <= PD_NPHASE; k++) HIPCHK(c, hipEventCreate(&c->pev[k]));
  c->prof = on != 0;
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
  if (hipMalloc((void**)&dm, 2 * sizeof(unsigned long long)) != hipSuccess) return PD_E_HIP;
  int rc = PD_OK;
  unsigned long long hm[2] = {0, 0};
  if (hipMemset(dm, 0, 2 * sizeof(unsigned long long)) != hipSuccess) rc = PD_E_HIP;
  if (rc == PD_OK) {
    hipLaunchKernelGGL(k_selftest_math, dim3(2048), dim3(256), 0, 0, n, seed, dm);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(hm, dm, sizeof hm, hipMemcpyDeviceToHost) != hipSuccess)
      rc = PD_E_HIP;
  }
  hipFree(dm);
  *mismatch_div = hm[0];
  *mismatch_sqrt = hm[1];
  return rc;
}

int pd_math_mode(pd_ctx* c, int* safe) {
  if (!c || !safe) return PD_E_ARG;
  uint32_t fl = 0;
  if (c->committed) {
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

const char* pd_last_error(pd_ctx* c) { return c ? c->err.c_str() : "null context"; }

void pd_destroy(pd_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) hipFree(p);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  for (int k = 0; k <= PD_NPHASE; k++)
    if (c->pev[k]) hipEventDestroy(c->pev[k]);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
