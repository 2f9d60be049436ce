// pd_kernels.hip -- CDNA4 (gfx950) kernels of the explicit step.
//
// One step (timeloop.h:395-797) is the launch sequence
//   k_position   Verlet position half + per-particle box of CurrPos
//   k_cellkey    the particle-pair box test on the lagged boxes
//                (timeloop.h:636-653), the wall prefilter, and the cell keys
//                of the nodes of jobs that rebuild their candidate lists
//   cell list    scan -> k_scatter -> k_build
//                (only in steps where some job rebuilds its candidate lists)
//   bond kernel  PD bond force + fused irreversible break, per-node row sums
//                (k_bond_lds: block tables staged in LDS; k_bond_t: unstaged),
//                then the node's self/wall/ext/torque/pair contact terms
//                added in the reference's order (add_contact)
//                and the Verlet velocity half (velocity_half)
//   k_reaction   wall reaction sum (jobs that report it, on its steps)
//   k_clamp      clamped nodes' displacement (pools with clamped nodes)
//   k_finalize   lagged boxes, wall motion, cell table cleared after a rebuild
// All floating point is FP64 with the reference's operation order; the build
// uses -ffp-contract=off so no multiply-add is fused.
#include "pd_device.h"

namespace pd {

__device__ __forceinline__ bool live(const DevInst& I) { return I.remaining != 0; }

// A record every lane of a wavefront reads at the same index, read through
// the scalar cache (the constant address space: the fields read this way are
// not written during the launch).
template <class T>
using kptr = const T __attribute__((address_space(4)))*;

template <class T>
__device__ __forceinline__ T ld_uniform(const T* base, uint32_t i) {
  static_assert(sizeof(T) % 4 == 0, "records of whole words");
  T v;
  uint32_t* w = reinterpret_cast<uint32_t*>(&v);
  const kptr<uint32_t> s = (kptr<uint32_t>)((kptr<T>)base + i);
#pragma unroll
  for (unsigned k = 0; k < sizeof(T) / 4; k++) w[k] = s[k];
  return v;
}

// The pool was halted earlier in this pd_step call by a candidate list
// longer than the capacity: the kernel does nothing (read once per
// wavefront; the flag only goes up during the list build, and a build that
// reads it clear just carries on).
__device__ __forceinline__ bool halted(const Dev& d) { return ld_uniform(&d.ctl->halt, 0u) != 0; }

// Does job I accumulate its wall reaction in this step?  On its last step and
// on the last step of a pd_step call, so the value read back afterwards is the
// reaction of the job's latest step, as Wall.reaction_* is (timeloop.h:397).
__device__ __forceinline__ bool reaction_step(const Dev& d, const DevInst& I) {
  return I.compute_reaction && I.allow_wall && (d.ctl->steps_left == 1 || I.remaining == 1);
}

// ---------------------------------------------------------------------------
// The bond kernel's neighbour copies of nodes [g0, g0+n) from U and, when a
// job is placed (pos), from X (after an upload; the position kernel keeps the
// displacement copies current).  XY/ZV are the reference positions the bond
// vectors are formed from and are written only at placement: a wheel reset
// moves X alone.
__global__ __launch_bounds__(256) void k_split(Dev d, uint64_t g0, uint64_t n, int pos) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t g = g0 + i;
  if (pos) {
    const double4 x = d.X[g];
    d.XY[g] = make_double2(x.x, x.y);
    d.ZV[g] = make_double2(x.z, x.w);
  }
  const double4 u = d.U[g];
  d.UXY[g] = make_double2(u.x, u.y);
  d.UZ[g] = u.z;
}

// The displacement's resident form is the split copy (UXY, UZ); U holds it
// only around host transfers (k_split after an upload, k_join before a
// read-back).
__device__ __forceinline__ double4 disp_of(const Dev& d, uint64_t g) {
  const double2 a = d.UXY[g];
  return make_double4(a.x, a.y, d.UZ[g], 0.0);
}

__device__ __forceinline__ void set_disp(const Dev& d, uint64_t g, double x, double y, double z) {
  d.UXY[g] = make_double2(x, y);
  d.UZ[g] = z;
}

// slot s of job's displacement maximum since its list build
__device__ __forceinline__ unsigned long long* du2_slot(const Dev& d, uint64_t job, uint32_t s) {
  return d.du2s + (job * kDuSlots + s) * kDuStride;
}

__global__ __launch_bounds__(256) void k_join(Dev d, uint64_t g0, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d.U[g0 + i] = disp_of(d, g0 + i);
}

// ---------------------------------------------------------------------------
// Verlet position half (timeloop.h:537-540; overloads.h:51-70,149-159):
//   disp += dt*vel + (dt*dt*0.5)*acc ; CurrPos = pos + disp
// plus the min/max of CurrPos that becomes the particle's lagged box at the
// end of the step (timeloop.h:779).
__global__ __launch_bounds__(256) void k_position(Dev d) {
  if (halted(d)) return;
  uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // the block's wall-contact flag of this step starts clear, and so does the
  // step's rebuild flag (set by k_cellkey)
  if (threadIdx.x == 0 && g < d.N) d.bflag[g >> 8] = 0;
  if (g == 0) {
    d.ctl->any_rebuild = 0;
    d.ctl->nrj = 0;
  }
  double cp[3] = {0, 0, 0};
  uint32_t p = 0xFFFFFFFFu, jid = 0xFFFFFFFFu;
  unsigned long long dkey = 0ull;
  bool movable = false, stepping = false;
  double dt = 0.0;
  if (g < d.N) {
    // the particle's and job's fields, through the scalar cache in
    // single-particle wavefronts
    p = d.pid[g];
    const uint32_t pu = __builtin_amdgcn_readfirstlane(p);
    if (__builtin_amdgcn_ballot_w64(p != pu) == 0) {
      const DevPart P = ld_uniform(d.part, pu);
      const DevInst I = ld_uniform(d.inst, P.inst);
      stepping = live(I);
      movable = P.movable != 0;
      jid = P.inst;
      dt = I.dt;
    } else {
      const DevPart& P = d.part[p];
      const DevInst& I = d.inst[P.inst];
      stepping = live(I);
      movable = P.movable != 0;
      jid = P.inst;
      dt = I.dt;
    }
    if (!stepping) p = jid = 0xFFFFFFFFu;
  }
  if (g < d.N && !stepping) {
    // frozen and void slots keep their displacement in the step's buffer
    d.UXY[g] = d.UXYo[g];
    d.UZ[g] = d.UZo[g];
  }
  if (stepping) {
    const double2 uo = d.UXYo[g];
    double4 u = make_double4(uo.x, uo.y, d.UZo[g], 0.0);
    const double4 x = d.X[g];
    const double h = (dt * dt) * 0.5;
    double v0 = d.V[3 * g], v1 = d.V[3 * g + 1], v2 = d.V[3 * g + 2];
    double a0 = d.A[3 * g], a1 = d.A[3 * g + 1], a2 = d.A[3 * g + 2];
    d.Vold[3 * g] = v0;
    d.Vold[3 * g + 1] = v1;
    d.Vold[3 * g + 2] = v2;
    u.x = u.x + ((dt * v0) + (h * a0));
    u.y = u.y + ((dt * v1) + (h * a1));
    u.z = u.z + ((dt * v2) + (h * a2));
    set_disp(d, g, u.x, u.y, u.z);
    if (!(in_fast_range(u.x) & in_fast_range(u.y) & in_fast_range(u.z))) atomicOr(d.flags, kFlagSafeMath);
    const double2 ub = d.UBXY[g];
    const double e0 = u.x - ub.x, e1 = u.y - ub.y, e2 = u.z - d.UBZ[g];
    dkey = (unsigned long long)__double_as_longlong((e0 * e0 + e1 * e1) + e2 * e2);
    cp[0] = x.x + u.x;
    cp[1] = x.y + u.y;
    cp[2] = x.z + u.z;
    if (!movable) p = 0xFFFFFFFFu;
  }
  // largest displacement since the candidate lists were built, per job: the
  // maximum over the wavefront, then over the block (a 256-node block
  // belongs to one job), then one atomic per block into the slot of the XCD
  // it runs on, only when it exceeds what the slot holds
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(dkey, off, 64);
    const uint32_t oj = __shfl_xor(jid, off, 64);
    dkey = o > dkey ? o : dkey;
    jid = oj < jid ? oj : jid;
  }
  __shared__ unsigned long long sdu[4];
  __shared__ uint32_t sjob[4];
  if ((threadIdx.x & 63) == 0) {
    sdu[threadIdx.x >> 6] = dkey;
    sjob[threadIdx.x >> 6] = jid;
  }
  // min/max of CurrPos per particle, as order-preserving keys.  Particles
  // hold contiguous node ranges, so a particle's nodes in a wavefront form
  // one run: a segmented reduction over the wavefront leaves each run's
  // extremes in the run's first lane.  The wavefront's first and last runs
  // may continue in the neighbouring wavefronts: they are merged across the
  // block's four wavefronts in LDS and published once per run of the block;
  // runs inside a wavefront publish directly.
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long mn[3], mx[3];
  for (int k = 0; k < 3; k++) {
    mn[k] = (p != 0xFFFFFFFFu) ? okey(cp[k]) : ~0ull;
    mx[k] = (p != 0xFFFFFFFFu) ? okey(cp[k]) : 0ull;
  }
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t po = __shfl_down(p, off, 64);
    const bool take = (lane + off < 64) && po == p;
    for (int k = 0; k < 3; k++) {
      const unsigned long long a = __shfl_down(mn[k], off, 64), b = __shfl_down(mx[k], off, 64);
      if (take) {
        mn[k] = a < mn[k] ? a : mn[k];
        mx[k] = b > mx[k] ? b : mx[k];
      }
    }
  }
  const uint32_t pup = __shfl_up(p, 1, 64);
  const uint32_t p0 = __shfl(p, 0, 64), plast = __shfl(p, 63, 64);
  const bool head = lane == 0 || pup != p;
  // per wavefront: entry 0 = its first run, entry 1 = its last run (void if
  // the wavefront is one run)
  __shared__ unsigned long long rmin[4][2][3], rmax[4][2][3];
  __shared__ uint32_t rp[4][2];
  if (head && p != 0xFFFFFFFFu && p != p0 && p != plast) {   // a run inside the wavefront
    for (int k = 0; k < 3; k++) {
      atomicMin(&d.lrtb_next[6 * p + k], mn[k]);
      atomicMax(&d.lrtb_next[6 * p + 3 + k], mx[k]);
    }
  }
  const int e = lane == 0 ? 0 : 1;
  if (lane == 0 || (head && p == plast && p != p0)) {
    rp[w][e] = p;
    for (int k = 0; k < 3; k++) {
      rmin[w][e][k] = mn[k];
      rmax[w][e][k] = mx[k];
    }
  }
  if (lane == 0 && p0 == plast) rp[w][1] = 0xFFFFFFFFu;
  __syncthreads();
  if (threadIdx.x == 0) {
    // the job's displacement maximum: one atomic per block into its XCD's slot
    unsigned long long m = 0ull;
    uint32_t job = 0xFFFFFFFFu;
    for (int v = 0; v < 4; v++)
      if (sjob[v] != 0xFFFFFFFFu) {
        job = sjob[v];
        m = sdu[v] > m ? sdu[v] : m;
      }
    if (job != 0xFFFFFFFFu) {
      unsigned long long* slot = du2_slot(d, job, blockIdx.x & (kDuSlots - 1));
      if (m > *slot) atomicMax(slot, m);
    }
    // the wavefronts' edge runs in node order, equal neighbours merged
    uint32_t cur = 0xFFFFFFFFu;
    unsigned long long cmn[3] = {~0ull, ~0ull, ~0ull}, cmx[3] = {0ull, 0ull, 0ull};
    for (int i = 0; i <= 8; i++) {
      const uint32_t q = i < 8 ? rp[i >> 1][i & 1] : 0xFFFFFFFEu;   // end marker
      if (i < 8 && (i & 1) && q == 0xFFFFFFFFu) continue;             // no separate last run
      if (q != cur) {
        if (cur != 0xFFFFFFFFu)
          for (int k = 0; k < 3; k++) {
            atomicMin(&d.lrtb_next[6 * cur + k], cmn[k]);
            atomicMax(&d.lrtb_next[6 * cur + 3 + k], cmx[k]);
          }
        cur = q == 0xFFFFFFFEu ? 0xFFFFFFFFu : q;
        for (int k = 0; k < 3; k++) {
          cmn[k] = ~0ull;
          cmx[k] = 0ull;
        }
      }
      if (i < 8 && cur != 0xFFFFFFFFu)
        for (int k = 0; k < 3; k++) {
          const unsigned long long a = rmin[i >> 1][i & 1][k], b = rmax[i >> 1][i & 1][k];
          cmn[k] = a < cmn[k] ? a : cmn[k];
          cmx[k] = b > cmx[k] ? b : cmx[k];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// Contact candidates come from per-node Verlet lists: every node j of the
// same job within rb = R_c (1 + 1e-9) + skin of node i when the list was
// built.  While no node of the job has moved more than skin/2 since (less a
// relative 1e-6 margin for rounding), every pair within R_c is in the lists,
// so the exact reference tests applied to the list members find exactly the
// reference's contact pairs.  Whether a job rebuilds its lists this step
// follows from the largest displacement since the build and the host's
// dirty flag; every kernel of the step that needs it evaluates it alike (its
// inputs do not change between k_position and k_finalize, which applies the
// resets).
__device__ __forceinline__ bool rebuilds(const Dev& d, uint32_t job) {
  const DevInst& I = d.inst[job];
  if (!live(I)) return false;
  unsigned long long m = 0ull;
#pragma unroll
  for (uint32_t s = 0; s < kDuSlots; s++) {
    const unsigned long long v = *du2_slot(d, job, s);
    m = v > m ? v : m;
  }
  return (I.dirty != 0) || (__longlong_as_double((long long)m) > I.du_lim2);
}

// The same for a node's job, evaluated once per wavefront through the scalar
// cache when the wavefront's nodes share one job (all but the padding slots
// at a job's end); the job records and slots are not written by the kernels
// that ask.
__device__ __forceinline__ bool rebuilds_node(const Dev& d, uint32_t job) {
  const uint32_t ju = __builtin_amdgcn_readfirstlane(job);
  if (__builtin_amdgcn_ballot_w64(job != ju) != 0) return rebuilds(d, job);
  const DevInst I = ld_uniform(d.inst, ju);
  if (!live(I)) return false;
  unsigned long long m = 0ull;
#pragma unroll
  for (uint32_t s = 0; s < kDuSlots; s++) {
    const unsigned long long v = ld_uniform(d.du2s, (ju * kDuSlots + s) * kDuStride);
    m = v > m ? v : m;
  }
  return (I.dirty != 0) || (__longlong_as_double((long long)m) > I.du_lim2);
}

// ---------------------------------------------------------------------------
// mean_CurrPos (particle2.h:94-103) of every particle of a job with torque on:
// pos + disp summed in node order, as the reference.
// The sums run one node after another, left to right, as the reference's
// loop: one wavefront per particle loads 64 consecutive nodes' values at
// once (the next 64 while the sums take these), then the sums take them lane
// after lane through scalar reads of the lanes (no memory wait between the
// additions).
__device__ __forceinline__ double lane_value(double v, uint32_t l) {
  const long long b = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, (int)l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), (int)l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ void add_lanes3(uint32_t cnt, double c0, double c1, double c2, double& m0, double& m1,
                                           double& m2) {
  for (uint32_t l = 0; l < cnt; l++) {
    m0 += lane_value(c0, l);
    m1 += lane_value(c1, l);
    m2 += lane_value(c2, l);
  }
}
__device__ __forceinline__ void currpos_of(const Dev& d, uint64_t g, uint64_t g1, double& c0, double& c1,
                                           double& c2) {
  c0 = c1 = c2 = 0.0;
  if (g < g1) {
    const double4 x = d.X[g];
    const double4 u = disp_of(d, g);
    c0 = x.x + u.x;
    c1 = x.y + u.y;
    c2 = x.z + u.z;
  }
}
__device__ __forceinline__ void sum_currpos(const Dev& d, uint64_t g0, uint64_t g1, double& m0, double& m1,
                                            double& m2) {
  double c0, c1, c2;
  currpos_of(d, g0 + threadIdx.x, g1, c0, c1, c2);
  for (uint64_t b = g0; b < g1; b += 64) {
    double n0, n1, n2;   // the next 64 nodes, loading meanwhile
    currpos_of(d, b + 64 + threadIdx.x, g1, n0, n1, n2);
    add_lanes3((uint32_t)(g1 - b < 64 ? g1 - b : 64), c0, c1, c2, m0, m1, m2);
    c0 = n0;
    c1 = n1;
    c2 = n2;
  }
}
__global__ __launch_bounds__(64) void k_centroid(Dev d) {
  if (halted(d)) return;
  const uint64_t q = blockIdx.x;
  if (q >= d.P) return;
  const DevPart& P = d.part[q];
  const DevInst& I = d.inst[P.inst];
  if (!live(I) || !I.enable_torque) return;
  double m0 = 0.0, m1 = 0.0, m2 = 0.0;
  sum_currpos(d, P.node_off, P.node_off + P.nnodes, m0, m1, m2);
  if (threadIdx.x == 0) {
    const double n = (double)P.nnodes;
    d.cent[3 * q] = m0 / n;
    d.cent[3 * q + 1] = m1 / n;
    d.cent[3 * q + 2] = m2 / n;
  }
}

// ---------------------------------------------------------------------------
// Wheel reset (timeloop.h:496-535) of a job's particle 0 (slot p0), before the
// position half of the step it precedes.  One workgroup:
//  * mean_CurrPos of particle 0 (particle2.h:94-103), one thread per
//    component summing pos + disp in node order;
//  * the bulk height: the largest pos_y + disp_y over the nodes [b0, b1)
//    (the job's particles 1.. in order) with |m_x - (pos_x + disp_x)| <=
//    wheel_rad.  The reference keeps the first node of largest value in that
//    order (a later one replaces it only when strictly greater), so each
//    thread keeps its first best in increasing node order and the threads'
//    results combine by (larger value, then smaller node index);
//  * with no such node, particle 1's pos[0]_y + disp[1]_y (node 1's
//    displacement, as the reference starts its maximum);
//  * move = ((bulk + wheel_rad) + Rc) - m_y is added to X.y of particle 0.
// Only X moves: the bond kernels read the reference positions from XY/ZV,
// which keep the values gen_xi formed the bond vectors from
// (particle2.h:241-257).  The move is returned in *out for the host's copy.
constexpr uint32_t kResetThreads = 256;
__global__ __launch_bounds__(kResetThreads) void k_partzero(Dev d, uint32_t p0, uint64_t b0, uint64_t b1,
                                                             double wheel_rad, double Rc, double* out) {
  __shared__ double sm[3];
  __shared__ double sv[kResetThreads];
  __shared__ uint64_t si[kResetThreads];
  __shared__ double smove;
  constexpr uint64_t kNone = ~0ull;
  const DevPart& P = d.part[p0];
  const uint32_t tid = threadIdx.x;
  if (tid < 3) {
    double m = 0.0;
    for (uint64_t g = P.node_off; g < P.node_off + P.nnodes; g++) {
      const double4 x = d.X[g];
      const double4 u = disp_of(d, g);
      m += tid == 0 ? (x.x + u.x) : tid == 1 ? (x.y + u.y) : (x.z + u.z);
    }
    sm[tid] = m / (double)P.nnodes;
  }
  __syncthreads();
  const double mx = sm[0];
  double best = 0.0;
  uint64_t bi = kNone;
  for (uint64_t g = b0 + tid; g < b1; g += kResetThreads) {
    const double4 x = d.X[g];
    const double4 u = disp_of(d, g);
    const double now_x = x.x + u.x, now_y = x.y + u.y;
    if (fabs(mx - now_x) <= wheel_rad && (bi == kNone || now_y > best)) {
      best = now_y;
      bi = g;
    }
  }
  sv[tid] = best;
  si[tid] = bi;
  __syncthreads();
  if (tid == 0) {
    for (uint32_t k = 1; k < kResetThreads; k++) {
      const uint64_t oi = si[k];
      if (oi == kNone) continue;
      if (bi == kNone || sv[k] > best || (sv[k] == best && oi < bi)) {
        best = sv[k];
        bi = oi;
      }
    }
    double top;
    if (bi != kNone) {
      top = best;
    } else {
      top = d.X[b0].y + disp_of(d, b0 + 1).y;
    }
    const double dest = top + wheel_rad + Rc;
    smove = dest - sm[1];
    *out = smove;
  }
  __syncthreads();
  const double move = smove;
  for (uint64_t g = P.node_off + tid; g < P.node_off + P.nnodes; g += kResetThreads) d.X[g].y += move;
}

// ---------------------------------------------------------------------------
// Pair box test on the lagged boxes, pair list in the reference's pair order
// restricted to one instance (timeloop.h:243-255, 636-653).
//
// In a step where the pair's job rebuilds its candidate lists, bit 3 marks
// both particles when their boxes of this step (min/max of the CurrPos the
// position kernel has just accumulated) come within rb (1 + 1e-6) of each
// other: a node of a particle with no self-contact and no such pair has no
// candidate within rb, so its list is empty and its cell walk is skipped.
// Until the next rebuild no node moves half a skin, so such a particle
// cannot come within R_c of another one before its lists are built again.
// Non-movable particles have no box of this step: their pairs count as near.
__device__ __forceinline__ void list_test(Dev& d, uint32_t i, uint32_t j, const DevInst& I) {
  bool near = !(d.part[i].movable && d.part[j].movable);
  if (!near) {
    double A[6], B[6];
    for (int k = 0; k < 6; k++) {
      A[k] = okey_inv(d.lrtb_next[6 * i + k]);
      B[k] = okey_inv(d.lrtb_next[6 * j + k]);
    }
    near = intersects_box_fext(A, B, I.rb * (1.0 + 1e-6));
  }
  if (near) {
    atomicOr(&d.pactive[i], 8u);
    atomicOr(&d.pactive[j], 8u);
  }
}

__device__ __forceinline__ void pair_test(Dev& d, uint64_t q) {
  uint32_t i = d.pairs[2 * q], j = d.pairs[2 * q + 1];
  const DevPart& Pi = d.part[i];
  const DevPart& Pj = d.part[j];
  const DevInst& I = d.inst[Pi.inst];
  if (!live(I) || !(Pi.movable || Pj.movable)) return;
  if (d.lists && rebuilds(d, Pi.inst)) list_test(d, i, j, I);
  if (!intersects_box_fext(d.lrtb + 6 * i, d.lrtb + 6 * j, I.Rc)) return;
  atomicAdd(&d.counters[3], 1ull);   // moved into the total by k_finalize
  // bit 1: the particle's nodes are contact candidates this step;
  // bit 0: the particle receives contact terms (only stoppable ones do)
  atomicOr(&d.pactive[i], 2u | (Pi.stoppable ? 1u : 0u));
  atomicOr(&d.pactive[j], 2u | (Pj.stoppable ? 1u : 0u));
}

// Wall prefilter of particle q on its lagged box (timeloop.h:561-570): bit 2
// of pactive marks the particles whose nodes test the wall planes this step.
__device__ __forceinline__ void wall_test(Dev& d, uint64_t q) {
  const DevPart& P = d.part[q];
  const DevInst& I = d.inst[P.inst];
  if (!live(I) || !(P.movable && P.stoppable && I.allow_wall)) return;
  if (!within_interior_fext(d.lrtb + 6 * q, I.wall, I.Rc)) atomicOr(&d.pactive[q], 4u);
}

__device__ __forceinline__ bool self_active(const DevPart& P, const DevInst& I) {
  return P.movable && I.enable_fracture && P.breakable && I.self_contact && P.stoppable;
}

// The node's cell in its job's grid (wrapped coordinates).
__device__ __forceinline__ void cell_of(const double* cp, const DevInst& I, uint32_t* w) {
  for (int k = 0; k < 3; k++) w[k] = wrap_cell(cp[k], I.cell, I.gn[k]);
}

// One launch per step for the independent per-step tests (thread q: pair q,
// particle q, job q) and the cell keys of the nodes of the jobs that rebuild
// their candidate lists.  The cell table (cstart) is zero when a rebuild
// starts: cleared at reservation, then by k_finalize after every rebuild.
__global__ __launch_bounds__(256) void k_cellkey(Dev d) {
  if (halted(d)) return;
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < d.npairs) pair_test(d, q);
  if (q < d.P) wall_test(d, q);
  if (!d.lists) return;   // no job of the pool can have contact: no lists
  if (q < d.ninst && rebuilds(d, (uint32_t)q)) {
    atomicOr(&d.ctl->any_rebuild, 1u);
    d.rjobs[atomicAdd(&d.ctl->nrj, 1u)] = (uint32_t)q;   // its cell blocks are cleared by k_finalize
  }
  const uint64_t g = q;
  if (g >= d.N) return;
  const uint32_t p = d.pid[g];
  const DevPart& P = d.part[p];
  if (!rebuilds_node(d, P.inst)) return;
  const DevInst& I = d.inst[P.inst];
  uint32_t key = kNoKey;
  if (!I.nl_bdry_only || d.brank[g] != kNoKey) {
    double4 x = d.X[g];
    double4 u = disp_of(d, g);
    double cp[3] = {x.x + u.x, x.y + u.y, x.z + u.z};
    uint32_t w[3];
    cell_of(cp, I, w);
    key = cell_key(I, w[0], w[1], w[2]);
    atomicAdd(&d.cstart[key], 1u);
    d.cpos[g] = make_double4(cp[0], cp[1], cp[2], 0.0);
  }
  d.ckey[g] = key;
}

// Exclusive scan of the cell counts of the rebuilding jobs, each job's
// cells on their own: cstart[k] becomes node0 + the number of the job's
// keyed nodes in cells before k, so a job's nodes in cell order fill the
// slots [node0, node0 + jkeyed[job]) of csorted / cps, inside its own node
// range.  Scan block b holds the keys of node block b, so it belongs to one
// job; blocks of jobs that do not rebuild are left alone (their counts are
// zero and nothing reads them).  Two phases: per-block scan (k_scan1) and
// add-back (k_scan3).
// 256 threads per scan block, four keys each (one 16-byte load).
constexpr int kScanThreads = kScanBlock / 4;
__global__ __launch_bounds__(kScanThreads) void k_scan1(Dev d) {
  if (d.ctl->halt || !d.ctl->any_rebuild) return;
  const uint32_t b = blockIdx.x;
  const uint32_t job = d.part[ld_uniform(d.pid, b * 256)].inst;
  if (!rebuilds_node(d, job)) return;
  // the thread's four counts, then an inclusive scan of the threads' totals
  // within each wavefront and the wavefronts' totals
  __shared__ uint32_t wsum[kScanThreads / 64];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint4* a = reinterpret_cast<uint4*>(d.cstart + (uint64_t)b * kScanBlock) + threadIdx.x;
  const uint4 v = *a;
  const uint32_t t = v.x + v.y + v.z + v.w;
  uint32_t x = t;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  for (uint32_t k = 0; k < wv; k++) x += wsum[k];
  const uint32_t e = x - t;   // exclusive prefix of the thread's first key
  *a = make_uint4(e, e + v.x, e + v.x + v.y, e + v.x + v.y + v.z);
  if (threadIdx.x == kScanThreads - 1) d.scan_tmp[b] = x;
}
// Add-back: each block sums the totals of its job's blocks before it itself
// (exact integer sums, L2-resident); the job's last block records the job's
// keyed node count.
__global__ __launch_bounds__(kScanThreads) void k_scan3(Dev d) {
  if (d.ctl->halt || !d.ctl->any_rebuild) return;
  const uint32_t b = blockIdx.x;
  const uint32_t job = d.part[ld_uniform(d.pid, b * 256)].inst;
  if (!rebuilds_node(d, job)) return;
  const DevInst& I = d.inst[job];
  const uint32_t b0 = I.gtb / kScanBlock;
  __shared__ uint32_t red[kScanThreads / 64];
  uint32_t part = 0;
  for (uint32_t j = b0 + threadIdx.x; j < b; j += kScanThreads) part += d.scan_tmp[j];
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  uint32_t prefix = 0;
  for (int w = 0; w < kScanThreads / 64; w++) prefix += red[w];
  const uint32_t o = prefix + (uint32_t)I.node0;
  uint4* a = reinterpret_cast<uint4*>(d.cstart + (uint64_t)b * kScanBlock) + threadIdx.x;
  uint4 v = *a;
  v.x += o;
  v.y += o;
  v.z += o;
  v.w += o;
  *a = v;
  reinterpret_cast<uint4*>(d.ccur + (uint64_t)b * kScanBlock)[threadIdx.x] = v;
  if (threadIdx.x == 0 && b == (uint32_t)((I.node1 - 1) / 256)) d.jkeyed[job] = prefix + d.scan_tmp[b];
}

// Counting sort of the keyed nodes into cell order (csorted, and cps: their
// CurrPos with the node index in the w word), plus the per-node work of a
// rebuild that needs no search: the displacement the list is built at (the
// reference of the skin test), and an empty list for the nodes that cannot
// have a candidate -- with nl_bdry_only the nodes outside the boundary list,
// and the nodes of a particle with no self-contact and no other particle
// within reach (bit 3 of pactive, pair_test).
__global__ __launch_bounds__(256) void k_scatter(Dev d) {
  if (halted(d) || !d.ctl->any_rebuild) return;
  uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= d.N) return;
  const uint32_t p = d.pid[g];
  const DevPart& P = d.part[p];
  // only the nodes of rebuilding jobs hold a key of this step
  if (!rebuilds_node(d, P.inst)) return;
  const DevInst& I = d.inst[P.inst];
  const double4 ui = disp_of(d, g);
  d.UBXY[g] = make_double2(ui.x, ui.y);
  d.UBZ[g] = ui.z;
  const uint32_t key = d.ckey[g];
  if (key != kNoKey) {
    const uint32_t slot = atomicAdd(&d.ccur[key], 1u);
    d.csorted[slot] = (uint32_t)g;
    double4 cp = d.cpos[g];
    cp.w = __longlong_as_double((long long)g);
    d.cps[slot] = cp;
  }
  if (key == kNoKey || (!self_active(P, I) && !(d.pactive[p] & 8u))) d.ccnt[g] = 0;
}

// Candidate list of node g: the nodes j != g of its job within rb, same
// particle ones only while self-contact is on, sorted by the reference's
// accumulation order (same particle first by node index, then the others by
// global node index = particle order, then node order).
//
// Two forms.  Without nl_bdry_only (k_build) a list entry, (other particle)
// << 31 | node index, is its own order key (particle order is node order),
// and the list is kept in registers: kMaxCand words, ascending, empty slots
// all ones; a candidate is passed through every slot with a compare-exchange
// (static indices, no private memory), dropped if equal to one -- the
// duplicate a wrapped grid axis of fewer than three cells gives -- and the
// pass stops once the value carried is empty in every lane.  In crushed
// regions lists fill up, and an insertion sort in private memory would be a
// long chain of dependent memory round trips.  With nl_bdry_only the order is
// the boundary lists' (k_build_bdry).
//
// A candidate's distance is taken between the CurrPos of this step that
// k_cellkey stored (cpos, and cps in cell-list order): it differs from the
// reference's (u_j - u_i) + (X_j - X_i) only by rounding, far inside the
// margins of rb.

// A node's list is longer than the pool's capacity: the pool halts (the
// host grows the lists and runs the step again) and the job records the
// length it needs.
__device__ __forceinline__ void list_overflow(const Dev& d, uint32_t job, uint32_t need) {
  atomicMax(&d.jneed[job], need);
  atomicMax(&d.ctl->need, need);
  atomicOr(&d.ctl->halt, 1u);
}

// Thread t builds the list of the node in cell-list slot t, so the lanes of a
// wavefront hold nodes of neighbouring cells of one job's grid and walk
// nearly the same cells.  The 27 cells around a node are 9 rows of three
// cells adjacent in key order: each row is one run of cell-list slots,
// [cstart[k - 1], cstart[k + 2]), unless the row wraps around its axis (then
// three runs of one cell).
__global__ __launch_bounds__(256) void k_build(Dev d) {
  if (halted(d) || !d.ctl->any_rebuild) return;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.N) return;
  // slot t of a rebuilding job's cell-ordered nodes lies in the job's node
  // range: the slot's own node tells the job
  const uint32_t jt = d.part[d.pid[t]].inst;
  if (!rebuilds_node(d, jt) || t - d.inst[jt].node0 >= d.jkeyed[jt]) return;
  const double4 ci = d.cps[t];
  const uint64_t g = (uint64_t)__double_as_longlong(ci.w);
  const uint32_t p = d.pid[g];
  const DevPart& P = d.part[p];
  const DevInst& I = d.inst[P.inst];
  if (I.nl_bdry_only) return;   // k_build_bdry
  const bool selfc = self_active(P, I);
  if (!selfc && !(d.pactive[p] & 8u)) return;   // empty list (k_scatter)
  const double cpi[3] = {ci.x, ci.y, ci.z};
  uint32_t w[3];
  cell_of(cpi, I, w);
  const bool xrun = I.gn[0] >= 3 && w[0] >= 1 && w[0] + 1 < I.gn[0];
  // candidates within rb (a superset filter: the contact terms apply the
  // exact tests, and rb's margins dwarf the rounding of its square)
  const double rb2 = I.rb * I.rb;
  constexpr uint32_t kChunk = 2;
  constexpr uint32_t kEmpty = ~0u;
  uint32_t* cl = d.clist + (g >> 6) * (uint64_t)(kSlice * d.ccap) + (g & 63);
  // passes of kMaxCand entries: pass k keeps the kMaxCand smallest entries
  // above the largest one of pass k-1
  uint32_t total = 0, lb = 0;
  bool more = true, first = true;
  while (more) {
    uint32_t wl[kMaxCand];
#pragma unroll
    for (int i = 0; i < kMaxCand; i++) wl[i] = kEmpty;
    bool over = false;
    // 9 rows of three cells; a row's bounds are loaded one row ahead.  A row
    // that wraps around the x axis is walked as three runs of one cell.
    uint32_t wy = wrap_add(w[1], -1, I.gn[1]), wz = wrap_add(w[2], -1, I.gn[2]);
    uint32_t row = cell_key(I, 0, wy, wz);
    uint32_t nb0 = xrun ? d.cstart[row + w[0] - 1] : 0, nb1 = xrun ? d.cstart[row + w[0] + 2] : 0;
    for (int r = 0; r < 9; r++) {
      const uint32_t cur = row, rb0 = nb0, rb1 = nb1;
      if (r < 8) {
        wy = wrap_add(w[1], (r + 1) % 3 - 1, I.gn[1]);
        wz = wrap_add(w[2], (r + 1) / 3 - 1, I.gn[2]);
        row = cell_key(I, 0, wy, wz);
        if (xrun) {
          nb0 = d.cstart[row + w[0] - 1];
          nb1 = d.cstart[row + w[0] + 2];
        }
      }
      for (int part = 0; part < (xrun ? 1 : 3); part++) {
        uint32_t b0 = rb0, b1 = rb1;
        if (!xrun) {
          const uint32_t k = cur + wrap_add(w[0], part - 1, I.gn[0]);
          b0 = d.cstart[k];
          b1 = d.cstart[k + 1];
        }
        // the run's records, kChunk loads in flight at a time
        for (uint32_t it = b0; it < b1; it += kChunk) {
          double4 cj[kChunk];
#pragma unroll
          for (uint32_t k = 0; k < kChunk; k++) cj[k] = d.cps[it + k < b1 ? it + k : b0];
#pragma unroll
          for (uint32_t k = 0; k < kChunk; k++) {
            // the candidate's list word, or kEmpty: jobs and particles hold
            // contiguous node ranges, so membership is a range test (wrapped
            // cells of a wide job may hold far nodes of it)
            const uint32_t j = (uint32_t)__double_as_longlong(cj[k].w);
            const bool same = j >= P.node_off && j < P.node_off + P.nnodes;
            uint32_t x = (same ? 0u : kBroken) | j;
            const double ex = cj[k].x - ci.x, ey = cj[k].y - ci.y, ez = cj[k].z - ci.z;
            const bool take = it + k < b1 && j != g && j >= I.node0 && j < I.node1 && (selfc || !same) &&
                              (first || x > lb) && (ex * ex + ey * ey) + ez * ez <= rb2;
            x = take ? x : kEmpty;
            // through the sorted registers
#pragma unroll
            for (int i0 = 0; i0 < kMaxCand; i0 += 8) {
              if (__builtin_amdgcn_ballot_w64(x != kEmpty) == 0) continue;   // nothing left to place
#pragma unroll
              for (int i = i0; i < i0 + 8; i++) {
                const uint32_t y = wl[i];
                x = x == y ? kEmpty : x;
                wl[i] = x < y ? x : y;
                x = x < y ? y : x;
              }
            }
            over |= x != kEmpty;   // a candidate fell off the end: another pass
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kMaxCand; i++)
      if (wl[i] != kEmpty) {
        if (total < d.ccap) cl[(uint64_t)total * kSlice] = wl[i];
        total++;
      }
    more = over;
    lb = wl[kMaxCand - 1];
    first = false;
  }
  if (total > d.ccap) {
    list_overflow(d, P.inst, total);
    total = d.ccap;
  }
  d.ccnt[g] = (uint8_t)total;
}

// Jobs with nl_bdry_only: candidates ordered by (other particle, particle
// order, rank in the boundary list), kept by insertion into private memory,
// in passes of kMaxCand as above.
__global__ __launch_bounds__(256) void k_build_bdry(Dev d) {
  if (halted(d) || !d.ctl->any_rebuild) return;
  uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= d.N) return;
  const uint32_t p = d.pid[g];
  const DevPart& P = d.part[p];
  if (!rebuilds_node(d, P.inst)) return;
  const DevInst& I = d.inst[P.inst];
  if (!I.nl_bdry_only || d.brank[g] == kNoKey) return;   // other form, or an empty list (k_scatter)
  const bool selfc = self_active(P, I);
  if (!selfc && !(d.pactive[p] & 8u)) return;           // empty list (k_scatter)
  const double4 ci = d.cpos[g];
  const double cpi[3] = {ci.x, ci.y, ci.z};
  uint32_t w[3];
  cell_of(cpi, I, w);
  uint32_t* cl = d.clist + (g >> 6) * (uint64_t)(kSlice * d.ccap) + (g & 63);
  // the 27 cells' keys, each once (an axis of fewer than three cells wraps
  // onto itself)
  uint32_t seen[27];
  int nseen = 0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const uint32_t key = cell_key(I, wrap_add(w[0], dx, I.gn[0]), wrap_add(w[1], dy, I.gn[1]),
                                      wrap_add(w[2], dz, I.gn[2]));
        bool dup = false;
        for (int q = 0; q < nseen; q++) dup |= (seen[q] == key);
        if (!dup) seen[nseen++] = key;
      }
  uint32_t total = 0;
  uint64_t lbkey = 0;
  bool more = true, first = true;
  while (more) {
    uint32_t cand[kMaxCand];
    uint64_t ckey[kMaxCand];   // order key: other particle, particle order, rank in the boundary list
    uint32_t ncand = 0;
    bool over = false;
    for (int q = 0; q < nseen; q++) {
      const uint32_t b0 = d.cstart[seen[q]], b1 = d.cstart[seen[q] + 1];
      for (uint32_t it = b0; it < b1; it++) {
        const uint32_t j = d.csorted[it];
        if (j == g) continue;
        if (j < I.node0 || j >= I.node1) continue;
        const bool same = j >= P.node_off && j < P.node_off + P.nnodes;
        if (same && !selfc) continue;
        const double4 cj = d.cps[it];
        const double dn = nrm3(cj.x - ci.x, cj.y - ci.y, cj.z - ci.z);
        if (!(dn <= I.rb)) continue;
        const uint64_t key2 =
            ((uint64_t)!same << 63) | ((uint64_t)d.part[d.pid[j]].node_off << 32) | (uint64_t)d.brank[j];
        if (!first && key2 <= lbkey) continue;   // placed by an earlier pass
        if (ncand == kMaxCand) {   // keep the kMaxCand smallest keys
          over = true;
          if (key2 > ckey[kMaxCand - 1]) continue;
          ncand--;
        }
        uint32_t w = ncand++;
        while (w > 0 && ckey[w - 1] > key2) {
          cand[w] = cand[w - 1];
          ckey[w] = ckey[w - 1];
          w--;
        }
        cand[w] = (same ? 0u : kBroken) | j;
        ckey[w] = key2;
      }
    }
    for (uint32_t k = 0; k < ncand; k++, total++)
      if (total < d.ccap) cl[(uint64_t)total * kSlice] = cand[k];
    more = over;
    if (ncand) lbkey = ckey[ncand - 1];
    first = false;
  }
  if (total > d.ccap) {
    list_overflow(d, P.inst, total);
    total = d.ccap;
  }
  d.ccnt[g] = (uint8_t)total;
}

// Is j an intact bond of node g's row?  Rows hold the post-break state of
// this step (the force loop above already erased this step's breaks), as
// std::find over NbdArr[i] sees it (contact.h:783, timeloop.h:552-557).
__device__ bool bonded_alive(const Dev& d, uint64_t g, uint32_t j, bool sorted) {
  const uint64_t t = (g & ~255ull) + d.node_lane[g];   // the node's slot
  const uint32_t* row = d.bond + d.soff[t >> 6] + (t & 63);
  uint32_t n = d.deg[g];
  if (sorted) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      uint32_t e = row[(uint64_t)mid * kSlice] & kIdxMask;
      if (e < j) lo = mid + 1;
      else hi = mid;
    }
    // duplicates of j share one stretch and break together
    for (uint32_t k = lo; k < n; k++) {
      uint32_t e = row[(uint64_t)k * kSlice];
      if ((e & kIdxMask) != j) break;
      if (!(e & kBroken)) return true;
    }
    return false;
  }
  for (uint32_t k = 0; k < n; k++) {
    uint32_t e = row[(uint64_t)k * kSlice];
    if (e == j) return true;
  }
  return false;
}

// ---------------------------------------------------------------------------
// Contact, wall and external force of node g, added to its PD sum f in the
// reference's order (timeloop.h:544-651): self-contact (q ascending), wall
// planes, external force, torque, then pair contact in pair order.
//
// The bond kernel calls it once the node's row is summed.  Everything it
// reads is final before the bond kernel starts (positions, the start-of-step
// velocities in Vold, candidate lists, lagged boxes, active pairs and wall
// prefilter bits, centroids) except the bond state, and of that it reads only
// node g's own row, which the same thread has just updated -- the post-break
// row std::find sees (contact.h:783).
//
// The candidate list of g holds every partner within rb, same particle
// first (q ascending), then the other particles by global node index (=
// particle order, then node order); the exact reference tests pick the
// pairs within R_c.
//
// The node's external force and its particle's contact flags are loaded
// before the row (ContactIn), so their latency hides behind the bond sum.
struct ContactIn {
  double e0, e1, e2;   // external force density
  uint32_t pa;         // pactive of the particle
};

__device__ __forceinline__ ContactIn contact_in(const Dev& d, uint64_t g, uint32_t p, bool movable) {
  ContactIn c{0.0, 0.0, 0.0, d.pactive[p]};
  if (movable) {
    c.e0 = d.E[3 * g];
    c.e1 = d.E[3 * g + 1];
    c.e2 = d.E[3 * g + 2];
  }
  return c;
}

__device__ __forceinline__ void add_contact(const Dev& d, uint64_t g, uint32_t p, const DevPart& P,
                                            const DevInst& I, const ContactIn& in, double& f0, double& f1,
                                            double& f2) {
  const double dt = I.dt;
  double f[3] = {f0, f1, f2};
  const bool selfc = self_active(P, I);
  const bool pairc = (in.pa & 1u) != 0;
  const bool wall = (in.pa & 4u) && (!I.nl_bdry_only || d.brank[g] != kNoKey);
  const uint32_t n = (selfc || pairc) ? (uint32_t)d.ccnt[g] : 0u;
  const uint32_t* cl = d.clist + (g >> 6) * (uint64_t)(kSlice * d.ccap) + (g & 63);
  double4 xi = make_double4(0.0, 0.0, 0.0, 0.0), ui = xi;
  double vi[3] = {0.0, 0.0, 0.0};
  if (n || wall || I.enable_torque) {
    xi = d.X[g];
    ui = disp_of(d, g);
  }
  if (n || wall) {
    vi[0] = d.Vold[3 * g];
    vi[1] = d.Vold[3 * g + 1];
    vi[2] = d.Vold[3 * g + 2];
  }
  const double cpi[3] = {xi.x + ui.x, xi.y + ui.y, xi.z + ui.z};

  // (b) self-contact, action only (contact.h:779-796 with is_self_contact=1),
  // over the leading same-particle entries
  uint32_t k = 0;
  for (; k < n; k++) {
    const uint32_t j = cl[(uint64_t)k * kSlice];
    if (j & kBroken) break;   // the other particles' entries start here
    if (!selfc) continue;
    const double4 xj = d.X[j];
    const double4 uj = disp_of(d, j);
    const double xx[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
    const double dd[3] = {(uj.x - ui.x) + xx[0], (uj.y - ui.y) + xx[1], (uj.z - ui.z) + xx[2]};
    const double dn = nrm3(dd[0], dd[1], dd[2]);
    if (!(dn <= I.Rc)) continue;   // the reference's distance test
    if (bonded_alive(d, g, j, P.rows_sorted)) continue;
    const double xin = nrm3(xx[0], xx[1], xx[2]);
    double len = I.Rc, K = I.Kn * I.Rc;
    if (xin < I.Rc) {
      len = xin;
      K = P.cnot;
    }
    const double rv[3] = {vi[0] - d.Vold[3 * j], vi[1] - d.Vold[3 * j + 1], vi[2] - d.Vold[3 * j + 2]};
    contact_terms(I, dt, kSelf, dd, dn, len, K, xj.w, xi.w, P.rho, rv, f);
  }

  // (c) wall contact (timeloop.h:561-570, contact.h:1060-1124)
  if (wall) {
    const uint32_t m = wall_node(I, dt, cpi, vi, xi.w, P.rho, f, nullptr);
    if (reaction_step(d, I)) {
      d.wmask[g] = (uint8_t)m;
      if (m) d.bflag[g >> 8] = 1;
    }
  }

  // (d) external force density (timeloop.h:581-597)
  if (P.movable) {
    if (I.gradient_extforce) {
      // ramp min(t / extforce_maxstep, 1) (timeloop.h:582-587)
      double sc = (double)I.t / (double)I.extforce_maxstep;
      if (sc > 1) sc = 1;
      f[0] += sc * in.e0;
      f[1] += sc * in.e1;
      f[2] += sc * in.e2;
    } else {
      f[0] += in.e0;
      f[1] += in.e1;
      f[2] += in.e2;
    }
  }

  // (d') external torque about the centroid, every particle (timeloop.h:601-633)
  if (I.enable_torque) {
    const uint32_t ta = P.taxis;
    double r[3] = {cpi[0] - d.cent[3 * p], cpi[1] - d.cent[3 * p + 1], cpi[2] - d.cent[3 * p + 2]};
    const double rn = nrm3(r[0], r[1], r[2]);
    const double rinv = rn != 0 ? 1 / rn : 0;
    r[ta] = 0;
    double rp[3];
    rp[ta] = 0;
    rp[(ta + 1) % 3] = -r[(ta + 2) % 3];
    rp[(ta + 2) % 3] = r[(ta + 1) % 3];
    const double pn = nrm3(rp[0], rp[1], rp[2]);
    if (pn > 0) {
      for (int c = 0; c < 3; c++) rp[c] = rp[c] / pn;
    } else {
      rp[(ta + 1) % 3] = 0;
      rp[(ta + 2) % 3] = 0;
    }
    const double sc = rinv * P.tval;
    for (int c = 0; c < 3; c++) f[c] += sc * rp[c];
  }

  // (e) inter-particle contact in pair order (timeloop.h:636-651) over the
  // remaining entries
  for (; pairc && k < n; k++) {
    const uint32_t j = cl[(uint64_t)k * kSlice] & kIdxMask;
    const uint32_t pj = d.pid[j];
    const DevPart& Q = d.part[pj];
    // pair (lo, hi) in the job's particle order, which is node order
    const bool below = Q.node_off < P.node_off;
    const uint32_t lo = below ? pj : p, hi = below ? p : pj;
    if (!(d.part[lo].movable || d.part[hi].movable)) continue;
    if (!intersects_box_fext(d.lrtb + 6 * lo, d.lrtb + 6 * hi, I.Rc)) continue;
    const double4 xj = d.X[j];
    const double4 uj = disp_of(d, j);
    double dd[3], rv[3];
    if (below) {   // the partner is P_C's node i, we are P_N's node j
      dd[0] = (ui.x - uj.x) + (xi.x - xj.x);
      dd[1] = (ui.y - uj.y) + (xi.y - xj.y);
      dd[2] = (ui.z - uj.z) + (xi.z - xj.z);
    } else {       // we are P_C's node i
      dd[0] = (uj.x - ui.x) + (xj.x - xi.x);
      dd[1] = (uj.y - ui.y) + (xj.y - xi.y);
      dd[2] = (uj.z - ui.z) + (xj.z - xi.z);
    }
    const double dn = nrm3(dd[0], dd[1], dd[2]);
    if (!(dn <= I.Rc)) continue;   // the reference's distance test
    const double vj[3] = {d.Vold[3 * j], d.Vold[3 * j + 1], d.Vold[3 * j + 2]};
    if (below) {   // reaction terms
      for (int c = 0; c < 3; c++) rv[c] = vj[c] - vi[c];
      contact_terms(I, dt, kReaction, dd, dn, I.Rc, I.Kn * I.Rc, xi.w, xj.w, Q.rho, rv, f);
    } else {
      for (int c = 0; c < 3; c++) rv[c] = vi[c] - vj[c];
      contact_terms(I, dt, kAction, dd, dn, I.Rc, I.Kn * I.Rc, xj.w, xi.w, P.rho, rv, f);
    }
  }
  f0 = f[0];
  f1 = f[1];
  f2 = f[2];
}

// Verlet velocity half (timeloop.h:662-682) of node g from its final force,
// after its contact terms; V and A of g are read and written by this thread
// only (the contact terms of every node read Vold).  A clamped node's
// velocity and acceleration are zeroed here and its displacement by k_clamp
// after the launch, since other nodes' contact terms still read U
// (timeloop.h:767-774).  The prescribed-velocity particle is updated in full
// and clamped after its prescription (k_pv_apply).
__device__ __forceinline__ void velocity_half(const Dev& d, uint64_t g, uint32_t p, const DevPart& P,
                                              const DevInst& I, double f0, double f1, double f2) {
  if (!P.movable) return;
  const bool vc = I.vc_on && p == I.vc_slot;
  if ((d.nflag[g] & 1u) && !vc) {
    for (int k = 0; k < 3; k++) d.V[3 * g + k] = d.A[3 * g + k] = 0.0;
    return;
  }
  const double hdt = 0.5 * I.dt;
  const double f[3] = {f0, f1, f2};
  for (int k = 0; k < 3; k++) {
    const double aold = d.A[3 * g + k];
    const double anew = P.rinv * f[k];
    if (vc) d.AFD[3 * g + k] = aold;
    d.A[3 * g + k] = anew;
    d.V[3 * g + k] = d.V[3 * g + k] + ((aold + anew) * hdt);
  }
}

// The end of node g's step in the bond kernel: contact terms, then the
// velocity half.  In a wavefront over one particle (all but the few that
// straddle two) the particle's and job's records are wave-uniform and are
// read through the scalar cache -- they are not written during the launch --
// instead of as a chain of per-lane loads.
__device__ __forceinline__ void finish_node(const Dev& d, uint64_t g, uint32_t p, const ContactIn& cin,
                                            double& f0, double& f1, double& f2) {
  const uint32_t pu = __builtin_amdgcn_readfirstlane(p);
  if (__builtin_amdgcn_ballot_w64(p != pu) == 0) {
    const DevPart Pu = ld_uniform(d.part, pu);
    const DevInst Iu = ld_uniform(d.inst, Pu.inst);
    add_contact(d, g, pu, Pu, Iu, cin, f0, f1, f2);
    velocity_half(d, g, pu, Pu, Iu, f0, f1, f2);
  } else {
    const DevPart& P = d.part[p];
    const DevInst& I = d.inst[P.inst];
    add_contact(d, g, p, P, I, cin, f0, f1, f2);
    velocity_half(d, g, p, P, I, f0, f1, f2);
  }
}

// What the row loop needs of the node's particle and job, through the
// scalar cache in single-particle wavefronts.
struct NodeHead {
  bool live, movable, frac;   // frac: breaks enabled for this particle
  double cnot, snot;
};

__device__ __forceinline__ NodeHead node_head(const Dev& d, uint32_t p) {
  const uint32_t pu = __builtin_amdgcn_readfirstlane(p);
  if (__builtin_amdgcn_ballot_w64(p != pu) == 0) {
    const DevPart P = ld_uniform(d.part, pu);
    const DevInst I = ld_uniform(d.inst, P.inst);
    return {live(I), P.movable != 0, I.enable_fracture && P.breakable, P.cnot, P.snot};
  }
  const DevPart& P = d.part[p];
  const DevInst& I = d.inst[P.inst];
  return {live(I), P.movable != 0, I.enable_fracture && P.breakable, P.cnot, P.snot};
}

// ---------------------------------------------------------------------------
// Bond-based PD force + critical-stretch break (get_peridynamic_force,
// particle2.h:107-193; remove_bonds, particle2.h:196-237).
//
// One thread per node sums its row in row order, exactly like the reference:
// no atomics, no reordering.  The 64 slots of a wavefront share one bond
// slice (entry k of lane l at soff + k*64 + l), so the k-th entries of their
// rows are one coalesced load; within each 256-node block the slots hold the
// nodes by decreasing row length, so a slice's width (its longest row)
// leaves few padding entries.  The stretch is tested against snot right after
// the term is summed; a broken entry stops contributing from the next step
// on, as the reference's erase after the force pass does.
//
// Per term (particle2.h:139-153):
//   e = (u_j - u_i) + xi,  s = (|e| - |xi|) / |xi|,  t = (cnot*s*V_j) * (e/|e|)
// Two to eight entries are processed per iteration as straight-line code, so
// their dependency chains interleave.  In the fast domain sqrt and division are the
// backend's own sequences without their range wrappers, the divisors'
// reciprocals refined from the square roots' own reciprocal estimates
// (bitwise equal to IEEE sqrt and division); whenever a lane leaves the
// domain, the whole batch uses the IEEE operators.  A dead
// (broken or padding) entry reads the node itself, gets unit lengths and a
// zero coefficient, and adds +-0 to a sum that is never -0: no effect.
//
// Workgroups are mapped so that each XCD (blocks are dealt round-robin over
// the 8 XCDs) runs one contiguous range of nodes: neighbouring tiles, which
// gather the same neighbour records, share that XCD's L2.
constexpr int kXcds = 8;

// High word of a double, and a double with its high word replaced.
__device__ __forceinline__ uint32_t hi_word(double v) {
  return (uint32_t)((unsigned long long)__double_as_longlong(v) >> 32);
}
__device__ __forceinline__ double with_hi(double v, uint32_t h) {
  const unsigned long long lo = (unsigned long long)__double_as_longlong(v) & 0xFFFFFFFFull;
  return __longlong_as_double((long long)(((unsigned long long)h << 32) | lo));
}
// |e|^2 in [2^-600, 2^600) <=> (high word - kN2Lo) < kN2Span as unsigned
// (0, denormals, NaN, inf and negative values all fall outside).
constexpr uint32_t kN2Lo = 0x1A700000u;                 // high word of 2^-600
constexpr uint32_t kN2Span = 0x65700000u - kN2Lo;       // high word of 2^600, minus kN2Lo
constexpr uint32_t kOneHi = 0x3FF00000u;                // high word of 1.0 (low word 0)

// The slice's bond entries as one uniform pointer per batch entry plus the
// lane's byte offset, so index loads and break stores take no vector address
// arithmetic.  T: uint32_t (pool node indices) or uint16_t (positions in the
// block's node table).
template <class T>
struct SliceRows {
  const char* base;   // entry 0 of the slice (wave-uniform)
  uint32_t lane_off;  // sizeof(T) * lane
  __device__ __forceinline__ T* at(uint32_t k) const {
    return (T*)(base + (uint64_t)k * (kSlice * sizeof(T)) + lane_off);
  }
};

// The k-th entries of the U rows-batches starting at k (pad past the width,
// W >= 1).  Every load is issued -- past the width it rereads entry W-1 and
// the value is replaced -- so the loop body has no load under a branch and
// the compiler can count outstanding loads exactly instead of draining them.
// The entries are read once per launch: non-temporal loads, which leave the
// caches to the neighbour records.
template <int U, class T>
__device__ __forceinline__ void bond_idx(const SliceRows<T>& R, uint32_t k, uint32_t W, uint32_t pad,
                                         uint32_t (&e)[U]) {
#pragma unroll
  for (int u = 0; u < U; u++) {
    const bool in = k + u < W;
    const uint32_t v = __builtin_nontemporal_load(R.at(in ? k + u : W - 1));
    e[u] = in ? v : pad;
  }
}

// Node j's neighbour record, read from the split copies (x,y | z,vol |
// ux,uy | uz).  OFF32: the pool spans fewer than 2^27 node slots, so each
// byte offset is a 32-bit shift and the loads take the array bases from
// scalar registers.
template <bool OFF32>
__device__ __forceinline__ void rec_u(const Dev& d, uint32_t j, double4& u) {
  double2 uxy;
  double uz;
  if (OFF32) {
    uxy = *(const double2*)((const char*)d.UXY + (j << 4));
    uz = *(const double*)((const char*)d.UZ + (j << 3));
  } else {
    uxy = d.UXY[j];
    uz = d.UZ[j];
  }
  u = make_double4(uxy.x, uxy.y, uz, 0.0);
}
template <bool OFF32>
__device__ __forceinline__ void rec(const Dev& d, uint32_t j, double4& x, double4& u) {
  double2 xy, zv;
  if (OFF32) {
    xy = *(const double2*)((const char*)d.XY + (j << 4));
    zv = *(const double2*)((const char*)d.ZV + (j << 4));
  } else {
    xy = d.XY[j];
    zv = d.ZV[j];
  }
  x = make_double4(xy.x, xy.y, zv.x, zv.y);
  rec_u<OFF32>(d, j, u);
}

// Where a bond kernel reads its entries and neighbour records from.
//
// GlobalRows: the u32 entries (pool node index, bit 31 = broken); records
// gathered from the split copies in HBM/L2.  A dead entry reads the node
// itself (branch-free; switching its lanes off for the loads measured slower).
template <bool OFF32>
struct GlobalRows {
  const Dev& d;
  SliceRows<uint32_t> R;
  SliceRows<uint16_t> R16;
  uint32_t g;
  static constexpr uint32_t kPadE = kPad;
  __device__ __forceinline__ const SliceRows<uint32_t>& rows() const { return R; }
  __device__ __forceinline__ bool alive(uint32_t e) const { return (int32_t)e >= 0; }
  template <int U>
  __device__ __forceinline__ void resolve(const uint32_t (&)[U], uint32_t (&)[U]) const {}
  template <int U>
  __device__ __forceinline__ void gather(const uint32_t (&e)[U], const uint32_t (&)[U], double4 (&xj)[U],
                                         double4 (&uj)[U]) const {
#pragma unroll
    for (int u = 0; u < U; u++) rec<OFF32>(d, alive(e[u]) ? e[u] : g, xj[u], uj[u]);
  }
  // both forms of the entry carry the broken flag
  __device__ __forceinline__ void broke(uint32_t k, uint32_t e) const {
    *R.at(k) = e | kBroken;
    *R16.at(k) = (uint16_t)kDead16;
  }
};

// LdsRows: the u16 entries (16 * position in the 256-node block's table of
// the nodes its rows reference, i.e. the byte offset of the node's 16-byte
// LDS records; kDead16 = broken); the block's table and the static part of
// those nodes' records (x, y, z, vol) staged in LDS once per launch, the
// displacement gathered from the split copies through the table.  A dead
// entry uses the node's own records.  The table lookup (resolve) runs one
// pipeline stage before the gather, so the displacement loads issue without
// waiting on LDS.
// One LDS object, the index table first, so that every table's offset fits
// the 16-bit immediate of an LDS load and no base register is needed.
struct LdsTables {
  uint32_t id[kLdsCap];   // 8 * pool node index: the byte offset of its UZ record
  double2 xy[kLdsCap];
  double2 zv[kLdsCap];
};
static_assert(sizeof(LdsTables) - sizeof(double2) * kLdsCap < 65536, "LDS table offsets exceed 16 bits");
__shared__ LdsTables s_tab;
struct LdsRows {
  const Dev& d;
  SliceRows<uint16_t> R16;
  SliceRows<uint32_t> R;
  uint32_t self16;   // 16 * the node's position in the table
  static constexpr uint32_t kPadE = kPad16;
  __device__ __forceinline__ const SliceRows<uint16_t>& rows() const { return R16; }
  __device__ __forceinline__ bool alive(uint32_t e) const { return e < kLive16; }
  __device__ __forceinline__ uint32_t off(uint32_t e) const { return e < kLive16 ? e : self16; }
  __device__ __forceinline__ static uint32_t id_of(uint32_t off16) {
    return *(const uint32_t*)((const char*)s_tab.id + (off16 >> 2));
  }
  template <int U>
  __device__ __forceinline__ void resolve(const uint32_t (&e)[U], uint32_t (&id8)[U]) const {
#pragma unroll
    for (int u = 0; u < U; u++) id8[u] = id_of(off(e[u]));
  }
  template <int U>
  __device__ __forceinline__ void gather(const uint32_t (&e)[U], const uint32_t (&id8)[U], double4 (&xj)[U],
                                         double4 (&uj)[U]) const {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const double2 uxy = *(const double2*)((const char*)d.UXY + (id8[u] << 1));
      const double uz = *(const double*)((const char*)d.UZ + id8[u]);
      uj[u] = make_double4(uxy.x, uxy.y, uz, 0.0);
      const uint32_t t = off(e[u]);
      const double2 xy = *(const double2*)((const char*)s_tab.xy + t);
      const double2 zv = *(const double2*)((const char*)s_tab.zv + t);
      xj[u] = make_double4(xy.x, xy.y, zv.x, zv.y);
    }
  }
  __device__ __forceinline__ void broke(uint32_t k, uint32_t e) const {
    *R.at(k) = (id_of(e) >> 3) | kBroken;
    *R16.at(k) = (uint16_t)kDead16;
  }
};

// Tolerance arithmetic: 1/sqrt(x) from the hardware estimate (relative error
// below 2^-24 on gfx950) refined by kTolNewton Newton-Raphson steps
// y <- y + y (1/2 - (x/2) y^2), each taken with two fused multiply-adds; one
// step leaves a relative error below 2^-47.  |e| = |e|^2 y_e and |xi| =
// |xi|^2 y_xi then carry that error, so the stretch (|e| - |xi|) y_xi is
// within ~1e-14 (absolute) of the correctly rounded one.  The squared norms,
// the stretch's difference and the accumulation of the terms are fused
// multiply-adds.
constexpr int kTolNewton = 1;
__device__ __forceinline__ double rsq_refined(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
#pragma unroll
  for (int k = 0; k < kTolNewton; k++) {
    const double t = __builtin_fma(-(h * y), y, 0.5);
    y = __builtin_fma(y, t, y);
  }
  return y;
}

// Terms of one batch, added to (f0, f1, f2) in row order; breaks recorded.
// TOL: the tolerance arithmetic above in the fast domain instead of the
// correctly rounded square roots and divisions (the IEEE branch is shared).
//
// Fast domain (every lane of the wave): a dead entry carries the node's own
// record, so with finite operands its e and xi are exactly +0 and |e|^2 =
// |xi|^2 = +0; giving both the high word of 1.0 makes them exactly 1.0, the
// stretch exactly +0 and the coefficient +-0 without a select on it.  The
// IEEE branch (safe mode, or a lane outside the domain) selects explicitly.
//
// STRIDE > 1 (bitwise arithmetic only): the U entries are k, k + STRIDE, ...
// of the row, and their terms are handed out in tout instead of added.
template <int U, bool TOL, class Src, int STRIDE = 1>
__device__ __forceinline__ void bond_terms(const Src& S, uint32_t k, const uint32_t (&e)[U], const double4 (&xj)[U],
                                           const double4 (&uj)[U], const double4& xi, const double4& ui, bool fast,
                                           bool can_break, double cnot, double snot, double& f0, double& f1,
                                           double& f2, double (*tout)[3] = nullptr) {
  static_assert(STRIDE == 1 || !TOL, "terms handed out only in bitwise arithmetic");
  bool alive[U];
  double ev[U][3], xv[U][3], n2[U], l2[U];
  uint32_t over = 0;
#pragma unroll
  for (int u = 0; u < U; u++) {
    alive[u] = S.alive(e[u]);
    xv[u][0] = xj[u].x - xi.x;   // xi (gen_xi)
    xv[u][1] = xj[u].y - xi.y;
    xv[u][2] = xj[u].z - xi.z;
    ev[u][0] = (uj[u].x - ui.x) + xv[u][0];
    ev[u][1] = (uj[u].y - ui.y) + xv[u][1];
    ev[u][2] = (uj[u].z - ui.z) + xv[u][2];
    // the tolerance arithmetic contracts the squared norms into fused multiply-adds
    const double a2 = TOL ? __builtin_fma(ev[u][2], ev[u][2], __builtin_fma(ev[u][1], ev[u][1], ev[u][0] * ev[u][0]))
                          : (ev[u][0] * ev[u][0] + ev[u][1] * ev[u][1]) + ev[u][2] * ev[u][2];
    const double b2 = TOL ? __builtin_fma(xv[u][2], xv[u][2], __builtin_fma(xv[u][1], xv[u][1], xv[u][0] * xv[u][0]))
                          : (xv[u][0] * xv[u][0] + xv[u][1] * xv[u][1]) + xv[u][2] * xv[u][2];
    n2[u] = with_hi(a2, alive[u] ? hi_word(a2) : kOneHi);
    l2[u] = with_hi(b2, alive[u] ? hi_word(b2) : kOneHi);
    over = max(over, hi_word(n2[u]) - kN2Lo);
  }
  double str[U], t[U][3];
  bool fused = false;   // t[u][0] holds the coefficient, added as fused multiply-adds
  if (__all(fast & (over < kN2Span))) {
    if (TOL) {
      // |e| |xi| and the coefficient from the refined reciprocal square roots
#pragma unroll
      for (int u = 0; u < U; u++) {
        const double re = rsq_refined(n2[u]), rl = rsq_refined(l2[u]);
        str[u] = __builtin_fma(n2[u], re, -(l2[u] * rl)) * rl;
        t[u][0] = ((cnot * str[u]) * xj[u].w) * re;
      }
      fused = true;
    } else {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const SqrtRcp sn = sqrt_rcp_core(n2[u]);
        const SqrtRcp sl = sqrt_rcp_core(l2[u]);
        str[u] = div_core(sn.s - sl.s, sl.r);
        double c = (cnot * str[u]) * xj[u].w;
#pragma unroll
        for (int q = 0; q < 3; q++) t[u][q] = c * div_core(ev[u][q], sn.r);
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const double a2 = (ev[u][0] * ev[u][0] + ev[u][1] * ev[u][1]) + ev[u][2] * ev[u][2];
      const double b2 = (xv[u][0] * xv[u][0] + xv[u][1] * xv[u][1]) + xv[u][2] * xv[u][2];
      double nn = sqrt(alive[u] ? a2 : 1.0);
      double L = sqrt(alive[u] ? b2 : 1.0);
      str[u] = (nn - L) / L;
      double c = (cnot * str[u]) * xj[u].w;
#pragma unroll
      for (int q = 0; q < 3; q++) t[u][q] = alive[u] ? c * (ev[u][q] / nn) : 0.0;
    }
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    if (STRIDE > 1) {
      tout[u][0] = t[u][0];
      tout[u][1] = t[u][1];
      tout[u][2] = t[u][2];
    } else if (TOL && fused) {
      f0 = __builtin_fma(t[u][0], ev[u][0], f0);
      f1 = __builtin_fma(t[u][0], ev[u][1], f1);
      f2 = __builtin_fma(t[u][0], ev[u][2], f2);
    } else {
      f0 += t[u][0];
      f1 += t[u][1];
      f2 += t[u][2];
    }
    if (alive[u] & can_break & (str[u] > snot)) S.broke(k + u * STRIDE, e[u]);
  }
}

// The rows of a wavefront's slice, W entries each, summed in row order.
// U: entries per batch.  PIPE: software pipelining -- the indices of batch
// k+2 and the neighbour records of batch k+1 are in flight while batch k is
// evaluated, so the index stream (HBM) and the gathers (L2 or LDS) overlap the
// FP64 arithmetic instead of stalling every batch; 2: the same, unrolled
// twice so the records alternate between two register sets.
template <int U, int PIPE, bool TOL, class Src>
__device__ __forceinline__ void bond_rows(const Src& S, uint32_t W, const double4& xi, const double4& ui, bool fast,
                                          bool can_break, double cnot, double snot, double& f0, double& f1,
                                          double& f2) {
  const auto& R = S.rows();
  const uint32_t pad = Src::kPadE;
  if (PIPE == 0) {
    for (uint32_t k = 0; k < W; k += U) {
      uint32_t e[U], a[U];
      double4 xj[U], uj[U];
      bond_idx<U>(R, k, W, pad, e);
      S.template resolve<U>(e, a);
      S.template gather<U>(e, a, xj, uj);
      bond_terms<U, TOL>(S, k, e, xj, uj, xi, ui, fast, can_break, cnot, snot, f0, f1, f2);
    }
  } else if (PIPE == 2 && W) {
    // batch k: records in (xa, ua); k+U: entries resolved, records loading
    // into (xb, ub); k+2U: entries loading, resolved after batch k's terms
    uint32_t ea[U], eb[U], aa[U], ab[U];
    double4 xa[U], ua[U], xb[U], ub[U];
    bond_idx<U>(R, 0, W, pad, ea);
    bond_idx<U>(R, U, W, pad, eb);
    S.template resolve<U>(ea, aa);
    S.template gather<U>(ea, aa, xa, ua);
    S.template resolve<U>(eb, ab);
    for (uint32_t k = 0; k < W; k += 2 * U) {
      uint32_t ec[U], ac[U];
      bond_idx<U>(R, k + 2 * U, W, pad, ec);
      S.template gather<U>(eb, ab, xb, ub);
      bond_terms<U, TOL>(S, k, ea, xa, ua, xi, ui, fast, can_break, cnot, snot, f0, f1, f2);
      if (k + U >= W) break;
      S.template resolve<U>(ec, ac);
      bond_idx<U>(R, k + 3 * U, W, pad, ea);
      S.template gather<U>(ec, ac, xa, ua);
      bond_terms<U, TOL>(S, k + U, eb, xb, ub, xi, ui, fast, can_break, cnot, snot, f0, f1, f2);
      S.template resolve<U>(ea, ab);   // batch k + 3U
#pragma unroll
      for (int u = 0; u < U; u++) {
        eb[u] = ea[u];   // batch k + 3U, resolved
        ea[u] = ec[u];   // batch k + 2U, its records already in xa/ua
      }
    }
  } else if (W) {
    uint32_t e0[U], e1[U], a1[U];
    double4 xa[U], ua[U];
    bond_idx<U>(R, 0, W, pad, e0);
    bond_idx<U>(R, U, W, pad, e1);
    S.template resolve<U>(e0, a1);
    S.template gather<U>(e0, a1, xa, ua);
    S.template resolve<U>(e1, a1);
    for (uint32_t k = 0; k < W; k += U) {
      uint32_t e2[U];
      double4 xb[U], ub[U];
      bond_idx<U>(R, k + 2 * U, W, pad, e2);
      S.template gather<U>(e1, a1, xb, ub);
      bond_terms<U, TOL>(S, k, e0, xa, ua, xi, ui, fast, can_break, cnot, snot, f0, f1, f2);
      S.template resolve<U>(e2, a1);
#pragma unroll
      for (int u = 0; u < U; u++) {
        e0[u] = e1[u];
        e1[u] = e2[u];
        xa[u] = xb[u];
        ua[u] = ub[u];
      }
    }
  }
}

// Workgroup -> node block: blocks are dealt round-robin over the XCDs, so
// each XCD runs one contiguous range of nodes and neighbouring tiles, which
// gather the same records, share that XCD's L2.
__device__ __forceinline__ uint64_t xcd_block() {
  const uint64_t nbp = gridDim.x;   // a multiple of kXcds
  return (blockIdx.x % kXcds) * (nbp / kXcds) + blockIdx.x / kXcds;
}

// The offset and width of slot t's slice, wave-uniform.
__device__ __forceinline__ void slice_of(const Dev& d, uint64_t t, uint64_t& base, uint32_t& W) {
  const uint64_t s = t >> 6;
  const uint64_t b0 = d.soff[s];
  const uint32_t blo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b0);
  const uint32_t bhi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b0 >> 32));
  base = ((uint64_t)bhi << 32) | (uint64_t)blo;
  W = (uint32_t)__builtin_amdgcn_readfirstlane((int)d.swid[s]);
}

// Bond kernel over the u32 entries, records gathered from HBM/L2.
// OFF32: see rec.
template <int U, int PIPE, bool OFF32, bool TOL = false>
__global__ __launch_bounds__(256) void k_bond_t(Dev d) {
  if (halted(d)) return;
  const uint64_t t = xcd_block() * blockDim.x + threadIdx.x;   // slot
  // padded node range: S slices of 64; whole wavefronts exit together
  if (t >= d.S * kSlice) return;
  const uint64_t g = (t & ~255ull) + d.lane_node[t];           // node
  const uint32_t p = d.pid[g];
  const NodeHead H = node_head(d, p);
  if (!H.live) return;             // frozen job or free slot: F keeps its value
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  const ContactIn cin = contact_in(d, g, p, H.movable);
  if (H.movable) {
    uint64_t base;
    uint32_t W;
    slice_of(d, t, base, W);
    const uint32_t lane = (uint32_t)(t & 63);
    const GlobalRows<OFF32> S{d,
                              {(const char*)(d.bond + base), lane * 4u},
                              {(const char*)(d.bloc + base), lane * 2u},
                              (uint32_t)g};
    // the node's own record from the same copies its neighbours read, so
    // each node's data is fetched from HBM once per launch
    double4 xi, ui;
    rec<false>(d, (uint32_t)g, xi, ui);
    bond_rows<U, PIPE, TOL>(S, W, xi, ui, (*d.flags & kFlagSafeMath) == 0, H.frac, H.cnot, H.snot, f0, f1, f2);
  }
  finish_node(d, g, p, cin, f0, f1, f2);
  d.F[3 * g] = f0;
  d.F[3 * g + 1] = f1;
  d.F[3 * g + 2] = f2;
}

// Bond kernel over the u16 entries: the workgroup's 256-node block first
// stages its node table and those nodes' static records in LDS (80 KB: two
// workgroups per CU), then every gather of x, y, z, vol is an LDS read and
// only the displacement is gathered from L2.  Used when every job of the
// pool has block tables of at most kLdsCap nodes.  The tables hold a CU to
// two workgroups (two waves per SIMD), so registers are allotted for two.
template <int U, int PIPE, bool TOL = false>
__global__ __launch_bounds__(256, 2) void k_bond_lds(Dev d) {
  if (halted(d)) return;
  const uint64_t blk = xcd_block();
  const uint64_t t = blk * blockDim.x + threadIdx.x;   // slot
  {
    uint32_t cnt = 0;
    if (blk < d.S / kBlockSlices) cnt = min(d.binfo[blk].x, kLdsCap);
    const uint32_t* tab = d.btab + blk * (uint64_t)kLdsCap;
    // kStage table entries per thread in flight: their index loads go out
    // together, then their gathers (a clamped index rereads entry i0); eight
    // cover a table in one round
    constexpr uint32_t kStage = 8;
    for (uint32_t i0 = threadIdx.x; i0 < cnt; i0 += kStage * blockDim.x) {
      uint32_t id[kStage];
      double2 xy[kStage], zv[kStage];
#pragma unroll
      for (uint32_t k = 0; k < kStage; k++) {
        const uint32_t i = i0 + k * blockDim.x;
        id[k] = tab[i < cnt ? i : i0];
      }
#pragma unroll
      for (uint32_t k = 0; k < kStage; k++) {
        xy[k] = d.XY[id[k]];
        zv[k] = d.ZV[id[k]];
      }
#pragma unroll
      for (uint32_t k = 0; k < kStage; k++) {
        const uint32_t i = i0 + k * blockDim.x;
        if (i < cnt) {
          s_tab.id[i] = id[k] << 3;
          s_tab.xy[i] = xy[k];
          s_tab.zv[i] = zv[k];
        }
      }
    }
    __syncthreads();
  }
  if (t >= d.S * kSlice) return;
  const uint64_t g = (t & ~255ull) + d.lane_node[t];   // node
  const uint32_t p = d.pid[g];
  const NodeHead H = node_head(d, p);
  if (!H.live) return;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  const ContactIn cin = contact_in(d, g, p, H.movable);
  if (H.movable) {
    uint64_t base;
    uint32_t W;
    slice_of(d, t, base, W);
    const uint32_t lane = (uint32_t)(t & 63);
    const uint32_t self = d.binfo[blk].y + (uint32_t)(g - blk * blockDim.x);
    const LdsRows S{d, {(const char*)(d.bloc + base), lane * 2u}, {(const char*)(d.bond + base), lane * 4u},
                    16 * self};
    const double2 xy = s_tab.xy[self], zv = s_tab.zv[self];
    const double4 xi = make_double4(xy.x, xy.y, zv.x, zv.y);
    double4 ui;
    rec_u<true>(d, (uint32_t)g, ui);
    bond_rows<U, PIPE, TOL>(S, W, xi, ui, (*d.flags & kFlagSafeMath) == 0, H.frac, H.cnot, H.snot, f0, f1, f2);
  }
  finish_node(d, g, p, cin, f0, f1, f2);
  d.F[3 * g] = f0;
  d.F[3 * g + 1] = f1;
  d.F[3 * g + 2] = f2;
}

// Entry s of every quad of lanes, broadcast to the quad (DPP quad_perm).
template <int SRC>
__device__ __forceinline__ double quad_lane(double v) {
  constexpr int ctrl = SRC * 0x55;   // every lane of the quad reads lane SRC
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, ctrl, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), ctrl, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Bond kernel for pools too small to fill the GPU (one wavefront per 64
// nodes leaves most SIMDs idle and each lane's row a long dependent chain):
// four lanes per node slot, lane s of the quad evaluating entries s, s + 4,
// s + 8, ... of the row, U per batch, with the next batch's entries and
// records loading while a batch is evaluated.  Every lane of the quad then
// adds the batch's terms in row order, entry 4m + s taken from lane s, so
// the sum is the one-lane row sum operation for operation; breaks are
// recorded by the lane that evaluated the entry, and the quad's first lane
// finishes the node.  The four lanes of a quad share their node, so they
// take the same branches and the exchanges always read active lanes.
template <int U>
__global__ __launch_bounds__(256) void k_bond_quad(Dev d) {
  if (halted(d)) return;
  const uint64_t T = xcd_block() * blockDim.x + threadIdx.x;
  const uint64_t t = T >> 2;                    // slot
  const uint32_t sub = (uint32_t)(T & 3);       // lane in the quad
  if (t >= d.S * kSlice) return;
  const uint64_t g = (t & ~255ull) + d.lane_node[t];
  const uint32_t p = d.pid[g];
  const NodeHead H = node_head(d, p);
  if (!H.live) return;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  const ContactIn cin = contact_in(d, g, p, H.movable);
  if (H.movable) {
    uint64_t base;
    uint32_t W;
    slice_of(d, t, base, W);   // a wavefront's 16 slots lie in one slice
    const uint32_t lane = (uint32_t)(t & 63);
    const GlobalRows<true> S{d,
                             {(const char*)(d.bond + base), lane * 4u},
                             {(const char*)(d.bloc + base), lane * 2u},
                             (uint32_t)g};
    double4 xi, ui;
    rec<true>(d, (uint32_t)g, xi, ui);
    const bool fast = (*d.flags & kFlagSafeMath) == 0;
    // entries k0 + sub + 4u of the batch at k0 (pad past the width)
    auto idx = [&](uint32_t k0, uint32_t(&e)[U]) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint32_t k = k0 + sub + 4u * u;
        const bool in = k < W;
        const uint32_t v = __builtin_nontemporal_load(S.R.at(in ? k : W - 1));
        e[u] = in ? v : GlobalRows<true>::kPadE;
      }
    };
    if (W) {
      uint32_t e0[U], e1[U];
      double4 xa[U], ua[U];
      idx(0, e0);
      idx(4 * U, e1);
      S.template gather<U>(e0, e0, xa, ua);
      for (uint32_t k = 0; k < W; k += 4 * U) {
        uint32_t e2[U];
        double4 xb[U], ub[U];
        idx(k + 8 * U, e2);
        S.template gather<U>(e1, e1, xb, ub);
        double tt[U][3];
        bond_terms<U, false, GlobalRows<true>, 4>(S, k + sub, e0, xa, ua, xi, ui, fast, H.frac, H.cnot, H.snot, f0,
                                                  f1, f2, tt);
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
          for (int q = 0; q < 3; q++) {
            const double t0 = quad_lane<0>(tt[u][q]), t1 = quad_lane<1>(tt[u][q]);
            const double t2 = quad_lane<2>(tt[u][q]), t3 = quad_lane<3>(tt[u][q]);
            double& f = q == 0 ? f0 : (q == 1 ? f1 : f2);
            f += t0;
            f += t1;
            f += t2;
            f += t3;
          }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          e0[u] = e1[u];
          e1[u] = e2[u];
          xa[u] = xb[u];
          ua[u] = ub[u];
        }
      }
    }
  }
  if (sub) return;
  finish_node(d, g, p, cin, f0, f1, f2);
  d.F[3 * g] = f0;
  d.F[3 * g + 1] = f1;
  d.F[3 * g + 2] = f2;
}

// Instantiated shapes of the bond kernel; the autotuner times them on the
// resident pool and keeps the fastest.  Every shape sums each row in row order,
// so all give the same bits.  They address node records with 32-bit offsets;
// pools of 2^27 node slots or more use bond_wide.  lds: the staged kernel,
// usable while every job's block tables fit kLdsCap.
struct BondShape {
  void (*fn)(Dev);
  int unroll, pipe, lds;   // pipe: 0 none, 1 pipelined, 2 pipelined and unrolled twice,
                           // 3 pipelined with four lanes per node slot (k_bond_quad)
  __host__ int lanes() const { return pipe == 3 ? 4 : 1; }   // threads per node slot
};
constexpr int kBondShapes = 15;
constexpr uint64_t kOff32MaxNodes = 1ull << 27;   // 16-byte records far below 4 GiB
__host__ inline const BondShape* bond_shapes() {
  static const BondShape t[kBondShapes] = {
      {k_bond_t<8, 0, true>, 8, 0, 0}, {k_bond_t<4, 0, true>, 4, 0, 0}, {k_bond_t<6, 0, true>, 6, 0, 0},
      {k_bond_t<3, 1, true>, 3, 1, 0}, {k_bond_t<4, 1, true>, 4, 1, 0}, {k_bond_t<5, 1, true>, 5, 1, 0},
      {k_bond_t<3, 2, true>, 3, 2, 0}, {k_bond_t<4, 2, true>, 4, 2, 0}, {k_bond_t<2, 2, true>, 2, 2, 0},
      {k_bond_lds<3, 2>, 3, 2, 1},     {k_bond_lds<4, 2>, 4, 2, 1},     {k_bond_lds<2, 2>, 2, 2, 1},
      {k_bond_lds<4, 1>, 4, 1, 1},     {k_bond_lds<5, 2>, 5, 2, 1},     {k_bond_quad<2>, 2, 3, 0}};
  return t;
}
// The same shapes in tolerance arithmetic (the default unstaged shape and the
// staged ones that fit the two-wave register budget; null: not built).
__host__ inline const BondShape* bond_shapes_tol() {
  static const BondShape t[kBondShapes] = {
      {nullptr, 8, 0, 0}, {nullptr, 4, 0, 0}, {nullptr, 6, 0, 0},
      {nullptr, 3, 1, 0}, {k_bond_t<4, 1, true, true>, 4, 1, 0}, {nullptr, 5, 1, 0},
      {nullptr, 3, 2, 0}, {nullptr, 4, 2, 0}, {nullptr, 2, 2, 0},
      {k_bond_lds<3, 2, true>, 3, 2, 1}, {k_bond_lds<4, 2, true>, 4, 2, 1}, {k_bond_lds<2, 2, true>, 2, 2, 1},
      {k_bond_lds<4, 1, true>, 4, 1, 1}, {nullptr, 5, 2, 1}, {nullptr, 2, 3, 0}};
  return t;
}
// The default shape with 64-bit record addressing, for pools too large for
// 32-bit offsets (never tuned).
__host__ inline BondShape bond_wide() { return {k_bond_t<4, 2, false>, 4, 2, 0}; }
// Shape index of bond_wide, selectable on any pool so that the 64-bit form
// can be checked on small ones.
constexpr int kWideShape = kBondShapes;

// ---------------------------------------------------------------------------
// Displacement of the clamped nodes (timeloop.h:767-774), after the bond
// kernel has zeroed their velocity and acceleration.  Launched only for pools
// with clamped nodes.
__global__ __launch_bounds__(256) void k_clamp(Dev d) {
  if (halted(d)) return;
  uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= d.N || !(d.nflag[g] & 1u)) return;
  const uint32_t p = d.pid[g];
  const DevPart& P = d.part[p];
  const DevInst& I = d.inst[P.inst];
  if (!live(I) || !P.movable || (I.vc_on && p == I.vc_slot)) return;
  set_disp(d, g, 0.0, 0.0, 0.0);
}

// Prescribed velocity (timeloop.h:685-764), one wavefront per job: the mean
// velocity and, for the angular part, mean_CurrPos of the particle, summed
// in node order like the reference.
__global__ __launch_bounds__(64) void k_pv_mean(Dev d) {
  if (halted(d)) return;
  const uint64_t q = blockIdx.x;
  if (q >= d.ninst) return;
  DevInst& I = d.inst[q];
  if (!live(I) || !I.vc_on || I.vc_slot == kNoKey) return;
  const DevPart& P = d.part[I.vc_slot];
  if (!P.movable) return;
  const uint64_t g0 = P.node_off, g1 = P.node_off + P.nnodes;
  double m[3] = {0, 0, 0}, c[3] = {0, 0, 0};
  for (uint64_t b = g0; b < g1; b += 64) {
    const uint64_t g = b + threadIdx.x;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (g < g1) {
      v0 = d.V[3 * g];
      v1 = d.V[3 * g + 1];
      v2 = d.V[3 * g + 2];
    }
    add_lanes3((uint32_t)(g1 - b < 64 ? g1 - b : 64), v0, v1, v2, m[0], m[1], m[2]);
  }
  if (I.vc_fix_ang) sum_currpos(d, g0, g1, c[0], c[1], c[2]);
  if (threadIdx.x == 0) {
    const double n = (double)P.nnodes;
    for (int k = 0; k < 3; k++) I.vc_mean[k] = m[k] / n;
    if (I.vc_fix_ang)
      for (int k = 0; k < 3; k++) I.vc_cent[k] = c[k] / n;
  }
}

// The prescription per node, the applied force density formed from the
// assigned velocity as the reference forms it, then the clamped nodes.  The
// reference's radial / tangential split (v_r, v_t, t_unit) feeds nothing.
__global__ __launch_bounds__(256) void k_pv_apply(Dev d) {
  if (halted(d)) return;
  uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= d.N) return;
  const uint32_t p = d.pid[g];
  const DevPart& P = d.part[p];
  const DevInst& I = d.inst[P.inst];
  if (!live(I) || !I.vc_on || p != I.vc_slot || !P.movable) return;
  const double* mean = I.vc_mean;
  double m[3] = {mean[0], mean[1], mean[2]};
  if (I.vc_fix_mean_x) m[0] = I.vc_mean_x;
  double v[3] = {d.V[3 * g], d.V[3 * g + 1], d.V[3 * g + 2]};
  double vr[3] = {v[0] - mean[0], v[1] - mean[1], v[2] - mean[2]};
  if (I.vc_fix_ang) {
    const double4 x = d.X[g];
    const double4 u = disp_of(d, g);
    const double rv[3] = {(x.x + u.x) - I.vc_cent[0], (x.y + u.y) - I.vc_cent[1], (x.z + u.z) - I.vc_cent[2]};
    const double r = nrm3(rv[0], rv[1], rv[2]);
    double ru[3] = {0, 0, 0};
    if (r > 0)
      for (int k = 0; k < 3; k++) ru[k] = rv[k] / r;
    vr[0] = (-(r * I.vc_omega)) * ru[1];
    vr[1] = (r * I.vc_omega) * ru[0];
    vr[2] = 0;
  }
  for (int k = 0; k < 3; k++) {
    const double pv = vr[k] + m[k];
    v[k] = pv;
    d.V[3 * g + k] = pv;
    d.AFD[3 * g + k] = P.rho * (((2 / I.dt) * (pv - v[k])) - d.AFD[3 * g + k]) - d.F[3 * g + k];
  }
  if (d.nflag[g] & 1u) {
    set_disp(d, g, 0.0, 0.0, 0.0);
    for (int k = 0; k < 3; k++) d.V[3 * g + k] = d.A[3 * g + k] = 0.0;
  }
}

// Lagged boxes of movable particles (timeloop.h:779), reset of the per-step
// flags, wall motion (timeloop.h:784-797).
__global__ __launch_bounds__(256) void k_finalize(Dev d) {
  if (halted(d)) return;
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < d.P) {
    // the box of this step if the particle's job stepped: every node of a
    // movable particle of a stepping job entered lrtb_next (the job's
    // remaining count is being decremented by another thread of this launch)
    if (d.part[q].movable && d.lrtb_next[6 * q] != ~0ull)
      for (int k = 0; k < 6; k++) d.lrtb[6 * q + k] = okey_inv(d.lrtb_next[6 * q + k]);
    for (int k = 0; k < 3; k++) {
      d.lrtb_next[6 * q + k] = ~0ull;
      d.lrtb_next[6 * q + 3 + k] = 0ull;
    }
    d.pactive[q] = 0;
  }
  if (q < d.ninst) {
    DevInst& I = d.inst[q];
    if (live(I)) {
      if (rebuilds(d, (uint32_t)q)) {
        atomicAdd(&d.counters[2], 1ull);
        // the lists built this step start from this step's displacement
        for (uint32_t s = 0; s < kDuSlots; s++) *du2_slot(d, q, s) = 0ull;
      }
      I.dirty = 0;
      for (int k = 0; k < 6; k++) I.wall[k] += I.speed[k] * I.dt;
      I.t += 1;
      I.remaining -= 1;
    }
  }
  if (q == 0) {
    if (d.ctl->steps_left) d.ctl->steps_left -= 1;
    // the step's pair computations count once the step is complete
    d.counters[0] += d.counters[3];
    d.counters[3] = 0ull;
  }
  // after a rebuild, the cell blocks of the jobs that rebuilt are cleared for
  // the next one (nothing reads them after k_build); the list and the flag
  // are reset by the next step's position kernel.  A job's blocks are those
  // of its node blocks, from its grid's first key.
  if (d.ctl->any_rebuild) {
    const uint32_t nr = ld_uniform(&d.ctl->nrj, 0u);
    for (uint32_t k = 0; k < nr; k++) {
      const DevInst& J = d.inst[ld_uniform(d.rjobs, k)];
      const uint32_t jb0 = J.gtb / kScanBlock, jb1 = (uint32_t)((J.node1 - 1) / 256);
      for (uint32_t b = jb0 + blockIdx.x; b <= jb1; b += gridDim.x)
        for (uint32_t i = threadIdx.x; i < (uint32_t)kScanBlock; i += blockDim.x) d.cstart[b * kScanBlock + i] = 0;
    }
  }
}

// Wall reaction of one step in the reference order (particle, node, plane,
// term), recomputed from the same device function as the force.  One
// wavefront per job walks its node range 64 nodes at a time: blocks without
// a wall contact are skipped through bflag (64 flags per load), the lanes
// evaluate their nodes' terms together, and the terms are then subtracted
// from the running sums one node after another in node order, so every
// component receives its terms in the reference's order.
// nl_bdry_only: particle by particle, each in boundary-list order
// (contact.h:1065-1069), one lane
__device__ void reaction_bdry_only(const Dev& d, const DevInst& I, uint64_t q) {
  double reac[18];
  for (int k = 0; k < 18; k++) reac[k] = 0.0;
  for (uint64_t g = I.node0; g < I.node1;) {
    const DevPart& P = d.part[d.pid[g]];
    for (uint32_t r = 0; r < P.nbdry; r++) {
      const uint32_t h = d.blist[P.boff + r];
      if (!d.wmask[h]) continue;
      const double4 x = d.X[h];
      const double4 u = disp_of(d, h);
      const double cp[3] = {x.x + u.x, x.y + u.y, x.z + u.z};
      const double v[3] = {d.Vold[3 * h], d.Vold[3 * h + 1], d.Vold[3 * h + 2]};
      double fdummy[3] = {0, 0, 0};
      wall_node(I, I.dt, cp, v, x.w, P.rho, fdummy, reac);
      d.wmask[h] = 0;
    }
    g = P.node_off + P.nnodes;
  }
  for (int k = 0; k < 18; k++) d.reaction[18 * q + k] = reac[k];
}

__global__ __launch_bounds__(64) void k_reaction(Dev d) {
  if (halted(d)) return;
  const uint64_t q = blockIdx.x;
  if (q >= d.ninst) return;
  const DevInst I = d.inst[q];
  if (!live(I) || !reaction_step(d, I)) return;
  const uint32_t lane = threadIdx.x;
  if (I.nl_bdry_only) {
    if (lane == 0) reaction_bdry_only(d, I, q);
    return;
  }
  double reac[18];
#pragma unroll
  for (int k = 0; k < 18; k++) reac[k] = 0.0;
  const uint64_t g0 = I.node0, g1 = I.node1;
  constexpr int plane[6] = {0, 3, 1, 4, 2, 5};   // x_min, x_max, y_min, y_max, z_min, z_max
  if (g1 > g0) {
    const uint64_t b0 = g0 >> 8, b1 = (g1 - 1) >> 8;
    for (uint64_t bb = b0; bb <= b1; bb += 64) {
      const uint64_t bl = bb + lane;
      uint64_t bm = __builtin_amdgcn_ballot_w64(bl <= b1 && d.bflag[bl] != 0);
      while (bm) {
        const uint64_t blk = bb + (uint64_t)__builtin_ctzll(bm);
        bm &= bm - 1;
        for (uint32_t sub = 0; sub < 256; sub += 64) {
          const uint64_t g = (blk << 8) + sub + lane;
          const bool on = g >= g0 && g < g1 && d.wmask[g] != 0;
          uint64_t cm = __builtin_amdgcn_ballot_w64(on);
          if (!cm) continue;
          WallSide w[6];
#pragma unroll
          for (int k = 0; k < 6; k++) w[k] = WallSide{};
          if (on) {
            const DevPart& P = d.part[d.pid[g]];
            const double4 x = d.X[g];
            const double4 u = disp_of(d, g);
            const double cp[3] = {x.x + u.x, x.y + u.y, x.z + u.z};
            const double v[3] = {d.Vold[3 * g], d.Vold[3 * g + 1], d.Vold[3 * g + 2]};
#pragma unroll
            for (int k = 0; k < 6; k++) {
              const int pl = plane[k], dir = pl % 3;
              w[k] = wall_side_eval(I, I.dt, cp[dir] - I.wall[pl], dir, v, x.w, P.rho, I.speed[pl]);
            }
            d.wmask[g] = 0;
          }
          // the nodes of this run in node order, their planes in the reference order
          while (cm) {
            const int l = __builtin_ctzll(cm);
            cm &= cm - 1;
#pragma unroll
            for (int k = 0; k < 6; k++) {
              const int pl = plane[k];
              WallSide b{};
              b.hit = __shfl((int)w[k].hit, l, 64) != 0;
              if (!b.hit) continue;
              b.dir = pl % 3;
              b.fric = __shfl((int)w[k].fric, l, 64) != 0;
              b.damp = __shfl((int)w[k].damp, l, 64) != 0;
              b.rn = __shfl(w[k].rn, l, 64);
              b.rf1 = __shfl(w[k].rf1, l, 64);
              b.rf2 = __shfl(w[k].rf2, l, 64);
              b.rd = __shfl(w[k].rd, l, 64);
              wall_side_react(b, reac + 3 * pl);
            }
          }
        }
      }
    }
  }
  if (lane == 0)
    for (int k = 0; k < 18; k++) d.reaction[18 * q + k] = reac[k];
}

// Back to the end of the last complete step after a halt (the host makes
// the other displacement buffer current): the halted step's box
// accumulators, pair flags, pair count, cell counts and list lengths asked
// for are cleared, and the halt itself.  With relist, every live job
// rebuilds its candidate lists at its next step (they were reallocated).
__global__ __launch_bounds__(256) void k_rollback(Dev d, uint32_t T, uint32_t relist) {
  const uint64_t n = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t q0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t q = q0; q < d.P; q += n) {
    for (int k = 0; k < 3; k++) {
      d.lrtb_next[6 * q + k] = ~0ull;
      d.lrtb_next[6 * q + 3 + k] = 0ull;
    }
    d.pactive[q] = 0;
  }
  for (uint64_t i = q0; i <= T; i += n) d.cstart[i] = 0;
  for (uint64_t q = q0; q < d.ninst; q += n) {
    d.jneed[q] = 0;
    if (relist && live(d.inst[q])) d.inst[q].dirty = 1;
  }
  if (q0 == 0) {
    d.counters[3] = 0ull;
    d.ctl->halt = 0;
    d.ctl->need = 0;
  }
}

// fill (graph-safe replacement of a memset)
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}


// NbdArr2conn (timeloop.h:20-44) of one particle (nodes [off, off + n)): its
// intact bonds as pairs (i, q), i < q, particle-relative, in row order -- node
// by node, each row in its own order.  k_conn_count counts each node's pairs,
// k_conn_scan turns the counts into offsets (one workgroup; the total goes to
// cnt[n]), k_conn_write writes each node's pairs at its offset.  A thread
// walks its row kConnBatch entries at a time, their loads in flight together.
constexpr uint32_t kConnThreads = 1024;
constexpr uint32_t kConnBatch = 8;
__device__ __forceinline__ const uint32_t* conn_row(const Dev& d, uint64_t g) {
  const uint64_t t = (g & ~255ull) + d.node_lane[g];   // the node's slot
  return d.bond + d.soff[t >> 6] + (t & 63);
}
__device__ __forceinline__ void conn_batch(const uint32_t* row, uint32_t k0, uint32_t deg, uint32_t (&e)[kConnBatch]) {
#pragma unroll
  for (uint32_t u = 0; u < kConnBatch; u++)
    e[u] = k0 + u < deg ? row[(uint64_t)(k0 + u) * kSlice] : kBroken;
}
__global__ __launch_bounds__(256) void k_conn_count(Dev d, uint64_t off, uint32_t n, uint32_t* cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t deg = d.deg[off + i];
  const uint32_t* row = conn_row(d, off + i);
  uint32_t c = 0;
  for (uint32_t k0 = 0; k0 < deg; k0 += kConnBatch) {
    uint32_t e[kConnBatch];
    conn_batch(row, k0, deg, e);
#pragma unroll
    for (uint32_t u = 0; u < kConnBatch; u++) c += !(e[u] & kBroken) && e[u] - (uint32_t)off > i;
  }
  cnt[i] = c;
}
__global__ __launch_bounds__(kConnThreads) void k_conn_scan(uint32_t* cnt, uint32_t n) {
  __shared__ uint32_t wsum[kConnThreads / 64];
  const uint32_t per = (n + kConnThreads - 1) / kConnThreads;
  const uint32_t i0 = min(n, threadIdx.x * per), i1 = min(n, i0 + per);
  uint32_t c = 0;
  for (uint32_t i = i0; i < i1; i++) c += cnt[i];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t x = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  for (uint32_t w = 0; w < wv; w++) x += wsum[w];
  uint32_t at = x - c;
  for (uint32_t i = i0; i < i1; i++) {
    const uint32_t v = cnt[i];
    cnt[i] = at;
    at += v;
  }
  if (threadIdx.x == kConnThreads - 1) cnt[n] = x;
}
__global__ __launch_bounds__(256) void k_conn_write(Dev d, uint64_t off, uint32_t n, const uint32_t* at0,
                                                    uint32_t* pairs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t deg = d.deg[off + i];
  const uint32_t* row = conn_row(d, off + i);
  uint64_t at = at0[i];
  for (uint32_t k0 = 0; k0 < deg; k0 += kConnBatch) {
    uint32_t e[kConnBatch];
    conn_batch(row, k0, deg, e);
#pragma unroll
    for (uint32_t u = 0; u < kConnBatch; u++) {
      const uint32_t q = e[u] - (uint32_t)off;
      if (!(e[u] & kBroken) && q > i) {
        pairs[2 * at] = i;
        pairs[2 * at + 1] = q;
        at++;
      }
    }
  }
}

// alive directed bonds (for counters and read-back)
__global__ __launch_bounds__(256) void k_count_alive(Dev d, uint64_t nent, unsigned long long* out) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // padding entries carry the broken bit, so they are not counted
  __shared__ uint32_t part[4];
  uint32_t alive = 0;
  for (uint64_t i = q; i < nent; i += (uint64_t)gridDim.x * blockDim.x)
    alive += (d.bond[i] & kBroken) ? 0u : 1u;
  for (int off = 32; off > 0; off >>= 1) alive += __shfl_down(alive, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = alive;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = (unsigned long long)part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(out, t);
  }
}

// ---------------------------------------------------------------------------
// Self-test of the fast-path arithmetic: random operands of the fast domain,
// div_core/sqrt_core against the IEEE operators, bit for bit.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double make_double(uint64_t mant, int e, int neg) {
  uint64_t b = ((uint64_t)(e + 1023) << 52) | (mant & 0xFFFFFFFFFFFFFull) | ((uint64_t)neg << 63);
  return __longlong_as_double((long long)b);
}
// Hard cases of correct rounding.  For an odd integer B of b bits let
// A = +-(2^b)^-1 mod B, lifted into [2^52, 2^53) by multiples of B.  Then
// A 2^b = n B +- 1, so the quotient A 2^(b-53) / B = n 2^-53 +- 2^-53 / B:
// for a quotient in [1, 2) and odd n that is a rounding midpoint missed by
// about 2^-106 relative (Euclid's inverse, no rounding anywhere).
__device__ uint64_t inv_mod_odd(uint64_t a, uint64_t m) {   // a^-1 mod m, gcd(a, m) = 1
  int64_t t = 0, nt = 1;
  uint64_t r = m, nr = a;
  while (nr != 0) {
    const uint64_t q = r / nr;
    const int64_t tt = t - (int64_t)q * nt;
    t = nt;
    nt = tt;
    const uint64_t rr = r - q * nr;
    r = nr;
    nr = rr;
  }
  return t < 0 ? (uint64_t)(t + (int64_t)m) : (uint64_t)t;
}
__device__ uint64_t midpoint_numerator(uint64_t B, int b, bool below) {
  uint64_t A = inv_mod_odd((1ull << b) % B, B);
  if (below) A = B - A;
  if (A < (1ull << 52)) {
    const uint64_t lifted = A + ((1ull << 52) - A + B - 1) / B * B;
    if (lifted < (1ull << 53)) A = lifted;   // else keep A: still A 2^b = n B +- 1
  }
  return A;
}

// Each sample i checks, bit for bit against the IEEE operators:
//  random operands spanning the fast domain -- div_core(a, rcp_core(b)),
//  sqrt_core(x), sqrt_rcp_core(x).s and division through its reciprocal;
//  hard operands -- quotients within ~2^-106 of a rounding midpoint, by
//  divisors with random, all-ones-like and near-power-of-two mantissas and
//  by computed roots (through sqrt_rcp_core), and square roots within
//  d^2 2^-107 of a midpoint: sqrt(X 2^54) for X = 2^52 + d (+1), d odd,
//  lies just below (above) the odd 54-bit integer 2^53 + d.
// mism[0] += division mismatches, mism[1] += sqrt mismatches,
// mism[2] += samples evaluated (n in total).
__global__ __launch_bounds__(256) void k_selftest_math(uint64_t n, uint64_t seed,
                                                       unsigned long long* mism) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bad_div = 0, bad_sqrt = 0, done = 0;
  for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    done++;
    uint64_t r0 = mix64(seed ^ (i * 4 + 0)), r1 = mix64(seed ^ (i * 4 + 1));
    uint64_t r2 = mix64(seed ^ (i * 4 + 2)), r3 = mix64(seed ^ (i * 4 + 3));
    // denominator in [2^-300, 2^301): |e| or |xi|
    int eb = (int)(r0 % 601) - 300;
    double b = make_double(r1, eb, 0);
    // numerator: a unit-vector component (exponent eb-652 .. eb) or a length
    // difference (exponent -352 .. 301), 1/32 of them zero of either sign
    int ea = ((r2 >> 40) & 1) ? eb - (int)(r0 >> 32) % 653 : (int)((r0 >> 20) % 654) - 352;
    double a = make_double(r3, ea, (int)(r2 >> 63));
    if ((r2 & 31) == 0) a = ((r2 >> 8) & 1) ? -0.0 : 0.0;
    double q = div_core(a, rcp_core(b));
    double qi = a / b;
    bool eq = (a == 0.0) ? (q == 0.0 && qi == 0.0)
                         : (__double_as_longlong(q) == __double_as_longlong(qi));
    bad_div += eq ? 0u : 1u;
    // sqrt argument in [2^-600, 2^601)
    const int ex = (int)(r3 % 1201) - 600;
    double x = make_double(r2 >> 3, ex, 0);
    const double sx = sqrt(x);
    bad_sqrt += (__double_as_longlong(sqrt_core(x)) == __double_as_longlong(sx)) ? 0u : 1u;
    // division by that root through sqrt_rcp_core's reciprocal: numerators
    // as above relative to the root (components of e, differences of lengths)
    const SqrtRcp sr = sqrt_rcp_core(x);
    bad_sqrt += (__double_as_longlong(sr.s) == __double_as_longlong(sx)) ? 0u : 1u;
    const int es = ex >> 1;
    const uint64_t r4 = mix64(seed ^ ~(i * 4 + 1));
    const int ea2 = es + 1 - (int)((r4 >> 32) % 653);
    double a2 = make_double(r4, ea2, (int)(r3 >> 63));
    if ((r4 & 31) == 0) a2 = ((r4 >> 8) & 1) ? -0.0 : 0.0;
    const double q2 = div_core(a2, sr.r);
    const double q2i = a2 / sx;
    const bool eq2 = (a2 == 0.0) ? (q2 == 0.0 && q2i == 0.0)
                                 : (__double_as_longlong(q2) == __double_as_longlong(q2i));
    bad_div += eq2 ? 0u : 1u;

    // --- hard cases
    const uint64_t r5 = mix64(seed ^ ~(i * 4 + 2)), r6 = mix64(seed ^ ~(i * 4 + 3));
    const uint64_t m53 = (1ull << 53) - 1;
    uint64_t B;
    switch (r5 & 3) {
      case 0: B = m53 - 2 * ((r5 >> 2) & 0xFFFF); break;           // all ones, nearly
      case 1: B = (1ull << 52) + 1 + 2 * ((r5 >> 2) & 0xFFFF); break; // just above 2^52
      default: B = (((r5 >> 2) & m53) | (1ull << 52) | 1ull) & m53; break;
    }
    const uint64_t A = midpoint_numerator(B, 53, (r5 >> 62) & 1);
    const int hb = (int)(r6 % 601) - 300;
    const int ha = hb - (int)((r6 >> 16) % 653);
    const double hbv = ldexp((double)B, hb - 52);
    // exponent ha whatever A's length (A = 1 for B = 2^53 - 1, where
    // 2^53 = 1 mod B): numerators stay inside the fast domain
    double hav = ldexp((double)A, ha - (63 - __builtin_clzll(A)));
    if ((r6 >> 63) & 1) hav = -hav;
    bad_div += (__double_as_longlong(div_core(hav, rcp_core(hbv))) ==
                __double_as_longlong(hav / hbv)) ? 0u : 1u;
    // by a computed root: B = the root's mantissa without trailing zeros
    const uint64_t sm = ((uint64_t)__double_as_longlong(sx) & ((1ull << 52) - 1)) | (1ull << 52);
    const int tz = __builtin_ctzll(sm);
    const uint64_t RB = sm >> tz;
    if (RB > 1) {
      const uint64_t RA = midpoint_numerator(RB, 53 - tz, (r6 >> 62) & 1);
      double ra = ldexp((double)RA, es + 1 - (int)((r6 >> 32) % 653) - (63 - __builtin_clzll(RA)));
      if ((r5 >> 61) & 1) ra = -ra;
      bad_div += (__double_as_longlong(div_core(ra, sr.r)) == __double_as_longlong(ra / sx)) ? 0u : 1u;
    }
    // square roots next to a midpoint
    const uint64_t dd = ((r5 >> 20) & 0xFFFFF) | 1ull;
    const uint64_t X = (1ull << 52) + dd + ((r5 >> 40) & 1);
    const int k2 = (int)((r6 >> 40) % 600) - 353;     // x = X 2^(54 + 2 k2), exponent in [-600, 600]
    const double hx = ldexp((double)X, 54 + 2 * k2);
    const double hs = sqrt(hx);
    bad_sqrt += (__double_as_longlong(sqrt_core(hx)) == __double_as_longlong(hs)) ? 0u : 1u;
    const SqrtRcp hr = sqrt_rcp_core(hx);
    bad_sqrt += (__double_as_longlong(hr.s) == __double_as_longlong(hs)) ? 0u : 1u;
  }
  if (bad_div) atomicAdd(&mism[0], (unsigned long long)bad_div);
  if (bad_sqrt) atomicAdd(&mism[1], (unsigned long long)bad_sqrt);
  if (done) atomicAdd(&mism[2], (unsigned long long)done);
}

}  // namespace pd
