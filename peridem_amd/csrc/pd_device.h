// pd_device.h -- device-side data model and per-term arithmetic of the
// explicit step.  Every formula follows the reference operation for
// operation (IEEE FP64, compiled with -ffp-contract=off): a term computed
// here is bitwise equal to the same term of the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pd {

constexpr uint32_t kBroken = 0x80000000u;   // high bit of a bond entry: bond erased
constexpr uint32_t kPad = 0xFFFFFFFFu;      // slice padding entry
constexpr uint32_t kIdxMask = 0x7FFFFFFFu;
constexpr int kSlice = 64;                  // nodes per bond slice (= wavefront width)
// Node blocks of the staged bond kernel: 4 slices = one 256-thread workgroup.
// Jobs occupy whole blocks.  Each block keeps the sorted table of the nodes
// its rows reference (its own nodes included), at most kLdsCap of them, and
// each bond entry also exists as a u16: 16 * its position in that table (the
// byte offset of the node's 16-byte LDS records), or kDead16.
constexpr int kBlockSlices = 4;
constexpr uint32_t kLdsCap = 2240;          // 2240 * (16 + 16 + 4) B = 80 KB of LDS per workgroup
constexpr uint32_t kLive16 = 16 * kLdsCap;  // u16 entries below this are live table offsets
constexpr uint32_t kDead16 = 0xFFFFu;       // u16 entry of an erased bond
constexpr uint32_t kPad16 = 0xFFFFu;        // u16 slice padding entry
static_assert(kLive16 <= kDead16, "table offsets must fit a u16 below the dead value");
// Contact candidate lists: a node's list is built in passes of kMaxCand
// entries (held in registers); the pool's list capacity (Dev::ccap, at least
// kMaxCand) grows on demand up to kMaxCap, the most a u8 count holds.
constexpr int kMaxCand = 48;
constexpr uint32_t kMaxCap = 255;
constexpr uint32_t kNoKey = 0xFFFFFFFFu;
// keys of the contact cell table per node slot: a job's grid gets this many
// cells per node slot it occupies (occupied cells of edge rb hold 1-2 nodes,
// so this covers a bounding box about 2-6 times the volume of the nodes)
constexpr uint32_t kCellsPerSlot = 4;
// the cell table is scanned in blocks of kScanBlock keys: one block per
// 256-node block, so each scan block belongs to one job
constexpr int kScanBlock = 1024;
static_assert(kScanBlock == 256 * kCellsPerSlot, "one scan block per node block");
// state bits of Dev::flags
constexpr uint32_t kFlagSafeMath = 1u;  // an operand left the fast-path range: use IEEE ops

// Per-particle parameters (ParticleN scalars, particle2.h:51-59).
struct DevPart {
  uint64_t node_off;
  uint32_t nnodes;
  uint32_t inst;
  double rho, rinv, cnot, snot;   // rinv = 1/rho as the reference forms it (timeloop.h:673)
  int32_t movable, breakable, stoppable, rows_sorted;
  double tval;                    // torque_val, torque_axis (particle2.h:57-59)
  uint32_t taxis;
  uint32_t nbdry;                 // boundary_nodes: blist[boff .. boff + nbdry) in list order
  uint64_t boff;
};

// The displacement maximum of a job since its candidate lists were built is
// kept in one slot per XCD (the position kernel's blocks update the slot of
// the XCD they run on), each slot in its own 128-byte line, so no single
// word takes the atomics of the whole device.
constexpr uint32_t kDuSlots = 8;
constexpr uint32_t kDuStride = 16;   // u64 per slot

// Per-job (independent system) parameters: Contact + RectWall
// (contact.h:104-183, 185-376), the run switches of run_timeloop, and the
// job's own clock.  A job with remaining == 0 is frozen (finished, or a free
// slot): every kernel leaves its nodes untouched.
struct DevInst {
  double Rc, Kn, mu, zeta;
  double cell;              // cell edge of the candidate search (= rb)
  double rb;                // candidate-list radius: Rc (1 + 1e-9) + skin
  double du_lim2;           // rebuild when some |u - u_build|^2 exceeds this
  uint32_t dirty;           // rebuild at the next step (set by the host)
  double dt;
  double wall[6];           // x_min, y_min, z_min, x_max, y_max, z_max
  double speed[6];
  uint64_t remaining;       // steps still to run
  uint64_t node0, node1;    // node range [node0, node1)
  uint32_t t;               // reference step index of the next step (timeloop.h:395)
  int32_t allow_wall, allow_damping, allow_friction, self_contact;
  int32_t enable_fracture, gradient_extforce, compute_reaction;
  uint32_t extforce_maxstep;
  int32_t enable_torque;
  int32_t nl_bdry_only;     // contact and wall terms on boundary nodes only (contact.h:751-760)
  uint32_t vc_slot;         // prescribed-velocity particle (timeloop.h:685-764), kNoKey if none
  int32_t vc_on, vc_fix_mean_x, vc_fix_ang, pad2;
  double vc_mean_x, vc_omega;
  double vc_mean[3], vc_cent[3];   // its mean velocity and mean CurrPos of this step
  uint32_t gn[3];           // contact cell grid: cells per axis (see cell_key)
  uint32_t gtb;             // its first key in the cell table
};

// Per-launch control words shared by the kernels of a step.
struct DevCtl {
  uint32_t steps_left;      // steps left in the current pd_step call
  uint32_t any_rebuild;     // some job rebuilds its candidate lists this step
  // A candidate list longer than the capacity halts the pool inside the step
  // that builds it: every later kernel of the call exits at once, so the
  // state is that of the end of the previous step (the displacement of the
  // halted step went to the other buffer) and the host grows the lists and
  // runs the step again.
  uint32_t halt;
  uint32_t need;            // longest list the halted build asked for
  uint32_t nrj;             // jobs that rebuild their lists in this step (listed in Dev::rjobs)
};

struct Dev {
  // node state, global node index g
  double4* X;        // x, y, z, vol
  double4* U;        // disp x, y, z, (unused): transfer form only (uploads, read-backs);
                     // the resident displacement is UXY, UZ
  // the bond kernel's neighbour copies of X and U in 16- and 8-byte records:
  // a wavefront's gather of one field then spans fewer cache lines than the
  // same field read out of 32-byte records
  double2* XY;       // x, y
  double2* ZV;       // z, vol
  double2* UXY;      // disp x, y: the resident displacement (written by the position kernel)
  double* UZ;        // disp z
  // The displacement is double-buffered: the position kernel reads the
  // previous step's buffer (UXYo, UZo) and writes the step's own (UXY, UZ),
  // which every later kernel of the step reads; the host alternates them.
  double2* UXYo;
  double* UZo;
  uint16_t* bloc;    // [entries] the bond entries as positions in their block's node table
  uint32_t* btab;    // [blocks * kLdsCap] per block: sorted pool indices of the nodes its rows reference
  uint2* binfo;      // [blocks] (table length, position of the block's first node in it)
  // the lanes of a 256-node block hold its nodes by decreasing row length:
  // slot t holds node (t & ~255) + lane_node[t]; node g sits at slot
  // (g & ~255) + node_lane[g].  Bond slices are laid out by slot.
  uint8_t* lane_node;
  uint8_t* node_lane;
  double* V;         // vel [3N]
  double* Vold;      // vel [3N] as the step's force phase reads it: copied by the
                     // position kernel, since the bond kernel writes V node by node
  double* A;         // acc [3N]
  double* F;         // force density [3N]
  double* AFD;       // applied_force_density of prescribed-velocity particles [3N]
  double* E;         // extforce [3N]
  uint32_t* pid;     // particle of node
  uint32_t* brank;   // position of the node in its particle's boundary list, kNoKey if absent
  uint32_t* blist;   // boundary lists, stored in the node range of their job
  uint8_t* nflag;    // bit0 clamped
  uint32_t* deg;     // row length at commit
  // bonds: sliced rows, 64 nodes per slice, entry k of lane l at
  // soff[s] + k*64 + l for k < swid[s]
  uint64_t* soff;    // [S]
  uint32_t* swid;    // [S]
  uint32_t* bond;
  // particles / instances
  DevPart* part;
  DevInst* inst;
  double* lrtb;      // [P*6] lagged boxes (get_minmax order)
  double* cent;      // [P*3] mean CurrPos of a particle (torque steps only)
  unsigned long long* lrtb_next;  // [P*6] ordered-key min/max accumulators
  uint32_t* pactive; // [P] particle has an active contact pair this step
  uint32_t* pairs;   // [npairs*2] particle pairs (i<j) of the same instance
  // per job, kDuSlots running maxima of |u - u_build|^2 since the list build
  // (bits of doubles >= 0), one per XCD, each in its own 128-byte line
  unsigned long long* du2s;
  unsigned long long* counters;   // [0] pair computations, [2] list rebuilds,
                                  // [3] pair computations of the running step
  uint32_t* jneed;   // [ninst] longest candidate list a job's halted build asked for
  uint32_t* flags;   // [1] kFlagSafeMath (sticky)
  DevCtl* ctl;
  // wall reaction (per instance, 6 planes x 3) and helpers
  double* reaction;
  uint8_t* wmask;    // [N] planes in contact this step (reaction steps only)
  uint8_t* bflag;    // [N/256+1] block has a wall contact
  // contact candidate lists (Verlet lists with a skin): entry k of node g at
  // clist[(g >> 6) * 64 * ccap + 64 k + (g & 63)], ccnt[g] entries
  uint32_t* clist;
  uint8_t* ccnt;
  double2* UBXY;     // [N] displacement when the node's list was built: x, y
  double* UBZ;       //     and z
  // contact cell list
  double4* cpos;     // [N] CurrPos of the nodes of rebuilding jobs, this step (k_cellkey)
  double4* cps;      // [N] the same in cell-list order, beside csorted (k_scatter)
  uint32_t* ckey;    // [N]
  uint32_t* cstart;  // [T+1]
  uint32_t* ccur;    // [T]
  uint32_t* csorted; // [N]
  uint32_t* scan_tmp;
  uint32_t* rjobs;   // [ninst] the jobs that rebuild their lists in this step (DevCtl::nrj of them)
  uint32_t* jkeyed;  // [ninst] nodes a rebuilding job put into its cells this step
  uint64_t N, S, P, npairs, ninst;   // N = S*64 node slots; P, ninst incl. the void slot
  uint32_t T;        // cell table size: kCellsPerSlot keys per node slot
  uint32_t lists;    // some job of the pool can have contact: candidate lists are kept
  uint32_t ccap;     // candidate list capacity per node (kMaxCand .. kMaxCap)
};

// ---------------------------------------------------------------------------
// Correctly rounded FP64 sqrt and division, guarded fast path.
//
// sqrt_core / rcp_core + div_core are the instruction sequences the AMDGPU
// backend emits for IEEE sqrt (v_rsq_f64 and two Newton-Raphson corrections)
// and IEEE division (v_rcp_f64, two Newton-Raphson steps, quotient and one
// remainder correction), without the exponent pre-scaling (v_ldexp,
// v_div_scale, v_div_fmas) and the special-value fix-up (v_cmp_class,
// v_div_fixup).  For operands inside the guarded ranges those wrappers are
// the identity, so the results are bitwise those of `sqrt(x)` and `a / b`;
// callers test sqrt_ok/div_ok and use the plain operators otherwise.  The
// refined reciprocal depends only on the denominator, so the three divisions
// e_k / |e| of one bond share it.
// Fast-path domain.  Node coordinates, reference bond lengths (checked at
// commit) and displacements (checked every step by the position kernel) are 0
// or of magnitude in [2^-600, 2^600]; per bond only |e|^2 is range-checked.
// Then every operand of the bond arithmetic is a multiple of 2^-652 and
// bounded by 2^302, far from the exponents where the backend's sqrt and
// division sequences rescale or fix up their operands (below 2^-767 for
// sqrt, numerators below 2^-969 or exponent gaps of 768 for division).
constexpr double kFastLo = 0x1p-600, kFastHi = 0x1p+600;
__host__ __device__ inline bool in_fast_range(double v) {
  double a = v < 0 ? -v : v;
  return (v == 0.0) || ((a >= kFastLo) && (a <= kFastHi));
}

__device__ __forceinline__ double sqrt_core(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return g;
}

struct Rcp {
  double nb;  // -b
  double r;   // refined reciprocal of b
};
__device__ __forceinline__ Rcp rcp_core(double b) {
  double r0 = __builtin_amdgcn_rcp(b);
  double e0 = __builtin_fma(-b, r0, 1.0);
  double r1 = __builtin_fma(r0, e0, r0);
  double e1 = __builtin_fma(-b, r1, 1.0);
  double r2 = __builtin_fma(r1, e1, r1);
  return {-b, r2};
}
// a / b for b > 0.  Equal to IEEE division except that a zero numerator may
// give +0 where IEEE gives -0; callers only use it where that cannot matter
// (a zero term added to a sum that is never -0).
__device__ __forceinline__ double div_core(double a, Rcp R) {
  double q = a * R.r;
  double rem = __builtin_fma(R.nb, q, a);
  return __builtin_fma(rem, R.r, q);
}

// sqrt(x) and the refined reciprocal of that root.  The reciprocal starts
// from the reciprocal square root that sqrt_core refines anyway (2h, already
// at least as close to 1/sqrt(x) as v_rcp_f64 is to 1/s) and takes the same
// two Newton-Raphson steps against the rounded root that rcp_core takes, so it
// ends in the same error class -- without the transcendental.  div_core with
// it is checked bit for bit against IEEE division by pd_selftest_math.
struct SqrtRcp {
  double s;   // correctly rounded sqrt(x)
  Rcp r;      // refined reciprocal of s
};
__device__ __forceinline__ SqrtRcp sqrt_rcp_core(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  const double r0 = h + h;
  const double e0 = __builtin_fma(-g, r0, 1.0);
  const double r1 = __builtin_fma(r0, e0, r0);
  const double e1 = __builtin_fma(-g, r1, 1.0);
  return {g, {-g, __builtin_fma(r1, e1, r1)}};
}

// Eigen 3.3 fixed-size 1x3 norm/dot order.
__device__ __forceinline__ double nrm3(double a, double b, double c) {
  return sqrt((a * a + b * b) + c * c);
}

// monotone map double -> uint64 for atomic min/max
__device__ __forceinline__ unsigned long long okey(double v) {
  unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double okey_inv(unsigned long long k) {
  unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  double d;
  __builtin_memcpy(&d, &b, 8);
  return d;
}

// Contact cells.  Each job has a grid of gn[0] x gn[1] x gn[2] cells of edge
// `cell` at keys gtb .. gtb + gn[0] gn[1] gn[2] - 1 of the cell table, x
// fastest, so the three cells of a row of neighbours along x are adjacent
// keys.  A position's cell is floor(p / cell) taken modulo gn per axis: the
// map is one to one on any gn consecutive cells, so a job whose nodes span
// at most gn cells per axis never puts two of its cells on one key; far-apart
// cells of a wider job may share one, and the distance test sorts their nodes
// apart.
__device__ __forceinline__ uint32_t wrap_cell(double p, double cell, uint32_t n) {
  double q = floor(p / cell);
  if (!(q > -0x1p52 && q < 0x1p52)) q = 0.0;   // NaN or beyond exact integers: any cell
  // q mod n in exact integer arithmetic on doubles: floor(q / n) may be one
  // off after rounding, which leaves m in [-n, 2n)
  const double dn = (double)n;
  double m = q - dn * floor(q / dn);
  if (m < 0.0) m += dn;
  if (m >= dn) m -= dn;
  return (uint32_t)m;
}
// w + dd (dd in -1..1) modulo n, w in [0, n)
__device__ __forceinline__ uint32_t wrap_add(uint32_t w, int dd, uint32_t n) {
  int64_t v = (int64_t)w + dd;
  if (v < 0) v += n;
  if (v >= (int64_t)n) v -= n;
  return (uint32_t)v;
}
__device__ __forceinline__ uint32_t cell_key(const DevInst& I, uint32_t wx, uint32_t wy, uint32_t wz) {
  return I.gtb + (wz * I.gn[1] + wy) * I.gn[0] + wx;
}

// intersects_box_fext<3> (contact.h:1143-1183): A extended by c.
__device__ __forceinline__ bool intersects_box_fext(const double* A, const double* B, double c) {
  double Aminx = A[0] - c, Aminy = A[1] - c, Aminz = A[2] - c;
  double Amaxx = A[3] + c, Amaxy = A[4] + c, Amaxz = A[5] + c;
  return (Aminx <= B[3]) && (Amaxx >= B[0]) && (Amaxy >= B[1]) && (Aminy <= B[4]) &&
         (Aminz <= B[5]) && (Amaxz >= B[2]);
}

// within_interior_fext (contact.h:1229-1243): the 3-D box read with 2-D indices.
__device__ __forceinline__ bool within_interior_fext(const double* A, const double* B, double c) {
  double AL = A[0] - c, AR = A[1] + c, AT = A[2] + c, AB = A[3] - c;
  return (B[0] < AL) && (AR < B[1]) && (B[2] > AT) && (B[3] < AB);
}

// wall_side_force_update<3> (contact.h:385-502) with segment_mass<3>
// (contact.h:14-39): the terms of one node and one plane, each as the
// reference forms it -- the force terms added to the node's force, and the
// same terms times the node volume subtracted from the wall reaction.
struct WallSide {
  bool hit, fric, damp;
  int dir;
  double fn, ff1, ff2, fd;   // force: normal, friction on the two tangents, damping
  double rn, rf1, rf2, rd;   // reaction terms (each subtracted)
};
__device__ __forceinline__ WallSide wall_side_eval(const DevInst& I, double dt, double dist, int dir,
                                                   const double* vel, double vol, double rho, double speed) {
  WallSide w{};
  w.dir = dir;
  const double Rc = I.Rc;
  if (!(fabs(dist) <= Rc)) return w;
  w.hit = true;
  double dd = fabs(dist);
  double Rd = Rc - dd;
  double area = ((M_PI / 3) * (Rd * Rd)) * (2 * Rc - dd);
  double q = Rc * Rc - dd * dd;
  double Im = (M_PI / 4) * (q * q);
  double cen = Im / area;
  double contrib = Rc - cen;
  double pc = contrib > 0 ? contrib : 0;
  double Fn = (I.Kn * pc) * area;
  double sg = dist > 0 ? 1.0 : -1.0;
  w.fn = sg * Fn;
  w.rn = (vol * sg) * Fn;
  double rv[3] = {vel[0], vel[1], vel[2]};
  rv[dir] -= speed;
  if (I.allow_friction) {
    int t1 = (dir + 1) % 3, t2 = (dir + 2) % 3;
    double p1 = rv[t1], p2 = rv[t2];
    double pn = sqrt(p1 * p1 + p2 * p2);
    if (pn > 0) {
      double u1 = p1 / pn, u2 = p2 / pn;
      double ff1 = ((-I.mu) * fabs(Fn)) * u1;
      double ff2 = ((-I.mu) * fabs(Fn)) * u2;
      if (pn < ((I.mu * fabs(Fn)) * dt) / rho) {
        ff1 = (((-pn) / rho) / dt) * u1;
        ff2 = (((-pn) / rho) / dt) * u2;
      }
      w.fric = true;
      w.ff1 = ff1;
      w.ff2 = ff2;
      w.rf1 = vol * ff1;
      w.rf2 = vol * ff2;
    }
  }
  if (I.allow_damping) {
    double dc = ((2 * I.zeta) * sqrt(I.Kn * rho)) * sqrt(fabs(area));
    double fd = (-dc) * rv[dir];
    w.damp = true;
    w.fd = fd;
    w.rd = vol * fd;
  }
  return w;
}
// The reaction terms of one plane subtracted in the reference's order.
__device__ __forceinline__ void wall_side_react(const WallSide& w, double* reac /* the plane's [3] */) {
  if (!w.hit) return;
  reac[w.dir] -= w.rn;
  if (w.fric) {
    reac[(w.dir + 1) % 3] -= w.rf1;
    reac[(w.dir + 2) % 3] -= w.rf2;
  }
  if (w.damp) reac[w.dir] -= w.rd;
}
// Adds the node's terms to f; when reac != nullptr also accumulates the wall
// reaction in the reference's order.  The arithmetic of wall_side_eval,
// written out once more so that the force path inside the bond kernel keeps
// its own register allocation; both are checked bit for bit against the
// oracle (forces at every parity checkpoint, reactions where they are
// formed).
__device__ __forceinline__ bool wall_side(const DevInst& I, double dt, double dist, int dir,
                                          const double* vel, double vol, double rho, double* f,
                                          double speed, double* reac) {
  const double Rc = I.Rc;
  if (!(fabs(dist) <= Rc)) return false;
  double dd = fabs(dist);
  double Rd = Rc - dd;
  double area = ((M_PI / 3) * (Rd * Rd)) * (2 * Rc - dd);
  double q = Rc * Rc - dd * dd;
  double Im = (M_PI / 4) * (q * q);
  double cen = Im / area;
  double contrib = Rc - cen;
  double pc = contrib > 0 ? contrib : 0;
  double Fn = (I.Kn * pc) * area;
  double sg = dist > 0 ? 1.0 : -1.0;
  f[dir] += sg * Fn;
  if (reac) reac[dir] -= (vol * sg) * Fn;
  double rv[3] = {vel[0], vel[1], vel[2]};
  rv[dir] -= speed;
  if (I.allow_friction) {
    int t1 = (dir + 1) % 3, t2 = (dir + 2) % 3;
    double p1 = rv[t1], p2 = rv[t2];
    double pn = sqrt(p1 * p1 + p2 * p2);
    if (pn > 0) {
      double u1 = p1 / pn, u2 = p2 / pn;
      double ff1 = ((-I.mu) * fabs(Fn)) * u1;
      double ff2 = ((-I.mu) * fabs(Fn)) * u2;
      if (pn < ((I.mu * fabs(Fn)) * dt) / rho) {
        ff1 = (((-pn) / rho) / dt) * u1;
        ff2 = (((-pn) / rho) / dt) * u2;
      }
      f[t1] += ff1;
      f[t2] += ff2;
      if (reac) {
        reac[t1] -= vol * ff1;
        reac[t2] -= vol * ff2;
      }
    }
  }
  if (I.allow_damping) {
    double dc = ((2 * I.zeta) * sqrt(I.Kn * rho)) * sqrt(fabs(area));
    double fd = (-dc) * rv[dir];
    f[dir] += fd;
    if (reac) reac[dir] -= vol * fd;
  }
  return true;
}

// The six planes in the reference order x_min, x_max, y_min, y_max, z_min,
// z_max (contact.h:1104-1121).  Returns the mask of planes in contact.
__device__ __forceinline__ uint32_t wall_node(const DevInst& I, double dt, const double* cp,
                                              const double* vel, double vol, double rho, double* f,
                                              double* reac /* [18] or null */) {
  uint32_t m = 0;
  const int plane[6] = {0, 3, 1, 4, 2, 5};
  for (int s = 0; s < 6; s++) {
    int pl = plane[s];
    int dir = pl % 3;
    if (wall_side(I, dt, cp[dir] - I.wall[pl], dir, vel, vol, rho, f, I.speed[pl],
                  reac ? reac + 3 * pl : nullptr))
      m |= 1u << pl;
  }
  return m;
}

// Contact roles of a node pair (update_contact_force_by_boundary,
// contact.h:746-942): ACTION = our node is P_C's node i; REACTION = our node
// is P_N's node j; SELF = self-contact (action only, contact.h:779-796).
enum Role { kAction = 0, kReaction = 1, kSelf = 2 };

// Apply the terms of one contact pair to our node's force f.  `d` is the
// reference's dir = (u_j - u_i) + (X_j - X_i) (P_C = i, P_N = j), `dn` its
// norm (already known to satisfy dn <= Rc), len/K the reference length and
// spring constant, VN = P_N.vol[j], VC = P_C.vol[i], rhoC = P_C.rho,
// rv = P_C.vel[i] - P_N.vel[j].
__device__ __forceinline__ void contact_terms(const DevInst& I, double dt, int role,
                                              const double* d, double dn, double len, double K,
                                              double VN, double VC, double rhoC,
                                              const double* rv, double* f) {
  if (!(dn < len)) return;
  double u[3] = {d[0] / dn, d[1] / dn, d[2] / dn};
  double s = (dn - len) / len;
  double fc[3];
  for (int k = 0; k < 3; k++) fc[k] = ((K * s) * VN) * u[k];
  double fcn = nrm3(fc[0], fc[1], fc[2]);
  if (role == kReaction) {
    for (int k = 0; k < 3; k++) f[k] -= (fc[k] / VN) * VC;
  } else {
    for (int k = 0; k < 3; k++) f[k] += fc[k];
  }
  if (I.allow_damping) {
    double pr = (rv[0] * u[0] + rv[1] * u[1]) + rv[2] * u[2];
    double dc = ((2 * I.zeta) * sqrt(K * rhoC)) * sqrt(VN);
    double fd[3];
    for (int k = 0; k < 3; k++) fd[k] = (-dc) * (pr * u[k]);
    if (role == kReaction) {
      double sV = sqrt(VN), sC = sqrt(VC);
      for (int k = 0; k < 3; k++) f[k] -= (fd[k] / sV) * sC;
    } else {
      for (int k = 0; k < 3; k++) f[k] += fd[k];
    }
  }
  if (I.allow_friction) {
    double pr = (rv[0] * u[0] + rv[1] * u[1]) + rv[2] * u[2];
    double tv[3];
    for (int k = 0; k < 3; k++) tv[k] = rv[k] - pr * u[k];
    double tn = nrm3(tv[0], tv[1], tv[2]);
    if (tn > 0) {
      double tu[3], ff[3];
      for (int k = 0; k < 3; k++) tu[k] = tv[k] / tn;
      for (int k = 0; k < 3; k++) ff[k] = ((-I.mu) * fcn) * tu[k];
      if (tn < ((I.mu * fcn) * dt) / rhoC)
        for (int k = 0; k < 3; k++) ff[k] = (((-tn) / rhoC) / dt) * tu[k];
      if (role == kReaction) {
        for (int k = 0; k < 3; k++) f[k] -= ff[k];
      } else {
        for (int k = 0; k < 3; k++) f[k] += ff[k];
      }
    }
  }
}

}  // namespace pd
