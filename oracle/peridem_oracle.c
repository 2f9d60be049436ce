/*
 * peridem_oracle.c -- CPU oracle for the PeriDEM explicit step.
 *
 * TEST INFRASTRUCTURE ONLY: what the HIP library's results are checked
 * against, and the CPU baseline of the benchmark.  Nothing in the product path links or calls it.
 *
 * A plain-C restatement of the reference algorithm, operation for operation
 * (IEEE FP64, no FMA contraction: build with -ffp-contract=off).  Each routine
 * cites the reference file:line it follows.  Eigen 3.3 fixed-size 1x3
 * reductions evaluate as (x0*x0 + x1*x1) + x2*x2, which nrm3()/dot3() follow.
 *
 * Pinning (what checks each function besides the HIP comparison):
 *   pd_force_particle  - two-node closed form, rigid translation (s = 0),
 *                        uniform dilation on a lattice (s = eps), momentum
 *                        balance, strict s > s0 threshold   (tests/test_oracle.py)
 *   remove_bonds       - irreversibility, stable row order, both directions
 *   contact_pair/self  - two-node normal/damping/friction closed forms, the
 *                        small-velocity friction cap, bonded-pair exclusion
 *   wall_side          - segment_mass closed form, sign per plane, reaction
 *   conn2nbr           - hand-made Connectivity incl. unsorted rows
 *   verlet             - free flight u = v t, constant-force parabola,
 *                        clamped nodes
 *   torque             - tangential direction and magnitude on a ring
 */
#include "peridem_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  uint32_t n, cap;
  uint32_t* nbr;  /* NbdArr[i] */
  double* xi;     /* xi[i][j], 3 per entry */
  double* xin;    /* xi_norm[i][j] */
  double* str;    /* stretch[i][j] */
} orow;

typedef struct {
  uint32_t n;
  double *pos, *disp, *cp, *vel, *acc, *force, *ext, *vol;
  orow* rows;
  uint32_t nclamped, nbdry;
  uint32_t *clamped, *bdry;
  uint64_t nconn;
  uint32_t* conn;
  double delta, rho, cnot, snot;
  int movable, breakable, stoppable;
  uint32_t taxis;
  double tval;
  uint64_t nbonds0; /* directed entries at commit */
} opart;

struct or_ctx {
  pd_run_desc run;
  double wall[6], speed[6];
  double reaction[6][3]; /* x_min, y_min, z_min, x_max, y_max, z_max */
  opart* P;
  uint32_t np, capp;
  uint64_t* gidx;
  uint64_t ntot;
  double (*lrtb)[6];
  double* gf;     /* giant_f, [ntot*3] */
  double* gfr;    /* per-rank buffers when nranks > 1 */
  uint32_t t;
  int committed, nranks;
  int prune;      /* contact candidates from a cell grid (or_set_prune) */
  pd_counters cnt;
  char err[256];
  /* prescribed velocity (timeloop.h:685-764): CFGV.particle_w_fixed_velocity,
   * fix_velocity_mean_x, prescribed_velocity_mean_x, fix_velocity_angular,
   * prescribed_angular_vel; TL.enable_velocity_constraint */
  int64_t vc_part;
  int vc_on, vc_fix_mean_x, vc_fix_ang;
  double vc_mean_x, vc_omega;
  double* afd;      /* applied_force_density of the constrained particle */
  double* aold;     /* scratch: its acc before the update */
};

static int fail(or_ctx* c, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  if (c) vsnprintf(c->err, sizeof c->err, fmt, ap);
  va_end(ap);
  return code;
}

/* Eigen 3.3 Matrix<double,1,3>::squaredNorm / dot order. */
static inline double nrm3(double a, double b, double c) { return sqrt((a * a + b * b) + c * c); }
static inline double dot3(const double* a, const double* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

void or_segment_mass3(double R, double d, double out[2]) {
  /* contact.h:14-39, dim == 3 branch; pow(x,2) is x*x. */
  double Rd = R - d;
  double Area = ((M_PI / 3) * (Rd * Rd)) * (2 * R - d);
  double q = R * R - d * d;
  double I = (M_PI / 4) * (q * q);
  out[0] = I / Area;
  out[1] = Area;
}

int or_create(or_ctx** out, const pd_run_desc* run) {
  if (!out || !run) return PD_E_ARG;
  or_ctx* c = (or_ctx*)calloc(1, sizeof(or_ctx));
  if (!c) return PD_E_NOMEM;
  c->run = *run;
  memcpy(c->wall, run->wall_lrtb, sizeof c->wall);
  memcpy(c->speed, run->wall_speed, sizeof c->speed);
  c->t = run->first_step ? run->first_step : 1;
  c->nranks = 1;
  c->vc_part = -1;
  *out = c;
  return PD_OK;
}

static double* dup3(const double* src, uint32_t n, int k) {
  double* d = (double*)calloc((size_t)n * k + 1, sizeof(double));
  if (d && src) memcpy(d, src, (size_t)n * k * sizeof(double));
  return d;
}

int or_add_particle(or_ctx* c, const pd_particle_desc* d, uint32_t* index) {
  if (!c || !d) return PD_E_ARG;
  if (c->committed) return fail(c, PD_E_STATE, "or_add_particle after commit");
  if (d->nnodes == 0 || !d->pos || !d->vol) return fail(c, PD_E_ARG, "particle needs nodes, pos, vol");
  if (d->instance != 0) return fail(c, PD_E_UNSUPPORTED, "oracle runs one instance");
  if (c->np == c->capp) {
    c->capp = c->capp ? 2 * c->capp : 8;
    c->P = (opart*)realloc(c->P, c->capp * sizeof(opart));
  }
  opart* p = &c->P[c->np];
  memset(p, 0, sizeof *p);
  uint32_t n = p->n = d->nnodes;
  p->pos = dup3(d->pos, n, 3);
  p->vol = dup3(d->vol, n, 1);
  p->disp = dup3(d->disp, n, 3);
  p->vel = dup3(d->vel, n, 3);
  p->acc = dup3(d->acc, n, 3);
  p->ext = dup3(d->extforce, n, 3);
  p->cp = dup3(NULL, n, 3);
  p->force = dup3(NULL, n, 3);
  p->nconn = d->nconn;
  p->conn = (uint32_t*)malloc((d->nconn * 2 + 1) * sizeof(uint32_t));
  if (d->nconn) memcpy(p->conn, d->conn, d->nconn * 2 * sizeof(uint32_t));
  p->nclamped = d->nclamped;
  p->clamped = (uint32_t*)malloc((d->nclamped + 1) * sizeof(uint32_t));
  if (d->nclamped) memcpy(p->clamped, d->clamped, d->nclamped * sizeof(uint32_t));
  p->nbdry = d->nbdry;
  p->bdry = (uint32_t*)malloc((d->nbdry + 1) * sizeof(uint32_t));
  if (d->nbdry) memcpy(p->bdry, d->bdry, d->nbdry * sizeof(uint32_t));
  p->delta = d->delta;
  p->rho = d->rho;
  p->cnot = d->cnot;
  p->snot = d->snot;
  p->movable = d->movable;
  p->breakable = d->breakable;
  p->stoppable = d->stoppable;
  p->taxis = d->torque_axis;
  p->tval = d->torque_val;
  for (uint32_t i = 0; i < d->nclamped; i++)
    if (d->clamped[i] >= n) return fail(c, PD_E_ARG, "clamped node out of range");
  for (uint32_t i = 0; i < d->nbdry; i++)
    if (d->bdry[i] >= n) return fail(c, PD_E_ARG, "boundary node out of range");
  if (index) *index = c->np;
  c->np++;
  return PD_OK;
}

static void row_push(orow* r, uint32_t j) {
  if (r->n == r->cap) {
    r->cap = r->cap ? 2 * r->cap : 16;
    r->nbr = (uint32_t*)realloc(r->nbr, r->cap * sizeof(uint32_t));
  }
  r->nbr[r->n++] = j;
}

/* conn2NArr (load_hdf5.h:27-37) then gen_xi (particle2.h:241-257). */
static int conn2nbr(or_ctx* c, opart* p) {
  p->rows = (orow*)calloc(p->n, sizeof(orow));
  for (uint64_t l = 0; l < p->nconn; l++) {
    uint32_t i = p->conn[2 * l], j = p->conn[2 * l + 1];
    if (i >= p->n || j >= p->n) return fail(c, PD_E_ARG, "Connectivity row %llu out of range", (unsigned long long)l);
    row_push(&p->rows[i], j);
    row_push(&p->rows[j], i);
  }
  for (uint32_t i = 0; i < p->n; i++) {
    orow* r = &p->rows[i];
    r->xi = (double*)malloc((3 * (size_t)r->n + 1) * sizeof(double));
    r->xin = (double*)malloc(((size_t)r->n + 1) * sizeof(double));
    r->str = (double*)calloc((size_t)r->n + 1, sizeof(double));
    for (uint32_t j = 0; j < r->n; j++) {
      uint32_t q = r->nbr[j];
      for (int k = 0; k < 3; k++) r->xi[3 * j + k] = p->pos[3 * q + k] - p->pos[3 * i + k];
      r->xin[j] = nrm3(r->xi[3 * j], r->xi[3 * j + 1], r->xi[3 * j + 2]);
    }
  }
  free(p->conn);
  p->conn = NULL;
  return PD_OK;
}

/* get_minmax<3> (contact.h:1186-1225). */
static void minmax(const opart* p, double out[6]) {
  for (int k = 0; k < 3; k++) out[k] = out[3 + k] = p->cp[k];
  for (uint32_t i = 0; i < p->n; i++)
    for (int k = 0; k < 3; k++) {
      double v = p->cp[3 * i + k];
      out[k] = (v < out[k]) ? v : out[k];             /* std::min */
      out[3 + k] = (out[3 + k] < v) ? v : out[3 + k]; /* std::max */
    }
}

int or_commit(or_ctx* c) {
  if (!c) return PD_E_ARG;
  if (c->committed) return fail(c, PD_E_STATE, "already committed");
  if (c->np == 0) return fail(c, PD_E_STATE, "no particles");
  c->gidx = (uint64_t*)malloc((c->np + 1) * sizeof(uint64_t));
  c->ntot = 0;
  for (uint32_t i = 0; i < c->np; i++) {
    c->gidx[i] = c->ntot;
    c->ntot += c->P[i].n;
    int rc = conn2nbr(c, &c->P[i]);
    if (rc) return rc;
    for (uint32_t v = 0; v < c->P[i].n; v++) c->P[i].nbonds0 += c->P[i].rows[v].n;
    c->cnt.total_bonds += c->P[i].nbonds0;
  }
  c->gidx[c->np] = c->ntot;
  c->gf = (double*)calloc(c->ntot * 3 + 1, sizeof(double));
  c->lrtb = (double(*)[6])calloc(c->np, sizeof(double[6]));
  /* timeloop.h:389-393: CurrPos = pos + disp; lrtb = minmax(CurrPos). */
  for (uint32_t i = 0; i < c->np; i++) {
    opart* p = &c->P[i];
    for (uint64_t k = 0; k < 3ull * p->n; k++) p->cp[k] = p->pos[k] + p->disp[k];
    minmax(p, c->lrtb[i]);
  }
  c->cnt.nodes = c->ntot;
  c->cnt.particles = c->np;
  c->committed = 1;
  return PD_OK;
}

/* get_peridynamic_force (particle2.h:107-193), row order, stretch stored. */
static void pd_force_particle(opart* p, double* gf) {
  for (uint32_t i = 0; i < p->n; i++) {
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    orow* r = &p->rows[i];
    const double* ui = p->disp + 3 * (size_t)i;
    for (uint32_t j = 0; j < r->n; j++) {
      uint32_t q = r->nbr[j];
      const double* uq = p->disp + 3 * (size_t)q;
      double e0 = (uq[0] - ui[0]) + r->xi[3 * j];
      double e1 = (uq[1] - ui[1]) + r->xi[3 * j + 1];
      double e2 = (uq[2] - ui[2]) + r->xi[3 * j + 2];
      double nn = nrm3(e0, e1, e2);
      double L = r->xin[j];
      double diff = nn - L;
      double s = diff / L;
      double cf = (p->cnot * s) * p->vol[q];
      v0 += cf * (e0 / nn);
      v1 += cf * (e1 / nn);
      v2 += cf * (e2 / nn);
      r->str[j] = s;
    }
    /* giant_f += temp_ft_i (timeloop.h:546-549) */
    gf[3 * (size_t)i] += v0;
    gf[3 * (size_t)i + 1] += v1;
    gf[3 * (size_t)i + 2] += v2;
  }
}

/* remove_bonds (particle2.h:196-237): stable erase of entries with s > snot. */
static void remove_bonds(opart* p) {
  for (uint32_t i = 0; i < p->n; i++) {
    orow* r = &p->rows[i];
    uint32_t w = 0;
    for (uint32_t j = 0; j < r->n; j++) {
      if (r->str[j] > p->snot) continue;
      r->nbr[w] = r->nbr[j];
      r->xi[3 * w] = r->xi[3 * j];
      r->xi[3 * w + 1] = r->xi[3 * j + 1];
      r->xi[3 * w + 2] = r->xi[3 * j + 2];
      r->xin[w] = r->xin[j];
      r->str[w] = r->str[j];
      w++;
    }
    r->n = w;
  }
}

static int row_has(const orow* r, uint32_t j) {
  for (uint32_t k = 0; k < r->n; k++)
    if (r->nbr[k] == j) return 1;
  return 0;
}

/* update_contact_force_by_boundary (contact.h:746-942): the terms of one
 * node pair, i of P_C = C and j of P_N = N, added to giant_f.  The distance
 * and self-contact tests come first, as in the reference's loop body. */
static void contact_pair(const or_ctx* c, const opart* C, const opart* N, double* gf, uint64_t ioff,
                         uint64_t joff, int self, uint32_t i, uint32_t j) {
  const pd_run_desc* R = &c->run;
  const double Rc = R->contact_rad, dt = R->dt;
  const double* Xi = C->pos + 3 * (size_t)i;
  const double* Ui = C->disp + 3 * (size_t)i;
  const double* Xj = N->pos + 3 * (size_t)j;
  const double* Uj = N->disp + 3 * (size_t)j;
  double xi[3], d[3];
  for (int k = 0; k < 3; k++) xi[k] = Xj[k] - Xi[k];
  for (int k = 0; k < 3; k++) d[k] = (Uj[k] - Ui[k]) + xi[k];
  double dn = nrm3(d[0], d[1], d[2]);
  if (!(dn <= Rc)) return;
  int do_compute = 1;
  double xin = nrm3(xi[0], xi[1], xi[2]);
  double len = Rc;
  double K = R->normal_stiffness * Rc;
  if (self) {
    if (i == j) do_compute = 0;
    else if (row_has(&C->rows[i], j)) do_compute = 0;
    else if (xin < Rc) {
      len = xin;
      K = N->cnot;
    } else
      len = Rc;
  }
  if (!(do_compute && dn < len)) return;
  double Vn = N->vol[j];
  double u[3] = {d[0] / dn, d[1] / dn, d[2] / dn};
  double s = (dn - len) / len;
  double fc[3];
  for (int k = 0; k < 3; k++) fc[k] = ((K * s) * Vn) * u[k];
  double fcn = nrm3(fc[0], fc[1], fc[2]);
  double* Fi = gf + 3 * (ioff + i);
  double* Fj = gf + 3 * (joff + j);
  if (C->stoppable)
    for (int k = 0; k < 3; k++) Fi[k] += fc[k];
  if (!self && N->stoppable)
    for (int k = 0; k < 3; k++) Fj[k] -= (fc[k] / Vn) * C->vol[i];
  if (R->allow_damping) {
    double rv[3];
    for (int k = 0; k < 3; k++) rv[k] = C->vel[3 * (size_t)i + k] - N->vel[3 * (size_t)j + k];
    double pr = dot3(rv, u);
    double dc = ((2 * R->damping_ratio) * sqrt(K * C->rho)) * sqrt(Vn);
    double fd[3];
    for (int k = 0; k < 3; k++) fd[k] = (-dc) * (pr * u[k]);
    if (C->stoppable)
      for (int k = 0; k < 3; k++) Fi[k] += fd[k];
    if (!self && N->stoppable)
      for (int k = 0; k < 3; k++) Fj[k] -= (fd[k] / sqrt(Vn)) * sqrt(C->vol[i]);
  }
  if (R->allow_friction) {
    double rv[3];
    for (int k = 0; k < 3; k++) rv[k] = C->vel[3 * (size_t)i + k] - N->vel[3 * (size_t)j + k];
    double pr = dot3(rv, u);
    double tv[3];
    for (int k = 0; k < 3; k++) tv[k] = rv[k] - pr * u[k];
    double tn = nrm3(tv[0], tv[1], tv[2]);
    if (tn > 0) {
      double tu[3], ff[3];
      for (int k = 0; k < 3; k++) tu[k] = tv[k] / tn;
      for (int k = 0; k < 3; k++) ff[k] = ((-R->friction_coefficient) * fcn) * tu[k];
      if (tn < ((R->friction_coefficient * fcn) * dt) / C->rho)
        for (int k = 0; k < 3; k++) ff[k] = (((-tn) / C->rho) / dt) * tu[k];
      if (C->stoppable)
        for (int k = 0; k < 3; k++) Fi[k] += ff[k];
      if (!self && N->stoppable)
        for (int k = 0; k < 3; k++) Fj[k] -= ff[k];
    }
  }
}

/* Candidate pruning (or_set_prune, off by default): the nodes of N are put
 * in cells of edge R_c (1 + 1e-6) by their current position, and node i
 * visits only the members of its 27 neighbouring cells, in the full loop's
 * own order (position in N's node list or boundary list).  Every pair within
 * R_c -- rounding moves computed distances by ~1e-17 m against a margin of
 * 1e-6 R_c -- is met, in the same order, so the sums are bitwise those of
 * the full loop; only pairs the distance test would reject are skipped. */
typedef struct {
  int64_t x, y, z;
  uint32_t q;
} ocell;

static int ocell_cmp(const void* a, const void* b) {
  const ocell *u = (const ocell*)a, *v = (const ocell*)b;
  if (u->x != v->x) return u->x < v->x ? -1 : 1;
  if (u->y != v->y) return u->y < v->y ? -1 : 1;
  if (u->z != v->z) return u->z < v->z ? -1 : 1;
  return u->q < v->q ? -1 : (u->q > v->q);
}

static int u32_cmp(const void* a, const void* b) {
  const uint32_t u = *(const uint32_t*)a, v = *(const uint32_t*)b;
  return u < v ? -1 : (u > v);
}

static void contact_pruned(const or_ctx* c, const opart* C, const opart* N, double* gf, uint64_t ioff,
                           uint64_t joff, int self) {
  const pd_run_desc* R = &c->run;
  const double cell = R->contact_rad * (1 + 1e-6);
  const uint32_t tni = R->nl_bdry_only ? C->nbdry : C->n;
  const uint32_t tnj = R->nl_bdry_only ? N->nbdry : N->n;
  if (!tni || !tnj) return;
  ocell* tab = (ocell*)malloc(tnj * sizeof(ocell));
  uint32_t* cand = (uint32_t*)malloc(tnj * sizeof(uint32_t));
  for (uint32_t qq = 0; qq < tnj; qq++) {
    const uint32_t j = R->nl_bdry_only ? N->bdry[qq] : qq;
    const double* X = N->pos + 3 * (size_t)j;
    const double* U = N->disp + 3 * (size_t)j;
    tab[qq].x = (int64_t)floor((X[0] + U[0]) / cell);
    tab[qq].y = (int64_t)floor((X[1] + U[1]) / cell);
    tab[qq].z = (int64_t)floor((X[2] + U[2]) / cell);
    tab[qq].q = qq;
  }
  qsort(tab, tnj, sizeof(ocell), ocell_cmp);
  for (uint32_t pp = 0; pp < tni; pp++) {
    const uint32_t i = R->nl_bdry_only ? C->bdry[pp] : pp;
    const double* X = C->pos + 3 * (size_t)i;
    const double* U = C->disp + 3 * (size_t)i;
    const int64_t cx = (int64_t)floor((X[0] + U[0]) / cell);
    const int64_t cy = (int64_t)floor((X[1] + U[1]) / cell);
    const int64_t cz = (int64_t)floor((X[2] + U[2]) / cell);
    uint32_t nc = 0;
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          const ocell key = {cx + dx, cy + dy, cz + dz, 0};
          size_t lo = 0, hi = tnj;   /* first entry >= key */
          while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (ocell_cmp(&tab[mid], &key) < 0) lo = mid + 1;
            else hi = mid;
          }
          for (; lo < tnj && tab[lo].x == key.x && tab[lo].y == key.y && tab[lo].z == key.z; lo++)
            cand[nc++] = tab[lo].q;
        }
    qsort(cand, nc, sizeof(uint32_t), u32_cmp);
    for (uint32_t k = 0; k < nc; k++) {
      const uint32_t j = R->nl_bdry_only ? N->bdry[cand[k]] : cand[k];
      contact_pair(c, C, N, gf, ioff, joff, self, i, j);
    }
  }
  free(cand);
  free(tab);
}

/* update_contact_force_by_boundary (contact.h:746-942).  C = P_C, N = P_N. */
static void contact(const or_ctx* c, const opart* C, const opart* N, double* gf,
                    uint64_t ioff, uint64_t joff, int self) {
  if (c->prune) {
    contact_pruned(c, C, N, gf, ioff, joff, self);
    return;
  }
  const pd_run_desc* R = &c->run;
  uint32_t tni = R->nl_bdry_only ? C->nbdry : C->n;
  uint32_t tnj = R->nl_bdry_only ? N->nbdry : N->n;
  for (uint32_t pp = 0; pp < tni; pp++) {
    uint32_t i = R->nl_bdry_only ? C->bdry[pp] : pp;
    for (uint32_t qq = 0; qq < tnj; qq++) {
      uint32_t j = R->nl_bdry_only ? N->bdry[qq] : qq;
      contact_pair(c, C, N, gf, ioff, joff, self, i, j);
    }
  }
}

/* wall_side_force_update<3> (contact.h:385-502). */
static void wall_side(const or_ctx* c, double dist, unsigned dir, const double* vel,
                      double vol, double rho, double* f, double speed, double* reac) {
  const pd_run_desc* R = &c->run;
  const double Rc = R->contact_rad;
  if (!(fabs(dist) <= Rc)) return;
  double ca[2];
  or_segment_mass3(Rc, fabs(dist), ca);
  double contrib = Rc - ca[0];
  double pc = contrib > 0 ? contrib : 0;
  double F = (R->normal_stiffness * pc) * ca[1];
  double sg = dist > 0 ? 1 : -1;
  f[dir] += sg * F;
  reac[dir] -= (vol * sg) * F;
  double rv[3] = {vel[0], vel[1], vel[2]};
  rv[dir] -= speed;
  if (R->allow_friction) {
    unsigned t1 = (dir + 1) % 3, t2 = (dir + 2) % 3;
    double p1 = rv[t1], p2 = rv[t2];
    double pn = sqrt(p1 * p1 + p2 * p2);
    if (pn > 0) {
      double u1 = p1 / pn, u2 = p2 / pn;
      double ff1 = ((-R->friction_coefficient) * fabs(F)) * u1;
      double ff2 = ((-R->friction_coefficient) * fabs(F)) * u2;
      if (pn < ((R->friction_coefficient * fabs(F)) * R->dt) / rho) {
        ff1 = (((-pn) / rho) / R->dt) * u1;
        ff2 = (((-pn) / rho) / R->dt) * u2;
      }
      f[t1] += ff1;
      f[t2] += ff2;
      reac[t1] -= vol * ff1;
      reac[t2] -= vol * ff2;
    }
  }
  if (R->allow_damping) {
    double dc = ((2 * R->damping_ratio) * sqrt(R->normal_stiffness * rho)) * sqrt(fabs(ca[1]));
    double fd = (-dc) * rv[dir];
    f[dir] += fd;
    reac[dir] -= vol * fd;
  }
}

/* update_wall_contact_force_by_boundary, dim 3 (contact.h:1060-1124). */
static void wall_contact(or_ctx* c, const opart* p, double* gf, double reac[6][3]) {
  uint32_t tn = c->run.nl_bdry_only ? p->nbdry : p->n;
  for (uint32_t pp = 0; pp < tn; pp++) {
    uint32_t i = c->run.nl_bdry_only ? p->bdry[pp] : pp;
    const double* cp = p->cp + 3 * (size_t)i;
    const double* v = p->vel + 3 * (size_t)i;
    double* f = gf + 3 * (size_t)i;
    double vol = p->vol[i];
    /* plane order x_min, x_max, y_min, y_max, z_min, z_max */
    wall_side(c, cp[0] - c->wall[0], 0, v, vol, p->rho, f, c->speed[0], reac[0]);
    wall_side(c, cp[0] - c->wall[3], 0, v, vol, p->rho, f, c->speed[3], reac[3]);
    wall_side(c, cp[1] - c->wall[1], 1, v, vol, p->rho, f, c->speed[1], reac[1]);
    wall_side(c, cp[1] - c->wall[4], 1, v, vol, p->rho, f, c->speed[4], reac[4]);
    wall_side(c, cp[2] - c->wall[2], 2, v, vol, p->rho, f, c->speed[2], reac[2]);
    wall_side(c, cp[2] - c->wall[5], 2, v, vol, p->rho, f, c->speed[5], reac[5]);
  }
}

/* within_interior_fext (contact.h:1229-1243): indexes the 3-D box as 2-D. */
static int within_interior_fext(const double* A, const double* B, double cc) {
  double AL = A[0] - cc, AR = A[1] + cc, AT = A[2] + cc, AB = A[3] - cc;
  return (B[0] < AL) && (AR < B[1]) && (B[2] > AT) && (B[3] < AB);
}

/* intersects_box_fext<3> (contact.h:1143-1183). */
static int intersects_box_fext(const double* A, const double* B, double cc) {
  double Aminx = A[0] - cc, Aminy = A[1] - cc, Aminz = A[2] - cc;
  double Amaxx = A[3] + cc, Amaxy = A[4] + cc, Amaxz = A[5] + cc;
  return (Aminx <= B[3]) && (Amaxx >= B[0]) && (Amaxy >= B[1]) && (Aminy <= B[4]) &&
         (Aminz <= B[5]) && (Amaxz >= B[2]);
}

/* external torque about the centroid (timeloop.h:601-633). */
static void torque(const opart* p, double* gf) {
  double m[3] = {0, 0, 0};
  for (uint32_t i = 0; i < p->n; i++)
    for (int k = 0; k < 3; k++) m[k] += (p->pos[3 * (size_t)i + k] + p->disp[3 * (size_t)i + k]);
  for (int k = 0; k < 3; k++) m[k] /= (double)p->n;
  unsigned ta = p->taxis;
  for (uint32_t i = 0; i < p->n; i++) {
    double r[3], rp[3];
    for (int k = 0; k < 3; k++) r[k] = p->cp[3 * (size_t)i + k] - m[k];
    double rn = nrm3(r[0], r[1], r[2]);
    double rinv = rn != 0 ? 1 / rn : 0;
    r[ta] = 0;
    rp[ta] = 0;
    rp[(ta + 1) % 3] = -r[(ta + 2) % 3];
    rp[(ta + 2) % 3] = r[(ta + 1) % 3];
    double pn = nrm3(rp[0], rp[1], rp[2]);
    if (pn > 0) {
      for (int k = 0; k < 3; k++) rp[k] /= pn;
    } else {
      rp[(ta + 1) % 3] = 0;
      rp[(ta + 2) % 3] = 0;
    }
    for (int k = 0; k < 3; k++) gf[3 * (size_t)i + k] += (rinv * p->tval) * rp[k];
  }
}

/* Work of one (emulated) rank: timeloop.h:544-651. */
static uint64_t rank_work(or_ctx* c, int rank, int nranks, double* gf, double reac[6][3]) {
  const pd_run_desc* R = &c->run;
  for (uint32_t i = rank; i < c->np; i += nranks) {
    opart* p = &c->P[i];
    double* g = gf + 3 * c->gidx[i];
    if (p->movable) {
      pd_force_particle(p, g);
      if (R->enable_fracture && p->breakable) {
        remove_bonds(p);
        if (R->self_contact) contact(c, p, p, gf, c->gidx[i], c->gidx[i], 1);
      }
      if (p->stoppable && R->allow_wall && !within_interior_fext(c->lrtb[i], c->wall, R->contact_rad))
        wall_contact(c, p, g, reac);
      if (R->gradient_extforce) {
        double sc = (double)c->t / (double)R->extforce_maxstep;
        if (sc > 1) sc = 1;
        for (uint64_t k = 0; k < 3ull * p->n; k++) g[k] += sc * p->ext[k];
      } else {
        for (uint64_t k = 0; k < 3ull * p->n; k++) g[k] += p->ext[k];
      }
    }
    if (R->enable_torque) torque(p, g);
  }
  uint64_t pw = 0, pidx = 0;
  for (uint32_t i = 0; i < c->np; i++)
    for (uint32_t j = i + 1; j < c->np; j++, pidx++) {
      if ((int)(pidx % (uint64_t)nranks) != rank) continue;
      if (!(c->P[i].movable || c->P[j].movable)) continue;
      if (!intersects_box_fext(c->lrtb[i], c->lrtb[j], R->contact_rad)) continue;
      contact(c, &c->P[i], &c->P[j], gf, c->gidx[i], c->gidx[j], 0);
      pw++;
    }
  return pw;
}

/* Prescribed velocity of one particle (timeloop.h:685-764), applied after
 * its velocity update and before its clamped nodes are reset.  The radial /
 * tangential split of the reference (v_r, v_t, t_unit) feeds nothing and is
 * left out; the applied force density is formed from the assigned velocity
 * exactly as the reference forms it. */
static void prescribe(or_ctx* c, opart* p, double dt) {
  const uint32_t n = p->n;
  double mean[3] = {0, 0, 0};
  for (uint32_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) mean[k] += p->vel[3 * (size_t)i + k];
  for (int k = 0; k < 3; k++) mean[k] /= (double)n;
  double m[3] = {mean[0], mean[1], mean[2]};
  if (c->vc_fix_mean_x) m[0] = c->vc_mean_x;
  double cen[3] = {0, 0, 0};
  if (c->vc_fix_ang) {
    for (uint32_t i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) cen[k] += (p->pos[3 * (size_t)i + k] + p->disp[3 * (size_t)i + k]);
    for (int k = 0; k < 3; k++) cen[k] /= (double)n;
  }
  for (uint32_t i = 0; i < n; i++) {
    double* v = p->vel + 3 * (size_t)i;
    double vr[3] = {v[0] - mean[0], v[1] - mean[1], v[2] - mean[2]};
    if (c->vc_fix_ang) {
      double rv[3];
      for (int k = 0; k < 3; k++) rv[k] = p->cp[3 * (size_t)i + k] - cen[k];
      double r = nrm3(rv[0], rv[1], rv[2]);
      double ru[3] = {0, 0, 0};
      if (r > 0)
        for (int k = 0; k < 3; k++) ru[k] = rv[k] / r;
      vr[0] = (-(r * c->vc_omega)) * ru[1];
      vr[1] = (r * c->vc_omega) * ru[0];
      vr[2] = 0;
    }
    double pv[3];
    for (int k = 0; k < 3; k++) pv[k] = vr[k] + m[k];
    for (int k = 0; k < 3; k++) {
      v[k] = pv[k];
      c->afd[3 * (size_t)i + k] =
          p->rho * (((2 / dt) * (pv[k] - v[k])) - c->aold[3 * (size_t)i + k]) - p->force[3 * (size_t)i + k];
    }
  }
}

static void one_step(or_ctx* c) {
  const pd_run_desc* R = &c->run;
  const double dt = R->dt;
  memset(c->reaction, 0, sizeof c->reaction);
  memset(c->gf, 0, c->ntot * 3 * sizeof(double));
  /* Verlet position half (timeloop.h:537-540, overloads.h:51-70,149-159). */
#pragma omp parallel for schedule(dynamic) if (c->nranks > 1)
  for (uint32_t i = 0; i < c->np; i++) {
    opart* p = &c->P[i];
    for (uint64_t k = 0; k < 3ull * p->n; k++) {
      p->disp[k] = p->disp[k] + ((dt * p->vel[k]) + (((dt * dt) * 0.5) * p->acc[k]));
      p->cp[k] = p->pos[k] + p->disp[k];
    }
  }
  /* bond-updates: directed entries of movable particles, broken ones included */
  for (uint32_t i = 0; i < c->np; i++)
    if (c->P[i].movable) c->cnt.bond_updates += c->P[i].nbonds0;
  uint64_t pw = 0;
  if (c->nranks <= 1) {
    pw = rank_work(c, 0, 1, c->gf, c->reaction);
  } else {
    int nr = c->nranks;
    double(*rr)[6][3] = (double(*)[6][3])calloc(nr, sizeof(double[6][3]));
    uint64_t* pwr = (uint64_t*)calloc(nr, sizeof(uint64_t));
#pragma omp parallel for schedule(static, 1) num_threads(nr)
    for (int r = 0; r < nr; r++) {
      double* g = c->gfr + (size_t)r * c->ntot * 3;
      memset(g, 0, c->ntot * 3 * sizeof(double));
      pwr[r] = rank_work(c, r, nr, g, rr[r]);
    }
    /* MPI_Allreduce(SUM) in rank order; rank 0's wall reaction is kept. */
#pragma omp parallel for num_threads(nr)
    for (uint64_t k = 0; k < c->ntot * 3; k++) {
      double s = c->gfr[k];
      for (int r = 1; r < nr; r++) s += c->gfr[(size_t)r * c->ntot * 3 + k];
      c->gf[k] = s;
    }
    memcpy(c->reaction, rr[0], sizeof c->reaction);
    for (int r = 0; r < nr; r++) pw += pwr[r];
    free(rr);
    free(pwr);
  }
  c->cnt.pair_computations += pw;
  /* Verlet velocity half (timeloop.h:662-781). */
#pragma omp parallel for schedule(dynamic) if (c->nranks > 1)
  for (uint32_t i = 0; i < c->np; i++) {
    opart* p = &c->P[i];
    const double* g = c->gf + 3 * c->gidx[i];
    memcpy(p->force, g, 3ull * p->n * sizeof(double));
    if (!p->movable) continue;
    double rinv = 1 / p->rho;
    const int vc = c->vc_on && (int64_t)i == c->vc_part;
    for (uint64_t k = 0; k < 3ull * p->n; k++) {
      double aold = p->acc[k];
      double anew = rinv * p->force[k];
      if (vc) c->aold[k] = aold;
      p->acc[k] = anew;
      p->vel[k] = p->vel[k] + ((aold + anew) * (0.5 * dt));
    }
    if (vc) prescribe(c, p, dt);
    for (uint32_t q = 0; q < p->nclamped; q++) {
      uint32_t cid = p->clamped[q];
      for (int k = 0; k < 3; k++) p->disp[3 * (size_t)cid + k] = p->vel[3 * (size_t)cid + k] = p->acc[3 * (size_t)cid + k] = 0;
    }
    minmax(p, c->lrtb[i]);
  }
  /* wall boundary update (timeloop.h:784-797). */
  for (int k = 0; k < 6; k++) c->wall[k] += c->speed[k] * dt;
  c->t++;
  c->cnt.steps++;
}

int or_step(or_ctx* c, uint32_t nsteps) {
  if (!c) return PD_E_ARG;
  if (!c->committed) return fail(c, PD_E_STATE, "or_step before commit");
  if (c->nranks > 1 && !c->gfr) c->gfr = (double*)malloc((size_t)c->nranks * c->ntot * 3 * sizeof(double));
  for (uint32_t s = 0; s < nsteps; s++) one_step(c);
  return PD_OK;
}

int or_set_prune(or_ctx* c, int on) {
  if (!c) return PD_E_ARG;
  c->prune = on != 0;
  return PD_OK;
}

int or_set_ranks(or_ctx* c, int nranks) {
  if (!c || nranks < 1) return PD_E_ARG;
  if (nranks != c->nranks) {
    free(c->gfr);
    c->gfr = NULL;
  }
  c->nranks = nranks;
  return PD_OK;
}

int or_set_flag(or_ctx* c, int flag, int value) {
  if (!c) return PD_E_ARG;
  switch (flag) {
    case PD_FLAG_ENABLE_FRACTURE: c->run.enable_fracture = value; break;
    case PD_FLAG_SELF_CONTACT: c->run.self_contact = value; break;
    case PD_FLAG_ALLOW_DAMPING: c->run.allow_damping = value; break;
    case PD_FLAG_ALLOW_FRICTION: c->run.allow_friction = value; break;
    case PD_FLAG_ENABLE_TORQUE: c->run.enable_torque = value; break;
    case PD_FLAG_COMPUTE_WALL_REACTION: c->run.compute_wall_reaction = value; break;
    case PD_FLAG_VELOCITY_CONSTRAINT: c->vc_on = value; break;
    default: return fail(c, PD_E_ARG, "unknown flag %d", flag);
  }
  return PD_OK;
}

int or_set_wall_speed(or_ctx* c, const double speed[6]) {
  if (!c || !speed) return PD_E_ARG;
  memcpy(c->speed, speed, sizeof c->speed);
  return PD_OK;
}

int or_set_wall(or_ctx* c, const double lrtb[6]) {
  if (!c || !lrtb) return PD_E_ARG;
  memcpy(c->wall, lrtb, sizeof c->wall);
  return PD_OK;
}

int or_set_particle_flag(or_ctx* c, uint32_t p, int which, int value) {
  if (!c || p >= c->np) return PD_E_ARG;
  switch (which) {
    case PD_PFLAG_MOVABLE: c->P[p].movable = value; break;
    case PD_PFLAG_STOPPABLE: c->P[p].stoppable = value; break;
    case PD_PFLAG_BREAKABLE: c->P[p].breakable = value; break;
    default: return fail(c, PD_E_ARG, "unknown particle flag %d", which);
  }
  return PD_OK;
}

int or_set_snot(or_ctx* c, uint32_t p, double snot) {
  if (!c || p >= c->np) return PD_E_ARG;
  c->P[p].snot = snot;
  return PD_OK;
}

/* Wheel reset (timeloop.h:496-535), applied before the position half of the
 * step it precedes.  Particle 0's mean current position m = mean_CurrPos()
 * (particle2.h:94-103: sum of pos + disp in node order, then / nnodes); the
 * bulk height is the largest pos_y + disp_y over the nodes of particles
 * 1.. whose |m_x - (pos_x + disp_x)| <= wheel_rad, the first such node in
 * particle and node order starting the running maximum; with no such node
 * it stays at particle 1's pos[0]_y + disp[1]_y (node 1's displacement, as
 * the reference initialises it).  Particle 0's reference positions move in y
 * by (bulk + wheel_rad + contact_rad) - m_y.  Its bond vectors xi keep their
 * values from gen_xi (particle2.h:241-257): only pos moves. */
int or_reset_partzero_y(or_ctx* c, double wheel_rad) {
  if (!c) return PD_E_ARG;
  if (!c->committed) return fail(c, PD_E_STATE, "reset_partzero_y before commit");
  if (c->np < 2 || c->P[1].n < 2)
    return fail(c, PD_E_ARG, "reset_partzero_y reads node 1 of particle 1");
  const opart* P0 = &c->P[0];
  double m[3] = {0.0, 0.0, 0.0};
  for (uint32_t i = 0; i < P0->n; i++)
    for (int k = 0; k < 3; k++) m[k] += (P0->pos[3 * (size_t)i + k] + P0->disp[3 * (size_t)i + k]);
  for (int k = 0; k < 3; k++) m[k] /= (double)P0->n;
  double top = c->P[1].pos[1] + c->P[1].disp[3 + 1];
  int init = 1;
  for (uint32_t q = 1; q < c->np; q++) {
    const opart* p = &c->P[q];
    for (uint32_t j = 0; j < p->n; j++) {
      const double now_x = p->pos[3 * (size_t)j] + p->disp[3 * (size_t)j];
      const double now_y = p->pos[3 * (size_t)j + 1] + p->disp[3 * (size_t)j + 1];
      if (fabs(m[0] - now_x) <= wheel_rad) {
        if (init) {
          top = now_y;
          init = 0;
        } else if (now_y > top) {
          top = now_y;
        }
      }
    }
  }
  const double dest = top + wheel_rad + c->run.contact_rad;
  const double move = dest - m[1];
  opart* W = &c->P[0];
  for (uint32_t j = 0; j < W->n; j++) W->pos[3 * (size_t)j + 1] += move;
  return PD_OK;
}

int or_get_state(or_ctx* c, uint32_t pi, double* disp, double* currpos, double* vel,
                 double* acc, double* force) {
  if (!c || pi >= c->np) return PD_E_ARG;
  opart* p = &c->P[pi];
  size_t b = 3ull * p->n * sizeof(double);
  if (disp) memcpy(disp, p->disp, b);
  if (currpos) memcpy(currpos, p->cp, b);
  if (vel) memcpy(vel, p->vel, b);
  if (acc) memcpy(acc, p->acc, b);
  if (force) memcpy(force, p->force, b);
  return PD_OK;
}

int or_set_state(or_ctx* c, uint32_t pi, const double* disp, const double* vel, const double* acc) {
  if (!c || pi >= c->np) return PD_E_ARG;
  opart* p = &c->P[pi];
  size_t b = 3ull * p->n * sizeof(double);
  if (disp) {
    memcpy(p->disp, disp, b);
    /* a resumed run starts from lrtb = minmax(pos + disp) (timeloop.h:389-393) */
    if (c->committed) {
      for (uint64_t k = 0; k < 3ull * p->n; k++) p->cp[k] = p->pos[k] + p->disp[k];
      minmax(p, c->lrtb[pi]);
    }
  }
  if (vel) memcpy(p->vel, vel, b);
  if (acc) memcpy(p->acc, acc, b);
  return PD_OK;
}

int or_get_connectivity(or_ctx* c, uint32_t pi, uint32_t* pairs, uint64_t cap, uint64_t* m) {
  if (!c || pi >= c->np || !m) return PD_E_ARG;
  if (!c->committed) return fail(c, PD_E_STATE, "before commit");
  opart* p = &c->P[pi];
  uint64_t cntp = 0;
  for (uint32_t i = 0; i < p->n; i++)
    for (uint32_t j = 0; j < p->rows[i].n; j++)
      if (i < p->rows[i].nbr[j]) cntp++;
  *m = cntp;
  if (!pairs) return PD_OK;
  if (cap < cntp) return fail(c, PD_E_ARG, "connectivity buffer too small");
  uint64_t w = 0;
  for (uint32_t i = 0; i < p->n; i++)
    for (uint32_t j = 0; j < p->rows[i].n; j++) {
      uint32_t q = p->rows[i].nbr[j];
      if (i < q) {
        pairs[2 * w] = i;
        pairs[2 * w + 1] = q;
        w++;
      }
    }
  return PD_OK;
}

int or_get_stretch(or_ctx* c, uint32_t pi, double* out, uint64_t cap, uint64_t* m) {
  if (!c || pi >= c->np || !m) return PD_E_ARG;
  opart* p = &c->P[pi];
  uint64_t n = 0;
  for (uint32_t i = 0; i < p->n; i++) n += p->rows[i].n;
  *m = n;
  if (!out) return PD_OK;
  if (cap < n) return PD_E_ARG;
  uint64_t w = 0;
  for (uint32_t i = 0; i < p->n; i++)
    for (uint32_t j = 0; j < p->rows[i].n; j++) out[w++] = p->rows[i].str[j];
  return PD_OK;
}

int or_get_wall(or_ctx* c, uint32_t instance, double lrtb[6], double reaction[18]) {
  if (!c || instance != 0) return PD_E_ARG;
  if (lrtb) memcpy(lrtb, c->wall, sizeof c->wall);
  if (reaction) memcpy(reaction, c->reaction, sizeof c->reaction);
  return PD_OK;
}

int or_get_counters(or_ctx* c, pd_counters* out) {
  if (!c || !out) return PD_E_ARG;
  *out = c->cnt;
  uint64_t alive = 0;
  for (uint32_t i = 0; i < c->np && c->committed; i++)
    for (uint32_t v = 0; v < c->P[i].n; v++) alive += c->P[i].rows[v].n;
  out->alive_bonds = alive;
  return PD_OK;
}

int or_set_velocity_constraint(or_ctx* c, uint32_t pi, int fix_mean_x, double mean_x, int fix_angular,
                               double angular_vel) {
  if (!c) return PD_E_ARG;
  if (!c->committed) return fail(c, PD_E_STATE, "or_set_velocity_constraint before commit");
  if (pi >= c->np) return fail(c, PD_E_ARG, "no particle %u", pi);
  free(c->afd);
  free(c->aold);
  c->afd = (double*)calloc(3 * (size_t)c->P[pi].n, sizeof(double));
  c->aold = (double*)calloc(3 * (size_t)c->P[pi].n, sizeof(double));
  c->vc_part = pi;
  c->vc_fix_mean_x = fix_mean_x;
  c->vc_mean_x = mean_x;
  c->vc_fix_ang = fix_angular;
  c->vc_omega = angular_vel;
  return PD_OK;
}

int or_get_applied_force(or_ctx* c, uint32_t pi, double* afd) {
  if (!c || !afd) return PD_E_ARG;
  if ((int64_t)pi != c->vc_part || !c->afd) return fail(c, PD_E_ARG, "particle %u has no prescribed velocity", pi);
  memcpy(afd, c->afd, 3 * (size_t)c->P[pi].n * sizeof(double));
  return PD_OK;
}

const char* or_last_error(or_ctx* c) { return c ? c->err : "null context"; }

void or_destroy(or_ctx* c) {
  if (!c) return;
  free(c->afd);
  free(c->aold);
  for (uint32_t i = 0; i < c->np; i++) {
    opart* p = &c->P[i];
    free(p->pos); free(p->disp); free(p->cp); free(p->vel); free(p->acc);
    free(p->force); free(p->ext); free(p->vol); free(p->clamped); free(p->bdry); free(p->conn);
    if (p->rows) {
      for (uint32_t v = 0; v < p->n; v++) {
        free(p->rows[v].nbr); free(p->rows[v].xi); free(p->rows[v].xin); free(p->rows[v].str);
      }
      free(p->rows);
    }
  }
  free(c->P); free(c->gidx); free(c->lrtb); free(c->gf); free(c->gfr);
  free(c);
}
