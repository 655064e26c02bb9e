// imdg_api.hip -- host side of the C ABI declared in include/imdg.h:
// context, device memory, table re-layout, replica geometry (make_box /
// init_cells semantics), launch sequencing and result download.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/imdg.h"
#include "imdg_internal.h"

namespace imdg {
template <typename R>
int step_grid_density(bool nt1, bool lt, size_t lds, int ncu, int lpa,
                      int threads, bool cub, bool ee);
template <typename R>
int step_grid_force(int mode, bool integrate, bool lt, bool hp, bool he,
                    size_t lds, int ncu, int lpa, int threads, bool cub,
                    bool ee);
template <typename R>
int force_threads(int mode);
template <typename R>
int step_grid_move(int ncu);
template <typename R>
hipError_t launch_density(const StepArgs<R>&, const TabDev<R>*, bool, int,
                          hipStream_t);
template <typename R>
hipError_t launch_force(const StepArgs<R>&, const TabDev<R>*, int, bool,
                        int, hipStream_t);
template <typename R>
hipError_t launch_move(const StepArgs<R>&, int, hipStream_t);
template <typename R>
hipError_t launch_finalize(const StepArgs<R>&, bool, bool, hipStream_t);
template <typename R>
hipError_t launch_copyrec(const StepArgs<R>&, int, hipStream_t);
template <typename R>
hipError_t launch_stress(const StepArgs<R>&, const TabDev<R>*, double*,
                         double*, int, hipStream_t);
void launch_build(const BuildArgs&, int, hipStream_t);
const char* noalloc_message();
void launch_check(const RepDev*, Ctl*, const int*, const double4*,
                  const double4*, const double4*, double*, int, hipStream_t);
}  // namespace imdg

using namespace imdg;

namespace {

enum { KC_BUILD = 0, KC_DENSITY, KC_FORCE, KC_MOVE, KC_FINAL, KC_N };
const char* kKernelNames[KC_N] = {"rebuild", "density", "force", "move",
                                  "finalize"};

struct HostTable {
  int ncols = 0;
  std::vector<double> begin, end, step, invstep;
  std::vector<int> len;
  std::vector<double> table;  // (maxsteps + 2) * ncols
  int maxsteps = 0;
  bool present = false;
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct imdg_ctx {
  imdg_config cfg;
  std::string err;
  int dev = 0;
  hipStream_t stream = nullptr;
  bool mixed = false;
  bool cubic = false;   // 4-point or spline tables (cubic segments)
  bool eeam = false;    // EEAM energy term present (tab[3])
  int xcd_remap = 1;
  int list_nt = -1;  // list words non-temporal: -1 by list size
                     // (IMDG_LIST_NT_BYTES), 0 never, 1 always (IMDG_LIST_NT)
  int tile_cap = IMDG_TILE_DEFAULT;  // records of a list-build tile (a
                                     // smaller cap sends more blocks to the
                                     // fallback)
  // tables
  HostTable tab[4];  // pair, rho, embed, EEAM energy term (emod)
  bool tables_set = false;
  double cellsz_tab = 0.0;
  // geometry
  int nrep = 0;
  std::vector<RepDev> rep;
  std::vector<int> natoms;      // per replica
  std::vector<long> in_off;     // input index of the first atom of a replica
  long natoms_tot = 0;
  int nslots = 0, nblk = 0, ncells = 0, nstile = 0;
  int kmax = 0;
  // device memory
  std::vector<DevBuf> bufs;
  RepDev* d_rep = nullptr;
  RepState* d_rst = nullptr;
  Ctl* d_ctl = nullptr;
  int *d_blk_rep = nullptr, *d_cell_rep = nullptr;
  double4 *d_pos[2] = {nullptr, nullptr}, *d_mom[2] = {nullptr, nullptr};
  int* d_aid[2] = {nullptr, nullptr};
  double4* d_xnbl = nullptr;
  int *d_cellid = nullptr, *d_cslot = nullptr, *d_acell = nullptr,
      *d_perm = nullptr;
  signed char* d_atype = nullptr;
  int *d_cell_cnt = nullptr, *d_cell_beg = nullptr, *d_cell_num = nullptr;
  int *d_cell_loc = nullptr, *d_tile_sum = nullptr, *d_tile_off = nullptr,
      *d_tile_rep = nullptr, *d_blkmax = nullptr;
  Tile* d_tiles = nullptr;
  uint32_t* d_nbr = nullptr;
  size_t nbr_bytes = 0;
  int list_k = 0;  // longest list built so far, as last read by the host
  int* d_nnb = nullptr;
  void *d_reca = nullptr, *d_recb = nullptr;
  void *d_rho = nullptr, *d_epot = nullptr, *d_frc = nullptr;
  void *d_eeP = nullptr, *d_eedm = nullptr;  // EEAM: P and 2E'(P) per slot
  double *d_partA = nullptr, *d_partB = nullptr, *d_maxd = nullptr;
  ThermoRec* d_ring = nullptr;
  void* d_tabs = nullptr;        // TabDev<R>[4]
  void* d_knots = nullptr;       // knots of all tables
  void* d_knots64 = nullptr;     // FP64 knots of embedding and EEAM term
  void* d_tab64 = nullptr;       // Tab64[2] (FP32 step kernels)
  // launch plan of the persistent step kernels
  int ncu = 0;
  int nchunk = 0;
  int part_chunk = 0;
  int knots_pair = 0, knots_rho = 0;   // first knot of the rho / embed tables
  int knots_embed = 0;
  int knots_joint = -1;                // interleaved (pair, rho), one species
  int knots_dpoly = -1;                // FP32 rho segment polynomials
  int knots_jointn = -1;               // FP32, several species, one grid:
  int knots_dpolyn = -1;               // force and rho segment records
  int knots_frho = -1;                 // per-type rho segments (LDS), see
  int knots_drho = -1;                 // StepArgs::frho / drho
  int knots_pairn = -1;                // StepArgs::pairn
  int* d_cmeta = nullptr;              // chunk metadata, 16-bit lists
  bool list16 = false;                 // lists in the 16-bit format
  int force_mode = 0;                  // 0 one species, 1 shared grid,
                                       // 2 several, 3 several on one grid
  int lpa_den = 0, lpa_frc = 0;        // lanes per slot of the two loops
  int thr_den = 0, thr_frc = 0;        // threads per block (0: default)
  int units_cap = 0;                   // wavefront partial slots allocated
  int lds_den_lo = 0, lds_den_n = 0;   // knots staged by the density loop
  int lds_frc_lo = 0, lds_frc_n = 0;   // knots staged by the force loop
  int grid_den = 0, grid_frc[2] = {0, 0}, grid_move = 0;
  int units_max = 0;
  // one captured step (rebuild chain + density + force), replayed by
  // imdg_run when graphs are on (IMDG_GRAPH=1) and profiling is off
  bool use_graph = false;
  // fused step: forces + integration in one launch (false: force launch,
  // then the move launch; IMDG_SPLIT_MOVE=1)
  bool split_move = false;
  // the last block of the force / move launch ends the step (true:
  // IMDG_FINAL_KERNEL=0) or a one-block launch after it does (default: a
  // release fence in every block of the launch costs more than the launch)
  bool final_inline = false;
  hipGraphExec_t step_graph = nullptr;
  // per-atom stress tensor (imdg_set_stress): computed after the last step
  // of every imdg_run and after every imdg_move_atoms while enabled
  bool stress_on = false;
  bool stress_valid = false;     // d_stress holds the last step's tensor
  double* d_stress = nullptr;    // 6 per slot: xx yy zz yz zx xy
  double* d_spart = nullptr;     // 6 per IMDG_BLOCK block
  bool atoms_set = false;
  long target_steps = 0;
  // profiling
  bool profiling = false;
  unsigned prof_mask = 0;  // kernel classes timed (bit k: class k)
  long prof_stride = 1;    // time every prof_stride-th launch of a class
  long kc_seen[KC_N] = {};  // launches of each class while profiling
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> event_pool;
  double kc_ms[KC_N] = {};
  long kc_n[KC_N] = {};
};

namespace {

int fail(imdg_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(ctx, call)                                                    \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(ctx, IMDG_ERR_HIP, "%s: %s (%s:%d)", #call,                \
                  hipGetErrorString(e_), __FILE__, __LINE__);                \
  } while (0)

// Lanes per slot of the neighbour loops: 1, 2 or 4.
int lpa_value(int v) { return v >= 4 ? 4 : (v >= 2 ? 2 : 1); }

// Drop the captured step (its kernel arguments hold buffer pointers).
void drop_graph(imdg_ctx* c) {
  if (c->step_graph) (void)hipGraphExecDestroy(c->step_graph);
  c->step_graph = nullptr;
}

int dalloc(imdg_ctx* c, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fail(c, IMDG_ERR_NOMEM, "hipMalloc(%zu): %s", bytes,
                hipGetErrorString(e));
  c->bufs.push_back({*p, bytes});
  return IMDG_OK;
}

void dfree(imdg_ctx* c, void* p) {
  for (auto it = c->bufs.begin(); it != c->bufs.end(); ++it)
    if (it->p == p) {
      hipFree(p);
      c->bufs.erase(it);
      return;
    }
}

void free_all(imdg_ctx* c) {
  for (auto& b : c->bufs) hipFree(b.p);
  c->bufs.clear();
  c->d_nbr = nullptr;
  c->nbr_bytes = 0;
  c->d_stress = nullptr;
  c->d_spart = nullptr;
  c->stress_valid = false;
}

void vec_prod(const double* u, const double* v, double* w) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}

double sprod(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

int copy_table(imdg_ctx* c, HostTable& h, const imdg_table* t, int ncols,
               const char* what) {
  h = HostTable();
  if (!t) return IMDG_OK;
  if (t->ncols != ncols)
    return fail(c, IMDG_ERR_NOTABLE, "%s table has %d columns, need %d", what,
                t->ncols, ncols);
  if (!t->begin || !t->end || !t->invstep || !t->len || !t->table)
    return fail(c, IMDG_ERR_ARG, "%s table has NULL arrays", what);
  h.ncols = ncols;
  h.maxsteps = t->maxsteps;
  h.begin.assign(t->begin, t->begin + ncols);
  h.end.assign(t->end, t->end + ncols);
  h.invstep.assign(t->invstep, t->invstep + ncols);
  // the spline coefficients use step itself (PAIR_INT_SP reads pt.step)
  h.step.resize(ncols);
  for (int k = 0; k < ncols; k++)
    h.step[k] = t->step ? t->step[k] : 1.0 / t->invstep[k];
  h.len.assign(t->len, t->len + ncols);
  for (int k = 0; k < ncols; k++)
    if (h.len[k] < 3 || h.len[k] > h.maxsteps)
      return fail(c, IMDG_ERR_NOTABLE, "%s column %d has length %d", what, k,
                  h.len[k]);
  h.table.assign(t->table, t->table + (size_t)(h.maxsteps + 2) * ncols);
  h.present = true;
  return IMDG_OK;
}

// Cubic segments k = 0..len-1 of column col for the 4-point or spline
// interpolation, as (c0, c1, c2, c3) per segment: value c0 + x (c1 + x (c2
// + x c3)) at fraction x of segment k, in FP64.
//  * 4-point (PAIR_INT3, src/potaccess.h:365-412): the Lagrange cubic
//    through rows k-1 .. k+2 at x = -1, 0, 1, 2; segment 0 is segment 1's
//    cubic continued (the reference reads it with k = 1 and x - 1).  The two
//    guard rows continue the last cubic (init_fourpoint,
//    src/imd_potential.c:1171-1187).
//  * spline (PAIR_INT_SP, src/potaccess.h:414-457): a p_k + b p_k+1 +
//    (a (a^2-1) y2_k + b (b^2-1) y2_k+1) h^2/6 with b = x, a = 1 - x, and the
//    second derivatives y2 of init_spline (src/imd_potential.c:1197-1246):
//    natural at the start; at the end zero slope (radial) or natural.
// Returns false if the column is too short (4 rows).
bool cubic_segments(const HostTable& h, int col, bool radial, int interp,
                    std::vector<double>& seg) {
  const int nc = h.ncols, n = h.len[col];
  if (n < 4) return false;
  std::vector<double> y(n + 2);
  for (int k = 0; k < n; k++) y[k] = h.table[(size_t)k * nc + col];
  seg.assign((size_t)4 * n, 0.0);
  if (interp == IMDG_INTERP_4POINT) {
    y[n] = 4 * y[n - 1] - 6 * y[n - 2] + 4 * y[n - 3] - y[n - 4];
    y[n + 1] = 10 * y[n - 1] - 20 * y[n - 2] + 15 * y[n - 3] - 4 * y[n - 4];
    for (int k = 1; k < n; k++) {
      double p0 = y[k - 1], p1 = y[k], p2 = y[k + 1], p3 = y[k + 2];
      double* q = &seg[(size_t)4 * k];
      q[0] = p1;
      q[1] = -p0 / 3 - p1 / 2 + p2 - p3 / 6;
      q[2] = (p0 + p2) / 2 - p1;
      q[3] = (p3 - p0) / 6 + (p1 - p2) / 2;
    }
    // segment 0: P1(x - 1) expanded in x
    const double* q1 = &seg[4];
    seg[0] = q1[0] - q1[1] + q1[2] - q1[3];
    seg[1] = q1[1] - 2 * q1[2] + 3 * q1[3];
    seg[2] = q1[2] - 3 * q1[3];
    seg[3] = q1[3];
    return true;
  }
  // spline: the tridiagonal sweep of init_spline, uniform step
  const double step = h.step[col];
  std::vector<double> y2(n + 1, 0.0), u(n, 0.0);
  for (int i = 1; i < n - 1; i++) {
    double p = 0.5 * y2[i - 1] + 2.0;
    y2[i] = -0.5 / p;
    double d = (y[i + 1] - 2 * y[i] + y[i - 1]) / step;
    u[i] = (6.0 * d / (2 * step) - 0.5 * u[i - 1]) / p;
  }
  double qn = 0.0, un = 0.0;
  if (radial) {
    qn = 0.5;
    un = (3.0 / step) * (y[n - 2] - y[n - 1]) / step;
  }
  y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
  for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
  y[n] = 2 * y[n - 1] - y[n - 2] + step * step * y2[n - 1];
  y2[n] = 2 * y2[n - 1] - y2[n - 2];
  const double h6 = step * step / 6;
  for (int k = 0; k < n; k++) {
    double* q = &seg[(size_t)4 * k];
    q[0] = y[k];
    q[1] = (y[k + 1] - y[k]) - h6 * (2 * y2[k] + y2[k + 1]);
    q[2] = 3 * h6 * y2[k];
    q[3] = h6 * (y2[k + 1] - y2[k]);
  }
  return true;
}

// Knots of IMD's 3-point interpolation: for rows k = 0..len of a column,
// (p0_k, dv_k) = (t[k], t[k+1] - t[k]); segment k reads knots k and k+1 and
// takes d2v = dv_k+1 - dv_k = t[k+2] - 2 t[k+1] + t[k] (src/potaccess.h:
// 345-349).  Differences in FP64, then rounded to R.  Tables are stored
// pair, rho, embed, so the radial tables form one contiguous knot range.
template <typename R>
int upload_tables(imdg_ctx* c) {
  typedef typename V2<R>::T R2;
  std::vector<R2> kn;
  TabDev<R> td[4];
  memset(td, 0, sizeof td);
  int first[5] = {0, 0, 0, 0, 0};
  for (int t = 0; t < 4; t++) {
    HostTable& h = c->tab[t];
    first[t] = (int)kn.size();
    td[t].ncols = h.present ? h.ncols : 0;
    if (!h.present) continue;
    for (int col = 0; col < h.ncols; col++) {
      td[t].off[col] = (int)kn.size();
      td[t].len[col] = h.len[col];
      td[t].begin[col] = (R)h.begin[col];
      td[t].end[col] = (R)h.end[col];
      td[t].invstep[col] = (R)h.invstep[col];
      td[t].twoistep[col] = (R)(2 * h.invstep[col]);
      {
        const double is = h.invstep[col], c0 = -h.begin[col] * is;
        td[t].ilo[col] = (R)(is - (double)(R)is);
        td[t].t0[col] = (R)c0;
        td[t].t0lo[col] = (R)(c0 - (double)(R)c0);
      }
      if (c->cubic) {
        std::vector<double> seg;
        if (!cubic_segments(h, col, t < 2, c->cfg.interpolation, seg))
          return fail(c, IMDG_ERR_NOTABLE,
                      "column %d of table %d is too short for the "
                      "interpolation", col, t);
        for (size_t k = 0; k < seg.size(); k += 2) {
          R2 q;
          q.x = (R)seg[k];
          q.y = (R)seg[k + 1];
          kn.push_back(q);
        }
        continue;
      }
      int nc = h.ncols;
      for (int k = 0; k <= h.len[col]; k++) {
        double p0 = h.table[(size_t)k * nc + col];
        double p1 = h.table[(size_t)(k + 1) * nc + col];
        R2 q;
        q.x = (R)p0;
        q.y = (R)(p1 - p0);
        kn.push_back(q);
      }
    }
  }
  first[4] = (int)kn.size();
  c->knots_pair = first[0];
  c->knots_rho = first[1];
  c->knots_embed = first[2];
  // one species with pair and rho on the same grid: also a segment table
  // for the force loop at an even knot index, for the len segments k, with
  // dv = p1 - p0 and d2v = p2 - 2 p1 + p0 as PAIR_INT2 forms them
  // (src/potaccess.h:347-349), in FP64 and then rounded once.
  //   FP64: S[k] = (phi p0, phi dv, phi d2v, rho dv), then rho d2v[k].
  //   FP32: Q[k] = (phi p0, phi A, phi B, phi A_lo), then (rho A, rho D)[k],
  //         with A = dv - d2v/2, B = d2v/2, D = d2v: value p0 + chi (A +
  //         B chi) and gradient / (2 istep) A + D chi = A + B (2 chi)
  //         (2B = D exactly in FP32); A_lo = dv - d2v/2 - A, the part of A
  //         that FP32 drops, enters the virial sum.  Segment len-1 is all
  //         zero: the reference evaluates it only at r^2 >= end, where the
  //         shifted radial tables vanish, so a clamp of the table argument
  //         replaces the cutoff tests.
  c->knots_joint = -1;
  const HostTable &hp = c->tab[0], &hr = c->tab[1];
  bool same = !c->cubic && !c->eeam && c->cfg.ntypes == 1 && hp.present &&
              hr.present &&
              hp.begin[0] == hr.begin[0] && hp.end[0] == hr.end[0] &&
              hp.invstep[0] == hr.invstep[0] && hp.len[0] == hr.len[0];
  if (const char* e = getenv("IMDG_JOINT_TABLE"))
    if (atoi(e) == 0) same = false;
  if (same) {
    if (kn.size() & 1) kn.push_back(kn.back());
    c->knots_joint = (int)kn.size();
    const int len = hp.len[0];
    auto at = [](const HostTable& h, int k) {
      return h.table[(size_t)k * h.ncols];
    };
    auto d2v = [&](const HostTable& h, int k) {
      double p0 = at(h, k), p1 = at(h, k + 1), p2 = at(h, k + 2);
      return p2 - 2 * p1 + p0;
    };
    if (sizeof(R) == 8) {
      for (int k = 0; k < len; k++) {
        R2 a, b;
        a.x = (R)at(hp, k);
        a.y = (R)(at(hp, k + 1) - at(hp, k));
        b.x = (R)d2v(hp, k);
        b.y = (R)(at(hr, k + 1) - at(hr, k));
        kn.push_back(a);
        kn.push_back(b);
      }
      for (int k = 0; k < len; k += 2) {
        R2 q;
        q.x = (R)d2v(hr, k);
        q.y = k + 1 < len ? (R)d2v(hr, k + 1) : (R)0;
        kn.push_back(q);
      }
    } else {
      std::vector<R2> rr;
      for (int k = 0; k < len; k++) {
        R2 a, b, r;
        a.x = a.y = b.x = b.y = r.x = r.y = (R)0;
        if (k < len - 1) {
          double dp = at(hp, k + 1) - at(hp, k), ddp = d2v(hp, k);
          double dr = at(hr, k + 1) - at(hr, k), ddr = d2v(hr, k);
          a.x = (R)at(hp, k);
          a.y = (R)(dp - 0.5 * ddp);
          b.x = (R)(0.5 * ddp);
          b.y = (R)((dp - 0.5 * ddp) - (double)a.y);
          r.x = (R)(dr - 0.5 * ddr);
          r.y = (R)ddr;
        }
        kn.push_back(a);
        kn.push_back(b);
        rr.push_back(r);
      }
      kn.insert(kn.end(), rr.begin(), rr.end());
    }
  }
  // FP32, one species: the density loop's rho table as segment polynomials
  // (p0, A, B) (value p0 + chi (A + B chi), A = dv - d2v/2, B = d2v/2),
  // three packed floats per segment, segment len-1 zero, like the force
  // loop's segment table
  c->knots_dpoly = -1;
  if (sizeof(R) == 4 && !c->cubic && !c->eeam && c->cfg.ntypes == 1 &&
      hr.present) {
    if (kn.size() & 1) kn.push_back(kn.back());
    c->knots_dpoly = (int)kn.size();
    const int len = hr.len[0];
    std::vector<R> f3(3 * (size_t)len + 1, (R)0);
    for (int k = 0; k < len - 1; k++) {
      const double* t = hr.table.data();
      const size_t nc = hr.ncols;
      double p0 = t[k * nc], p1 = t[(k + 1) * nc], p2 = t[(k + 2) * nc];
      double dv = p1 - p0, d2v = p2 - 2 * p1 + p0;
      f3[3 * k] = (R)p0;
      f3[3 * k + 1] = (R)(dv - 0.5 * d2v);
      f3[3 * k + 2] = (R)(0.5 * d2v);
    }
    for (size_t k = 0; k + 1 < f3.size(); k += 2) {
      R2 x;
      x.x = f3[k];
      x.y = f3[k + 1];
      kn.push_back(x);
    }
  }
  // FP32, several species, every pair and rho column on one grid: per type
  // pair p = it*nt + jt and segment k a force record of 32 B, (phi p0, phi
  // A, phi B, phi A_lo) of pair column p and (A, D) of rho columns p (density
  // at an it-atom due to a jt-atom) and jt*nt + it, and a density record
  // (rho p0, A, B, 0) of rho column p; segment len-1 zero, as for one
  // species (the table argument is clamped instead of the cutoff tests)
  c->knots_jointn = c->knots_dpolyn = -1;
  c->knots_frho = c->knots_drho = c->knots_pairn = -1;
  const int nt = c->cfg.ntypes;
  bool grid_n = sizeof(R) == 4 && !c->cubic && !c->eeam && nt > 1 &&
                hp.present && hr.present;
  for (int col = 0; grid_n && col < nt * nt; col++)
    grid_n = hp.begin[col] == hp.begin[0] && hp.end[col] == hp.end[0] &&
             hp.invstep[col] == hp.invstep[0] && hp.len[col] == hp.len[0] &&
             hr.begin[col] == hp.begin[0] && hr.end[col] == hp.end[0] &&
             hr.invstep[col] == hp.invstep[0] && hr.len[col] == hp.len[0];
  if (const char* e = getenv("IMDG_JOINT_TABLE"))
    if (atoi(e) == 0) grid_n = false;
  if (grid_n) {
    const int len = hp.len[0];
    // (p0, A, B, D) of segment k of column col, zero for k = len-1
    auto seg = [&](const HostTable& h, int col, int k, double* q) {
      q[0] = q[1] = q[2] = q[3] = 0.0;
      if (k >= len - 1) return;
      const size_t nc = h.ncols;
      double p0 = h.table[k * nc + col], p1 = h.table[(k + 1) * nc + col];
      double p2 = h.table[(k + 2) * nc + col];
      double dv = p1 - p0, d2v = p2 - 2 * p1 + p0;
      q[0] = p0;
      q[1] = dv - 0.5 * d2v;
      q[2] = 0.5 * d2v;
      q[3] = d2v;
    };
    auto push4 = [&](double a, double b, double cc, double d) {
      R2 x, y;
      x.x = (R)a;
      x.y = (R)b;
      y.x = (R)cc;
      y.y = (R)d;
      kn.push_back(x);
      kn.push_back(y);
    };
    if (kn.size() & 1) kn.push_back(kn.back());
    c->knots_jointn = (int)kn.size();
    for (int it = 0; it < nt; it++)
      for (int jt = 0; jt < nt; jt++)
        for (int k = 0; k < len; k++) {
          double ph[4], rc[4], rt[4];
          seg(hp, it * nt + jt, k, ph);
          seg(hr, it * nt + jt, k, rc);
          seg(hr, jt * nt + it, k, rt);
          push4(ph[0], ph[1], ph[2], ph[1] - (double)(R)ph[1]);
          push4(rc[1], rc[3], rt[1], rt[3]);
        }
    c->knots_dpolyn = (int)kn.size();
    for (int p = 0; p < nt * nt; p++)
      for (int k = 0; k < len; k++) {
        double rc[4];
        seg(hr, p, k, rc);
        push4(rc[0], rc[1], rc[2], 0.0);
      }
    // rho column it*nt + jt the same for every it (the density at an atom
    // due to a jt-atom, the usual EAM form): the rho parts of the records
    // above once per type, small enough for LDS, with the same FP32 values
    // (the loops that read them give bit-identical results)
    bool by_jt = true;
    const size_t rows = hr.table.size() / hr.ncols;
    for (int it = 1; by_jt && it < nt; it++)
      for (int jt = 0; by_jt && jt < nt; jt++)
        for (size_t k = 0; by_jt && k < rows; k++)
          by_jt = hr.table[k * hr.ncols + it * nt + jt] ==
                  hr.table[k * hr.ncols + jt];
    if (const char* e = getenv("IMDG_RHO_BY_TYPE"))
      if (atoi(e) == 0) by_jt = false;
    if (by_jt) {
      // the pair halves of the force records again, 16 B apart (the loop
      // that takes rho from LDS reads only them)
      c->knots_pairn = (int)kn.size();
      for (int p = 0; p < nt * nt; p++)
        for (int k = 0; k < len; k++) {
          double ph[4];
          seg(hp, p, k, ph);
          push4(ph[0], ph[1], ph[2], ph[1] - (double)(R)ph[1]);
        }
      c->knots_frho = (int)kn.size();
      for (int t = 0; t < nt; t++)
        for (int k = 0; k < len; k++) {
          double rc[4];
          seg(hr, t, k, rc);
          R2 x;
          x.x = (R)rc[1];
          x.y = (R)rc[3];
          kn.push_back(x);
        }
      c->knots_drho = (int)kn.size();
      std::vector<R> f3;
      for (int t = 0; t < nt; t++)
        for (int k = 0; k < len; k++) {
          double rc[4];
          seg(hr, t, k, rc);
          f3.push_back((R)rc[0]);
          f3.push_back((R)rc[1]);
          f3.push_back((R)rc[2]);
        }
      if (f3.size() & 1) f3.push_back((R)0);
      for (size_t k = 0; k < f3.size(); k += 2) {
        R2 x;
        x.x = f3[k];
        x.y = f3[k + 1];
        kn.push_back(x);
      }
    }
  }
  // the FP32 step kernels evaluate the embedding function and the EEAM
  // energy term (once per atom) in FP64: their knots or cubic segments again
  // in FP64
  if (sizeof(R) == 4) {
    std::vector<double2> k64;
    Tab64 t64[2];
    memset(t64, 0, sizeof t64);
    for (int q = 0; q < 2; q++) {
      HostTable& h = c->tab[2 + q];
      if (!h.present) continue;
      for (int col = 0; col < h.ncols && col < IMDG_MAXTYPES; col++) {
        t64[q].off[col] = (int)k64.size();
        t64[q].len[col] = h.len[col];
        t64[q].begin[col] = h.begin[col];
        t64[q].end[col] = h.end[col];
        t64[q].invstep[col] = h.invstep[col];
        t64[q].twoistep[col] = 2 * h.invstep[col];
        if (c->cubic) {
          std::vector<double> seg;
          if (!cubic_segments(h, col, false, c->cfg.interpolation, seg))
            return fail(c, IMDG_ERR_NOTABLE,
                        "column %d of table %d is too short for the "
                        "interpolation", col, 2 + q);
          for (size_t k = 0; k < seg.size(); k += 2)
            k64.push_back(make_double2(seg[k], seg[k + 1]));
          continue;
        }
        const int nc = h.ncols;
        for (int k = 0; k <= h.len[col]; k++) {
          double p0 = h.table[(size_t)k * nc + col];
          double p1 = h.table[(size_t)(k + 1) * nc + col];
          k64.push_back(make_double2(p0, p1 - p0));
        }
      }
    }
    int rc64 = dalloc(c, &c->d_knots64, k64.size() * sizeof(double2));
    if (rc64) return rc64;
    if (!k64.empty())
      HIPCHK(c, hipMemcpy(c->d_knots64, k64.data(),
                          k64.size() * sizeof(double2), hipMemcpyHostToDevice));
    rc64 = dalloc(c, &c->d_tab64, sizeof t64);
    if (rc64) return rc64;
    HIPCHK(c, hipMemcpy(c->d_tab64, t64, sizeof t64, hipMemcpyHostToDevice));
  }
  int rc = dalloc(c, &c->d_knots, kn.size() * sizeof(R2));
  if (rc) return rc;
  HIPCHK(c, hipMemcpy(c->d_knots, kn.data(), kn.size() * sizeof(R2),
                      hipMemcpyHostToDevice));
  rc = dalloc(c, &c->d_tabs, sizeof td);
  if (rc) return rc;
  HIPCHK(c, hipMemcpy(c->d_tabs, td, sizeof td, hipMemcpyHostToDevice));
  return IMDG_OK;
}

// Launch plan of the persistent step kernels: which knots each neighbour
// loop stages in LDS (all knots of the tables it reads, when they fit next
// to the kernel's static LDS) and the resident grid of each kernel.
template <typename R>
int plan_steps(imdg_ctx* c) {
  typedef typename V2<R>::T R2;
  if (!c->ncu) {
    HIPCHK(c, hipDeviceGetAttribute(&c->ncu,
                                    hipDeviceAttributeMultiprocessorCount,
                                    c->dev));
  }
  bool nt1 = c->cfg.ntypes == 1;
  bool hp = c->tab[0].present, he = c->tab[1].present;
  const bool cub = c->cubic;
  c->force_mode = !nt1 ? (c->knots_jointn >= 0 ? 3 : 2)
                       : (c->knots_joint >= 0 ? 1 : 0);
  int mode = c->force_mode;
  // default shapes (what imdg_autotune picks on the 2M-atom EAM crystal):
  // four lanes per atom and full blocks, except the FP64 force loop, whose
  // registers leave one lane per atom and 512 threads fastest
  const bool fp32 = sizeof(R) == 4;
  if (c->lpa_den <= 0) c->lpa_den = 4;
  if (c->lpa_frc <= 0) c->lpa_frc = fp32 ? 4 : 1;
  if (c->thr_den <= 0) c->thr_den = IMDG_STEP_THREADS;
  if (c->thr_frc <= 0) c->thr_frc = force_threads<R>(mode);
  int td = c->thr_den, tf = c->thr_frc;
  int no_lds = 0;
  if (const char* e = getenv("IMDG_TABLES_IN_LDS")) no_lds = (atoi(e) == 0);
  // cubic segments, EEAM: the generic loops with the tables in HBM/L2
  if (cub || c->eeam) no_lds = 1;
  // density: the rho knots (the several-species segment records stay in
  // HBM/L2: nt^2 columns)
  c->lds_den_lo = c->knots_rho;
  c->lds_den_n = he ? c->knots_embed - c->knots_rho : 0;
  if (c->knots_dpoly >= 0) {  // three floats per segment
    c->lds_den_lo = c->knots_dpoly;
    c->lds_den_n = (3 * c->tab[1].len[0] + 1) / 2;
  }
  const int nty = c->cfg.ntypes;
  if (c->knots_drho >= 0) {  // three floats per segment and type
    c->lds_den_lo = c->knots_drho;
    c->lds_den_n = (3 * nty * c->tab[1].len[0] + 1) / 2;
  }
  c->grid_den = 0;
  if (he && !no_lds && (c->knots_dpolyn < 0 || c->knots_drho >= 0))
    c->grid_den = step_grid_density<R>(nt1, true, c->lds_den_n * sizeof(R2),
                                       c->ncu, c->lpa_den, td, cub, c->eeam);
  if (he && c->grid_den == 0) {
    c->lds_den_n = 0;
    c->grid_den =
        step_grid_density<R>(nt1, false, 0, c->ncu, c->lpa_den, td, cub,
                             c->eeam);
  }
  // force: pair and rho knots (or the interleaved table)
  if (mode == 1) {
    int len = c->tab[0].len[0];
    c->lds_frc_lo = c->knots_joint;
    c->lds_frc_n = sizeof(R) == 8 ? 2 * len + (len + 1) / 2 : 3 * len;
  } else if (mode == 3) {
    c->lds_frc_lo = c->knots_frho;
    c->lds_frc_n = nty * c->tab[1].len[0];
  } else {
    c->lds_frc_lo = hp ? c->knots_pair : c->knots_rho;
    c->lds_frc_n = c->knots_embed - c->lds_frc_lo;
  }
  for (int integ = 0; integ < 2; integ++) {
    int g = 0;
    if (!no_lds && (mode != 3 || c->knots_frho >= 0))
      g = step_grid_force<R>(mode, integ, true, hp, he,
                             c->lds_frc_n * sizeof(R2), c->ncu, c->lpa_frc,
                             tf, cub, c->eeam);
    if (g == 0) {
      c->lds_frc_n = 0;
      g = step_grid_force<R>(mode, integ, false, hp, he, 0, c->ncu,
                             c->lpa_frc, tf, cub, c->eeam);
      if (integ == 1)  // keep both variants on the same table placement
        c->grid_frc[0] = step_grid_force<R>(mode, 0, false, hp, he, 0,
                                            c->ncu, c->lpa_frc, tf, cub,
                                            c->eeam);
    }
    c->grid_frc[integ] = g;
  }
  c->grid_move = step_grid_move<R>(c->ncu);
  if (c->grid_frc[0] == 0 || c->grid_frc[1] == 0 || c->grid_move == 0 ||
      (he && c->grid_den == 0))
    return fail(c, IMDG_ERR_HIP, "cannot size the persistent step kernels");
  int units = std::max(std::max(c->grid_den * (td / IMDG_WAVE),
                                c->grid_move * IMDG_STEP_WAVES),
                       std::max(c->grid_frc[0], c->grid_frc[1]) *
                           (tf / IMDG_WAVE));
  c->units_max = units;
  if (c->units_cap > 0 && units > c->units_cap)
    return fail(c, IMDG_ERR_HIP, "step kernels need %d partial slots, %d "
                "allocated", units, c->units_cap);
  return IMDG_OK;
}

template <typename R>
StepArgs<R> step_args(imdg_ctx* c) {
  typedef typename V4<R>::T R4;
  StepArgs<R> a;
  memset(&a, 0, sizeof a);
  a.rep = c->d_rep;
  a.rst = c->d_rst;
  a.ctl = c->d_ctl;
  a.blk_rep = c->d_blk_rep;
  for (int k = 0; k < 2; k++) {
    a.pos[k] = c->d_pos[k];
    a.mom[k] = c->d_mom[k];
  }
  a.xnbl = c->d_xnbl;
  a.acell = c->d_acell;
  a.atype = c->d_atype;
  a.rec_a = (R4*)c->d_reca;
  a.rec_b = (R4*)c->d_recb;
  a.nbr = c->d_nbr;
  a.nnb = c->d_nnb;
  a.knots = (const typename V2<R>::T*)c->d_knots;
  a.knots64 = (const double2*)c->d_knots64;
  a.cmeta = c->d_cmeta;
  a.list16 = c->list16 ? 1 : 0;
  // list words read per step: slots x the longest list of the last build
  // the host has seen (the capacity before any), at the format's entry size
  const int klist = c->list_k > 0 ? std::min(c->list_k, c->kmax) : c->kmax;
  const size_t list_read =
      (size_t)c->nslots * klist * (c->list16 ? 2 : 4);
  a.list_nt = c->list_nt < 0 ? (list_read > IMDG_LIST_NT_BYTES ? 1 : 0)
                             : c->list_nt;
  a.type_bit = (c->mixed && c->cfg.ntypes == 2 && c->tab[1].present &&
                c->force_mode == 3 && !c->cubic && !c->eeam)
                   ? 1
                   : 0;
  a.tab64 = (const Tab64*)c->d_tab64;
  a.joint = c->knots_joint;
  a.dpoly = c->knots_dpoly;
  a.jointn = c->knots_jointn;
  a.dpolyn = c->knots_dpolyn;
  a.frho = c->knots_frho;
  a.pairn = c->knots_pairn;
  a.drho = c->knots_drho;
  a.cubic = c->cubic ? 1 : 0;
  a.eeam = c->eeam ? 1 : 0;
  a.eeP = (R*)c->d_eeP;
  a.eedm = (R*)c->d_eedm;
  a.final_inline = c->final_inline ? 1 : 0;
  a.lpa_den = c->lpa_den;
  a.lpa_frc = c->lpa_frc;
  a.rho = (R*)c->d_rho;
  a.epot = (R*)c->d_epot;
  a.frc = (R4*)c->d_frc;
  a.partA = c->d_partA;
  a.partB = c->d_partB;
  a.ring = c->d_ring;
  a.nrep = c->nrep;
  a.nchunk = c->nchunk;
  a.part_chunk = c->part_chunk;
  a.unitsA = c->grid_den;  // one partial per block
  a.thr_den = c->thr_den;
  a.thr_frc = c->thr_frc;
  a.kmax = c->kmax;
  a.ntypes = c->cfg.ntypes;
  a.ensemble = c->cfg.ensemble;
  a.dt = c->cfg.timestep;
  a.has_pair = c->tab[0].present;
  a.has_eam = c->tab[1].present;
  a.xcd_remap = c->xcd_remap;
  return a;
}

BuildArgs build_args(imdg_ctx* c) {
  BuildArgs b;
  memset(&b, 0, sizeof b);
  b.rep = c->d_rep;
  b.rst = c->d_rst;
  b.ctl = c->d_ctl;
  b.blk_rep = c->d_blk_rep;
  b.cell_rep = c->d_cell_rep;
  for (int k = 0; k < 2; k++) {
    b.pos[k] = c->d_pos[k];
    b.mom[k] = c->d_mom[k];
    b.aid[k] = c->d_aid[k];
  }
  b.xnbl = c->d_xnbl;
  b.rec_a = c->d_reca;
  b.mixed = c->mixed ? 1 : 0;
  b.cellid = c->d_cellid;
  b.cslot = c->d_cslot;
  b.acell = c->d_acell;
  b.atype = c->d_atype;
  b.perm = c->d_perm;
  b.cell_cnt = c->d_cell_cnt;
  b.cell_beg = c->d_cell_beg;
  b.cell_num = c->d_cell_num;
  b.cell_loc = c->d_cell_loc;
  b.tile_sum = c->d_tile_sum;
  b.tile_off = c->d_tile_off;
  b.tile_rep = c->d_tile_rep;
  b.blkmax = c->d_blkmax;
  b.tiles = c->d_tiles;
  b.nbr = c->d_nbr;
  b.nnb = c->d_nnb;
  b.nrep = c->nrep;
  b.nslots = c->nslots;
  b.nblk = c->nblk;
  b.ncells = c->ncells;
  b.nstile = c->nstile;
  b.kmax = c->kmax;
  b.tile_cap = c->tile_cap;
  b.cmeta = c->d_cmeta;
  b.list16 = c->list16 ? 1 : 0;
  return b;
}

hipEvent_t get_event(imdg_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

struct ProfScope {
  imdg_ctx* c;
  int kc;
  hipEvent_t a = nullptr, b = nullptr;
  bool on;
  ProfScope(imdg_ctx* c_, int kc_)
      : c(c_), kc(kc_),
        on(c_->profiling && ((c_->prof_mask >> kc_) & 1) &&
           (c_->kc_seen[kc_]++ % c_->prof_stride) == 0) {
    if (on) {
      a = get_event(c);
      b = get_event(c);
      hipEventRecord(a, c->stream);
    }
  }
  ~ProfScope() {
    if (on) {
      hipEventRecord(b, c->stream);
      c->pending.push_back({kc, {a, b}});
    }
  }
};

void collect_profile(imdg_ctx* c) {
  for (auto& p : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.second.first, p.second.second) ==
        hipSuccess) {
      c->kc_ms[p.first] += ms;
      c->kc_n[p.first] += 1;
    }
    c->event_pool.push_back(p.second.first);
    c->event_pool.push_back(p.second.second);
  }
  c->pending.clear();
}

int grid_build(imdg_ctx* c) {
  int g = (c->nslots + 255) / 256;
  return std::max(1, std::min(g, 2048));
}

int enqueue_build(imdg_ctx* c) {
  ProfScope ps(c, KC_BUILD);
  launch_build(build_args(c), grid_build(c), c->stream);
  HIPCHK(c, hipGetLastError());
  return IMDG_OK;
}

template <typename R>
hipError_t enqueue_forces_t(imdg_ctx* c, bool integrate, bool store) {
  StepArgs<R> a = step_args<R>(c);
  a.store = store ? 1 : 0;
  const TabDev<R>* t = (const TabDev<R>*)c->d_tabs;
  bool nt1 = c->cfg.ntypes == 1;
  hipError_t e;
  {
    // the force loop gathers (u, 2F') records: from the density loop with
    // EAM, a copy of the (u, type) records without
    ProfScope ps(c, KC_DENSITY);
    if (a.has_eam) {
      a.lds_lo = c->lds_den_lo;
      a.lds_n = c->lds_den_n;
      e = launch_density<R>(a, t, nt1, c->grid_den, c->stream);
    } else {
      e = launch_copyrec<R>(a, c->nslots, c->stream);
    }
    if (e != hipSuccess) return e;
  }
  a.lds_lo = c->lds_frc_lo;
  a.lds_n = c->lds_frc_n;
  int grid = c->grid_frc[integrate ? 1 : 0];
  a.unitsB = grid;
  {
    ProfScope ps(c, KC_FORCE);
    e = launch_force<R>(a, t, c->force_mode, integrate, grid, c->stream);
  }
  if (e == hipSuccess && !c->final_inline) {
    ProfScope ps(c, KC_FINAL);
    e = launch_finalize<R>(a, true, integrate, c->stream);
  }
  return e;
}

template <typename R>
hipError_t enqueue_move_t(imdg_ctx* c) {
  StepArgs<R> a = step_args<R>(c);
  a.unitsB = c->grid_move;
  hipError_t e;
  {
    ProfScope ps(c, KC_MOVE);
    e = launch_move<R>(a, c->grid_move, c->stream);
  }
  if (e == hipSuccess && !c->final_inline) {
    ProfScope ps(c, KC_FINAL);
    e = launch_finalize<R>(a, false, true, c->stream);
  }
  return e;
}

// store: write rho, epot and frc (always without integration: the move
// launch reads frc).
int enqueue_forces(imdg_ctx* c, bool integrate, bool store = true) {
  store = store || !integrate;
  if (c->mixed)
    HIPCHK(c, enqueue_forces_t<float>(c, integrate, store));
  else
    HIPCHK(c, enqueue_forces_t<double>(c, integrate, store));
  HIPCHK(c, hipGetLastError());
  return IMDG_OK;
}

int enqueue_move(imdg_ctx* c) {
  if (c->mixed)
    HIPCHK(c, enqueue_move_t<float>(c));
  else
    HIPCHK(c, enqueue_move_t<double>(c));
  HIPCHK(c, hipGetLastError());
  return IMDG_OK;
}

template <typename R>
hipError_t enqueue_stress_t(imdg_ctx* c) {
  StepArgs<R> a = step_args<R>(c);
  return launch_stress<R>(a, (const TabDev<R>*)c->d_tabs, c->d_stress,
                          c->d_spart, c->nblk, c->stream);
}

// The stress tensor of the step just enqueued (never part of a captured
// step: it runs after the step's last kernel).
int enqueue_stress(imdg_ctx* c) {
  int rc;
  if (!c->d_stress) {
    if ((rc = dalloc(c, (void**)&c->d_stress,
                     sizeof(double) * 6 * (size_t)c->nslots)))
      return rc;
    if ((rc = dalloc(c, (void**)&c->d_spart,
                     sizeof(double) * 6 * (size_t)c->nblk)))
      return rc;
  }
  if (c->mixed)
    HIPCHK(c, enqueue_stress_t<float>(c));
  else
    HIPCHK(c, enqueue_stress_t<double>(c));
  c->stress_valid = true;
  return IMDG_OK;
}

int read_ctl(imdg_ctx* c, Ctl* h) {
  HIPCHK(c, hipMemcpyAsync(h, c->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->list_k = h->kmax_seen;  // the list policy's size estimate
  return IMDG_OK;
}

size_t list_bytes(imdg_ctx* c, int kmax) {
  return (size_t)(c->nslots / IMDG_WAVE) * kmax * IMDG_WAVE * 4;
}

// Grow the list after an overflow and rebuild; *fixed = 1 if it had to.
int fix_overflow(imdg_ctx* c, int* fixed) {
  *fixed = 0;
  Ctl h;
  int rc = read_ctl(c, &h);
  if (rc) return rc;
  if (!h.frozen) return IMDG_OK;
  int newk = std::max(c->kmax + IMDG_LIST_PAD, (int)ceil(h.kmax_seen * 1.2));
  newk = (newk + IMDG_LIST_PAD - 1) / IMDG_LIST_PAD * IMDG_LIST_PAD;
  drop_graph(c);
  dfree(c, c->d_nbr);
  c->d_nbr = nullptr;
  c->kmax = newk;
  size_t bytes = list_bytes(c, newk);
  if ((rc = dalloc(c, (void**)&c->d_nbr, bytes))) return rc;
  c->nbr_bytes = bytes;
  int zero2[2] = {0, 0};  // frozen, overflow
  HIPCHK(c, hipMemcpyAsync(&c->d_ctl->frozen, zero2, 2 * sizeof(int),
                           hipMemcpyHostToDevice, c->stream));
  if ((rc = enqueue_build(c))) return rc;
  *fixed = 1;
  return IMDG_OK;
}

int ensure_ready(imdg_ctx* c) {
  if (!c->atoms_set) return fail(c, IMDG_ERR_STATE, "no atoms set");
  if (!c->d_nbr) {
    size_t bytes = list_bytes(c, c->kmax);
    int rc = dalloc(c, (void**)&c->d_nbr, bytes);
    if (rc) return rc;
    c->nbr_bytes = bytes;
  }
  return IMDG_OK;
}

// Run the build chain until it fits (calc_forces / make_nblist path).
int build_until_fits(imdg_ctx* c) {
  int rc = enqueue_build(c);
  if (rc) return rc;
  for (int tries = 0; tries < 8; tries++) {
    int fixed = 0;
    if ((rc = fix_overflow(c, &fixed))) return rc;
    if (!fixed) return IMDG_OK;
  }
  return fail(c, IMDG_ERR_NOMEM, "neighbour list keeps overflowing");
}

// Device copies of the per-slot arrays needed to map results back.
struct SlotView {
  Ctl ctl;
  std::vector<double4> pos, mom;
  std::vector<int> aid, acell, nnb;
};

int read_slots(imdg_ctx* c, SlotView& v, bool state) {
  int rc = read_ctl(c, &v.ctl);
  if (rc) return rc;
  size_t ns = c->nslots;
  v.aid.resize(ns);
  HIPCHK(c, hipMemcpy(v.aid.data(), c->d_aid[v.ctl.rec_cur], sizeof(int) * ns,
                      hipMemcpyDeviceToHost));
  if (state) {
    v.pos.resize(ns);
    v.mom.resize(ns);
    HIPCHK(c, hipMemcpy(v.pos.data(), c->d_pos[v.ctl.pos_cur],
                        sizeof(double4) * ns, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(v.mom.data(), c->d_mom[v.ctl.pos_cur],
                        sizeof(double4) * ns, hipMemcpyDeviceToHost));
  }
  return IMDG_OK;
}

// Mean time of `reps` launches of one neighbour loop (the density loop, or
// the force loop without integration) on the current state, after one
// untimed launch.  Both only recompute per-step outputs from the positions
// and the list; the force loop's per-atom energies build on the density
// loop's, so the caller finishes with one regular force evaluation.
template <typename R>
int time_loop(imdg_ctx* c, bool density, int reps, double* ms,
              bool integrate = false) {
  StepArgs<R> a = step_args<R>(c);
  const TabDev<R>* t = (const TabDev<R>*)c->d_tabs;
  bool nt1 = c->cfg.ntypes == 1;
  auto launch = [&]() -> hipError_t {
    if (density) {
      a.lds_lo = c->lds_den_lo;
      a.lds_n = c->lds_den_n;
      return launch_density<R>(a, t, nt1, c->grid_den, c->stream);
    }
    a.lds_lo = c->lds_frc_lo;
    a.lds_n = c->lds_frc_n;
    const int grid = c->grid_frc[integrate ? 1 : 0];
    a.unitsB = grid;
    return launch_force<R>(a, t, c->force_mode, integrate, grid, c->stream);
  };
  hipEvent_t e0 = get_event(c), e1 = get_event(c);
  hipError_t e = launch();
  if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
  for (int k = 0; k < reps && e == hipSuccess; k++) e = launch();
  if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t_ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t_ms, e0, e1);
  c->event_pool.push_back(e0);
  c->event_pool.push_back(e1);
  HIPCHK(c, e);
  *ms = t_ms / std::max(1, reps);
  return IMDG_OK;
}

int plan(imdg_ctx* c) {
  return c->mixed ? plan_steps<float>(c) : plan_steps<double>(c);
}

// Try every lanes-per-slot x block-size variant of the density loop, then
// of the force loop, on the current state; keep the fastest of each.
template <typename R>
int autotune_t(imdg_ctx* c, int reps, imdg_tuning* out) {
  static const int lanes[3] = {1, 2, 4};
  const int fmax = force_threads<R>(c->force_mode);
  const int thr_d[3] = {256, 512, 1024};
  const int thr_f[4] = {256, 512, 768, 1024};
  int rc;
  double best_d = 0.0, best_f = 0.0;
  if (c->tab[1].present) {
    int bl = c->lpa_den, bt = c->thr_den;
    best_d = 1e300;
    for (int li = 0; li < 3; li++)
      for (int ti = 0; ti < 3; ti++) {
        c->lpa_den = lanes[li];
        c->thr_den = thr_d[ti];
        if (plan(c) != IMDG_OK) continue;  // not launchable this way
        double ms;
        if ((rc = time_loop<R>(c, true, reps, &ms))) return rc;
        if (ms < best_d) {
          best_d = ms;
          bl = lanes[li];
          bt = thr_d[ti];
        }
      }
    c->lpa_den = bl;
    c->thr_den = bt;
    if ((rc = plan(c))) return rc;
    double ms;
    if ((rc = time_loop<R>(c, true, 0, &ms))) return rc;  // records for B
  }
  {
    // the step's force launch integrates as well (unless the move is a
    // launch of its own): time that variant, on a copy of the state it
    // changes (positions, momenta, (u, type) records), put back afterwards
    const bool integ = !c->split_move;
    const size_t ns = c->nslots, rb = (c->mixed ? 16 : 32) * ns;
    void *sp = nullptr, *sm = nullptr, *sr = nullptr;
    int pc = 0;
    if (integ) {
      Ctl h;
      if ((rc = read_ctl(c, &h))) return rc;
      pc = h.pos_cur;
      if (hipMalloc(&sp, 32 * ns) != hipSuccess ||
          hipMalloc(&sm, 32 * ns) != hipSuccess ||
          hipMalloc(&sr, rb) != hipSuccess) {
        hipFree(sp);
        hipFree(sm);
        hipFree(sr);
        return fail(c, IMDG_ERR_NOMEM, "autotune: no room for a state copy");
      }
      HIPCHK(c, hipMemcpyAsync(sp, c->d_pos[pc], 32 * ns,
                               hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(sm, c->d_mom[pc], 32 * ns,
                               hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(sr, c->d_reca, rb, hipMemcpyDeviceToDevice,
                               c->stream));
    }
    int bl = c->lpa_frc, bt = c->thr_frc;
    best_f = 1e300;
    for (int li = 0; li < 3 && rc == IMDG_OK; li++)
      for (int ti = 0; ti < 4; ti++) {
        if (thr_f[ti] > fmax) continue;
        c->lpa_frc = lanes[li];
        c->thr_frc = thr_f[ti];
        if (plan(c) != IMDG_OK) continue;
        double ms;
        if ((rc = time_loop<R>(c, false, reps, &ms, integ))) break;
        if (ms < best_f) {
          best_f = ms;
          bl = lanes[li];
          bt = thr_f[ti];
        }
      }
    if (integ) {
      hipMemcpyAsync(c->d_pos[pc], sp, 32 * ns, hipMemcpyDeviceToDevice,
                     c->stream);
      hipMemcpyAsync(c->d_mom[pc], sm, 32 * ns, hipMemcpyDeviceToDevice,
                     c->stream);
      hipMemcpyAsync(c->d_reca, sr, rb, hipMemcpyDeviceToDevice, c->stream);
      hipStreamSynchronize(c->stream);
      hipFree(sp);
      hipFree(sm);
      hipFree(sr);
    }
    if (rc) return rc;
    c->lpa_frc = bl;
    c->thr_frc = bt;
  }
  if ((rc = plan(c))) return rc;
  c->err.clear();
  if (out) {
    out->density_lanes = c->lpa_den;
    out->density_threads = c->thr_den;
    out->force_lanes = c->lpa_frc;
    out->force_threads = c->thr_frc;
    out->density_ms = best_d;
    out->force_ms = best_f;
  }
  return IMDG_OK;
}

}  // namespace

extern "C" {

long imdg_struct_size(const char* name) {
  if (!name) return -1;
  static const struct {
    const char* n;
    long size;
  } tab[] = {{"imdg_config", (long)sizeof(imdg_config)},
             {"imdg_table", (long)sizeof(imdg_table)},
             {"imdg_replica", (long)sizeof(imdg_replica)},
             {"imdg_thermo", (long)sizeof(imdg_thermo)},
             {"imdg_info", (long)sizeof(imdg_info)},
             {"imdg_tuning", (long)sizeof(imdg_tuning)}};
  for (const auto& e : tab)
    if (!strcmp(e.n, name)) return e.size;
  return -1;
}

const char* imdg_last_error(const imdg_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  // a context-less call (imdg_check_nblist_noalloc) keeps its message per
  // host thread
  const char* m = imdg::noalloc_message();
  return (m && m[0]) ? m : "null context";
}

int imdg_ctx_create(const imdg_config* cfg, imdg_ctx** out) {
  if (!cfg || !out) return IMDG_ERR_ARG;
  *out = nullptr;
  if (cfg->precision != IMDG_FP64 && cfg->precision != IMDG_MIXED)
    return IMDG_ERR_ARG;
  if (cfg->ensemble != IMDG_NVE && cfg->ensemble != IMDG_NVT)
    return IMDG_ERR_ARG;
  if (cfg->ntypes < 1 || cfg->ntypes > IMDG_MAXTYPES) return IMDG_ERR_ARG;
  if (!(cfg->timestep > 0) || !(cfg->nbl_margin >= 0)) return IMDG_ERR_ARG;
  if (cfg->interpolation < IMDG_INTERP_3POINT ||
      cfg->interpolation > IMDG_INTERP_SPLINE)
    return IMDG_ERR_ARG;
  imdg_ctx* c = new imdg_ctx();
  c->cfg = *cfg;
  c->cubic = cfg->interpolation != IMDG_INTERP_3POINT;
  if (!(c->cfg.nbl_size > 0)) c->cfg.nbl_size = 1.1;
  c->mixed = cfg->precision == IMDG_MIXED;
  const char* remap = getenv("IMDG_XCD_REMAP");
  c->xcd_remap = remap ? atoi(remap) != 0 : 1;
  const char* graph = getenv("IMDG_GRAPH");
  c->use_graph = graph ? atoi(graph) != 0 : false;
  if (const char* e = getenv("IMDG_SPLIT_MOVE")) c->split_move = atoi(e) != 0;
  if (const char* e = getenv("IMDG_TILE_LIMIT"))
    c->tile_cap = std::max(0, std::min(atoi(e), IMDG_TILE_CAP));
  if (const char* e = getenv("IMDG_LIST_NT")) c->list_nt = atoi(e) != 0;
  if (const char* e = getenv("IMDG_FINAL_KERNEL"))
    c->final_inline = atoi(e) == 0;
  if (cfg->lanes_per_atom > 0)
    c->lpa_den = c->lpa_frc = lpa_value(cfg->lanes_per_atom);
  if (const char* e = getenv("IMDG_LPA_DENSITY")) c->lpa_den = lpa_value(atoi(e));
  if (const char* e = getenv("IMDG_LPA_FORCE")) c->lpa_frc = lpa_value(atoi(e));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 ||
      cfg->device < 0 || cfg->device >= ndev) {
    delete c;
    return IMDG_ERR_HIP;
  }
  c->dev = cfg->device;
  if (hipSetDevice(c->dev) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) !=
          hipSuccess) {
    delete c;
    return IMDG_ERR_HIP;
  }
  *out = c;
  return IMDG_OK;
}

int imdg_ctx_destroy(imdg_ctx* c) {
  if (!c) return IMDG_ERR_ARG;
  hipSetDevice(c->dev);
  if (c->stream) hipStreamSynchronize(c->stream);
  collect_profile(c);
  drop_graph(c);
  for (auto e : c->event_pool) hipEventDestroy(e);
  free_all(c);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return IMDG_OK;
}

int imdg_set_tables(imdg_ctx* c, const imdg_table* pair, const imdg_table* rho,
                    const imdg_table* embed) {
  if (!c) return IMDG_ERR_ARG;
  if ((rho == nullptr) != (embed == nullptr))
    return fail(c, IMDG_ERR_NOTABLE, "rho and embed go together");
  if (!pair && !rho) return fail(c, IMDG_ERR_NOTABLE, "no tables");
  if (c->atoms_set) return fail(c, IMDG_ERR_STATE, "set tables before atoms");
  int nt = c->cfg.ntypes, rc;
  if ((rc = copy_table(c, c->tab[0], pair, nt * nt, "pair"))) return rc;
  if ((rc = copy_table(c, c->tab[1], rho, nt * nt, "rho"))) return rc;
  if ((rc = copy_table(c, c->tab[2], embed, nt, "embed"))) return rc;
  // IMD's cellsz: the largest end of the radial tables
  // (src/imd_potential.c:367, 406)
  c->cellsz_tab = 0.0;
  for (int t = 0; t < 2; t++)
    for (int k = 0; k < c->tab[t].ncols; k++)
      c->cellsz_tab = std::max(c->cellsz_tab, c->tab[t].end[k]);
  c->tables_set = true;
  return IMDG_OK;
}

int imdg_set_eeam_table(imdg_ctx* c, const imdg_table* emod) {
  if (!c) return IMDG_ERR_ARG;
  if (c->atoms_set) return fail(c, IMDG_ERR_STATE, "set tables before atoms");
  if (emod && !c->tab[1].present)
    return fail(c, IMDG_ERR_NOTABLE, "EEAM needs the EAM tables");
  return copy_table(c, c->tab[3], emod, c->cfg.ntypes, "eeam");
}

int imdg_set_atoms(imdg_ctx* c, int nrep, const imdg_replica* reps,
                   const int* natoms, const double* x, const double* p,
                   const int* type, const double* mass) {
  if (!c) return IMDG_ERR_ARG;
  if (!c->tables_set) return fail(c, IMDG_ERR_STATE, "set tables first");
  if (nrep < 1 || !reps || !natoms || !x || !type || !mass)
    return fail(c, IMDG_ERR_ARG, "bad set_atoms arguments");
  hipSetDevice(c->dev);
  drop_graph(c);
  free_all(c);
  c->atoms_set = false;
  c->nrep = nrep;
  c->rep.assign(nrep, RepDev());
  c->natoms.assign(natoms, natoms + nrep);
  c->in_off.assign(nrep, 0);
  long tot = 0;
  int slot = 0, cell = 0, stile = 0;
  double kest = 0.0;
  for (int r = 0; r < nrep; r++) {
    RepDev& R = c->rep[r];
    memset(&R, 0, sizeof R);
    if (natoms[r] < 0) return fail(c, IMDG_ERR_ARG, "negative atom count");
    const double* bx[3] = {reps[r].box_x, reps[r].box_y, reps[r].box_z};
    for (int d = 0; d < 3; d++)
      for (int e = 0; e < 3; e++) R.box[d][e] = bx[d][e];
    // make_box (src/imd_geom_3d.c:52-104)
    vec_prod(R.box[1], R.box[2], R.tbox[0]);
    vec_prod(R.box[2], R.box[0], R.tbox[1]);
    vec_prod(R.box[0], R.box[1], R.tbox[2]);
    double vol = sprod(R.box[0], R.tbox[0]);
    if (vol == 0) return fail(c, IMDG_ERR_GEOM, "replica %d: flat box", r);
    double height[3];
    for (int d = 0; d < 3; d++)
      for (int e = 0; e < 3; e++) R.tbox[d][e] /= vol;
    for (int d = 0; d < 3; d++) height[d] = 1.0 / sprod(R.tbox[d], R.tbox[d]);
    // init_cells: skin added once, grid (src/imd_geom_3d.c:122-144)
    double cs = sqrt(c->cellsz_tab) + c->cfg.nbl_margin;
    R.cellsz = cs * cs;
    for (int d = 0; d < 3; d++) {
      R.pbc[d] = reps[r].pbc[d] ? 1 : 0;
      double scale = sqrt(R.cellsz / height[d]);
      R.gdim[d] = (int)(1.0 / scale);
      if (R.gdim[d] < 1 || (R.pbc[d] && R.gdim[d] < 3))
        return fail(c, IMDG_ERR_GEOM,
                    "replica %d: %d cells along axis %d (periodic axes need "
                    "at least 3 cells of the list radius)",
                    r, R.gdim[d], d);
      if (R.gdim[d] > 4096)
        return fail(c, IMDG_ERR_GEOM, "replica %d: %d cells along axis %d", r,
                    R.gdim[d], d);
    }
    double h = 0.5 * c->cfg.nbl_margin;
    R.skin2 = h * h;
    R.temperature = reps[r].temperature;
    R.isq_tau_eta = reps[r].isq_tau_eta;
    for (int d = 0; d < 3; d++)
      for (int e = 0; e < 3; e++) R.cvec[d][e] = R.box[d][e] / R.gdim[d];
    for (int ox = -1; ox <= 1; ox++)
      for (int oy = -1; oy <= 1; oy++)
        for (int oz = -1; oz <= 1; oz++) {
          int code = (ox + 1) * 9 + (oy + 1) * 3 + (oz + 1);
          for (int e = 0; e < 3; e++)
            R.offs[code][e] =
                ox * R.cvec[0][e] + oy * R.cvec[1][e] + oz * R.cvec[2][e];
        }
    for (int e = 0; e < 3; e++) R.offs[IMDG_CODE_PAD][e] = IMDG_PAD_SHIFT;
    R.atom_off = slot;
    R.natoms = natoms[r];
    R.nslots = ((natoms[r] + IMDG_BLOCK - 1) / IMDG_BLOCK) * IMDG_BLOCK;
    if (R.nslots == 0) R.nslots = IMDG_BLOCK;
    R.cell_off = cell;
    R.ncells = R.gdim[0] * R.gdim[1] * R.gdim[2];
    R.blk_off = slot / IMDG_BLOCK;
    R.nblk = R.nslots / IMDG_BLOCK;
    R.stile_off = stile;
    R.nstile = (R.ncells + IMDG_SCAN_TILE - 1) / IMDG_SCAN_TILE;
    stile += R.nstile;
    R.nactive = 3 * natoms[r];
    c->in_off[r] = tot;
    tot += natoms[r];
    slot += R.nslots;
    cell += R.ncells;
    double dens = natoms[r] / fabs(vol);
    double rl = sqrt(R.cellsz);
    kest = std::max(kest, dens * 4.18879020478639 * rl * rl * rl);
  }
  if (slot >= (1 << IMDG_IDX_BITS))
    return fail(c, IMDG_ERR_ARG, "too many atoms for 27-bit list indices");
  c->natoms_tot = tot;
  c->nslots = slot;
  c->nblk = slot / IMDG_BLOCK;
  c->ncells = cell;
  c->nstile = stile;
  c->kmax = std::max(16, (int)ceil(kest * c->cfg.nbl_size * 1.25));
  c->kmax = (c->kmax + IMDG_LIST_PAD - 1) / IMDG_LIST_PAD * IMDG_LIST_PAD;

  int rc;
#define ALLOC(ptr, bytes)                                        \
  if ((rc = dalloc(c, (void**)&(ptr), (bytes)))) return rc
  size_t ns = c->nslots, rs = c->mixed ? sizeof(float) : sizeof(double);
  ALLOC(c->d_rep, sizeof(RepDev) * nrep);
  ALLOC(c->d_rst, sizeof(RepState) * nrep);
  ALLOC(c->d_ctl, sizeof(Ctl));
  ALLOC(c->d_blk_rep, sizeof(int) * c->nblk);
  ALLOC(c->d_cell_rep, sizeof(int) * c->ncells);
  for (int k = 0; k < 2; k++) {
    ALLOC(c->d_pos[k], sizeof(double4) * ns);
    ALLOC(c->d_mom[k], sizeof(double4) * ns);
    ALLOC(c->d_aid[k], sizeof(int) * ns);
  }
  ALLOC(c->d_xnbl, sizeof(double4) * ns);
  ALLOC(c->d_cellid, sizeof(int) * ns);
  ALLOC(c->d_cslot, sizeof(int) * ns);
  ALLOC(c->d_acell, sizeof(int) * ns);
  ALLOC(c->d_atype, ns);
  ALLOC(c->d_perm, sizeof(int) * ns);
  ALLOC(c->d_cell_cnt, sizeof(int) * c->ncells);
  ALLOC(c->d_cell_beg, sizeof(int) * c->ncells);
  ALLOC(c->d_cell_num, sizeof(int) * c->ncells);
  ALLOC(c->d_cell_loc, sizeof(int) * c->ncells);
  ALLOC(c->d_tile_sum, sizeof(int) * c->nstile);
  ALLOC(c->d_tile_off, sizeof(int) * c->nstile);
  ALLOC(c->d_tile_rep, sizeof(int) * c->nstile);
  ALLOC(c->d_blkmax, sizeof(int) * c->nblk);
  ALLOC(c->d_tiles, sizeof(Tile) * c->nblk);
  ALLOC(c->d_nnb, sizeof(int) * ns);
  ALLOC(c->d_reca, rs * 4 * ns);
  ALLOC(c->d_recb, rs * 4 * ns);
  ALLOC(c->d_rho, rs * ns);
  ALLOC(c->d_epot, rs * ns);
  c->eeam = c->tab[3].present;
  c->d_eeP = c->d_eedm = nullptr;
  if (c->eeam) {
    ALLOC(c->d_eeP, rs * ns);
    ALLOC(c->d_eedm, rs * ns);
  }
  ALLOC(c->d_frc, rs * 4 * ns);
  ALLOC(c->d_maxd, sizeof(double) * c->nblk);
  ALLOC(c->d_ring, sizeof(ThermoRec) * IMDG_RING * nrep);
  if ((rc = c->mixed ? upload_tables<float>(c) : upload_tables<double>(c)))
    return rc;
  c->nchunk = c->nslots / IMDG_CHUNK;
  c->part_chunk = nrep > 1 ? 1 : 0;
  c->units_cap = 0;  // partials are allocated below, for any plan
  if ((rc = c->mixed ? plan_steps<float>(c) : plan_steps<double>(c)))
    return rc;
  // the 16-bit list format where only the one-grid FP32 loops read the
  // list: their walks (and the stress kernel's) decode it
  c->list16 = c->mixed && c->tab[1].present && !c->cubic && !c->eeam &&
              (c->cfg.ntypes == 1
                   ? (c->force_mode == 1 && c->knots_dpoly >= 0)
                   : (c->force_mode == 3 && c->knots_dpolyn >= 0));
  if (const char* e = getenv("IMDG_LIST16"))
    if (atoi(e) == 0) c->list16 = false;
  c->d_cmeta = nullptr;
  if (c->list16) ALLOC(c->d_cmeta, sizeof(int) * IMDG_META * c->nchunk);
  {
    // room for any launch plan: at most 32 resident wavefronts per CU
    c->units_cap = std::max(c->units_max, 32 * c->ncu);
    size_t units = std::max(c->nchunk, c->units_cap);
    ALLOC(c->d_partA, sizeof(double) * units);
    ALLOC(c->d_partB, sizeof(double) * PB_N * units);
  }
#undef ALLOC

  // host staging in slot order
  std::vector<double4> hpos(ns), hmom(ns);
  std::vector<int> haid(ns, -1), hblk(c->nblk), hcrep(c->ncells),
      htrep(c->nstile);
  for (size_t k = 0; k < ns; k++) {
    hpos[k] = make_double4(0, 0, 0, -1.0);
    hmom[k] = make_double4(0, 0, 0, 1.0);
  }
  for (int r = 0; r < nrep; r++) {
    RepDev& R = c->rep[r];
    for (int b = 0; b < R.nblk; b++) hblk[R.blk_off + b] = r;
    for (int q = 0; q < R.ncells; q++) hcrep[R.cell_off + q] = r;
    for (int q = 0; q < R.nstile; q++) htrep[R.stile_off + q] = r;
    for (int k = 0; k < R.natoms; k++) {
      long in = c->in_off[r] + k;
      int t = type[in];
      if (t < 0 || t >= c->cfg.ntypes)
        return fail(c, IMDG_ERR_ARG, "atom %ld has type %d", in, t);
      if (!(mass[in] > 0))
        return fail(c, IMDG_ERR_ARG, "atom %ld has mass %g", in, mass[in]);
      double q[3] = {x[3 * in], x[3 * in + 1], x[3 * in + 2]};
      // back_into_box (src/imd_geom_3d.c:1083-1110)
      for (int d = 0; d < 3; d++) {
        if (R.pbc[d] != 1) continue;
        double i = floor(q[0] * R.tbox[d][0] + q[1] * R.tbox[d][1] +
                         q[2] * R.tbox[d][2]);
        q[0] -= i * R.box[d][0];
        q[1] -= i * R.box[d][1];
        q[2] -= i * R.box[d][2];
      }
      int s = R.atom_off + k;
      hpos[s] = make_double4(q[0], q[1], q[2], (double)t);
      hmom[s] = make_double4(p ? p[3 * in] : 0.0, p ? p[3 * in + 1] : 0.0,
                             p ? p[3 * in + 2] : 0.0, mass[in]);
      haid[s] = (int)in;
    }
  }
  std::vector<RepState> hrst(nrep);
  for (int r = 0; r < nrep; r++) {
    memset(&hrst[r], 0, sizeof(RepState));
    hrst[r].eta = reps[r].eta;
    hrst[r].need_rebuild = 1;
  }
  Ctl hc;
  memset(&hc, 0, sizeof hc);
  hc.any_rebuild = 1;
  hipStream_t st = c->stream;
  HIPCHK(c, hipMemcpyAsync(c->d_rep, c->rep.data(), sizeof(RepDev) * nrep,
                           hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->d_rst, hrst.data(), sizeof(RepState) * nrep,
                           hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->d_ctl, &hc, sizeof hc, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->d_blk_rep, hblk.data(), sizeof(int) * c->nblk,
                           hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->d_cell_rep, hcrep.data(),
                           sizeof(int) * c->ncells, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->d_tile_rep, htrep.data(),
                           sizeof(int) * c->nstile, hipMemcpyHostToDevice, st));
  for (int k = 0; k < 2; k++) {
    HIPCHK(c, hipMemcpyAsync(c->d_pos[k], hpos.data(), sizeof(double4) * ns,
                             hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(c->d_mom[k], hmom.data(), sizeof(double4) * ns,
                             hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(c->d_aid[k], haid.data(), sizeof(int) * ns,
                             hipMemcpyHostToDevice, st));
  }
  HIPCHK(c, hipMemcpyAsync(c->d_xnbl, hpos.data(), sizeof(double4) * ns,
                           hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemsetAsync(c->d_cell_cnt, 0, sizeof(int) * c->ncells, st));
  HIPCHK(c, hipMemsetAsync(c->d_cell_beg, 0, sizeof(int) * c->ncells, st));
  HIPCHK(c, hipMemsetAsync(c->d_cell_num, 0, sizeof(int) * c->ncells, st));
  HIPCHK(c, hipMemsetAsync(c->d_tiles, 0, sizeof(Tile) * c->nblk, st));
  HIPCHK(c, hipMemsetAsync(c->d_acell, 0, sizeof(int) * ns, st));
  HIPCHK(c, hipMemsetAsync(c->d_atype, 0, ns, st));
  HIPCHK(c, hipMemsetAsync(c->d_nnb, 0, sizeof(int) * ns, st));
  HIPCHK(c, hipMemsetAsync(c->d_perm, 0, sizeof(int) * ns, st));
  HIPCHK(c, hipMemsetAsync(c->d_reca, 0, rs * 4 * ns, st));
  HIPCHK(c, hipMemsetAsync(c->d_recb, 0, rs * 4 * ns, st));
  HIPCHK(c, hipMemsetAsync(c->d_rho, 0, rs * ns, st));
  HIPCHK(c, hipMemsetAsync(c->d_epot, 0, rs * ns, st));
  if (c->eeam) {
    HIPCHK(c, hipMemsetAsync(c->d_eeP, 0, rs * ns, st));
    HIPCHK(c, hipMemsetAsync(c->d_eedm, 0, rs * ns, st));
  }
  HIPCHK(c, hipMemsetAsync(c->d_frc, 0, rs * 4 * ns, st));
  HIPCHK(c, hipMemsetAsync(c->d_ring, 0, sizeof(ThermoRec) * IMDG_RING * nrep,
                           st));
  HIPCHK(c, hipStreamSynchronize(st));
  c->atoms_set = true;
  c->target_steps = 0;
  return ensure_ready(c);
}

int imdg_set_nvt(imdg_ctx* c, int rep, double temperature, double eta,
                 double isq_tau_eta) {
  if (!c || !c->atoms_set || rep < 0 || rep >= c->nrep) return IMDG_ERR_ARG;
  c->rep[rep].temperature = temperature;
  c->rep[rep].isq_tau_eta = isq_tau_eta;
  HIPCHK(c, hipMemcpyAsync(&c->d_rep[rep], &c->rep[rep], sizeof(RepDev),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(&c->d_rst[rep].eta, &eta, sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return IMDG_OK;
}

int imdg_calc_forces(imdg_ctx* c, int steps) {
  (void)steps;
  if (!c) return IMDG_ERR_ARG;
  int rc = ensure_ready(c);
  if (rc) return rc;
  c->stress_valid = false;
  if ((rc = build_until_fits(c))) return rc;
  if ((rc = enqueue_forces(c, false))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_profile(c);
  return IMDG_OK;
}

int imdg_move_atoms(imdg_ctx* c) {
  if (!c) return IMDG_ERR_ARG;
  int rc = ensure_ready(c);
  if (rc) return rc;
  if ((rc = enqueue_move(c))) return rc;
  if (c->stress_on && (rc = enqueue_stress(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_profile(c);
  c->target_steps += 1;
  return IMDG_OK;
}

int imdg_check_nblist(imdg_ctx* c, int* invalid) {
  if (!c) return IMDG_ERR_ARG;
  int rc = ensure_ready(c);
  if (rc) return rc;
  launch_check(c->d_rep, c->d_ctl, c->d_blk_rep, c->d_pos[0], c->d_pos[1],
               c->d_xnbl, c->d_maxd, c->nblk, c->stream);
  HIPCHK(c, hipGetLastError());
  std::vector<double> maxd(c->nblk);
  std::vector<RepState> st(c->nrep);
  HIPCHK(c, hipMemcpyAsync(maxd.data(), c->d_maxd, sizeof(double) * c->nblk,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(st.data(), c->d_rst, sizeof(RepState) * c->nrep,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int any = 0;
  for (int r = 0; r < c->nrep; r++) {
    double m = 0.0;
    for (int b = 0; b < c->rep[r].nblk; b++)
      m = std::max(m, maxd[c->rep[r].blk_off + b]);
    st[r].need_rebuild = (m > c->rep[r].skin2) ? 1 : st[r].need_rebuild;
    any |= st[r].need_rebuild;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_rst, st.data(), sizeof(RepState) * c->nrep,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(&c->d_ctl->any_rebuild, &any, sizeof(int),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (invalid) *invalid = any;
  return IMDG_OK;
}

static int flag_all_for_rebuild(imdg_ctx* c) {
  std::vector<RepState> st(c->nrep);
  HIPCHK(c, hipMemcpyAsync(st.data(), c->d_rst, sizeof(RepState) * c->nrep,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& s : st) s.need_rebuild = 1;
  int one = 1;
  HIPCHK(c, hipMemcpyAsync(c->d_rst, st.data(), sizeof(RepState) * c->nrep,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(&c->d_ctl->any_rebuild, &one, sizeof(int),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return IMDG_OK;
}

int imdg_make_nblist(imdg_ctx* c) {
  if (!c) return IMDG_ERR_ARG;
  int rc = ensure_ready(c);
  if (rc) return rc;
  c->stress_valid = false;  // the rebuild renumbers the slots
  if ((rc = flag_all_for_rebuild(c))) return rc;
  if ((rc = build_until_fits(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_profile(c);
  return IMDG_OK;
}

int imdg_deallocate_nblist(imdg_ctx* c) {
  if (!c) return IMDG_ERR_ARG;
  if (!c->atoms_set) return IMDG_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  dfree(c, c->d_nbr);
  c->d_nbr = nullptr;
  c->nbr_bytes = 0;
  return flag_all_for_rebuild(c);
}

// Capture one fused step into a graph: the launches only, no memory
// operations (nothing in the step clears memory with a memset).
static int capture_step(imdg_ctx* c) {
  hipGraph_t g = nullptr;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_build(c);
  if (!rc) rc = enqueue_forces(c, !c->split_move);
  if (!rc && c->split_move) rc = enqueue_move(c);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIPCHK(c, e);
  e = hipGraphInstantiate(&c->step_graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHK(c, e);
  return IMDG_OK;
}

int imdg_run(imdg_ctx* c, long nsteps) {
  if (!c || nsteps < 0) return IMDG_ERR_ARG;
  int rc = ensure_ready(c);
  if (rc) return rc;
  if (c->use_graph && !c->profiling && nsteps > 0) {
    if (!c->step_graph && (rc = capture_step(c))) return rc;
    for (long s = 0; s < nsteps; s++)
      HIPCHK(c, hipGraphLaunch(c->step_graph, c->stream));
  } else {
    for (long s = 0; s < nsteps; s++) {
      if ((rc = enqueue_build(c))) return rc;
      // per-atom outputs only on the last step: nothing reads them between
      if ((rc = enqueue_forces(c, !c->split_move, s == nsteps - 1))) return rc;
      if (c->split_move && (rc = enqueue_move(c))) return rc;
    }
  }
  if (nsteps > 0) {
    c->stress_valid = false;
    if (c->stress_on && (rc = enqueue_stress(c))) return rc;
  }
  c->target_steps += nsteps;
  return IMDG_OK;
}

int imdg_get_eeam(imdg_ctx* c, double* P, double* dm) {
  if (!c || !c->atoms_set) return IMDG_ERR_ARG;
  if (!c->eeam) return fail(c, IMDG_ERR_STATE, "no EEAM table set");
  hipSetDevice(c->dev);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  SlotView v;
  int rc = read_slots(c, v, false);
  if (rc) return rc;
  size_t ns = c->nslots;
  std::vector<double> hp(ns), hd(ns);
  if (c->mixed) {
    std::vector<float> t(ns);
    HIPCHK(c, hipMemcpy(t.data(), c->d_eeP, ns * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ns; k++) hp[k] = t[k];
    HIPCHK(c, hipMemcpy(t.data(), c->d_eedm, ns * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ns; k++) hd[k] = t[k];
  } else {
    HIPCHK(c, hipMemcpy(hp.data(), c->d_eeP, ns * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hd.data(), c->d_eedm, ns * 8, hipMemcpyDeviceToHost));
  }
  for (int r = 0; r < c->nrep; r++) {
    const RepDev& R = c->rep[r];
    for (int k = 0; k < R.natoms; k++) {
      int s = R.atom_off + k;
      long in = v.aid[s];
      if (in < 0 || in >= c->natoms_tot)
        return fail(c, IMDG_ERR_STATE, "slot %d has no atom", s);
      if (P) P[in] = hp[s];
      if (dm) dm[in] = hd[s];
    }
  }
  return IMDG_OK;
}

int imdg_set_stress(imdg_ctx* c, int enable) {
  if (!c) return IMDG_ERR_ARG;
  c->stress_on = enable != 0;
  if (!c->stress_on) c->stress_valid = false;
  return IMDG_OK;
}

int imdg_get_stress(imdg_ctx* c, double* per_atom, double* total) {
  if (!c || !c->atoms_set) return IMDG_ERR_ARG;
  hipSetDevice(c->dev);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!c->stress_valid)
    return fail(c, IMDG_ERR_STATE,
                "no stress tensor of the last step (imdg_set_stress, then "
                "imdg_run or imdg_move_atoms)");
  SlotView v;
  int rc = read_slots(c, v, false);
  if (rc) return rc;
  if (v.ctl.frozen)
    return fail(c, IMDG_ERR_STATE, "list overflow pending (imdg_sync first)");
  if (total) {
    std::vector<double> sp((size_t)6 * c->nblk);
    HIPCHK(c, hipMemcpy(sp.data(), c->d_spart, sizeof(double) * sp.size(),
                        hipMemcpyDeviceToHost));
    for (int r = 0; r < c->nrep; r++) {
      const RepDev& R = c->rep[r];
      for (int q = 0; q < 6; q++) {
        double t = 0.0;
        for (int b = 0; b < R.nblk; b++) t += sp[(size_t)(R.blk_off + b) * 6 + q];
        total[(size_t)6 * r + q] = t;
      }
    }
  }
  if (per_atom) {
    std::vector<double> st((size_t)6 * c->nslots);
    HIPCHK(c, hipMemcpy(st.data(), c->d_stress, sizeof(double) * st.size(),
                        hipMemcpyDeviceToHost));
    for (int r = 0; r < c->nrep; r++) {
      const RepDev& R = c->rep[r];
      for (int k = 0; k < R.natoms; k++) {
        int s = R.atom_off + k;
        long in = v.aid[s];
        if (in < 0 || in >= c->natoms_tot)
          return fail(c, IMDG_ERR_STATE, "slot %d has no atom", s);
        for (int q = 0; q < 6; q++)
          per_atom[(size_t)6 * in + q] = st[(size_t)6 * s + q];
      }
    }
  }
  return IMDG_OK;
}

int imdg_sync(imdg_ctx* c) {
  if (!c) return IMDG_ERR_ARG;
  hipSetDevice(c->dev);
  for (int guard = 0; guard < 64; guard++) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    collect_profile(c);
    Ctl h;
    int rc = read_ctl(c, &h);
    if (rc) return rc;
    if (!h.frozen) {
      if (h.steps_done != c->target_steps)
        return fail(c, IMDG_ERR_STATE, "device did %ld steps, expected %ld",
                    h.steps_done, c->target_steps);
      return IMDG_OK;
    }
    int fixed = 0;
    if ((rc = fix_overflow(c, &fixed))) return rc;
    long missing = c->target_steps - h.steps_done;
    c->target_steps = h.steps_done;
    if ((rc = imdg_run(c, missing))) return rc;
  }
  return fail(c, IMDG_ERR_NOMEM, "neighbour list keeps overflowing");
}

int imdg_get_atoms(imdg_ctx* c, double* x, double* p, double* f, double* epot,
                   double* rho, double* df) {
  if (!c || !c->atoms_set) return IMDG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  SlotView v;
  int rc = read_slots(c, v, true);
  if (rc) return rc;
  size_t ns = c->nslots;
  std::vector<double> hf(4 * ns), he(ns), hr(ns), hb(4 * ns);
  if (c->mixed) {
    std::vector<float> t4(4 * ns), t(ns);
    HIPCHK(c, hipMemcpy(t4.data(), c->d_frc, 4 * ns * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < 4 * ns; k++) hf[k] = t4[k];
    HIPCHK(c, hipMemcpy(t4.data(), c->d_recb, 4 * ns * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < 4 * ns; k++) hb[k] = t4[k];
    HIPCHK(c, hipMemcpy(t.data(), c->d_epot, ns * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ns; k++) he[k] = t[k];
    HIPCHK(c, hipMemcpy(t.data(), c->d_rho, ns * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ns; k++) hr[k] = t[k];
  } else {
    HIPCHK(c, hipMemcpy(hf.data(), c->d_frc, 4 * ns * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hb.data(), c->d_recb, 4 * ns * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(he.data(), c->d_epot, ns * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hr.data(), c->d_rho, ns * 8, hipMemcpyDeviceToHost));
  }
  bool eam = c->tab[1].present;
  for (int r = 0; r < c->nrep; r++) {
    const RepDev& R = c->rep[r];
    for (int k = 0; k < R.natoms; k++) {
      int s = R.atom_off + k;
      long in = v.aid[s];
      if (in < 0 || in >= c->natoms_tot)
        return fail(c, IMDG_ERR_STATE, "slot %d has no atom", s);
      if (x) {
        x[3 * in] = v.pos[s].x;
        x[3 * in + 1] = v.pos[s].y;
        x[3 * in + 2] = v.pos[s].z;
      }
      if (p) {
        p[3 * in] = v.mom[s].x;
        p[3 * in + 1] = v.mom[s].y;
        p[3 * in + 2] = v.mom[s].z;
      }
      if (f) {
        f[3 * in] = hf[4 * s];
        f[3 * in + 1] = hf[4 * s + 1];
        f[3 * in + 2] = hf[4 * s + 2];
      }
      if (epot) epot[in] = he[s];
      if (rho) rho[in] = eam ? hr[s] : 0.0;
      if (df) df[in] = eam ? hb[4 * s + 3] : 0.0;
    }
  }
  return IMDG_OK;
}

int imdg_get_thermo(imdg_ctx* c, int rep, imdg_thermo* out) {
  if (!c || !c->atoms_set || rep < 0 || rep >= c->nrep || !out)
    return IMDG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Ctl h;
  int rc = read_ctl(c, &h);
  if (rc) return rc;
  RepState st;
  HIPCHK(c, hipMemcpy(&st, &c->d_rst[rep], sizeof st, hipMemcpyDeviceToHost));
  memset(out, 0, sizeof *out);
  out->step = h.steps_done;
  out->epot = st.epot;
  out->virial = st.virial;
  out->eta = st.eta;
  out->natoms = c->rep[rep].natoms;
  if (h.steps_done > 0) {
    ThermoRec tr;
    long slot = (h.steps_done - 1) % IMDG_RING;
    HIPCHK(c, hipMemcpy(&tr, &c->d_ring[slot * c->nrep + rep], sizeof tr,
                        hipMemcpyDeviceToHost));
    out->ekin = tr.ekin;
    out->max_disp2 = tr.max_disp2;
    out->rebuilt = tr.rebuilt;
  }
  return IMDG_OK;
}

int imdg_get_thermo_history(imdg_ctx* c, long first, long count,
                            imdg_thermo* out) {
  if (!c || !c->atoms_set || first < 0 || count < 0 || !out)
    return IMDG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Ctl h;
  int rc = read_ctl(c, &h);
  if (rc) return rc;
  if (first + count > h.steps_done || first < h.steps_done - IMDG_RING)
    return fail(c, IMDG_ERR_ARG, "steps [%ld,%ld) not in the ring (done %ld)",
                first, first + count, h.steps_done);
  std::vector<ThermoRec> ring((size_t)IMDG_RING * c->nrep);
  HIPCHK(c, hipMemcpy(ring.data(), c->d_ring, sizeof(ThermoRec) * ring.size(),
                      hipMemcpyDeviceToHost));
  for (long s = 0; s < count; s++)
    for (int r = 0; r < c->nrep; r++) {
      const ThermoRec& t = ring[((first + s) % IMDG_RING) * c->nrep + r];
      imdg_thermo& o = out[s * c->nrep + r];
      o.step = t.step;
      o.epot = t.epot;
      o.ekin = t.ekin;
      o.virial = t.virial;
      o.eta = t.eta;
      o.max_disp2 = t.max_disp2;
      o.rebuilt = t.rebuilt;
      o.natoms = t.natoms;
    }
  return IMDG_OK;
}

int imdg_get_cells(imdg_ctx* c, int* cell3) {
  if (!c || !c->atoms_set || !cell3) return IMDG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  SlotView v;
  int rc = read_slots(c, v, false);
  if (rc) return rc;
  size_t ns = c->nslots;
  std::vector<int> hc(ns);
  HIPCHK(c, hipMemcpy(hc.data(), c->d_acell, sizeof(int) * ns,
                      hipMemcpyDeviceToHost));
  for (int r = 0; r < c->nrep; r++) {
    const RepDev& R = c->rep[r];
    for (int k = 0; k < R.natoms; k++) {
      int s = R.atom_off + k;
      int lc = hc[s] - R.cell_off;
      long in = v.aid[s];
      cell3[3 * in + 2] = lc % R.gdim[2];
      cell3[3 * in + 1] = (lc / R.gdim[2]) % R.gdim[1];
      cell3[3 * in + 0] = lc / (R.gdim[1] * R.gdim[2]);
    }
  }
  return IMDG_OK;
}

int imdg_get_pairs(imdg_ctx* c, long* out, long cap, long* npairs) {
  if (!c || !c->atoms_set) return IMDG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!c->d_nbr) return fail(c, IMDG_ERR_STATE, "no list");
  SlotView v;
  int rc = read_slots(c, v, false);
  if (rc) return rc;
  size_t ns = c->nslots;
  std::vector<int> hn(ns), hc(ns);
  std::vector<uint32_t> hl(c->nbr_bytes / 4);
  HIPCHK(c, hipMemcpy(hn.data(), c->d_nnb, sizeof(int) * ns,
                      hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(hc.data(), c->d_acell, sizeof(int) * ns,
                      hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(hl.data(), c->d_nbr, c->nbr_bytes, hipMemcpyDeviceToHost));
  std::vector<int> meta;
  if (c->list16) {
    meta.resize((size_t)IMDG_META * c->nchunk);
    HIPCHK(c, hipMemcpy(meta.data(), c->d_cmeta, sizeof(int) * meta.size(),
                        hipMemcpyDeviceToHost));
  }
  const uint16_t* hl16 = reinterpret_cast<const uint16_t*>(hl.data());
  long np = 0;
  for (int r = 0; r < c->nrep; r++) {
    const RepDev& R = c->rep[r];
    for (int k = 0; k < R.natoms; k++) {
      int s = R.atom_off + k;
      long ii = v.aid[s] - c->in_off[r];
      int lc = hc[s] - R.cell_off;
      int ci[3] = {lc / (R.gdim[1] * R.gdim[2]), (lc / R.gdim[2]) % R.gdim[1],
                   lc % R.gdim[2]};
      const int* cm = c->list16 ? &meta[(size_t)(s >> 6) * IMDG_META] : nullptr;
      const bool w16 = cm && cm[IMDG_META_WIDE] == 0;
      for (int m = 0; m < hn[s]; m++) {
        int idx, code;
        if (w16) {
          size_t at = (size_t)(s >> 6) * c->kmax * IMDG_WAVE * 2 +
                      ((size_t)(m >> 2) * IMDG_WAVE + (s & 63)) * IMDG_GROUP +
                      (m & 3);
          uint32_t e = hl16[at];
          code = (int)(e >> IMDG_K16_BITS);
          idx = cm[code] + (int)(e & IMDG_K16_MASK);
        } else {
          size_t at = (((size_t)(s >> 6) * (c->kmax >> 2) + (m >> 2)) *
                           IMDG_WAVE + (s & 63)) * IMDG_GROUP + (m & 3);
          uint32_t e = hl[at];
          idx = (int)(e & IMDG_IDX_MASK);
          code = (int)(e >> IMDG_IDX_BITS);
        }
        if (code == IMDG_CODE_PAD) continue;
        int j = idx;
        long jj = v.aid[j] - c->in_off[r];
        int o[3] = {code / 9 - 1, (code / 3) % 3 - 1, code % 3 - 1}, sh[3];
        for (int d = 0; d < 3; d++) {
          int n = ci[d] + o[d];
          sh[d] = n < 0 ? -1 : (n >= R.gdim[d] ? 1 : 0);
        }
        if (ii > jj) continue;  // each unordered pair once
        if (ii == jj &&
            (sh[0] < 0 || (sh[0] == 0 && (sh[1] < 0 || (sh[1] == 0 && sh[2] < 0)))))
          continue;
        if (np < cap) {
          long* ob = out + 6 * np;
          ob[0] = r;
          ob[1] = ii;
          ob[2] = jj;
          ob[3] = sh[0];
          ob[4] = sh[1];
          ob[5] = sh[2];
        }
        np++;
      }
    }
  }
  if (npairs) *npairs = np;
  return IMDG_OK;
}

int imdg_get_grid(imdg_ctx* c, int rep, int* gdim3) {
  if (!c || !c->atoms_set || rep < 0 || rep >= c->nrep || !gdim3)
    return IMDG_ERR_ARG;
  for (int d = 0; d < 3; d++) gdim3[d] = c->rep[rep].gdim[d];
  return IMDG_OK;
}

int imdg_get_info(imdg_ctx* c, imdg_info* info) {
  if (!c || !info) return IMDG_ERR_ARG;
  memset(info, 0, sizeof *info);
  info->nrep = c->nrep;
  info->natoms = c->natoms_tot;
  info->nslots = c->nslots;
  info->kmax = c->kmax;
  info->lanes_per_atom = c->lpa_frc;
  info->list_bytes = (double)c->nbr_bytes;
  if (!c->atoms_set) return IMDG_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Ctl h;
  int rc = read_ctl(c, &h);
  if (rc) return rc;
  RepState st;
  HIPCHK(c, hipMemcpy(&st, c->d_rst, sizeof st, hipMemcpyDeviceToHost));
  info->kmax_seen = h.kmax_seen;
  info->steps_done = h.steps_done;
  info->is_short = h.is_short;
  if (h.is_short) {
    // reported once: the flag covers the work since the previous call
    static const int zero = 0;
    HIPCHK(c, hipMemcpyAsync(&c->d_ctl->is_short, &zero, sizeof(int),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  info->nbl_count = st.nbl_count;
  info->tiles_fallback = h.tiles_fallback;
  info->exact_pairs = h.exact_pairs;
  info->tables_in_lds =
      (c->lds_den_n > 0 ? 1 : 0) | (c->lds_frc_n > 0 ? 2 : 0);
  std::vector<int> hn(c->nslots);
  HIPCHK(c, hipMemcpy(hn.data(), c->d_nnb, sizeof(int) * c->nslots,
                      hipMemcpyDeviceToHost));
  double tot = 0.0;
  for (int r = 0; r < c->nrep; r++)
    for (int k = 0; k < c->rep[r].natoms; k++) tot += hn[c->rep[r].atom_off + k];
  info->mean_k = c->natoms_tot ? tot / c->natoms_tot : 0.0;
  return IMDG_OK;
}

int imdg_profile(imdg_ctx* c, int enable) {
  if (!c) return IMDG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_profile(c);
  c->profiling = enable != 0;
  c->prof_mask = enable == 1 ? ~0u : (unsigned)enable;
  // IMDG_PROFILE_STRIDE=k: time every k-th launch of each class (an event
  // pair makes the stream wait at both ends of the timed launch)
  c->prof_stride = 1;
  if (const char* e = getenv("IMDG_PROFILE_STRIDE"))
    c->prof_stride = std::max(1, atoi(e));
  for (int k = 0; k < KC_N; k++) {
    c->kc_ms[k] = 0;
    c->kc_n[k] = 0;
    c->kc_seen[k] = 0;
  }
  return IMDG_OK;
}

int imdg_get_tuning(imdg_ctx* c, imdg_tuning* out) {
  if (!c || !out) return IMDG_ERR_ARG;
  memset(out, 0, sizeof *out);
  out->density_lanes = c->lpa_den;
  out->density_threads = c->thr_den;
  out->force_lanes = c->lpa_frc;
  out->force_threads = c->thr_frc;
  return IMDG_OK;
}

int imdg_set_tuning(imdg_ctx* c, const imdg_tuning* t) {
  if (!c || !t) return IMDG_ERR_ARG;
  auto lanes_ok = [](int v) { return v == 1 || v == 2 || v == 4; };
  auto thr_ok = [](int v) {
    return v >= IMDG_WAVE && v <= IMDG_STEP_THREADS && v % IMDG_WAVE == 0;
  };
  if (!lanes_ok(t->density_lanes) || !lanes_ok(t->force_lanes) ||
      !thr_ok(t->density_threads) || !thr_ok(t->force_threads))
    return fail(c, IMDG_ERR_ARG, "invalid tuning (lanes 1/2/4, threads a "
                "multiple of 64 up to %d)", IMDG_STEP_THREADS);
  int old[4] = {c->lpa_den, c->thr_den, c->lpa_frc, c->thr_frc};
  c->lpa_den = t->density_lanes;
  c->thr_den = t->density_threads;
  c->lpa_frc = t->force_lanes;
  c->thr_frc = t->force_threads;
  if (!c->atoms_set) return IMDG_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (plan(c) != IMDG_OK) {
    std::string why = c->err;
    c->lpa_den = old[0];
    c->thr_den = old[1];
    c->lpa_frc = old[2];
    c->thr_frc = old[3];
    int rc = plan(c);
    if (rc) return rc;
    return fail(c, IMDG_ERR_ARG, "tuning not launchable: %s", why.c_str());
  }
  drop_graph(c);
  return IMDG_OK;
}

int imdg_autotune(imdg_ctx* c, int reps, imdg_tuning* chosen) {
  if (!c || reps < 1) return IMDG_ERR_ARG;
  int rc = ensure_ready(c);
  if (rc) return rc;
  if ((rc = build_until_fits(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  bool prof = c->profiling;
  c->profiling = false;
  rc = c->mixed ? autotune_t<float>(c, reps, chosen)
                : autotune_t<double>(c, reps, chosen);
  c->profiling = prof;
  if (rc) return rc;
  drop_graph(c);
  // leave forces and energies as calc_forces would
  if ((rc = enqueue_forces(c, false))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return IMDG_OK;
}

int imdg_kernel_times(imdg_ctx* c, const char** names, double* ms,
                      long* launches, int n) {
  if (!c) return IMDG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_profile(c);
  for (int k = 0; k < KC_N && k < n; k++) {
    if (names) names[k] = kKernelNames[k];
    if (ms) ms[k] = c->kc_ms[k];
    if (launches) launches[k] = c->kc_n[k];
  }
  return KC_N;
}

}  // extern "C"
