/*
 * imd_oracle.c -- CPU restatement of IMD's neighbour-list pair/EAM path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * imd_oracle.h for the reference lines each part follows.  Build with
 * -ffp-contract=off so every product and sum rounds separately, as the
 * reference's generic x86-64 build does.
 */
#include "imd_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PSTEP 50 /* src/config.h:251 */

/* ------------------------------------------------------------------------ */
/* tables                                                                    */
/* ------------------------------------------------------------------------ */

static int table_alloc_info(orc_table *pt, int ncols) {
  memset(pt, 0, sizeof(*pt));
  pt->ncols = ncols;
  pt->begin = calloc(ncols, sizeof(double));
  pt->end = calloc(ncols, sizeof(double));
  pt->step = calloc(ncols, sizeof(double));
  pt->invstep = calloc(ncols, sizeof(double));
  pt->len = calloc(ncols, sizeof(int));
  if (!pt->begin || !pt->end || !pt->step || !pt->invstep || !pt->len)
    return ORC_ERR_NOMEM;
  return ORC_OK;
}

void orc_table_free(orc_table *pt) {
  if (!pt) return;
  free(pt->begin);
  free(pt->end);
  free(pt->step);
  free(pt->invstep);
  free(pt->len);
  free(pt->table);
  free(pt->table2);
  memset(pt, 0, sizeof(*pt));
}

int orc_table_nrows(const orc_table *pt) { return pt->maxsteps + 2; }

/* init_threepoint (src/imd_potential.c:1256-1272): continue the last
   interpolation parabola by two rows. */
static int init_threepoint(orc_table *pt) {
  int nc = pt->ncols;
  for (int col = 0; col < nc; col++) {
    double *y = pt->table + col;
    int n = pt->len[col];
    if (n < 3) return ORC_ERR_FORMAT;
    y[n * nc] = 3 * y[(n - 1) * nc] - 3 * y[(n - 2) * nc] + y[(n - 3) * nc];
    y[(n + 1) * nc] =
        6 * y[(n - 1) * nc] - 8 * y[(n - 2) * nc] + 3 * y[(n - 3) * nc];
  }
  return ORC_OK;
}

/* format 1 (src/imd_potential.c:297-376): rows "r2 v0 v1 ..." */
static int read_format1(orc_table *pt, FILE *f, int radial, double *cellsz) {
  int nc = pt->ncols, npot = 0;
  double r2 = 0.0, r2_start = 0.0, val;
  pt->maxsteps = PSTEP;
  pt->table = malloc(sizeof(double) * nc * pt->maxsteps);
  if (!pt->table) return ORC_ERR_NOMEM;
  for (;;) {
    if (npot > 0 && npot % PSTEP == 0) {
      pt->maxsteps += PSTEP;
      double *t = realloc(pt->table, sizeof(double) * nc * pt->maxsteps);
      if (!t) return ORC_ERR_NOMEM;
      pt->table = t;
    }
    if (fscanf(f, "%lf", &r2) != 1) break;
    if (npot == 0) r2_start = r2;
    for (int i = 0; i < nc; i++) {
      if (fscanf(f, "%lf", &val) != 1) return ORC_ERR_FORMAT;
      pt->table[npot * nc + i] = val;
      if (val != 0.0) {
        pt->end[i] = r2;
        pt->len[i] = npot + 1;
      }
    }
    npot++;
  }
  if (npot < 2) return ORC_ERR_FORMAT;
  double r2_step = (r2 - r2_start) / (npot - 1);
  for (int i = 0; i < nc; i++) {
    pt->begin[i] = r2_start;
    pt->step[i] = r2_step;
    pt->invstep[i] = 1.0 / r2_step;
    double delta = pt->table[(npot - 1) * nc + i];
    if (radial && delta != 0.0)
      for (int k = 0; k < npot; k++) pt->table[k * nc + i] -= delta;
    if (radial && pt->end[i] > *cellsz) *cellsz = pt->end[i];
  }
  double *t = realloc(pt->table, sizeof(double) * nc * (pt->maxsteps + 2));
  if (!t) return ORC_ERR_NOMEM;
  pt->table = t;
  return ORC_OK;
}

/* format 2 (src/imd_potential.c:394-462): per column "begin end step",
   then the values column after column. */
static int read_format2(orc_table *pt, FILE *f, int radial, double *cellsz) {
  int nc = pt->ncols;
  pt->maxsteps = 0;
  for (int i = 0; i < nc; i++) {
    if (fscanf(f, "%lf %lf %lf", &pt->begin[i], &pt->end[i], &pt->step[i]) !=
        3)
      return ORC_ERR_FORMAT;
    if (radial && pt->end[i] > *cellsz) *cellsz = pt->end[i];
    pt->invstep[i] = 1.0 / pt->step[i];
    double numstep = 1 + (pt->end[i] - pt->begin[i]) / pt->step[i];
    pt->len[i] = (int)(numstep + 0.49);
    if (pt->len[i] > pt->maxsteps) pt->maxsteps = pt->len[i];
  }
  pt->table = calloc((size_t)nc * (pt->maxsteps + 2), sizeof(double));
  if (!pt->table) return ORC_ERR_NOMEM;
  for (int i = 0; i < nc; i++)
    for (int k = 0; k < pt->len[i]; k++) {
      double val;
      if (fscanf(f, "%lf", &val) != 1) return ORC_ERR_FORMAT;
      pt->table[k * nc + i] = val;
    }
  if (radial)
    for (int i = 0; i < nc; i++) {
      double delta = pt->table[(pt->len[i] - 1) * nc + i];
      if (delta != 0.0)
        for (int k = 0; k < pt->len[i]; k++) pt->table[k * nc + i] -= delta;
    }
  return ORC_OK;
}

/* read_pot_table (src/imd_potential.c:161-282) */
int orc_table_read(orc_table *pt, const char *path, int ncols, int radial,
                   int default_format, double *cellsz) {
  char buf[1024];
  int have_header = 0, have_format = 0, end_header = 0;
  int format = default_format, size = ncols, rc;
  FILE *f = fopen(path, "r");
  if (!f) return ORC_ERR_IO;
  do {
    if (!fgets(buf, sizeof buf, f)) {
      fclose(f);
      return ORC_ERR_FORMAT;
    }
    if (buf[0] == '#') {
      have_header = 1;
      end_header = (buf[1] == 'E');
      if (buf[1] == 'F') {
        if (sscanf(buf + 2, "%d%d", &format, &size) != 2 || size != ncols ||
            (format != 1 && format != 2)) {
          fclose(f);
          return ORC_ERR_FORMAT;
        }
        have_format = 1;
      }
    } else if (have_header) {
      fclose(f);
      return ORC_ERR_FORMAT;
    } else {
      end_header = 1;
    }
  } while (!end_header);
  if (have_header && !have_format) {
    fclose(f);
    return ORC_ERR_FORMAT;
  }
  if (!have_header) rewind(f);
  if ((rc = table_alloc_info(pt, ncols)) != ORC_OK) {
    fclose(f);
    return rc;
  }
  rc = (format == 1) ? read_format1(pt, f, radial, cellsz)
                     : read_format2(pt, f, radial, cellsz);
  fclose(f);
  if (rc == ORC_OK) rc = init_threepoint(pt);
  if (rc != ORC_OK) orc_table_free(pt);
  return rc;
}

/* init_fourpoint (src/imd_potential.c:1171-1187): continue the last
   4-point polynomial by two rows. */
static int init_fourpoint(orc_table *pt) {
  int nc = pt->ncols;
  for (int col = 0; col < nc; col++) {
    double *y = pt->table + col;
    int n = pt->len[col];
    if (n < 4) return ORC_ERR_FORMAT;
    y[n * nc] = 4 * y[(n - 1) * nc] - 6 * y[(n - 2) * nc] +
                4 * y[(n - 3) * nc] - y[(n - 4) * nc];
    y[(n + 1) * nc] = 10 * y[(n - 1) * nc] - 20 * y[(n - 2) * nc] +
                      15 * y[(n - 3) * nc] - 4 * y[(n - 4) * nc];
  }
  return ORC_OK;
}

/* init_spline (src/imd_potential.c:1197-1246): second derivatives of the
   cubic spline through the len rows of each column (tridiagonal sweep,
   natural at the start; at the end first derivative zero for radial
   tables, natural otherwise), then one guard row of values and of second
   derivatives. */
static int init_spline(orc_table *pt, int radial) {
  int nc = pt->ncols, size = pt->maxsteps + 2;
  free(pt->table2);
  pt->table2 = calloc((size_t)nc * size, sizeof(double));
  double *u = calloc(size, sizeof(double));
  if (!pt->table2 || !u) {
    free(u);
    return ORC_ERR_NOMEM;
  }
  for (int col = 0; col < nc; col++) {
    double *y2 = pt->table2 + col, *y = pt->table + col;
    int n = pt->len[col];
    double step = pt->step[col], p, qn, un;
    if (n < 3) {
      free(u);
      return ORC_ERR_FORMAT;
    }
    y2[0] = u[0] = 0;
    for (int i = 1; i < n - 1; i++) {
      p = 0.5 * y2[(i - 1) * nc] + 2.0;
      y2[i * nc] = -0.5 / p;
      u[i] = (y[(i + 1) * nc] - 2 * y[i * nc] + y[(i - 1) * nc]) / step;
      u[i] = (6.0 * u[i] / (2 * step) - 0.5 * u[i - 1]) / p;
    }
    if (radial) {
      qn = 0.5;
      un = (3.0 / step) * (y[(n - 2) * nc] - y[(n - 1) * nc]) / step;
    } else {
      qn = un = 0.0;
    }
    y2[(n - 1) * nc] = (un - qn * u[n - 2]) / (qn * y2[(n - 2) * nc] + 1.0);
    for (int k = n - 2; k >= 0; k--)
      y2[k * nc] = y2[k * nc] * y2[(k + 1) * nc] + u[k];
    y[n * nc] = 2 * y[(n - 1) * nc] - y[(n - 2) * nc] +
                step * step * y2[(n - 1) * nc];
    y2[n * nc] = 2 * y2[(n - 1) * nc] - y2[(n - 2) * nc];
  }
  free(u);
  return ORC_OK;
}

int orc_table_set_interp(orc_table *pt, int interp, int radial) {
  int rc;
  if (interp == ORC_INTERP_4POINT)
    rc = init_fourpoint(pt);
  else if (interp == ORC_INTERP_SPLINE)
    rc = init_spline(pt, radial);
  else if (interp == ORC_INTERP_3POINT)
    rc = init_threepoint(pt);
  else
    return ORC_ERR_ARG;
  if (rc == ORC_OK) pt->interp = interp;
  return rc;
}

/* Common argument handling of the access macros: r2 clamped to end, then
   to begin (is_short), scaled to the table index t >= 0. */
static inline double tab_arg(const orc_table *pt, int col, double r2,
                             int *is_short) {
  double r2a = (r2 < pt->end[col]) ? r2 : pt->end[col];
  r2a = r2a - pt->begin[col];
  if (r2a < 0) {
    r2a = 0;
    *is_short = 1;
  }
  return r2a * pt->invstep[col];
}

/* PAIR_INT2 / VAL_FUNC2 / DERIV_FUNC2 (src/potaccess.h:323-354, 465-495,
   591-621): 3-point on rows k, k+1, k+2. */
static void int2(const orc_table *pt, int col, int inc, double t, double *pot,
                 double *grad) {
  int k = (int)t;
  double chi = t - k;
  const double *ptr = pt->table + (long)k * inc + col;
  double p0 = ptr[0], p1 = ptr[inc], p2 = ptr[2 * inc];
  double dv = p1 - p0;
  double d2v = p2 - 2 * p1 + p0;
  if (pot) *pot = p0 + chi * dv + 0.5 * chi * (chi - 1) * d2v;
  if (grad) *grad = 2 * pt->invstep[col] * (dv + (chi - 0.5) * d2v);
}

/* PAIR_INT3 / VAL_FUNC3 / DERIV_FUNC3 (src/potaccess.h:365-412, 502-532,
   628-660): 4-point on rows k-1 .. k+2, k = max(trunc(t), 1). */
static void int3(const orc_table *pt, int col, int inc, double t, double *pot,
                 double *grad) {
  int k = (int)t;
  if (k < 1) k = 1;
  double chi = t - k;
  const double *ptr = pt->table + (long)(k - 1) * inc + col;
  double p0 = ptr[0], p1 = ptr[inc], p2 = ptr[2 * inc], p3 = ptr[3 * inc];
  if (pot) {
    double fac0 = -(1.0 / 6.0) * chi * (chi - 1.0) * (chi - 2.0);
    double fac1 = 0.5 * (chi * chi - 1.0) * (chi - 2.0);
    double fac2 = -0.5 * chi * (chi + 1.0) * (chi - 2.0);
    double fac3 = (1.0 / 6.0) * chi * (chi * chi - 1.0);
    *pot = fac0 * p0 + fac1 * p1 + fac2 * p2 + fac3 * p3;
  }
  if (grad) {
    double dfac0 = -(1.0 / 6.0) * ((3.0 * chi - 6.0) * chi + 2.0);
    double dfac1 = 0.5 * ((3.0 * chi - 4.0) * chi - 1.0);
    double dfac2 = -0.5 * ((3.0 * chi - 2.0) * chi - 2.0);
    double dfac3 = 1.0 / 6.0 * (3.0 * chi * chi - 1.0);
    *grad = 2 * pt->invstep[col] *
            (dfac0 * p0 + dfac1 * p1 + dfac2 * p2 + dfac3 * p3);
  }
}

/* PAIR_INT_SP / VAL_FUNC_SP / DERIV_FUNC_SP (src/potaccess.h:414-457,
   540-576, 667-702): cubic spline on rows k, k+1 and their second
   derivatives. */
static void int_sp(const orc_table *pt, int col, int inc, double t,
                   double *pot, double *grad) {
  int k = (int)t;
  double b = t - k, a = 1.0 - b;
  double step = pt->step[col], istep = pt->invstep[col];
  long q = (long)k * inc + col;
  double p1 = pt->table[q], d21 = pt->table2[q];
  q += inc;
  double p2 = pt->table[q], d22 = pt->table2[q];
  double a2 = a * a - 1, b2 = b * b - 1, st6 = step / 6;
  if (pot) *pot = a * p1 + b * p2 + (a * a2 * d21 + b * b2 * d22) * st6 * step;
  if (grad)
    *grad = 2 * ((p2 - p1) * istep +
                 ((3 * b2 + 2) * d22 - (3 * a2 + 2) * d21) * st6);
}

static void tab_eval(const orc_table *pt, int col, int inc, double r2,
                     int *is_short, double *pot, double *grad) {
  double t = tab_arg(pt, col, r2, is_short);
  if (pt->interp == ORC_INTERP_4POINT)
    int3(pt, col, inc, t, pot, grad);
  else if (pt->interp == ORC_INTERP_SPLINE)
    int_sp(pt, col, inc, t, pot, grad);
  else
    int2(pt, col, inc, t, pot, grad);
}

void orc_pair_int2(const orc_table *pt, int col, int inc, double r2,
                   double *pot, double *grad, int *is_short) {
  tab_eval(pt, col, inc, r2, is_short, pot, grad);
}

double orc_val_func2(const orc_table *pt, int col, int inc, double r2,
                     int *is_short) {
  double v;
  tab_eval(pt, col, inc, r2, is_short, &v, NULL);
  return v;
}

double orc_deriv_func2(const orc_table *pt, int col, int inc, double r2,
                       int *is_short) {
  double g;
  tab_eval(pt, col, inc, r2, is_short, NULL, &g);
  return g;
}

/* ------------------------------------------------------------------------ */
/* system                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n, cap;
  int *a;
} ivec;

static int ivec_push(ivec *v, int x) {
  if (v->n == v->cap) {
    int nc = v->cap ? 2 * v->cap : 16;
    int *a = realloc(v->a, sizeof(int) * nc);
    if (!a) return ORC_ERR_NOMEM;
    v->a = a;
    v->cap = nc;
  }
  v->a[v->n++] = x;
  return ORC_OK;
}

struct orc_sys {
  /* geometry (src/imd_geom_3d.c:52-104) */
  double box[3][3], tbox[3][3], height[3], volume;
  int pbc[3];
  double nbl_margin, cellsz; /* cellsz: squared list radius incl. skin */
  int gdim[3], cdim[3], nall, ncells;
  int ensemble, ntypes;
  double timestep, temperature, eta, isq_tau_eta;
  const orc_table *pair, *rho_t, *embed;
  const orc_table *emod; /* EEAM energy modification term E(P), or NULL */
  /* real atoms, caller order */
  int n;
  long *id;
  int *type;
  double *mass, *x, *p, *x_nbl;
  int *cell3;
  /* slots: 0..n-1 real atoms, then buffer copies in creation order */
  int nslot, slotcap;
  double *sx;   /* 3 per slot */
  int *stype, *ssrc, *sroot, *sshift; /* sshift: 3 per slot */
  double *svec; /* shift vector added when the slot was copied, 3 per slot */
  double *sf, *sepot, *srho, *sdf;
  double *sP, *sdm; /* EEAM: P = sum rho^2 and 2E'(P) (EAM_P, EAM_DM) */
  /* per-slot stress tensor xx yy zz yz zx xy, when do_press is set */
  double *spress;
  int do_press;
  /* cells */
  ivec *cells;
  int *inner;      /* inner cell indices, IMD order */
  int (*nq)[14];   /* stencil per inner cell */
  /* half list (src/imd_forces_nbl.c:25-43) */
  int *tl, *tb, *tl_atom;
  long tbn, tbcap;
  int have_valid_nbl, nbl_count, is_short;
  double tot_pot_energy, virial, tot_kin_energy;
  /* sum of |terms| of the virial: the scale of its rounding error (a
     tolerance for reduced-precision results, not part of IMD) */
  double virial_abs;
  /* threads of the timing variant (orc_set_threads); 1 = IMD's serial
     order.  tacc: per-thread copies of the scattered accumulators. */
  int nthreads, tused;
  double *tacc;
  long tacc_cap;
};

static void vec_prod(const double *u, const double *v, double *w) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}

static double sprod(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

orc_sys *orc_new(int ntypes, const double *box9, const int *pbc3,
                 double nbl_margin, double cellsz_tab, int ensemble,
                 double timestep) {
  orc_sys *s = calloc(1, sizeof(orc_sys));
  if (!s) return NULL;
  s->ntypes = ntypes;
  s->nthreads = 1;
  for (int d = 0; d < 3; d++) {
    for (int e = 0; e < 3; e++) s->box[d][e] = box9[3 * d + e];
    s->pbc[d] = pbc3[d];
  }
  s->nbl_margin = nbl_margin;
  s->ensemble = ensemble;
  s->timestep = timestep;
  /* make_box: dual vectors, volume, heights */
  vec_prod(s->box[1], s->box[2], s->tbox[0]);
  vec_prod(s->box[2], s->box[0], s->tbox[1]);
  vec_prod(s->box[0], s->box[1], s->tbox[2]);
  s->volume = sprod(s->box[0], s->tbox[0]);
  if (s->volume == 0) {
    free(s);
    return NULL;
  }
  for (int d = 0; d < 3; d++)
    for (int e = 0; e < 3; e++) s->tbox[d][e] /= s->volume;
  for (int d = 0; d < 3; d++) s->height[d] = 1.0 / sprod(s->tbox[d], s->tbox[d]);
  if (s->volume < 0) s->volume = -s->volume;
  /* init_cells: add the skin once, size the grid (tol = 1, cpu_dim = 1) */
  s->cellsz = sqrt(cellsz_tab) + nbl_margin;
  s->cellsz = s->cellsz * s->cellsz;
  for (int d = 0; d < 3; d++) {
    double cs = sqrt(s->cellsz / s->height[d]);
    s->gdim[d] = (int)(1.0 / cs);
    if (s->gdim[d] < 1) {
      free(s);
      return NULL;
    }
    s->cdim[d] = s->gdim[d] + 2;
  }
  s->nall = s->cdim[0] * s->cdim[1] * s->cdim[2];
  s->ncells = s->gdim[0] * s->gdim[1] * s->gdim[2];
  s->cells = calloc(s->nall, sizeof(ivec));
  s->inner = malloc(sizeof(int) * s->ncells);
  s->nq = malloc(sizeof(int[14]) * s->ncells);
  if (!s->cells || !s->inner || !s->nq) {
    orc_free(s);
    return NULL;
  }
  /* make_cell_lists, NBL version (src/imd_geom_3d.c:823-949) */
  int l = 0;
  for (int i = 1; i < s->cdim[0] - 1; i++)
    for (int j = 1; j < s->cdim[1] - 1; j++)
      for (int k = 1; k < s->cdim[2] - 1; k++) {
        int np = (i * s->cdim[1] + j) * s->cdim[2] + k;
        s->inner[l] = np;
        s->nq[l][0] = np;
        int nn = 1;
        for (int a = 0; a <= 1; a++)
          for (int b = -a; b <= 1; b++)
            for (int c = (a == 0 ? -b : -a); c <= 1; c++) {
              int r = i + a, t = j + b, u = k + c;
              int qq = (r * s->cdim[1] + t) * s->cdim[2] + u;
              if (qq == np) continue;
              int g[3] = {r - 1, t - 1, u - 1}, ok = 1;
              for (int d = 0; d < 3; d++) {
                int ip = (g[d] < 0) ? -1 : (g[d] > s->gdim[d] - 1 ? 1 : 0);
                if (!(s->pbc[d] == 1 || ip == 0)) ok = 0;
              }
              s->nq[l][nn++] = ok ? qq : -1;
            }
        l++;
      }
  return s;
}

void orc_free(orc_sys *s) {
  if (!s) return;
  if (s->cells)
    for (int c = 0; c < s->nall; c++) free(s->cells[c].a);
  free(s->cells);
  free(s->inner);
  free(s->nq);
  free(s->id);
  free(s->type);
  free(s->mass);
  free(s->x);
  free(s->p);
  free(s->x_nbl);
  free(s->cell3);
  free(s->sx);
  free(s->stype);
  free(s->ssrc);
  free(s->sroot);
  free(s->sshift);
  free(s->svec);
  free(s->sf);
  free(s->sepot);
  free(s->spress);
  free(s->sP);
  free(s->sdm);
  free(s->srho);
  free(s->sdf);
  free(s->tl);
  free(s->tb);
  free(s->tl_atom);
  free(s->tacc);
  free(s);
}

int orc_set_tables(orc_sys *s, const orc_table *pair, const orc_table *rho,
                   const orc_table *embed) {
  if ((rho == NULL) != (embed == NULL)) return ORC_ERR_ARG;
  int nt2 = s->ntypes * s->ntypes;
  if (pair && pair->ncols != nt2) return ORC_ERR_ARG;
  if (rho && rho->ncols != nt2) return ORC_ERR_ARG;
  if (embed && embed->ncols != s->ntypes) return ORC_ERR_ARG;
  s->pair = pair;
  s->rho_t = rho;
  s->embed = embed;
  return ORC_OK;
}

int orc_set_eeam(orc_sys *s, const orc_table *emod) {
  if (emod && (!s->rho_t || emod->ncols != s->ntypes)) return ORC_ERR_ARG;
  s->emod = emod;
  return ORC_OK;
}

void orc_set_nvt(orc_sys *s, double temperature, double eta,
                 double isq_tau_eta) {
  s->temperature = temperature;
  s->eta = eta;
  s->isq_tau_eta = isq_tau_eta;
}

/* back_into_box (src/imd_geom_3d.c:1083-1110) */
static void back_into_box(const orc_sys *s, double *pos) {
  for (int d = 0; d < 3; d++) {
    if (s->pbc[d] != 1) continue;
    double i = floor(sprod(pos, s->tbox[d]));
    pos[0] -= i * s->box[d][0];
    pos[1] -= i * s->box[d][1];
    pos[2] -= i * s->box[d][2];
  }
}

int orc_set_atoms(orc_sys *s, int n, const long *id, const int *type,
                  const double *mass, const double *x, const double *p) {
  s->n = n;
  s->id = malloc(sizeof(long) * n);
  s->type = malloc(sizeof(int) * n);
  s->mass = malloc(sizeof(double) * n);
  s->x = malloc(sizeof(double) * 3 * n);
  s->p = malloc(sizeof(double) * 3 * n);
  s->x_nbl = malloc(sizeof(double) * 3 * n);
  s->cell3 = malloc(sizeof(int) * 3 * n);
  if (!s->id || !s->type || !s->mass || !s->x || !s->p || !s->x_nbl ||
      !s->cell3)
    return ORC_ERR_NOMEM;
  for (int i = 0; i < n; i++) {
    if (type[i] < 0 || type[i] >= s->ntypes || !(mass[i] > 0))
      return ORC_ERR_ARG;
    s->id[i] = id[i];
    s->type[i] = type[i];
    s->mass[i] = mass[i];
    double pos[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    back_into_box(s, pos);
    for (int d = 0; d < 3; d++) {
      s->x[3 * i + d] = pos[d];
      s->p[3 * i + d] = p ? p[3 * i + d] : 0.0;
    }
  }
  s->have_valid_nbl = 0;
  return ORC_OK;
}

/* cell_coord (src/imd_geom_3d.c:1054-1074) */
static void cell_coord(const orc_sys *s, const double *x, int *c) {
  for (int d = 0; d < 3; d++) {
    c[d] = (int)(s->gdim[d] * (x[0] * s->tbox[d][0] + x[1] * s->tbox[d][1] +
                               x[2] * s->tbox[d][2]));
    if (c[d] >= s->gdim[d])
      c[d] = s->gdim[d] - 1;
    else if (c[d] < 0)
      c[d] = 0;
  }
}

/* fix_cells + do_boundaries: wrap in x, then y, then z, and re-bin every
   atom (src/imd_fix_cells_3d.c:36-201, src/imd_main_3d.c:1972-2065).
   Within a cell atoms are kept in caller order. */
static void fix_cells(orc_sys *s) {
  for (int i = 0; i < s->n; i++) {
    double *x = s->x + 3 * i;
    for (int d = 0; d < 3; d++) {
      if (s->pbc[d] != 1) continue;
      double f = -floor(x[0] * s->tbox[d][0] + x[1] * s->tbox[d][1] +
                        x[2] * s->tbox[d][2]);
      x[0] += f * s->box[d][0];
      x[1] += f * s->box[d][1];
      x[2] += f * s->box[d][2];
    }
  }
  for (int c = 0; c < s->nall; c++) s->cells[c].n = 0;
  for (int i = 0; i < s->n; i++) {
    int *c = s->cell3 + 3 * i;
    cell_coord(s, s->x + 3 * i, c);
    int q = ((c[0] + 1) * s->cdim[1] + (c[1] + 1)) * s->cdim[2] + (c[2] + 1);
    ivec_push(&s->cells[q], i);
  }
  s->have_valid_nbl = 0;
}

static int slot_reserve(orc_sys *s, int need) {
  if (need <= s->slotcap) return ORC_OK;
  int nc = need + need / 2 + 64;
#define GROW(ptr, T, m)                                   \
  do {                                                    \
    T *q_ = realloc(ptr, sizeof(T) * (size_t)(m) * nc);   \
    if (!q_) return ORC_ERR_NOMEM;                        \
    ptr = q_;                                             \
  } while (0)
  GROW(s->sx, double, 3);
  GROW(s->stype, int, 1);
  GROW(s->ssrc, int, 1);
  GROW(s->sroot, int, 1);
  GROW(s->sshift, int, 3);
  GROW(s->svec, double, 3);
  GROW(s->sf, double, 3);
  GROW(s->sepot, double, 1);
  GROW(s->srho, double, 1);
  GROW(s->sdf, double, 1);
  GROW(s->sP, double, 1);
  GROW(s->sdm, double, 1);
  GROW(s->spress, double, 6);
#undef GROW
  s->slotcap = nc;
  return ORC_OK;
}

static int cidx(const orc_sys *s, int i, int j, int k) {
  return (i * s->cdim[1] + j) * s->cdim[2] + k;
}

/* copy_cell (src/imd_comm_force_3d.c:726-782): the destination buffer cell
   receives shifted copies; dir/sgn record the integer image shift. */
static int copy_cell(orc_sys *s, int from, int to, const double *v, int dir,
                     int sgn) {
  ivec *f = &s->cells[from], *t = &s->cells[to];
  t->n = 0;
  for (int m = 0; m < f->n; m++) {
    int src = f->a[m];
    if (slot_reserve(s, s->nslot + 1) != ORC_OK) return ORC_ERR_NOMEM;
    int q = s->nslot++;
    s->ssrc[q] = src;
    s->sroot[q] = s->sroot[src];
    s->stype[q] = s->stype[src];
    for (int d = 0; d < 3; d++) {
      s->svec[3 * q + d] = v[d];
      s->sshift[3 * q + d] = s->sshift[3 * src + d];
      s->sx[3 * q + d] = s->sx[3 * src + d] + v[d];
    }
    s->sshift[3 * q + dir] += sgn;
    if (ivec_push(t, q) != ORC_OK) return ORC_ERR_NOMEM;
  }
  return ORC_OK;
}

/* send_cells, serial branch (src/imd_comm_force_3d.c:222-397): z faces both
   ways, y faces both ways, x only into the high buffer (AR). */
static int build_buffers(orc_sys *s) {
  if (slot_reserve(s, s->n) != ORC_OK) return ORC_ERR_NOMEM;
  s->nslot = s->n;
  for (int i = 0; i < s->n; i++) {
    s->ssrc[i] = -1;
    s->sroot[i] = i;
    s->stype[i] = s->type[i];
    for (int d = 0; d < 3; d++) {
      s->sx[3 * i + d] = s->x[3 * i + d];
      s->sshift[3 * i + d] = 0;
      s->svec[3 * i + d] = 0.0;
    }
  }
  double up[3] = {0, 0, 0}, dn[3] = {0, 0, 0}, no[3] = {0, 0, 0},
         so[3] = {0, 0, 0}, ea[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) {
    if (s->pbc[2] == 1) {
      up[d] = s->box[2][d];
      dn[d] = -s->box[2][d];
    }
    if (s->pbc[1] == 1) {
      no[d] = s->box[1][d];
      so[d] = -s->box[1][d];
    }
    if (s->pbc[0] == 1) ea[d] = s->box[0][d];
  }
  int cx = s->cdim[0], cy = s->cdim[1], cz = s->cdim[2], rc;
  int zs = s->pbc[2] == 1, ys = s->pbc[1] == 1, xs = s->pbc[0] == 1;
  for (int i = 1; i < cx - 1; i++)
    for (int j = 1; j < cy - 1; j++) {
      if ((rc = copy_cell(s, cidx(s, i, j, 1), cidx(s, i, j, cz - 1), up, 2,
                          zs)))
        return rc;
      if ((rc = copy_cell(s, cidx(s, i, j, cz - 2), cidx(s, i, j, 0), dn, 2,
                          -zs)))
        return rc;
    }
  for (int i = 1; i < cx - 1; i++)
    for (int k = 0; k < cz; k++) {
      if ((rc = copy_cell(s, cidx(s, i, 1, k), cidx(s, i, cy - 1, k), no, 1,
                          ys)))
        return rc;
      if ((rc = copy_cell(s, cidx(s, i, cy - 2, k), cidx(s, i, 0, k), so, 1,
                          -ys)))
        return rc;
    }
  for (int j = 0; j < cy; j++)
    for (int k = 0; k < cz; k++)
      if ((rc = copy_cell(s, cidx(s, 1, j, k), cidx(s, cx - 1, j, k), ea, 0,
                          xs)))
        return rc;
  return ORC_OK;
}

/* refresh buffer positions in creation order (every calc_forces) */
static void refresh_buffers(orc_sys *s) {
  for (int i = 0; i < s->n; i++)
    for (int d = 0; d < 3; d++) s->sx[3 * i + d] = s->x[3 * i + d];
  for (int q = s->n; q < s->nslot; q++) {
    int src = s->ssrc[q];
    for (int d = 0; d < 3; d++)
      s->sx[3 * q + d] = s->sx[3 * src + d] + s->svec[3 * q + d];
  }
}

/* make_nblist (src/imd_forces_nbl.c:136-273) */
/* make_nblist on several threads (orc_set_threads): a counting pass and a
   filling pass over the cells, each split over threads, with the same
   per-atom entry order as the serial build, so the list is identical. */
static int nblist_cell(const orc_sys *s, int c, int i, int *out) {
  int c1 = s->inner[c], k = 0;
  const ivec *p = &s->cells[c1];
  int ia = p->a[i];
  double d1[3] = {s->sx[3 * ia], s->sx[3 * ia + 1], s->sx[3 * ia + 2]};
  for (int m = 0; m < 14; m++) {
    int c2 = s->nq[c][m];
    if (c2 < 0) continue;
    int jstart = (c2 == c1) ? i + 1 : 0;
    const ivec *q = &s->cells[c2];
    for (int j = jstart; j < q->n; j++) {
      int js = q->a[j];
      double d[3] = {s->sx[3 * js] - d1[0], s->sx[3 * js + 1] - d1[1],
                     s->sx[3 * js + 2] - d1[2]};
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (r2 < s->cellsz) {
        if (out) out[k] = js;
        k++;
      }
    }
  }
  return k;
}

static int make_nblist_threaded(orc_sys *s) {
  int T = s->nthreads, nc = s->ncells;
  int *first = malloc(sizeof(int) * (nc + 1));
  if (!first) return ORC_ERR_NOMEM;
  first[0] = 0;
  for (int c = 0; c < nc; c++) first[c + 1] = first[c] + s->cells[s->inner[c]].n;
#pragma omp parallel for num_threads(T) schedule(dynamic, 16)
  for (int c = 0; c < nc; c++) {
    const ivec *p = &s->cells[s->inner[c]];
    for (int i = 0; i < p->n; i++) {
      s->tl_atom[first[c] + i] = p->a[i];
      s->tl[first[c] + i + 1] = nblist_cell(s, c, i, NULL);
    }
  }
  long tn = 0;
  for (int n = 0; n < s->n; n++) {
    tn += s->tl[n + 1];
    if (tn > 2147483647L) {
      free(first);
      return ORC_ERR_NOMEM;
    }
    s->tl[n + 1] = (int)tn;
  }
  if (tn > s->tbcap) {
    int *t = realloc(s->tb, sizeof(int) * (tn > 0 ? tn : 1));
    if (!t) {
      free(first);
      return ORC_ERR_NOMEM;
    }
    s->tb = t;
    s->tbcap = tn;
  }
#pragma omp parallel for num_threads(T) schedule(dynamic, 16)
  for (int c = 0; c < nc; c++) {
    const ivec *p = &s->cells[s->inner[c]];
    for (int i = 0; i < p->n; i++)
      nblist_cell(s, c, i, s->tb + s->tl[first[c] + i]);
  }
  free(first);
  s->tbn = tn;
  s->have_valid_nbl = 1;
  s->nbl_count++;
  return ORC_OK;
}

static int make_nblist(orc_sys *s) {
  for (int i = 0; i < 3 * s->n; i++) s->x_nbl[i] = s->x[i];
  free(s->tl);
  free(s->tl_atom);
  s->tl = malloc(sizeof(int) * (s->n + 1));
  s->tl_atom = malloc(sizeof(int) * (s->n + 1));
  if (!s->tl || !s->tl_atom) return ORC_ERR_NOMEM;
  s->tl[0] = 0;
  if (s->nthreads > 1) return make_nblist_threaded(s);
  long tn = 0;
  int n = 0;
  s->tl[0] = 0;
  for (int c = 0; c < s->ncells; c++) {
    int c1 = s->inner[c];
    ivec *p = &s->cells[c1];
    for (int i = 0; i < p->n; i++) {
      int ia = p->a[i];
      double d1[3] = {s->sx[3 * ia], s->sx[3 * ia + 1], s->sx[3 * ia + 2]};
      for (int m = 0; m < 14; m++) {
        int c2 = s->nq[c][m];
        if (c2 < 0) continue;
        int jstart = (c2 == c1) ? i + 1 : 0;
        ivec *q = &s->cells[c2];
        for (int j = jstart; j < q->n; j++) {
          int js = q->a[j];
          double d[3] = {s->sx[3 * js] - d1[0], s->sx[3 * js + 1] - d1[1],
                         s->sx[3 * js + 2] - d1[2]};
          double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          if (r2 < s->cellsz) {
            if (tn == s->tbcap) {
              long nc = s->tbcap ? 2 * s->tbcap : 1 << 16;
              int *t = realloc(s->tb, sizeof(int) * nc);
              if (!t) return ORC_ERR_NOMEM;
              s->tb = t;
              s->tbcap = nc;
            }
            s->tb[tn++] = js;
          }
        }
      }
      s->tl_atom[n] = ia;
      s->tl[++n] = (int)tn;
    }
  }
  s->tbn = tn;
  s->have_valid_nbl = 1;
  s->nbl_count++;
  return ORC_OK;
}

/* calc_forces, PAIR + EAM2 branches (src/imd_forces_nbl.c:281-1323) */
/* STRESS_TENS of one pair (src/imd_forces_nbl.c:558-581, 1283-1318): half
   of the pair force, h = fc / 2, gives -d_a h_b to the neighbour's tensor
   and to the atom's running sum ps, in the order xx yy zz yz zx xy. */
static void press_pair(double *pj, double *ps, const double *d,
                       const double *fc) {
  double h[3] = {fc[0] * 0.5, fc[1] * 0.5, fc[2] * 0.5};
  double t[6] = {d[0] * h[0], d[1] * h[1], d[2] * h[2],
                 d[1] * h[2], d[2] * h[0], d[0] * h[1]};
  for (int k = 0; k < 6; k++) {
    ps[k] -= t[k];
    pj[k] -= t[k];
  }
}

/* the kinetic part p (x) p / m with the new momenta
   (src/imd_integrate.c:411-431, 1085-1094) */
static void press_kin(double *pi, const double *p, double m) {
  pi[0] += p[0] * p[0] / m;
  pi[1] += p[1] * p[1] / m;
  pi[2] += p[2] * p[2] / m;
  pi[3] += p[1] * p[2] / m;
  pi[4] += p[2] * p[0] / m;
  pi[5] += p[0] * p[1] / m;
}

void orc_set_press(orc_sys *s, int on) { s->do_press = on; }

void orc_get_eeam(const orc_sys *s, double *P, double *dm) {
  for (int i = 0; i < s->n; i++) {
    if (P) P[i] = s->sP ? s->sP[i] : 0.0;
    if (dm) dm[i] = s->sdm ? s->sdm[i] : 0.0;
  }
}

void orc_get_press(const orc_sys *s, double *out) {
  for (int i = 0; i < s->n; i++)
    for (int k = 0; k < 6; k++)
      out[6 * i + k] = s->spress ? s->spress[6 * i + k] : 0.0;
}

/* Timing variant of the force loops on several cores (orc_set_threads):
   the same half list and the same pair, density, embedding and EAM terms
   as orc_calc_forces, with loops A and B split over threads by atom.  The
   scatter to j goes into per-thread copies of the accumulators, which are
   summed after each loop, so the sums are taken in another order than
   IMD's serial loop and differ from it in the last bits.  Used for pair
   and EAM runs without stress or EEAM. */
void orc_set_threads(orc_sys *s, int n) { s->nthreads = n > 1 ? n : 1; }

static int tacc_reserve(orc_sys *s, long per_thread) {
  long need = (long)s->nthreads * per_thread;
  if (need > s->tacc_cap) {
    double *t = realloc(s->tacc, (size_t)need * sizeof(double));
    if (!t) return ORC_ERR_NOMEM;
    s->tacc = t;
    s->tacc_cap = need;
  }
  return ORC_OK;
}

static int forces_threaded(orc_sys *s) {
  int nt = s->ntypes, inc = nt * nt, ns = s->nslot, T = s->nthreads;
  const orc_table *pp = s->pair, *rt = s->rho_t, *et = s->embed;
  long blk = 5L * ns;
  int rc;
  if ((rc = tacc_reserve(s, blk))) return rc;
  double epot = 0.0, vir = 0.0, vabs = 0.0;
  int shrt = 0;
  /* loop A: pair forces and host electron density */
#pragma omp parallel num_threads(T) reduction(+ : epot, vir, vabs) \
    reduction(| : shrt)
  {
    int t = omp_get_thread_num();
    if (t == 0) s->tused = omp_get_num_threads();
    double *af = s->tacc + t * blk, *ae = af + 3L * ns, *ar = ae + ns;
    memset(af, 0, (size_t)blk * sizeof(double));
#pragma omp for schedule(static)
    for (int n = 0; n < s->n; n++) {
      int ia = s->tl_atom[n];
      const double *d1 = s->sx + 3 * ia;
      double ff[3] = {0, 0, 0}, ee = 0.0, eam_r = 0.0;
      int it = s->stype[ia];
      for (int m = s->tl[n]; m < s->tl[n + 1]; m++) {
        int js = s->tb[m];
        double d[3] = {s->sx[3 * js] - d1[0], s->sx[3 * js + 1] - d1[1],
                       s->sx[3 * js + 2] - d1[2]};
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        int jt = s->stype[js], col = it * nt + jt, col2 = jt * nt + it;
        if (pp && r2 <= pp->end[col]) {
          double pot, grad;
          orc_pair_int2(pp, col, inc, r2, &pot, &grad, &shrt);
          epot += pot;
          double fc[3] = {d[0] * grad, d[1] * grad, d[2] * grad};
          af[3 * js] -= fc[0];
          af[3 * js + 1] -= fc[1];
          af[3 * js + 2] -= fc[2];
          ff[0] += fc[0];
          ff[1] += fc[1];
          ff[2] += fc[2];
          pot *= 0.5;
          ee += pot;
          ae[js] += pot;
          vir -= r2 * grad;
          vabs += fabs(r2 * grad);
        }
        if (rt) {
          double rho_h = 0.0;
          if (r2 < rt->end[col]) {
            rho_h = orc_val_func2(rt, col, inc, r2, &shrt);
            eam_r += rho_h;
          }
          if (it == jt) {
            if (r2 < rt->end[col]) ar[js] += rho_h;
          } else if (r2 < rt->end[col2]) {
            ar[js] += orc_val_func2(rt, col2, inc, r2, &shrt);
          }
        }
      }
      af[3 * ia] += ff[0];
      af[3 * ia + 1] += ff[1];
      af[3 * ia + 2] += ff[2];
      ae[ia] += ee;
      ar[ia] += eam_r;
    }
  }
  int used = s->tused;
#pragma omp parallel for num_threads(T) schedule(static)
  for (int q = 0; q < ns; q++) {
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, e = 0.0, r = 0.0;
    for (int t = 0; t < used; t++) {
      const double *b = s->tacc + t * blk;
      f0 += b[3L * q];
      f1 += b[3L * q + 1];
      f2 += b[3L * q + 2];
      e += b[3L * ns + q];
      r += b[4L * ns + q];
    }
    s->sf[3 * q] = f0;
    s->sf[3 * q + 1] = f1;
    s->sf[3 * q + 2] = f2;
    s->sepot[q] = e;
    s->srho[q] = r;
  }
  if (rt) {
    /* add_rho: buffer copies back to their sources (x, then y, then z) */
    for (int q = ns - 1; q >= s->n; q--) s->srho[s->ssrc[q]] += s->srho[q];
    /* embedding energy and 2F' */
    double eemb = 0.0;
#pragma omp parallel for num_threads(T) schedule(dynamic, 64) \
    reduction(+ : eemb)
    for (int c = 0; c < s->ncells; c++) {
      ivec *p = &s->cells[s->inner[c]];
      int idummy = 0;
      for (int i = 0; i < p->n; i++) {
        int ia = p->a[i];
        double pot, df;
        orc_pair_int2(et, s->stype[ia], nt, s->srho[ia], &pot, &df, &idummy);
        s->sdf[ia] = df;
        s->sepot[ia] += pot;
        eemb += pot;
      }
    }
    epot += eemb;
    /* copy_dF: a buffer slot may be a copy of an earlier buffer slot, so
       each takes the value of its real root atom (the same value the
       chain of copies carries), which keeps the copies independent */
#pragma omp parallel for num_threads(T) schedule(static)
    for (int q = s->n; q < ns; q++) s->sdf[q] = s->sdf[s->sroot[q]];
    /* loop B: EAM forces */
    long fblk = 3L * ns;
    double vir2 = 0.0, vabs2 = 0.0;
#pragma omp parallel num_threads(T) reduction(+ : vir2, vabs2) \
    reduction(| : shrt)
    {
      int t = omp_get_thread_num();
      if (t == 0) s->tused = omp_get_num_threads();
      double *af = s->tacc + t * fblk;
      memset(af, 0, (size_t)fblk * sizeof(double));
#pragma omp for schedule(static)
      for (int n = 0; n < s->n; n++) {
        int ia = s->tl_atom[n];
        const double *d1 = s->sx + 3 * ia;
        double ff[3] = {0, 0, 0};
        int it = s->stype[ia];
        for (int m = s->tl[n]; m < s->tl[n + 1]; m++) {
          int js = s->tb[m];
          double d[3] = {s->sx[3 * js] - d1[0], s->sx[3 * js + 1] - d1[1],
                         s->sx[3 * js + 2] - d1[2]};
          double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          int jt = s->stype[js], col1 = jt * nt + it, col2 = it * nt + jt;
          if (r2 < rt->end[col1] || r2 < rt->end[col2]) {
            double rho_i_s = orc_deriv_func2(rt, col1, inc, r2, &shrt);
            double rho_j_s = (col1 == col2)
                                 ? rho_i_s
                                 : orc_deriv_func2(rt, col2, inc, r2, &shrt);
            double grad = 0.5 * (s->sdf[ia] * rho_j_s + s->sdf[js] * rho_i_s);
            double fc[3] = {d[0] * grad, d[1] * grad, d[2] * grad};
            af[3 * js] -= fc[0];
            af[3 * js + 1] -= fc[1];
            af[3 * js + 2] -= fc[2];
            ff[0] += fc[0];
            ff[1] += fc[1];
            ff[2] += fc[2];
            double w = d[0] * fc[0] + d[1] * fc[1] + d[2] * fc[2];
            vir2 -= w;
            vabs2 += fabs(w);
          }
        }
        af[3 * ia] += ff[0];
        af[3 * ia + 1] += ff[1];
        af[3 * ia + 2] += ff[2];
      }
    }
    used = s->tused;
#pragma omp parallel for num_threads(T) schedule(static)
    for (int q = 0; q < ns; q++) {
      for (int t = 0; t < used; t++) {
        const double *b = s->tacc + t * fblk;
        s->sf[3 * q] += b[3L * q];
        s->sf[3 * q + 1] += b[3L * q + 1];
        s->sf[3 * q + 2] += b[3L * q + 2];
      }
    }
    vir += vir2;
    vabs += vabs2;
  }
  /* add_forces: buffer copies back to their sources (x, then y, then z) */
  for (int q = ns - 1; q >= s->n; q--) {
    int src = s->ssrc[q];
    s->sf[3 * src] += s->sf[3 * q];
    s->sf[3 * src + 1] += s->sf[3 * q + 1];
    s->sf[3 * src + 2] += s->sf[3 * q + 2];
    s->sepot[src] += s->sepot[q];
  }
  s->tot_pot_energy = epot;
  s->virial = vir;
  s->virial_abs = vabs;
  s->is_short |= shrt;
  return ORC_OK;
}

int orc_calc_forces(orc_sys *s, int steps) {
  (void)steps;
  int rc, nt = s->ntypes, inc = nt * nt;
  if (!s->have_valid_nbl) {
    fix_cells(s);
    if ((rc = build_buffers(s))) return rc;
    if ((rc = make_nblist(s))) return rc;
  } else {
    refresh_buffers(s);
  }
  s->tot_pot_energy = 0.0;
  s->virial = 0.0;
  s->virial_abs = 0.0;
  for (int q = 0; q < s->nslot; q++) {
    s->sf[3 * q] = s->sf[3 * q + 1] = s->sf[3 * q + 2] = 0.0;
    s->sepot[q] = 0.0;
    s->srho[q] = 0.0;
    s->sP[q] = 0.0;
    if (s->do_press)
      for (int k = 0; k < 6; k++) s->spress[6 * q + k] = 0.0;
  }
  if (s->nthreads > 1 && !s->do_press && !s->emod) return forces_threaded(s);
  const orc_table *pp = s->pair, *rt = s->rho_t, *et = s->embed;
  int is_short = 0;
  /* loop A: pair forces and host electron density */
  for (int n = 0; n < s->n; n++) {
    int ia = s->tl_atom[n];
    const double *d1 = s->sx + 3 * ia;
    double ff[3] = {0, 0, 0}, ee = 0.0, eam_r = 0.0, eam_p = 0.0;
    double ps[6] = {0, 0, 0, 0, 0, 0};
    int it = s->stype[ia];
    for (int m = s->tl[n]; m < s->tl[n + 1]; m++) {
      int js = s->tb[m];
      double d[3] = {s->sx[3 * js] - d1[0], s->sx[3 * js + 1] - d1[1],
                     s->sx[3 * js + 2] - d1[2]};
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      int jt = s->stype[js], col = it * nt + jt, col2 = jt * nt + it;
      if (pp && r2 <= pp->end[col]) {
        double pot, grad;
        orc_pair_int2(pp, col, inc, r2, &pot, &grad, &is_short);
        s->tot_pot_energy += pot;
        double fc[3] = {d[0] * grad, d[1] * grad, d[2] * grad};
        s->sf[3 * js] -= fc[0];
        s->sf[3 * js + 1] -= fc[1];
        s->sf[3 * js + 2] -= fc[2];
        ff[0] += fc[0];
        ff[1] += fc[1];
        ff[2] += fc[2];
        pot *= 0.5;
        ee += pot;
        s->sepot[js] += pot;
        s->virial -= r2 * grad;
        s->virial_abs += fabs(r2 * grad);
        if (s->do_press) press_pair(s->spress + 6 * js, ps, d, fc);
      }
      if (rt) {
        double rho_h = 0.0;
        if (r2 < rt->end[col]) {
          rho_h = orc_val_func2(rt, col, inc, r2, &is_short);
          eam_r += rho_h;
          eam_p += rho_h * rho_h; /* EEAM (src/imd_forces_nbl.c:591-593) */
        }
        if (it == jt) {
          if (r2 < rt->end[col]) {
            s->srho[js] += rho_h;
            s->sP[js] += rho_h * rho_h;
          }
        } else if (r2 < rt->end[col2]) {
          rho_h = orc_val_func2(rt, col2, inc, r2, &is_short);
          s->srho[js] += rho_h;
          s->sP[js] += rho_h * rho_h;
        }
      }
    }
    s->sf[3 * ia] += ff[0];
    s->sf[3 * ia + 1] += ff[1];
    s->sf[3 * ia + 2] += ff[2];
    s->sepot[ia] += ee;
    s->srho[ia] += eam_r;
    s->sP[ia] += eam_p;
    if (s->do_press)
      for (int k = 0; k < 6; k++) s->spress[6 * ia + k] += ps[k];
  }
  if (rt) {
    /* add_rho: buffer copies back to their sources (x, then y, then z) */
    for (int q = s->nslot - 1; q >= s->n; q--) {
      s->srho[s->ssrc[q]] += s->srho[q];
      s->sP[s->ssrc[q]] += s->sP[q];
    }
    /* embedding energy and 2F' (src/imd_forces_nbl.c:1079-1112) */
    int idummy = 0;
    for (int c = 0; c < s->ncells; c++) {
      ivec *p = &s->cells[s->inner[c]];
      for (int i = 0; i < p->n; i++) {
        int ia = p->a[i];
        double pot, df;
        orc_pair_int2(et, s->stype[ia], nt, s->srho[ia], &pot, &df, &idummy);
        s->sdf[ia] = df;
        s->sepot[ia] += pot;
        s->tot_pot_energy += pot;
        if (s->emod) {
          /* EEAM: E(P) and EAM_DM = 2E'(P) (src/imd_forces_nbl.c:1090-1095) */
          orc_pair_int2(s->emod, s->stype[ia], nt, s->sP[ia], &pot, &df,
                        &idummy);
          s->sdm[ia] = df;
          s->sepot[ia] += pot;
          s->tot_pot_energy += pot;
        }
      }
    }
    /* copy_dF into buffer cells (creation order) */
    for (int q = s->n; q < s->nslot; q++) {
      s->sdf[q] = s->sdf[s->ssrc[q]];
      s->sdm[q] = s->sdm[s->ssrc[q]];
    }
    /* loop B: EAM forces */
    for (int n = 0; n < s->n; n++) {
      int ia = s->tl_atom[n];
      const double *d1 = s->sx + 3 * ia;
      double ff[3] = {0, 0, 0};
      double ps[6] = {0, 0, 0, 0, 0, 0};
      int it = s->stype[ia];
      for (int m = s->tl[n]; m < s->tl[n + 1]; m++) {
        int js = s->tb[m];
        double d[3] = {s->sx[3 * js] - d1[0], s->sx[3 * js + 1] - d1[1],
                       s->sx[3 * js + 2] - d1[2]};
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        int jt = s->stype[js], col1 = jt * nt + it, col2 = it * nt + jt;
        if (r2 < rt->end[col1] || r2 < rt->end[col2]) {
          double rho_i = 0.0, rho_j = 0.0, rho_i_s, rho_j_s;
          if (s->emod) {
            /* EEAM: values too (src/imd_forces_nbl.c:1191-1199) */
            orc_pair_int2(rt, col1, inc, r2, &rho_i, &rho_i_s, &is_short);
            if (col1 == col2) {
              rho_j = rho_i;
              rho_j_s = rho_i_s;
            } else {
              orc_pair_int2(rt, col2, inc, r2, &rho_j, &rho_j_s, &is_short);
            }
          } else {
            rho_i_s = orc_deriv_func2(rt, col1, inc, r2, &is_short);
            rho_j_s = (col1 == col2)
                          ? rho_i_s
                          : orc_deriv_func2(rt, col2, inc, r2, &is_short);
          }
          double grad = 0.5 * (s->sdf[ia] * rho_j_s + s->sdf[js] * rho_i_s);
          if (s->emod)
            grad += (s->sdm[ia] * rho_j * rho_j_s + s->sdm[js] * rho_i * rho_i_s);
          double fc[3] = {d[0] * grad, d[1] * grad, d[2] * grad};
          s->sf[3 * js] -= fc[0];
          s->sf[3 * js + 1] -= fc[1];
          s->sf[3 * js + 2] -= fc[2];
          ff[0] += fc[0];
          ff[1] += fc[1];
          ff[2] += fc[2];
          s->virial -= d[0] * fc[0] + d[1] * fc[1] + d[2] * fc[2];
          s->virial_abs += fabs(d[0] * fc[0] + d[1] * fc[1] + d[2] * fc[2]);
          if (s->do_press) press_pair(s->spress + 6 * js, ps, d, fc);
        }
      }
      s->sf[3 * ia] += ff[0];
      s->sf[3 * ia + 1] += ff[1];
      s->sf[3 * ia + 2] += ff[2];
      if (s->do_press)
        for (int k = 0; k < 6; k++) s->spress[6 * ia + k] += ps[k];
    }
  }
  /* add_forces: buffer copies back to their sources (x, then y, then z) */
  for (int q = s->nslot - 1; q >= s->n; q--) {
    int src = s->ssrc[q];
    s->sf[3 * src] += s->sf[3 * q];
    s->sf[3 * src + 1] += s->sf[3 * q + 1];
    s->sf[3 * src + 2] += s->sf[3 * q + 2];
    s->sepot[src] += s->sepot[q];
    /* add_forces carries the stress too (src/imd_comm_force_3d.c:916-921) */
    if (s->do_press)
      for (int k = 0; k < 6; k++) s->spress[6 * src + k] += s->spress[6 * q + k];
  }
  s->is_short |= is_short;
  return ORC_OK;
}

/* move_atoms_nve (src/imd_integrate.c:32-497) and move_atoms_nvt
   (src/imd_integrate.c:891-1147), without restrictions (all ones). */
void orc_move_atoms(orc_sys *s) {
  double dt = s->timestep;
  if (s->ensemble == ORC_NVE) {
    double ek = 0.0;
    /* one thread unless orc_set_threads asked for more; one thread keeps
       IMD's serial order (the reduction starts from 0 and adds once) */
#pragma omp parallel for if (s->nthreads > 1) num_threads(s->nthreads) \
    schedule(static) reduction(+ : ek)
    for (int c = 0; c < s->ncells; c++) {
      ivec *cl = &s->cells[s->inner[c]];
      for (int a = 0; a < cl->n; a++) {
        int i = cl->a[a];
        double *p = s->p + 3 * i, *x = s->x + 3 * i, *f = s->sf + 3 * i;
        double k1 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        p[0] += dt * f[0];
        p[1] += dt * f[1];
        p[2] += dt * f[2];
        double k2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        ek += (k1 + k2) / (4 * s->mass[i]);
        double tmp = dt / s->mass[i];
        x[0] += tmp * p[0];
        x[1] += tmp * p[1];
        x[2] += tmp * p[2];
        if (s->do_press) press_kin(s->spress + 6 * i, p, s->mass[i]);
      }
    }
    s->tot_kin_energy = ek;
  } else {
    double e1 = 0.0, e2 = 0.0;
    double reibung = 1.0 - s->eta * dt / 2.0;
    double eins_d_reib = 1.0 / (1.0 + s->eta * dt / 2.0);
#pragma omp parallel for if (s->nthreads > 1) num_threads(s->nthreads) \
    schedule(static) reduction(+ : e1, e2)
    for (int c = 0; c < s->ncells; c++) {
      ivec *cl = &s->cells[s->inner[c]];
      for (int a = 0; a < cl->n; a++) {
        int i = cl->a[a];
        double *p = s->p + 3 * i, *x = s->x + 3 * i, *f = s->sf + 3 * i;
        e1 += (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) / s->mass[i];
        p[0] = (p[0] * reibung + dt * f[0]) * eins_d_reib;
        p[1] = (p[1] * reibung + dt * f[1]) * eins_d_reib;
        p[2] = (p[2] * reibung + dt * f[2]) * eins_d_reib;
        e2 += (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) / s->mass[i];
        double tmp = dt / s->mass[i];
        x[0] += tmp * p[0];
        x[1] += tmp * p[1];
        x[2] += tmp * p[2];
        if (s->do_press) press_kin(s->spress + 6 * i, p, s->mass[i]);
      }
    }
    s->tot_kin_energy = (e1 + e2) / 4.0;
    double ttt = 3.0 * s->n * s->temperature;
    s->eta += dt * (e2 / ttt - 1.0) * s->isq_tau_eta;
  }
}

/* check_nblist (src/imd_forces_nbl.c:2007-2037) */
void orc_check_nblist(orc_sys *s) {
  double max1 = 0.0;
#pragma omp parallel for if (s->nthreads > 1) num_threads(s->nthreads) \
    schedule(static) reduction(max : max1)
  for (int i = 0; i < s->n; i++) {
    double d[3] = {s->x[3 * i] - s->x_nbl[3 * i],
                   s->x[3 * i + 1] - s->x_nbl[3 * i + 1],
                   s->x[3 * i + 2] - s->x_nbl[3 * i + 2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 > max1) max1 = r2;
  }
  double h = 0.5 * s->nbl_margin;
  if (max1 > h * h) s->have_valid_nbl = 0;
}

void orc_invalidate_nblist(orc_sys *s) { s->have_valid_nbl = 0; }

int orc_run(orc_sys *s, int nsteps, double *thermo) {
  for (int st = 0; st < nsteps; st++) {
    int before = s->nbl_count, rc;
    if ((rc = orc_calc_forces(s, st))) return rc;
    orc_move_atoms(s);
    orc_check_nblist(s);
    if (thermo) {
      double *t = thermo + 5 * st;
      t[0] = s->tot_pot_energy;
      t[1] = s->tot_kin_energy;
      t[2] = s->virial;
      t[3] = s->eta;
      t[4] = (double)(s->nbl_count - before);
    }
  }
  return ORC_OK;
}

void orc_get_atoms(const orc_sys *s, double *x, double *p, double *f,
                   double *epot, double *rho, double *df) {
  for (int i = 0; i < s->n; i++) {
    for (int d = 0; d < 3; d++) {
      if (x) x[3 * i + d] = s->x[3 * i + d];
      if (p) p[3 * i + d] = s->p[3 * i + d];
      if (f) f[3 * i + d] = s->sf ? s->sf[3 * i + d] : 0.0;
    }
    if (epot) epot[i] = s->sepot ? s->sepot[i] : 0.0;
    if (rho) rho[i] = s->srho ? s->srho[i] : 0.0;
    if (df) df[i] = s->sdf ? s->sdf[i] : 0.0;
  }
}

double orc_get_scalar(const orc_sys *s, int which) {
  switch (which) {
    case 0: return s->tot_pot_energy;
    case 1: return s->virial;
    case 2: return s->tot_kin_energy;
    case 3: return s->eta;
    case 4: return s->cellsz;
    case 5: return s->volume;
    case 6: return s->is_short;
    case 7: return s->nbl_count;
    case 8: return s->n;
    case 9: return s->virial_abs;
    default: return 0.0;
  }
}

void orc_get_grid(const orc_sys *s, int *gdim3) {
  for (int d = 0; d < 3; d++) gdim3[d] = s->gdim[d];
}

void orc_get_cells(const orc_sys *s, int *cell3) {
  memcpy(cell3, s->cell3, sizeof(int) * 3 * s->n);
}

long orc_get_pairs(const orc_sys *s, long *out, long cap) {
  long np = 0;
  if (!s->have_valid_nbl && s->nbl_count == 0) return 0;
  for (int n = 0; n < s->n; n++) {
    int ia = s->tl_atom[n];
    for (int m = s->tl[n]; m < s->tl[n + 1]; m++) {
      int js = s->tb[m], jr = s->sroot[js];
      const int *sh = s->sshift + 3 * js;
      if (np < cap) {
        long *o = out + 5 * np;
        if (s->id[ia] < s->id[jr]) {
          o[0] = s->id[ia];
          o[1] = s->id[jr];
          o[2] = sh[0];
          o[3] = sh[1];
          o[4] = sh[2];
        } else {
          o[0] = s->id[jr];
          o[1] = s->id[ia];
          o[2] = -sh[0];
          o[3] = -sh[1];
          o[4] = -sh[2];
        }
      }
      np++;
    }
  }
  return np;
}
