/*
 * host_potential.c -- IMD potential tables.
 *
 * read_pot_table follows src/imd_potential.c:161-282: an optional header of
 * '#' lines, "#F format ncols" selecting format 1 (rows "r2 v0 v1 ...",
 * src/imd_potential.c:297-376) or format 2 (per column "begin end step",
 * then the values column after column, src/imd_potential.c:394-462),
 * "#E" ending it; files without a header use format 2 when EAM tables are
 * read and format 1 otherwise (src/config.h:58-63).  Radial tables (in r^2)
 * are shifted so that each column ends at zero, and the list radius cellsz
 * grows to their largest end.  init_threepoint appends the two guard rows
 * continuing the last parabola (src/imd_potential.c:1256-1272).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "imd_host.h"

#define PSTEP 50 /* table growth step of format 1 (src/config.h:251) */

pot_table_t pair_pot, embed_pot, rho_h_tab, emod_pot;
double cellsz = 0.0;

void free_pot_table(pot_table_t *pt) {
  free(pt->begin);
  free(pt->end);
  free(pt->step);
  free(pt->invstep);
  free(pt->len);
  free(pt->table);
  memset(pt, 0, sizeof *pt);
}

static int alloc_info(pot_table_t *pt, int ncols) {
  memset(pt, 0, sizeof *pt);
  pt->ncols = ncols;
  pt->begin = calloc(ncols, sizeof(double));
  pt->end = calloc(ncols, sizeof(double));
  pt->step = calloc(ncols, sizeof(double));
  pt->invstep = calloc(ncols, sizeof(double));
  pt->len = calloc(ncols, sizeof(int));
  return (pt->begin && pt->end && pt->step && pt->invstep && pt->len) ? 0 : 1;
}

static int format1(pot_table_t *pt, FILE *f, int radial) {
  int nc = pt->ncols, npot = 0;
  double r2 = 0.0, r2_start = 0.0, val;
  pt->maxsteps = PSTEP;
  pt->table = malloc(sizeof(double) * nc * pt->maxsteps);
  if (!pt->table) return 1;
  for (;;) {
    if (npot > 0 && npot % PSTEP == 0) {
      pt->maxsteps += PSTEP;
      double *t = realloc(pt->table, sizeof(double) * nc * pt->maxsteps);
      if (!t) return 1;
      pt->table = t;
    }
    if (fscanf(f, "%lf", &r2) != 1) break;
    if (npot == 0) r2_start = r2;
    for (int i = 0; i < nc; i++) {
      if (fscanf(f, "%lf", &val) != 1) return 2;
      pt->table[(size_t)npot * nc + i] = val;
      if (val != 0.0) {
        pt->end[i] = r2;
        pt->len[i] = npot + 1;
      }
    }
    npot++;
  }
  if (npot < 3) return 2;
  double r2_step = (r2 - r2_start) / (npot - 1);
  for (int i = 0; i < nc; i++) {
    pt->begin[i] = r2_start;
    pt->step[i] = r2_step;
    pt->invstep[i] = 1.0 / r2_step;
    double delta = pt->table[(size_t)(npot - 1) * nc + i];
    if (radial && delta != 0.0)
      for (int k = 0; k < npot; k++) pt->table[(size_t)k * nc + i] -= delta;
    if (radial && pt->end[i] > cellsz) cellsz = pt->end[i];
  }
  double *t = realloc(pt->table, sizeof(double) * nc * (pt->maxsteps + 2));
  if (!t) return 1;
  pt->table = t;
  return 0;
}

static int format2(pot_table_t *pt, FILE *f, int radial) {
  int nc = pt->ncols;
  pt->maxsteps = 0;
  for (int i = 0; i < nc; i++) {
    if (fscanf(f, "%lf %lf %lf", &pt->begin[i], &pt->end[i], &pt->step[i]) != 3)
      return 2;
    if (radial && pt->end[i] > cellsz) cellsz = pt->end[i];
    pt->invstep[i] = 1.0 / pt->step[i];
    double numstep = 1 + (pt->end[i] - pt->begin[i]) / pt->step[i];
    pt->len[i] = (int)(numstep + 0.49);
    if (numstep - pt->len[i] >= 0.1 || pt->len[i] - numstep >= 0.1)
      imd_warning("numstep = %f rounded to %d", numstep, pt->len[i]);
    if (pt->len[i] > pt->maxsteps) pt->maxsteps = pt->len[i];
  }
  pt->table = calloc((size_t)nc * (pt->maxsteps + 2), sizeof(double));
  if (!pt->table) return 1;
  for (int i = 0; i < nc; i++)
    for (int k = 0; k < pt->len[i]; k++) {
      double val;
      if (fscanf(f, "%lf", &val) != 1) return 2;
      pt->table[(size_t)k * nc + i] = val;
    }
  if (radial)
    for (int i = 0; i < nc; i++) {
      double delta = pt->table[(size_t)(pt->len[i] - 1) * nc + i];
      if (delta != 0.0)
        for (int k = 0; k < pt->len[i]; k++)
          pt->table[(size_t)k * nc + i] -= delta;
    }
  return 0;
}

static int threepoint(pot_table_t *pt) {
  int nc = pt->ncols;
  for (int col = 0; col < nc; col++) {
    double *y = pt->table + col;
    int n = pt->len[col];
    if (n < 3) return 2;
    y[(size_t)n * nc] =
        3 * y[(size_t)(n - 1) * nc] - 3 * y[(size_t)(n - 2) * nc] + y[(size_t)(n - 3) * nc];
    y[(size_t)(n + 1) * nc] = 6 * y[(size_t)(n - 1) * nc] -
                              8 * y[(size_t)(n - 2) * nc] +
                              3 * y[(size_t)(n - 3) * nc];
  }
  return 0;
}

int read_pot_table(pot_table_t *pt, const char *filename, int ncols,
                   int radial) {
  char buf[1024];
  int have_header = 0, have_format = 0, end_header = 0;
  int format = have_eam ? 2 : 1, size = ncols, rc;
  FILE *f = fopen(filename, "r");
  if (!f) return 1;
  do {
    if (!fgets(buf, sizeof buf, f)) {
      fclose(f);
      return 2;
    }
    if (buf[0] == '#') {
      have_header = 1;
      end_header = (buf[1] == 'E');
      if (buf[1] == 'F') {
        if (sscanf(buf + 2, "%d%d", &format, &size) != 2 || size != ncols ||
            (format != 1 && format != 2)) {
          fclose(f);
          return 3;
        }
        have_format = 1;
      }
    } else if (have_header) {
      fclose(f);
      return 3;
    } else {
      end_header = 1;
    }
  } while (!end_header);
  if (have_header && !have_format) {
    fclose(f);
    return 3;
  }
  if (!have_header) {
    imd_warning("File %s has no header", filename);
    rewind(f);
  }
  if (alloc_info(pt, ncols)) {
    fclose(f);
    return 1;
  }
  rc = (format == 1) ? format1(pt, f, radial) : format2(pt, f, radial);
  fclose(f);
  if (!rc) rc = threepoint(pt);
  if (rc) free_pot_table(pt);
  return rc;
}

/* setup_potentials (src/imd_potential.c:43-131): pair table with ntypes^2
   columns, EAM embedding (ntypes columns) and density (ntypes^2), and with
   EEAM the energy term E(P) (ntypes columns, src/imd_potential.c:82-85). */
int setup_potentials(void) {
  int rc;
  cellsz = 0.0;
  if (have_potfile) {
    rc = read_pot_table(&pair_pot, potfilename, ntypes * ntypes, 1);
    if (rc) imd_error("cannot read potential table %s (%d)", potfilename, rc);
  }
  if (have_eam) {
    rc = read_pot_table(&embed_pot, eam2_emb_E_filename, ntypes, 0);
    if (rc)
      imd_error("cannot read embedding table %s (%d)", eam2_emb_E_filename, rc);
    rc = read_pot_table(&rho_h_tab, eam2_at_rho_filename, ntypes * ntypes, 1);
    if (rc)
      imd_error("cannot read density table %s (%d)", eam2_at_rho_filename, rc);
  }
  if (have_eeam) {
    if (!have_eam) imd_error("eeam_energy_file needs the EAM tables");
    rc = read_pot_table(&emod_pot, eeam_mod_E_filename, ntypes, 0);
    if (rc)
      imd_error("cannot read EEAM table %s (%d)", eeam_mod_E_filename, rc);
  }
  return 0;
}
