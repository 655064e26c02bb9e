/*
 * host_io.c -- configurations, velocity initialisation and thermo output.
 *
 * read_atoms: IMD's configuration format (src/imd_io_3d.c:44-520): header
 *   lines "#F A n t m np nv nd" (numbers of number/type/mass/position/
 *   velocity/data columns, src/imd_io.c:3369-3497), "#X/#Y/#Z" box vectors
 *   (src/imd_io.c:3318-3360) and "#E"; ASCII rows
 *   "number type [mass] x y z [vx vy vz] [data...]"; the type is taken
 *   modulo ntypes, the mass from the row or the masses key, p = m v, and no
 *   velocity columns request Maxwell initialisation.
 * generate_fcc: sites at (x + 0.5) a / 2 for integer x, y, z with even sum
 *   in box_param unit cells of edge box_unit (src/imd_generate.c:306-486).
 * maxwell: momenta N(0, sqrt(kT m)) per component from drand48 with the
 *   polar method, then the mean momentum removed (src/imd_maxwell.c:44-323),
 *   in IMD's cell traversal order: inner cells x-major, then y, then z
 *   (make_cell_lists), atoms of a cell in input order, cells from
 *   init_cells' grid and cell_coord after back_into_box, as read_atoms
 *   fills them (src/imd_io_3d.c:318-333, 425-442).  drand48 is seeded
 *   once, when the atoms are read or generated (src/imd_io_3d.c:87,
 *   src/imd_generate.c:80), so a later phase continues the stream.
 * write_eng_file: one line per eng_int steps with time, Epot per atom,
 *   temperature 2 Ekin / nactive, pressure T/(V/N) + virial/(3V), volume per
 *   atom, for NVT eta * tau_eta, and with the stress tensor on the six
 *   components of tot_presstens / volume (src/imd_io.c:2262-2431); the
 *   header is written when the file is started, rows are appended.
 * write_config(n, steps): the checkpoint <outfiles>.%05d.chkpt (n >= 0),
 *   <outfiles>-final.chkpt (n == -1) or <outfiles>-interm.chkpt (n == -2)
 *   with IMD's header (#F A 1 1 1 3 3 nd, #C, #X/#Y/#Z, ##PBC, ## lines,
 *   #E; src/imd_io.c:3152-3300) and rows "number type mass x y z vx vy vz
 *   Epot [eam_rho]" (src/imd_io_3d.c:696-948), then the iteration file
 *   (src/imd_io.c:171-181).  Numbers are written with 17 significant
 *   digits, so a restart reads back the same doubles.
 * write_config_press(n): <outfiles>.%05d.press with rows "number type mass
 *   x y z P_xx P_yy P_zz P_yz P_zx P_xy" (src/imd_io.c:1047-1240).
 * write_itr_file(n, steps, ""): <outfiles>.%05d.itr, -final.itr or
 *   -interm.itr with "# checkpoint n", startstep = steps + 1, the box
 *   vectors and, for NVT with tau_eta, starttemp and eta
 *   (src/imd_io_3d.c:949-1095); it is a parameter file.
 */
#define _XOPEN_SOURCE 600
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "imd_host.h"

long natoms = 0;
int nactive = 0;
double *atom_x = NULL, *atom_p = NULL, *atom_mass = NULL;
double *atom_epot = NULL, *atom_rho = NULL, *atom_press = NULL;
double *atom_eeam_p = NULL;
double tot_presstens[6];
int *atom_type = NULL;
long *atom_number = NULL;
double tot_pot_energy = 0.0, tot_kin_energy = 0.0, virial = 0.0, volume = 0.0;
int nbl_count = 0;

static FILE *eng_file = NULL;

static void write_box_header(FILE *f, int steps);

static int grow_atoms(long cap) {
  atom_x = realloc(atom_x, sizeof(double) * 3 * cap);
  atom_p = realloc(atom_p, sizeof(double) * 3 * cap);
  atom_mass = realloc(atom_mass, sizeof(double) * cap);
  atom_type = realloc(atom_type, sizeof(int) * cap);
  atom_number = realloc(atom_number, sizeof(long) * cap);
  atom_epot = realloc(atom_epot, sizeof(double) * cap);
  atom_rho = realloc(atom_rho, sizeof(double) * cap);
  atom_press = realloc(atom_press, sizeof(double) * 6 * cap);
  atom_eeam_p = realloc(atom_eeam_p, sizeof(double) * cap);
  return (atom_x && atom_p && atom_mass && atom_type && atom_number &&
          atom_epot && atom_rho && atom_press && atom_eeam_p) ? 0 : 1;
}

/* the volume of make_box (src/imd_geom_3d.c:52-104) */
void make_box(void) {
  double c[3] = {box_y[1] * box_z[2] - box_y[2] * box_z[1],
                 box_y[2] * box_z[0] - box_y[0] * box_z[2],
                 box_y[0] * box_z[1] - box_y[1] * box_z[0]};
  volume = fabs(box_x[0] * c[0] + box_x[1] * c[1] + box_x[2] * c[2]);
}

int read_atoms(const char *filename) {
  char line[4096];
  int have_format = 0, n_num = 1, n_type = 1, n_mass = 1, n_pos = 3,
      n_vel = 3, n_data = 0;
  char fmt = 'A';
  FILE *f = fopen(filename, "r");
  if (!f) imd_error("cannot open input file %s", filename);
  long pos_after_header = 0;
  while (fgets(line, sizeof line, f)) {
    if (line[0] != '#') break;
    if (line[1] == 'F') {
      char s[16];
      if (sscanf(line + 2, "%15s %d %d %d %d %d %d", s, &n_num, &n_type,
                 &n_mass, &n_pos, &n_vel, &n_data) < 7)
        imd_error("Format line of file %s corrupt", filename);
      fmt = s[0];
      have_format = 1;
    } else if (line[1] == 'X') {
      if (box_from_header || box_x[0] == 0.0)
        sscanf(line + 2, "%lf %lf %lf", &box_x[0], &box_x[1], &box_x[2]);
    } else if (line[1] == 'Y') {
      if (box_from_header || box_y[1] == 0.0)
        sscanf(line + 2, "%lf %lf %lf", &box_y[0], &box_y[1], &box_y[2]);
    } else if (line[1] == 'Z') {
      if (box_from_header || box_z[2] == 0.0)
        sscanf(line + 2, "%lf %lf %lf", &box_z[0], &box_z[1], &box_z[2]);
    } else if (line[1] == 'E') {
      pos_after_header = ftell(f);
      break;
    }
  }
  if (!have_format) imd_error("configuration %s has no #F header", filename);
  if (fmt != 'A')
    imd_error("configuration format %c not supported (ASCII 'A' only)", fmt);
  if (n_num != 1 || n_type != 1 || n_pos != 3)
    imd_error("configuration %s: need number, type and 3 positions", filename);
  fseek(f, pos_after_header, SEEK_SET);
  long cap = 1024;
  natoms = 0;
  if (grow_atoms(cap)) imd_error("cannot allocate atoms");
  int n_items = n_num + n_type + n_mass + n_pos + n_vel + n_data;
  int missing_vel = (n_vel == 0);
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    double d[32];
    int num, typ;
    char *s = line, *e;
    num = (int)strtol(s, &e, 10);
    if (e == s) continue; /* empty line */
    s = e;
    typ = (int)strtol(s, &e, 10);
    if (e == s) imd_error("incorrect line in configuration file.");
    s = e;
    int nd = 0;
    while (nd < 32) {
      double v = strtod(s, &e);
      if (e == s) break;
      d[nd++] = v;
      s = e;
    }
    int p = nd + 2;
    if (p < n_items && p != n_items - n_vel)
      imd_error("incorrect line in configuration file.");
    if (natoms == cap) {
      cap *= 2;
      if (grow_atoms(cap)) imd_error("cannot allocate atoms");
    }
    int k = 0;
    int t = ((typ % ntypes) + ntypes) % ntypes;
    double m = n_mass ? d[k++] : masses[t];
    if (!(m > 0)) imd_error("Mass zero or negative.");
    atom_number[natoms] = num;
    atom_type[natoms] = t;
    atom_mass[natoms] = m;
    for (int c = 0; c < 3; c++) atom_x[3 * natoms + c] = d[k++];
    if (n_vel && p == n_items) {
      for (int c = 0; c < 3; c++) atom_p[3 * natoms + c] = d[k++] * m;
    } else {
      for (int c = 0; c < 3; c++) atom_p[3 * natoms + c] = 0.0;
      missing_vel = 1;
    }
    natoms++;
  }
  fclose(f);
  if (missing_vel) do_maxwell = 1;
  nactive = (int)(3 * natoms);
  make_box();
  srand48(seed);
  return 0;
}

int generate_fcc(void) {
  int nx = box_param[0], ny = box_param[1], nz = box_param[2];
  if (nx <= 0 || ny <= 0 || nz <= 0) imd_error("box_param not set");
  double a = box_unit;
  memset(box_x, 0, sizeof box_x);
  memset(box_y, 0, sizeof box_y);
  memset(box_z, 0, sizeof box_z);
  box_x[0] = nx * a;
  box_y[1] = ny * a;
  box_z[2] = nz * a;
  long cap = 4L * nx * ny * nz;
  if (grow_atoms(cap)) imd_error("cannot allocate atoms");
  natoms = 0;
  for (int x = 0; x < 2 * nx; x++)
    for (int y = 0; y < 2 * ny; y++)
      for (int z = 0; z < 2 * nz; z++) {
        if ((x + y + z) % 2) continue;
        atom_x[3 * natoms] = (x + 0.5) * a / 2;
        atom_x[3 * natoms + 1] = (y + 0.5) * a / 2;
        atom_x[3 * natoms + 2] = (z + 0.5) * a / 2;
        atom_type[natoms] = 0;
        atom_mass[natoms] = masses[0];
        atom_number[natoms] = natoms;
        atom_p[3 * natoms] = atom_p[3 * natoms + 1] = atom_p[3 * natoms + 2] = 0;
        natoms++;
      }
  nactive = (int)(3 * natoms);
  make_box();
  srand48(seed);
  do_maxwell = 1;
  return 0;
}

/* polar (Box-Muller) method, Knuth vol. 2, 3rd ed., p. 122 */
static double gaussian(double sigma) {
  double x, y, r2;
  do {
    x = -1 + 2 * drand48();
    y = -1 + 2 * drand48();
    r2 = x * x + y * y;
  } while (r2 > 1.0 || r2 == 0);
  return sigma * y * sqrt(-2.0 * log(r2) / r2);
}

/* Atom indices in IMD's cell traversal order: the grid of init_cells
   (src/imd_geom_3d.c:113-170: the skin added to the tables' list radius,
   gdim = (int)(1 / sqrt(cellsz / height)) per direction), positions put
   back into the box along periodic directions (back_into_box), cell_coord
   (src/imd_geom_3d.c:1054-1074: (int)(gdim (x . tbox)), clamped), inner
   cells x-major, then y, then z, atoms of a cell in input order (a stable
   counting sort).  NULL when the box or the radius leaves no grid; the
   caller then keeps input order. */
static long *cell_order(void) {
  const double *b[3] = {box_x, box_y, box_z};
  double t[3][3], vol;
  for (int d = 0; d < 3; d++) {
    const double *u = b[(d + 1) % 3], *v = b[(d + 2) % 3];
    t[d][0] = u[1] * v[2] - u[2] * v[1];
    t[d][1] = u[2] * v[0] - u[0] * v[2];
    t[d][2] = u[0] * v[1] - u[1] * v[0];
  }
  vol = box_x[0] * t[0][0] + box_x[1] * t[0][1] + box_x[2] * t[0][2];
  if (vol == 0 || natoms == 0 || !(cellsz > 0)) return NULL;
  int g[3];
  double rl = sqrt(cellsz) + nbl_margin;
  for (int d = 0; d < 3; d++) {
    for (int e = 0; e < 3; e++) t[d][e] /= vol;
    double h = 1.0 / (t[d][0] * t[d][0] + t[d][1] * t[d][1] + t[d][2] * t[d][2]);
    g[d] = (int)(1.0 / sqrt(rl * rl / h));
    if (g[d] < 1) return NULL;
  }
  long ncell = (long)g[0] * g[1] * g[2];
  long *cell = malloc(sizeof(long) * natoms);
  long *first = calloc(ncell + 1, sizeof(long));
  long *order = malloc(sizeof(long) * natoms);
  if (!cell || !first || !order) {
    free(cell);
    free(first);
    free(order);
    return NULL;
  }
  for (long i = 0; i < natoms; i++) {
    double x[3] = {atom_x[3 * i], atom_x[3 * i + 1], atom_x[3 * i + 2]};
    for (int d = 0; d < 3; d++) {
      if (pbc_dirs[d] != 1) continue;
      double f = -floor(x[0] * t[d][0] + x[1] * t[d][1] + x[2] * t[d][2]);
      for (int e = 0; e < 3; e++) x[e] += f * b[d][e];
    }
    int c[3];
    for (int d = 0; d < 3; d++) {
      c[d] = (int)(g[d] * (x[0] * t[d][0] + x[1] * t[d][1] + x[2] * t[d][2]));
      if (c[d] < 0) c[d] = 0;
      if (c[d] > g[d] - 1) c[d] = g[d] - 1;
    }
    cell[i] = ((long)c[0] * g[1] + c[1]) * g[2] + c[2];
    first[cell[i] + 1]++;
  }
  for (long k = 0; k < ncell; k++) first[k + 1] += first[k];
  for (long i = 0; i < natoms; i++) order[first[cell[i]]++] = i;
  free(cell);
  free(first);
  return order;
}

void maxwell(double temp) {
  double tot[3] = {0, 0, 0};
  long *order = cell_order();
  for (long n = 0; n < natoms; n++) {
    long i = order ? order[n] : n;
    double s = sqrt(temp * atom_mass[i]);
    for (int c = 0; c < 3; c++) {
      atom_p[3 * i + c] = gaussian(s);
      tot[c] += atom_p[3 * i + c];
    }
  }
  free(order);
  for (int c = 0; c < 3; c++) tot[c] = natoms ? tot[c] / natoms : 0.0;
  for (long i = 0; i < natoms; i++)
    for (int c = 0; c < 3; c++) atom_p[3 * i + c] -= tot[c];
}

/* <outfiles>.%05d.<ext>, <outfiles>-final.<ext> or <outfiles>-interm.<ext> */
static void out_name(char *dst, size_t cap, int fzhlr, const char *ext) {
  if (fzhlr >= 0)
    snprintf(dst, cap, "%s.%05d.%s", outfilename, fzhlr, ext);
  else if (fzhlr == -1)
    snprintf(dst, cap, "%s-final.%s", outfilename, ext);
  else
    snprintf(dst, cap, "%s-interm.%s", outfilename, ext);
}

int write_itr_file(int fzhlr, int steps, const char *suffix) {
  char fname[IMD_STRLEN + 32];
  char ext[16];
  snprintf(ext, sizeof ext, "%sitr", suffix ? suffix : "");
  out_name(fname, sizeof fname, fzhlr, ext);
  FILE *out = fopen(fname, "w");
  if (!out) imd_error("Cannot write iteration file.");
  fprintf(out, "# checkpoint %d\n", fzhlr);
  fprintf(out, "startstep \t%d\n", steps + 1);
  fprintf(out, "box_x \t%.16f %.16f %.16f\n", box_x[0], box_x[1], box_x[2]);
  fprintf(out, "box_y \t%.16f %.16f %.16f\n", box_y[0], box_y[1], box_y[2]);
  fprintf(out, "box_z \t%.16f %.16f %.16f\n", box_z[0], box_z[1], box_z[2]);
  if (ensemble == IMDG_NVT && isq_tau_eta > 0) {
    fprintf(out, "starttemp \t%.17g\n", temperature);
    fprintf(out, "eta \t%.17g\n", eta);
  }
  fclose(out);
  return 0;
}

int write_config(int fzhlr, int steps) {
  char fname[IMD_STRLEN + 32];
  out_name(fname, sizeof fname, fzhlr, "chkpt");
  FILE *f = fopen(fname, "w");
  if (!f) imd_error("Cannot open output file %s", fname);
  int nd = 1 + (have_eam ? 1 : 0) + (have_eeam ? 1 : 0);
  fprintf(f, "#F A 1 1 1 3 3 %d\n", nd);
  fprintf(f, "#C number type mass x y z vx vy vz Epot%s%s\n",
          have_eam ? " eam_rho" : "", have_eeam ? " eam_p" : "");
  write_box_header(f, steps);
  for (long i = 0; i < natoms; i++) {
    double m = atom_mass[i];
    fprintf(f, "%ld %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g",
            atom_number[i], atom_type[i], m, atom_x[3 * i], atom_x[3 * i + 1],
            atom_x[3 * i + 2], atom_p[3 * i] / m, atom_p[3 * i + 1] / m,
            atom_p[3 * i + 2] / m, atom_epot[i]);
    if (have_eam) fprintf(f, " %.17g", atom_rho[i]);
    if (have_eeam) fprintf(f, " %.17g", atom_eeam_p[i]);
    fputc('\n', f);
  }
  if (fclose(f)) imd_error("Cannot write output file %s", fname);
  write_itr_file(fzhlr, steps, "");
  return 0;
}

/* the header lines after #F and #C of IMD's configuration-like files */
static void write_box_header(FILE *f, int steps) {
  time_t now = time(NULL);
  fprintf(f, "#X \t%.16e %.16e %.16e\n", box_x[0], box_x[1], box_x[2]);
  fprintf(f, "#Y \t%.16e %.16e %.16e\n", box_y[0], box_y[1], box_y[2]);
  fprintf(f, "#Z \t%.16e %.16e %.16e\n", box_z[0], box_z[1], box_z[2]);
  fprintf(f, "##PBC %d %d %d\n", pbc_dirs[0], pbc_dirs[1], pbc_dirs[2]);
  fprintf(f, "## Generated on %s", ctime(&now));
  fprintf(f, "## by %s after step %d\n", progname, steps);
  fprintf(f, "#E\n");
}

int write_config_press(int nr) {
  char fname[IMD_STRLEN + 32];
  snprintf(fname, sizeof fname, "%s.%05d.press", outfilename, nr);
  FILE *f = fopen(fname, "w");
  if (!f) imd_error("Cannot open output file %s", fname);
  fprintf(f, "#F A 1 1 1 3 0 6\n");
  fprintf(f, "#C number type mass x y z P_xx P_yy P_zz P_yz P_zx P_xy\n");
  write_box_header(f, nr * press_int);
  for (long i = 0; i < natoms; i++) {
    const double *P = atom_press + 6 * i;
    fprintf(f, "%ld %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g "
            "%.17g\n", atom_number[i], atom_type[i], atom_mass[i],
            atom_x[3 * i], atom_x[3 * i + 1], atom_x[3 * i + 2], P[0], P[1],
            P[2], P[3], P[4], P[5]);
  }
  if (fclose(f)) imd_error("Cannot write output file %s", fname);
  return 0;
}

int write_eng_file_header(void) {
  char fname[IMD_STRLEN + 8];
  snprintf(fname, sizeof fname, "%s.eng", outfilename);
  FILE *fl = fopen(fname, "w");
  if (!fl) imd_error("Cannot open properties file.");
  fprintf(fl, "# time Epot temperature pressure volume%s%s\n",
          ensemble == IMDG_NVT ? " eta*tau_eta" : "",
          stress_tens ? " Press_xx Press_yy Press_zz Press_yz Press_xz Press_xy"
                      : "");
  fclose(fl);
  return 0;
}

int write_eng_file(int steps) {
  if (!eng_file) {
    char fname[IMD_STRLEN + 8];
    snprintf(fname, sizeof fname, "%s.eng", outfilename);
    eng_file = fopen(fname, "a");
    if (!eng_file) imd_error("Cannot open properties file.");
  }
  double Epot = tot_pot_energy / natoms;
  double Temp = 2.0 * tot_kin_energy / nactive;
  double vol = volume / natoms;
  double pressure = Temp / vol + virial / (3 * volume);
  fprintf(eng_file, "%e %.18e %e %e %e", steps * timestep, Epot, Temp,
          pressure, vol);
  if (ensemble == IMDG_NVT) fprintf(eng_file, " %e", eta * tau_eta);
  if (stress_tens)
    for (int k = 0; k < 6; k++)
      fprintf(eng_file, " %e", tot_presstens[k] / volume);
  fprintf(eng_file, "\n");
  fflush(eng_file);
  return 0;
}
