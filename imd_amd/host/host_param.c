/*
 * host_param.c -- IMD parameter (tag) files for the force path.
 *
 * Syntax as getparamfile (src/imd_param.c:251-3520): one "key value..."
 * per line, '#' starts a comment, separators are blanks, tabs and '=';
 * "simulation n" opens phase n and reading stops at the first later phase
 * (src/imd_param.c:304-318).  Unknown keys are reported and ignored
 * (src/imd_param.c:3507-3509).  Required keys are checked like
 * check_parameters_complete (src/imd_param.c:3525-3547).
 *
 * read_parameters(file, phase) follows src/imd_param.c:3880-3979: a later
 * phase first writes <outfiles>-final.itr, re-reads the parameter file
 * through its own section and then that iteration file; phase 1 reads the
 * itrname file (keeping startstep); with -r n the iteration file
 * <outfiles>.n.itr (or .%05d.itr) is read and the configuration becomes
 * <outfiles>.n.chkpt (or .%05d.chkpt); a fresh phase 1 deletes the .eng
 * file it will append to.  read_command_line takes -p file and -r n, both
 * also written without the blank (src/imd_param.c:3829-3866).
 */
#define _GNU_SOURCE
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <unistd.h>

#include "imd_host.h"

int ensemble = -1;
int ntypes = 0;
double masses[IMD_MAXTYPES];
double box_x[3], box_y[3], box_z[3];
int box_param[3];
double box_unit = 1.0;
int pbc_dirs[3] = {1, 1, 1};
double timestep = 0.0;
double temperature = 0.0;
double eta = 0.0, tau_eta = 0.0, isq_tau_eta = 0.0;
double nbl_margin = 0.4, nbl_size = 1.1; /* src/globals.h:419-423 */
int steps_min = 0, steps_max = 0, eng_int = 0, checkpt_int = 0;
int watch_int = 0, stop_int = 0;
int press_int = 0, stress_tens = 0;
int imdrestart = 0, simulation = 1;
int do_maxwell = 0, box_from_header = 0;
long seed = 0;
double maxwalltime = 0.0;
char infilename[IMD_STRLEN], outfilename[IMD_STRLEN] = "imd";
char itrfilename[IMD_STRLEN], paramfilename[IMD_STRLEN];
char progname[IMD_STRLEN] = "imd_gpu";
char potfilename[IMD_STRLEN], eam2_emb_E_filename[IMD_STRLEN];
char eam2_at_rho_filename[IMD_STRLEN], eeam_mod_E_filename[IMD_STRLEN];
int gpu_precision = IMDG_MIXED, gpu_device = 0;
int gpu_lanes = 0, gpu_autotune = 0;
int interpolation = IMDG_INTERP_3POINT;
int have_potfile = 0, have_eam = 0, have_eeam = 0;

void imd_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "Error: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
  fflush(stderr);
  exit(2);
}

void imd_warning(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "WARNING: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
}

void init_parameters(void) {
  ensemble = -1;
  ntypes = 0;
  memset(masses, 0, sizeof masses);
  memset(box_x, 0, sizeof box_x);
  memset(box_y, 0, sizeof box_y);
  memset(box_z, 0, sizeof box_z);
  memset(box_param, 0, sizeof box_param);
  box_unit = 1.0;
  pbc_dirs[0] = pbc_dirs[1] = pbc_dirs[2] = 1;
  timestep = 0.0;
  temperature = 0.0;
  eta = tau_eta = isq_tau_eta = 0.0;
  nbl_margin = 0.4;
  nbl_size = 1.1;
  steps_min = steps_max = eng_int = checkpt_int = 0;
  watch_int = stop_int = 0;
  press_int = stress_tens = 0;
  do_maxwell = box_from_header = 0;
  seed = 0;
  maxwalltime = 0.0;
  infilename[0] = potfilename[0] = eam2_emb_E_filename[0] = 0;
  itrfilename[0] = 0;
  eam2_at_rho_filename[0] = eeam_mod_E_filename[0] = 0;
  strcpy(outfilename, "imd");
  gpu_precision = IMDG_MIXED;
  gpu_device = 0;
  gpu_lanes = 0;
  gpu_autotune = 0;
  interpolation = IMDG_INTERP_3POINT;
  have_potfile = have_eam = have_eeam = 0;
}

static int curline;

static char *next_tok(void) { return strtok(NULL, " =\t\r\n"); }

static void get_str(const char *key, char *dst) {
  char *t = next_tok();
  if (!t) imd_error("parameter for %s missing in line %d", key, curline);
  strncpy(dst, t, IMD_STRLEN - 1);
  dst[IMD_STRLEN - 1] = 0;
}

static void get_ints(const char *key, int *dst, int n) {
  for (int i = 0; i < n; i++) {
    char *t = next_tok();
    if (!t)
      imd_error("parameter for %s missing in line %d\ninteger vector of dim %d "
                "expected", key, curline, n);
    dst[i] = atoi(t);
  }
}

static void get_reals(const char *key, double *dst, int n) {
  for (int i = 0; i < n; i++) {
    char *t = next_tok();
    if (!t)
      imd_error("parameter for %s missing in line %d\nreal vector of dim %d "
                "expected", key, curline, n);
    dst[i] = atof(t);
  }
}

/* one parameter (or iteration) file through phase `phase`; returns 1 when
   the end of the file was reached (no later phase follows) */
static int getparamfile(const char *paramfile, int phase) {
  char buf[1024];
  int finished = 1;
  FILE *pf = fopen(paramfile, "r");
  if (!pf) imd_error("Could not open parameter file %s", paramfile);
  curline = 0;
  while (fgets(buf, sizeof buf, pf)) {
    curline++;
    char *c = strchr(buf, '#');
    if (c) *c = 0;
    char *key = strtok(buf, " =\t\r\n");
    if (!key) continue;
    if (!strcasecmp(key, "simulation")) {
      int ph;
      get_ints(key, &ph, 1);
      if (phase < ph) {
        finished = 0;
        break;
      }
    } else if (!strcasecmp(key, "ensemble")) {
      char s[IMD_STRLEN];
      get_str(key, s);
      if (!strcasecmp(s, "nve"))
        ensemble = IMDG_NVE;
      else if (!strcasecmp(s, "nvt"))
        ensemble = IMDG_NVT;
      else
        imd_error("ensemble %s is not supported by the GPU force path", s);
    } else if (!strcasecmp(key, "maxsteps")) {
      get_ints(key, &steps_max, 1);
    } else if (!strcasecmp(key, "startstep")) {
      get_ints(key, &steps_min, 1);
    } else if (!strcasecmp(key, "timestep")) {
      get_reals(key, &timestep, 1);
    } else if (!strcasecmp(key, "checkpt_int")) {
      get_ints(key, &checkpt_int, 1);
    } else if (!strcasecmp(key, "eng_int")) {
      get_ints(key, &eng_int, 1);
    } else if (!strcasecmp(key, "watch_int")) {
      get_ints(key, &watch_int, 1);
    } else if (!strcasecmp(key, "stop_int")) {
      get_ints(key, &stop_int, 1);
    } else if (!strcasecmp(key, "press_int")) {
      /* .press files of the per-atom stress tensor (src/imd_param.c:2224) */
      get_ints(key, &press_int, 1);
    } else if (!strcasecmp(key, "stress_tens")) {
      /* what IMD's STRESS_TENS build option turns on: the stress tensor on
         output steps and its six columns in the .eng file */
      get_ints(key, &stress_tens, 1);
    } else if (!strcasecmp(key, "itrname")) {
      get_str(key, itrfilename);
    } else if (!strcasecmp(key, "coordname")) {
      get_str(key, infilename);
    } else if (!strcasecmp(key, "outfiles")) {
      get_str(key, outfilename);
    } else if (!strcasecmp(key, "potfile") ||
               !strcasecmp(key, "core_potential_file")) {
      get_str(key, potfilename);
      have_potfile = 1;
    } else if (!strcasecmp(key, "embedding_energy_file")) {
      get_str(key, eam2_emb_E_filename);
      have_eam = 1;
    } else if (!strcasecmp(key, "atomic_e-density_file")) {
      get_str(key, eam2_at_rho_filename);
      have_eam = 1;
    } else if (!strcasecmp(key, "eeam_energy_file")) {
      /* EEAM energy term E(P) (src/imd_param.c:2615-2619) */
      get_str(key, eeam_mod_E_filename);
      have_eeam = 1;
    } else if (!strcasecmp(key, "ntypes")) {
      int nt;
      get_ints(key, &nt, 1);
      if (ntypes && nt != ntypes)
        imd_error("ntypes must be constant during a simulation");
      if (nt < 1 || nt > IMD_MAXTYPES)
        imd_error("ntypes must be between 1 and %d", IMD_MAXTYPES);
      ntypes = nt;
    } else if (!strcasecmp(key, "total_types")) {
      int tt;
      get_ints(key, &tt, 1);
    } else if (!strcasecmp(key, "masses")) {
      if (ntypes == 0)
        imd_error("specify parameter ntypes before parameter masses");
      get_reals(key, masses, ntypes);
    } else if (!strcasecmp(key, "pbc_dirs")) {
      get_ints(key, pbc_dirs, 3);
    } else if (!strcasecmp(key, "box_x")) {
      get_reals(key, box_x, 3);
    } else if (!strcasecmp(key, "box_y")) {
      get_reals(key, box_y, 3);
    } else if (!strcasecmp(key, "box_z")) {
      get_reals(key, box_z, 3);
    } else if (!strcasecmp(key, "box_param")) {
      get_ints(key, box_param, 3);
    } else if (!strcasecmp(key, "box_unit")) {
      get_reals(key, &box_unit, 1);
    } else if (!strcasecmp(key, "box_from_header")) {
      get_ints(key, &box_from_header, 1);
    } else if (!strcasecmp(key, "starttemp")) {
      get_reals(key, &temperature, 1);
    } else if (!strcasecmp(key, "do_maxwell")) {
      get_ints(key, &do_maxwell, 1);
    } else if (!strcasecmp(key, "seed")) {
      int s;
      get_ints(key, &s, 1);
      seed = s;
      if (seed > 0) seed = -seed;
    } else if (!strcasecmp(key, "eta")) {
      get_reals(key, &eta, 1);
    } else if (!strcasecmp(key, "tau_eta")) {
      get_reals(key, &tau_eta, 1);
      if (tau_eta == 0.0) imd_error("tau_eta is zero.");
      isq_tau_eta = 1.0 / (tau_eta * tau_eta);
    } else if (!strcasecmp(key, "isq_tau_eta")) {
      get_reals(key, &isq_tau_eta, 1);
      if (isq_tau_eta == 0.0) tau_eta = 0.0;
    } else if (!strcasecmp(key, "nbl_margin")) {
      get_reals(key, &nbl_margin, 1);
    } else if (!strcasecmp(key, "nbl_size")) {
      get_reals(key, &nbl_size, 1);
    } else if (!strcasecmp(key, "maxwalltime")) {
      get_reals(key, &maxwalltime, 1);
    } else if (!strcasecmp(key, "gpu_precision")) {
      char s[IMD_STRLEN];
      get_str(key, s);
      if (!strcasecmp(s, "mixed"))
        gpu_precision = IMDG_MIXED;
      else if (!strcasecmp(s, "fp64"))
        gpu_precision = IMDG_FP64;
      else
        imd_error("gpu_precision must be mixed or fp64");
    } else if (!strcasecmp(key, "gpu_device")) {
      get_ints(key, &gpu_device, 1);
    } else if (!strcasecmp(key, "gpu_lanes")) {
      /* lanes per atom of the neighbour loops (1, 2 or 4; 0: default) */
      get_ints(key, &gpu_lanes, 1);
      if (gpu_lanes != 0 && gpu_lanes != 1 && gpu_lanes != 2 && gpu_lanes != 4)
        imd_error("gpu_lanes must be 0, 1, 2 or 4");
    } else if (!strcasecmp(key, "interpolation")) {
      /* what IMD's FOURPOINT / SPLINE build options select
         (src/potaccess.h:24-36): threepoint (default), fourpoint, spline */
      char v[IMD_STRLEN];
      get_str(key, v);
      if (!strcasecmp(v, "threepoint"))
        interpolation = IMDG_INTERP_3POINT;
      else if (!strcasecmp(v, "fourpoint"))
        interpolation = IMDG_INTERP_4POINT;
      else if (!strcasecmp(v, "spline"))
        interpolation = IMDG_INTERP_SPLINE;
      else
        imd_error("interpolation must be threepoint, fourpoint or spline");
    } else if (!strcasecmp(key, "gpu_autotune")) {
      /* time every launch shape once before the main loop (imdg_autotune) */
      get_ints(key, &gpu_autotune, 1);
    } else if (!strcasecmp(key, "binary_output") ||
               !strcasecmp(key, "debug_potential")) {
      /* accepted; the output and steering they control is not part of
         the force path */
    } else {
      imd_warning("****** Unknown TAG %s ignored ******", key);
    }
  }
  fclose(pf);
  return finished;
}

/* the first of "<out>.<n>.<ext>" and "<out>.%05d.<ext>" that exists */
static void numbered_file(char *dst, int n, const char *ext) {
  snprintf(dst, IMD_STRLEN, "%s.%d.%s", outfilename, n, ext);
  FILE *f = fopen(dst, "r");
  if (!f) {
    snprintf(dst, IMD_STRLEN, "%s.%05d.%s", outfilename, n, ext);
    f = fopen(dst, "r");
    if (!f) imd_error("file %s not found", dst);
  }
  fclose(f);
}

int read_parameters(const char *paramfile, int phase) {
  char fname[IMD_STRLEN];
  /* the iteration file that carries the state into the next phase */
  if (phase > 1) write_itr_file(-1, steps_max, "");
  int finished = getparamfile(paramfile, phase);
  if (phase == 1 && itrfilename[0]) {
    int keep = steps_min;
    getparamfile(itrfilename, 1);
    steps_min = keep;
  }
  if (phase > 1) {
    snprintf(itrfilename, sizeof itrfilename, "%s-final.itr", outfilename);
    getparamfile(itrfilename, 1);
  }
  check_parameters_complete();
  if (imdrestart) {
    numbered_file(fname, imdrestart, "itr");
    getparamfile(fname, 1);
    numbered_file(infilename, imdrestart, "chkpt");
    printf("Restarting from %s.\n", infilename);
  } else if (phase == 1) {
    /* a fresh run: the .eng file is appended to, so start it anew */
    snprintf(fname, sizeof fname, "%s.eng", outfilename);
    unlink(fname);
  }
  return finished;
}

static void usage(void) {
  fprintf(stderr, "usage: %s [-r nnn] [-p paramfile]\n", progname);
}

int read_command_line(int argc, char **argv) {
  strncpy(progname, argv[0], IMD_STRLEN - 1);
  while (argc > 1 && argv[1][0] == '-') {
    char opt = argv[1][1];
    const char *val = NULL;
    if (opt == 'r' || opt == 'p') {
      if (argv[1][2] == 0) {
        if (argc > 2) {
          val = argv[2];
          argc--;
          argv++;
        }
      } else {
        val = argv[1] + 2;
      }
      if (val && opt == 'r') imdrestart = atoi(val);
      if (val && opt == 'p') {
        strncpy(paramfilename, val, IMD_STRLEN - 1);
        paramfilename[IMD_STRLEN - 1] = 0;
      }
    } else {
      printf("Illegal option %s \n", argv[1]);
      usage();
      exit(-1);
    }
    argv++;
    argc--;
  }
  if (!paramfilename[0]) {
    usage();
    exit(2);
  }
  return 0;
}

int check_parameters_complete(void) {
  if (ensemble < 0) imd_error("missing or unknown ensemble parameter.");
  if (timestep == 0.0) imd_error("timestep is missing or zero.");
  if (ntypes == 0) imd_error("ntypes is missing or zero.");
  if (!have_potfile && !have_eam) imd_error("no potential files given");
  if (have_eam && (!eam2_emb_E_filename[0] || !eam2_at_rho_filename[0]))
    imd_error("EAM needs embedding_energy_file and atomic_e-density_file");
  if (ensemble == IMDG_NVT && isq_tau_eta == 0.0)
    imd_warning("NVT without tau_eta: eta stays constant");
  if (press_int > 0) stress_tens = 1;
  for (int t = 0; t < ntypes; t++)
    if (!(masses[t] > 0.0) && !infilename[0])
      imd_error("masses are missing");
  return 0;
}
