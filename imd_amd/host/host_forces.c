/*
 * host_forces.c -- IMD's force-module entry points on the MI355X library.
 *
 * calc_forces(steps), make_nblist(), check_nblist(), deallocate_nblist()
 * (src/prototypes.h:173, 196-198) and move_atoms_nve / move_atoms_nvt
 * behind the move_atoms pointer chosen by the ensemble key
 * (src/globals.h:1444, src/imd_param.c:377-391) keep their names and their
 * effect on the globals (tot_pot_energy, virial, tot_kin_energy, eta,
 * nbl_count); the work runs in libimdg through include/imdg.h.
 * main_loop runs IMD's step sequence (src/imd_main_3d.c:155-870:
 * calc_forces, move_atoms, then per step the checkpoint every checkpt_int
 * steps, the .eng row every eng_int steps, the .press file every press_int
 * steps, the "write" file test every
 * watch_int steps, the "stop" file test every stop_int steps and the
 * maxwalltime test (src/imd_main_3d.c:690-815, 1892-1965), check_nblist)
 * with the steps between two such output steps fused on the device; the
 * wall time is tested at the end of each fused run (at most 2048 steps).
 * A phase ends by writing <outfiles>-final.itr (src/imd_main_3d.c:877-881).
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "imd_host.h"

static imdg_ctx *ctx = NULL;
static double t_start = -1.0;
double main_loop_seconds = 0.0;

void (*move_atoms)(void) = NULL;

imdg_ctx *imd_gpu_context(void) { return ctx; }

static void check(int rc, const char *what) {
  if (rc != IMDG_OK)
    imd_error("%s failed (%d): %s", what, rc,
              ctx ? imdg_last_error(ctx) : "no context");
}

static imdg_table as_imdg(const pot_table_t *pt) {
  imdg_table t;
  t.ncols = pt->ncols;
  t.maxsteps = pt->maxsteps;
  t.begin = pt->begin;
  t.end = pt->end;
  t.step = pt->step;
  t.invstep = pt->invstep;
  t.len = pt->len;
  t.table = pt->table;
  return t;
}

static double wall(void);

int imd_gpu_init(void) {
  if (t_start < 0) t_start = wall();
  imdg_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.precision = gpu_precision;
  cfg.ensemble = ensemble;
  cfg.ntypes = ntypes;
  cfg.timestep = timestep;
  cfg.nbl_margin = nbl_margin;
  cfg.nbl_size = nbl_size;
  cfg.device = gpu_device;
  cfg.lanes_per_atom = gpu_lanes;
  cfg.interpolation = interpolation;
  check(imdg_ctx_create(&cfg, &ctx), "imdg_ctx_create");
  imdg_table tp, tr, te;
  if (have_potfile) tp = as_imdg(&pair_pot);
  if (have_eam) {
    tr = as_imdg(&rho_h_tab);
    te = as_imdg(&embed_pot);
  }
  check(imdg_set_tables(ctx, have_potfile ? &tp : NULL, have_eam ? &tr : NULL,
                        have_eam ? &te : NULL),
        "imdg_set_tables");
  if (have_eeam) {
    imdg_table tm = as_imdg(&emod_pot);
    check(imdg_set_eeam_table(ctx, &tm), "imdg_set_eeam_table");
  }
  imdg_replica rep;
  memset(&rep, 0, sizeof rep);
  memcpy(rep.box_x, box_x, sizeof box_x);
  memcpy(rep.box_y, box_y, sizeof box_y);
  memcpy(rep.box_z, box_z, sizeof box_z);
  memcpy(rep.pbc, pbc_dirs, sizeof pbc_dirs);
  rep.temperature = temperature;
  rep.eta = eta;
  rep.isq_tau_eta = isq_tau_eta;
  int n = (int)natoms;
  check(imdg_set_atoms(ctx, 1, &rep, &n, atom_x, atom_p, atom_type, atom_mass),
        "imdg_set_atoms");
  if (gpu_autotune) {
    /* builds the list, times the launch shapes, leaves the state as is */
    imdg_tuning t;
    check(imdg_autotune(ctx, 3, &t), "imdg_autotune");
    printf("gpu_autotune: density %d lanes x %d threads, force %d lanes x %d "
           "threads\n", t.density_lanes, t.density_threads, t.force_lanes,
           t.force_threads);
  }
  if (stress_tens) check(imdg_set_stress(ctx, 1), "imdg_set_stress");
  move_atoms = (ensemble == IMDG_NVT) ? move_atoms_nvt : move_atoms_nve;
  return 0;
}

void imd_gpu_finish(void) {
  if (ctx) imdg_ctx_destroy(ctx);
  ctx = NULL;
}

void imd_gpu_reconfigure(void) {
  /* a new phase may change the ensemble, time step, thermostat, list
     parameters or box: carry the atoms over into a new context */
  fetch_atoms();
  /* main_loop's Maxwell step (src/imd_main_3d.c:87-91): not on a restart */
  if (!imdrestart && do_maxwell) maxwell(temperature);
  do_maxwell = 0;
  make_box();
  imd_gpu_finish();
  imd_gpu_init();
}

/* IMD's warnings after a force evaluation with a pair closer than a
   table's begin (src/imd_forces_nbl.c:982, 1323): is_short stays set from
   the pair loop into the EAM loop there, so an EAM system prints both */
static void warn_short(const imdg_info *info, int steps) {
  if (!info->is_short) return;
  fprintf(stderr, "Short distance, pair, step %d!\n", steps);
  if (have_eam) fprintf(stderr, "\n Short distance, EAM, step %d!\n", steps);
}

static void pull_thermo(int steps) {
  imdg_thermo th;
  check(imdg_get_thermo(ctx, 0, &th), "imdg_get_thermo");
  tot_pot_energy = th.epot;
  virial = th.virial;
  eta = th.eta;
  if (th.step > 0) tot_kin_energy = th.ekin;
  imdg_info info;
  check(imdg_get_info(ctx, &info), "imdg_get_info");
  nbl_count = (int)info.nbl_count;
  if (steps >= 0) warn_short(&info, steps);
}

void calc_forces(int steps) {
  check(imdg_calc_forces(ctx, steps), "calc_forces");
  pull_thermo(steps);
}

void make_nblist(void) { check(imdg_make_nblist(ctx), "make_nblist"); }

void check_nblist(void) {
  check(imdg_check_nblist(ctx, NULL), "check_nblist");
}

void deallocate_nblist(void) {
  check(imdg_deallocate_nblist(ctx), "deallocate_nblist");
}

void move_atoms_nve(void) {
  check(imdg_move_atoms(ctx), "move_atoms");
  pull_thermo(-1);
}

void move_atoms_nvt(void) {
  check(imdg_move_atoms(ctx), "move_atoms");
  pull_thermo(-1);
}

void fetch_atoms(void) {
  check(imdg_get_atoms(ctx, atom_x, atom_p, NULL, atom_epot,
                       have_eam ? atom_rho : NULL, NULL),
        "imdg_get_atoms");
  if (have_eeam)
    check(imdg_get_eeam(ctx, atom_eeam_p, NULL), "imdg_get_eeam");
}

static double wall(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* empty file "write": write <outfiles>-interm.chkpt (src/imd_main_3d.c:1892-1912) */
static void check_write(int steps) {
  FILE *t = fopen("write", "r");
  if (!t) return;
  int empty = fgetc(t) == EOF;
  fclose(t);
  if (!empty) return;
  remove("write");
  printf("Write file found after %d steps\n", steps);
  fetch_atoms();
  write_config(-2, steps);
}

/* file "stop": write <outfiles>-final.chkpt and end (src/imd_main_3d.c:1918-1940) */
static int check_stop(int steps) {
  FILE *t = fopen("stop", "r");
  if (!t) return 0;
  fclose(t);
  unlink("stop");
  printf("Stop file found after %d steps\n", steps);
  fetch_atoms();
  write_config(-1, steps);
  steps_max = steps;
  return 1;
}

/* src/imd_main_3d.c:1946-1965 */
static int check_walltime(int steps) {
  if (wall() - t_start <= maxwalltime) return 0;
  printf("Maximal allowed walltime reached after %d steps\n", steps);
  fetch_atoms();
  write_config(-1, steps);
  steps_max = steps;
  return 1;
}

/* the outputs and tests at the end of step `steps`, in IMD's order, with
   tot_pot_energy, tot_kin_energy, virial and eta of that step; returns 1
   when the run is to end */
static int step_outputs(int steps) {
  int eng = eng_int > 0 && steps % eng_int == 0;
  int press = press_int > 0 && steps % press_int == 0;
  if (checkpt_int > 0 && steps % checkpt_int == 0) {
    fetch_atoms();
    write_config(steps / checkpt_int, steps);
  }
  if (stress_tens && eng)
    check(imdg_get_stress(ctx, NULL, tot_presstens), "imdg_get_stress");
  if (eng) write_eng_file(steps);
  if (press) {
    /* src/imd_main_3d.c:733-738 */
    fetch_atoms();
    check(imdg_get_stress(ctx, atom_press, tot_presstens), "imdg_get_stress");
    write_config_press(steps / press_int);
  }
  if (watch_int > 0 && steps % watch_int == 0) check_write(steps);
  if (stop_int > 0 && steps % stop_int == 0 && check_stop(steps)) return 1;
  if (maxwalltime > 0 && check_walltime(steps)) return 1;
  return 0;
}

static int next_multiple(int steps, int every, int last) {
  if (every <= 0) return last;
  int m = ((steps + every - 1) / every) * every;
  return m < last ? m : last;
}

int main_loop(void) {
  const char *unfused = getenv("IMD_GPU_UNFUSED");
  int finished = 0;
  long done_steps = 0;
  double t0 = wall();
  if (unfused && atoi(unfused)) {
    /* IMD's own call sequence, one entry point at a time */
    for (int steps = steps_min; steps <= steps_max; steps++) {
      calc_forces(steps);
      move_atoms();
      done_steps++;
      if (step_outputs(steps)) {
        finished = 1;
        break;
      }
      check_nblist();
    }
  } else {
    imdg_thermo hist[1];
    int steps = steps_min;
    long done = 0;
    {
      imdg_info info;
      check(imdg_get_info(ctx, &info), "imdg_get_info");
      done = info.steps_done;
    }
    while (steps <= steps_max) {
      /* run through the next output step, at most 2048 steps at a time */
      int last = steps_max;
      last = next_multiple(steps, eng_int, last);
      last = next_multiple(steps, checkpt_int, last);
      last = next_multiple(steps, watch_int, last);
      last = next_multiple(steps, stop_int, last);
      last = next_multiple(steps, press_int, last);
      if (last - steps + 1 > 2048) last = steps + 2047;
      long n = last - steps + 1;
      check(imdg_run(ctx, n), "imdg_run");
      check(imdg_sync(ctx), "imdg_sync");
      /* every output interval ends a fused run: the outputs all belong to
         its last step */
      check(imdg_get_thermo_history(ctx, done + n - 1, 1, hist),
            "thermo history");
      tot_pot_energy = hist[0].epot;
      tot_kin_energy = hist[0].ekin;
      virial = hist[0].virial;
      eta = hist[0].eta;
      {
        /* the device flags a short distance per fused run: the warning
           names the run's last step */
        imdg_info info;
        check(imdg_get_info(ctx, &info), "imdg_get_info");
        warn_short(&info, last);
      }
      if (step_outputs(last)) finished = 1;
      done += n;
      done_steps += n;
      steps = last + 1;
      if (finished) break;
    }
  }
  double el = wall() - t0;
  main_loop_seconds += el;
  imdg_info info;
  check(imdg_get_info(ctx, &info), "imdg_get_info");
  nbl_count = (int)info.nbl_count;
  printf("Did %ld steps in %.3f s: %.4e atom-steps/s, %ld list builds\n",
         done_steps, el, el > 0 ? natoms * (double)done_steps / el : 0.0,
         (long)nbl_count);
  /* the phase is done: the restart flag is cleared and the state is kept
     for the next phase */
  imdrestart = 0;
  write_itr_file(-1, steps_max, "");
  printf("End of simulation %d\n", simulation);
  return finished;
}
