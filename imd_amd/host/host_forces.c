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
 * calc_forces, move_atoms, write_eng_file every eng_int steps,
 * check_nblist) with the steps between two outputs fused on the device.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "imd_host.h"

static imdg_ctx *ctx = NULL;

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

int imd_gpu_init(void) {
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
  move_atoms = (ensemble == IMDG_NVT) ? move_atoms_nvt : move_atoms_nve;
  return 0;
}

void imd_gpu_finish(void) {
  if (ctx) imdg_ctx_destroy(ctx);
  ctx = NULL;
}

static void pull_thermo(void) {
  imdg_thermo th;
  check(imdg_get_thermo(ctx, 0, &th), "imdg_get_thermo");
  tot_pot_energy = th.epot;
  virial = th.virial;
  eta = th.eta;
  if (th.step > 0) tot_kin_energy = th.ekin;
  imdg_info info;
  check(imdg_get_info(ctx, &info), "imdg_get_info");
  nbl_count = (int)info.nbl_count;
}

void calc_forces(int steps) {
  check(imdg_calc_forces(ctx, steps), "calc_forces");
  pull_thermo();
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
  pull_thermo();
}

void move_atoms_nvt(void) {
  check(imdg_move_atoms(ctx), "move_atoms");
  pull_thermo();
}

void fetch_atoms(void) {
  check(imdg_get_atoms(ctx, atom_x, atom_p, NULL, NULL, NULL, NULL),
        "imdg_get_atoms");
}

static double wall(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

int main_loop(void) {
  const char *unfused = getenv("IMD_GPU_UNFUSED");
  long total = (long)steps_max - steps_min + 1;
  double t0 = wall();
  if (total <= 0) return 0;
  if (unfused && atoi(unfused)) {
    /* IMD's own call sequence, one entry point at a time */
    for (int steps = steps_min; steps <= steps_max; steps++) {
      calc_forces(steps);
      move_atoms();
      if (eng_int > 0 && steps % eng_int == 0) write_eng_file(steps);
      if (checkpt_int > 0 && steps % checkpt_int == 0) {
        char fn[IMD_STRLEN + 32];
        fetch_atoms();
        snprintf(fn, sizeof fn, "%s.%d.chkpt", outfilename, steps / checkpt_int);
        write_config(fn);
      }
      check_nblist();
    }
  } else {
    imdg_thermo *hist = malloc(sizeof(imdg_thermo) * 2048);
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
      if (eng_int > 0) {
        int ne = ((steps + eng_int - 1) / eng_int) * eng_int;
        if (ne < last) last = ne;
      }
      if (checkpt_int > 0) {
        int nc = ((steps + checkpt_int - 1) / checkpt_int) * checkpt_int;
        if (nc < last) last = nc;
      }
      if (last - steps + 1 > 2048) last = steps + 2047;
      long n = last - steps + 1;
      check(imdg_run(ctx, n), "imdg_run");
      check(imdg_sync(ctx), "imdg_sync");
      check(imdg_get_thermo_history(ctx, done, n, hist), "thermo history");
      for (long k = 0; k < n; k++) {
        int s = steps + (int)k;
        if (eng_int > 0 && s % eng_int == 0) {
          tot_pot_energy = hist[k].epot;
          tot_kin_energy = hist[k].ekin;
          virial = hist[k].virial;
          eta = hist[k].eta;
          write_eng_file(s);
        }
      }
      tot_pot_energy = hist[n - 1].epot;
      tot_kin_energy = hist[n - 1].ekin;
      virial = hist[n - 1].virial;
      eta = hist[n - 1].eta;
      if (checkpt_int > 0 && last % checkpt_int == 0) {
        char fn[IMD_STRLEN + 32];
        fetch_atoms();
        snprintf(fn, sizeof fn, "%s.%d.chkpt", outfilename, last / checkpt_int);
        write_config(fn);
      }
      done += n;
      steps = last + 1;
      if (maxwalltime > 0 && wall() - t0 > maxwalltime) {
        imd_warning("maxwalltime reached after step %d", last);
        break;
      }
    }
    free(hist);
  }
  double el = wall() - t0;
  imdg_info info;
  check(imdg_get_info(ctx, &info), "imdg_get_info");
  nbl_count = (int)info.nbl_count;
  printf("Did %ld steps in %.3f s: %.4e atom-steps/s, %ld list builds\n",
         total, el, natoms * (double)total / el, (long)nbl_count);
  return 0;
}
