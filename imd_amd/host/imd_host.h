/*
 * imd_host.h -- the C host of the MI355X neighbour-list path.
 *
 * It keeps IMD's process-global program model for the force module: the
 * same parameter-file keys (src/imd_param.c:251-3520, subset below), the
 * same potential-table formats read with read_pot_table semantics
 * (src/imd_potential.c:161-462, src/imd_potential.c:1256-1272), the same
 * configuration format (src/imd_io_3d.c:44-520), and the same entry points
 * calc_forces(int), make_nblist(), check_nblist(), deallocate_nblist() and
 * the move_atoms function pointer (src/prototypes.h:173, 196-198;
 * src/globals.h:1444).  Instead of the CPU loops of src/imd_forces_nbl.c the
 * entry points call the HIP library through include/imdg.h.
 *
 * Errors follow IMD's convention: imd_error() prints the message and exits
 * with status 2 (src/imd_misc.c:78-98); the library functions themselves
 * return status codes.
 */
#ifndef IMD_HOST_H
#define IMD_HOST_H

#include "imdg.h"

#define IMD_MAXTYPES 8
#define IMD_STRLEN 256

/* pot_table_t (src/types.h:416-428) */
typedef struct {
  double *begin, *end, *step, *invstep;
  int *len;
  int ncols;
  int maxsteps;
  double *table; /* (maxsteps + 2) * ncols, table[k*ncols + col] */
} pot_table_t;

/* ------------------------------------------------------------------ */
/* globals (the subset of src/globals.h the force path reads)          */
/* ------------------------------------------------------------------ */
extern int ensemble;              /* IMDG_NVE / IMDG_NVT */
extern int ntypes;
extern double masses[IMD_MAXTYPES];
extern double box_x[3], box_y[3], box_z[3];
extern int box_param[3];
extern double box_unit;
extern int pbc_dirs[3];
extern double timestep;
extern double temperature;        /* starttemp, k_B T */
extern double eta, tau_eta, isq_tau_eta;
extern double nbl_margin, nbl_size;
extern int steps_min, steps_max, eng_int, checkpt_int;
extern int watch_int, stop_int;
extern int press_int, stress_tens;  /* .press interval; STRESS_TENS on */
extern int imdrestart, simulation;  /* -r n; current phase */
extern int do_maxwell, box_from_header;
extern long seed;
extern double maxwalltime;
extern char infilename[IMD_STRLEN], outfilename[IMD_STRLEN];
extern char itrfilename[IMD_STRLEN], paramfilename[IMD_STRLEN];
extern char progname[IMD_STRLEN];
extern char potfilename[IMD_STRLEN], eam2_emb_E_filename[IMD_STRLEN];
extern char eam2_at_rho_filename[IMD_STRLEN], eeam_mod_E_filename[IMD_STRLEN];
extern int gpu_precision, gpu_device, gpu_lanes, gpu_autotune;
extern int interpolation;  /* IMDG_INTERP_* */
extern int have_potfile, have_eam, have_eeam;

extern pot_table_t pair_pot, embed_pot, rho_h_tab, emod_pot;
extern double cellsz;

/* atoms, in input order */
extern long natoms;
extern int nactive;
extern double *atom_x, *atom_p, *atom_mass;
extern double *atom_epot, *atom_rho; /* per-atom Epot and eam_rho */
extern double *atom_eeam_p;          /* EEAM: eam_p (EAM_P) */
extern double *atom_press;           /* 6 per atom: xx yy zz yz zx xy */
extern double tot_presstens[6];      /* sum over the atoms */
extern int *atom_type;
extern long *atom_number;

/* thermodynamic results */
extern double tot_pot_energy, tot_kin_energy, virial, volume;
extern int nbl_count;
extern double main_loop_seconds;     /* wall time spent in main_loop */

extern void (*move_atoms)(void);

/* ------------------------------------------------------------------ */
/* functions                                                           */
/* ------------------------------------------------------------------ */
void imd_error(const char *fmt, ...);
void imd_warning(const char *fmt, ...);

/* parameters: defaults, then one simulation phase of a tag file;
   read_parameters returns 1 when no later phase follows */
void init_parameters(void);
int  read_command_line(int argc, char **argv);
int  read_parameters(const char *paramfile, int phase);
int  check_parameters_complete(void);

/* tables; returns 0 or a non-zero error code */
int  read_pot_table(pot_table_t *pt, const char *filename, int ncols,
                    int radial);
void free_pot_table(pot_table_t *pt);
int  setup_potentials(void);

/* configuration */
int  read_atoms(const char *filename);
int  generate_fcc(void);
void make_box(void);  /* volume from the box vectors */
void maxwell(double temp);
/* checkpoint number n (-1: final, -2: intermediate) after step `steps`,
   with its iteration file */
int  write_config(int fzhlr, int steps);
int  write_itr_file(int fzhlr, int steps, const char *suffix);
/* <outfiles>.%05d.press: per-atom stress tensor (atom_press) */
int  write_config_press(int nr);
int  write_eng_file_header(void);
int  write_eng_file(int steps);

/* force module entry points (IMD names and semantics) */
int  imd_gpu_init(void);
void imd_gpu_finish(void);
void imd_gpu_reconfigure(void);  /* between phases */
void calc_forces(int steps);
void make_nblist(void);
void check_nblist(void);
void deallocate_nblist(void);
void move_atoms_nve(void);
void move_atoms_nvt(void);
void fetch_atoms(void);

/* main loop of one phase: IMD's per-step sequence, fused on the device
   between output steps; returns 1 when the simulation is to end (stop
   file, wall time) */
int  main_loop(void);

imdg_ctx *imd_gpu_context(void);

#endif
