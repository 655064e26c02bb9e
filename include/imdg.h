/*
 * imdg.h -- C ABI of the MI355X neighbour-list pair/EAM force path.
 *
 * This is the drop-in boundary for IMD's force module.  IMD selects its
 * force loop at link time (src/Makefile:1144-1147: the "nbl" target token
 * links src/imd_forces_nbl.c) and calls it through process-global entry
 * points without arguments.  A host that links this library instead keeps
 * those names and semantics and forwards to the functions below:
 *
 *   IMD entry point (reference)                         -> this ABI
 *   void calc_forces(int)   src/prototypes.h:173,
 *                           src/imd_forces_nbl.c:281     imdg_calc_forces
 *   void make_nblist(void)  src/prototypes.h:196,
 *                           src/imd_forces_nbl.c:136     imdg_make_nblist
 *   void check_nblist(void) src/prototypes.h:197,
 *                           src/imd_forces_nbl.c:2007    imdg_check_nblist
 *   void deallocate_nblist(void) src/prototypes.h:198,
 *                           src/imd_forces_nbl.c:56      imdg_deallocate_nblist
 *   void (*move_atoms)(void) src/globals.h:1444, set from
 *        "ensemble" (src/imd_param.c:377-444): move_atoms_nve
 *        src/imd_integrate.c:32, move_atoms_nvt :891     imdg_move_atoms
 *   read_pot_table result pot_table_t src/types.h:416-428
 *                           (src/imd_potential.c:161)    imdg_set_tables
 *   cell_array/ORT/IMPULS/SORTE/MASSE globals
 *                           src/types.h:181-313          imdg_set_atoms,
 *                                                        imdg_get_atoms
 *   main_loop step: calc_forces, move_atoms, check_nblist
 *                           src/imd_main_3d.c:405,559,769 imdg_run
 *
 * Conventions.  Every call returns IMDG_OK (0) or an error code and never
 * exits; imdg_last_error() gives the message.  A context owns all device
 * memory and one HIP stream; host arrays are only borrowed for the duration
 * of a call.  A context is not thread-safe; separate contexts are
 * independent.  Units and conventions are IMD's: positions in length units,
 * momenta p = m v, temperatures as k_B T in energy units, tables in r^2.
 */
#ifndef IMDG_H
#define IMDG_H

#ifdef __cplusplus
extern "C" {
#endif

#define IMDG_OK            0
#define IMDG_ERR_ARG       1   /* invalid argument or state */
#define IMDG_ERR_HIP       2   /* HIP runtime failure */
#define IMDG_ERR_NOMEM     3   /* device or host allocation failed */
#define IMDG_ERR_GEOM      4   /* box too small for the cell grid */
#define IMDG_ERR_NOTABLE   5   /* tables missing or inconsistent */
#define IMDG_ERR_STATE     6   /* call order violated */

#define IMDG_FP64   0   /* all force arithmetic in FP64 */
#define IMDG_MIXED  1   /* FP32 forces/tables, FP64 integration, Kahan sums */

#define IMDG_NVE    0
#define IMDG_NVT    1

typedef struct imdg_ctx imdg_ctx;

typedef struct {
  int    precision;      /* IMDG_FP64 or IMDG_MIXED */
  int    ensemble;       /* IMDG_NVE or IMDG_NVT (IMD key "ensemble") */
  int    ntypes;         /* IMD key "ntypes" (1..8) */
  double timestep;       /* IMD key "timestep" */
  double nbl_margin;     /* IMD key "nbl_margin" (list skin) */
  double nbl_size;       /* IMD key "nbl_size": list capacity headroom */
  int    device;         /* HIP device ordinal */
  int    lanes_per_atom; /* lanes per atom of both neighbour loops: 1, 2
                            or 4; 0 = the default (see imdg_tuning) */
  int    interpolation;  /* IMDG_INTERP_*: IMD's table interpolation, a
                            build option there (src/potaccess.h:24-36) */
} imdg_config;

/* Table interpolation.  3-point is IMD's default (PAIR_INT2, VAL_FUNC2,
   DERIV_FUNC2); FOURPOINT uses PAIR_INT3 and friends
   (src/potaccess.h:365-412), SPLINE PAIR_INT_SP and friends
   (src/potaccess.h:414-457).  For the last two the library derives the
   guard rows (init_fourpoint, src/imd_potential.c:1171-1187) and the spline
   second derivatives (init_spline, src/imd_potential.c:1197-1246; radial
   tables with zero slope at the end, the embedding function natural) from
   the first len rows of each column, so the tables are passed as read. */
#define IMDG_INTERP_3POINT 0
#define IMDG_INTERP_4POINT 1
#define IMDG_INTERP_SPLINE 2

/* A read IMD function table (pot_table_t, src/types.h:416-428), after
   read_pot_table has shifted radial columns and appended the two guard rows
   (src/imd_potential.c:1256-1272).  Values are table[k*ncols + col] for
   k < maxsteps + 2.  Radial tables are functions of r^2. */
typedef struct {
  int           ncols;
  int           maxsteps;
  const double *begin, *end, *step, *invstep;
  const int    *len;
  const double *table;
} imdg_table;

/* One independent system (a replica).  IMD keys box_x/box_y/box_z,
   pbc_dirs, starttemp (NVT target), eta, tau_eta (isq_tau_eta = 1/tau^2). */
typedef struct {
  double box_x[3], box_y[3], box_z[3];
  int    pbc[3];
  double temperature;
  double eta;
  double isq_tau_eta;
} imdg_replica;

/* Per-replica thermodynamic record of one step (src/imd_io.c:2262-2540
   reads these globals): tot_pot_energy, tot_kin_energy, virial, eta. */
typedef struct {
  long   step;          /* steps completed on the device when recorded */
  double epot;          /* tot_pot_energy */
  double ekin;          /* tot_kin_energy */
  double virial;        /* virial (pair -r^2 grad, EAM -d.F) */
  double eta;           /* Nose-Hoover eta after this step */
  double max_disp2;     /* max |x - x_nbl|^2 after this step */
  int    rebuilt;       /* the list was rebuilt at the start of this step */
  int    natoms;
} imdg_thermo;

typedef struct {
  int    nrep;
  long   natoms;        /* real atoms over all replicas */
  long   nslots;        /* padded atom slots on the device */
  int    kmax;          /* list capacity per atom */
  int    kmax_seen;     /* largest list length built so far */
  long   nbl_count;     /* list builds of replica 0 */
  long   steps_done;    /* device step counter */
  int    is_short;      /* some r^2 fell below a table's begin since the
                           previous imdg_get_info (which clears it): IMD's
                           is_short of calc_forces
                           (src/imd_forces_nbl.c:982, 1323) */
  int    lanes_per_atom;/* lanes per atom of the force loop */
  double list_bytes;    /* bytes of list storage */
  double mean_k;        /* mean full-list length per real atom */
  int    tiles_fallback;/* blocks whose neighbour tile did not fit in LDS
                           (their neighbours are gathered from HBM) */
  long   exact_pairs;   /* list candidates (so far) within the FP32 screening band
                           of the list radius, decided by IMD's FP64
                           expression r^2 < cellsz
                           (src/imd_forces_nbl.c:249-261) */
  int    tables_in_lds; /* the step loops read their table segments from
                           LDS: bit 0 the density loop, bit 1 the force
                           loop (otherwise from HBM/L2) */
} imdg_info;

/* lifecycle */
int  imdg_ctx_create(const imdg_config *cfg, imdg_ctx **out);
int  imdg_ctx_destroy(imdg_ctx *ctx);
const char *imdg_last_error(const imdg_ctx *ctx);

/* setup_potentials (src/imd_potential.c:43-131): pair has ntypes^2 columns,
   rho ntypes^2 (column it*ntypes+jt = density at an it-atom due to a
   jt-atom), embed ntypes columns (functions of rho).  pair may be NULL for
   EAM without a pair term; rho and embed are both NULL for a pair-only
   build (imd_nbl_nve). */
int  imdg_set_tables(imdg_ctx *ctx, const imdg_table *pair,
                     const imdg_table *rho, const imdg_table *embed);
/* EEAM (src/imd_forces_nbl.c:591-608, 1090-1095, 1180-1207): the energy
   modification term E(P) of the second density P_i = sum_j rho_ij(r^2)^2,
   ntypes columns, not radial (read_pot_table(&emod_pot, ...),
   src/imd_potential.c:82-85; IMD key eeam_energy_file).  The force gains
   2E'(P_i) rho_j rho'_j + 2E'(P_j) rho_i rho'_i per pair.  After
   imdg_set_tables, before imdg_set_atoms; NULL removes it. */
int  imdg_set_eeam_table(imdg_ctx *ctx, const imdg_table *emod);

/* read_atoms (src/imd_io_3d.c:44-520) for nrep replicas.  Arrays hold the
   atoms of replica 0, then replica 1, ... (natoms[r] each): x and p 3 per
   atom, type 0..ntypes-1, mass per atom.  Positions are mapped back into
   the box (back_into_box, src/imd_geom_3d.c:1083-1110).  The grid is sized
   like init_cells (src/imd_geom_3d.c:113-212) from the largest radial table
   end plus nbl_margin; periodic directions need at least 3 cells. */
int  imdg_set_atoms(imdg_ctx *ctx, int nrep, const imdg_replica *reps,
                    const int *natoms, const double *x, const double *p,
                    const int *type, const double *mass);

/* Nose-Hoover state of one replica (IMD keys starttemp, eta, tau_eta). */
int  imdg_set_nvt(imdg_ctx *ctx, int rep, double temperature, double eta,
                  double isq_tau_eta);

/* calc_forces(steps): rebuild the list first if it is invalid (fix_cells,
   make_nblist), then pair forces, host density, embedding and EAM forces,
   per-atom energies, tot_pot_energy and virial.  Synchronous. */
int  imdg_calc_forces(imdg_ctx *ctx, int steps);
/* move_atoms for the configured ensemble, using the last forces. */
int  imdg_move_atoms(imdg_ctx *ctx);
/* check_nblist: marks replicas whose max displacement since the last list
   exceeds nbl_margin/2; *invalid (may be NULL) gets 1 if any is marked. */
int  imdg_check_nblist(imdg_ctx *ctx, int *invalid);
/* make_nblist after fix_cells, for every replica, now. */
int  imdg_make_nblist(imdg_ctx *ctx);
/* deallocate_nblist: frees the list storage and invalidates the list. */
int  imdg_deallocate_nblist(imdg_ctx *ctx);

/* nsteps of main_loop (calc_forces, move_atoms, check_nblist) fused on the
   device: no host synchronisation per step, list rebuilds decided on the
   device.  Returns after the steps are enqueued; imdg_sync waits. */
int  imdg_run(imdg_ctx *ctx, long nsteps);
int  imdg_sync(imdg_ctx *ctx);

/* Results.  Per-atom arrays are in input order (as given to set_atoms),
   any pointer may be NULL. */
int  imdg_get_atoms(imdg_ctx *ctx, double *x, double *p, double *f,
                    double *epot, double *rho, double *df);
int  imdg_get_thermo(imdg_ctx *ctx, int rep, imdg_thermo *out);
/* Thermo records of steps [first, first+count) for every replica:
   out[(s*nrep + r)] ; steps must still be in the device ring (last 4096). */
int  imdg_get_thermo_history(imdg_ctx *ctx, long first, long count,
                             imdg_thermo *out);
/* Global cell coordinates (3 ints per atom, input order) of the last
   rebuild (cell_coord, src/imd_geom_3d.c:1054-1074). */
int  imdg_get_cells(imdg_ctx *ctx, int *cell3);
/* Canonical half-pair set of the current list: 6 longs per pair
   (replica, i_lo, i_hi, sx, sy, sz) with i_lo < i_hi input indices within
   the replica and (sx,sy,sz) the image shift of i_hi relative to i_lo in
   box vectors.  *npairs receives the count; at most cap pairs written. */
int  imdg_get_pairs(imdg_ctx *ctx, long *out, long cap, long *npairs);
/* Grid dimensions (3 ints) of a replica. */
int  imdg_get_grid(imdg_ctx *ctx, int rep, int *gdim3);

/* Per-atom stress tensor, IMD's STRESS_TENS with do_press_calc on the
   steps whose tensor is read (PRESSTENS: src/imd_forces_nbl.c:346-352,
   558-581, 1283-1318; src/imd_integrate.c:411-431, 1085-1094; summed over
   the atoms like calc_tot_presstens, src/imd_main_3d.c:2069-2100).  While
   enabled, the tensor of the last step of every imdg_run and of every
   imdg_move_atoms is computed after that step: per atom -1/2 sum_j
   (d d^T) g over its neighbours, g the pair's force/distance factor, plus
   p p^T / m with the momenta after the step.  Energy units (divide the sum
   by the volume for a pressure tensor). */
int  imdg_set_stress(imdg_ctx *ctx, int enable);
/* 6 doubles per atom (input order) and 6 per replica (the sums), in the
   order xx yy zz yz zx xy; either pointer may be NULL.  IMDG_ERR_STATE if
   no tensor of the current state was computed. */
int  imdg_get_stress(imdg_ctx *ctx, double *per_atom, double *total);

/* EEAM per-atom P (EAM_P, from the last step's force evaluation) and
   2E'(P) (EAM_DM), input order; either pointer may be NULL.
   IMDG_ERR_STATE without an EEAM table. */
int  imdg_get_eeam(imdg_ctx *ctx, double *P, double *dm);
int  imdg_get_info(imdg_ctx *ctx, imdg_info *info);

/* Kernel timing with HIP events on the context's stream.  enable=1 starts
   recording the duration of every launch of each kernel class, any other
   non-zero value only of the classes whose bit (1 << class index, in the
   order imdg_kernel_times reports them) is set, 0 stops; names and
   accumulated milliseconds/launch counts come back from imdg_kernel_times
   (returns the number of classes, writes at most n). */
int  imdg_profile(imdg_ctx *ctx, int enable);
int  imdg_kernel_times(imdg_ctx *ctx, const char **names, double *ms,
                       long *launches, int n);

/* Launch shape of the two neighbour loops (no IMD counterpart: IMD runs
   them as plain loops over atoms).  Each atom's list is walked by `lanes`
   lanes of a wavefront, lane s taking entries s, s+lanes, ... (so every
   lane handles about K/lanes neighbours, and the records one gather reads
   for an atom are consecutive list entries); a block of `threads` threads
   holds threads/lanes atoms at a time.  Results differ between shapes only
   by the order of floating-point sums. */
typedef struct {
  int    density_lanes;   /* 1, 2 or 4 */
  int    density_threads; /* 256, 512 or 1024 */
  int    force_lanes;     /* 1, 2 or 4 */
  int    force_threads;   /* 256, 512, or up to the force loop's maximum
                             (768 mixed, 512 FP64) */
  double density_ms;      /* imdg_autotune: mean time of the chosen density
                             launch (0 otherwise) */
  double force_ms;        /* imdg_autotune: same for the force launch */
} imdg_tuning;

/* Fields are 0 until imdg_set_atoms has planned the launches (0 = the
   default shape for the context's precision). */
int  imdg_get_tuning(imdg_ctx *ctx, imdg_tuning *out);
/* Takes effect for the next launch; IMDG_ERR_ARG (and no change) for a
   shape the loops cannot be launched with. */
int  imdg_set_tuning(imdg_ctx *ctx, const imdg_tuning *t);
/* Times every lanes x threads shape of the density loop and then of the
   force loop on the current state (reps launches each, the list is built
   first if needed) and keeps the fastest of each.  The state does not
   advance; forces and energies are left as imdg_calc_forces leaves them.
   *chosen (may be NULL) receives the result. */
int  imdg_autotune(imdg_ctx *ctx, int reps, imdg_tuning *chosen);

/* check_nblist on caller-owned device arrays, allocating nothing.  No
   counterpart in the reference (IMD's check_nblist works on its own cell
   arrays): it is for callers that cannot allocate device memory, such as
   code under HIP graph capture or a host loop that keeps its positions on
   the device.  x and x_nbl are device pointers to n positions, 3 doubles
   each (now and at the last list build); the kernels leave
   max |x - x_nbl|^2 in *max_disp2 and IMD's decision max > (nbl_margin /
   2)^2 in *invalid (device pointers; either may be NULL, not both), on
   `stream` (a hipStream_t, NULL for the null stream), without host
   synchronisation.  The reduction goes through a fixed device array of
   65,536 doubles per GPU: n above 16,777,216 (65,536 blocks of 256 atoms)
   is rejected with IMDG_ERR_ARG before anything is launched.  Calls that
   overlap on one GPU share that array and race.  imdg_last_error(NULL)
   gives the message of a failed call on the calling host thread. */
int  imdg_check_nblist_noalloc(int device, const double *x,
                               const double *x_nbl, long n,
                               double nbl_margin, double *max_disp2,
                               int *invalid, void *stream);

/* sizeof() of a public struct by name ("imdg_config", "imdg_table",
   "imdg_replica", "imdg_thermo", "imdg_info", "imdg_tuning"), -1 for an
   unknown name: lets
   a binding in another language check its layouts against the library's.
   No IMD counterpart (IMD links its force module statically). */
long imdg_struct_size(const char *name);

#ifdef __cplusplus
}
#endif
#endif
