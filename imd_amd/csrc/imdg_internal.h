// imdg_internal.h -- device-side data layout of the MI355X neighbour-list
// pair/EAM path.  Shared by the kernels (imdg_kernels.hip) and the host side
// of the C ABI (imdg_api.hip).
//
// Layout in HBM (one context, all replicas concatenated):
//  * atoms live in "slots"; every replica owns a run of slots padded to a
//    multiple of IMDG_BLOCK so a thread block never spans two replicas;
//    padding slots carry type -1 and take part in nothing;
//  * after every list rebuild the atoms of a replica are sorted by cell
//    (z fastest, then y, then x; inside a cell by input index), so a block
//    of 256 consecutive slots is a compact column of cells and neighbour
//    gathers hit L2;
//  * pos[2]  double4 (x, y, z, type)   -- double buffered: the fused force
//    kernel reads pos[cur] and writes the integrated positions to pos[1-cur];
//  * mom[2]  double4 (px, py, pz, mass) and aid[2] (input index) -- double
//    buffered only for the rebuild permutation;
//  * nbr     uint32 full neighbour list, entry = index | image-code << 27,
//    padded per atom to a multiple of 4 with entries that point at the
//    atom itself through the far-away code IMDG_CODE_PAD, and stored in
//    groups of four per lane: entries 4c..4c+3 of slot i form one uint4 at
//    uint4 index ((i / 64) * (kmax / 4) + c) * 64 + i % 64, so a wavefront
//    reads 1 KiB contiguous per group and every lane gets four neighbours
//    from one load;
//  * tables as per-segment coefficients (p0, dv, d2v) of IMD's 3-point
//    interpolation, precomputed in FP64 and rounded to the force precision.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define IMDG_BLOCK 256
#define IMDG_WAVE 64
#define IMDG_MAXTYPES 8
#define IMDG_MAXCOLS (IMDG_MAXTYPES * IMDG_MAXTYPES)
#define IMDG_RING 4096
#define IMDG_IDX_BITS 27
#define IMDG_IDX_MASK ((1u << IMDG_IDX_BITS) - 1u)
#define IMDG_CODE_NOSHIFT 13
#define IMDG_CODE_PAD 27      // list padding: shift far beyond any cutoff
#define IMDG_NCODES 28
#define IMDG_PAD_SHIFT 1.0e6
#define IMDG_GROUP 4          // list entries per lane and load

// number of FP64 block partials written by the force/move kernels
#define PB_EPAIR 0
#define PB_VIRIAL 1
#define PB_KE 2      // NVE: sum (k1+k2)/(4m)
#define PB_E1 3      // NVT: sum k1/m
#define PB_E2 4      // NVT: sum k2/m
#define PB_MAXD 5    // max displacement^2 (max, not sum)
#define PB_N 6

// done counters (last-block-done pattern) per kernel kind
#define DC_PERMUTE 0
#define DC_LIST 1
#define DC_FINAL 2
#define DC_N 4

struct RepDev {
  double box[3][3];       // rows: box_x, box_y, box_z
  double tbox[3][3];      // dual vectors (make_box)
  double shift[IMDG_NCODES][3];  // image shift of code (sx+1)*9+(sy+1)*3+sz+1,
                                 // code IMDG_CODE_PAD: list padding
  double cellsz;          // list radius squared, skin included
  double skin2;           // (nbl_margin / 2)^2
  double temperature;     // NVT target kT
  double isq_tau_eta;     // 1 / tau_eta^2
  int gdim[3];
  int pbc[3];
  int atom_off;           // first slot
  int natoms;             // real atoms (slots atom_off .. atom_off+natoms-1)
  int nslots;             // padded slot count
  int cell_off;           // first cell
  int ncells;
  int blk_off;            // first logical block
  int nblk;
  int nactive;            // 3 * natoms
};

struct RepState {
  double eta;
  double epot, virial;    // from the last force evaluation
  int need_rebuild;
  int rebuilt;            // list rebuilt in the current step
  long nbl_count;
};

struct ThermoRec {
  double epot, ekin, virial, eta, max_disp2;
  long step;
  int rebuilt, natoms;
};

struct Ctl {
  int pos_cur;            // current position buffer
  int rec_cur;            // current momentum/index buffer
  int any_rebuild;        // some replica needs a list rebuild
  int frozen;             // list overflow: every kernel is a no-op until
                          // the host grows the list and re-runs the build
  int overflow;           // set by the list build, turned into frozen by
                          // its last block (frozen never changes mid-kernel)
  int kmax_seen;
  int is_short;
  unsigned int done[DC_N];
  long steps_done;
};

// Table of one interpolated function set on the device.
template <typename R>
struct Seg {
  R p0, dv, d2v, pad;
};

template <typename R>
struct TabDev {
  const Seg<R>* seg;      // all columns, column c at seg + off[c]
  int ncols;
  int off[IMDG_MAXCOLS];
  int len[IMDG_MAXCOLS];  // usable segments (rows of the IMD column)
  R begin[IMDG_MAXCOLS], end[IMDG_MAXCOLS], invstep[IMDG_MAXCOLS];
  R twoistep[IMDG_MAXCOLS];
};

// Everything the per-step kernels need, passed by value.
template <typename R>
struct StepArgs {
  const RepDev* rep;
  RepState* rst;
  Ctl* ctl;
  const int* blk_rep;     // logical block -> replica
  double4* pos[2];
  double4* mom[2];
  int* aid[2];
  double4* xnbl;
  const uint32_t* nbr;
  const int* nnb;
  const Seg<R>* seg;      // segment coefficients of all tables (global)
  R* rho;
  R* df;                  // 2 F'(rho)
  R* epot;
  R* frc;                 // 4 per slot (fx, fy, fz, 0)
  double* partA;          // [nblk] embedding energy partials
  double* partB;          // [nblk * PB_N]
  ThermoRec* ring;        // [IMDG_RING * nrep]
  int nrep;
  int nblk;
  int kmax;
  int ntypes;
  int ensemble;
  double dt;
  int has_pair, has_eam;
  int xcd_remap;          // map blocks to XCD-contiguous atom runs
};

struct BuildArgs {
  const RepDev* rep;
  RepState* rst;
  Ctl* ctl;
  const int* blk_rep;
  const int* cell_rep;
  double4* pos[2];
  double4* mom[2];
  int* aid[2];
  double4* xnbl;
  int* cellid;            // cell of slot (binning order)
  int* cslot;             // slot inside the cell from the atomic counter
  int* acell;             // cell of slot after the permutation
  int* perm;              // new slot -> old slot
  int* cell_cnt;          // atomic counters (zero outside a rebuild)
  int* cell_beg;
  int* cell_num;
  uint32_t* nbr;
  int* nnb;
  int nrep;
  int nslots;
  int ncells;
  int kmax;
};
