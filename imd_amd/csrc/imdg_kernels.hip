// imdg_kernels.hip -- CDNA4 (gfx950) kernels of the neighbour-list pair/EAM
// path: cell binning and sort, per-block neighbour tiles, full neighbour
// list build, host density with embedding, pair + EAM forces fused with the
// leapfrog step and the skin test, compensated FP64 reductions.
//
// Reference semantics (paths relative to the reference root):
//   binning / wrap     src/imd_fix_cells_3d.c:36-201, src/imd_main_3d.c:1972-2065,
//                      cell_coord src/imd_geom_3d.c:1054-1074
//   list               make_nblist src/imd_forces_nbl.c:136-273 (half list over
//                      the 14-cell stencil of src/imd_geom_3d.c:895-899 with
//                      buffer-cell images src/imd_comm_force_3d.c:222-397);
//                      built here as a full list whose r^2 is evaluated in the
//                      same direction and operation order as the reference, so
//                      the pair set is identical
//   density/embedding  src/imd_forces_nbl.c:586-611, 1079-1112
//   pair + EAM force   src/imd_forces_nbl.c:490-584, 1117-1322
//   tables             PAIR_INT2/VAL_FUNC2/DERIV_FUNC2 src/potaccess.h:323-354,
//                      465-495, 591-621
//   integrators        move_atoms_nve src/imd_integrate.c:122-358,
//                      move_atoms_nvt src/imd_integrate.c:949-1141
//   skin test          check_nblist src/imd_forces_nbl.c:2007-2037
#include <hip/hip_runtime.h>
#include <math.h>

#include "imdg_internal.h"

namespace imdg {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of blocks (a contiguous, spatially compact run of
// cell-sorted atoms) so neighbour data stays inside that XCD's L2.
__device__ __forceinline__ int logical_block() {
  int nb = gridDim.x, b = blockIdx.x;
  int q = nb >> 3, rem = nb & 7, x = b & 7, idx = b >> 3;
  return (x < rem) ? x * (q + 1) + idx : rem * (q + 1) + (x - rem) * q + idx;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, IMDG_WAVE);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, IMDG_WAVE));
  return v;
}

// Error-free sum of two doubles (Knuth TwoSum).
__device__ __forceinline__ void two_sum(double a, double b, double& s,
                                        double& e) {
#pragma clang fp contract(off)
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Neumaier-compensated accumulation.
__device__ __forceinline__ void kahan_add(double& s, double& c, double x) {
#pragma clang fp contract(off)
  double t = s + x;
  if (fabs(s) >= fabs(x))
    c += (s - t) + x;
  else
    c += (x - t) + s;
  s = t;
}

// Block-wide reduction of N values (sum except the indices flagged in
// maxmask), result valid in thread 0.
template <int N>
__device__ __forceinline__ void block_reduce(double (&v)[N], unsigned maxmask,
                                             double* lds /* NWAVE*N */) {
#pragma unroll
  for (int q = 0; q < N; q++)
    v[q] = ((maxmask >> q) & 1) ? wave_max(v[q]) : wave_sum(v[q]);
  int w = threadIdx.x / IMDG_WAVE, l = threadIdx.x % IMDG_WAVE;
  if (l == 0) {
#pragma unroll
    for (int q = 0; q < N; q++) lds[w * N + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < N; q++) {
      double a = lds[q];
#pragma unroll
      for (int k = 1; k < IMDG_NWAVE; k++)
        a = ((maxmask >> q) & 1) ? fmax(a, lds[k * N + q]) : a + lds[k * N + q];
      v[q] = a;
    }
  }
}

template <typename R>
__device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }

// Segment index and fraction of IMD's 3-point table access
// (src/potaccess.h:329-340): argument clamped to [begin, end].
template <typename R>
__device__ __forceinline__ int seg_locate(R x, R begin, R end, R invstep,
                                          int len, R& chi) {
  R a = rmin(x, end) - begin;
  a = a < (R)0 ? (R)0 : a;
  R t = a * invstep;
  int k = (int)t;
  chi = t - (R)k;
  return k < len - 1 ? k : len - 1;
}

// value p0 + chi (dv + (chi-1)/2 d2v) and twice the argument-derivative
// 2 istep (dv + (chi - 1/2) d2v) of IMD's 3-point interpolation
template <typename R>
__device__ __forceinline__ R seg_val(const Seg<R>& c, R chi) {
  return c.p0 + chi * (c.dv + (R)0.5 * (chi - (R)1) * c.d2v);
}
template <typename R>
__device__ __forceinline__ R seg_grad(const Seg<R>& c, R chi, R twoistep) {
  return twoistep * (c.dv + (chi - (R)0.5) * c.d2v);
}

template <typename R>
__device__ __forceinline__ void tab_val_grad(const Seg<R>* seg, int len,
                                             R begin, R end, R invstep,
                                             R twoistep, R x, R& val, R& grad,
                                             int& is_short) {
  R chi;
  int k = seg_locate(x, begin, end, invstep, len, chi);
  is_short |= (x < begin);
  Seg<R> c = seg[k];
  val = seg_val(c, chi);
  grad = seg_grad(c, chi, twoistep);
}

// Entry m of slot i in the grouped list layout (4 entries per lane).
__device__ __forceinline__ size_t nbr_index(int i, int m, int kmax) {
  return (((size_t)(i >> 6) * (kmax >> 2) + (m >> 2)) * IMDG_WAVE + (i & 63)) *
             IMDG_GROUP +
         (m & 3);
}

// First uint4 of slot i's list.
__device__ __forceinline__ const uint4* list_of(const uint32_t* nbr, int i,
                                                int kmax) {
  return reinterpret_cast<const uint4*>(nbr) +
         ((size_t)(i >> 6) * (kmax >> 2)) * IMDG_WAVE + (i & 63);
}

__device__ __forceinline__ void cell_coords(const RepDev& R, int cell,
                                            int& cx, int& cy, int& cz) {
  int lc = cell - R.cell_off;
  cz = lc % R.gdim[2];
  int q = lc / R.gdim[2];
  cy = q % R.gdim[1];
  cx = q / R.gdim[1];
}

// Centre of a cell: sum_d (c_d + 1/2) box_d / gdim_d.
__device__ __forceinline__ void cell_centre(const RepDev& R, int cell,
                                            double* c) {
  int cx, cy, cz;
  cell_coords(R, cell, cx, cy, cz);
  double fx = cx + 0.5, fy = cy + 0.5, fz = cz + 0.5;
#pragma unroll
  for (int e = 0; e < 3; e++)
    c[e] = fx * R.cvec[0][e] + fy * R.cvec[1][e] + fz * R.cvec[2][e];
}

// ---------------------------------------------------------------------------
// rebuild: wrap + bin, scan, place, per-cell order, permute, tiles, list
// ---------------------------------------------------------------------------

// fix_cells/do_boundaries + cell_coord, for replicas that need a rebuild.
__global__ void __launch_bounds__(256) k_bin(BuildArgs a) {
#pragma clang fp contract(off)
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  double4* pos = a.pos[a.ctl->pos_cur];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nslots;
       i += gridDim.x * blockDim.x) {
    int r = a.blk_rep[i / IMDG_BLOCK];
    if (!a.rst[r].need_rebuild) continue;
    const RepDev& R = a.rep[r];
    if (i - R.atom_off >= R.natoms) continue;
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; d++) {
      if (R.pbc[d] != 1) continue;
      double f = -floor(x[0] * R.tbox[d][0] + x[1] * R.tbox[d][1] +
                        x[2] * R.tbox[d][2]);
      x[0] += f * R.box[d][0];
      x[1] += f * R.box[d][1];
      x[2] += f * R.box[d][2];
    }
    pos[i] = make_double4(x[0], x[1], x[2], p.w);
    int c[3];
    for (int d = 0; d < 3; d++) {
      c[d] = (int)(R.gdim[d] * (x[0] * R.tbox[d][0] + x[1] * R.tbox[d][1] +
                                x[2] * R.tbox[d][2]));
      if (c[d] >= R.gdim[d])
        c[d] = R.gdim[d] - 1;
      else if (c[d] < 0)
        c[d] = 0;
    }
    int cell = R.cell_off + (c[0] * R.gdim[1] + c[1]) * R.gdim[2] + c[2];
    a.cellid[i] = cell;
    a.cslot[i] = atomicAdd(&a.cell_cnt[cell], 1);
  }
}

// Exclusive scan of the cell counts of one replica (one block each), and
// reset of the counters.
__global__ void __launch_bounds__(1024) k_scan(BuildArgs a) {
  __shared__ int sums[1024];
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  int r = blockIdx.x;
  if (!a.rst[r].need_rebuild) return;
  const RepDev& R = a.rep[r];
  int n = R.ncells, base = R.cell_off, t = threadIdx.x;
  int per = (n + 1023) / 1024;
  int b = t * per, e = min(b + per, n);
  int s = 0;
  for (int c = b; c < e; c++) s += a.cell_cnt[base + c];
  sums[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = (t >= o) ? sums[t - o] : 0;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  int run = R.atom_off + sums[t] - s;
  for (int c = b; c < e; c++) {
    int k = a.cell_cnt[base + c];
    a.cell_beg[base + c] = run;
    a.cell_num[base + c] = k;
    a.cell_cnt[base + c] = 0;
    run += k;
  }
}

__global__ void __launch_bounds__(256) k_place(BuildArgs a) {
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nslots;
       i += gridDim.x * blockDim.x) {
    int r = a.blk_rep[i / IMDG_BLOCK];
    if (!a.rst[r].need_rebuild) continue;
    const RepDev& R = a.rep[r];
    if (i - R.atom_off >= R.natoms) continue;
    a.perm[a.cell_beg[a.cellid[i]] + a.cslot[i]] = i;
  }
}

// Order the atoms of a cell by input index (deterministic list order).
__global__ void __launch_bounds__(256) k_cellsort(BuildArgs a) {
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  const int* aid = a.aid[a.ctl->rec_cur];
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < a.ncells;
       c += gridDim.x * blockDim.x) {
    int r = a.cell_rep[c];
    if (!a.rst[r].need_rebuild) continue;
    int b = a.cell_beg[c], n = a.cell_num[c];
    int* p = a.perm + b;
    for (int k = 1; k < n; k++) {
      int v = p[k], key = aid[v], m = k - 1;
      while (m >= 0 && aid[p[m]] > key) {
        p[m + 1] = p[m];
        m--;
      }
      p[m + 1] = v;
    }
  }
}

// Gather every slot into the other buffers (new cell-sorted order for the
// rebuilt replicas, identity for the others), write the cell-relative
// records of the moved atoms, then flip the buffers.
__global__ void __launch_bounds__(256) k_permute(BuildArgs a) {
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  int pc = a.ctl->pos_cur, rc = a.ctl->rec_cur;
  const double4* pin = a.pos[pc];
  double4* pout = a.pos[pc ^ 1];
  const double4* min_ = a.mom[rc];
  double4* mout = a.mom[rc ^ 1];
  const int* ain = a.aid[rc];
  int* aout = a.aid[rc ^ 1];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nslots;
       i += gridDim.x * blockDim.x) {
    int r = a.blk_rep[i / IMDG_BLOCK];
    const RepDev& R = a.rep[r];
    bool moved = a.rst[r].need_rebuild && (i - R.atom_off < R.natoms);
    int old = moved ? a.perm[i] : i;
    double4 p = pin[old];
    pout[i] = p;
    mout[i] = min_[old];
    aout[i] = ain[old];
    if (moved) {
      int cell = a.cellid[old];
      a.acell[i] = cell;
      a.atype[i] = (signed char)p.w;
      a.xnbl[i] = p;
      double c[3];
      cell_centre(R, cell, c);
      if (a.mixed)
        reinterpret_cast<float4*>(a.rec_a)[i] =
            make_float4((float)(p.x - c[0]), (float)(p.y - c[1]),
                        (float)(p.z - c[2]), (float)p.w);
      else
        reinterpret_cast<double4*>(a.rec_a)[i] =
            make_double4(p.x - c[0], p.y - c[1], p.z - c[2], p.w);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned t = atomicAdd(&a.ctl->done[DC_PERMUTE], 1u);
    if (t == gridDim.x - 1) {
      a.ctl->pos_cur = pc ^ 1;
      a.ctl->rec_cur = rc ^ 1;
      a.ctl->done[DC_PERMUTE] = 0;
      __threadfence();
    }
  }
}

// Tiles: for every block of a rebuilt replica, the slot runs of all cells
// adjacent (27-stencil, periodic wrap, nothing beyond an open face) to the
// cells of the block's atoms, sorted and merged.
#define TILE_MAXRANGE 96
__global__ void __launch_bounds__(64) k_tiles(BuildArgs a) {
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < a.nblk;
       b += gridDim.x * blockDim.x) {
    int r = a.blk_rep[b];
    if (!a.rst[r].need_rebuild) continue;
    const RepDev& R = a.rep[r];
    Tile& T = a.tiles[b];
    int s0 = b * IMDG_BLOCK;
    int s1 = min(s0 + IMDG_BLOCK, R.atom_off + R.natoms) - 1;
    T.nrun = 0;
    T.natoms = 0;
    T.ok = 1;
    if (s1 < s0) continue;
    int g0 = R.gdim[0], g1 = R.gdim[1], g2 = R.gdim[2];
    int lc0 = a.acell[s0] - R.cell_off, lc1 = a.acell[s1] - R.cell_off;
    int rb[TILE_MAXRANGE], re[TILE_MAXRANGE], nr = 0;
    bool ok = true;
    // walk the block's cells column by column
    for (int col = lc0 / g2; col <= lc1 / g2 && ok; col++) {
      int z0 = (col == lc0 / g2) ? lc0 % g2 : 0;
      int z1 = (col == lc1 / g2) ? lc1 % g2 : g2 - 1;
      int cx = col / g1, cy = col % g1;
      for (int ox = -1; ox <= 1 && ok; ox++) {
        int nx = cx + ox;
        if (nx < 0 || nx >= g0) {
          if (R.pbc[0] != 1) continue;
          nx = (nx + g0) % g0;
        }
        for (int oy = -1; oy <= 1 && ok; oy++) {
          int ny = cy + oy;
          if (ny < 0 || ny >= g1) {
            if (R.pbc[1] != 1) continue;
            ny = (ny + g1) % g1;
          }
          int za = z0 - 1, zb = z1 + 1, p[2], q[2], np = 0;
          if (R.pbc[2] == 1) {
            if (zb - za + 1 >= g2) {
              p[0] = 0; q[0] = g2 - 1; np = 1;
            } else if (za < 0) {
              p[0] = za + g2; q[0] = g2 - 1; p[1] = 0; q[1] = zb; np = 2;
            } else if (zb >= g2) {
              p[0] = za; q[0] = g2 - 1; p[1] = 0; q[1] = zb - g2; np = 2;
            } else {
              p[0] = za; q[0] = zb; np = 1;
            }
          } else {
            p[0] = za < 0 ? 0 : za;
            q[0] = zb >= g2 ? g2 - 1 : zb;
            np = 1;
          }
          int cbase = R.cell_off + (nx * g1 + ny) * g2;
          for (int k = 0; k < np; k++) {
            if (nr >= TILE_MAXRANGE) {
              ok = false;
              break;
            }
            int cf = cbase + p[k], cl = cbase + q[k];
            rb[nr] = a.cell_beg[cf];
            re[nr] = a.cell_beg[cl] + a.cell_num[cl];
            if (re[nr] > rb[nr]) nr++;
          }
        }
      }
    }
    // sort ranges by first slot, merge overlapping or touching ones
    for (int k = 1; k < nr; k++) {
      int vb = rb[k], ve = re[k], m = k - 1;
      while (m >= 0 && rb[m] > vb) {
        rb[m + 1] = rb[m];
        re[m + 1] = re[m];
        m--;
      }
      rb[m + 1] = vb;
      re[m + 1] = ve;
    }
    int nrun = 0, tot = 0;
    for (int k = 0; k < nr && ok; k++) {
      if (nrun > 0 && rb[k] <= T.run[nrun - 1].x + T.run[nrun - 1].y) {
        int4& u = T.run[nrun - 1];
        int end = max(u.x + u.y, re[k]);
        tot += end - (u.x + u.y);
        u.y = end - u.x;
      } else {
        if (nrun >= IMDG_MAXRUN) {
          ok = false;
          break;
        }
        T.run[nrun] = make_int4(rb[k], re[k] - rb[k], tot, 0);
        tot += re[k] - rb[k];
        nrun++;
      }
    }
    if (!ok || tot > IMDG_TILE_CAP) {
      T.ok = 0;
      T.nrun = 0;
      T.natoms = 0;
      atomicAdd(&a.ctl->tiles_fallback, 1);
    } else {
      T.nrun = nrun;
      T.natoms = tot;
    }
  }
}

// Image of x shifted by s box vectors, added z, then y, then x: the order in
// which send_cells builds buffer-cell copies (copy_cell adds the box vector,
// src/imd_comm_force_3d.c:754-756).
__device__ __forceinline__ void image_of(const RepDev& R, double x0, double x1,
                                        double x2, int sx, int sy, int sz,
                                        double& y0, double& y1, double& y2) {
#pragma clang fp contract(off)
  y0 = x0;
  y1 = x1;
  y2 = x2;
  if (sz) {
    double s = (double)sz;
    y0 = y0 + s * R.box[2][0];
    y1 = y1 + s * R.box[2][1];
    y2 = y2 + s * R.box[2][2];
  }
  if (sy) {
    double s = (double)sy;
    y0 = y0 + s * R.box[1][0];
    y1 = y1 + s * R.box[1][1];
    y2 = y2 + s * R.box[1][2];
  }
  if (sx) {
    double s = (double)sx;
    y0 = y0 + s * R.box[0][0];
    y1 = y1 + s * R.box[0][1];
    y2 = y2 + s * R.box[0][2];
  }
}

// Position of a slot inside the block's tile (runs sorted by first slot).
__device__ __forceinline__ int tile_index(const int4* runs, int nrun, int j) {
  int lo = 0, hi = nrun - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (runs[mid].x <= j)
      lo = mid;
    else
      hi = mid - 1;
  }
  return runs[lo].z + (j - runs[lo].x);
}

// Slot of a tile position (runs sorted by first tile position).
__device__ __forceinline__ int tile_slot(const int4* runs, int nrun, int t) {
  int lo = 0, hi = nrun - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (runs[mid].z <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return runs[lo].x + (t - runs[lo].z);
}

// Full neighbour list.  For each cell offset o of the 27, the reference
// evaluates a pair from the atom whose cell has the other in its half
// stencil (o lexicographically positive: the stencil of
// src/imd_geom_3d.c:895-899) as (image of the partner) - (owner); the same
// expression decides membership here, so each unordered pair is in the
// list of both atoms exactly when the reference has it in its half list.
// One block per atom block; its tile's FP64 positions are staged in LDS.
__global__ void __launch_bounds__(IMDG_BLOCK) k_list(BuildArgs a) {
#pragma clang fp contract(off)
  __shared__ double4 tpos[IMDG_TILE_CAP];
  __shared__ int4 truns[IMDG_MAXRUN];
  __shared__ int tinfo[2];
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  int b = blockIdx.x;
  int r = a.blk_rep[b];
  const RepDev& R = a.rep[r];
  bool flagged = a.rst[r].need_rebuild != 0;
  const double4* pos = a.pos[a.ctl->pos_cur];
  int kmax = a.kmax, seen = 0;
  if (flagged) {
    const Tile& T = a.tiles[b];
    if (threadIdx.x == 0) {
      tinfo[0] = T.nrun;
      tinfo[1] = T.ok;
    }
    for (int k = threadIdx.x; k < T.nrun; k += blockDim.x) truns[k] = T.run[k];
    __syncthreads();
    int nrun = tinfo[0], tok = tinfo[1];
    if (tok)
      for (int k = 0; k < nrun; k++) {
        int4 u = truns[k];
        for (int t = threadIdx.x; t < u.y; t += blockDim.x)
          tpos[u.z + t] = pos[u.x + t];
      }
    __syncthreads();
    int i = b * IMDG_BLOCK + threadIdx.x;
    if (i - R.atom_off < R.natoms) {
      int cx, cy, cz;
      cell_coords(R, a.acell[i], cx, cy, cz);
      int g0 = R.gdim[0], g1 = R.gdim[1], g2 = R.gdim[2];
      double4 xi = pos[i];
      double cs = R.cellsz;
      int cnt = 0;
      for (int ox = -1; ox <= 1; ox++) {
        int n0 = cx + ox, sx = 0;
        if (n0 < 0) {
          if (R.pbc[0] != 1) continue;
          n0 += g0;
          sx = -1;
        } else if (n0 >= g0) {
          if (R.pbc[0] != 1) continue;
          n0 -= g0;
          sx = 1;
        }
        for (int oy = -1; oy <= 1; oy++) {
          int n1 = cy + oy, sy = 0;
          if (n1 < 0) {
            if (R.pbc[1] != 1) continue;
            n1 += g1;
            sy = -1;
          } else if (n1 >= g1) {
            if (R.pbc[1] != 1) continue;
            n1 -= g1;
            sy = 1;
          }
          for (int oz = -1; oz <= 1; oz++) {
            int n2 = cz + oz, sz = 0;
            if (n2 < 0) {
              if (R.pbc[2] != 1) continue;
              n2 += g2;
              sz = -1;
            } else if (n2 >= g2) {
              if (R.pbc[2] != 1) continue;
              n2 -= g2;
              sz = 1;
            }
            bool self = (ox == 0 && oy == 0 && oz == 0);
            bool owner =
                ox > 0 || (ox == 0 && (oy > 0 || (oy == 0 && oz >= 0)));
            uint32_t code = (uint32_t)((ox + 1) * 9 + (oy + 1) * 3 + (oz + 1));
            int nc = R.cell_off + (n0 * g1 + n1) * g2 + n2;
            int jb = a.cell_beg[nc], je = jb + a.cell_num[nc];
            int tb = (tok && je > jb) ? tile_index(truns, nrun, jb) : jb;
            for (int j = jb; j < je; j++) {
              if (self && j == i) continue;
              double4 xj = tok ? tpos[tb + (j - jb)] : pos[j];
              double d0, d1, d2;
              if (owner) {
                double y0, y1, y2;
                image_of(R, xj.x, xj.y, xj.z, sx, sy, sz, y0, y1, y2);
                d0 = y0 - xi.x;
                d1 = y1 - xi.y;
                d2 = y2 - xi.z;
              } else {
                double y0, y1, y2;
                image_of(R, xi.x, xi.y, xi.z, -sx, -sy, -sz, y0, y1, y2);
                d0 = y0 - xj.x;
                d1 = y1 - xj.y;
                d2 = y2 - xj.z;
              }
              double r2 = d0 * d0 + d1 * d1 + d2 * d2;
              if (r2 < cs) {
                if (cnt < kmax)
                  a.nbr[nbr_index(i, cnt, kmax)] =
                      (uint32_t)(tb + (j - jb)) | (code << IMDG_IDX_BITS);
                cnt++;
              }
            }
          }
        }
      }
      // pad to a whole group with far-away copies of the atom itself
      int self_idx = tok ? tile_index(truns, nrun, i) : i;
      int pend = (cnt + IMDG_GROUP - 1) & ~(IMDG_GROUP - 1);
      for (int q = cnt; q < pend && q < kmax; q++)
        a.nbr[nbr_index(i, q, kmax)] =
            (uint32_t)self_idx | ((uint32_t)IMDG_CODE_PAD << IMDG_IDX_BITS);
      a.nnb[i] = cnt < kmax ? cnt : kmax;
      seen = cnt;
    }
  }
  // block max of the list length, then one atomic per wave
  for (int o = 32; o > 0; o >>= 1)
    seen = max(seen, __shfl_xor(seen, o, IMDG_WAVE));
  if ((threadIdx.x % IMDG_WAVE) == 0 && seen > 0) {
    atomicMax(&a.ctl->kmax_seen, seen);
    if (seen > kmax) atomicExch(&a.ctl->overflow, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned t = atomicAdd(&a.ctl->done[DC_LIST], 1u);
    if (t == gridDim.x - 1) {
      __threadfence();
      if (atomicAdd(&a.ctl->overflow, 0) != 0) {
        // keep the rebuild flags: the host grows the list and builds again
        a.ctl->frozen = 1;
      } else {
        for (int q = 0; q < a.nrep; q++)
          if (a.rst[q].need_rebuild) {
            a.rst[q].rebuilt = 1;
            a.rst[q].nbl_count += 1;
            a.rst[q].need_rebuild = 0;
          }
        a.ctl->any_rebuild = 0;
      }
      a.ctl->done[DC_LIST] = 0;
      __threadfence();
    }
  }
}

// ---------------------------------------------------------------------------
// per-step kernels
// ---------------------------------------------------------------------------

// Cell-offset vectors of the block's replica in the force precision.
template <typename R>
__device__ __forceinline__ void load_offsets(const RepDev& Rp, R* off) {
  for (int k = threadIdx.x; k < 3 * IMDG_NCODES; k += blockDim.x)
    off[k] = (R)((&Rp.offs[0][0])[k]);
}

// Copy the block's tile of records into LDS (coalesced run copies).
template <typename R4>
__device__ __forceinline__ int stage_tile(const Tile* tiles, int b,
                                          const R4* __restrict__ src, R4* dst,
                                          int4* runs, int* nrun_out) {
  const Tile& T = tiles[b];
  int nrun = T.nrun, ok = T.ok;
  for (int k = threadIdx.x; k < nrun; k += blockDim.x) runs[k] = T.run[k];
  __syncthreads();
  if (ok)
    for (int k = 0; k < nrun; k++) {
      int4 u = runs[k];
      for (int t = threadIdx.x; t < u.y; t += blockDim.x)
        dst[u.z + t] = src[u.x + t];
    }
  __syncthreads();
  *nrun_out = nrun;
  return ok;
}

// Separation of a listed neighbour: (u_j - u_i) + offset(code).
template <typename R, typename R4>
__device__ __forceinline__ void separation(const R4& ui, const R4& uj,
                                           const R* off, uint32_t code,
                                           R& dx, R& dy, R& dz) {
  const R* o = off + 3 * code;
  dx = (uj.x - ui.x) + o[0];
  dy = (uj.y - ui.y) + o[1];
  dz = (uj.z - ui.z) + o[2];
}

// Host density, embedding energy and 2F'(rho) (src/imd_forces_nbl.c:586-611,
// 1079-1112).  One thread per atom, full list, neighbour records gathered
// from the block's LDS tile, four neighbours per step.
template <typename R, bool NT1>
__global__ void __launch_bounds__(IMDG_BLOCK)
    k_density(StepArgs<R> a, const TabDev<R>* __restrict__ tabs) {
  typedef typename V4<R>::T R4;
  __shared__ R4 trec[IMDG_TILE_CAP];
  __shared__ int4 truns[IMDG_MAXRUN];
  __shared__ R off[3 * IMDG_NCODES];
  __shared__ double red[IMDG_NWAVE];
  if (a.ctl->frozen) return;
  int lb = a.xcd_remap ? logical_block() : (int)blockIdx.x;
  int r = a.blk_rep[lb];
  const RepDev& Rp = a.rep[r];
  load_offsets(Rp, off);
  const R4* __restrict__ reca = a.rec_a;
  int nrun;
  bool tok = stage_tile(a.tiles, lb, reca, trec, truns, &nrun) != 0;
  const R4* __restrict__ src = tok ? (const R4*)trec : reca;
  int i = lb * IMDG_BLOCK + threadIdx.x;
  bool active = (i - Rp.atom_off) < Rp.natoms;
  const Seg<R>* __restrict__ segs = a.seg;
  const TabDev<R>& tr = tabs[1];
  const TabDev<R>& te = tabs[2];
  double emb = 0.0;
  int is_short = 0;
  if (active) {
    R4 ui = reca[i];
    int it = NT1 ? 0 : (int)ui.w;
    int ng = (a.nnb[i] + 3) >> 2;
    R rho = (R)0;
    const uint4* __restrict__ nb = list_of(a.nbr, i, a.kmax);
    if (NT1) {
      const Seg<R>* __restrict__ seg = segs + tr.off[0];
      int len = tr.len[0];
      R b = tr.begin[0], e = tr.end[0], is = tr.invstep[0];
      for (int g = 0; g < ng; g++) {
        uint4 e4 = nb[(size_t)g * IMDG_WAVE];
        uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
        R r2[4], chi[4];
        int k[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          R4 uj = src[en[u] & IMDG_IDX_MASK];
          R dx, dy, dz;
          separation(ui, uj, off, en[u] >> IMDG_IDX_BITS, dx, dy, dz);
          r2[u] = dx * dx + dy * dy + dz * dz;
          k[u] = seg_locate(r2[u], b, e, is, len, chi[u]);
        }
        Seg<R> c[4];
#pragma unroll
        for (int u = 0; u < 4; u++) c[u] = seg[k[u]];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          bool in = r2[u] < e;
          rho += in ? seg_val(c[u], chi[u]) : (R)0;
          is_short |= (in && r2[u] < b);
        }
      }
    } else {
      int nt = a.ntypes;
      for (int g = 0; g < ng; g++) {
        uint4 e4 = nb[(size_t)g * IMDG_WAVE];
        uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
        for (int u = 0; u < 4; u++) {
          R4 uj = src[en[u] & IMDG_IDX_MASK];
          R dx, dy, dz;
          separation(ui, uj, off, en[u] >> IMDG_IDX_BITS, dx, dy, dz);
          R r2 = dx * dx + dy * dy + dz * dz;
          int col = it * nt + (int)uj.w;
          R e = tr.end[col], b = tr.begin[col], chi;
          int k = seg_locate(r2, b, e, tr.invstep[col], tr.len[col], chi);
          R v = seg_val(segs[tr.off[col] + k], chi);
          bool in = r2 < e;
          rho += in ? v : (R)0;
          is_short |= (in && r2 < b);
        }
      }
    }
    // embedding: PAIR_INT2 on embed_pot, column = type, argument rho
    R F, dF;
    int idummy = 0;
    tab_val_grad(segs + te.off[it], te.len[it], te.begin[it], te.end[it],
                 te.invstep[it], te.twoistep[it], rho, F, dF, idummy);
    R4 rb;
    rb.x = ui.x;
    rb.y = ui.y;
    rb.z = ui.z;
    rb.w = dF;
    a.rec_b[i] = rb;
    a.rho[i] = rho;
    a.epot[i] = F;
    emb = (double)F;
  }
  double v[1] = {emb};
  block_reduce<1>(v, 0u, red);
  if (threadIdx.x == 0) a.partA[lb] = v[0];
  if (is_short) atomicOr(&a.ctl->is_short, 1);
}

// Finalisation by the last block of a step kernel: compensated sums of the
// block partials of every replica, thermostat, skin decision, thermo record.
template <typename R>
__device__ void finalize_step(StepArgs<R>& a, bool forces, bool moved) {
  __shared__ double lred[IMDG_NWAVE][PB_N + 1][2];
  int w = threadIdx.x / IMDG_WAVE, l = threadIdx.x % IMDG_WAVE;
  bool any = false;
  for (int r = 0; r < a.nrep; r++) {
    const RepDev& Rp = a.rep[r];
    // each thread: Neumaier sums over a strided subset of the blocks
    double s[PB_N + 1], c[PB_N + 1];
#pragma unroll
    for (int q = 0; q <= PB_N; q++) s[q] = c[q] = 0.0;
    for (int b = Rp.blk_off + threadIdx.x; b < Rp.blk_off + Rp.nblk;
         b += blockDim.x) {
      const double* pb = a.partB + (size_t)b * PB_N;
#pragma unroll
      for (int q = 0; q < PB_N; q++) {
        if (q == PB_MAXD)
          s[q] = fmax(s[q], pb[q]);
        else
          kahan_add(s[q], c[q], pb[q]);
      }
      if (forces && a.has_eam) kahan_add(s[PB_N], c[PB_N], a.partA[b]);
    }
    // combine across the block in a fixed order (TwoSum keeps the error)
#pragma unroll
    for (int q = 0; q <= PB_N; q++) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        double so = __shfl_xor(s[q], o, IMDG_WAVE);
        double co = __shfl_xor(c[q], o, IMDG_WAVE);
        if (q == PB_MAXD) {
          s[q] = fmax(s[q], so);
        } else {
          double t, e;
          two_sum(s[q], so, t, e);
          s[q] = t;
          c[q] = c[q] + co + e;
        }
      }
      if (l == 0) {
        lred[w][q][0] = s[q];
        lred[w][q][1] = c[q];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot[PB_N + 1];
#pragma unroll
      for (int q = 0; q <= PB_N; q++) {
        double ss = lred[0][q][0], cc = lred[0][q][1];
#pragma unroll
        for (int k = 1; k < IMDG_NWAVE; k++) {
          if (q == PB_MAXD) {
            ss = fmax(ss, lred[k][q][0]);
          } else {
            double t, e;
            two_sum(ss, lred[k][q][0], t, e);
            ss = t;
            cc = cc + lred[k][q][1] + e;
          }
        }
        tot[q] = (q == PB_MAXD) ? ss : ss + cc;
      }
      RepState& st = a.rst[r];
      if (forces) {
        st.epot = tot[PB_EPAIR] + tot[PB_N];
        st.virial = tot[PB_VIRIAL];
      }
      if (moved) {
        double ekin;
        if (a.ensemble == 0) {
          ekin = tot[PB_KE];
        } else {
          ekin = (tot[PB_E1] + tot[PB_E2]) / 4.0;
          double ttt = Rp.nactive * Rp.temperature;
          st.eta += a.dt * (tot[PB_E2] / ttt - 1.0) * Rp.isq_tau_eta;
        }
        int need = tot[PB_MAXD] > Rp.skin2;
        st.need_rebuild = need;
        any |= need;
        long step = a.ctl->steps_done;
        ThermoRec& tr = a.ring[(size_t)(step % IMDG_RING) * a.nrep + r];
        tr.epot = st.epot;
        tr.ekin = ekin;
        tr.virial = st.virial;
        tr.eta = st.eta;
        tr.max_disp2 = tot[PB_MAXD];
        tr.step = step;
        tr.rebuilt = st.rebuilt;
        tr.natoms = Rp.natoms;
        st.rebuilt = 0;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (moved) {
      a.ctl->any_rebuild = any ? 1 : 0;
      a.ctl->steps_done += 1;
    }
    __threadfence();
  }
}

template <typename R>
__device__ __forceinline__ void last_block_finalize(StepArgs<R>& a,
                                                    bool forces, bool moved) {
  __shared__ int is_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned t = atomicAdd(&a.ctl->done[DC_FINAL], 1u);
    is_last = (t == gridDim.x - 1);
    if (is_last) a.ctl->done[DC_FINAL] = 0;
  }
  __syncthreads();
  if (is_last) {
    __threadfence();
    finalize_step(a, forces, moved);
  }
}

// Leapfrog on momenta (src/imd_integrate.c:213-217, 329-358 and
// 1020-1056), FP64, separately rounded like the reference.
__device__ __forceinline__ void leapfrog(int ensemble, double dt, double eta,
                                         double fx, double fy, double fz,
                                         double4& p, double& x0, double& x1,
                                         double& x2, double& ke, double& e1,
                                         double& e2) {
#pragma clang fp contract(off)
  double m = p.w;
  if (ensemble == 0) {
    double k1 = p.x * p.x + p.y * p.y + p.z * p.z;
    p.x += dt * fx;
    p.y += dt * fy;
    p.z += dt * fz;
    double k2 = p.x * p.x + p.y * p.y + p.z * p.z;
    ke = (k1 + k2) / (4 * m);
    e1 = e2 = 0.0;
  } else {
    double reib = 1.0 - eta * dt / 2.0;
    double einsd = 1.0 / (1.0 + eta * dt / 2.0);
    e1 = (p.x * p.x + p.y * p.y + p.z * p.z) / m;
    p.x = (p.x * reib + dt * fx) * einsd;
    p.y = (p.y * reib + dt * fy) * einsd;
    p.z = (p.z * reib + dt * fz) * einsd;
    e2 = (p.x * p.x + p.y * p.y + p.z * p.z) / m;
    ke = 0.0;
  }
  double tmp = dt / m;
  x0 += tmp * p.x;
  x1 += tmp * p.y;
  x2 += tmp * p.z;
}

__device__ __forceinline__ double disp2(double4 a, double4 b) {
#pragma clang fp contract(off)
  double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
  return d0 * d0 + d1 * d1 + d2 * d2;
}

// Integrate slot i, refresh its cell-relative record and return the
// partials (kinetic terms, displacement).
template <typename R>
__device__ __forceinline__ void integrate_slot(StepArgs<R>& a,
                                               const RepDev& Rp, int r, int i,
                                               double fx, double fy, double fz,
                                               double* v) {
  typedef typename V4<R>::T R4;
  int pc = a.ctl->pos_cur;
  double4* pos = a.pos[pc];
  double4* mom = a.mom[pc];
  double4 xi = pos[i];
  double4 p = mom[i];
  double x0 = xi.x, x1 = xi.y, x2 = xi.z, ke, e1, e2;
  leapfrog(a.ensemble, a.dt, a.rst[r].eta, fx, fy, fz, p, x0, x1, x2, ke, e1,
           e2);
  mom[i] = p;
  double4 xn = make_double4(x0, x1, x2, xi.w);
  pos[i] = xn;
  double c[3];
  cell_centre(Rp, a.acell[i], c);
  R4 ra;
  ra.x = (R)(x0 - c[0]);
  ra.y = (R)(x1 - c[1]);
  ra.z = (R)(x2 - c[2]);
  ra.w = (R)xi.w;
  a.rec_a[i] = ra;
  v[PB_KE] = ke;
  v[PB_E1] = e1;
  v[PB_E2] = e2;
  v[PB_MAXD] = disp2(xn, a.xnbl[i]);
}

// Pair + EAM forces (src/imd_forces_nbl.c:490-584, 1117-1322) as a gather
// over the full list from the block's LDS tile of (u, 2F') records; with
// INTEGRATE the leapfrog step, the skin test and the step's reductions
// follow in the same thread (calc_forces, move_atoms and check_nblist of one
// main_loop iteration in one launch).
template <typename R, bool NT1, bool INTEGRATE, bool HP, bool HE>
__global__ void __launch_bounds__(IMDG_BLOCK)
    k_force(StepArgs<R> a, const TabDev<R>* __restrict__ tabs) {
  typedef typename V4<R>::T R4;
  __shared__ R4 trec[IMDG_TILE_CAP];
  __shared__ int4 truns[IMDG_MAXRUN];
  __shared__ R off[3 * IMDG_NCODES];
  __shared__ double red[IMDG_NWAVE * PB_N];
  if (a.ctl->frozen) return;
  int lb = a.xcd_remap ? logical_block() : (int)blockIdx.x;
  int r = a.blk_rep[lb];
  const RepDev& Rp = a.rep[r];
  load_offsets(Rp, off);
  // the gathered records (rec_b): (u, 2F') from the density loop with EAM,
  // a copy of (u, type) without; the integration below rewrites rec_a
  const R4* __restrict__ recg = a.rec_b;
  int nrun;
  bool tok = stage_tile(a.tiles, lb, recg, trec, truns, &nrun) != 0;
  const R4* __restrict__ src = tok ? (const R4*)trec : recg;
  int i = lb * IMDG_BLOCK + threadIdx.x;
  bool active = (i - Rp.atom_off) < Rp.natoms;
  const Seg<R>* __restrict__ segs = a.seg;
  const signed char* __restrict__ atype = a.atype;
  const TabDev<R>& tp = tabs[0];
  const TabDev<R>& tr = tabs[1];
  double v[PB_N] = {0, 0, 0, 0, 0, 0};
  int is_short = 0;
  R fx = 0, fy = 0, fz = 0;
  if (active) {
    R4 ui = recg[i];
    int it = NT1 ? 0 : (int)atype[i];
    int ng = (a.nnb[i] + 3) >> 2;
    R ep = 0, vir = 0;
    R dFi = HE ? ui.w : (R)0;
    const uint4* __restrict__ nb = list_of(a.nbr, i, a.kmax);
    if (NT1) {
      const Seg<R>* __restrict__ ps = segs + tp.off[0];
      const Seg<R>* __restrict__ rs = segs + tr.off[0];
      int plen = tp.len[0], rlen = tr.len[0];
      R pb = tp.begin[0], pe = tp.end[0], pis = tp.invstep[0],
        p2is = tp.twoistep[0];
      R rb = tr.begin[0], re = tr.end[0], ris = tr.invstep[0],
        r2is = tr.twoistep[0];
      for (int g = 0; g < ng; g++) {
        uint4 e4 = nb[(size_t)g * IMDG_WAVE];
        uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
        R dx[4], dy[4], dz[4], r2[4], cp[4], cr[4], dfj[4];
        int kp[4], kr[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          R4 uj = src[en[u] & IMDG_IDX_MASK];
          separation(ui, uj, off, en[u] >> IMDG_IDX_BITS, dx[u], dy[u],
                     dz[u]);
          dfj[u] = uj.w;
          r2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
          if constexpr (HP) kp[u] = seg_locate(r2[u], pb, pe, pis, plen, cp[u]);
          if constexpr (HE) kr[u] = seg_locate(r2[u], rb, re, ris, rlen, cr[u]);
        }
        Seg<R> sp[4], sr[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if constexpr (HP) sp[u] = ps[kp[u]];
          if constexpr (HE) sr[u] = rs[kr[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          R gg = (R)0;
          if constexpr (HP) {
            bool inp = r2[u] <= pe;
            gg += inp ? seg_grad(sp[u], cp[u], p2is) : (R)0;
            ep += inp ? seg_val(sp[u], cp[u]) : (R)0;
            is_short |= (inp && r2[u] < pb);
          }
          if constexpr (HE) {
            bool inr = r2[u] < re;
            gg += inr ? (R)0.5 * (dFi + dfj[u]) * seg_grad(sr[u], cr[u], r2is)
                      : (R)0;
            is_short |= (inr && r2[u] < rb);
          }
          fx += dx[u] * gg;
          fy += dy[u] * gg;
          fz += dz[u] * gg;
          vir += r2[u] * gg;
        }
      }
    } else {
      int nt = a.ntypes;
      for (int g = 0; g < ng; g++) {
        uint4 e4 = nb[(size_t)g * IMDG_WAVE];
        uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
        for (int u = 0; u < 4; u++) {
          int jidx = (int)(en[u] & IMDG_IDX_MASK);
          R4 uj = src[jidx];
          R dx, dy, dz;
          separation(ui, uj, off, en[u] >> IMDG_IDX_BITS, dx, dy, dz);
          R r2 = dx * dx + dy * dy + dz * dz;
          // the list holds tile positions; the runs map them to slots
          int js = tok ? tile_slot(truns, nrun, jidx) : jidx;
          int jt = (int)atype[js];
          int col = it * nt + jt, colT = jt * nt + it;
          R g = 0;
          if (HP && r2 <= tp.end[col]) {
            R pot, gp;
            tab_val_grad(segs + tp.off[col], tp.len[col], tp.begin[col],
                         tp.end[col], tp.invstep[col], tp.twoistep[col], r2,
                         pot, gp, is_short);
            ep += pot;
            g += gp;
          }
          if (HE && (r2 < tr.end[colT] || r2 < tr.end[col])) {
            // rho'_i from column jt*nt+it, rho'_j from it*nt+jt
            // (src/imd_forces_nbl.c:1156-1203)
            R chi;
            int k = seg_locate(r2, tr.begin[colT], tr.end[colT],
                               tr.invstep[colT], tr.len[colT], chi);
            is_short |= (r2 < tr.begin[colT]);
            R rpi = seg_grad(segs[tr.off[colT] + k], chi, tr.twoistep[colT]);
            R rpj = rpi;
            if (colT != col) {
              k = seg_locate(r2, tr.begin[col], tr.end[col], tr.invstep[col],
                             tr.len[col], chi);
              is_short |= (r2 < tr.begin[col]);
              rpj = seg_grad(segs[tr.off[col] + k], chi, tr.twoistep[col]);
            }
            g += (R)0.5 * (dFi * rpj + uj.w * rpi);
          }
          fx += dx * g;
          fy += dy * g;
          fz += dz * g;
          vir += r2 * g;
        }
      }
    }
    // per-atom energy: half of every pair term plus F(rho_i)
    R ei = (R)0.5 * ep + (HE ? a.epot[i] : (R)0);
    a.epot[i] = ei;
    R4 fv;
    fv.x = fx;
    fv.y = fy;
    fv.z = fz;
    fv.w = (R)0;
    a.frc[i] = fv;
    v[PB_EPAIR] = 0.5 * (double)ep;
    v[PB_VIRIAL] = -0.5 * (double)vir;
    if (INTEGRATE)
      integrate_slot(a, Rp, r, i, (double)fx, (double)fy, (double)fz, v);
  }
  block_reduce<PB_N>(v, 1u << PB_MAXD, red);
  if (threadIdx.x == 0) {
    double* pbk = a.partB + (size_t)lb * PB_N;
#pragma unroll
    for (int q = 0; q < PB_N; q++) pbk[q] = v[q];
  }
  if (is_short) atomicOr(&a.ctl->is_short, 1);
  last_block_finalize(a, true, INTEGRATE);
}

// move_atoms + check_nblist on the stored forces (the unfused entry point).
template <typename R>
__global__ void __launch_bounds__(IMDG_BLOCK) k_move(StepArgs<R> a) {
  typedef typename V4<R>::T R4;
  __shared__ double red[IMDG_NWAVE * PB_N];
  if (a.ctl->frozen) return;
  int lb = a.xcd_remap ? logical_block() : (int)blockIdx.x;
  int r = a.blk_rep[lb];
  const RepDev& Rp = a.rep[r];
  int i = lb * IMDG_BLOCK + threadIdx.x;
  bool active = (i - Rp.atom_off) < Rp.natoms;
  double v[PB_N] = {0, 0, 0, 0, 0, 0};
  if (active) {
    R4 f = a.frc[i];
    integrate_slot(a, Rp, r, i, (double)f.x, (double)f.y, (double)f.z, v);
  }
  block_reduce<PB_N>(v, 1u << PB_MAXD, red);
  if (threadIdx.x == 0) {
    double* pbk = a.partB + (size_t)lb * PB_N;
    // keep the force-kernel sums of this step (energy, virial)
    pbk[PB_KE] = v[PB_KE];
    pbk[PB_E1] = v[PB_E1];
    pbk[PB_E2] = v[PB_E2];
    pbk[PB_MAXD] = v[PB_MAXD];
  }
  last_block_finalize(a, false, true);
}

// Pair-only systems: the force loop gathers rec_b, which holds a copy of the
// (u, type) records (rec_a is rewritten by the integration in that loop).
template <typename R>
__global__ void __launch_bounds__(256) k_copyrec(StepArgs<R> a, int nslots) {
  if (a.ctl->frozen) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslots) a.rec_b[i] = a.rec_a[i];
}

// check_nblist alone: max displacement per block.
__global__ void __launch_bounds__(IMDG_BLOCK)
    k_check(const RepDev* rep, Ctl* ctl, const int* blk_rep,
            const double4* pos0, const double4* pos1, const double4* xnbl,
            double* maxd) {
  __shared__ double red[IMDG_NWAVE];
  if (ctl->frozen) return;
  int lb = blockIdx.x;
  int r = blk_rep[lb];
  const RepDev& Rp = rep[r];
  int i = lb * IMDG_BLOCK + threadIdx.x;
  const double4* pos = ctl->pos_cur ? pos1 : pos0;
  double v[1] = {0.0};
  if ((i - Rp.atom_off) < Rp.natoms) v[0] = disp2(pos[i], xnbl[i]);
  block_reduce<1>(v, 1u, red);
  if (threadIdx.x == 0) maxd[lb] = v[0];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <typename R>
void launch_density(const StepArgs<R>& a, const TabDev<R>* tabs, bool nt1,
                    int nblk, hipStream_t st) {
  if (nt1)
    hipLaunchKernelGGL((k_density<R, true>), dim3(nblk), dim3(IMDG_BLOCK), 0,
                       st, a, tabs);
  else
    hipLaunchKernelGGL((k_density<R, false>), dim3(nblk), dim3(IMDG_BLOCK), 0,
                       st, a, tabs);
}

template <typename R, bool INTEGRATE>
void launch_force_i(const StepArgs<R>& a, const TabDev<R>* tabs, bool nt1,
                    int nblk, hipStream_t st) {
  dim3 g(nblk), b(IMDG_BLOCK);
  if (!nt1) {
    if (a.has_pair && a.has_eam)
      hipLaunchKernelGGL((k_force<R, false, INTEGRATE, true, true>), g, b, 0,
                         st, a, tabs);
    else if (a.has_pair)
      hipLaunchKernelGGL((k_force<R, false, INTEGRATE, true, false>), g, b, 0,
                         st, a, tabs);
    else
      hipLaunchKernelGGL((k_force<R, false, INTEGRATE, false, true>), g, b, 0,
                         st, a, tabs);
  } else if (a.has_pair && a.has_eam) {
    hipLaunchKernelGGL((k_force<R, true, INTEGRATE, true, true>), g, b, 0, st,
                       a, tabs);
  } else if (a.has_pair) {
    hipLaunchKernelGGL((k_force<R, true, INTEGRATE, true, false>), g, b, 0, st,
                       a, tabs);
  } else {
    hipLaunchKernelGGL((k_force<R, true, INTEGRATE, false, true>), g, b, 0, st,
                       a, tabs);
  }
}

template <typename R>
void launch_force(const StepArgs<R>& a, const TabDev<R>* tabs, bool nt1,
                  bool integrate, int nblk, hipStream_t st) {
  if (integrate)
    launch_force_i<R, true>(a, tabs, nt1, nblk, st);
  else
    launch_force_i<R, false>(a, tabs, nt1, nblk, st);
}

template <typename R>
void launch_move(const StepArgs<R>& a, int nblk, hipStream_t st) {
  hipLaunchKernelGGL((k_move<R>), dim3(nblk), dim3(IMDG_BLOCK), 0, st, a);
}

template void launch_density<float>(const StepArgs<float>&,
                                    const TabDev<float>*, bool, int,
                                    hipStream_t);
template void launch_density<double>(const StepArgs<double>&,
                                     const TabDev<double>*, bool, int,
                                     hipStream_t);
template void launch_force<float>(const StepArgs<float>&,
                                  const TabDev<float>*, bool, bool, int,
                                  hipStream_t);
template void launch_force<double>(const StepArgs<double>&,
                                   const TabDev<double>*, bool, bool, int,
                                   hipStream_t);
template void launch_move<float>(const StepArgs<float>&, int, hipStream_t);
template void launch_move<double>(const StepArgs<double>&, int, hipStream_t);

template <typename R>
void launch_copyrec(const StepArgs<R>& a, int nslots, hipStream_t st) {
  hipLaunchKernelGGL((k_copyrec<R>), dim3((nslots + 255) / 256), dim3(256), 0,
                     st, a, nslots);
}
template void launch_copyrec<float>(const StepArgs<float>&, int, hipStream_t);
template void launch_copyrec<double>(const StepArgs<double>&, int,
                                     hipStream_t);

void launch_build(const BuildArgs& b, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_bin, dim3(grid), dim3(256), 0, st, b);
  hipLaunchKernelGGL(k_scan, dim3(b.nrep), dim3(1024), 0, st, b);
  hipLaunchKernelGGL(k_place, dim3(grid), dim3(256), 0, st, b);
  hipLaunchKernelGGL(k_cellsort, dim3(grid), dim3(256), 0, st, b);
  hipLaunchKernelGGL(k_permute, dim3(grid), dim3(256), 0, st, b);
  int gt = (b.nblk + 63) / 64;
  hipLaunchKernelGGL(k_tiles, dim3(gt < 1 ? 1 : gt), dim3(64), 0, st, b);
  hipLaunchKernelGGL(k_list, dim3(b.nblk), dim3(IMDG_BLOCK), 0, st, b);
}

void launch_check(const RepDev* rep, Ctl* ctl, const int* blk_rep,
                  const double4* pos0, const double4* pos1,
                  const double4* xnbl, double* maxd, int nblk,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_check, dim3(nblk), dim3(IMDG_BLOCK), 0, st, rep, ctl,
                     blk_rep, pos0, pos1, xnbl, maxd);
}

}  // namespace imdg
