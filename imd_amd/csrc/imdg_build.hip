// imdg_build.hip -- CDNA4 (gfx950) kernels of the list rebuild: wrap and
// cell binning, per-replica scan, cell sort by input index, permutation of
// the state into cell order, per-block neighbour tiles and the full
// neighbour list.  Every kernel of the chain is launched every step and
// returns at its first instructions unless the skin test of the previous
// step (or the host) asked for a rebuild, so the fused loop never waits for
// the host.
//
// Reference semantics (paths relative to the reference root):
//   binning / wrap     src/imd_fix_cells_3d.c:36-201, src/imd_main_3d.c:1972-2065,
//                      cell_coord src/imd_geom_3d.c:1054-1074
//   list               make_nblist src/imd_forces_nbl.c:136-273 (half list over
//                      the 14-cell stencil of src/imd_geom_3d.c:895-899 with
//                      buffer-cell images src/imd_comm_force_3d.c:222-397);
//                      built here as a full list whose r^2 is decided in the
//                      same direction and operation order as the reference, so
//                      the pair set is identical
//   skin test          check_nblist src/imd_forces_nbl.c:2007-2037
#include "imdg_device.h"

namespace imdg {

// Blocks of the per-slot rebuild kernels cover IMDG_BLOCK slots of one
// replica, so the replica's geometry is wave-uniform (scalar loads).
__device__ __forceinline__ int block_replica(const int* blk_rep) {
  return __builtin_amdgcn_readfirstlane(blk_rep[blockIdx.x]);
}

// fix_cells/do_boundaries + cell_coord, for replicas that need a rebuild.
// The atoms of a block are counted per cell in an LDS hash table first, so
// the global counters see one atomic per distinct cell of the block (the
// slot of an atom inside its cell is base + its rank in the block; the
// order inside a cell is fixed later by k_cellsort).
#define BIN_HASH 256

// Flag bit of a block's list-length maximum (blkmax): its tile exceeded the
// tile cap and the list build read its candidates from HBM/L2.
#define IMDG_BLK_FALLBACK (1 << 30)

__global__ void __launch_bounds__(IMDG_BLOCK) k_bin(BuildArgs a) {
#pragma clang fp contract(off)
  __shared__ int hkey[BIN_HASH], hcnt[BIN_HASH], hbase[BIN_HASH];
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  int r = block_replica(a.blk_rep);
  if (!a.rst[r].need_rebuild) return;
  const RepDev& R = a.rep[r];
  for (int e = threadIdx.x; e < BIN_HASH; e += IMDG_BLOCK) {
    hkey[e] = -1;
    hcnt[e] = 0;
  }
  __syncthreads();
  int i = blockIdx.x * IMDG_BLOCK + threadIdx.x;
  bool active = i - R.atom_off < R.natoms;
  int h = -1, rank = 0;
  if (active) {
    double4* pos = a.pos[a.ctl->pos_cur];
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
    // at most IMDG_BLOCK distinct keys in BIN_HASH entries: probing ends
    h = (int)(((unsigned)cell * 2654435761u) >> 24) & (BIN_HASH - 1);
    for (;;) {
      int old = atomicCAS(&hkey[h], -1, cell);
      if (old == -1 || old == cell) break;
      h = (h + 1) & (BIN_HASH - 1);
    }
    rank = atomicAdd(&hcnt[h], 1);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < BIN_HASH; e += IMDG_BLOCK)
    if (hkey[e] >= 0) hbase[e] = atomicAdd(&a.cell_cnt[hkey[e]], hcnt[e]);
  __syncthreads();
  if (active) a.cslot[i] = hbase[h] + rank;
}

// Exclusive scan of the cell counts, two levels in one launch: every block
// scans one tile of IMDG_SCAN_TILE cells of a replica (cell_loc: offset
// inside the tile, tile_sum: the tile's total); the last block to finish
// scans the tile totals of every rebuilt replica (tile_off: first slot of
// the tile's first cell).
__global__ void __launch_bounds__(IMDG_SCAN_TILE) k_scan(BuildArgs a) {
  __shared__ int wsum[IMDG_SCAN_TILE / IMDG_WAVE];
  __shared__ int is_last;
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  int t = blockIdx.x;
  int r = __builtin_amdgcn_readfirstlane(a.tile_rep[t]);
  const RepDev& R = a.rep[r];
  int lane = threadIdx.x % IMDG_WAVE, w = threadIdx.x / IMDG_WAVE;
  if (a.rst[r].need_rebuild) {
    int lc = (t - R.stile_off) * IMDG_SCAN_TILE + threadIdx.x;
    int cell = R.cell_off + lc;
    int n = lc < R.ncells ? a.cell_cnt[cell] : 0;
    // inclusive scan inside the wave, then over the wave totals
    int v = n;
#pragma unroll
    for (int o = 1; o < IMDG_WAVE; o <<= 1) {
      int y = __shfl_up(v, o, IMDG_WAVE);
      if (lane >= o) v += y;
    }
    if (lane == IMDG_WAVE - 1) wsum[w] = v;
    __syncthreads();
    if (w == 0) {
      int s = lane < IMDG_SCAN_TILE / IMDG_WAVE ? wsum[lane] : 0;
#pragma unroll
      for (int o = 1; o < IMDG_WAVE; o <<= 1) {
        int y = __shfl_up(s, o, IMDG_WAVE);
        if (lane >= o) s += y;
      }
      if (lane < IMDG_SCAN_TILE / IMDG_WAVE) wsum[lane] = s;
    }
    __syncthreads();
    int before = (w > 0 ? wsum[w - 1] : 0) + v - n;
    if (lc < R.ncells) a.cell_loc[cell] = before;
    if (threadIdx.x == IMDG_SCAN_TILE - 1)
      a.tile_sum[t] = wsum[IMDG_SCAN_TILE / IMDG_WAVE - 1];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned d = atomicAdd(&a.ctl->done[DC_SCAN], 1u);
    is_last = (d == gridDim.x - 1);
    if (is_last) a.ctl->done[DC_SCAN] = 0;
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  // one wavefront per replica: exclusive scan of its tile totals
  for (int q = w; q < a.nrep; q += IMDG_SCAN_TILE / IMDG_WAVE) {
    if (!a.rst[q].need_rebuild) continue;
    const RepDev& Q = a.rep[q];
    int run = Q.atom_off;
    for (int k0 = 0; k0 < Q.nstile; k0 += IMDG_WAVE) {
      int k = k0 + lane;
      int n = k < Q.nstile ? __atomic_load_n(&a.tile_sum[Q.stile_off + k],
                                              __ATOMIC_RELAXED)
                           : 0;
      int v = n;
#pragma unroll
      for (int o = 1; o < IMDG_WAVE; o <<= 1) {
        int y = __shfl_up(v, o, IMDG_WAVE);
        if (lane >= o) v += y;
      }
      if (k < Q.nstile) a.tile_off[Q.stile_off + k] = run + v - n;
      run += __shfl(v, IMDG_WAVE - 1, IMDG_WAVE);
    }
  }
}

// First slot of a cell of a rebuilt replica (after k_scan).
__device__ __forceinline__ int cell_first_slot(const BuildArgs& a,
                                               const RepDev& R, int cell) {
  int lc = cell - R.cell_off;
  return a.tile_off[R.stile_off + lc / IMDG_SCAN_TILE] + a.cell_loc[cell];
}

__global__ void __launch_bounds__(IMDG_BLOCK) k_place(BuildArgs a) {
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  int r = block_replica(a.blk_rep);
  if (!a.rst[r].need_rebuild) return;
  const RepDev& R = a.rep[r];
  int i = blockIdx.x * IMDG_BLOCK + threadIdx.x;
  if (i - R.atom_off >= R.natoms) return;
  int cell = a.cellid[i];
  a.perm[cell_first_slot(a, R, cell) + a.cslot[i]] = i;
}

// Per cell: first slot and count, counter reset, and the atoms of the cell
// ordered by input index (deterministic list order).  32 lanes per cell:
// each lane takes one atom and writes it at its rank among the cell's input
// indices (unique); a cell with more atoms is sorted by its first lane.
#define SORT_LANES 32

__global__ void __launch_bounds__(256) k_cellsort(BuildArgs a) {
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  const int* aid = a.aid[a.ctl->rec_cur];
  int seg = threadIdx.x / SORT_LANES, l = threadIdx.x % SORT_LANES;
  int base = (threadIdx.x % IMDG_WAVE) - l;  // first lane of the segment
  for (int c = blockIdx.x * (blockDim.x / SORT_LANES) + seg; c < a.ncells;
       c += gridDim.x * (blockDim.x / SORT_LANES)) {
    int r = a.cell_rep[c];
    if (!a.rst[r].need_rebuild) continue;  // the whole segment skips
    const RepDev& R = a.rep[r];
    int b = cell_first_slot(a, R, c), n = a.cell_cnt[c];
    int* p = a.perm + b;
    if (l == 0) {
      a.cell_beg[c] = b;
      a.cell_num[c] = n;
      a.cell_cnt[c] = 0;
    }
    if (n <= SORT_LANES) {
      int v = l < n ? p[l] : 0;
      int key = l < n ? aid[v] : 0x7fffffff;
      int rank = 0;
      for (int m = 0; m < n; m++) {
        int km = __shfl(key, base + m, IMDG_WAVE);
        rank += km < key;
      }
      if (l < n) p[rank] = v;
    } else if (l == 0) {
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
}

// Tiles: for every block of a rebuilt replica, the slot runs of all cells
// adjacent (27-stencil, periodic wrap, nothing beyond an open face) to the
// cells of the block's atoms, sorted and merged.  One wavefront per block
// (the first of the block's permutation launch): lane l handles neighbour
// column l % 9 of the block's column l / 9 (at most 7 columns; more falls
// back to a tile-less list), producing at most two slot ranges (a z range
// split by the periodic wrap); the wavefront sorts and merges them.
#define TILE_MAXRANGE 126
#define TILE_MAXCOL 7

// LDS rows of one wavefront's tile construction.
struct TileLds {
  int rb[TILE_MAXRANGE], re[TILE_MAXRANGE];
  int4 run[IMDG_MAXRUN];
};

__device__ void build_tile(const BuildArgs& a, int b, TileLds& S);

// Gather every slot into the other buffers (new cell-sorted order for the
// rebuilt replicas, identity for the others) and write the cell-relative
// records of the moved atoms.  The buffers are flipped by k_list_done once
// the list fits (on an overflow the whole chain runs again from the old
// buffers).
__global__ void __launch_bounds__(IMDG_BLOCK) k_permute(BuildArgs a) {
  __shared__ TileLds tl;
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  int pc = a.ctl->pos_cur, rc = a.ctl->rec_cur;
  const double4* pin = a.pos[pc];
  double4* pout = a.pos[pc ^ 1];
  const double4* min_ = a.mom[rc];
  double4* mout = a.mom[rc ^ 1];
  const int* ain = a.aid[rc];
  int* aout = a.aid[rc ^ 1];
  int r = block_replica(a.blk_rep);
  const RepDev& R = a.rep[r];
  int i = blockIdx.x * IMDG_BLOCK + threadIdx.x;
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
  // the block's tile for the list build (needs the cell ranges only)
  if (a.rst[r].need_rebuild && threadIdx.x < IMDG_WAVE)
    build_tile(a, blockIdx.x, tl);
}

// The tile of 128-slot block b of a rebuilt replica, built by one whole
// wavefront (lane l); the cells of the block's first and last slots are
// those the permutation gives them.
__device__ void build_tile(const BuildArgs& a, int b, TileLds& S) {
  int l = threadIdx.x % IMDG_WAVE;
  int r = __builtin_amdgcn_readfirstlane(a.blk_rep[b]);
  const RepDev& R = a.rep[r];
  Tile& T = a.tiles[b];
  int s0 = b * IMDG_BLOCK;
  int s1 = min(s0 + IMDG_BLOCK, R.atom_off + R.natoms) - 1;
  if (s1 < s0) {
    if (l == 0) {
      T.nrun = 0;
      T.natoms = 0;
      T.ok = 1;
    }
    return;
  }
  int g0 = R.gdim[0], g1 = R.gdim[1], g2 = R.gdim[2];
  int lc0 = a.cellid[a.perm[s0]] - R.cell_off;
  int lc1 = a.cellid[a.perm[s1]] - R.cell_off;
  int col0 = lc0 / g2, col1 = lc1 / g2, ncol = col1 - col0 + 1;
  bool ok = ncol <= TILE_MAXCOL;
  int nmine = 0, mb[2] = {0, 0}, me[2] = {0, 0};
  if (ok && l < 9 * ncol) {
    int col = col0 + l / 9;
    int z0 = (col == col0) ? lc0 % g2 : 0;
    int z1 = (col == col1) ? lc1 % g2 : g2 - 1;
    int cx = col / g1, cy = col % g1;
    int ox = (l % 9) / 3 - 1, oy = (l % 9) % 3 - 1;
    int nx = cx + ox, ny = cy + oy;
    bool valid = true;
    if (nx < 0 || nx >= g0) {
      if (R.pbc[0] != 1)
        valid = false;
      else
        nx = (nx + g0) % g0;
    }
    if (ny < 0 || ny >= g1) {
      if (R.pbc[1] != 1)
        valid = false;
      else
        ny = (ny + g1) % g1;
    }
    if (valid) {
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
        int cf = cbase + p[k], cl = cbase + q[k];
        int rb = a.cell_beg[cf], re = a.cell_beg[cl] + a.cell_num[cl];
        if (re > rb) {
          mb[nmine] = rb;
          me[nmine] = re;
          nmine++;
        }
      }
    }
  }
  // compact the ranges of all lanes into the wavefront's LDS rows
  int pre = nmine;
#pragma unroll
  for (int o = 1; o < IMDG_WAVE; o <<= 1) {
    int y = __shfl_up(pre, o, IMDG_WAVE);
    if (l >= o) pre += y;
  }
  int nr = __shfl(pre, IMDG_WAVE - 1, IMDG_WAVE);
  pre -= nmine;
  for (int k = 0; k < nmine; k++) {
    S.rb[pre + k] = mb[k];
    S.re[pre + k] = me[k];
  }
  wave_lds_sync();
  // sort the ranges by first slot (rank sort, one range per lane; ties by
  // position), then merge overlapping or touching ranges into runs: a run
  // starts where a range begins after the furthest end of all ranges before
  // it (wave prefix max of the ends)
  int* rb = S.rb;
  int* re = S.re;
  if (nr > IMDG_WAVE) ok = false;  // more ranges than lanes: no tile
  int vb = 0, ve = 0, rank = 0;
  bool mine = ok && l < nr;
  if (mine) {
    vb = rb[l];
    ve = re[l];
    for (int m = 0; m < nr; m++) {
      int bm = rb[m];
      rank += (bm < vb || (bm == vb && m < l)) ? 1 : 0;
    }
  }
  wave_lds_sync();
  if (mine) {
    rb[rank] = vb;
    re[rank] = ve;
  }
  wave_lds_sync();
  int sb = 0, se = 0;
  if (mine) {
    sb = rb[l];
    se = re[l];
  }
  // prefix max of the ends over the lanes before this one
  int pmax = mine ? se : -1;
#pragma unroll
  for (int o = 1; o < IMDG_WAVE; o <<= 1) {
    int y = __shfl_up(pmax, o, IMDG_WAVE);
    if (l >= o) pmax = max(pmax, y);
  }
  int prev = __shfl_up(pmax, 1, IMDG_WAVE);
  bool start = mine && (l == 0 || sb > prev);
  unsigned long long starts = __ballot(start);
  int nrun = __popcll(starts);
  // run of this lane: number of starts at or below it, minus one
  int runid = __popcll(starts & (~0ull >> (63 - l))) - 1;
  bool last_of_run =
      mine && (l == nr - 1 || (l < IMDG_WAVE - 1 && ((starts >> (l + 1)) & 1ull)));
  if (nrun > IMDG_MAXRUN) ok = false;
  // run begin (at its start lane) and end (the prefix max at its last lane)
  int4* run = S.run;
  if (ok && start) run[runid].x = sb;
  if (ok && last_of_run) run[runid].z = pmax;
  wave_lds_sync();
  // run lengths and their exclusive prefix sums (tile positions)
  int len = 0;
  if (ok && l < nrun) len = run[l].z - run[l].x;
  int pos = len;
#pragma unroll
  for (int o = 1; o < IMDG_WAVE; o <<= 1) {
    int y = __shfl_up(pos, o, IMDG_WAVE);
    if (l >= o) pos += y;
  }
  int tot = __shfl(pos, IMDG_WAVE - 1, IMDG_WAVE);
  pos -= len;
  ok = ok && tot <= a.tile_cap;
  if (ok && l < nrun) T.run[l] = make_int4(run[l].x, len, pos, 0);
  if (l == 0) {
    if (!ok) {
      // counted by k_list_done (one counter word updated by every block
      // would serialise the blocks)
      T.ok = 0;
      T.nrun = 0;
      T.natoms = 0;
    } else {
      T.ok = 1;
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

// FP32 screening band of the list build: a candidate whose FP32 r^2 from
// the cell-relative records lies within this relative distance of the list
// radius is decided by the reference's FP64 expression instead.  The FP32
// r^2 is off by less than 2e-6 relative (records and offsets rounded once,
// |u| below a cell diagonal), so everything outside the band is decided
// exactly as the FP64 expression would decide it.
#define IMDG_SCREEN_BAND 1.0e-4

// The reference's decision for one candidate: r^2 of (image of the partner)
// - (owner) in FP64 with separate roundings, owner being the atom whose
// half stencil holds the other's cell.  Rare (candidates near the radius),
// kept out of line.
__device__ __noinline__ bool exact_pair(const RepDev& R,
                                        const double4* __restrict__ pos,
                                        int i, int j, int sx, int sy, int sz,
                                        bool owner) {
#pragma clang fp contract(off)
  double4 xi = pos[i], xj = pos[j];
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
  return r2 < R.cellsz;
}

// Cell-relative position of slot j in FP32.
template <bool MIXED>
__device__ __forceinline__ float3 rec_xyz(const BuildArgs& a, int j) {
  if (MIXED) {
    const float* p = reinterpret_cast<const float*>(a.rec_a) + 4 * (size_t)j;
    return make_float3(p[0], p[1], p[2]);
  }
  const double* p = reinterpret_cast<const double*>(a.rec_a) + 4 * (size_t)j;
  return make_float3((float)p[0], (float)p[1], (float)p[2]);
}

// What one lane carries through the candidate loops of its atom.
struct ListLane {
  int i;
  float3 ui;
  float lo, hi;
  char* lst;  // the lane's first entry (group 0) in the chunk's format
  int cnt, kmax;
  int cx, cy, cz;  // the atom's cell (image shifts of the FP64 decision)
  uint32_t g0, g1, g2, g3;  // entries of the group being filled
};

// Append entry e.  The four entries of a group (slots m%4 of group m/4)
// are collected in registers and written with one store when the group is
// full: 8 B in the 16-bit format (code << 11 | slot - base of the code;
// groups of 64 lanes 512 B apart), 16 B in the 32-bit one (slot | code <<
// 27; 1 KiB apart).  One scattered 2- or 4-byte store per entry would cost
// an L2 request each.  Entries past kmax are only counted: the list has
// overflowed and is built again, longer.
template <bool W16>
__device__ __forceinline__ void list_put(ListLane& L, uint32_t e) {
  if (W16) {
    // shift e in at the top of the 64-bit register (g1:g0): after four
    // appends g0 = e0 | e1 << 16, g1 = e2 | e3 << 16
    L.g0 = __builtin_amdgcn_alignbit(L.g1, L.g0, 16);
    L.g1 = __builtin_amdgcn_alignbit(e, L.g1, 16);
  } else {
    L.g0 = L.g1;
    L.g1 = L.g2;
    L.g2 = L.g3;
    L.g3 = e;
  }
  L.cnt++;
  if ((L.cnt & 3) == 0 && L.cnt <= L.kmax) {
    const unsigned g = (unsigned)(L.cnt >> 2) - 1u;
    if (W16)
      *reinterpret_cast<uint2*>(L.lst + (g << 9)) = make_uint2(L.g0, L.g1);
    else
      *reinterpret_cast<uint4*>(L.lst + (g << 10)) =
          make_uint4(L.g0, L.g1, L.g2, L.g3);
  }
}

// Image shift of neighbour cell offset o along an axis: -1 / +1 when the
// cell index wraps below 0 / beyond the grid.
__device__ __forceinline__ int wrap_shift(int c, int o, int g) {
  int n = c + o;
  return n < 0 ? -1 : (n >= g ? 1 : 0);
}

// The block's LDS tile of FP32 positions, addressed as LDS (32-bit
// addresses; a generic pointer would be indexed in 64 bits).
typedef __attribute__((address_space(3))) const float3 lds_float3;
typedef __attribute__((address_space(3))) const char lds_char;
typedef __attribute__((address_space(3))) const float lds_float;

// Candidates of one neighbour cell, four at a time (their records in
// flight together): FP32 screen, FP64 decision inside the band (one branch
// per four candidates; the band is rarely hit), append.  The separation is
// u_j + (o - u_i), with o - u_i formed once per cell: two roundings of a
// few Angstrom, far inside the screening band.  Candidate t of the cell is
// entry ebase + t.  SELF: the atom's own cell (code 13), where the atom is
// a candidate of itself.
template <bool MIXED, bool TOK, bool W16, bool SELF>
__device__ __forceinline__ void list_cell(
    const BuildArgs& a, const RepDev& R, const double4* __restrict__ pos,
    lds_float3* trec, ListLane& L, int jb, int jn, int tb, float o0,
    float o1, float o2, bool owner, int code, uint32_t ebase) {
  const float w0 = o0 - L.ui.x, w1 = o1 - L.ui.y, w2 = o2 - L.ui.z;
  // the cell's run in the tile: four records per pass at immediate offsets
  // from one address (up to three records past the run: inside the tile
  // array of tile_cap + 4 records, and masked out below)
  lds_char* run = (lds_char*)trec + __umul24((unsigned)tb, 12u);
  for (int t0 = 0; t0 < jn; t0 += 4) {
    float3 uj[4];
    lds_float* q = (lds_float*)(run + __umul24((unsigned)t0, 12u));
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (TOK) {
        uj[u] = make_float3(q[3 * u], q[3 * u + 1], q[3 * u + 2]);
      } else {
        uj[u] = rec_xyz<MIXED>(a, jb + min(t0 + u, jn - 1));
      }
    }
    bool in[4], band[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int j = jb + t0 + u;
      bool valid = (t0 + u < jn) && !(SELF && j == L.i);
      float d0 = uj[u].x + w0;
      float d1 = uj[u].y + w1;
      float d2 = uj[u].z + w2;
      float r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const bool inside = r2 < L.lo, near = r2 < L.hi;
      in[u] = valid & inside;
      band[u] = (valid & near) ^ in[u];  // near but not inside
    }
    if (band[0] || band[1] || band[2] || band[3]) {
      atomicAdd(&a.ctl->exact_pairs,
                (int)band[0] + (int)band[1] + (int)band[2] + (int)band[3]);
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (band[u])
          in[u] = exact_pair(R, pos, L.i, jb + t0 + u,
                             wrap_shift(L.cx, code / 9 - 1, R.gdim[0]),
                             wrap_shift(L.cy, (code / 3) % 3 - 1, R.gdim[1]),
                             wrap_shift(L.cz, code % 3 - 1, R.gdim[2]),
                             owner);
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (in[u]) list_put<W16>(L, ebase + (uint32_t)(t0 + u));
  }
}

// The candidate loops of one atom: the 9 neighbour columns (ox, oy) in a
// rolled loop, the 3 cells of a column (oz) unrolled, code
// k = (ox+1)*9 + (oy+1)*3 + (oz+1).  First slot and count of a column's
// cells are loaded together; a cell beyond an open face has count 0.  Tile
// positions: a search for the column's centre cell; its z neighbours
// follow it in slot order (cells are sorted z fastest) and so in the tile,
// whose runs are laid out in slot order without gaps over the block's
// neighbour cells, unless the column wraps through a periodic z face.  cb:
// the chunk's base slot per code (16-bit format).
template <bool MIXED, bool TOK, bool W16>
__device__ __forceinline__ void list_atom(const BuildArgs& a, const RepDev& R,
                                          const double4* __restrict__ pos,
                                          lds_float3* trec,
                                          const int4* truns, int nrun,
                                          const float* off, const int* cb,
                                          ListLane& L) {
  const int g0 = R.gdim[0], g1 = R.gdim[1], g2 = R.gdim[2];
  const bool p0 = R.pbc[0] == 1, p1 = R.pbc[1] == 1, p2 = R.pbc[2] == 1;
#pragma unroll 1
  for (int col = 0; col < 9; col++) {
    int n0 = L.cx + col / 3 - 1, n1 = L.cy + col % 3 - 1;
    const int sx = n0 < 0 ? -1 : (n0 >= g0 ? 1 : 0);
    const int sy = n1 < 0 ? -1 : (n1 >= g1 ? 1 : 0);
    const bool cvalid = ((sx == 0) | p0) & ((sy == 0) | p1);
    n0 -= sx * g0;
    n1 -= sy * g1;
    const int cbase = R.cell_off + (n0 * g1 + n1) * g2;
    int jb[3], jn[3], tb[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      int n2 = L.cz + q - 1;
      const int sz = n2 < 0 ? -1 : (n2 >= g2 ? 1 : 0);
      const bool valid = cvalid & ((sz == 0) | p2);
      n2 -= sz * g2;
      const int nc = valid ? cbase + n2 : R.cell_off;
      jb[q] = a.cell_beg[nc];
      jn[q] = valid ? a.cell_num[nc] : 0;
      tb[q] = 0;
    }
    if (TOK) {
      tb[1] = tile_index(truns, nrun, jb[1]);
#pragma unroll
      for (int q = 0; q < 3; q += 2) {
        const int n2 = L.cz + q - 1;
        if (n2 < 0 || n2 >= g2)
          tb[q] = jn[q] > 0 ? tile_index(truns, nrun, jb[q]) : 0;
        else
          tb[q] = tb[1] + (jb[q] - jb[1]);
      }
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int code = col * 3 + q;
      const uint32_t ebase =
          W16 ? ((uint32_t)code << IMDG_K16_BITS) | (uint32_t)(jb[q] - cb[code])
              : (uint32_t)jb[q] | ((uint32_t)code << IMDG_IDX_BITS);
      if (q == 1 && col == 4)  // code 13 (q is a constant here)
        list_cell<MIXED, TOK, W16, true>(
            a, R, pos, trec, L, jb[q], jn[q], tb[q], off[3 * code],
            off[3 * code + 1], off[3 * code + 2], true, code, ebase);
      else
        list_cell<MIXED, TOK, W16, false>(
            a, R, pos, trec, L, jb[q], jn[q], tb[q], off[3 * code],
            off[3 * code + 1], off[3 * code + 2], code >= 13, code, ebase);
    }
  }
  // pad to a multiple of IMDG_LIST_PAD entries with far-away copies of the
  // atom itself (kmax is a multiple of IMDG_LIST_PAD, so the padded list
  // fits whenever the list does; the padding completes the last group)
  const int cnt = L.cnt;
  if (cnt <= L.kmax) {
    const uint32_t pad =
        W16 ? ((uint32_t)IMDG_CODE_PAD << IMDG_K16_BITS) |
                  (uint32_t)(L.i - cb[IMDG_CODE_PAD])
            : (uint32_t)L.i | ((uint32_t)IMDG_CODE_PAD << IMDG_IDX_BITS);
    const int mend = ((cnt + IMDG_LIST_PAD - 1) / IMDG_LIST_PAD) * IMDG_LIST_PAD;
    for (int m = cnt; m < mend; m++) list_put<W16>(L, pad);
    L.cnt = cnt;
  }
}

// Full neighbour list.  For each cell offset o of the 27, the reference
// evaluates a pair from the atom whose cell has the other in its half
// stencil (o lexicographically positive: the stencil of
// src/imd_geom_3d.c:895-899) as (image of the partner) - (owner); the same
// expression decides membership here (directly, or after the FP32 screen
// for candidates far from the list radius), so each unordered pair is in
// the list of both atoms exactly when the reference has it in its half
// list.  One block per 128-slot block; the FP32 positions of its tile are
// staged in LDS (12 B each, tile_cap + 4 records of dynamic LDS: the cap
// trades the blocks that fall back to loads from HBM/L2 against the blocks
// resident per CU).  The format of a chunk (wavefront) is
// uniform: 16-bit entries when every neighbour cell run of every code lies
// within IMDG_K16_SPAN slots of the chunk's base for that code.
template <bool MIXED>
__global__ void __launch_bounds__(IMDG_BLOCK) k_list(BuildArgs a) {
  extern __shared__ __align__(16) unsigned char list_lds[];
  float3* trec = reinterpret_cast<float3*>(list_lds);
  __shared__ int4 truns[IMDG_MAXRUN];
  __shared__ float off[3 * IMDG_NCODES];
  __shared__ int tinfo[2];
  __shared__ int wmax[IMDG_NWAVE];
  __shared__ int cbase[IMDG_NWAVE][IMDG_META];
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  int b = blockIdx.x;
  int r = a.blk_rep[b];
  const RepDev& R = a.rep[r];
  bool flagged = a.rst[r].need_rebuild != 0;
  // the permuted positions (k_list_done flips the buffers afterwards)
  const double4* pos = a.pos[a.ctl->pos_cur ^ 1];
  int kmax = a.kmax, seen = 0;
  if (flagged) {
    const Tile& T = a.tiles[b];
    if (threadIdx.x == 0) {
      tinfo[0] = T.nrun;
      tinfo[1] = T.ok;
    }
    for (int k = threadIdx.x; k < T.nrun; k += blockDim.x) truns[k] = T.run[k];
    for (int k = threadIdx.x; k < 3 * IMDG_NCODES; k += blockDim.x)
      off[k] = (float)((&R.offs[0][0])[k]);
    __syncthreads();
    const int nrun = tinfo[0], tok = tinfo[1];
    if (tok) {
      int tot = T.natoms;
      // all of a thread's loads in flight together: 8 records per pass
      int k = 0;
      for (int t0 = threadIdx.x; t0 < tot; t0 += 8 * IMDG_BLOCK) {
        float3 v[8];
        int t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          t[u] = t0 + u * IMDG_BLOCK;
          if (t[u] < tot) {
            while (t[u] >= truns[k].z + truns[k].y) k++;
            v[u] = rec_xyz<MIXED>(a, truns[k].x + (t[u] - truns[k].z));
          }
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
          if (t[u] < tot) trec[t[u]] = v[u];
      }
    }
    __syncthreads();
    const int i = b * IMDG_BLOCK + threadIdx.x;
    const bool mine = i - R.atom_off < R.natoms;
    int cx = 0, cy = 0, cz = 0;
    if (mine) cell_coords(R, a.acell[i], cx, cy, cz);
    // 16-bit format: the chunk's (wavefront's) base slot per code, and
    // whether every neighbour cell run lies within IMDG_K16_SPAN slots of
    // it (a lane without an atom has no runs)
    bool w16 = false;
    const int w = threadIdx.x / IMDG_WAVE, lane = threadIdx.x & 63;
    if (a.list16) {
      const int g0 = R.gdim[0], g1 = R.gdim[1], g2 = R.gdim[2];
      const bool p0 = R.pbc[0] == 1, p1 = R.pbc[1] == 1, p2 = R.pbc[2] == 1;
      const int chunk = i >> 6;
      bool bad = false;
#pragma unroll
      for (int k = 0; k < 27; k++) {
        int n0 = cx + k / 9 - 1, n1 = cy + (k / 3) % 3 - 1, n2 = cz + k % 3 - 1;
        int sx = n0 < 0 ? -1 : (n0 >= g0 ? 1 : 0);
        int sy = n1 < 0 ? -1 : (n1 >= g1 ? 1 : 0);
        int sz = n2 < 0 ? -1 : (n2 >= g2 ? 1 : 0);
        bool valid = mine & ((sx == 0) | p0) & ((sy == 0) | p1) & ((sz == 0) | p2);
        int nc = valid ? R.cell_off + ((n0 - sx * g0) * g1 + n1 - sy * g1) * g2 +
                             n2 - sz * g2
                       : R.cell_off;
        int jb = a.cell_beg[nc], jn = valid ? a.cell_num[nc] : 0;
        int lo = wave_min_i(jn > 0 ? jb : 0x7fffffff);
        if (lo == 0x7fffffff) lo = 0;
        bad |= jn > 0 && (jb < lo || jb + jn - lo > IMDG_K16_SPAN);
        if (lane == 0) cbase[w][k] = lo;
      }
      const int wide = __ballot(bad) != 0ull ? 1 : 0;
      if (lane == 0) {
        cbase[w][IMDG_CODE_PAD] = chunk * IMDG_CHUNK;
        cbase[w][IMDG_META_WIDE] = wide;
      }
      wave_lds_sync();
      if (lane <= IMDG_META_WIDE)
        a.cmeta[(size_t)chunk * IMDG_META + lane] = cbase[w][lane];
      w16 = __builtin_amdgcn_readfirstlane(wide) == 0;
    }
    if (mine) {
      ListLane L;
      L.i = i;
      L.ui = rec_xyz<MIXED>(a, i);
      const float cs = (float)R.cellsz;
      L.lo = cs * (float)(1.0 - IMDG_SCREEN_BAND);
      L.hi = cs * (float)(1.0 + IMDG_SCREEN_BAND);
      L.cnt = 0;
      L.kmax = kmax;
      L.g0 = L.g1 = L.g2 = L.g3 = 0u;
      L.cx = cx;
      L.cy = cy;
      L.cz = cz;
      const int* cb = cbase[w];
      if (w16) {
        L.lst = const_cast<char*>(list16_of(a.nbr, i, kmax));
        if (tok)
          list_atom<MIXED, true, true>(a, R, pos, (lds_float3*)trec, truns,
                                       nrun, off, cb, L);
        else
          list_atom<MIXED, false, true>(a, R, pos, (lds_float3*)trec, truns,
                                        nrun, off, cb, L);
      } else {
        L.lst = reinterpret_cast<char*>(
            a.nbr + (((size_t)(i >> 6) * (kmax >> 2)) * IMDG_WAVE + (i & 63)) *
                        IMDG_GROUP);
        if (tok)
          list_atom<MIXED, true, false>(a, R, pos, (lds_float3*)trec, truns,
                                        nrun, off, cb, L);
        else
          list_atom<MIXED, false, false>(a, R, pos, (lds_float3*)trec, truns,
                                         nrun, off, cb, L);
      }
      a.nnb[i] = L.cnt < kmax ? L.cnt : kmax;
      seen = L.cnt;
    }
  }
    // block max of the list length (k_list_done reduces the blocks: one
  // atomic word updated by every block would serialise them)
  for (int o = 32; o > 0; o >>= 1)
    seen = max(seen, __shfl_xor(seen, o, IMDG_WAVE));
  if ((threadIdx.x % IMDG_WAVE) == 0) wmax[threadIdx.x / IMDG_WAVE] = seen;
  __syncthreads();
  // with bit 30: the block's tile fell back to HBM/L2 loads
  if (threadIdx.x == 0) {
    int m = wmax[0];
    for (int k = 1; k < IMDG_NWAVE; k++) m = max(m, wmax[k]);
    a.blkmax[b] = m | (flagged && !tinfo[1] ? IMDG_BLK_FALLBACK : 0);
  }
}

// End of a rebuild (one block): the longest list; on an overflow freeze the
// state (the host grows the list and runs the chain again from the old
// buffers), otherwise clear the rebuild flags and flip the buffers that
// k_permute filled.
__global__ void __launch_bounds__(1024) k_list_done(BuildArgs a) {
  __shared__ int wmax[1024 / IMDG_WAVE], wfb[1024 / IMDG_WAVE];
  if (!a.ctl->any_rebuild || a.ctl->frozen) return;
  int m = 0, fb = 0;
  for (int b = threadIdx.x; b < a.nblk; b += blockDim.x)
    if (a.rst[a.blk_rep[b]].need_rebuild) {
      const int v = a.blkmax[b];
      m = max(m, v & ~IMDG_BLK_FALLBACK);
      fb += v & IMDG_BLK_FALLBACK ? 1 : 0;
    }
  for (int o = 32; o > 0; o >>= 1) {
    m = max(m, __shfl_xor(m, o, IMDG_WAVE));
    fb += __shfl_xor(fb, o, IMDG_WAVE);
  }
  if ((threadIdx.x % IMDG_WAVE) == 0) {
    wmax[threadIdx.x / IMDG_WAVE] = m;
    wfb[threadIdx.x / IMDG_WAVE] = fb;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 1; k < (int)(blockDim.x / IMDG_WAVE); k++) {
    m = max(m, wmax[k]);
    fb += wfb[k];
  }
  m = max(m, wmax[0]);
  a.ctl->tiles_fallback += fb;
  if (m > a.ctl->kmax_seen) a.ctl->kmax_seen = m;
  if (m > a.kmax) {
    // keep the rebuild flags: the host grows the list and builds again
    a.ctl->overflow = 1;
    a.ctl->frozen = 1;
  } else {
    for (int q = 0; q < a.nrep; q++)
      if (a.rst[q].need_rebuild) {
        a.rst[q].rebuilt = 1;
        a.rst[q].nbl_count += 1;
        a.rst[q].need_rebuild = 0;
      }
    a.ctl->pos_cur ^= 1;
    a.ctl->rec_cur ^= 1;
    a.ctl->any_rebuild = 0;
  }
}

// check_nblist alone: max displacement^2 per block of IMDG_BLOCK slots.
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
  double v = 0.0;
  if ((i - Rp.atom_off) < Rp.natoms) v = disp2(pos[i], xnbl[i]);
  v = wave_max(v);
  if ((threadIdx.x % IMDG_WAVE) == 0) red[threadIdx.x / IMDG_WAVE] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int k = 1; k < IMDG_NWAVE; k++) m = fmax(m, red[k]);
    maxd[lb] = m;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

void launch_build(const BuildArgs& b, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_bin, dim3(b.nblk), dim3(IMDG_BLOCK), 0, st, b);
  hipLaunchKernelGGL(k_scan, dim3(b.nstile), dim3(IMDG_SCAN_TILE), 0, st, b);
  hipLaunchKernelGGL(k_place, dim3(b.nblk), dim3(IMDG_BLOCK), 0, st, b);
  hipLaunchKernelGGL(k_cellsort, dim3(grid), dim3(256), 0, st, b);
  hipLaunchKernelGGL(k_permute, dim3(b.nblk), dim3(IMDG_BLOCK), 0, st, b);
  const size_t tile_lds = (size_t)(b.tile_cap + 4) * sizeof(float3);
  if (b.mixed)
    hipLaunchKernelGGL((k_list<true>), dim3(b.nblk), dim3(IMDG_BLOCK),
                       tile_lds, st, b);
  else
    hipLaunchKernelGGL((k_list<false>), dim3(b.nblk), dim3(IMDG_BLOCK),
                       tile_lds, st, b);
  hipLaunchKernelGGL(k_list_done, dim3(1), dim3(1024), 0, st, b);
}

void launch_check(const RepDev* rep, Ctl* ctl, const int* blk_rep,
                  const double4* pos0, const double4* pos1,
                  const double4* xnbl, double* maxd, int nblk,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_check, dim3(nblk), dim3(IMDG_BLOCK), 0, st, rep, ctl,
                     blk_rep, pos0, pos1, xnbl, maxd);
}

}  // namespace imdg
