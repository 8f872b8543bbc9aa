// worklist.h -- device enumeration of the flattened cell-pair work list,
// shared by the pair and triplet kernels.
#pragma once
#include "pcf_internal.h"

namespace pcf {

// ---------------------------------------------------------------------------
// Work list
// ---------------------------------------------------------------------------
// Tiles are 64 consecutive points of one (cx,cy) column (they may span a few
// z-cells), so every lane of a wave holds a point; a tile's j-range in a
// neighbour column covers the z-cells within reach of any of its cells.  The
// tiles whose first point lies in cell c are emitted by S threads (cell, s):
// S = 1 takes every neighbour column, S = 2Kx+1 one x offset (dx = s - Kx),
// S = (2Kx+1)(2Ky+1) one column (dx, dy), so the dependent index loads of a
// cell spread over S threads instead of running as one serial chain.  The
// list is ordered by (cell, s, tile, remaining offsets, z segment).
template <bool FILL>
__device__ inline int enumerate_cell(const WorkListArgs& a, int cell, int s, int S,
                                     PairItem* out) {
  const int ny = 2 * a.K[1] + 1;
  int dxl = -a.K[0], dxh = a.K[0], dyl = -a.K[1], dyh = a.K[1];
  if (S > 1) {
    if (S == (2 * a.K[0] + 1) * ny) {
      dxl = dxh = s / ny - a.K[0];
      dyl = dyh = s % ny - a.K[1];
    } else {
      dxl = dxh = s - a.K[0];
    }
  }
  const int ib = a.x_cell_start[cell];
  const int ie = a.x_cell_start[cell + 1];
  if (ie <= ib) return 0;
  const GridGeom& g = a.g;
  const int Pz = g.P[2], Py = g.P[1], Px = g.P[0];
  const int cz = cell % Pz;
  const int cy = (cell / Pz) % Py;
  const int cx = cell / (Pz * Py);
  const int colbase = cell - cz;
  const int col_start = a.x_cell_start[colbase];
  const int col_end = a.x_cell_start[colbase + Pz];
  const float reach2 = a.reach * a.reach;
  int count = 0;
  // First tile boundary at or after ib (tiles start at col_start + 64k).
  int i0 = col_start + ((ib - col_start + kTile - 1) / kTile) * kTile;
  for (; i0 < ie; i0 += kTile) {
    const int ni = min(kTile, col_end - i0);
    // The ends of cells cz .. cz + kCellWin - 1 of the column, read together
    // (independent loads, not a chain of dependent ones): the cell of a
    // tile point p >= i0 is cz plus the number of those ends <= p, while
    // that count stays below kCellWin (beyond it the walk reads on).
    constexpr int kCellWin = 8;
    int cend[kCellWin];
#pragma unroll
    for (int t = 0; t < kCellWin; ++t)
      cend[t] = cz + t + 1 < Pz ? a.x_cell_start[colbase + cz + t + 1] : 0x7fffffff;
    auto cell_of_point = [&](int p) {
      int c = cz;
#pragma unroll
      for (int t = 0; t < kCellWin; ++t) c += cend[t] <= p ? 1 : 0;
      while (c + 1 < Pz && c - cz >= kCellWin && a.x_cell_start[colbase + c + 1] <= p) ++c;
      return c;
    };
    // z-cell of the tile's last point.
    const int zb = cell_of_point(i0 + ni - 1);
    // z cells of the first and last point of each 16-point block (-1: no
    // such block), for the items' step windows; with sub-slab grids also
    // their sub-slab keys (the column's points are sorted by key, so the
    // block's points all lie between the two).
    int bcf[4], bcl[4], bkf[4], bkl[4];
    if (FILL) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int f = i0 + 16 * b;
        bkf[b] = bkl[b] = 0;
        if (f >= i0 + ni) {
          bcf[b] = bcl[b] = -1;
          continue;
        }
        const int l = min(f + 15, i0 + ni - 1);
        bcf[b] = cell_of_point(f);
        bcl[b] = cell_of_point(l);
        if (a.y_fine_start) {
          bkf[b] = zkey(a.x_pts[f].z, g, a.zq);
          bkl[b] = zkey(a.x_pts[l].z, g, a.zq);
        }
      }
    }
    for (int dx = dxl; dx <= dxh; ++dx) {
      int ccx = cx + dx, kx = 0;
      if (ccx < 0) {
        if (!g.bpc) continue;
        ccx += Px;
        kx = -1;
      } else if (ccx >= Px) {
        if (!g.bpc) continue;
        ccx -= Px;
        kx = 1;
      }
      const float gx = (float)max(0, abs(dx) - 1) * g.cs[0];
      for (int dy = dyl; dy <= dyh; ++dy) {
        int ccy = cy + dy, ky = 0;
        if (ccy < 0) {
          if (!g.bpc) continue;
          ccy += Py;
          ky = -1;
        } else if (ccy >= Py) {
          if (!g.bpc) continue;
          ccy -= Py;
          ky = 1;
        }
        const float gy = (float)max(0, abs(dy) - 1) * g.cs[1];
        const float rho2 = gx * gx + gy * gy;
        if (rho2 >= reach2) continue;
        // Sphere: the z reach shrinks with the lateral gap.  Cylinder (2ani):
        // the z and perpendicular ranges are independent.
        const float hz = a.cylinder ? a.reach : sqrtf(reach2 - rho2);
        const int Kz = min(a.K[2], (int)ceilf(hz * g.inv_cs[2]));
        // The same reach in sub-slabs, plus one for the rounding of the keys.
        const int Kf = (int)ceilf(hz * 1.00001f * g.inv_cs[2] * (float)a.zq) + 1;
        const int col = (ccx * Py + ccy) * Pz;
        const int zlo = cz - Kz, zhi = zb + Kz;
        // Up to three z segments: wrapped below (kz=-1), inside, wrapped above.
        // zlo > -Pz and zhi < 2*Pz because Kz < Pz.
        int seg_lo[3], seg_hi[3], seg_k[3], nseg = 0;
        if (g.bpc) {
          if (zlo < 0) { seg_lo[nseg] = zlo + Pz; seg_hi[nseg] = Pz - 1; seg_k[nseg++] = -1; }
          seg_lo[nseg] = max(zlo, 0); seg_hi[nseg] = min(zhi, Pz - 1); seg_k[nseg++] = 0;
          if (zhi > Pz - 1) { seg_lo[nseg] = 0; seg_hi[nseg] = zhi - Pz; seg_k[nseg++] = 1; }
        } else {
          seg_lo[0] = max(zlo, 0); seg_hi[0] = min(zhi, Pz - 1); seg_k[0] = 0; nseg = 1;
        }
        for (int s = 0; s < nseg; ++s) {
          if (seg_hi[s] < seg_lo[s]) continue;
          int j0 = a.y_cell_start[col + seg_lo[s]];
          int j1 = a.y_cell_start[col + seg_hi[s] + 1];
          if (j1 <= j0) continue;
          int flag = 0;
          if (a.auto_) {
            // Count each unordered pair once: j > i in cell order.
            if (j1 <= i0 + 1) continue;
            if (j0 < i0 + ni) {
              j0 = max(j0, i0 + 1);
              flag = kCodeMask;
            }
          }
          if (a.map) {
            j0 = a.map[j0];
            j1 = a.map[j1];
            if (j1 <= j0) continue;
          }
          if (FILL) {
            PairItem it;
            it.i0 = i0;
            it.ni = ni;
            it.j0 = j0;
            it.nj = j1 - j0;
            it.code = ((kx + 1) * 9 + (ky + 1) * 3 + (seg_k[s] + 1)) | flag;
            // Step windows: the z cells within Kz of block b's cells (the
            // tile's own rule, per block), in this segment's frame; with
            // sub-slab grids the sub-slabs within Kf of the block's.
            const int nsteps = (it.nj + 15) >> 4;
            it.ws = it.we = 0u;
            if (nsteps > 255) {
              it.code |= kCodeNoWin;
            } else {
              for (int b = 0; b < 4; ++b) {
                if (bcf[b] < 0) continue;
                int w0, w1;
                if (a.y_fine_start) {
                  const int zq = a.zq;
                  const int lo = max(seg_lo[s] * zq, bkf[b] - Kf - seg_k[s] * Pz * zq);
                  const int hi = min(seg_hi[s] * zq + zq - 1, bkl[b] + Kf - seg_k[s] * Pz * zq);
                  if (lo > hi) continue;
                  w0 = a.y_fine_start[col * zq + lo];
                  w1 = a.y_fine_start[col * zq + hi + 1];
                } else {
                  const int lo = max(seg_lo[s], bcf[b] - Kz - seg_k[s] * Pz);
                  const int hi = min(seg_hi[s], bcl[b] + Kz - seg_k[s] * Pz);
                  if (lo > hi) continue;
                  w0 = a.y_cell_start[col + lo];
                  w1 = a.y_cell_start[col + hi + 1];
                }
                if (a.auto_) w0 = max(w0, i0 + 1);
                if (w1 <= w0) continue;
                if (a.map) {
                  w0 = a.map[w0];
                  w1 = a.map[w1];
                }
                const int sb = min(max((w0 - j0) >> 4, 0), nsteps);
                const int eb = min(max((w1 - j0 + 15) >> 4, 0), nsteps);
                if (eb <= sb) continue;
                it.ws |= (unsigned)sb << (8 * b);
                it.we |= (unsigned)eb << (8 * b);
              }
            }
            out[count] = it;
          }
          ++count;
        }
      }
    }
  }
  return count;
}


// Threads per cell (S above) for a work list over `cells` cells in total (all
// jobs of a batch): the finest split whose count array stays below 2^23
// entries.
inline int worklist_split(const WorkListArgs& a, int64_t cells) {
  const int nx = 2 * a.K[0] + 1, ny = 2 * a.K[1] + 1;
  const int64_t lim = (int64_t)1 << 23;
  if (cells * nx * ny <= lim) return nx * ny;
  if (cells * nx <= lim) return nx;
  return 1;
}

}  // namespace pcf
